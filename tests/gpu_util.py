"""Device-memory helpers of the GPU tests (torch is only the allocator)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def require_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def dev(a, dtype=None):
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(a).cuda()


def P(t):
    return None if t is None else t.data_ptr()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def empty(shape, dtype=torch.float64):
    return torch.empty(shape, dtype=dtype, device="cuda")


def zeros(shape, dtype=torch.float64):
    return torch.zeros(shape, dtype=dtype, device="cuda")


def relerr(a, b):
    """max|a-b| / max|b| over finite entries; inf if the non-finite patterns differ."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    fa, fb = np.isfinite(a), np.isfinite(b)
    if not np.array_equal(fa, fb) or not np.array_equal(a[~fb], b[~fb]):
        return np.inf
    if not fb.any():
        return 0.0
    scale = max(np.abs(b[fb]).max(), 1e-300)
    return np.abs(a[fb] - b[fb]).max() / scale
