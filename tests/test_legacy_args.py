"""Every legacy wrapper checks the arrays its kernel always dereferences and
returns HS_EINVAL before any launch when one is NULL (masks stay nullable:
NULL means uniform, as in the reference).  Host-side only: the other
operands are fake addresses that must never reach a kernel, so this test
runs only where no GPU is present (there a launch attempt would report a
HIP error code instead of HS_EINVAL, and the test would fail)."""
import ctypes as C
import os

import pytest

from hotspin_amd import _lib

HS_EINVAL = 10001
F = 0x1000  # fake device address: only ever compared with NULL


def _no_gpu():
    try:
        import torch
        return not torch.cuda.is_available()
    except Exception:
        return True


pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH) or not _no_gpu(),
                                reason="host-side check: needs the built library and no GPU")

# (name, arguments as a list with 'P' = a required pointer, other values literal)
D = 1.0
CASES = [
    ("exchange6Async", ["P"] * 6 + [None, None, 2, 2, 2, 0, 0, 0, D, D, D, D, None]),
    ("llbar_torque_async", ["P"] * 9 + [None, None, 8]),
    ("llbar_local00nc_async", ["P"] * 6 + [None, None, None, None, D, D, D, None, 8]),
    ("llbar_local02nc_async", ["P"] * 9 + [None, None, None, None, D, D, D, None, 8]),
    ("llbar_local02c_async", ["P"] * 9 + [None, None, None, None, D, D, D, None, 8]),
    ("long_field_async", ["P"] * 6 + [None] * 5 + [D] * 5 + [8, None]),
    ("temperature_scaleAnizNoise", ["P"] * 3 + [None] * 5 + [D] * 4 + [None, 8]),
    ("normalizeAsync", ["P"] * 3 + [None, 8]),
    ("QinterAsync", ["P"] * 3 + [None, D, 8, None]),
    ("addAsync", ["P"] * 3 + [None, 8]),
    ("linearCombination3Async", ["P", "P", D, "P", D, "P", D, None, 8]),
    ("vecMaddAsync", ["P"] * 6 + [None, None, None, D, D, D, None, 8]),
    ("dotAsync", ["P"] * 7 + [None, 8]),
    ("tensSYMMVecMul", ["P"] * 6 + [None] * 6 + [D, 2, 2, 2, None]),
    ("copyPad3DAsync", ["P", 4, 4, 4, "P", 2, 2, 2, 3, None]),
    ("partialMaxDiffAsync", ["P", "P", "P", 4, 256, 8, None]),
    ("partialMaxNorm3SqDiffAsync", ["P"] * 7 + [4, 256, 8, None]),
]


@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_null_required_operand_is_refused_before_launch(name, args):
    L = _lib.lib()
    f = getattr(L, name)
    slots = [i for i, a in enumerate(args) if a == "P"]
    for hole in slots:
        call = [(None if i == hole else F) if a == "P" else a for i, a in enumerate(args)]
        L.hs_clear_error()
        f(*call)
        code = L.hs_last_error()
        msg = L.hs_last_error_string().decode()
        L.hs_clear_error()
        assert code == HS_EINVAL, (name, hole, code, msg)


def test_empty_calls_accept_null():
    """N = 0 launches nothing, so NULL operands are fine."""
    L = _lib.lib()
    L.hs_clear_error()
    L.addAsync(None, None, None, None, 0)
    L.llbar_torque_async(*([None] * 11), 0)
    assert L.hs_last_error() == 0


def test_demag_plan_constructors_refuse_a_bad_mesh():
    """hs_demag_create and hs_demag_create_newell return NULL with HS_EINVAL
    for a missing or empty mesh before any device call."""
    import ctypes as C
    L = _lib.lib()
    cell = (C.c_double * 3)(2e-9, 2e-9, 2e-9)
    pbc = (C.c_int * 3)(0, 0, 0)
    for bad in (None, (C.c_int * 3)(4, 0, 8)):
        L.hs_clear_error()
        assert not L.hs_demag_create(bad, cell, pbc, 6, None)
        assert L.hs_last_error() == HS_EINVAL
        L.hs_clear_error()
        assert not L.hs_demag_create_newell(bad, cell, pbc, None)
        assert L.hs_last_error() == HS_EINVAL
    L.hs_clear_error()
