"""hotspin_amd: MI355X-native hot path of hotspin (LLB-r RK step incl. demag).

The compute lives in libhotspin_amd.so (hand-written HIP for gfx950) behind
the C ABI of include/libhotspin.h and include/hotspin_amd.h; this package is
the host-side loader and interface.
"""
from ._lib import HotspinError, lib  # noqa: F401

__all__ = ["HotspinError", "lib"]
