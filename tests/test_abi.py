"""The C-ABI library loads and exports every symbol include/*.h declares (no GPU calls)."""
import os
import subprocess

import pytest

from hotspin_amd import _lib

LEGACY = [
    "exchange6Async", "llbar_torque_async", "llbar_local00nc_async", "llbar_local02nc_async",
    "llbar_local02c_async", "long_field_async", "temperature_scaleAnizNoise", "brillouinAsync",
    "normalizeAsync", "QinterAsync", "Qspat_async", "addAsync", "linearCombination2Async",
    "linearCombination3Async", "addMaddAsync", "madd1Async", "madd2Async", "maddAsync", "vecMaddAsync",
    "cmaddAsync", "mulAsync", "tensSYMMVecMul", "divAsync", "dotMaskAsync", "dotAsync", "dotSignAsync",
    "copyPad3DAsync", "copyUnPad3DAsync", "zeroArrayAsync", "partialSumAsync", "partialSDotAsync",
    "partialMaxAsync", "partialMinAsync", "partialMaxAbsAsync", "partialMaxDiffAsync", "partialMaxSumAsync",
    "partialMaxNorm3SqAsync", "partialMaxNorm3SqDiffAsync",
]


def test_headers_declare_the_reference_symbols():
    d = _lib.declarations()
    for name in LEGACY:
        assert name in d, name
    # exchange6Async keeps the reference's 19 arguments (exchange6.h:20-28)
    assert len(d["exchange6Async"][1]) == 19
    assert len(d["long_field_async"][1]) == 18
    assert len(d["tensSYMMVecMul"][1]) == 17


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    out = subprocess.check_output(["nm", "-D", "--defined-only", _lib.LIB_PATH]).decode()
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in _lib.declarations() if n not in exported]
    assert not missing, missing
    L = _lib.lib()  # loads and binds every declaration
    for name in _lib.declarations():
        assert hasattr(L, name)
    assert b"gfx950" in L.hs_version()


def test_library_is_built_for_gfx950_only():
    """The offload bundle of our own build names gfx950 and no other target."""
    import re
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    blob = open(_lib.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-f]+)", blob))
    assert targets == {b"gfx950"}, targets
