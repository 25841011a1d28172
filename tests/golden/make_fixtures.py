"""Derive the small committed fixtures from the reference's golden dumps.

Reads tests/test-demag/{B_ref,B_ref_pbc}.dump and
tests/test-exchange/{hex_ref,hex_ref_pbc}.dump under the reference checkout
(default /root/reference) and writes tests/golden/golden_dumps.npz holding, per
dump: a seeded sample of 2048 cells (internal index and the 3 float32
components), the per-component sums and sums of squares over all cells (as
float64), the mesh header and the CRC stored in the file.

Usage: python tests/golden/make_fixtures.py [reference_root]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from hotspin_amd import dump  # noqa: E402

DUMPS = {
    "B_ref": "tests/test-demag/B_ref.dump",
    "B_ref_pbc": "tests/test-demag/B_ref_pbc.dump",
    "hex_ref": "tests/test-exchange/hex_ref.dump",
    "hex_ref_pbc": "tests/test-exchange/hex_ref_pbc.dump",
}
NSAMPLE = 2048


def main(root="/root/reference"):
    out = {}
    rng = np.random.default_rng(20261202)
    for name, rel in DUMPS.items():
        fr = dump.read(os.path.join(root, rel))
        d = fr.data.reshape(3, -1)
        n = d.shape[1]
        idx = np.sort(rng.choice(n, NSAMPLE, replace=False)).astype(np.int64)
        out[name + "_idx"] = idx
        out[name + "_val"] = d[:, idx].astype(np.float32)
        out[name + "_sum"] = d.astype(np.float64).sum(axis=1)
        out[name + "_sumsq"] = (d.astype(np.float64) ** 2).sum(axis=1)
        out[name + "_sumabs"] = np.abs(d.astype(np.float64)).sum(axis=1)
        out[name + "_mesh"] = np.array(fr.mesh_size, np.int64)
        out[name + "_step"] = np.array(fr.mesh_step, np.float64)
        out[name + "_crc"] = np.array([fr.crc], np.uint64)
    path = os.path.join(HERE, "golden_dumps.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(*sys.argv[1:])
