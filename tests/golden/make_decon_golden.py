"""Writes the golden fixtures of the PSF deconvolution from the reference's data files.

The reference's deconvolution (kkdecon_acoustic/kk_decon_interp_psf_minewindow.cpp)
reads a migrated image (kkdecon_acoustic/input/migration_dupp.bin) and the
migrated image of a lattice of point scatterers, the PSF image
(kkdecon_acoustic/input/migration_dupsf.bin), both 845 x 702 float32, ix-major,
with the point scatterers at (15 + 31 i, 15 + 31 j).  A 310 x 279 crop of both
starting at (248, 217), a multiple of the 31-cell lattice so that the crop is a
valid deconvolution input of its own, is stored as raw float32 together with
the md5 of the full files.  The reference ships no deconvolved output, so no
expected output is stored.  Run once where the reference is mounted:

    python tests/golden/make_decon_golden.py /root/reference
"""
import hashlib
import json
import os
import sys

import numpy as np

NX, NZ = 845, 702
X0, Z0, CX, CZ = 248, 217, 310, 279


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    meta = {"shape": [NX, NZ], "origin": [X0, Z0], "crop": [CX, CZ]}
    for name in ("migration_dupp", "migration_dupsf"):
        path = os.path.join(ref, "kkdecon_acoustic", "input", name + ".bin")
        a = np.fromfile(path, np.float32).reshape(NX, NZ)
        meta[name + "_md5"] = hashlib.md5(open(path, "rb").read()).hexdigest()
        np.ascontiguousarray(a[X0:X0 + CX, Z0:Z0 + CZ]).tofile(os.path.join(here, f"decon_{name[10:]}_crop.f32"))
    with open(os.path.join(here, "decon_crop.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
