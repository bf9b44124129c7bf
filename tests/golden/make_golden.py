"""Writes the golden fixtures of the smoothing rule from the reference's data files.

The reference ships the true model (input/acc_vp.bin, 845 x 702 floats, ix-major)
and the migration model its own ini_model_mine produced from it
(kkdecon_acoustic/input/ini_vp.bin, window 5).  Three 48 x 48 crops of both
(top-left corner, a middle patch, bottom-right corner) are stored as raw float32
together with the md5 of the full files.  Run once where the reference is mounted:

    python tests/golden/make_golden.py /root/reference
"""
import hashlib
import json
import os
import sys

import numpy as np

NX, NZ, C = 845, 702, 48
CROPS = {"tl": (0, 0), "mid": (400, 300), "br": (NX - C, NZ - C)}


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    acc_p = os.path.join(ref, "input", "acc_vp.bin")
    ini_p = os.path.join(ref, "kkdecon_acoustic", "input", "ini_vp.bin")
    acc = np.fromfile(acc_p, np.float32).reshape(NX, NZ)
    ini = np.fromfile(ini_p, np.float32).reshape(NX, NZ)
    meta = {"acc_vp_md5": hashlib.md5(open(acc_p, "rb").read()).hexdigest(),
            "ini_vp_md5": hashlib.md5(open(ini_p, "rb").read()).hexdigest(),
            "shape": [NX, NZ], "window": 5, "crop": C, "crops": CROPS}
    for name, (x0, z0) in CROPS.items():
        acc[x0:x0 + C, z0:z0 + C].tofile(os.path.join(here, f"acc_vp_{name}.f32"))
        ini[x0:x0 + C, z0:z0 + C].tofile(os.path.join(here, f"ini_vp_{name}.f32"))
    with open(os.path.join(here, "ini_vp_crops.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
