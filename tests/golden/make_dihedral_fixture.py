"""Extracts the reference-held simulated dihedral sweeps into a small JSON fixture.

Source (read as plain data, run only in a container that holds the reference):
evaluation/simulation/dihedral-{5,7,10,15}cm — 601 poses (yaw -75°..+75° in 0.25° steps) ×
256 range bins, complex64, chirp 0 / channel 0 of libhussar frames. The fixture keeps, per
pose, the peak bin and the peak magnitude, which is what these MC-noisy outputs pin.

Usage: python tests/golden/make_dihedral_fixture.py /root/reference
"""
import json
import os
import sys

import numpy as np


def main(ref_root: str) -> None:
    out = {"source": "evaluation/simulation/dihedral-*cm", "poses": 601, "yaw_first_deg": -75.0,
           "yaw_step_deg": 0.25, "bins": 256, "sizes_cm": {}}
    for s in (5, 7, 10, 15):
        a = np.fromfile(os.path.join(ref_root, "evaluation", "simulation", f"dihedral-{s}cm"),
                        np.complex64).reshape(601, 256)
        mag = np.abs(a)
        out["sizes_cm"][str(s)] = {"peak_bin": mag.argmax(1).astype(int).tolist(),
                                   "peak_abs": [round(float(x), 4) for x in mag.max(1)]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dihedral_fixture.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
