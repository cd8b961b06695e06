#!/usr/bin/env python3
"""Write the full-size fixture of the 1,728,000-ion molten NaCl (60^3 rocksalt
cells, SPME mesh 360^3, order 8, rc 12 A, alpha 0.28/A, padding 0.5 A) from
the oracle (oracle/dlpx_oracle.c, FP64).

The oracle takes a few minutes at this size, so the fixture keeps what the
GPU test compares: the forces of every 1000th atom, the energies, virials,
stress and the 20-slot buffer, with a checksum of the generated positions
(the test regenerates the system from the same seed).

Run in the build container:
    python tests/golden/make_large_fixture.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "dl-poly_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402
from dlpx import systems  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "large", "nacl_1728k_strided.npz")
STRIDE = 1000


def system():
    return systems.nacl(ncells=60, kvec=(360, 360, 360), padding=0.5)


def main():
    s = system()
    o = oracle.two_body(s)
    idx = np.arange(0, s.natms, STRIDE)
    np.savez_compressed(OUT, idx=idx, frc=o["frc"][idx], frc_rms=np.sqrt((o["frc"] ** 2).mean()),
                        pos_checksum=np.array([s.pos.sum(), (s.pos ** 2).sum()]),
                        energies=np.array([o["engcpe"], o["vircpe"], o["engsrp"], o["virsrp"]]),
                        stress=o["stress"], buffer=o["buffer"], natms=s.natms)
    print(OUT)


if __name__ == "__main__":
    main()
