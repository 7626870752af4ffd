"""Writes the FEM comparison fixtures used by the manuscript-metric GPU
tests from the reference's bundled FEM data (reference_data/annulus/
annulus_traction_fem_n196.npz, reference_data/superellipse/
superellipse_mixed_fem_n384.npz): the final displacement u (2, n, n), the
deformation gradient F (2, 2, n, n), and the active/valid masks.  Run once
in a checkout where the reference sits at /root/reference."""

from pathlib import Path

import numpy as np

REF = Path("/root/reference/reference_data")
OUT = Path(__file__).resolve().parent


def main() -> None:
    for sub, name in (("annulus", "annulus_traction_fem_n196.npz"), ("superellipse", "superellipse_mixed_fem_n384.npz")):
        src = np.load(REF / sub / name, allow_pickle=False)
        np.savez_compressed(OUT / name, u=src["u"], F=src["F"], active=src["active"], valid=src["valid"])


if __name__ == "__main__":
    main()
