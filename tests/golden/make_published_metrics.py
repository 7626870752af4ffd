"""Writes tests/golden/published_metrics.json: the subset of the reference's
published manuscript metrics (validation/manuscript_metrics.json of
JingsenFeng/curved-boundary-vlbm) that the end-to-end GPU tests reproduce,
plus, under "release", the values of the reference's release re-run
(validation/release_verification.json, full_benchmark_verification.cases),
which it made after its frozen-source initial-boundary fix
(validation/initialization_buffer_fix.json): the rotated-ellipsoid rows, the
only published case that fix changes beyond rounding, and the tube rows.

Run from the repository root with the reference checked out at
/root/reference:  python tests/golden/make_published_metrics.py
The output is data only (numbers copied from that JSON file)."""

from __future__ import annotations

import json
import sys
from pathlib import Path

REFERENCE = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
OUT = Path(__file__).resolve().parent / "published_metrics.json"

TRACTION_KEYS = {
    "shell": ("rel_l2_u", "rel_l2_F", "rel_l2_sigma", "rms_v", "max_abs_v", "min_J", "max_J"),
    "annulus": ("rel_l2_u", "rel_l2_F", "rel_l2_F_minus_I", "max_abs_u_error", "max_abs_F_error", "rms_v",
                "max_abs_v", "min_J", "max_J"),
    "superellipse": ("rel_l2_u", "rel_l2_F", "rel_l2_F_minus_I", "max_abs_u_error", "max_abs_F_error", "rms_v",
                     "max_abs_v", "min_J", "max_J"),
}
DIRICHLET_KEYS = ("steps", "time", "rel_l2_u", "rel_l2_v", "rel_l2_F", "rel_l2_F_minus_I", "rel_l2_sigma", "min_J",
                  "max_J")


def main() -> None:
    src = json.loads((REFERENCE / "validation" / "manuscript_metrics.json").read_text())
    out = {"source": "validation/manuscript_metrics.json of the reference (manuscript runs, FP64, RTX 5090)"}
    for case, keys in TRACTION_KEYS.items():
        entry = src["traction_cases"][case]
        row = {k: entry["metrics"][k] for k in keys}
        row["steps"] = entry["steps"]
        out[case] = row
    for case in ("affine_annulus", "ellipsoid"):
        out[case] = {str(r["n"]): {k: r[k] for k in DIRICHLET_KEYS if k in r} for r in src["dirichlet_cases"][case]}
    entry = src["traction_cases"]["tube10"]
    out["tube10"] = {"steps": entry["steps"], "time": entry["time"], **entry["metrics"]}
    rel = json.loads((REFERENCE / "validation" / "release_verification.json").read_text())
    cases = rel["full_benchmark_verification"]["cases"]
    out["release"] = {"source": "validation/release_verification.json of the reference "
                                "(full_benchmark_verification.cases, re-run after the initial-boundary fix)"}
    for case in ("ellipsoid", "tube10"):
        out["release"][case] = {
            str(r.get("n", r.get("radius_nodes"))): {"steps": r["steps"], "time": r["time"],
                                                     **{k: v["value"] for k, v in r["metrics"].items()}}
            for r in cases[case]["rows"]}
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
