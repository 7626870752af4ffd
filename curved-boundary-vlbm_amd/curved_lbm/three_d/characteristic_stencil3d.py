"""Fixed-radius geometry weights for the D3Q6 characteristic traction closure.

For every traction cut link this builds, on the host and once per geometry:

* a weighted quadratic least-squares fit over at most 48 active nodes within
  radius sqrt(9.5) of the cut point (lattice units), whose rows 1..3 give the
  state gradient at the cut point;
* a predictor of the state at the cut point from at most 16 samples: two
  trilinear probes along the inward normal combined by linear extrapolation,
  else one probe, else a 1/2/3-node extrapolation along the link.

The output layout is the reference's (characteristic_stencil3d.py:19-171):
``indices`` (64, L) int32 with -1 for unused slots and ``weights``
(4, 64, L): row 0 the predictor, rows 1..3 the x/y/z gradient.  The weights
come from the threaded C++ builder of libvlbm (include/vlbm_geom.h), which
selects the same samples in the same slots as the reference and computes the
weights to rounding.
"""

from __future__ import annotations

import numpy as np

MAX_GRADIENT_SAMPLES = 48
MAX_PROBE_SAMPLES = 16
MAX_CHARACTERISTIC_SAMPLES = MAX_GRADIENT_SAMPLES + MAX_PROBE_SAMPLES


def _raise_for_status(status, batch_size):
    """The errors of the batched NumPy build, from per-link statuses: the
    first batch holding a failure decides; rank/condition failures (1) are
    checked before invalid weights (2)."""
    for start in range(0, status.size, batch_size):
        st = status[start:start + batch_size]
        if np.any(st == 1):
            bad = (np.flatnonzero(st == 1) + start)[:8]
            raise ValueError(
                f"local_f needs a full-rank quadratic stencil within three cells; cut links {bad.tolist()}"
            )
        if np.any(st == 2):
            raise ValueError(f"invalid local_f gradient weights near link {start + np.flatnonzero(st == 2)[0]}")


def build_characteristic_weights(geometry, tables, *, batch_size=1024, link_mask=None, builder="native"):
    """Characteristic predictor/gradient weights for the cut links.

    ``link_mask`` restricts the fit to the selected links; unselected links
    get index -1 and zero weights.  ``batch_size`` groups the links whose
    errors are reported together (the first group holding a failure
    raises), as in the reference.  The only builder is the native one.
    """
    if builder != "native":
        raise ValueError("characteristic builder must be 'native'")
    L = int(geometry.n_links)
    sel = np.arange(L) if link_mask is None else np.flatnonzero(np.asarray(link_mask, dtype=bool))
    if link_mask is not None and np.asarray(link_mask).shape != (L,):
        raise ValueError("link_mask must select one entry per cut link")
    spacing = np.array((geometry.dx, geometry.dy, geometry.dz))
    if not np.allclose(spacing, spacing[0], rtol=1e-12, atol=0.0):
        raise ValueError("local_f requires an isotropic reference lattice")
    from .. import _native

    link = {name: np.asarray(getattr(geometry, "link_" + name))[sel]
            for name in ("i", "j", "k", "q", "xb", "yb", "zb", "nx", "ny", "nz", "eta")}
    idx, w, modes, status = _native.geom3d_char_weights(
        tables["active_id"], link["i"], link["j"], link["k"], link["q"], link["xb"], link["yb"], link["zb"],
        link["nx"], link["ny"], link["nz"], link["eta"], float(spacing[0]))
    _raise_for_status(status, batch_size)
    if link_mask is None:
        indices, weights, all_modes = idx, w, modes
    else:
        indices = np.full((MAX_CHARACTERISTIC_SAMPLES, L), -1, np.int32)
        weights = np.zeros((4, MAX_CHARACTERISTIC_SAMPLES, L), np.float64)
        all_modes = np.zeros(L, np.int32)
        indices[:, sel], weights[:, :, sel], all_modes[sel] = idx, w, modes
    out = {
        "indices": np.ascontiguousarray(indices),
        "weights": np.ascontiguousarray(weights),
        "probe_modes": all_modes,
        "radius": 3,
        "gradient_samples": MAX_GRADIENT_SAMPLES,
        "max_samples": MAX_CHARACTERISTIC_SAMPLES,
    }
    if link_mask is not None:
        out["link_mask"] = np.asarray(link_mask, dtype=bool)
    return out
