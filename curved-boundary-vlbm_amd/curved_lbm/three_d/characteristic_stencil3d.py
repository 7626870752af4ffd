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
(4, 64, L): row 0 the predictor, rows 1..3 the x/y/z gradient.  Sample
selection order (stable sort by distance over the 7x7x7 offsets in
lexicographic order), Gaussian weights exp(-r^2/4.5) + 1e-4, the 1e-12
Tikhonov shift, probe distances and acceptance rules follow that file so the
weights agree with the reference to rounding.
"""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np

MAX_GRADIENT_SAMPLES = 48
MAX_PROBE_SAMPLES = 16
MAX_CHARACTERISTIC_SAMPLES = MAX_GRADIENT_SAMPLES + MAX_PROBE_SAMPLES

_OFFSETS = np.stack(np.meshgrid(np.arange(-3, 4), np.arange(-3, 4), np.arange(-3, 4), indexing="ij"), -1)
_OFFSETS = _OFFSETS.reshape(-1, 3).astype(np.int32)  # lexicographic, first axis slowest
_CORNERS = np.stack(np.meshgrid((0, 1), (0, 1), (0, 1), indexing="ij"), -1).reshape(-1, 3).astype(np.int32)
_DIRS = np.array(((1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1)))
_PROBE_PAIRS = ((0.75, 1.75), (1.0, 2.0), (1.25, 2.5), (1.5, 3.0))
_PROBE_SINGLE = (0.75, 1.0, 1.25, 1.5, 2.0)


def _subset(geometry, ids):
    sub = SimpleNamespace(
        active=geometry.active, dx=geometry.dx, dy=geometry.dy, dz=geometry.dz, n_links=len(ids)
    )
    for name in ("i", "j", "k", "xb", "yb", "zb", "nx", "ny", "nz", "q", "eta"):
        setattr(sub, "link_" + name, np.asarray(getattr(geometry, "link_" + name))[ids])
    return sub


def _quadratic_basis(xi: np.ndarray) -> np.ndarray:
    x, y, z = xi[..., 0], xi[..., 1], xi[..., 2]
    return np.stack(
        (np.ones_like(x), x, y, z, 0.5 * x * x, x * y, x * z, 0.5 * y * y, y * z, 0.5 * z * z), axis=-1
    )


def _gradient_fit(ctr, xb, shape, active_id, dx, start):
    count = ctr.shape[0]
    nodes = ctr[:, None, :] + _OFFSETS[None, :, :]
    inside = np.all((nodes >= 0) & (nodes < shape), axis=2)
    safe = np.clip(nodes, 0, shape - 1)
    ids = active_id[safe[..., 0], safe[..., 1], safe[..., 2]]
    xi = nodes - xb[:, None, :]
    r2 = np.sum(xi * xi, axis=2)
    dist = np.where(inside & (ids >= 0) & (r2 <= 9.5), r2, np.inf)
    order = np.argsort(dist, axis=1, kind="stable")[:, :MAX_GRADIENT_SAMPLES]
    rows = np.arange(count)[:, None]
    r2 = dist[rows, order]
    xi = xi[rows, order]
    chosen = np.isfinite(r2)
    A = _quadratic_basis(xi)
    W = np.where(chosen, np.exp(-r2 / (2 * 1.5**2)) + 1e-4, 0.0)
    M = np.einsum("bmi,bm,bmj->bij", A, W, A)
    cond = np.linalg.cond(M)
    ok = (np.sum(chosen, axis=1) >= 10) & np.isfinite(cond) & (cond <= 1e8)
    if not np.all(ok):
        bad = (np.flatnonzero(~ok) + start)[:8]
        raise ValueError(
            f"local_f needs a full-rank quadratic stencil within three cells; cut links {bad.tolist()}"
        )
    coef = np.linalg.solve(M + 1e-12 * np.eye(10), np.swapaxes(A * W[:, :, None], 1, 2))
    grads = coef[:, 1:4, :] / dx
    ok = np.all(np.isfinite(grads), axis=(1, 2))
    ok &= np.all(np.abs(np.sum(grads, axis=2)) * dx < 1e-6, axis=1)
    ok &= np.all(np.sum(np.abs(grads), axis=2) * dx <= 12.0, axis=1)
    if not np.all(ok):
        raise ValueError(f"invalid local_f gradient weights near link {start + np.flatnonzero(~ok)[0]}")
    return np.where(chosen, ids[rows, order], -1), grads


def _probe(point, ctr, shape, active_id):
    lower = np.floor(point).astype(np.int32)
    frac = point - lower
    grid = lower[:, None, :] + _CORNERS[None, :, :]
    in_grid = np.all((grid >= 0) & (grid < shape), axis=2)
    near = np.all(np.abs(grid - ctr[:, None, :]) <= 3, axis=2)
    clipped = np.clip(grid, 0, shape - 1)
    pp = active_id[clipped[..., 0], clipped[..., 1], clipped[..., 2]]
    ww = np.prod(np.where(_CORNERS[None, :, :] == 1, frac[:, None, :], 1 - frac[:, None, :]), axis=2)
    ok = np.all(in_grid & near & (pp >= 0), axis=1)
    return pp, ww, ok


def _predictor(ctr, xb, normal, shape, active_id, link_q, link_eta):
    count = ctr.shape[0]
    done = np.zeros(count, bool)
    pp_out = np.full((count, MAX_PROBE_SAMPLES), -1, np.int32)
    ww_out = np.zeros((count, MAX_PROBE_SAMPLES))
    mode = np.zeros(count, np.int32)
    for h1, h2 in _PROBE_PAIRS:
        p1, w1, ok1 = _probe(xb - h1 * normal, ctr, shape, active_id)
        p2, w2, ok2 = _probe(xb - h2 * normal, ctr, shape, active_id)
        take = ~done & ok1 & ok2
        pp_out[take, :8] = p1[take]
        pp_out[take, 8:] = p2[take]
        ww_out[take, :8] = h2 / (h2 - h1) * w1[take]
        ww_out[take, 8:] = -h1 / (h2 - h1) * w2[take]
        mode[take] = 2
        done |= take
    for h in _PROBE_SINGLE:
        pp, ww, ok = _probe(xb - h * normal, ctr, shape, active_id)
        take = ~done & ok
        pp_out[take, :8] = pp[take]
        ww_out[take, :8] = ww[take]
        mode[take] = 1
        done |= take
    for row in np.flatnonzero(~done):
        eta = link_eta[row]
        step = _DIRS[link_q[row]]
        chain = []
        for k in range(3):
            node = ctr[row] - k * step
            if np.any(node < 0) or np.any(node >= shape):
                break
            p = active_id[tuple(node)]
            if p < 0:
                break
            chain.append(p)
        if len(chain) == 3:
            ww = (0.5 * (eta + 1) * (eta + 2), -eta * (eta + 2), 0.5 * eta * (eta + 1))
        elif len(chain) == 2:
            ww = (1 + eta, -eta)
        else:
            ww = (1.0,)
        pp_out[row, : len(chain)] = chain
        ww_out[row, : len(chain)] = ww
        mode[row] = -len(chain)
    return pp_out, ww_out, mode


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
    get index -1 and zero weights.  ``builder="native"`` runs the threaded
    C++ builder of libvlbm (the same samples, slots and modes; weights equal
    to rounding), ``"numpy"`` the batched NumPy restatement below.
    """
    if builder not in ("native", "numpy"):
        raise ValueError("characteristic builder must be 'native' or 'numpy'")
    if link_mask is not None:
        mask = np.asarray(link_mask, dtype=bool)
        if mask.shape != (geometry.n_links,):
            raise ValueError("link_mask must select one entry per cut link")
        ids = np.flatnonzero(mask)
        fit = build_characteristic_weights(_subset(geometry, ids), tables, batch_size=batch_size, builder=builder)
        indices = np.full((MAX_CHARACTERISTIC_SAMPLES, geometry.n_links), -1, np.int32)
        weights = np.zeros((4, MAX_CHARACTERISTIC_SAMPLES, geometry.n_links), np.float64)
        modes = np.zeros(geometry.n_links, np.int32)
        indices[:, ids] = fit["indices"]
        weights[:, :, ids] = fit["weights"]
        modes[ids] = fit["probe_modes"]
        fit.update(indices=indices, weights=weights, probe_modes=modes, link_mask=mask)
        return fit
    nlinks = geometry.n_links
    spacing = np.array((geometry.dx, geometry.dy, geometry.dz))
    if not np.allclose(spacing, spacing[0], rtol=1e-12, atol=0.0):
        raise ValueError("local_f requires an isotropic reference lattice")
    dx = spacing[0]
    centers = np.column_stack((geometry.link_i, geometry.link_j, geometry.link_k))
    cuts = np.column_stack((geometry.link_xb, geometry.link_yb, geometry.link_zb))
    normals = np.column_stack((geometry.link_nx, geometry.link_ny, geometry.link_nz))
    shape = np.array(geometry.active.shape)
    active_id = tables["active_id"]
    indices = np.full((MAX_CHARACTERISTIC_SAMPLES, nlinks), -1, np.int32)
    weights = np.zeros((4, MAX_CHARACTERISTIC_SAMPLES, nlinks), np.float64)
    modes = np.zeros(nlinks, np.int32)
    link_q = np.asarray(geometry.link_q)
    link_eta = np.asarray(geometry.link_eta)
    if builder == "native":
        from .. import _native

        indices, weights, modes, status = _native.geom3d_char_weights(
            active_id, geometry.link_i, geometry.link_j, geometry.link_k, link_q, geometry.link_xb,
            geometry.link_yb, geometry.link_zb, geometry.link_nx, geometry.link_ny, geometry.link_nz, link_eta, dx)
        _raise_for_status(status, batch_size)
        nlinks = 0
    for start in range(0, nlinks, batch_size):
        stop = min(start + batch_size, nlinks)
        ctr = centers[start:stop]
        xb = cuts[start:stop] / dx - 0.5
        ids, grads = _gradient_fit(ctr, xb, shape, active_id, dx, start)
        indices[:MAX_GRADIENT_SAMPLES, start:stop] = ids.T
        weights[1:, :MAX_GRADIENT_SAMPLES, start:stop] = np.transpose(grads, (1, 2, 0))
        pp, ww, mode = _predictor(
            ctr, xb, normals[start:stop], shape, active_id, link_q[start:stop], link_eta[start:stop]
        )
        indices[MAX_GRADIENT_SAMPLES:, start:stop] = pp.T
        weights[0, MAX_GRADIENT_SAMPLES:, start:stop] = ww.T
        modes[start:stop] = mode
    return {
        "indices": np.ascontiguousarray(indices),
        "weights": np.ascontiguousarray(weights),
        "probe_modes": modes,
        "radius": 3,
        "gradient_samples": MAX_GRADIENT_SAMPLES,
        "max_samples": MAX_CHARACTERISTIC_SAMPLES,
    }
