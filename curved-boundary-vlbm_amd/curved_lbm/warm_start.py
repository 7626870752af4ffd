"""Warm start of the 3D curved solver from a predicted equilibrium.

A predicted displacement u and deformation gradient F given on a coarser
lattice of the same box (the lattice the surrogate network works on) are
carried to the solver's lattice and installed through the reference's own
initialisation entry point, ``initialize_from_numpy`` (reference
curved_boundary_warp3d.py:4216-4236): v = 0, F and u from the prediction,
populations f = feq(U) - dt/(2Q) B, frozen-source boundary reconstruction at
the start time.  The solver clock then starts at the end of the load ramp,
so the boundary values are the full load.  The solver runs to the same
equilibrium residual as from rest; the prediction only changes where it
starts.  If the installed state has a non-positive Jacobian, non-finite
values, or a first-window residual above the plain start's limit, the run
falls back to the plain start from rest.

Carrying a field between lattices: the coarse field is first extended into
the inactive cells next to the body (each sweep assigns to every inactive
cell with active neighbours the mean of those neighbours), then upsampled
by trilinear interpolation on cell centres (node i of a lattice of spacing
h sits at (i + 1/2) h, so with factor r fine node I samples coarse position
(I + 1/2) / r - 1/2, clamped to the box).  Torch does this arithmetic; it is
part of the learned path's plumbing, not of the time step.
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch
import torch.nn.functional as tf


def dense_state(solver) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """(u (3, nx, ny, nz), F (9, nx, ny, nz), active mask (nx, ny, nz)) of a
    3D solver's current state on its device, in float64; inactive cells hold
    zeros (F too)."""
    solver.refresh_current()
    dev = solver.device
    shape = (solver.nx, solver.ny, solver.nz)
    idx = torch.from_numpy(
        (solver.active_i.astype(np.int64) * shape[1] + solver.active_j) * shape[2] + solver.active_k).to(dev)
    u = torch.zeros((3, math.prod(shape)), dtype=torch.float64, device=dev)
    F = torch.zeros((9, math.prod(shape)), dtype=torch.float64, device=dev)
    with torch.cuda.stream(solver._stream):
        u[:, idx] = solver.u.tensor
        F[:, idx] = solver.U.tensor[3:12]
    mask = torch.zeros(math.prod(shape), dtype=torch.bool, device=dev)
    mask[idx] = True
    torch.cuda.synchronize(dev)
    return u.view(3, *shape), F.view(9, *shape), mask.view(shape)


def extend_into_inactive(field: torch.Tensor, mask: torch.Tensor, sweeps: int = 3) -> torch.Tensor:
    """Field (C, X, Y, Z) with its values extended `sweeps` cells into the
    inactive region: each sweep gives every not-yet-filled cell that has
    filled face neighbours the mean of those neighbours."""
    filled = mask.clone().to(field.dtype)
    out = field * filled
    for _ in range(int(sweeps)):
        num = torch.zeros_like(out)
        cnt = torch.zeros_like(filled)
        for axis in (1, 2, 3):
            for shift in (1, -1):
                num += _shift(out, shift, axis)
                cnt += _shift(filled[None], shift, axis)[0]
        new = (filled == 0) & (cnt > 0)
        out = torch.where(new[None], num / cnt.clamp(min=1.0)[None], out)
        filled = torch.where(new, torch.ones_like(filled), filled)
    return out


def _shift(x: torch.Tensor, shift: int, axis: int) -> torch.Tensor:
    """x shifted by one cell along `axis` with zero fill (no wrap)."""
    out = torch.zeros_like(x)
    n = x.shape[axis]
    src = [slice(None)] * x.dim()
    dst = [slice(None)] * x.dim()
    if shift > 0:
        src[axis], dst[axis] = slice(0, n - 1), slice(1, n)
    else:
        src[axis], dst[axis] = slice(1, n), slice(0, n - 1)
    out[tuple(dst)] = x[tuple(src)]
    return out


def upsample(field: torch.Tensor, fine_shape: tuple[int, int, int]) -> torch.Tensor:
    """Trilinear interpolation of a cell-centred field (C, X, Y, Z) onto a
    lattice of the same box with shape fine_shape (integer factors)."""
    if tuple(field.shape[1:]) == tuple(fine_shape):
        return field
    for a, b in zip(field.shape[1:], fine_shape):
        if b % a:
            raise ValueError(f"fine lattice {fine_shape} is not an integer refinement of {tuple(field.shape[1:])}")
    return tf.interpolate(field[None], size=tuple(fine_shape), mode="trilinear", align_corners=False)[0]


def carry_to_lattice(u: torch.Tensor, F: torch.Tensor, mask: torch.Tensor, fine_shape) -> tuple[torch.Tensor, torch.Tensor]:
    """u (3, ...) and F (9, ...) of a coarse lattice carried to fine_shape."""
    both = torch.cat([u, F], dim=0).to(torch.float64)
    both = extend_into_inactive(both, mask)
    # inactive cells that no sweep reached hold the rest state
    reach = extend_into_inactive(torch.ones_like(mask, dtype=torch.float64)[None], mask)[0] > 0
    eye = torch.eye(3, dtype=both.dtype, device=both.device).reshape(9, 1, 1, 1)
    both[3:] = torch.where(reach[None], both[3:], eye.expand_as(both[3:]))
    fine = upsample(both, fine_shape)
    return fine[:3], fine[3:]


def transform(u: torch.Tensor, F: torch.Tensor, perm, sign) -> tuple[torch.Tensor, torch.Tensor]:
    """The fields moved by the signed axis permutation R, (R v)_a =
    sign[a] v[perm[a]], about the box centre: u'(x) = R u(R^T x) and
    F'(x) = R F(R^T x) R^T (cell-centred lattices map onto themselves)."""
    def move(f):
        g = f.permute(0, 1 + perm[0], 1 + perm[1], 1 + perm[2])
        flips = [1 + a for a in range(3) if sign[a] < 0]
        return g.flip(flips) if flips else g

    uu = move(u)
    u2 = torch.stack([sign[a] * uu[perm[a]] for a in range(3)])
    FF = move(F).reshape(3, 3, *uu.shape[1:])
    F2 = torch.stack([torch.stack([sign[a] * sign[b] * FF[perm[a], perm[b]] for b in range(3)]) for a in range(3)])
    return u2, F2.reshape(9, *uu.shape[1:])


def _is_cube_group(group) -> bool:
    """Whether group is the full 48-element symmetry group of the cube."""
    import itertools

    full = {(tuple(p), tuple(s)) for p in itertools.permutations(range(3)) for s in itertools.product((1, -1), repeat=3)}
    return len(group) == 48 and {(tuple(p), tuple(s)) for p, s in group} == full


def _symmetrize_cube(u: torch.Tensor, F: torch.Tensor, mask: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """The average over the cube group in nine moves instead of 48: every
    element is uniquely a sign flip after an axis permutation, so the group
    average is the average over the eight flips of the average over the six
    permutations, and the flip average is three successive averages over
    {identity, flip of one axis} (the flips commute)."""
    m = mask[None].to(u.dtype)
    for perm, sign in (((1, 0, 2), (1, 1, 1)), ((1, 2, 0), (1, 1, 1)), ((0, 1, 2), (-1, 1, 1)),
                       ((0, 1, 2), (1, -1, 1)), ((0, 1, 2), (1, 1, -1))):
        mm, _ = transform(m.expand(3, *mask.shape), m.expand(9, *mask.shape), perm, sign)
        if not torch.equal(mm[0].abs(), m[0]):
            raise ValueError(f"the active cells are not invariant under {perm}, {sign}")
    import itertools

    su, sF = torch.zeros_like(u), torch.zeros_like(F)
    for perm in itertools.permutations(range(3)):
        a, b = transform(u, F, perm, (1, 1, 1))
        su += a
        sF += b
    u, F = su / 6.0, sF / 6.0
    for axis in range(3):
        sign = tuple(-1 if k == axis else 1 for k in range(3))
        a, b = transform(u, F, (0, 1, 2), sign)
        u, F = 0.5 * (u + a), 0.5 * (F + b)
    return u, F


def symmetrize(u: torch.Tensor, F: torch.Tensor, mask: torch.Tensor, group) -> tuple[torch.Tensor, torch.Tensor]:
    """Average of the fields over a symmetry group of the problem (a list
    of (perm, sign) closed under composition, the identity included).  The
    discrete equilibrium of a problem with that symmetry is invariant, so
    the average keeps it and removes the error components that excite the
    antisymmetric modes.  Raises if the active mask is not invariant."""
    if _is_cube_group(group):
        return _symmetrize_cube(u, F, mask)
    su, sF = torch.zeros_like(u), torch.zeros_like(F)
    m = mask[None].to(u.dtype)
    for perm, sign in group:
        mm, _ = transform(m.expand(3, *mask.shape), m.expand(9, *mask.shape), perm, sign)
        if not torch.equal(mm[0].abs(), m[0]):
            raise ValueError(f"the active cells are not invariant under {perm}, {sign}")
        a, b = transform(u, F, perm, sign)
        su += a
        sF += b
    return su / len(group), sF / len(group)


def _axis_pairs(x: torch.Tensor, axis: int):
    """(lower, upper) views of x along a lattice axis (0, 1, 2) of its last
    three dimensions: lower[..., i] = x[..., i], upper[..., i] = x[..., i + 1]."""
    n = x.shape[x.dim() - 3 + axis]
    return x.narrow(x.dim() - 3 + axis, 0, n - 1), x.narrow(x.dim() - 3 + axis, 1, n - 1)


def integrate_displacement(F: torch.Tensor, mask: torch.Tensor, spacing, anchors, u0: torch.Tensor | None = None,
                           *, tol: float = 1e-10, max_iter: int = 4000) -> tuple[torch.Tensor, dict]:
    """Displacement u (3, X, Y, Z) whose lattice differences match F - I in
    the least-squares sense, pinned at the Dirichlet wall points.

    Unknowns: u on the active cells.  Equations, all of weight one: for each
    pair of face-adjacent active cells a, b = a + e_d,
    u(b) - u(a) = h_d (G(a) + G(b)) e_d / 2 with G = F - I (exact for a
    linear field); for each anchor (a cell a, a wall point x_w and the wall
    displacement u_w there), u(a) + G(a) (x_w - x_a) = u_w.  The normal
    equations (a graph Laplacian plus the anchor diagonal) are solved by
    Jacobi-preconditioned conjugate gradients from u0 (zero by default) to a
    relative residual tol.  anchors = (i, j, k, xw (3, n), uw (3, n)) in
    lattice indices and physical coordinates.  The three components share
    the matrix and are solved as one block-diagonal system.
    """
    dev, dt = F.device, torch.float64
    h = [float(s) for s in spacing]
    act = mask.to(dt)
    G = (F.to(dt) - torch.eye(3, dtype=dt, device=dev).reshape(9, 1, 1, 1)).reshape(3, 3, *mask.shape)
    rhs = torch.zeros((3,) + tuple(mask.shape), dtype=dt, device=dev)
    diag = torch.zeros(tuple(mask.shape), dtype=dt, device=dev)
    pairs = []
    for d in range(3):
        lo, hi = _axis_pairs(act, d)
        pm = lo * hi
        pairs.append(pm)
        glo, ghi = _axis_pairs(G[:, d], d)
        g = 0.5 * h[d] * (glo + ghi) * pm
        rlo, rhi = _axis_pairs(rhs, d)
        rlo -= g
        rhi += g
        dlo, dhi = _axis_pairs(diag, d)
        dlo += pm
        dhi += pm
    ai, aj, ak, xw, uw = anchors
    ai, aj, ak = (torch.as_tensor(np.asarray(a), dtype=torch.int64, device=dev) for a in (ai, aj, ak))
    xw = torch.as_tensor(np.asarray(xw), dtype=dt, device=dev)
    uw = torch.as_tensor(np.asarray(uw), dtype=dt, device=dev)
    if ai.numel():
        xa = torch.stack([(ai.to(dt) + 0.5) * h[0], (aj.to(dt) + 0.5) * h[1], (ak.to(dt) + 0.5) * h[2]])
        Ga = G[:, :, ai, aj, ak]                                    # (3, 3, n)
        target = uw - (Ga * (xw - xa)[None]).sum(1)                 # (3, n)
        flat = (ai * mask.shape[1] + aj) * mask.shape[2] + ak
        diag.view(-1).index_add_(0, flat, torch.ones_like(flat, dtype=dt))
        rhs.view(3, -1).index_add_(1, flat, target)

    def apply(u):
        out = diag * u
        for d in range(3):
            ulo, uhi = _axis_pairs(u, d)
            olo, ohi = _axis_pairs(out, d)
            olo -= pairs[d] * uhi
            ohi -= pairs[d] * ulo
        return out

    pre = torch.where(diag > 0, 1.0 / diag.clamp(min=1e-300), torch.zeros_like(diag)) * act
    u = torch.zeros_like(rhs) if u0 is None else u0.to(dt) * act
    r = (rhs - apply(u)) * act
    z = pre * r
    p = z.clone()
    rz = (r * z).sum()
    bnorm = float(torch.linalg.norm(rhs * act)) or 1.0
    it, res = 0, float(torch.linalg.norm(r)) / bnorm
    while it < max_iter and res > tol:
        Ap = apply(p)
        alpha = rz / (p * Ap).sum()
        u += alpha * p
        r -= alpha * Ap
        it += 1
        if it % 25 == 0 or it == max_iter:
            res = float(torch.linalg.norm(r)) / bnorm
        z = pre * r
        rz_new = (r * z).sum()
        p = z + (rz_new / rz) * p
        rz = rz_new
    res = float(torch.linalg.norm(r)) / bnorm
    return u * act, {"iterations": it, "relative_residual": res}


def install(solver, u_dense: torch.Tensor, F_dense: torch.Tensor, t_start: float) -> None:
    """Initialise the solver at rest velocity with the given u and F (dense
    device fields on the solver lattice) and set its clock to t_start."""
    ai = torch.from_numpy(solver.active_i.astype(np.int64)).to(u_dense.device)
    aj = torch.from_numpy(solver.active_j.astype(np.int64)).to(u_dense.device)
    ak = torch.from_numpy(solver.active_k.astype(np.int64)).to(u_dense.device)
    U = torch.zeros((12, solver.n_active), dtype=torch.float64, device=u_dense.device)
    U[3:12] = F_dense[:, ai, aj, ak]
    disp = u_dense[:, ai, aj, ak].to(torch.float64)
    # initialize_from_numpy's path (cbw3d:4216-4236) with the active-node
    # values handed over on the device
    torch.cuda.synchronize(u_dense.device)
    solver._initialize_from_active(U, disp, initial_boundary_time=float(t_start))
    solver.time = float(t_start)


def run_warm(solver, u_dense, F_dense, *, t_start: float, tol: float, check_every: int = 200,
             max_steps: int = 10**6, max_initial_residual: float = 1e-2, plain=None) -> dict:
    """Warm-started run to the equilibrium residual with the fallback.

    Installs the prediction, checks the installed state (finite, min J > 0)
    and the residual over the first check window (limit 1e-2 by default:
    from rest, the first windows of the shell and tube loads of the families
    move at 2e-4 and 4e-3); if either check fails the
    solver is re-initialised from rest (``plain()`` may supply the plain
    schedule, by default ``initialize_identity`` with the clock at 0) and run
    to the residual, not stopping before t_start (the end of the load ramp)
    as a plain run does not.  Returns the run record with ``warm`` telling which start
    was used and ``wall_s`` the time from install to equilibrium."""
    torch.cuda.synchronize(solver.device)
    t0 = time.perf_counter()
    install(solver, u_dense, F_dense, t_start)
    red = solver._refreshed_reductions()
    ok = not red["nonfinite"] and red["min_J"] > 0.0 and not solver.invalid_state()
    reason = None if ok else f"installed state invalid (min J {red['min_J']:.3g}, nonfinite {red['nonfinite']})"
    out = None
    if ok:
        out = solver.run_to_equilibrium(tol, check_every=check_every, max_steps=solver.steps + max_steps,
                                        first_check_limit=max_initial_residual)
        if out.get("first_check_failed"):
            reason = f"first-window residual {out['history'][0][1]:.3g} > {max_initial_residual:.3g}"
            ok = False
    if not ok:
        if plain is not None:
            out = plain(solver)
        else:
            solver.initialize_identity()
            out = solver.run_to_equilibrium(tol, check_every=check_every, min_time=t_start, max_steps=max_steps)
    torch.cuda.synchronize(solver.device)
    out = dict(out)
    out.update(warm=ok, fallback_reason=reason, wall_s=time.perf_counter() - t0)
    return out
