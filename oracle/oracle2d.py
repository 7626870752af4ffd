"""Python harness of the CPU FP64 D2Q4x6 oracle (oracle/vlbm2d_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's CPU-baseline
leg, never by the product package.

``OracleSolver2D`` takes the constructor arguments of
``curved_lbm.two_d.curved_boundary_warp.WarpCurvedBoundaryLBM2D`` and a
``CurvedBoundaryGeometry`` (its cut links), rebuilds the normal-probe / LSQ
weights with its own NumPy restatement of the reference's builder
(oracle/stencils2d.py) rather than taking the product's, then advances the
dense reference state with the C restatement of the reference step.  State arrays are numpy
arrays in the reference layout (components, nx, ny).
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import oracle3d as _o3
from . import stencils2d

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "libvlbm2d_oracle.so"
_dp = C.c_void_p


class _State(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("n_links", C.c_int32),
        ("model", C.c_int32),
        ("lam", C.c_double), ("mu", C.c_double), ("p0", C.c_double), ("p1", C.c_double),
        ("p2", C.c_double), ("j_floor", C.c_double),
        ("dt", C.c_double), ("c", C.c_double), ("omega", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
        ("source_mode", C.c_int32), ("source_params", C.c_double * 6),
        ("recon", C.c_int32), ("normal_solver", C.c_int32),
        ("disp_interval", C.c_int32), ("disp_sweeps", C.c_int32),
        ("disp_relax", C.c_double), ("disp_weight", C.c_double), ("disp_boundary_id", C.c_int32),
        ("compat_interval", C.c_int32), ("compat_blend", C.c_double), ("compat_interior_only", C.c_int32),
        ("use_dyn", C.c_int32),
        ("collision_model", C.c_int32), ("ghost_omega", C.c_double), ("filter_strength", C.c_double),
        ("box", C.c_int32), ("side_kind", C.c_int32 * 4), ("side_mode", C.c_int32 * 4),
        ("side_params", C.c_double * 32), ("length_x", C.c_double), ("length_y", C.c_double),
        ("f0", _dp), ("f1", _dp), ("fpost", _dp),
        ("U", _dp), ("u", _dp), ("u_star", _dp), ("u_next", _dp),
        ("bc_diag", _dp), ("bc_rhs", _dp), ("P", _dp), ("sigma", _dp), ("invalid", _dp),
        ("active", _dp), ("link_i", _dp), ("link_j", _dp), ("link_q", _dp), ("link_bid", _dp),
        ("link_eta", _dp), ("link_xb", _dp), ("link_yb", _dp), ("link_nx", _dp), ("link_ny", _dp),
        ("lsq_i", _dp), ("lsq_j", _dp), ("lsq_wb", _dp), ("lsq_wt", _dp), ("lsq_wx", _dp), ("lsq_wy", _dp),
        ("lsq_valid", _dp),
        ("kind_by_id", _dp), ("mode_by_id", _dp), ("params_by_id", _dp),
        ("dyn_tx", _dp), ("dyn_ty", _dp),
    ]


_lib = None


def build(quiet: bool = True) -> Path:
    return _o3.build(quiet)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        S = C.POINTER(_State)
        lib.or2d_step.argtypes = [S, C.c_double, C.c_int64, C.c_int32]
        lib.or2d_refresh.argtypes = [S]
        lib.or2d_apply_boundary.argtypes = [S, C.c_double, C.c_int32]
        lib.or2d_piola.argtypes = [_dp, C.c_int] + [C.c_double] * 6 + [_dp]
        lib.or2d_piola.restype = None
        lib.or2d_solve_normal.argtypes = [S, _dp] + [C.c_double] * 4 + [_dp]
        lib.or2d_solve_normal.restype = None
        lib.or2d_boundary_value.argtypes = [C.c_int, C.c_int, _dp] + [C.c_double] * 5
        lib.or2d_boundary_value.restype = C.c_double
        lib.or2d_du.argtypes = [S, _dp, _dp]
        lib.or2d_kinetic_filter.argtypes = [S, _dp]
        lib.or2d_kinetic_filter.restype = None
        lib.or2d_du.restype = None
        lib.or2d_equilibrium.argtypes = [S, _dp, _dp]
        lib.or2d_equilibrium.restype = None
        for name in ("or2d_step", "or2d_refresh", "or2d_apply_boundary", "or2d_threads"):
            getattr(lib, name).restype = C.c_int
        _lib = lib
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def boundary_value(comp: int, mode: int, params, x: float, y: float, nx: float, ny: float, t: float) -> float:
    p = np.zeros(8)
    vals = tuple(params)[:8]
    p[: len(vals)] = vals
    return float(load().or2d_boundary_value(int(comp), int(mode), _p(p), x, y, nx, ny, t))


def piola(F: np.ndarray, material) -> np.ndarray:
    """P of one 2x2 F (row-major) with the oracle's constitutive map."""
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(4))
    P = np.zeros(4)
    load().or2d_piola(_p(F), int(material.id), float(material.lam), float(material.mu),
                      *map(float, material.params), 1.0e-12, _p(P))
    return P.reshape(2, 2)


class OracleSolver2D:
    """Dense CPU FP64 oracle with the reference solver's constructor and methods."""

    def __init__(self, geometry, *, material, boundary, lattice_speed=5.0, collision_omega=2.0,
                 boundary_reconstruction="bfl", jacobian_floor=1.0e-12, source_mode="none",
                 source_vector=(0.0,) * 6, damping_gamma=0.0, local_compatibility_interval=0,
                 local_compatibility_blend=0.05, local_compatibility_interior_only=True,
                 local_displacement_interval=0, local_displacement_sweeps=1, local_displacement_relax=0.85,
                 local_displacement_boundary_weight=20.0, local_displacement_boundary_id=2,
                 normal_jacobian=None, neumann_solver=None, collision_model="bgk", ghost_omega=None,
                 kinetic_filter_strength=0.0, threads: int | None = None, **_ignored):
        from curved_lbm.two_d import curved_boundary_warp as cb
        from curved_lbm.two_d.vector_nonlinear_elastic_warp import source_id

        self.lib = load()
        g = geometry
        self.geometry = g
        self.nx, self.ny = int(g.nx), int(g.ny)
        self.dx, self.dy = g.dx, g.dy
        self.material = material
        self.lattice_speed = float(lattice_speed)
        self.collision_omega = float(collision_omega)
        self.dt = self.dx / self.lattice_speed
        self.time = 0.0
        self.steps = 0
        self.jacobian_floor = float(jacobian_floor)
        self.source_mode = source_id(source_mode)
        self.source_params = np.zeros(6)
        if self.source_mode == 1:
            vals = tuple(source_vector)[:6]
            self.source_params[: len(vals)] = vals
        elif self.source_mode == 2:
            self.source_params[0] = float(damping_gamma)
        rid = cb.boundary_reconstruction_id(boundary_reconstruction)
        solver = cb.neumann_solver_name(neumann_solver, material)
        nsm = 2 if solver == "energy" else cb.normal_jacobian_id(normal_jacobian)
        kind, mode, params = boundary.tables(cb.MAX_BOUNDARY_IDS)
        sh = (self.nx, self.ny)
        z = lambda c: np.zeros((c,) + sh)  # noqa: E731
        self.f0, self.f1, self.fpost = z(24), z(24), z(24)
        self.U, self.u, self.u_star, self.u_next = z(6), z(2), z(2), z(2)
        self.bc_diag, self.bc_rhs, self.P, self.sigma = np.zeros(sh), z(2), z(4), z(4)
        self.invalid = np.zeros(1, dtype=np.int32)
        L = g.n_links
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        if L:
            xs = (np.arange(self.nx) + 0.5) * self.dx
            ys = (np.arange(self.ny) + 0.5) * self.dy
            lsq = stencils2d.boundary_lsq_weights(
                active=i32(g.active), x_coords=xs, y_coords=ys, link_i=g.link_i, link_j=g.link_j,
                link_xb=g.link_xb, link_yb=g.link_yb, link_nx=g.link_nx, link_ny=g.link_ny, dx=self.dx,
                max_stencil=np.asarray(g.lsq_i).shape[1])
        else:
            lsq = (g.lsq_i, g.lsq_j, g.lsq_wb, g.lsq_wt, g.lsq_wx, g.lsq_wy, g.lsq_valid)
        self.lsq = dict(zip(("lsq_i", "lsq_j", "lsq_wb", "lsq_wt", "lsq_wx", "lsq_wy", "lsq_valid"), lsq))
        self._tabs = dict(
            active=i32(g.active), link_i=i32(g.link_i), link_j=i32(g.link_j), link_q=i32(g.link_q),
            link_bid=i32(g.link_boundary_id), link_eta=f64(g.link_eta), link_xb=f64(g.link_xb),
            link_yb=f64(g.link_yb), link_nx=f64(g.link_nx), link_ny=f64(g.link_ny),
            lsq_i=i32(lsq[0]), lsq_j=i32(lsq[1]), lsq_wb=f64(lsq[2]), lsq_wt=f64(lsq[3]),
            lsq_wx=f64(lsq[4]), lsq_wy=f64(lsq[5]), lsq_valid=i32(lsq[6]),
            kind_by_id=i32(kind), mode_by_id=i32(mode), params_by_id=f64(params),
        )
        self.dyn_tx = np.zeros(max(L, 1))
        self.dyn_ty = np.zeros(max(L, 1))
        s = _State()
        s.nx, s.ny, s.n_links = self.nx, self.ny, L
        s.model, s.lam, s.mu = material.id, material.lam, material.mu
        s.p0, s.p1, s.p2 = material.params
        s.j_floor = self.jacobian_floor
        s.dt, s.c, s.omega, s.dx, s.dy = self.dt, self.lattice_speed, self.collision_omega, self.dx, self.dy
        s.source_mode = self.source_mode
        for k in range(6):
            s.source_params[k] = self.source_params[k]
        s.recon, s.normal_solver = rid, nsm
        s.disp_interval = max(0, int(local_displacement_interval))
        s.disp_sweeps = int(local_displacement_sweeps)
        s.disp_relax = float(local_displacement_relax)
        s.disp_weight = float(local_displacement_boundary_weight)
        s.disp_boundary_id = int(local_displacement_boundary_id)
        s.compat_interval = max(0, int(local_compatibility_interval))
        s.compat_blend = float(local_compatibility_blend)
        s.compat_interior_only = int(bool(local_compatibility_interior_only))
        s.use_dyn = 0
        if collision_model not in ("bgk", "mrt_ghost"):
            raise ValueError("collision_model must be 'bgk' or 'mrt_ghost'")
        s.collision_model = 1 if collision_model == "mrt_ghost" else 0
        s.ghost_omega = float(self.collision_omega if ghost_omega is None else ghost_omega)
        s.filter_strength = float(kinetic_filter_strength)
        for name in ("f0", "f1", "fpost", "U", "u", "u_star", "u_next", "bc_diag", "bc_rhs", "P", "sigma",
                     "invalid"):
            setattr(s, name, _p(getattr(self, name)))
        for name, arr in self._tabs.items():
            setattr(s, name, _p(arr))
        s.dyn_tx, s.dyn_ty = _p(self.dyn_tx), _p(self.dyn_ty)
        self._s = s
        self.initialize_identity()

    # -- reference-style initialisation (cbw2d:3842-3954) --
    def initialize_identity(self, *, init_order: int = 1) -> None:
        U = np.zeros((6, self.nx, self.ny))
        U[2] = 1.0
        U[5] = 1.0
        self.initialize_from_numpy(U=U, init_order=init_order)

    def initialize_from_numpy(self, *, U, displacement=None, init_order=1, apply_initial_boundaries=True,
                              initial_boundary_time=0.0) -> None:
        from curved_lbm.two_d import vector_nonlinear_elastic_warp as ve

        U = np.asarray(U, dtype=np.float64)
        disp = np.zeros((2, self.nx, self.ny)) if displacement is None else np.asarray(displacement, np.float64)
        act = self._tabs["active"] != 0
        if init_order == 1:
            f = np.zeros((24, self.nx, self.ny))
            Ua = np.ascontiguousarray(U[:, act].T)
            fa = np.zeros((Ua.shape[0], 24))
            for k in range(Ua.shape[0]):
                self.lib.or2d_equilibrium(C.byref(self._s), _p(Ua[k]), _p(fa[k]))
            src = ve._source_np(U, self.source_mode, self.source_params)
            fa -= (0.5 * self.dt / 4.0) * np.tile(src[:, act].T, (1, 4))
            f[:, act] = fa.T
        elif init_order == 2:
            f = ve._flatten_pop(ve._second_order_populations_np(
                U, self.material, self.lattice_speed, dx=self.dx, dy=self.dy, dt=self.dt, periodic_x=False,
                periodic_y=False, source_mode=self.source_mode, source_params=self.source_params))
        elif init_order == 3:
            f = ve._flatten_pop(ve._second_order_populations_masked_np(
                U, act, self.material, self.lattice_speed, dx=self.dx, dy=self.dy, dt=self.dt,
                periodic_x=False, periodic_y=False, source_mode=self.source_mode,
                source_params=self.source_params))
        else:
            raise ValueError("init_order must be 1, 2, or 3")
        f[:, ~act] = 0.0
        us = disp.copy()
        us[:, ~act] = 0.0
        us = us - 0.5 * self.dt * U[:2]
        us[:, ~act] = 0.0
        self.f0[...] = f
        self.f1[...] = 0.0
        self.fpost[...] = 0.0
        self.u_star[...] = us
        self.time = 0.0
        self.steps = 0
        self.refresh_current()
        if apply_initial_boundaries:
            self.apply_boundary_reconstruction(float(initial_boundary_time), in_place=True)
            if self._s.filter_strength > 0.0:
                self.lib.or2d_kinetic_filter(C.byref(self._s), _p(self.f0))
            self.refresh_current()

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise ValueError("unsupported configuration for the oracle")

    def refresh_current(self) -> None:
        self._check(self.lib.or2d_refresh(C.byref(self._s)))

    def apply_boundary_reconstruction(self, t: float, *, in_place: bool = False) -> None:
        self._check(self.lib.or2d_apply_boundary(C.byref(self._s), float(t), int(in_place)))

    def set_dynamic_traction(self, tx, ty) -> None:
        self.dyn_tx[: len(tx)] = tx
        self.dyn_ty[: len(ty)] = ty
        self._s.use_dyn = 1

    def run(self, n: int) -> None:
        n = int(n)
        if n <= 0:
            return
        self._check(self.lib.or2d_step(C.byref(self._s), self.time, self.steps, n))
        if self._s.f0 != _p(self.f0):
            self.f0, self.f1 = self.f1, self.f0
        self._s.f0, self._s.f1 = _p(self.f0), _p(self.f1)
        for _ in range(n):
            self.time += self.dt
        self.steps += n

    def step(self) -> None:
        self.run(1)

    def run_until(self, t_end: float) -> None:
        n = 0
        t = self.time
        while t < float(t_end) - 0.5 * self.dt:
            t += self.dt
            n += 1
        self.run(n)
        self.refresh_current()

    def du(self) -> np.ndarray:
        out = np.zeros((4, self.nx, self.ny))
        u = np.ascontiguousarray(self.u)
        self.lib.or2d_du(C.byref(self._s), _p(u), _p(out))
        return out

    # -- accessors --
    def displacement(self) -> np.ndarray:
        return self.u.copy()

    def system_state(self) -> np.ndarray:
        return self.U.copy()

    def deformation_gradient(self) -> np.ndarray:
        U = self.U
        return np.stack([np.stack([U[2], U[3]]), np.stack([U[4], U[5]])])

    def deformation_jacobian(self) -> np.ndarray:
        return self.U[2] * self.U[5] - self.U[3] * self.U[4]

    def active_mask(self) -> np.ndarray:
        return self._tabs["active"] != 0

    def invalid_state(self) -> bool:
        if int(self.invalid[0]) != 0:
            return True
        return not np.isfinite(self.U[:, self.active_mask()]).all()

    def numpy_field(self, name: str) -> np.ndarray:
        return getattr(self, name).copy()


class OracleBox2D(OracleSolver2D):
    """CPU FP64 oracle of the rectangular-box solver
    (vector_nonlinear_elastic_warp.py:1896-2352): every node active, no cut
    links, periodic axes wrapped in the stream, half-way faces
    (_apply_boundary_side_kernel, :1685-1820)."""

    def __init__(self, nx, ny, *, material, boundaries=None, length_x=1.0, length_y=1.0, lattice_speed=5.0,
                 collision_omega=1.8, jacobian_floor=1.0e-12, source_mode="none", source_vector=(0.0,) * 6,
                 damping_gamma=0.0, threads: int | None = None, **_ignored):
        from curved_lbm.two_d import box_solver as bx
        from curved_lbm.two_d import curved_boundary_warp as cb

        empty_i, empty_f = np.zeros(0, dtype=np.int32), np.zeros(0)
        g = cb.CurvedBoundaryGeometry(
            nx=int(nx), ny=int(ny), length_x=float(length_x), length_y=float(length_y),
            active=np.ones((int(nx), int(ny)), dtype=np.int32), link_i=empty_i, link_j=empty_i, link_q=empty_i,
            link_boundary_id=empty_i, link_eta=empty_f, link_xb=empty_f, link_yb=empty_f, link_nx=empty_f,
            link_ny=empty_f, label="rectangle", lsq_i=np.zeros((0, 24), dtype=np.int32),
            lsq_j=np.zeros((0, 24), dtype=np.int32), lsq_wb=np.zeros((0, 24)), lsq_wt=np.zeros((0, 24)),
            lsq_wx=np.zeros((0, 24)), lsq_wy=np.zeros((0, 24)), lsq_valid=np.zeros(0, dtype=np.int32))
        super().__init__(g, material=material, boundary=cb.CurvedBoundarySpec(kind="neumann"),
                         lattice_speed=lattice_speed, collision_omega=collision_omega,
                         jacobian_floor=jacobian_floor, source_mode=source_mode, source_vector=source_vector,
                         damping_gamma=damping_gamma, threads=threads)
        self.length_x, self.length_y = float(length_x), float(length_y)
        self._bx = bx
        self.boundaries = boundaries or bx.RectangularBoundarySet.periodic()
        self._set_sides(self.boundaries)
        self.initialize_identity()

    def _set_sides(self, b) -> None:
        s = self._s
        s.box = 1
        for k in range(4):
            s.side_kind[k] = int(b.side_kind[k])
            s.side_mode[k] = int(b.value_mode[k])
        for k, v in enumerate(np.asarray(b.params, dtype=np.float64).reshape(32)):
            s.side_params[k] = float(v)
        s.length_x, s.length_y = self.length_x, self.length_y

    def set_boundaries(self, boundaries) -> None:
        boundaries.validate_periodic_pairs()
        self.boundaries = boundaries
        self._set_sides(boundaries)

    def initialize_from_numpy(self, *, U, displacement=None, init_order=1, enforce_dirichlet_tangent=True,
                              apply_initial_boundaries=True, initial_boundary_time=0.0) -> None:
        from curved_lbm.two_d import vector_nonlinear_elastic_warp as ve

        if not hasattr(self, "_bx"):  # the base constructor's identity start, before the faces are set
            return super().initialize_from_numpy(U=U, displacement=displacement, init_order=init_order,
                                                 apply_initial_boundaries=False)
        U = np.asarray(U, dtype=np.float64)
        disp = np.zeros((2, self.nx, self.ny)) if displacement is None else np.asarray(displacement, np.float64)
        if enforce_dirichlet_tangent:
            U = self._bx._enforce_dirichlet_tangent_columns(U, disp, self.boundaries, dx=self.dx, dy=self.dy)
        super().initialize_from_numpy(U=U, displacement=disp, init_order=1, apply_initial_boundaries=False)
        if init_order == 2:
            f = ve._second_order_populations_np(
                U, self.material, self.lattice_speed, dx=self.dx, dy=self.dy, dt=self.dt,
                periodic_x=self.boundaries.periodic_x, periodic_y=self.boundaries.periodic_y,
                source_mode=self.source_mode, source_params=self.source_params)
            self.f0[...] = f.reshape((24, self.nx, self.ny))
            self.refresh_current()
        elif init_order != 1:
            raise ValueError("init_order must be 1 or 2")
        if apply_initial_boundaries:
            self.apply_boundary_reconstruction(float(initial_boundary_time), in_place=True)
            self.refresh_current()
            self.u_star[...] = disp - 0.5 * self.dt * self.U[:2]
            self.refresh_current()
