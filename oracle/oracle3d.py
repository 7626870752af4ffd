"""Python harness of the CPU FP64 oracle (oracle/vlbm3d_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.

``OracleSolver3D`` takes the same constructor arguments as
``curved_lbm.three_d.curved_boundary_warp3d.WarpCurvedBoundaryLBM3D`` and
builds the same host tables (active ids, neighbours, cut links,
characteristic weights) with the product's geometry code, then advances the
state with the C restatement of the reference step.  State arrays are numpy
arrays in the reference layout.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from . import stencils3d

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "libvlbm3d_oracle.so"

_dp = C.c_void_p


class _State(C.Structure):
    _fields_ = [
        ("n_active", C.c_int32), ("n_links", C.c_int32),
        ("model", C.c_int32),
        ("lam", C.c_double), ("mu", C.c_double), ("p0", C.c_double), ("p1", C.c_double),
        ("p2", C.c_double), ("j_floor", C.c_double),
        ("dt", C.c_double), ("c", C.c_double), ("omega", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("source_mode", C.c_int32), ("source_params", C.c_double * 12),
        ("recon", C.c_int32), ("normal_solver", C.c_int32),
        ("disp_interval", C.c_int32), ("disp_sweeps", C.c_int32),
        ("disp_relax", C.c_double), ("disp_weight", C.c_double),
        ("disp_boundary_id", C.c_int32), ("compat_interval", C.c_int32),
        ("compat_blend", C.c_double), ("has_char", C.c_int32),
        ("f0", _dp), ("f1", _dp), ("U", _dp), ("u", _dp), ("u_star", _dp), ("u_next", _dp),
        ("bc_diag", _dp), ("bc_rhs", _dp), ("invalid", _dp),
        ("neighbor", _dp), ("active_i", _dp), ("active_j", _dp), ("active_k", _dp),
        ("link_p", _dp), ("link_up", _dp), ("link_q", _dp), ("link_bid", _dp),
        ("link_eta", _dp), ("link_xb", _dp), ("link_yb", _dp), ("link_zb", _dp),
        ("link_nx", _dp), ("link_ny", _dp), ("link_nz", _dp),
        ("char_indices", _dp), ("char_weights", _dp), ("char_state", _dp),
        ("tan_p", _dp), ("tan_w1", _dp), ("tan_w2", _dp), ("tan_valid", _dp),
        ("kind_by_id", _dp), ("comp_kind_by_id", _dp), ("mode_by_id", _dp), ("params_by_id", _dp),
        ("dyn_enabled", _dp), ("dyn_traction", _dp),
        ("lkind", _dp), ("lcomp_kind", _dp), ("lbv", _dp), ("lFb", _dp), ("lPb", _dp),
        ("box", C.c_int32), ("box_n", C.c_int32 * 3), ("side_kind", C.c_int32 * 6), ("side_mode", C.c_int32 * 6),
        ("side_params", C.c_double * 96), ("box_length", C.c_double * 3), ("box_speed", C.c_double * 3),
    ]


_lib = None


def build(quiet: bool = True) -> Path:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        for name in ("or3d_refresh", "or3d_step", "or3d_apply_boundary", "or3d_set_ustar",
                     "or3d_init_populations", "or3d_threads"):
            getattr(lib, name).restype = C.c_int
        lib.or3d_step.argtypes = [C.POINTER(_State), C.c_double, C.c_int64, C.c_int]
        lib.or3d_apply_boundary.argtypes = [C.POINTER(_State), C.c_double, C.c_int]
        lib.or3d_refresh.argtypes = [C.POINTER(_State)]
        lib.or3d_set_ustar.argtypes = [C.POINTER(_State), _dp]
        lib.or3d_init_populations.argtypes = [C.POINTER(_State), _dp]
        lib.or3d_piola.argtypes = [_dp, C.c_int] + [C.c_double] * 6 + [_dp]
        lib.or3d_piola.restype = None
        lib.or3d_du_axis_all.argtypes = [C.POINTER(_State), _dp, _dp]
        lib.or3d_du_axis_all.restype = None
        lib.or3d_value_probe.argtypes = [_dp, C.c_int, C.c_int] + [C.c_double] * 7 + [C.c_int]
        lib.or3d_value_probe.restype = C.c_double
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def piola(F: np.ndarray, material, j_floor: float = 1e-12) -> np.ndarray:
    """Oracle constitutive map for a single 3x3 F."""
    lib = load()
    Fi = np.ascontiguousarray(np.asarray(F, np.float64).reshape(9))
    P = np.zeros(9)
    lib.or3d_piola(_p(Fi), material.id, material.lam, material.mu, material.param0, material.param1,
                   material.param2, j_floor, _p(P))
    return P.reshape(3, 3)


def boundary_value(mode, params, comp, x, y, z, n, t, *, displacement=False) -> float:
    """Oracle boundary value (or, with ``displacement``, the Dirichlet
    anchor displacement) of one value mode at one point."""
    prm = np.zeros(16)
    vals = np.asarray(params, np.float64).ravel()[:16]
    prm[: vals.size] = vals
    return float(load().or3d_value_probe(_p(prm), int(mode), int(comp), float(x), float(y), float(z),
                                         float(n[0]), float(n[1]), float(n[2]), float(t),
                                         1 if displacement else 0))


class OracleSolver3D:
    """The reference 3D sparse step on the CPU with the reference's API subset
    used by the tests: initialize_*, step/run, apply_boundary_reconstruction,
    refresh_current, and numpy state attributes f0, f1, U, u, u_star, link_*."""

    def __init__(self, geometry, *, material, boundary, lattice_speed=5.0, collision_omega=1.8,
                 boundary_reconstruction="bfl", jacobian_floor=1e-12, source_mode="none",
                 source_vector=(0.0,) * 12, damping_gamma=0.0, local_displacement_interval=0,
                 local_displacement_sweeps=1, local_displacement_relax=0.85,
                 local_displacement_boundary_weight=20.0, local_displacement_boundary_id=-1,
                 local_compatibility_interval=0, local_compatibility_blend=0.1, normal_jacobian=None,
                 neumann_solver=None, check_linearized_cfl=True, threads: int | None = None,
                 tables=None, characteristic_stencil=None, tangent_lsq=None):
        from curved_lbm.three_d import curved_boundary_warp3d as cb

        self.cb = cb
        self.geometry = geometry
        self.material = material
        self.lattice_speed = float(lattice_speed)
        self.collision_omega = float(collision_omega)
        self.dx, self.dy, self.dz = geometry.dx, geometry.dy, geometry.dz
        self.dt = self.dx / self.lattice_speed
        self.jacobian_floor = float(jacobian_floor)
        self.recon = cb._recon_id(boundary_reconstruction)
        if check_linearized_cfl and material.linearized_stability_ratio(self.lattice_speed) >= 1.0:
            raise ValueError("linearized D3Q6 stability ratio must be below one")
        nj = cb._normal_jacobian_for_material(material, normal_jacobian)
        solver = cb.NORMAL_SOLVER_ENERGY if cb.neumann_solver_name(neumann_solver, material) == "energy" else nj
        self.source_mode = cb.source_id(source_mode)
        sp = np.zeros(12)
        if self.source_mode == cb.SOURCE_CONSTANT:
            vals = tuple(source_vector)[:12]
            sp[: len(vals)] = vals
        elif self.source_mode == cb.SOURCE_DAMPING:
            sp[0] = float(damping_gamma)
        self.source_params = sp
        if threads is not None:
            os.environ["OMP_NUM_THREADS"] = str(int(threads))
        tables = cb._sparse_tables_from_geometry(geometry) if tables is None else tables
        self.tables = tables
        n = int(tables["active_i"].size)
        L = geometry.n_links
        self.n_active = n
        kind, mode, params = boundary.tables(cb.MAX_BOUNDARY_IDS)
        ck = boundary.component_kind_table(cb.MAX_BOUNDARY_IDS)
        lids = np.clip(np.asarray(geometry.link_boundary_id, np.int32), 0, cb.MAX_BOUNDARY_IDS - 1)
        neu_rows = ck[:, lids] == cb.BC_NEUMANN
        self.characteristic_stencil = characteristic_stencil
        if characteristic_stencil is None and self.recon == cb.BOUNDARY_RECON_LOCAL_F and L and np.any(neu_rows):
            self.characteristic_stencil = stencils3d.characteristic_weights(
                geometry, tables, link_mask=np.any(neu_rows, axis=0))
        i32, f64 = np.int32, np.float64
        a = self._arrays = {}

        def keep(name, arr, dtype):
            a[name] = np.ascontiguousarray(np.asarray(arr, dtype=dtype))
            return a[name]

        self.f0 = keep("f0", np.zeros((72, n)), f64)
        self.f1 = keep("f1", np.zeros((72, n)), f64)
        self.U = keep("U", np.zeros((12, n)), f64)
        self.u = keep("u", np.zeros((3, n)), f64)
        self.u_star = keep("u_star", np.zeros((3, n)), f64)
        keep("u_next", np.zeros((3, n)), f64)
        keep("bc_diag", np.zeros((3, n)), f64)
        keep("bc_rhs", np.zeros((3, n)), f64)
        self.invalid = keep("invalid", np.zeros(1), i32)
        self.neighbor = keep("neighbor", tables["neighbor"], i32)
        for k in ("active_i", "active_j", "active_k"):
            keep(k, tables[k], i32)
        self.link_p = keep("link_p", tables["link_p"], i32)
        keep("link_up", tables["link_upstream"], i32)
        keep("link_q", geometry.link_q, i32)
        keep("link_bid", geometry.link_boundary_id, i32)
        for k in ("eta", "xb", "yb", "zb", "nx", "ny", "nz"):
            keep("link_" + k, getattr(geometry, "link_" + k), f64)
        if self.characteristic_stencil is not None:
            keep("char_indices", self.characteristic_stencil["indices"], i32)
            keep("char_weights", self.characteristic_stencil["weights"], f64)
            self.char_state = keep("char_state", np.zeros((4, 12, L)), f64)
        else:
            keep("char_indices", np.zeros(1), i32)
            keep("char_weights", np.zeros(1), f64)
            self.char_state = keep("char_state", np.zeros(1), f64)
        self.tangent_lsq = None
        if self.recon == cb.BOUNDARY_RECON_COMPAT_BFL:
            if tangent_lsq is None:
                tangent_lsq = stencils3d.tangent_weights(geometry, tables)
            self.tangent_lsq = tangent_lsq
            for k in ("lsq_p", "lsq_wt1", "lsq_wt2"):
                if np.asarray(tangent_lsq[k]).shape != (L, 32):
                    raise ValueError(f"{k} must have shape (n_links, 32)")
            keep("tan_p", tangent_lsq["lsq_p"], i32)
            keep("tan_w1", tangent_lsq["lsq_wt1"], f64)
            keep("tan_w2", tangent_lsq["lsq_wt2"], f64)
            keep("tan_valid", tangent_lsq["lsq_valid"], i32)
        else:
            for k in ("tan_p", "tan_valid"):
                keep(k, np.zeros(1), i32)
            for k in ("tan_w1", "tan_w2"):
                keep(k, np.zeros(1), f64)
        keep("kind_by_id", kind, i32)
        keep("comp_kind_by_id", ck, i32)
        keep("mode_by_id", mode, i32)
        keep("params_by_id", params, f64)
        self.dyn_enabled = keep("dyn_enabled", np.zeros(max(L, 1)), i32)
        self.dyn_traction = keep("dyn_traction", np.zeros((3, max(L, 1))), f64)
        self.link_kind = keep("lkind", np.zeros(max(L, 1)), i32)
        self.link_component_kind = keep("lcomp_kind", np.zeros((3, max(L, 1))), i32)
        self.link_bv = keep("lbv", np.zeros((3, max(L, 1))), f64)
        self.link_Fb = keep("lFb", np.zeros((9, max(L, 1))), f64)
        self.link_Pb = keep("lPb", np.zeros((9, max(L, 1))), f64)

        s = self._s = _State()
        s.n_active, s.n_links = n, L
        s.model, s.lam, s.mu = material.id, material.lam, material.mu
        s.p0, s.p1, s.p2, s.j_floor = material.param0, material.param1, material.param2, self.jacobian_floor
        s.dt, s.c, s.omega = self.dt, self.lattice_speed, self.collision_omega
        s.dx, s.dy, s.dz = self.dx, self.dy, self.dz
        s.source_mode = self.source_mode
        for k in range(12):
            s.source_params[k] = sp[k]
        s.recon, s.normal_solver = self.recon, solver
        s.disp_interval = int(max(0, local_displacement_interval))
        s.disp_sweeps = int(local_displacement_sweeps)
        s.disp_relax = float(local_displacement_relax)
        s.disp_weight = float(local_displacement_boundary_weight)
        s.disp_boundary_id = int(local_displacement_boundary_id)
        s.compat_interval = int(max(0, local_compatibility_interval))
        s.compat_blend = float(local_compatibility_blend)
        s.has_char = 1 if self.characteristic_stencil is not None else 0
        for name, arr in a.items():
            setattr(s, name, _p(arr))
        self._lib = load()
        self.time = 0.0
        self.steps = 0
        self.initialize_identity()

    # population buffers alternate like the reference's f0/f1 swap
    def _sync_pointers(self):
        if self._s.f0 != self.f0.ctypes.data:
            self.f0, self.f1 = self.f1, self.f0

    def active_coordinates(self):
        t = self.tables
        return ((t["active_i"] + 0.5) * self.dx, (t["active_j"] + 0.5) * self.dy, (t["active_k"] + 0.5) * self.dz)

    def refresh_current(self):
        self._lib.or3d_refresh(C.byref(self._s))

    def apply_boundary_reconstruction(self, t, *, in_place=False):
        self._lib.or3d_apply_boundary(C.byref(self._s), float(t), 1 if in_place else 0)

    def initialize_identity(self):
        U = np.zeros((12, self.n_active))
        U[3] = U[7] = U[11] = 1.0
        self.initialize_from_numpy(U=U, displacement=np.zeros((3, self.n_active)))

    def initialize_from_numpy(self, *, U, displacement=None, apply_initial_boundaries=True,
                              initial_boundary_time=0.0):
        U = np.ascontiguousarray(np.asarray(U, np.float64))
        disp = np.zeros((3, self.n_active)) if displacement is None else np.ascontiguousarray(displacement, np.float64)
        self._lib.or3d_init_populations(C.byref(self._s), _p(U))
        self.f1[...] = 0.0
        self.u_star[...] = disp - 0.5 * self.dt * U[:3]
        self.time = 0.0
        self.steps = 0
        self.refresh_current()
        if apply_initial_boundaries:
            self.apply_boundary_reconstruction(initial_boundary_time, in_place=True)
            self.refresh_current()
            self._lib.or3d_set_ustar(C.byref(self._s), _p(disp))
            self.refresh_current()

    def run(self, nsteps):
        self._lib.or3d_step(C.byref(self._s), float(self.time), int(self.steps), int(nsteps))
        for _ in range(int(nsteps)):
            self.time += self.dt
        self.steps += int(nsteps)
        self._sync_pointers()

    def step(self):
        self.run(1)

    def du_axis_all(self, u):
        u = np.ascontiguousarray(np.asarray(u, np.float64))
        out = np.zeros((3, 3, self.n_active))
        self._lib.or3d_du_axis_all(C.byref(self._s), _p(u), _p(out))
        return out

    def rms_v(self):
        self.refresh_current()
        return float(np.sqrt(np.mean(np.sum(self.U[:3] ** 2, axis=0))))


class OracleBox3D(OracleSolver3D):
    """CPU FP64 oracle of the rectangular-box solver
    (vector_nonlinear_elastic_warp3d.py:2305-2744): every node active, no
    cut links, periodic axes wrapped in the neighbour table, half-way faces
    (_apply_boundary_side_kernel, :1999-2200).  lattice_speed is one speed or
    the three per-axis speeds of one common dt (:2342-2361, the anisotropic
    equilibrium of :1240-1367 and the face-normal speed of :2123-2127)."""

    def __init__(self, nx, ny, nz, *, material, boundaries=None, length_x=1.0, length_y=1.0, length_z=1.0,
                 lattice_speed=5.0, collision_omega=1.8, jacobian_floor=1e-12, source_mode="none",
                 source_vector=(0.0,) * 12, damping_gamma=0.0, threads: int | None = None):
        from curved_lbm.three_d import box_solver3d as bx
        from curved_lbm.three_d import curved_boundary_warp3d as cb

        ei, ef = np.zeros(0, dtype=np.int32), np.zeros(0)
        g = cb.CurvedBoundaryGeometry3D(
            nx=int(nx), ny=int(ny), nz=int(nz), length_x=float(length_x), length_y=float(length_y),
            length_z=float(length_z), active=np.ones((int(nx), int(ny), int(nz)), dtype=bool), link_i=ei,
            link_j=ei, link_k=ei, link_q=ei, link_boundary_id=ei, link_eta=ef, link_xb=ef, link_yb=ef, link_zb=ef,
            link_nx=ef, link_ny=ef, link_nz=ef, label="box")
        self._bx, self._cb = bx, cb
        self.boundaries = boundaries or bx.BoxBoundarySet.periodic()
        self.length = (float(length_x), float(length_y), float(length_z))
        tables = cb._sparse_tables_from_geometry(g)
        self._box_shape = (int(nx), int(ny), int(nz))
        tables["neighbor"] = self._neighbor(tables, self.boundaries)
        from curved_lbm.three_d import vector_nonlinear_elastic_warp3d as ve3

        self.lattice_speeds = ve3._speed_tuple(lattice_speed)
        h = (float(length_x) / int(nx), float(length_y) / int(ny), float(length_z) / int(nz))
        dts = [h[a] / self.lattice_speeds[a] for a in range(3)]
        if max(abs(d - dts[0]) for d in dts[1:]) > 1e-12 * max(1.0, abs(dts[0])):
            raise ValueError("directional spacings and lattice speeds must define one common dt")
        super().__init__(g, material=material, boundary=cb.CurvedBoundarySpec3D(kind="neumann"),
                         lattice_speed=self.lattice_speeds[0], collision_omega=collision_omega,
                         boundary_reconstruction="bfl", jacobian_floor=jacobian_floor, source_mode=source_mode,
                         source_vector=source_vector, damping_gamma=damping_gamma, threads=threads, tables=tables,
                         check_linearized_cfl=False)
        self._set_faces(self.boundaries)
        self.initialize_identity()

    def _neighbor(self, tables, b):
        return self._cb.periodic_neighbor_table(tables, self._box_shape, [b.periodic_x, b.periodic_y, b.periodic_z])

    def _set_faces(self, b) -> None:
        s = self._s
        s.box = 1
        for a in range(3):
            s.box_n[a] = self._box_shape[a]
            s.box_length[a] = self.length[a]
            s.box_speed[a] = self.lattice_speeds[a]
        for k in range(6):
            s.side_kind[k] = int(b.side_kind[k])
            s.side_mode[k] = int(b.value_mode[k])
        for k, v in enumerate(np.asarray(b.params, dtype=np.float64).reshape(96)):
            s.side_params[k] = float(v)

    def set_boundaries(self, boundaries) -> None:
        boundaries.validate_periodic_pairs()
        self.boundaries = boundaries
        self.neighbor[...] = self._neighbor(self.tables, boundaries)
        self._set_faces(boundaries)

    def initialize_from_numpy(self, *, U, displacement=None, init_order=1, enforce_dirichlet_tangent=True,
                              apply_initial_boundaries=True, initial_boundary_time=0.0):
        if not hasattr(self, "_s") or not getattr(self._s, "box", 0):
            return super().initialize_from_numpy(U=U, displacement=displacement, apply_initial_boundaries=False)
        from curved_lbm.three_d import vector_nonlinear_elastic_warp3d as ve3

        shape = self._box_shape
        n = self.n_active
        U = np.asarray(U, dtype=np.float64).reshape((12,) + shape)
        disp = np.zeros((3,) + shape) if displacement is None else np.asarray(displacement, np.float64).reshape(
            (3,) + shape)
        if enforce_dirichlet_tangent:
            U = self._bx._enforce_dirichlet_tangent_columns(U, disp, self.boundaries, dx=self.dx, dy=self.dy,
                                                            dz=self.dz)
        super().initialize_from_numpy(U=U.reshape(12, n), displacement=disp.reshape(3, n),
                                      apply_initial_boundaries=False)
        if init_order == 2:
            b = self.boundaries
            f = ve3._second_order_populations_np(
                U, self.material, self.lattice_speeds, dx=self.dx, dy=self.dy, dz=self.dz, dt=self.dt,
                periodic_x=b.periodic_x, periodic_y=b.periodic_y, periodic_z=b.periodic_z,
                source_mode=self.source_mode, source_params=self.source_params)
            self.f0[...] = f.reshape(72, n)
            self.refresh_current()
        elif init_order != 1:
            raise ValueError("init_order must be 1 or 2")
        if apply_initial_boundaries:
            self.apply_boundary_reconstruction(float(initial_boundary_time), in_place=True)
            self.refresh_current()
            self._lib.or3d_set_ustar(C.byref(self._s), _p(np.ascontiguousarray(disp.reshape(3, n))))
            self.refresh_current()
