"""Shared 2D cases for the HIP parity tests: (geometry, solver kwargs,
optional initial state).  Kwargs are valid for both the HIP solver and the
oracle."""

import numpy as np

from curved_lbm.two_d import curved_boundary_warp as cb
from curved_lbm.two_d import vector_nonlinear_elastic_warp as ve


def material(name="neo_hooke"):
    if name == "yeoh":
        return cb.WarpHyperelasticMaterial.from_poisson("yeoh", 0.2, yeoh_c2=0.05, yeoh_c3=0.01)
    if name == "sugiyama_svk":
        return cb.WarpHyperelasticMaterial(model="sugiyama_svk", lam=2.0 / 3.0, mu=1.0, param0=0.3)
    return cb.WarpHyperelasticMaterial.from_poisson(name, 0.2, mu=1.0)


def settings(**over):
    """The reference's common traction settings (cases/traction_settings.py)
    with gamma = 10 damping and all-active F repair."""
    kw = dict(lattice_speed=10.0, collision_omega=1.8, boundary_reconstruction="local_f", source_mode="damping",
              damping_gamma=10.0, local_compatibility_interval=1, local_compatibility_blend=0.5,
              local_compatibility_interior_only=False, local_displacement_interval=1, local_displacement_sweeps=1,
              local_displacement_relax=0.85, local_displacement_boundary_weight=100.0,
              local_displacement_boundary_id=cb.BOUNDARY_ID_INNER)
    kw.update(over)
    return kw


def annulus_spec(mode=cb.CURVED_VALUE_RADIAL_TRACTION_LINEAR_RAMP, params=(0.05, 0.05)):
    """Outer boundary (id 1) Neumann with the given traction mode, inner
    boundary (id 2) clamped Dirichlet (zero velocity)."""
    return cb.CurvedBoundarySpec(kind="neumann", value_mode=mode, params=params,
                                 id_kinds={cb.BOUNDARY_ID_OUTER: "neumann", cb.BOUNDARY_ID_INNER: "dirichlet"},
                                 id_value_modes={cb.BOUNDARY_ID_OUTER: mode, cb.BOUNDARY_ID_INNER: cb.CURVED_VALUE_ZERO},
                                 id_params={cb.BOUNDARY_ID_OUTER: params, cb.BOUNDARY_ID_INNER: ()})


def annulus(n, *, model="neo_hooke", mode=cb.CURVED_VALUE_RADIAL_TRACTION_LINEAR_RAMP, params=(0.05, 0.05),
            ri=0.2, ro=0.45, **over):
    g = cb.CurvedBoundaryGeometry.annulus(nx=n, ny=n, inner_radius=ri, outer_radius=ro)
    kw = settings(**over)
    kw.update(material=material(model), boundary=annulus_spec(mode, params))
    return g, kw, None


def circle_dirichlet(n, recon, affine=False):
    g = cb.CurvedBoundaryGeometry.circle(nx=n, ny=n, radius=0.38)
    if affine:
        spec = cb.CurvedBoundarySpec(kind="dirichlet", value_mode=ve.VALUE_AFFINE_VELOCITY,
                                     params=(0.01, 0.002, 0.0, -0.003, -0.01, 0.0, 0.004))
    else:
        spec = cb.CurvedBoundarySpec(kind="dirichlet", value_mode=ve.VALUE_CONSTANT, params=(0.002, -0.001))
    kw = settings(boundary_reconstruction=recon, local_displacement_boundary_id=-1,
                  local_compatibility_interior_only=True, source_mode="none")
    kw.update(material=material(), boundary=spec)
    return g, kw, None


def circle_traction(n):
    g = cb.CurvedBoundaryGeometry.circle(nx=n, ny=n, radius=0.38)
    spec = cb.CurvedBoundarySpec(kind="neumann", value_mode=cb.CURVED_VALUE_PIOLA_NORMAL,
                                 params=(0.02, 0.005, 0.005, -0.01))
    kw = settings(source_mode="constant", source_vector=(0.01, -0.02, 0.0, 0.001, 0.0, 0.0),
                  local_displacement_interval=0)
    kw.update(material=material(), boundary=spec)
    rng = np.random.default_rng(2)
    U = np.zeros((6, n, n))
    U[2] = 1.0
    U[5] = 1.0
    U[:, :, :] += 1e-3 * rng.standard_normal((6, n, n))
    return g, kw, dict(U=U, displacement=np.zeros((2, n, n)))
