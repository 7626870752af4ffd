"""MI355X-native curved-boundary vectorial lattice Boltzmann solvers.

Drop-in for the ``curved_lbm`` package of JingsenFeng/curved-boundary-vlbm:
``curved_lbm.three_d.curved_boundary_warp3d`` (D3Q6x12).  Time steps run in
libvlbm.so (HIP, gfx950); this package holds the host-side geometry
preprocessing and the solver classes.
"""

__version__ = "0.1.0"
