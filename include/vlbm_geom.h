/* libvlbm host-side geometry preprocessing: per-link stencil weights of the
 * curved-boundary closures, built once per geometry on the host (threads
 * over links).  Plain host pointers; nothing here touches the GPU, so these
 * entry points also run where no HIP device is present.
 *
 * Reference interfaces replaced (reference = JingsenFeng/curved-boundary-vlbm):
 *   vlbm_geom3d_char_weights  build_characteristic_weights
 *                             (src/curved_lbm/three_d/characteristic_stencil3d.py:19-171)
 *   vlbm_geom3d_tangent_weights  _build_sparse_tangent_lsq_weights
 *                             (src/curved_lbm/three_d/curved_boundary_warp3d.py:3725-3849)
 *   vlbm_geom2d_lsq_weights   _build_boundary_lsq_weights with _local_quadratic_fit
 *                             (src/curved_lbm/two_d/curved_boundary_warp.py:332-606)
 *
 * Sample selection, slot assignment, probe modes and validity follow those
 * functions exactly; the weights agree with them to rounding (the
 * least-squares systems are solved by LU with partial pivoting).
 */
#ifndef VLBM_GEOM_H
#define VLBM_GEOM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VLBM_GEOM3D_GRADIENT_SAMPLES 48
#define VLBM_GEOM3D_PROBE_SAMPLES 16

/* Characteristic stencils of the n_links cut links of a D3Q6 lattice.
 * active_id: (nx, ny, nz) active-node index or -1.  Per link: node (li, lj,
 * lk), outgoing direction lq (+x,+y,+z,-x,-y,-z), cut point (xb, yb, zb) in
 * physical units, outward normal (lnx, lny, lnz), fraction eta; dx the
 * (isotropic) spacing.  Outputs: indices (64, n_links) (-1 unused; slots
 * 0-47 gradient samples, 48-63 predictor samples), weights (4, 64, n_links)
 * (row 0 predictor, rows 1-3 d/dx, d/dy, d/dz), modes (n_links) (2 two
 * probes, 1 one probe, -k a k-node link extrapolation), status (n_links):
 * 0 ok, 1 no full-rank quadratic stencil (fewer than 10 samples or
 * cond > 1e8), 2 invalid gradient weights.  nthreads <= 0 picks the host
 * thread count.  Returns 0, or -1 for invalid arguments. */
int vlbm_geom3d_char_weights(const int32_t *active_id, int32_t nx, int32_t ny, int32_t nz, int64_t n_links,
                             const int32_t *li, const int32_t *lj, const int32_t *lk, const int32_t *lq,
                             const double *xb, const double *yb, const double *zb, const double *lnx,
                             const double *lny, const double *lnz, const double *eta, double dx, int32_t nthreads,
                             int32_t *indices, double *weights, int32_t *modes, uint8_t *status);

/* Surface-tangent derivative stencils of the 3D compat_bfl closure
 * (replaces _build_sparse_tangent_lsq_weights, curved_boundary_warp3d.py:
 * 3725-3849): per link, up to max_stencil active nodes (ascending distance)
 * and the weights of d/dt1 and d/dt2; (n_links, max_stencil) row-major
 * outputs out_p (active-node index, -1 unused), wt1, wt2, and valid
 * (n_links).  dx, dy, dz the spacings.  Returns 0, or -1 for invalid
 * arguments. */
int vlbm_geom3d_tangent_weights(const int32_t *active_id, int32_t nx, int32_t ny, int32_t nz, int64_t n_links,
                                const int32_t *li, const int32_t *lj, const int32_t *lk, const double *xb,
                                const double *yb, const double *zb, const double *lnx, const double *lny,
                                const double *lnz, double dx, double dy, double dz, int32_t max_stencil,
                                int32_t nthreads, int32_t *out_p, double *wt1, double *wt2, int32_t *valid);

/* Boundary stencils of the n_links cut links of a D2Q4 lattice.  active:
 * (nx, ny) 0/1; x_coords (nx), y_coords (ny) node coordinates; per link:
 * node (li, lj), cut point (xb, yb), outward normal (lnx, lny); dx the
 * spacing; max_stencil slots per link.  Outputs (n_links, max_stencil):
 * sample node out_i/out_j (-1 unused), value weights w, tangential wt and
 * Cartesian wx/wy derivative weights; valid (n_links).  Returns 0, or -1
 * for invalid arguments. */
int vlbm_geom2d_lsq_weights(const int32_t *active, int32_t nx, int32_t ny, const double *x_coords,
                            const double *y_coords, int64_t n_links, const int32_t *li, const int32_t *lj,
                            const double *xb, const double *yb, const double *lnx, const double *lny, double dx,
                            int32_t max_stencil, int32_t nthreads, int32_t *out_i, int32_t *out_j, double *w,
                            double *wt, double *wx, double *wy, int32_t *valid);

#ifdef __cplusplus
}
#endif
#endif
