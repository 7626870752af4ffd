/* libvlbm: MI355X (gfx950) HIP implementation of the D2Q4x6 vectorial
 * lattice Boltzmann time step on curved (level-set) domains.
 *
 * C ABI only: plain pointers and sizes.  Every device pointer in a
 * vlbm2d_desc is owned by the caller and must stay valid until
 * vlbm2d_destroy.  Fields are dense (components, nx, ny) planes with node
 * index i*ny + j, the reference's (components, nx, ny, 1) arrays.
 * Populations (24 planes, index q*6 + a, directions E, N, W, S) and the
 * record (12 planes) are held with adjacent planes paired: element
 * (plane, node) sits at ((plane / 2) * nx*ny + node) * 2 + plane % 2, so a
 * node's planes 2k and 2k + 1 are one 16-byte (FP64) or 8-byte (FP32) pair;
 * a population buffer is opaque and vlbm2d_populations_to_f64 / _from_f64
 * convert to and from (24, nx, ny) double planes.  Only active nodes are
 * read or written: entries of inactive nodes keep the value the caller gave
 * them (zero in the Python layer).
 *
 * Every entry point returns 0 on success and a negative code on failure;
 * vlbm_last_error() (declared in vlbm3d.h) then describes the failure.
 * Work is enqueued on the descriptor's stream and is asynchronous unless
 * stated.
 *
 * Reference interfaces replaced (reference = JingsenFeng/curved-boundary-vlbm,
 * cbw2d = src/curved_lbm/two_d/curved_boundary_warp.py):
 *   vlbm2d_create          device-array setup in WarpCurvedBoundaryLBM2D.__init__ (cbw2d:3621-3840)
 *   vlbm2d_step            WarpCurvedBoundaryLBM2D.step (cbw2d:4180-4256), repeated nsteps times
 *   vlbm2d_refresh         refresh_current (cbw2d:4258-4288) -> _refresh_masked_kernel (cbw2d:3128-3210)
 *   vlbm2d_apply_boundary  apply_boundary_reconstruction (cbw2d:4020-4034)
 *                          -> _apply_curved_boundary_links_kernel (cbw2d:2064-3098)
 *   vlbm2d_init_populations the init_order=1 branch of initialize_from_numpy (cbw2d:3873-3876)
 *   vlbm2d_set_ustar       u_star0 = u0 - dt/2 v0 of initialize_from_numpy (cbw2d:3920-3922)
 *   vlbm2d_stress          the P / sigma planes written by the collide, refresh and
 *                          repair kernels (cbw2d:1808-1818, 3197-3210, 3602-3615)
 *   vlbm2d_post_collision  the fpost array left by _collide_masked_kernel and
 *                          _local_compatibility_repair_kernel (cbw2d:1692-1843, 3522-3615)
 *   vlbm2d_reduce          host reductions of summary()/invalid_state() (cbw2d:4307-4341)
 *   vlbm2d_kinetic_filter  _kinetic_ghost_filter_kernel (cbw2d:3100-3124) as launched by
 *                          initialize_from_numpy (cbw2d:3947-3953)
 *   vlbm2d_set_box_sides   WarpNonlinearElasticLBM2D.set_boundaries
 *                          (src/curved_lbm/two_d/vector_nonlinear_elastic_warp.py:1970-1979)
 * With desc.box the same step, refresh and boundary calls run the box solver
 * WarpNonlinearElasticLBM2D (vector_nonlinear_elastic_warp.py:1896-2352):
 * vlbm2d_step its step(), vlbm2d_apply_boundary (f_src may equal f_dst) its
 * apply_boundary_reconstruction (_apply_boundary_side_kernel, :1685-1820).
 *   vlbm2d_destroy         object lifetime end
 */
#ifndef VLBM2D_H
#define VLBM2D_H

#include <stdint.h>

#include "vlbm3d.h" /* vlbm_last_error, VLBM_PREC_*, VLBM_RED_* */

#ifdef __cplusplus
extern "C" {
#endif

#define VLBM2D_ABI_VERSION 8
#define VLBM2D_MAX_IDS 8
#define VLBM2D_MAX_LSQ 24
#define VLBM2D_RECORD 12

typedef struct vlbm2d_handle vlbm2d_handle;

typedef struct {
    uint32_t struct_size; /* sizeof(vlbm2d_desc), checked by vlbm2d_create */
    int32_t abi_version;  /* VLBM2D_ABI_VERSION */
    int32_t nx, ny;
    int32_t n_active;     /* length of active_list */
    int32_t n_links;
    int32_t precision;    /* VLBM_PREC_F64 or VLBM_PREC_F32 (float planes hold f - f_rest) */
    /* material (WarpHyperelasticMaterial) */
    int32_t model;
    double lam, mu, p0, p1, p2, j_floor;
    /* lattice and collision */
    double dt, c, omega, dx, dy;
    int32_t collision_model; /* 0 bgk, 1 mrt_ghost (moment-space collide, cbw2d:1847-2041) */
    double ghost_omega;      /* mrt_ghost relaxation rate of the ghost moment, (0, 2] */
    double kinetic_filter;   /* strength in [0, 1] of the ghost filter after every step (cbw2d:3100-3124) */
    int32_t source_mode;
    double source_params[6];
    int32_t recon;         /* reference reconstruction id 0..13: 0 bfl, 1 qbfl, 2 lsq_bfl, 3 regularized_bfl,
                              4 regularized_nee, 5 pair_nee, 6 same_nee, 7 poly_ibb, 8 lowq_reg_bfl,
                              9 compat_bfl, 10 compat_nee, 11 local_char2, 12 local_f, 13 local_f_bfl */
    int32_t normal_solver; /* 0 finite-difference Newton, 1 acoustic Newton, 2 energy */
    /* local kinematics */
    int32_t disp_interval, disp_sweeps;
    double disp_relax, disp_weight;
    int32_t compat_interval;
    double compat_blend;
    int32_t compat_interior_only;
    int32_t n_anchor_links; /* length of anchor_ids */
    /* state (device) */
    double *U, *u, *u_star, *u_next; /* (6|2, nx, ny) */
    void *record;                    /* (12, nx, ny) deferred-collision record, in the population
                                        precision and the paired-plane layout (float: A - omega
                                        f_rest on the A planes) */
    int32_t *invalid;                /* (1) */
    double *ghost;                   /* (2, 6, nx, ny) ghost moments of the pre-collision populations,
                                        two step parities (mrt_ghost only; else may be NULL) */
    double *reduce_scratch;          /* (>= 8 * 1025) */
    /* tables (device) */
    const int32_t *active;           /* (nx, ny) 0/1 */
    const int32_t *active_list;      /* (n_active) node index of every active node, ascending */
    const int32_t *anchor_off;       /* (n_active + 1) CSR offsets, in active_list order */
    const int32_t *anchor_ids;       /* Dirichlet-anchor links of each node, ascending */
    const int32_t *link_p, *link_q, *link_bid; /* node index, outgoing direction, boundary id */
    const double *link_eta, *link_xb, *link_yb, *link_nx, *link_ny;
    const int32_t *lsq_p;            /* (24, L) node index of each stencil sample, -1 if unused */
    const double *lsq_wb, *lsq_wx, *lsq_wy; /* (24, L) value, d/dx, d/dy weights */
    const double *lsq_wt;            /* (24, L) tangential-derivative weights (compat_bfl) */
    const int32_t *lsq_valid;        /* (L) */
    double *link_state;              /* (19, L) scratch: F at the cut point, dU/dx, dU/dy, and for
                                        compat_bfl the tangential derivative of u and sum |w_t| */
    const int32_t *kind_by_id;       /* (8) */
    const int32_t *mode_by_id;       /* (8) */
    const double *params_by_id;      /* (8, 8) */
    const double *dyn_tx, *dyn_ty;   /* (L) per-link nominal traction (dynamic override) */
    int32_t use_dyn;                 /* initial state of the dynamic override */
    /* rectangular box solver (WarpNonlinearElasticLBM2D): every node active, no cut links;
       faces left, right, bottom, top */
    int32_t box;                     /* 1 = box solver */
    int32_t box_side_kind[4];        /* 0 periodic (both sides of an axis), 1 Dirichlet, 2 Neumann */
    int32_t box_value_mode[4];       /* value mode of each face (0..6) */
    double box_params[32];           /* 8 params per face */
    double length_x, length_y;       /* box lengths (face coordinates) */
    void *stream;                    /* hipStream_t; NULL = null stream */
} vlbm2d_desc;

int vlbm2d_create(const vlbm2d_desc *desc, vlbm2d_handle **out);
int vlbm2d_destroy(vlbm2d_handle *h);
/* Advance nsteps steps from (time, steps).  f0 holds the current
 * populations and f1 is scratch; the buffers alternate every step, so the
 * newest populations are in f0 when nsteps is even and in f1 when odd. */
int vlbm2d_step(vlbm2d_handle *h, void *f0, void *f1, double time, int64_t steps, int32_t nsteps);
/* Moments of f into U, u = u_star + dt/2 v, invalid flag. */
int vlbm2d_refresh(vlbm2d_handle *h, const void *f);
/* Cut-link reconstruction at time t reading f_src (as post-collision
 * populations) and writing the incoming link populations of f_dst. */
int vlbm2d_apply_boundary(vlbm2d_handle *h, const void *f_src, void *f_dst, double t);
/* f = feq(U) - dt/(2Q) B(U) from the U planes at active nodes. */
int vlbm2d_init_populations(vlbm2d_handle *h, void *f);
/* u_star = disp - dt/2 v at active nodes, disp (2, nx, ny) device. */
int vlbm2d_set_ustar(vlbm2d_handle *h, const double *disp);
/* First Piola stress P and Cauchy stress sigma (4, nx, ny) each from the F of U. */
int vlbm2d_stress(vlbm2d_handle *h, double *P, double *sigma);
/* Post-collision populations of the last step: fpost = (1 - omega) f_prev +
 * record terms (+ the ghost-moment term of mrt_ghost), with f_prev the
 * populations that step started from. */
int vlbm2d_post_collision(vlbm2d_handle *h, const void *f_prev, void *fpost);
/* Replaces the four faces of a box solver (kinds, value modes, 4 x 8 params). */
int vlbm2d_set_box_sides(vlbm2d_handle *h, const int32_t *kind, const int32_t *mode, const double *params);
/* Kinetic ghost filter of the descriptor's strength, in place on f at the
 * active nodes (no-op at strength 0).  vlbm2d_step applies it after every
 * step by itself. */
int vlbm2d_kinetic_filter(vlbm2d_handle *h, void *f);
/* Enables (1) or disables (0) the per-link traction override on Neumann links. */
int vlbm2d_set_dynamic_traction(vlbm2d_handle *h, int32_t enabled);
/* Replays blocks of `steps` fused steps (an even count; 0 disables) from a
 * captured HIP graph inside vlbm2d_step, with the step time kept on the
 * device; used when the displacement and F-repair intervals are 0 or 1 and
 * timing is off.  The graph holds the descriptor's pointers and the f0/f1
 * pair of the call that captured it (one per buffer order is cached).  No
 * reference counterpart: the reference launches every kernel from Python. */
int vlbm2d_set_graph_steps(vlbm2d_handle *h, int32_t steps);
/* Synchronous reductions over the active nodes (VLBM_RED_* slots). */
int vlbm2d_reduce(vlbm2d_handle *h, const double *u_prev, double *out);
/* Conversion between the stored layout and precision and whole
 * (24, nx, ny) double planes. */
int vlbm2d_populations_to_f64(vlbm2d_handle *h, const void *f, double *dst);
int vlbm2d_populations_from_f64(vlbm2d_handle *h, const double *src, void *f);
/* Per-kernel timing, slots VLBM_SLOT_* of vlbm3d.h (no reference counterpart). */
int vlbm2d_timing_enable(vlbm2d_handle *h, int32_t enable);
int vlbm2d_timing_read(vlbm2d_handle *h, double *ms, int32_t *counts, int32_t nslots);

/* Batches: many 2D lattices stepped together, one launch per kernel per
 * step over the work of all selected shells (bitwise the results of stepping
 * each alone).  No reference counterpart (the reference steps one solver per
 * geometry, cbw2d:4163-4258).  Requirements: the same population precision,
 * BGK without the kinetic filter, a uniform schedule (intervals 0 or 1, one
 * sweep), no box solvers.  f0[k] / f1[k] are shell k's population buffers;
 * the populations are in f0 while vlbm2d_batch_parity() is 0 and in f1 while
 * it is 1 (an odd step count flips it).  vlbm2d_batch_select sets the
 * shells that step (ids strictly increasing) with their current population
 * buffers (f0: the populations, f1: scratch) and resets the parity to 0.
 * times: the start time of each selected shell (host array, selection
 * order). */
typedef struct vlbm2d_batch vlbm2d_batch;
int vlbm2d_batch_create(vlbm2d_handle *const *handles, void *const *f0, void *const *f1, int32_t n, void *stream,
                        vlbm2d_batch **out);
int vlbm2d_batch_select(vlbm2d_batch *b, const int32_t *ids, void *const *f0, void *const *f1, int32_t count);
int vlbm2d_batch_set_graph_steps(vlbm2d_batch *b, int32_t steps);
int vlbm2d_batch_parity(const vlbm2d_batch *b);
int vlbm2d_batch_step(vlbm2d_batch *b, const double *times, int32_t nsteps);
int vlbm2d_batch_destroy(vlbm2d_batch *b);
/* Block shapes of the batched launches: threads per block of the node
 * kernels (64, 128, 256 or 512) and of the cut-link kernel (32, 64 or 128).
 * vlbm2d_batch_autotune times `steps` batched steps of the selected shells
 * with every pair (HIP events on the batch stream), keeps the fastest and
 * returns it (ms_out, optional: 12 per-step times, node-block major).  It
 * advances the shells' state: tune before initialising them. */
int vlbm2d_batch_autotune(vlbm2d_batch *b, int32_t steps, int32_t *nb_out, int32_t *nl_out, float *ms_out);
int vlbm2d_batch_set_blocks(vlbm2d_batch *b, int32_t nb, int32_t nl);
#ifdef __cplusplus
}
#endif
#endif
