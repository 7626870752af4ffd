/* libvlbm: MI355X (gfx950) HIP implementation of the D3Q6x12 vectorial
 * lattice Boltzmann time step on curved domains.
 *
 * C ABI only: plain pointers and sizes.  Every device pointer in a
 * vlbm3d_desc is owned by the caller (the Python layer allocates them as
 * device tensors) and must stay valid until vlbm3d_destroy.  Fields are
 * stored as (components, n_active) planes.  Populations (72 planes, index
 * q*12 + a, directions (+x, +y, +z, -x, -y, -z)) and the record (24 planes)
 * are held in 16-byte groups of adjacent planes: FP64 element (plane, node)
 * sits at ((plane / 2) * n_active + node) * 2 + plane % 2 (pairs), FP32 at
 * ((plane / 4) * n_active + node) * 4 + plane % 4 (quads).  A population
 * buffer is therefore opaque; vlbm3d_populations_to_f64 / _from_f64 convert
 * to and from (72, n_active) double planes.
 *
 * Every entry point returns 0 on success and a negative code on failure;
 * vlbm_last_error() then describes the failure.  Work is enqueued on the
 * stream given in the descriptor and is asynchronous unless stated.
 *
 * Reference interfaces replaced (reference = JingsenFeng/curved-boundary-vlbm,
 * cbw3d = src/curved_lbm/three_d/curved_boundary_warp3d.py):
 *   vlbm3d_create          device-array setup in WarpSparseCurvedBoundaryLBM3D.__init__ (cbw3d:3859-4071)
 *   vlbm3d_step            WarpSparseCurvedBoundaryLBM3D.step (cbw3d:4549-4589), repeated nsteps times
 *   vlbm3d_refresh         refresh_current (cbw3d:4401-4425) -> _sparse_refresh_kernel (cbw3d:2110-2189)
 *   vlbm3d_apply_boundary  apply_boundary_reconstruction (cbw3d:4339-4399)
 *   vlbm3d_init_populations the population part of _initialize_from_active (cbw3d:4251-4262)
 *   vlbm3d_set_ustar       _sparse_set_ustar launch (cbw3d:2097-2106, 4270-4278)
 *   vlbm3d_reduce          host reductions of summary()/invalid_state()/rms_v
 *                          (cbw3d:4667-4769; cases/spherical_shell.py:351-361)
 *   vlbm3d_destroy         object lifetime end (no reference counterpart beyond GC)
 */
#ifndef VLBM3D_H
#define VLBM3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VLBM3D_ABI_VERSION 9
#define VLBM3D_MAX_IDS 16
#define VLBM3D_NPRM 16 /* parameters per boundary id: params_by_id is (MAX_IDS, NPRM) */
#define VLBM3D_MAX_CHAR 64
#define VLBM3D_RECORD 24

/* Precision of the population planes. */
#define VLBM_PREC_F64 0
#define VLBM_PREC_F32 1 /* f stored as float deviation from the rest state */

typedef struct vlbm3d_handle vlbm3d_handle;

typedef struct {
    uint32_t struct_size; /* sizeof(vlbm3d_desc), checked by vlbm3d_create */
    int32_t abi_version;  /* VLBM3D_ABI_VERSION */
    int32_t n_active;
    int32_t n_links;
    int32_t precision;    /* VLBM_PREC_* */
    /* material (WarpHyperelasticMaterial3D) */
    int32_t model;
    double lam, mu, p0, p1, p2, j_floor;
    /* lattice and collision */
    double dt, c, omega, dx, dy, dz;
    int32_t source_mode;
    double source_params[12];
    int32_t recon;         /* 0 bfl, 1 halfway, 2 compat_bfl, 3 local_f, 4 local_f_bfl */
    int32_t normal_solver; /* 0 finite-difference Newton, 1 acoustic Newton, 2 energy */
    /* local kinematics */
    int32_t disp_interval, disp_sweeps;
    double disp_relax, disp_weight;
    int32_t disp_boundary_id;
    int32_t compat_interval;
    double compat_blend;
    int32_t has_char;
    int32_t n_anchor_links; /* length of node_link_ids */
    int32_t n_char_links;   /* number of traction links with a characteristic stencil */
    int32_t n_char_nodes;   /* number of distinct stencil sample nodes */
    /* state (device) */
    double *U, *u, *u_star, *u_next; /* (12,n), (3,n), (3,n), (3,n) */
    void *record;                    /* (24,n) deferred-collision record, in the population precision and
                                        the grouped-plane layout (float records hold A - omega f_rest on
                                        the A planes) */
    int32_t *invalid;                /* (1) */
    double *reduce_scratch;          /* (>= 8 * 1024) */
    /* tables (device) */
    const int32_t *neighbor;         /* (6,n); -1 where the neighbour is missing */
    const int32_t *active_i, *active_j, *active_k;
    const int32_t *node_link_off;    /* (n+1) CSR offsets into node_link_ids */
    const int32_t *node_link_ids;    /* links of each node, ascending link order */
    const int32_t *link_p, *link_up, *link_q, *link_bid;
    const double *link_eta, *link_xb, *link_yb, *link_zb, *link_nx, *link_ny, *link_nz;
    /* characteristic stencils of the traction links, compacted: sample
     * lists hold indices into char_nodes, -1 padded after the last sample */
    const int32_t *char_link;        /* (Lc) link index of each stencil */
    const int32_t *char_nodes;       /* (Ns) active node of each sample slot */
    double *char_pack;               /* (Ns,12) node-major copy of U at the sample nodes */
    const int32_t *char_pack_slot;   /* (n) sample slot of each active node, -1 if none */
    const int32_t *char_gidx;        /* (48,Lc) gradient samples (into char_nodes) */
    const double *char_gw;           /* (3,48,Lc) x/y/z gradient weights */
    const int32_t *char_pidx;        /* (16,Lc) predictor samples (into char_nodes) */
    const double *char_pw;           /* (16,Lc) predictor weights */
    double *char_state;              /* (4,12,L) predictor and gradient of U */
    /* compat_bfl surface-tangent stencils (may be null for other reconstructions) */
    const int32_t *tan_p;            /* (32,L) active node of each sample, -1 if unused */
    const double *tan_w1, *tan_w2;   /* (32,L) d/dt1 and d/dt2 weights */
    const int32_t *tan_valid;        /* (L) */
    const int32_t *kind_by_id;       /* (16) */
    const int32_t *comp_kind_by_id;  /* (3,16) */
    const int32_t *mode_by_id;       /* (16) */
    const double *params_by_id;      /* (MAX_IDS, NPRM) row-major */
    const int32_t *dyn_enabled;      /* (L) */
    const double *dyn_traction;      /* (3,L) */
    int32_t *lkind;                  /* (L) */
    int32_t *lcomp_kind;             /* (3,L) */
    double *lbv, *lFb, *lPb;         /* (3,L), (9,L), (9,L) */
    /* rectangular box solver (WarpNonlinearElasticLBM3D): every node of a box_n box active in
       (i, j, k) order with k fastest, n_links = 0, periodic wrap in the neighbour table; faces
       left, right, bottom, top, front, back */
    int32_t box;                     /* 1 = box solver */
    int32_t box_n[3];
    int32_t box_side_kind[6];        /* 0 periodic (both faces of an axis), 1 Dirichlet, 2 Neumann */
    int32_t box_value_mode[6];       /* box value modes 0..5 */
    double box_params[96];           /* 16 params per face */
    double box_length[3];
    double box_speed[3];             /* lattice speeds along x, y, z (one dt: dx/cx = dy/cy = dz/cz), as
                                        vnew3d:2342-2361; a zero entry means c.  The curved solver
                                        ignores them. */
    void *stream;                    /* hipStream_t; NULL = null stream */
} vlbm3d_desc;

/* Reduction slots filled by vlbm3d_reduce. */
#define VLBM_RED_SUM_V2 0   /* sum over active nodes of |v|^2 (rms_v = sqrt(sum / n)) */
#define VLBM_RED_MAX_V 1    /* max |v| */
#define VLBM_RED_MIN_J 2    /* min det F */
#define VLBM_RED_MAX_J 3    /* max det F */
#define VLBM_RED_MAX_U 4    /* max |u| */
#define VLBM_RED_NONFINITE 5 /* count of non-finite entries of U and u */
#define VLBM_RED_INVALID 6  /* device invalid flag */
#define VLBM_RED_SUM_DU2 7  /* sum over active nodes of |u - u_prev|^2 (0 when u_prev is NULL) */
#define VLBM_RED_COUNT 8

int vlbm3d_create(const vlbm3d_desc *desc, vlbm3d_handle **out);
int vlbm3d_destroy(vlbm3d_handle *h);
/* Advance nsteps steps from (time, steps).  f0 holds the current populations
 * and f1 is scratch; the buffers alternate every step, so on return the
 * newest populations are in f0 when nsteps is even and in f1 when odd.
 * f0/f1 are double planes (precision F64) or float planes (precision F32). */
int vlbm3d_step(vlbm3d_handle *h, void *f0, void *f1, double time, int64_t steps, int32_t nsteps);
/* Replays blocks of `steps` fused steps (even, 0 = off) from captured HIP
 * graphs whenever a vlbm3d_step call has a uniform schedule (sweep and F
 * update every step or never); the boundary time is then kept on the device.
 * Results are bitwise those of direct launches.  No reference counterpart
 * (the reference launches every kernel from the host, cbw3d:4549-4589). */
int vlbm3d_set_graph_steps(vlbm3d_handle *h, int32_t steps);
/* Moments of f into U, u = u_star + dt/2 v, invalid flag. */
int vlbm3d_refresh(vlbm3d_handle *h, const void *f);
/* Boundary reconstruction at time t.  in_place = 0: reads f_src, writes the
 * incoming link populations of f_dst.  in_place = 1: copies f_src into f_dst
 * (scratch) and writes the link populations into f_src, reading the copy. */
int vlbm3d_apply_boundary(vlbm3d_handle *h, void *f_src, void *f_dst, double t, int32_t in_place);
/* f = feq(U) - dt/(2Q) B(U) from the U planes. */
int vlbm3d_init_populations(vlbm3d_handle *h, void *f);
/* u_star = disp - dt/2 v from a (3,n) device array. */
int vlbm3d_set_ustar(vlbm3d_handle *h, const double *disp);
/* Synchronous reductions over U and u (see VLBM_RED_*); out has 8 slots.
 * u_prev (nullable, (3,n) device) is an earlier copy of u for the
 * displacement-rate equilibrium residual. */
int vlbm3d_reduce(vlbm3d_handle *h, const double *u_prev, double *out);
/* The reductions of vlbm3d_reduce with no allocation, copy or wait: no
 * counterpart in the reference; for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops).  The 8 results (VLBM_RED_*) are
 * written to the device array out_dev on the solver stream and are valid once
 * that stream has completed.  The per-block partials live in a static device
 * array of 1024 rows x 8 doubles (64 KiB per GPU, held until the process
 * ends); the partial grid is capped at 1024 blocks, so every lattice fits.
 * All calls on one GPU share that array: calls that overlap in time on one
 * GPU (different streams) race. */
int vlbm3d_reduce_noalloc(vlbm3d_handle *h, const double *u_prev, double *out_dev);
/* Converts populations between the stored layout and precision and
 * (72, n_active) double planes. */
int vlbm3d_populations_to_f64(vlbm3d_handle *h, const void *f, double *dst);
int vlbm3d_populations_from_f64(vlbm3d_handle *h, const double *src, void *f);
/* Per-kernel timing with HIP events on the solver stream (for measurement).
 * vlbm3d_timing_enable(h, 1) starts recording (and clears earlier marks);
 * vlbm3d_timing_read synchronises the stream and returns, per slot
 * VLBM_SLOT_*, the summed kernel milliseconds and launch counts since the
 * last enable/read.  No reference counterpart (the reference reports only
 * wall-clock elapsed seconds per case, cases/reproduce.py). */
#define VLBM_SLOT_COLLIDE 0
#define VLBM_SLOT_SWEEP 1
#define VLBM_SLOT_COMMIT 2
#define VLBM_SLOT_RECORD 3
#define VLBM_SLOT_STREAM 4
#define VLBM_SLOT_GATHER 5
#define VLBM_SLOT_PREPARE 6
#define VLBM_SLOT_LINKS 7
#define VLBM_SLOT_STEP 8
#define VLBM_NSLOTS 9
int vlbm3d_timing_enable(vlbm3d_handle *h, int32_t enable);
int vlbm3d_timing_read(vlbm3d_handle *h, double *ms, int32_t *counts, int32_t nslots);
const char *vlbm_last_error(void);
int vlbm_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
