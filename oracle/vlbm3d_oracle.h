/* CPU FP64 restatement of the reference D3Q6x12 sparse curved-boundary step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle for the HIP
 * path and the CPU baseline timed by bench.py.  It is never linked into,
 * loaded by or called from the product library.
 *
 * Every routine restates one kernel or device function of the reference,
 * src/curved_lbm/three_d/curved_boundary_warp3d.py (abbreviated cbw3d) and
 * src/curved_lbm/three_d/vector_nonlinear_elastic_warp3d.py (vnew3d), in
 * sequential C with OpenMP loops where the reference launch is data-parallel.
 * Array layouts are the reference's: fields are (components, n_active)
 * row-major, populations (72, n_active) with index q*12+a.
 */
#ifndef VLBM3D_ORACLE_H
#define VLBM3D_ORACLE_H
#include <stdint.h>

#define OR_MAX_IDS 16
#define OR_NPRM 16 /* parameters per boundary id */
#define OR_MAX_CHAR 64
#define OR_MAX_TAN 32

typedef struct {
    int32_t n_active;
    int32_t n_links;
    /* material */
    int32_t model;
    double lam, mu, p0, p1, p2, j_floor;
    /* lattice / collision */
    double dt, c, omega, dx, dy, dz;
    int32_t source_mode;
    double source_params[12];
    int32_t recon;
    int32_t normal_solver; /* 0 FD, 1 acoustic, 2 energy */
    /* kinematics */
    int32_t disp_interval, disp_sweeps;
    double disp_relax, disp_weight;
    int32_t disp_boundary_id;
    int32_t compat_interval;
    double compat_blend;
    int32_t has_char;
    /* state (owned by the caller) */
    double *f0, *f1, *U, *u, *u_star, *u_next, *bc_diag, *bc_rhs;
    int32_t *invalid;
    /* tables */
    const int32_t *neighbor; /* (6, n) */
    const int32_t *active_i, *active_j, *active_k;
    const int32_t *link_p, *link_up, *link_q, *link_bid;
    const double *link_eta, *link_xb, *link_yb, *link_zb, *link_nx, *link_ny, *link_nz;
    const int32_t *char_indices; /* (64, L) */
    const double *char_weights;  /* (4, 64, L) */
    double *char_state;          /* (4, 12, L) */
    const int32_t *tan_p;        /* (L, 32) compat_bfl tangent stencil nodes, -1 unused */
    const double *tan_w1, *tan_w2; /* (L, 32) d/dt1, d/dt2 weights */
    const int32_t *tan_valid;    /* (L) */
    const int32_t *kind_by_id;   /* (16) */
    const int32_t *comp_kind_by_id; /* (3, 16) */
    const int32_t *mode_by_id;   /* (16) */
    const double *params_by_id;  /* (16, 16) */
    const int32_t *dyn_enabled;  /* (L) */
    const double *dyn_traction;  /* (3, L) */
    int32_t *lkind;              /* (L) */
    int32_t *lcomp_kind;         /* (3, L) */
    double *lbv, *lFb, *lPb;     /* (3, L), (9, L), (9, L) */
    /* rectangular box solver: every node of box_n active (k fastest), no cut
     * links, periodic wrap in the neighbour table; faces left, right,
     * bottom, top, front, back */
    int32_t box;
    int32_t box_n[3];
    int32_t side_kind[6], side_mode[6];
    double side_params[96];
    double box_length[3];
    double box_speed[3];  /* lattice speeds along x, y, z; a zero entry means c */
} or3d_state;

int or3d_refresh(or3d_state *s);
int or3d_set_ustar(or3d_state *s, const double *disp);
int or3d_init_populations(or3d_state *s, const double *Uinit);
int or3d_apply_boundary(or3d_state *s, double t, int in_place);
/* Runs nsteps reference steps starting at (time, steps).  Like the reference
 * it swaps the f0/f1 pointers after every step: on return s->f0 points at the
 * newest populations, which is the caller's f1 buffer when nsteps is odd. */
int or3d_step(or3d_state *s, double time, int64_t steps, int nsteps);
/* Direct access to device-function restatements, for the unit tests. */
void or3d_piola(const double *F, int model, double lam, double mu, double p0, double p1,
                double p2, double jf, double *P);
void or3d_du_axis_all(const or3d_state *s, const double *u, double *out); /* (3,3,n) */
double or3d_value_probe(const double *prm, int mode, int comp, double x, double y, double z, double nx,
                        double ny, double nz, double t, int want_disp);
int or3d_threads(void);
#endif
