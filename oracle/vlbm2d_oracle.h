/* CPU FP64 restatement of the reference D2Q4x6 curved-boundary time step
 * (JingsenFeng/curved-boundary-vlbm, src/curved_lbm/two_d/curved_boundary_warp.py).
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle for the HIP library and the
 * CPU baseline of the benchmark.  The product package never links it.
 *
 * Arrays are dense (components, nx, ny) planes, index c*nx*ny + i*ny + j,
 * exactly the reference's (components, nx, ny, 1) Warp arrays; populations
 * use index q*6 + a with directions E, N, W, S.
 */
#ifndef VLBM2D_ORACLE_H
#define VLBM2D_ORACLE_H

#include <stdint.h>

typedef struct {
    int32_t nx, ny, n_links;
    int32_t model;
    double lam, mu, p0, p1, p2, j_floor;
    double dt, c, omega, dx, dy;
    int32_t source_mode;
    double source_params[6];
    int32_t recon;         /* reference boundary_reconstruction id (0, 2, 12, 13 provided) */
    int32_t normal_solver; /* 0 finite-difference Newton, 1 acoustic Newton, 2 energy */
    int32_t disp_interval, disp_sweeps;
    double disp_relax, disp_weight;
    int32_t disp_boundary_id;
    int32_t compat_interval;
    double compat_blend;
    int32_t compat_interior_only;
    int32_t use_dyn;
    int32_t collision_model;           /* 0 bgk, 1 mrt_ghost */
    double ghost_omega;                /* mrt_ghost relaxation of the ghost moment */
    double filter_strength;            /* kinetic ghost filter after each step, [0, 1] */
    int32_t box;                       /* 1: rectangular box solver (no cut links) */
    int32_t side_kind[4], side_mode[4]; /* left, right, bottom, top: 0 periodic, 1 Dirichlet, 2 Neumann */
    double side_params[32];
    double length_x, length_y;
    /* state */
    double *f0, *f1, *fpost;           /* (24, nx, ny) */
    double *U, *u, *u_star, *u_next;   /* (6|2, nx, ny) */
    double *bc_diag, *bc_rhs;          /* (nx, ny), (2, nx, ny) */
    double *P, *sigma;                 /* (4, nx, ny) */
    int32_t *invalid;                  /* (1) */
    /* tables */
    const int32_t *active;             /* (nx, ny) */
    const int32_t *link_i, *link_j, *link_q, *link_bid;
    const double *link_eta, *link_xb, *link_yb, *link_nx, *link_ny;
    const int32_t *lsq_i, *lsq_j;      /* (L, 24) */
    const double *lsq_wb, *lsq_wt, *lsq_wx, *lsq_wy;
    const int32_t *lsq_valid;
    const int32_t *kind_by_id, *mode_by_id; /* (8) */
    const double *params_by_id;        /* (8, 8) */
    const double *dyn_tx, *dyn_ty;     /* (L) */
} or2d_state;

/* 0 on success, -1 for an unsupported configuration. */
int or2d_step(or2d_state *s, double time, int64_t steps, int32_t nsteps);
int or2d_refresh(or2d_state *s);
int or2d_apply_boundary(or2d_state *s, double t, int32_t in_place);
void or2d_piola(const double *F, int model, double lam, double mu, double p0, double p1, double p2,
                double jf, double *P);
void or2d_solve_normal(const or2d_state *s, const double *Fg, double nx, double ny, double tx,
                       double ty, double *Fb);
double or2d_boundary_value(int comp, int mode, const double *p, double x, double y, double nx,
                           double ny, double t);
void or2d_du(const or2d_state *s, const double *u, double *out); /* (4, nx, ny): dux/dx, dux/dy, duy/dx, duy/dy */
void or2d_equilibrium(const or2d_state *s, const double *U, double *f); /* U (6) -> f (24) */
int or2d_threads(void);
void or2d_kinetic_filter(or2d_state *s, double *f); /* in place on (24, nx, ny) */

#endif
