/* CPU FP64 restatement of the reference D3Q6x12 sparse curved-boundary step.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle and CPU baseline); see the header.
 * cbw3d = src/curved_lbm/three_d/curved_boundary_warp3d.py
 * vnew3d = src/curved_lbm/three_d/vector_nonlinear_elastic_warp3d.py
 *
 * Arithmetic follows the reference expression order so that results agree to
 * rounding; the constitutive map is evaluated once per node instead of once
 * per (q, a) pair, which yields identical values because it is a pure
 * function of F.
 */
#include "vlbm3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NC 12
#define NQ 6
#define FQ 72

enum { M_SVK = 0, M_NH = 1, M_LOGNH = 2, M_MR = 3, M_YEOH = 4, M_GENT = 5, M_SPHNH = 6 };
enum { BC_DIR = 1, BC_NEU = 2 };
enum { SRC_NONE = 0, SRC_CONST = 1, SRC_DAMP = 2 };
enum { R_BFL = 0, R_HALF = 1, R_COMPAT = 2, R_LOCALF = 3, R_LOCALF_BFL = 4 };
enum { NJ_FD = 0, NJ_AC = 1, NS_ENERGY = 2 };
enum {
    V_ZERO = 0, V_CONST = 1, V_AFFINE_VEL = 4, V_RAMP = 200, V_PIOLA_N = 201, V_NORMAL_T = 202,
    V_SINE2_NT = 203, V_SINE2_HOLD = 204, V_AFFINE_NH = 205, V_ZPULL = 206, V_ZTWIST = 207,
    V_LINEAR_NT = 208, V_XPULL = 209, V_XTWIST = 210
};


static const double PI_D = 3.141592653589793238462643383279502884;

/* Tube end loads about the axis through (p[0], p[1]) of the transverse plane
 * (a, b); comp 0, 1 transverse, 2 axial (cbw3d:1100-1132, 2384-2398).  The
 * x-axis modes 209, 210 use them in the frame (a, b, axial) = (y, z, x). */
static double tube_pull(const double *p, int comp, double a, double b, double t) {
    double active = p[4], denom = p[5];
    if (denom <= 0.0) denom = 2.0 * active;
    double scale = 1.0;
    if (active > 0.0 && t < active) scale = sin(PI_D * t / denom) * sin(PI_D * t / denom);
    if (comp == 0) return -scale * p[3] * (b - p[1]);
    if (comp == 1) return scale * p[3] * (a - p[0]);
    return scale * p[2];
}
static double twist_velocity(const double *p, int comp, double a, double b, double t) {
    double active = p[4];
    if (active <= 0.0 || t >= active) return 0.0;
    double phase = PI_D * t / active;
    double sv = 0.5 * (1.0 - cos(phase));
    double sdot = 0.5 * PI_D / active * sin(phase);
    double theta = p[3] * sv, thetadot = p[3] * sdot;
    double da = a - p[0], db = b - p[1];
    if (comp == 0) return thetadot * (-sin(theta) * da - cos(theta) * db);
    if (comp == 1) return thetadot * (cos(theta) * da - sin(theta) * db);
    return p[2] * sdot;
}
static double twist_disp(const double *p, int comp, double a, double b, double time) {
    double active = p[4], sv = 1.0;
    if (active > 0.0 && time < active) sv = 0.5 * (1.0 - cos(PI_D * time / active));
    double theta = p[3] * sv;
    double da = a - p[0], db = b - p[1];
    double c = cos(theta), ss = sin(theta);
    if (comp == 0) return (c - 1.0) * da - ss * db;
    if (comp == 1) return ss * da + (c - 1.0) * db;
    return p[2] * sv;
}
static inline int tube_comp(int comp) { return comp == 0 ? 2 : comp - 1; }

static inline int dir_x(int q) { return q == 0 ? 1 : (q == 3 ? -1 : 0); }
static inline int dir_y(int q) { return q == 1 ? 1 : (q == 4 ? -1 : 0); }
static inline int dir_z(int q) { return q == 2 ? 1 : (q == 5 ? -1 : 0); }
static inline int opp(int q) { return q < 3 ? q + 3 : q - 3; }

/* vnew3d.py:774-789 */
static inline double det3(const double *F) {
    return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
           F[2] * (F[3] * F[7] - F[4] * F[6]);
}

/* vnew3d.py:793-1022 (_piola_mat), F and P row-major 3x3. */
void or3d_piola(const double *F, int model, double lam, double mu, double p0, double p1,
                double p2, double jf, double *P) {
    const double F11 = F[0], F12 = F[1], F13 = F[2], F21 = F[3], F22 = F[4], F23 = F[5],
                 F31 = F[6], F32 = F[7], F33 = F[8];
    if (model == M_SVK) {
        double C00 = F11 * F11 + F21 * F21 + F31 * F31, C01 = F11 * F12 + F21 * F22 + F31 * F32;
        double C02 = F11 * F13 + F21 * F23 + F31 * F33, C11 = F12 * F12 + F22 * F22 + F32 * F32;
        double C12 = F12 * F13 + F22 * F23 + F32 * F33, C22 = F13 * F13 + F23 * F23 + F33 * F33;
        double E00 = 0.5 * (C00 - 1.0), E01 = 0.5 * C01, E02 = 0.5 * C02;
        double E11 = 0.5 * (C11 - 1.0), E12 = 0.5 * C12, E22 = 0.5 * (C22 - 1.0);
        double trE = E00 + E11 + E22;
        double S00 = lam * trE + 2.0 * mu * E00, S01 = 2.0 * mu * E01, S02 = 2.0 * mu * E02;
        double S11 = lam * trE + 2.0 * mu * E11, S12 = 2.0 * mu * E12, S22 = lam * trE + 2.0 * mu * E22;
        P[0] = F11 * S00 + F12 * S01 + F13 * S02;
        P[1] = F11 * S01 + F12 * S11 + F13 * S12;
        P[2] = F11 * S02 + F12 * S12 + F13 * S22;
        P[3] = F21 * S00 + F22 * S01 + F23 * S02;
        P[4] = F21 * S01 + F22 * S11 + F23 * S12;
        P[5] = F21 * S02 + F22 * S12 + F23 * S22;
        P[6] = F31 * S00 + F32 * S01 + F33 * S02;
        P[7] = F31 * S01 + F32 * S11 + F33 * S12;
        P[8] = F31 * S02 + F32 * S12 + F33 * S22;
        return;
    }
    double J = det3(F);
    if (J < jf) J = jf;
    double H[9];
    H[0] = (F22 * F33 - F23 * F32) / J;
    H[1] = (F23 * F31 - F21 * F33) / J;
    H[2] = (F21 * F32 - F22 * F31) / J;
    H[3] = (F13 * F32 - F12 * F33) / J;
    H[4] = (F11 * F33 - F13 * F31) / J;
    H[5] = (F12 * F31 - F11 * F32) / J;
    H[6] = (F12 * F23 - F13 * F22) / J;
    H[7] = (F13 * F21 - F11 * F23) / J;
    H[8] = (F11 * F22 - F12 * F21) / J;
    double logJ = log(J);
    if (model == M_NH || model == M_LOGNH) {
        double vol = model == M_NH ? 0.5 * lam * (J * J - 1.0) : lam * logJ;
        for (int k = 0; k < 9; ++k) P[k] = mu * (F[k] - H[k]) + vol * H[k];
        return;
    }
    double C00 = F11 * F11 + F21 * F21 + F31 * F31, C01 = F11 * F12 + F21 * F22 + F31 * F32;
    double C02 = F11 * F13 + F21 * F23 + F31 * F33, C11 = F12 * F12 + F22 * F22 + F32 * F32;
    double C12 = F12 * F13 + F22 * F23 + F32 * F33, C22 = F13 * F13 + F23 * F23 + F33 * F33;
    double I1 = C00 + C11 + C22;
    double jm23 = exp((-2.0 / 3.0) * logJ);
    double G1[9];
    for (int k = 0; k < 9; ++k) G1[k] = jm23 * (2.0 * F[k] - (2.0 / 3.0) * I1 * H[k]);
    if (model == M_SPHNH) {
        double vol = 0.5 * lam * (J * J - 1.0), sc = 0.5 * mu;
        for (int k = 0; k < 9; ++k) P[k] = sc * G1[k] + vol * H[k];
        return;
    }
    double trC2 = C00 * C00 + C11 * C11 + C22 * C22 + 2.0 * (C01 * C01 + C02 * C02 + C12 * C12);
    double I2 = 0.5 * (I1 * I1 - trC2);
    double jm43 = exp((-4.0 / 3.0) * logJ);
    double Cm[9] = {C00, C01, C02, C01, C11, C12, C02, C12, C22};
    double FC[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            FC[3 * i + j] = F[3 * i] * Cm[j] + F[3 * i + 1] * Cm[3 + j] + F[3 * i + 2] * Cm[6 + j];
    double vol = lam * logJ;
    if (model == M_MR) {
        for (int k = 0; k < 9; ++k) {
            double G2 = jm43 * (2.0 * (I1 * F[k] - FC[k]) - (4.0 / 3.0) * I2 * H[k]);
            P[k] = p0 * G1[k] + p1 * G2 + vol * H[k];
        }
        return;
    }
    if (model == M_YEOH) {
        double A = jm23 * I1 - 3.0;
        double sc = p0 + 2.0 * p1 * A + 3.0 * p2 * A * A;
        for (int k = 0; k < 9; ++k) P[k] = sc * G1[k] + vol * H[k];
        return;
    }
    if (model == M_GENT) {
        double A = jm23 * I1 - 3.0;
        double denom = 1.0 - A / p0;
        if (denom < jf) denom = jf;
        double sc = 0.5 * mu / denom;
        for (int k = 0; k < 9; ++k) P[k] = sc * G1[k] + vol * H[k];
        return;
    }
    for (int k = 0; k < 9; ++k) P[k] = 0.0;
}

#define PIOLA(F, P) or3d_piola((F), s->model, s->lam, s->mu, s->p0, s->p1, s->p2, s->j_floor, (P))

/* vnew3d.py:1067-1133 (_flux_comp) with P precomputed. */
static inline double flux_comp(int axis, int a, const double *U, const double *P) {
    if (a < 3) return -P[3 * a + axis];
    int i = (a - 3) / 3, j = (a - 3) % 3;
    return j == axis ? -U[i] : 0.0;
}

/* vnew3d.py:1137-1236 (_equilibrium_comp) with P precomputed. */
static inline double equilibrium_comp(int q, int a, const double *U, const double *P, double c) {
    double sx = dir_x(q), sy = dir_y(q), sz = dir_z(q);
    double phix = flux_comp(0, a, U, P), phiy = flux_comp(1, a, U, P), phiz = flux_comp(2, a, U, P);
    return (1.0 / 6.0) * (U[a] + (3.0 / c) * (sx * phix + sy * phiy + sz * phiz));
}

/* Lattice speed along an axis: the box's per-axis speeds (vnew3d:2342-2361),
 * else the single c. */
static inline double axis_speed(const or3d_state *s, int axis) {
    return (s->box && s->box_speed[axis] > 0.0) ? s->box_speed[axis] : s->c;
}

/* vnew3d.py:1240-1367 (_equilibrium_anisotropic_comp): equal speeds take
 * the isotropic form, otherwise each axis's flux is divided by its speed. */
static inline double equilibrium_aniso(const or3d_state *s, int q, int a, const double *U, const double *P) {
    const double cx = axis_speed(s, 0), cy = axis_speed(s, 1), cz = axis_speed(s, 2);
    if (cx == cy && cx == cz) return equilibrium_comp(q, a, U, P, cx);
    double sx = dir_x(q), sy = dir_y(q), sz = dir_z(q);
    double phix = flux_comp(0, a, U, P), phiy = flux_comp(1, a, U, P), phiz = flux_comp(2, a, U, P);
    return (1.0 / 6.0) * (U[a] + 3.0 * (sx * phix / cx + sy * phiy / cy + sz * phiz / cz));
}

/* vnew3d.py:1372-1390 */
static inline double source_comp(int a, int mode, const double *params, double vx, double vy,
                                 double vz) {
    if (mode == SRC_CONST) return params[a];
    if (mode == SRC_DAMP) {
        if (a == 0) return -params[0] * vx;
        if (a == 1) return -params[0] * vy;
        if (a == 2) return -params[0] * vz;
    }
    return 0.0;
}

#define IDX(a, p, n) ((size_t)(a) * (size_t)(n) + (size_t)(p))

/* Moments + half-step source shared by cbw3d.py:2110-2189 and :2193-2280. */
static void node_moments(const or3d_state *s, const double *f, int p, double *U) {
    const int n = s->n_active;
    for (int a = 0; a < NC; ++a) U[a] = 0.0;
    for (int q = 0; q < NQ; ++q)
        for (int a = 0; a < NC; ++a) U[a] += f[IDX(q * 12 + a, p, n)];
    double half_dt = 0.5 * s->dt;
    if (s->source_mode == SRC_CONST) {
        for (int a = 0; a < NC; ++a) U[a] += half_dt * s->source_params[a];
    } else if (s->source_mode == SRC_DAMP) {
        double denom = 1.0 + half_dt * s->source_params[0];
        U[0] = U[0] / denom;
        U[1] = U[1] / denom;
        U[2] = U[2] / denom;
    }
}

/* cbw3d.py:2110-2189 (_sparse_refresh_kernel) */
int or3d_refresh(or3d_state *s) {
    const int n = s->n_active;
    int bad = 0;
    s->invalid[0] = 0;
#pragma omp parallel for reduction(| : bad) schedule(static)
    for (int p = 0; p < n; ++p) {
        double U[NC];
        node_moments(s, s->f0, p, U);
        for (int a = 0; a < NC; ++a) s->U[IDX(a, p, n)] = U[a];
        double half_dt = 0.5 * s->dt;
        for (int k = 0; k < 3; ++k) s->u[IDX(k, p, n)] = s->u_star[IDX(k, p, n)] + half_dt * U[k];
        if (det3(U + 3) <= s->j_floor) bad = 1;
    }
    if (bad) s->invalid[0] = 1;
    return 0;
}

/* cbw3d.py:2097-2106 (_sparse_set_ustar) */
int or3d_set_ustar(or3d_state *s, const double *disp) {
    const int n = s->n_active;
    for (int k = 0; k < 3; ++k)
        for (int p = 0; p < n; ++p)
            s->u_star[IDX(k, p, n)] = disp[IDX(k, p, n)] - 0.5 * s->dt * s->U[IDX(k, p, n)];
    return 0;
}

/* vnew3d.py:535-560 (_equilibrium_np, _source_np) as used by cbw3d.py:4252-4253:
 * f = feq(U) - 0.5 dt / Q * B(U). */
int or3d_init_populations(or3d_state *s, const double *Uin) {
    const int n = s->n_active;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; ++p) {
        double U[NC], F[9], P[9];
        for (int a = 0; a < NC; ++a) U[a] = Uin[IDX(a, p, n)];
        for (int k = 0; k < 9; ++k) F[k] = U[3 + k];
        PIOLA(F, P);
        double B[NC];
        for (int a = 0; a < NC; ++a) B[a] = 0.0;
        if (s->source_mode == SRC_CONST)
            for (int a = 0; a < NC; ++a) B[a] = s->source_params[a];
        else if (s->source_mode == SRC_DAMP)
            for (int a = 0; a < 3; ++a) B[a] = -s->source_params[0] * U[a];
        for (int q = 0; q < NQ; ++q) {
            int sd[3] = {dir_x(q), dir_y(q), dir_z(q)};
            for (int a = 0; a < NC; ++a) {
                /* directional flux: sum over axes of s[axis] * flux[axis] / c */
                double dfl = 0.0;
                for (int ax = 0; ax < 3; ++ax) dfl += sd[ax] * flux_comp(ax, a, U, P) / axis_speed(s, ax);
                double f = (1.0 / 6.0) * (U[a] + 3.0 * dfl);
                f -= 0.5 * s->dt / 6.0 * B[a];
                s->f0[IDX(q * 12 + a, p, n)] = f;
            }
        }
    }
    return 0;
}

/* cbw3d.py:2193-2316 (_sparse_collide_inplace_kernel) */
static void collide_inplace(or3d_state *s) {
    const int n = s->n_active;
    int bad = 0;
#pragma omp parallel for reduction(| : bad) schedule(static)
    for (int p = 0; p < n; ++p) {
        double U[NC];
        node_moments(s, s->f0, p, U);
        for (int a = 0; a < NC; ++a) s->U[IDX(a, p, n)] = U[a];
        double half_dt = 0.5 * s->dt;
        for (int k = 0; k < 3; ++k) {
            double uk = s->u_star[IDX(k, p, n)] + half_dt * U[k];
            s->u[IDX(k, p, n)] = uk;
            s->u_star[IDX(k, p, n)] = uk + half_dt * U[k];
        }
        if (det3(U + 3) <= s->j_floor) bad = 1;
        double P[9];
        PIOLA(U + 3, P);
        for (int q = 0; q < NQ; ++q)
            for (int a = 0; a < NC; ++a) {
                double feq = s->box ? equilibrium_aniso(s, q, a, U, P) : equilibrium_comp(q, a, U, P, s->c);
                double old = s->f0[IDX(q * 12 + a, p, n)];
                double st = source_comp(a, s->source_mode, s->source_params, U[0], U[1], U[2]);
                s->f0[IDX(q * 12 + a, p, n)] =
                    old + s->omega * (feq - old) + s->dt * (2.0 - s->omega) * (1.0 / 12.0) * st;
            }
    }
    if (bad) s->invalid[0] = 1;
}

/* cbw3d.py:2334-2347 */
static double sine2_integral(double time, double active, double denom) {
    if (active <= 0.0) return time;
    if (denom <= 0.0) denom = active;
    double t = time > 0.0 ? time : 0.0;
    if (t < active) return 0.5 * t - denom * sin(2.0 * PI_D * t / denom) / (4.0 * PI_D);
    double prefix = 0.5 * active - denom * sin(2.0 * PI_D * active / denom) / (4.0 * PI_D);
    return prefix + (t - active);
}

#define PRM(bid, k) (s->params_by_id[(bid) * OR_NPRM + (k)])

/* cbw3d.py:2351-2399 (_local_dirichlet_displacement_comp); mode 208 (linear
 * ramp, a traction mode) has no displacement anchor and yields zero. */
static double dirichlet_disp(const or3d_state *s, int mode, int bid, int comp, double x, double y,
                             double z, double time) {
    if (mode == V_ZERO) return 0.0;
    if (mode == V_CONST) return time * PRM(bid, comp);
    if (mode == V_RAMP) return PRM(bid, comp) * sine2_integral(time, PRM(bid, 3), PRM(bid, 3));
    if (mode == V_SINE2_HOLD) {
        double active = PRM(bid, 3), denom = PRM(bid, 4);
        if (denom <= 0.0) denom = 2.0 * active;
        return PRM(bid, comp) * sine2_integral(time, active, denom);
    }
    if (mode == V_AFFINE_VEL) {
        double at = time, active = PRM(bid, 12);
        if (active > 0.0 && at > active) at = active;
        double out = PRM(bid, 0) * x + PRM(bid, 1) * y + PRM(bid, 2) * z + PRM(bid, 9);
        if (comp == 1) out = PRM(bid, 3) * x + PRM(bid, 4) * y + PRM(bid, 5) * z + PRM(bid, 10);
        else if (comp == 2) out = PRM(bid, 6) * x + PRM(bid, 7) * y + PRM(bid, 8) * z + PRM(bid, 11);
        return at * out;
    }
    if (mode == V_ZTWIST) return twist_disp(&PRM(bid, 0), comp, x, y, time);
    if (mode == V_XTWIST) return twist_disp(&PRM(bid, 0), tube_comp(comp), y, z, time);
    return 0.0;
}

/* cbw3d.py:2403-2469 (bc clear + build; links accumulated in link order). */
static void build_bc(or3d_state *s, double time) {
    const int n = s->n_active, L = s->n_links;
    memset(s->bc_diag, 0, sizeof(double) * 3 * (size_t)n);
    memset(s->bc_rhs, 0, sizeof(double) * 3 * (size_t)n);
    double w2 = s->disp_weight * s->disp_weight;
    for (int ell = 0; ell < L; ++ell) {
        int bid = s->link_bid[ell];
        int force_all = 0;
        if (s->disp_boundary_id >= 0) {
            if (bid != s->disp_boundary_id) continue;
            force_all = 1;
        }
        int p = s->link_p[ell];
        if (p < 0) continue;
        double x = s->link_xb[ell], y = s->link_yb[ell], z = s->link_zb[ell];
        double dbx = x - ((double)s->active_i[p] + 0.5) * s->dx;
        double dby = y - ((double)s->active_j[p] + 0.5) * s->dy;
        double dbz = z - ((double)s->active_k[p] + 0.5) * s->dz;
        int mode = s->mode_by_id[bid];
        double ub0 = dirichlet_disp(s, mode, bid, 0, x, y, z, time);
        double ub1 = dirichlet_disp(s, mode, bid, 1, x, y, z, time);
        double ub2 = dirichlet_disp(s, mode, bid, 2, x, y, z, time);
        const double *U = s->U;
#define UU(a) U[IDX(a, p, n)]
        if (force_all || s->comp_kind_by_id[0 * OR_MAX_IDS + bid] == BC_DIR) {
            double rhs = ub0 - ((UU(3) - 1.0) * dbx + UU(4) * dby + UU(5) * dbz);
            s->bc_diag[IDX(0, p, n)] += w2;
            s->bc_rhs[IDX(0, p, n)] += w2 * rhs;
        }
        if (force_all || s->comp_kind_by_id[1 * OR_MAX_IDS + bid] == BC_DIR) {
            double rhs = ub1 - (UU(6) * dbx + (UU(7) - 1.0) * dby + UU(8) * dbz);
            s->bc_diag[IDX(1, p, n)] += w2;
            s->bc_rhs[IDX(1, p, n)] += w2 * rhs;
        }
        if (force_all || s->comp_kind_by_id[2 * OR_MAX_IDS + bid] == BC_DIR) {
            double rhs = ub2 - (UU(9) * dbx + UU(10) * dby + (UU(11) - 1.0) * dbz);
            s->bc_diag[IDX(2, p, n)] += w2;
            s->bc_rhs[IDX(2, p, n)] += w2 * rhs;
        }
#undef UU
    }
}

/* cbw3d.py:2473-2527 (_sparse_local_displacement_sweep) */
static void disp_sweep(or3d_state *s) {
    const int n = s->n_active;
    const double *U = s->U;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; ++p) {
        double d0 = s->bc_diag[IDX(0, p, n)], d1 = s->bc_diag[IDX(1, p, n)],
               d2 = s->bc_diag[IDX(2, p, n)];
        double r0 = s->bc_rhs[IDX(0, p, n)], r1 = s->bc_rhs[IDX(1, p, n)],
               r2 = s->bc_rhs[IDX(2, p, n)];
        for (int q = 0; q < NQ; ++q) {
            int pn = s->neighbor[IDX(q, p, n)];
            if (pn < 0) continue;
            double ddx = (double)dir_x(q) * s->dx, ddy = (double)dir_y(q) * s->dy,
                   ddz = (double)dir_z(q) * s->dz;
#define A(a, node) U[IDX(a, node, n)]
            double g0 = 0.5 * ((A(3, p) + A(3, pn) - 2.0) * ddx + (A(4, p) + A(4, pn)) * ddy +
                               (A(5, p) + A(5, pn)) * ddz);
            double g1 = 0.5 * ((A(6, p) + A(6, pn)) * ddx + (A(7, p) + A(7, pn) - 2.0) * ddy +
                               (A(8, p) + A(8, pn)) * ddz);
            double g2 = 0.5 * ((A(9, p) + A(9, pn)) * ddx + (A(10, p) + A(10, pn)) * ddy +
                               (A(11, p) + A(11, pn) - 2.0) * ddz);
#undef A
            r0 += s->u[IDX(0, pn, n)] - g0;
            r1 += s->u[IDX(1, pn, n)] - g1;
            r2 += s->u[IDX(2, pn, n)] - g2;
            d0 += 1.0;
            d1 += 1.0;
            d2 += 1.0;
        }
        double dd[3] = {d0, d1, d2}, rr[3] = {r0, r1, r2};
        for (int k = 0; k < 3; ++k) {
            double uk = s->u[IDX(k, p, n)];
            s->u_next[IDX(k, p, n)] = uk;
            if (dd[k] > 0.0) s->u_next[IDX(k, p, n)] = uk + s->disp_relax * (rr[k] / dd[k] - uk);
        }
    }
}

/* cbw3d.py:2531-2540 (_sparse_local_displacement_commit) */
static void disp_commit(or3d_state *s) {
    const int n = s->n_active;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; ++p)
        for (int k = 0; k < 3; ++k) {
            double delta = s->u_next[IDX(k, p, n)] - s->u[IDX(k, p, n)];
            s->u[IDX(k, p, n)] = s->u_next[IDX(k, p, n)];
            s->u_star[IDX(k, p, n)] = s->u_star[IDX(k, p, n)] + delta;
        }
}

/* cbw3d.py:2544-2571 (_sparse_local_du_axis) */
static double du_axis(const or3d_state *s, const double *u, int comp, int p, int qp, int qm,
                      double h) {
    const int n = s->n_active;
    int pp = s->neighbor[IDX(qp, p, n)], pm = s->neighbor[IDX(qm, p, n)];
#define UC(node) u[IDX(comp, node, n)]
    if (pp >= 0 && pm >= 0) return (UC(pp) - UC(pm)) / (2.0 * h);
    if (pp >= 0) {
        int pp2 = s->neighbor[IDX(qp, pp, n)];
        if (pp2 >= 0) return (-3.0 * UC(p) + 4.0 * UC(pp) - UC(pp2)) / (2.0 * h);
        return (UC(pp) - UC(p)) / h;
    }
    if (pm >= 0) {
        int pm2 = s->neighbor[IDX(qm, pm, n)];
        if (pm2 >= 0) return (3.0 * UC(p) - 4.0 * UC(pm) + UC(pm2)) / (2.0 * h);
        return (UC(p) - UC(pm)) / h;
    }
#undef UC
    return 0.0;
}

void or3d_du_axis_all(const or3d_state *s, const double *u, double *out) {
    const int n = s->n_active;
    double h[3] = {s->dx, s->dy, s->dz};
    for (int c = 0; c < 3; ++c)
        for (int A = 0; A < 3; ++A)
            for (int p = 0; p < n; ++p) out[((size_t)c * 3 + A) * n + p] = du_axis(s, u, c, p, A, A + 3, h[A]);
}

/* cbw3d.py:2575-2672 (_sparse_local_repair_F) */
static void repair_F(or3d_state *s) {
    const int n = s->n_active;
    double beta = s->compat_blend;
    if (beta > 1.0) beta = 1.0;
    if (beta < 0.0) beta = 0.0;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; ++p) {
        double o[9], t[9], nn[9];
        for (int k = 0; k < 9; ++k) o[k] = s->U[IDX(3 + k, p, n)];
        double h[3] = {s->dx, s->dy, s->dz};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                t[3 * i + j] = (i == j ? 1.0 : 0.0) + du_axis(s, s->u, i, p, j, j + 3, h[j]);
        /* the reference adds 1.0 to the diagonal derivative: 1 + du */
        for (int k = 0; k < 9; ++k) nn[k] = o[k] + beta * (t[k] - o[k]);
        if (det3(nn) <= s->j_floor) continue;
        double oldP[9], newP[9];
        PIOLA(o, oldP);
        PIOLA(nn, newP);
        double ff = 0.5 / s->c, sixth = 1.0 / 6.0;
        for (int q = 0; q < NQ; ++q) {
            double sx = dir_x(q), sy = dir_y(q), sz = dir_z(q);
            int base = q * 12;
            for (int r = 0; r < 3; ++r)
                s->f0[IDX(base + r, p, n)] =
                    s->f0[IDX(base + r, p, n)] -
                    ff * (sx * (newP[3 * r] - oldP[3 * r]) + sy * (newP[3 * r + 1] - oldP[3 * r + 1]) +
                          sz * (newP[3 * r + 2] - oldP[3 * r + 2]));
            for (int k = 0; k < 9; ++k)
                s->f0[IDX(base + 3 + k, p, n)] = s->f0[IDX(base + 3 + k, p, n)] + sixth * (nn[k] - o[k]);
        }
        for (int k = 0; k < 9; ++k) s->U[IDX(3 + k, p, n)] = nn[k];
    }
}

/* cbw3d.py:2320-2330 (_sparse_stream_kernel) after _sparse_clear */
static void stream_push(or3d_state *s) {
    const int n = s->n_active;
    memset(s->f1, 0, sizeof(double) * FQ * (size_t)n);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < NQ; ++q) {
            int nid = s->neighbor[IDX(q, p, n)];
            if (nid >= 0)
                for (int a = 0; a < NC; ++a) s->f1[IDX(q * 12 + a, nid, n)] = s->f0[IDX(q * 12 + a, p, n)];
        }
}

/* cbw3d.py:2676-2700 (_sparse_characteristic_gather) */
static void char_gather(or3d_state *s) {
    const int n = s->n_active, L = s->n_links;
#pragma omp parallel for schedule(static)
    for (int ell = 0; ell < L; ++ell) {
        if (s->char_indices[IDX(0, ell, L)] < 0) continue;
        for (int a = 0; a < NC; ++a) {
            double value = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
            for (int m = 0; m < OR_MAX_CHAR; ++m) {
                int p = s->char_indices[IDX(m, ell, L)];
                if (p < 0) continue;
                double uu = s->U[IDX(a, p, n)];
                value += s->char_weights[((size_t)0 * OR_MAX_CHAR + m) * L + ell] * uu;
                gx += s->char_weights[((size_t)1 * OR_MAX_CHAR + m) * L + ell] * uu;
                gy += s->char_weights[((size_t)2 * OR_MAX_CHAR + m) * L + ell] * uu;
                gz += s->char_weights[((size_t)3 * OR_MAX_CHAR + m) * L + ell] * uu;
            }
            s->char_state[((size_t)0 * NC + a) * L + ell] = value;
            s->char_state[((size_t)1 * NC + a) * L + ell] = gx;
            s->char_state[((size_t)2 * NC + a) * L + ell] = gy;
            s->char_state[((size_t)3 * NC + a) * L + ell] = gz;
        }
    }
}

/* cbw3d.py:996-1133 (_curved_value_comp) plus mode 208: nominal traction
 * amp * min(t / t_ramp, 1) * n with params (amp, t_ramp). */
static double curved_value(const or3d_state *s, int mode, int bid, int comp, double x, double y,
                           double z, double nx, double ny, double nz, double t) {
    double nv[3] = {nx, ny, nz};
    if (mode == V_ZERO) return 0.0;
    if (mode == V_CONST) return PRM(bid, comp);
    if (mode == V_RAMP) {
        double scale = 1.0, active = PRM(bid, 3);
        if (active > 0.0 && t < active) scale = sin(PI_D * t / active) * sin(PI_D * t / active);
        return PRM(bid, comp) * scale;
    }
    if (mode == V_AFFINE_VEL) {
        double active = PRM(bid, 12);
        if (active > 0.0 && t > active) return 0.0;
        double out = PRM(bid, 0) * x + PRM(bid, 1) * y + PRM(bid, 2) * z + PRM(bid, 9);
        if (comp == 1) out = PRM(bid, 3) * x + PRM(bid, 4) * y + PRM(bid, 5) * z + PRM(bid, 10);
        else if (comp == 2) out = PRM(bid, 6) * x + PRM(bid, 7) * y + PRM(bid, 8) * z + PRM(bid, 11);
        return out;
    }
    if (mode == V_SINE2_HOLD) {
        double active = PRM(bid, 3), denom = PRM(bid, 4);
        if (denom <= 0.0) denom = 2.0 * active;
        double scale = 1.0;
        if (active > 0.0 && t < active) scale = sin(PI_D * t / denom) * sin(PI_D * t / denom);
        return PRM(bid, comp) * scale;
    }
    if (mode == V_PIOLA_N)
        return PRM(bid, 3 * comp + 0) * nx + PRM(bid, 3 * comp + 1) * ny + PRM(bid, 3 * comp + 2) * nz;
    if (mode == V_AFFINE_NH) {
        double tt = t, active = PRM(bid, 11);
        if (active > 0.0 && tt > active) tt = active;
        double F[9], P[9];
        for (int k = 0; k < 9; ++k) F[k] = (k % 4 == 0 ? 1.0 : 0.0) + tt * PRM(bid, k);
        F[1] = tt * PRM(bid, 1);
        F[2] = tt * PRM(bid, 2);
        F[3] = tt * PRM(bid, 3);
        F[5] = tt * PRM(bid, 5);
        F[6] = tt * PRM(bid, 6);
        F[7] = tt * PRM(bid, 7);
        or3d_piola(F, M_NH, PRM(bid, 9), PRM(bid, 10), 0.0, 0.0, 0.0, s->j_floor, P);
        return P[3 * comp] * nx + P[3 * comp + 1] * ny + P[3 * comp + 2] * nz;
    }
    if (mode == V_NORMAL_T) return PRM(bid, 0) * nv[comp];
    if (mode == V_SINE2_NT) {
        double amp = PRM(bid, 0), active = PRM(bid, 1), denom = PRM(bid, 2);
        if (denom <= 0.0) denom = 2.0 * active;
        double scale = 1.0;
        if (active > 0.0 && t < active) scale = sin(PI_D * t / denom) * sin(PI_D * t / denom);
        return amp * scale * nv[comp];
    }
    if (mode == V_LINEAR_NT) {
        double amp = PRM(bid, 0), active = PRM(bid, 1), scale = 1.0;
        if (active > 0.0 && t < active) scale = t / active;
        return amp * scale * nv[comp];
    }
    if (mode == V_ZPULL) return tube_pull(&PRM(bid, 0), comp, x, y, t);
    if (mode == V_ZTWIST) return twist_velocity(&PRM(bid, 0), comp, x, y, t);
    if (mode == V_XPULL) return tube_pull(&PRM(bid, 0), tube_comp(comp), y, z, t);
    if (mode == V_XTWIST) return twist_velocity(&PRM(bid, 0), tube_comp(comp), y, z, t);
    return 0.0;
}

/* Test access: one boundary value (want_disp = 0, cbw3d.py:996-1133) or
 * Dirichlet anchor displacement (want_disp = 1, cbw3d.py:2351-2399) of value
 * mode `mode` with the 16 parameters `prm`, evaluated at boundary id 0. */
double or3d_value_probe(const double *prm, int mode, int comp, double x, double y, double z, double nx,
                        double ny, double nz, double t, int want_disp) {
    or3d_state st;
    double table[OR_MAX_IDS * OR_NPRM];
    memset(&st, 0, sizeof(st));
    memset(table, 0, sizeof(table));
    memcpy(table, prm, sizeof(double) * OR_NPRM);
    st.params_by_id = table;
    st.j_floor = 1e-12;
    if (want_disp) return dirichlet_disp(&st, mode, 0, comp, x, y, z, t);
    return curved_value(&st, mode, 0, comp, x, y, z, nx, ny, nz, t);
}

static inline void normalize3(double *v) {
    double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (r < 1.0e-30) r = 1.0e-30;
    v[0] = v[0] / r;
    v[1] = v[1] / r;
    v[2] = v[2] / r;
}

static inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* cbw3d.py:1137-1156 (_make_F_from_h) */
static void make_F(const double *h, const double *n, const double *t1, const double *t2,
                   const double *g1, const double *g2, double *F) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F[3 * i + j] = h[i] * n[j] + g1[i] * t1[j] + g2[i] * t2[j];
}

/* cbw3d.py:1160-1202 (_normal_residual_h) */
static void normal_residual(const or3d_state *s, const double *h, const double *n, const double *t1,
                            const double *t2, const double *g1, const double *g2, const double *T,
                            double *r) {
    double F[9], P[9];
    make_F(h, n, t1, t2, g1, g2, F);
    PIOLA(F, P);
    for (int i = 0; i < 3; ++i) r[i] = P[3 * i] * n[0] + P[3 * i + 1] * n[1] + P[3 * i + 2] * n[2] - T[i];
}

/* cbw3d.py:1206-1246 (_neo_hooke_acoustic_normal_3d) */
static void acoustic_normal(const double *F, const double *n, double lam, double mu, double jf,
                            double *Qn) {
    double J = det3(F);
    if (J < jf) J = jf;
    double F11 = F[0], F12 = F[1], F13 = F[2], F21 = F[3], F22 = F[4], F23 = F[5], F31 = F[6],
           F32 = F[7], F33 = F[8];
    double H11 = (F22 * F33 - F23 * F32) / J, H12 = (F23 * F31 - F21 * F33) / J,
           H13 = (F21 * F32 - F22 * F31) / J, H21 = (F13 * F32 - F12 * F33) / J,
           H22 = (F11 * F33 - F13 * F31) / J, H23 = (F12 * F31 - F11 * F32) / J,
           H31 = (F12 * F23 - F13 * F22) / J, H32 = (F13 * F21 - F11 * F23) / J,
           H33 = (F11 * F22 - F12 * F21) / J;
    double beta = 0.5 * lam * (J * J - 1.0) - mu;
    double coeff = lam * J * J - beta;
    double n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    double v[3] = {H11 * n[0] + H12 * n[1] + H13 * n[2], H21 * n[0] + H22 * n[1] + H23 * n[2],
                   H31 * n[0] + H32 * n[1] + H33 * n[2]};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Qn[3 * i + j] = (i == j ? mu * n2 : 0.0) + coeff * v[i] * v[j];
}

/* cbw3d.py:1250-1329 (_energy_Fb_neo_hooke_3d); returns 0 and a zero matrix on failure. */
static int energy_Fb(const double *F, const double *n, const double *T, int k0, int k1, int k2,
                     double lam, double mu, double jf, double *out) {
    double cof[9] = {F[4] * F[8] - F[5] * F[7], F[5] * F[6] - F[3] * F[8], F[3] * F[7] - F[4] * F[6],
                     F[2] * F[7] - F[1] * F[8], F[0] * F[8] - F[2] * F[6], F[1] * F[6] - F[0] * F[7],
                     F[1] * F[5] - F[2] * F[4], F[2] * F[3] - F[0] * F[5], F[0] * F[4] - F[1] * F[3]};
    double a[3], hold[3];
    for (int i = 0; i < 3; ++i) {
        a[i] = cof[3 * i] * n[0] + cof[3 * i + 1] * n[1] + cof[3 * i + 2] * n[2];
        hold[i] = F[3 * i] * n[0] + F[3 * i + 1] * n[1] + F[3 * i + 2] * n[2];
    }
    for (int k = 0; k < 9; ++k) out[k] = 0.0;
    if (k0 == BC_NEU && k1 == BC_NEU && k2 == BC_NEU) {
        double a2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        if (a2 > 1.0e-28) {
            double area = sqrt(a2);
            double e[3] = {a[0] / area, a[1] / area, a[2] / area};
            double tau = T[0] * e[0] + T[1] * e[1] + T[2] * e[2];
            double aa = mu + 0.5 * lam * a2, bb = mu + 0.5 * lam;
            double disc = sqrt(tau * tau + 4.0 * aa * bb);
            double sv = (tau + disc) / (2.0 * aa);
            if (tau < 0.0) sv = 2.0 * bb / (disc - tau);
            if (area * sv > jf) {
                double h[3];
                for (int i = 0; i < 3; ++i) h[i] = (T[i] - tau * e[i]) / mu + sv * e[i];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) out[3 * i + j] = F[3 * i + j] + (h[i] - hold[i]) * n[j];
                return 1;
            }
        }
        return 0;
    }
    double an[3] = {0, 0, 0}, tn[3] = {0, 0, 0}, fixed[3] = {hold[0], hold[1], hold[2]};
    int kk[3] = {k0, k1, k2};
    for (int i = 0; i < 3; ++i)
        if (kk[i] == BC_NEU) {
            an[i] = a[i];
            tn[i] = T[i];
            fixed[i] = 0.0;
        }
    double a2 = an[0] * an[0] + an[1] * an[1] + an[2] * an[2];
    double d = a[0] * fixed[0] + a[1] * fixed[1] + a[2] * fixed[2];
    if (a2 > 1.0e-28) {
        double aa = mu + 0.5 * lam * a2, bb = (mu + 0.5 * lam) * a2;
        double cc = mu * d + (an[0] * tn[0] + an[1] * tn[1] + an[2] * tn[2]);
        double disc = sqrt(cc * cc + 4.0 * aa * bb);
        double J = (cc + disc) / (2.0 * aa);
        if (cc < 0.0) J = 2.0 * bb / (disc - cc);
        if (J > jf) {
            double area = sqrt(a2);
            double e[3] = {an[0] / area, an[1] / area, an[2] / area};
            double te = tn[0] * e[0] + tn[1] * e[1] + tn[2] * e[2];
            double h[3];
            for (int i = 0; i < 3; ++i) h[i] = fixed[i] + (tn[i] - te * e[i]) / mu + ((J - d) / area) * e[i];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) out[3 * i + j] = F[3 * i + j] + (h[i] - hold[i]) * n[j];
            return 1;
        }
    } else if (a2 == 0.0 && d > jf) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                out[3 * i + j] = F[3 * i + j] + (fixed[i] + tn[i] / mu - hold[i]) * n[j];
        return 1;
    }
    return 0;
}

/* cbw3d.py:1333-1515 (_solve_Fb_arbitrary_normal) and :1519-1737
 * (_solve_Fb_componentwise_normal); `kinds` NULL selects the all-Neumann form. */
static void solve_Fb(const or3d_state *s, const double *Fin, double nx, double ny, double nz,
                     const double *T, const int *kinds, double *Fb) {
    double n[3] = {nx, ny, nz};
    normalize3(n);
    int k0 = kinds ? kinds[0] : BC_NEU, k1 = kinds ? kinds[1] : BC_NEU, k2 = kinds ? kinds[2] : BC_NEU;
    if (s->normal_solver == NS_ENERGY && s->model == M_NH) {
        double E[9];
        energy_Fb(Fin, n, T, k0, k1, k2, s->lam, s->mu, s->j_floor, E);
        if (det3(E) > s->j_floor) {
            memcpy(Fb, E, sizeof(double) * 9);
            return;
        }
    }
    double ref[3] = {1.0, 0.0, 0.0};
    if (fabs(n[0]) >= 0.75) {
        ref[0] = 0.0;
        ref[1] = 1.0;
    }
    double t1[3], t2[3];
    cross3(n, ref, t1);
    normalize3(t1);
    cross3(n, t1, t2);
    double h[3], g1[3], g2[3];
    for (int i = 0; i < 3; ++i) {
        h[i] = Fin[3 * i] * n[0] + Fin[3 * i + 1] * n[1] + Fin[3 * i + 2] * n[2];
        g1[i] = Fin[3 * i] * t1[0] + Fin[3 * i + 1] * t1[1] + Fin[3 * i + 2] * t1[2];
        g2[i] = Fin[3 * i] * t2[0] + Fin[3 * i + 1] * t2[1] + Fin[3 * i + 2] * t2[2];
    }
    const double eps = 1.0e-6;
    int neu[3] = {k0 == BC_NEU, k1 == BC_NEU, k2 == BC_NEU};
    for (int it = 0; it < 16; ++it) {
        double r[3];
        normal_residual(s, h, n, t1, t2, g1, g2, T, r);
        double rr[3] = {r[0], r[1], r[2]};
        if (kinds)
            for (int i = 0; i < 3; ++i)
                if (!neu[i]) rr[i] = 0.0;
        double rnorm = sqrt(rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2]);
        if (rnorm < 1.0e-10) break;
        double M[9];
        if (s->normal_solver != NJ_FD && s->model == M_NH) {
            double F[9];
            make_F(h, n, t1, t2, g1, g2, F);
            acoustic_normal(F, n, s->lam, s->mu, s->j_floor, M);
        } else {
            double rp[3][3];
            for (int col = 0; col < 3; ++col) {
                double hp[3] = {h[0], h[1], h[2]};
                hp[col] = h[col] + eps;
                normal_residual(s, hp, n, t1, t2, g1, g2, T, rp[col]);
            }
            for (int i = 0; i < 3; ++i)
                for (int col = 0; col < 3; ++col) M[3 * i + col] = (rp[col][i] - r[i]) / eps;
        }
        if (kinds)
            for (int i = 0; i < 3; ++i)
                if (!neu[i])
                    for (int col = 0; col < 3; ++col) M[3 * i + col] = (i == col) ? 1.0 : 0.0;
        double a = M[0], b = M[1], c = M[2], d = M[3], e = M[4], f = M[5], g = M[6], hh = M[7], ii = M[8];
        double det = a * (e * ii - f * hh) - b * (d * ii - f * g) + c * (d * hh - e * g);
        if (fabs(det) < 1.0e-14) break;
        double inv[9] = {(e * ii - f * hh) / det, (c * hh - b * ii) / det, (b * f - c * e) / det,
                         (f * g - d * ii) / det,  (a * ii - c * g) / det,  (c * d - a * f) / det,
                         (d * hh - e * g) / det,  (b * g - a * hh) / det,  (a * e - b * d) / det};
        double delta[3];
        for (int i = 0; i < 3; ++i) delta[i] = -(inv[3 * i] * rr[0] + inv[3 * i + 1] * rr[1] + inv[3 * i + 2] * rr[2]);
        double step = 1.0;
        for (int ls = 0; ls < 12; ++ls) {
            double cnd[3];
            for (int i = 0; i < 3; ++i) cnd[i] = (!kinds || neu[i]) ? h[i] + step * delta[i] : h[i];
            double F[9];
            make_F(cnd, n, t1, t2, g1, g2, F);
            if (det3(F) > s->j_floor) {
                h[0] = cnd[0];
                h[1] = cnd[1];
                h[2] = cnd[2];
                break;
            }
            step = 0.5 * step;
        }
    }
    make_F(h, n, t1, t2, g1, g2, Fb);
}

/* cbw3d.py:2888-3151 (_sparse_prepare_links_gpu) for the bfl, halfway,
 * compat_bfl, local_f and local_f_bfl reconstructions. */
static void prepare_links(or3d_state *s, double t) {
    const int n = s->n_active, L = s->n_links;
#pragma omp parallel for schedule(dynamic, 64)
    for (int ell = 0; ell < L; ++ell) {
        int bid = s->link_bid[ell];
        int kind = s->kind_by_id[bid], mode = s->mode_by_id[bid];
        int kk[3] = {s->comp_kind_by_id[0 * OR_MAX_IDS + bid], s->comp_kind_by_id[1 * OR_MAX_IDS + bid],
                     s->comp_kind_by_id[2 * OR_MAX_IDS + bid]};
        double x = s->link_xb[ell], y = s->link_yb[ell], z = s->link_zb[ell];
        double nx = s->link_nx[ell], ny = s->link_ny[ell], nz = s->link_nz[ell];
        double T[3];
        for (int k = 0; k < 3; ++k) T[k] = curved_value(s, mode, bid, k, x, y, z, nx, ny, nz, t);
        if (s->dyn_enabled[ell] != 0)
            for (int k = 0; k < 3; ++k) T[k] = s->dyn_traction[IDX(k, ell, L)];
        s->lkind[ell] = kind;
        for (int k = 0; k < 3; ++k) {
            s->lcomp_kind[IDX(k, ell, L)] = kk[k];
            s->lbv[IDX(k, ell, L)] = T[k];
        }
        double Fb[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Pb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (kk[0] == BC_NEU || kk[1] == BC_NEU || kk[2] == BC_NEU) {
            int p = s->link_p[ell];
            double F[9];
            for (int k = 0; k < 9; ++k) F[k] = s->U[IDX(3 + k, p, n)];
            if (s->recon == R_LOCALF_BFL || s->recon == R_COMPAT) {
                int up = s->link_up[ell];
                if (up >= 0) {
                    double eta = s->link_eta[ell];
                    for (int k = 0; k < 9; ++k)
                        F[k] = s->U[IDX(3 + k, p, n)] + eta * (s->U[IDX(3 + k, p, n)] - s->U[IDX(3 + k, up, n)]);
                }
            }
            /* cbw3d.py:3005-3049: geometry-compatible F of compat_bfl. */
            if (s->recon == R_COMPAT && s->tan_valid[ell] != 0) {
                double nv[3] = {nx, ny, nz};
                double r = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
                if (r < 1.0e-30) r = 1.0e-30;
                for (int c = 0; c < 3; ++c) nv[c] /= r;
                double ref[3] = {1.0, 0.0, 0.0};
                if (fabs(nv[0]) >= 0.75) {
                    ref[0] = 0.0;
                    ref[1] = 1.0;
                }
                double t1[3] = {nv[1] * ref[2] - nv[2] * ref[1], nv[2] * ref[0] - nv[0] * ref[2],
                                nv[0] * ref[1] - nv[1] * ref[0]};
                double r1 = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
                if (r1 < 1.0e-30) r1 = 1.0e-30;
                for (int c = 0; c < 3; ++c) t1[c] /= r1;
                double t2[3] = {nv[1] * t1[2] - nv[2] * t1[1], nv[2] * t1[0] - nv[0] * t1[2],
                                nv[0] * t1[1] - nv[1] * t1[0]};
                double du1[3] = {0, 0, 0}, du2[3] = {0, 0, 0}, wt_abs = 0.0;
                for (int m = 0; m < OR_MAX_TAN; ++m) {
                    int pp = s->tan_p[(size_t)ell * OR_MAX_TAN + m];
                    if (pp < 0) continue;
                    double w1 = s->tan_w1[(size_t)ell * OR_MAX_TAN + m], w2 = s->tan_w2[(size_t)ell * OR_MAX_TAN + m];
                    for (int c = 0; c < 3; ++c) {
                        du1[c] += w1 * s->u[IDX(c, pp, n)];
                        du2[c] += w2 * s->u[IDX(c, pp, n)];
                    }
                    wt_abs += fabs(w1) + fabs(w2);
                }
                if (wt_abs > 0.0)
                    for (int row = 0; row < 3; ++row) {
                        double h = F[3 * row] * nv[0] + F[3 * row + 1] * nv[1] + F[3 * row + 2] * nv[2];
                        double g1 = t1[row] + du1[row], g2 = t2[row] + du2[row];
                        for (int c = 0; c < 3; ++c) F[3 * row + c] = h * nv[c] + g1 * t1[c] + g2 * t2[c];
                    }
            }
            if (s->recon == R_LOCALF)
                for (int k = 0; k < 9; ++k) F[k] = s->char_state[((size_t)0 * NC + 3 + k) * L + ell];
            if (kk[0] == BC_NEU && kk[1] == BC_NEU && kk[2] == BC_NEU)
                solve_Fb(s, F, nx, ny, nz, T, NULL, Fb);
            else
                solve_Fb(s, F, nx, ny, nz, T, kk, Fb);
            PIOLA(Fb, Pb);
        }
        for (int k = 0; k < 9; ++k) {
            s->lFb[IDX(k, ell, L)] = Fb[k];
            s->lPb[IDX(k, ell, L)] = Pb[k];
        }
    }
}

/* cbw3d.py:3155-3241 (_sparse_apply_links) */
static void apply_links(or3d_state *s, const double *fpost, double *f1) {
    const int n = s->n_active, L = s->n_links;
#pragma omp parallel for schedule(static)
    for (int ell = 0; ell < L; ++ell) {
        int p = s->link_p[ell], qo = s->link_q[ell], qi = opp(qo);
        double eta = s->link_eta[ell];
        if (s->recon == R_HALF) eta = 0.5;
        double din[3] = {(double)dir_x(qi), (double)dir_y(qi), (double)dir_z(qi)};
        int up = s->link_up[ell];
        double v[3] = {s->lbv[IDX(0, ell, L)], s->lbv[IDX(1, ell, L)], s->lbv[IDX(2, ell, L)]};
        for (int a = 0; a < NC; ++a) {
            int row = a < 3 ? a : (a - 3) / 3;
            int kind = s->lcomp_kind[IDX(row, ell, L)];
            double D = 1.0, S = 0.0;
            if (kind == BC_DIR) {
                if (a < 3) {
                    D = -1.0;
                    S = v[a] / 3.0;
                } else {
                    int col = (a - 3) - row * 3;
                    S = -din[col] * v[row] / s->c;
                }
            } else {
                if (a < 3) {
                    S = -(din[0] * s->lPb[IDX(3 * a, ell, L)] + din[1] * s->lPb[IDX(3 * a + 1, ell, L)] +
                          din[2] * s->lPb[IDX(3 * a + 2, ell, L)]) / s->c;
                } else {
                    D = -1.0;
                    S = s->lFb[IDX(a - 3, ell, L)] / 3.0;
                }
            }
            double fo = fpost[IDX(qo * 12 + a, p, n)], fi = fpost[IDX(qi * 12 + a, p, n)];
            double val;
            if (eta >= 0.5) {
                double alpha = 1.0 / (2.0 * eta);
                val = alpha * (D * fo + S) + (1.0 - alpha) * fi;
            } else {
                double upv = fo;
                if (up >= 0) upv = fpost[IDX(qo * 12 + a, up, n)];
                val = D * (2.0 * eta * fo + (1.0 - 2.0 * eta) * upv) + S;
            }
            f1[IDX(qi * 12 + a, p, n)] = val;
        }
    }
}

static inline double det_state(const double *U) {
    return U[3] * (U[7] * U[11] - U[8] * U[10]) - U[4] * (U[6] * U[11] - U[8] * U[9]) +
           U[5] * (U[6] * U[10] - U[7] * U[9]);
}

/* cbw3d.py:2729-2752 (_state_flux_jvp3) */
static void flux_jvp3(const or3d_state *s, int axis, const double *U, const double *W, double *out) {
    double scale = 1.0;
    for (int a = 0; a < NC; ++a)
        if (fabs(W[a]) > scale) scale = fabs(W[a]);
    double eps = 1.0e-6 / scale;
    double Fp[9], Fm[9], Pp[9], Pm[9];
    for (int k = 0; k < 9; ++k) {
        Fp[k] = U[3 + k] + eps * W[3 + k];
        Fm[k] = U[3 + k] - eps * W[3 + k];
    }
    PIOLA(Fp, Pp);
    PIOLA(Fm, Pm);
    for (int a = 0; a < NC; ++a) out[a] = 0.0;
    for (int i = 0; i < 3; ++i) {
        out[i] = -(Pp[3 * i + axis] - Pm[3 * i + axis]) / (2.0 * eps);
        out[3 + 3 * i + axis] = -W[i];
    }
}

/* cbw3d.py:2756-2809 (_local_characteristic_population3) */
static void char_population(const or3d_state *s, const double *Ub, const double *Gx, const double *Gy,
                            const double *Gz, const double *d, double eta, double *outp) {
    const double dt = s->dt, c = s->c;
    double shift[NC], Uc[NC], tmp[NC];
    for (int a = 0; a < NC; ++a) shift[a] = eta * dt * c * (d[0] * Gx[a] + d[1] * Gy[a] + d[2] * Gz[a]);
    double chi = 1.0;
    for (int guard = 0; guard < 10; ++guard) {
        for (int a = 0; a < NC; ++a) tmp[a] = Ub[a] + chi * shift[a];
        if (det_state(tmp) > 10.0 * s->j_floor) break;
        chi = 0.5 * chi;
    }
    for (int a = 0; a < NC; ++a) Uc[a] = Ub[a] + chi * shift[a];
    double B[NC];
    for (int a = 0; a < NC; ++a) B[a] = source_comp(a, s->source_mode, s->source_params, Uc[0], Uc[1], Uc[2]);
    double j0[NC], j1[NC], j2[NC], w[NC];
    flux_jvp3(s, 0, Uc, Gx, j0);
    flux_jvp3(s, 1, Uc, Gy, j1);
    flux_jvp3(s, 2, Uc, Gz, j2);
    for (int a = 0; a < NC; ++a)
        w[a] = B[a] - j0[a] - j1[a] - j2[a] + c * (d[0] * Gx[a] + d[1] * Gy[a] + d[2] * Gz[a]);
    int axis = 0;
    if (fabs(d[1]) > 0.5) axis = 1;
    if (fabs(d[2]) > 0.5) axis = 2;
    double Aw[NC], P[9], flux[NC];
    flux_jvp3(s, axis, Uc, w, Aw);
    PIOLA(Uc + 3, P);
    for (int a = 0; a < NC; ++a) flux[a] = 0.0;
    for (int i = 0; i < 3; ++i) {
        flux[i] = -P[3 * i + axis];
        flux[3 + 3 * i + axis] = -Uc[i];
    }
    double da = d[axis];
    for (int a = 0; a < NC; ++a)
        outp[a] = (Uc[a] + 3.0 * da * flux[a] / c - dt / s->omega * (w[a] + 3.0 * da * Aw[a] / c) +
                   dt * (1.0 / s->omega - 0.5) * B[a]) / 6.0;
}

/* cbw3d.py:2813-2884 (_sparse_apply_characteristic) */
static void apply_characteristic(or3d_state *s, double *f1) {
    const int n = s->n_active, L = s->n_links;
    int bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(dynamic, 64)
    for (int ell = 0; ell < L; ++ell) {
        int kk[3] = {s->lcomp_kind[IDX(0, ell, L)], s->lcomp_kind[IDX(1, ell, L)], s->lcomp_kind[IDX(2, ell, L)]};
        if (kk[0] != BC_NEU && kk[1] != BC_NEU && kk[2] != BC_NEU) continue;
        int p = s->link_p[ell], qi = opp(s->link_q[ell]);
        double Ub[NC], Gx[NC], Gy[NC], Gz[NC];
        for (int a = 0; a < NC; ++a) {
            Gx[a] = s->char_state[((size_t)1 * NC + a) * L + ell];
            Gy[a] = s->char_state[((size_t)2 * NC + a) * L + ell];
            Gz[a] = s->char_state[((size_t)3 * NC + a) * L + ell];
            if (a < 3) {
                Ub[a] = s->U[IDX(a, p, n)];
                if (kk[a] == BC_DIR) Ub[a] = s->lbv[IDX(a, ell, L)];
            } else {
                Ub[a] = s->lFb[IDX(a - 3, ell, L)];
            }
        }
        double d[3] = {(double)dir_x(qi), (double)dir_y(qi), (double)dir_z(qi)};
        if (det_state(Ub) <= s->j_floor) bad += 1;
        double inc[NC];
        char_population(s, Ub, Gx, Gy, Gz, d, s->link_eta[ell], inc);
        for (int a = 0; a < NC; ++a) {
            int row = a < 3 ? a : (a - 3) / 3;
            if (kk[row] == BC_NEU) f1[IDX(qi * 12 + a, p, n)] = inc[a];
        }
    }
    s->invalid[0] += bad;
}

/* cbw3d.py:4339-4399 (apply_boundary_reconstruction) */
/* vnew3d:1393-1452 (_boundary_value_comp): face datum of the box. */
static double box_value(int comp, int mode, const double *p, double x, double y, double z, double t) {
    if (mode == 1) return p[comp];
    if (mode == 2) {
        double scale = 1.0;
        if (p[3] > 0.0 && t < p[3]) scale = sin(PI_D * t / p[3]) * sin(PI_D * t / p[3]);
        return p[comp] * scale;
    }
    if (mode == 3) {
        double scale = 1.0, denom = p[4];
        if (p[3] > 0.0 && t < p[3]) {
            if (denom <= 0.0) denom = 2.0 * p[3];
            scale = sin(PI_D * t / denom) * sin(PI_D * t / denom);
        }
        return p[comp] * scale;
    }
    if (mode == 4) {
        if (p[12] > 0.0 && t > p[12]) return 0.0;
        return p[3 * comp] * x + p[3 * comp + 1] * y + p[3 * comp + 2] * z + p[9 + comp];
    }
    if (mode == 5) return (comp == 0 && t < p[1]) ? p[0] * sin(PI_D * t / p[1]) : 0.0;
    return 0.0;
}

/* vnew3d:1455-1766 (_normal_residual_from_col, _det_from_col,
 * _solve_normal_column): face F of an axis-aligned Neumann face. */
static void col_F(const double *Fin, int axis, const double *x, double *G) {
    for (int k = 0; k < 9; ++k) G[k] = Fin[k];
    for (int r = 0; r < 3; ++r) G[3 * r + axis] = x[r];
}
static void col_residual(const or3d_state *s, const double *Fin, int axis, double sign, const double *T,
                         const double *x, double *r) {
    double G[9], P[9];
    col_F(Fin, axis, x, G);
    or3d_piola(G, s->model, s->lam, s->mu, s->p0, s->p1, s->p2, s->j_floor, P);
    for (int row = 0; row < 3; ++row) r[row] = sign * P[3 * row + axis] - T[row];
}
static void solve_normal_column(const or3d_state *s, const double *Fin, int axis, double sign, const double *T,
                                double *Fout) {
    double x[3] = {Fin[axis], Fin[3 + axis], Fin[6 + axis]};
    const double eps = 1.0e-6;
    for (int it = 0; it < 16; ++it) {
        double r[3], rp[3][3];
        col_residual(s, Fin, axis, sign, T, x, r);
        if (sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) < 1.0e-10) break;
        for (int c = 0; c < 3; ++c) {
            double xc[3] = {x[0], x[1], x[2]};
            xc[c] = x[c] + eps;
            col_residual(s, Fin, axis, sign, T, xc, rp[c]);
        }
        double a = (rp[0][0] - r[0]) / eps, b = (rp[1][0] - r[0]) / eps, c = (rp[2][0] - r[0]) / eps;
        double d = (rp[0][1] - r[1]) / eps, e = (rp[1][1] - r[1]) / eps, f = (rp[2][1] - r[1]) / eps;
        double g = (rp[0][2] - r[2]) / eps, h = (rp[1][2] - r[2]) / eps, ii = (rp[2][2] - r[2]) / eps;
        double det = a * (e * ii - f * h) - b * (d * ii - f * g) + c * (d * h - e * g);
        if (fabs(det) < 1.0e-14) break;
        double i00 = (e * ii - f * h) / det, i01 = (c * h - b * ii) / det, i02 = (b * f - c * e) / det;
        double i10 = (f * g - d * ii) / det, i11 = (a * ii - c * g) / det, i12 = (c * d - a * f) / det;
        double i20 = (d * h - e * g) / det, i21 = (b * g - a * h) / det, i22 = (a * e - b * d) / det;
        double d0 = -(i00 * r[0] + i01 * r[1] + i02 * r[2]);
        double d1 = -(i10 * r[0] + i11 * r[1] + i12 * r[2]);
        double d2 = -(i20 * r[0] + i21 * r[1] + i22 * r[2]);
        double step = 1.0;
        for (int ls = 0; ls < 12; ++ls) {
            double cand[3] = {x[0] + step * d0, x[1] + step * d1, x[2] + step * d2}, G[9];
            col_F(Fin, axis, cand, G);
            if (det3(G) > s->j_floor) {
                x[0] = cand[0];
                x[1] = cand[1];
                x[2] = cand[2];
                break;
            }
            step = 0.5 * step;
        }
    }
    col_F(Fin, axis, x, Fout);
}

/* vnew3d:1999-2200 (_apply_boundary_side_kernel), faces in order left,
 * right, bottom, top, front, back. */
static void apply_sides(or3d_state *s, const double *fpost, double *f1, double tb) {
    const int nx = s->box_n[0], ny = s->box_n[1], nz = s->box_n[2], n = s->n_active;
    for (int side = 0; side < 6; ++side) {
        const int kind = s->side_kind[side];
        if (kind == 0) continue;
        const int axis = side / 2, hi = side & 1;
        const double c = axis_speed(s, axis); /* speed along the face normal, vnew3d:2123-2127 */
        const int nn = axis == 0 ? ny * nz : (axis == 1 ? nx * nz : nx * ny);
        const int q_out = hi ? axis : axis + 3, q_in = hi ? axis + 3 : axis;
        const double outward = hi ? 1.0 : -1.0;
        const double *p = s->side_params + 16 * side;
        for (int t = 0; t < nn; ++t) {
            int i, j, k;
            if (axis == 0) { i = hi ? nx - 1 : 0; j = t % ny; k = t / ny; }
            else if (axis == 1) { j = hi ? ny - 1 : 0; i = t % nx; k = t / nx; }
            else { k = hi ? nz - 1 : 0; i = t % nx; j = t / nx; }
            const int node = (i * ny + j) * nz + k;
            double xb[3] = {(i + 0.5) * s->dx, (j + 0.5) * s->dy, (k + 0.5) * s->dz};
            xb[axis] = hi ? s->box_length[axis] : 0.0;
            double v[3];
            for (int cc = 0; cc < 3; ++cc)
                v[cc] = box_value(cc, s->side_mode[side], p, xb[0], xb[1], xb[2], tb);
            double fo[12];
            for (int a = 0; a < 12; ++a) fo[a] = fpost[IDX(q_out * 12 + a, node, n)];
            double val[12];
            if (kind == BC_DIR) {
                for (int cc = 0; cc < 3; ++cc) val[cc] = -fo[cc] + v[cc] / 3.0;
                for (int r = 0; r < 3; ++r)
                    for (int cc = 0; cc < 3; ++cc)
                        val[3 + 3 * r + cc] = fo[3 + 3 * r + cc] + (cc == axis ? outward : 0.0) * v[r] / c;
            } else {
                double Fn[9], Fb[9];
                for (int a = 0; a < 9; ++a) Fn[a] = s->U[IDX(3 + a, node, n)];
                solve_normal_column(s, Fn, axis, outward, v, Fb);
                for (int cc = 0; cc < 3; ++cc) val[cc] = fo[cc] + v[cc] / c;
                for (int a = 0; a < 9; ++a) val[3 + a] = -fo[3 + a] + Fb[a] / 3.0;
            }
            for (int a = 0; a < 12; ++a) f1[IDX(q_in * 12 + a, node, n)] = val[a];
        }
    }
}

int or3d_apply_boundary(or3d_state *s, double t, int in_place) {
    if (s->box) {
        if (in_place) {
            memcpy(s->f1, s->f0, sizeof(double) * FQ * (size_t)s->n_active);
            apply_sides(s, s->f1, s->f0, t);
        } else {
            apply_sides(s, s->f0, s->f1, t);
        }
        return 0;
    }
    if (s->n_links == 0) return 0;
    if (s->has_char) char_gather(s);
    prepare_links(s, t);
    if (in_place) memcpy(s->f1, s->f0, sizeof(double) * FQ * (size_t)s->n_active);
    double *src = in_place ? s->f1 : s->f0;
    double *tgt = in_place ? s->f0 : s->f1;
    apply_links(s, src, tgt);
    if (s->has_char) apply_characteristic(s, tgt);
    return 0;
}

/* cbw3d.py:4549-4589 (step) */
int or3d_step(or3d_state *s, double time, int64_t steps, int nsteps) {
    for (int k = 0; k < nsteps; ++k) {
        s->invalid[0] = 0;
        collide_inplace(s);
        double tb = time + 0.5 * s->dt;
        if (s->disp_interval > 0 && steps % s->disp_interval == 0) {
            build_bc(s, tb);
            for (int sw = 0; sw < s->disp_sweeps; ++sw) {
                disp_sweep(s);
                disp_commit(s);
            }
        }
        if (s->compat_interval > 0 && steps % s->compat_interval == 0) repair_F(s);
        stream_push(s);
        if (s->box) {
            apply_sides(s, s->f0, s->f1, tb);
        } else if (s->n_links > 0) {
            if (s->has_char) char_gather(s);
            prepare_links(s, tb);
            apply_links(s, s->f0, s->f1);
            if (s->has_char) apply_characteristic(s, s->f1);
        }
        double *tmp = s->f0;
        s->f0 = s->f1;
        s->f1 = tmp;
        time += s->dt;
        steps += 1;
    }
    return 0;
}

int or3d_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
