// libvlbm: D2Q4x6 vectorial LBM time step for MI355X (gfx950).
//
// Storage is the reference's dense (components, nx, ny) layout, node index
// p = i*ny + j, so the +-y neighbours of a node are adjacent in memory and
// the +-x neighbours one row away.  Every per-node kernel runs over the
// compacted list of active nodes, so the inactive part of the box is never
// read or written.
//
// One step (cbw2d:4180-4256) runs as
//   k2_moments  moments of f0 (U, u, u*), J flag                     first step of a call only
//   k2_sweep    local displacement sweep into u_next                 per node   [if due]
//   k2_commit   intermediate sweep commits                           per node   [sweeps > 1]
//   k2_record   F repair + deferred-collision record + final commit  per node
//   k2_links    boundary value, predictor, normal-traction Newton,
//               BFL / characteristic populations of every cut link   per link
//   k2_stream   pull stream fused with the next step's moments       per node
// Post-collision populations are never materialised: k2_record stores 12
// doubles per node from which fpost[q][a] = (1 - omega) f[q][a] + A[a] +
// s_q B[axis(q)][a] follows exactly (including the equilibrium change of the
// F repair), and k2_stream / k2_links apply it while pulling.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/vlbm2d.h"
#include "vlbm2d_device.h"
#include "vlbm_common.h"

using namespace vlbm2;
using vlbm_common::BLOCK;
using vlbm_common::fail;
using vlbm_common::fail_msg;
using vlbm_common::TimedSlot;
using vlbm_common::Timing;
using vlbm_common::xcd_block;
using vlbm_common::xcd_grid;

namespace {

constexpr int LBLOCK = 64;
constexpr int RED_BLOCKS = 1024;

// Record layout (12 planes): A[6], B_v[r*2+axis] (4), B_F[r] (2).
enum { R_A = 0, R_BV = 6, R_BF = 10 };

struct Params2 {
    int nx, ny, na, L;
    size_t N;
    Material2 m;
    double dt, c, omega, dx, dy;
    int src_mode;
    double src[6];
    int recon, solver;
    double disp_relax, w2;
    double blend;
    int interior_only;
    int use_dyn;
    double gw;    // mrt_ghost: 0.25 (omega - omega_ghost), 0 for bgk
    double filt;  // kinetic ghost filter strength, 0 = off
    int per_x, per_y;  // periodic axes (rectangular box solver)
    const double *clock;  // device time of the step start (captured step loops), or nullptr
};

// Boundary time of a step: the host value, or clock + dt/2 when the launch
// belongs to a captured step loop that keeps the time on the device.
__device__ __forceinline__ double step_time(double tb, const Params2 &P) {
    return P.clock ? P.clock[0] + 0.5 * P.dt : tb;
}

struct Links2 {
    const int32_t *lp, *lq, *lbid;
    const double *eta, *xb, *yb, *nx, *ny;
    const int32_t *lsq_p;
    const double *wb, *wx, *wy;
    const int32_t *valid;
    const int32_t *kind_by_id, *mode_by_id;
    const double *prm_by_id;
    const double *dyn_tx, *dyn_ty;
    const int32_t *act;
};

// Element offsets of population and record planes: adjacent planes paired,
// pair (plane / 2) * N + p holds planes 2k and 2k + 1 side by side, so the six
// values of one direction at a node are three aligned 16-byte (FP64) or
// 8-byte (FP32) pairs, which the pull stream moves as one access each.  The
// reference's (24, nx, ny) view exists only through the conversion kernels.
template <typename T>
__device__ __forceinline__ size_t pofs2(const T *, int plane, size_t p, size_t N) {
    return (((size_t)(plane >> 1) * N + p) << 1) + (size_t)(plane & 1);
}
typedef float float_pair2 __attribute__((ext_vector_type(2)));
typedef double double_pair2 __attribute__((ext_vector_type(2)));

// Population access.  double planes hold f; float planes hold f - f_rest
// where f_rest[q][a] = 1/4 on the diagonal of F (a = 2, 5), 0 elsewhere.
__device__ __forceinline__ double rest(int a) { return (a == 2 || a == 5) ? 0.25 : 0.0; }
__device__ __forceinline__ double ld(const double *f, size_t i, int) { return f[i]; }
__device__ __forceinline__ double ld(const float *f, size_t i, int a) { return (double)f[i] + rest(a); }
__device__ __forceinline__ void st(double *f, size_t i, int, double v) { f[i] = v; }
__device__ __forceinline__ void st(float *f, size_t i, int a, double v) { f[i] = (float)(v - rest(a)); }
__device__ __forceinline__ double ld_nt(const double *f, size_t i, int) { return __builtin_nontemporal_load(f + i); }
__device__ __forceinline__ double ld_nt(const float *f, size_t i, int a) {
    return (double)__builtin_nontemporal_load(f + i) + rest(a);
}
// Store v and return the value as stored (rounded to the plane precision).
__device__ __forceinline__ double st_nt(double *f, size_t i, int, double v) {
    __builtin_nontemporal_store(v, f + i);
    return v;
}
__device__ __forceinline__ double st_nt(float *f, size_t i, int a, double v) {
    const float x = (float)(v - rest(a));
    __builtin_nontemporal_store(x, f + i);
    return (double)x + rest(a);
}
// Planes (plane, plane + 1) at node p, plane even (components a, a + 1).
__device__ __forceinline__ void ld2_nt(const double *f, int plane, size_t p, size_t N, int, double &x0, double &x1) {
    const double_pair2 v = __builtin_nontemporal_load((const double_pair2 *)(f + pofs2(f, plane, p, N)));
    x0 = v.x;
    x1 = v.y;
}
__device__ __forceinline__ void ld2_nt(const float *f, int plane, size_t p, size_t N, int a, double &x0, double &x1) {
    const float_pair2 v = __builtin_nontemporal_load((const float_pair2 *)(f + pofs2(f, plane, p, N)));
    x0 = (double)v.x + rest(a);
    x1 = (double)v.y + rest(a + 1);
}
// Store planes (plane, plane + 1); v0 / v1 return the values as stored.
__device__ __forceinline__ void st2_nt(double *f, int plane, size_t p, size_t N, int, double &v0, double &v1) {
    double_pair2 x;
    x.x = v0;
    x.y = v1;
    __builtin_nontemporal_store(x, (double_pair2 *)(f + pofs2(f, plane, p, N)));
}
__device__ __forceinline__ void st2_nt(float *f, int plane, size_t p, size_t N, int a, double &v0, double &v1) {
    float_pair2 x;
    x.x = (float)(v0 - rest(a));
    x.y = (float)(v1 - rest(a + 1));
    __builtin_nontemporal_store(x, (float_pair2 *)(f + pofs2(f, plane, p, N)));
    v0 = (double)x.x + rest(a);
    v1 = (double)x.y + rest(a + 1);
}

// Records are stored in the precision of the populations; float records
// hold A - omega f_rest on the A planes (rest state subtracted).
__device__ __forceinline__ double rld(const double *r, size_t i, int, double) { return r[i]; }
__device__ __forceinline__ double rld(const float *r, size_t i, int a, double w) {
    return (double)r[i] + w * rest(a);
}
__device__ __forceinline__ double rld_b(const double *r, size_t i) { return r[i]; }
__device__ __forceinline__ double rld_b(const float *r, size_t i) { return (double)r[i]; }
__device__ __forceinline__ void rst_a(double *r, size_t i, int, double, double v) { r[i] = v; }
__device__ __forceinline__ void rst_a(float *r, size_t i, int a, double w, double v) {
    r[i] = (float)(v - w * rest(a));
}
__device__ __forceinline__ void rst_b(double *r, size_t i, double v) { r[i] = v; }
__device__ __forceinline__ void rst_b(float *r, size_t i, double v) { r[i] = (float)v; }
// Record planes (k, k + 1), k even, at node p in one access; fa0 / fa1 the
// A-plane index whose rest part a float record subtracts, or -1 for a B plane.
__device__ __forceinline__ void rst2(double *r, int k, size_t p, size_t N, int, int, double, double v0, double v1) {
    double_pair2 x;
    x.x = v0;
    x.y = v1;
    *(double_pair2 *)(r + pofs2(r, k, p, N)) = x;
}
__device__ __forceinline__ void rst2(float *r, int k, size_t p, size_t N, int fa0, int fa1, double w, double v0,
                                     double v1) {
    float_pair2 x;
    x.x = (float)(fa0 >= 0 ? v0 - w * rest(fa0) : v0);
    x.y = (float)(fa1 >= 0 ? v1 - w * rest(fa1) : v1);
    *(float_pair2 *)(r + pofs2(r, k, p, N)) = x;
}

// Record planes (k, k + 1), k even, at node p in one access: A planes
// (rld2, rest part added back for float records) and B planes (rldb2).
__device__ __forceinline__ void rld2(const double *r, int k, size_t p, size_t N, double, double &x0, double &x1) {
    const double_pair2 v = *(const double_pair2 *)(r + pofs2(r, k, p, N));
    x0 = v.x;
    x1 = v.y;
}
__device__ __forceinline__ void rld2(const float *r, int k, size_t p, size_t N, double w, double &x0, double &x1) {
    const float_pair2 v = *(const float_pair2 *)(r + pofs2(r, k, p, N));
    x0 = (double)v.x + w * rest(k - R_A);
    x1 = (double)v.y + w * rest(k - R_A + 1);
}
__device__ __forceinline__ void rldb2(const double *r, int k, size_t p, size_t N, double &x0, double &x1) {
    const double_pair2 v = *(const double_pair2 *)(r + pofs2(r, k, p, N));
    x0 = v.x;
    x1 = v.y;
}
__device__ __forceinline__ void rldb2(const float *r, int k, size_t p, size_t N, double &x0, double &x1) {
    const float_pair2 v = *(const float_pair2 *)(r + pofs2(r, k, p, N));
    x0 = (double)v.x;
    x1 = (double)v.y;
}

// fpost[q][a] at node p from its pre-collision value fv and the node's record.
template <typename R>
__device__ __forceinline__ double fpost_v(double fv, const R *rec, size_t N, int q, int a, int p, double om1,
                                          double w) {
    const int axis = q & 1;
    const double s = q < 2 ? 1.0 : -1.0;
    double b = 0.0;
    if (a < 2) {
        b = rld_b(rec, pofs2(rec, R_BV + 2 * a + axis, p, N));
    } else if (((a - 2) & 1) == axis) {
        b = rld_b(rec, pofs2(rec, R_BF + (a - 2) / 2, p, N));
    }
    return om1 * fv + rld(rec, pofs2(rec, R_A + a, p, N), a, w) + s * b;
}

// fpost_v with the record's entries already loaded: Av = A[a], B[r] =
// B_V[2 r + axis(q)], B[2 + r] = B_F[r].
__device__ __forceinline__ double fpost_b(double fv, double Av, const double *B, int q, int a, double om1) {
    const int axis = q & 1;
    const double s = q < 2 ? 1.0 : -1.0;
    double b = 0.0;
    if (a < 2) {
        b = B[a];
    } else if (((a - 2) & 1) == axis) {
        b = B[2 + (a - 2) / 2];
    }
    return om1 * fv + Av + s * b;
}

// Ghost part of the mrt_ghost collision (cbw2d:1847-2041).  Relaxing the
// ghost moment m3 = fE - fN + fW - fS with omega_ghost instead of omega adds
// 0.25 (omega - omega_ghost) m3 to fpost on E, W and subtracts it on N, S;
// gh holds m3 of the pre-collision populations, (6, N).
__device__ __forceinline__ double ghost_term(const double *gh, size_t N, int q, int a, int p, double gw) {
    const double g = gw * gh[(size_t)a * N + p];
    return (q & 1) ? -g : g;
}

__global__ void k2_reset_invalid(int32_t *invalid) { invalid[0] = 0; }
__global__ void k2_set_clock(double *clock, double t) { clock[0] = t; }
// time += dt, the same FP64 operation the host performs after each step.
__global__ void k2_advance_clock(double *clock, double dt) { clock[0] = clock[0] + dt; }

// Moments with the half-step source; u = u* + dt/2 v and (collide only)
// u* = u + dt/2 v; J flag.  cbw2d:1692-1790 and :3128-3195.
// With gout (mrt_ghost) also the ghost moment m3 of every component.
template <typename T, bool COLLIDE>
__device__ __forceinline__ void k2_moments_body(const T *__restrict__ f, double *__restrict__ U,
                                                    double *__restrict__ u, double *__restrict__ us,
                                                    int32_t *__restrict__ invalid,
                                                    const int32_t *__restrict__ list, double *__restrict__ gout,
                                                    Params2 P, int t) {
    if (t >= P.na) return;
    const int p = list[t];
    const size_t N = P.N;
    double Um[NC], Gm[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) Um[a] = Gm[a] = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int a = 0; a < NC; ++a) {
            const double x = ld(f, pofs2(f, q * 6 + a, p, N), a);
            Um[a] += x;
            Gm[a] += (q & 1) ? -x : x;
        }
    if (gout)
#pragma unroll
        for (int a = 0; a < NC; ++a) gout[(size_t)a * N + p] = Gm[a];
    const double hdt = 0.5 * P.dt;
    if (P.src_mode == SRC_CONST) {
#pragma unroll
        for (int a = 0; a < NC; ++a) Um[a] += hdt * P.src[a];
    } else if (P.src_mode == SRC_DAMP) {
        const double den = 1.0 + hdt * P.src[0];
        Um[0] = Um[0] / den;
        Um[1] = Um[1] / den;
    }
#pragma unroll
    for (int a = 0; a < NC; ++a) U[a * N + p] = Um[a];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const double uc = us[c * N + p] + hdt * Um[c];
        u[c * N + p] = uc;
        if (COLLIDE) us[c * N + p] = uc + hdt * Um[c];
    }
    if (Um[2] * Um[5] - Um[3] * Um[4] <= P.m.jf) atomicOr(invalid, 1);
}
template <typename T, bool COLLIDE>
__global__ __launch_bounds__(BLOCK) void k2_moments(const T *__restrict__ f, double *__restrict__ U,
                                                    double *__restrict__ u, double *__restrict__ us,
                                                    int32_t *__restrict__ invalid,
                                                    const int32_t *__restrict__ list, double *__restrict__ gout,
                                                    Params2 P) {
    k2_moments_body<T, COLLIDE>(f, U, u, us, invalid, list, gout, P, xcd_block() * BLOCK + threadIdx.x);
}

// Per active node, which lattice neighbours along the axes are on the
// lattice and active: first (+-x, +-y) and second (+-2x, +-2y) neighbours,
// without periodic wrap.  Geometry only; computed once, so the sweep and the
// record kernels test bits instead of loading the dense mask.
enum { NB_XP = 1, NB_XM = 2, NB_YP = 4, NB_YM = 8, NB_XP2 = 16, NB_XM2 = 32, NB_YP2 = 64, NB_YM2 = 128,
       NB_NEAR = 15 };
__global__ __launch_bounds__(BLOCK) void k2_nbmask(const int32_t *__restrict__ list, const int32_t *__restrict__ act,
                                                   uint8_t *__restrict__ nbm, Params2 P) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= P.na) return;
    const int p = list[t];
    const int ny = P.ny, nx = P.nx;
    const int i = p / ny, j = p - i * ny;
    int b = 0;
    if (i + 1 < nx && act[p + ny] != 0) b |= NB_XP;
    if (i > 0 && act[p - ny] != 0) b |= NB_XM;
    if (j + 1 < ny && act[p + 1] != 0) b |= NB_YP;
    if (j > 0 && act[p - 1] != 0) b |= NB_YM;
    if (i + 2 < nx && act[p + 2 * ny] != 0) b |= NB_XP2;
    if (i > 1 && act[p - 2 * ny] != 0) b |= NB_XM2;
    if (j + 2 < ny && act[p + 2] != 0) b |= NB_YP2;
    if (j > 1 && act[p - 2] != 0) b |= NB_YM2;
    nbm[t] = (uint8_t)b;
}

// One relaxed nearest-neighbour compatibility sweep of u (cbw2d:3421-3490)
// with the Dirichlet cut-point anchors of the node's own links
// (cbw2d:3288-3417) summed in link order, so the sums are deterministic.
__device__ __forceinline__ void k2_sweep_body(const double *__restrict__ U, const double *__restrict__ u,
                                                  double *__restrict__ un, const int32_t *__restrict__ list,
                                                  const uint8_t *__restrict__ nbm, const int32_t *__restrict__ aoff,
                                                  const int32_t *__restrict__ aids, Links2 T, double time,
                                                  Params2 P, int t) {
    if (t >= P.na) return;
    const int p = list[t];
    const int ny = P.ny;
    const int i = p / ny, j = p - i * ny;
    const size_t N = P.N;
    const double *U2 = U + 2 * N, *U3 = U + 3 * N, *U4 = U + 4 * N, *U5 = U + 5 * N;
    const double *u0 = u, *u1 = u + N;
    double diag = 0.0, rx = 0.0, ry = 0.0;
    const int l0 = aoff[t], l1 = aoff[t + 1];
    if (l1 > l0) time = step_time(time, P);
    for (int l = l0; l < l1; ++l) {
        const int ell = aids[l];
        const int bid = T.lbid[ell];
        const int mode = T.mode_by_id[bid];
        const double *prm = T.prm_by_id + bid * 8;
        const double xb = T.xb[ell], yb = T.yb[ell], nxb = T.nx[ell], nyb = T.ny[ell];
        double ub0 = 0.0, ub1 = 0.0;
        if (mode == V_CONST) {
            ub0 = time * value2(0, mode, prm, xb, yb, nxb, nyb, 0.0);
            ub1 = time * value2(1, mode, prm, xb, yb, nxb, nyb, 0.0);
        } else if (mode == V_AFFINE_VEL) {
            double at = time;
            if (prm[6] > 0.0 && at > prm[6]) at = prm[6];
            ub0 = at * value2(0, mode, prm, xb, yb, nxb, nyb, 0.0);
            ub1 = at * value2(1, mode, prm, xb, yb, nxb, nyb, 0.0);
        }
        const double dbx = xb - ((double)i + 0.5) * P.dx;
        const double dby = yb - ((double)j + 0.5) * P.dy;
        const double tx = ub0 - ((U2[p] - 1.0) * dbx + U3[p] * dby);
        const double ty = ub1 - (U4[p] * dbx + (U5[p] - 1.0) * dby);
        diag += P.w2;
        rx += P.w2 * tx;
        ry += P.w2 * ty;
    }
    const double dx = P.dx, dy = P.dy;
    const int nb = nbm[t];
    if (nb & NB_XP) {
        const int q = p + ny;
        const double gx = 0.5 * dx * ((U2[p] - 1.0) + (U2[q] - 1.0));
        const double gy = 0.5 * dx * (U4[p] + U4[q]);
        rx += u0[q] - gx;
        ry += u1[q] - gy;
        diag += 1.0;
    }
    if (nb & NB_XM) {
        const int q = p - ny;
        const double gx = 0.5 * dx * ((U2[q] - 1.0) + (U2[p] - 1.0));
        const double gy = 0.5 * dx * (U4[q] + U4[p]);
        rx += u0[q] + gx;
        ry += u1[q] + gy;
        diag += 1.0;
    }
    if (nb & NB_YP) {
        const int q = p + 1;
        const double gx = 0.5 * dy * (U3[p] + U3[q]);
        const double gy = 0.5 * dy * ((U5[p] - 1.0) + (U5[q] - 1.0));
        rx += u0[q] - gx;
        ry += u1[q] - gy;
        diag += 1.0;
    }
    if (nb & NB_YM) {
        const int q = p - 1;
        const double gx = 0.5 * dy * (U3[q] + U3[p]);
        const double gy = 0.5 * dy * ((U5[q] - 1.0) + (U5[p] - 1.0));
        rx += u0[q] + gx;
        ry += u1[q] + gy;
        diag += 1.0;
    }
    const double a0 = u0[p], a1 = u1[p];
    if (diag > 0.0) {
        un[p] = a0 + P.disp_relax * (rx / diag - a0);
        un[N + p] = a1 + P.disp_relax * (ry / diag - a1);
    } else {
        un[p] = a0;
        un[N + p] = a1;
    }
}
__global__ __launch_bounds__(BLOCK) void k2_sweep(const double *__restrict__ U, const double *__restrict__ u,
                                                  double *__restrict__ un, const int32_t *__restrict__ list,
                                                  const uint8_t *__restrict__ nbm, const int32_t *__restrict__ aoff,
                                                  const int32_t *__restrict__ aids, Links2 T, double time,
                                                  Params2 P) {
    k2_sweep_body(U, u, un, list, nbm, aoff, aids, T, time, P, xcd_block() * BLOCK + threadIdx.x);
}

__device__ __forceinline__ void k2_commit_body(const double *__restrict__ un, double *__restrict__ u,
                                                   double *__restrict__ us, const int32_t *__restrict__ list,
                                                   int na, size_t N, int t) {
    if (t >= na) return;
    const int p = list[t];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const size_t k = (size_t)c * N + p;
        const double delta = un[k] - u[k];
        u[k] = un[k];
        us[k] = us[k] + delta;
    }
}
__global__ __launch_bounds__(BLOCK) void k2_commit(const double *__restrict__ un, double *__restrict__ u,
                                                   double *__restrict__ us, const int32_t *__restrict__ list,
                                                   int na, size_t N) {
    k2_commit_body(un, u, us, list, na, N, xcd_block() * BLOCK + threadIdx.x);
}

// Centred / second-order one-sided / first-order one-sided derivative of one
// displacement component plane along x (axis 0) or y (axis 1) over active
// nodes (cbw2d:3213-3272).
__device__ __forceinline__ double du2(const double *uc, int nb, int p, int axis, const Params2 &P) {
    const int s = axis == 0 ? P.ny : 1;
    const double h = axis == 0 ? P.dx : P.dy;
    const bool pp = (nb & (axis == 0 ? NB_XP : NB_YP)) != 0, pm = (nb & (axis == 0 ? NB_XM : NB_YM)) != 0;
    const bool pp2 = (nb & (axis == 0 ? NB_XP2 : NB_YP2)) != 0, pm2 = (nb & (axis == 0 ? NB_XM2 : NB_YM2)) != 0;
    if (pp && pm) return (uc[p + s] - uc[p - s]) / (2.0 * h);
    if (pp) {
        if (pp2) return (-3.0 * uc[p] + 4.0 * uc[p + s] - uc[p + 2 * s]) / (2.0 * h);
        return (uc[p + s] - uc[p]) / h;
    }
    if (pm) {
        if (pm2) return (3.0 * uc[p] - 4.0 * uc[p - s] + uc[p - 2 * s]) / (2.0 * h);
        return (uc[p] - uc[p - s]) / h;
    }
    return 0.0;
}

// Deformation-gradient update F <- F + alpha (I + grad u - F) with the
// halving J guard (cbw2d:3522-3615), folded with the BGK collision of
// cbw2d:1820-1843 into the per-node record, plus the last displacement
// commit.  ug is the displacement the repair differentiates (the sweep's
// u_next when a sweep ran, else u).
template <typename R, int LAW = -1>
__device__ __forceinline__ void k2_record_body(double *__restrict__ U, const double *__restrict__ ug,
                                                   double *__restrict__ u, double *__restrict__ us,
                                                   R *__restrict__ rec, const int32_t *__restrict__ list,
                                                   const uint8_t *__restrict__ nbm, int repair, int commit,
                                                   int32_t *__restrict__ reset_invalid, Params2 P, int t) {
    if (t >= P.na) return;
    // LAW >= 0 fixes the constitutive law at compile time (the launch checks
    // it is the solver's), so the inlined Piola maps keep only its code.
    if (LAW >= 0) P.m.model = LAW;
    const int p = list[t];
    const int ny = P.ny, nx = P.nx;
    const int i = p / ny, j = p - i * ny;
    const size_t N = P.N;
    // Every load before the first store (a load behind a store would wait
    // for the store: the vector-memory counter drains in order).
    const double v0 = U[p], v1 = U[N + p];
    double Fo[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) Fo[k] = U[(2 + k) * N + p];
    double ugc[2] = {0.0, 0.0}, uc[2] = {0.0, 0.0}, usc[2] = {0.0, 0.0};
    if (commit)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const size_t k = (size_t)c * N + p;
            ugc[c] = ug[k];
            uc[c] = u[k];
            usc[c] = us[k];
        }
    double Po[4], dP[4] = {0.0, 0.0, 0.0, 0.0}, dF[4] = {0.0, 0.0, 0.0, 0.0};
    piola2(Fo, P.m, Po);
    bool do_rep = repair != 0;
    const int nb = nbm[t];
    // interior_only: all four nearest neighbours on the lattice and active
    if (do_rep && P.interior_only && (nb & NB_NEAR) != NB_NEAR) do_rep = false;
    if (do_rep) {
        const double tt[4] = {1.0 + du2(ug, nb, p, 0, P), du2(ug, nb, p, 1, P), du2(ug + N, nb, p, 0, P),
                              1.0 + du2(ug + N, nb, p, 1, P)};
        double alpha = P.blend;
        alpha = alpha < 0.0 ? 0.0 : (alpha > 1.0 ? 1.0 : alpha);
        double Fn[4];
        for (int g = 0; g < 10; ++g) {
#pragma unroll
            for (int k = 0; k < 4; ++k) Fn[k] = Fo[k] + alpha * (tt[k] - Fo[k]);
            if (Fn[0] * Fn[3] - Fn[1] * Fn[2] > P.m.jf) break;
            alpha = 0.5 * alpha;
        }
        double Pn[4];
        piola2(Fn, P.m, Pn);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            dP[k] = Pn[k] - Po[k];
            dF[k] = Fn[k] - Fo[k];
            U[(2 + k) * N + p] = Fn[k];
        }
    }
    if (commit) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const size_t k = (size_t)c * N + p;
            const double delta = ugc[c] - uc[c];
            u[k] = ugc[c];
            us[k] = usc[c] + delta;
        }
    }
    const double w = P.omega;
    const double st = P.dt * (2.0 - w) * 0.125;
    const double w4 = 0.25 * w, w2c = w / (2.0 * P.c), h2c = 0.5 / P.c;
    double rv[12];
    rv[R_A + 0] = w4 * v0 + st * source2(0, P.src_mode, P.src, v0, v1);
    rv[R_A + 1] = w4 * v1 + st * source2(1, P.src_mode, P.src, v0, v1);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        rv[R_A + 2 + k] = w4 * Fo[k] + st * source2(2 + k, P.src_mode, P.src, v0, v1) + 0.25 * dF[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) rv[R_BV + k] = -w2c * Po[k] - h2c * dP[k];
    rv[R_BF + 0] = -w2c * v0;
    rv[R_BF + 1] = -w2c * v1;
#pragma unroll
    for (int k = 0; k < 12; k += 2) rst2(rec, k, p, N, k < R_BV ? k : -1, k + 1 < R_BV ? k + 1 : -1, w, rv[k], rv[k + 1]);
    // Clears the invalid flag ahead of the stream kernel's fused collide
    // (the link rows never set it).
    if (reset_invalid && t == 0) reset_invalid[0] = 0;
}
template <typename R, int LAW>
__global__ __launch_bounds__(BLOCK) void k2_record(double *__restrict__ U, const double *__restrict__ ug,
                                                   double *__restrict__ u, double *__restrict__ us,
                                                   R *__restrict__ rec, const int32_t *__restrict__ list,
                                                   const uint8_t *__restrict__ nbm, int repair, int commit,
                                                   int32_t *__restrict__ reset_invalid, Params2 P) {
    k2_record_body<R, LAW>(U, ug, u, us, rec, list, nbm, repair, commit, reset_invalid, P, xcd_block() * BLOCK + threadIdx.x);
}

// Quadratic one-sided reconstruction of U[comp] along the inward link
// coordinate (cbw2d:1443-1480).
__device__ inline double interp_link(const double *U, const int32_t *act, size_t N, int comp, int p, int i, int j,
                                     int qo, double sv, const Params2 &P) {
    const double *Uc = U + comp * N;
    const double v0 = Uc[p];
    const int i1 = i - cx(qo), j1 = j - cy(qo);
    if (i1 >= 0 && i1 < P.nx && j1 >= 0 && j1 < P.ny && act[i1 * P.ny + j1] != 0) {
        const double v1 = Uc[i1 * P.ny + j1];
        const int i2 = i - 2 * cx(qo), j2 = j - 2 * cy(qo);
        if (i2 >= 0 && i2 < P.nx && j2 >= 0 && j2 < P.ny && act[i2 * P.ny + j2] != 0) {
            const double v2 = Uc[i2 * P.ny + j2];
            const double l0 = 0.5 * (sv + 1.0) * (sv + 2.0);
            const double l1 = -sv * (sv + 2.0);
            const double l2 = 0.5 * sv * (sv + 1.0);
            return l0 * v0 + l1 * v1 + l2 * v2;
        }
        return (sv + 1.0) * v0 - sv * v1;
    }
    return v0;
}

// Normal-probe value and LSQ gradients of U at every cut point
// (cbw2d:1202-1221 as used at :2141-2160 and :2491-2503): 32 lanes per
// link, lane m loads stencil sample m; the sums then run over m in stencil
// order, one thread per (link, output).  Output cs (19, L): rows 0-3 the
// value of F at the cut point, rows 4-9 d/dx and 10-15 d/dy of U, and with
// `tangent` rows 16-17 the tangential derivative of u and row 18 the sum of
// |w_t| (compat_bfl, cbw2d:2221-2230).
constexpr int GL = 8;  // links per block of 256 threads
__device__ __forceinline__ void k2_gather_body(const double *__restrict__ U, const double *__restrict__ u,
                                                 const int32_t *__restrict__ lsq_p, const double *__restrict__ wb,
                                                 const double *__restrict__ wx, const double *__restrict__ wy,
                                                 const double *__restrict__ wt, int tangent, double *__restrict__ cs,
                                                 int L, size_t N, int blk) {
    __shared__ double sU[8][GL * 32];
    __shared__ double sW[4][GL * 32];
    const int t = threadIdx.x;
    const int ll = t >> 5, m = t & 31;
    const int ell = blk * GL + ll;
    double u8[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, w4[4] = {0.0, 0.0, 0.0, 0.0};
    if (ell < L && m < MAX_LSQ) {
        const int pp = lsq_p[(size_t)m * L + ell];
        if (pp >= 0) {
#pragma unroll
            for (int a = 0; a < 6; ++a) u8[a] = U[a * N + pp];
            w4[0] = wb[(size_t)m * L + ell];
            w4[1] = wx[(size_t)m * L + ell];
            w4[2] = wy[(size_t)m * L + ell];
            if (tangent) {
                u8[6] = u[pp];
                u8[7] = u[N + pp];
                w4[3] = wt[(size_t)m * L + ell];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) sU[a][t] = u8[a];
#pragma unroll
    for (int k = 0; k < 4; ++k) sW[k][t] = w4[k];
    __syncthreads();
    const int nout = tangent ? 19 : 16;
    if (t >= GL * nout) return;
    const int l2 = t / nout, v = t - l2 * nout;
    const int e2 = blk * GL + l2;
    if (e2 >= L) return;
    double acc = 0.0;
    if (v < 16) {
        const int set = v < 4 ? 0 : (v < 10 ? 1 : 2);
        const int comp = v < 4 ? 2 + v : (v < 10 ? v - 4 : v - 10);
        for (int k = 0; k < MAX_LSQ; ++k) acc += sW[set][l2 * 32 + k] * sU[comp][l2 * 32 + k];
    } else if (v < 18) {
        for (int k = 0; k < MAX_LSQ; ++k) acc += sW[3][l2 * 32 + k] * sU[6 + (v - 16)][l2 * 32 + k];
    } else {
        for (int k = 0; k < MAX_LSQ; ++k) acc += fabs(sW[3][l2 * 32 + k]);
    }
    cs[(size_t)v * L + e2] = acc;
}
__global__ __launch_bounds__(256) void k2_gather(const double *__restrict__ U, const double *__restrict__ u,
                                                 const int32_t *__restrict__ lsq_p, const double *__restrict__ wb,
                                                 const double *__restrict__ wx, const double *__restrict__ wy,
                                                 const double *__restrict__ wt, int tangent, double *__restrict__ cs,
                                                 int L, size_t N) {
    k2_gather_body(U, u, lsq_p, wb, wx, wy, wt, tangent, cs, L, N, blockIdx.x);
}

// Incoming populations of every cut link (cbw2d:2064-3098).  use_rec
// selects post-collision sources (time step: pre-collision f plus this
// step's records and, for mrt_ghost, the ghost term of gh) or populations
// read as they are (standalone reconstruction).  EXT instantiates the
// regularized, NEE, quadratic and polynomial variants (recon_ext), so the
// kernel of the other reconstructions carries none of their code.
template <typename T, bool EXT, int LAW = -1>
__device__ __forceinline__ void k2_links_body(const T *__restrict__ fsrc, T *__restrict__ fdst,
                                                   const T *__restrict__ rec, const double *__restrict__ U,
                                                   const double *__restrict__ cs, Links2 Tb, int use_rec,
                                                   const double *__restrict__ gh, double time, Params2 P, int ell) {
    if (ell >= P.L) return;
    if (LAW >= 0) P.m.model = LAW;
    const int r = P.recon;
    const size_t N = P.N;
    const int L = P.L;
    const int p = Tb.lp[ell], qo = Tb.lq[ell], qi = (qo + 2) & 3;
    const int ny = P.ny;
    const int i = p / ny, j = p - i * ny;
    int bid = Tb.lbid[ell];
    if (bid < 0 || bid >= MAX_IDS) bid = 0;
    const int kind = Tb.kind_by_id[bid], mode = Tb.mode_by_id[bid];
    const double *prm = Tb.prm_by_id + bid * 8;
    const double eta = Tb.eta[ell], xb = Tb.xb[ell], yb = Tb.yb[ell], nxb = Tb.nx[ell], nyb = Tb.ny[ell];
    const double dxi = (double)cx(qi), dyi = (double)cy(qi);
    time = step_time(time, P);
    double bv0 = value2(0, mode, prm, xb, yb, nxb, nyb, time);
    double bv1 = value2(1, mode, prm, xb, yb, nxb, nyb, time);
    if (kind == BC_NEU && P.use_dyn) {
        bv0 = Tb.dyn_tx[ell];
        bv1 = Tb.dyn_ty[ell];
    }
    const int iu = i - cx(qo), ju = j - cy(qo);
    const bool up_ok = iu >= 0 && iu < P.nx && ju >= 0 && ju < ny && Tb.act[iu * ny + ju] != 0;
    const int pu = up_ok ? iu * ny + ju : p;
    const bool valid = Tb.valid[ell] != 0;
    const double om1 = 1.0 - P.omega;
    // post-collision population q, component a of node nd
    auto fps = [&](int q, int a, int nd) -> double {
        const size_t k = pofs2(fsrc, q * 6 + a, nd, N);
        if (!use_rec) return ld(fsrc, k, a);
        double v = fpost_v(ld(fsrc, k, a), rec, N, q, a, nd, om1, P.omega);
        if (gh) v += ghost_term(gh, N, q, a, nd, P.gw);
        return v;
    };
    // Boundary F predictor (cbw2d:2162-2250): one-link extrapolation for the
    // bfl family, else the quadratic one-sided interpolation at the cut
    // point; the LSQ value for lsq_bfl, local_f and local_char2.
    double Fg[4];
    if (r == R_BFL || r == R_REG_BFL || r == R_LSQ_BFL || r == R_LOCALF_BFL) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double uc = U[(2 + k) * N + p];
            Fg[k] = up_ok ? uc + eta * (uc - U[(2 + k) * N + pu]) : uc;
        }
        if (r == R_LSQ_BFL && valid)
#pragma unroll
            for (int k = 0; k < 4; ++k) Fg[k] = cs[(size_t)k * L + ell];
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            Fg[k] = ((r == R_LOCALF || r == R_CHAR2) && valid) ? cs[(size_t)k * L + ell]
                                                              : interp_link(U, Tb.act, N, 2 + k, p, i, j, qo, eta, P);
    }
    {
        // Geometry-compatible predictor (cbw2d:2206-2250): F n from the link
        // predictor, F t = t + d_t u (traction links: the LSQ tangential
        // derivative of u; Dirichlet links: the time integral of an affine
        // velocity's tangential derivative, zero for other modes).
        if (recon_compat(r) && valid) {
            const double tanx = -nyb, tany = nxb;
            double dut0 = 0.0, dut1 = 0.0, wt_abs = 1.0;
            if (kind == BC_NEU) {
                dut0 = cs[(size_t)16 * L + ell];
                dut1 = cs[(size_t)17 * L + ell];
                wt_abs = cs[(size_t)18 * L + ell];
            } else if (mode == V_AFFINE_VEL) {
                double at = time;
                if (prm[6] > 0.0 && at > prm[6]) at = prm[6];
                dut0 = at * (prm[0] * tanx + prm[1] * tany);
                dut1 = at * (prm[3] * tanx + prm[4] * tany);
            }
            if (wt_abs > 0.0) {
                const double hn0 = Fg[0] * nxb + Fg[1] * nyb, hn1 = Fg[2] * nxb + Fg[3] * nyb;
                const double gt0 = tanx + dut0, gt1 = tany + dut1;
                Fg[0] = hn0 * nxb + gt0 * tanx;
                Fg[1] = hn0 * nyb + gt0 * tany;
                Fg[2] = hn1 * nxb + gt1 * tanx;
                Fg[3] = hn1 * nyb + gt1 * tany;
            }
        }
    }
    double Fb[4] = {Fg[0], Fg[1], Fg[2], Fg[3]};
    if (kind == BC_NEU) solve_normal2(Fg, nxb, nyb, bv0, bv1, P.solver, P.m, Fb);

    double node[6], Pn[4];
    if constexpr (EXT) {
#pragma unroll
        for (int k = 0; k < 6; ++k) node[k] = U[(size_t)k * N + p];
        piola2(node + 2, P.m, Pn);
    }
    // fpost - feq(st) of node nd with the D2Q4 ghost moment removed
    // (cbw2d:1517-1643).
    auto hreg = [&](int q, int a, int nd, const double *stt, const double *Ps) -> double {
        double h[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) h[k] = fps(k, a, nd) - feq2(k, a, stt, Ps, P.c);
        const double corr = 0.25 * (h[0] - h[1] + h[2] - h[3]);
        return (q == 0 || q == 2) ? h[q] - corr : h[q] + corr;
    };
    if constexpr (EXT) {
        if (r == R_RNEE) {
            // cbw2d:2276-2356: ghost state mirrored through the boundary datum.
            const double sm = 2.0 * eta - 1.0;
            double m[6], g[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) m[k] = interp_link(U, Tb.act, N, k, p, i, j, qo, sm, P);
            if (kind == BC_DIR) {
                g[0] = 2.0 * bv0 - m[0];
                g[1] = 2.0 * bv1 - m[1];
            } else {
                g[0] = 2.0 * interp_link(U, Tb.act, N, 0, p, i, j, qo, eta, P) - m[0];
                g[1] = 2.0 * interp_link(U, Tb.act, N, 1, p, i, j, qo, eta, P) - m[1];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) g[2 + k] = 2.0 * Fb[k] - m[2 + k];
            if (g[2] * g[5] - g[3] * g[4] <= P.m.jf)
#pragma unroll
                for (int k = 0; k < 4; ++k) g[2 + k] = Fb[k];
            double Pg[4];
            piola2(g + 2, P.m, Pg);
            for (int a = 0; a < 6; ++a)
                st(fdst, pofs2(fdst, qi * 6 + a, p, N), a, feq2(qi, a, g, Pg, P.c) + hreg(qi, a, p, node, Pn));
            return;
        }
        if (r == R_PAIR_NEE || r == R_SAME_NEE || r == R_COMPAT_NEE) {
            // cbw2d:2358-2454: boundary equilibrium + regularized neq of the
            // node, the outgoing direction's with the half-way parity (pair,
            // compat) or the incoming direction's (same).
            double b[6] = {bv0, bv1, Fb[0], Fb[1], Fb[2], Fb[3]};
            if (kind == BC_NEU) {
                b[0] = interp_link(U, Tb.act, N, 0, p, i, j, qo, eta, P);
                b[1] = interp_link(U, Tb.act, N, 1, p, i, j, qo, eta, P);
            }
            double Pbb[4];
            piola2(b + 2, P.m, Pbb);
            for (int a = 0; a < 6; ++a) {
                const double Dp = (kind == BC_DIR) ? (a < 2 ? -1.0 : 1.0) : (a >= 2 ? -1.0 : 1.0);
                const double h = r == R_SAME_NEE ? hreg(qi, a, p, node, Pn) : Dp * hreg(qo, a, p, node, Pn);
                st(fdst, pofs2(fdst, qi * 6 + a, p, N), a, feq2(qi, a, b, Pbb, P.c) + h);
            }
            return;
        }
    }
    double Pb[4];
    piola2(Fb, P.m, Pb);

    // Characteristic closure of the traction rows (cbw2d:2456-2886, 3034-3097).
    const bool use_char = !EXT && (r == R_LOCALF || r == R_CHAR2) && kind == BC_NEU && valid;
    bool char_ok = false;
    double cv[6], Acw[6], Pc[4];
    if (use_char) {
        cv[0] = U[p];
        cv[1] = U[N + p];
#pragma unroll
        for (int k = 0; k < 4; ++k) cv[2 + k] = Fb[k];
        double gx[6], gy[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            gx[a] = cs[(size_t)(4 + a) * L + ell];
            gy[a] = cs[(size_t)(10 + a) * L + ell];
        }
        const double shx = -eta * P.dt * P.c * (double)cx(qo);
        const double shy = -eta * P.dt * P.c * (double)cy(qo);
        double d[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) d[a] = shx * gx[a] + shy * gy[a];
        double sc = 1.0;
        for (int g = 0; g < 10; ++g) {
            const double t11 = cv[2] + sc * d[2], t12 = cv[3] + sc * d[3];
            const double t21 = cv[4] + sc * d[4], t22 = cv[5] + sc * d[5];
            if (t11 * t22 - t12 * t21 > 10.0 * P.m.jf) break;
            sc = 0.5 * sc;
        }
#pragma unroll
        for (int a = 0; a < 6; ++a) cv[a] = cv[a] + sc * d[a];
        double jx[6], jy[6], cw[6];
        flux_jvp2(0, cv, gx, P.m, jx);
        flux_jvp2(1, cv, gy, P.m, jy);
#pragma unroll
        for (int a = 0; a < 6; ++a)
            cw[a] = source2(a, P.src_mode, P.src, cv[0], cv[1]) - jx[a] - jy[a] + P.c * (dxi * gx[a] + dyi * gy[a]);
        if (cv[2] * cv[5] - cv[3] * cv[4] > P.m.jf) {
            char_ok = true;
            piola2(cv + 2, P.m, Pc);
            double ax[6], ay[6];
            flux_jvp2(0, cv, cw, P.m, ax);
            flux_jvp2(1, cv, cw, P.m, ay);
#pragma unroll
            for (int a = 0; a < 6; ++a) Acw[a] = 0.25 * (cw[a] + (2.0 / P.c) * (dxi * ax[a] + dyi * ay[a]));
        }
    }
    // regularized_bfl / lowq_reg_bfl sources (traction links)
    const bool use_reg = EXT && kind == BC_NEU && (r == R_REG_BFL || (r == R_LOWQ_REG_BFL && eta < 0.5));
    double nodeu[6], Pu[4];
    if (EXT && use_reg && eta < 0.5 && up_ok) {
#pragma unroll
        for (int k = 0; k < 6; ++k) nodeu[k] = U[(size_t)k * N + pu];
        piola2(nodeu + 2, P.m, Pu);
    }
    // quadratic one-sided interpolation of fpost[qo][a] along the link
    // (cbw2d:1484-1513)
    auto interp_fp = [&](int a, double sv) -> double {
        const double v0 = fps(qo, a, p);
        if (!up_ok) return v0;
        const double v1 = fps(qo, a, pu);
        const int i2 = i - 2 * cx(qo), j2 = j - 2 * cy(qo);
        if (i2 >= 0 && i2 < P.nx && j2 >= 0 && j2 < ny && Tb.act[i2 * ny + j2] != 0) {
            const double v2 = fps(qo, a, i2 * ny + j2);
            const double l0 = 0.5 * (sv + 1.0) * (sv + 2.0);
            const double l1 = -sv * (sv + 2.0);
            const double l2 = 0.5 * sv * (sv + 1.0);
            return l0 * v0 + l1 * v1 + l2 * v2;
        }
        return (sv + 1.0) * v0 - sv * v1;
    };
    // All six values first, then the stores: a load behind a store would wait
    // for the store (the vector-memory counter drains in order).
    double vals[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        double val;
        if (char_ok) {
            const double geq = feq2(qi, a, cv, Pc, P.c);
            const double src = source2(a, P.src_mode, P.src, cv[0], cv[1]);
            val = geq - P.dt / P.omega * Acw[a] + P.dt * (1.0 / P.omega - 0.5) * 0.25 * src;
        } else {
            double D, S;
            if (kind == BC_DIR) {
                if (a < 2) {
                    D = -1.0;
                    S = 0.5 * (a == 0 ? bv0 : bv1);
                } else {
                    const int k = a - 2;
                    D = 1.0;
                    S = -((k & 1) == 0 ? dxi : dyi) * (k < 2 ? bv0 : bv1) / P.c;
                }
            } else {
                if (a < 2) {
                    D = 1.0;
                    S = -(dxi * Pb[2 * a] + dyi * Pb[2 * a + 1]) / P.c;
                } else {
                    D = -1.0;
                    S = 0.5 * Fb[a - 2];
                }
            }
            double fo, fi;
            if (EXT && use_reg) {
                fo = feq2(qo, a, node, Pn, P.c) + hreg(qo, a, p, node, Pn);
                fi = feq2(qi, a, node, Pn, P.c) + hreg(qi, a, p, node, Pn);
            } else {
                fo = fps(qo, a, p);
                fi = fps(qi, a, p);
            }
            if (EXT && r == R_POLY && up_ok && eta > 0.25 && eta < 0.95) {
                // cbw2d:2977-2993: incoming polynomial through s = -1, 0, eta at s = 1
                const double gin_b = D * interp_fp(a, eta) + S;
                const double g1 = fps(qi, a, pu);
                const double wb = 2.0 / (eta * (eta + 1.0));
                const double w0 = -2.0 * (1.0 - eta) / eta;
                const double w1 = (1.0 - eta) / (1.0 + eta);
                val = wb * gin_b + w0 * fi + w1 * g1;
            } else if (EXT && r == R_QBFL && eta < 0.5) {
                // cbw2d:2994-3000: quadratic interpolation at the BFL point s = 2 eta - 1
                val = D * interp_fp(a, 2.0 * eta - 1.0) + S;
            } else if (eta >= 0.5) {
                const double alpha = 1.0 / (2.0 * eta);
                val = alpha * (D * fo + S) + (1.0 - alpha) * fi;
            } else {
                double upv = fo;
                if (up_ok) upv = (EXT && use_reg) ? feq2(qo, a, nodeu, Pu, P.c) + hreg(qo, a, pu, nodeu, Pu)
                                                  : fps(qo, a, pu);
                val = D * (2.0 * eta * fo + (1.0 - 2.0 * eta) * upv) + S;
            }
        }
        vals[a] = val;
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) st(fdst, pofs2(fdst, qi * 6 + a, p, N), a, vals[a]);
}
template <typename T, bool EXT, int LAW = -1>
__global__ __launch_bounds__(LBLOCK) void k2_links(const T *__restrict__ fsrc, T *__restrict__ fdst,
                                                   const T *__restrict__ rec, const double *__restrict__ U,
                                                   const double *__restrict__ cs, Links2 Tb, int use_rec,
                                                   const double *__restrict__ gh, double time, Params2 P) {
    k2_links_body<T, EXT, LAW>(fsrc, fdst, rec, U, cs, Tb, use_rec, gh, time, P, blockIdx.x * LBLOCK + threadIdx.x);
}

// Index along a periodic axis of n nodes (one step off either end wraps).
__device__ __forceinline__ int wrap_i(int i, int n, int periodic) {
    if (!periodic) return i;
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// Faces of the rectangular box solver: kind (0 periodic, 1 Dirichlet
// velocity, 2 Neumann traction), value mode and 8 params per side (left,
// right, bottom, top), and the box lengths.
struct Box2 {
    int kind[4], mode[4];
    double prm[4][8];
    double lx, ly;
};

// Half-way face reconstruction of the rectangular box solver
// (vnew2d:1685-1820): one thread per (side, face node); Dirichlet faces
// bounce back with the velocity datum, Neumann faces impose the traction on
// the velocity rows and the face F of the column solve on the F rows.  The
// four sides read and write disjoint populations, so f_src may equal f_dst
// (the in-place initial reconstruction).  use_rec as in k2_links.
template <typename T>
__global__ __launch_bounds__(LBLOCK) void k2_box_sides(const T *fsrc, T *fdst, const T *__restrict__ rec,
                                                       const double *__restrict__ U, int use_rec, Box2 B,
                                                       double time, Params2 P) {
    const int m = P.nx > P.ny ? P.nx : P.ny;
    const int t = blockIdx.x * LBLOCK + threadIdx.x;
    if (t >= 4 * m) return;
    const int side = t / m, sidx = t - side * m;
    const int kind = B.kind[side];
    if (kind == 0) return;
    const int nn = side < 2 ? P.ny : P.nx;
    if (sidx >= nn) return;
    int i = 0, j = sidx, nxn = -1, nyn = 0, qi = 0, qo = 2, axis = 0;
    double sign = -1.0;
    if (side == 1) {
        i = P.nx - 1;
        nxn = 1;
        qi = 2;
        qo = 0;
        sign = 1.0;
    } else if (side == 2) {
        i = sidx;
        j = 0;
        nxn = 0;
        nyn = -1;
        qi = 1;
        qo = 3;
        axis = 1;
    } else if (side == 3) {
        i = sidx;
        j = P.ny - 1;
        nxn = 0;
        nyn = 1;
        qi = 3;
        qo = 1;
        axis = 1;
        sign = 1.0;
    }
    const size_t N = P.N;
    const int p = i * P.ny + j;
    double xb = (i + 0.5) * P.dx, yb = (j + 0.5) * P.dy;
    if (side == 0) xb = 0.0;
    else if (side == 1) xb = B.lx;
    else if (side == 2) yb = 0.0;
    else yb = B.ly;
    time = step_time(time, P);
    const double *prm = B.prm[side];
    const double v0 = value2(0, B.mode[side], prm, xb, yb, 0.0, 0.0, time);
    const double v1 = value2(1, B.mode[side], prm, xb, yb, 0.0, 0.0, time);
    const double om1 = 1.0 - P.omega;
    double fo[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const size_t k = pofs2(fsrc, qo * 6 + a, p, N);
        fo[a] = use_rec ? fpost_v(ld(fsrc, k, a), rec, N, qo, a, p, om1, P.omega) : ld(fsrc, k, a);
    }
    double val[6];
    if (kind == BC_DIR) {
        val[0] = -fo[0] + 0.5 * v0;
        val[1] = -fo[1] + 0.5 * v1;
        val[2] = fo[2] + (double)nxn * v0 / P.c;
        val[3] = fo[3] + (double)nyn * v0 / P.c;
        val[4] = fo[4] + (double)nxn * v1 / P.c;
        val[5] = fo[5] + (double)nyn * v1 / P.c;
    } else {
        double F[4] = {U[2 * N + p], U[3 * N + p], U[4 * N + p], U[5 * N + p]};
        solve_normal_column2(F, axis, sign, v0, v1, P.m);
        val[0] = fo[0] + v0 / P.c;
        val[1] = fo[1] + v1 / P.c;
#pragma unroll
        for (int k = 0; k < 4; ++k) val[2 + k] = -fo[2 + k] + 0.5 * F[k];
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) st(fdst, pofs2(fdst, qi * 6 + a, p, N), a, val[a]);
}

// Bit q of upm[t]: the upstream node of population q at the active node
// list[t] lies on the lattice (after the periodic wrap) and is active.  It
// depends on the geometry and the periodic axes only, so it is computed once
// and after the periodicity changes, and the stream kernel loads it beside
// the node index instead of four dependent mask loads per node.
__global__ __launch_bounds__(BLOCK) void k2_upmask(const int32_t *__restrict__ list, const int32_t *__restrict__ act,
                                                   uint8_t *__restrict__ upm, Params2 P) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= P.na) return;
    const int p = list[t];
    const int ny = P.ny;
    const int i = p / ny, j = p - i * ny;
    int bits = 0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int iu = wrap_i(i - cx(q), P.nx, P.per_x), ju = wrap_i(j - cy(q), ny, P.per_y);
        if (iu >= 0 && iu < P.nx && ju >= 0 && ju < ny && act[iu * ny + ju] != 0) bits |= 1 << q;
    }
    upm[t] = (uint8_t)bits;
}

// Pull stream of the post-collision populations; entries whose upstream
// node is inactive or off the lattice belong to cut links and were written
// by k2_links.  With NEXT the kernel also performs the moment part of the
// next step's collide (cbw2d:1731-1790).
//
// MODE bit 0 (mrt_ghost): the sources carry the ghost term of gin, and with
// NEXT the ghost moment of the new populations is written to gout.  MODE
// bit 1: the kinetic ghost filter of cbw2d:3100-3124 (m3 scaled by
// 1 - strength through +-m3 strength / 4 on E, W and N, S) is applied to all
// four populations of the node, cut-link ones included, before they are
// stored and summed.
// Directions per load group of the 2D pull: all four measured fastest
// (1024^2 superellipse FP64 step 114.3 -> 111.0 us, 2 directions 111.8).
constexpr int STREAM_GROUP2 = 4;
template <typename T, bool NEXT, int MODE>
__device__ __forceinline__ void k2_stream_body(const T *__restrict__ f0, T *__restrict__ f1,
                                                   const T *__restrict__ rec, const int32_t *__restrict__ list,
                                                   const uint8_t *__restrict__ upm, double *__restrict__ U,
                                                   double *__restrict__ u, double *__restrict__ us,
                                                   int32_t *__restrict__ invalid, const double *__restrict__ gin,
                                                   double *__restrict__ gout, Params2 P, int t) {
    constexpr bool MRT = (MODE & 1) != 0, FILT = (MODE & 2) != 0;
    if (t >= P.na) return;
    const int p = list[t];
    const int up = upm[t];
    const int ny = P.ny;
    const int i = p / ny, j = p - i * ny;
    const size_t N = P.N;
    const double om1 = 1.0 - P.omega;
    double Um[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) Um[a] = 0.0;
    double uso[2] = {0.0, 0.0};
    if constexpr (MODE == 0) {
        // Every direction's upstream node first, then the directions in
        // groups whose loads all precede the group's stores (the vector-memory
        // counter drains in order over loads and stores alike); a missing
        // upstream node reads node p instead, unused.
        int srcs[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int iu = wrap_i(i - cx(q), P.nx, P.per_x), ju = wrap_i(j - cy(q), ny, P.per_y);
            srcs[q] = ((up >> q) & 1) ? iu * ny + ju : -1;
        }
        if (NEXT)
#pragma unroll
            for (int c = 0; c < 2; ++c) uso[c] = us[c * N + p];
#pragma unroll
        for (int q0 = 0; q0 < NQ; q0 += STREAM_GROUP2) {
            double fv[STREAM_GROUP2][NC], A[STREAM_GROUP2][NC], B[STREAM_GROUP2][4];
#pragma unroll
            for (int g = 0; g < STREAM_GROUP2; ++g) {
                const int q = q0 + g, axis = q & 1;
                const int src = srcs[q] >= 0 ? srcs[q] : p;
                // components in pairs (a, a + 1): one access each
#pragma unroll
                for (int h = 0; h < NC / 2; ++h)
                    ld2_nt(f0, q * 6 + 2 * h, src, N, 2 * h, fv[g][2 * h], fv[g][2 * h + 1]);
#pragma unroll
                for (int h = 0; h < NC / 2; ++h) rld2(rec, R_A + 2 * h, src, N, P.omega, A[g][2 * h], A[g][2 * h + 1]);
#pragma unroll
                for (int r = 0; r < 2; ++r) B[g][r] = rld_b(rec, pofs2(rec, R_BV + 2 * r + axis, src, N));
                rldb2(rec, R_BF, src, N, B[g][2], B[g][3]);
            }
#pragma unroll
            for (int g = 0; g < STREAM_GROUP2; ++g) {
                const int q = q0 + g;
                if (srcs[q] >= 0) {
#pragma unroll
                    for (int h = 0; h < NC / 2; ++h) {
                        double x0 = fpost_b(fv[g][2 * h], A[g][2 * h], B[g], q, 2 * h, om1);
                        double x1 = fpost_b(fv[g][2 * h + 1], A[g][2 * h + 1], B[g], q, 2 * h + 1, om1);
                        st2_nt(f1, q * 6 + 2 * h, p, N, 2 * h, x0, x1);
                        if (NEXT) {
                            Um[2 * h] += x0;
                            Um[2 * h + 1] += x1;
                        }
                    }
                } else if (NEXT) {
#pragma unroll
                    for (int a = 0; a < NC; ++a) Um[a] += ld(f1, pofs2(f1, q * 6 + a, p, N), a);
                }
            }
        }
    } else {
        double fq[NQ][NC];
        bool okq[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int iu = wrap_i(i - cx(q), P.nx, P.per_x), ju = wrap_i(j - cy(q), ny, P.per_y);
            okq[q] = ((up >> q) & 1) != 0;
            if (okq[q]) {
                const int src = iu * ny + ju;
#pragma unroll
                for (int a = 0; a < NC; ++a) {
                    double x = fpost_v(ld_nt(f0, pofs2(f0, q * 6 + a, src, N), a), rec, N, q, a, src, om1, P.omega);
                    if (MRT) x += ghost_term(gin, N, q, a, src, P.gw);
                    fq[q][a] = x;
                }
            } else {
#pragma unroll
                for (int a = 0; a < NC; ++a) fq[q][a] = ld(f1, pofs2(f1, q * 6 + a, p, N), a);
            }
        }
        if (FILT) {
#pragma unroll
            for (int a = 0; a < NC; ++a) {
                const double m3 = fq[0][a] - fq[1][a] + fq[2][a] - fq[3][a];
                const double quarter = 0.25 * (-P.filt * m3);
                fq[0][a] = fq[0][a] + quarter;
                fq[2][a] = fq[2][a] + quarter;
                fq[1][a] = fq[1][a] - quarter;
                fq[3][a] = fq[3][a] - quarter;
            }
        }
        double Gm[NC];
#pragma unroll
        for (int a = 0; a < NC; ++a) Gm[a] = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int a = 0; a < NC; ++a) {
                const double x = (FILT || okq[q]) ? st_nt(f1, pofs2(f1, q * 6 + a, p, N), a, fq[q][a]) : fq[q][a];
                Um[a] += x;
                Gm[a] += (q & 1) ? -x : x;
            }
        if (MRT && NEXT)
#pragma unroll
            for (int a = 0; a < NC; ++a) gout[(size_t)a * N + p] = Gm[a];
    }
    if (NEXT) {
        const double hdt = 0.5 * P.dt;
        if (P.src_mode == SRC_CONST) {
#pragma unroll
            for (int a = 0; a < NC; ++a) Um[a] += hdt * P.src[a];
        } else if (P.src_mode == SRC_DAMP) {
            const double den = 1.0 + hdt * P.src[0];
            Um[0] = Um[0] / den;
            Um[1] = Um[1] / den;
        }
#pragma unroll
        for (int a = 0; a < NC; ++a) U[a * N + p] = Um[a];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const double uc = (MODE == 0 ? uso[c] : us[c * N + p]) + hdt * Um[c];
            u[c * N + p] = uc;
            us[c * N + p] = uc + hdt * Um[c];
        }
        if (Um[2] * Um[5] - Um[3] * Um[4] <= P.m.jf) atomicOr(invalid, 1);
    }
}
template <typename T, bool NEXT, int MODE>
__global__ __launch_bounds__(BLOCK) void k2_stream(const T *__restrict__ f0, T *__restrict__ f1,
                                                   const T *__restrict__ rec, const int32_t *__restrict__ list,
                                                   const uint8_t *__restrict__ upm, double *__restrict__ U,
                                                   double *__restrict__ u, double *__restrict__ us,
                                                   int32_t *__restrict__ invalid, const double *__restrict__ gin,
                                                   double *__restrict__ gout, Params2 P) {
    k2_stream_body<T, NEXT, MODE>(f0, f1, rec, list, upm, U, u, us, invalid, gin, gout, P, xcd_block() * BLOCK + threadIdx.x);
}

// f = feq(U) - dt/(2Q) B(U) at active nodes (cbw2d:3873-3876).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k2_init_pop(const double *__restrict__ U, T *__restrict__ f,
                                                     const int32_t *__restrict__ list, Params2 P) {
    const int t = xcd_block() * BLOCK + threadIdx.x;
    if (t >= P.na) return;
    const int p = list[t];
    const size_t N = P.N;
    double Uv[NC], Pf[4];
#pragma unroll
    for (int a = 0; a < NC; ++a) Uv[a] = U[a * N + p];
    piola2(Uv + 2, P.m, Pf);
    const double cs = 0.5 * P.dt / 4.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int a = 0; a < NC; ++a)
            st(f, pofs2(f, q * 6 + a, p, N), a, feq2(q, a, Uv, Pf, P.c) - cs * source2(a, P.src_mode, P.src, Uv[0], Uv[1]));
}

__global__ __launch_bounds__(BLOCK) void k2_set_ustar(const double *__restrict__ disp, const double *__restrict__ U,
                                                      double *__restrict__ us, const int32_t *__restrict__ list,
                                                      double dt, int na, size_t N) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= na) return;
    const int p = list[t];
#pragma unroll
    for (int c = 0; c < 2; ++c) us[c * N + p] = disp[c * N + p] - 0.5 * dt * U[c * N + p];
}

__global__ __launch_bounds__(BLOCK) void k2_stress(const double *__restrict__ U, double *__restrict__ Pout,
                                                   double *__restrict__ sig, const int32_t *__restrict__ list,
                                                   Params2 P) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= P.na) return;
    const int p = list[t];
    const size_t N = P.N;
    double F[4], Pf[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) F[k] = U[(2 + k) * N + p];
    piola2(F, P.m, Pf);
    const double Js = clamp_j(F[0] * F[3] - F[1] * F[2], P.m.jf);
#pragma unroll
    for (int k = 0; k < 4; ++k) Pout[k * N + p] = Pf[k];
    sig[0 * N + p] = (Pf[0] * F[0] + Pf[1] * F[1]) / Js;
    sig[1 * N + p] = (Pf[0] * F[2] + Pf[1] * F[3]) / Js;
    sig[2 * N + p] = (Pf[2] * F[0] + Pf[3] * F[1]) / Js;
    sig[3 * N + p] = (Pf[2] * F[2] + Pf[3] * F[3]) / Js;
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k2_post(const T *__restrict__ f, T *__restrict__ fp,
                                                 const T *__restrict__ rec, const int32_t *__restrict__ list,
                                                 Params2 P) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= P.na) return;
    const int p = list[t];
    const double om1 = 1.0 - P.omega;
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        double fv[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) fv[q] = ld(f, pofs2(f, q * 6 + a, p, P.N), a);
        const double g = P.gw * (fv[0] - fv[1] + fv[2] - fv[3]);  // mrt_ghost term, 0 for bgk
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            st(fp, pofs2(fp, q * 6 + a, p, P.N), a,
               fpost_v(fv[q], rec, P.N, q, a, p, om1, P.omega) + ((q & 1) ? -g : g));
    }
}

// Kinetic ghost filter in place (cbw2d:3100-3124).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k2_filter(T *__restrict__ f, const int32_t *__restrict__ list, Params2 P) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= P.na) return;
    const int p = list[t];
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        double fv[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) fv[q] = ld(f, pofs2(f, q * 6 + a, p, P.N), a);
        const double quarter = 0.25 * (-P.filt * (fv[0] - fv[1] + fv[2] - fv[3]));
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            st(f, pofs2(f, q * 6 + a, p, P.N), a, (q & 1) ? fv[q] - quarter : fv[q] + quarter);
    }
}

template <typename T>
__global__ void k2_copy(const T *__restrict__ s, T *__restrict__ d, size_t count) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < count) d[i] = s[i];
}
// Conversions between the (24, nx, ny) double planes and the stored pairs.
template <typename T>
__global__ void k2_to_f64(const T *__restrict__ f, double *__restrict__ d, size_t N) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= (size_t)FQ * N) return;
    const int plane = (int)(i / N);
    d[i] = ld(f, pofs2(f, plane, i - (size_t)plane * N, N), plane % 6);
}
template <typename T>
__global__ void k2_from_f64(const double *__restrict__ s, T *__restrict__ f, size_t N) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= (size_t)FQ * N) return;
    const int plane = (int)(i / N);
    st(f, pofs2(f, plane, i - (size_t)plane * N, N), plane % 6, s[i]);
}

// Deterministic two-pass reductions over the active nodes.
constexpr int NRED = 7;
__device__ __forceinline__ void red_combine(double *acc, const double *o) {
    acc[0] += o[0];
    acc[1] = fmax(acc[1], o[1]);
    acc[2] = fmin(acc[2], o[2]);
    acc[3] = fmax(acc[3], o[3]);
    acc[4] = fmax(acc[4], o[4]);
    acc[5] += o[5];
    acc[6] += o[6];
}
__device__ inline void block_reduce(double *acc, double (*sh)[NRED]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < NRED; ++k) sh[t][k] = acc[k];
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (t < s) red_combine(sh[t], sh[t + s]);
        __syncthreads();
    }
}
__global__ __launch_bounds__(BLOCK) void k2_reduce_partial(const double *__restrict__ U, const double *__restrict__ u,
                                                           const double *__restrict__ uprev,
                                                           const int32_t *__restrict__ list,
                                                           double *__restrict__ part, int na, size_t N) {
    __shared__ double sh[BLOCK][NRED];
    double acc[NRED] = {0.0, 0.0, INFINITY, -INFINITY, 0.0, 0.0, 0.0};
    for (int t = blockIdx.x * BLOCK + threadIdx.x; t < na; t += gridDim.x * BLOCK) {
        const int p = list[t];
        double Uv[NC];
        double nonfinite = 0.0;
#pragma unroll
        for (int a = 0; a < NC; ++a) {
            Uv[a] = U[(size_t)a * N + p];
            if (!isfinite(Uv[a])) nonfinite += 1.0;
        }
        double uu = 0.0, du = 0.0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const double x = u[(size_t)c * N + p];
            if (!isfinite(x)) nonfinite += 1.0;
            uu += x * x;
            if (uprev) {
                const double dd = x - uprev[(size_t)c * N + p];
                du += dd * dd;
            }
        }
        const double v2 = Uv[0] * Uv[0] + Uv[1] * Uv[1];
        const double J = Uv[2] * Uv[5] - Uv[3] * Uv[4];
        const double o[NRED] = {v2, sqrt(v2), J, J, sqrt(uu), nonfinite, du};
        red_combine(acc, o);
    }
    block_reduce(acc, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NRED; ++k) part[blockIdx.x * 8 + k] = sh[0][k];
}
__global__ __launch_bounds__(BLOCK) void k2_reduce_final(double *__restrict__ part, int nparts,
                                                         const int32_t *__restrict__ invalid,
                                                         double *__restrict__ out) {
    __shared__ double sh[BLOCK][NRED];
    double acc[NRED] = {0.0, 0.0, INFINITY, -INFINITY, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nparts; b += BLOCK) red_combine(acc, part + b * 8);
    block_reduce(acc, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) out[k] = sh[0][k];
        out[6] = (double)invalid[0];
        out[7] = sh[0][6];
    }
}

}  // namespace

// A captured loop of `steps` fused steps (an even count, so the population
// buffers end where they started) for one (f0, f1) pair.
struct StepGraph {
    const void *f0 = nullptr, *f1 = nullptr;
    hipGraphExec_t exec = nullptr;
};

struct vlbm2d_handle {
    vlbm2d_desc d;
    Params2 P;
    Links2 T;
    Box2 box{};  // faces of the rectangular box solver (d.box != 0)
    hipStream_t stream;
    Timing timing;
    int graph_steps = 0;
    double *clock = nullptr;
    hipStream_t cap = nullptr;
    StepGraph graphs[2];
    uint8_t *upmask = nullptr;  // (n_active) upstream-active bits, see k2_upmask
    uint8_t *nbmask = nullptr;  // (n_active) neighbour-active bits, see k2_nbmask
    bool upmask_ok = false;
};

static void drop_graphs(vlbm2d_handle *h) {
    for (auto &g : h->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = StepGraph{};
    }
}

static Params2 make_params2(const vlbm2d_desc &d) {
    Params2 P{};
    P.nx = d.nx;
    P.ny = d.ny;
    P.na = d.n_active;
    P.L = d.n_links;
    P.N = (size_t)d.nx * (size_t)d.ny;
    P.m = Material2{d.model, d.lam, d.mu, d.p0, d.p1, d.p2, d.j_floor};
    P.dt = d.dt;
    P.c = d.c;
    P.omega = d.omega;
    P.dx = d.dx;
    P.dy = d.dy;
    P.src_mode = d.source_mode;
    for (int k = 0; k < 6; ++k) P.src[k] = d.source_params[k];
    P.recon = d.recon;
    P.solver = d.normal_solver;
    P.disp_relax = d.disp_relax;
    P.w2 = d.disp_weight * d.disp_weight;
    P.blend = d.compat_blend;
    P.interior_only = d.compat_interior_only;
    P.use_dyn = d.use_dyn;
    P.gw = d.collision_model == 1 ? 0.25 * (d.omega - d.ghost_omega) : 0.0;
    P.filt = d.kinetic_filter;
    P.per_x = d.box && d.box_side_kind[0] == 0 && d.box_side_kind[1] == 0;
    P.per_y = d.box && d.box_side_kind[2] == 0 && d.box_side_kind[3] == 0;
    return P;
}

static Box2 make_box2(const vlbm2d_desc &d) {
    Box2 B{};
    for (int s = 0; s < 4; ++s) {
        B.kind[s] = d.box_side_kind[s];
        B.mode[s] = d.box_value_mode[s];
        for (int k = 0; k < 8; ++k) B.prm[s][k] = d.box_params[s * 8 + k];
    }
    B.lx = d.length_x;
    B.ly = d.length_y;
    return B;
}

// Side kinds of a box: 0..2 each, periodic on both sides of an axis or neither.
static const char *box_sides_error(const int32_t *kind) {
    for (int s = 0; s < 4; ++s)
        if (kind[s] < 0 || kind[s] > 2) return "box side kind must be 0 (periodic), 1 (Dirichlet) or 2 (Neumann)";
    if ((kind[0] == 0) != (kind[1] == 0)) return "x-periodic boundaries must be set on both left and right";
    if ((kind[2] == 0) != (kind[3] == 0)) return "y-periodic boundaries must be set on both bottom and top";
    return nullptr;
}

static Links2 make_links2(const vlbm2d_desc &d) {
    Links2 T{};
    T.lp = d.link_p;
    T.lq = d.link_q;
    T.lbid = d.link_bid;
    T.eta = d.link_eta;
    T.xb = d.link_xb;
    T.yb = d.link_yb;
    T.nx = d.link_nx;
    T.ny = d.link_ny;
    T.lsq_p = d.lsq_p;
    T.wb = d.lsq_wb;
    T.wx = d.lsq_wx;
    T.wy = d.lsq_wy;
    T.valid = d.lsq_valid;
    T.kind_by_id = d.kind_by_id;
    T.mode_by_id = d.mode_by_id;
    T.prm_by_id = d.params_by_id;
    T.dyn_tx = d.dyn_tx;
    T.dyn_ty = d.dyn_ty;
    T.act = d.active;
    return T;
}

#define CHECK_LAUNCH(where) VLBM_CHECK_LAUNCH(where)

template <typename T>
static void launch_box2(vlbm2d_handle *h, const Params2 &P, hipStream_t st, const T *src, T *dst, double t,
                        int use_rec) {
    const int m = P.nx > P.ny ? P.nx : P.ny;
    hipLaunchKernelGGL(k2_box_sides<T>, dim3((4 * m + LBLOCK - 1) / LBLOCK), dim3(LBLOCK), 0, st, src, dst,
                       (const T *)h->d.record, (const double *)h->d.U, use_rec, h->box, t, P);
}

template <typename T>
static int run_links2(vlbm2d_handle *h, const T *src, T *dst, double t, int use_rec) {
    const vlbm2d_desc &d = h->d;
    if (d.box) {
        launch_box2<T>(h, h->P, h->stream, src, dst, t, use_rec);
        CHECK_LAUNCH("k2_box_sides");
        return 0;
    }
    if (d.n_links == 0) return 0;
    const bool gather = recon_gather(d.recon);
    if (gather) {
        TimedSlot tg(h->timing, h->stream, VLBM_SLOT_GATHER);
        const int gg = (d.n_links + GL - 1) / GL;
        hipLaunchKernelGGL(k2_gather, dim3(gg), dim3(256), 0, h->stream, (const double *)d.U, (const double *)d.u,
                           d.lsq_p, d.lsq_wb, d.lsq_wx, d.lsq_wy, d.lsq_wt, (int)recon_compat(d.recon),
                           d.link_state, d.n_links, h->P.N);
        CHECK_LAUNCH("k2_gather");
    }
    TimedSlot ts(h->timing, h->stream, VLBM_SLOT_LINKS);
    const int gl = (d.n_links + LBLOCK - 1) / LBLOCK;
    if (recon_ext(d.recon))
        hipLaunchKernelGGL((k2_links<T, true>), dim3(gl), dim3(LBLOCK), 0, h->stream, src, dst, (const T *)d.record,
                           (const double *)d.U, (const double *)d.link_state, h->T, use_rec, (const double *)nullptr,
                           t, h->P);
    else
        hipLaunchKernelGGL((k2_links<T, false>), dim3(gl), dim3(LBLOCK), 0, h->stream, src, dst, (const T *)d.record,
                           (const double *)d.U, (const double *)d.link_state, h->T, use_rec, (const double *)nullptr,
                           t, h->P);
    CHECK_LAUNCH("k2_links");
    return 0;
}

// Gather and cut-link kernels of one step on stream `st`.
template <typename T>
static int launch_links2(vlbm2d_handle *h, const Params2 &P, hipStream_t st, const T *f0, T *f1, double tb,
                         const double *gin) {
    const vlbm2d_desc &d = h->d;
    if (d.box) {
        TimedSlot tl(h->timing, st, VLBM_SLOT_LINKS);
        launch_box2<T>(h, P, st, f0, f1, tb, 1);
        CHECK_LAUNCH("k2_box_sides");
        return 0;
    }
    if (d.n_links == 0) return 0;
    const bool gather = recon_gather(d.recon);
    if (gather) {
        TimedSlot tg(h->timing, st, VLBM_SLOT_GATHER);
        const int gg = (d.n_links + GL - 1) / GL;
        hipLaunchKernelGGL(k2_gather, dim3(gg), dim3(256), 0, st, (const double *)d.U, (const double *)d.u,
                           d.lsq_p, d.lsq_wb, d.lsq_wx, d.lsq_wy, d.lsq_wt, (int)recon_compat(d.recon),
                           d.link_state, d.n_links, P.N);
        CHECK_LAUNCH("k2_gather");
    }
    TimedSlot tl(h->timing, st, VLBM_SLOT_LINKS);
    const int gl = (d.n_links + LBLOCK - 1) / LBLOCK;
    const bool nh = d.model == M_NH;
    auto kern = recon_ext(d.recon) ? (nh ? k2_links<T, true, M_NH> : k2_links<T, true, -1>)
                                   : (nh ? k2_links<T, false, M_NH> : k2_links<T, false, -1>);
    hipLaunchKernelGGL(kern, dim3(gl), dim3(LBLOCK), 0, st, f0, f1, (const T *)d.record, (const double *)d.U,
                       (const double *)d.link_state, h->T, 1, gin, tb, P);
    CHECK_LAUNCH("k2_links");
    return 0;
}

template <typename T, bool NEXT, int MODE>
static void launch_stream2_t(vlbm2d_handle *h, const Params2 &P, const T *f0, T *f1, const double *gin,
                             double *gout) {
    const vlbm2d_desc &d = h->d;
    hipLaunchKernelGGL((k2_stream<T, NEXT, MODE>), dim3(xcd_grid(d.n_active)), dim3(BLOCK), 0, h->stream, f0, f1,
                       (const T *)d.record, d.active_list, (const uint8_t *)h->upmask, d.U, d.u, d.u_star, d.invalid,
                       gin, gout, P);
}

// mode bit 0: mrt_ghost (gin/gout ghost moments), bit 1: kinetic filter.
template <typename T>
static int launch_stream2(vlbm2d_handle *h, const Params2 &P, bool next, const T *f0, T *f1, const double *gin,
                          double *gout) {
    const int mode = (P.gw != 0.0 ? 1 : 0) | (P.filt > 0.0 ? 2 : 0);
    switch (mode | (next ? 4 : 0)) {
    case 0: launch_stream2_t<T, false, 0>(h, P, f0, f1, gin, gout); break;
    case 1: launch_stream2_t<T, false, 1>(h, P, f0, f1, gin, gout); break;
    case 2: launch_stream2_t<T, false, 2>(h, P, f0, f1, gin, gout); break;
    case 3: launch_stream2_t<T, false, 3>(h, P, f0, f1, gin, gout); break;
    case 4: launch_stream2_t<T, true, 0>(h, P, f0, f1, gin, gout); break;
    case 5: launch_stream2_t<T, true, 1>(h, P, f0, f1, gin, gout); break;
    case 6: launch_stream2_t<T, true, 2>(h, P, f0, f1, gin, gout); break;
    default: launch_stream2_t<T, true, 3>(h, P, f0, f1, gin, gout); break;
    }
    CHECK_LAUNCH("k2_stream");
    return 0;
}

// Ghost-moment buffer of step parity `par` (two (6, N) halves; mrt_ghost only).
static double *ghost_buf(const vlbm2d_handle *h, int par) {
    if (h->P.gw == 0.0) return nullptr;
    return h->d.ghost + (size_t)(par & 1) * NC * h->P.N;
}

// Launches of one step from (time, steps): sweeps, record, links and the
// stream (with the next step's collide when `next`).  P carries the clock
// pointer inside captured step loops.
template <typename T>
static int launch_step2(vlbm2d_handle *h, const Params2 &P, T *f0, T *f1, double time, int64_t steps, bool next,
                        int par) {
    const vlbm2d_desc &d = h->d;
    const int gn = xcd_grid(d.n_active);
    TimedSlot whole(h->timing, h->stream, VLBM_SLOT_STEP);
    const double tb = time + 0.5 * d.dt;
    const bool disp = d.disp_interval > 0 && steps % d.disp_interval == 0;
    const bool compat = d.compat_interval > 0 && steps % d.compat_interval == 0;
    hipStream_t st = h->stream;
    if (disp) {
        for (int s = 0; s < d.disp_sweeps; ++s) {
            TimedSlot ts(h->timing, st, VLBM_SLOT_SWEEP);
            hipLaunchKernelGGL(k2_sweep, dim3(gn), dim3(BLOCK), 0, st, (const double *)d.U, (const double *)d.u,
                               d.u_next, d.active_list, (const uint8_t *)h->nbmask, d.anchor_off, d.anchor_ids, h->T, tb, P);
            CHECK_LAUNCH("k2_sweep");
            if (s + 1 < d.disp_sweeps) {
                TimedSlot tc(h->timing, st, VLBM_SLOT_COMMIT);
                hipLaunchKernelGGL(k2_commit, dim3(gn), dim3(BLOCK), 0, st, (const double *)d.u_next, d.u, d.u_star,
                                   d.active_list, d.n_active, P.N);
                CHECK_LAUNCH("k2_commit");
            }
        }
    }
    {
        TimedSlot ts(h->timing, st, VLBM_SLOT_RECORD);
        auto kern = d.model == M_NH ? k2_record<T, M_NH> : k2_record<T, -1>;
        hipLaunchKernelGGL(kern, dim3(gn), dim3(BLOCK), 0, st, d.U,
                           disp ? (const double *)d.u_next : (const double *)d.u, d.u, d.u_star, (T *)d.record,
                           d.active_list, (const uint8_t *)h->nbmask, (int)compat, (int)disp,
                           next ? d.invalid : (int32_t *)nullptr,
                           P);
        CHECK_LAUNCH("k2_record");
    }
    const double *gin = ghost_buf(h, par);
    int rc = launch_links2<T>(h, P, st, f0, f1, tb, gin);
    if (rc) return rc;
    TimedSlot ts(h->timing, st, VLBM_SLOT_STREAM);
    return launch_stream2<T>(h, P, next, f0, f1, gin, ghost_buf(h, par + 1));
}

// Captures `graph_steps` fused steps with the time kept on the device (the
// schedule must be the same every step: intervals 0 or 1).
template <typename T>
static int capture_steps2(vlbm2d_handle *h, T *f0, T *f1, hipGraphExec_t *out) {
    hipError_t e;
    if (!h->cap) {
        e = hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking);
        if (e != hipSuccess) return fail("vlbm2d: capture stream", e);
    }
    if (!h->clock) {
        e = hipMalloc((void **)&h->clock, sizeof(double));
        if (e != hipSuccess) return fail("vlbm2d: clock", e);
    }
    Params2 P = h->P;
    P.clock = h->clock;
    hipStream_t keep = h->stream;
    const bool timing = h->timing.on;
    h->timing.on = false;
    h->stream = h->cap;
    e = hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        h->stream = keep;
        h->timing.on = timing;
        return fail("vlbm2d: begin capture", e);
    }

    int rc = 0;
    for (int k = 0; k < h->graph_steps && rc == 0; ++k) {
        rc = (k % 2 == 0) ? launch_step2<T>(h, P, f0, f1, 0.0, 0, true, 0)
                          : launch_step2<T>(h, P, f1, f0, 0.0, 0, true, 1);
        if (rc == 0) hipLaunchKernelGGL(k2_advance_clock, dim3(1), dim3(1), 0, h->cap, h->clock, h->d.dt);
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(h->cap, &graph);
    h->stream = keep;
    h->timing.on = timing;
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail("vlbm2d: end capture", e);
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail("vlbm2d: instantiate", e);
    return 0;
}

// One library call runs nsteps reference steps; the collide of every step
// after the first is fused into the previous step's stream kernel, the cut
// links run before the stream kernel, and the last step streams without the
// fused collide so that U, u and u* end as the reference leaves them.  With
// graph_steps > 0, blocks of fused steps are replayed from a captured graph.
template <typename T>
static int run_steps2(vlbm2d_handle *h, T *f0, T *f1, double time, int64_t steps, int nsteps) {
    const vlbm2d_desc &d = h->d;
    const int gn = xcd_grid(d.n_active);
    if (nsteps <= 0) return 0;
    hipLaunchKernelGGL(k2_reset_invalid, dim3(1), dim3(1), 0, h->stream, d.invalid);
    {
        TimedSlot ts(h->timing, h->stream, VLBM_SLOT_COLLIDE);
        hipLaunchKernelGGL((k2_moments<T, true>), dim3(gn), dim3(BLOCK), 0, h->stream, (const T *)f0, d.U, d.u,
                           d.u_star, d.invalid, d.active_list, ghost_buf(h, 0), h->P);
        CHECK_LAUNCH("k2_moments");
    }
    int k = 0;
    const bool uniform = d.disp_interval <= 1 && d.compat_interval <= 1;
    const int K = h->graph_steps;
    if (K > 0 && uniform && !h->timing.on && nsteps - 1 >= K) {
        StepGraph *g = nullptr;
        for (auto &c : h->graphs)
            if (c.exec && c.f0 == f0 && c.f1 == f1) g = &c;
        if (!g) {
            g = !h->graphs[0].exec ? &h->graphs[0] : &h->graphs[1];
            if (g->exec) (void)hipGraphExecDestroy(g->exec);
            *g = StepGraph{};
            int rc = capture_steps2<T>(h, f0, f1, &g->exec);
            if (rc) return rc;
            g->f0 = f0;
            g->f1 = f1;
        }
        hipLaunchKernelGGL(k2_set_clock, dim3(1), dim3(1), 0, h->stream, h->clock, time);
        CHECK_LAUNCH("k2_set_clock");
        for (; k + K <= nsteps - 1; k += K) {
            hipError_t e = hipGraphLaunch(g->exec, h->stream);
            if (e != hipSuccess) return fail("vlbm2d: graph launch", e);
            for (int j = 0; j < K; ++j) time += d.dt;
            steps += K;
        }
    }
    for (; k < nsteps; ++k) {
        const bool next = k + 1 < nsteps;
        int rc = launch_step2<T>(h, h->P, f0, f1, time, steps, next, k & 1);
        if (rc) return rc;
        T *tmp = f0;
        f0 = f1;
        f1 = tmp;
        time += d.dt;
        steps += 1;
    }
    return 0;
}

namespace {
// ---- batches: many lattices, one launch per kernel per step ---------------
//
// Every shell of a batch keeps its own handle, buffers and parameters; a
// batch step issues each kernel once over the work items of all active
// shells.  Block b of a batched launch (after the XCD remap) belongs to the
// shell j with start[j] <= b < start[j + 1] of that kernel's block table and
// works on the shell's local block b - start[j], running the same per-node /
// per-link code as a single-lattice launch, so every shell's results are
// bitwise those of stepping it alone.  The boundary time of each shell lives
// in its own device clock (k2b_advance_clocks adds the shell's dt after every
// step, the operation the host performs).
struct Job2 {
    Params2 P;  // P.clock -> the shell's clock
    Links2 T;
    double *U, *u, *u_star, *u_next;
    void *record;
    int32_t *invalid;
    const int32_t *list, *aoff, *aids;
    const uint8_t *nbm, *upm;
    const int32_t *lsq_p;
    const double *wb, *wx, *wy, *wt;
    double *link_state;
    void *f[2];  // populations: step parity k reads f[k & 1] and writes f[(k + 1) & 1]
    int disp, compat, tangent;
};

enum { BK_NODE = 0, BK_SWEEP, BK_GATHER, BK_LINK, BK_COUNT };

__device__ __forceinline__ int find_job(const int *start, int njobs, int b) {
    int lo = 0, hi = njobs;  // start[lo] <= b < start[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <typename T, int NB>
__global__ __launch_bounds__(NB) void k2b_moments(const Job2 *__restrict__ jobs, const int *__restrict__ start,
                                                  int njobs, int parity) {
    const int b = xcd_block();
    const int j = find_job(start, njobs, b);
    const Job2 &J = jobs[j];
    k2_moments_body<T, true>((const T *)J.f[parity], J.U, J.u, J.u_star, J.invalid, J.list, nullptr, J.P,
                             (b - start[j]) * NB + threadIdx.x);
}

template <int NB>
__global__ __launch_bounds__(NB) void k2b_sweep(const Job2 *__restrict__ jobs, const int *__restrict__ start,
                                                int njobs) {
    const int b = xcd_block();
    const int j = find_job(start, njobs, b);
    const Job2 &J = jobs[j];
    if (!J.disp) return;  // padding blocks past the table land on the last shell
    k2_sweep_body(J.U, J.u, J.u_next, J.list, J.nbm, J.aoff, J.aids, J.T, 0.0, J.P, (b - start[j]) * NB + threadIdx.x);
}

template <typename T, int NB, int LAW>
__global__ __launch_bounds__(NB) void k2b_record(const Job2 *__restrict__ jobs, const int *__restrict__ start,
                                                 int njobs, int next) {
    const int b = xcd_block();
    const int j = find_job(start, njobs, b);
    const Job2 &J = jobs[j];
    k2_record_body<T, LAW>(J.U, J.disp ? J.u_next : J.u, J.u, J.u_star, (T *)J.record, J.list, J.nbm, J.compat, J.disp,
                      next ? J.invalid : (int32_t *)nullptr, J.P, (b - start[j]) * NB + threadIdx.x);
}

__global__ __launch_bounds__(256) void k2b_gather(const Job2 *__restrict__ jobs, const int *__restrict__ start,
                                                  int njobs) {
    const int b = blockIdx.x;
    const int j = find_job(start, njobs, b);
    const Job2 &J = jobs[j];
    k2_gather_body(J.U, J.u, J.lsq_p, J.wb, J.wx, J.wy, J.wt, J.tangent, J.link_state, J.P.L, J.P.N, b - start[j]);
}

template <typename T, bool EXT, int NL, int LAW>
__global__ __launch_bounds__(NL) void k2b_links(const Job2 *__restrict__ jobs, const int *__restrict__ start,
                                                int njobs, int parity) {
    const int b = blockIdx.x;
    const int j = find_job(start, njobs, b);
    const Job2 &J = jobs[j];
    k2_links_body<T, EXT, LAW>((const T *)J.f[parity], (T *)J.f[parity ^ 1], (const T *)J.record, J.U, J.link_state, J.T,
                          1, nullptr, 0.0, J.P, (b - start[j]) * NL + threadIdx.x);
}

template <typename T, bool NEXT, int NB>
__global__ __launch_bounds__(NB) void k2b_stream(const Job2 *__restrict__ jobs, const int *__restrict__ start,
                                                 int njobs, int parity) {
    const int b = xcd_block();
    const int j = find_job(start, njobs, b);
    const Job2 &J = jobs[j];
    k2_stream_body<T, NEXT, 0>((const T *)J.f[parity], (T *)J.f[parity ^ 1], (const T *)J.record, J.list, J.upm,
                               J.U, J.u, J.u_star, J.invalid, nullptr, nullptr, J.P, (b - start[j]) * NB + threadIdx.x);
}

__global__ void k2b_reset_invalid(const Job2 *__restrict__ jobs, int njobs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < njobs) jobs[j].invalid[0] = 0;
}

__global__ void k2b_advance_clocks(double *__restrict__ clocks, const double *__restrict__ dts, int njobs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < njobs) clocks[j] = clocks[j] + dts[j];
}

}  // namespace

struct vlbm2d_batch {
    std::vector<vlbm2d_handle *> all;  // every shell, in creation order
    std::vector<void *> f0, f1;        // their population buffers at parity 0
    int precision = 0;
    bool ext = false;
    bool nh = false;  // every shell neo-Hooke: record and link kernels specialised on that law
    hipStream_t stream = nullptr;
    // device tables of the active shells
    Job2 *jobs = nullptr;
    int *start = nullptr;  // BK_COUNT tables of njobs + 1 block offsets
    double *clocks = nullptr, *dts = nullptr;
    int njobs = 0, cap = 0;
    int nblk[BK_COUNT] = {0, 0, 0, 0};
    int parity = 0;  // populations are in f[parity] of every shell
    int nb = 256, nl = 64;    // threads per block of the node kernels and of the link kernel
    std::vector<int> active;  // indices into `all`
    bool gather = false;
    // captured step loops of the active set (graph_steps fused steps)
    int graph_steps = 0;
    hipStream_t capst = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // one per starting parity
    void drop_graphs() {
        for (auto &g : graph) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }
    void release() {
        drop_graphs();
        if (capst) (void)hipStreamDestroy(capst);
        for (void *ptr : {(void *)jobs, (void *)start, (void *)clocks, (void *)dts})
            if (ptr) (void)hipFree(ptr);
        jobs = nullptr;
        start = nullptr;
        clocks = dts = nullptr;
        capst = nullptr;
    }
};

namespace {

// Builds the job and block tables of the shells listed in b->active.
int batch_tables(vlbm2d_batch *b) {
    const int n = (int)b->active.size();
    if (n > b->cap) {
        for (void *ptr : {(void *)b->jobs, (void *)b->start, (void *)b->clocks, (void *)b->dts})
            if (ptr) (void)hipFree(ptr);
        b->jobs = nullptr;
        b->start = nullptr;
        b->clocks = b->dts = nullptr;
        hipError_t e = hipMalloc((void **)&b->jobs, sizeof(Job2) * (size_t)n);
        if (e == hipSuccess) e = hipMalloc((void **)&b->start, sizeof(int) * (size_t)(n + 1) * BK_COUNT);
        if (e == hipSuccess) e = hipMalloc((void **)&b->clocks, sizeof(double) * (size_t)n);
        if (e == hipSuccess) e = hipMalloc((void **)&b->dts, sizeof(double) * (size_t)n);
        if (e != hipSuccess) return fail("vlbm2d_batch: tables", e);
        b->cap = n;
    }
    std::vector<Job2> jobs((size_t)n);
    std::vector<int> start((size_t)(n + 1) * BK_COUNT, 0);
    std::vector<double> dts((size_t)n);
    int tot[BK_COUNT] = {0, 0, 0, 0};
    b->gather = false;
    for (int k = 0; k < n; ++k) {
        const int s = b->active[k];
        vlbm2d_handle *h = b->all[s];
        const vlbm2d_desc &d = h->d;
        Job2 &J = jobs[k];
        J.P = h->P;
        J.P.clock = b->clocks + k;
        J.T = h->T;
        J.U = d.U;
        J.u = d.u;
        J.u_star = d.u_star;
        J.u_next = d.u_next;
        J.record = d.record;
        J.invalid = d.invalid;
        J.list = d.active_list;
        J.aoff = d.anchor_off;
        J.aids = d.anchor_ids;
        J.nbm = h->nbmask;
        J.upm = h->upmask;
        J.lsq_p = d.lsq_p;
        J.wb = d.lsq_wb;
        J.wx = d.lsq_wx;
        J.wy = d.lsq_wy;
        J.wt = d.lsq_wt;
        J.link_state = d.link_state;
        J.f[0] = b->f0[s];
        J.f[1] = b->f1[s];
        J.disp = d.disp_interval == 1;
        J.compat = d.compat_interval == 1;
        J.tangent = recon_compat(d.recon) ? 1 : 0;
        const bool g = d.n_links > 0 && recon_gather(d.recon);
        b->gather = b->gather || g;
        const int NBk = b->nb, NLk = b->nl;
        const int nb[BK_COUNT] = {(d.n_active + NBk - 1) / NBk, J.disp ? (d.n_active + NBk - 1) / NBk : 0,
                                  g ? (d.n_links + GL - 1) / GL : 0, (d.n_links + NLk - 1) / NLk};
        for (int t = 0; t < BK_COUNT; ++t) {
            start[(size_t)t * (n + 1) + k] = tot[t];
            tot[t] += nb[t];
        }
        dts[k] = d.dt;
    }
    for (int t = 0; t < BK_COUNT; ++t) {
        start[(size_t)t * (n + 1) + n] = tot[t];
        b->nblk[t] = tot[t];
    }
    b->njobs = n;
    hipError_t e = hipMemcpyAsync(b->jobs, jobs.data(), sizeof(Job2) * (size_t)n, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->dts, dts.data(), sizeof(double) * (size_t)n, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail("vlbm2d_batch: upload", e);
    b->drop_graphs();
    return 0;
}

// Node-kernel grids rounded up to whole XCD groups; the extra blocks find
// the last shell and their threads fall past its node count.
inline int batch_grid(int blocks) { return ((blocks + vlbm_common::NXCD - 1) / vlbm_common::NXCD) * vlbm_common::NXCD; }

template <typename T, int NB, int NL>
int batch_step_launches_t(vlbm2d_batch *b, hipStream_t st, int parity, bool next) {
    const int n = b->njobs;
    const int *S = b->start;
    // node grids are rounded up to whole XCD groups; the padding blocks map
    // to the last shell past its node count (or, in the sweep, skip it)
    if (b->nblk[BK_SWEEP] > 0) {
        hipLaunchKernelGGL(k2b_sweep<NB>, dim3(batch_grid(b->nblk[BK_SWEEP])), dim3(NB), 0, st, b->jobs,
                           S + 1 * (n + 1), n);
        CHECK_LAUNCH("k2b_sweep");
    }
    auto rec = b->nh ? k2b_record<T, NB, M_NH> : k2b_record<T, NB, -1>;
    hipLaunchKernelGGL(rec, dim3(batch_grid(b->nblk[BK_NODE])), dim3(NB), 0, st, b->jobs, S, n, (int)next);
    CHECK_LAUNCH("k2b_record");
    if (b->gather && b->nblk[BK_GATHER] > 0) {
        hipLaunchKernelGGL(k2b_gather, dim3(b->nblk[BK_GATHER]), dim3(256), 0, st, b->jobs, S + 2 * (n + 1), n);
        CHECK_LAUNCH("k2b_gather");
    }
    if (b->nblk[BK_LINK] > 0) {
        auto links = b->ext ? (b->nh ? k2b_links<T, true, NL, M_NH> : k2b_links<T, true, NL, -1>)
                            : (b->nh ? k2b_links<T, false, NL, M_NH> : k2b_links<T, false, NL, -1>);
        hipLaunchKernelGGL(links, dim3(b->nblk[BK_LINK]), dim3(NL), 0, st, b->jobs, S + 3 * (n + 1), n, parity);
        CHECK_LAUNCH("k2b_links");
    }
    if (next)
        hipLaunchKernelGGL((k2b_stream<T, true, NB>), dim3(batch_grid(b->nblk[BK_NODE])), dim3(NB), 0, st, b->jobs, S,
                           n, parity);
    else
        hipLaunchKernelGGL((k2b_stream<T, false, NB>), dim3(batch_grid(b->nblk[BK_NODE])), dim3(NB), 0, st, b->jobs,
                           S, n, parity);
    CHECK_LAUNCH("k2b_stream");
    hipLaunchKernelGGL(k2b_advance_clocks, dim3((n + 255) / 256), dim3(256), 0, st, b->clocks, b->dts, n);
    CHECK_LAUNCH("k2b_advance_clocks");
    return 0;
}

// Block sizes the autotuner chooses from.
constexpr int TUNE_NB[4] = {64, 128, 256, 512};
constexpr int TUNE_NL[3] = {32, 64, 128};

template <typename T, int NB>
int batch_step_nl(vlbm2d_batch *b, hipStream_t st, int parity, bool next) {
    switch (b->nl) {
    case 32: return batch_step_launches_t<T, NB, 32>(b, st, parity, next);
    case 128: return batch_step_launches_t<T, NB, 128>(b, st, parity, next);
    default: return batch_step_launches_t<T, NB, 64>(b, st, parity, next);
    }
}

template <typename T>
int batch_step_launches(vlbm2d_batch *b, hipStream_t st, int parity, bool next) {
    switch (b->nb) {
    case 64: return batch_step_nl<T, 64>(b, st, parity, next);
    case 128: return batch_step_nl<T, 128>(b, st, parity, next);
    case 512: return batch_step_nl<T, 512>(b, st, parity, next);
    default: return batch_step_nl<T, 256>(b, st, parity, next);
    }
}

template <typename T>
int batch_moments(vlbm2d_batch *b, hipStream_t st) {
    const int n = b->njobs;
    const dim3 g(batch_grid(b->nblk[BK_NODE]));
    switch (b->nb) {
    case 64: hipLaunchKernelGGL((k2b_moments<T, 64>), g, dim3(64), 0, st, b->jobs, b->start, n, b->parity); break;
    case 128: hipLaunchKernelGGL((k2b_moments<T, 128>), g, dim3(128), 0, st, b->jobs, b->start, n, b->parity); break;
    case 512: hipLaunchKernelGGL((k2b_moments<T, 512>), g, dim3(512), 0, st, b->jobs, b->start, n, b->parity); break;
    default: hipLaunchKernelGGL((k2b_moments<T, 256>), g, dim3(256), 0, st, b->jobs, b->start, n, b->parity); break;
    }
    CHECK_LAUNCH("k2b_moments");
    return 0;
}

template <typename T>
int batch_capture(vlbm2d_batch *b, int parity, hipGraphExec_t *out) {
    hipError_t e;
    if (!b->capst) {
        e = hipStreamCreateWithFlags(&b->capst, hipStreamNonBlocking);
        if (e != hipSuccess) return fail("vlbm2d_batch: capture stream", e);
    }
    e = hipStreamBeginCapture(b->capst, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail("vlbm2d_batch: begin capture", e);
    int rc = 0;
    for (int k = 0; k < b->graph_steps && rc == 0; ++k) rc = batch_step_launches<T>(b, b->capst, (parity + k) & 1, true);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(b->capst, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail("vlbm2d_batch: end capture", e);
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail("vlbm2d_batch: instantiate", e);
    return 0;
}

// nsteps steps of every active shell from its clock; as run_steps2 for one
// lattice: the first collide, fused collides inside, the last step without.
template <typename T>
int batch_run(vlbm2d_batch *b, int nsteps) {
    hipStream_t st = b->stream;
    const int n = b->njobs;
    if (n == 0 || nsteps <= 0) return 0;
    hipLaunchKernelGGL(k2b_reset_invalid, dim3((n + 255) / 256), dim3(256), 0, st, b->jobs, n);
    CHECK_LAUNCH("k2b_reset_invalid");
    if (int rc = batch_moments<T>(b, st)) return rc;
    int k = 0;
    const int K = b->graph_steps;
    if (K > 0 && nsteps - 1 >= K) {
        hipGraphExec_t &g = b->graph[b->parity];
        if (!g) {
            int rc = batch_capture<T>(b, b->parity, &g);
            if (rc) return rc;
        }
        for (; k + K <= nsteps - 1; k += K) {
            hipError_t e = hipGraphLaunch(g, st);
            if (e != hipSuccess) return fail("vlbm2d_batch: graph launch", e);
        }
    }
    for (; k < nsteps; ++k) {
        int rc = batch_step_launches<T>(b, st, (b->parity + k) & 1, k + 1 < nsteps);
        if (rc) return rc;
    }
    if (nsteps & 1) b->parity ^= 1;
    return 0;
}

}  // namespace

extern "C" {

int vlbm2d_create(const vlbm2d_desc *desc, vlbm2d_handle **out) {
    if (!desc || !out) return fail_msg("vlbm2d_create: null argument");
    if (desc->struct_size != sizeof(vlbm2d_desc))
        return fail_msg("vlbm2d_create: descriptor size mismatch (" + std::to_string(desc->struct_size) +
                        " != " + std::to_string(sizeof(vlbm2d_desc)) + ")");
    if (desc->abi_version != VLBM2D_ABI_VERSION) return fail_msg("vlbm2d_create: ABI version mismatch");
    if (desc->nx < 4 || desc->ny < 4) return fail_msg("vlbm2d_create: nx and ny must be at least four");
    if (desc->n_active <= 0 || (size_t)desc->n_active > (size_t)desc->nx * desc->ny)
        return fail_msg("vlbm2d_create: n_active out of range");
    if ((size_t)desc->nx * desc->ny * FQ >= (size_t)1 << 40) return fail_msg("vlbm2d_create: lattice too large");
    if ((size_t)desc->nx * desc->ny >= (size_t)1 << 31) return fail_msg("vlbm2d_create: lattice exceeds int32 indexing");
    if (desc->n_links < 0) return fail_msg("vlbm2d_create: n_links must be non-negative");
    if (desc->precision != VLBM_PREC_F64 && desc->precision != VLBM_PREC_F32)
        return fail_msg("vlbm2d_create: unknown precision");
    if (desc->recon < R_BFL || desc->recon > R_LOCALF_BFL)
        return fail_msg("vlbm2d_create: unknown reconstruction id (0..13)");
    if (desc->disp_sweeps < 1 || desc->disp_sweeps > 8) return fail_msg("vlbm2d_create: disp_sweeps must be in [1,8]");
    if (desc->collision_model != 0 && desc->collision_model != 1)
        return fail_msg("vlbm2d_create: collision_model must be 0 (bgk) or 1 (mrt_ghost)");
    if (desc->collision_model == 1) {
        if (!(desc->ghost_omega > 0.0 && desc->ghost_omega <= 2.0))
            return fail_msg("vlbm2d_create: ghost_omega must be in (0,2]");
        if (!desc->ghost) return fail_msg("vlbm2d_create: ghost is null (mrt_ghost)");
    }
    if (!(desc->kinetic_filter >= 0.0 && desc->kinetic_filter <= 1.0))
        return fail_msg("vlbm2d_create: kinetic_filter must be in [0,1]");
    if (desc->box) {
        if (desc->n_links != 0) return fail_msg("vlbm2d_create: a box has no cut links (n_links must be 0)");
        if (desc->n_active != desc->nx * desc->ny) return fail_msg("vlbm2d_create: every box node must be active");
        if (const char *e = box_sides_error(desc->box_side_kind)) return fail_msg(std::string("vlbm2d_create: ") + e);
        if (!(desc->length_x > 0.0 && desc->length_y > 0.0)) return fail_msg("vlbm2d_create: box lengths must be positive");
    }
    const void *req[] = {desc->U, desc->u, desc->u_star, desc->u_next, desc->record, desc->invalid,
                         desc->reduce_scratch, desc->active, desc->active_list, desc->anchor_off,
                         desc->kind_by_id, desc->mode_by_id, desc->params_by_id};
    for (const void *ptr : req)
        if (!ptr) return fail_msg("vlbm2d_create: a required device pointer is null");
    if (desc->n_links > 0) {
        const void *lreq[] = {desc->link_p, desc->link_q, desc->link_bid, desc->link_eta, desc->link_xb,
                              desc->link_yb, desc->link_nx, desc->link_ny, desc->lsq_p, desc->lsq_wb,
                              desc->lsq_wx, desc->lsq_wy, desc->lsq_valid, desc->dyn_tx, desc->dyn_ty,
                              desc->link_state};
        for (const void *ptr : lreq)
            if (!ptr) return fail_msg("vlbm2d_create: a required link pointer is null");
        if (recon_compat(desc->recon) && !desc->lsq_wt) return fail_msg("vlbm2d_create: lsq_wt is null");
    }
    if (desc->n_anchor_links > 0 && !desc->anchor_ids) return fail_msg("vlbm2d_create: anchor_ids is null");
    uint8_t *upmask = nullptr, *nbmask = nullptr;
    hipError_t ae = hipMalloc((void **)&upmask, (size_t)desc->n_active);
    if (ae != hipSuccess) return fail("vlbm2d_create: upstream mask", ae);
    ae = hipMalloc((void **)&nbmask, (size_t)desc->n_active);
    if (ae != hipSuccess) {
        (void)hipFree(upmask);
        return fail("vlbm2d_create: neighbour mask", ae);
    }
    vlbm2d_handle *h = new vlbm2d_handle;
    h->upmask = upmask;
    h->nbmask = nbmask;
    h->d = *desc;
    h->P = make_params2(*desc);
    h->T = make_links2(*desc);
    h->box = make_box2(*desc);
    h->stream = (hipStream_t)desc->stream;
    *out = h;
    return 0;
}

int vlbm2d_destroy(vlbm2d_handle *h) {
    if (!h) return 0;
    drop_graphs(h);
    if (h->clock) (void)hipFree(h->clock);
    if (h->upmask) (void)hipFree(h->upmask);
    if (h->nbmask) (void)hipFree(h->nbmask);
    if (h->cap) (void)hipStreamDestroy(h->cap);
    h->timing.destroy();
    delete h;
    return 0;
}

int vlbm2d_set_graph_steps(vlbm2d_handle *h, int32_t steps) {
    if (!h) return fail_msg("vlbm2d_set_graph_steps: null handle");
    if (steps < 0 || steps % 2) return fail_msg("vlbm2d_set_graph_steps: steps must be a non-negative even count");
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail("vlbm2d_set_graph_steps: sync", e);
    drop_graphs(h);
    h->graph_steps = steps;
    return 0;
}

int vlbm2d_timing_enable(vlbm2d_handle *h, int32_t enable) {
    if (!h) return fail_msg("vlbm2d_timing_enable: null handle");
    return h->timing.enable(h->stream, enable);
}

int vlbm2d_timing_read(vlbm2d_handle *h, double *ms, int32_t *counts, int32_t nslots) {
    if (!h || !ms || !counts) return fail_msg("vlbm2d_timing_read: null argument");
    return h->timing.read(h->stream, ms, counts, nslots);
}

int vlbm2d_step(vlbm2d_handle *h, void *f0, void *f1, double time, int64_t steps, int32_t nsteps) {
    if (!h || !f0 || !f1) return fail_msg("vlbm2d_step: null argument");
    if (nsteps < 0) return fail_msg("vlbm2d_step: nsteps must be non-negative");
    if (!h->upmask_ok) {
        // On the handle's stream ahead of every launch (and graph) that reads it.
        hipLaunchKernelGGL(k2_upmask, dim3((h->d.n_active + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, h->stream,
                           h->d.active_list, h->d.active, h->upmask, h->P);
        CHECK_LAUNCH("k2_upmask");
        hipLaunchKernelGGL(k2_nbmask, dim3((h->d.n_active + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, h->stream,
                           h->d.active_list, h->d.active, h->nbmask, h->P);
        CHECK_LAUNCH("k2_nbmask");
        h->upmask_ok = true;
    }
    if (h->d.precision == VLBM_PREC_F64)
        return run_steps2<double>(h, (double *)f0, (double *)f1, time, steps, nsteps);
    return run_steps2<float>(h, (float *)f0, (float *)f1, time, steps, nsteps);
}

int vlbm2d_refresh(vlbm2d_handle *h, const void *f) {
    if (!h || !f) return fail_msg("vlbm2d_refresh: null argument");
    const vlbm2d_desc &d = h->d;
    hipLaunchKernelGGL(k2_reset_invalid, dim3(1), dim3(1), 0, h->stream, d.invalid);
    const int gn = xcd_grid(d.n_active);
    if (d.precision == VLBM_PREC_F64)
        hipLaunchKernelGGL((k2_moments<double, false>), dim3(gn), dim3(BLOCK), 0, h->stream, (const double *)f, d.U,
                           d.u, d.u_star, d.invalid, d.active_list, (double *)nullptr, h->P);
    else
        hipLaunchKernelGGL((k2_moments<float, false>), dim3(gn), dim3(BLOCK), 0, h->stream, (const float *)f, d.U,
                           d.u, d.u_star, d.invalid, d.active_list, (double *)nullptr, h->P);
    CHECK_LAUNCH("k2_refresh");
    return 0;
}

int vlbm2d_apply_boundary(vlbm2d_handle *h, const void *f_src, void *f_dst, double t) {
    if (!h || !f_src || !f_dst) return fail_msg("vlbm2d_apply_boundary: null argument");
    if (f_src == f_dst && !h->d.box) return fail_msg("vlbm2d_apply_boundary: source and target must differ");
    if (h->d.precision == VLBM_PREC_F64) return run_links2<double>(h, (const double *)f_src, (double *)f_dst, t, 0);
    return run_links2<float>(h, (const float *)f_src, (float *)f_dst, t, 0);
}

int vlbm2d_init_populations(vlbm2d_handle *h, void *f) {
    if (!h || !f) return fail_msg("vlbm2d_init_populations: null argument");
    const int gn = xcd_grid(h->d.n_active);
    if (h->d.precision == VLBM_PREC_F64)
        hipLaunchKernelGGL(k2_init_pop<double>, dim3(gn), dim3(BLOCK), 0, h->stream, (const double *)h->d.U,
                           (double *)f, h->d.active_list, h->P);
    else
        hipLaunchKernelGGL(k2_init_pop<float>, dim3(gn), dim3(BLOCK), 0, h->stream, (const double *)h->d.U,
                           (float *)f, h->d.active_list, h->P);
    CHECK_LAUNCH("k2_init_pop");
    return 0;
}

int vlbm2d_set_ustar(vlbm2d_handle *h, const double *disp) {
    if (!h || !disp) return fail_msg("vlbm2d_set_ustar: null argument");
    const int g = (h->d.n_active + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k2_set_ustar, dim3(g), dim3(BLOCK), 0, h->stream, disp, (const double *)h->d.U, h->d.u_star,
                       h->d.active_list, h->d.dt, h->d.n_active, h->P.N);
    CHECK_LAUNCH("k2_set_ustar");
    return 0;
}

int vlbm2d_stress(vlbm2d_handle *h, double *P, double *sigma) {
    if (!h || !P || !sigma) return fail_msg("vlbm2d_stress: null argument");
    const int g = (h->d.n_active + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k2_stress, dim3(g), dim3(BLOCK), 0, h->stream, (const double *)h->d.U, P, sigma,
                       h->d.active_list, h->P);
    CHECK_LAUNCH("k2_stress");
    return 0;
}

int vlbm2d_post_collision(vlbm2d_handle *h, const void *f_prev, void *fpost) {
    if (!h || !f_prev || !fpost) return fail_msg("vlbm2d_post_collision: null argument");
    const int g = (h->d.n_active + BLOCK - 1) / BLOCK;
    if (h->d.precision == VLBM_PREC_F64)
        hipLaunchKernelGGL(k2_post<double>, dim3(g), dim3(BLOCK), 0, h->stream, (const double *)f_prev,
                           (double *)fpost, (const double *)h->d.record, h->d.active_list, h->P);
    else
        hipLaunchKernelGGL(k2_post<float>, dim3(g), dim3(BLOCK), 0, h->stream, (const float *)f_prev,
                           (float *)fpost, (const float *)h->d.record, h->d.active_list, h->P);
    CHECK_LAUNCH("k2_post");
    return 0;
}

int vlbm2d_set_box_sides(vlbm2d_handle *h, const int32_t *kind, const int32_t *mode, const double *params) {
    if (!h || !kind || !mode || !params) return fail_msg("vlbm2d_set_box_sides: null argument");
    if (!h->d.box) return fail_msg("vlbm2d_set_box_sides: not a box solver");
    if (const char *e = box_sides_error(kind)) return fail_msg(std::string("vlbm2d_set_box_sides: ") + e);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail("vlbm2d_set_box_sides: sync", e);
    drop_graphs(h);  // the faces are captured kernel arguments
    for (int s = 0; s < 4; ++s) {
        h->d.box_side_kind[s] = kind[s];
        h->d.box_value_mode[s] = mode[s];
    }
    for (int k = 0; k < 32; ++k) h->d.box_params[k] = params[k];
    const int use_dyn = h->P.use_dyn;
    h->P = make_params2(h->d);
    h->P.use_dyn = use_dyn;
    h->box = make_box2(h->d);
    h->upmask_ok = false;  // the periodic axes may have changed
    return 0;
}

int vlbm2d_kinetic_filter(vlbm2d_handle *h, void *f) {
    if (!h || !f) return fail_msg("vlbm2d_kinetic_filter: null argument");
    if (h->P.filt <= 0.0) return 0;
    const int g = (h->d.n_active + BLOCK - 1) / BLOCK;
    if (h->d.precision == VLBM_PREC_F64)
        hipLaunchKernelGGL(k2_filter<double>, dim3(g), dim3(BLOCK), 0, h->stream, (double *)f, h->d.active_list, h->P);
    else
        hipLaunchKernelGGL(k2_filter<float>, dim3(g), dim3(BLOCK), 0, h->stream, (float *)f, h->d.active_list, h->P);
    CHECK_LAUNCH("k2_filter");
    return 0;
}

int vlbm2d_set_dynamic_traction(vlbm2d_handle *h, int32_t enabled) {
    if (!h) return fail_msg("vlbm2d_set_dynamic_traction: null handle");
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail("vlbm2d_set_dynamic_traction: sync", e);
    drop_graphs(h);  // the flag is a captured kernel argument
    h->P.use_dyn = enabled != 0;
    h->d.use_dyn = enabled != 0;
    return 0;
}

int vlbm2d_reduce(vlbm2d_handle *h, const double *u_prev, double *out) {
    if (!h || !out) return fail_msg("vlbm2d_reduce: null argument");
    const vlbm2d_desc &d = h->d;
    int nb = (d.n_active + BLOCK - 1) / BLOCK;
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    double *part = d.reduce_scratch;
    double *res = d.reduce_scratch + (size_t)RED_BLOCKS * 8;
    hipLaunchKernelGGL(k2_reduce_partial, dim3(nb), dim3(BLOCK), 0, h->stream, (const double *)d.U,
                       (const double *)d.u, u_prev, d.active_list, part, d.n_active, h->P.N);
    CHECK_LAUNCH("k2_reduce_partial");
    hipLaunchKernelGGL(k2_reduce_final, dim3(1), dim3(BLOCK), 0, h->stream, part, nb, (const int32_t *)d.invalid, res);
    CHECK_LAUNCH("k2_reduce_final");
    hipError_t e = hipMemcpyAsync(out, res, sizeof(double) * VLBM_RED_COUNT, hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return fail("vlbm2d_reduce: copy", e);
    e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail("vlbm2d_reduce: sync", e);
    return 0;
}

int vlbm2d_populations_to_f64(vlbm2d_handle *h, const void *f, double *dst) {
    if (!h || !f || !dst) return fail_msg("vlbm2d_populations_to_f64: null argument");
    const size_t count = (size_t)FQ * h->P.N;
    const int g = (int)((count + BLOCK - 1) / BLOCK);
    if (h->d.precision == VLBM_PREC_F64)
        hipLaunchKernelGGL(k2_to_f64<double>, dim3(g), dim3(BLOCK), 0, h->stream, (const double *)f, dst, h->P.N);
    else
        hipLaunchKernelGGL(k2_to_f64<float>, dim3(g), dim3(BLOCK), 0, h->stream, (const float *)f, dst, h->P.N);
    CHECK_LAUNCH("vlbm2d_populations_to_f64");
    return 0;
}

int vlbm2d_populations_from_f64(vlbm2d_handle *h, const double *src, void *f) {
    if (!h || !f || !src) return fail_msg("vlbm2d_populations_from_f64: null argument");
    const size_t count = (size_t)FQ * h->P.N;
    const int g = (int)((count + BLOCK - 1) / BLOCK);
    if (h->d.precision == VLBM_PREC_F64)
        hipLaunchKernelGGL(k2_from_f64<double>, dim3(g), dim3(BLOCK), 0, h->stream, src, (double *)f, h->P.N);
    else
        hipLaunchKernelGGL(k2_from_f64<float>, dim3(g), dim3(BLOCK), 0, h->stream, src, (float *)f, h->P.N);
    CHECK_LAUNCH("vlbm2d_populations_from_f64");
    return 0;
}


int vlbm2d_batch_create(vlbm2d_handle *const *handles, void *const *f0, void *const *f1, int32_t n, void *stream,
                        vlbm2d_batch **out) {
    if (!handles || !f0 || !f1 || !out) return fail_msg("vlbm2d_batch_create: null argument");
    if (n < 1 || n > 65536) return fail_msg("vlbm2d_batch_create: n must be in [1, 65536]");
    const int prec = handles[0] ? handles[0]->d.precision : 0;
    bool ext = false, nh = true;
    for (int k = 0; k < n; ++k) {
        const vlbm2d_handle *h = handles[k];
        if (!h || !f0[k] || !f1[k]) return fail_msg("vlbm2d_batch_create: null handle or populations");
        const vlbm2d_desc &d = h->d;
        if (d.precision != prec) return fail_msg("vlbm2d_batch_create: every shell needs the same precision");
        if (d.box) return fail_msg("vlbm2d_batch_create: box solvers are not batched");
        if (h->P.gw != 0.0 || h->P.filt > 0.0)
            return fail_msg("vlbm2d_batch_create: batched shells use BGK without the kinetic filter");
        if (d.disp_interval > 1 || d.compat_interval > 1 || d.disp_sweeps != 1)
            return fail_msg("vlbm2d_batch_create: batched shells need a uniform schedule (intervals 0 or 1, one sweep)");
        ext = ext || recon_ext(d.recon);
        nh = nh && d.model == M_NH;
        for (int j = 0; j < k; ++j)
            if (handles[j] == h) return fail_msg("vlbm2d_batch_create: a shell appears twice");
    }
    vlbm2d_batch *b = new vlbm2d_batch;
    b->precision = prec;
    b->ext = ext;
    b->nh = nh;
    b->stream = (hipStream_t)stream;
    for (int k = 0; k < n; ++k) {
        vlbm2d_handle *h = handles[k];
        b->all.push_back(h);
        b->f0.push_back(f0[k]);
        b->f1.push_back(f1[k]);
        b->active.push_back(k);
        if (!h->upmask_ok) {
            hipLaunchKernelGGL(k2_upmask, dim3((h->d.n_active + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, b->stream,
                               h->d.active_list, h->d.active, h->upmask, h->P);
            hipLaunchKernelGGL(k2_nbmask, dim3((h->d.n_active + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, b->stream,
                               h->d.active_list, h->d.active, h->nbmask, h->P);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) {
                delete b;
                return fail("vlbm2d_batch_create: masks", e);
            }
            h->upmask_ok = true;
        }
    }
    for (int k = 0; k < n; ++k) {
        hipError_t e = hipStreamSynchronize(b->all[k]->stream);
        if (e != hipSuccess) {
            delete b;
            return fail("vlbm2d_batch_create: sync", e);
        }
    }
    const int rc = batch_tables(b);
    if (rc) {
        b->release();
        delete b;
        return rc;
    }
    *out = b;
    return 0;
}

int vlbm2d_batch_destroy(vlbm2d_batch *b) {
    if (!b) return 0;
    (void)hipStreamSynchronize(b->stream);
    b->release();
    delete b;
    return 0;
}

int vlbm2d_batch_select(vlbm2d_batch *b, const int32_t *ids, void *const *f0, void *const *f1, int32_t count) {
    if (!b || (count > 0 && (!ids || !f0 || !f1))) return fail_msg("vlbm2d_batch_select: null argument");
    if (count < 0 || count > (int)b->all.size()) return fail_msg("vlbm2d_batch_select: count out of range");
    std::vector<int> act;
    for (int k = 0; k < count; ++k) {
        if (ids[k] < 0 || ids[k] >= (int)b->all.size()) return fail_msg("vlbm2d_batch_select: shell id out of range");
        if (k > 0 && ids[k] <= ids[k - 1]) return fail_msg("vlbm2d_batch_select: ids must be strictly increasing");
        if (!f0[k] || !f1[k]) return fail_msg("vlbm2d_batch_select: null populations");
        act.push_back(ids[k]);
    }
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail("vlbm2d_batch_select: sync", e);
    for (int k = 0; k < count; ++k) {
        b->f0[ids[k]] = f0[k];
        b->f1[ids[k]] = f1[k];
    }
    b->active = act;
    b->parity = 0;
    return batch_tables(b);
}

int vlbm2d_batch_set_graph_steps(vlbm2d_batch *b, int32_t steps) {
    if (!b) return fail_msg("vlbm2d_batch_set_graph_steps: null batch");
    if (steps < 0 || steps > 4096) return fail_msg("vlbm2d_batch_set_graph_steps: steps must be in [0, 4096]");
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail("vlbm2d_batch_set_graph_steps: sync", e);
    b->drop_graphs();
    b->graph_steps = steps;
    return 0;
}

int vlbm2d_batch_parity(const vlbm2d_batch *b) { return b ? b->parity : -1; }

int vlbm2d_batch_step(vlbm2d_batch *b, const double *times, int32_t nsteps) {
    if (!b || (b->njobs > 0 && !times)) return fail_msg("vlbm2d_batch_step: null argument");
    if (nsteps < 0) return fail_msg("vlbm2d_batch_step: nsteps must be non-negative");
    if (b->njobs == 0 || nsteps == 0) return 0;
    hipError_t e = hipMemcpyAsync(b->clocks, times, sizeof(double) * (size_t)b->njobs, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail("vlbm2d_batch_step: clocks", e);
    if (b->precision == VLBM_PREC_F64) return batch_run<double>(b, nsteps);
    return batch_run<float>(b, nsteps);
}

int vlbm2d_batch_autotune(vlbm2d_batch *b, int32_t steps, int32_t *nb_out, int32_t *nl_out, float *ms_out) {
    if (!b || !nb_out || !nl_out) return fail_msg("vlbm2d_batch_autotune: null argument");
    if (steps < 2 || steps > 1000) return fail_msg("vlbm2d_batch_autotune: steps must be in [2, 1000]");
    if (b->njobs == 0) return fail_msg("vlbm2d_batch_autotune: no shell selected");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) return fail("vlbm2d_batch_autotune: events", e);
    std::vector<double> zeros((size_t)b->njobs, 0.0);
    float best = -1.0f;
    int bnb = b->nb, bnl = b->nl, rc = 0;
    const int keep_graph = b->graph_steps;
    b->graph_steps = 0;
    for (int i = 0; i < 4 && rc == 0; ++i)
        for (int k = 0; k < 3 && rc == 0; ++k) {
            b->nb = TUNE_NB[i];
            b->nl = TUNE_NL[k];
            rc = batch_tables(b);
            if (rc) break;
            e = hipMemcpy(b->clocks, zeros.data(), sizeof(double) * zeros.size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                rc = fail("vlbm2d_batch_autotune: clocks", e);
                break;
            }
            rc = b->precision == VLBM_PREC_F64 ? batch_run<double>(b, 1) : batch_run<float>(b, 1);
            if (rc) break;
            (void)hipEventRecord(e0, b->stream);
            rc = b->precision == VLBM_PREC_F64 ? batch_run<double>(b, steps) : batch_run<float>(b, steps);
            (void)hipEventRecord(e1, b->stream);
            if (rc) break;
            e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) {
                rc = fail("vlbm2d_batch_autotune: timing", e);
                break;
            }
            if (ms_out) ms_out[3 * i + k] = ms / (float)steps;
            if (best < 0.0f || ms < best) {
                best = ms;
                bnb = b->nb;
                bnl = b->nl;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->graph_steps = keep_graph;
    if (rc) return rc;
    b->nb = bnb;
    b->nl = bnl;
    *nb_out = bnb;
    *nl_out = bnl;
    return batch_tables(b);
}

int vlbm2d_batch_set_blocks(vlbm2d_batch *b, int32_t nb, int32_t nl) {
    if (!b) return fail_msg("vlbm2d_batch_set_blocks: null batch");
    if ((nb != 64 && nb != 128 && nb != 256 && nb != 512) || (nl != 32 && nl != 64 && nl != 128))
        return fail_msg("vlbm2d_batch_set_blocks: node blocks must be 64, 128, 256 or 512 and link blocks 32, 64 or 128");
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail("vlbm2d_batch_set_blocks: sync", e);
    b->nb = nb;
    b->nl = nl;
    return batch_tables(b);
}

}  // extern "C"
