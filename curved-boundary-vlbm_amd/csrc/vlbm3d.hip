// libvlbm: D3Q6x12 vectorial LBM time step for MI355X (gfx950).
//
// Storage is the active-node list of the reference (cbw3d:3683-3722): every
// field is a (components, n_active) plane, nodes in lexicographic (i, j, k)
// order with k fastest, so the +-z neighbours of a node are adjacent in memory
// and +-y / +-x neighbours sit one row / one plane away.
//
// One step (cbw3d:4549-4589) runs as
//   k_collide   moments of f0 (U, u, u*), J flag                      per node
//   k_sweep     local displacement sweep into u_next                  per node   [if due]
//   k_commit    intermediate sweep commits                            per node   [sweeps > 1]
//   k_record    F repair + deferred-collision record + final commit   per node
//   k_gather    characteristic state of the traction links            per (link, comp)
//   k_prepare   boundary values + normal-traction Newton              per link
//   k_stream    pull stream: f1[q,p] = fpost(f0, record)[q, p - e_q]  per node
//   k_links     BFL + characteristic populations of cut links         per link
// Post-collision populations are never materialised: k_record stores 24
// doubles per node from which fpost[q][a] = (1 - omega) f[q][a] + A[a] +
// s_q B[axis(q)][a] follows exactly, and k_stream / k_links apply it while
// pulling.  Populations therefore cross HBM once per step in each direction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/vlbm3d.h"
#include "vlbm3d_device.h"

static_assert(vlbm::MAX_IDS == VLBM3D_MAX_IDS && vlbm::NPRM == VLBM3D_NPRM,
              "boundary-id table shape must match the C ABI");
#include "vlbm_common.h"

using namespace vlbm;
using vlbm_common::BLOCK;
using vlbm_common::fail;
using vlbm_common::fail_msg;
using vlbm_common::Timing;
using vlbm_common::TimedSlot;
using vlbm_common::xcd_block;
using vlbm_common::xcd_grid;

std::string &vlbm_common::last_error() {
    static thread_local std::string err;
    return err;
}

namespace {

constexpr int RED_BLOCKS = 1024;

// Record layout (24 planes): A_v[3], A_F[9], B_v[r*3+axis] (9), B_F[3].
enum { R_AV = 0, R_AF = 3, R_BV = 12, R_BF = 21, RECORD_PLANES = 24 };

struct Params {
    int n, L;
    Material m;
    double dt, c, omega, dx, dy, dz;
    int src_mode;
    double src[12];
    int recon, solver;
    double disp_relax, w2;
    int disp_bid;
    double blend;
    const double *clock;  // device time of the step start (captured step loops), or nullptr
    // Lattice speed along each axis (all c except for an anisotropic box,
    // vnew3d:1240-1367) and cr[a] = c / cs[a], the factor that turns the
    // record's F-flux planes (kept with 1/c) into the axis's 1/cs[a]; cr is
    // exactly 1 for equal speeds, so isotropic results do not change.
    double cs[3], cr[3];
};

// Boundary time of a step: the host value, or clock + dt/2 when the launch
// belongs to a captured step loop that keeps the time on the device.
__device__ __forceinline__ double step_time(double tb, const Params &P) {
    return P.clock ? P.clock[0] + 0.5 * P.dt : tb;
}

// Element offsets of population and record planes in 16-byte groups: FP64
// planes paired, pair (plane / 2) * n + node holding planes 2k and 2k + 1
// side by side; FP32 planes in fours (below).  The twelve values of one
// direction at a node (and the record's A planes) are six FP64 pairs or three
// FP32 quads, so the pull stream moves 16 bytes per lane and access (single
// floats stream at about two thirds of the rate of pairs, single doubles at
// 0.93 of it; FP32 quads took the FP32 pull from 146 to 137 us).  The
// reference's (72, n) view of the populations exists only through the
// conversion kernels.
template <typename T>
__device__ __forceinline__ size_t pofs(const T *, int plane, size_t node, size_t n) {
    return (((size_t)(plane >> 1) * n + node) << 1) + (size_t)(plane & 1);
}
// FP32 planes in fours: quad (plane / 4) * n + node holds planes 4k .. 4k + 3
// (16 bytes, like an FP64 pair); each even/odd pair stays adjacent inside it.
__device__ __forceinline__ size_t pofs(const float *, int plane, size_t node, size_t n) {
    return (((size_t)(plane >> 2) * n + node) << 2) + (size_t)(plane & 3);
}
typedef float float_pair __attribute__((ext_vector_type(2)));
typedef float float_quad __attribute__((ext_vector_type(4)));
typedef double double_pair __attribute__((ext_vector_type(2)));

// Population access.  double planes hold f; float planes hold f - f_rest
// where f_rest[q][a] = 1/6 on the diagonal of F (a = 3, 7, 11), 0 elsewhere.
__device__ __forceinline__ double rest(int a) { return (a == 3 || a == 7 || a == 11) ? (1.0 / 6.0) : 0.0; }
__device__ __forceinline__ double ld(const double *f, size_t i, int) { return f[i]; }
__device__ __forceinline__ double ld(const float *f, size_t i, int a) { return (double)f[i] + rest(a); }
__device__ __forceinline__ void st(double *f, size_t i, int, double v) { f[i] = v; }
__device__ __forceinline__ void st(float *f, size_t i, int a, double v) { f[i] = (float)(v - rest(a)); }

// Streaming variants: populations are touched once per step, so they bypass
// the caches' retention and leave L2 to the records and neighbour tables.
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
// Planes (a, a + 1), a even, at one node: one 8-byte access for float pairs.
__device__ __forceinline__ void ld2_nt(const double *f, int plane, size_t node, size_t n, int, double &x0,
                                       double &x1) {
    const double_pair v = __builtin_nontemporal_load((const double_pair *)(f + pofs(f, plane, node, n)));
    x0 = v.x;
    x1 = v.y;
}
__device__ __forceinline__ void ld2_nt(const float *f, int plane, size_t node, size_t n, int a, double &x0,
                                       double &x1) {
    const float_pair v = __builtin_nontemporal_load((const float_pair *)(f + pofs(f, plane, node, n)));
    x0 = (double)v.x + rest(a);
    x1 = (double)v.y + rest(a + 1);
}
// Cached pair accesses for the scattered link kernel: planes (a, a + 1),
// a even, at one node, and a pair of entries at element i (i even).
__device__ __forceinline__ void ld2(const double *f, int plane, size_t node, size_t n, int, double &x0, double &x1) {
    const double_pair v = *(const double_pair *)(f + pofs(f, plane, node, n));
    x0 = v.x;
    x1 = v.y;
}
__device__ __forceinline__ void ld2(const float *f, int plane, size_t node, size_t n, int a, double &x0, double &x1) {
    const float_pair v = *(const float_pair *)(f + pofs(f, plane, node, n));
    x0 = (double)v.x + rest(a);
    x1 = (double)v.y + rest(a + 1);
}
__device__ __forceinline__ void st2(double *f, size_t i, int, double v0, double v1) {
    double_pair x;
    x.x = v0;
    x.y = v1;
    *(double_pair *)(f + i) = x;
}
__device__ __forceinline__ void st2(float *f, size_t i, int a, double v0, double v1) {
    float_pair x;
    x.x = (float)(v0 - rest(a));
    x.y = (float)(v1 - rest(a + 1));
    *(float_pair *)(f + i) = x;
}
// Store planes (a, a + 1), a even; returns the values as stored.
__device__ __forceinline__ void st2_nt(double *f, int plane, size_t node, size_t n, int, double &v0, double &v1) {
    double_pair x;
    x.x = v0;
    x.y = v1;
    __builtin_nontemporal_store(x, (double_pair *)(f + pofs(f, plane, node, n)));
}
__device__ __forceinline__ void st2_nt(float *f, int plane, size_t node, size_t n, int a, double &v0, double &v1) {
    float_pair x;
    x.x = (float)(v0 - rest(a));
    x.y = (float)(v1 - rest(a + 1));
    __builtin_nontemporal_store(x, (float_pair *)(f + pofs(f, plane, node, n)));
    v0 = (double)x.x + rest(a);
    v1 = (double)x.y + rest(a + 1);
}

// Records are stored in the precision of the populations.  double records
// hold the values themselves; float records hold A - omega f_rest on the A
// planes (rest state subtracted, as for float populations), so that their
// rounding error scales with the deviation from rest.
__device__ __forceinline__ double rld(const double *r, size_t i, int, double) { return r[i]; }
__device__ __forceinline__ double rld(const float *r, size_t i, int a, double w) {
    return (double)r[i] + w * rest(a);
}
__device__ __forceinline__ double rld_b(const double *r, size_t i) { return r[i]; }
__device__ __forceinline__ double rld_b(const float *r, size_t i) { return (double)r[i]; }
// Records are written once per step and read by other kernels, so the
// stores are non-temporal: they do not displace the neighbour data that the
// writing kernel itself reuses from L2.
__device__ __forceinline__ void rst_a(double *r, size_t i, int, double, double v) {
    __builtin_nontemporal_store(v, r + i);
}
__device__ __forceinline__ void rst_a(float *r, size_t i, int a, double w, double v) {
    __builtin_nontemporal_store((float)(v - w * rest(a)), r + i);
}
__device__ __forceinline__ void rst_b(double *r, size_t i, double v) { __builtin_nontemporal_store(v, r + i); }
__device__ __forceinline__ void rst_b(float *r, size_t i, double v) { __builtin_nontemporal_store((float)v, r + i); }
// Record planes (k, k + 1), k even, at one node; `fa` for each the A-plane
// index whose rest part the float record subtracts, or -1 for a B plane.
__device__ __forceinline__ void rst2(double *r, int k, size_t node, size_t n, int, int, double, double v0,
                                     double v1) {
    double_pair x;
    x.x = v0;
    x.y = v1;
    __builtin_nontemporal_store(x, (double_pair *)(r + pofs(r, k, node, n)));
}
__device__ __forceinline__ void rst2(float *r, int k, size_t node, size_t n, int fa0, int fa1, double w, double v0,
                                     double v1) {
    float_pair x;
    x.x = (float)(fa0 >= 0 ? v0 - w * rest(fa0) : v0);
    x.y = (float)(fa1 >= 0 ? v1 - w * rest(fa1) : v1);
    __builtin_nontemporal_store(x, (float_pair *)(r + pofs(r, k, node, n)));
}
// A planes (a, a + 1), a even, of the record at one node.
__device__ __forceinline__ void rld2(const double *r, int a, size_t node, size_t n, double, double &x0, double &x1) {
    const double_pair v = *(const double_pair *)(r + pofs(r, a, node, n));
    x0 = v.x;
    x1 = v.y;
}
__device__ __forceinline__ void rld2(const float *r, int a, size_t node, size_t n, double w, double &x0, double &x1) {
    const float_pair v = *(const float_pair *)(r + pofs(r, a, node, n));
    x0 = (double)v.x + w * rest(a);
    x1 = (double)v.y + w * rest(a + 1);
}
// Per-node state touched once by a kernel (no neighbour reuse).
__device__ __forceinline__ double nt_ld(const double *a, size_t i) { return __builtin_nontemporal_load(a + i); }
__device__ __forceinline__ void nt_st(double *a, size_t i, double v) { __builtin_nontemporal_store(v, a + i); }

// fpost[q][a] at `node` from its pre-collision value fv and the node's
// record; fs = Params::cr[q % 3] scales the F-flux term to the speed of q's axis.
template <typename R>
__device__ __forceinline__ double fpost_v(double fv, const R *rec, size_t n, int q, int a, int node, double om1,
                                          double w, double fs) {
    const int axis = q % 3;
    const double s = q < 3 ? 1.0 : -1.0;
    double b = 0.0;
    if (a < 3) {
        b = rld_b(rec, pofs(rec, R_BV + 3 * a + axis, node, n));
    } else if ((a - 3) % 3 == axis) {
        b = rld_b(rec, pofs(rec, R_BF + (a - 3) / 3, node, n)) * fs;
    }
    return om1 * fv + rld(rec, pofs(rec, a, node, n), a, w) + s * b;
}

// fpost_v with the record's A value already loaded (Av).
template <typename R>
__device__ __forceinline__ double fpost_a(double fv, double Av, const R *rec, size_t n, int q, int a, int node,
                                          double om1, double fs) {
    const int axis = q % 3;
    const double s = q < 3 ? 1.0 : -1.0;
    double b = 0.0;
    if (a < 3) {
        b = rld_b(rec, pofs(rec, R_BV + 3 * a + axis, node, n));
    } else if ((a - 3) % 3 == axis) {
        b = rld_b(rec, pofs(rec, R_BF + (a - 3) / 3, node, n)) * fs;
    }
    return om1 * fv + Av + s * b;
}

// fpost_a with the record's flux entries of q's axis already loaded:
// B[r] = B_V[3 r + axis(q)], B[3 + r] = B_F[r].
__device__ __forceinline__ double fpost_b(double fv, double Av, const double *B, int q, int a, double om1,
                                          double fs) {
    const int axis = q % 3;
    const double s = q < 3 ? 1.0 : -1.0;
    double b = 0.0;
    if (a < 3) {
        b = B[a];
    } else if ((a - 3) % 3 == axis) {
        b = B[3 + (a - 3) / 3] * fs;
    }
    return om1 * fv + Av + s * b;
}

// fpost[q][a] at node m from pre-collision f and the node's record.
template <typename T>
__device__ __forceinline__ double fpost(const T *f, const T *rec, size_t n, int q, int a, int node, double om1,
                                        double w, double fs) {
    return fpost_v(ld(f, pofs(f, q * 12 + a, node, n), a), rec, n, q, a, node, om1, w, fs);
}

__global__ void k_reset_invalid(int32_t *invalid) { invalid[0] = 0; }
__global__ void k_set_clock(double *clock, double t) { clock[0] = t; }
// time += dt, the same FP64 operation the host performs after each step.
__global__ void k_advance_clock(double *clock, double dt) { clock[0] = clock[0] + dt; }

// Moments with the half-step source; u = u* + dt/2 v and (collide only)
// u* = u + dt/2 v.  cbw3d:2110-2189 and :2193-2280.
template <typename T, bool COLLIDE>
__global__ __launch_bounds__(BLOCK) void k_collide(const T *__restrict__ f, double *__restrict__ U,
                                                   double *__restrict__ u, double *__restrict__ us,
                                                   int32_t *__restrict__ invalid, Params P) {
    const int p = xcd_block() * BLOCK + threadIdx.x;
    if (p >= P.n) return;
    const size_t n = P.n;
    double Um[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) Um[a] = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int a = 0; a < NC; ++a) Um[a] += ld(f, pofs(f, q * 12 + a, p, n), a);
    const double hdt = 0.5 * P.dt;
    if (P.src_mode == SRC_CONST) {
#pragma unroll
        for (int a = 0; a < NC; ++a) Um[a] += hdt * P.src[a];
    } else if (P.src_mode == SRC_DAMP) {
        const double denom = 1.0 + hdt * P.src[0];
        Um[0] = Um[0] / denom;
        Um[1] = Um[1] / denom;
        Um[2] = Um[2] / denom;
    }
#pragma unroll
    for (int a = 0; a < NC; ++a) U[a * n + p] = Um[a];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double uk = us[k * n + p] + hdt * Um[k];
        u[k * n + p] = uk;
        if (COLLIDE) us[k * n + p] = uk + hdt * Um[k];
    }
    if (det3(Um + 3) <= P.m.jf) atomicOr(invalid, 1);
}

// One relaxed nearest-neighbour compatibility sweep of u (cbw3d:2473-2527)
// with the Dirichlet anchors of the node's own cut links gathered in link
// order (cbw3d:2413-2469), so the sums are deterministic.  The anchors come
// from a compacted per-node list built once at set-up (only links that
// contribute a component; per entry the boundary id with a component mask,
// the cut point and its offset from the node), so an anchored node loads its
// entries directly instead of walking link -> id -> coordinates.
__global__ __launch_bounds__(BLOCK) void k_sweep(
    const double *__restrict__ U, const double *__restrict__ u, double *__restrict__ un,
    const int32_t *__restrict__ nbr, const int32_t *__restrict__ aoff, const int32_t *__restrict__ akey,
    const double *__restrict__ adat, const double *__restrict__ prm_by_id, double time, Params P) {
    const int p = xcd_block() * BLOCK + threadIdx.x;
    if (p >= P.n) return;
    const size_t n = P.n;
    double d[3] = {0.0, 0.0, 0.0}, r[3] = {0.0, 0.0, 0.0};
    const int l0 = aoff[p], l1 = aoff[p + 1];
    if (l1 > l0) {
        time = step_time(time, P);
        double F[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = U[(3 + k) * n + p];
        for (int t = l0; t < l1; ++t) {
            const int key = akey[t];
            const int bid = key & 0xff, mask = (key >> 8) & 7, mode = key >> 12;
            const double *e = adat + (size_t)6 * t;
            const double x = e[0], y = e[1], z = e[2];
            const double db[3] = {e[3], e[4], e[5]};
            const double *prm = prm_by_id + bid * NPRM;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (!((mask >> c) & 1)) continue;
                const double ub = dirichlet_disp(mode, prm, c, x, y, z, time);
                const double g = (F[3 * c] - (c == 0 ? 1.0 : 0.0)) * db[0] +
                                 (F[3 * c + 1] - (c == 1 ? 1.0 : 0.0)) * db[1] +
                                 (F[3 * c + 2] - (c == 2 ? 1.0 : 0.0)) * db[2];
                d[c] += P.w2;
                r[c] += P.w2 * (ub - g);
            }
        }
    }
    double Fp[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Fp[k] = U[(3 + k) * n + p];
    // Only the column of F along the link axis enters: the reference's
    // off-axis terms are multiplied by a zero spacing and add exact zeros.
    // The neighbours' values are loaded from an index that falls back to p
    // and a missing neighbour's terms are dropped by selects, so no load
    // waits behind a branch.
    int pn[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) pn[q] = nbr[q * n + p];
    double Fq[NQ][3], uq[NQ][3];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const size_t src = pn[q] >= 0 ? pn[q] : p;
        const int j = q % 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Fq[q][c] = U[(3 + 3 * c + j) * n + src];
            uq[q][c] = u[c * n + src];
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const bool live = pn[q] >= 0;
        const int j = q % 3;
        const double h = (q < 3 ? 1.0 : -1.0) * (j == 0 ? P.dx : (j == 1 ? P.dy : P.dz));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double s = Fp[3 * c + j] + Fq[q][c];
            if (j == c) s -= 2.0;
            const double g = 0.5 * (s * h);
            r[c] = live ? r[c] + (uq[q][c] - g) : r[c];
            d[c] = live ? d[c] + 1.0 : d[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double uc = u[c * n + p];
        nt_st(un, (size_t)c * n + p, d[c] > 0.0 ? uc + P.disp_relax * (r[c] / d[c] - uc) : uc);
    }
}

__global__ __launch_bounds__(BLOCK) void k_commit(const double *__restrict__ un, double *__restrict__ u,
                                                  double *__restrict__ us, int n) {
    const int p = xcd_block() * BLOCK + threadIdx.x;
    if (p >= n) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t i = (size_t)c * n + p;
        const double delta = un[i] - u[i];
        u[i] = un[i];
        us[i] = us[i] + delta;
    }
}

// Centred / second-order one-sided / first-order derivative of u along one
// axis (cbw3d:2544-2571).  pp/pm are the neighbours along +axis/-axis and
// pp2/pm2 the next ones (second-neighbour table).  Every value any case
// might use is loaded from an index that falls back to p, and the case only
// selects the numerator and the spacing of the one division: with no branch
// around a load, the loads of all nine derivatives issue together.
__device__ __forceinline__ double du_axis(const double *uc, int p, int pp, int pm, int pp2, int pm2, double h) {
    const bool fwd = pp >= 0 && pm < 0 && pp2 >= 0, bwd = pm >= 0 && pp < 0 && pm2 >= 0;
    const double u0 = uc[p], up = uc[pp >= 0 ? pp : p], um = uc[pm >= 0 ? pm : p];
    const double up2 = uc[fwd ? pp2 : p], um2 = uc[bwd ? pm2 : p];
    const double central = up - um, fwd2 = -3.0 * u0 + 4.0 * up - up2, fwd1 = up - u0;
    const double bwd2 = 3.0 * u0 - 4.0 * um + um2, bwd1 = u0 - um;
    double num = 0.0, den = 1.0;
    if (pp >= 0 && pm >= 0) {
        num = central;
        den = 2.0 * h;
    } else if (pp >= 0) {
        num = pp2 >= 0 ? fwd2 : fwd1;
        den = pp2 >= 0 ? 2.0 * h : h;
    } else if (pm >= 0) {
        num = pm2 >= 0 ? bwd2 : bwd1;
        den = pm2 >= 0 ? 2.0 * h : h;
    }
    return num / den;
}

// Neighbour table of the stream kernel: the neighbour table with each cut
// link's missing entry (node link_p, direction link_q) replaced by
// -(ell + 2), so the fused collide finds the link's incoming population in
// the link-major copy that k_links writes (one 96-byte record instead of
// twelve population planes); -1 stays a missing neighbour without a link.
__global__ __launch_bounds__(BLOCK) void k_nbr_pull(const int32_t *__restrict__ nbr, const int32_t *__restrict__ lp,
                                                    const int32_t *__restrict__ lq, int32_t *__restrict__ out,
                                                    int32_t *__restrict__ bad, int n, int L) {
    const int ell = blockIdx.x * BLOCK + threadIdx.x;
    if (ell >= L) return;
    const size_t i = (size_t)lq[ell] * n + lp[ell];
    if (nbr[i] >= 0)
        atomicOr(bad, 1);
    else
        out[i] = -(ell + 2);
}

// Second neighbours: nbr2[q][p] = nbr[q][nbr[q][p]] (or -1).
__global__ __launch_bounds__(BLOCK) void k_nbr2(const int32_t *__restrict__ nbr, int32_t *__restrict__ nbr2,
                                                int n) {
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n) return;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int a = nbr[(size_t)q * n + p];
        nbr2[(size_t)q * n + p] = a >= 0 ? nbr[(size_t)q * n + a] : -1;
    }
}

// Deformation-gradient update F <- F + theta (I + grad u - F) with the J
// guard (cbw3d:2575-2672), folded with the BGK collision of cbw3d:2281-2316
// into the per-node record, plus the last displacement commit.
// The post-repair state goes to the U planes when write_U is set and, for
// the characteristic sample nodes, straight into the gather's packed records.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_record(double *__restrict__ U, const double *__restrict__ ugrad,
                                                  double *__restrict__ u, double *__restrict__ us,
                                                  R *__restrict__ rec, const int32_t *__restrict__ nbr,
                                                  const int32_t *__restrict__ nbr2,
                                                  const int32_t *__restrict__ pack_slot, double *__restrict__ pack,
                                                  int repair, int commit, int write_U,
                                                  int32_t *__restrict__ reset_invalid, Params P) {
    const int p = xcd_block() * BLOCK + threadIdx.x;
    if (p >= P.n) return;
    const size_t n = P.n;
    // Every load comes before the first store: the vector-memory counter
    // drains in order over loads and stores, so a load behind a store would
    // wait for the store.
    double v[3], Fo[9];
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = U[k * n + p];
#pragma unroll
    for (int k = 0; k < 9; ++k) Fo[k] = U[(3 + k) * n + p];
    const int sl = pack_slot ? pack_slot[p] : -1;
    double un[3] = {0.0, 0.0, 0.0}, uo[3] = {0.0, 0.0, 0.0}, uso[3] = {0.0, 0.0, 0.0};
    if (commit)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t i = (size_t)c * n + p;
            un[c] = ugrad[i];
            uo[c] = nt_ld(u, i);
            uso[c] = nt_ld(us, i);
        }
    double Po[9], dP[9], dF[9], Fnew[9];
    bool repaired = false;
    piola(Fo, P.m, Po);
#pragma unroll
    for (int k = 0; k < 9; ++k) Fnew[k] = Fo[k];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        dP[k] = 0.0;
        dF[k] = 0.0;
    }
    if (repair) {
        double beta = P.blend;
        beta = beta < 0.0 ? 0.0 : (beta > 1.0 ? 1.0 : beta);
        const double hh[3] = {P.dx, P.dy, P.dz};
        int nb[NQ], nb2[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            nb[q] = nbr[q * n + p];
            nb2[q] = nbr2[q * n + p];
        }
        double Fn[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double t =
                    (i == j ? 1.0 : 0.0) + du_axis(ugrad + i * n, p, nb[j], nb[j + 3], nb2[j], nb2[j + 3], hh[j]);
                Fn[3 * i + j] = Fo[3 * i + j] + beta * (t - Fo[3 * i + j]);
            }
        if (det3(Fn) > P.m.jf) {
            double Pn[9];
            piola(Fn, P.m, Pn);
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                dP[k] = Pn[k] - Po[k];
                dF[k] = Fn[k] - Fo[k];
                Fnew[k] = Fn[k];
            }
            repaired = true;
        }
    }
    if (repaired && write_U)
#pragma unroll
        for (int k = 0; k < 9; ++k) nt_st(U, (size_t)(3 + k) * n + p, Fnew[k]);
    if (sl >= 0) {
        double *pr = pack + (size_t)sl * NC;
#pragma unroll
        for (int k = 0; k < 3; ++k) pr[k] = v[k];
#pragma unroll
        for (int k = 0; k < 9; ++k) pr[3 + k] = Fnew[k];
    }
    if (commit) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t i = (size_t)c * n + p;
            const double delta = un[c] - uo[c];
            nt_st(u, i, un[c]);
            nt_st(us, i, uso[c] + delta);
        }
    }
    const double w = P.omega;
    const double st = P.dt * (2.0 - w) * (1.0 / 12.0);
    const double w6 = w / 6.0, w2c = w / (2.0 * P.c);
    double rv[RECORD_PLANES];
#pragma unroll
    for (int a = 0; a < 3; ++a) rv[R_AV + a] = w6 * v[a] + st * source_comp(a, P.src_mode, P.src, v[0], v[1], v[2]);
#pragma unroll
    for (int k = 0; k < 9; ++k)
        rv[R_AF + k] = w6 * Fo[k] + st * source_comp(3 + k, P.src_mode, P.src, v[0], v[1], v[2]) + dF[k] / 6.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const double ca = P.cs[k % 3];  // plane 3 r + axis: the speed of its axis
        rv[R_BV + k] = -(w / (2.0 * ca)) * Po[k] - (0.5 / ca) * dP[k];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) rv[R_BF + i] = -w2c * v[i];
#pragma unroll
    for (int k = 0; k < RECORD_PLANES; k += 2)
        rst2(rec, k, p, n, k < R_BV ? k : -1, k + 1 < R_BV ? k + 1 : -1, w, rv[k], rv[k + 1]);
    // Clears the invalid flag ahead of the stream kernel's fused collide.
    if (reset_invalid && p == 0) reset_invalid[0] = 0;
}

// Directions per load group of the pull: 3 (214 VGPRs FP64, two waves per
// SIMD) measured fastest; 1, 2 and 6 are 2-5% slower.
constexpr int STREAM_GROUP = 3;

// Pull stream of the post-collision populations; the entries whose upstream
// node is missing belong to cut links and were written by k_links.  With
// NEXT the kernel also performs the moment part of the next step's collide
// (cbw3d:2213-2280): U = moments of the new populations with the half-step
// source, u = u* + dt/2 v, u* = u + dt/2 v, and the J check.
template <typename T, bool NEXT>
__global__ __launch_bounds__(BLOCK) void k_stream(const T *__restrict__ f0, T *__restrict__ f1,
                                                  const T *__restrict__ rec, const T *__restrict__ flink,
                                                  const int32_t *__restrict__ nbr, double *__restrict__ U,
                                                  double *__restrict__ u, double *__restrict__ us,
                                                  int32_t *__restrict__ invalid, Params P) {
    const int p = xcd_block() * BLOCK + threadIdx.x;
    if (p >= P.n) return;
    const size_t n = P.n;
    const double om1 = 1.0 - P.omega;
    double Um[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) Um[a] = 0.0;
    // All six upstream indices first, then the directions in groups whose
    // loads are all issued before the group's stores: the vector-memory
    // counter drains in order over loads and stores alike, so a load issued
    // after a store cannot be waited for without waiting for the store.
    int srcs[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) srcs[q] = nbr[opp(q) * n + p];
    double uso[3] = {0.0, 0.0, 0.0};
    if (NEXT)
#pragma unroll
        for (int k = 0; k < 3; ++k) uso[k] = nt_ld(us, (size_t)k * n + p);
#pragma unroll
    for (int q0 = 0; q0 < NQ; q0 += STREAM_GROUP) {
        // components in pairs (a, a + 1): one 8-byte access each for float
        // planes; a missing upstream node reads node p instead, unused
        double fv[STREAM_GROUP][NC], A[STREAM_GROUP][NC], B[STREAM_GROUP][6];
#pragma unroll
        for (int g = 0; g < STREAM_GROUP; ++g) {
            const int q = q0 + g, axis = q % 3;
            const size_t src = srcs[q] >= 0 ? (size_t)srcs[q] : (size_t)p;
            if constexpr (sizeof(T) == 4) {
                // FP32: the direction's twelve values and the record's A planes as three quads each
#pragma unroll
                for (int h = 0; h < NC / 4; ++h) {
                    const float_quad v = __builtin_nontemporal_load(
                        (const float_quad *)((const float *)f0 + pofs((const float *)f0, q * 12 + 4 * h, src, n)));
                    const float_quad r = *(const float_quad *)((const float *)rec +
                                                               pofs((const float *)rec, 4 * h, src, n));
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        fv[g][4 * h + e] = (double)v[e] + rest(4 * h + e);
                        A[g][4 * h + e] = (double)r[e] + P.omega * rest(4 * h + e);
                    }
                }
            } else {
#pragma unroll
                for (int h = 0; h < NC / 2; ++h)
                    ld2_nt(f0, q * 12 + 2 * h, src, n, 2 * h, fv[g][2 * h], fv[g][2 * h + 1]);
#pragma unroll
                for (int h = 0; h < NC / 2; ++h) rld2(rec, 2 * h, src, n, P.omega, A[g][2 * h], A[g][2 * h + 1]);
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                B[g][r] = rld_b(rec, pofs(rec, R_BV + 3 * r + axis, src, n));
                B[g][3 + r] = rld_b(rec, pofs(rec, R_BF + r, src, n));
            }
        }
#pragma unroll
        for (int g = 0; g < STREAM_GROUP; ++g) {
            const int q = q0 + g, src = srcs[q];
            if (src >= 0) {
#pragma unroll
                for (int h = 0; h < NC / 2; ++h) {
                    double x0 = fpost_b(fv[g][2 * h], A[g][2 * h], B[g], q, 2 * h, om1, P.cr[q % 3]);
                    double x1 = fpost_b(fv[g][2 * h + 1], A[g][2 * h + 1], B[g], q, 2 * h + 1, om1, P.cr[q % 3]);
                    st2_nt(f1, q * 12 + 2 * h, p, n, 2 * h, x0, x1);
                    if (NEXT) {
                        Um[2 * h] += x0;
                        Um[2 * h + 1] += x1;
                    }
                }
            } else if (NEXT) {
                // the cut link's population, written by k_links: from its
                // link-major copy, else (box faces) from the population planes
                if (src <= -2) {
                    const T *fl = flink + (size_t)(-2 - src) * NC;
#pragma unroll
                    for (int a = 0; a < NC; ++a) Um[a] += ld(fl, (size_t)a, a);
                } else {
#pragma unroll
                    for (int a = 0; a < NC; ++a) Um[a] += ld(f1, pofs(f1, q * 12 + a, p, n), a);
                }
            }
        }
    }
    if (NEXT) {
        const double hdt = 0.5 * P.dt;
        if (P.src_mode == SRC_CONST) {
#pragma unroll
            for (int a = 0; a < NC; ++a) Um[a] += hdt * P.src[a];
        } else if (P.src_mode == SRC_DAMP) {
            const double denom = 1.0 + hdt * P.src[0];
            Um[0] = Um[0] / denom;
            Um[1] = Um[1] / denom;
            Um[2] = Um[2] / denom;
        }
#pragma unroll
        for (int a = 0; a < NC; ++a) nt_st(U, (size_t)a * n + p, Um[a]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double uk = uso[k] + hdt * Um[k];
            nt_st(u, (size_t)k * n + p, uk);
            nt_st(us, (size_t)k * n + p, uk + hdt * Um[k]);
        }
        if (det3(Um + 3) <= P.m.jf) atomicOr(invalid, 1);
    }
}

// Node-major copy of U at the characteristic sample nodes: one 96-byte
// record per sample instead of twelve plane-strided loads in the gather.
__global__ __launch_bounds__(BLOCK) void k_pack(const double *__restrict__ U, const int32_t *__restrict__ nodes,
                                                double *__restrict__ pack, int n, int ns) {
    const int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= ns) return;
    const int p = nodes[s];
#pragma unroll
    for (int a = 0; a < NC; ++a) pack[(size_t)s * NC + a] = U[(size_t)a * n + p];
}

// Characteristic state at each traction link: predictor value and gradient
// of every state component (cbw3d:2676-2700).  GLANES lanes share a link,
// each accumulating NC / GLANES components from the packed sample records;
// the group's table loads coalesce and each sample costs one 96-byte line.
constexpr int GLANES = 2;
constexpr int GC = NC / GLANES;  // components per lane
static_assert(NC % GLANES == 0, "components must split evenly over the lanes of a link");
constexpr int GATHER_BLOCK = 256;
__global__ __launch_bounds__(GATHER_BLOCK) void k_gather(const double *__restrict__ pack, const int32_t *__restrict__ clink,
                                                  const int32_t *__restrict__ gidx, const double *__restrict__ gw,
                                                  const int32_t *__restrict__ pidx, const double *__restrict__ pw,
                                                  double *__restrict__ cs, int L, int Lc) {
    constexpr int NG = MAX_CHAR - 16;
    const int t = xcd_block() * GATHER_BLOCK + threadIdx.x;
    const int c = t / GLANES;
    const int a0 = GC * (t % GLANES);
    if (c >= Lc) return;
    double val[GC], gx[GC], gy[GC], gz[GC];
#pragma unroll
    for (int k = 0; k < GC; ++k) val[k] = gx[k] = gy[k] = gz[k] = 0.0;
    // Samples are consumed in stencil order up to the first empty slot; the
    // indices, weights and records of GU slots are loaded before any is
    // used, so each lane keeps GU sample loads in flight (four rather than
    // eight: the short stencils stop earlier).  The sixteen value samples'
    // loads, free of branches, are all issued together.
    constexpr int GU = 4;
    static_assert(NG % GU == 0, "gradient stencil length must be a multiple of the unroll");
    // No load sits behind a branch: an empty slot reads row 0 of the pack and
    // its terms are dropped by selects; the next chunk's indices are
    // requested before this chunk's records, so the index -> record chain
    // costs one round trip per chunk instead of two.
    bool alive = true;
    int nidx[GU];
#pragma unroll
    for (int k = 0; k < GU; ++k) nidx[k] = gidx[(size_t)k * Lc + c];
    for (int m0 = 0; m0 < NG && alive; m0 += GU) {
        int sidx[GU];
#pragma unroll
        for (int k = 0; k < GU; ++k) sidx[k] = nidx[k];
        if (m0 + GU < NG)
#pragma unroll
            for (int k = 0; k < GU; ++k) nidx[k] = gidx[(size_t)(m0 + GU + k) * Lc + c];
        bool use[GU];
#pragma unroll
        for (int k = 0; k < GU; ++k) {
            alive = alive && sidx[k] >= 0;
            use[k] = alive;
        }
        double w1[GU], w2[GU], w3[GU], x[GU][GC];
#pragma unroll
        for (int k = 0; k < GU; ++k) {
            const int m = m0 + k;
            w1[k] = gw[((size_t)0 * NG + m) * Lc + c];
            w2[k] = gw[((size_t)1 * NG + m) * Lc + c];
            w3[k] = gw[((size_t)2 * NG + m) * Lc + c];
            const double *rp = pack + (size_t)(use[k] ? sidx[k] : 0) * NC + a0;
#pragma unroll
            for (int j = 0; j < GC; ++j) x[k][j] = rp[j];
        }
#pragma unroll
        for (int k = 0; k < GU; ++k)
#pragma unroll
            for (int j = 0; j < GC; ++j) {
                gx[j] = use[k] ? gx[j] + w1[k] * x[k][j] : gx[j];
                gy[j] = use[k] ? gy[j] + w2[k] * x[k][j] : gy[j];
                gz[j] = use[k] ? gz[j] + w3[k] * x[k][j] : gz[j];
            }
    }
    for (int m0 = 0; m0 < 16; m0 += GU) {
        int sidx[GU];
#pragma unroll
        for (int k = 0; k < GU; ++k) sidx[k] = pidx[(size_t)(m0 + k) * Lc + c];
        double w0[GU], x[GU][GC];
#pragma unroll
        for (int k = 0; k < GU; ++k) {
            w0[k] = pw[(size_t)(m0 + k) * Lc + c];
            const double *rp = pack + (size_t)(sidx[k] >= 0 ? sidx[k] : 0) * NC + a0;
#pragma unroll
            for (int j = 0; j < GC; ++j) x[k][j] = rp[j];
        }
#pragma unroll
        for (int k = 0; k < GU; ++k)
#pragma unroll
            for (int j = 0; j < GC; ++j) val[j] = sidx[k] >= 0 ? val[j] + w0[k] * x[k][j] : val[j];
    }
    const int ell = clink[c];
#pragma unroll
    for (int k = 0; k < GC; ++k) {
        const int a = a0 + k;
        cs[((size_t)0 * NC + a) * L + ell] = val[k];
        cs[((size_t)1 * NC + a) * L + ell] = gx[k];
        cs[((size_t)2 * NC + a) * L + ell] = gy[k];
        cs[((size_t)3 * NC + a) * L + ell] = gz[k];
    }
}

struct LinkTables {
    const int32_t *lp, *lup, *lq, *lbid;
    const double *eta, *xb, *yb, *zb, *nx, *ny, *nz;
    const int32_t *kind_by_id, *comp_kind, *mode_by_id;
    const double *prm_by_id;
    const int32_t *dyn_en;
    const double *dyn_t;
    int32_t *lkind, *lck;
    double *lbv, *lFb, *lPb;
    double *cs;
    const int32_t *tp, *tvalid;  // compat_bfl tangent stencils (MAX_TAN, L)
    const double *tw1, *tw2;
};

// Boundary values, F predictor and the normal-traction solve per link
// (cbw3d:2888-3151).
// One wave per block, as for the link kernel below.
constexpr int PREP_BLOCK = 64;
__global__ __launch_bounds__(PREP_BLOCK) void k_prepare(const double *__restrict__ U, const double *__restrict__ uu,
                                                        LinkTables T, double t, Params P) {
    const int ell = blockIdx.x * PREP_BLOCK + threadIdx.x;
    if (ell >= P.L) return;
    t = step_time(t, P);
    const size_t n = P.n, L = P.L;
    const int bid = T.lbid[ell];
    const int mode = T.mode_by_id[bid];
    const int kk[3] = {T.comp_kind[0 * MAX_IDS + bid], T.comp_kind[1 * MAX_IDS + bid], T.comp_kind[2 * MAX_IDS + bid]};
    const double x = T.xb[ell], y = T.yb[ell], z = T.zb[ell];
    const double nv[3] = {T.nx[ell], T.ny[ell], T.nz[ell]};
    const double *prm = T.prm_by_id + bid * NPRM;
    double tr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) tr[c] = curved_value(mode, prm, c, x, y, z, nv, t, P.m.jf);
    if (T.dyn_en[ell] != 0)
#pragma unroll
        for (int c = 0; c < 3; ++c) tr[c] = T.dyn_t[c * L + ell];
    // The link's tables are written after its last load (a load behind a
    // store would wait for the store).
    const int lkind = T.kind_by_id[bid];
    double Fb[9] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0};
    double Pb[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const bool any_neu = kk[0] == BC_NEU || kk[1] == BC_NEU || kk[2] == BC_NEU;
    if (any_neu) {
        const int p = T.lp[ell];
        double F[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = U[(3 + k) * n + p];
        if (P.recon == R_LOCALF_BFL || P.recon == R_COMPAT) {
            const int up = T.lup[ell];
            if (up >= 0) {
                const double eta = T.eta[ell];
#pragma unroll
                for (int k = 0; k < 9; ++k) F[k] = F[k] + eta * (F[k] - U[(3 + k) * n + up]);
            }
        }
        if (P.recon == R_COMPAT && T.tvalid[ell] != 0) {
            // Geometry-compatible boundary F (cbw3d:3005-3049): F n kept, the
            // tangential images replaced by t_i + d u / d t_i from the
            // tangent stencils.
            double nn[3] = {nv[0], nv[1], nv[2]};
            double r = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            if (r < 1.0e-30) r = 1.0e-30;
#pragma unroll
            for (int c = 0; c < 3; ++c) nn[c] = nn[c] / r;
            const bool alt = fabs(nn[0]) >= 0.75;
            const double ref[3] = {alt ? 0.0 : 1.0, alt ? 1.0 : 0.0, 0.0};
            double t1[3] = {nn[1] * ref[2] - nn[2] * ref[1], nn[2] * ref[0] - nn[0] * ref[2],
                            nn[0] * ref[1] - nn[1] * ref[0]};
            double r1 = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
            if (r1 < 1.0e-30) r1 = 1.0e-30;
#pragma unroll
            for (int c = 0; c < 3; ++c) t1[c] = t1[c] / r1;
            const double t2[3] = {nn[1] * t1[2] - nn[2] * t1[1], nn[2] * t1[0] - nn[0] * t1[2],
                                  nn[0] * t1[1] - nn[1] * t1[0]};
            double du1[3] = {0.0, 0.0, 0.0}, du2[3] = {0.0, 0.0, 0.0}, wt_abs = 0.0;
            for (int m = 0; m < MAX_TAN; ++m) {
                const int pp = T.tp[(size_t)m * L + ell];
                if (pp < 0) continue;
                const double w1 = T.tw1[(size_t)m * L + ell], w2 = T.tw2[(size_t)m * L + ell];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double uc = uu[(size_t)c * n + pp];
                    du1[c] += w1 * uc;
                    du2[c] += w2 * uc;
                }
                wt_abs += fabs(w1) + fabs(w2);
            }
            if (wt_abs > 0.0) {
#pragma unroll
                for (int row = 0; row < 3; ++row) {
                    const double h = F[3 * row] * nn[0] + F[3 * row + 1] * nn[1] + F[3 * row + 2] * nn[2];
                    const double g1 = t1[row] + du1[row], g2 = t2[row] + du2[row];
#pragma unroll
                    for (int c = 0; c < 3; ++c) F[3 * row + c] = h * nn[c] + g1 * t1[c] + g2 * t2[c];
                }
            }
        }
        if (P.recon == R_LOCALF)
#pragma unroll
            for (int k = 0; k < 9; ++k) F[k] = T.cs[(size_t)(3 + k) * L + ell];
        const bool all_neu = kk[0] == BC_NEU && kk[1] == BC_NEU && kk[2] == BC_NEU;
        solve_Fb(F, nv, tr, kk, !all_neu, P.solver, P.m, Fb);
        piola(Fb, P.m, Pb);
    }
    T.lkind[ell] = lkind;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T.lck[c * L + ell] = kk[c];
        T.lbv[c * L + ell] = tr[c];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        T.lFb[k * L + ell] = Fb[k];
        T.lPb[k * L + ell] = Pb[k];
    }
}

// Incoming populations of the cut links: BFL with the population-pair
// identities (cbw3d:3155-3241), then the characteristic closure on the
// traction rows (cbw3d:2813-2884).  `use_rec` selects post-collision sources
// (time step) or the raw populations (standalone reconstruction).
// One wave per block: the one or two waves per SIMD that the link count gives
// spread more evenly over the CUs than 128-thread blocks do (75 000 links:
// 4.6 blocks per CU instead of 2.3).
constexpr int LINK_BLOCK = 64;
template <typename T>
__global__ __launch_bounds__(LINK_BLOCK) void k_links(const T *__restrict__ fsrc, T *__restrict__ fdst,
                                                      T *__restrict__ flink,
                                                      const T *__restrict__ rec, const double *__restrict__ U,
                                                      LinkTables Tb, int32_t *__restrict__ invalid, int use_rec,
                                                      int has_char, Params P) {
    const int ell = blockIdx.x * LINK_BLOCK + threadIdx.x;
    if (ell >= P.L) return;
    const size_t n = P.n, L = P.L;
    const int p = Tb.lp[ell], qo = Tb.lq[ell], qi = opp(qo);
    double eta = Tb.eta[ell];
    if (P.recon == R_HALF) eta = 0.5;
    const double din[3] = {(double)dir_comp(qi, 0), (double)dir_comp(qi, 1), (double)dir_comp(qi, 2)};
    const int up = Tb.lup[ell];
    const double v[3] = {Tb.lbv[ell], Tb.lbv[L + ell], Tb.lbv[2 * L + ell]};
    const int kk[3] = {Tb.lck[ell], Tb.lck[L + ell], Tb.lck[2 * L + ell]};
    const double om1 = 1.0 - P.omega;
    const bool char_rows = has_char && (kk[0] == BC_NEU || kk[1] == BC_NEU || kk[2] == BC_NEU);
    double outv[NC];
#pragma unroll
    for (int h = 0; h < NC / 2; ++h) {
        const int a0 = 2 * h;
        bool need[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int a = a0 + e, row = a < 3 ? a : (a - 3) / 3;
            // Traction rows are overwritten by the characteristic closure below.
            need[e] = !(char_rows && kk[row] == BC_NEU);
        }
        outv[a0] = outv[a0 + 1] = 0.0;
        if (!need[0] && !need[1]) continue;
        // The pair's populations in both directions and its record entries
        // at p (and upstream of p), one access per pair; qo and qi share an
        // axis, so the record's flux entry is the same for both.
        const bool upstream = eta < 0.5 && up >= 0;
        double fo[2], fi[2], fu[2] = {0.0, 0.0};
        if (use_rec) {
            double vo[2], vi[2], A[2];
            ld2(fsrc, qo * 12 + a0, p, n, a0, vo[0], vo[1]);
            ld2(fsrc, qi * 12 + a0, p, n, a0, vi[0], vi[1]);
            rld2(rec, a0, p, n, P.omega, A[0], A[1]);
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                fo[e] = fpost_a(vo[e], A[e], rec, n, qo, a0 + e, p, om1, P.cr[qo % 3]);
                fi[e] = fpost_a(vi[e], A[e], rec, n, qi, a0 + e, p, om1, P.cr[qi % 3]);
            }
            if (upstream) {
                double vu[2], Au[2];
                ld2(fsrc, qo * 12 + a0, up, n, a0, vu[0], vu[1]);
                rld2(rec, a0, up, n, P.omega, Au[0], Au[1]);
#pragma unroll
                for (int e = 0; e < 2; ++e) fu[e] = fpost_a(vu[e], Au[e], rec, n, qo, a0 + e, up, om1, P.cr[qo % 3]);
            }
        } else {
            ld2(fsrc, qo * 12 + a0, p, n, a0, fo[0], fo[1]);
            ld2(fsrc, qi * 12 + a0, p, n, a0, fi[0], fi[1]);
            if (upstream) ld2(fsrc, qo * 12 + a0, up, n, a0, fu[0], fu[1]);
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            if (!need[e]) continue;
            const int a = a0 + e, row = a < 3 ? a : (a - 3) / 3;
            double D = 1.0, S = 0.0;
            if (kk[row] == BC_DIR) {
                if (a < 3) {
                    D = -1.0;
                    S = v[a] / 3.0;
                } else {
                    const int col = (a - 3) - row * 3;
                    S = -din[col] * v[row] / P.c;
                }
            } else {
                if (a < 3) {
                    S = -(din[0] * Tb.lPb[(3 * a) * L + ell] + din[1] * Tb.lPb[(3 * a + 1) * L + ell] +
                          din[2] * Tb.lPb[(3 * a + 2) * L + ell]) / P.c;
                } else {
                    D = -1.0;
                    S = Tb.lFb[(a - 3) * L + ell] / 3.0;
                }
            }
            double val;
            if (eta >= 0.5) {
                const double alpha = 1.0 / (2.0 * eta);
                val = alpha * (D * fo[e] + S) + (1.0 - alpha) * fi[e];
            } else {
                const double upv = upstream ? fu[e] : fo[e];
                val = D * (2.0 * eta * fo[e] + (1.0 - 2.0 * eta) * upv) + S;
            }
            outv[a] = val;
        }
    }
    if (char_rows) {
        double Ub[NC], Gx[NC], Gy[NC], Gz[NC], inc[NC];
#pragma unroll
        for (int a = 0; a < NC; ++a) {
            Gx[a] = Tb.cs[((size_t)1 * NC + a) * L + ell];
            Gy[a] = Tb.cs[((size_t)2 * NC + a) * L + ell];
            Gz[a] = Tb.cs[((size_t)3 * NC + a) * L + ell];
            if (a < 3)
                Ub[a] = kk[a] == BC_DIR ? v[a] : U[(size_t)a * n + p];
            else
                Ub[a] = Tb.lFb[(a - 3) * L + ell];
        }
        if (invalid && det_state(Ub) <= P.m.jf) atomicAdd(invalid, 1);
        char_population(Ub, Gx, Gy, Gz, din, Tb.eta[ell], P.dt, P.c, P.omega, P.src_mode, P.src, P.m, inc);
#pragma unroll
        for (int a = 0; a < NC; ++a) {
            const int row = a < 3 ? a : (a - 3) / 3;
            if (kk[row] == BC_NEU) outv[a] = inc[a];
        }
    }
#pragma unroll
    for (int h = 0; h < NC / 2; ++h) {
        st2(fdst, pofs(fdst, qi * 12 + 2 * h, p, n), 2 * h, outv[2 * h], outv[2 * h + 1]);
        st2(flink, (size_t)ell * NC + 2 * h, 2 * h, outv[2 * h], outv[2 * h + 1]);
    }
}

// f = feq(U) - dt/(2Q) B(U) (vnew3d:535-560 as used by cbw3d:4251-4253).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_init_pop(const double *__restrict__ U, T *__restrict__ f, Params P) {
    const int p = xcd_block() * BLOCK + threadIdx.x;
    if (p >= P.n) return;
    const size_t n = P.n;
    double Uv[NC], Pf[9];
#pragma unroll
    for (int a = 0; a < NC; ++a) Uv[a] = U[a * n + p];
    piola(Uv + 3, P.m, Pf);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int axis = q % 3;
        const double s = q < 3 ? 1.0 : -1.0;
#pragma unroll
        for (int a = 0; a < NC; ++a) {
            double flux = 0.0;
            if (a < 3)
                flux = -Pf[3 * a + axis];
            else if ((a - 3) % 3 == axis)
                flux = -Uv[(a - 3) / 3];
            const double B = source_comp(a, P.src_mode, P.src, Uv[0], Uv[1], Uv[2]);
            double fv = (1.0 / 6.0) * (Uv[a] + 3.0 * (s * flux / P.cs[axis]));
            fv -= 0.5 * P.dt / 6.0 * B;
            st(f, pofs(f, q * 12 + a, p, n), a, fv);
        }
    }
}

__global__ void k_set_ustar(const double *__restrict__ disp, const double *__restrict__ U,
                            double *__restrict__ us, double dt, int n) {
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) us[(size_t)c * n + p] = disp[(size_t)c * n + p] - 0.5 * dt * U[(size_t)c * n + p];
}

template <typename T>
__global__ void k_copy_pop(const T *__restrict__ s, T *__restrict__ d, size_t count) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < count) d[i] = s[i];
}

// Conversions between the (72, n) double planes and the stored pairs.
template <bool TO_PLANES>
__global__ void k_pop_layout(const double *__restrict__ s, double *__restrict__ d, int n) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= (size_t)FQ * n) return;
    const int plane = (int)(i / n);
    const size_t k = pofs(s, plane, i - (size_t)plane * n, n);
    if (TO_PLANES)
        d[i] = s[k];
    else
        d[k] = s[i];
}
__global__ void k_pop_to_f64(const float *__restrict__ f, double *__restrict__ d, int n) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= (size_t)FQ * n) return;
    const int plane = (int)(i / n);
    d[i] = ld(f, pofs(f, plane, i - (size_t)plane * n, n), plane % 12);
}
__global__ void k_pop_from_f64(const double *__restrict__ s, float *__restrict__ f, int n) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= (size_t)FQ * n) return;
    const int plane = (int)(i / n);
    st(f, pofs(f, plane, i - (size_t)plane * n, n), plane % 12, s[i]);
}

// Deterministic two-pass reductions over U and u.
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

__global__ __launch_bounds__(BLOCK) void k_reduce_partial(const double *__restrict__ U, const double *__restrict__ u,
                                                          const double *__restrict__ uprev,
                                                          double *__restrict__ part, int n) {
    __shared__ double sh[BLOCK][NRED];
    double acc[NRED] = {0.0, 0.0, INFINITY, -INFINITY, 0.0, 0.0, 0.0};
    for (int p = blockIdx.x * BLOCK + threadIdx.x; p < n; p += gridDim.x * BLOCK) {
        double Uv[NC];
        double nonfinite = 0.0;
#pragma unroll
        for (int a = 0; a < NC; ++a) {
            Uv[a] = U[(size_t)a * n + p];
            if (!isfinite(Uv[a])) nonfinite += 1.0;
        }
        double uu = 0.0, du = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double x = u[(size_t)c * n + p];
            if (!isfinite(x)) nonfinite += 1.0;
            uu += x * x;
            if (uprev) {
                const double dd = x - uprev[(size_t)c * n + p];
                du += dd * dd;
            }
        }
        const double v2 = Uv[0] * Uv[0] + Uv[1] * Uv[1] + Uv[2] * Uv[2];
        const double J = det3(Uv + 3);
        const double o[NRED] = {v2, sqrt(v2), J, J, sqrt(uu), nonfinite, du};
        red_combine(acc, o);
    }
    block_reduce(acc, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NRED; ++k) part[blockIdx.x * 8 + k] = sh[0][k];
}

__global__ __launch_bounds__(BLOCK) void k_reduce_final(double *__restrict__ part, int nparts,
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

// Faces of the rectangular box solver (left, right, bottom, top, front,
// back): kind 0 periodic / 1 Dirichlet velocity / 2 Neumann traction, value
// mode and 16 params each, the lattice and the box lengths.
struct Box3 {
    int kind[6], mode[6];
    double prm[6][16];
    int n[3];
    double len[3];
};

// Half-way face reconstruction of the box solver (vnew3d:1999-2200): one
// thread per (face, face node); the box's nodes are all active in (i, j, k)
// order with k fastest.  Dirichlet faces bounce back with the velocity
// datum (velocity rows -f + v/3, F rows f + n v / c); Neumann faces impose
// the traction on the velocity rows (f + T / c) and the face F of the column
// solve on the F rows (-f + F_b / 3).  Faces read and write disjoint
// populations.
template <typename T>
__global__ __launch_bounds__(128) void k_box_sides3(const T *fsrc, T *fdst, const T *__restrict__ rec,
                                                     const double *__restrict__ U, int use_rec, Box3 B,
                                                     double time, Params P) {
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    int m = ny * nz;
    if (nx * nz > m) m = nx * nz;
    if (nx * ny > m) m = nx * ny;
    const long long t = (long long)blockIdx.x * 128 + threadIdx.x;
    if (t >= 6LL * m) return;
    time = step_time(time, P);
    const int side = (int)(t / m), s = (int)(t - (long long)side * m);
    const int kind = B.kind[side];
    if (kind == 0) return;
    const int axis = side / 2;
    const int nn = axis == 0 ? ny * nz : (axis == 1 ? nx * nz : nx * ny);
    if (s >= nn) return;
    int i, j, k;
    if (axis == 0) {
        i = (side & 1) ? nx - 1 : 0;
        j = s % ny;
        k = s / ny;
    } else if (axis == 1) {
        j = (side & 1) ? ny - 1 : 0;
        i = s % nx;
        k = s / nx;
    } else {
        k = (side & 1) ? nz - 1 : 0;
        i = s % nx;
        j = s / nx;
    }
    const int outward = (side & 1) ? 1 : -1;     // normal component along the axis
    const int qo = (side & 1) ? axis : axis + 3; // outgoing direction (towards the face)
    const int qi = opp(qo);
    const double sign = (double)outward;
    const size_t n = P.n;
    const int p = (i * ny + j) * nz + k;
    double xb[3] = {(i + 0.5) * P.dx, (j + 0.5) * P.dy, (k + 0.5) * P.dz};
    xb[axis] = (side & 1) ? B.len[axis] : 0.0;
    const double *prm = B.prm[side];
    double v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = box_value3(c, B.mode[side], prm, xb[0], xb[1], xb[2], time);
    const double om1 = 1.0 - P.omega;
    double fo[12];
#pragma unroll
    for (int a = 0; a < 12; ++a)
        fo[a] = use_rec ? fpost(fsrc, rec, n, qo, a, p, om1, P.omega, P.cr[axis])
                        : ld(fsrc, pofs(fsrc, qo * 12 + a, p, n), a);
    const double cn = P.cs[axis];  // speed along the face normal (vnew3d:2123-2127)
    double val[12];
    if (kind == BC_DIR) {
#pragma unroll
        for (int c = 0; c < 3; ++c) val[c] = -fo[c] + v[c] / 3.0;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                val[3 + 3 * r + c] = fo[3 + 3 * r + c] + (c == axis ? (double)outward : 0.0) * v[r] / cn;
    } else {
        double F[9];
#pragma unroll
        for (int a = 0; a < 9; ++a) F[a] = U[(size_t)(3 + a) * n + p];
        solve_normal_column3(F, axis, sign, v, P.m);
#pragma unroll
        for (int c = 0; c < 3; ++c) val[c] = fo[c] + v[c] / cn;
#pragma unroll
        for (int a = 0; a < 9; ++a) val[3 + a] = -fo[3 + a] + F[a] / 3.0;
    }
#pragma unroll
    for (int a = 0; a < 12; ++a) st(fdst, pofs(fdst, qi * 12 + a, p, n), a, val[a]);
}

}  // namespace

struct vlbm3d_handle {
    vlbm3d_desc d;
    Params P;
    LinkTables T;
    Box3 box{};  // faces of the rectangular box solver (d.box != 0)
    hipStream_t stream;
    int device = 0;  // device current at create
    Timing timing;
    // Captured step loops: graph_steps fused steps per graph, replayed on
    // the solver stream; the time lives in `clock` on the device.
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        void *f0 = nullptr, *f1 = nullptr;
    };
    int graph_steps = 0;
    hipStream_t cap = nullptr;
    double *clock = nullptr;
    StepGraph graphs[2];
    void drop_graphs() {
        for (auto &g : graphs) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g = StepGraph{};
        }
    }
    // Library-owned tables built at create: the compacted anchor list of
    // the displacement sweep and the second-neighbour table.
    int32_t *anc_off = nullptr, *anc_key = nullptr, *nbr2 = nullptr;
    double *anc_dat = nullptr;
    // Stream kernel's neighbour table and the link-major copy of the cut
    // links' populations (L x 12 in the population precision).
    int32_t *nbr_pull = nullptr;
    void *flink = nullptr;
    void release() {
        drop_graphs();
        if (cap) (void)hipStreamDestroy(cap);
        cap = nullptr;
        for (void *ptr : {(void *)anc_off, (void *)anc_key, (void *)nbr2, (void *)anc_dat, (void *)clock,
                          (void *)nbr_pull, flink})
            if (ptr) (void)hipFree(ptr);
        anc_off = anc_key = nbr2 = nbr_pull = nullptr;
        anc_dat = nullptr;
        flink = nullptr;
        clock = nullptr;
    }
};

// Compacted Dirichlet anchors per node (cbw3d:2413-2469): for every link of
// the node, in ascending link order, that contributes a component (all three
// when it carries the sweep's boundary id, else the components whose kind is
// Dirichlet): key = boundary id | component mask << 8 | value mode << 12 and
// (x_b, y_b, z_b, x_b - x_p, y_b - y_p, z_b - z_p) with x_p = (i + 1/2) dx as
// in the sweep.
static int build_anchor_tables(vlbm3d_handle *h) {
    const vlbm3d_desc &d = h->d;
    const size_t n = (size_t)d.n_active, L = (size_t)d.n_links, na = (size_t)std::max(d.n_anchor_links, 0);
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail("vlbm3d_create: sync", e);
    std::vector<int32_t> off(n + 1), ids(na), bid(L), ai(n), aj(n), ak(n), ck(3 * MAX_IDS), mode(MAX_IDS);
    std::vector<double> xb(L), yb(L), zb(L);
    e = hipMemcpy(off.data(), d.node_link_off, (n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess && na) e = hipMemcpy(ids.data(), d.node_link_ids, na * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess && L) {
        e = hipMemcpy(bid.data(), d.link_bid, L * sizeof(int32_t), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(xb.data(), d.link_xb, L * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(yb.data(), d.link_yb, L * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(zb.data(), d.link_zb, L * sizeof(double), hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess) e = hipMemcpy(ai.data(), d.active_i, n * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(aj.data(), d.active_j, n * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ak.data(), d.active_k, n * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(ck.data(), d.comp_kind_by_id, ck.size() * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(mode.data(), d.mode_by_id, mode.size() * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail("vlbm3d_create: anchor table copy", e);
    std::vector<int32_t> aoff(n + 1, 0), key;
    std::vector<double> dat;
    for (size_t p = 0; p < n; ++p) {
        aoff[p] = (int32_t)key.size();
        if (off[p] < 0 || off[p + 1] < off[p] || (size_t)off[p + 1] > na)
            return fail_msg("vlbm3d_create: node_link_off is not a CSR offset table");
        const double xp = ((double)ai[p] + 0.5) * d.dx, yp = ((double)aj[p] + 0.5) * d.dy,
                     zp = ((double)ak[p] + 0.5) * d.dz;
        for (int t = off[p]; t < off[p + 1]; ++t) {
            const int ell = ids[t];
            if (ell < 0 || (size_t)ell >= L) return fail_msg("vlbm3d_create: node_link_ids entry out of range");
            const int b = bid[ell];
            if (b < 0 || b >= MAX_IDS) return fail_msg("vlbm3d_create: link boundary id out of range");
            int mask = 0;
            if (d.disp_boundary_id >= 0) {
                if (b == d.disp_boundary_id) mask = 7;
            } else {
                for (int c = 0; c < 3; ++c)
                    if (ck[c * MAX_IDS + b] == BC_DIR) mask |= 1 << c;
            }
            if (!mask) continue;
            if (mode[b] < 0 || mode[b] >= (1 << 19)) return fail_msg("vlbm3d_create: value mode out of range");
            key.push_back(b | (mask << 8) | (mode[b] << 12));
            const double x = xb[ell], y = yb[ell], z = zb[ell];
            const double ent[6] = {x, y, z, x - xp, y - yp, z - zp};
            dat.insert(dat.end(), ent, ent + 6);
        }
    }
    aoff[n] = (int32_t)key.size();
    const size_t ne = std::max<size_t>(key.size(), 1);
    e = hipMalloc((void **)&h->anc_off, (n + 1) * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&h->anc_key, ne * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&h->anc_dat, ne * 6 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&h->nbr2, (size_t)NQ * n * sizeof(int32_t));
    if (e == hipSuccess) e = hipMemcpy(h->anc_off, aoff.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && !key.empty())
        e = hipMemcpy(h->anc_key, key.data(), key.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && !dat.empty())
        e = hipMemcpy(h->anc_dat, dat.data(), dat.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail("vlbm3d_create: anchor tables", e);
    hipLaunchKernelGGL(k_nbr2, dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, h->stream,
                       (const int32_t *)d.neighbor, h->nbr2, (int)n);
    e = hipGetLastError();
    if (e != hipSuccess) return fail("vlbm3d_create: k_nbr2", e);
    // stream kernel's neighbour table and the link-major population copy
    const size_t esz = d.precision == VLBM_PREC_F32 ? sizeof(float) : sizeof(double);
    e = hipMalloc((void **)&h->nbr_pull, (size_t)NQ * n * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&h->flink, std::max<size_t>(L, 1) * NC * esz);
    int32_t *bad = nullptr;
    if (e == hipSuccess) e = hipMalloc((void **)&bad, sizeof(int32_t));
    if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(int32_t), h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h->nbr_pull, d.neighbor, (size_t)NQ * n * sizeof(int32_t), hipMemcpyDeviceToDevice,
                           h->stream);
    if (e == hipSuccess && L > 0) {
        hipLaunchKernelGGL(k_nbr_pull, dim3((unsigned)((L + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, h->stream,
                           (const int32_t *)d.neighbor, d.link_p, d.link_q, h->nbr_pull, bad, (int)n, (int)L);
        e = hipGetLastError();
    }
    int32_t hbad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (bad) (void)hipFree(bad);
    if (e != hipSuccess) return fail("vlbm3d_create: stream neighbour table", e);
    if (hbad) return fail_msg("vlbm3d_create: a cut link's neighbour is present in the neighbour table");
    return 0;
}

static Params make_params(const vlbm3d_desc &d) {
    Params P{};
    P.n = d.n_active;
    P.L = d.n_links;
    P.m = Material{d.model, d.lam, d.mu, d.p0, d.p1, d.p2, d.j_floor};
    P.dt = d.dt;
    P.c = d.c;
    P.omega = d.omega;
    P.dx = d.dx;
    P.dy = d.dy;
    P.dz = d.dz;
    P.src_mode = d.source_mode;
    for (int k = 0; k < 12; ++k) P.src[k] = d.source_params[k];
    P.recon = d.recon;
    P.solver = d.normal_solver;
    P.disp_relax = d.disp_relax;
    P.w2 = d.disp_weight * d.disp_weight;
    P.disp_bid = d.disp_boundary_id;
    P.blend = d.compat_blend;
    for (int a = 0; a < 3; ++a) {
        P.cs[a] = (d.box && d.box_speed[a] > 0.0) ? d.box_speed[a] : d.c;
        P.cr[a] = d.c / P.cs[a];
    }
    return P;
}

static LinkTables make_tables(const vlbm3d_desc &d) {
    LinkTables T{};
    T.lp = d.link_p;
    T.lup = d.link_up;
    T.lq = d.link_q;
    T.lbid = d.link_bid;
    T.eta = d.link_eta;
    T.xb = d.link_xb;
    T.yb = d.link_yb;
    T.zb = d.link_zb;
    T.nx = d.link_nx;
    T.ny = d.link_ny;
    T.nz = d.link_nz;
    T.kind_by_id = d.kind_by_id;
    T.comp_kind = d.comp_kind_by_id;
    T.mode_by_id = d.mode_by_id;
    T.prm_by_id = d.params_by_id;
    T.dyn_en = d.dyn_enabled;
    T.dyn_t = d.dyn_traction;
    T.lkind = d.lkind;
    T.lck = d.lcomp_kind;
    T.lbv = d.lbv;
    T.lFb = d.lFb;
    T.lPb = d.lPb;
    T.cs = d.char_state;
    T.tp = d.tan_p;
    T.tw1 = d.tan_w1;
    T.tw2 = d.tan_w2;
    T.tvalid = d.tan_valid;
    return T;
}

#define CHECK_LAUNCH(where) VLBM_CHECK_LAUNCH(where)

// use_rec: the sources are pre-collision populations plus the per-node
// records of this step, and k_record has already filled the gather pack.
// The kernels go to stream `st`; `invalid` (may be null) receives the
// characteristic closure's J flags.
template <typename T>
static int run_boundary(vlbm3d_handle *h, hipStream_t st, const T *src, T *dst, double t, int use_rec,
                        int32_t *invalid, const Params &P) {
    const vlbm3d_desc &d = h->d;
    if (d.box) {
        const Box3 &B = h->box;
        long long m = (long long)B.n[1] * B.n[2];
        if ((long long)B.n[0] * B.n[2] > m) m = (long long)B.n[0] * B.n[2];
        if ((long long)B.n[0] * B.n[1] > m) m = (long long)B.n[0] * B.n[1];
        TimedSlot ts(h->timing, st, VLBM_SLOT_LINKS);
        hipLaunchKernelGGL((k_box_sides3<T>), dim3((unsigned)((6 * m + 127) / 128)), dim3(128), 0, st, src, dst,
                           (const T *)d.record, (const double *)d.U, use_rec, B, t, P);
        CHECK_LAUNCH("k_box_sides3");
        return 0;
    }
    if (d.n_links == 0) return 0;
    if (d.has_char && d.n_char_links > 0) {
        TimedSlot ts(h->timing, st, VLBM_SLOT_GATHER);
        if (!use_rec) {
            const int gp = (d.n_char_nodes + BLOCK - 1) / BLOCK;
            hipLaunchKernelGGL(k_pack, dim3(gp), dim3(BLOCK), 0, st, (const double *)d.U, d.char_nodes,
                               d.char_pack, d.n_active, d.n_char_nodes);
            CHECK_LAUNCH("k_pack");
        }
        const int g = xcd_grid((int)((size_t)d.n_char_links * GLANES), GATHER_BLOCK);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(GATHER_BLOCK), 0, st, (const double *)d.char_pack, d.char_link,
                           d.char_gidx, d.char_gw, d.char_pidx, d.char_pw, d.char_state, d.n_links,
                           d.n_char_links);
        CHECK_LAUNCH("k_gather");
    }
    {
        TimedSlot ts(h->timing, st, VLBM_SLOT_PREPARE);
        hipLaunchKernelGGL(k_prepare, dim3((d.n_links + PREP_BLOCK - 1) / PREP_BLOCK), dim3(PREP_BLOCK), 0, st, (const double *)d.U, (const double *)d.u, h->T, t,
                           P);
        CHECK_LAUNCH("k_prepare");
    }
    TimedSlot ts(h->timing, st, VLBM_SLOT_LINKS);
    hipLaunchKernelGGL((k_links<T>), dim3((d.n_links + LINK_BLOCK - 1) / LINK_BLOCK), dim3(LINK_BLOCK), 0, st, src,
                       dst, (T *)h->flink, (const T *)d.record,
                       (const double *)d.U, h->T, invalid, use_rec, d.has_char, P);
    CHECK_LAUNCH("k_links");
    return 0;
}

template <typename T>
static int launch_stream(vlbm3d_handle *h, bool next, const T *f0, T *f1, const Params &P) {
    const vlbm3d_desc &d = h->d;
    const int gn = xcd_grid(d.n_active);
    if (next)
        hipLaunchKernelGGL((k_stream<T, true>), dim3(gn), dim3(BLOCK), 0, h->stream, f0, f1, (const T *)d.record,
                           (const T *)h->flink, (const int32_t *)h->nbr_pull, d.U, d.u, d.u_star, d.invalid, P);
    else
        hipLaunchKernelGGL((k_stream<T, false>), dim3(gn), dim3(BLOCK), 0, h->stream, f0, f1, (const T *)d.record,
                           (const T *)h->flink, d.neighbor, d.U, d.u, d.u_star, d.invalid, P);
    CHECK_LAUNCH("k_stream");
    return 0;
}

// Launches of one step from (time, steps) on the solver stream: sweeps,
// record, boundary and the stream kernel (with the next step's collide when
// `next`).  `last` marks the call's last step; P carries the clock pointer
// inside captured step loops.
template <typename T>
static int launch_step3(vlbm3d_handle *h, const Params &P, T *f0, T *f1, double time, int64_t steps, bool next,
                        bool last) {
    const vlbm3d_desc &d = h->d;
    const int gn = xcd_grid(d.n_active);
    TimedSlot whole(h->timing, h->stream, VLBM_SLOT_STEP);
    const double tb = time + 0.5 * d.dt;
    const bool disp = d.disp_interval > 0 && steps % d.disp_interval == 0;
    const bool compat = d.compat_interval > 0 && steps % d.compat_interval == 0;
    if (disp) {
        for (int s = 0; s < d.disp_sweeps; ++s) {
            TimedSlot ts(h->timing, h->stream, VLBM_SLOT_SWEEP);
            hipLaunchKernelGGL(k_sweep, dim3(gn), dim3(BLOCK), 0, h->stream, (const double *)d.U,
                               (const double *)d.u, d.u_next, d.neighbor, (const int32_t *)h->anc_off,
                               (const int32_t *)h->anc_key, (const double *)h->anc_dat, d.params_by_id, tb, P);
            CHECK_LAUNCH("k_sweep");
            if (s + 1 < d.disp_sweeps) {
                TimedSlot tc(h->timing, h->stream, VLBM_SLOT_COMMIT);
                hipLaunchKernelGGL(k_commit, dim3(gn), dim3(BLOCK), 0, h->stream, (const double *)d.u_next, d.u,
                                   d.u_star, d.n_active);
                CHECK_LAUNCH("k_commit");
            }
        }
    }
    {
        TimedSlot ts(h->timing, h->stream, VLBM_SLOT_RECORD);
        // Within a call only the gather pack consumes the post-repair F of
        // the local_f closure; the U planes get it on the call's last step.
        const bool pack = d.has_char && d.n_char_links > 0;
        const int write_U = (last || d.recon != R_LOCALF || !pack) ? 1 : 0;
        // The flag is cleared here when the stream kernel performs the next
        // collide; the links then report only on the call's last step, where
        // the reference's flag holds the last collide's and the last
        // closure's checks.
        hipLaunchKernelGGL(k_record<T>, dim3(gn), dim3(BLOCK), 0, h->stream, d.U,
                           disp ? (const double *)d.u_next : (const double *)d.u, d.u, d.u_star, (T *)d.record,
                           d.neighbor, (const int32_t *)h->nbr2, pack ? d.char_pack_slot : (const int32_t *)nullptr,
                           d.char_pack, (int)compat, (int)disp, write_U, last ? (int32_t *)nullptr : d.invalid, P);
        CHECK_LAUNCH("k_record");
    }
    int32_t *link_flag = next ? (int32_t *)nullptr : d.invalid;
    int rc = run_boundary<T>(h, h->stream, f0, f1, tb, 1, link_flag, P);
    if (rc) return rc;
    TimedSlot ts(h->timing, h->stream, VLBM_SLOT_STREAM);
    return launch_stream<T>(h, next, f0, f1, P);
}

// Captures graph_steps fused steps (never a call's last step) with the time
// kept on the device; the schedule must be the same every step.
template <typename T>
static int capture_steps3(vlbm3d_handle *h, T *f0, T *f1, hipGraphExec_t *out) {
    hipError_t e;
    if (!h->cap) {
        e = hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking);
        if (e != hipSuccess) return fail("vlbm3d: capture stream", e);
    }
    if (!h->clock) {
        e = hipMalloc((void **)&h->clock, sizeof(double));
        if (e != hipSuccess) return fail("vlbm3d: clock", e);
    }
    Params P = h->P;
    P.clock = h->clock;
    hipStream_t keep = h->stream;
    const bool timing = h->timing.on;
    h->timing.on = false;
    h->stream = h->cap;
    e = hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        h->stream = keep;
        h->timing.on = timing;
        return fail("vlbm3d: begin capture", e);
    }
    int rc = 0;
    for (int k = 0; k < h->graph_steps && rc == 0; ++k) {
        rc = (k % 2 == 0) ? launch_step3<T>(h, P, f0, f1, 0.0, 0, true, false)
                          : launch_step3<T>(h, P, f1, f0, 0.0, 0, true, false);
        if (rc == 0) hipLaunchKernelGGL(k_advance_clock, dim3(1), dim3(1), 0, h->cap, h->clock, h->d.dt);
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(h->cap, &graph);
    h->stream = keep;
    h->timing.on = timing;
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail("vlbm3d: end capture", e);
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail("vlbm3d: instantiate", e);
    return 0;
}

// One library call runs nsteps reference steps.  The collide of the first
// step is a kernel of its own; the collide of every later step is fused into
// the previous step's stream kernel (moments of the freshly written
// populations), so the populations are read once per step.  The cut-link
// kernel therefore runs before the stream kernel, which reads the link rows
// back to complete each node's moments.  The last step of the call streams
// without the fused collide, so U, u and u* end as the reference leaves them
// (post-repair state, committed displacement).  With graph_steps = K > 0 and
// a uniform schedule, blocks of K fused steps are replayed from a captured
// graph (an even K keeps f0 / f1 where they started).
template <typename T>
static int run_steps(vlbm3d_handle *h, T *f0, T *f1, double time, int64_t steps, int nsteps) {
    const vlbm3d_desc &d = h->d;
    const int gn = xcd_grid(d.n_active);
    if (nsteps <= 0) return 0;
    hipLaunchKernelGGL(k_reset_invalid, dim3(1), dim3(1), 0, h->stream, d.invalid);
    {
        TimedSlot ts(h->timing, h->stream, VLBM_SLOT_COLLIDE);
        hipLaunchKernelGGL((k_collide<T, true>), dim3(gn), dim3(BLOCK), 0, h->stream, (const T *)f0, d.U, d.u,
                           d.u_star, d.invalid, h->P);
        CHECK_LAUNCH("k_collide");
    }
    int k = 0;
    const bool uniform = d.disp_interval <= 1 && d.compat_interval <= 1;
    const int K = h->graph_steps;
    if (K > 0 && uniform && !h->timing.on && nsteps - 1 >= K) {
        vlbm3d_handle::StepGraph *g = nullptr;
        for (auto &c : h->graphs)
            if (c.exec && c.f0 == (void *)f0 && c.f1 == (void *)f1) g = &c;
        if (!g) {
            g = !h->graphs[0].exec ? &h->graphs[0] : &h->graphs[1];
            if (g->exec) (void)hipGraphExecDestroy(g->exec);
            *g = vlbm3d_handle::StepGraph{};
            int rc = capture_steps3<T>(h, f0, f1, &g->exec);
            if (rc) return rc;
            g->f0 = f0;
            g->f1 = f1;
        }
        hipLaunchKernelGGL(k_set_clock, dim3(1), dim3(1), 0, h->stream, h->clock, time);
        CHECK_LAUNCH("k_set_clock");
        for (; k + K <= nsteps - 1; k += K) {
            hipError_t e = hipGraphLaunch(g->exec, h->stream);
            if (e != hipSuccess) return fail("vlbm3d: graph launch", e);
            for (int j = 0; j < K; ++j) time += d.dt;
            steps += K;
        }
    }
    for (; k < nsteps; ++k) {
        const bool next = k + 1 < nsteps;
        int rc = launch_step3<T>(h, h->P, f0, f1, time, steps, next, !next);
        if (rc) return rc;
        T *tmp = f0;
        f0 = f1;
        f1 = tmp;
        time += d.dt;
        steps += 1;
    }
    return 0;
}

// Device scratch of the allocation-free reduction (vlbm3d_reduce_noalloc):
// the per-block partial rows of k_reduce_partial, at most RED_BLOCKS rows of
// 8 doubles (64 KiB).  Each GPU the library runs on holds its own copy until
// the process ends.  Host code never passes `scratch` itself to a kernel (on
// the host it is only a stand-in symbol); it uses the device address looked
// up once per device below.
constexpr int SCRATCH_ROWS = RED_BLOCKS;
static __device__ __align__(256) double scratch[SCRATCH_ROWS * 8];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on `dev`, looked up on the first call for that
// device with the device current (the caller's current device is restored).
int scratch_address(int dev, double **out) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail_msg("vlbm3d_reduce_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return fail("vlbm3d_reduce_noalloc: get device", e);
        if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return fail("vlbm3d_reduce_noalloc: set device", e);
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) return fail("vlbm3d_reduce_noalloc: scratch address", e);
        g_scratch_ptr[dev] = (double *)ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" {

const char *vlbm_last_error(void) { return vlbm_common::last_error().c_str(); }
int vlbm_abi_version(void) { return VLBM3D_ABI_VERSION; }

int vlbm3d_create(const vlbm3d_desc *desc, vlbm3d_handle **out) {
    if (!desc || !out) return fail_msg("vlbm3d_create: null argument");
    if (desc->struct_size != sizeof(vlbm3d_desc))
        return fail_msg("vlbm3d_create: descriptor size mismatch (" + std::to_string(desc->struct_size) +
                        " != " + std::to_string(sizeof(vlbm3d_desc)) + ")");
    if (desc->abi_version != VLBM3D_ABI_VERSION) return fail_msg("vlbm3d_create: ABI version mismatch");
    if (desc->n_active <= 0) return fail_msg("vlbm3d_create: n_active must be positive");
    if (desc->n_links < 0) return fail_msg("vlbm3d_create: n_links must be non-negative");
    if (desc->precision != VLBM_PREC_F64 && desc->precision != VLBM_PREC_F32)
        return fail_msg("vlbm3d_create: unknown precision");
    if (desc->recon == R_COMPAT && desc->n_links > 0 && (!desc->tan_p || !desc->tan_w1 || !desc->tan_w2 || !desc->tan_valid))
        return fail_msg("vlbm3d_create: compat_bfl needs the tangent stencil tables");
    if (desc->recon < 0 || desc->recon > R_LOCALF_BFL) return fail_msg("vlbm3d_create: unknown reconstruction");
    if (desc->box) {
        const int *bn = desc->box_n;
        if (bn[0] < 1 || bn[1] < 1 || bn[2] < 1 || (long long)bn[0] * bn[1] * bn[2] != desc->n_active)
            return fail_msg("vlbm3d_create: a box needs every node of box_n active");
        if (desc->n_links != 0) return fail_msg("vlbm3d_create: a box has no cut links (n_links must be 0)");
        for (int s = 0; s < 6; ++s)
            if (desc->box_side_kind[s] < 0 || desc->box_side_kind[s] > 2)
                return fail_msg("vlbm3d_create: box side kind must be 0 (periodic), 1 (Dirichlet) or 2 (Neumann)");
        for (int a = 0; a < 3; ++a)
            if ((desc->box_side_kind[2 * a] == 0) != (desc->box_side_kind[2 * a + 1] == 0))
                return fail_msg("vlbm3d_create: periodic box faces must come in pairs");
    }
    if (desc->disp_sweeps < 1 || desc->disp_sweeps > 8) return fail_msg("vlbm3d_create: disp_sweeps must be in [1,8]");
    if ((size_t)desc->n_active * FQ >= (size_t)1 << 40) return fail_msg("vlbm3d_create: lattice too large");
    const void *req[] = {desc->U, desc->u, desc->u_star, desc->u_next, desc->record, desc->invalid,
                         desc->reduce_scratch, desc->neighbor, desc->active_i, desc->active_j, desc->active_k,
                         desc->node_link_off, desc->kind_by_id, desc->comp_kind_by_id, desc->mode_by_id,
                         desc->params_by_id};
    for (const void *ptr : req)
        if (!ptr) return fail_msg("vlbm3d_create: a required device pointer is null");
    if (desc->n_links > 0) {
        const void *lreq[] = {desc->link_p, desc->link_up, desc->link_q, desc->link_bid, desc->link_eta,
                              desc->link_xb, desc->link_yb, desc->link_zb, desc->link_nx, desc->link_ny,
                              desc->link_nz, desc->dyn_enabled, desc->dyn_traction, desc->lkind,
                              desc->lcomp_kind, desc->lbv, desc->lFb, desc->lPb};
        for (const void *ptr : lreq)
            if (!ptr) return fail_msg("vlbm3d_create: a required link pointer is null");
        if (desc->n_anchor_links > 0 && !desc->node_link_ids)
            return fail_msg("vlbm3d_create: node_link_ids is null");
        if (desc->has_char && (!desc->char_state ||
                               (desc->n_char_links > 0 && (!desc->char_link || !desc->char_gidx || !desc->char_gw ||
                                                           !desc->char_pidx || !desc->char_pw || !desc->char_nodes ||
                                                           !desc->char_pack || desc->n_char_nodes <= 0))))
            return fail_msg("vlbm3d_create: characteristic tables are null");
    }
    vlbm3d_handle *h = new vlbm3d_handle;
    h->d = *desc;
    h->P = make_params(*desc);
    h->T = make_tables(*desc);
    for (int s = 0; s < 6; ++s) {
        h->box.kind[s] = desc->box_side_kind[s];
        h->box.mode[s] = desc->box_value_mode[s];
        for (int k = 0; k < 16; ++k) h->box.prm[s][k] = desc->box_params[16 * s + k];
    }
    for (int a = 0; a < 3; ++a) {
        h->box.n[a] = desc->box_n[a];
        h->box.len[a] = desc->box_length[a];
    }
    h->stream = (hipStream_t)desc->stream;
    if (hipGetDevice(&h->device) != hipSuccess) h->device = 0;
    const int rc = build_anchor_tables(h);
    if (rc) {
        h->release();
        delete h;
        return rc;
    }
    *out = h;
    return 0;
}

int vlbm3d_destroy(vlbm3d_handle *h) {
    if (!h) return 0;
    (void)hipStreamSynchronize(h->stream);
    h->release();
    h->timing.destroy();
    delete h;
    return 0;
}

int vlbm3d_timing_enable(vlbm3d_handle *h, int32_t enable) {
    if (!h) return fail_msg("vlbm3d_timing_enable: null handle");
    return h->timing.enable(h->stream, enable);
}

int vlbm3d_timing_read(vlbm3d_handle *h, double *ms, int32_t *counts, int32_t nslots) {
    if (!h || !ms || !counts) return fail_msg("vlbm3d_timing_read: null argument");
    return h->timing.read(h->stream, ms, counts, nslots);
}

int vlbm3d_set_graph_steps(vlbm3d_handle *h, int32_t steps) {
    if (!h) return fail_msg("vlbm3d_set_graph_steps: null handle");
    if (steps < 0 || steps > 4096) return fail_msg("vlbm3d_set_graph_steps: steps must be in [0, 4096]");
    if (steps % 2) return fail_msg("vlbm3d_set_graph_steps: steps must be even");
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail("vlbm3d_set_graph_steps: sync", e);
    h->drop_graphs();
    h->graph_steps = steps;
    return 0;
}

int vlbm3d_step(vlbm3d_handle *h, void *f0, void *f1, double time, int64_t steps, int32_t nsteps) {
    if (!h || !f0 || !f1) return fail_msg("vlbm3d_step: null argument");
    if (nsteps < 0) return fail_msg("vlbm3d_step: nsteps must be non-negative");
    if (h->d.precision == VLBM_PREC_F64)
        return run_steps<double>(h, (double *)f0, (double *)f1, time, steps, nsteps);
    return run_steps<float>(h, (float *)f0, (float *)f1, time, steps, nsteps);
}

int vlbm3d_refresh(vlbm3d_handle *h, const void *f) {
    if (!h || !f) return fail_msg("vlbm3d_refresh: null argument");
    const vlbm3d_desc &d = h->d;
    hipLaunchKernelGGL(k_reset_invalid, dim3(1), dim3(1), 0, h->stream, d.invalid);
    const int gn = xcd_grid(d.n_active);
    if (d.precision == VLBM_PREC_F64)
        hipLaunchKernelGGL((k_collide<double, false>), dim3(gn), dim3(BLOCK), 0, h->stream, (const double *)f, d.U,
                           d.u, d.u_star, d.invalid, h->P);
    else
        hipLaunchKernelGGL((k_collide<float, false>), dim3(gn), dim3(BLOCK), 0, h->stream, (const float *)f, d.U,
                           d.u, d.u_star, d.invalid, h->P);
    CHECK_LAUNCH("k_refresh");
    return 0;
}

int vlbm3d_apply_boundary(vlbm3d_handle *h, void *f_src, void *f_dst, double t, int32_t in_place) {
    if (!h || !f_src || !f_dst) return fail_msg("vlbm3d_apply_boundary: null argument");
    const vlbm3d_desc &d = h->d;
    if (d.n_links == 0 && !d.box) return 0;
    const size_t count = (size_t)FQ * d.n_active;
    const int gc = (int)((count + BLOCK - 1) / BLOCK);
    if (d.precision == VLBM_PREC_F64) {
        double *s = (double *)f_src, *t2 = (double *)f_dst;
        if (in_place) {
            hipLaunchKernelGGL(k_copy_pop<double>, dim3(gc), dim3(BLOCK), 0, h->stream, s, t2, count);
            CHECK_LAUNCH("k_copy_pop");
            return run_boundary<double>(h, h->stream, t2, s, t, 0, d.invalid, h->P);
        }
        return run_boundary<double>(h, h->stream, s, t2, t, 0, d.invalid, h->P);
    }
    float *s = (float *)f_src, *t2 = (float *)f_dst;
    if (in_place) {
        hipLaunchKernelGGL(k_copy_pop<float>, dim3(gc), dim3(BLOCK), 0, h->stream, s, t2, count);
        CHECK_LAUNCH("k_copy_pop");
        return run_boundary<float>(h, h->stream, t2, s, t, 0, d.invalid, h->P);
    }
    return run_boundary<float>(h, h->stream, s, t2, t, 0, d.invalid, h->P);
}

int vlbm3d_init_populations(vlbm3d_handle *h, void *f) {
    if (!h || !f) return fail_msg("vlbm3d_init_populations: null argument");
    const int gn = xcd_grid(h->d.n_active);
    if (h->d.precision == VLBM_PREC_F64)
        hipLaunchKernelGGL(k_init_pop<double>, dim3(gn), dim3(BLOCK), 0, h->stream, (const double *)h->d.U,
                           (double *)f, h->P);
    else
        hipLaunchKernelGGL(k_init_pop<float>, dim3(gn), dim3(BLOCK), 0, h->stream, (const double *)h->d.U,
                           (float *)f, h->P);
    CHECK_LAUNCH("k_init_pop");
    return 0;
}

int vlbm3d_set_ustar(vlbm3d_handle *h, const double *disp) {
    if (!h || !disp) return fail_msg("vlbm3d_set_ustar: null argument");
    const int g = (h->d.n_active + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_set_ustar, dim3(g), dim3(BLOCK), 0, h->stream, disp, (const double *)h->d.U, h->d.u_star,
                       h->d.dt, h->d.n_active);
    CHECK_LAUNCH("k_set_ustar");
    return 0;
}

int vlbm3d_reduce(vlbm3d_handle *h, const double *u_prev, double *out) {
    if (!h || !out) return fail_msg("vlbm3d_reduce: null argument");
    const vlbm3d_desc &d = h->d;
    int nb = (d.n_active + BLOCK - 1) / BLOCK;
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    double *part = d.reduce_scratch;
    double *res = d.reduce_scratch + (size_t)RED_BLOCKS * 8;
    hipLaunchKernelGGL(k_reduce_partial, dim3(nb), dim3(BLOCK), 0, h->stream, (const double *)d.U,
                       (const double *)d.u, u_prev, part, d.n_active);
    CHECK_LAUNCH("k_reduce_partial");
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(BLOCK), 0, h->stream, part, nb, (const int32_t *)d.invalid, res);
    CHECK_LAUNCH("k_reduce_final");
    hipError_t e = hipMemcpyAsync(out, res, sizeof(double) * VLBM_RED_COUNT, hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return fail("vlbm3d_reduce: copy", e);
    e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail("vlbm3d_reduce: sync", e);
    return 0;
}

int vlbm3d_reduce_noalloc(vlbm3d_handle *h, const double *u_prev, double *out_dev) {
    if (!h || !out_dev) return fail_msg("vlbm3d_reduce_noalloc: null argument");
    const vlbm3d_desc &d = h->d;
    int nb = (d.n_active + BLOCK - 1) / BLOCK;
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;  // k_reduce_partial strides over the rest
    if (nb > SCRATCH_ROWS)
        return fail_msg("vlbm3d_reduce_noalloc: " + std::to_string(nb) + " partial rows exceed the scratch (" +
                        std::to_string(SCRATCH_ROWS) + ")");
    double *part = nullptr;
    const int rc = scratch_address(h->device, &part);
    if (rc) return rc;
    hipLaunchKernelGGL(k_reduce_partial, dim3(nb), dim3(BLOCK), 0, h->stream, (const double *)d.U,
                       (const double *)d.u, u_prev, part, d.n_active);
    CHECK_LAUNCH("k_reduce_partial");
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(BLOCK), 0, h->stream, part, nb, (const int32_t *)d.invalid,
                       out_dev);
    CHECK_LAUNCH("k_reduce_final");
    return 0;
}

int vlbm3d_populations_to_f64(vlbm3d_handle *h, const void *f, double *dst) {
    if (!h || !f || !dst) return fail_msg("vlbm3d_populations_to_f64: null argument");
    const size_t count = (size_t)FQ * h->d.n_active;
    const int g = (int)((count + BLOCK - 1) / BLOCK);
    if (h->d.precision == VLBM_PREC_F64)
        hipLaunchKernelGGL(k_pop_layout<true>, dim3(g), dim3(BLOCK), 0, h->stream, (const double *)f, dst,
                           h->d.n_active);
    else
        hipLaunchKernelGGL(k_pop_to_f64, dim3(g), dim3(BLOCK), 0, h->stream, (const float *)f, dst, h->d.n_active);
    CHECK_LAUNCH("populations_to_f64");
    return 0;
}

int vlbm3d_populations_from_f64(vlbm3d_handle *h, const double *src, void *f) {
    if (!h || !f || !src) return fail_msg("vlbm3d_populations_from_f64: null argument");
    const size_t count = (size_t)FQ * h->d.n_active;
    const int g = (int)((count + BLOCK - 1) / BLOCK);
    if (h->d.precision == VLBM_PREC_F64)
        hipLaunchKernelGGL(k_pop_layout<false>, dim3(g), dim3(BLOCK), 0, h->stream, src, (double *)f, h->d.n_active);
    else
        hipLaunchKernelGGL(k_pop_from_f64, dim3(g), dim3(BLOCK), 0, h->stream, src, (float *)f, h->d.n_active);
    CHECK_LAUNCH("populations_from_f64");
    return 0;
}

}  // extern "C"
