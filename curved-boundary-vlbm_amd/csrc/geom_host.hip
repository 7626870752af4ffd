// Host-side geometry preprocessing of libvlbm: the per-link stencil weights
// that the curved-boundary kernels consume, built once per geometry.
//
//   vlbm_geom3d_char_weights  characteristic traction closure stencils of the
//                             D3Q6x12 solver (characteristic_stencil3d.py:19-171
//                             of the reference): weighted quadratic LSQ
//                             gradients over <= 48 nodes and a normal-probe /
//                             link-extrapolation predictor over <= 16 nodes.
//   vlbm_geom3d_tangent_weights  the two surface-tangent derivative stencils
//                             of the D3Q6x12 compat_bfl closure
//                             (_build_sparse_tangent_lsq_weights,
//                             curved_boundary_warp3d.py:3725-3849).
//   vlbm_geom2d_lsq_weights   normal-probe value weights and quadratic LSQ
//                             tangential / Cartesian derivative weights of the
//                             D2Q4x6 solver (curved_boundary_warp.py:332-606).
//
// Links are independent, so they are split over host threads.  The
// arithmetic follows the reference's NumPy expressions operation by
// operation (no contraction into FMAs), so sample selection, index slots and
// validity flags agree exactly; the least-squares solves use LU with partial
// pivoting and Jacobi eigen/singular values, which agree with LAPACK to
// rounding.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <limits>
#include <thread>
#include <vector>

#include "vlbm_geom.h"

#pragma clang fp contract(off)

namespace {

int thread_count(int32_t requested, int64_t work) {
    int n = requested > 0 ? requested : (int)std::thread::hardware_concurrency();
    if (n < 1) n = 1;
    if (n > 64) n = 64;
    if (work < 64) n = 1;
    return n;
}

// Runs fn(begin, end) over [0, n) on `threads` threads in contiguous chunks.
template <typename Fn>
void parallel_for(int64_t n, int threads, Fn fn) {
    if (threads <= 1) {
        fn((int64_t)0, n);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([=] { fn(b, e); });
    }
    for (auto &th : pool) th.join();
}

// Solves (M) X = B in place for a dense n x n M (row-major) and n x nrhs B
// (row-major) by LU with partial pivoting.  Returns false on a zero pivot.
bool lu_solve(double *M, int n, double *B, int nrhs) {
    for (int c = 0; c < n; ++c) {
        int piv = c;
        double best = std::fabs(M[c * n + c]);
        for (int r = c + 1; r < n; ++r) {
            const double v = std::fabs(M[r * n + c]);
            if (v > best) {
                best = v;
                piv = r;
            }
        }
        if (!(best > 0.0)) return false;
        if (piv != c) {
            for (int k = 0; k < n; ++k) std::swap(M[c * n + k], M[piv * n + k]);
            for (int k = 0; k < nrhs; ++k) std::swap(B[c * nrhs + k], B[piv * nrhs + k]);
        }
        const double d = M[c * n + c];
        for (int r = c + 1; r < n; ++r) {
            const double f = M[r * n + c] / d;
            M[r * n + c] = f;
            for (int k = c + 1; k < n; ++k) M[r * n + k] -= f * M[c * n + k];
            for (int k = 0; k < nrhs; ++k) B[r * nrhs + k] -= f * B[c * nrhs + k];
        }
    }
    for (int c = n - 1; c >= 0; --c) {
        const double d = M[c * n + c];
        for (int k = 0; k < nrhs; ++k) {
            double s = B[c * nrhs + k];
            for (int j = c + 1; j < n; ++j) s -= M[c * n + j] * B[j * nrhs + k];
            B[c * nrhs + k] = s / d;
        }
    }
    return true;
}

// Eigenvalues of a symmetric n x n matrix (cyclic Jacobi); returns the
// 2-norm condition number max|lambda| / min|lambda| (inf when singular, NaN
// when the input is not finite).  For the symmetric positive semi-definite
// normal matrices here this is NumPy's cond(M) (ratio of singular values).
double sym_cond(const double *Min, int n) {
    double A[16 * 16];
    for (int k = 0; k < n * n; ++k) {
        if (!std::isfinite(Min[k])) return std::numeric_limits<double>::quiet_NaN();
        A[k] = Min[k];
    }
    // Sweeps until no off-diagonal entry is above rounding relative to its
    // diagonal pair (entries below that are set to zero).
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (apq == 0.0) continue;
                const double app = A[p * n + p], aqq = A[q * n + q];
                if (std::fabs(apq) <= 1e-18 * std::sqrt(std::fabs(app * aqq))) {
                    A[p * n + q] = A[q * n + p] = 0.0;
                    continue;
                }
                rotated = true;
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
            }
        if (!rotated) break;
    }
    double lo = std::numeric_limits<double>::infinity(), hi = 0.0;
    for (int i = 0; i < n; ++i) {
        const double v = std::fabs(A[i * n + i]);
        lo = std::min(lo, v);
        hi = std::max(hi, v);
    }
    if (lo == 0.0) return std::numeric_limits<double>::infinity();
    return hi / lo;
}

// Singular values of an m x n matrix (m >= n, row-major) by one-sided
// Jacobi; writes them (unsorted) to sv.
void singular_values(const double *Ain, int m, int n, double *sv) {
    std::vector<double> A(Ain, Ain + (size_t)m * n);
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double a = 0.0, b = 0.0, g = 0.0;
                for (int i = 0; i < m; ++i) {
                    const double x = A[i * n + p], y = A[i * n + q];
                    a += x * x;
                    b += y * y;
                    g += x * y;
                }
                if (std::fabs(g) <= 1e-15 * std::sqrt(a * b) || g == 0.0) continue;
                rotated = true;
                const double zeta = (b - a) / (2.0 * g);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (int i = 0; i < m; ++i) {
                    const double x = A[i * n + p], y = A[i * n + q];
                    A[i * n + p] = c * x - s * y;
                    A[i * n + q] = s * x + c * y;
                }
            }
        if (!rotated) break;
    }
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < m; ++i) s += A[i * n + j] * A[i * n + j];
        sv[j] = std::sqrt(s);
    }
}

// ---------------------------------------------------------------- 3D -----

constexpr int G3 = VLBM_GEOM3D_GRADIENT_SAMPLES;  // 48
constexpr int P3 = VLBM_GEOM3D_PROBE_SAMPLES;     // 16
constexpr int S3 = G3 + P3;

struct Grid3 {
    const int32_t *id;
    int nx, ny, nz;
    bool inside(int i, int j, int k) const { return i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz; }
    int32_t at(int i, int j, int k) const { return id[((size_t)i * ny + j) * nz + k]; }
};

// Trilinear probe at `pt` (lattice units): the 8 corner nodes in (x, y, z)
// lexicographic order, their weights, and whether every corner is on the
// grid, within three cells of the link node and active
// (characteristic_stencil3d.py:89-100 as restated in the host builder).
bool probe3(const Grid3 &g, const double pt[3], const int ctr[3], int32_t pp[8], double ww[8]) {
    int lower[3];
    double frac[3];
    for (int a = 0; a < 3; ++a) {
        const double fl = std::floor(pt[a]);
        lower[a] = (int)fl;
        frac[a] = pt[a] - (double)lower[a];
    }
    bool ok = true;
    for (int c = 0; c < 8; ++c) {
        const int cb[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
        int node[3];
        bool in_grid = true, near = true;
        for (int a = 0; a < 3; ++a) {
            node[a] = lower[a] + cb[a];
            const int lim = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
            if (node[a] < 0 || node[a] >= lim) in_grid = false;
            if (std::abs(node[a] - ctr[a]) > 3) near = false;
        }
        int cl[3];
        for (int a = 0; a < 3; ++a) {
            const int lim = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
            cl[a] = std::min(std::max(node[a], 0), lim - 1);
        }
        pp[c] = g.at(cl[0], cl[1], cl[2]);
        double w = 1.0;
        for (int a = 0; a < 3; ++a) w = w * (cb[a] == 1 ? frac[a] : 1.0 - frac[a]);
        ww[c] = w;
        if (!(in_grid && near && pp[c] >= 0)) ok = false;
    }
    return ok;
}

struct Links3 {
    const int32_t *li, *lj, *lk, *lq;
    const double *xb, *yb, *zb, *nx, *ny, *nz, *eta;
};

// status: 0 ok, 1 stencil rank / condition, 2 invalid gradient weights.
int build_link3(const Grid3 &g, const Links3 &Lk, int64_t ell, int64_t L, double dx, int32_t *indices,
                double *weights, int32_t *modes) {
    const int ctr[3] = {Lk.li[ell], Lk.lj[ell], Lk.lk[ell]};
    const double xb[3] = {Lk.xb[ell] / dx - 0.5, Lk.yb[ell] / dx - 0.5, Lk.zb[ell] / dx - 0.5};
    // Candidates: the 7x7x7 offsets in lexicographic order, kept when on the
    // grid, active and within r^2 <= 9.5; stable order by r^2.
    struct Cand {
        double r2, xi[3];
        int32_t id;
        int order;
    };
    // The G3 nearest candidates in stable order: an insertion goes after
    // every kept entry with r^2 <= its own, so equal distances keep the
    // offset order.
    Cand cand[G3];
    int nc = 0;
    for (int o = 0; o < 343; ++o) {
        const int off[3] = {o / 49 - 3, (o / 7) % 7 - 3, o % 7 - 3};
        const int node[3] = {ctr[0] + off[0], ctr[1] + off[1], ctr[2] + off[2]};
        if (!g.inside(node[0], node[1], node[2])) continue;
        const int32_t id = g.at(node[0], node[1], node[2]);
        if (id < 0) continue;
        Cand c;
        for (int a = 0; a < 3; ++a) c.xi[a] = (double)node[a] - xb[a];
        c.r2 = c.xi[0] * c.xi[0] + c.xi[1] * c.xi[1];
        c.r2 = c.r2 + c.xi[2] * c.xi[2];
        if (!(c.r2 <= 9.5)) continue;
        c.id = id;
        c.order = o;
        if (nc == G3 && !(c.r2 < cand[G3 - 1].r2)) continue;
        int pos = nc < G3 ? nc : G3 - 1;
        while (pos > 0 && c.r2 < cand[pos - 1].r2) {
            cand[pos] = cand[pos - 1];
            --pos;
        }
        cand[pos] = c;
        if (nc < G3) ++nc;
    }
    const int m = nc;
    double A[G3][10], W[G3];
    for (int s = 0; s < m; ++s) {
        const double x = cand[s].xi[0], y = cand[s].xi[1], z = cand[s].xi[2];
        const double row[10] = {1.0, x, y, z, 0.5 * x * x, x * y, x * z, 0.5 * y * y, y * z, 0.5 * z * z};
        for (int b = 0; b < 10; ++b) A[s][b] = row[b];
        W[s] = std::exp(-cand[s].r2 / (2 * 1.5 * 1.5)) + 1e-4;
    }
    double M[100];
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j) {
            double acc = 0.0;
            for (int s = 0; s < m; ++s) acc += A[s][i] * W[s] * A[s][j];
            M[i * 10 + j] = acc;
        }
    const double cond = sym_cond(M, 10);
    if (!(m >= 10 && std::isfinite(cond) && cond <= 1e8)) return 1;
    for (int i = 0; i < 10; ++i) M[i * 10 + i] += 1e-12;
    double B[10 * G3];
    for (int i = 0; i < 10; ++i)
        for (int s = 0; s < m; ++s) B[i * m + s] = A[s][i] * W[s];
    if (!lu_solve(M, 10, B, m)) return 1;
    double grads[3][G3];
    for (int r = 0; r < 3; ++r) {
        double sum = 0.0, sabs = 0.0;
        bool finite = true;
        for (int s = 0; s < m; ++s) {
            grads[r][s] = B[(1 + r) * m + s] / dx;
            finite = finite && std::isfinite(grads[r][s]);
            sum += grads[r][s];
            sabs += std::fabs(grads[r][s]);
        }
        if (!finite || !(std::fabs(sum) * dx < 1e-6) || !(sabs * dx <= 12.0)) return 2;
    }
    for (int s = 0; s < G3; ++s) {
        indices[(size_t)s * L + ell] = s < m ? cand[s].id : -1;
        for (int r = 0; r < 3; ++r) weights[((size_t)(1 + r) * S3 + s) * L + ell] = s < m ? grads[r][s] : 0.0;
        weights[((size_t)0 * S3 + s) * L + ell] = 0.0;
    }
    // Predictor: two probes on the inward normal (linear extrapolation),
    // else one probe, else a 1/2/3-node extrapolation along the link.
    const double nrm[3] = {Lk.nx[ell], Lk.ny[ell], Lk.nz[ell]};
    int32_t pp[P3];
    double ww[P3];
    for (int s = 0; s < P3; ++s) {
        pp[s] = -1;
        ww[s] = 0.0;
    }
    int mode = 0;
    const double pairs[4][2] = {{0.75, 1.75}, {1.0, 2.0}, {1.25, 2.5}, {1.5, 3.0}};
    for (int t = 0; t < 4 && mode == 0; ++t) {
        const double h1 = pairs[t][0], h2 = pairs[t][1];
        double p1[3], p2[3];
        for (int a = 0; a < 3; ++a) {
            p1[a] = xb[a] - h1 * nrm[a];
            p2[a] = xb[a] - h2 * nrm[a];
        }
        int32_t q1[8], q2[8];
        double w1[8], w2[8];
        const bool ok1 = probe3(g, p1, ctr, q1, w1);
        const bool ok2 = probe3(g, p2, ctr, q2, w2);
        if (ok1 && ok2) {
            const double c1 = h2 / (h2 - h1), c2 = -h1 / (h2 - h1);
            for (int c = 0; c < 8; ++c) {
                pp[c] = q1[c];
                ww[c] = c1 * w1[c];
                pp[8 + c] = q2[c];
                ww[8 + c] = c2 * w2[c];
            }
            mode = 2;
        }
    }
    const double singles[5] = {0.75, 1.0, 1.25, 1.5, 2.0};
    for (int t = 0; t < 5 && mode == 0; ++t) {
        double p1[3];
        for (int a = 0; a < 3; ++a) p1[a] = xb[a] - singles[t] * nrm[a];
        int32_t q1[8];
        double w1[8];
        if (probe3(g, p1, ctr, q1, w1)) {
            for (int c = 0; c < 8; ++c) {
                pp[c] = q1[c];
                ww[c] = w1[c];
            }
            mode = 1;
        }
    }
    if (mode == 0) {
        static const int dirs[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
        const int q = Lk.lq[ell];
        const int *st = dirs[(q >= 0 && q < 6) ? q : 0];
        int32_t chain[3];
        int len = 0;
        for (int k = 0; k < 3; ++k) {
            const int node[3] = {ctr[0] - k * st[0], ctr[1] - k * st[1], ctr[2] - k * st[2]};
            if (!g.inside(node[0], node[1], node[2])) break;
            const int32_t id = g.at(node[0], node[1], node[2]);
            if (id < 0) break;
            chain[len++] = id;
        }
        const double eta = Lk.eta[ell];
        double cw[3] = {1.0, 0.0, 0.0};
        if (len == 3) {
            cw[0] = 0.5 * (eta + 1) * (eta + 2);
            cw[1] = -eta * (eta + 2);
            cw[2] = 0.5 * eta * (eta + 1);
        } else if (len == 2) {
            cw[0] = 1 + eta;
            cw[1] = -eta;
        }
        for (int k = 0; k < len; ++k) {
            pp[k] = chain[k];
            ww[k] = cw[k];
        }
        mode = -len;
    }
    for (int s = 0; s < P3; ++s) {
        indices[(size_t)(G3 + s) * L + ell] = pp[s];
        weights[((size_t)0 * S3 + G3 + s) * L + ell] = ww[s];
        for (int r = 1; r < 4; ++r) weights[((size_t)r * S3 + G3 + s) * L + ell] = 0.0;
    }
    modes[ell] = mode;
    return 0;
}

// Tangent-derivative stencil of one 3D cut link: weighted quadratic LSQ
// over at most `S` active nodes within r^2 <= (radius + 1/4)^2 of the cut
// point (radius 3, then 4, then 5), Gaussian weights exp(-r^2/(2*1.75^2)) +
// 1e-4, rank >= 10 (singular values above 1e-10), cond <= 1e9, and the
// derivative rows along t1, t2 of the normal's basis (t1 = n x e_x, or n x e_y
// when |n_x| >= 0.75; t2 = n x t1).  Returns the sample count, 0 when no
// radius gives an acceptable fit.
int tangent_link3(const Grid3 &g, int i0, int j0, int k0, double xb, double yb, double zb, const double nin[3],
                  double dx, double dy, double dz, int S, int32_t *op, double *ow1, double *ow2) {
    double nrm = std::sqrt(nin[0] * nin[0] + nin[1] * nin[1] + nin[2] * nin[2]);
    if (!(nrm > std::numeric_limits<double>::epsilon())) nrm = std::numeric_limits<double>::epsilon();
    const double n[3] = {nin[0] / nrm, nin[1] / nrm, nin[2] / nrm};
    const double ref[3] = {std::fabs(n[0]) < 0.75 ? 1.0 : 0.0, std::fabs(n[0]) < 0.75 ? 0.0 : 1.0, 0.0};
    double t1[3] = {n[1] * ref[2] - n[2] * ref[1], n[2] * ref[0] - n[0] * ref[2], n[0] * ref[1] - n[1] * ref[0]};
    double r1 = std::sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    if (!(r1 > std::numeric_limits<double>::epsilon())) r1 = std::numeric_limits<double>::epsilon();
    for (double &v : t1) v = v / r1;
    const double t2[3] = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
    struct C {
        double r2;
        int32_t id;
        double rx, ry, rz;
    };
    std::vector<C> cand;
    std::vector<double> A, wd, M(100), B;
    for (int radius = 3; radius <= 5; ++radius) {
        cand.clear();
        const double r2max = (double(radius) + 0.25) * (double(radius) + 0.25);
        for (int ii = std::max(0, i0 - radius); ii < std::min(g.nx, i0 + radius + 1); ++ii)
            for (int jj = std::max(0, j0 - radius); jj < std::min(g.ny, j0 + radius + 1); ++jj)
                for (int kk = std::max(0, k0 - radius); kk < std::min(g.nz, k0 + radius + 1); ++kk) {
                    const int32_t id = g.at(ii, jj, kk);
                    if (id < 0) continue;
                    const double rx = ((double(ii) + 0.5) * dx - xb) / dx;
                    const double ry = ((double(jj) + 0.5) * dy - yb) / dy;
                    const double rz = ((double(kk) + 0.5) * dz - zb) / dz;
                    double r2 = rx * rx + ry * ry;
                    r2 = r2 + rz * rz;
                    if (r2 <= r2max) cand.push_back({r2, id, rx, ry, rz});
                }
        std::stable_sort(cand.begin(), cand.end(), [](const C &a, const C &b) { return a.r2 < b.r2; });
        if (cand.size() < 10) continue;
        const int m = std::min<int>((int)cand.size(), S);
        A.assign((size_t)m * 10, 0.0);
        wd.assign(m, 0.0);
        for (int s = 0; s < m; ++s) {
            const double x = cand[s].rx, y = cand[s].ry, z = cand[s].rz;
            const double row[10] = {1.0, x, y, z, 0.5 * x * x, x * y, x * z, 0.5 * y * y, y * z, 0.5 * z * z};
            for (int b = 0; b < 10; ++b) A[(size_t)s * 10 + b] = row[b];
            wd[s] = std::exp(-0.5 * cand[s].r2 / (1.75 * 1.75)) + 1.0e-4;
        }
        double sv[10];
        singular_values(A.data(), m, 10, sv);
        int rank = 0;
        for (double v : sv) rank += v > 1.0e-10 ? 1 : 0;
        if (rank < 10) continue;
        for (int i = 0; i < 10; ++i)
            for (int j = 0; j < 10; ++j) {
                double acc = 0.0;
                for (int s = 0; s < m; ++s) acc += A[(size_t)s * 10 + i] * (A[(size_t)s * 10 + j] * wd[s]);
                M[i * 10 + j] = acc;
            }
        const double cond = sym_cond(M.data(), 10);
        if (!std::isfinite(cond) || cond > 1.0e9) continue;
        for (int i = 0; i < 10; ++i) M[i * 10 + i] += 1.0e-12;
        B.assign((size_t)10 * m, 0.0);
        for (int i = 0; i < 10; ++i)
            for (int s = 0; s < m; ++s) B[(size_t)i * m + s] = A[(size_t)s * 10 + i] * wd[s];
        if (!lu_solve(M.data(), 10, B.data(), m)) continue;
        const double g1[3] = {t1[0] / dx, t1[1] / dy, t1[2] / dz}, g2[3] = {t2[0] / dx, t2[1] / dy, t2[2] / dz};
        const double h = std::max(dx, std::max(dy, dz));
        bool ok = true;
        double s1 = 0.0, s2 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int s = 0; s < m; ++s) {
            const double w1 = g1[0] * B[(size_t)1 * m + s] + g1[1] * B[(size_t)2 * m + s] + g1[2] * B[(size_t)3 * m + s];
            const double w2 = g2[0] * B[(size_t)1 * m + s] + g2[1] * B[(size_t)2 * m + s] + g2[2] * B[(size_t)3 * m + s];
            ok = ok && std::isfinite(w1) && std::isfinite(w2);
            s1 += w1;
            s2 += w2;
            a1 += std::fabs(w1);
            a2 += std::fabs(w2);
            ow1[s] = w1;
            ow2[s] = w2;
            op[s] = cand[s].id;
        }
        if (!ok || std::fabs(s1) * h > 1.0e-6 || std::fabs(s2) * h > 1.0e-6 || a1 * h > 12.0 || a2 * h > 12.0) {
            for (int s = 0; s < m; ++s) {
                op[s] = -1;
                ow1[s] = ow2[s] = 0.0;
            }
            continue;
        }
        return m;
    }
    return 0;
}

// ---------------------------------------------------------------- 2D -----

struct Grid2 {
    const int32_t *act;
    int nx, ny;
    bool active(int i, int j) const { return act[(size_t)i * ny + j] != 0; }
};

struct Node2 {
    int i, j;
    double w;
};

// Bilinear probe (curved_boundary_warp.py:369-385 restated): false when a
// corner is off the lattice or inactive.
bool bilinear2(const Grid2 &g, double x0g, double y0g, double dx, double xp, double yp, Node2 out[4]) {
    const double rx = (xp - x0g) / dx, ry = (yp - y0g) / dx;
    const double fi = std::floor(rx), fj = std::floor(ry);
    if (!(fi >= -1e9 && fi <= 1e9 && fj >= -1e9 && fj <= 1e9)) return false;
    const int i0 = (int)fi, j0 = (int)fj;
    const double sx = rx - (double)i0, sy = ry - (double)j0;
    if (i0 < 0 || j0 < 0 || i0 + 1 >= g.nx || j0 + 1 >= g.ny) return false;
    out[0] = {i0, j0, (1.0 - sx) * (1.0 - sy)};
    out[1] = {i0 + 1, j0, sx * (1.0 - sy)};
    out[2] = {i0, j0 + 1, (1.0 - sx) * sy};
    out[3] = {i0 + 1, j0 + 1, sx * sy};
    for (int c = 0; c < 4; ++c)
        if (!g.active(out[c].i, out[c].j)) return false;
    return true;
}

struct Cand2 {
    double r2;
    int i, j;
    double rx, ry;
};

// Weighted quadratic LSQ fit around the cut point (curved_boundary_warp.py:
// 221-248 restated): false when rank-deficient or ill-conditioned.
bool quadratic_fit2(const Grid2 &g, const double *xc, const double *yc, int i0, int j0, double xb, double yb,
                    double dx, double sigma, int max_stencil, std::vector<Cand2> &cand, std::vector<double> &coeff) {
    cand.clear();
    for (int ii = std::max(0, i0 - 3); ii < std::min(g.nx, i0 + 4); ++ii)
        for (int jj = std::max(0, j0 - 3); jj < std::min(g.ny, j0 + 4); ++jj) {
            if (!g.active(ii, jj)) continue;
            const double rx = (xc[ii] - xb) / dx, ry = (yc[jj] - yb) / dx;
            const double r2 = rx * rx + ry * ry;
            if (r2 <= 9.5) cand.push_back({r2, ii, jj, rx, ry});
        }
    std::stable_sort(cand.begin(), cand.end(), [](const Cand2 &a, const Cand2 &b) { return a.r2 < b.r2; });
    if ((int)cand.size() < 6) return false;
    if ((int)cand.size() > max_stencil) cand.resize(max_stencil);
    const int m = (int)cand.size();
    std::vector<double> A((size_t)m * 6), wd(m);
    for (int s = 0; s < m; ++s) {
        const double rx = cand[s].rx, ry = cand[s].ry;
        const double row[6] = {1.0, rx, ry, 0.5 * rx * rx, rx * ry, 0.5 * ry * ry};
        for (int b = 0; b < 6; ++b) A[(size_t)s * 6 + b] = row[b];
        wd[s] = std::exp(-0.5 * cand[s].r2 / (sigma * sigma)) + 1.0e-4;
    }
    double sv[6];
    singular_values(A.data(), m, 6, sv);
    int rank = 0;
    for (int b = 0; b < 6; ++b) rank += sv[b] > 1.0e-10 ? 1 : 0;
    if (rank < 6) return false;
    double M[36];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double acc = 0.0;
            for (int s = 0; s < m; ++s) acc += A[(size_t)s * 6 + i] * (A[(size_t)s * 6 + j] * wd[s]);
            M[i * 6 + j] = acc;
        }
    const double cond = sym_cond(M, 6);
    if (!std::isfinite(cond) || cond > 1.0e8) return false;
    for (int i = 0; i < 6; ++i) M[i * 6 + i] += 1.0e-12;
    coeff.assign((size_t)6 * m, 0.0);
    for (int i = 0; i < 6; ++i)
        for (int s = 0; s < m; ++s) coeff[(size_t)i * m + s] = A[(size_t)s * 6 + i] * wd[s];
    return lu_solve(M, 6, coeff.data(), m);
}

struct Out2 {
    int32_t *oi, *oj;
    double *w, *wt, *wx, *wy;
    int32_t *valid;
    int64_t L;
    int S;
    int32_t &I(int64_t l, int m) const { return oi[l * S + m]; }
    int32_t &J(int64_t l, int m) const { return oj[l * S + m]; }
};

// Sample slots: existing (i, j) slots are reused, new nodes appended; false
// (with the slots assigned so far kept) when the stencil would overflow
// (curved_boundary_warp.py:299-318 restated).
bool place2(const Out2 &o, int64_t ell, const std::vector<Cand2> &cand, double *const *dst,
            const std::vector<double> *const *src, int ncols) {
    int nxt = 0;
    for (int m = 0; m < o.S; ++m)
        if (o.I(ell, m) >= 0 && o.J(ell, m) >= 0) ++nxt;
    for (size_t m = 0; m < cand.size(); ++m) {
        int slot = -1;
        for (int s = 0; s < o.S; ++s)
            if (o.I(ell, s) == cand[m].i && o.J(ell, s) == cand[m].j && o.I(ell, s) >= 0 && o.J(ell, s) >= 0) {
                slot = s;
                break;
            }
        if (slot < 0) {
            if (nxt >= o.S) return false;
            slot = nxt++;
            o.I(ell, slot) = cand[m].i;
            o.J(ell, slot) = cand[m].j;
        }
        for (int c = 0; c < ncols; ++c) dst[c][ell * o.S + slot] = (*src[c])[m];
    }
    return true;
}

void build_link2(const Grid2 &g, const double *xc, const double *yc, const int32_t *li, const int32_t *lj,
                 const double *lxb, const double *lyb, const double *lnx, const double *lny, double dx,
                 const Out2 &o, int64_t ell) {
    const double x0g = g.nx > 0 ? xc[0] : 0.5 * dx, y0g = g.ny > 0 ? yc[0] : 0.5 * dx;
    const double xb = lxb[ell], yb = lyb[ell], nxb = lnx[ell], nyb = lny[ell];
    const int i0 = li[ell], j0 = lj[ell];
    // Value weights, accumulated per node in first-insertion order.
    std::vector<Node2> acc;
    auto add = [&](int i, int j, double w) {
        for (auto &a : acc)
            if (a.i == i && a.j == j) {
                a.w = a.w + w;
                return;
            }
        acc.push_back({i, j, 0.0 + w});
    };
    bool ok = false;
    const double pairs[4][2] = {{0.75, 1.75}, {1.0, 2.0}, {1.25, 2.5}, {1.5, 3.0}};
    for (int t = 0; t < 4 && !ok; ++t) {
        const double h1 = pairs[t][0], h2 = pairs[t][1];
        Node2 w1[4], w2[4];
        if (!bilinear2(g, x0g, y0g, dx, xb - h1 * dx * nxb, yb - h1 * dx * nyb, w1)) continue;
        if (!bilinear2(g, x0g, y0g, dx, xb - h2 * dx * nxb, yb - h2 * dx * nyb, w2)) continue;
        acc.clear();
        const double c1 = h2 / (h2 - h1), c2 = -h1 / (h2 - h1);
        for (auto &n : w1) add(n.i, n.j, c1 * n.w);
        for (auto &n : w2) add(n.i, n.j, c2 * n.w);
        ok = true;
    }
    const double singles[5] = {0.75, 1.0, 1.25, 1.5, 2.0};
    for (int t = 0; t < 5 && !ok; ++t) {
        Node2 w1[4];
        if (!bilinear2(g, x0g, y0g, dx, xb - singles[t] * dx * nxb, yb - singles[t] * dx * nyb, w1)) continue;
        acc.clear();
        for (auto &n : w1) add(n.i, n.j, n.w);
        ok = true;
    }
    if (!ok) return;
    std::vector<Node2> items;
    for (auto &a : acc)
        if (std::fabs(a.w) > 1.0e-14) items.push_back(a);
    std::stable_sort(items.begin(), items.end(), [&](const Node2 &a, const Node2 &b) {
        const int da = (a.i - i0) * (a.i - i0) + (a.j - j0) * (a.j - j0);
        const int db = (b.i - i0) * (b.i - i0) + (b.j - j0) * (b.j - j0);
        return da < db;
    });
    if ((int)items.size() > o.S) return;
    double sabs = 0.0;
    bool finite = true;
    for (auto &it : items) {
        finite = finite && std::isfinite(it.w);
        sabs += std::fabs(it.w);
    }
    if (items.empty() || !finite || sabs > 3.5) return;
    for (size_t m = 0; m < items.size(); ++m) {
        o.I(ell, (int)m) = items[m].i;
        o.J(ell, (int)m) = items[m].j;
        o.w[ell * o.S + m] = items[m].w;
    }
    std::vector<Cand2> cand;
    std::vector<double> coeff;
    if (quadratic_fit2(g, xc, yc, i0, j0, xb, yb, dx, 1.5, o.S, cand, coeff)) {
        const int m = (int)cand.size();
        const double tx = -nyb, ty = nxb;
        std::vector<double> wt(m), wx(m), wy(m);
        for (int s = 0; s < m; ++s) {
            wt[s] = (tx / dx) * coeff[(size_t)1 * m + s] + (ty / dx) * coeff[(size_t)2 * m + s];
            wx[s] = (1.0 / dx) * coeff[(size_t)1 * m + s];
            wy[s] = (1.0 / dx) * coeff[(size_t)2 * m + s];
        }
        auto sums = [&](const std::vector<double> &v, double &sum, double &sab, bool &fin) {
            sum = 0.0;
            sab = 0.0;
            fin = true;
            for (double x : v) {
                fin = fin && std::isfinite(x);
                sum += x;
                sab += std::fabs(x);
            }
        };
        double s0, a0;
        bool f0;
        sums(wt, s0, a0, f0);
        if (f0 && std::fabs(s0) * dx <= 1.0e-6 && a0 * dx <= 8.0) {
            double *dst[1] = {o.wt};
            const std::vector<double> *src[1] = {&wt};
            if (!place2(o, ell, cand, dst, src, 1))
                for (int s = 0; s < o.S; ++s) o.wt[ell * o.S + s] = 0.0;
        }
        double s1, a1, s2, a2;
        bool f1, f2;
        sums(wx, s1, a1, f1);
        sums(wy, s2, a2, f2);
        if (f1 && f2 && std::fabs(s1) * dx <= 1.0e-6 && std::fabs(s2) * dx <= 1.0e-6 && a1 * dx <= 12.0 &&
            a2 * dx <= 12.0) {
            double *dst[2] = {o.wx, o.wy};
            const std::vector<double> *src[2] = {&wx, &wy};
            if (!place2(o, ell, cand, dst, src, 2))
                for (int s = 0; s < o.S; ++s) {
                    o.wx[ell * o.S + s] = 0.0;
                    o.wy[ell * o.S + s] = 0.0;
                }
        }
    }
    o.valid[ell] = 1;
}

}  // namespace

extern "C" {

int vlbm_geom3d_char_weights(const int32_t *active_id, int32_t nx, int32_t ny, int32_t nz, int64_t n_links,
                             const int32_t *li, const int32_t *lj, const int32_t *lk, const int32_t *lq,
                             const double *xb, const double *yb, const double *zb, const double *lnx,
                             const double *lny, const double *lnz, const double *eta, double dx, int32_t nthreads,
                             int32_t *indices, double *weights, int32_t *modes, uint8_t *status) {
    if (!active_id || nx <= 0 || ny <= 0 || nz <= 0 || n_links < 0 || !(dx > 0.0)) return -1;
    if (n_links > 0 && (!li || !lj || !lk || !lq || !xb || !yb || !zb || !lnx || !lny || !lnz || !eta || !indices ||
                        !weights || !modes || !status))
        return -1;
    const Grid3 g{active_id, nx, ny, nz};
    const Links3 Lk{li, lj, lk, lq, xb, yb, zb, lnx, lny, lnz, eta};
    parallel_for(n_links, thread_count(nthreads, n_links), [&](int64_t b, int64_t e) {
        for (int64_t l = b; l < e; ++l) {
            for (int s = 0; s < S3; ++s) {
                indices[(size_t)s * n_links + l] = -1;
                for (int r = 0; r < 4; ++r) weights[((size_t)r * S3 + s) * n_links + l] = 0.0;
            }
            modes[l] = 0;
            status[l] = (uint8_t)build_link3(g, Lk, l, n_links, dx, indices, weights, modes);
        }
    });
    return 0;
}

int vlbm_geom3d_tangent_weights(const int32_t *active_id, int32_t nx, int32_t ny, int32_t nz, int64_t n_links,
                                const int32_t *li, const int32_t *lj, const int32_t *lk, const double *xb,
                                const double *yb, const double *zb, const double *lnx, const double *lny,
                                const double *lnz, double dx, double dy, double dz, int32_t max_stencil,
                                int32_t nthreads, int32_t *out_p, double *wt1, double *wt2, int32_t *valid) {
    if (!active_id || nx <= 0 || ny <= 0 || nz <= 0 || n_links < 0 || max_stencil < 10 || !(dx > 0.0) ||
        !(dy > 0.0) || !(dz > 0.0))
        return -1;
    if (n_links > 0 && (!li || !lj || !lk || !xb || !yb || !zb || !lnx || !lny || !lnz || !out_p || !wt1 || !wt2 ||
                        !valid))
        return -1;
    const Grid3 g{active_id, nx, ny, nz};
    const int S = max_stencil;
    parallel_for(n_links, thread_count(nthreads, n_links), [&](int64_t b, int64_t e) {
        for (int64_t l = b; l < e; ++l) {
            int32_t *op = out_p + l * S;
            double *w1 = wt1 + l * S, *w2 = wt2 + l * S;
            for (int s = 0; s < S; ++s) {
                op[s] = -1;
                w1[s] = w2[s] = 0.0;
            }
            const double nv[3] = {lnx[l], lny[l], lnz[l]};
            valid[l] = tangent_link3(g, li[l], lj[l], lk[l], xb[l], yb[l], zb[l], nv, dx, dy, dz, S, op, w1, w2) > 0;
        }
    });
    return 0;
}

int vlbm_geom2d_lsq_weights(const int32_t *active, int32_t nx, int32_t ny, const double *x_coords,
                            const double *y_coords, int64_t n_links, const int32_t *li, const int32_t *lj,
                            const double *xb, const double *yb, const double *lnx, const double *lny, double dx,
                            int32_t max_stencil, int32_t nthreads, int32_t *out_i, int32_t *out_j, double *w,
                            double *wt, double *wx, double *wy, int32_t *valid) {
    if (!active || nx <= 0 || ny <= 0 || !x_coords || !y_coords || n_links < 0 || max_stencil <= 0) return -1;
    if (n_links > 0 && (!li || !lj || !xb || !yb || !lnx || !lny || !out_i || !out_j || !w || !wt || !wx || !wy ||
                        !valid))
        return -1;
    const Grid2 g{active, nx, ny};
    const Out2 o{out_i, out_j, w, wt, wx, wy, valid, n_links, max_stencil};
    parallel_for(n_links, thread_count(nthreads, n_links), [&](int64_t b, int64_t e) {
        for (int64_t l = b; l < e; ++l) {
            for (int s = 0; s < max_stencil; ++s) {
                o.I(l, s) = -1;
                o.J(l, s) = -1;
                w[l * max_stencil + s] = wt[l * max_stencil + s] = 0.0;
                wx[l * max_stencil + s] = wy[l * max_stencil + s] = 0.0;
            }
            valid[l] = 0;
            build_link2(g, x_coords, y_coords, li, lj, xb, yb, lnx, lny, dx, o, l);
        }
    });
    return 0;
}

}  // extern "C"
