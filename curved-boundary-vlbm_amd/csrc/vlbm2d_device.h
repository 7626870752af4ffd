// Device-side constitutive, boundary-value and traction-solve functions for
// the D2Q4x6 solver (gfx950).  2x2 matrices are row-major double[4]; states
// are U = [vx, vy, F11, F12, F21, F22].
//
// Reference behaviour restated here (reference file abbreviations:
// cbw2d = src/curved_lbm/two_d/curved_boundary_warp.py,
// vnew2d = src/curved_lbm/two_d/vector_nonlinear_elastic_warp.py):
//   piola2          vnew2d:938-1052   (_piola_vec, all seven plane-strain laws)
//   flux / feq      vnew2d:1055-1168  (_state_comp, _flux_x/y_comp, _equilibrium_comp)
//   source2         vnew2d:1250-1280  (_source_comp)
//   value2          vnew2d:1171-1247 + cbw2d:936-984 (+ linear-ramp radial mode 105)
//   solve_normal2   cbw2d:987-1192    (_normal_traction_residual_arbitrary,
//                                      _neo_hooke_acoustic_normal_2d, _solve_normal_image_arbitrary)
//   flux_jvp2       cbw2d:1224-1365   (_jvp_eps6, _flux_x/y_jvp_comp)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vlbm2 {

constexpr int NC = 6;
constexpr int NQ = 4;
constexpr int FQ = 24;
constexpr int MAX_IDS = 8;
constexpr int MAX_LSQ = 24;

enum { M_SVK = 0, M_NH = 1, M_LOGNH = 2, M_MR = 3, M_YEOH = 4, M_GENT = 5, M_SUGI = 6 };
enum { BC_DIR = 1, BC_NEU = 2 };
enum { SRC_NONE = 0, SRC_CONST = 1, SRC_DAMP = 2 };
enum {
    R_BFL = 0, R_QBFL = 1, R_LSQ_BFL = 2, R_REG_BFL = 3, R_RNEE = 4, R_PAIR_NEE = 5, R_SAME_NEE = 6, R_POLY = 7,
    R_LOWQ_REG_BFL = 8, R_COMPAT_BFL = 9, R_COMPAT_NEE = 10, R_CHAR2 = 11, R_LOCALF = 12, R_LOCALF_BFL = 13
};
// Reconstructions handled by the extended link kernel (regularized, NEE,
// quadratic and polynomial bounce-back forms).
__host__ __device__ inline bool recon_ext(int r) {
    return r == R_QBFL || r == R_REG_BFL || r == R_RNEE || r == R_PAIR_NEE || r == R_SAME_NEE || r == R_POLY ||
           r == R_LOWQ_REG_BFL || r == R_COMPAT_NEE;
}
// Reconstructions that read the LSQ gather of the cut points, and those of
// them with the geometry-compatible tangential projection.
__host__ __device__ inline bool recon_gather(int r) {
    return r == R_LOCALF || r == R_LSQ_BFL || r == R_COMPAT_BFL || r == R_CHAR2 || r == R_COMPAT_NEE;
}
__host__ __device__ inline bool recon_compat(int r) { return r == R_COMPAT_BFL || r == R_COMPAT_NEE || r == R_CHAR2; }
enum { NJ_FD = 0, NJ_AC = 1, NS_ENERGY = 2 };
enum {
    V_ZERO = 0, V_CONST = 1, V_RAMP = 2, V_SHEAR_X = 3, V_RICKER_Y = 4, V_SINE2_HOLD = 5,
    V_AFFINE_VEL = 6, CV_PIOLA_N = 101, CV_RADIAL = 102, CV_RADIAL_SINE2_HOLD = 103,
    CV_RADIAL_SINE2_PULSE = 104, CV_RADIAL_LINEAR = 105
};

constexpr double PI_D = 3.141592653589793238462643383279502884;

struct Material2 {
    int model;
    double lam, mu, p0, p1, p2, jf;
};

// Lattice directions E, N, W, S.
__device__ __forceinline__ int cx(int q) { return q == 0 ? 1 : (q == 2 ? -1 : 0); }
__device__ __forceinline__ int cy(int q) { return q == 1 ? 1 : (q == 3 ? -1 : 0); }

__device__ __forceinline__ double clamp_j(double j, double f) { return j < f ? f : j; }

// First Piola stress of the 2D laws, evaluated without FMA contraction (each
// product and sum rounded, as the FP64 oracle is built): the characteristic
// closure differentiates it by central differences with step ~1e-6.
__device__ __forceinline__ void piola2(const double *F, const Material2 &m, double *P) {
#pragma clang fp contract(off)
    const double F11 = F[0], F12 = F[1], F21 = F[2], F22 = F[3];
    const double lam = m.lam, mu = m.mu;
    if (m.model == M_SVK) {
        const double C00 = F11 * F11 + F21 * F21, C01 = F11 * F12 + F21 * F22, C11 = F12 * F12 + F22 * F22;
        const double E00 = 0.5 * (C00 - 1.0), E01 = 0.5 * C01, E11 = 0.5 * (C11 - 1.0);
        const double trE = E00 + E11;
        const double S00 = lam * trE + 2.0 * mu * E00, S01 = 2.0 * mu * E01, S11 = lam * trE + 2.0 * mu * E11;
        P[0] = F11 * S00 + F12 * S01;
        P[1] = F11 * S01 + F12 * S11;
        P[2] = F21 * S00 + F22 * S01;
        P[3] = F21 * S01 + F22 * S11;
        return;
    }
    const double J = clamp_j(F11 * F22 - F12 * F21, m.jf);
    const double H[4] = {F22 / J, -F21 / J, -F12 / J, F11 / J};
    if (m.model == M_NH) {
        const double vol = 0.5 * lam * (J * J - 1.0);
#pragma unroll
        for (int k = 0; k < 4; ++k) P[k] = mu * (F[k] - H[k]) + vol * H[k];
        return;
    }
    const double logJ = log(J);
    const double vol = lam * logJ;
    if (m.model != M_MR && m.model != M_SUGI && m.model != M_YEOH && m.model != M_GENT) {
#pragma unroll
        for (int k = 0; k < 4; ++k) P[k] = mu * (F[k] - H[k]) + vol * H[k];
        return;
    }
    const double F2 = F11 * F11 + F12 * F12 + F21 * F21 + F22 * F22;
    const double I1 = F2 + 1.0, I2 = J * J + F2;
    const double jm23 = exp((-2.0 / 3.0) * logJ), jm43 = exp((-4.0 / 3.0) * logJ);
    const double h1 = (-2.0 / 3.0) * I1, h2 = 2.0 * J * J - (4.0 / 3.0) * I2;
    double G1[4], G2[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        G1[k] = jm23 * (2.0 * F[k] + h1 * H[k]);
        G2[k] = jm43 * (2.0 * F[k] + h2 * H[k]);
    }
    if (m.model == M_SUGI) {
        const double A = jm23 * I1 - 3.0;
        const double c2 = -0.5 * mu, c3 = (m.p0 + 2.0 * mu) / 8.0;
        const double sc = mu + 2.0 * c3 * A;
#pragma unroll
        for (int k = 0; k < 4; ++k) P[k] = sc * G1[k] + c2 * G2[k] + vol * H[k];
        return;
    }
    if (m.model == M_MR) {
#pragma unroll
        for (int k = 0; k < 4; ++k) P[k] = m.p0 * G1[k] + m.p1 * G2[k] + vol * H[k];
        return;
    }
    const double A = jm23 * I1 - 3.0;
    double sc;
    if (m.model == M_YEOH) {
        sc = m.p0 + 2.0 * m.p1 * A + 3.0 * m.p2 * A * A;
    } else {
        double den = 1.0 - A / m.p0;
        if (den < m.jf) den = m.jf;
        sc = 0.5 * mu / den;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) P[k] = sc * G1[k] + vol * H[k];
}

// Flux of component a along x (axis 0) or y (axis 1): G_x = (-P col 0,
// -v x e_x), G_y = (-P col 1, -v x e_y).
__device__ __forceinline__ double flux2(int axis, int a, const double *U, const double *P) {
    if (a < 2) return -P[2 * a + axis];
    const int k = a - 2;  // F index, row k/2, column k%2
    return (k % 2 == axis) ? -U[k / 2] : 0.0;
}

__device__ __forceinline__ double feq2(int q, int a, const double *U, const double *P, double c) {
    const double ex = (double)cx(q), ey = (double)cy(q);
    return 0.25 * (U[a] + (2.0 / c) * (ex * flux2(0, a, U, P) + ey * flux2(1, a, U, P)));
}

__device__ __forceinline__ double source2(int a, int mode, const double *sp, double vx, double vy) {
    if (mode == SRC_CONST) return sp[a];
    if (mode == SRC_DAMP) {
        if (a == 0) return -sp[0] * vx;
        if (a == 1) return -sp[0] * vy;
    }
    return 0.0;
}

__device__ __forceinline__ double smooth5(double z) {
    z = z < 0.0 ? 0.0 : (z > 1.0 ? 1.0 : z);
    return z * z * z * (10.0 - 15.0 * z + 6.0 * z * z);
}

// Boundary datum of component comp: a velocity for Dirichlet links, a
// nominal traction for Neumann links.  Mode 105 is nominal traction
// amp * min(t / t_ramp, 1) * n with params (amp, t_ramp).
__device__ __forceinline__ double value2(int comp, int mode, const double *p, double x, double y, double nxb,
                                double nyb, double t) {
    switch (mode) {
    case CV_PIOLA_N:
        return comp == 0 ? p[0] * nxb + p[1] * nyb : p[2] * nxb + p[3] * nyb;
    case CV_RADIAL:
        return comp == 0 ? p[0] * nxb : p[0] * nyb;
    case CV_RADIAL_SINE2_HOLD: {
        double scale = 1.0;
        if (p[1] > 0.0 && t < p[1]) {
            double den = p[2];
            if (den <= 0.0) den = 2.0 * p[1];
            const double s = sin(PI_D * t / den);
            scale = s * s;
        }
        return comp == 0 ? p[0] * scale * nxb : p[0] * scale * nyb;
    }
    case CV_RADIAL_SINE2_PULSE: {
        double scale = 0.0;
        if (p[1] > 0.0 && t < p[1]) {
            const double s = sin(PI_D * t / p[1]);
            scale = s * s;
        }
        return comp == 0 ? p[0] * scale * nxb : p[0] * scale * nyb;
    }
    case CV_RADIAL_LINEAR: {
        double scale = 1.0;
        if (p[1] > 0.0 && t < p[1]) scale = t / p[1];
        return comp == 0 ? p[0] * scale * nxb : p[0] * scale * nyb;
    }
    case V_CONST:
        return comp == 0 ? p[0] : p[1];
    case V_RAMP: {
        double scale = 1.0;
        if (p[2] > 0.0 && t < p[2]) {
            const double s = sin(PI_D * t / p[2]);
            scale = s * s;
        }
        return comp == 0 ? p[0] * scale : p[1] * scale;
    }
    case V_SINE2_HOLD: {
        double scale = 1.0;
        if (p[2] > 0.0 && t < p[2]) {
            double den = p[3];
            if (den <= 0.0) den = 2.0 * p[2];
            const double s = sin(PI_D * t / den);
            scale = s * s;
        }
        return comp == 0 ? p[0] * scale : p[1] * scale;
    }
    case V_SHEAR_X:
        return (comp == 0 && t < p[1]) ? p[0] * sin(PI_D * t / p[1]) : 0.0;
    case V_RICKER_Y: {
        if (comp != 1 || t > p[3]) return 0.0;
        const double z = (t - p[2]) / p[1];
        const double pref = 2.0 * p[0] / (sqrt(3.0 * p[1]) * sqrt(sqrt(PI_D)));
        double value = pref * (1.0 - z * z) * exp(-0.5 * z * z);
        if (p[6] > 0.0) {
            double layer = p[6] * p[5];
            if (layer < 1.0e-30) layer = 1.0e-30;
            double window = smooth5(y / layer) * smooth5((p[4] - y) / layer);
            if (p[7] > 1.0e-14) window = window / p[7];
            value = value * window;
        }
        return value;
    }
    case V_AFFINE_VEL:
        if (p[6] > 0.0 && t > p[6]) return 0.0;
        return comp == 0 ? p[0] * x + p[1] * y + p[2] : p[3] * x + p[4] * y + p[5];
    default:
        return 0.0;
    }
}

// Traction solve of an axis-aligned face of the rectangular box solver
// (vnew2d:1284-1448, _normal_residual + _solve_normal_column): Newton with a
// forward-difference Jacobian (eps 1e-6) on the F column of the normal
// axis, residual sign * P column - T below 1e-10, at most 12 iterations,
// each step halved up to 10 times until J > j_floor.  F holds the node's F
// on entry and the face F on exit.
__device__ inline void box_residual2(const double *F, int axis, double sign, double tx, double ty,
                                     const Material2 &m, double *r) {
    double P[4];
    piola2(F, m, P);
    r[0] = sign * P[axis] - tx;
    r[1] = sign * P[2 + axis] - ty;
}
__device__ inline void solve_normal_column2(double *F, int axis, double sign, double tx, double ty,
                                            const Material2 &m) {
    const double Fin[4] = {F[0], F[1], F[2], F[3]};
    double x0 = Fin[axis], x1 = Fin[2 + axis];
    const double eps = 1.0e-6;
    for (int it = 0; it < 12; ++it) {
        F[axis] = x0;
        F[2 + axis] = x1;
        double r[2];
        box_residual2(F, axis, sign, tx, ty, m, r);
        if (sqrt(r[0] * r[0] + r[1] * r[1]) < 1.0e-10) break;
        double Fp[4] = {F[0], F[1], F[2], F[3]}, rp0[2], rp1[2];
        Fp[axis] = F[axis] + eps;
        box_residual2(Fp, axis, sign, tx, ty, m, rp0);
        Fp[axis] = F[axis];
        Fp[2 + axis] = F[2 + axis] + eps;
        box_residual2(Fp, axis, sign, tx, ty, m, rp1);
        const double a = (rp0[0] - r[0]) / eps, c = (rp0[1] - r[1]) / eps;
        const double b = (rp1[0] - r[0]) / eps, d = (rp1[1] - r[1]) / eps;
        const double det = a * d - b * c;
        if (fabs(det) < 1.0e-14) break;
        const double d0 = (-r[0] * d + b * r[1]) / det, d1 = (c * r[0] - a * r[1]) / det;
        double step = 1.0;
        for (int ls = 0; ls < 10; ++ls) {
            const double c0 = x0 + step * d0, c1 = x1 + step * d1;
            double T[4] = {Fin[0], Fin[1], Fin[2], Fin[3]};
            T[axis] = c0;
            T[2 + axis] = c1;
            if (T[0] * T[3] - T[1] * T[2] > m.jf) {
                x0 = c0;
                x1 = c1;
                break;
            }
            step = 0.5 * step;
        }
    }
    F[axis] = x0;
    F[2 + axis] = x1;
}

__device__ __forceinline__ void normal_residual2(double hn0, double hn1, double gt0, double gt1, double nxb, double nyb,
                                        double tx, double ty, const Material2 &m, double *r) {
    const double t1 = -nyb, t2 = nxb;
    const double F[4] = {hn0 * nxb + gt0 * t1, hn0 * nyb + gt0 * t2, hn1 * nxb + gt1 * t1, hn1 * nyb + gt1 * t2};
    double P[4];
    piola2(F, m, P);
    r[0] = P[0] * nxb + P[1] * nyb - tx;
    r[1] = P[2] * nxb + P[3] * nyb - ty;
}

// Normal image of the boundary F: keep F t, solve P(F) n = T for h = F n
// (closed-form energy minimiser for neo-Hooke when selected, else Newton
// with the acoustic tensor or forward differences, 16 iterations, tol 1e-11,
// 12 line-search halvings keeping J above the floor).
__device__ __forceinline__ void solve_normal2(const double *Fg, double nxb, double nyb, double tx, double ty, int solver,
                                     const Material2 &m, double *Fb) {
    const double jf = m.jf;
    if (solver == NS_ENERGY && m.model == M_NH) {
        const double a0 = Fg[3] * nxb - Fg[2] * nyb;
        const double a1 = -Fg[1] * nxb + Fg[0] * nyb;
        const double a2 = a0 * a0 + a1 * a1;
        if (a2 > 1.0e-28) {
            const double area = sqrt(a2);
            const double e0 = a0 / area, e1 = a1 / area;
            const double tau = tx * e0 + ty * e1;
            const double aa = m.mu + 0.5 * m.lam * a2;
            const double bb = m.mu + 0.5 * m.lam;
            const double disc = sqrt(tau * tau + 4.0 * aa * bb);
            double sr = (tau + disc) / (2.0 * aa);
            if (tau < 0.0) sr = 2.0 * bb / (disc - tau);
            if (area * sr > jf) {
                const double h0 = (tx - tau * e0) / m.mu + sr * e0;
                const double h1 = (ty - tau * e1) / m.mu + sr * e1;
                const double d0 = h0 - (Fg[0] * nxb + Fg[1] * nyb);
                const double d1 = h1 - (Fg[2] * nxb + Fg[3] * nyb);
                Fb[0] = Fg[0] + d0 * nxb;
                Fb[1] = Fg[1] + d0 * nyb;
                Fb[2] = Fg[2] + d1 * nxb;
                Fb[3] = Fg[3] + d1 * nyb;
                return;
            }
        }
    }
    const double tanx = -nyb, tany = nxb;
    double hn0 = Fg[0] * nxb + Fg[1] * nyb, hn1 = Fg[2] * nxb + Fg[3] * nyb;
    const double gt0 = Fg[0] * tanx + Fg[1] * tany, gt1 = Fg[2] * tanx + Fg[3] * tany;
    const double eps = 1.0e-6;
    for (int it = 0; it < 16; ++it) {
        double r[2];
        normal_residual2(hn0, hn1, gt0, gt1, nxb, nyb, tx, ty, m, r);
        if (sqrt(r[0] * r[0] + r[1] * r[1]) < 1.0e-11) break;
        double a, b, c, d;
        if (solver != NJ_FD && m.model == M_NH) {
            const double F11 = hn0 * nxb + gt0 * tanx, F12 = hn0 * nyb + gt0 * tany;
            const double F21 = hn1 * nxb + gt1 * tanx, F22 = hn1 * nyb + gt1 * tany;
            const double J = clamp_j(F11 * F22 - F12 * F21, jf);
            const double H00 = F22 / J, H01 = -F21 / J, H10 = -F12 / J, H11 = F11 / J;
            const double beta = 0.5 * m.lam * (J * J - 1.0) - m.mu;
            const double coeff = m.lam * J * J - beta;
            const double n2 = nxb * nxb + nyb * nyb;
            const double v0 = H00 * nxb + H01 * nyb, v1 = H10 * nxb + H11 * nyb;
            a = m.mu * n2 + coeff * v0 * v0;
            b = coeff * v0 * v1;
            c = coeff * v1 * v0;
            d = m.mu * n2 + coeff * v1 * v1;
        } else {
            double rp0[2], rp1[2];
            normal_residual2(hn0 + eps, hn1, gt0, gt1, nxb, nyb, tx, ty, m, rp0);
            normal_residual2(hn0, hn1 + eps, gt0, gt1, nxb, nyb, tx, ty, m, rp1);
            a = (rp0[0] - r[0]) / eps;
            c = (rp0[1] - r[1]) / eps;
            b = (rp1[0] - r[0]) / eps;
            d = (rp1[1] - r[1]) / eps;
        }
        const double det = a * d - b * c;
        if (fabs(det) < 1.0e-14) break;
        const double d0 = (-r[0] * d + b * r[1]) / det;
        const double d1 = (c * r[0] - a * r[1]) / det;
        double step = 1.0;
        for (int ls = 0; ls < 12; ++ls) {
            const double c0 = hn0 + step * d0, c1 = hn1 + step * d1;
            const double F11 = c0 * nxb + gt0 * tanx, F12 = c0 * nyb + gt0 * tany;
            const double F21 = c1 * nxb + gt1 * tanx, F22 = c1 * nyb + gt1 * tany;
            if (F11 * F22 - F12 * F21 > jf) {
                hn0 = c0;
                hn1 = c1;
                break;
            }
            step = 0.5 * step;
        }
    }
    Fb[0] = hn0 * nxb + gt0 * tanx;
    Fb[1] = hn0 * nyb + gt0 * tany;
    Fb[2] = hn1 * nxb + gt1 * tanx;
    Fb[3] = hn1 * nyb + gt1 * tany;
}

// Central-difference directional derivative of the axis flux of every
// component at U along W, step 1e-6 / max(1, max|W|).
__device__ __forceinline__ void flux_jvp2(int axis, const double *U, const double *W, const Material2 &m, double *out) {
    double sc = fabs(W[0]);
#pragma unroll
    for (int k = 1; k < 6; ++k) sc = fabs(W[k]) > sc ? fabs(W[k]) : sc;
    double eps = 1.0e-6;
    if (sc > 1.0) eps = eps / sc;
    double Up[6], Um[6], Pp[4], Pm[4];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        Up[k] = U[k] + eps * W[k];
        Um[k] = U[k] - eps * W[k];
    }
    piola2(Up + 2, m, Pp);
    piola2(Um + 2, m, Pm);
#pragma unroll
    for (int a = 0; a < 6; ++a) out[a] = (flux2(axis, a, Up, Pp) - flux2(axis, a, Um, Pm)) / (2.0 * eps);
}

}  // namespace vlbm2
