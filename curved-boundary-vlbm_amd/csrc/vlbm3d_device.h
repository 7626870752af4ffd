// Device-side constitutive, boundary-value and traction-solve functions for
// the D3Q6x12 solver (gfx950).  Matrices are row-major double[9].
//
// Reference behaviour restated here (reference file abbreviations:
// cbw3d = src/curved_lbm/three_d/curved_boundary_warp3d.py,
// vnew3d = src/curved_lbm/three_d/vector_nonlinear_elastic_warp3d.py):
//   piola            vnew3d:793-1022   (_piola_mat, all seven material laws)
//   curved_value     cbw3d:996-1133    (_curved_value_comp) + linear-ramp mode 208
//   dirichlet_disp   cbw3d:2334-2399   (_local_sine2_integral, _local_dirichlet_displacement_comp)
//   solve_Fb         cbw3d:1333-1737   (_solve_Fb_arbitrary_normal / _componentwise_normal)
//   acoustic_normal  cbw3d:1206-1246
//   energy_Fb        cbw3d:1250-1329
//   char_population  cbw3d:2729-2809   (_state_flux_jvp3, _local_characteristic_population3)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vlbm {

constexpr int NC = 12;
constexpr int NQ = 6;
constexpr int FQ = 72;
constexpr int MAX_IDS = 16;  // boundary ids
constexpr int NPRM = 16;     // parameters per boundary id (row stride of params_by_id)
constexpr int MAX_CHAR = 64;
constexpr int MAX_TAN = 32;  // compat_bfl tangent stencil length

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

// The x-axis tube modes are the z-axis ones in the cyclically relabelled
// frame (a, b, axial) = (y, z, x): component comp of the x form is
// component tube_comp(comp) of the z form evaluated at (y, z).
__host__ __device__ __forceinline__ int tube_comp(int comp) { return comp == 0 ? 2 : comp - 1; }

constexpr double PI_D = 3.141592653589793238462643383279502884;

struct Material {
    int model;
    double lam, mu, p0, p1, p2, jf;
};

__device__ __forceinline__ int dir_comp(int q, int axis) {
    return (q == axis) ? 1 : ((q == axis + 3) ? -1 : 0);
}
__device__ __forceinline__ int opp(int q) { return q < 3 ? q + 3 : q - 3; }

__device__ __forceinline__ double det3(const double *F) {
    return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
           F[2] * (F[3] * F[7] - F[4] * F[6]);
}

// First Piola stress of the seven compressible laws; J is clamped to jf for
// every law except SVK.
__device__ inline void piola(const double *F, const Material &m, double *P) {
    const double F11 = F[0], F12 = F[1], F13 = F[2], F21 = F[3], F22 = F[4], F23 = F[5],
                 F31 = F[6], F32 = F[7], F33 = F[8];
    if (m.model == M_SVK) {
        const double C00 = F11 * F11 + F21 * F21 + F31 * F31, C01 = F11 * F12 + F21 * F22 + F31 * F32;
        const double C02 = F11 * F13 + F21 * F23 + F31 * F33, C11 = F12 * F12 + F22 * F22 + F32 * F32;
        const double C12 = F12 * F13 + F22 * F23 + F32 * F33, C22 = F13 * F13 + F23 * F23 + F33 * F33;
        const double E00 = 0.5 * (C00 - 1.0), E11 = 0.5 * (C11 - 1.0), E22 = 0.5 * (C22 - 1.0);
        const double trE = E00 + E11 + E22;
        const double S00 = m.lam * trE + 2.0 * m.mu * E00, S01 = 2.0 * m.mu * (0.5 * C01);
        const double S02 = 2.0 * m.mu * (0.5 * C02), S11 = m.lam * trE + 2.0 * m.mu * E11;
        const double S12 = 2.0 * m.mu * (0.5 * C12), S22 = m.lam * trE + 2.0 * m.mu * E22;
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
    if (J < m.jf) J = m.jf;
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
    if (m.model == M_NH) {
        const double vol = 0.5 * m.lam * (J * J - 1.0);
#pragma unroll
        for (int k = 0; k < 9; ++k) P[k] = m.mu * (F[k] - H[k]) + vol * H[k];
        return;
    }
    const double logJ = log(J);
    if (m.model == M_LOGNH) {
        const double vol = m.lam * logJ;
#pragma unroll
        for (int k = 0; k < 9; ++k) P[k] = m.mu * (F[k] - H[k]) + vol * H[k];
        return;
    }
    const double C00 = F11 * F11 + F21 * F21 + F31 * F31, C01 = F11 * F12 + F21 * F22 + F31 * F32;
    const double C02 = F11 * F13 + F21 * F23 + F31 * F33, C11 = F12 * F12 + F22 * F22 + F32 * F32;
    const double C12 = F12 * F13 + F22 * F23 + F32 * F33, C22 = F13 * F13 + F23 * F23 + F33 * F33;
    const double I1 = C00 + C11 + C22;
    const double jm23 = exp((-2.0 / 3.0) * logJ);
    double G1[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) G1[k] = jm23 * (2.0 * F[k] - (2.0 / 3.0) * I1 * H[k]);
    if (m.model == M_SPHNH) {
        const double vol = 0.5 * m.lam * (J * J - 1.0), sc = 0.5 * m.mu;
#pragma unroll
        for (int k = 0; k < 9; ++k) P[k] = sc * G1[k] + vol * H[k];
        return;
    }
    const double vol = m.lam * logJ;
    if (m.model == M_MR) {
        const double trC2 = C00 * C00 + C11 * C11 + C22 * C22 + 2.0 * (C01 * C01 + C02 * C02 + C12 * C12);
        const double I2 = 0.5 * (I1 * I1 - trC2);
        const double jm43 = exp((-4.0 / 3.0) * logJ);
        const double Cm[9] = {C00, C01, C02, C01, C11, C12, C02, C12, C22};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int k = 3 * i + j;
                const double FC = F[3 * i] * Cm[j] + F[3 * i + 1] * Cm[3 + j] + F[3 * i + 2] * Cm[6 + j];
                const double G2 = jm43 * (2.0 * (I1 * F[k] - FC) - (4.0 / 3.0) * I2 * H[k]);
                P[k] = m.p0 * G1[k] + m.p1 * G2 + vol * H[k];
            }
        return;
    }
    double sc = 0.0;
    if (m.model == M_YEOH) {
        const double A = jm23 * I1 - 3.0;
        sc = m.p0 + 2.0 * m.p1 * A + 3.0 * m.p2 * A * A;
    } else if (m.model == M_GENT) {
        const double A = jm23 * I1 - 3.0;
        double denom = 1.0 - A / m.p0;
        if (denom < m.jf) denom = m.jf;
        sc = 0.5 * m.mu / denom;
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) P[k] = 0.0;
        return;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) P[k] = sc * G1[k] + vol * H[k];
}

__device__ __forceinline__ double source_comp(int a, int mode, const double *prm, double vx,
                                              double vy, double vz) {
    if (mode == SRC_CONST) return prm[a];
    if (mode == SRC_DAMP) {
        if (a == 0) return -prm[0] * vx;
        if (a == 1) return -prm[0] * vy;
        if (a == 2) return -prm[0] * vz;
    }
    return 0.0;
}

__device__ inline double sine2_integral(double time, double active, double denom) {
    if (active <= 0.0) return time;
    if (denom <= 0.0) denom = active;
    const double t = time > 0.0 ? time : 0.0;
    if (t < active) return 0.5 * t - denom * sin(2.0 * PI_D * t / denom) / (4.0 * PI_D);
    const double prefix = 0.5 * active - denom * sin(2.0 * PI_D * active / denom) / (4.0 * PI_D);
    return prefix + (t - active);
}

// Exact finite twist about the axis through (prm[0], prm[1]) of the
// transverse plane (a, b) plus axial stretch, as a displacement: comp 0, 1
// transverse, 2 axial (cbw3d:2384-2398).
__device__ inline double twist_disp(const double *prm, int comp, double a, double b, double time) {
    const double active = prm[4];
    double sv = 1.0;
    if (active > 0.0 && time < active) sv = 0.5 * (1.0 - cos(PI_D * time / active));
    const double theta = prm[3] * sv, da = a - prm[0], db = b - prm[1];
    const double c = cos(theta), s = sin(theta);
    if (comp == 0) return (c - 1.0) * da - s * db;
    if (comp == 1) return s * da + (c - 1.0) * db;
    return prm[2] * sv;
}

// Prescribed boundary displacement used by the Dirichlet anchors.
__device__ inline double dirichlet_disp(int mode, const double *prm, int comp, double x, double y,
                                        double z, double time) {
    switch (mode) {
    case V_CONST: return time * prm[comp];
    case V_RAMP: return prm[comp] * sine2_integral(time, prm[3], prm[3]);
    case V_SINE2_HOLD: {
        const double active = prm[3];
        double denom = prm[4];
        if (denom <= 0.0) denom = 2.0 * active;
        return prm[comp] * sine2_integral(time, active, denom);
    }
    case V_AFFINE_VEL: {
        double at = time;
        if (prm[12] > 0.0 && at > prm[12]) at = prm[12];
        const int r = comp * 3;
        return at * (prm[r] * x + prm[r + 1] * y + prm[r + 2] * z + prm[9 + comp]);
    }
    case V_ZTWIST: return twist_disp(prm, comp, x, y, time);
    case V_XTWIST: return twist_disp(prm, tube_comp(comp), y, z, time);
    default: return 0.0;
    }
}

__device__ inline double sine2_scale(double t, double active, double denom) {
    if (denom <= 0.0) denom = 2.0 * active;
    if (active > 0.0 && t < active) {
        const double s = sin(PI_D * t / denom);
        return s * s;
    }
    return 1.0;
}

// Boundary value (velocity for Dirichlet rows, nominal traction for Neumann
// Tube end loads about the axis through (prm[0], prm[1]) of the transverse
// plane (a, b); comp 0, 1 transverse, 2 axial (cbw3d:1100-1132): the
// pull-torsion traction and the velocity of the exact finite twist.
__device__ inline double tube_pull(const double *prm, int comp, double a, double b, double t) {
    const double scale = sine2_scale(t, prm[4], prm[5]);
    if (comp == 0) return -scale * prm[3] * (b - prm[1]);
    if (comp == 1) return scale * prm[3] * (a - prm[0]);
    return scale * prm[2];
}
__device__ inline double twist_velocity(const double *prm, int comp, double a, double b, double t) {
    const double active = prm[4];
    if (active <= 0.0 || t >= active) return 0.0;
    const double phase = PI_D * t / active;
    const double sv = 0.5 * (1.0 - cos(phase)), sdot = 0.5 * PI_D / active * sin(phase);
    const double theta = prm[3] * sv, thetadot = prm[3] * sdot;
    const double da = a - prm[0], db = b - prm[1];
    if (comp == 0) return thetadot * (-sin(theta) * da - cos(theta) * db);
    if (comp == 1) return thetadot * (cos(theta) * da - sin(theta) * db);
    return prm[2] * sdot;
}

// rows) of component comp at a cut point.
__device__ inline double curved_value(int mode, const double *prm, int comp, double x, double y,
                                      double z, const double *nv, double t, double jf) {
    switch (mode) {
    case V_CONST: return prm[comp];
    case V_RAMP: {
        double scale = 1.0;
        if (prm[3] > 0.0 && t < prm[3]) {
            const double s = sin(PI_D * t / prm[3]);
            scale = s * s;
        }
        return prm[comp] * scale;
    }
    case V_AFFINE_VEL: {
        if (prm[12] > 0.0 && t > prm[12]) return 0.0;
        const int r = comp * 3;
        return prm[r] * x + prm[r + 1] * y + prm[r + 2] * z + prm[9 + comp];
    }
    case V_SINE2_HOLD: return prm[comp] * sine2_scale(t, prm[3], prm[4]);
    case V_PIOLA_N: return prm[3 * comp] * nv[0] + prm[3 * comp + 1] * nv[1] + prm[3 * comp + 2] * nv[2];
    case V_AFFINE_NH: {
        double tt = t;
        if (prm[11] > 0.0 && tt > prm[11]) tt = prm[11];
        double F[9], P[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = tt * prm[k];
        F[0] = 1.0 + tt * prm[0];
        F[4] = 1.0 + tt * prm[4];
        F[8] = 1.0 + tt * prm[8];
        Material nh{M_NH, prm[9], prm[10], 0.0, 0.0, 0.0, jf};
        piola(F, nh, P);
        return P[3 * comp] * nv[0] + P[3 * comp + 1] * nv[1] + P[3 * comp + 2] * nv[2];
    }
    case V_NORMAL_T: return prm[0] * nv[comp];
    case V_SINE2_NT: return prm[0] * sine2_scale(t, prm[1], prm[2]) * nv[comp];
    case V_LINEAR_NT: {
        double scale = 1.0;
        if (prm[1] > 0.0 && t < prm[1]) scale = t / prm[1];
        return prm[0] * scale * nv[comp];
    }
    case V_ZPULL: return tube_pull(prm, comp, x, y, t);
    case V_ZTWIST: return twist_velocity(prm, comp, x, y, t);
    case V_XPULL: return tube_pull(prm, tube_comp(comp), y, z, t);
    case V_XTWIST: return twist_velocity(prm, tube_comp(comp), y, z, t);
    default: return 0.0;
    }
}

// Face datum of the rectangular box solver (vnew3d:1393-1452); its value
// modes are numbered apart from the curved ones: 0 zero, 1 constant,
// 2 sin^2 ramp, 3 sin^2 hold, 4 affine velocity (params A row-major, b,
// active time at 12), 5 shear velocity x.
__device__ inline double box_value3(int comp, int mode, const double *p, double x, double y, double z, double t) {
    switch (mode) {
    case 1: return p[comp];
    case 2: {
        double scale = 1.0;
        if (p[3] > 0.0 && t < p[3]) {
            const double s = sin(PI_D * t / p[3]);
            scale = s * s;
        }
        return p[comp] * scale;
    }
    case 3: {
        double scale = 1.0;
        if (p[3] > 0.0 && t < p[3]) {
            double den = p[4];
            if (den <= 0.0) den = 2.0 * p[3];
            const double s = sin(PI_D * t / den);
            scale = s * s;
        }
        return p[comp] * scale;
    }
    case 4:
        if (p[12] > 0.0 && t > p[12]) return 0.0;
        return p[3 * comp] * x + p[3 * comp + 1] * y + p[3 * comp + 2] * z + p[9 + comp];
    case 5: return (comp == 0 && t < p[1]) ? p[0] * sin(PI_D * t / p[1]) : 0.0;
    default: return 0.0;
    }
}

// Traction solve of an axis-aligned box face (vnew3d:1455-1766): Newton with
// a forward-difference Jacobian (eps 1e-6) on the F column of the normal
// axis, residual sign * P column - T below 1e-10, at most 16 iterations,
// each step halved up to 12 times until det F > j_floor.  F (row-major)
// holds the node's F on entry and the face F on exit.
__device__ inline void box_residual3(const double *F, int axis, double sign, const double *T, const Material &m,
                                     double *r) {
    double P[9];
    piola(F, m, P);
#pragma unroll
    for (int row = 0; row < 3; ++row) r[row] = sign * P[3 * row + axis] - T[row];
}
__device__ inline void solve_normal_column3(double *F, int axis, double sign, const double *T, const Material &m) {
    double Fin[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Fin[k] = F[k];
    double x[3] = {Fin[axis], Fin[3 + axis], Fin[6 + axis]};
    const double eps = 1.0e-6;
    auto with = [&](const double *col, double *G) {
#pragma unroll
        for (int k = 0; k < 9; ++k) G[k] = Fin[k];
#pragma unroll
        for (int row = 0; row < 3; ++row) G[3 * row + axis] = col[row];
    };
    for (int it = 0; it < 16; ++it) {
        double G[9], r[3], rp[3][3];
        with(x, G);
        box_residual3(G, axis, sign, T, m, r);
        if (sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) < 1.0e-10) break;
        for (int c = 0; c < 3; ++c) {
            double xc[3] = {x[0], x[1], x[2]};
            xc[c] = x[c] + eps;
            with(xc, G);
            box_residual3(G, axis, sign, T, m, rp[c]);
        }
        // J[row][c] = d r[row] / d x[c]
        const double a = (rp[0][0] - r[0]) / eps, b = (rp[1][0] - r[0]) / eps, c = (rp[2][0] - r[0]) / eps;
        const double d = (rp[0][1] - r[1]) / eps, e = (rp[1][1] - r[1]) / eps, f = (rp[2][1] - r[1]) / eps;
        const double g = (rp[0][2] - r[2]) / eps, h = (rp[1][2] - r[2]) / eps, ii = (rp[2][2] - r[2]) / eps;
        const double det = a * (e * ii - f * h) - b * (d * ii - f * g) + c * (d * h - e * g);
        if (fabs(det) < 1.0e-14) break;
        const double i00 = (e * ii - f * h) / det, i01 = (c * h - b * ii) / det, i02 = (b * f - c * e) / det;
        const double i10 = (f * g - d * ii) / det, i11 = (a * ii - c * g) / det, i12 = (c * d - a * f) / det;
        const double i20 = (d * h - e * g) / det, i21 = (b * g - a * h) / det, i22 = (a * e - b * d) / det;
        const double d0 = -(i00 * r[0] + i01 * r[1] + i02 * r[2]);
        const double d1 = -(i10 * r[0] + i11 * r[1] + i12 * r[2]);
        const double d2 = -(i20 * r[0] + i21 * r[1] + i22 * r[2]);
        double step = 1.0;
        for (int ls = 0; ls < 12; ++ls) {
            const double cand[3] = {x[0] + step * d0, x[1] + step * d1, x[2] + step * d2};
            with(cand, G);
            if (det3(G) > m.jf) {
                x[0] = cand[0];
                x[1] = cand[1];
                x[2] = cand[2];
                break;
            }
            step = 0.5 * step;
        }
    }
    with(x, F);
}

__device__ __forceinline__ void normalize3(double *v) {
    double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (r < 1.0e-30) r = 1.0e-30;
    v[0] /= r;
    v[1] /= r;
    v[2] /= r;
}

__device__ __forceinline__ void make_F(const double *h, const double *n, const double *t1,
                                       const double *t2, const double *g1, const double *g2,
                                       double *F) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[3 * i + j] = h[i] * n[j] + g1[i] * t1[j] + g2[i] * t2[j];
}

__device__ inline void normal_residual(const double *h, const double *n, const double *t1,
                                       const double *t2, const double *g1, const double *g2,
                                       const double *T, const Material &m, double *r) {
    double F[9], P[9];
    make_F(h, n, t1, t2, g1, g2, F);
    piola(F, m, P);
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] = P[3 * i] * n[0] + P[3 * i + 1] * n[1] + P[3 * i + 2] * n[2] - T[i];
}

// Neo-Hooke acoustic tensor Q(n) = mu |n|^2 I + (lam J^2 - beta) (Hn)(Hn)^T.
__device__ inline void acoustic_normal(const double *F, const double *n, const Material &m, double *Qn) {
    double J = det3(F);
    if (J < m.jf) J = m.jf;
    const double F11 = F[0], F12 = F[1], F13 = F[2], F21 = F[3], F22 = F[4], F23 = F[5],
                 F31 = F[6], F32 = F[7], F33 = F[8];
    const double H11 = (F22 * F33 - F23 * F32) / J, H12 = (F23 * F31 - F21 * F33) / J,
                 H13 = (F21 * F32 - F22 * F31) / J, H21 = (F13 * F32 - F12 * F33) / J,
                 H22 = (F11 * F33 - F13 * F31) / J, H23 = (F12 * F31 - F11 * F32) / J,
                 H31 = (F12 * F23 - F13 * F22) / J, H32 = (F13 * F21 - F11 * F23) / J,
                 H33 = (F11 * F22 - F12 * F21) / J;
    const double beta = 0.5 * m.lam * (J * J - 1.0) - m.mu;
    const double coeff = m.lam * J * J - beta;
    const double n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    const double v[3] = {H11 * n[0] + H12 * n[1] + H13 * n[2], H21 * n[0] + H22 * n[1] + H23 * n[2],
                         H31 * n[0] + H32 * n[1] + H33 * n[2]};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Qn[3 * i + j] = (i == j ? m.mu * n2 : 0.0) + coeff * v[i] * v[j];
}

// Closed-form minimiser of W(G + h n^T) - T.h over the Neumann rows of the
// normal image (quadratic-volumetric neo-Hooke).  Returns false when the
// root is degenerate or below the Jacobian floor.
__device__ inline bool energy_Fb(const double *F, const double *n, const double *T, const int *kk,
                                 const Material &m, double *out) {
    const double cof[9] = {F[4] * F[8] - F[5] * F[7], F[5] * F[6] - F[3] * F[8], F[3] * F[7] - F[4] * F[6],
                           F[2] * F[7] - F[1] * F[8], F[0] * F[8] - F[2] * F[6], F[1] * F[6] - F[0] * F[7],
                           F[1] * F[5] - F[2] * F[4], F[2] * F[3] - F[0] * F[5], F[0] * F[4] - F[1] * F[3]};
    double a[3], hold[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        a[i] = cof[3 * i] * n[0] + cof[3 * i + 1] * n[1] + cof[3 * i + 2] * n[2];
        hold[i] = F[3 * i] * n[0] + F[3 * i + 1] * n[1] + F[3 * i + 2] * n[2];
    }
    double h[3];
    if (kk[0] == BC_NEU && kk[1] == BC_NEU && kk[2] == BC_NEU) {
        const double a2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        if (!(a2 > 1.0e-28)) return false;
        const double area = sqrt(a2);
        const double e[3] = {a[0] / area, a[1] / area, a[2] / area};
        const double tau = T[0] * e[0] + T[1] * e[1] + T[2] * e[2];
        const double aa = m.mu + 0.5 * m.lam * a2, bb = m.mu + 0.5 * m.lam;
        const double disc = sqrt(tau * tau + 4.0 * aa * bb);
        double sv = (tau + disc) / (2.0 * aa);
        if (tau < 0.0) sv = 2.0 * bb / (disc - tau);
        if (!(area * sv > m.jf)) return false;
#pragma unroll
        for (int i = 0; i < 3; ++i) h[i] = (T[i] - tau * e[i]) / m.mu + sv * e[i];
    } else {
        double an[3] = {0.0, 0.0, 0.0}, tn[3] = {0.0, 0.0, 0.0}, fixed[3] = {hold[0], hold[1], hold[2]};
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if (kk[i] == BC_NEU) {
                an[i] = a[i];
                tn[i] = T[i];
                fixed[i] = 0.0;
            }
        const double a2 = an[0] * an[0] + an[1] * an[1] + an[2] * an[2];
        const double d = a[0] * fixed[0] + a[1] * fixed[1] + a[2] * fixed[2];
        if (a2 > 1.0e-28) {
            const double aa = m.mu + 0.5 * m.lam * a2, bb = (m.mu + 0.5 * m.lam) * a2;
            const double cc = m.mu * d + (an[0] * tn[0] + an[1] * tn[1] + an[2] * tn[2]);
            const double disc = sqrt(cc * cc + 4.0 * aa * bb);
            double J = (cc + disc) / (2.0 * aa);
            if (cc < 0.0) J = 2.0 * bb / (disc - cc);
            if (!(J > m.jf)) return false;
            const double area = sqrt(a2);
            const double e[3] = {an[0] / area, an[1] / area, an[2] / area};
            const double te = tn[0] * e[0] + tn[1] * e[1] + tn[2] * e[2];
#pragma unroll
            for (int i = 0; i < 3; ++i) h[i] = fixed[i] + (tn[i] - te * e[i]) / m.mu + ((J - d) / area) * e[i];
        } else if (a2 == 0.0 && d > m.jf) {
#pragma unroll
            for (int i = 0; i < 3; ++i) h[i] = fixed[i] + tn[i] / m.mu;
        } else {
            return false;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) out[3 * i + j] = F[3 * i + j] + (h[i] - hold[i]) * n[j];
    return true;
}

// Solve P(F_b) n = T for the normal image h = F_b n (rows with Neumann data
// only when `componentwise`), keeping the tangential images F t1, F t2.
__device__ inline void solve_Fb(const double *Fin, const double *nin, const double *T, const int *kk,
                                bool componentwise, int solver, const Material &m, double *Fb) {
    double n[3] = {nin[0], nin[1], nin[2]};
    normalize3(n);
    if (solver == NS_ENERGY && m.model == M_NH) {
        double E[9];
        const int all_neu[3] = {BC_NEU, BC_NEU, BC_NEU};
        if (!energy_Fb(Fin, n, T, componentwise ? kk : all_neu, m, E)) {
#pragma unroll
            for (int k = 0; k < 9; ++k) E[k] = 0.0;
        }
        if (det3(E) > m.jf) {
#pragma unroll
            for (int k = 0; k < 9; ++k) Fb[k] = E[k];
            return;
        }
    }
    double ref[3] = {1.0, 0.0, 0.0};
    if (fabs(n[0]) >= 0.75) {
        ref[0] = 0.0;
        ref[1] = 1.0;
    }
    double t1[3] = {n[1] * ref[2] - n[2] * ref[1], n[2] * ref[0] - n[0] * ref[2], n[0] * ref[1] - n[1] * ref[0]};
    normalize3(t1);
    const double t2[3] = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
    double h[3], g1[3], g2[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        h[i] = Fin[3 * i] * n[0] + Fin[3 * i + 1] * n[1] + Fin[3 * i + 2] * n[2];
        g1[i] = Fin[3 * i] * t1[0] + Fin[3 * i + 1] * t1[1] + Fin[3 * i + 2] * t1[2];
        g2[i] = Fin[3 * i] * t2[0] + Fin[3 * i + 1] * t2[1] + Fin[3 * i + 2] * t2[2];
    }
    bool neu[3] = {true, true, true};
    if (componentwise)
        for (int i = 0; i < 3; ++i) neu[i] = kk[i] == BC_NEU;
    const double eps = 1.0e-6;
    for (int it = 0; it < 16; ++it) {
        double r[3];
        normal_residual(h, n, t1, t2, g1, g2, T, m, r);
        double rr[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) rr[i] = neu[i] ? r[i] : 0.0;
        if (sqrt(rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2]) < 1.0e-10) break;
        double M[9];
        if (solver != NJ_FD && m.model == M_NH) {
            double F[9];
            make_F(h, n, t1, t2, g1, g2, F);
            acoustic_normal(F, n, m, M);
        } else {
            for (int col = 0; col < 3; ++col) {
                double hp[3] = {h[0], h[1], h[2]}, rp[3];
                hp[col] = h[col] + eps;
                normal_residual(hp, n, t1, t2, g1, g2, T, m, rp);
#pragma unroll
                for (int i = 0; i < 3; ++i) M[3 * i + col] = (rp[i] - r[i]) / eps;
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if (!neu[i]) {
                M[3 * i] = i == 0 ? 1.0 : 0.0;
                M[3 * i + 1] = i == 1 ? 1.0 : 0.0;
                M[3 * i + 2] = i == 2 ? 1.0 : 0.0;
            }
        const double a = M[0], b = M[1], c = M[2], d = M[3], e = M[4], f = M[5], g = M[6], hh = M[7], ii = M[8];
        const double det = a * (e * ii - f * hh) - b * (d * ii - f * g) + c * (d * hh - e * g);
        if (fabs(det) < 1.0e-14) break;
        const double inv[9] = {(e * ii - f * hh) / det, (c * hh - b * ii) / det, (b * f - c * e) / det,
                               (f * g - d * ii) / det,  (a * ii - c * g) / det,  (c * d - a * f) / det,
                               (d * hh - e * g) / det,  (b * g - a * hh) / det,  (a * e - b * d) / det};
        double delta[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) delta[i] = -(inv[3 * i] * rr[0] + inv[3 * i + 1] * rr[1] + inv[3 * i + 2] * rr[2]);
        double step = 1.0;
        for (int ls = 0; ls < 12; ++ls) {
            double cnd[3], F[9];
#pragma unroll
            for (int i = 0; i < 3; ++i) cnd[i] = neu[i] ? h[i] + step * delta[i] : h[i];
            make_F(cnd, n, t1, t2, g1, g2, F);
            if (det3(F) > m.jf) {
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

__device__ __forceinline__ double det_state(const double *U) {
    return U[3] * (U[7] * U[11] - U[8] * U[10]) - U[4] * (U[6] * U[11] - U[8] * U[9]) +
           U[5] * (U[6] * U[10] - U[7] * U[9]);
}

// Centred finite-difference directional derivative of the axis flux,
// step 1e-6 / max(1, |W|_inf).
__device__ inline void flux_jvp3(int axis, const double *U, const double *W, const Material &m,
                                 double *out) {
    double scale = 1.0;
#pragma unroll
    for (int a = 0; a < NC; ++a) scale = fmax(scale, fabs(W[a]));
    const double eps = 1.0e-6 / scale;
    double Fp[9], Fm[9], Pp[9], Pm[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        Fp[k] = U[3 + k] + eps * W[3 + k];
        Fm[k] = U[3 + k] - eps * W[3 + k];
    }
    piola(Fp, m, Pp);
    piola(Fm, m, Pm);
#pragma unroll
    for (int a = 0; a < NC; ++a) out[a] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        out[i] = -(Pp[3 * i + axis] - Pm[3 * i + axis]) / (2.0 * eps);
        out[3 + 3 * i + axis] = -W[i];
    }
}

// Incoming population of a traction link from the characteristic closure:
// f = feq(Uc) - dt/omega feq'(Uc) w + dt (1/omega - 1/2) B(Uc) / 6.
__device__ inline void char_population(const double *Ub, const double *Gx, const double *Gy,
                                       const double *Gz, const double *d, double eta, double dt,
                                       double c, double omega, int src_mode, const double *src_prm,
                                       const Material &m, double *outp) {
    double shift[NC], Uc[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) shift[a] = eta * dt * c * (d[0] * Gx[a] + d[1] * Gy[a] + d[2] * Gz[a]);
    double chi = 1.0;
    for (int guard = 0; guard < 10; ++guard) {
#pragma unroll
        for (int a = 0; a < NC; ++a) Uc[a] = Ub[a] + chi * shift[a];
        if (det_state(Uc) > 10.0 * m.jf) break;
        chi = 0.5 * chi;
    }
#pragma unroll
    for (int a = 0; a < NC; ++a) Uc[a] = Ub[a] + chi * shift[a];
    double B[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) B[a] = source_comp(a, src_mode, src_prm, Uc[0], Uc[1], Uc[2]);
    double w[NC], j0[NC];
    flux_jvp3(0, Uc, Gx, m, j0);
#pragma unroll
    for (int a = 0; a < NC; ++a) w[a] = B[a] - j0[a];
    flux_jvp3(1, Uc, Gy, m, j0);
#pragma unroll
    for (int a = 0; a < NC; ++a) w[a] = w[a] - j0[a];
    flux_jvp3(2, Uc, Gz, m, j0);
#pragma unroll
    for (int a = 0; a < NC; ++a) w[a] = w[a] - j0[a] + c * (d[0] * Gx[a] + d[1] * Gy[a] + d[2] * Gz[a]);
    int axis = 0;
    if (fabs(d[1]) > 0.5) axis = 1;
    if (fabs(d[2]) > 0.5) axis = 2;
    double Aw[NC], P[9], flux[NC];
    flux_jvp3(axis, Uc, w, m, Aw);
    piola(Uc + 3, m, P);
#pragma unroll
    for (int a = 0; a < NC; ++a) flux[a] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        flux[i] = -P[3 * i + axis];
        flux[3 + 3 * i + axis] = -Uc[i];
    }
    const double da = d[axis];
#pragma unroll
    for (int a = 0; a < NC; ++a)
        outp[a] = (Uc[a] + 3.0 * da * flux[a] / c - dt / omega * (w[a] + 3.0 * da * Aw[a] / c) +
                   dt * (1.0 / omega - 0.5) * B[a]) / 6.0;
}

}  // namespace vlbm
