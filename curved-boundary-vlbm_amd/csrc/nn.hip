// Convolution kernels of the warm-start network on gfx950: BF16 operands on
// the MFMA units, FP32 accumulation.  The layout is described in
// include/vlbm_nn.h.
//
// Forward (and backward-data) convolution as an implicit GEMM: M = voxels,
// N = 32 output channels, K = 27 taps x 32 input channels.  One
// v_mfma_f32_16x16x32_bf16 consumes one tap (K = 32 input channels) for 16
// voxels and 16 output channels.  Lane l of a wave supplies the A row
// l % 16 (a voxel) with input channels 8 (l / 16) .. +7, which is one
// 16-byte load from the channels-last activations, and the B column l % 16
// (an output channel) with the same 8 input channels of that tap's weights,
// one 16-byte LDS read.  A wave computes 64 voxels x 32 channels (eight
// accumulators), a workgroup of four waves 256 voxels, with the layer's
// 54 KiB of weights staged once in LDS.  The accumulator holds, in lane l,
// output channel l % 16 of voxels 4 (l / 16) .. +3, whose addresses come
// from the lanes that loaded them (shuffle).
//
// Weight gradient: K = voxels.  Chunks of 32 voxels are staged in LDS
// channel-major (dz as [out][voxel] and, for all 27 taps, the shifted input
// as [tap][in][voxel]) so that each lane's 8 consecutive voxels are one
// 16-byte LDS read; each wave owns one 16 x 16 (out, in) block for all 27
// taps in registers over the workgroup's chunks, and the per-workgroup
// partial sums are reduced in a fixed order by a second kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/vlbm_nn.h"
#include "vlbm_common.h"

using vlbm_common::fail;
using vlbm_common::fail_msg;

namespace {

constexpr int CH = VLBM_NN_CH;
constexpr int TAPS = VLBM_NN_TAPS;
constexpr int NW = VLBM_NN_WEIGHTS;
constexpr int TCH = VLBM_NN_TARGET_CH;
constexpr int NB = 256;          // threads per workgroup
constexpr int VOX_PER_WAVE = 64;  // forward: 4 M-blocks of 16
constexpr int WG_CHUNK = 32;      // weight gradient: voxels per staged chunk
constexpr int MAX_WGRAD_PARTIALS = 1024;
constexpr int MAX_LOSS_PARTIALS = 4096;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

struct Geo {
    int nx, ny, nz;
    int Yp, Zp;
    long long S;      // voxels per padded sample
    long long V;      // interior voxels over the batch
    long long cells;  // interior voxels per sample
};

Geo make_geo(const vlbm_nn_grid &g) {
    Geo G;
    G.nx = g.nx;
    G.ny = g.ny;
    G.nz = g.nz;
    G.Yp = g.ny + 2;
    G.Zp = g.nz + 2;
    G.S = (long long)(g.nx + 2) * G.Yp * G.Zp;
    G.cells = (long long)g.nx * g.ny * g.nz;
    G.V = G.cells * g.batch;
    return G;
}

// Padded voxel of interior cell v (flattened over the batch).
__device__ __forceinline__ long long padded(const Geo &G, long long v) {
    const long long b = v / G.cells;
    long long r = v - b * G.cells;
    const int z = (int)(r % G.nz);
    r /= G.nz;
    const int y = (int)(r % G.ny);
    const int x = (int)(r / G.ny);
    return b * G.S + ((long long)(x + 1) * G.Yp + (y + 1)) * G.Zp + (z + 1);
}

__device__ __forceinline__ int tap_offset(const Geo &G, int t) {
    const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
    return (dx * G.Yp + dy) * G.Zp + dz;
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 u) { return __builtin_bit_cast(bf16x8, u); }

enum { EPI_HIDDEN = 0, EPI_OUTPUT = 1, EPI_BWD = 2 };

template <int EPI>
__global__ __launch_bounds__(NB) void k_conv(const uint16_t *__restrict__ in, const uint16_t *__restrict__ w,
                                             const float *__restrict__ bias, const uint16_t *__restrict__ zprev,
                                             uint16_t *__restrict__ out16, uint16_t *__restrict__ out16b,
                                             float *__restrict__ out32, Geo G) {
    __shared__ uint4 ws[NW / 8];
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(w);
        for (int i = threadIdx.x; i < NW / 8; i += NB) ws[i] = src[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const long long base = ((long long)blockIdx.x * (NB / 64) + wave) * VOX_PER_WAVE;
    long long addr[4];
    int valid[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        const long long v = base + mb * 16 + row;
        valid[mb] = v < G.V;
        addr[mb] = padded(G, valid[mb] ? v : 0);
    }
    floatx4 acc[4][2];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) acc[mb][nb] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint4 *inv = reinterpret_cast<const uint4 *>(in);
    for (int t = 0; t < TAPS; ++t) {
        const int off = tap_offset(G, t);
        uint4 a[4];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) a[mb] = inv[(addr[mb] + off) * (CH / 8) + kg];
        const bf16x8 b0 = as_bf16x8(ws[(t * CH + row) * (CH / 8) + kg]);
        const bf16x8 b1 = as_bf16x8(ws[(t * CH + 16 + row) * (CH / 8) + kg]);
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const bf16x8 av = as_bf16x8(a[mb]);
            acc[mb][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b0, acc[mb][0], 0, 0, 0);
            acc[mb][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b1, acc[mb][1], 0, 0, 0);
        }
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int src = 4 * kg + r;  // voxel row held by this lane's accumulator slot r
            const long long p = __shfl(addr[mb], src, 64);
            const int ok = __shfl(valid[mb], src, 64);
            if (!ok) continue;
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const int co = nb * 16 + row;
                const float s = acc[mb][nb][r];
                const long long idx = p * CH + co;
                if (EPI == EPI_HIDDEN) {
                    const float z = s + bias[co];
                    out16[idx] = f2bf(z);
                    out16b[idx] = f2bf(silu(z));
                } else if (EPI == EPI_OUTPUT) {
                    out32[idx] = s + bias[co];
                } else {
                    out16[idx] = f2bf(s * dsilu(bf2f(zprev[idx])));
                }
            }
        }
    }
}

// LDS row of a staged tile: 32 voxels of one channel plus 8 of padding.
constexpr int TROW = WG_CHUNK + 8;

__global__ __launch_bounds__(NB) void k_wgrad(const uint16_t *__restrict__ in, const uint16_t *__restrict__ dz,
                                              float *__restrict__ partials, Geo G, long long nchunks) {
    __shared__ __align__(16) uint16_t dzt[CH * TROW];
    __shared__ __align__(16) uint16_t int_[TAPS * CH * TROW];
    __shared__ long long vaddr[WG_CHUNK];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const int cb = wave >> 1, ib = wave & 1;
    floatx4 acc[TAPS];
#pragma unroll
    for (int t = 0; t < TAPS; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    float dbacc = 0.0f;
    const uint4 *dzv = reinterpret_cast<const uint4 *>(dz);
    const uint4 *inv = reinterpret_cast<const uint4 *>(in);
    for (long long c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const long long v0 = c * WG_CHUNK;
        if (threadIdx.x < WG_CHUNK) {
            const long long v = v0 + threadIdx.x;
            vaddr[threadIdx.x] = v < G.V ? padded(G, v) : -1;
        }
        __syncthreads();
        // dz tile, transposed: 16 voxel pairs x 4 channel groups
        if (threadIdx.x < 64) {
            const int pr = threadIdx.x >> 2, q = threadIdx.x & 3;
            const long long p0 = vaddr[2 * pr], p1 = vaddr[2 * pr + 1];
            const uint4 zero = {0u, 0u, 0u, 0u};
            const uint4 u0 = p0 >= 0 ? dzv[p0 * (CH / 8) + q] : zero;
            const uint4 u1 = p1 >= 0 ? dzv[p1 * (CH / 8) + q] : zero;
            const uint16_t *h0 = reinterpret_cast<const uint16_t *>(&u0);
            const uint16_t *h1 = reinterpret_cast<const uint16_t *>(&u1);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t word = (uint32_t)h0[e] | ((uint32_t)h1[e] << 16);
                *reinterpret_cast<uint32_t *>(&dzt[(8 * q + e) * TROW + 2 * pr]) = word;
            }
        }
        // shifted input tiles of the 27 taps, transposed
        for (int item = threadIdx.x; item < TAPS * 16 * 4; item += NB) {
            const int q = item & 3, pr = (item >> 2) & 15, t = item >> 6;
            const int off = tap_offset(G, t);
            const long long p0 = vaddr[2 * pr], p1 = vaddr[2 * pr + 1];
            const uint4 zero = {0u, 0u, 0u, 0u};
            const uint4 u0 = p0 >= 0 ? inv[(p0 + off) * (CH / 8) + q] : zero;
            const uint4 u1 = p1 >= 0 ? inv[(p1 + off) * (CH / 8) + q] : zero;
            const uint16_t *h0 = reinterpret_cast<const uint16_t *>(&u0);
            const uint16_t *h1 = reinterpret_cast<const uint16_t *>(&u1);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t word = (uint32_t)h0[e] | ((uint32_t)h1[e] << 16);
                *reinterpret_cast<uint32_t *>(&int_[(t * CH + 8 * q + e) * TROW + 2 * pr]) = word;
            }
        }
        __syncthreads();
        const bf16x8 av = as_bf16x8(*reinterpret_cast<const uint4 *>(&dzt[(cb * 16 + row) * TROW + 8 * kg]));
#pragma unroll
        for (int t = 0; t < TAPS; ++t) {
            const bf16x8 bv =
                as_bf16x8(*reinterpret_cast<const uint4 *>(&int_[(t * CH + ib * 16 + row) * TROW + 8 * kg]));
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
        }
        if (threadIdx.x < CH) {
            float s = 0.0f;
            for (int j = 0; j < WG_CHUNK; ++j) s += bf2f(dzt[threadIdx.x * TROW + j]);
            dbacc += s;
        }
        __syncthreads();
    }
    float *part = partials + (size_t)blockIdx.x * VLBM_NN_WGRAD_PARTIAL;
#pragma unroll
    for (int t = 0; t < TAPS; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = cb * 16 + 4 * kg + r, ci = ib * 16 + row;
            part[(t * CH + co) * CH + ci] = acc[t][r];
        }
    if (threadIdx.x < CH) part[NW + threadIdx.x] = dbacc;
}

__global__ __launch_bounds__(NB) void k_wgrad_reduce(const float *__restrict__ partials, int nparts,
                                                     float *__restrict__ dw, float *__restrict__ db) {
    const int i = blockIdx.x * NB + threadIdx.x;
    if (i >= VLBM_NN_WGRAD_PARTIAL) return;
    float s = 0.0f;
    for (int b = 0; b < nparts; ++b) s += partials[(size_t)b * VLBM_NN_WGRAD_PARTIAL + i];
    if (i < NW)
        dw[i] = s;
    else
        db[i - NW] = s;
}

__global__ __launch_bounds__(NB) void k_loss(const float *__restrict__ y, const float *__restrict__ target,
                                             const float *__restrict__ chw, int n_out, float inv_norm,
                                             uint16_t *__restrict__ dz, double *__restrict__ partials, Geo G) {
    __shared__ double red[NB];
    double l = 0.0;
    for (long long v = (long long)blockIdx.x * NB + threadIdx.x; v < G.V; v += (long long)gridDim.x * NB) {
        const long long p = padded(G, v);
        const float wv = target[p * TCH + (TCH - 1)];
        for (int c = 0; c < CH; ++c) {
            float g = 0.0f;
            if (c < n_out) {
                const float d = y[p * CH + c] - target[p * TCH + c];
                l += (double)wv * chw[c] * (double)d * d;
                g = 2.0f * wv * chw[c] * d * inv_norm;
            }
            dz[p * CH + c] = f2bf(g);
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = NB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(NB) void k_loss_final(const double *__restrict__ partials, int n, double scale,
                                                   double *__restrict__ out) {
    __shared__ double red[NB];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += NB) s += partials[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = NB / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}

__global__ __launch_bounds__(NB) void k_pack(const float *__restrict__ master, uint16_t *__restrict__ w,
                                             uint16_t *__restrict__ wt) {
    const int i = blockIdx.x * NB + threadIdx.x;
    if (i >= NW) return;
    const int t = i / (CH * CH), o = (i / CH) % CH, c = i % CH;
    const uint16_t h = f2bf(master[i]);
    w[i] = h;
    wt[((TAPS - 1 - t) * CH + c) * CH + o] = h;
}

__global__ __launch_bounds__(NB) void k_adam(int n, float *__restrict__ p, float *__restrict__ m,
                                             float *__restrict__ v, const float *__restrict__ g, float lr, float b1,
                                             float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * NB + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

int check_grid(const vlbm_nn_grid *g, const char *where) {
    if (!g) return fail_msg(std::string(where) + ": null grid");
    if (g->batch < 1 || g->nx < 1 || g->ny < 1 || g->nz < 1)
        return fail_msg(std::string(where) + ": grid sizes must be positive");
    if ((long long)g->batch * (g->nx + 2) * (g->ny + 2) * (g->nz + 2) * CH >= (1ll << 40))
        return fail_msg(std::string(where) + ": grid too large");
    return 0;
}

}  // namespace

#define CHECK_LAUNCH(where) VLBM_CHECK_LAUNCH(where)

extern "C" {

int vlbm_nn_conv_forward(const vlbm_nn_grid *g, const uint16_t *in, const uint16_t *w, const float *bias,
                         int32_t mode, uint16_t *z_out, uint16_t *a_out, float *y_out, void *stream) {
    if (int rc = check_grid(g, "vlbm_nn_conv_forward")) return rc;
    if (!in || !w || !bias) return fail_msg("vlbm_nn_conv_forward: null input");
    const Geo G = make_geo(*g);
    const long long nblk = (G.V + NB - 1) / NB;
    hipStream_t st = (hipStream_t)stream;
    if (mode == VLBM_NN_HIDDEN) {
        if (!z_out || !a_out) return fail_msg("vlbm_nn_conv_forward: hidden layer needs z_out and a_out");
        hipLaunchKernelGGL(k_conv<EPI_HIDDEN>, dim3((unsigned)nblk), dim3(NB), 0, st, in, w, bias, nullptr, z_out,
                           a_out, nullptr, G);
    } else if (mode == VLBM_NN_OUTPUT) {
        if (!y_out) return fail_msg("vlbm_nn_conv_forward: output layer needs y_out");
        hipLaunchKernelGGL(k_conv<EPI_OUTPUT>, dim3((unsigned)nblk), dim3(NB), 0, st, in, w, bias, nullptr, nullptr,
                           nullptr, y_out, G);
    } else {
        return fail_msg("vlbm_nn_conv_forward: unknown mode");
    }
    CHECK_LAUNCH("k_conv");
    return 0;
}

int vlbm_nn_conv_backward_data(const vlbm_nn_grid *g, const uint16_t *dz, const uint16_t *wt,
                               const uint16_t *z_prev, uint16_t *dz_prev, void *stream) {
    if (int rc = check_grid(g, "vlbm_nn_conv_backward_data")) return rc;
    if (!dz || !wt || !z_prev || !dz_prev) return fail_msg("vlbm_nn_conv_backward_data: null argument");
    const Geo G = make_geo(*g);
    const long long nblk = (G.V + NB - 1) / NB;
    hipLaunchKernelGGL(k_conv<EPI_BWD>, dim3((unsigned)nblk), dim3(NB), 0, (hipStream_t)stream, dz, wt, nullptr,
                       z_prev, dz_prev, nullptr, nullptr, G);
    CHECK_LAUNCH("k_conv<bwd>");
    return 0;
}

int vlbm_nn_conv_backward_weight(const vlbm_nn_grid *g, const uint16_t *in, const uint16_t *dz, float *partials,
                                 int32_t n_partials, float *dw, float *db, void *stream) {
    if (int rc = check_grid(g, "vlbm_nn_conv_backward_weight")) return rc;
    if (!in || !dz || !partials || !dw || !db) return fail_msg("vlbm_nn_conv_backward_weight: null argument");
    if (n_partials < 1 || n_partials > MAX_WGRAD_PARTIALS)
        return fail_msg("vlbm_nn_conv_backward_weight: n_partials must be in [1, 1024]");
    const Geo G = make_geo(*g);
    const long long nchunks = (G.V + WG_CHUNK - 1) / WG_CHUNK;
    const int nwg = (int)(nchunks < n_partials ? nchunks : n_partials);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_wgrad, dim3(nwg), dim3(NB), 0, st, in, dz, partials, G, nchunks);
    CHECK_LAUNCH("k_wgrad");
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((VLBM_NN_WGRAD_PARTIAL + NB - 1) / NB), dim3(NB), 0, st, partials, nwg,
                       dw, db);
    CHECK_LAUNCH("k_wgrad_reduce");
    return 0;
}

int vlbm_nn_loss(const vlbm_nn_grid *g, const float *y, const float *target, const float *chan_w, int32_t n_out,
                 float inv_norm, uint16_t *dz, double *partials, int32_t n_partials, double *loss_dev, void *stream) {
    if (int rc = check_grid(g, "vlbm_nn_loss")) return rc;
    if (!y || !target || !chan_w || !dz || !partials || !loss_dev) return fail_msg("vlbm_nn_loss: null argument");
    if (n_out < 1 || n_out > TCH - 1) return fail_msg("vlbm_nn_loss: n_out must be in [1, 15]");
    if (n_partials < 1 || n_partials > MAX_LOSS_PARTIALS)
        return fail_msg("vlbm_nn_loss: n_partials must be in [1, 4096]");
    const Geo G = make_geo(*g);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_loss, dim3(n_partials), dim3(NB), 0, st, y, target, chan_w, n_out, inv_norm, dz, partials,
                       G);
    CHECK_LAUNCH("k_loss");
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(NB), 0, st, partials, n_partials, (double)inv_norm, loss_dev);
    CHECK_LAUNCH("k_loss_final");
    return 0;
}

int vlbm_nn_pack_weights(const float *master, uint16_t *w, uint16_t *wt, void *stream) {
    if (!master || !w || !wt) return fail_msg("vlbm_nn_pack_weights: null argument");
    hipLaunchKernelGGL(k_pack, dim3((NW + NB - 1) / NB), dim3(NB), 0, (hipStream_t)stream, master, w, wt);
    CHECK_LAUNCH("k_pack");
    return 0;
}

int vlbm_nn_adam(int32_t n, float *param, float *m, float *v, const float *grad, float lr, float b1, float b2,
                 float eps, float c1, float c2, void *stream) {
    if (n < 0) return fail_msg("vlbm_nn_adam: negative size");
    if (n == 0) return 0;
    if (!param || !m || !v || !grad) return fail_msg("vlbm_nn_adam: null argument");
    hipLaunchKernelGGL(k_adam, dim3((n + NB - 1) / NB), dim3(NB), 0, (hipStream_t)stream, n, param, m, v, grad, lr,
                       b1, b2, eps, c1, c2);
    CHECK_LAUNCH("k_adam");
    return 0;
}

}  // extern "C"
