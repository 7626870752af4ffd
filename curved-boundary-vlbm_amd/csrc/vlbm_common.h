// Host-side helpers shared by the 2D and 3D halves of libvlbm: the error
// string behind vlbm_last_error(), the per-kernel HIP-event timing, and the
// XCD-aware launch geometry.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

namespace vlbm_common {

// Thread-local text of the last failure (defined once in vlbm3d.hip).
std::string &last_error();

inline int fail(const char *where, hipError_t e) {
    last_error() = std::string(where) + ": " + hipGetErrorString(e);
    return -2;
}
inline int fail_msg(const std::string &m) {
    last_error() = m;
    return -1;
}

// Node-kernel block: one wave.  On the 128^3 shell the whole step is 1.4%
// faster than with 256-thread blocks (512: 3% slower), bitwise the same.
constexpr int BLOCK = 64;
constexpr int NXCD = 8;

// Grid sizes rounded up to a multiple of the XCD count; block b of the launch
// is mapped to logical block (b % 8) * (G / 8) + b / 8 so that each XCD walks
// one contiguous slab of nodes and neighbour rows stay in its own L2.
__device__ __forceinline__ int xcd_block() {
    const int G = gridDim.x;
    const int per = G / NXCD;
    const int b = blockIdx.x;
    return (b % NXCD) * per + b / NXCD;
}
inline int xcd_grid(int n, int block = BLOCK) {
    int g = (n + block - 1) / block;
    return ((g + NXCD - 1) / NXCD) * NXCD;
}

// Optional per-kernel timing with HIP events recorded on the launch stream.
struct Timing {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    hipEvent_t get() {
        if (used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
    void destroy() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
        pool.clear();
        marks.clear();
        used = 0;
    }
    int enable(hipStream_t s, int en) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) return fail("timing_enable: sync", e);
        marks.clear();
        used = 0;
        on = en != 0;
        return 0;
    }
    int read(hipStream_t s, double *ms, int32_t *counts, int nslots) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) return fail("timing_read: sync", e);
        for (int k = 0; k < nslots; ++k) {
            ms[k] = 0.0;
            counts[k] = 0;
        }
        for (auto &m : marks) {
            float t = 0.0f;
            e = hipEventElapsedTime(&t, m.second.first, m.second.second);
            if (e != hipSuccess) return fail("timing_read: elapsed", e);
            if (m.first >= 0 && m.first < nslots) {
                ms[m.first] += t;
                counts[m.first] += 1;
            }
        }
        marks.clear();
        used = 0;
        return 0;
    }
};

// RAII bracket around one launch: records a begin/end event pair on the
// stream when timing is enabled.
struct TimedSlot {
    Timing &tm;
    hipStream_t stream;
    int slot;
    hipEvent_t b = nullptr;
    TimedSlot(Timing &t, hipStream_t s, int sl) : tm(t), stream(s), slot(sl) {
        if (tm.on) {
            b = tm.get();
            if (b && hipEventRecord(b, stream) != hipSuccess) b = nullptr;
        }
    }
    ~TimedSlot() {
        if (b) {
            hipEvent_t e = tm.get();
            if (e && hipEventRecord(e, stream) == hipSuccess) tm.marks.push_back({slot, {b, e}});
        }
    }
};

}  // namespace vlbm_common

#define VLBM_CHECK_LAUNCH(where)                                      \
    do {                                                              \
        hipError_t e_ = hipGetLastError();                            \
        if (e_ != hipSuccess) return vlbm_common::fail((where), e_);  \
    } while (0)
