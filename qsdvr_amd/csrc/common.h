// Internal helpers shared by the gfx950 kernels of the QSDVR render path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

#include "qsdvr_hip.h"

namespace qsdvr {

// ------------------------------------------------------------ status ----
void set_error(const char *fmt, ...);

#define QSDVR_REQUIRE(cond, ...)                  \
    do {                                          \
        if (!(cond)) {                            \
            ::qsdvr::set_error(__VA_ARGS__);      \
            return QSDVR_EINVAL;                  \
        }                                         \
    } while (0)

#define QSDVR_HIP(call)                                                     \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) {                                             \
            ::qsdvr::set_error("%s failed: %s", #call, hipGetErrorString(e_)); \
            return QSDVR_EHIP;                                              \
        }                                                                   \
    } while (0)

#define QSDVR_LAUNCHED()                                                    \
    do {                                                                    \
        hipError_t e_ = hipGetLastError();                                  \
        if (e_ != hipSuccess) {                                             \
            ::qsdvr::set_error("kernel launch failed: %s", hipGetErrorString(e_)); \
            return QSDVR_EHIP;                                              \
        }                                                                   \
    } while (0)

inline hipStream_t as_stream(qsdvr_stream s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_1d(int64_t n, int block) {
    return static_cast<unsigned>((n + block - 1) / block);
}

constexpr int kWave = 64;

// Workspace carving: slices at 256-byte aligned addresses out of one caller
// buffer, whatever the buffer's own alignment (the size queries add 256
// bytes for the first slice's padding; with base == nullptr `used` is the
// size from an aligned base).
struct Carver {
    char *base;
    size_t cap;
    size_t used = 0;
    template <typename T>
    T *take(size_t count) {
        const uintptr_t b0 = reinterpret_cast<uintptr_t>(base);
        size_t off = (size_t)(((b0 + used + 255) & ~uintptr_t(255)) - b0);
        size_t bytes = count * sizeof(T);
        used = off + bytes;
        if (base == nullptr || used > cap) return nullptr;
        return reinterpret_cast<T *>(base + off);
    }
};

// ---------------------------------------------------- device helpers ----
struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 load3(const float *p, int64_t i) {
    const float *q = p + 3 * i;
    return f3{q[0], q[1], q[2]};
}

__device__ __forceinline__ void store3(float *p, int64_t i, f3 v) {
    float *q = p + 3 * i;
    q[0] = v.x;
    q[1] = v.y;
    q[2] = v.z;
}

// Workgroups are dealt round-robin over the 8 XCDs (block b runs on XCD
// b % 8, each XCD with its own L2).  xcd_block() renumbers the launch so
// that XCD x takes one contiguous range of logical blocks (the first
// nb % 8 XCDs one block more than the others): neighbouring work items --
// e.g. the samples of neighbouring rays, which gather the same grid rows --
// then share an L2.  A bijection on [0, gridDim.x).
__device__ __forceinline__ int64_t xcd_block() {
    const int64_t nb = gridDim.x, b = blockIdx.x;
    const int64_t q = nb / 8, r = nb % 8, x = b % 8, k = b / 8;
    return x * q + (x < r ? x : r) + k;
}

// Hardware reciprocal, reciprocal square root and square root (v_rcp_f32,
// v_rsq_f32, v_sqrt_f32: about 1 ulp).  The library is built with IEEE
// division and square root for the bit-exact ray generation; the per-sample
// shading and compositing math, held to relative tolerances, uses these.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Hardware float atomic add (global_atomic_add_f32, no return value).
__device__ __forceinline__ void atomic_addf(float *p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// SH constants of models/render_torch.py:181-185 (the float32 roundings of
// the closed-form real-SH normalisations 1/(2 sqrt(pi)), sqrt(3/(4 pi)),
// sqrt(15/(4 pi)), sqrt(5/(16 pi)), sqrt(15/(16 pi))).
#define QSDVR_SH_K0 0.2820947917f
#define QSDVR_SH_K1 0.4886025119f
#define QSDVR_SH_K4 1.0925484305f
#define QSDVR_SH_K6 0.3153915652f
#define QSDVR_SH_K8 0.5462742152f

// Degree-3 constants (closed forms, same y-pole real-SH family, unsigned):
// k9 = sqrt(35/(32 pi)), k10 = sqrt(105/(4 pi)), k11 = sqrt(21/(32 pi)),
// k12 = sqrt(7/(16 pi)), k14 = sqrt(105/(16 pi)).
#define QSDVR_SH_K9 0.5900435899f
#define QSDVR_SH_K10 2.8906114426f
#define QSDVR_SH_K11 0.4570457995f
#define QSDVR_SH_K12 0.3731763326f
#define QSDVR_SH_K14 1.4453057213f

}  // namespace qsdvr
