// Deterministic (sorted-segment) gradient scatter: sort the adjoint's
// gradient records by cell with rocPRIM's radix sort and sum every run of
// equal keys in a fixed order (one wavefront per run).  See det_scatter.h
// for the record formats and why the appearance runs go in 8 parity passes.
#include "det_scatter.h"
#include "shade.h"

#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace qsdvr {

namespace {

// smallest b with 2^b > n
int bits_above(uint64_t n) {
    int b = 1;
    while ((1ull << b) <= n) ++b;
    return b;
}

size_t scan_temp_bytes(int64_t n) {
    size_t t = 0;
    rocprim::exclusive_scan(nullptr, t, (const uint32_t *)nullptr, (uint32_t *)nullptr, 0u,
                            (size_t)n, rocprim::plus<uint32_t>(), (hipStream_t)0);
    return t;
}

size_t sort_temp_bytes(int64_t n, int end_bit) {
    size_t t = 0;
    rocprim::radix_sort_pairs(nullptr, t, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                              (const uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0u,
                              (unsigned)end_bit, (hipStream_t)0);
    return t;
}

size_t temp_bytes_for(int64_t n_rays, int64_t cap, int R) {
    const int cbits = det_cell_bits(R);
    const int qbits = bits_above((uint64_t)R * R * R);
    size_t t = scan_temp_bytes(n_rays + 1);
    t = std::max(t, sort_temp_bytes(2 * cap, qbits));
    t = std::max(t, sort_temp_bytes(cap, cbits + 4));
    return t;
}

template <typename T>
T *carve(Carver &c, size_t n) {
    return c.take<T>(n);
}

}  // namespace

size_t det_workspace_bytes(int64_t n_rays, int64_t cap, int C, int R) {
    Carver c{nullptr, 0};
    carve<uint32_t>(c, n_rays + 1);
    carve<uint32_t>(c, n_rays + 1);
    for (int i = 0; i < 4; ++i) carve<uint32_t>(c, 2 * cap);
    carve<float4>(c, 4 * cap);
    for (int i = 0; i < 4; ++i) carve<uint32_t>(c, cap);
    carve<float>(c, cap * det_app_stride(C));
    carve<uint32_t>(c, 16);
    carve<uint32_t>(c, 1);
    carve<char>(c, temp_bytes_for(n_rays, cap, R));
    return c.used + 256;
}

int det_carve(void *ws, size_t bytes, int64_t n_rays, int64_t cap, int C, int R, DetBuffers &b) {
    QSDVR_REQUIRE(R >= 1 && R <= 512, "deterministic scatter: reso %d out of range (1..512)", R);
    if (2 * cap >= 0xFFFFFFFFll) {
        set_error("deterministic scatter: %lld sample slots exceed the 32-bit record index",
                  (long long)cap);
        return QSDVR_ERANGE;
    }
    Carver c{static_cast<char *>(ws), bytes};
    b.n_rays = n_rays;
    b.cap = cap;
    b.C = C;
    b.R = R;
    b.cbits = det_cell_bits(R);
    b.counts = carve<uint32_t>(c, n_rays + 1);
    b.offsets = carve<uint32_t>(c, n_rays + 1);
    b.qkey = carve<uint32_t>(c, 2 * cap);
    b.qkey_s = carve<uint32_t>(c, 2 * cap);
    b.qidx = carve<uint32_t>(c, 2 * cap);
    b.qidx_s = carve<uint32_t>(c, 2 * cap);
    b.qval = carve<float4>(c, 4 * cap);
    b.akey = carve<uint32_t>(c, cap);
    b.akey_s = carve<uint32_t>(c, cap);
    b.aidx = carve<uint32_t>(c, cap);
    b.aidx_s = carve<uint32_t>(c, cap);
    b.aval = carve<float>(c, cap * det_app_stride(C));
    b.bounds = carve<uint32_t>(c, 16);
    b.total = carve<uint32_t>(c, 1);
    b.temp_bytes = temp_bytes_for(n_rays, cap, R);
    b.temp = carve<char>(c, b.temp_bytes);
    if (ws == nullptr || b.temp == nullptr) {
        set_error("deterministic scatter: workspace of %zu bytes is too small (%zu needed)", bytes,
                  det_workspace_bytes(n_rays, cap, C, R));
        return QSDVR_EWORKSPACE;
    }
    return QSDVR_OK;
}

__global__ void __launch_bounds__(256) det_count_kernel(const float4 *__restrict__ ray_state,
                                                        int64_t n_rays, int K,
                                                        uint32_t *__restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n_rays) return;
    uint32_t n = 0;
    if (i < n_rays) {
        const uint32_t cnt = __float_as_uint(ray_state[2 * i].w);
        n = cnt >= 2 ? (K > 0 ? (uint32_t)K : cnt) : 0u;
    }
    counts[i] = n;
}

int det_offsets(DetBuffers &b, const float4 *ray_state, int K, hipStream_t s,
                int64_t *total_host) {
    const int64_t n = b.n_rays + 1;
    det_count_kernel<<<grid_1d(n, 256), 256, 0, s>>>(ray_state, b.n_rays, K, b.counts);
    QSDVR_LAUNCHED();
    size_t tb = b.temp_bytes;
    QSDVR_HIP(rocprim::exclusive_scan(b.temp, tb, b.counts, b.offsets, 0u, (size_t)n,
                                      rocprim::plus<uint32_t>(), s));
    uint32_t total = 0;
    QSDVR_HIP(hipMemcpyAsync(&total, b.offsets + b.n_rays, sizeof(uint32_t),
                             hipMemcpyDeviceToHost, s));
    QSDVR_HIP(hipStreamSynchronize(s));
    *total_host = total;
    if ((int64_t)total > b.cap) {
        set_error("deterministic scatter: %u sdf samples exceed the workspace capacity %lld",
                  total, (long long)b.cap);
        return QSDVR_EWORKSPACE;
    }
    return QSDVR_OK;
}

// First position >= from in the sorted keys [.., n) whose key differs from k
// (n if none).  Wave-uniform arguments; all 64 lanes call it.
__device__ __forceinline__ int64_t run_end(const uint32_t *__restrict__ key, int64_t from,
                                           int64_t n, uint32_t k) {
    const int lane = threadIdx.x % kWave;
    for (;;) {
        if (from >= n) return n;
        const int64_t p = from + lane;
        const bool stop = p >= n || key[p] != k;
        const unsigned long long m = __ballot(stop);
        if (m != 0) return from + (__ffsll((long long)m) - 1);
        from += kWave;
    }
}

// Quadric runs.  One thread per sorted position; the thread at the head of
// a run of at most 64 records sums it alone, in order, and adds the sum to
// its cell.  Longer runs (a cell crossed by many rays: full-frame batches)
// are summed by the whole wave, lanes striding over the records and a
// fixed butterfly combining them.
__global__ void __launch_bounds__(256) det_quadric_reduce_kernel(
    const uint32_t *__restrict__ key, const uint32_t *__restrict__ idx,
    const float4 *__restrict__ qval, int64_t n, uint32_t invalid, float *__restrict__ gq) {
    const int lane = threadIdx.x % kWave;
    const int64_t p0 = ((int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * kWave;
    if (p0 >= n) return;
    const int64_t p = p0 + lane;
    const uint32_t k = p < n ? key[p] : invalid;
    const bool head = p < n && k != invalid && (p == 0 || key[p - 1] != k);
    const bool long_run = head && p + kWave < n && key[p + kWave] == k;
    if (head && !long_run) {
        float acc[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        for (int64_t i = p; i < n && key[i] == k; ++i) {
            const uint32_t r = idx[i];
            const float4 a = qval[2 * (int64_t)r], c = qval[2 * (int64_t)r + 1];
            acc[0] += a.x;
            acc[1] += a.y;
            acc[2] += a.z;
            acc[3] += a.w;
            acc[4] += c.x;
            acc[5] += c.y;
            acc[6] += c.z;
        }
        float *dst = gq + 7 * (int64_t)k;
#pragma unroll
        for (int j = 0; j < 7; ++j) dst[j] += acc[j];
    }
    unsigned long long longs = __ballot(long_run);
    while (longs != 0) {
        const int h = __ffsll((long long)longs) - 1;
        longs &= longs - 1;
        const uint32_t hk = __shfl(k, h, kWave);
        const int64_t start = p0 + h;
        const int64_t end = run_end(key, start + kWave, n, hk);
        float acc[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        for (int64_t i = start + lane; i < end; i += kWave) {
            const uint32_t r = idx[i];
            const float4 a = qval[2 * (int64_t)r], c = qval[2 * (int64_t)r + 1];
            acc[0] += a.x;
            acc[1] += a.y;
            acc[2] += a.z;
            acc[3] += a.w;
            acc[4] += c.x;
            acc[5] += c.y;
            acc[6] += c.z;
        }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
#pragma unroll
            for (int j = 0; j < 7; ++j) acc[j] += __shfl_xor(acc[j], o, kWave);
        }
        if (lane < 7) {
            float v = acc[0];
#pragma unroll
            for (int j = 1; j < 7; ++j) v = lane == j ? acc[j] : v;
            gq[7 * (int64_t)hk + lane] += v;
        }
    }
}

// Class boundaries in the sorted appearance keys: bounds[c] = first
// position whose key is >= c << cbits (c = 0..8; bounds[8] starts the
// invalid keys).
__global__ void det_class_bounds_kernel(const uint32_t *__restrict__ key, int64_t n, int cbits,
                                        uint32_t *__restrict__ bounds) {
    const int c = threadIdx.x;
    if (c > 8) return;
    const uint32_t target = (uint32_t)c << cbits;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (key[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    bounds[c] = (uint32_t)lo;
}

// Appearance runs of parity class `cls`.  Waves stride over the class's
// sorted positions 64 at a time and sum every run that starts in their
// chunk: a group of lanes (one lane per channel; two groups per wave when
// C <= 32) takes one run, walks its records in order -- each lane derives
// the record's 8 trilinear weights from its local point -- and adds the
// 8 vertex rows to the gradient grid.
template <int C>
__global__ void __launch_bounds__(256) det_app_reduce_kernel(
    const uint32_t *__restrict__ key, const uint32_t *__restrict__ idx,
    const float *__restrict__ aval, const uint32_t *__restrict__ bounds, int cls, int cbits,
    int R, float *__restrict__ gdata) {
    static_assert(C <= kWave, "one lane per channel");
    constexpr int AP = det_app_stride(C);
    constexpr int NG = C <= kWave / 2 ? 2 : 1;  // runs in flight per wave
    constexpr int GW = kWave / NG;
    const int64_t lo = bounds[cls], hi = bounds[cls + 1];
    const int lane = threadIdx.x % kWave;
    const int grp = lane / GW, ch = lane % GW;
    const uint32_t cmask = (1u << cbits) - 1u;
    const int64_t RR = (int64_t)R * R;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    for (int64_t p0 = lo + wave * kWave; p0 < hi; p0 += nwaves * kWave) {
        const int64_t p = p0 + lane;
        const uint32_t k = p < hi ? key[p] : 0xFFFFFFFFu;
        const bool boundary = p < hi && (p == lo || key[p - 1] != k);
        const unsigned long long bmask = __ballot(boundary);
        if (bmask == 0) continue;  // inside a run that started in an earlier chunk
        const int last = 63 - __clzll((long long)bmask);
        const int64_t tail_end = run_end(key, p0 + kWave, hi, __shfl(k, last, kWave));
        unsigned long long heads = bmask;
        while (heads != 0) {
            int hs[2] = {-1, -1};
            hs[0] = __ffsll((long long)heads) - 1;
            heads &= heads - 1;
            if (NG == 2 && heads != 0) {
                hs[1] = __ffsll((long long)heads) - 1;
                heads &= heads - 1;
            }
            const int my = hs[grp];
            const uint32_t hk = __shfl(k, my < 0 ? 0 : my, kWave);
            if (my >= 0) {
                const unsigned long long later = bmask & ~((2ull << my) - 1);
                const int64_t start = p0 + my;
                const int64_t end = later != 0 ? p0 + (__ffsll((long long)later) - 1) : tail_end;
                float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
                for (int64_t i = start; i < end; ++i) {
                    const float *rec = aval + (int64_t)idx[i] * AP;
                    const float4 lp = *reinterpret_cast<const float4 *>(rec);
                    const float g = ch < C ? rec[4 + ch] : 0.0f;
                    float w[8];
                    trilinear_weights(f3{lp.x, lp.y, lp.z}, w);
#pragma unroll
                    for (int v = 0; v < 8; ++v) acc[v] += w[v] * g;
                }
                if (ch < C) {
                    const int64_t cell = hk & cmask;
#pragma unroll
                    for (int v = 0; v < 8; ++v) {
                        const int64_t row =
                            cell + (v & 1) + ((v >> 1) & 1) * (int64_t)R + (v >> 2) * RR;
                        gdata[row * C + ch] += acc[v];
                    }
                }
            }
        }
    }
}

int det_reduce(DetBuffers &b, int64_t T, float *gq, float *gdata, hipStream_t s) {
    if (T == 0) return QSDVR_OK;
    const uint64_t R3 = (uint64_t)b.R * b.R * b.R;
    const int qbits = bits_above(R3);
    size_t tb = b.temp_bytes;
    QSDVR_HIP(rocprim::radix_sort_pairs(b.temp, tb, b.qkey, b.qkey_s, b.qidx, b.qidx_s,
                                        (size_t)(2 * T), 0u, (unsigned)qbits, s));
    tb = b.temp_bytes;
    QSDVR_HIP(rocprim::radix_sort_pairs(b.temp, tb, b.akey, b.akey_s, b.aidx, b.aidx_s, (size_t)T,
                                        0u, (unsigned)(b.cbits + 4), s));
    det_quadric_reduce_kernel<<<grid_1d(2 * T, 256), 256, 0, s>>>(b.qkey_s, b.qidx_s, b.qval,
                                                                 2 * T, (uint32_t)R3, gq);
    QSDVR_LAUNCHED();
    det_class_bounds_kernel<<<1, 64, 0, s>>>(b.akey_s, T, b.cbits, b.bounds);
    QSDVR_LAUNCHED();
    // each class holds about T/8 records; the waves stride over its range
    const int64_t chunks = (T / 8 + kWave - 1) / kWave;
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((chunks + 3) / 4, 4096));
    for (int cls = 0; cls < 8; ++cls) {
        if (b.C == 32)
            det_app_reduce_kernel<32><<<blocks, 256, 0, s>>>(b.akey_s, b.aidx_s, b.aval, b.bounds,
                                                             cls, b.cbits, b.R, gdata);
        else
            det_app_reduce_kernel<53><<<blocks, 256, 0, s>>>(b.akey_s, b.aidx_s, b.aval, b.bounds,
                                                             cls, b.cbits, b.R, gdata);
        QSDVR_LAUNCHED();
    }
    return QSDVR_OK;
}

}  // namespace qsdvr
