// Deterministic (sorted-segment) gradient scatter: sort the adjoint's
// gradient records by cell with rocPRIM's radix sort and sum every run of
// equal keys in a fixed order (one wavefront per run).  See det_scatter.h
// for the record formats and why the appearance runs go in 8 parity passes.
#include "det_scatter.h"
#include "quadric.h"
#include "shade.h"

#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace qsdvr {

namespace {

// waves of 64 records in the quadric segmented sum (head / tail sums)
int64_t seg_waves(int64_t n) { return (n + kWave - 1) / kWave; }

// smallest b with 2^b > n
int bits_above(uint64_t n) {
    int b = 1;
    while ((1ull << b) <= n) ++b;
    return b;
}

// Temporary-storage size queries.  A failed query leaves 0; the checked
// call that later uses the storage then reports the error.
size_t scan_temp_bytes(int64_t n) {
    size_t t = 0;
    (void)rocprim::exclusive_scan(nullptr, t, (const uint32_t *)nullptr, (uint32_t *)nullptr, 0u,
                            (size_t)n, rocprim::plus<uint32_t>(), (hipStream_t)0);
    return t;
}

size_t sort_temp_bytes(int64_t n, int end_bit) {
    size_t t = 0;
    (void)rocprim::radix_sort_pairs(nullptr, t, (const uint32_t *)nullptr, (uint32_t *)nullptr,
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
    carve<float>(c, 7 * seg_waves(2 * cap));
    carve<float>(c, 7 * seg_waves(2 * cap));
    carve<char>(c, temp_bytes_for(n_rays, cap, R));
    return c.used + 256;
}

int det_carve(void *ws, size_t bytes, int64_t n_rays, int64_t cap, int C, int R, DetBuffers &b) {
    QSDVR_REQUIRE(R >= 2 && R <= 512 && R % 2 == 0,
                  "deterministic scatter: reso %d must be even and in 2..512", R);
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
    b.qhead = carve<float>(c, 7 * seg_waves(2 * cap));
    b.qtail = carve<float>(c, 7 * seg_waves(2 * cap));
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

// Sort keys of the quadric runs: the cell itself.
struct CellKeys {
    __device__ __forceinline__ int64_t cell(uint32_t k) const { return k; }
};

// What GridInterpolationBackward sorts with its keys: the sample's point and
// its index (so that the run sums read the points in sorted order).
struct alignas(16) PointRec {
    float x, y, z;
    uint32_t i;
};

// Where the record at sorted position `pos` of a quadric run gets its 7
// coefficient gradients from: the adjoint's precomputed records (through
// the sorted record indices), or (op surface) the closed forms of the
// SampleSDF / SampleNormal adjoints at the sample's point and upstream
// gradient, sorted along with the keys, with the cell's quadric.
struct QuadricRecords {
    const float4 *qval;
    const uint32_t *idx;
    __device__ __forceinline__ void add(int64_t pos, int64_t, float acc[7]) const {
        const int64_t r = idx[pos];
        const float4 a = qval[2 * r], c = qval[2 * r + 1];
        acc[0] += a.x;
        acc[1] += a.y;
        acc[2] += a.z;
        acc[3] += a.w;
        acc[4] += c.x;
        acc[5] += c.y;
        acc[6] += c.z;
    }
};

// SampleSDF record: local point and upstream gradient
struct SdfSamples {
    const float4 *pay;
    const float *grid;
    float inv_reso;
    __device__ __forceinline__ void add(int64_t pos, int64_t cell, float acc[7]) const {
        const float4 p = pay[pos];
        quadric_sdf_grad(load_quadric(grid, cell), f3{p.x, p.y, p.z}, p.w, inv_reso, acc);
    }
};

// SampleNormal record: gathered through the sorted sample index (sorting
// the 24-B point and upstream gradient with the keys measured slower than
// gathering them)
struct NormalSamples {
    const uint32_t *idx;
    const float *grid, *pts, *up;
    __device__ __forceinline__ void add(int64_t pos, int64_t cell, float acc[7]) const {
        const int64_t r = idx[pos];
        quadric_normal_grad(load_quadric(grid, cell), load3(pts, r), load3(up, r), acc);
    }
};

// SampleNormal record on GridInterpolationBackward's sort of the same
// samples: the point in sorted order, the upstream gradient gathered
// through the sample index sorted with it
struct NormalPresorted {
    const PointRec *pay;
    const float *grid, *up;
    __device__ __forceinline__ void add(int64_t pos, int64_t cell, float acc[7]) const {
        const float4 p = *reinterpret_cast<const float4 *>(pay + pos);
        quadric_normal_grad(load_quadric(grid, cell), f3{p.x, p.y, p.z},
                            load3(up, __float_as_uint(p.w)), acc);
    }
};

// Quadric runs: a deterministic segmented sum in two passes.  Pass 1: wave
// w takes sorted positions [64 w, 64 w + 64), one record per lane, and a
// segmented inclusive scan over equal keys (Hillis-Steele, a fixed tree
// order) leaves each segment's sum in its last lane.  A run that lies
// inside the wave is added to its cell there; the part of a run that enters
// from the previous wave is stored as the wave's head sum, the part of a
// run that starts in the wave and leaves it as its tail sum.  Pass 2: the
// wave where such a run starts adds its tail sum and the head sums of the
// waves the run covers, in order, to the cell.  The same input therefore
// always gives the same bits.
// `keys` maps a key to its cell (CellKeys, or GridInterpolationBackward's
// ClassKeys when a run sum reuses that sort).
template <class Src, class Keys>
__global__ void __launch_bounds__(256) quadric_seg_kernel(const uint32_t *__restrict__ key,
                                                          Src src, int64_t n, uint32_t invalid,
                                                          Keys keys,
                                                          float *__restrict__ gq,
                                                          float *__restrict__ head,
                                                          float *__restrict__ tail) {
    const int lane = threadIdx.x % kWave;
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int64_t p0 = w * kWave;
    if (p0 >= n) return;  // wave-uniform
    const int64_t p = p0 + lane;
    // Every lane requests its key and the wave's two neighbouring keys at
    // once (positions clamped to the records), then its record and quadric
    // (cell 0 for an invalid key, whose sums are dropped): two memory round
    // trips, where lane-masked requests are each waited for in turn.
    const uint32_t kraw = key[p < n ? p : n - 1];
    const uint32_t kleft = key[p0 > 0 ? p0 - 1 : 0];
    const uint32_t kright = key[p0 + kWave < n ? p0 + kWave : n - 1];
    const uint32_t k = p < n ? kraw : invalid;
    float v[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    src.add(p < n ? p : n - 1, k != invalid ? keys.cell(k) : 0, v);
    if (k == invalid) {
#pragma unroll
        for (int j = 0; j < 7; ++j) v[j] = 0.0f;
    }
    const uint32_t kp = __shfl_up(k, 1, kWave), kn = __shfl_down(k, 1, kWave);
    const bool start = lane == 0 || kp != k;
    const bool last = lane == kWave - 1 || kn != k;
    const unsigned long long starts = __ballot(start);
    bool f = start;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        float pv[7];
#pragma unroll
        for (int j = 0; j < 7; ++j) pv[j] = __shfl_up(v[j], d, kWave);
        const bool pf = __shfl_up((int)f, d, kWave) != 0;
        if (lane >= d && !f) {
#pragma unroll
            for (int j = 0; j < 7; ++j) v[j] += pv[j];
            f = pf;
        }
    }
    const uint32_t k0 = __shfl(k, 0, kWave), k63 = __shfl(k, kWave - 1, kWave);
    const bool cont_in = p0 > 0 && kleft == k0;
    const bool cont_out = p0 + kWave < n && kright == k63;
    if (last && k != invalid) {
        // the segment holds lane 0 when no segment starts in lanes 1..lane
        const bool has0 = (starts & ((2ull << lane) - 2ull)) == 0;
        float *dst;
        if (has0 && cont_in) dst = head + 7 * w;
        else if (lane == kWave - 1 && cont_out) dst = tail + 7 * w;
        else dst = nullptr;
        if (dst != nullptr) {
#pragma unroll
            for (int j = 0; j < 7; ++j) dst[j] = v[j];
        } else {
            float *g = gq + 7 * keys.cell(k);
#pragma unroll
            for (int j = 0; j < 7; ++j) g[j] += v[j];
        }
    }
}

// Pass 2: one thread per wave of pass 1.
template <class Keys>
__global__ void __launch_bounds__(256) quadric_seg_join_kernel(const uint32_t *__restrict__ key,
                                                               int64_t n, uint32_t invalid,
                                                               Keys keys,
                                                               const float *__restrict__ head,
                                                               const float *__restrict__ tail,
                                                               float *__restrict__ gq) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t p0 = w * kWave;
    if (p0 + kWave >= n) return;  // the last wave has no run leaving it
    const uint32_t k = key[p0 + kWave - 1];
    if (k == invalid || key[p0 + kWave] != k) return;
    if (p0 > 0 && key[p0] == k && key[p0 - 1] == k) return;  // a middle piece of an earlier run
    float t[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) t[j] = tail[7 * w + j];
    for (int64_t u = w + 1;; ++u) {
#pragma unroll
        for (int j = 0; j < 7; ++j) t[j] += head[7 * u + j];
        const int64_t q = (u + 1) * kWave;
        if (q >= n || key[q] != k) break;
    }
    float *g = gq + 7 * keys.cell(k);
#pragma unroll
    for (int j = 0; j < 7; ++j) g[j] += t[j];
}

// Both passes over n sorted records; head and tail hold 7 floats per wave.
template <class Src, class Keys = CellKeys>
int quadric_seg_sum(const uint32_t *key, Src src, int64_t n, uint32_t invalid, float *gq,
                    float *head, float *tail, hipStream_t s, Keys keys = Keys{}) {
    if (n == 0) return QSDVR_OK;
    quadric_seg_kernel<<<grid_1d(n, 256), 256, 0, s>>>(key, src, n, invalid, keys, gq, head,
                                                       tail);
    QSDVR_LAUNCHED();
    const int64_t waves = (n + kWave - 1) / kWave;
    quadric_seg_join_kernel<<<grid_1d(waves, 256), 256, 0, s>>>(key, n, invalid, keys, head,
                                                                tail, gq);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

// Class boundaries in the sorted appearance keys: bounds[c] = first
// position whose key is >= c * stride (c = 0..8; bounds[8] starts the
// invalid keys).
__global__ void det_class_bounds_kernel(const uint32_t *__restrict__ key, int64_t n,
                                        uint32_t stride, uint32_t *__restrict__ bounds) {
    const int c = threadIdx.x;
    if (c > 8) return;
    const uint32_t target = (uint32_t)c * stride;
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
// Appearance record r: local point and gradient row, from the adjoint's
// records or (op surface) GridInterpolationBackward's point list and
// upstream gradient rows.
template <int C>
struct AppRecords {
    const float *aval;
    __device__ __forceinline__ f3 local(uint32_t r) const {
        const float4 lp = *reinterpret_cast<const float4 *>(aval + (int64_t)r * det_app_stride(C));
        return f3{lp.x, lp.y, lp.z};
    }
    __device__ __forceinline__ float grad(uint32_t r, int ch) const {
        return aval[(int64_t)r * det_app_stride(C) + 4 + ch];
    }
    // channels 4q .. 4q + 3 (16-B aligned: the stride and the row offset are
    // multiples of 4 floats)
    __device__ __forceinline__ float4 grad4(uint32_t r, int q) const {
        return *reinterpret_cast<const float4 *>(aval + (int64_t)r * det_app_stride(C) + 4 + 4 * q);
    }
};

template <int C, class Src>
__global__ void __launch_bounds__(256) det_app_reduce_kernel(
    const uint32_t *__restrict__ key, const uint32_t *__restrict__ idx, Src src,
    const uint32_t *__restrict__ bounds, int cls, int cbits, int R, float *__restrict__ gdata) {
    static_assert(C <= kWave, "one lane per channel");
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
#pragma unroll 4
                for (int64_t i = start; i < end; ++i) {
                    const uint32_t r = idx[i];
                    const float g = ch < C ? src.grad(r, ch) : 0.0f;
                    float w[8];
                    trilinear_weights(src.local(r), w);
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

// det_app_reduce_kernel for C = 32 with 8 lanes per run, each lane four
// channels (16-B row loads): 8 runs in flight per wavefront instead of 2,
// for the short runs of random ray batches.  Each (vertex, channel) sum
// still runs over the run's records in order, so the result is bitwise
// that of det_app_reduce_kernel<32>.  Needs the rows and gdata 16-B aligned.
template <class Src>
__global__ void __launch_bounds__(256) det_app_reduce32x4_kernel(
    const uint32_t *__restrict__ key, const uint32_t *__restrict__ idx, Src src,
    const uint32_t *__restrict__ bounds, int cls, int cbits, int R, float *__restrict__ gdata) {
    constexpr int C = 32, GW = 8, NG = kWave / GW;
    const int64_t lo = bounds[cls], hi = bounds[cls + 1];
    const int lane = threadIdx.x % kWave;
    const int grp = lane / GW, q = lane % GW;
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
            // group g takes the g-th of the remaining run heads (wave-uniform)
            int my = -1;
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const int h = heads != 0 ? __ffsll((long long)heads) - 1 : -1;
                if (g == grp) my = h;
                if (heads != 0) heads &= heads - 1;
            }
            const uint32_t hk = __shfl(k, my < 0 ? 0 : my, kWave);
            if (my >= 0) {
                const unsigned long long later = bmask & ~((2ull << my) - 1);
                const int64_t start = p0 + my;
                const int64_t end = later != 0 ? p0 + (__ffsll((long long)later) - 1) : tail_end;
                float4 acc[8];
#pragma unroll
                for (int v = 0; v < 8; ++v) acc[v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll 4
                for (int64_t i = start; i < end; ++i) {
                    const uint32_t r = idx[i];
                    const float4 g = src.grad4(r, q);
                    float w[8];
                    trilinear_weights(src.local(r), w);
#pragma unroll
                    for (int v = 0; v < 8; ++v) {
                        acc[v].x += w[v] * g.x;
                        acc[v].y += w[v] * g.y;
                        acc[v].z += w[v] * g.z;
                        acc[v].w += w[v] * g.w;
                    }
                }
                const int64_t cell = hk & cmask;
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    const int64_t row = cell + (v & 1) + ((v >> 1) & 1) * (int64_t)R + (v >> 2) * RR;
                    float4 *dst = reinterpret_cast<float4 *>(gdata + row * C + 4 * q);
                    float4 o = *dst;
                    o.x += acc[v].x;
                    o.y += acc[v].y;
                    o.z += acc[v].z;
                    o.w += acc[v].w;
                    *dst = o;
                }
            }
        }
    }
}

// The op surface's appearance runs (GridInterpolationBackward): as
// det_app_reduce_kernel, with the run's records fetched eight at a time --
// every lane of the group loads the next 8 (point, sample index) records,
// which the sort left in order, and then issues the 8 upstream rows they
// index together, so a run costs two memory round trips per 8 records
// instead of one dependent chain per record.  Each (vertex, channel) sum
// runs over the run's records in order.  The walk loop is uniform across
// the wavefront (groups whose run is done issue masked loads).
template <int C>
__global__ void __launch_bounds__(256) op_app_reduce_kernel(
    const uint32_t *__restrict__ key, const PointRec *__restrict__ pay,
    const float *__restrict__ up, const uint32_t *__restrict__ bounds, int cls, ClassKeys ck,
    float *__restrict__ gdata) {
    static_assert(C <= kWave, "one lane per channel");
    constexpr int NG = C <= kWave / 2 ? 2 : 1;  // runs in flight per wave
    constexpr int GW = kWave / NG;
    constexpr int B = 8;                        // records per fetch
    const int64_t lo = bounds[cls], hi = bounds[cls + 1];
    const int lane = threadIdx.x % kWave;
    const int grp = lane / GW, ch = lane % GW;
    const int64_t R = ck.R, RR = (int64_t)R * R;
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
            int64_t start = 0, end = 0;
            if (my >= 0) {
                const unsigned long long later = bmask & ~((2ull << my) - 1);
                start = p0 + my;
                end = later != 0 ? p0 + (__ffsll((long long)later) - 1) : tail_end;
            }
            float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            for (int64_t i = start;; i += B) {
                if (!__any(i < end)) break;
                // unconditional loads (a finished lane reads the class's
                // first record, a lane past the channels the row's last),
                // so that all of them issue before the first wait
                float4 rec[B];
#pragma unroll
                for (int b = 0; b < B; ++b)
                    rec[b] = *reinterpret_cast<const float4 *>(pay + (i + b < end ? i + b : lo));
                __builtin_amdgcn_sched_barrier(0);  // keep the 8 record loads together
                const int cc = ch < C ? ch : C - 1;
                float g[B];
#pragma unroll
                for (int b = 0; b < B; ++b)
                    g[b] = up[(int64_t)__float_as_uint(rec[b].w) * C + cc];
                __builtin_amdgcn_sched_barrier(0);  // and the 8 row loads
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    // a select, not a branch (which would let the compiler
                    // sink the record's loads into it)
                    const bool ok = i + b < end && ch < C;
                    float w[8];
                    trilinear_weights(f3{rec[b].x, rec[b].y, rec[b].z}, w);
#pragma unroll
                    for (int v = 0; v < 8; ++v) {
                        const float t = acc[v] + w[v] * g[b];
                        acc[v] = ok ? t : acc[v];
                    }
                }
            }
            if (my >= 0 && ch < C) {
                const int64_t cell = ck.cell(hk);
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    const int64_t row = cell + (v & 1) + ((v >> 1) & 1) * R + (v >> 2) * RR;
                    gdata[row * C + ch] += acc[v];
                }
            }
        }
    }
}

// op_app_reduce_kernel for C = 32 with one run per wavefront and lane =
// (channel quad q, record slot b): each step takes 8 records of the run --
// one 16-B load brings their (point, index) records, one more their full
// 128-B upstream rows -- and lane (q, b) adds record b's 8 vertex weights
// times its 4 channels to 32 running sums (records b, b + 8, ... of the run,
// two sets of 8 per round trip).
// A memory instruction costs the address unit the same however few lanes
// or addresses it carries, so this needs 2 per 8 records where a lane per
// channel needs 2 per record.  At the run's end a reduce-scatter over the 8
// slots (three DPP exchange-and-add steps) leaves vertex b's 4 channels in
// lane (q, b), and one read-modify-write per lane adds the 8 vertex rows.
// Runs of a frame are long (a median of about 33 samples per cell at
// 128^3, 800x800), so few slots idle.  A fixed order: bitwise reproducible.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
}

__global__ void __launch_bounds__(256) op_app_reduce32_kernel(
    const uint32_t *__restrict__ key, const PointRec *__restrict__ pay,
    const float4 *__restrict__ up, const uint32_t *__restrict__ bounds, int cls, ClassKeys ck,
    float *__restrict__ gdata) {
    constexpr int kHalfMirror = 0x141, kSwap2 = 0x4E, kSwap1 = 0xB1;  // DPP controls
    const int64_t lo = bounds[cls], hi = bounds[cls + 1];
    const int lane = threadIdx.x % kWave;
    const int q = lane >> 3, b = lane & 7;
    const int64_t R = ck.R, RR = (int64_t)R * R;
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
            const int h = __ffsll((long long)heads) - 1;
            heads &= heads - 1;
            const int64_t start = p0 + h;
            const int64_t end = heads != 0 ? p0 + (__ffsll((long long)heads) - 1) : tail_end;
            const uint32_t hk = __shfl(k, h, kWave);
            // lane (q, b)'s quad of vertex row b, read now: no other run of
            // this class touches the cell's rows, and the read is then done
            // by the time the sums are
            const int64_t row = ck.cell(hk) + (b & 1) + ((b >> 1) & 1) * R + (b >> 2) * RR;
            float4 *dst = reinterpret_cast<float4 *>(gdata + row * 32) + q;
            float4 o = *dst;
            float acc[8][4];
#pragma unroll
            for (int v = 0; v < 8; ++v)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[v][c] = 0.0f;
            // 16 records a step (slot b takes i0 + b and i0 + 8 + b), the next
            // step's records loaded while this step's rows are; a slot past the
            // run's end reads the run's first record and adds nothing (its row
            // is zeroed)
            auto rec_at = [&](int64_t i) {
                return *reinterpret_cast<const float4 *>(pay + (i < end ? i : start));
            };
            float4 na = rec_at(start + b), nb = rec_at(start + 8 + b);
            for (int64_t i0 = start; i0 < end; i0 += 16) {
                const float4 ra = na, rb = nb;
                float4 ga = up[(int64_t)__float_as_uint(ra.w) * 8 + q];
                float4 gb = up[(int64_t)__float_as_uint(rb.w) * 8 + q];
                na = rec_at(i0 + 16 + b);
                nb = rec_at(i0 + 24 + b);
                if (i0 + b >= end) ga = float4{0.0f, 0.0f, 0.0f, 0.0f};
                if (i0 + 8 + b >= end) gb = float4{0.0f, 0.0f, 0.0f, 0.0f};
                float wa[8], wb[8];
                trilinear_weights(f3{ra.x, ra.y, ra.z}, wa);
                trilinear_weights(f3{rb.x, rb.y, rb.z}, wb);
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    acc[v][0] += wa[v] * ga.x;
                    acc[v][1] += wa[v] * ga.y;
                    acc[v][2] += wa[v] * ga.z;
                    acc[v][3] += wa[v] * ga.w;
                }
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    acc[v][0] += wb[v] * gb.x;
                    acc[v][1] += wb[v] * gb.y;
                    acc[v][2] += wb[v] * gb.z;
                    acc[v][3] += wb[v] * gb.w;
                }
            }
            // reduce-scatter over the 8 slots: split the vertices by bit 2
            // with slot 7 - b, by bit 1 with slot b ^ 2, by bit 0 with b ^ 1
            const bool h2 = (b & 4) != 0, h1 = (b & 2) != 0, h0 = (b & 1) != 0;
            float k1[4][4], k2[2][4], tot[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float keep = h2 ? acc[4 + j][c] : acc[j][c];
                    const float send = h2 ? acc[j][c] : acc[4 + j][c];
                    k1[j][c] = keep + dpp_f<kHalfMirror>(send);
                }
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float keep = h1 ? k1[2 + j][c] : k1[j][c];
                    const float send = h1 ? k1[j][c] : k1[2 + j][c];
                    k2[j][c] = keep + dpp_f<kSwap2>(send);
                }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float keep = h0 ? k2[1][c] : k2[0][c];
                const float send = h0 ? k2[0][c] : k2[1][c];
                tot[c] = keep + dpp_f<kSwap1>(send);
            }
            o.x += tot[0];
            o.y += tot[1];
            o.z += tot[2];
            o.w += tot[3];
            *dst = o;
        }
    }
}

// GridInterpolationForward (kernel/render.cu:317-323, the lerp order of
// models/render_torch.py:128-143) over the samples in GridInterpolation's
// sorted order, C = 32: a wavefront takes one run of a cell's samples at a
// time with lane = (channel quad q, sample slot b).  Lane (q, b) loads quad q
// of the cell's vertex row b -- one 16-B load per lane brings all 8 rows --
// and ds_swizzle broadcasts give each lane quad q of the 8 rows; then every
// 8 samples cost one 16-B load of their (point, index) records and one
// 16-B store per lane, against 8 row loads per 8 samples when each sample
// gathers its own rows (the address unit's cost is per instruction).  The
// outputs are the per-sample gather's, bit for bit.  Waves take contiguous
// ranges of sorted positions (neighbouring cells share rows in one XCD's
// L2), request each chunk's keys one chunk ahead and find run heads with a
// lane shuffle.  Samples with an invalid index (sorted last) get NaN.
template <int BCAST>
__device__ __forceinline__ float4 bcast8(float4 v) {
    // bit-mode swizzle within 32 lanes: lane (l & 0x18) | BCAST
    constexpr int pat = 0x18 | (BCAST << 5);
    return float4{__int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v.x), pat)),
                  __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v.y), pat)),
                  __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v.z), pat)),
                  __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v.w), pat))};
}

__device__ __forceinline__ float4 lerp4s(float4 a, float4 b, float w) {
    return float4{lerp_ref(a.x, b.x, w), lerp_ref(a.y, b.y, w), lerp_ref(a.z, b.z, w),
                  lerp_ref(a.w, b.w, w)};
}

__global__ void __launch_bounds__(256) op_interp_fwd_runs32_kernel(
    const uint32_t *__restrict__ key, const PointRec *__restrict__ pay,
    const float4 *__restrict__ data, const uint32_t *__restrict__ bounds, ClassKeys ck,
    int64_t n, int64_t chunks_per_wave, float4 *__restrict__ out) {
    const int64_t hi = bounds[8];  // valid samples: sorted positions [0, hi)
    const int lane = threadIdx.x % kWave;
    const int q = lane >> 3, b = lane & 7;
    const int64_t R = ck.R, RR = R * R;
    const int64_t wave = xcd_block() * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int64_t c0 = wave * chunks_per_wave;
    const int64_t pend = std::min<int64_t>((c0 + chunks_per_wave) * kWave, hi);
    auto key_at = [&](int64_t p) { return p < hi ? key[p] : 0xFFFFFFFFu; };
    int64_t p0 = c0 * kWave;
    // the key before the range, and this chunk's keys; each chunk's keys
    // are requested one chunk ahead
    uint32_t prev = p0 > 0 && p0 < pend ? key[p0 - 1] : 0xFFFFFFFFu;
    uint32_t k = p0 < pend ? key_at(p0 + lane) : 0xFFFFFFFFu;
    for (; p0 < pend; p0 += kWave) {
        // (also past the range's end: its last run may continue into the next range)
        const uint32_t kn = p0 + kWave < hi ? key_at(p0 + kWave + lane) : 0xFFFFFFFFu;
        const int64_t p = p0 + lane;
        uint32_t left = __shfl_up(k, 1, kWave);
        if (lane == 0) left = prev;
        prev = __shfl(k, kWave - 1, kWave);
        const bool boundary = p < hi && (p == 0 || left != k);
        const unsigned long long bmask = __ballot(boundary);
        if (bmask != 0) {
            const int last = 63 - __clzll((long long)bmask);
            // the last run's end: in the next chunk's keys, else further on
            const uint32_t lk = __shfl(k, last, kWave);
            const unsigned long long nx = __ballot(kn != lk);
            const int64_t tail_end = nx != 0 ? std::min<int64_t>(p0 + kWave + (__ffsll((long long)nx) - 1), hi)
                                             : run_end(key, p0 + 2 * kWave, hi, lk);
            // The chunk's runs one after another, each run's row quad and first
            // records requested while the run before it is interpolated.
            struct RunIn {
                int64_t start, end;
                float4 mine, na, nb;
            };
            unsigned long long heads = bmask;
            auto fetch = [&]() {
                RunIn r;
                const int h = __ffsll((long long)heads) - 1;
                heads &= heads - 1;
                r.start = p0 + h;
                r.end = heads != 0 ? p0 + (__ffsll((long long)heads) - 1) : tail_end;
                const int64_t cell = ck.cell(__shfl(k, h, kWave));
                const int64_t row = cell + (b & 1) + ((b >> 1) & 1) * R + (b >> 2) * RR;
                r.na = *reinterpret_cast<const float4 *>(pay + std::min(r.start + b, r.end - 1));
                r.nb = *reinterpret_cast<const float4 *>(pay + std::min(r.start + 8 + b, r.end - 1));
                r.mine = data[row * 8 + q];
                return r;
            };
            RunIn cur = fetch();
            for (;;) {
                const bool more = heads != 0;
                RunIn nxt = cur;
                if (more) nxt = fetch();
                const int64_t start = cur.start, end = cur.end;
                // 16 samples a round trip (slot b takes i0 + b and i0 + 8 + b),
                // the next round's records requested before this round's stores
                auto rec_at = [&](int64_t i) {
                    return *reinterpret_cast<const float4 *>(pay + (i < end ? i : start));
                };
                const float4 mine = cur.mine;
                const float4 v0 = bcast8<0>(mine), v1 = bcast8<1>(mine), v2 = bcast8<2>(mine),
                             v3 = bcast8<3>(mine), v4 = bcast8<4>(mine), v5 = bcast8<5>(mine),
                             v6 = bcast8<6>(mine), v7 = bcast8<7>(mine);
                auto interp = [&](float4 r) {
                    const float4 a00 = lerp4s(v0, v1, r.x), a01 = lerp4s(v2, v3, r.x);
                    const float4 a0 = lerp4s(a00, a01, r.y);
                    const float4 a10 = lerp4s(v4, v5, r.x), a11 = lerp4s(v6, v7, r.x);
                    const float4 a1 = lerp4s(a10, a11, r.y);
                    return lerp4s(a0, a1, r.z);
                };
                float4 na = cur.na, nb = cur.nb;
                for (int64_t i0 = start; i0 < end; i0 += 16) {
                    const float4 ra = na, rb = nb;
                    if (i0 + 16 < end) {
                        na = rec_at(i0 + 16 + b);
                        nb = rec_at(i0 + 24 + b);
                    }
                    const float4 oa = interp(ra), ob = interp(rb);
                    if (i0 + b < end) out[(int64_t)__float_as_uint(ra.w) * 8 + q] = oa;
                    if (i0 + 8 + b < end) out[(int64_t)__float_as_uint(rb.w) * 8 + q] = ob;
                }
                if (!more) break;
                cur = nxt;
            }
        }
        k = kn;
    }
    // the invalid samples (normally none): NaN rows, 8 samples a wave step
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
    const float nan = __builtin_nanf("");
    for (int64_t i = hi + wave * 8 + b; i < n; i += nwaves * 8) {
        const uint32_t idx = pay[i].i;
        out[(int64_t)idx * 8 + q] = float4{nan, nan, nan, nan};
    }
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

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
    int st = quadric_seg_sum(b.qkey_s, QuadricRecords{b.qval, b.qidx_s}, 2 * T, (uint32_t)R3, gq,
                             b.qhead, b.qtail, s);
    if (st != QSDVR_OK) return st;
    det_class_bounds_kernel<<<1, 64, 0, s>>>(b.akey_s, T, 1u << b.cbits, b.bounds);
    QSDVR_LAUNCHED();
    // each class holds about T/8 records; the waves stride over its range
    const int64_t chunks = (T / 8 + kWave - 1) / kWave;
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((chunks + 3) / 4, 4096));
    for (int cls = 0; cls < 8; ++cls) {
        if (b.C == 32 && aligned16(b.aval) && aligned16(gdata))
            det_app_reduce32x4_kernel<<<blocks, 256, 0, s>>>(b.akey_s, b.aidx_s,
                                                            AppRecords<32>{b.aval}, b.bounds, cls,
                                                            b.cbits, b.R, gdata);
        else if (b.C == 32)
            det_app_reduce_kernel<32><<<blocks, 256, 0, s>>>(b.akey_s, b.aidx_s,
                                                             AppRecords<32>{b.aval}, b.bounds, cls,
                                                             b.cbits, b.R, gdata);
        else
            det_app_reduce_kernel<53><<<blocks, 256, 0, s>>>(b.akey_s, b.aidx_s,
                                                             AppRecords<53>{b.aval}, b.bounds, cls,
                                                             b.cbits, b.R, gdata);
        QSDVR_LAUNCHED();
    }
    return QSDVR_OK;
}

// ------------------------------------- op-surface sorted scatters ----
// GridInterpolationBackward, SampleSDFBackward and SampleNormalBackward as
// sorted-segment sums.  GridInterpolationBackward sorts (point, sample
// index) records by ClassKeys and SampleSDFBackward its 16-B records (point,
// upstream) by cell, so that their run sums read the points in order;
// SampleNormalBackward sorts sample indices by cell and gathers its records
// in sorted order, or sums its runs on the sort of a GridInterpolationBackward
// over the same samples.

// ClassKeys key and (point, index) record of sample i
__global__ void __launch_bounds__(256) op_interp_keys_kernel(const int32_t *__restrict__ idx,
                                                             const float *__restrict__ pts,
                                                             int64_t n, ClassKeys ck,
                                                             uint32_t *__restrict__ key,
                                                             PointRec *__restrict__ pay) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t cell = idx[i], R = ck.R;
    uint32_t k = 8u * ck.h3;
    if (cell >= 0 && cell < R * R * R) k = ck.key(cell % R, (cell / R) % R, cell / (R * R));
    key[i] = k;
    const f3 p = load3(pts, i);
    *reinterpret_cast<float4 *>(pay + i) = float4{p.x, p.y, p.z, __uint_as_float((uint32_t)i)};
}

// key of sample i: its cell, invalid `cells`; the record (point and
// upstream gradient) is packed next to it and sorted with it, so that the
// run sums read their records in order instead of gathering them
__global__ void __launch_bounds__(256) op_sdf_keys_kernel(const int32_t *__restrict__ idx,
                                                          const float *__restrict__ pts,
                                                          const float *__restrict__ up,
                                                          int64_t n, int64_t cells,
                                                          uint32_t *__restrict__ key,
                                                          float4 *__restrict__ pay) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t cell = idx[i];
    key[i] = (cell >= 0 && cell < cells) ? (uint32_t)cell : (uint32_t)cells;
    const f3 p = load3(pts, i);
    pay[i] = float4{p.x, p.y, p.z, up[i]};
}

// key of sample i: its cell, invalid `cells`; value: i
__global__ void __launch_bounds__(256) op_cell_idx_kernel(const int32_t *__restrict__ idx,
                                                          int64_t n, int64_t cells,
                                                          uint32_t *__restrict__ key,
                                                          uint32_t *__restrict__ val) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t cell = idx[i];
    key[i] = (cell >= 0 && cell < cells) ? (uint32_t)cell : (uint32_t)cells;
    val[i] = (uint32_t)i;
}

namespace {

// val / val_s hold sample indices or SampleSDF's 16-B records
struct OpBuffers {
    uint32_t *key, *key_s, *bounds;
    char *val, *val_s;
    float *head, *tail;  // quadric segment sums: 7 floats per wave of 64 records
    void *temp;
    size_t temp_bytes;
};

template <typename V>
size_t sort_temp_bytes_v(int64_t n, int end_bit) {
    size_t t = 0;
    (void)rocprim::radix_sort_pairs(nullptr, t, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                    (const V *)nullptr, (V *)nullptr, (size_t)n, 0u,
                                    (unsigned)end_bit, (hipStream_t)0);
    return t;
}

size_t op_temp_bytes(int64_t n, int64_t cells) {
    const int qb = bits_above((uint64_t)cells);
    return std::max({sort_temp_bytes(n, qb), sort_temp_bytes_v<float4>(n, qb),
                     sort_temp_bytes_v<PointRec>(n, qb)});
}

size_t op_carve(void *ws, size_t bytes, int64_t n, int64_t cells, OpBuffers &b) {
    Carver c{static_cast<char *>(ws), bytes};
    b.key = c.take<uint32_t>(n);
    b.key_s = c.take<uint32_t>(n);
    b.val = c.take<char>(n * sizeof(float4));
    b.val_s = c.take<char>(n * sizeof(float4));
    b.bounds = c.take<uint32_t>(16);
    b.head = c.take<float>(7 * seg_waves(n));
    b.tail = c.take<float>(7 * seg_waves(n));
    b.temp_bytes = op_temp_bytes(n, cells);
    b.temp = c.take<char>(b.temp_bytes);
    return c.used + 256;
}

int op_prepare(void *ws, size_t bytes, int64_t n, int64_t cells, OpBuffers &b, const char *who) {
    if (n >= 0xFFFFFFFFll || cells < 1 || cells >= 0x7FFFFFFFll) {
        set_error("%s: %lld samples / %lld cells exceed the 32-bit sort keys", who, (long long)n,
                  (long long)cells);
        return QSDVR_ERANGE;
    }
    const size_t need = op_carve(nullptr, 0, n, cells, b);
    if (ws == nullptr || bytes < need) {
        set_error("%s: workspace of %zu bytes is too small (%zu needed)", who, bytes, need);
        return QSDVR_EWORKSPACE;
    }
    op_carve(ws, bytes, n, cells, b);
    return QSDVR_OK;
}

// GridInterpolation's sort: ClassKeys keys of the samples, sorted with
// their (point, index) records, and bounds[0..8] of the parity classes in
// the sorted keys (bounds[8] starts the invalid samples).
int interp_sort(const float *point, const int32_t *index, int64_t n, int R, OpBuffers &b,
                hipStream_t s) {
    const ClassKeys ck = class_keys(R);
    PointRec *pay = reinterpret_cast<PointRec *>(b.val), *pay_s = reinterpret_cast<PointRec *>(b.val_s);
    op_interp_keys_kernel<<<grid_1d(n, 256), 256, 0, s>>>(index, point, n, ck, b.key, pay);
    QSDVR_LAUNCHED();
    size_t tb = b.temp_bytes;
    QSDVR_HIP(rocprim::radix_sort_pairs(b.temp, tb, b.key, b.key_s, pay, pay_s, (size_t)n, 0u,
                                        (unsigned)bits_above((uint64_t)R * R * R), s));
    det_class_bounds_kernel<<<1, 64, 0, s>>>(b.key_s, n, ck.h3, b.bounds);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

// GridInterpolationBackward's sums on that sort: 8 passes, one per cell
// parity class (cells of one class share no vertex row).
int interp_class_sums(float *grad_grid, const float *grad_out, int64_t n, int R, int channels,
                      const OpBuffers &b, hipStream_t s) {
    const ClassKeys ck = class_keys(R);
    const PointRec *pay_s = reinterpret_cast<const PointRec *>(b.val_s);
    const int64_t chunks = (n / 8 + kWave - 1) / kWave;
    const unsigned blocks =
        (unsigned)std::max<int64_t>(1, std::min<int64_t>((chunks + 3) / 4, 4096));
    const bool quads = channels == 32 && aligned16(grad_out) && aligned16(grad_grid);
    for (int cls = 0; cls < 8; ++cls) {
        // (one run per wavefront or per half: a frame's samples sorted by cell
        // make long, uneven runs, and a wavefront of 8 runs waits for its
        // longest)
        if (quads)
            op_app_reduce32_kernel<<<blocks, 256, 0, s>>>(
                b.key_s, pay_s, reinterpret_cast<const float4 *>(grad_out), b.bounds, cls, ck,
                grad_grid);
        else if (channels == 32)
            op_app_reduce_kernel<32><<<blocks, 256, 0, s>>>(b.key_s, pay_s, grad_out, b.bounds,
                                                           cls, ck, grad_grid);
        else
            op_app_reduce_kernel<53><<<blocks, 256, 0, s>>>(b.key_s, pay_s, grad_out, b.bounds,
                                                           cls, ck, grad_grid);
        QSDVR_LAUNCHED();
    }
    return QSDVR_OK;
}

}  // namespace

}  // namespace qsdvr

using namespace qsdvr;

extern "C" {

size_t qsdvr_sorted_scatter_workspace_size(int64_t n, int64_t cell_count) {
    if (n < 0 || cell_count < 1 || n >= 0xFFFFFFFFll || cell_count >= 0x7FFFFFFFll) return 0;
    OpBuffers b;
    return op_carve(nullptr, 0, n, cell_count, b);
}

int qsdvr_grid_interpolation_backward_sorted(float *grad_grid, const float *point,
                                             const int32_t *index, const float *grad_out,
                                             int64_t n, int32_t vertex_reso, int32_t channels,
                                             void *workspace, size_t workspace_bytes,
                                             qsdvr_stream stream) {
    QSDVR_REQUIRE(n >= 0 && vertex_reso >= 3 && (vertex_reso - 1) % 2 == 0 &&
                      (channels == 32 || channels == 53),
                  "qsdvr_grid_interpolation_backward_sorted: bad sizes (even cell reso, "
                  "channels 32 or 53)");
    if (n == 0) return QSDVR_OK;
    QSDVR_REQUIRE(grad_grid && point && index && grad_out,
                  "qsdvr_grid_interpolation_backward_sorted: null pointer");
    const int R = vertex_reso - 1;
    OpBuffers b;
    int st = op_prepare(workspace, workspace_bytes, n, (int64_t)R * R * R, b,
                        "qsdvr_grid_interpolation_backward_sorted");
    if (st != QSDVR_OK) return st;
    hipStream_t s = as_stream(stream);
    st = interp_sort(point, index, n, R, b, s);
    if (st != QSDVR_OK) return st;
    return interp_class_sums(grad_grid, grad_out, n, R, channels, b, s);
}

int qsdvr_grid_interpolation_forward_sorted(const float *data_grid, const float *point,
                                            const int32_t *index, float *out, int64_t n,
                                            int32_t vertex_reso, int32_t channels,
                                            void *workspace, size_t workspace_bytes,
                                            qsdvr_stream stream) {
    QSDVR_REQUIRE(n >= 0 && vertex_reso >= 3 && (vertex_reso - 1) % 2 == 0 && channels == 32,
                  "qsdvr_grid_interpolation_forward_sorted: bad sizes (even cell reso, "
                  "channels 32)");
    if (n == 0) return QSDVR_OK;
    QSDVR_REQUIRE(data_grid && point && index && out,
                  "qsdvr_grid_interpolation_forward_sorted: null pointer");
    QSDVR_REQUIRE(aligned16(data_grid) && aligned16(out),
                  "qsdvr_grid_interpolation_forward_sorted: data_grid and out must be 16-B "
                  "aligned");
    const int R = vertex_reso - 1;
    OpBuffers b;
    int st = op_prepare(workspace, workspace_bytes, n, (int64_t)R * R * R, b,
                        "qsdvr_grid_interpolation_forward_sorted");
    if (st != QSDVR_OK) return st;
    hipStream_t s = as_stream(stream);
    st = interp_sort(point, index, n, R, b, s);
    if (st != QSDVR_OK) return st;
    // 4 waves a block; each wave a contiguous range of 64-position chunks
    const int64_t chunks = (n + kWave - 1) / kWave;
    const int64_t waves = std::max<int64_t>(1, std::min<int64_t>(chunks, 65536));
    const int64_t per_wave = (chunks + waves - 1) / waves;
    const unsigned blocks = (unsigned)((waves + 3) / 4);
    op_interp_fwd_runs32_kernel<<<blocks, 256, 0, s>>>(
        b.key_s, reinterpret_cast<const PointRec *>(b.val_s),
        reinterpret_cast<const float4 *>(data_grid), b.bounds, class_keys(R), n, per_wave,
        reinterpret_cast<float4 *>(out));
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_grid_interpolation_backward_presorted(float *grad_grid, const float *grad_out,
                                                int64_t n, int32_t vertex_reso,
                                                int32_t channels, void *workspace,
                                                size_t workspace_bytes, qsdvr_stream stream) {
    QSDVR_REQUIRE(n >= 0 && vertex_reso >= 3 && (vertex_reso - 1) % 2 == 0 &&
                      (channels == 32 || channels == 53),
                  "qsdvr_grid_interpolation_backward_presorted: bad sizes (even cell reso, "
                  "channels 32 or 53)");
    if (n == 0) return QSDVR_OK;
    QSDVR_REQUIRE(grad_grid && grad_out,
                  "qsdvr_grid_interpolation_backward_presorted: null pointer");
    const int R = vertex_reso - 1;
    OpBuffers b;
    int st = op_prepare(workspace, workspace_bytes, n, (int64_t)R * R * R, b,
                        "qsdvr_grid_interpolation_backward_presorted");
    if (st != QSDVR_OK) return st;
    return interp_class_sums(grad_grid, grad_out, n, R, channels, b, as_stream(stream));
}

int qsdvr_sample_sdf_backward_sorted(const float *grad_out, const float *sdf_grid,
                                     const float *point, const int32_t *index, float *grad_grid,
                                     int64_t n, int64_t cell_count, float reso, void *workspace,
                                     size_t workspace_bytes, qsdvr_stream stream) {
    QSDVR_REQUIRE(n >= 0 && cell_count > 0 && reso != 0.0f,
                  "qsdvr_sample_sdf_backward_sorted: bad sizes");
    if (n == 0) return QSDVR_OK;
    QSDVR_REQUIRE(grad_out && sdf_grid && point && index && grad_grid,
                  "qsdvr_sample_sdf_backward_sorted: null pointer");
    hipStream_t s = as_stream(stream);
    OpBuffers b;
    int st = op_prepare(workspace, workspace_bytes, n, cell_count, b,
                        "qsdvr_sample_sdf_backward_sorted");
    if (st != QSDVR_OK) return st;
    float4 *pay = reinterpret_cast<float4 *>(b.val), *pay_s = reinterpret_cast<float4 *>(b.val_s);
    op_sdf_keys_kernel<<<grid_1d(n, 256), 256, 0, s>>>(index, point, grad_out, n, cell_count,
                                                       b.key, pay);
    QSDVR_LAUNCHED();
    size_t tb = b.temp_bytes;
    QSDVR_HIP(rocprim::radix_sort_pairs(b.temp, tb, b.key, b.key_s, pay, pay_s, (size_t)n, 0u,
                                        (unsigned)bits_above((uint64_t)cell_count), s));
    return quadric_seg_sum(b.key_s, SdfSamples{pay_s, sdf_grid, 1.0f / reso}, n,
                           (uint32_t)cell_count, grad_grid, b.head, b.tail, s);
}

int qsdvr_sample_normal_backward_sorted(const float *grad_out, const float *sdf_grid,
                                        const float *point, const int32_t *index,
                                        float *grad_grid, int64_t n, int64_t cell_count,
                                        void *workspace, size_t workspace_bytes,
                                        qsdvr_stream stream) {
    QSDVR_REQUIRE(n >= 0 && cell_count > 0, "qsdvr_sample_normal_backward_sorted: bad sizes");
    if (n == 0) return QSDVR_OK;
    QSDVR_REQUIRE(grad_out && sdf_grid && point && index && grad_grid,
                  "qsdvr_sample_normal_backward_sorted: null pointer");
    hipStream_t s = as_stream(stream);
    OpBuffers b;
    int st = op_prepare(workspace, workspace_bytes, n, cell_count, b,
                        "qsdvr_sample_normal_backward_sorted");
    if (st != QSDVR_OK) return st;
    uint32_t *val = reinterpret_cast<uint32_t *>(b.val), *val_s = reinterpret_cast<uint32_t *>(b.val_s);
    op_cell_idx_kernel<<<grid_1d(n, 256), 256, 0, s>>>(index, n, cell_count, b.key, val);
    QSDVR_LAUNCHED();
    size_t tb = b.temp_bytes;
    QSDVR_HIP(rocprim::radix_sort_pairs(b.temp, tb, b.key, b.key_s, val, val_s, (size_t)n, 0u,
                                        (unsigned)bits_above((uint64_t)cell_count), s));
    return quadric_seg_sum(b.key_s, NormalSamples{val_s, sdf_grid, point, grad_out}, n,
                           (uint32_t)cell_count, grad_grid, b.head, b.tail, s);
}

int qsdvr_sample_normal_backward_presorted(const float *grad_out, const float *sdf_grid,
                                           float *grad_grid, int64_t n, int32_t reso,
                                           void *workspace, size_t workspace_bytes,
                                           qsdvr_stream stream) {
    QSDVR_REQUIRE(n >= 0 && reso >= 2 && reso % 2 == 0,
                  "qsdvr_sample_normal_backward_presorted: bad sizes (even reso)");
    if (n == 0) return QSDVR_OK;
    QSDVR_REQUIRE(grad_out && sdf_grid && grad_grid,
                  "qsdvr_sample_normal_backward_presorted: null pointer");
    const int64_t cells = (int64_t)reso * reso * reso;
    OpBuffers b;
    int st = op_prepare(workspace, workspace_bytes, n, cells, b,
                        "qsdvr_sample_normal_backward_presorted");
    if (st != QSDVR_OK) return st;
    const PointRec *pay_s = reinterpret_cast<const PointRec *>(b.val_s);
    return quadric_seg_sum(b.key_s, NormalPresorted{pay_s, sdf_grid, grad_out}, n,
                           (uint32_t)cells, grad_grid, b.head, b.tail, as_stream(stream),
                           class_keys(reso));
}

}  // extern "C"
