// Op-surface appearance kernels for gfx950: GridInterpolation, Shader and
// RayAggregate, forward and backward.
// Reference semantics: models/render_torch.py:117-195 (the oracle), with the
// reference kernels of kernel/render.cu:98-363 as the operator surface.
#include "shade.h"

namespace qsdvr {

// ------------------------------------------------- GridInterpolation ----
// One lane per (sample, channel): a wavefront reads whole vertex rows
// (C consecutive floats) per gather, so every row load is one coalesced
// burst and every output row is written contiguously.
template <int CFIX>
__global__ void __launch_bounds__(256) interp_fwd_kernel(const float *__restrict__ data,
                                                         const float *__restrict__ pts,
                                                         const int32_t *__restrict__ idx,
                                                         float *__restrict__ out, int64_t n,
                                                         int R, int Cdyn) {
    const int C = CFIX > 0 ? CFIX : Cdyn;
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n * C) return;
    int64_t s = id / C;
    int c = (int)(id - s * C);
    int64_t cell = idx[s];
    int64_t cells = (int64_t)R * R * R;
    if (cell < 0 || cell >= cells) {
        out[id] = __builtin_nanf("");
        return;
    }
    f3 p = load3(pts, s);
    const int64_t RR = (int64_t)R * R;
    const float *b = data + cell * C + c;
    auto D = [&](int64_t off) { return b[off * C]; };
    float a00 = lerp_ref(D(0), D(1), p.x);
    float a01 = lerp_ref(D(R), D(R + 1), p.x);
    float a0 = lerp_ref(a00, a01, p.y);
    float a10 = lerp_ref(D(RR), D(RR + 1), p.x);
    float a11 = lerp_ref(D(RR + R), D(RR + R + 1), p.x);
    float a1 = lerp_ref(a10, a11, p.y);
    out[id] = lerp_ref(a0, a1, p.z);
}

// C = 32: lane = (sample slot, channel quad), 8 lanes per sample with one
// 16-B load per vertex row, 8 samples per wave step and ITER steps per wave;
// the same lerp order per channel as interp_fwd_kernel.  The loads carry no
// branches (a sample past the end reads the last sample, an invalid cell
// reads cell 0 and stores NaN), so that a lane issues the indices of all its
// steps, then the points and rows of all its steps, before it waits.
__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float w) {
    return float4{lerp_ref(a.x, b.x, w), lerp_ref(a.y, b.y, w), lerp_ref(a.z, b.z, w),
                  lerp_ref(a.w, b.w, w)};
}

template <int ITER>
__global__ void __launch_bounds__(256) interp_fwd32_kernel(const float4 *__restrict__ data,
                                                           const float *__restrict__ pts,
                                                           const int32_t *__restrict__ idx,
                                                           float4 *__restrict__ out, int64_t n,
                                                           int R) {
    const int lane = threadIdx.x % kWave, q = lane & 7, slot = lane >> 3;
    const int64_t wave = (xcd_block() * blockDim.x + threadIdx.x) / kWave;
    const int64_t cells = (int64_t)R * R * R, RR = (int64_t)R * R;
    int64_t s[ITER], cell[ITER];
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
        const int64_t si = (wave * ITER + it) * 8 + slot;
        s[it] = si < n ? si : n - 1;
        cell[it] = idx[s[it]];
    }
    f3 p[ITER];
    float4 v[ITER][8];
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
        p[it] = load3(pts, s[it]);
        const bool ok = cell[it] >= 0 && cell[it] < cells;
        const float4 *b = data + (ok ? cell[it] : 0) * 8 + q;
        const int64_t off[8] = {0, 1, R, R + 1, RR, RR + 1, RR + R, RR + R + 1};
#pragma unroll
        for (int k = 0; k < 8; ++k) v[it][k] = b[off[k] * 8];
    }
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
        const int64_t si = (wave * ITER + it) * 8 + slot;
        const float4 a00 = lerp4(v[it][0], v[it][1], p[it].x);
        const float4 a01 = lerp4(v[it][2], v[it][3], p[it].x);
        const float4 a0 = lerp4(a00, a01, p[it].y);
        const float4 a10 = lerp4(v[it][4], v[it][5], p[it].x);
        const float4 a11 = lerp4(v[it][6], v[it][7], p[it].x);
        const float4 a1 = lerp4(a10, a11, p[it].y);
        float4 r = lerp4(a0, a1, p[it].z);
        if (!(cell[it] >= 0 && cell[it] < cells)) {
            const float nan = __builtin_nanf("");
            r = float4{nan, nan, nan, nan};
        }
        if (si < n) out[si * 8 + q] = r;
    }
}

template <int CFIX>
__global__ void __launch_bounds__(256) interp_bwd_kernel(float *__restrict__ gdata,
                                                         const float *__restrict__ pts,
                                                         const int32_t *__restrict__ idx,
                                                         const float *__restrict__ up, int64_t n,
                                                         int R, int Cdyn) {
    const int C = CFIX > 0 ? CFIX : Cdyn;
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n * C) return;
    int64_t s = id / C;
    int c = (int)(id - s * C);
    int64_t cell = idx[s];
    if (cell < 0 || cell >= (int64_t)R * R * R) return;
    f3 p = load3(pts, s);
    float w[8];
    trilinear_weights(p, w);
    float g = up[id];
    const int64_t RR = (int64_t)R * R;
    float *b = gdata + cell * C + c;
    const int64_t off[8] = {0, 1, R, R + 1, RR, RR + 1, RR + R, RR + R + 1};
#pragma unroll
    for (int k = 0; k < 8; ++k) atomic_addf(b + off[k] * C, w[k] * g);
}

// ------------------------------------------------------------- Shader ----
// Lane = sample.  A sample's appearance row (C = 32 or 53 floats) is read
// and its gradient row written through LDS: the block's rows are one
// contiguous span, moved with coalesced 16-B accesses and stored in LDS
// with an odd row pitch, so that each lane then reads or writes its own
// row without bank conflicts (a lane-per-row access of global memory would
// touch 64 cache lines per instruction).
constexpr int kShadeBlock = 256;

template <int C, int P>
__device__ __forceinline__ void lds_put(float *lds, int e, float v) {
    const int r = e / C;
    lds[r * P + (e - r * C)] = v;
}

template <int C, int P>
__device__ __forceinline__ float lds_get(const float *lds, int e) {
    const int r = e / C;
    return lds[r * P + (e - r * C)];
}

// rows [row0, row0 + nrows) of the contiguous [*, C] array src -> LDS
template <int C, int P>
__device__ __forceinline__ void stage_rows_in(const float *__restrict__ src, int64_t row0,
                                              int nrows, float *lds) {
    const float *g = src + row0 * C;
    const int total = nrows * C;
    int done = 0;
    if (((uintptr_t)g & 15u) == 0) {
        const int n4 = total / 4;
        const float4 *g4 = reinterpret_cast<const float4 *>(g);
        for (int e4 = threadIdx.x; e4 < n4; e4 += blockDim.x) {
            const float4 v = g4[e4];
            lds_put<C, P>(lds, 4 * e4 + 0, v.x);
            lds_put<C, P>(lds, 4 * e4 + 1, v.y);
            lds_put<C, P>(lds, 4 * e4 + 2, v.z);
            lds_put<C, P>(lds, 4 * e4 + 3, v.w);
        }
        done = 4 * n4;
    }
    for (int e = done + threadIdx.x; e < total; e += blockDim.x) lds_put<C, P>(lds, e, g[e]);
}

// LDS -> rows [row0, row0 + nrows) of the contiguous [*, C] array dst
template <int C, int P>
__device__ __forceinline__ void stage_rows_out(float *__restrict__ dst, int64_t row0, int nrows,
                                               const float *lds) {
    float *g = dst + row0 * C;
    const int total = nrows * C;
    int done = 0;
    if (((uintptr_t)g & 15u) == 0) {
        const int n4 = total / 4;
        float4 *g4 = reinterpret_cast<float4 *>(g);
        for (int e4 = threadIdx.x; e4 < n4; e4 += blockDim.x)
            g4[e4] = float4{lds_get<C, P>(lds, 4 * e4 + 0), lds_get<C, P>(lds, 4 * e4 + 1),
                            lds_get<C, P>(lds, 4 * e4 + 2), lds_get<C, P>(lds, 4 * e4 + 3)};
        done = 4 * n4;
    }
    for (int e = done + threadIdx.x; e < total; e += blockDim.x) g[e] = lds_get<C, P>(lds, e);
}

template <int NC>
__global__ void __launch_bounds__(kShadeBlock) shader_fwd_kernel(const float *__restrict__ normal,
                                                                 const float *__restrict__ view,
                                                                 const float *__restrict__ data,
                                                                 float *__restrict__ out,
                                                                 float *__restrict__ spec,
                                                                 int64_t n) {
    constexpr int C = 3 * NC + 5;
    constexpr int P = C % 2 == 1 ? C : C + 1;
    __shared__ float rows[kShadeBlock * P];
    const int64_t row0 = (int64_t)blockIdx.x * kShadeBlock;
    const int nrows = (int)min((int64_t)kShadeBlock, n - row0);
    stage_rows_in<C, P>(data, row0, nrows, rows);
    __syncthreads();
    const int t = threadIdx.x;
    if (t >= nrows) return;
    const int64_t i = row0 + t;
    float d[C];
#pragma unroll
    for (int k = 0; k < C; ++k) d[k] = rows[t * P + k];
    ShadeOut o = shade<NC>(load3(normal, i), load3(view, i), d);
    store3(out, i, o.rgb);
    store3(spec, i, o.spec);
}

template <int NC>
__global__ void __launch_bounds__(kShadeBlock) shader_bwd_kernel(const float *__restrict__ up,
                                                                 const float *__restrict__ normal,
                                                                 const float *__restrict__ view,
                                                                 const float *__restrict__ data,
                                                                 const float *__restrict__ spec,
                                                                 float *__restrict__ gnormal,
                                                                 float *__restrict__ gdata,
                                                                 int64_t n) {
    constexpr int C = 3 * NC + 5;
    constexpr int P = C % 2 == 1 ? C : C + 1;
    __shared__ float rows[kShadeBlock * P];
    const int64_t row0 = (int64_t)blockIdx.x * kShadeBlock;
    const int nrows = (int)min((int64_t)kShadeBlock, n - row0);
    stage_rows_in<C, P>(data, row0, nrows, rows);
    __syncthreads();
    const int t = threadIdx.x;
    const bool live = t < nrows;
    const int64_t i = row0 + t;
    float d[C], dd[C];
    if (live) {
#pragma unroll
        for (int k = 0; k < C; ++k) d[k] = rows[t * P + k];
        f3 gn = shade_vjp<NC>(load3(normal, i), load3(view, i), d, load3(spec, i), load3(up, i),
                              dd);
        store3(gnormal, i, gn);
    }
    __syncthreads();
    if (live) {
#pragma unroll
        for (int k = 0; k < C; ++k) rows[t * P + k] = dd[k];
    }
    __syncthreads();
    stage_rows_out<C, P>(gdata, row0, nrows, rows);
}

// -------------------------------------------------------- RayAggregate ----
// NeuS logistic compositing (render_torch.py:155-167).
__device__ __forceinline__ bool ray_ok(int4 r, int64_t sdf_total, int64_t render_total) {
    uint32_t sc = (uint32_t)r.x, so = (uint32_t)r.y, rc = (uint32_t)r.z, ro = (uint32_t)r.w;
    if (rc == 0) return true;
    return (int64_t)so + rc + 1 <= sdf_total && (int64_t)ro + rc <= render_total && sc >= rc;
}

// One wavefront per ray, lanes over 64 consecutive samples at a time (the
// reference walks each ray serially in one thread, kernel/render.cu:223-253).
// T_i is the carried transmittance times an exclusive product scan of
// (1 - alpha) over the lanes; the colour is each lane's sum, reduced over
// the wave in a fixed order at the end.
__global__ void __launch_bounds__(256) aggregate_fwd_kernel(
    float *__restrict__ out, const float *__restrict__ rgb, const float *__restrict__ sdf,
    const int32_t *__restrict__ ray, float *__restrict__ alpha, float *__restrict__ T_out,
    int64_t rays, int64_t sdf_total, int64_t render_total, float s) {
    const int lane = threadIdx.x % kWave;
    const int64_t id = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (id >= rays) return;  // whole wavefront
    const int4 r = reinterpret_cast<const int4 *>(ray)[id];
    if (!ray_ok(r, sdf_total, render_total)) {
        const float q = __builtin_nanf("");
        if (lane == 0) store3(out, id, f3{q, q, q});
        return;
    }
    const uint32_t sc = (uint32_t)r.x, so = (uint32_t)r.y, rc = (uint32_t)r.z, ro = (uint32_t)r.w;
    f3 col{0.0f, 0.0f, 0.0f};
    if (sc != 0 && rc != 0) {
        float T = 1.0f;
        for (uint32_t c0 = 0; c0 < rc; c0 += kWave) {
            const uint32_t i = c0 + lane;
            const bool valid = i < rc;
            // every lane requests its sdf values and colour at once, at positions
            // clamped to the ray (a lane past the end computes values no valid
            // lane reads): lane-masked requests are each waited for in turn
            // (sdf_j, sdf_{j+1}) as one 8-B request at j = min(i, rc - 1): a lane
            // at i = rc takes sdf_rc from its second half
            typedef float f2a __attribute__((ext_vector_type(2), aligned(4)));
            const f2a sp =
                *reinterpret_cast<const f2a *>(sdf + (int64_t)so + (valid ? i : rc - 1));
            const f3 c = load3(rgb, (int64_t)ro + (valid ? i : rc - 1));
            const float e0 = i <= rc ? expf(-s * (valid ? sp.x : sp.y)) : 0.0f;
            float e1 = __shfl_down(e0, 1, kWave);
            if (lane == kWave - 1 && i + 1 <= rc) e1 = expf(-s * sp.y);
            const float a = valid ? fminf(fmaxf((e1 - e0) / (1.0f + e1), 0.0f), 10.0f) : 0.0f;
            float p = 1.0f - a;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const float u = __shfl_up(p, o, kWave);
                if (lane >= o) p *= u;
            }
            const float up = __shfl_up(p, 1, kWave);
            const float Ti = T * (lane == 0 ? 1.0f : up);
            if (valid) {
                alpha[(int64_t)ro + i] = a;
                T_out[(int64_t)ro + i] = Ti;
                col.x += c.x * a * Ti;
                col.y += c.y * a * Ti;
                col.z += c.z * a * Ti;
            }
            T *= __shfl(p, kWave - 1, kWave);
        }
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        col.x += __shfl_xor(col.x, o, kWave);
        col.y += __shfl_xor(col.y, o, kWave);
        col.z += __shfl_xor(col.z, o, kWave);
    }
    if (lane == 0) store3(out, id, col);
}

// Adjoint with the saved alpha and T.  dL/dalpha_k = T_k (g.rgb_k - rho_k)
// where rho is the suffix colour seen after sample k, rho_{k-1} =
// alpha_k (g.rgb_k) + (1 - alpha_k) rho_k -- no division by (1 - alpha).
// One wavefront per ray, chunks of 64 samples from the back: rho over a
// chunk is a reverse scan of the affine maps x -> (1 - alpha_k) x +
// alpha_k g.rgb_k applied to the carried rho of the later chunks.  The
// clamp passes the gradient wherever the pre-clamp value is in [0, 10]
// (torch.clamp semantics).
__global__ void __launch_bounds__(256) aggregate_bwd_kernel(
    const float *__restrict__ up, const float *__restrict__ rgb, const float *__restrict__ sdf,
    const int32_t *__restrict__ ray, const float *__restrict__ alpha,
    const float *__restrict__ T_in, float *__restrict__ grgb, float *__restrict__ gsdf,
    int64_t rays, int64_t sdf_total, int64_t render_total, float s) {
    const int lane = threadIdx.x % kWave;
    const int64_t id = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (id >= rays) return;  // whole wavefront
    const int4 r = reinterpret_cast<const int4 *>(ray)[id];
    const uint32_t sc = (uint32_t)r.x, so = (uint32_t)r.y, rc = (uint32_t)r.z, ro = (uint32_t)r.w;
    if (sc == 0 || rc == 0 || !ray_ok(r, sdf_total, render_total)) return;
    const f3 g = load3(up, id);
    float rho_after = 0.0f;  // rho at the last sample of the chunk
    float ds0_next = 0.0f;   // d alpha / d sdf of the first sample of the later chunk
    for (int64_t c0 = ((int64_t)rc - 1) / kWave * kWave; c0 >= 0; c0 -= kWave) {
        const int64_t k = c0 + lane;
        const bool valid = k < (int64_t)rc;
        float a = 0.0f, T = 0.0f, gc = 0.0f, e0 = 0.0f, e1 = 0.0f;
        if (valid) {
            a = alpha[(int64_t)ro + k];
            T = T_in[(int64_t)ro + k];
            const f3 c = load3(rgb, (int64_t)ro + k);
            gc = g.x * c.x + g.y * c.y + g.z * c.z;
            e0 = expf(-s * sdf[(int64_t)so + k]);
            e1 = expf(-s * sdf[(int64_t)so + k + 1]);
        }
        // inclusive suffix composition of the maps of lanes l..63
        float A = valid ? 1.0f - a : 1.0f, B = valid ? a * gc : 0.0f;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const float A2 = __shfl_down(A, o, kWave), B2 = __shfl_down(B, o, kWave);
            if (lane + o < kWave) {
                B = A * B2 + B;
                A = A * A2;
            }
        }
        const float An = __shfl_down(A, 1, kWave), Bn = __shfl_down(B, 1, kWave);
        const float rho = lane == kWave - 1 ? rho_after : An * rho_after + Bn;
        const float da = T * (gc - rho);
        const float inv = 1.0f / (1.0f + e1);
        const float pre = (e1 - e0) * inv;
        const float pass = (valid && pre >= 0.0f && pre <= 10.0f) ? da : 0.0f;
        // d alpha_k / d sdf_k and d alpha_k / d sdf_{k+1}; where the clamp
        // passes, e0 <= e1 < inf, so every factor below is at most 1 (the
        // product e1 (1 + e0) alone would overflow), and a clamped or
        // non-finite alpha contributes exactly 0
        const float d_s0 = pass != 0.0f ? pass * s * e0 * inv : 0.0f;
        const float d_s1 = pass != 0.0f ? -pass * s * (e1 * inv) * ((1.0f + e0) * inv) : 0.0f;
        float nxt = __shfl_down(d_s0, 1, kWave);
        if (lane == kWave - 1) nxt = ds0_next;
        if (valid) {
            gsdf[(int64_t)so + k + 1] = nxt + d_s1;
            const float w = a * T;
            store3(grgb, (int64_t)ro + k, f3{g.x * w, g.y * w, g.z * w});
        }
        ds0_next = __shfl(d_s0, 0, kWave);
        rho_after = __shfl(A * rho_after + B, 0, kWave);
    }
    if (lane == 0) gsdf[so] = ds0_next;
}

}  // namespace qsdvr

using namespace qsdvr;

extern "C" {

int qsdvr_grid_interpolation_forward(const float *data_grid, const float *point,
                                     const int32_t *index, float *out, int64_t n,
                                     int32_t vertex_reso, int32_t channels, qsdvr_stream stream) {
    QSDVR_REQUIRE(n >= 0 && vertex_reso >= 2 && channels > 0 && channels <= 1024,
                  "qsdvr_grid_interpolation_forward: bad sizes");
    if (n == 0) return QSDVR_OK;
    QSDVR_REQUIRE(data_grid && point && index && out,
                  "qsdvr_grid_interpolation_forward: null pointer");
    int R = vertex_reso - 1;
    int64_t threads = n * channels;
    hipStream_t s = as_stream(stream);
    const bool aligned16 = ((uintptr_t)data_grid % 16 == 0) && ((uintptr_t)out % 16 == 0);
    if (channels == 32 && aligned16) {
        constexpr int kIter = 2;  // samples per wave: 8 kIter
        const int64_t waves = (n + 8 * kIter - 1) / (8 * kIter);
        interp_fwd32_kernel<kIter><<<grid_1d(waves * kWave, 256), 256, 0, s>>>(
            reinterpret_cast<const float4 *>(data_grid), point, index, reinterpret_cast<float4 *>(out),
            n, R);
    } else if (channels == 32)
        interp_fwd_kernel<32><<<grid_1d(threads, 256), 256, 0, s>>>(data_grid, point, index, out,
                                                                    n, R, 32);
    else
        interp_fwd_kernel<0><<<grid_1d(threads, 256), 256, 0, s>>>(data_grid, point, index, out, n,
                                                                   R, channels);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_grid_interpolation_backward(float *grad_grid, const float *point, const int32_t *index,
                                      const float *grad_out, int64_t n, int32_t vertex_reso,
                                      int32_t channels, qsdvr_stream stream) {
    QSDVR_REQUIRE(n >= 0 && vertex_reso >= 2 && channels > 0 && channels <= 1024,
                  "qsdvr_grid_interpolation_backward: bad sizes");
    if (n == 0) return QSDVR_OK;
    QSDVR_REQUIRE(grad_grid && point && index && grad_out,
                  "qsdvr_grid_interpolation_backward: null pointer");
    int R = vertex_reso - 1;
    int64_t threads = n * channels;
    hipStream_t s = as_stream(stream);
    if (channels == 32)
        interp_bwd_kernel<32><<<grid_1d(threads, 256), 256, 0, s>>>(grad_grid, point, index,
                                                                    grad_out, n, R, 32);
    else
        interp_bwd_kernel<0><<<grid_1d(threads, 256), 256, 0, s>>>(grad_grid, point, index,
                                                                   grad_out, n, R, channels);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_shader_forward(float *out, const float *normal, const float *view_dir, const float *data,
                         float *specular, int64_t n, int32_t channels, qsdvr_stream stream) {
    QSDVR_REQUIRE(n >= 0, "qsdvr_shader_forward: negative n");
    QSDVR_REQUIRE(channels == 32 || channels == 53,
                  "qsdvr_shader_forward: channels must be 32 (SH degree <= 2) or 53 (degree 3)");
    if (n == 0) return QSDVR_OK;
    QSDVR_REQUIRE(out && normal && view_dir && data && specular,
                  "qsdvr_shader_forward: null pointer");
    hipStream_t s = as_stream(stream);
    if (channels == 32)
        shader_fwd_kernel<9><<<grid_1d(n, kShadeBlock), kShadeBlock, 0, s>>>(
            normal, view_dir, data, out, specular, n);
    else
        shader_fwd_kernel<16><<<grid_1d(n, kShadeBlock), kShadeBlock, 0, s>>>(
            normal, view_dir, data, out, specular, n);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_shader_backward(const float *grad_out, const float *normal, const float *view_dir,
                          const float *data, const float *specular, float *normal_grad,
                          float *data_grad, int64_t n, int32_t channels, qsdvr_stream stream) {
    QSDVR_REQUIRE(n >= 0, "qsdvr_shader_backward: negative n");
    QSDVR_REQUIRE(channels == 32 || channels == 53,
                  "qsdvr_shader_backward: channels must be 32 or 53");
    if (n == 0) return QSDVR_OK;
    QSDVR_REQUIRE(grad_out && normal && view_dir && data && specular && normal_grad && data_grad,
                  "qsdvr_shader_backward: null pointer");
    hipStream_t s = as_stream(stream);
    if (channels == 32)
        shader_bwd_kernel<9><<<grid_1d(n, kShadeBlock), kShadeBlock, 0, s>>>(
            grad_out, normal, view_dir, data, specular, normal_grad, data_grad, n);
    else
        shader_bwd_kernel<16><<<grid_1d(n, kShadeBlock), kShadeBlock, 0, s>>>(
            grad_out, normal, view_dir, data, specular, normal_grad, data_grad, n);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_ray_aggregate_forward(float *out, const float *rgb, const float *sdf,
                                const int32_t *ray_list, float *alpha, float *T,
                                int64_t ray_count, int64_t sdf_total, int64_t render_total,
                                float logistic_coef, qsdvr_stream stream) {
    QSDVR_REQUIRE(ray_count >= 0 && sdf_total >= 0 && render_total >= 0,
                  "qsdvr_ray_aggregate_forward: negative size");
    if (ray_count == 0) return QSDVR_OK;
    QSDVR_REQUIRE(out && ray_list, "qsdvr_ray_aggregate_forward: null pointer");
    QSDVR_REQUIRE(render_total == 0 || (rgb && sdf && alpha && T),
                  "qsdvr_ray_aggregate_forward: null sample list");
    aggregate_fwd_kernel<<<grid_1d(ray_count, 4), 256, 0, as_stream(stream)>>>(
        out, rgb, sdf, ray_list, alpha, T, ray_count, sdf_total, render_total, logistic_coef);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_ray_aggregate_backward(const float *grad_out, const float *rgb, const float *sdf,
                                 const int32_t *ray_list, const float *alpha, const float *T,
                                 float *rgb_grad, float *sdf_grad, int64_t ray_count,
                                 int64_t sdf_total, int64_t render_total, float logistic_coef,
                                 qsdvr_stream stream) {
    QSDVR_REQUIRE(ray_count >= 0 && sdf_total >= 0 && render_total >= 0,
                  "qsdvr_ray_aggregate_backward: negative size");
    if (ray_count == 0) return QSDVR_OK;
    QSDVR_REQUIRE(grad_out && ray_list, "qsdvr_ray_aggregate_backward: null pointer");
    QSDVR_REQUIRE(render_total == 0 || (rgb && sdf && alpha && T && rgb_grad && sdf_grad),
                  "qsdvr_ray_aggregate_backward: null sample list");
    aggregate_bwd_kernel<<<grid_1d(ray_count, 4), 256, 0, as_stream(stream)>>>(
        grad_out, rgb, sdf, ray_list, alpha, T, rgb_grad, sdf_grad, ray_count, sdf_total,
        render_total, logistic_coef);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

}  // extern "C"
