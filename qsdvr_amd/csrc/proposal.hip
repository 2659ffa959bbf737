// Proposal network for gfx950: a per-point MLP PE(39) -> 64 -> 64 -> 64 -> 1
// with bfloat16 weights on the matrix cores (v_mfma_f32_16x16x32_bf16, fp32
// accumulation), its training kernels, and the inverse-CDF fine sampler.
// No counterpart in the reference.
//
// Layout trick: every layer is computed transposed, H^T(out x points) =
// W^T ... so the MFMA result of one layer is already in the B-operand layout
// the next layer needs.  The D tile of v_mfma_f32_16x16x32 gives lane
// (j + 16 g) the 4 outputs 4g..4g+3 of point j; the next layer's B operand
// wants 8 K values per lane.  Since a dot product does not care about the
// order of K, K slot s of lane group g in K block kb is defined as feature
// 16 (2 kb + s/4) + 4 g + (s mod 4): exactly the registers the lane already
// holds.  The weight fragments are packed once, host-side order-free, into
// that permuted order (qsdvr_proposal_pack).
#include "common.h"

#include <math.h>

#include <algorithm>

namespace qsdvr {

constexpr int kBins = 64;
constexpr int kHid = 64;
constexpr int kPE = 39;
constexpr int kParams = 64 * 39 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1;  // 10945
// offsets into the float32 master vector
constexpr int oW1 = 0, ob1 = oW1 + 64 * 39, oW2 = ob1 + 64, ob2 = oW2 + 64 * 64,
              oW3 = ob2 + 64, ob3 = oW3 + 64 * 64, ow4 = ob3 + 64, ob4 = ow4 + 64;

// packed inference buffer: A fragments [layer 3][mt 4][kb 2][lane 64][8 bf16],
// then fp32 b1, b2, b3, w4 (bf16-rounded), b4
constexpr int kFragBf16 = 3 * 4 * 2 * 64 * 8;
constexpr int kPackedFloats = 4 * 64 + 4;
constexpr size_t kPackedBytes = kFragBf16 * 2 + kPackedFloats * 4;

typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned short bf16_bits(float f) {
    return __builtin_bit_cast(unsigned short, (__bf16)f);
}

__device__ __forceinline__ float bf16_round(float f) { return (float)(__bf16)f; }

// canonical PE index of chain slot s (0..15) of axis a, or -1 (zero pad):
// slot 0 = coordinate, 1 + 2k = sin(2^k pi c), 2 + 2k = cos(2^k pi c)
__host__ __device__ __forceinline__ int pe_index(int axis, int s) {
    if (axis >= 3 || s >= 13) return -1;
    if (s == 0) return axis;
    const int k = (s - 1) >> 1;
    return (s & 1) ? 3 + 6 * k + axis : 3 + 6 * k + 3 + axis;
}

// input feature held in K slot s8 of lane group g, K block kb
__host__ __device__ __forceinline__ int hidden_index(int kb, int g, int s8) {
    return 16 * (2 * kb + (s8 >> 2)) + 4 * g + (s8 & 3);
}

__global__ void __launch_bounds__(256) pack_kernel(const float *__restrict__ theta,
                                                   unsigned short *__restrict__ frag,
                                                   float *__restrict__ tail) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id < kFragBf16) {
        const int s8 = id & 7, lane = (id >> 3) & 63, kb = (id >> 9) & 1, mt = (id >> 10) & 3,
                  layer = id >> 12;
        const int out = 16 * mt + (lane & 15), g = lane >> 4;
        float v = 0.0f;
        if (layer == 0) {
            const int f = pe_index(g, 8 * kb + s8);
            if (f >= 0) v = theta[oW1 + out * kPE + f];
        } else {
            const int f = hidden_index(kb, g, s8);
            v = theta[(layer == 1 ? oW2 : oW3) + out * kHid + f];
        }
        frag[id] = bf16_bits(v);
    } else if (id < kFragBf16 + kPackedFloats) {
        const int t = id - kFragBf16;
        float v;
        if (t < 64) v = theta[ob1 + t];
        else if (t < 128) v = theta[ob2 + t - 64];
        else if (t < 192) v = theta[ob3 + t - 128];
        else if (t < 256) v = bf16_round(theta[ow4 + t - 192]);
        else v = t == 256 ? theta[ob4] : 0.0f;
        tail[t] = v;
    }
}

typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef short v2s __attribute__((ext_vector_type(2)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

// two floats -> one dword of two bfloat16, a in the low half (round to
// nearest even, one v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned pk_bf16(float a, float b) {
    const v2f f = {a, b};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f, v2bf));
}

// relu of two packed bfloat16 (v_pk_max_i16 with 0): read as int16, a
// bfloat16 is negative exactly when its sign bit is set.  Rounding keeps
// the sign, so this equals rounding the float relu (up to the sign of 0).
__device__ __forceinline__ unsigned pk_relu_bf16(unsigned x) {
    const v2s z = {0, 0};
    return __builtin_bit_cast(unsigned, __builtin_elementwise_max(__builtin_bit_cast(v2s, x), z));
}

__device__ __forceinline__ v8s pack8(const float v[8]) {
    const v4u r = {pk_bf16(v[0], v[1]), pk_bf16(v[2], v[3]), pk_bf16(v[4], v[5]),
                   pk_bf16(v[6], v[7])};
    return __builtin_bit_cast(v8s, r);
}

__device__ __forceinline__ v4f mfma(v8s a, v8s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                   __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

// One hidden layer on two 16-point tiles: D[t][mt] = W^T[mt, :] x B[t] + bias.
// Each weight fragment is read from LDS once for both tiles, and the two
// tiles' MFMA chains are independent.
__device__ __forceinline__ void layer2(const v8s *__restrict__ A /*[4][2][64]*/, const v8s B[2][2],
                                       const float *__restrict__ bias, int lane, v4f D[2][4], int zofs) {
    const int g = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const v4f b = {bias[16 * mt + 4 * g + 0], bias[16 * mt + 4 * g + 1],
                       bias[16 * mt + 4 * g + 2], bias[16 * mt + 4 * g + 3]};
        v4f acc0 = b, acc1 = b;
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            const v8s w = A[(mt * 2 + kb) * 64 + lane + zofs];
            acc0 = mfma(w, B[0][kb], acc0);
            acc1 = mfma(w, B[1][kb], acc1);
        }
        D[0][mt] = acc0;
        D[1][mt] = acc1;
    }
}

// Positional-encoding chain of one axis at the coordinate o + d t (o, d: the
// axis's components of the ray origin and direction; lane group 3 passes 0,
// 0): [p, sin(pi p), cos(pi p), ..., sin(32 pi p), cos(32 pi p), 0, 0, 0].
// Slots 13-15 and the whole chain of lane group 3 meet zero weights in the
// packed fragments, so only their finiteness matters.
__device__ __forceinline__ void pe_chain(float o, float d, float t, float chain[16]) {
    const float c = o + d * t;
    float sn, cs;
    sincospif(c, &sn, &cs);
    chain[0] = c;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        chain[1 + 2 * k] = sn;
        chain[2 + 2 * k] = cs;
        const float s2 = 2.0f * sn * cs, c2 = 1.0f - 2.0f * sn * sn;
        sn = s2;
        cs = c2;
    }
    chain[13] = chain[14] = chain[15] = 0.0f;
}

// relu + bf16 of a layer output into the next layer's B operand (inference:
// the pair is rounded first and the relu applied to the packed halves)
__device__ __forceinline__ void to_operand_packed(const v4f D[4], v8s B[2]) {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        v4u r;
#pragma unroll
        for (int p = 0; p < 4; ++p)
            r[p] = pk_relu_bf16(pk_bf16(D[2 * kb + (p >> 1)][2 * (p & 1)],
                                        D[2 * kb + (p >> 1)][2 * (p & 1) + 1]));
        B[kb] = __builtin_bit_cast(v8s, r);
    }
}

// relu + bf16 of a layer output into the next layer's B operand, and the
// activations at their bf16 values (training)
__device__ __forceinline__ void to_operand(const v4f D[4], v8s B[2], float act[16]) {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        float v[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            float x = fmaxf(D[2 * kb + (s >> 2)][s & 3], 0.0f);
            v[s] = x;
            act[8 * kb + s] = bf16_round(x);
        }
        B[kb] = pack8(v);
    }
}

constexpr int kMlpBlock = 256;

// Logits of the 64 bin-centre points of every ray (one wavefront per ray,
// four 16-point tiles).  With x0..h3 non-null (training) the activations
// are also written in canonical [point][64] order.
template <bool TRAIN>
__global__ void __launch_bounds__(kMlpBlock) mlp_fwd_kernel(
    const float4 *__restrict__ ray_state, int64_t n_rays, float interval,
    const unsigned short *__restrict__ packed, float *__restrict__ logits,
    float *__restrict__ x0s, float *__restrict__ x1s, float *__restrict__ x2s,
    float *__restrict__ h3s) {
    __shared__ v8s sA[3 * 4 * 2 * 64];
    __shared__ float sT[kPackedFloats];
    {
        const v8s *src = reinterpret_cast<const v8s *>(packed);
        for (int i = threadIdx.x; i < 3 * 4 * 2 * 64; i += kMlpBlock) sA[i] = src[i];
        const float *tail = reinterpret_cast<const float *>(packed + kFragBf16);
        for (int i = threadIdx.x; i < kPackedFloats; i += kMlpBlock) sT[i] = tail[i];
    }
    __syncthreads();
    const int lane = threadIdx.x % kWave, g = lane >> 4, j = lane & 15;
    const int64_t wave0 = (int64_t)blockIdx.x * (kMlpBlock / kWave) + threadIdx.x / kWave;
    const int64_t nwaves = (int64_t)gridDim.x * (kMlpBlock / kWave);
    constexpr bool train = TRAIN;
    for (int64_t r = wave0; r < n_rays; r += nwaves) {
        const float4 sa = ray_state[2 * r], sb = ray_state[2 * r + 1];
        const uint32_t count = __float_as_uint(sa.w);
        if (!train && count < 2) {  // no segment: logits unused, zero them
            logits[r * kBins + lane] = 0.0f;
            continue;
        }
        const float L = count >= 2 ? (float)(count - 1) * interval : 0.0f;
        // this lane group's axis of the ray (group 3: none)
        const float po = g == 0 ? sa.x : (g == 1 ? sa.y : (g == 2 ? sa.z : 0.0f));
        const float pd = g == 0 ? sb.x : (g == 1 ? sb.y : (g == 2 ? sb.z : 0.0f));
#pragma unroll 1
        for (int tp = 0; tp < kBins / 32; ++tp) {
            // two 16-point tiles: bins 32 tp + j and 32 tp + 16 + j
            v8s B[2][2];
            int64_t pt[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int bin = 32 * tp + 16 * u + j;
                const float t = ((float)bin + 0.5f) * (1.0f / kBins) * L;
                float chain[16];
                pe_chain(po, pd, t, chain);
                B[u][0] = pack8(chain);
                B[u][1] = pack8(chain + 8);
                pt[u] = r * kBins + bin;
                if (train) {
#pragma unroll
                    for (int s = 0; s < 13; ++s) {
                        const int f = pe_index(g, s);
                        if (f >= 0) x0s[pt[u] * 64 + f] = bf16_round(chain[s]);
                    }
                }
            }
            v4f D[2][4];
            float act[16];
            // opaque zero offset: keeps the weight-fragment reads inside the
            // loop rather than hoisted into ~100 registers for the kernel
            int zofs;
            __asm__ volatile("v_mov_b32 %0, 0" : "=v"(zofs));
            layer2(sA, B, sT, lane, D, zofs);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (train) {
                    to_operand(D[u], B[u], act);
#pragma unroll
                    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                        for (int s = 0; s < 8; ++s)
                            x1s[pt[u] * 64 + hidden_index(kb, g, s)] = act[8 * kb + s];
                } else {
                    to_operand_packed(D[u], B[u]);
                }
            }
            layer2(sA + 4 * 2 * 64, B, sT + 64, lane, D, zofs);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (train) {
                    to_operand(D[u], B[u], act);
#pragma unroll
                    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                        for (int s = 0; s < 8; ++s)
                            x2s[pt[u] * 64 + hidden_index(kb, g, s)] = act[8 * kb + s];
                } else {
                    to_operand_packed(D[u], B[u]);
                }
            }
            layer2(sA + 2 * 4 * 2 * 64, B, sT + 128, lane, D, zofs);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                float z = 0.0f;
#pragma unroll
                for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int f = 16 * mt + 4 * g + q;
                        // inference: relu as med3(x, 0, FLT_MAX), one instruction
                        // (fmaxf adds a NaN-quieting one in IEEE mode); equal
                        // to relu for every finite x
                        const float h = train ? fmaxf(D[u][mt][q], 0.0f)
                                              : __builtin_amdgcn_fmed3f(D[u][mt][q], 0.0f,
                                                                        3.40282347e38f);
                        if (train) h3s[pt[u] * 64 + f] = h;
                        z += sT[192 + f] * h;
                    }
                z += __shfl_xor(z, 16, kWave);
                z += __shfl_xor(z, 32, kWave);
                if (g == 0) logits[pt[u]] = z + sT[256];
            }
        }
    }
}

// Backward, one lane per point and one wavefront per ray: the composited
// log-likelihood against the bin targets, then dpre3 -> dpre2 -> dpre1 in
// float32 (weights at their bfloat16 values, read from LDS).
__global__ void __launch_bounds__(256) mlp_bwd_point_kernel(
    const float *__restrict__ logits, const float *__restrict__ target, int64_t n_rays,
    const float *__restrict__ theta, const float *__restrict__ x1s,
    const float *__restrict__ x2s, const float *__restrict__ h3s, float *__restrict__ dp1,
    float *__restrict__ dp2, float *__restrict__ dp3, float *__restrict__ dz_out,
    float *__restrict__ loss) {
    __shared__ float W2[64 * 64], W3[64 * 64], w4[64];
    for (int i = threadIdx.x; i < 64 * 64; i += 256) {
        W2[i] = bf16_round(theta[oW2 + i]);
        W3[i] = bf16_round(theta[oW3 + i]);
    }
    if (threadIdx.x < 64) w4[threadIdx.x] = bf16_round(theta[ow4 + threadIdx.x]);
    __syncthreads();
    const int64_t pt = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t ray = pt / kBins;
    if (ray >= n_rays) return;  // whole wavefronts (one ray = one wave)
    // density reading of the logits, composited over the ray's 64 bins
    // (lane = bin): d = softplus(z), a = 1 - e^-d, w = e^{-sum_{c<b} d_c} a;
    // loss = -sum q log w - (1 - sum q) sum d;
    // dL/dz_b = sigmoid(z_b) [(1 - sum_{c<=b} q_c) - q_b / expm1(d_b)]
    const float z = logits[pt], q = target[pt];
    const int lane = threadIdx.x & (kWave - 1);
    const float d = z > 15.0f ? z : log1pf(expf(z));
    float Sd = d, Cq = q;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const float a = __shfl_up(Sd, o, kWave), b = __shfl_up(Cq, o, kWave);
        if (lane >= o) {
            Sd += a;
            Cq += b;
        }
    }
    const float Stot = __shfl(Sd, kWave - 1, kWave), Qtot = __shfl(Cq, kWave - 1, kWave);
    float l = 0.0f;
    if (q > 0.0f) l = -q * (-(Sd - d) + logf(fmaxf(-expm1f(-d), 1e-30f)));
    if (lane == kWave - 1) l += (1.0f - Qtot) * Stot;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, kWave);
    if (lane == 0 && loss) loss[ray] = l;
    const float sig = 1.0f / (1.0f + expf(-z));
    const float dz = sig * ((1.0f - Cq) - (q > 0.0f ? q / fmaxf(expm1f(d), 1e-30f) : 0.0f));
    dz_out[pt] = dz;
    // d2[i] = [x2 > 0] sum_o W3[o][i] dpre3[o];  d1 likewise through W2.
    // Loops run over o at runtime and over i fully unrolled, so the register
    // arrays are only ever indexed with compile-time indices.
    const float *h3 = h3s + pt * 64;
    float d2[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) d2[i] = 0.0f;
    for (int o = 0; o < 64; ++o) {
        const float d = h3[o] > 0.0f ? dz * w4[o] : 0.0f;
        dp3[pt * 64 + o] = d;
#pragma unroll
        for (int i = 0; i < 64; ++i) d2[i] += W3[o * 64 + i] * d;
    }
    const float *x2 = x2s + pt * 64;
#pragma unroll
    for (int i = 0; i < 64; ++i) dp2[pt * 64 + i] = x2[i] > 0.0f ? d2[i] : 0.0f;
    float d1[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) d1[i] = 0.0f;
    for (int o = 0; o < 64; ++o) {
        const float d = dp2[pt * 64 + o];
#pragma unroll
        for (int i = 0; i < 64; ++i) d1[i] += W2[o * 64 + i] * d;
    }
    const float *x1 = x1s + pt * 64;
#pragma unroll
    for (int i = 0; i < 64; ++i) dp1[pt * 64 + i] = x1[i] > 0.0f ? d1[i] : 0.0f;
}

// Partial weight gradients of one layer over a chunk of points:
// part[o * 64 + i] = sum_pt dpre[pt][o] x[pt][i], part[4096 + o] = sum dpre,
// as a (WIDTH x chunk) x (chunk x 64) product on the matrix cores in float32
// (v_mfma_f32_16x16x4_f32, one wavefront per chunk, K = 4 points per step).
// Lane t supplies A[t % 16][t / 16] = dpre[pt0 + t / 16][16 mt + t % 16] and
// B[t / 16][t % 16] = x[pt0 + t / 16][16 nt + t % 16] and holds
// D[4 (t / 16) + i][t % 16].  WIDTH = 64 (hidden layers) or 1 (the output
// layer: dpre = dz [npts], one 16-row tile whose rows 1..15 are zero).
constexpr int kWgradChunk = 1024;

template <int WIDTH>
__global__ void __launch_bounds__(kWave) mlp_wgrad_kernel(const float *__restrict__ dpre,
                                                          const float *__restrict__ x,
                                                          int64_t npts, float *__restrict__ part) {
    constexpr int MT = WIDTH == 64 ? 4 : 1;
    const int t = threadIdx.x, r = t >> 4, m = t & 15;
    const int64_t c0 = (int64_t)blockIdx.x * kWgradChunk;
    const int64_t c1 = min(c0 + kWgradChunk, npts);
    v4f acc[MT][4];
    float bsum[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        bsum[mt] = 0.0f;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }
    for (int64_t p0 = c0; p0 < c1; p0 += 4) {
        const int64_t pt = p0 + r;
        const bool ok = pt < c1;
        float av[MT], bv[4];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            if (WIDTH == 64) av[mt] = ok ? dpre[pt * 64 + 16 * mt + m] : 0.0f;
            else av[mt] = ok && m == 0 ? dpre[pt] : 0.0f;
            bsum[mt] += av[mt];
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) bv[nt] = ok ? x[pt * 64 + 16 * nt + m] : 0.0f;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[mt], bv[nt], acc[mt][nt], 0,
                                                                   0, 0);
    }
    float *dst = part + (int64_t)blockIdx.x * (64 * 64 + 64);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int o = 16 * mt + 4 * r + i;
                if (o < WIDTH) dst[o * 64 + 16 * nt + m] = acc[mt][nt][i];
            }
        // bias: the column sums, gathered over the four point rows of a step
        float v = bsum[mt];
        v += __shfl_xor(v, 16, kWave);
        v += __shfl_xor(v, 32, kWave);
        if (t < 16 && 16 * mt + t < WIDTH) dst[4096 + 16 * mt + t] = v;
    }
}

// Sum the per-chunk partials into the float32 gradient vector (fixed order).
__global__ void __launch_bounds__(256) mlp_reduce_kernel(const float *__restrict__ p1,
                                                         const float *__restrict__ p2,
                                                         const float *__restrict__ p3,
                                                         const float *__restrict__ p4,
                                                         int nchunks, float *__restrict__ grad) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= kParams) return;
    const float *src;
    int off;
    if (id < ob1) { src = p1; off = (id / kPE) * 64 + id % kPE; }
    else if (id < oW2) { src = p1; off = 4096 + (id - ob1); }
    else if (id < ob2) { src = p2; off = id - oW2; }
    else if (id < oW3) { src = p2; off = 4096 + (id - ob2); }
    else if (id < ob3) { src = p3; off = id - oW3; }
    else if (id < ow4) { src = p3; off = 4096 + (id - ob3); }
    else if (id < ob4) { src = p4; off = id - ow4; }
    else { src = p4; off = 4096; }
    float s = 0.0f;
    for (int c = 0; c < nchunks; ++c) s += src[(int64_t)c * (64 * 64 + 64) + off];
    grad[id] = s;
}

__device__ __forceinline__ void adam1(float &p, float g, float &m, float &v, float lr, float b1,
                                      float b2, float eps, float c1, float c2) {
    const float mm = b1 * m + (1.0f - b1) * g;
    const float vv = b2 * v + (1.0f - b2) * g * g;
    m = mm;
    v = vv;
    p -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

// Bias-corrected Adam over n parameters (float4 body when `vec`, scalar
// tail); grid-stride, 64-bit indices (appearance grids exceed 2^31 floats).
__global__ void __launch_bounds__(256) adam_kernel(float *__restrict__ theta,
                                                   const float *__restrict__ grad,
                                                   float *__restrict__ m, float *__restrict__ v,
                                                   int64_t n, int vec, float lr, float b1, float b2,
                                                   float eps, float c1, float c2) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n4 = vec ? n / 4 : 0;
    for (int64_t i = tid; i < n4; i += stride) {
        float4 p = reinterpret_cast<float4 *>(theta)[i];
        const float4 g = reinterpret_cast<const float4 *>(grad)[i];
        float4 mm = reinterpret_cast<float4 *>(m)[i];
        float4 vv = reinterpret_cast<float4 *>(v)[i];
        adam1(p.x, g.x, mm.x, vv.x, lr, b1, b2, eps, c1, c2);
        adam1(p.y, g.y, mm.y, vv.y, lr, b1, b2, eps, c1, c2);
        adam1(p.z, g.z, mm.z, vv.z, lr, b1, b2, eps, c1, c2);
        adam1(p.w, g.w, mm.w, vv.w, lr, b1, b2, eps, c1, c2);
        reinterpret_cast<float4 *>(theta)[i] = p;
        reinterpret_cast<float4 *>(m)[i] = mm;
        reinterpret_cast<float4 *>(v)[i] = vv;
    }
    for (int64_t i = 4 * n4 + tid; i < n; i += stride)
        adam1(theta[i], grad[i], m[i], v[i], lr, b1, b2, eps, c1, c2);
}

// Inverse-CDF fine samples, one wavefront per ray (lane = bin): t = 0, t = L
// and K-2 interior samples at u = (k + 0.5)/(K-2) of the mixture pdf
// (1-eps) w^g/sum(w^g) + eps/64, w the composited proposal weights of the
// bins and g the sharpening exponent.
__device__ __forceinline__ float wave_scan_incl(float v, int lane) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const float u = __shfl_up(v, o, kWave);
        if (lane >= o) v += u;
    }
    return v;
}

__global__ void __launch_bounds__(256) sampler_kernel(const float *__restrict__ logits,
                                                      const float4 *__restrict__ ray_state,
                                                      int64_t n_rays, float interval, int K,
                                                      float eps, float gamma,
                                                      float *__restrict__ tlist) {
    __shared__ float s_lo[4][kBins], s_pdf[4][kBins];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int64_t r = (int64_t)blockIdx.x * 4 + wv;
    if (r >= n_rays) return;  // whole wavefront
    // the ray's count and its logits requested together (one round trip): a
    // ray with no segment (count < 2) runs the same code on its unused logits
    // and writes all its samples at the origin
    const uint32_t count = __float_as_uint(ray_state[2 * r].w);
    const float z = logits[r * kBins + lane];
    float *out = tlist + r * K;
    const bool empty = count < 2;
    const float L = empty ? 0.0f : (float)(count - 1) * interval;
    const float d = z > 15.0f ? z : log1pf(expf(z));
    const float S = wave_scan_incl(d, lane);
    const float w1 = expf(-(S - d)) * -expm1f(-d);
    const float wb = gamma == 1.0f ? w1 : (gamma == 2.0f ? w1 * w1 : powf(w1, gamma));
    float tot = wb;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, kWave);
    const float pdf = tot > 1e-30f ? (1.0f - eps) * wb / tot + eps / kBins : 1.0f / kBins;
    const float hi = wave_scan_incl(pdf, lane);
    s_lo[wv][lane] = hi - pdf;
    s_pdf[wv][lane] = pdf;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int nm = K - 2;
    for (int k = lane; k < nm; k += kWave) {
        const float u = ((float)k + 0.5f) / (float)nm;
        int lo = 0;  // last bin with cdf_lo <= u
#pragma unroll
        for (int step = kBins / 2; step > 0; step >>= 1)
            if (lo + step < kBins && s_lo[wv][lo + step] <= u) lo += step;
        const float frac = fminf(fmaxf((u - s_lo[wv][lo]) / s_pdf[wv][lo], 0.0f), 1.0f);
        out[1 + k] = empty ? 0.0f : ((float)lo + frac) * (1.0f / kBins) * L;
    }
    if (lane == 0) {
        out[0] = 0.0f;
        out[K - 1] = L;
    }
}

}  // namespace qsdvr

using namespace qsdvr;

extern "C" {

int32_t qsdvr_proposal_param_count(void) { return kParams; }

size_t qsdvr_proposal_packed_bytes(void) { return kPackedBytes; }

int qsdvr_proposal_pack(const float *theta, void *packed, qsdvr_stream stream) {
    QSDVR_REQUIRE(theta && packed, "qsdvr_proposal_pack: null pointer argument");
    unsigned short *frag = (unsigned short *)packed;
    float *tail = (float *)(frag + kFragBf16);
    pack_kernel<<<grid_1d(kFragBf16 + kPackedFloats, 256), 256, 0, as_stream(stream)>>>(theta, frag,
                                                                                     tail);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_proposal_forward(const float *ray_state, int64_t n_rays, float interval,
                           const void *packed, float *logits, float *x0, float *x1, float *x2,
                           float *h3, qsdvr_stream stream) {
    QSDVR_REQUIRE(n_rays >= 0 && interval > 0.0f, "qsdvr_proposal_forward: bad sizes");
    if (n_rays == 0) return QSDVR_OK;
    QSDVR_REQUIRE(ray_state && packed && logits, "qsdvr_proposal_forward: null pointer argument");
    const bool train = x0 || x1 || x2 || h3;
    QSDVR_REQUIRE(!train || (x0 && x1 && x2 && h3),
                  "qsdvr_proposal_forward: training needs all four activation buffers");
    int64_t waves = n_rays;
    unsigned blocks = (unsigned)std::min<int64_t>((waves + 3) / 4, 256 * 32);
    if (train)
        mlp_fwd_kernel<true><<<blocks, kMlpBlock, 0, as_stream(stream)>>>(
            reinterpret_cast<const float4 *>(ray_state), n_rays, interval,
            (const unsigned short *)packed, logits, x0, x1, x2, h3);
    else
        mlp_fwd_kernel<false><<<blocks, kMlpBlock, 0, as_stream(stream)>>>(
            reinterpret_cast<const float4 *>(ray_state), n_rays, interval,
            (const unsigned short *)packed, logits, x0, x1, x2, h3);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

size_t qsdvr_proposal_workspace_size(int64_t n_rays) {
    if (n_rays < 0) return 0;
    const int64_t npts = n_rays * kBins;
    const int64_t nch = (npts + kWgradChunk - 1) / kWgradChunk;
    return (size_t)(3 * npts * 64 + npts) * 4 + (size_t)(4 * nch * (64 * 64 + 64)) * 4 + 8 * 256;
}

int qsdvr_proposal_backward(const float *logits, const float *target, int64_t n_rays,
                            const float *theta, const float *x0, const float *x1, const float *x2,
                            const float *h3, void *workspace, size_t workspace_bytes, float *grad,
                            float *loss, qsdvr_stream stream) {
    QSDVR_REQUIRE(n_rays > 0, "qsdvr_proposal_backward: n_rays must be positive");
    QSDVR_REQUIRE(logits && target && theta && x0 && x1 && x2 && h3 && grad,
                  "qsdvr_proposal_backward: null pointer argument");
    const int64_t npts = n_rays * kBins;
    const int64_t nch = (npts + kWgradChunk - 1) / kWgradChunk;
    Carver cv{(char *)workspace, workspace_bytes};
    float *dp1 = cv.take<float>(npts * 64), *dp2 = cv.take<float>(npts * 64);
    float *dp3 = cv.take<float>(npts * 64), *dz = cv.take<float>(npts);
    float *pt1 = cv.take<float>(nch * 4160), *pt2 = cv.take<float>(nch * 4160);
    float *pt3 = cv.take<float>(nch * 4160), *pt4 = cv.take<float>(nch * 4160);
    if (!dp1 || !dp2 || !dp3 || !dz || !pt1 || !pt2 || !pt3 || !pt4) {
        set_error("qsdvr_proposal_backward: workspace too small (need %zu bytes)",
                  qsdvr_proposal_workspace_size(n_rays));
        return QSDVR_EWORKSPACE;
    }
    hipStream_t s = as_stream(stream);
    mlp_bwd_point_kernel<<<(unsigned)((npts + 255) / 256), 256, 0, s>>>(
        logits, target, n_rays, theta, x1, x2, h3, dp1, dp2, dp3, dz, loss);
    QSDVR_LAUNCHED();
    mlp_wgrad_kernel<64><<<(unsigned)nch, kWave, 0, s>>>(dp1, x0, npts, pt1);
    mlp_wgrad_kernel<64><<<(unsigned)nch, kWave, 0, s>>>(dp2, x1, npts, pt2);
    mlp_wgrad_kernel<64><<<(unsigned)nch, kWave, 0, s>>>(dp3, x2, npts, pt3);
    mlp_wgrad_kernel<1><<<(unsigned)nch, kWave, 0, s>>>(dz, h3, npts, pt4);
    QSDVR_LAUNCHED();
    mlp_reduce_kernel<<<grid_1d(kParams, 256), 256, 0, s>>>(pt1, pt2, pt3, pt4, (int)nch, grad);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_adam_step(float *theta, const float *grad, float *m, float *v, int64_t n, float lr,
                    float beta1, float beta2, float eps, int32_t step, qsdvr_stream stream) {
    QSDVR_REQUIRE(theta && grad && m && v && n > 0 && step >= 1,
                  "qsdvr_adam_step: bad arguments");
    const float c1 = (float)(1.0 / (1.0 - pow((double)beta1, (double)step)));
    const float c2 = (float)(1.0 / (1.0 - pow((double)beta2, (double)step)));
    const int vec = ((reinterpret_cast<uintptr_t>(theta) | reinterpret_cast<uintptr_t>(grad) |
                      reinterpret_cast<uintptr_t>(m) | reinterpret_cast<uintptr_t>(v)) & 15) == 0;
    const int64_t work = vec ? (n + 3) / 4 : n;
    const unsigned blocks = (unsigned)std::min<int64_t>((work + 255) / 256, 4096);
    adam_kernel<<<blocks, 256, 0, as_stream(stream)>>>(theta, grad, m, v, n, vec, lr, beta1,
                                                       beta2, eps, c1, c2);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_proposal_sample(const float *logits, const float *ray_state, int64_t n_rays,
                          float interval, int32_t K, float eps, float gamma, float *tlist,
                          qsdvr_stream stream) {
    QSDVR_REQUIRE(n_rays >= 0 && K >= 3 && K <= 4096 && eps >= 0.0f && eps <= 1.0f &&
                      gamma > 0.0f && gamma <= 16.0f,
                  "qsdvr_proposal_sample: bad sizes");
    if (n_rays == 0) return QSDVR_OK;
    QSDVR_REQUIRE(logits && ray_state && tlist, "qsdvr_proposal_sample: null pointer argument");
    sampler_kernel<<<(unsigned)((n_rays + 3) / 4), 256, 0, as_stream(stream)>>>(
        logits, reinterpret_cast<const float4 *>(ray_state), n_rays, interval, K, eps, gamma,
        tlist);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

}  // extern "C"
