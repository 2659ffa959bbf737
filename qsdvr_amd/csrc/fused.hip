// Fused plain renderer for gfx950: ray generation + SDF + normal +
// trilinear appearance + SH shading + NeuS compositing in one pass, and its
// recompute adjoint.  No reference counterpart as one kernel; it computes
// exactly the composition of kernel/camera.cu:12-96 (bit-exact sample
// positions and cell indices) with the math of models/render_torch.py:77-167.
//
// Work mapping: one lane per ray; a wavefront owns an 8x8 pixel tile, so
// the 64 rays of a wave march side by side through the same few cells and
// their vertex-row gathers hit the same cache lines.  Blocks are remapped so
// that each XCD (workgroups are dealt round-robin over the 8 XCDs) renders
// one contiguous slab of tiles and keeps its L2 footprint local.
//
// Adjoint: the forward keeps nothing per sample.  The backward walks every
// ray front to back again, rebuilding e, alpha, T and the prefix
// sum_{i<=k} w_i rgb_i; the suffix the alpha adjoint needs is
// saved_colour - prefix, and 1/(1 - alpha_k) = (1 + e_{k+1}) / (1 + e_k)
// exactly (no division by a vanishing 1 - alpha).  d alpha_j / d sdf_{j+1} =
// -s e_{j+1} (1 + e_j) / (1 + e_{j+1})^2 is evaluated as the product of
// e_{j+1} / (1 + e_{j+1}) and (1 + e_j) / (1 + e_{j+1}), both at most 1
// where alpha is not clamped: e_{j+1} (1 + e_j) alone overflows float32
// where the SDF is below about -1.1 / s (s = 40: non-metric SDF values of
// the gyroid shell reach that).
#include "det_scatter.h"
#include "quadric.h"
#include "raygen.h"
#include "shade.h"

#include <limits.h>

#include <algorithm>
#include <rocprim/device/device_radix_sort.hpp>

namespace qsdvr {

constexpr int kFusedBlock = 256;
constexpr int kWavesPerBlock = kFusedBlock / kWave;

struct RaySetup {
    int64_t out;       // output row of the ray, -1 = inactive lane
    int cam;
    f3 ori, dir;
    uint32_t renders;  // render samples (sdf samples = renders + 1), 0 = empty
    f3 center;         // camera centre (depth reference)
    const float *tl;   // explicit sample positions t[K] along the ray, or nullptr (uniform)
};

__device__ __forceinline__ int logical_block() {
    // XCD-aware remap: hardware deals block b to XCD b % 8.
    const unsigned nb = gridDim.x, b = blockIdx.x;
    if (nb % 8 != 0) return (int)b;
    return (int)((b % 8) * (nb / 8) + b / 8);
}

// kTileStride: SCATTER launches visit the tiles of a frame in the order
// t -> t * kTileStride mod tiles (a bijection: the stride is prime and
// divides no tile count used here, checked below), so that the heavy tiles
// of an object are spread over the launch instead of arriving together, and
// waves in flight at once touch distant parts of the grids.
constexpr int64_t kTileStride = 7919;

template <int WPB = kWavesPerBlock, bool SCATTER = false>
__device__ __forceinline__ RaySetup setup_ray(const qsdvr_camera *__restrict__ cams, int n_cams,
                                              uint32_t W, uint32_t H,
                                              const int32_t *__restrict__ ray_pixel,
                                              int64_t n_rays, float interval,
                                              const float4 *__restrict__ ray_state = nullptr,
                                              const float *__restrict__ tlist = nullptr,
                                              int K = 0,
                                              const uint32_t *__restrict__ order = nullptr) {
    RaySetup r;
    r.out = -1;
    r.renders = 0;
    r.tl = nullptr;
    const int lane = threadIdx.x % kWave;
    // with a tile order, launch block b takes tile order[b] (blocks are dealt
    // round-robin to the XCDs, so each XCD gets every 8th tile of the order)
    const int64_t wave = order != nullptr ? (int64_t)blockIdx.x * WPB + threadIdx.x / kWave
                                          : (int64_t)logical_block() * WPB + threadIdx.x / kWave;
    int cam = 0;
    uint32_t pix = 0;
    if (ray_pixel == nullptr) {
        const int64_t tx = (W + 7) / 8, ty = (H + 7) / 8, tiles = tx * ty;
        if (wave >= tiles * n_cams) return r;
        const int64_t g = order != nullptr ? (int64_t)order[wave] : wave;
        if (g >= tiles * n_cams) return r;
        cam = (int)(g / tiles);
        int64_t t = g - (int64_t)cam * tiles;
        if (SCATTER && order == nullptr && tiles % kTileStride != 0) t = (t * kTileStride) % tiles;
        uint32_t px = (uint32_t)(t % tx) * 8 + (lane % 8);
        uint32_t py = (uint32_t)(t / tx) * 8 + (lane / 8);
        if (px >= W || py >= H) return r;
        pix = py * W + px;
        r.out = (int64_t)cam * W * H + pix;
    } else {
        int64_t i = wave * kWave + lane;
        if (i >= n_rays) return r;
        int64_t code = (uint32_t)ray_pixel[i];
        cam = (int)(code / ((int64_t)W * H));
        if (cam >= n_cams) return r;
        pix = (uint32_t)(code - (int64_t)cam * W * H);
        r.out = i;
    }
    const qsdvr_camera &c = cams[cam];
    if (c.width != W || c.height != H) return r;  // inconsistent camera: lane stays idle
    r.cam = cam;
    r.center = f3{c.translation[0], c.translation[1], c.translation[2]};
    uint32_t count;
    if (ray_state != nullptr) {
        // precomputed by qsdvr_ray_state: (origin, count bits), (dir, 0)
        const float4 a = ray_state[2 * r.out], b = ray_state[2 * r.out + 1];
        r.ori = f3{a.x, a.y, a.z};
        r.dir = f3{b.x, b.y, b.z};
        count = __float_as_uint(a.w);
    } else {
        RayCutResult cut = ray_cut_pixel(c, pix, interval);
        r.ori = cut.origin;
        r.dir = cut.dir;
        count = cut.count;
    }
    if (count < 2) return r;
    if (tlist != nullptr) {
        r.tl = tlist + r.out * (int64_t)K;
        r.renders = (uint32_t)(K - 1);
    } else {
        r.renders = count - 1;
    }
    return r;
}

// Weighted sum of the 8 vertex rows of a cell (render_torch.py:128-143 as
// sum_k w_k row_k).  Rows: base + dx + dy R + dz R^2.
template <int C>
__device__ __forceinline__ void gather_interp(const float *__restrict__ data, int64_t base,
                                              int R, const float w[8], float d[C]) {
#pragma unroll
    for (int c = 0; c < C; ++c) d[c] = 0.0f;
    const int64_t RR = (int64_t)R * R;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int64_t row = base + (k & 1) + ((k >> 1) & 1) * (int64_t)R + (k >> 2) * RR;
        const float *src = data + row * C;
        if constexpr (C % 4 == 0) {
            const float4 *s4 = reinterpret_cast<const float4 *>(src);
#pragma unroll
            for (int q = 0; q < C / 4; ++q) {
                float4 v = s4[q];
                d[4 * q + 0] += w[k] * v.x;
                d[4 * q + 1] += w[k] * v.y;
                d[4 * q + 2] += w[k] * v.z;
                d[4 * q + 3] += w[k] * v.w;
            }
        } else {
            // rows of C floats are only 4-B aligned: 16-B loads at 4-B alignment
            // (global_load_dwordx4 needs dword alignment), then the tail
            typedef float f4a __attribute__((ext_vector_type(4), aligned(4)));
#pragma unroll
            for (int q = 0; q < C / 4; ++q) {
                const f4a v = *reinterpret_cast<const f4a *>(src + 4 * q);
                d[4 * q + 0] += w[k] * v.x;
                d[4 * q + 1] += w[k] * v.y;
                d[4 * q + 2] += w[k] * v.z;
                d[4 * q + 3] += w[k] * v.w;
            }
#pragma unroll
            for (int q = C / 4 * 4; q < C; ++q) d[q] += w[k] * src[q];
        }
    }
}

// Cells produced by the ray walk always lie in [0, R)^3 (every sample is
// strictly inside the +-0.9999 cube); the clamp only guards degenerate camera
// input (NaN / inf) against out-of-bounds gathers.
__device__ __forceinline__ void clamp_cell(CellPoint &cp, int R) {
    if ((unsigned)cp.ix >= (unsigned)R || (unsigned)cp.iy >= (unsigned)R ||
        (unsigned)cp.iz >= (unsigned)R) {
        cp.ix = min(max(cp.ix, 0), R - 1);
        cp.iy = min(max(cp.iy, 0), R - 1);
        cp.iz = min(max(cp.iz, 0), R - 1);
        cp.cell = cp.ix + cp.iy * R + cp.iz * R * R;
        cp.local = f3{0.0f, 0.0f, 0.0f};
    }
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
    return v;
}

// DPP lane moves within rows of 16 lanes (row_shr:D = 0x110 + D, row_shl:1
// = 0x101, quad_perm, mirrors); a lane whose source lies outside its row
// reads 0
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}

// Wave-wide min / max returned in a scalar register: DPP steps within each
// row of 16 lanes (quad_perm xor 1 and xor 2, half-row mirror, row mirror)
// leave every lane with its row's extreme; the four rows are read with
// readlane.  Every lane of the wave must be active.
__device__ __forceinline__ int wave_min_s(int v) {
    v = min(v, dpp_i<0xB1>(v));
    v = min(v, dpp_i<0x4E>(v));
    v = min(v, dpp_i<0x141>(v));
    v = min(v, dpp_i<0x140>(v));
    const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return min(min(a, b), min(c, d));
}

__device__ __forceinline__ int wave_max_s(int v) {
    v = max(v, dpp_i<0xB1>(v));
    v = max(v, dpp_i<0x4E>(v));
    v = max(v, dpp_i<0x141>(v));
    v = max(v, dpp_i<0x140>(v));
    const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return max(max(a, b), max(c, d));
}

// Orders this wave's LDS accesses for the compiler only.  A wave's LDS
// instructions execute in issue order, so LDS data written by one lane is
// visible to a later read by another lane of the same wave without
// s_barrier (each wave below works in its own LDS region).
__device__ __forceinline__ void wave_lds_order() { __asm__ volatile("" ::: "memory"); }

// Appearance staging box: the vertex rows of the box of cells that a wave's
// samples fall in at one step, copied to LDS once by the whole wave (64
// lanes x 16 B per load) and read there by each lane, in place of 64
// per-lane 16-B gathers per sample.  kBoxV vertices of C = 32 channels per
// wave, each row padded by one quad (9 KB): a lane's 8 reads of a row share
// one address register (the quad is the instruction's offset), and lanes
// reading different vertices at one quad fall in different banks.
constexpr int kBoxV = 64;

struct BoxView {
    const float4 *lds;  // nullptr: gather from global memory
    int vb;             // the lane's cell-origin vertex in the box
    int ex, sxy;        // box strides in vertices
};

// gather_interp from the staged box (same operation order).
template <int C>
__device__ __forceinline__ void box_interp(const BoxView &bx, const float w[8], float d[C]) {
    constexpr int Q = C / 4, QP = Q + 1;
#pragma unroll
    for (int c = 0; c < C; ++c) d[c] = 0.0f;
    const float4 *p0 = bx.lds + bx.vb * QP;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float4 *p = p0 + ((k & 1) + ((k >> 1) & 1) * bx.ex + (k >> 2) * bx.sxy) * QP;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const float4 val = p[q];
            d[4 * q + 0] += w[k] * val.x;
            d[4 * q + 1] += w[k] * val.y;
            d[4 * q + 2] += w[k] * val.z;
            d[4 * q + 3] += w[k] * val.w;
        }
    }
}

// Stage the box of the cells of the lanes with ``need`` set (their cell
// ``rp``): returns the lane's view of it, or a null view (gather from global
// memory) when the box has more than kBoxV vertices, and the box's vertex
// count in nv_out.  Called by the whole wave with every lane active.
template <int C>
__device__ __forceinline__ BoxView stage_box(float4 *box, const float *__restrict__ data, int R,
                                             bool need, const CellPoint &rp, int lane,
                                             int &nv_out) {
    constexpr int Q = C / 4, QP = Q + 1;
    const int lx = wave_min_s(need ? rp.ix : INT_MAX), hx = wave_max_s(need ? rp.ix : INT_MIN);
    const int ly = wave_min_s(need ? rp.iy : INT_MAX), hy = wave_max_s(need ? rp.iy : INT_MIN);
    const int lz = wave_min_s(need ? rp.iz : INT_MAX), hz = wave_max_s(need ? rp.iz : INT_MIN);
    const int ex = hx - lx + 2, ey = hy - ly + 2, ez = hz - lz + 2;
    const int sxy = ex * ey, nv = sxy * ez;
    nv_out = nv;
    if (nv > kBoxV) return BoxView{nullptr, 0, 0, 0};
    // box vertex v = vx + vy ex + vz sxy (v < 64, sxy >= 4, ex >= 2): the
    // quotients from float reciprocals are exact, since (v + 0.5) / sxy lies
    // at least 0.5 / sxy from an integer
    const float rsxy = __builtin_amdgcn_rcpf((float)sxy), rex = __builtin_amdgcn_rcpf((float)ex);
    const int64_t RR = (int64_t)R * R;
    const float4 *__restrict__ src =
        reinterpret_cast<const float4 *>(data) + ((int64_t)lx + (int64_t)ly * R + lz * RR) * Q;
    wave_lds_order();  // the previous step's box reads are issued before these writes
    for (int i = lane; i < nv * Q; i += kWave) {
        const int v = i / Q, q = i % Q;
        const int vz = (int)(((float)v + 0.5f) * rsxy), r2 = v - vz * sxy;
        const int vy = (int)(((float)r2 + 0.5f) * rex), vx = r2 - vy * ex;
        box[v * QP + q] = src[(vx + vy * R + vz * (int)RR) * Q + q];
    }
    wave_lds_order();
    return BoxView{box, (rp.ix - lx) + (rp.iy - ly) * ex + (rp.iz - lz) * sxy, ex, sxy};
}

// 53-channel rows (212 B, only 4-B aligned in the grid) are staged the same
// way into 56-float LDS rows: 13 16-B chunks (unaligned global loads, aligned
// LDS stores) and a dword per row, for boxes of up to kBoxV53 vertices
// (10.5 KB per wave).
constexpr int kBoxV53 = 48;
constexpr int kRow53 = 56;  // LDS floats per staged 53-channel row

template <int C>
__device__ __forceinline__ BoxView stage_box53(float4 *box, const float *__restrict__ data,
                                               int R, bool need, const CellPoint &rp, int lane,
                                               int &nv_out) {
    static_assert(C == 53, "53-channel staging");
    const int lx = wave_min_s(need ? rp.ix : INT_MAX), hx = wave_max_s(need ? rp.ix : INT_MIN);
    const int ly = wave_min_s(need ? rp.iy : INT_MAX), hy = wave_max_s(need ? rp.iy : INT_MIN);
    const int lz = wave_min_s(need ? rp.iz : INT_MAX), hz = wave_max_s(need ? rp.iz : INT_MIN);
    if (hx < lx) {  // no lane needs rows
        nv_out = 0;
        return BoxView{nullptr, 0, 0, 0};
    }
    const int ex = hx - lx + 2, ey = hy - ly + 2, ez = hz - lz + 2;
    const int sxy = ex * ey, nv = sxy * ez;
    nv_out = nv;
    if (nv > kBoxV53) return BoxView{nullptr, 0, 0, 0};
    const float rsxy = __builtin_amdgcn_rcpf((float)sxy), rex = __builtin_amdgcn_rcpf((float)ex);
    const int64_t RR = (int64_t)R * R;
    const float *__restrict__ src = data + ((int64_t)lx + (int64_t)ly * R + lz * RR) * C;
    float *dst = reinterpret_cast<float *>(box);
    typedef float f4a __attribute__((ext_vector_type(4), aligned(4)));
    wave_lds_order();  // the previous step's box reads are issued before these writes
    for (int i = lane; i < nv * 14; i += kWave) {
        const int v = i / 14, k = i - 14 * v;
        const int vz = (int)(((float)v + 0.5f) * rsxy), r2 = v - vz * sxy;
        const int vy = (int)(((float)r2 + 0.5f) * rex), vx = r2 - vy * ex;
        const float *row = src + (int64_t)(vx + vy * R + vz * (int)RR) * C;
        if (k < 13) {
            const f4a val = *reinterpret_cast<const f4a *>(row + 4 * k);
            *reinterpret_cast<float4 *>(dst + v * kRow53 + 4 * k) =
                float4{val.x, val.y, val.z, val.w};
        } else {
            dst[v * kRow53 + 52] = row[52];
        }
    }
    wave_lds_order();
    return BoxView{box, (rp.ix - lx) + (rp.iy - ly) * ex + (rp.iz - lz) * sxy, ex, sxy};
}

// gather_interp<53> from the staged box (same operation order).
__device__ __forceinline__ void box_interp53(const BoxView &bx, const float w[8], float d[53]) {
#pragma unroll
    for (int c = 0; c < 53; ++c) d[c] = 0.0f;
    const float *p0 = reinterpret_cast<const float *>(bx.lds) + bx.vb * kRow53;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float *p = p0 + ((k & 1) + ((k >> 1) & 1) * bx.ex + (k >> 2) * bx.sxy) * kRow53;
#pragma unroll
        for (int q = 0; q < 13; ++q) {
            const float4 val = *reinterpret_cast<const float4 *>(p + 4 * q);
            d[4 * q + 0] += w[k] * val.x;
            d[4 * q + 1] += w[k] * val.y;
            d[4 * q + 2] += w[k] * val.z;
            d[4 * q + 3] += w[k] * val.w;
        }
        d[52] += w[k] * p[52];
    }
}

struct SampleEval {
    CellPoint rp;     // render sample cell / local point
    f3 normal;
    f3 rgb, spec;
    float t;          // distance from the camera centre
};

// Cell and local point of the cube position ori + dir * t (listed samples).
__device__ __forceinline__ CellPoint point_at(const RaySetup &ray, float t, float half_reso, int R) {
    const float qx = ((ray.ori.x + ray.dir.x * t) + 1.0f) * half_reso;
    const float qy = ((ray.ori.y + ray.dir.y * t) + 1.0f) * half_reso;
    const float qz = ((ray.ori.z + ray.dir.z * t) + 1.0f) * half_reso;
    return to_cell(qx, qy, qz, R);
}

// Sample positions of a listed ray.  AHEAD: held ahead of their use, so at
// step j a = t_j and b = t_{j+1} are in registers and t_{j+2} is in flight,
// and a step's quadric and row gathers never wait on a load from the list
// (each lane's list is a row of its own, so a load per step puts a memory
// round trip in front of every gather of the step).  Otherwise each
// position is loaded where it is used, which needs no registers: the
// 53-channel forward then keeps 4 waves per SIMD.  Uniform rays read none.
template <bool AHEAD>
struct TWindow {
    float a = 0.0f, b = 0.0f, c = 0.0f;
    // AHEAD: the list the window reads, and its last index.  A lane without
    // a list (an idle lane) reads entry 0 of `spare` instead, and every lane
    // advances every step (past the ray's end the last entry repeats): the
    // loads are then not lane-masked and write the window's registers
    // themselves, where a masked load is merged into them with copies that
    // wait for every outstanding load and atomic.
    const float *src = nullptr;
    uint32_t last = 0;
    // spare: any readable list entry (the kernel's list argument)
    __device__ __forceinline__ void start(const RaySetup &ray, const float *spare) {
        if (AHEAD) {
            src = ray.tl != nullptr ? ray.tl : spare;
            last = ray.tl != nullptr ? ray.renders : 0u;
            a = src[0];
            b = src[min(1u, last)];
            c = src[min(2u, last)];
        }
    }
    // step j done: t_{j+1}, t_{j+2} move down and t_{j+3} is requested
    __device__ __forceinline__ void advance(int j) {
        if (AHEAD) {
            a = b;
            b = c;
            c = src[min((uint32_t)(j + 3), last)];
        }
    }
    // t_j and t_{j+1} at step j
    __device__ __forceinline__ float t0(const RaySetup &ray, int j) const {
        return AHEAD ? a : (ray.tl ? ray.tl[j] : 0.0f);
    }
    __device__ __forceinline__ float t1(const RaySetup &ray, int j) const {
        return AHEAD ? b : (ray.tl ? ray.tl[j + 1] : 0.0f);
    }
    __device__ __forceinline__ float t2(const RaySetup &ray, int j) const {
        return AHEAD ? c : (ray.tl ? ray.tl[j + 2] : 0.0f);
    }
    // position of render sample j (midpoint of t_j and t_{j+1})
    __device__ __forceinline__ float mid(const RaySetup &ray, int j) const {
        return 0.5f * (t0(ray, j) + t1(ray, j));
    }
};

// Render sample j of the ray (cell clamped as clamp_cell does); sl is its
// position on a listed ray (TWindow::mid), unused for uniform rays.
__device__ __forceinline__ CellPoint render_point_t(const RaySetup &ray, uint32_t j, float sl,
                                                    float interval, float half_reso, int R) {
    CellPoint rp = ray.tl ? point_at(ray, sl, half_reso, R)
                          : render_sample(ray.ori, ray.dir, j, interval, half_reso, R);
    clamp_cell(rp, R);
    return rp;
}

// Render sample j at its cell point rp (quadric q): normal, appearance (from
// the staged box when bx.lds is set, else from global memory), shading and
// distance.
template <int NC>
__device__ __forceinline__ void eval_render_at(const RaySetup &ray, uint32_t j, float sl,
                                               float interval, const CellPoint &rp,
                                               const Quadric &q, int R,
                                               const float *__restrict__ data, const BoxView &bx,
                                               float w[8], float d[3 * NC + 5], SampleEval &e) {
    constexpr int C = 3 * NC + 5;
    const float sj = ray.tl ? sl : ((float)j + 0.5f) * interval;
    e.rp = rp;
    e.normal = quadric_normal(q, e.rp.local);
    trilinear_weights(e.rp.local, w);
    if constexpr (C == 32) {
        if (bx.lds != nullptr) box_interp<C>(bx, w, d);
        else gather_interp<C>(data, e.rp.cell, R, w, d);
    } else if constexpr (C == 53) {
        if (bx.lds != nullptr) box_interp53(bx, w, d);
        else gather_interp<C>(data, e.rp.cell, R, w, d);
    } else {
        gather_interp<C>(data, e.rp.cell, R, w, d);
    }
    ShadeOut so = shade<NC>(e.normal, ray.dir, d);
    e.rgb = so.rgb;
    e.spec = so.spec;
    const float px = ray.ori.x + ray.dir.x * sj - ray.center.x;
    const float py = ray.ori.y + ray.dir.y * sj - ray.center.y;
    const float pz = ray.ori.z + ray.dir.z * sj - ray.center.z;
    e.t = fast_sqrt(px * px + py * py + pz * pz);
}

// Render sample j with per-lane gathers from global memory (sl: position on
// a listed ray, TWindow::mid).
template <int NC>
__device__ __forceinline__ void eval_render_sample_t(const RaySetup &ray, uint32_t j, float sl,
                                                     float interval, float half_reso, int R,
                                                     const float *__restrict__ qgrid,
                                                     const float *__restrict__ data, float w[8],
                                                     float d[3 * NC + 5], SampleEval &e) {
    constexpr int C = 3 * NC + 5;
    const float sj = ray.tl ? sl : ((float)j + 0.5f) * interval;
    e.rp = ray.tl ? point_at(ray, sj, half_reso, R)
                  : render_sample(ray.ori, ray.dir, j, interval, half_reso, R);
    clamp_cell(e.rp, R);
    const Quadric q = load_quadric(qgrid, e.rp.cell);
    e.normal = quadric_normal(q, e.rp.local);
    trilinear_weights(e.rp.local, w);
    gather_interp<C>(data, e.rp.cell, R, w, d);
    ShadeOut so = shade<NC>(e.normal, ray.dir, d);
    e.rgb = so.rgb;
    e.spec = so.spec;
    const float px = ray.ori.x + ray.dir.x * sj - ray.center.x;
    const float py = ray.ori.y + ray.dir.y * sj - ray.center.y;
    const float pz = ray.ori.z + ray.dir.z * sj - ray.center.z;
    e.t = fast_sqrt(px * px + py * py + pz * pz);
}

// Render sample j, a listed ray's position read from its list.
template <int NC>
__device__ __forceinline__ void eval_render_sample(const RaySetup &ray, uint32_t j, float interval,
                                                   float half_reso, int R,
                                                   const float *__restrict__ qgrid,
                                                   const float *__restrict__ data, float w[8],
                                                   float d[3 * NC + 5], SampleEval &e) {
    const float sl = ray.tl ? 0.5f * (ray.tl[j] + ray.tl[j + 1]) : 0.0f;
    eval_render_sample_t<NC>(ray, j, sl, interval, half_reso, R, qgrid, data, w, d, e);
}

// Cell point of sdf sample j (as sdf_at computes it); tj is its position on
// a listed ray (unused otherwise).
__device__ __forceinline__ CellPoint ray_sdf_point_t(const RaySetup &ray, uint32_t j, float tj,
                                                     float interval, float half_reso, int R) {
    CellPoint cp = ray.tl ? point_at(ray, tj, half_reso, R)
                          : sdf_sample(ray.ori, ray.dir, j, interval, half_reso, R);
    clamp_cell(cp, R);
    return cp;
}

__device__ __forceinline__ CellPoint ray_sdf_point(const RaySetup &ray, uint32_t j, float interval,
                                                   float half_reso, int R) {
    return ray_sdf_point_t(ray, j, ray.tl ? ray.tl[j] : 0.0f, interval, half_reso, R);
}

// SDF at sample j; tj is its position on a listed ray (unused otherwise).
__device__ __forceinline__ float sdf_at_t(const RaySetup &ray, uint32_t j, float tj,
                                          float interval, float half_reso, int R, float inv_reso,
                                          const float *__restrict__ qgrid, CellPoint &cp) {
    cp = ray.tl ? point_at(ray, tj, half_reso, R)
                : sdf_sample(ray.ori, ray.dir, j, interval, half_reso, R);
    clamp_cell(cp, R);
    return quadric_sdf(load_quadric(qgrid, cp.cell), cp.local, inv_reso);
}

__device__ __forceinline__ float sdf_at(const RaySetup &ray, uint32_t j, float interval,
                                        float half_reso, int R, float inv_reso,
                                        const float *__restrict__ qgrid, CellPoint &cp) {
    return sdf_at_t(ray, j, ray.tl ? ray.tl[j] : 0.0f, interval, half_reso, R, inv_reso, qgrid,
                    cp);
}

// LISTED: the rays carry explicit sample lists (tlist); the uniform-ray
// instantiation drops the list code and its registers.  C = 32: at most 128
// registers per lane, 4 waves per SIMD (as many as the staging LDS allows).
// COST: record each tile's expected backward cost in tile_cost (frame
// launches only).  Only the costed entry point instantiates it: at 53
// channels the store alone moves the compiler to a schedule with fewer
// registers that measured 47% slower.
template <int NC, bool BINS, bool LISTED, bool COST = false>
__global__ void __launch_bounds__(kFusedBlock, NC == 9 ? 4 : 1) fused_fwd_kernel(
    const qsdvr_camera *__restrict__ cams, int n_cams, uint32_t W, uint32_t H,
    const int32_t *__restrict__ ray_pixel,
    int64_t n_rays, const float *__restrict__ qgrid, const float *__restrict__ data, int R,
    float interval, float s, float *__restrict__ color, float *__restrict__ depth,
    const float4 *__restrict__ ray_state, const float *__restrict__ tlist, int K,
    float *__restrict__ bins, uint32_t *__restrict__ tile_cost) {
    constexpr int C = 3 * NC + 5;
    extern __shared__ float4 fwd_lds[];
    RaySetup ray = setup_ray(cams, n_cams, W, H, ray_pixel, n_rays, interval, ray_state,
                             LISTED ? tlist : nullptr, K);
    if constexpr (!LISTED) ray.tl = nullptr;
    const bool live = ray.out >= 0 && ray.renders > 0;
    // Tile waves (rays of one 8x8 pixel tile) stage their appearance rows in
    // LDS; waves of listed rays gather per lane.  The step loop runs
    // wave-uniformly (lanes past their last sample idle) for the staging.
    const bool stage = (C == 32 || C == 53) && ray_pixel == nullptr;
    const int lane = threadIdx.x % kWave;
    float4 *box = fwd_lds + (threadIdx.x / kWave) * (C == 32 ? kBoxV * (C / 4 + 1)
                                                             : kBoxV53 * kRow53 / 4);
    const float half_reso = (float)R / 2.0f, inv_reso = 1.0f / (float)R;
    f3 col{0.0f, 0.0f, 0.0f};
    float dep = 0.0f;
    const int steps = wave_max(live ? (int)ray.renders : 0);
    CellPoint cp;
    TWindow<LISTED && C == 32> tw;
    tw.start(ray, tlist);
    float e0 = live ? expf(-s * sdf_at_t(ray, 0, tw.t0(ray, 0), interval, half_reso, R, inv_reso,
                                         qgrid, cp))
                    : 1.0f;
    // QAHEAD (32 channels): the quadric of sdf sample j + 1 is loaded a step
    // ahead of its use, so a step's alpha never waits on a load.
    constexpr bool QAHEAD = C == 32;
    Quadric qn;
    if (QAHEAD && live) {
        cp = ray_sdf_point_t(ray, 1, tw.t1(ray, 0), interval, half_reso, R);
        qn = load_quadric(qgrid, cp.cell);
    }
    float T = 1.0f;
    // 64 equal bins over the ray's sampled segment (training targets).
    // The bin index never decreases along the ray, so each bin is summed
    // in a register and stored once (the caller zeroes the bins).
    int bin = 0;
    float bin_sum = 0.0f;
    // The tile's expected backward cost, for qsdvr_fused_backward_ordered:
    // 64 per step with an appearance sample, plus the step's box vertices,
    // plus 4 times them when the box overflows (the adjoint then needs more
    // than one round).  Fitted to measured per-tile backward times.
    [[maybe_unused]] uint32_t cost = 0;
    for (int j = 0; j < steps; ++j) {
        const bool act = live && j < (int)ray.renders;
        float a = 0.0f, wt = 0.0f, e1 = 0.0f;
        if (act) {
            if constexpr (QAHEAD) {
                e1 = expf(-s * quadric_sdf(qn, cp.local, inv_reso));
            } else {
                e1 = expf(-s * sdf_at_t(ray, j + 1, tw.t1(ray, j), interval, half_reso, R,
                                        inv_reso, qgrid, cp));
            }
            a = fmaxf((e1 - e0) * fast_rcp(1.0f + e1), 0.0f);
            wt = a * T;
        }
        if constexpr (QAHEAD) {
            // sdf sample j + 2, for the next step, requested by every lane (past
            // its ray's end a lane loads a clamped cell's quadric and never uses
            // it), so that the load is not lane-masked
            const CellPoint cn = ray_sdf_point_t(ray, j + 2, tw.t2(ray, j), interval, half_reso, R);
            if (act && (uint32_t)(j + 2) <= ray.renders) cp = cn;
            qn = load_quadric(qgrid, cn.cell);
        }
        // A sample of weight exactly zero (alpha clamped to 0 where the SDF
        // rises along the ray, or T underflowed) adds nothing: its
        // appearance gather and shading are skipped.
        const bool need = wt != 0.0f;
        if (__ballot(need) != 0) {
            if constexpr (COST) cost += 64;
            CellPoint rp;
            rp.ix = rp.iy = rp.iz = 0;
            rp.cell = 0;
            Quadric qr;  // the render sample's quadric, requested before the row staging
            if (need) {
                rp = render_point_t(ray, (uint32_t)j, tw.mid(ray, j), interval, half_reso, R);
                qr = load_quadric(qgrid, rp.cell);
            }
            BoxView bx{nullptr, 0, 0, 0};
            if constexpr (C == 32) {
                if (stage) {
                    int nv = 0;
                    bx = stage_box<C>(box, data, R, need, rp, lane, nv);
                    if constexpr (COST) cost += (uint32_t)nv + (nv > kBoxV ? 4u * (uint32_t)nv : 0u);
                }
            } else if constexpr (C == 53) {
                if (stage) {
                    int nv = 0;
                    bx = stage_box53<C>(box, data, R, need, rp, lane, nv);
                }
            }
            if (need) {
                float w[8], d[C];
                SampleEval ev;
                eval_render_at<NC>(ray, (uint32_t)j, tw.mid(ray, j), interval, rp, qr, R, data, bx,
                                   w, d, ev);
                col.x += ev.rgb.x * wt;
                col.y += ev.rgb.y * wt;
                col.z += ev.rgb.z * wt;
                dep += ev.t * wt;
            }
        }
        if (act) {
            if constexpr (BINS) {
                const int b = min((int)(((float)j + 0.5f) / (float)ray.renders * 64.0f), 63);
                if (b != bin) {
                    bins[ray.out * 64 + bin] = bin_sum;
                    bin = b;
                    bin_sum = 0.0f;
                }
                bin_sum += wt;
            }
            T = T * (1.0f - a);
            e0 = e1;
        }
        tw.advance(j);
    }
    if constexpr (COST) {
        if (ray_pixel == nullptr && lane == 0) {
            const int64_t tw_ = (int64_t)logical_block() * kWavesPerBlock + threadIdx.x / kWave;
            if (tw_ < (int64_t)((W + 7) / 8) * ((H + 7) / 8) * n_cams) tile_cost[tw_] = cost;
        }
    }
    if (ray.out < 0) return;
    if constexpr (BINS) {
        if (live) bins[ray.out * 64 + bin] = bin_sum;
    }
    store3(color, ray.out, col);
    if (depth) depth[ray.out] = dep;
}

// ---------------------------------------------------------------------------
// Listed rays (random batches over many poses) in the deterministic adjoint:
// one wavefront per ray, lanes over its samples, 64 at a time.  A
// lane-per-ray walk leaves too few waves to hide the latency of the
// per-sample adjoint at batch sizes like 65,536 rays (1,024 waves); here
// every sample is a lane.  (The forward of such batches stays lane per ray:
// its random row gathers already run at the memory system's rate, and the
// wave form measured no faster.)

// Ray i of the list, set up by every lane of the wave from the precomputed
// ray state (qsdvr_ray_state).
__device__ __forceinline__ RaySetup setup_ray_wave(const qsdvr_camera *__restrict__ cams,
                                                   int n_cams, uint32_t W, uint32_t H,
                                                   const int32_t *__restrict__ ray_pixel,
                                                   int64_t i, float interval,
                                                   const float4 *__restrict__ ray_state) {
    RaySetup r;
    r.out = -1;
    r.renders = 0;
    r.tl = nullptr;
    const int64_t code = (uint32_t)ray_pixel[i];
    const int cam = (int)(code / ((int64_t)W * H));
    if (cam >= n_cams) return r;
    const uint32_t pix = (uint32_t)(code - (int64_t)cam * W * H);
    const qsdvr_camera &c = cams[cam];
    if (c.width != W || c.height != H) return r;
    r.out = i;
    r.cam = cam;
    r.center = f3{c.translation[0], c.translation[1], c.translation[2]};
    const float4 a = ray_state[2 * i], b = ray_state[2 * i + 1];
    r.ori = f3{a.x, a.y, a.z};
    r.dir = f3{b.x, b.y, b.z};
    const uint32_t count = __float_as_uint(a.w);
    r.renders = count >= 2 ? count - 1 : 0;
    return r;
}

// Exclusive product over lanes < lane of x (all lanes call it).
__device__ __forceinline__ float wave_excl_product(float x, float &total) {
    const int lane = threadIdx.x % kWave;
    float p = x;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const float u = __shfl_up(p, o, kWave);
        if (lane >= o) p *= u;
    }
    total = __shfl(p, kWave - 1, kWave);
    const float up = __shfl_up(p, 1, kWave);
    return lane == 0 ? 1.0f : up;
}

__device__ __forceinline__ void scatter_quadric(float *__restrict__ g, int32_t cell,
                                                const float v[7], int k_count) {
    float *dst = g + 7 * (int64_t)cell;
    for (int k = 0; k < k_count; ++k) atomic_addf(dst + k, v[k]);
}

template <int NC>
__global__ void __launch_bounds__(kFusedBlock) fused_bwd_kernel(
    const qsdvr_camera *__restrict__ cams, int n_cams, uint32_t W, uint32_t H,
    const int32_t *__restrict__ ray_pixel,
    int64_t n_rays, const float *__restrict__ qgrid, const float *__restrict__ data, int R,
    float interval, float s, const float *__restrict__ color, const float *__restrict__ depth,
    const float *__restrict__ gcolor, const float *__restrict__ gdepth,
    float *__restrict__ gq, float *__restrict__ gdata, const float4 *__restrict__ ray_state,
    const float *__restrict__ tlist, int K) {
    constexpr int C = 3 * NC + 5;
    RaySetup ray = setup_ray(cams, n_cams, W, H, ray_pixel, n_rays, interval, ray_state, tlist, K);
    if (ray.out < 0 || ray.renders == 0) return;
    const float half_reso = (float)R / 2.0f, inv_reso = 1.0f / (float)R;
    const f3 Cs = load3(color, ray.out);
    const f3 gC = load3(gcolor, ray.out);
    const float Ds = depth ? depth[ray.out] : 0.0f;
    const float gD = gdepth ? gdepth[ray.out] : 0.0f;
    const int64_t RR = (int64_t)R * R;

    CellPoint cp0, cp1;
    float sdf0 = sdf_at(ray, 0, interval, half_reso, R, inv_reso, qgrid, cp0);
    float e0 = expf(-s * sdf0);
    float T = 1.0f;
    f3 pre{0.0f, 0.0f, 0.0f};
    float pre_d = 0.0f;
    float carry = 0.0f;  // gradient of sdf sample j from alpha_{j-1}
    for (uint32_t j = 0; j < ray.renders; ++j) {
        float sdf1 = sdf_at(ray, j + 1, interval, half_reso, R, inv_reso, qgrid, cp1);
        float e1 = expf(-s * sdf1);
        float inv1 = fast_rcp(1.0f + e1);
        float praw = (e1 - e0) * inv1;
        float a = fmaxf(praw, 0.0f);
        const float wt = a * T;
        // The render sample is needed for the prefix (weight wt) and for
        // dL/dalpha (when the clamp passes, times T); otherwise it adds
        // exactly nothing and is not evaluated.
        const bool need = praw >= 0.0f && T != 0.0f;
        float w[8], d[C];
        SampleEval ev;
        float here = 0.0f;
        if (need) {
            eval_render_sample<NC>(ray, j, interval, half_reso, R, qgrid, data, w, d, ev);
            pre.x += ev.rgb.x * wt;
            pre.y += ev.rgb.y * wt;
            pre.z += ev.rgb.z * wt;
            pre_d += ev.t * wt;
            here = gC.x * ev.rgb.x + gC.y * ev.rgb.y + gC.z * ev.rgb.z + gD * ev.t;
        }
        // dL/dalpha_j = T_j (gC.rgb_j + gD t_j) - (gC.S_j + gD SD_j) / (1 - alpha_j)
        float g_sdf0 = carry, g_sdf1 = 0.0f;
        if (praw >= 0.0f) {
            float one_over = (1.0f + e1) * fast_rcp(1.0f + e0);
            float suffix = gC.x * (Cs.x - pre.x) + gC.y * (Cs.y - pre.y) + gC.z * (Cs.z - pre.z) +
                           gD * (Ds - pre_d);
            float da = T * here - suffix * one_over;
            g_sdf0 += da * s * e0 * inv1;
            g_sdf1 = -da * s * (e1 * inv1) * ((1.0f + e0) * inv1);
        }
        // sdf sample j is final: scatter its quadric gradient
        {
            float gq7[7] = {0, 0, 0, 0, 0, 0, 0};
            quadric_sdf_grad(load_quadric(qgrid, cp0.cell), cp0.local, g_sdf0, inv_reso, gq7);
            scatter_quadric(gq, cp0.cell, gq7, 7);
        }
        carry = g_sdf1;
        // rgb_j gradient -> shader adjoint -> normal and appearance grads
        // (all exactly zero when wt == 0)
        if (need && wt != 0.0f) {
            f3 grgb{gC.x * wt, gC.y * wt, gC.z * wt};
            float dd[C];
            f3 gn = shade_vjp<NC>(ev.normal, ray.dir, d, ev.spec, grgb, dd);
            {
                float gq7[7] = {0, 0, 0, 0, 0, 0, 0};
                quadric_normal_grad(load_quadric(qgrid, ev.rp.cell), ev.rp.local, gn, gq7);
                scatter_quadric(gq, ev.rp.cell, gq7, 6);
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int64_t row = (int64_t)ev.rp.cell + (k & 1) + ((k >> 1) & 1) * (int64_t)R +
                                    (k >> 2) * RR;
                float *dst = gdata + row * C;
#pragma unroll
                for (int c = 0; c < C; ++c) atomic_addf(dst + c, w[k] * dd[c]);
            }
        }
        T = T * (1.0f - a);
        e0 = e1;
        cp0 = cp1;
    }
    {
        float gq7[7] = {0, 0, 0, 0, 0, 0, 0};
        quadric_sdf_grad(load_quadric(qgrid, cp0.cell), cp0.local, carry, inv_reso, gq7);
        scatter_quadric(gq, cp0.cell, gq7, 7);
    }
}

// ------------------------------------------ deterministic-scatter adjoint ----
// The same recompute adjoint, lane = ray, writing gradient records (keyed by
// cell) instead of adding into the grids; det_scatter.hip sorts and sums
// them in a fixed order.  Slot q = offsets[ray] + j holds sdf sample j and
// render sample j of the ray: quadric records 2q (sdf) and 2q + 1 (normal),
// appearance record q.  The last slot of a ray has no render sample and
// carries invalid keys.
struct DetRecordOut {
    const uint32_t *offsets;
    uint32_t *qkey, *qidx;
    float4 *qval;
    uint32_t *akey, *aidx;
    float *aval;
    uint32_t q_invalid;  // R^3
    int cbits;           // appearance key = (parity class << cbits) | cell
};

__device__ __forceinline__ void det_put_quadric(const DetRecordOut &o, uint32_t rec, uint32_t key,
                                                const float g[7]) {
    o.qkey[rec] = key;
    o.qidx[rec] = rec;
    o.qval[2 * (int64_t)rec] = make_float4(g[0], g[1], g[2], g[3]);
    o.qval[2 * (int64_t)rec + 1] = make_float4(g[4], g[5], g[6], 0.0f);
}

template <int NC>
__global__ void __launch_bounds__(kFusedBlock) fused_bwd_record_kernel(
    const qsdvr_camera *__restrict__ cams, int n_cams, uint32_t W, uint32_t H,
    const int32_t *__restrict__ ray_pixel,
    int64_t n_rays, const float *__restrict__ qgrid, const float *__restrict__ data, int R,
    float interval, float s, const float *__restrict__ color, const float *__restrict__ depth,
    const float *__restrict__ gcolor, const float *__restrict__ gdepth,
    const float4 *__restrict__ ray_state, const float *__restrict__ tlist, int K,
    DetRecordOut o) {
    constexpr int C = 3 * NC + 5;
    constexpr int AP = 4 + (C + 3) / 4 * 4;
    RaySetup ray = setup_ray(cams, n_cams, W, H, ray_pixel, n_rays, interval, ray_state, tlist, K);
    if (ray.out < 0 || ray.renders == 0) return;
    const uint32_t base = o.offsets[ray.out];
    const float half_reso = (float)R / 2.0f, inv_reso = 1.0f / (float)R;
    const f3 Cs = load3(color, ray.out);
    const f3 gC = load3(gcolor, ray.out);
    const float Ds = depth ? depth[ray.out] : 0.0f;
    const float gD = gdepth ? gdepth[ray.out] : 0.0f;

    CellPoint cp0, cp1;
    float sdf0 = sdf_at(ray, 0, interval, half_reso, R, inv_reso, qgrid, cp0);
    float e0 = expf(-s * sdf0);
    float T = 1.0f;
    f3 pre{0.0f, 0.0f, 0.0f};
    float pre_d = 0.0f;
    float carry = 0.0f;
    for (uint32_t j = 0; j < ray.renders; ++j) {
        const uint32_t q = base + j;
        float sdf1 = sdf_at(ray, j + 1, interval, half_reso, R, inv_reso, qgrid, cp1);
        float e1 = expf(-s * sdf1);
        float inv1 = fast_rcp(1.0f + e1);
        float praw = (e1 - e0) * inv1;
        float a = fmaxf(praw, 0.0f);
        const float wt = a * T;
        const bool need = praw >= 0.0f && T != 0.0f;  // as in fused_bwd_kernel
        float w[8], d[C];
        SampleEval ev;
        float here = 0.0f;
        if (need) {
            eval_render_sample<NC>(ray, j, interval, half_reso, R, qgrid, data, w, d, ev);
            pre.x += ev.rgb.x * wt;
            pre.y += ev.rgb.y * wt;
            pre.z += ev.rgb.z * wt;
            pre_d += ev.t * wt;
            here = gC.x * ev.rgb.x + gC.y * ev.rgb.y + gC.z * ev.rgb.z + gD * ev.t;
        }
        float g_sdf0 = carry, g_sdf1 = 0.0f;
        if (praw >= 0.0f) {
            float one_over = (1.0f + e1) * fast_rcp(1.0f + e0);
            float suffix = gC.x * (Cs.x - pre.x) + gC.y * (Cs.y - pre.y) + gC.z * (Cs.z - pre.z) +
                           gD * (Ds - pre_d);
            float da = T * here - suffix * one_over;
            g_sdf0 += da * s * e0 * inv1;
            g_sdf1 = -da * s * (e1 * inv1) * ((1.0f + e0) * inv1);
        }
        {
            float gq7[7] = {0, 0, 0, 0, 0, 0, 0};
            quadric_sdf_grad(load_quadric(qgrid, cp0.cell), cp0.local, g_sdf0, inv_reso, gq7);
            det_put_quadric(o, 2 * q, (uint32_t)cp0.cell, gq7);
        }
        carry = g_sdf1;
        if (!(need && wt != 0.0f)) {
            // zero-weight render sample: no normal or appearance gradient
            const float zero7[7] = {0, 0, 0, 0, 0, 0, 0};
            det_put_quadric(o, 2 * q + 1, o.q_invalid, zero7);
            o.akey[q] = 8u << o.cbits;
            o.aidx[q] = q;
        } else {
        f3 grgb{gC.x * wt, gC.y * wt, gC.z * wt};
        float dd[C];
        f3 gn = shade_vjp<NC>(ev.normal, ray.dir, d, ev.spec, grgb, dd);
        {
            float gq7[7] = {0, 0, 0, 0, 0, 0, 0};
            quadric_normal_grad(load_quadric(qgrid, ev.rp.cell), ev.rp.local, gn, gq7);
            det_put_quadric(o, 2 * q + 1, (uint32_t)ev.rp.cell, gq7);
        }
        {
            const uint32_t cls = (uint32_t)((ev.rp.ix & 1) | ((ev.rp.iy & 1) << 1) |
                                            ((ev.rp.iz & 1) << 2));
            o.akey[q] = (cls << o.cbits) | (uint32_t)ev.rp.cell;
            o.aidx[q] = q;
            float4 *dst = reinterpret_cast<float4 *>(o.aval + (int64_t)q * AP);
            dst[0] = make_float4(ev.rp.local.x, ev.rp.local.y, ev.rp.local.z, 0.0f);
#pragma unroll
            for (int c = 0; c < (AP - 4) / 4; ++c) {
                const int c0 = 4 * c;
                dst[1 + c] = make_float4(dd[c0], c0 + 1 < C ? dd[c0 + 1] : 0.0f,
                                         c0 + 2 < C ? dd[c0 + 2] : 0.0f,
                                         c0 + 3 < C ? dd[c0 + 3] : 0.0f);
            }
        }
        }
        T = T * (1.0f - a);
        e0 = e1;
        cp0 = cp1;
    }
    {
        const uint32_t q = base + ray.renders;
        float gq7[7] = {0, 0, 0, 0, 0, 0, 0};
        quadric_sdf_grad(load_quadric(qgrid, cp0.cell), cp0.local, carry, inv_reso, gq7);
        det_put_quadric(o, 2 * q, (uint32_t)cp0.cell, gq7);
        const float zero7[7] = {0, 0, 0, 0, 0, 0, 0};
        det_put_quadric(o, 2 * q + 1, o.q_invalid, zero7);
        o.akey[q] = 8u << o.cbits;
        o.aidx[q] = q;
    }
}

// fused_bwd_record_kernel for listed rays with uniform samples, one
// wavefront per ray: lane l of chunk c0 takes
// render sample j = c0 + l and finalises sdf sample j.  T is the carried
// transmittance times an exclusive product scan of (1 - alpha); the prefix
// sum_{i<=j} w_i rgb_i an inclusive sum scan; sdf sample j's gradient is
// its own alpha_j term plus lane j-1's alpha_{j-1} term.  The records are
// those of the lane-per-ray kernel, in the same slots.
__device__ __forceinline__ float wave_incl_sum(float x) {
    const int lane = threadIdx.x % kWave;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const float u = __shfl_up(x, o, kWave);
        if (lane >= o) x += u;
    }
    return x;
}

template <int NC>
__global__ void __launch_bounds__(kFusedBlock) fused_bwd_record_ray_kernel(
    const qsdvr_camera *__restrict__ cams, int n_cams, uint32_t W, uint32_t H,
    const int32_t *__restrict__ ray_pixel, int64_t n_rays, const float *__restrict__ qgrid,
    const float *__restrict__ data, int R, float interval, float s,
    const float *__restrict__ color, const float *__restrict__ depth,
    const float *__restrict__ gcolor, const float *__restrict__ gdepth,
    const float4 *__restrict__ ray_state, DetRecordOut o) {
    constexpr int C = 3 * NC + 5;
    constexpr int AP = 4 + (C + 3) / 4 * 4;
    const int lane = threadIdx.x % kWave;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    if (i >= n_rays) return;  // wave-uniform
    const RaySetup ray = setup_ray_wave(cams, n_cams, W, H, ray_pixel, i, interval, ray_state);
    if (ray.out < 0 || ray.renders == 0) return;
    const uint32_t base = o.offsets[ray.out];
    const float half_reso = (float)R / 2.0f, inv_reso = 1.0f / (float)R;
    const f3 Cs = load3(color, ray.out);
    const f3 gC = load3(gcolor, ray.out);
    const float Ds = depth ? depth[ray.out] : 0.0f;
    const float gD = gdepth ? gdepth[ray.out] : 0.0f;
    CellPoint cp;
    float e_prev = expf(-s * sdf_at(ray, 0, interval, half_reso, R, inv_reso, qgrid, cp));
    float T = 1.0f, g1_prev = 0.0f;
    f3 pre{0.0f, 0.0f, 0.0f};
    float pre_d = 0.0f;
    for (uint32_t c0 = 0; c0 < ray.renders; c0 += kWave) {
        const uint32_t j = c0 + (uint32_t)lane;
        const bool act = j < ray.renders;
        const uint32_t q = base + j;
        CellPoint cp0, cp1;
        cp0.cell = 0;
        float e1 = 0.0f;
        if (act) {
            cp0 = ray_sdf_point(ray, j, interval, half_reso, R);
            e1 = expf(-s * sdf_at(ray, j + 1, interval, half_reso, R, inv_reso, qgrid, cp1));
        }
        const float e_up = __shfl_up(e1, 1, kWave);
        const float e0 = lane == 0 ? e_prev : e_up;
        const float inv1 = fast_rcp(1.0f + e1);
        const float praw = act ? (e1 - e0) * inv1 : -1.0f;
        const float a = fmaxf(praw, 0.0f);
        float chunk;
        const float Tj = T * wave_excl_product(1.0f - a, chunk);
        const float wt = a * Tj;
        const bool need = praw >= 0.0f && Tj != 0.0f;  // as in fused_bwd_record_kernel
        float w[8], d[C];
        SampleEval ev;
        float here = 0.0f;
        f3 rw{0.0f, 0.0f, 0.0f};
        float tw = 0.0f;
        if (need) {
            eval_render_sample<NC>(ray, j, interval, half_reso, R, qgrid, data, w, d, ev);
            rw = f3{ev.rgb.x * wt, ev.rgb.y * wt, ev.rgb.z * wt};
            tw = ev.t * wt;
            here = gC.x * ev.rgb.x + gC.y * ev.rgb.y + gC.z * ev.rgb.z + gD * ev.t;
        }
        const f3 P{pre.x + wave_incl_sum(rw.x), pre.y + wave_incl_sum(rw.y),
                   pre.z + wave_incl_sum(rw.z)};
        const float Pd = pre_d + wave_incl_sum(tw);
        float g0 = 0.0f, g1 = 0.0f;
        if (praw >= 0.0f) {
            const float one_over = (1.0f + e1) * fast_rcp(1.0f + e0);
            const float suffix = gC.x * (Cs.x - P.x) + gC.y * (Cs.y - P.y) +
                                 gC.z * (Cs.z - P.z) + gD * (Ds - Pd);
            const float da = Tj * here - suffix * one_over;
            g0 = da * s * e0 * inv1;
            g1 = -da * s * (e1 * inv1) * ((1.0f + e0) * inv1);
        }
        const float g1_up = __shfl_up(g1, 1, kWave);
        const float g_sdf = g0 + (lane == 0 ? g1_prev : g1_up);
        if (act) {
            float gq7[7] = {0, 0, 0, 0, 0, 0, 0};
            quadric_sdf_grad(load_quadric(qgrid, cp0.cell), cp0.local, g_sdf, inv_reso, gq7);
            det_put_quadric(o, 2 * q, (uint32_t)cp0.cell, gq7);
            if (!(need && wt != 0.0f)) {
                const float zero7[7] = {0, 0, 0, 0, 0, 0, 0};
                det_put_quadric(o, 2 * q + 1, o.q_invalid, zero7);
                o.akey[q] = 8u << o.cbits;
                o.aidx[q] = q;
            } else {
                const f3 grgb{gC.x * wt, gC.y * wt, gC.z * wt};
                float dd[C];
                const f3 gn = shade_vjp<NC>(ev.normal, ray.dir, d, ev.spec, grgb, dd);
                float gq6[7] = {0, 0, 0, 0, 0, 0, 0};
                quadric_normal_grad(load_quadric(qgrid, ev.rp.cell), ev.rp.local, gn, gq6);
                det_put_quadric(o, 2 * q + 1, (uint32_t)ev.rp.cell, gq6);
                const uint32_t cls = (uint32_t)((ev.rp.ix & 1) | ((ev.rp.iy & 1) << 1) |
                                                ((ev.rp.iz & 1) << 2));
                o.akey[q] = (cls << o.cbits) | (uint32_t)ev.rp.cell;
                o.aidx[q] = q;
                float4 *dst = reinterpret_cast<float4 *>(o.aval + (int64_t)q * AP);
                dst[0] = make_float4(ev.rp.local.x, ev.rp.local.y, ev.rp.local.z, 0.0f);
#pragma unroll
                for (int c = 0; c < (AP - 4) / 4; ++c) {
                    const int cc = 4 * c;
                    dst[1 + c] = make_float4(dd[cc], cc + 1 < C ? dd[cc + 1] : 0.0f,
                                             cc + 2 < C ? dd[cc + 2] : 0.0f,
                                             cc + 3 < C ? dd[cc + 3] : 0.0f);
                }
            }
        }
        // carries into the next chunk: the last lane of this one (lane 63
        // unless this is the ray's last chunk)
        const int last = (int)min((uint32_t)(kWave - 1), ray.renders - 1 - c0);
        T *= chunk;
        e_prev = __shfl(e1, last, kWave);
        g1_prev = __shfl(g1, last, kWave);
        pre = f3{__shfl(P.x, last, kWave), __shfl(P.y, last, kWave), __shfl(P.z, last, kWave)};
        pre_d = __shfl(Pd, last, kWave);
    }
    if (lane == 0) {  // the last sdf sample: only alpha_{renders-1}'s term
        const uint32_t q = base + ray.renders;
        const CellPoint cpl = ray_sdf_point(ray, ray.renders, interval, half_reso, R);
        float gq7[7] = {0, 0, 0, 0, 0, 0, 0};
        quadric_sdf_grad(load_quadric(qgrid, cpl.cell), cpl.local, g1_prev, inv_reso, gq7);
        det_put_quadric(o, 2 * q, (uint32_t)cpl.cell, gq7);
        const float zero7[7] = {0, 0, 0, 0, 0, 0, 0};
        det_put_quadric(o, 2 * q + 1, o.q_invalid, zero7);
        o.akey[q] = 8u << o.cbits;
        o.aidx[q] = q;
    }
}


// ------------------------------------------------ tile-reduced adjoint ----
// One wavefront per block.  At every step j the 64 rays of the tile touch a
// small box of cells (typically 2x2x2 to 3x3x3).  The appearance gradient
// of the step is a small dense product: box[v][ch] = sum over lanes l of
// W[l][v] * G[l][ch], where G[l] is lane l's gradient of its interpolated
// appearance row and W[l][v] its trilinear weight of box vertex v (8
// non-zeros per lane).  G and W are staged transposed in LDS and multiplied
// on the matrix cores in float32 (v_mfma_f32_16x16x4_f32); the result tile
// is flushed straight from the accumulators with one global atomic per
// touched (vertex, channel), 16 consecutive channels of 4 vertices per
// instruction.  (The previous scheme -- per-lane LDS read-modify-write of
// 8 x C words per sample -- spent most of its time in LDS traffic; one
// global atomic per sample and channel was ~40x slower still.)  The 13
// quadric gradients per sample (7 at the sdf sample's cell, 6 at the render
// sample's) are carried per lane while its ray stays in a cell, summed over
// runs of equal cells along each 8-lane pixel row (DPP scans) and added
// with LDS atomics into a 4x4x4-cell window
// that goes to the grid with global atomics only when the rays leave it:
// per-lane LDS atomics without the merge serialised on the same words, and
// per-step global atomics held up the next step's gathers.  Steps whose
// appearance box exceeds kCapV vertices (about 4% of the
// uniform steps at 800x800, 12% of proposal-guided steps, mostly silhouette
// tiles) are split into rounds of boxes that fit: slabs of the tight box
// along its longest axis, or 3x3x3-cell boxes (4x4x4 = 64 vertices).
constexpr int kCapV = 64;
static_assert(kCapV >= 64, "a fallback round's 3x3x3-cell box has 64 vertices");

// One step of a segmented inclusive scan over the 8-lane row: lane x adds
// the partial sum of lane x - D unless a run starts in (x - D, x].
template <int N, int D>
__device__ __forceinline__ void run_scan_step(float v[7], bool &f, int x) {
    float p[N];
#pragma unroll
    for (int k = 0; k < N; ++k) p[k] = dpp_f<0x110 + D>(v[k]);
    const bool pf = dpp_i<0x110 + D>((int)f) != 0;
    if (x >= D && !f) {
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] += p[k];
        f = pf;
    }
}

// Sums the first N of v[] over runs of equal `key` (key < 0: no value)
// along each 8-lane pixel row.  Returns the key whose sum this lane now
// holds in v[] (the last lane of each run), or -1.  Equal keys in different
// rows stay separate sums (at most one per row): the window adds them with
// LDS atomics, which measured faster than merging them across rows with a
// shuffle butterfly.  Must be called by all 64 lanes.
template <int N>
__device__ __forceinline__ int row_run_merge(float v[7], int key, int lane) {
    const int x = lane & 7;
    const int left = dpp_i<0x111>(key);
    const bool head = x == 0 || key < 0 || left != key;
    const bool next_head = dpp_i<0x101>((int)head) != 0;
    bool f = head;
    run_scan_step<N, 1>(v, f, x);
    run_scan_step<N, 2>(v, f, x);
    run_scan_step<N, 4>(v, f, x);
    return key >= 0 && (x == 7 || next_head) ? key : -1;
}

// Quadric gradient box: a 4x4x4-cell window of the quadric gradient grid
// held in LDS by one wave ([coefficient][cell], 7 x 64 floats).  The merged
// quadric sums of the tile's rays add into it with LDS float atomics; when
// a sum falls outside, the window is written back with one global atomic
// per nonzero entry and moved to put that cell at (1,1,1).  A step's global atomics -- whose completion
// the next step's gathers would wait on -- become rare window write-backs.
constexpr int kQBox = 4;
constexpr int kQBoxCells = kQBox * kQBox * kQBox;
static_assert(kQBoxCells == kWave, "one window cell per lane");

struct QuadricBox {
    int bx = 0, by = 0, bz = 0;
    bool live = false;
};

__device__ __forceinline__ void qbox_writeback(float *__restrict__ qb, QuadricBox &b, int R,
                                               int lane, float *__restrict__ gq) {
    if (!b.live) return;
    const int cx = b.bx + (lane & 3), cy = b.by + ((lane >> 2) & 3), cz = b.bz + (lane >> 4);
    const bool in = cx >= 0 && cy >= 0 && cz >= 0 && cx < R && cy < R && cz < R;
    const int64_t cell = (int64_t)cx + (int64_t)cy * R + (int64_t)cz * R * R;
    wave_lds_order();
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const float v = qb[k * kWave + lane];
        qb[k * kWave + lane] = 0.0f;
        if (in && v != 0.0f) atomic_addf(gq + 7 * cell + k, v);
    }
    wave_lds_order();
    b.live = false;
}

// Adds v[] (the merged sum of cell `tail` at (tx, ty, tz), tail < 0: none)
// to the window.  Must be called by all 64 lanes.
__device__ __forceinline__ void qbox_add(float *__restrict__ qb, QuadricBox &b, const float v[7],
                                         int tail, int tx, int ty, int tz, int R, int lane,
                                         float *__restrict__ gq) {
    const bool has = tail >= 0;
    bool in = has && b.live && (unsigned)(tx - b.bx) < (unsigned)kQBox &&
              (unsigned)(ty - b.by) < (unsigned)kQBox && (unsigned)(tz - b.bz) < (unsigned)kQBox;
    const unsigned long long out = __ballot(has && !in);
    if (out != 0) {
        const int leader = __ffsll((long long)out) - 1;
        const int lx = __builtin_amdgcn_readlane(tx, leader);
        const int ly = __builtin_amdgcn_readlane(ty, leader);
        const int lz = __builtin_amdgcn_readlane(tz, leader);
        qbox_writeback(qb, b, R, lane, gq);
        b.bx = lx - 1;
        b.by = ly - 1;
        b.bz = lz - 1;
        b.live = true;
        in = has && (unsigned)(tx - b.bx) < (unsigned)kQBox &&
             (unsigned)(ty - b.by) < (unsigned)kQBox && (unsigned)(tz - b.bz) < (unsigned)kQBox;
    }
    if (in) {
        // LDS float atomics: the row merge leaves at most one lane per cell
        // and pixel row
        const int slot = (tx - b.bx) + kQBox * (ty - b.by) + kQBox * kQBox * (tz - b.bz);
#pragma unroll
        for (int k = 0; k < 7; ++k)
            if (v[k] != 0.0f)
                __hip_atomic_fetch_add(qb + k * kWave + slot, v[k], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
    } else if (has) {  // spread wider than the window: straight to the grid
        float *dst = gq + 7 * (int64_t)tail;
#pragma unroll
        for (int k = 0; k < 7; ++k)
            if (v[k] != 0.0f) atomic_addf(dst + k, v[k]);
    }
    wave_lds_order();
}

// Per-lane carry of quadric gradients along the ray: contributions to the
// lane's current cell are summed in registers, and the sum goes to the
// window when the ray moves on to another cell (the samples of a
// ray walk its cells in order, so a cell it has left does not come back;
// were it to, it would only get two separate additions).  Dense samples --
// several per cell, as the proposal places them -- then cost one merge
// per cell instead of one per sample.
struct QuadricCarry {
    int cell = -1;
    int ix = 0, iy = 0, iz = 0;
    float acc[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
};

// Hands the carried sums of the lanes in `flush` to the window (merged
// first over equal cells).  Must be called by all 64 lanes.
__device__ __forceinline__ void carry_flush(QuadricCarry &q, bool flush, float *__restrict__ qb,
                                            QuadricBox &b, int R, int lane,
                                            float *__restrict__ gq) {
    float f[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) f[k] = flush ? q.acc[k] : 0.0f;
    const int tail = row_run_merge<7>(f, flush ? q.cell : -1, lane);
    // a merged group's cell is its lane's own carried cell
    qbox_add(qb, b, f, tail, q.ix, q.iy, q.iz, R, lane, gq);
}

// Adds the first N of g[] at cell `cp` to the carry of the active lanes.
// Must be called by all 64 lanes.
template <int N>
__device__ __forceinline__ void carry_add(QuadricCarry &q, bool active, const CellPoint &cp,
                                          const float g[7], float *__restrict__ qb,
                                          QuadricBox &b, int R, int lane,
                                          float *__restrict__ gq) {
    const bool flush = active && q.cell >= 0 && q.cell != cp.cell;
    if (__ballot(flush) != 0) carry_flush(q, flush, qb, b, R, lane, gq);
    if (active) {
        if (q.cell != cp.cell) {
            q.cell = cp.cell;
            q.ix = cp.ix;
            q.iy = cp.iy;
            q.iz = cp.iz;
#pragma unroll
            for (int k = 0; k < 7; ++k) q.acc[k] = 0.0f;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) q.acc[k] += g[k];
    }
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Four K steps of a 16x16 float32 tile product: lane t supplies
// A[t % 16][t / 16] and B[t / 16][t % 16] of each 16x16x4 step and holds
// D[4 (t / 16) + i][t % 16] for i < 4.
__device__ __forceinline__ floatx4 mfma1(float a, float b, floatx4 d) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, d, 0, 0, 0);
}

__device__ __forceinline__ floatx4 mfma4(float4 a, float4 b, floatx4 d) {
    d = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.y, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, b.z, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, b.w, d, 0, 0, 0);
    return d;
}

// LDS layout of the transposed operands.  Gradient rows: [channel][64
// lanes + 4 pad] (row stride a multiple of 4 floats, so rows stay float4
// aligned and 16 consecutive rows start on 16 different bank groups).
// Weights: [vertex][32 lanes of one half-wave] with the float4 column groups
// XOR-swizzled by the row, element (row, col) at
// row * 32 + (col ^ 4 ((row / 2) % 8)), which keeps the fragment reads of 16
// consecutive rows on 16 different bank groups with no padding.
constexpr int kGradStride = kWave + 4;
constexpr int kHalf = kWave / 2;
__device__ __forceinline__ int swz(int row, int col) {
    return row * kHalf + (col ^ (((row >> 1) & 7) << 2));
}

template <int NC>
constexpr int tile_channels() {
    return (3 * NC + 5 + 15) / 16 * 16;  // appearance channels padded to 16-wide tiles
}

// Gradient rows staged in LDS at a time: all of them up to 32 channels; 32
// (one weight staging's two 16-channel blocks) above, so that the 53-channel
// adjoint fits 18.3 KB of LDS per wave (8 waves per CU) instead of 27 KB.
template <int NC>
constexpr int tile_grad_rows() {
    return tile_channels<NC>() < 32 ? tile_channels<NC>() : 32;
}

// At most 256 registers per lane (2 waves per SIMD, 8 per CU, which the
// 18.25 KiB of LDS per wave also allows: 146 of the CU's 160 KiB): occupancy
// hides the latency of the per-sample gathers.
template <int NC, bool LISTED>
__global__ void __launch_bounds__(kWave, 2) fused_bwd_tile_kernel(
    const qsdvr_camera *__restrict__ cams, int n_cams, uint32_t W, uint32_t H,
    const int32_t *__restrict__ ray_pixel,
    int64_t n_rays, const float *__restrict__ qgrid, const float *__restrict__ data, int R,
    float interval, float s, const float *__restrict__ color, const float *__restrict__ depth,
    const float *__restrict__ gcolor, const float *__restrict__ gdepth,
    float *__restrict__ gq, float *__restrict__ gdata, const float4 *__restrict__ ray_state,
    const float *__restrict__ tlist, int K, const uint32_t *__restrict__ tile_order) {
    constexpr int C = 3 * NC + 5;
    constexpr int CP = tile_channels<NC>();
    constexpr int GR = tile_grad_rows<NC>();
    extern __shared__ float4 lds4[];
    float *gt = reinterpret_cast<float *>(lds4);  // [GR][68]: gradient rows, channel-major
    float *wt = gt + GR * kGradStride;            // [kCapV][32]: trilinear weights, vertex-major
    float *qb = wt + kCapV * kHalf;               // [7][64]: quadric gradient window
    const int lane = threadIdx.x;
    for (int i = lane; i < GR * kGradStride + kCapV * kHalf + 7 * kQBoxCells; i += kWave)
        gt[i] = 0.0f;

    // tiles in scattered order: measured 7% faster here (12% on listed
    // samples) than raster order; the forward keeps raster order, whose
    // neighbouring tiles share their staged rows in L2.  With a tile order
    // (qsdvr_fused_backward_ordered) the tiles run in that order instead.
    RaySetup ray = setup_ray<1, true>(cams, n_cams, W, H, ray_pixel, n_rays, interval, ray_state,
                                      LISTED ? tlist : nullptr, K, tile_order);
    if constexpr (!LISTED) ray.tl = nullptr;
    const bool live = ray.out >= 0 && ray.renders > 0;
    const int steps = wave_max(live ? (int)ray.renders + 1 : 0);
    if (steps == 0) return;
    __syncthreads();

    const float half_reso = (float)R / 2.0f, inv_reso = 1.0f / (float)R;
    const int64_t RR = (int64_t)R * R;
    f3 Cs{0, 0, 0}, gC{0, 0, 0};
    float Ds = 0.0f, gD = 0.0f;
    CellPoint cp0, cp1;
    cp0.ix = cp0.iy = cp0.iz = 0;
    cp0.cell = 0;
    float e0 = 1.0f;
    if (live) {
        Cs = load3(color, ray.out);
        gC = load3(gcolor, ray.out);
        Ds = depth ? depth[ray.out] : 0.0f;
        gD = gdepth ? gdepth[ray.out] : 0.0f;
    }
    // QAHEAD: the quadrics of sdf samples j and j + 1 stay in registers, the
    // latter loaded a step ahead of its use, so neither the step's alpha nor
    // the sdf sample's gradient waits on a quadric load.  (53 channels have
    // no registers to spare for them: loaded where used.)
    constexpr bool QAHEAD = C == 32;
    Quadric qs0, qs1;
    TWindow<LISTED && C == 32> tw;
    tw.start(ray, tlist);
    if (live) {
        cp0 = ray_sdf_point_t(ray, 0, tw.t0(ray, 0), interval, half_reso, R);
        qs0 = load_quadric(qgrid, cp0.cell);
        if constexpr (QAHEAD) {
            cp1 = ray_sdf_point_t(ray, 1, tw.t1(ray, 0), interval, half_reso, R);
            qs1 = load_quadric(qgrid, cp1.cell);
        }
        e0 = expf(-s * quadric_sdf(qs0, cp0.local, inv_reso));
    }
    float T = 1.0f;
    f3 pre{0.0f, 0.0f, 0.0f};
    float pre_d = 0.0f;
    float carry = 0.0f;
    QuadricCarry qcarry;
    QuadricBox qbox;

    for (int j = 0; j < steps; ++j) {
        const bool act_r = live && j < (int)ray.renders;   // render sample j
        const bool act_s = live && j <= (int)ray.renders;  // sdf sample j (final gradient)
        float dd[C], w[8];
        float gs7[7] = {0, 0, 0, 0, 0, 0, 0}, gn7[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < C; ++c) dd[c] = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) w[k] = 0.0f;
        CellPoint rp = cp0;
        float g_sdf0 = carry, g_sdf1 = 0.0f;
        bool app = false;  // render sample j of nonzero weight: normal + appearance gradients
        if (act_r) {
            float sdf1;  // sdf sample j + 1
            if constexpr (QAHEAD)
                sdf1 = quadric_sdf(qs1, cp1.local, inv_reso);
            else
                sdf1 = sdf_at_t(ray, j + 1, tw.t1(ray, j), interval, half_reso, R, inv_reso, qgrid,
                                cp1);
            float e1 = expf(-s * sdf1);
            float inv1 = fast_rcp(1.0f + e1);
            float praw = (e1 - e0) * inv1;
            float a = fmaxf(praw, 0.0f);
            const float wt = a * T;
            // the render sample is evaluated only where it adds to the prefix
            // or to dL/dalpha; a zero-weight sample has zero gradients.  (Its
            // vertex rows are gathered per lane: staging them in LDS as the
            // forward does measured slower here, at 2 waves per SIMD.)
            if (praw >= 0.0f && T != 0.0f) {
                float d[C];
                SampleEval ev;
                eval_render_sample_t<NC>(ray, j, tw.mid(ray, j), interval, half_reso, R, qgrid,
                                         data, w, d, ev);
                pre.x += ev.rgb.x * wt;
                pre.y += ev.rgb.y * wt;
                pre.z += ev.rgb.z * wt;
                pre_d += ev.t * wt;
                float one_over = (1.0f + e1) * fast_rcp(1.0f + e0);
                float suffix = gC.x * (Cs.x - pre.x) + gC.y * (Cs.y - pre.y) +
                               gC.z * (Cs.z - pre.z) + gD * (Ds - pre_d);
                float here = gC.x * ev.rgb.x + gC.y * ev.rgb.y + gC.z * ev.rgb.z + gD * ev.t;
                float da = T * here - suffix * one_over;
                g_sdf0 += da * s * e0 * inv1;
                g_sdf1 = -da * s * (e1 * inv1) * ((1.0f + e0) * inv1);
                if (wt != 0.0f) {
                    f3 grgb{gC.x * wt, gC.y * wt, gC.z * wt};
                    f3 gn = shade_vjp<NC>(ev.normal, ray.dir, d, ev.spec, grgb, dd);
                    quadric_normal_grad(load_quadric(qgrid, ev.rp.cell), ev.rp.local, gn, gn7);
                    rp = ev.rp;
                    app = true;
                } else {
#pragma unroll
                    for (int k = 0; k < 8; ++k) w[k] = 0.0f;
                }
            } else if (praw >= 0.0f) {
                // T == 0: dL/dalpha = -suffix / (1 - alpha)
                float one_over = (1.0f + e1) * fast_rcp(1.0f + e0);
                float suffix = gC.x * (Cs.x - pre.x) + gC.y * (Cs.y - pre.y) +
                               gC.z * (Cs.z - pre.z) + gD * (Ds - pre_d);
                float da = -suffix * one_over;
                g_sdf0 += da * s * e0 * inv1;
                g_sdf1 = -da * s * (e1 * inv1) * ((1.0f + e0) * inv1);
            }
            T = T * (1.0f - a);
            e0 = e1;
        }
        tw.advance(j);
        const CellPoint sc = cp0;  // sdf sample j
        if (act_s)
            quadric_sdf_grad(QAHEAD ? qs0 : load_quadric(qgrid, sc.cell), sc.local, g_sdf0,
                             inv_reso, gs7);
        if (act_r) {
            carry = g_sdf1;
            cp0 = cp1;
            if constexpr (QAHEAD) qs0 = qs1;
        }
        if constexpr (QAHEAD) {
            // sdf sample j + 2, for the next step, requested by every lane (past
            // its ray's end a lane loads a clamped cell's quadric and never uses
            // it): the load then writes the loop-carried registers themselves,
            // and the step's end does not wait for it (a lane-masked load is
            // merged into them with copies that wait for every outstanding
            // load and atomic of the step)
            const CellPoint cn =
                ray_sdf_point_t(ray, j + 2, tw.t1(ray, j + 1), interval, half_reso, R);
            if (act_r && (uint32_t)(j + 2) <= ray.renders) cp1 = cn;
            qs1 = load_quadric(qgrid, cn.cell);
        }

        // quadric gradients of the sdf sample and of the render sample
        carry_add<7>(qcarry, act_s, sc, gs7, qb, qbox, R, lane, gq);
        if (__ballot(app) == 0) continue;  // no appearance or normal gradient this step
        carry_add<6>(qcarry, app, rp, gn7, qb, qbox, R, lane, gq);
        // stage this step's gradient row once, transposed: gt[c][lane]
        // (more than 32 channels: 32 at a time, below)
        if constexpr (GR == CP) {
#pragma unroll
            for (int c = 0; c < C; ++c) gt[c * kGradStride + lane] = dd[c];
        }
        bool pend_r = app;  // appearance gradient not yet accumulated
        // Rounds: the tight box of all pending cells when it fits in kCapV
        // vertices (the common case: one round); otherwise a part of it that
        // fits (below).  Every round takes at least one pending lane, so the
        // loop ends.
        for (;;) {
            int lx = INT_MAX, ly = INT_MAX, lz = INT_MAX, hx = INT_MIN, hy = INT_MIN, hz = INT_MIN;
            if (pend_r) {
                lx = rp.ix; ly = rp.iy; lz = rp.iz;
                hx = rp.ix; hy = rp.iy; hz = rp.iz;
            }
            lx = wave_min_s(lx); ly = wave_min_s(ly); lz = wave_min_s(lz);
            hx = wave_max_s(hx); hy = wave_max_s(hy); hz = wave_max_s(hz);
            if (hx < lx) break;  // nothing pending (wave-uniform)
            const int cx = hx - lx + 1, cy = hy - ly + 1, cz = hz - lz + 1;  // cells
            // the tight box holds every pending lane: this round is the last
            const bool whole = (cx + 1) * (cy + 1) * (cz + 1) <= kCapV;
            if (!whole) {
                // Oversize: the slab of the tight box at its low end along the
                // longest axis whose cross-section leaves room for two vertex
                // layers, with as many cell layers as fit; else the 3x3x3-cell
                // box around the first pending lane.  Either holds a pending lane.
                const int sx = (cy + 1) * (cz + 1), sy = (cx + 1) * (cz + 1),
                          sz = (cx + 1) * (cy + 1);
                const bool fx = 2 * sx <= kCapV, fy = 2 * sy <= kCapV, fz = 2 * sz <= kCapV;
                if (fz && cz >= (fx ? cx : 0) && cz >= (fy ? cy : 0)) {
                    hz = lz + kCapV / sz - 2;
                } else if (fy && cy >= (fx ? cx : 0)) {
                    hy = ly + kCapV / sy - 2;
                } else if (fx) {
                    hx = lx + kCapV / sx - 2;
                } else {
                    const unsigned long long m = __ballot(pend_r);
                    const int leader = __ffsll((long long)m) - 1;
                    const int ax = __shfl(rp.ix, leader, kWave);
                    const int ay = __shfl(rp.iy, leader, kWave);
                    const int az = __shfl(rp.iz, leader, kWave);
                    lx = ax - 1; hx = ax + 1;
                    ly = ay - 1; hy = ay + 1;
                    lz = az - 1; hz = az + 1;
                }
            }
            const int ex = hx - lx + 2, ey = hy - ly + 2, ez = hz - lz + 2;
            const int sxy = ex * ey;
            const bool in_r = pend_r && rp.ix >= lx && rp.ix <= hx && rp.iy >= ly &&
                              rp.iy <= hy && rp.iz >= lz && rp.iz <= hz;
            const int vb = (rp.ix - lx) + (rp.iy - ly) * ex + (rp.iz - lz) * sxy;
            const int nv = sxy * ez;
            int wofs[8];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                wofs[k] = swz(vb + (k & 1) + ((k >> 1) & 1) * ex + (k >> 2) * sxy, lane & (kHalf - 1));
            // box-vertex offsets in the global grids, vertex v held by lane v
            int vofs = 0;
            if (lane < nv) {
                // exact quotients (lane < 64, ex >= 2, sxy >= 4), as in stage_box
                const int vz = (int)(((float)lane + 0.5f) * fast_rcp((float)sxy));
                const int r2 = lane - vz * sxy;
                const int vy = (int)(((float)r2 + 0.5f) * fast_rcp((float)ex)), vx = r2 - vy * ex;
                vofs = vx + vy * R + vz * (int)RR;
            }
            const int64_t vbase = (int64_t)lx + (int64_t)ly * R + (int64_t)lz * RR;
            float *const vrow = gdata + vbase * C;  // the box's first vertex row
            wave_lds_order();
            // Appearance: box[v][ch] = sum over lanes l of W[l][v] * G[l][ch],
            // a (box vertices x 64) x (64 x C) product on the matrix cores
            // (v_mfma_f32_16x16x4_f32, float32 throughout), one 16-channel
            // block at a time.  K runs over the lanes in two halves (the
            // weight buffer holds one half-wave's columns); within a half,
            // K block kb of the 16x16x4 tile takes lanes {kb, 8 + kb,
            // 16 + kb, 24 + kb}, so every lane reads its 8 operands of a
            // tile as 2 float4.
            if (__ballot(in_r) != 0) {
                // lane id re-read here so that the operand addresses below are
                // not hoisted out of the step loop (they would stay live
                // through the register-heavy sample evaluation)
                int ln;
                __asm__ volatile("v_mov_b32 %0, %1" : "=v"(ln) : "v"(lane));
                const int g = ln >> 4, m = ln & 15;
                const int k0 = (2 * g) ^ ((ln >> 1) & 7), k1 = k0 ^ 1;
                // wave-uniform: the row-block tests below branch on a scalar
                const int nmb = __builtin_amdgcn_readfirstlane((nv + 15) >> 4);
                // NBG channel blocks share one staging of the weight columns
                constexpr int NBG = CP / 16 >= 2 ? 2 : 1;
                static_assert(GR == CP || GR == 16 * NBG, "one gradient-row staging per weight staging");
#pragma unroll
                for (int nb0 = 0; nb0 < CP / 16; nb0 += NBG) {
                    if constexpr (GR < CP) {
                        // channels 16 nb0 .. 16 nb0 + 31 of the gradient rows (the
                        // previous ones' operand reads are issued before)
                        wave_lds_order();
#pragma unroll
                        for (int c = 0; c < GR; ++c)
                            gt[c * kGradStride + lane] =
                                16 * nb0 + c < C ? dd[16 * nb0 + c < C ? 16 * nb0 + c : 0] : 0.0f;
                        wave_lds_order();
                    }
                    floatx4 acc[NBG][4];
#pragma unroll
                    for (int q = 0; q < NBG; ++q)
#pragma unroll
                        for (int mb = 0; mb < 4; ++mb) acc[q][mb] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
                    // K in two halves: the weight columns of lanes 0-31, then 32-63
                    for (int h = 0; h < 2; ++h) {
                        const bool mine = in_r && (lane >> 5) == h;
                        if (mine) {
#pragma unroll
                            for (int k = 0; k < 8; ++k) wt[wofs[k]] = w[k];
                        }
                        wave_lds_order();
                        // gradient fragments of the NBG channel blocks, then
                        // each weight fragment read once for all of them
                        float4 b0[NBG], b1[NBG];
#pragma unroll
                        for (int q = 0; q < NBG; ++q) {
                            const float4 *sb = reinterpret_cast<const float4 *>(
                                gt + (16 * (GR < CP ? q : nb0 + q) + m) * kGradStride +
                                kHalf * h + 8 * g);
                            b0[q] = sb[0];
                            b1[q] = sb[1];
                        }
#pragma unroll
                        for (int mb = 0; mb < 4; ++mb) {
                            if (mb < nmb) {
                                const float4 *sa =
                                    reinterpret_cast<const float4 *>(wt + (16 * mb + m) * kHalf);
                                const float4 a0 = sa[k0], a1 = sa[k1];
                                // the NBG accumulation chains interleaved
#pragma unroll
                                for (int q = 0; q < NBG; ++q) acc[q][mb] = mfma1(a0.x, b0[q].x, acc[q][mb]);
#pragma unroll
                                for (int q = 0; q < NBG; ++q) acc[q][mb] = mfma1(a0.y, b0[q].y, acc[q][mb]);
#pragma unroll
                                for (int q = 0; q < NBG; ++q) acc[q][mb] = mfma1(a0.z, b0[q].z, acc[q][mb]);
#pragma unroll
                                for (int q = 0; q < NBG; ++q) acc[q][mb] = mfma1(a0.w, b0[q].w, acc[q][mb]);
#pragma unroll
                                for (int q = 0; q < NBG; ++q) acc[q][mb] = mfma1(a1.x, b1[q].x, acc[q][mb]);
#pragma unroll
                                for (int q = 0; q < NBG; ++q) acc[q][mb] = mfma1(a1.y, b1[q].y, acc[q][mb]);
#pragma unroll
                                for (int q = 0; q < NBG; ++q) acc[q][mb] = mfma1(a1.z, b1[q].z, acc[q][mb]);
#pragma unroll
                                for (int q = 0; q < NBG; ++q) acc[q][mb] = mfma1(a1.w, b1[q].w, acc[q][mb]);
                            }
                        }
                        wave_lds_order();
                        if (mine) {
#pragma unroll
                            for (int k = 0; k < 8; ++k) wt[wofs[k]] = 0.0f;
                        }
                    }
#pragma unroll
                    for (int mb = 0; mb < 4; ++mb) {
                        if (mb < nmb) {
                            // the block's 4 vertex offsets fetched together,
                            // ahead of its atomics
                            int vo[4];
#pragma unroll
                            for (int i = 0; i < 4; ++i) vo[i] = __shfl(vofs, 16 * mb + 4 * g + i, kWave);
#pragma unroll
                            for (int i = 0; i < 4; ++i) {
                                // one branch per touched vertex, its channel
                                // blocks' atomics addressed from one 32-bit offset
                                const int v = 16 * mb + 4 * g + i;
                                bool touched = false;
#pragma unroll
                                for (int q = 0; q < NBG; ++q) touched |= acc[q][mb][i] != 0.0f;
                                if (v < nv && touched) {
                                    const uint32_t o = (uint32_t)(vo[i] * C + 16 * nb0 + m);
#pragma unroll
                                    for (int q = 0; q < NBG; ++q)
                                        if (16 * (nb0 + q) + m < C)
                                            atomic_addf(vrow + o + 16 * q, acc[q][mb][i]);
                                }
                            }
                        }
                    }
                }
            }
            wave_lds_order();
            if (whole) break;
            pend_r = pend_r && !in_r;
        }
    }
    carry_flush(qcarry, qcarry.cell >= 0, qb, qbox, R, lane, gq);  // the last carried cells
    qbox_writeback(qb, qbox, R, lane, gq);
}

template <int NC>
constexpr size_t tile_lds_bytes() {
    return (size_t)(tile_grad_rows<NC>() * kGradStride + kCapV * kHalf + 7 * kQBoxCells) *
           sizeof(float);
}

template <int NC>
constexpr size_t fwd_lds_bytes() {
    return 3 * NC + 5 == 32 ? (size_t)kWavesPerBlock * kBoxV * ((3 * NC + 5) / 4 + 1) * 16
                            : (size_t)kWavesPerBlock * kBoxV53 * kRow53 * 4;
}

static int64_t fused_waves(uint32_t W, uint32_t H, int n_cams, const int32_t *ray_pixel,
                           int64_t n_rays) {
    if (ray_pixel == nullptr) {
        int64_t tx = (W + 7) / 8, ty = (H + 7) / 8;
        return tx * ty * n_cams;
    }
    return (n_rays + kWave - 1) / kWave;
}

}  // namespace qsdvr

using namespace qsdvr;

static int fused_check(const qsdvr_camera *cams, int32_t n_cams, int32_t width, int32_t height,
                       const int32_t *ray_pixel, int64_t n_rays, const float *sdf_grid,
                       const float *data_grid, int32_t reso, int32_t channels, float interval) {
    QSDVR_REQUIRE(cams && sdf_grid && data_grid, "qsdvr_fused: null pointer argument");
    QSDVR_REQUIRE(n_cams > 0 && n_rays >= 0, "qsdvr_fused: bad camera / ray counts");
    QSDVR_REQUIRE(width > 0 && height > 0, "qsdvr_fused: empty image");
    QSDVR_REQUIRE(reso > 0 && reso <= 1290, "qsdvr_fused: reso %d out of range", reso);
    QSDVR_REQUIRE(channels == 32 || channels == 53, "qsdvr_fused: channels must be 32 or 53");
    QSDVR_REQUIRE(interval > 0.0f, "qsdvr_fused: interval must be > 0");
    QSDVR_REQUIRE((int64_t)n_cams * width * height <= 0x7FFFFFFFll,
                  "qsdvr_fused: n_cams*W*H exceeds the 31-bit ray code");
    if (ray_pixel == nullptr)
        QSDVR_REQUIRE(n_rays == (int64_t)n_cams * width * height,
                      "qsdvr_fused: n_rays must be n_cams*W*H when ray_pixel is NULL");
    return QSDVR_OK;
}

// Per-ray state for the listed-sample and proposal paths: (origin, count
// as bits), (dir, 0) -- the RayCut result of every ray, bit for bit.
__global__ void __launch_bounds__(256) ray_state_kernel(const qsdvr_camera *__restrict__ cams,
                                                        int n_cams, uint32_t W, uint32_t H,
                                                        const int32_t *__restrict__ ray_pixel,
                                                        int64_t n_rays, float interval,
                                                        float4 *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_rays) return;
    int64_t code = ray_pixel ? (int64_t)(uint32_t)ray_pixel[i] : i;
    const int cam = (int)(code / ((int64_t)W * H));
    float4 a = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
    float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (cam < n_cams && cams[cam].width == W && cams[cam].height == H) {
        const uint32_t pix = (uint32_t)(code - (int64_t)cam * W * H);
        RayCutResult cut = ray_cut_pixel(cams[cam], pix, interval);
        a = make_float4(cut.origin.x, cut.origin.y, cut.origin.z, __uint_as_float(cut.count));
        b = make_float4(cut.dir.x, cut.dir.y, cut.dir.z, 0.0f);
    }
    out[2 * i] = a;
    out[2 * i + 1] = b;
}

// Compositing weights of the plain uniform renderer summed into 64 equal
// bins per ray, from the SDF alone (the proposal's training targets): one
// wavefront per ray, lanes over 64 consecutive render samples at a time.
// alpha_j needs sdf samples j and j+1 (the next lane's; lane 63 evaluates its
// own), T_j is the carried transmittance times an exclusive product scan of
// (1 - alpha) over the lanes, and each lane adds w_j = alpha_j T_j to its
// bin with an LDS atomic.  Bins as in fused_fwd_kernel<.., true>; T is
// multiplied in scan order instead of sample order, so the weights agree
// with the sequential walk to float32 rounding.
__global__ void __launch_bounds__(256) plain_bins_kernel(const float4 *__restrict__ ray_state,
                                                         int64_t n_rays,
                                                         const float *__restrict__ qgrid, int R,
                                                         float interval, float s,
                                                         float *__restrict__ bins) {
    __shared__ float lb[256 / kWave][64];
    const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
    const float half_reso = (float)R / 2.0f, inv_reso = 1.0f / (float)R;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / kWave) + wv; r < n_rays; r += nwaves) {
        const float4 a4 = ray_state[2 * r], b4 = ray_state[2 * r + 1];
        const uint32_t count = __float_as_uint(a4.w);
        lb[wv][lane] = 0.0f;
        __builtin_amdgcn_wave_barrier();
        if (count >= 2) {
            const f3 ori{a4.x, a4.y, a4.z}, dir{b4.x, b4.y, b4.z};
            const uint32_t renders = count - 1;
            float T = 1.0f;
            for (uint32_t c0 = 0; c0 < renders; c0 += kWave) {
                const uint32_t j = c0 + lane;
                const bool valid = j < renders;
                float e = 0.0f, e_next = 0.0f;
                if (j <= renders) {
                    CellPoint cp = sdf_sample(ori, dir, j, interval, half_reso, R);
                    clamp_cell(cp, R);
                    e = expf(-s * quadric_sdf(load_quadric(qgrid, cp.cell), cp.local, inv_reso));
                }
                e_next = __shfl_down(e, 1, kWave);
                if (lane == kWave - 1 && j + 1 <= renders) {
                    CellPoint cp = sdf_sample(ori, dir, j + 1, interval, half_reso, R);
                    clamp_cell(cp, R);
                    e_next = expf(-s * quadric_sdf(load_quadric(qgrid, cp.cell), cp.local, inv_reso));
                }
                const float a = valid ? fmaxf((e_next - e) * fast_rcp(1.0f + e_next), 0.0f) : 0.0f;
                // inclusive product scan of (1 - a) over the lanes
                float p = 1.0f - a;
#pragma unroll
                for (int o = 1; o < kWave; o <<= 1) {
                    const float u = __shfl_up(p, o, kWave);
                    if (lane >= o) p *= u;
                }
                const float up = __shfl_up(p, 1, kWave);  // every lane shuffles
                const float excl = lane == 0 ? 1.0f : up;
                const float w = a * (T * excl);
                if (valid && w != 0.0f) {
                    const int b = min((int)(((float)j + 0.5f) / (float)renders * 64.0f), 63);
                    __hip_atomic_fetch_add(&lb[wv][b], w, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                T *= __shfl(p, kWave - 1, kWave);
            }
        }
        __builtin_amdgcn_wave_barrier();
        bins[r * 64 + lane] = lb[wv][lane];
        __builtin_amdgcn_wave_barrier();
    }
}

static int fused_forward_impl(const qsdvr_camera *cams, int32_t n_cams, int32_t width,
                              int32_t height, const int32_t *ray_pixel, int64_t n_rays,
                              const float *ray_state, const float *tlist, int32_t K,
                              const float *sdf_grid, const float *data_grid, int32_t reso,
                              int32_t channels, float interval, float logistic_coef,
                              float *color, float *depth, float *bins, hipStream_t s,
                              uint32_t *tile_cost = nullptr) {
    int st = fused_check(cams, n_cams, width, height, ray_pixel, n_rays, sdf_grid, data_grid,
                         reso, channels, interval);
    if (st != QSDVR_OK) return st;
    QSDVR_REQUIRE(color != nullptr, "qsdvr_fused_forward: null color");
    QSDVR_REQUIRE(tlist == nullptr || (ray_state != nullptr && K >= 2 && K <= 4096),
                  "qsdvr_fused_forward: listed samples need the ray state and 2 <= K <= 4096");
    if (n_rays == 0) return QSDVR_OK;
    const float4 *rs = reinterpret_cast<const float4 *>(ray_state);
    int64_t waves = fused_waves(width, height, n_cams, ray_pixel, n_rays);
    unsigned blocks = (unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
#define QSDVR_FWD(NC, BINS, LISTED, COST)                                                    \
    fused_fwd_kernel<NC, BINS, LISTED, COST><<<blocks, kFusedBlock, fwd_lds_bytes<NC>(), s>>>( \
        cams, n_cams, (uint32_t)width, (uint32_t)height, ray_pixel, n_rays, sdf_grid, data_grid, \
        reso, interval, logistic_coef, color, depth, rs, tlist, K, bins, tile_cost)
#define QSDVR_FWD_NC(NC)                                  \
    do {                                                  \
        if (tile_cost != nullptr) {                       \
            if (tlist != nullptr) QSDVR_FWD(NC, false, true, true); \
            else QSDVR_FWD(NC, false, false, true);       \
        } else if (tlist != nullptr) {                    \
            if (bins) QSDVR_FWD(NC, true, true, false);   \
            else QSDVR_FWD(NC, false, true, false);       \
        } else {                                          \
            if (bins) QSDVR_FWD(NC, true, false, false);  \
            else QSDVR_FWD(NC, false, false, false);      \
        }                                                 \
    } while (0)
    if (channels == 32) QSDVR_FWD_NC(9);
    else QSDVR_FWD_NC(16);
#undef QSDVR_FWD_NC
#undef QSDVR_FWD
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

static int fused_backward_impl(const qsdvr_camera *cams, int32_t n_cams, int32_t width,
                               int32_t height, const int32_t *ray_pixel, int64_t n_rays,
                               const float *ray_state, const float *tlist, int32_t K,
                               const float *sdf_grid, const float *data_grid, int32_t reso,
                               int32_t channels, float interval, float logistic_coef,
                               const float *color, const float *depth, const float *grad_color,
                               const float *grad_depth, float *sdf_grid_grad,
                               float *data_grid_grad, int32_t scatter_mode, hipStream_t s,
                               const uint32_t *tile_order = nullptr) {
    int st = fused_check(cams, n_cams, width, height, ray_pixel, n_rays, sdf_grid, data_grid,
                         reso, channels, interval);
    if (st != QSDVR_OK) return st;
    QSDVR_REQUIRE(color && grad_color && sdf_grid_grad && data_grid_grad,
                  "qsdvr_fused_backward: null pointer argument");
    QSDVR_REQUIRE((grad_depth == nullptr) || (depth != nullptr),
                  "qsdvr_fused_backward: grad_depth needs the saved depth");
    QSDVR_REQUIRE(scatter_mode == 0 || scatter_mode == 1,
                  "qsdvr_fused_backward: scatter_mode must be 0 or 1");
    QSDVR_REQUIRE(tlist == nullptr || (ray_state != nullptr && K >= 2 && K <= 4096),
                  "qsdvr_fused_backward: listed samples need the ray state and 2 <= K <= 4096");
    if (n_rays == 0) return QSDVR_OK;
    const float4 *rs = reinterpret_cast<const float4 *>(ray_state);
    const float *dsave = grad_depth ? depth : nullptr;
    int64_t waves = fused_waves(width, height, n_cams, ray_pixel, n_rays);
    if (scatter_mode == 1) {
        unsigned blocks = (unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
        if (channels == 32)
            fused_bwd_kernel<9><<<blocks, kFusedBlock, 0, s>>>(
                cams, n_cams, (uint32_t)width, (uint32_t)height, ray_pixel, n_rays, sdf_grid,
                data_grid, reso, interval, logistic_coef, color, dsave, grad_color, grad_depth,
                sdf_grid_grad, data_grid_grad, rs, tlist, K);
        else
            fused_bwd_kernel<16><<<blocks, kFusedBlock, 0, s>>>(
                cams, n_cams, (uint32_t)width, (uint32_t)height, ray_pixel, n_rays, sdf_grid,
                data_grid, reso, interval, logistic_coef, color, dsave, grad_color, grad_depth,
                sdf_grid_grad, data_grid_grad, rs, tlist, K);
        QSDVR_LAUNCHED();
        return QSDVR_OK;
    }
    unsigned blocks = (unsigned)waves;
#define QSDVR_TILE(NC, LISTED)                                                               \
    fused_bwd_tile_kernel<NC, LISTED><<<blocks, kWave, tile_lds_bytes<NC>(), s>>>(           \
        cams, n_cams, (uint32_t)width, (uint32_t)height, ray_pixel, n_rays, sdf_grid,        \
        data_grid, reso, interval, logistic_coef, color, dsave, grad_color, grad_depth,      \
        sdf_grid_grad, data_grid_grad, rs, tlist, K, tile_order)
    if (channels == 32) {
        if (tlist != nullptr) QSDVR_TILE(9, true);
        else QSDVR_TILE(9, false);
    } else {
        if (tlist != nullptr) QSDVR_TILE(16, true);
        else QSDVR_TILE(16, false);
    }
#undef QSDVR_TILE
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

extern "C" {

int qsdvr_fused_forward(const qsdvr_camera *cams, int32_t n_cams, int32_t width, int32_t height,
                        const int32_t *ray_pixel, int64_t n_rays, const float *sdf_grid,
                        const float *data_grid, int32_t reso, int32_t channels, float interval,
                        float logistic_coef, float *color, float *depth, qsdvr_stream stream) {
    return fused_forward_impl(cams, n_cams, width, height, ray_pixel, n_rays, nullptr, nullptr, 0,
                              sdf_grid, data_grid, reso, channels, interval, logistic_coef, color,
                              depth, nullptr, as_stream(stream));
}

int qsdvr_fused_forward_bins(const qsdvr_camera *cams, int32_t n_cams, int32_t width,
                             int32_t height, const int32_t *ray_pixel, int64_t n_rays,
                             const float *sdf_grid, const float *data_grid, int32_t reso,
                             int32_t channels, float interval, float logistic_coef, float *color,
                             float *depth, float *bins, qsdvr_stream stream) {
    QSDVR_REQUIRE(bins != nullptr, "qsdvr_fused_forward_bins: null bins");
    return fused_forward_impl(cams, n_cams, width, height, ray_pixel, n_rays, nullptr, nullptr, 0,
                              sdf_grid, data_grid, reso, channels, interval, logistic_coef, color,
                              depth, bins, as_stream(stream));
}

int qsdvr_fused_backward(const qsdvr_camera *cams, int32_t n_cams, int32_t width, int32_t height,
                         const int32_t *ray_pixel, int64_t n_rays, const float *sdf_grid,
                         const float *data_grid, int32_t reso, int32_t channels, float interval,
                         float logistic_coef, const float *color, const float *depth,
                         const float *grad_color, const float *grad_depth, float *sdf_grid_grad,
                         float *data_grid_grad, int32_t scatter_mode, qsdvr_stream stream) {
    return fused_backward_impl(cams, n_cams, width, height, ray_pixel, n_rays, nullptr, nullptr, 0,
                               sdf_grid, data_grid, reso, channels, interval, logistic_coef,
                               color, depth, grad_color, grad_depth, sdf_grid_grad,
                               data_grid_grad, scatter_mode, as_stream(stream));
}

int qsdvr_ray_state(const qsdvr_camera *cams, int32_t n_cams, int32_t width, int32_t height,
                    const int32_t *ray_pixel, int64_t n_rays, float interval, float *ray_state,
                    qsdvr_stream stream) {
    QSDVR_REQUIRE(cams && ray_state, "qsdvr_ray_state: null pointer argument");
    QSDVR_REQUIRE(n_cams > 0 && width > 0 && height > 0 && n_rays >= 0 && interval > 0.0f,
                  "qsdvr_ray_state: bad sizes");
    QSDVR_REQUIRE(ray_pixel != nullptr || n_rays == (int64_t)n_cams * width * height,
                  "qsdvr_ray_state: n_rays must be n_cams*W*H when ray_pixel is NULL");
    if (n_rays == 0) return QSDVR_OK;
    ray_state_kernel<<<grid_1d(n_rays, 256), 256, 0, as_stream(stream)>>>(
        cams, n_cams, (uint32_t)width, (uint32_t)height, ray_pixel, n_rays, interval,
        reinterpret_cast<float4 *>(ray_state));
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_plain_bins(const float *ray_state, int64_t n_rays, const float *sdf_grid,
                     int32_t reso, float interval, float logistic_coef, float *bins,
                     qsdvr_stream stream) {
    QSDVR_REQUIRE(n_rays >= 0 && reso > 0 && reso <= 1290 && interval > 0.0f,
                  "qsdvr_plain_bins: bad sizes");
    if (n_rays == 0) return QSDVR_OK;
    QSDVR_REQUIRE(ray_state && sdf_grid && bins, "qsdvr_plain_bins: null pointer argument");
    const int64_t blocks = std::min<int64_t>((n_rays + 3) / 4, 8192);
    plain_bins_kernel<<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(
        reinterpret_cast<const float4 *>(ray_state), n_rays, sdf_grid, reso, interval,
        logistic_coef, bins);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_fused_forward_listed(const qsdvr_camera *cams, int32_t n_cams, int32_t width,
                               int32_t height, const int32_t *ray_pixel, int64_t n_rays,
                               const float *ray_state, const float *tlist, int32_t K,
                               const float *sdf_grid, const float *data_grid, int32_t reso,
                               int32_t channels, float interval, float logistic_coef,
                               float *color, float *depth, qsdvr_stream stream) {
    QSDVR_REQUIRE(ray_state != nullptr, "qsdvr_fused_forward_listed: null ray state");
    return fused_forward_impl(cams, n_cams, width, height, ray_pixel, n_rays, ray_state, tlist, K,
                              sdf_grid, data_grid, reso, channels, interval, logistic_coef, color,
                              depth, nullptr, as_stream(stream));
}

int qsdvr_fused_backward_listed(const qsdvr_camera *cams, int32_t n_cams, int32_t width,
                                int32_t height, const int32_t *ray_pixel, int64_t n_rays,
                                const float *ray_state, const float *tlist, int32_t K,
                                const float *sdf_grid, const float *data_grid, int32_t reso,
                                int32_t channels, float interval, float logistic_coef,
                                const float *color, const float *depth, const float *grad_color,
                                const float *grad_depth, float *sdf_grid_grad,
                                float *data_grid_grad, int32_t scatter_mode,
                                qsdvr_stream stream) {
    QSDVR_REQUIRE(ray_state != nullptr, "qsdvr_fused_backward_listed: null ray state");
    return fused_backward_impl(cams, n_cams, width, height, ray_pixel, n_rays, ray_state, tlist,
                               K, sdf_grid, data_grid, reso, channels, interval, logistic_coef,
                               color, depth, grad_color, grad_depth, sdf_grid_grad,
                               data_grid_grad, scatter_mode, as_stream(stream));
}

// ------------------------------------------ cost-ordered tile backward ----
// The tile adjoint's time per 8x8 tile varies about 4x across a frame (tiles
// whose appearance boxes overflow need several rounds per step), and a launch
// in a fixed order leaves its heaviest tiles to finish last.  The costed
// forward records each tile's expected backward cost; the ordered backward
// sorts the tiles by it (heaviest first) and runs them in that order.

}  // extern "C"

namespace {

__global__ void iota_kernel(uint32_t *__restrict__ v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

size_t order_temp_bytes(int64_t n) {
    size_t t = 0;
    (void)rocprim::radix_sort_pairs_desc(nullptr, t, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                         (const uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n,
                                         0u, 32u, (hipStream_t)0);
    return t;
}

struct OrderBuffers {
    uint32_t *keys_out, *ids, *order;
    void *temp;
    size_t temp_bytes;
};

size_t order_carve(void *ws, size_t bytes, int64_t n, OrderBuffers &b) {
    Carver c{static_cast<char *>(ws), bytes};
    b.keys_out = c.take<uint32_t>(n);
    b.ids = c.take<uint32_t>(n);
    b.order = c.take<uint32_t>(n);
    b.temp_bytes = order_temp_bytes(n);
    b.temp = c.take<char>(b.temp_bytes);
    return c.used + 256;
}

int64_t frame_tiles(int32_t n_cams, int32_t width, int32_t height) {
    return (int64_t)((width + 7) / 8) * ((height + 7) / 8) * n_cams;
}

}  // namespace

extern "C" {

size_t qsdvr_fused_order_workspace_size(int32_t n_cams, int32_t width, int32_t height) {
    if (n_cams <= 0 || width <= 0 || height <= 0) return 0;
    OrderBuffers b;
    return order_carve(nullptr, 0, frame_tiles(n_cams, width, height), b);
}

int qsdvr_fused_forward_costed(const qsdvr_camera *cams, int32_t n_cams, int32_t width,
                               int32_t height, const float *ray_state, const float *tlist,
                               int32_t K, const float *sdf_grid, const float *data_grid,
                               int32_t reso, int32_t channels, float interval,
                               float logistic_coef, float *color, float *depth,
                               uint32_t *tile_cost, qsdvr_stream stream) {
    QSDVR_REQUIRE(tile_cost != nullptr, "qsdvr_fused_forward_costed: null tile_cost");
    QSDVR_REQUIRE(n_cams > 0 && width > 0 && height > 0,
                  "qsdvr_fused_forward_costed: bad frame size");
    return fused_forward_impl(cams, n_cams, width, height, nullptr, (int64_t)n_cams * width * height,
                              ray_state, tlist, K, sdf_grid, data_grid, reso, channels, interval,
                              logistic_coef, color, depth, nullptr, as_stream(stream), tile_cost);
}

int qsdvr_fused_backward_ordered(const qsdvr_camera *cams, int32_t n_cams, int32_t width,
                                 int32_t height, const float *ray_state, const float *tlist,
                                 int32_t K, const float *sdf_grid, const float *data_grid,
                                 int32_t reso, int32_t channels, float interval,
                                 float logistic_coef, const float *color, const float *depth,
                                 const float *grad_color, const float *grad_depth,
                                 float *sdf_grid_grad, float *data_grid_grad,
                                 const uint32_t *tile_cost, void *workspace,
                                 size_t workspace_bytes, qsdvr_stream stream) {
    QSDVR_REQUIRE(tile_cost != nullptr, "qsdvr_fused_backward_ordered: null tile_cost");
    QSDVR_REQUIRE(n_cams > 0 && width > 0 && height > 0,
                  "qsdvr_fused_backward_ordered: bad frame size");
    const int64_t n = frame_tiles(n_cams, width, height);
    OrderBuffers b;
    const size_t need = order_carve(nullptr, 0, n, b);
    if (workspace == nullptr || workspace_bytes < need) {
        set_error("qsdvr_fused_backward_ordered: workspace of %zu bytes is too small (%zu needed)",
                  workspace_bytes, need);
        return QSDVR_EWORKSPACE;
    }
    order_carve(workspace, workspace_bytes, n, b);
    hipStream_t s = as_stream(stream);
    iota_kernel<<<grid_1d(n, 256), 256, 0, s>>>(b.ids, n);
    QSDVR_LAUNCHED();
    size_t tb = b.temp_bytes;
    QSDVR_HIP(rocprim::radix_sort_pairs_desc(b.temp, tb, tile_cost, b.keys_out, b.ids, b.order,
                                             (size_t)n, 0u, 32u, s));
    return fused_backward_impl(cams, n_cams, width, height, nullptr, (int64_t)n_cams * width * height,
                               ray_state, tlist, K, sdf_grid, data_grid, reso, channels, interval,
                               logistic_coef, color, depth, grad_color, grad_depth, sdf_grid_grad,
                               data_grid_grad, 0, s, b.order);
}

size_t qsdvr_fused_det_workspace_size(int64_t n_rays, int64_t sample_capacity, int32_t channels,
                                      int32_t reso) {
    if (n_rays < 0 || sample_capacity < 0 || reso < 1 || reso > 512) return 0;
    return det_workspace_bytes(n_rays, sample_capacity, channels, reso);
}

int qsdvr_fused_backward_det(const qsdvr_camera *cams, int32_t n_cams, int32_t width,
                             int32_t height, const int32_t *ray_pixel, int64_t n_rays,
                             const float *ray_state, const float *tlist, int32_t K,
                             const float *sdf_grid, const float *data_grid, int32_t reso,
                             int32_t channels, float interval, float logistic_coef,
                             const float *color, const float *depth, const float *grad_color,
                             const float *grad_depth, float *sdf_grid_grad,
                             float *data_grid_grad, int64_t sample_capacity, void *workspace,
                             size_t workspace_bytes, qsdvr_stream stream) {
    int st = fused_check(cams, n_cams, width, height, ray_pixel, n_rays, sdf_grid, data_grid,
                         reso, channels, interval);
    if (st != QSDVR_OK) return st;
    QSDVR_REQUIRE(ray_state != nullptr, "qsdvr_fused_backward_det: null ray state");
    QSDVR_REQUIRE(color && grad_color && sdf_grid_grad && data_grid_grad,
                  "qsdvr_fused_backward_det: null pointer argument");
    QSDVR_REQUIRE((grad_depth == nullptr) || (depth != nullptr),
                  "qsdvr_fused_backward_det: grad_depth needs the saved depth");
    QSDVR_REQUIRE(tlist == nullptr || (K >= 2 && K <= 4096),
                  "qsdvr_fused_backward_det: listed samples need 2 <= K <= 4096");
    QSDVR_REQUIRE(reso <= 512, "qsdvr_fused_backward_det: reso %d above 512", reso);
    QSDVR_REQUIRE(sample_capacity >= 0 && n_rays < 0xFFFFFFFFll,
                  "qsdvr_fused_backward_det: bad sizes");
    if (n_rays == 0) return QSDVR_OK;
    hipStream_t s = as_stream(stream);
    DetBuffers b;
    st = det_carve(workspace, workspace_bytes, n_rays, sample_capacity, channels, reso, b);
    if (st != QSDVR_OK) return st;
    const float4 *rs = reinterpret_cast<const float4 *>(ray_state);
    int64_t total = 0;
    st = det_offsets(b, rs, tlist ? K : 0, s, &total);
    if (st != QSDVR_OK) return st;
    if (total == 0) return QSDVR_OK;
    const uint32_t R3 = (uint32_t)((uint64_t)reso * reso * reso);
    DetRecordOut o{b.offsets, b.qkey, b.qidx, b.qval, b.akey, b.aidx, b.aval, R3, b.cbits};
    const float *dsave = grad_depth ? depth : nullptr;
    const int64_t waves = fused_waves(width, height, n_cams, ray_pixel, n_rays);
    const unsigned blocks = (unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    if (ray_pixel != nullptr && tlist == nullptr) {
        // listed rays with uniform samples: one wavefront per ray
        const unsigned rblocks = (unsigned)((n_rays + kWavesPerBlock - 1) / kWavesPerBlock);
        if (channels == 32)
            fused_bwd_record_ray_kernel<9><<<rblocks, kFusedBlock, 0, s>>>(
                cams, n_cams, (uint32_t)width, (uint32_t)height, ray_pixel, n_rays, sdf_grid,
                data_grid, reso, interval, logistic_coef, color, dsave, grad_color, grad_depth, rs,
                o);
        else
            fused_bwd_record_ray_kernel<16><<<rblocks, kFusedBlock, 0, s>>>(
                cams, n_cams, (uint32_t)width, (uint32_t)height, ray_pixel, n_rays, sdf_grid,
                data_grid, reso, interval, logistic_coef, color, dsave, grad_color, grad_depth, rs,
                o);
    } else if (channels == 32)
        fused_bwd_record_kernel<9><<<blocks, kFusedBlock, 0, s>>>(
            cams, n_cams, (uint32_t)width, (uint32_t)height, ray_pixel, n_rays, sdf_grid,
            data_grid, reso, interval, logistic_coef, color, dsave, grad_color, grad_depth, rs,
            tlist, K, o);
    else
        fused_bwd_record_kernel<16><<<blocks, kFusedBlock, 0, s>>>(
            cams, n_cams, (uint32_t)width, (uint32_t)height, ray_pixel, n_rays, sdf_grid,
            data_grid, reso, interval, logistic_coef, color, dsave, grad_color, grad_depth, rs,
            tlist, K, o);
    QSDVR_LAUNCHED();
    return det_reduce(b, total, sdf_grid_grad, data_grid_grad, s);
}

}  // extern "C"
