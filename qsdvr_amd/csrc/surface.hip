// Quadric surface grid for gfx950: LayerToGrid forward/backward and the
// SampleSDF / SampleNormal ops with their adjoints.
// Reference semantics: kernel/surface.cu:3-375, models/render_torch.py:30-100.
#include "quadric.h"

namespace qsdvr {

// ------------------------------------------------------- LayerToGrid ----
// Recurrences of kernel/surface.cu:161-178 (== render_torch.py:39-75), one
// thread per axis, float32 in the reference's order (contraction off):
//   d[0] = off, d[i] = (2 a[i-1] + 2 a[i]) + d[i-1]
//   c[0] = 0,   c[i] = ((3 a[i-1] + a[i]) + 2 d[i-1]) + c[i-1]
__global__ void layer_recurrence_kernel(const float *__restrict__ layers,  // x|y|z|off
                                        float *__restrict__ rec, int R) {
#pragma clang fp contract(off)
    int axis = threadIdx.x;
    if (axis >= 3) return;
    const float *a = layers + axis * R;
    float *d = rec + axis * R;          // X3 / Y4 / Z5
    float *c = rec + (3 + axis) * R;    // X6 / Y6 / Z6
    d[0] = layers[3 * R + axis];
    c[0] = 0.0f;
    for (int i = 1; i < R; ++i) d[i] = (2.0f * a[i - 1] + 2.0f * a[i]) + d[i - 1];
    for (int i = 1; i < R; ++i) c[i] = ((3.0f * a[i - 1] + a[i]) + 2.0f * d[i - 1]) + c[i - 1];
}

// One thread per output float (coalesced stores of the [R^3,7] grid).
__global__ void __launch_bounds__(256) layer_broadcast_kernel(const float *__restrict__ layers,
                                                              const float *__restrict__ rec,
                                                              float *__restrict__ out, int R) {
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t total = (int64_t)R * R * R * 7;
    if (id >= total) return;
    int64_t cell = id / 7;
    int k = (int)(id - cell * 7);
    int x = (int)(cell % R), y = (int)((cell / R) % R), z = (int)(cell / ((int64_t)R * R));
    float v;
    switch (k) {
        case 0: v = layers[x]; break;
        case 1: v = layers[R + y]; break;
        case 2: v = layers[2 * R + z]; break;
        case 3: v = rec[x]; break;
        case 4: v = rec[R + y]; break;
        case 5: v = rec[2 * R + z]; break;
        default: v = rec[3 * R + x] + rec[4 * R + y] + rec[5 * R + z] + layers[3 * R + 3]; break;
    }
    out[id] = v;
}

// Backward pass A: one block per z slice.  For every (z,y) row the row is
// staged in LDS; per-x sums over y of k = 0,3,6 stay in registers, per-row
// sums over x of k = 1,4,6 are block-reduced, per-slice sums of k = 2,5,6
// are reduced at the end.  Fixed summation order => deterministic.
constexpr int kL2GBlock = 256;
constexpr int kMaxReso = 512;

__device__ __forceinline__ float block_sum(float v, float *red) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    int wave = threadIdx.x / kWave;
    __syncthreads();
    if (threadIdx.x % kWave == 0) red[wave] = v;
    __syncthreads();
    float s = 0.0f;
    for (int w = 0; w < kL2GBlock / kWave; ++w) s += red[w];
    return s;
}

__global__ void __launch_bounds__(kL2GBlock) layer_grad_slice_kernel(
    const float *__restrict__ grad, int R, float *__restrict__ part_x,  // [R z][R x][3]
    float *__restrict__ part_y,                                         // [R z][R y][3]
    float *__restrict__ sum_z) {                                        // [R z][3]
    extern __shared__ float row[];  // 7R floats
    __shared__ float red[kL2GBlock / kWave];
    const int z = blockIdx.x;
    float ax[2][3] = {{0, 0, 0}, {0, 0, 0}};
    float az[3] = {0, 0, 0};
    for (int y = 0; y < R; ++y) {
        const float *src = grad + ((int64_t)z * R + y) * R * 7;
        __syncthreads();
        for (int i = threadIdx.x; i < 7 * R; i += kL2GBlock) row[i] = src[i];
        __syncthreads();
        float ay[3] = {0, 0, 0};
        for (int h = 0; h < 2; ++h) {
            int x = threadIdx.x + h * kL2GBlock;
            if (x < R) {
                const float *g = row + 7 * x;
                ax[h][0] += g[0];
                ax[h][1] += g[3];
                ax[h][2] += g[6];
                ay[0] += g[1];
                ay[1] += g[4];
                ay[2] += g[6];
                az[0] += g[2];
                az[1] += g[5];
                az[2] += g[6];
            }
        }
        for (int k = 0; k < 3; ++k) {
            float s = block_sum(ay[k], red);
            if (threadIdx.x == 0) part_y[((int64_t)z * R + y) * 3 + k] = s;
        }
    }
    for (int h = 0; h < 2; ++h) {
        int x = threadIdx.x + h * kL2GBlock;
        if (x < R)
            for (int k = 0; k < 3; ++k) part_x[((int64_t)z * R + x) * 3 + k] = ax[h][k];
    }
    for (int k = 0; k < 3; ++k) {
        float s = block_sum(az[k], red);
        if (threadIdx.x == 0) sum_z[z * 3 + k] = s;
    }
}

// Backward pass B (single block): finish the per-axis sums over z, run the
// reverse recurrences of kernel/surface.cu:301-334 and write the gradients
// of the layers and of offset (offset[3] = sum of all G gradients).
__global__ void __launch_bounds__(kL2GBlock) layer_grad_finish_kernel(
    const float *__restrict__ part_x, const float *__restrict__ part_y,
    const float *__restrict__ sum_z, int R, float *__restrict__ out) {  // x|y|z|off
    extern __shared__ float sh[];  // [3 axes][3 kinds][R]
    for (int i = threadIdx.x; i < 3 * R; i += kL2GBlock) {
        int v = i / 3, k = i % 3;  // v = x or y index, k = kind (A, D, G)
        float sx = 0.0f, sy = 0.0f;
        for (int z = 0; z < R; ++z) {
            sx += part_x[((int64_t)z * R + v) * 3 + k];
            sy += part_y[((int64_t)z * R + v) * 3 + k];
        }
        sh[(0 * 3 + k) * R + v] = sx;
        sh[(1 * 3 + k) * R + v] = sy;
        sh[(2 * 3 + k) * R + v] = sum_z[v * 3 + k];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int axis = threadIdx.x;
        float *ga = sh + (axis * 3 + 0) * R;
        float *gd = sh + (axis * 3 + 1) * R;
        float *gc = sh + (axis * 3 + 2) * R;
        for (int i = R - 1; i > 0; --i) {
            ga[i] += gc[i];
            ga[i - 1] += 3.0f * gc[i];
            gd[i - 1] += 2.0f * gc[i];
            gc[i - 1] += gc[i];
        }
        for (int i = R - 1; i > 0; --i) {
            ga[i] += 2.0f * gd[i];
            ga[i - 1] += 2.0f * gd[i];
            gd[i - 1] += gd[i];
        }
        out[3 * R + axis] = gd[0];
    }
    if (threadIdx.x == 3) {
        float s = 0.0f;
        for (int z = 0; z < R; ++z) s += sum_z[z * 3 + 2];
        out[3 * R + 3] = s;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * R; i += kL2GBlock) {
        int axis = i / R, v = i % R;
        out[axis * R + v] = sh[(axis * 3 + 0) * R + v];
    }
}

// --------------------------------------------------- SDF and normals ----
// SampleSDF / SampleNormal forward: SPL samples per lane, lane-strided so
// that every load instruction stays coalesced; the indices of all SPL
// samples are loaded first, then their quadrics and points, so a wave makes
// two memory round trips for SPL samples.
constexpr int kSampleSPL = 4;

__global__ void __launch_bounds__(256) sample_sdf_fwd_kernel(float *__restrict__ out,
                                                             const float *__restrict__ grid,
                                                             const float *__restrict__ pts,
                                                             const int32_t *__restrict__ idx,
                                                             int64_t n, int64_t cells,
                                                             float inv_reso) {
    const int64_t base = xcd_block() * blockDim.x * kSampleSPL + threadIdx.x;
    // every lane loads its indices (a lane past the end reads the last one
    // and marks it invalid afterwards): a lane-masked load would be merged
    // with a wait before the next one is issued
    int64_t c[kSampleSPL];
#pragma unroll
    for (int u = 0; u < kSampleSPL; ++u) {
        const int64_t i = base + (int64_t)u * blockDim.x;
        c[u] = (int64_t)idx[i < n ? i : n - 1];
    }
    // the quadrics and points likewise (an invalid sample reads cell 0 and the
    // last point, and gets NaN below): a masked block also pulls the first
    // arithmetic on its loads in with a wait
    Quadric q[kSampleSPL];
    f3 p[kSampleSPL];
#pragma unroll
    for (int u = 0; u < kSampleSPL; ++u) {
        const int64_t i = base + (int64_t)u * blockDim.x;
        if (i >= n) c[u] = -1;
        q[u] = load_quadric(grid, c[u] >= 0 && c[u] < cells ? c[u] : 0);
        p[u] = load3(pts, i < n ? i : n - 1);
    }
#pragma unroll
    for (int u = 0; u < kSampleSPL; ++u) {
        const int64_t i = base + (int64_t)u * blockDim.x;
        if (i >= n) continue;
        out[i] = (c[u] < 0 || c[u] >= cells) ? __builtin_nanf("")
                                             : quadric_sdf(q[u], p[u], inv_reso);
    }
}

__global__ void __launch_bounds__(256) sample_sdf_bwd_kernel(const float *__restrict__ up,
                                                             const float *__restrict__ grid,
                                                             const float *__restrict__ pts,
                                                             const int32_t *__restrict__ idx,
                                                             float *__restrict__ ggrid, int64_t n,
                                                             int64_t cells, float inv_reso) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t c = idx[i];
    if (c < 0 || c >= cells) return;
    float g[7] = {0, 0, 0, 0, 0, 0, 0};
    quadric_sdf_grad(load_quadric(grid, c), load3(pts, i), up[i], inv_reso, g);
    float *dst = ggrid + 7 * c;
#pragma unroll
    for (int k = 0; k < 7; ++k) atomic_addf(dst + k, g[k]);
}

__global__ void __launch_bounds__(256) sample_normal_fwd_kernel(float *__restrict__ out,
                                                                const float *__restrict__ grid,
                                                                const float *__restrict__ pts,
                                                                const int32_t *__restrict__ idx,
                                                                int64_t n, int64_t cells) {
    const int64_t base = xcd_block() * blockDim.x * kSampleSPL + threadIdx.x;
    // every lane loads its indices (a lane past the end reads the last one
    // and marks it invalid afterwards): a lane-masked load would be merged
    // with a wait before the next one is issued
    int64_t c[kSampleSPL];
#pragma unroll
    for (int u = 0; u < kSampleSPL; ++u) {
        const int64_t i = base + (int64_t)u * blockDim.x;
        c[u] = (int64_t)idx[i < n ? i : n - 1];
    }
    // the quadrics and points likewise (an invalid sample reads cell 0 and the
    // last point, and gets NaN below): a masked block also pulls the first
    // arithmetic on its loads in with a wait
    Quadric q[kSampleSPL];
    f3 p[kSampleSPL];
#pragma unroll
    for (int u = 0; u < kSampleSPL; ++u) {
        const int64_t i = base + (int64_t)u * blockDim.x;
        if (i >= n) c[u] = -1;
        q[u] = load_quadric(grid, c[u] >= 0 && c[u] < cells ? c[u] : 0);
        p[u] = load3(pts, i < n ? i : n - 1);
    }
#pragma unroll
    for (int u = 0; u < kSampleSPL; ++u) {
        const int64_t i = base + (int64_t)u * blockDim.x;
        if (i >= n) continue;
        f3 r;
        if (c[u] < 0 || c[u] >= cells) {
            const float nan = __builtin_nanf("");
            r = f3{nan, nan, nan};
        } else {
            r = quadric_normal(q[u], p[u]);
        }
        store3(out, i, r);
    }
}

__global__ void __launch_bounds__(256) sample_normal_bwd_kernel(const float *__restrict__ up,
                                                                const float *__restrict__ grid,
                                                                const float *__restrict__ pts,
                                                                const int32_t *__restrict__ idx,
                                                                float *__restrict__ ggrid,
                                                                int64_t n, int64_t cells) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t c = idx[i];
    if (c < 0 || c >= cells) return;
    float g[7] = {0, 0, 0, 0, 0, 0, 0};
    quadric_normal_grad(load_quadric(grid, c), load3(pts, i), load3(up, i), g);
    float *dst = ggrid + 7 * c;
#pragma unroll
    for (int k = 0; k < 6; ++k) atomic_addf(dst + k, g[k]);
}

}  // namespace qsdvr

using namespace qsdvr;

extern "C" {

size_t qsdvr_layer_to_grid_workspace_size(int32_t reso) {
    if (reso <= 0) return 0;
    size_t R = (size_t)reso;
    // layers (3R+4), recurrences (6R), part_x and part_y (3R^2 each), sum_z (3R)
    return (3 * R + 4 + 6 * R + 6 * R * R + 3 * R) * sizeof(float) + 8 * 256;
}

int qsdvr_layer_to_grid_forward(float *grid_out, const float *x_layer, const float *y_layer,
                                const float *z_layer, const float *offset, int32_t reso,
                                int layers_on_host, void *workspace, size_t workspace_bytes,
                                qsdvr_stream stream) {
    QSDVR_REQUIRE(grid_out && x_layer && y_layer && z_layer && offset,
                  "qsdvr_layer_to_grid_forward: null pointer argument");
    QSDVR_REQUIRE(reso > 0 && reso <= kMaxReso, "qsdvr_layer_to_grid_forward: reso %d not in [1, %d]",
                  reso, kMaxReso);
    Carver cv{(char *)workspace, workspace_bytes};
    float *layers = cv.take<float>(3 * (size_t)reso + 4);
    float *rec = cv.take<float>(6 * (size_t)reso);
    if (!layers || !rec) {
        set_error("qsdvr_layer_to_grid_forward: workspace too small (need %zu bytes)",
                  qsdvr_layer_to_grid_workspace_size(reso));
        return QSDVR_EWORKSPACE;
    }
    hipStream_t s = as_stream(stream);
    hipMemcpyKind kind = layers_on_host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    size_t l = (size_t)reso * sizeof(float);
    QSDVR_HIP(hipMemcpyAsync(layers, x_layer, l, kind, s));
    QSDVR_HIP(hipMemcpyAsync(layers + reso, y_layer, l, kind, s));
    QSDVR_HIP(hipMemcpyAsync(layers + 2 * reso, z_layer, l, kind, s));
    QSDVR_HIP(hipMemcpyAsync(layers + 3 * reso, offset, 4 * sizeof(float), kind, s));
    layer_recurrence_kernel<<<1, 64, 0, s>>>(layers, rec, reso);
    QSDVR_LAUNCHED();
    int64_t total = (int64_t)reso * reso * reso * 7;
    layer_broadcast_kernel<<<grid_1d(total, 256), 256, 0, s>>>(layers, rec, grid_out, reso);
    QSDVR_LAUNCHED();
    if (layers_on_host) QSDVR_HIP(hipStreamSynchronize(s));  // host buffers may be reused
    return QSDVR_OK;
}

int qsdvr_layer_to_grid_backward(const float *grad_grid, float *gx, float *gy, float *gz,
                                 float *goffset, int32_t reso, int outputs_on_host,
                                 void *workspace, size_t workspace_bytes, qsdvr_stream stream) {
    QSDVR_REQUIRE(grad_grid && gx && gy && gz && goffset,
                  "qsdvr_layer_to_grid_backward: null pointer argument");
    QSDVR_REQUIRE(reso > 0 && reso <= kMaxReso,
                  "qsdvr_layer_to_grid_backward: reso %d not in [1, %d]", reso, kMaxReso);
    Carver cv{(char *)workspace, workspace_bytes};
    size_t R = (size_t)reso;
    float *out = cv.take<float>(3 * R + 4);
    float *rec = cv.take<float>(6 * R);  // keeps the layout of the forward workspace
    float *part_x = cv.take<float>(3 * R * R);
    float *part_y = cv.take<float>(3 * R * R);
    float *sum_z = cv.take<float>(3 * R);
    (void)rec;
    if (!out || !part_x || !part_y || !sum_z) {
        set_error("qsdvr_layer_to_grid_backward: workspace too small (need %zu bytes)",
                  qsdvr_layer_to_grid_workspace_size(reso));
        return QSDVR_EWORKSPACE;
    }
    hipStream_t s = as_stream(stream);
    layer_grad_slice_kernel<<<reso, kL2GBlock, 7 * R * sizeof(float), s>>>(grad_grid, reso, part_x,
                                                                          part_y, sum_z);
    QSDVR_LAUNCHED();
    layer_grad_finish_kernel<<<1, kL2GBlock, 9 * R * sizeof(float), s>>>(part_x, part_y, sum_z,
                                                                        reso, out);
    QSDVR_LAUNCHED();
    hipMemcpyKind kind = outputs_on_host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    size_t l = R * sizeof(float);
    QSDVR_HIP(hipMemcpyAsync(gx, out, l, kind, s));
    QSDVR_HIP(hipMemcpyAsync(gy, out + R, l, kind, s));
    QSDVR_HIP(hipMemcpyAsync(gz, out + 2 * R, l, kind, s));
    QSDVR_HIP(hipMemcpyAsync(goffset, out + 3 * R, 4 * sizeof(float), kind, s));
    if (outputs_on_host) QSDVR_HIP(hipStreamSynchronize(s));
    return QSDVR_OK;
}

#define QSDVR_CHECK_SAMPLES(name)                                                        \
    QSDVR_REQUIRE(n >= 0 && cell_count > 0, name ": bad sizes n=%lld cells=%lld",         \
                  (long long)n, (long long)cell_count);                                  \
    if (n == 0) return QSDVR_OK;

int qsdvr_sample_sdf_forward(float *out, const float *sdf_grid, const float *point,
                             const int32_t *index, int64_t n, int64_t cell_count, float reso,
                             qsdvr_stream stream) {
    QSDVR_CHECK_SAMPLES("qsdvr_sample_sdf_forward");
    QSDVR_REQUIRE(out && sdf_grid && point && index, "qsdvr_sample_sdf_forward: null pointer");
    QSDVR_REQUIRE(reso != 0.0f, "qsdvr_sample_sdf_forward: reso must be nonzero");
    sample_sdf_fwd_kernel<<<grid_1d(n, 256 * kSampleSPL), 256, 0, as_stream(stream)>>>(
        out, sdf_grid, point, index, n, cell_count, 1.0f / reso);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_sample_sdf_backward(const float *grad_out, const float *sdf_grid, const float *point,
                              const int32_t *index, float *grad_grid, int64_t n,
                              int64_t cell_count, float reso, qsdvr_stream stream) {
    QSDVR_CHECK_SAMPLES("qsdvr_sample_sdf_backward");
    QSDVR_REQUIRE(grad_out && sdf_grid && point && index && grad_grid,
                  "qsdvr_sample_sdf_backward: null pointer");
    QSDVR_REQUIRE(reso != 0.0f, "qsdvr_sample_sdf_backward: reso must be nonzero");
    sample_sdf_bwd_kernel<<<grid_1d(n, 256), 256, 0, as_stream(stream)>>>(
        grad_out, sdf_grid, point, index, grad_grid, n, cell_count, 1.0f / reso);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_sample_normal_forward(float *out, const float *sdf_grid, const float *point,
                                const int32_t *index, int64_t n, int64_t cell_count,
                                qsdvr_stream stream) {
    QSDVR_CHECK_SAMPLES("qsdvr_sample_normal_forward");
    QSDVR_REQUIRE(out && sdf_grid && point && index, "qsdvr_sample_normal_forward: null pointer");
    sample_normal_fwd_kernel<<<grid_1d(n, 256 * kSampleSPL), 256, 0, as_stream(stream)>>>(
        out, sdf_grid, point, index, n, cell_count);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

int qsdvr_sample_normal_backward(const float *grad_out, const float *sdf_grid,
                                 const float *point, const int32_t *index, float *grad_grid,
                                 int64_t n, int64_t cell_count, qsdvr_stream stream) {
    QSDVR_CHECK_SAMPLES("qsdvr_sample_normal_backward");
    QSDVR_REQUIRE(grad_out && sdf_grid && point && index && grad_grid,
                  "qsdvr_sample_normal_backward: null pointer");
    sample_normal_bwd_kernel<<<grid_1d(n, 256), 256, 0, as_stream(stream)>>>(
        grad_out, sdf_grid, point, index, grad_grid, n, cell_count);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

}  // extern "C"
