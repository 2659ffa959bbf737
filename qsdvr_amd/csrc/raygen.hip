// Ray generation for gfx950: RayCut (per-pixel cut + device exclusive scan)
// and GenerateRayPoints (one wavefront per ray, lane-contiguous stores).
// Reference semantics: kernel/camera.cu:12-146.
#include "raygen.h"

#include <mutex>
#include <string.h>

namespace qsdvr {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// ------------------------------------------------------------ RayCut ----
__global__ void __launch_bounds__(256) ray_cut_kernel(qsdvr_camera cam, float interval,
                                                      uint32_t P, int32_t *__restrict__ ray,
                                                      float *__restrict__ origin,
                                                      float *__restrict__ dir) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= P) return;
    RayCutResult r = ray_cut_pixel(cam, id, interval);
    store3(dir, id, r.dir);
    store3(origin, id, r.origin);
    ray[4 * (int64_t)id] = (int32_t)r.count;
}

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ void masked_counts(const int32_t *ray, const uint8_t *mask, int64_t i,
                                              int64_t P, uint32_t &c, uint32_t &r) {
    c = 0;
    r = 0;
    if (i < P && (mask == nullptr || mask[i])) {
        uint32_t k = (uint32_t)ray[4 * i];
        c = k < 2 ? 0u : k;                 // kernel/camera.cu:117-120
        r = c > 1 ? c - 1 : 0u;             // kernel/camera.cu:124
    }
}

// Inclusive block scan of 64-bit pairs (wave shuffles + one LDS pass).
__device__ __forceinline__ void block_scan2(uint64_t &a, uint64_t &b, uint64_t *lds_a,
                                            uint64_t *lds_b, uint64_t &tot_a, uint64_t &tot_b) {
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        uint64_t ua = __shfl_up(a, o, kWave);
        uint64_t ub = __shfl_up(b, o, kWave);
        if (lane >= o) {
            a += ua;
            b += ub;
        }
    }
    if (lane == kWave - 1) {
        lds_a[wave] = a;
        lds_b[wave] = b;
    }
    __syncthreads();
    uint64_t pa = 0, pb = 0;
    tot_a = 0;
    tot_b = 0;
    for (int w = 0; w < kScanBlock / kWave; ++w) {
        if (w < wave) {
            pa += lds_a[w];
            pb += lds_b[w];
        }
        tot_a += lds_a[w];
        tot_b += lds_b[w];
    }
    a += pa;
    b += pb;
    __syncthreads();
}

__global__ void __launch_bounds__(kScanBlock) scan_reduce_kernel(const int32_t *__restrict__ ray,
                                                                 const uint8_t *__restrict__ mask,
                                                                 int64_t P,
                                                                 uint64_t *__restrict__ sums) {
    __shared__ uint64_t la[kScanBlock / kWave], lb[kScanBlock / kWave];
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint64_t a = 0, b = 0;
    for (int k = 0; k < kScanItems; ++k) {
        uint32_t c, r;
        masked_counts(ray, mask, base + k, P, c, r);
        a += c;
        b += r;
    }
    uint64_t ta, tb;
    block_scan2(a, b, la, lb, ta, tb);
    if (threadIdx.x == 0) {
        sums[2 * blockIdx.x] = ta;
        sums[2 * blockIdx.x + 1] = tb;
    }
}

// Single block: exclusive scan of the per-tile sums in place, totals last.
__global__ void __launch_bounds__(kScanBlock) scan_sums_kernel(uint64_t *__restrict__ sums,
                                                               int64_t nb,
                                                               uint64_t *__restrict__ totals) {
    __shared__ uint64_t la[kScanBlock / kWave], lb[kScanBlock / kWave];
    uint64_t carry_a = 0, carry_b = 0;
    for (int64_t s = 0; s < nb; s += kScanBlock) {
        int64_t i = s + threadIdx.x;
        uint64_t a = i < nb ? sums[2 * i] : 0, b = i < nb ? sums[2 * i + 1] : 0;
        uint64_t ea = a, eb = b;
        uint64_t ta, tb;
        block_scan2(a, b, la, lb, ta, tb);
        if (i < nb) {
            sums[2 * i] = carry_a + a - ea;
            sums[2 * i + 1] = carry_b + b - eb;
        }
        carry_a += ta;
        carry_b += tb;
    }
    if (threadIdx.x == 0) {
        totals[0] = carry_a;
        totals[1] = carry_b;
    }
}

__global__ void __launch_bounds__(kScanBlock) scan_write_kernel(int32_t *__restrict__ ray,
                                                                const uint8_t *__restrict__ mask,
                                                                int64_t P,
                                                                const uint64_t *__restrict__ sums) {
    __shared__ uint64_t la[kScanBlock / kWave], lb[kScanBlock / kWave];
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t c[kScanItems], r[kScanItems];
    uint64_t a = 0, b = 0;
    for (int k = 0; k < kScanItems; ++k) {
        masked_counts(ray, mask, base + k, P, c[k], r[k]);
        a += c[k];
        b += r[k];
    }
    uint64_t ea = a, eb = b, ta, tb;
    block_scan2(a, b, la, lb, ta, tb);
    uint64_t oa = sums[2 * blockIdx.x] + a - ea;
    uint64_t ob = sums[2 * blockIdx.x + 1] + b - eb;
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + k;
        if (i < P && (mask == nullptr || mask[i])) {
            int4 row = make_int4((int32_t)c[k], (int32_t)(uint32_t)oa, (int32_t)r[k],
                                 (int32_t)(uint32_t)ob);
            reinterpret_cast<int4 *>(ray)[i] = row;
        }
        oa += c[k];
        ob += r[k];
    }
}

// --------------------------------------------------- GenerateRayPoints ----
constexpr int kGenBlock = 256;

__global__ void __launch_bounds__(kGenBlock) generate_points_kernel(
    float *__restrict__ sdf_point, int32_t *__restrict__ sdf_index,
    float *__restrict__ render_point, int32_t *__restrict__ render_index,
    float *__restrict__ view_dir, const int32_t *__restrict__ ray,
    const float *__restrict__ origin, const float *__restrict__ dirs, int64_t rays,
    int64_t sdf_total, int64_t render_total, float interval, int R) {
    const int lane = threadIdx.x % kWave;
    const int64_t rid = (int64_t)blockIdx.x * (kGenBlock / kWave) + threadIdx.x / kWave;
    if (rid >= rays) return;
    const int4 info = reinterpret_cast<const int4 *>(ray)[rid];
    const f3 ori = load3(origin, rid);
    const f3 dir = load3(dirs, rid);
    const float half_reso = (float)R / 2.0f;
    const uint32_t sc = (uint32_t)info.x, so = (uint32_t)info.y;
    const uint32_t rc = (uint32_t)info.z, ro = (uint32_t)info.w;
    if ((int64_t)so + sc <= sdf_total) {
        for (uint32_t j = lane; j < sc; j += kWave) {
            CellPoint cp = sdf_sample(ori, dir, j, interval, half_reso, R);
            store3(sdf_point, (int64_t)so + j, cp.local);
            sdf_index[(int64_t)so + j] = cp.cell;
        }
    }
    if ((int64_t)ro + rc <= render_total) {
        for (uint32_t j = lane; j < rc; j += kWave) {
            CellPoint cp = render_sample(ori, dir, j, interval, half_reso, R);
            store3(render_point, (int64_t)ro + j, cp.local);
            render_index[(int64_t)ro + j] = cp.cell;
            store3(view_dir, (int64_t)ro + j, dir);
        }
    }
}

// RayCut's workspace: device copy of the host mask + per-tile scan sums +
// the two totals, each 256-byte aligned.
constexpr size_t ray_cut_workspace_bytes(int64_t pixel_count) {
    return ((size_t)pixel_count + 255) / 256 * 256 +
           (size_t)((pixel_count + kScanTile - 1) / kScanTile + 1) * 16 + 512 + 256;
}

// ------------------------------------------- allocation-free RayCut ----
// qsdvr_ray_cut_noalloc keeps RayCut's workspace in this array instead of a
// caller's buffer: 8.07 MiB, enough for 2^23 pixels (3840x2160 fits), held
// by every GPU the library runs on until the process ends.  Calls on one
// GPU share it, so calls that overlap on one GPU race.  Nothing in it needs
// to start at zero: every word a call reads was written earlier in the
// same call.
constexpr int64_t kNoAllocMaxPixels = int64_t(1) << 23;
constexpr size_t kNoAllocScratchBytes = ray_cut_workspace_bytes(kNoAllocMaxPixels);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (the array's host-side name is not a device address).
int scratch_on_current_device(void **out) {
    int dev = -1;
    QSDVR_HIP(hipGetDevice(&dev));
    QSDVR_REQUIRE(dev >= 0 && dev < kMaxDevices, "qsdvr_ray_cut_noalloc: device %d out of range",
                  dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        QSDVR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        QSDVR_REQUIRE(p != nullptr, "qsdvr_ray_cut_noalloc: no device address for the workspace");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return QSDVR_OK;
}

}  // namespace qsdvr

using namespace qsdvr;

extern "C" {

int qsdvr_abi_version(void) { return QSDVR_ABI_VERSION; }

const char *qsdvr_last_error(void) { return g_err; }

size_t qsdvr_ray_cut_workspace_size(int64_t pixel_count) {
    if (pixel_count < 0) return 0;
    return ray_cut_workspace_bytes(pixel_count);
}

int qsdvr_ray_cut(const qsdvr_camera *cam, const uint8_t *mask_host, int32_t *ray_list,
                  float *origin, float *dir, float interval, int64_t pixel_count,
                  void *workspace, size_t workspace_bytes, int64_t *sdf_total,
                  int64_t *render_total, qsdvr_stream stream) {
    QSDVR_REQUIRE(cam && ray_list && origin && dir && sdf_total && render_total,
                  "qsdvr_ray_cut: null pointer argument");
    QSDVR_REQUIRE(pixel_count >= 0 && pixel_count <= 0xFFFFFFFFll,
                  "qsdvr_ray_cut: pixel_count %lld out of range", (long long)pixel_count);
    QSDVR_REQUIRE((int64_t)cam->width * cam->height == pixel_count,
                  "qsdvr_ray_cut: pixel_count %lld != width*height %u*%u",
                  (long long)pixel_count, cam->width, cam->height);
    QSDVR_REQUIRE(interval > 0.0f, "qsdvr_ray_cut: interval must be > 0");
    QSDVR_REQUIRE(cam->fx != 0.0f && cam->fy != 0.0f, "qsdvr_ray_cut: zero focal length");
    if (pixel_count == 0) {
        *sdf_total = 0;
        *render_total = 0;
        return QSDVR_OK;
    }
    Carver cv{(char *)workspace, workspace_bytes};
    uint8_t *mask_dev = cv.take<uint8_t>(pixel_count);
    int64_t nb = (pixel_count + kScanTile - 1) / kScanTile;
    uint64_t *sums = cv.take<uint64_t>(2 * nb);
    uint64_t *totals = cv.take<uint64_t>(2);
    if (!mask_dev || !sums || !totals) {
        set_error("qsdvr_ray_cut: workspace of %zu bytes too small (need %zu)", workspace_bytes,
                  qsdvr_ray_cut_workspace_size(pixel_count));
        return QSDVR_EWORKSPACE;
    }
    hipStream_t s = as_stream(stream);
    const uint8_t *mask = nullptr;
    if (mask_host) {
        QSDVR_HIP(hipMemcpyAsync(mask_dev, mask_host, pixel_count, hipMemcpyHostToDevice, s));
        mask = mask_dev;
    }
    ray_cut_kernel<<<grid_1d(pixel_count, 256), 256, 0, s>>>(*cam, interval, (uint32_t)pixel_count,
                                                            ray_list, origin, dir);
    QSDVR_LAUNCHED();
    scan_reduce_kernel<<<(unsigned)nb, kScanBlock, 0, s>>>(ray_list, mask, pixel_count, sums);
    QSDVR_LAUNCHED();
    scan_sums_kernel<<<1, kScanBlock, 0, s>>>(sums, nb, totals);
    QSDVR_LAUNCHED();
    scan_write_kernel<<<(unsigned)nb, kScanBlock, 0, s>>>(ray_list, mask, pixel_count, sums);
    QSDVR_LAUNCHED();
    uint64_t host_totals[2];
    QSDVR_HIP(hipMemcpyAsync(host_totals, totals, sizeof(host_totals), hipMemcpyDeviceToHost, s));
    QSDVR_HIP(hipStreamSynchronize(s));
    if (host_totals[0] > 0xFFFFFFFFull || host_totals[1] > 0xFFFFFFFFull) {
        set_error("qsdvr_ray_cut: sample total %llu exceeds the 32-bit RayInfo offsets",
                  (unsigned long long)host_totals[0]);
        return QSDVR_ERANGE;
    }
    *sdf_total = (int64_t)host_totals[0];
    *render_total = (int64_t)host_totals[1];
    return QSDVR_OK;
}

int64_t qsdvr_ray_cut_noalloc_max_pixels(void) { return kNoAllocMaxPixels; }

int qsdvr_ray_cut_noalloc(const qsdvr_camera *cam, const uint8_t *mask_host, int32_t *ray_list,
                          float *origin, float *dir, float interval, int64_t pixel_count,
                          int64_t *sdf_total, int64_t *render_total, qsdvr_stream stream) {
    if (pixel_count < 0 || pixel_count > kNoAllocMaxPixels) {
        set_error("qsdvr_ray_cut_noalloc: %lld pixels exceed the static workspace (at most %lld)",
                  (long long)pixel_count, (long long)kNoAllocMaxPixels);
        return QSDVR_ERANGE;
    }
    void *ws = nullptr;
    int st = scratch_on_current_device(&ws);
    if (st != QSDVR_OK) return st;
    return qsdvr_ray_cut(cam, mask_host, ray_list, origin, dir, interval, pixel_count, ws,
                         kNoAllocScratchBytes, sdf_total, render_total, stream);
}

int qsdvr_generate_ray_points(float *sdf_point, int32_t *sdf_index, float *render_point,
                              int32_t *render_index, float *view_dir, const int32_t *ray_list,
                              const float *origin, const float *dir, int64_t ray_count,
                              int64_t sdf_total, int64_t render_total, float interval,
                              int32_t base_reso, qsdvr_stream stream) {
    QSDVR_REQUIRE(ray_count >= 0 && sdf_total >= 0 && render_total >= 0,
                  "qsdvr_generate_ray_points: negative size");
    QSDVR_REQUIRE(base_reso > 0 && base_reso <= 1290,
                  "qsdvr_generate_ray_points: base_reso %d out of range", base_reso);
    QSDVR_REQUIRE(interval > 0.0f, "qsdvr_generate_ray_points: interval must be > 0");
    if (ray_count == 0) return QSDVR_OK;
    QSDVR_REQUIRE(ray_list && origin && dir, "qsdvr_generate_ray_points: null ray input");
    QSDVR_REQUIRE((sdf_total == 0 || (sdf_point && sdf_index)) &&
                      (render_total == 0 || (render_point && render_index && view_dir)),
                  "qsdvr_generate_ray_points: null output");
    hipStream_t s = as_stream(stream);
    unsigned blocks = grid_1d(ray_count, kGenBlock / kWave);
    generate_points_kernel<<<blocks, kGenBlock, 0, s>>>(sdf_point, sdf_index, render_point,
                                                        render_index, view_dir, ray_list, origin,
                                                        dir, ray_count, sdf_total, render_total,
                                                        interval, base_reso);
    QSDVR_LAUNCHED();
    return QSDVR_OK;
}

}  // extern "C"
