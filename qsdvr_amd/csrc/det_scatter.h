// Deterministic (sorted-segment) gradient scatter of the fused adjoint.
//
// The adjoint kernel writes one record per gradient contribution instead of
// adding it to the gradient grids:
//   * quadric records, two per sdf-sample slot: the sdf sample's 7
//     coefficient gradients and the render sample's 6 normal gradients,
//     keyed by cell;
//   * appearance records, one per render sample: the local point and the
//     gradient of the interpolated row (C channels), keyed by
//     (cell parity class, cell).
// Records are sorted by key (stable LSD radix sort, so equal keys keep slot
// order) and every run of equal keys is summed by one wavefront in a fixed
// order.  Quadric runs own their cell and add to it directly.  Appearance
// runs add to the 8 vertex rows of their cell in 8 passes, one per parity
// class (x%2, y%2, z%2): cells of one class share no vertex, so each pass
// is a plain read-modify-write without atomics, and every vertex receives
// its contributions in class order.  The result is bitwise reproducible
// run to run.
#pragma once

#include "common.h"

namespace qsdvr {

// Appearance record: float4 (local x, y, z, 0) then C gradient channels,
// padded to a multiple of 4 floats.
__host__ __device__ constexpr int det_app_stride(int C) { return 4 + (C + 3) / 4 * 4; }

// Bits of a cell index in [0, R^3) (at least 1).
inline int det_cell_bits(int R) {
    const uint64_t n = (uint64_t)R * R * R;
    int b = 1;
    while ((1ull << b) < n) ++b;
    return b;
}

// GridInterpolationBackward's sort keys.  A sample in cell (x, y, z) of an
// R^3 grid (R even, h = R / 2) gets the key cls h^3 + x' + h y' + h^2 z', with
// cls = (x & 1) | (y & 1) << 1 | (z & 1) << 2 the cell's parity class and
// (x', y', z') = (x, y, z) >> 1; an invalid sample gets R^3.  The keys of one
// class are contiguous and ordered as their cells, and a key needs no more
// bits than the cell index (a class field above the cell bits would add
// 3 bits and an invalid flag a fourth, one more radix pass at R = 128).
struct ClassKeys {
    uint32_t h, hh, h3;  // R / 2, (R / 2)^2, (R / 2)^3
    int R;
    __host__ __device__ __forceinline__ uint32_t key(int64_t x, int64_t y, int64_t z) const {
        const uint32_t cls = (uint32_t)((x & 1) | ((y & 1) << 1) | ((z & 1) << 2));
        return cls * h3 + (uint32_t)(x >> 1) + h * ((uint32_t)(y >> 1) + h * (uint32_t)(z >> 1));
    }
    __host__ __device__ __forceinline__ int64_t cell(uint32_t k) const {
        uint32_t cls = 0;
#pragma unroll
        for (uint32_t c = 1; c < 8; ++c) cls += k >= c * h3 ? 1u : 0u;
        uint32_t c = k - cls * h3;
        const uint32_t z = c / hh;
        c -= z * hh;
        const uint32_t y = c / h, x = c - y * h;
        return (int64_t)(2 * x + (cls & 1)) +
               (int64_t)R * ((int64_t)(2 * y + ((cls >> 1) & 1)) + (int64_t)R * (2 * z + (cls >> 2)));
    }
};

inline ClassKeys class_keys(int R) {
    const uint32_t h = (uint32_t)(R / 2);
    return ClassKeys{h, h * h, h * h * h, R};
}

struct DetBuffers {
    int64_t n_rays = 0, cap = 0;  // rays, sdf-sample slot capacity
    int C = 0, R = 0, cbits = 0;
    uint32_t *counts = nullptr;   // [n_rays + 1] sdf samples per ray
    uint32_t *offsets = nullptr;  // [n_rays + 1] exclusive scan of counts
    uint32_t *qkey = nullptr, *qkey_s = nullptr, *qidx = nullptr, *qidx_s = nullptr;  // [2 cap]
    float4 *qval = nullptr;       // [2 cap][2]: 7 coefficient gradients + pad
    uint32_t *akey = nullptr, *akey_s = nullptr, *aidx = nullptr, *aidx_s = nullptr;  // [cap]
    float *aval = nullptr;        // [cap][det_app_stride(C)]
    uint32_t *bounds = nullptr;   // [9]: class boundaries in the sorted appearance keys
    uint32_t *total = nullptr;    // [1] host-visible copy target is the caller's
    float *qhead = nullptr, *qtail = nullptr;  // [waves of 2 cap][7] quadric segment sums
    void *temp = nullptr;         // rocPRIM temporary storage
    size_t temp_bytes = 0;
};

// Bytes of workspace for n_rays rays and at most `cap` sdf samples in all.
size_t det_workspace_bytes(int64_t n_rays, int64_t cap, int C, int R);

// Carves the workspace; QSDVR_EWORKSPACE if it is too small.
int det_carve(void *ws, size_t bytes, int64_t n_rays, int64_t cap, int C, int R, DetBuffers &b);

// counts[i] = sdf samples of ray i from the ray state (origin, count bits,
// dir, 0) -- count if >= 2 else 0, or K for listed samples -- then the
// exclusive scan into offsets; returns the total in *total_host (one
// device-to-host copy and stream synchronisation).
int det_offsets(DetBuffers &b, const float4 *ray_state, int K, hipStream_t s,
                int64_t *total_host);

// Sorts the T slots' records and adds the run sums into the gradient grids.
int det_reduce(DetBuffers &b, int64_t T, float *gq, float *gdata, hipStream_t s);

}  // namespace qsdvr
