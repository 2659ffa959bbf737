// Bit-exact ray generation arithmetic (kernel/camera.cu semantics), shared by
// the op-surface kernels and the fused renderer.
//
// Every function here is compiled with floating-point contraction off, so
// each +, -, * and / is one IEEE round-to-nearest float32 operation in the
// reference's association order; division and sqrt are the correctly rounded
// HIP defaults (-fhip-fp32-correctly-rounded-divide-sqrt).
#pragma once

#include "common.h"

namespace qsdvr {

struct RayCutResult {
    f3 dir;
    f3 origin;
    uint32_t count;
};

// kernel/camera.cu:3-10: ((u + cx) / fx) * z, rotation rows summed left to
// right, then translation.
__host__ __device__ __forceinline__ f3 camera_to_world(const qsdvr_camera &c, float u, float v, float z) {
#pragma clang fp contract(off)
    float x = ((u + c.cx) / c.fx) * z;
    float y = ((v + c.cy) / c.fy) * z;
    f3 p;
    p.x = ((c.rotation[0] * x + c.rotation[1] * y) + c.rotation[2] * z) + c.translation[0];
    p.y = ((c.rotation[3] * x + c.rotation[4] * y) + c.rotation[5] * z) + c.translation[1];
    p.z = ((c.rotation[6] * x + c.rotation[7] * y) + c.rotation[8] * z) + c.translation[2];
    return p;
}

__host__ __device__ __forceinline__ bool in_cube(f3 p) {
    const float b = 0.9999f;  // kernel/camera.cu:46
    return p.x < b && p.x > -b && p.y < b && p.y > -b && p.z < b && p.z > -b;
}

// Candidate step j of a ray: near + (dir*j)*interval (kernel/camera.cu:45).
__host__ __device__ __forceinline__ f3 cut_step(f3 n, f3 dir, float interval, uint32_t j) {
#pragma clang fp contract(off)
    float fj = (float)j;
    return f3{n.x + (dir.x * fj) * interval, n.y + (dir.y * fj) * interval,
              n.z + (dir.z * fj) * interval};
}

// The reference's walk over the candidate steps (kernel/camera.cu:43-54):
// count of steps inside and the first one.  The inside set is contiguous
// (each coordinate of a step is monotone in j), so the walk stops at the
// first outside step after the inside run.
__host__ __device__ __forceinline__ uint32_t cut_walk(f3 n, f3 dir, float interval, uint32_t cand,
                                                      f3 &ori) {
    uint32_t count = 0;
    ori = f3{0.0f, 0.0f, 0.0f};
    for (uint32_t j = 0; j < cand; ++j) {
        const f3 p = cut_step(n, dir, interval, j);
        if (in_cube(p)) {
            if (count == 0) ori = p;
            ++count;
        } else if (count != 0) {
            break;
        }
    }
    return count;
}

// The same count and first inside step without walking from the near plane:
// a float64 slab test gives the real-valued inside range (lo, hi) of j, and
// the exact float32 predicate is evaluated on the steps within 2 of its ends
// only.  A float32 step lies within 5e-7 of its real value (three roundings
// of magnitudes below 4), so no inside step lies outside
// [floor(lo) - 2, ceil(hi) + 2] as long as every axis moves at least 1e-4
// per step (the error is then under 0.005 steps); a ray that grazes a face
// more slowly takes the walk.  Bit-identical to cut_walk.
__host__ __device__ __forceinline__ uint32_t cut_closed_form(f3 n, f3 dir, float interval,
                                                             uint32_t cand, f3 &ori) {
    ori = f3{0.0f, 0.0f, 0.0f};
    if (cand == 0) return 0;
    const double b = (double)0.9999f;
    double lo = 0.0, hi = (double)cand - 1.0;
    const float na[3] = {n.x, n.y, n.z}, da[3] = {dir.x, dir.y, dir.z};
    for (int a = 0; a < 3; ++a) {
        if (da[a] == 0.0f) {  // the coordinate stays exactly n: inside for every j, or none
            if (!(na[a] < 0.9999f && na[a] > -0.9999f)) return 0;
            continue;
        }
        const double step = (double)da[a] * (double)interval;
        if (step < 1e-4 && step > -1e-4) return cut_walk(n, dir, interval, cand, ori);
        const double t1 = (-b - (double)na[a]) / step, t2 = (b - (double)na[a]) / step;
        lo = t1 < t2 ? (t1 > lo ? t1 : lo) : (t2 > lo ? t2 : lo);
        hi = t1 < t2 ? (t2 < hi ? t2 : hi) : (t1 < hi ? t1 : hi);
    }
    if (lo > hi + 4.0) return 0;  // no step within 2 of an inside real j
    const double flo = lo - 2.0 < 0.0 ? 0.0 : lo - 2.0;
    const double fhi = hi + 2.0 > (double)cand - 1.0 ? (double)cand - 1.0 : hi + 2.0;
    if (fhi < flo) return 0;
    uint32_t j0 = (uint32_t)flo;                  // floor (non-negative)
    uint32_t j1 = (uint32_t)fhi;
    if ((double)j1 < fhi) ++j1;                   // ceil
    if (j1 > cand - 1) j1 = cand - 1;
    while (j0 <= j1 && !in_cube(cut_step(n, dir, interval, j0))) ++j0;
    if (j0 > j1) return 0;
    while (j1 > j0 && !in_cube(cut_step(n, dir, interval, j1))) --j1;
    ori = cut_step(n, dir, interval, j0);
    return j1 - j0 + 1;
}

// Ray of pixel `pix` (row-major) and its in-cube step count / first inside
// step (kernel/camera.cu:12-58; the count from cut_closed_form).
__host__ __device__ __forceinline__ RayCutResult ray_cut_pixel(const qsdvr_camera &c, uint32_t pix,
                                                               float interval) {
#pragma clang fp contract(off)
    float u = (float)(pix % c.width) + 0.5f;
    float v = (float)(pix / c.width) + 0.5f;
    f3 n = camera_to_world(c, u, v, c.near_depth);
    f3 f = camera_to_world(c, u, v, c.far_depth);
    f3 d{f.x - n.x, f.y - n.y, f.z - n.z};
    float nrm = sqrtf((d.x * d.x + d.y * d.y) + d.z * d.z);
    f3 dir{d.x / nrm, d.y / nrm, d.z / nrm};
    float end;
    if (dir.x != 0.0f)
        end = (f.x - n.x) / dir.x;
    else if (dir.y != 0.0f)
        end = (f.y - n.y) / dir.y;
    else
        end = (f.z - n.z) / dir.z;
    uint32_t cand = (uint32_t)(end / interval);
    f3 ori;
    const uint32_t count = cut_closed_form(n, dir, interval, cand, ori);
    return RayCutResult{dir, ori, count};
}

struct CellPoint {
    f3 local;
    int32_t cell;  // x + y*R + z*R*R
    int ix, iy, iz;
};

__device__ __forceinline__ CellPoint to_cell(float qx, float qy, float qz, int R) {
#pragma clang fp contract(off)
    int x0 = (int)floorf(qx);
    int y0 = (int)floorf(qy);
    int z0 = (int)floorf(qz);
    CellPoint r;
    r.local = f3{qx * 2.0f - ((float)(2 * x0) + 1.0f), qy * 2.0f - ((float)(2 * y0) + 1.0f),
                 qz * 2.0f - ((float)(2 * z0) + 1.0f)};
    r.cell = (int32_t)((uint32_t)x0 + (uint32_t)y0 * (uint32_t)R + (uint32_t)z0 * (uint32_t)(R * R));
    r.ix = x0;
    r.iy = y0;
    r.iz = z0;
    return r;
}

// sdf sample j: ((ori + (dir*j)*interval) + 1) * (R/2)   (kernel/camera.cu:73-78)
__device__ __forceinline__ CellPoint sdf_sample(f3 ori, f3 dir, uint32_t j, float interval,
                                                float half_reso, int R) {
#pragma clang fp contract(off)
    float fj = (float)j;
    float qx = ((ori.x + (dir.x * fj) * interval) + 1.0f) * half_reso;
    float qy = ((ori.y + (dir.y * fj) * interval) + 1.0f) * half_reso;
    float qz = ((ori.z + (dir.z * fj) * interval) + 1.0f) * half_reso;
    return to_cell(qx, qy, qz, R);
}

// render sample j: ((ori + dir*((j+0.5)*interval)) + 1) * (R/2)   (camera.cu:86-91)
__device__ __forceinline__ CellPoint render_sample(f3 ori, f3 dir, uint32_t j, float interval,
                                                   float half_reso, int R) {
#pragma clang fp contract(off)
    float s = ((float)j + 0.5f) * interval;
    float qx = ((ori.x + dir.x * s) + 1.0f) * half_reso;
    float qy = ((ori.y + dir.y * s) + 1.0f) * half_reso;
    float qz = ((ori.z + dir.z * s) + 1.0f) * half_reso;
    return to_cell(qx, qy, qz, R);
}

}  // namespace qsdvr
