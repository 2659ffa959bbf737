// Quadric cell evaluation shared by the op-surface kernels and the fused
// renderer.  A cell stores (A,B,C,D,E,F,G): f(p) = A x^2 + B y^2 + C z^2 +
// D x + E y + F z + G over the local point p in [-1,1)^3
// (kernel/surface.cu:54-146, models/render_torch.py:77-100).
#pragma once

#include "common.h"

namespace qsdvr {

struct Quadric {
    float a, b, c, d, e, f, g;
};

__device__ __forceinline__ Quadric load_quadric(const float *grid, int64_t cell) {
    const float *q = grid + 7 * cell;
    return Quadric{q[0], q[1], q[2], q[3], q[4], q[5], q[6]};
}

// sdf = f / ||grad f|| / reso, with a = A x etc. and
// f = ((a+D) x + (b+E) y) + (c+F) z + G   (render_torch.py:84-90)
__device__ __forceinline__ float quadric_sdf(const Quadric &q, f3 p, float inv_reso) {
    float a0 = q.a * p.x, a1 = q.b * p.y, a2 = q.c * p.z;
    float f = (a0 + q.d) * p.x + (a1 + q.e) * p.y + (a2 + q.f) * p.z + q.g;
    float gx = 2.0f * a0 + q.d, gy = 2.0f * a1 + q.e, gz = 2.0f * a2 + q.f;
    return f * fast_rsqrt(gx * gx + gy * gy + gz * gz) * inv_reso;
}

// Adds d sdf / d(A..G) * upstream to g[7] (closed form of the quotient rule):
// k = 1/||grad f||, q = f k^3, s = upstream / reso;
// dA = s x (x k - 2 g_x q), dD = s (x k - g_x q), ..., dG = s k.
// A zero upstream adds nothing, also where grad f vanishes (k infinite):
// the renderers give such a sample alpha 0, so its upstream is 0, and
// 0 * inf must not turn the cell's gradient into NaN.
__device__ __forceinline__ void quadric_sdf_grad(const Quadric &q, f3 p, float upstream,
                                                 float inv_reso, float g[7]) {
    if (upstream == 0.0f) return;
    float a0 = q.a * p.x, a1 = q.b * p.y, a2 = q.c * p.z;
    float f = (a0 + q.d) * p.x + (a1 + q.e) * p.y + (a2 + q.f) * p.z + q.g;
    float gx = 2.0f * a0 + q.d, gy = 2.0f * a1 + q.e, gz = 2.0f * a2 + q.f;
    float k = fast_rsqrt(gx * gx + gy * gy + gz * gz);
    float qq = f * k * k * k;
    float s = upstream * inv_reso;
    float tx = p.x * k - gx * qq, ty = p.y * k - gy * qq, tz = p.z * k - gz * qq;
    g[0] += s * p.x * (p.x * k - 2.0f * gx * qq);
    g[1] += s * p.y * (p.y * k - 2.0f * gy * qq);
    g[2] += s * p.z * (p.z * k - 2.0f * gz * qq);
    g[3] += s * tx;
    g[4] += s * ty;
    g[5] += s * tz;
    g[6] += s * k;
}

// n = grad f / max(||grad f||, 1e-12)   (F.normalize, render_torch.py:100)
__device__ __forceinline__ f3 quadric_normal(const Quadric &q, f3 p, float *inv_len = nullptr) {
    float gx = 2.0f * q.a * p.x + q.d, gy = 2.0f * q.b * p.y + q.e, gz = 2.0f * q.c * p.z + q.f;
    // 1 / max(||g||, 1e-12): rsq(0) = inf clamps to 1e12
    float k = fminf(fast_rsqrt(gx * gx + gy * gy + gz * gz), 1e12f);
    if (inv_len) *inv_len = k;
    return f3{gx * k, gy * k, gz * k};
}

// Adds the adjoint of the normal map to g[0..5]:
// h = (nbar - n (n . nbar)) / ||grad f||; dA = 2x h_x, dD = h_x, ...
__device__ __forceinline__ void quadric_normal_grad(const Quadric &q, f3 p, f3 nbar, float g[7]) {
    float gx = 2.0f * q.a * p.x + q.d, gy = 2.0f * q.b * p.y + q.e, gz = 2.0f * q.c * p.z + q.f;
    float len2 = gx * gx + gy * gy + gz * gz;
    float hx, hy, hz;
    if (len2 > 1e-24f) {  // ||grad f|| > 1e-12
        float k = fast_rsqrt(len2);
        float nx = gx * k, ny = gy * k, nz = gz * k;
        float dn = nx * nbar.x + ny * nbar.y + nz * nbar.z;
        hx = (nbar.x - nx * dn) * k;
        hy = (nbar.y - ny * dn) * k;
        hz = (nbar.z - nz * dn) * k;
    } else {
        hx = nbar.x * 1e12f;
        hy = nbar.y * 1e12f;
        hz = nbar.z * 1e12f;
    }
    g[0] += 2.0f * p.x * hx;
    g[1] += 2.0f * p.y * hy;
    g[2] += 2.0f * p.z * hz;
    g[3] += hx;
    g[4] += hy;
    g[5] += hz;
}

}  // namespace qsdvr
