// Appearance: trilinear weights, real-SH reflection shading and their
// adjoints, shared by the op-surface kernels and the fused renderer.
//
// Channel layout of an appearance row with NC = number of SH coefficients
// (9 for degree <= 2, 16 for degree 3): channel 3k+c = coefficient k of
// colour c, then diffuse RGB at 3NC..3NC+2, metallic at 3NC+3, ao at 3NC+4
// (models/render_torch.py:144-152 for NC = 9).
#pragma once

#include "common.h"

namespace qsdvr {

// Basis terms Y_k(v) (with their constants) for k < NC.
// k < 9 : models/render_torch.py:174-195.  That table is the unsigned real SH
//         of degree <= 2 in the order m = -l..l, with the pole on y: the
//         z-pole polynomials under (x, y, z) -> (z, x, y).
// k >= 9: degree 3 of the same family, m = -3..3:
//         x(3z^2 - x^2), xyz, x(4y^2 - x^2 - z^2), y(2y^2 - 3x^2 - 3z^2),
//         z(4y^2 - x^2 - z^2), y(z^2 - x^2), z(z^2 - 3x^2), with the
//         normalisations sqrt(35/32pi), sqrt(105/4pi), sqrt(21/32pi),
//         sqrt(7/16pi), sqrt(21/32pi), sqrt(105/16pi), sqrt(35/32pi).
template <int NC>
__device__ __forceinline__ void sh_basis(f3 v, float Y[NC]) {
    const float x = v.x, y = v.y, z = v.z;
    const float zz = z * z, xx = x * x;
    Y[0] = QSDVR_SH_K0;
    Y[1] = QSDVR_SH_K1 * x;
    Y[2] = QSDVR_SH_K1 * y;
    Y[3] = QSDVR_SH_K1 * z;
    Y[4] = QSDVR_SH_K4 * (x * z);
    Y[5] = QSDVR_SH_K4 * (x * y);
    Y[6] = QSDVR_SH_K6 * (2.0f * y * y - zz - xx);
    Y[7] = QSDVR_SH_K4 * (y * z);
    Y[8] = QSDVR_SH_K8 * (zz - xx);
    if constexpr (NC > 9) {
        const float yy = y * y;
        Y[9] = QSDVR_SH_K9 * (x * (3.0f * zz - xx));
        Y[10] = QSDVR_SH_K10 * (x * y * z);
        Y[11] = QSDVR_SH_K11 * (x * (4.0f * yy - xx - zz));
        Y[12] = QSDVR_SH_K12 * (y * (2.0f * yy - 3.0f * xx - 3.0f * zz));
        Y[13] = QSDVR_SH_K11 * (z * (4.0f * yy - xx - zz));
        Y[14] = QSDVR_SH_K14 * (y * (zz - xx));
        Y[15] = QSDVR_SH_K9 * (z * (zz - 3.0f * xx));
    }
}

// vbar += sum_k gc[k] * grad Y_k(v)   (gc already includes the constants'
// product with the coefficients: gc[k] = K_k * (coef_k . upstream)).
template <int NC>
__device__ __forceinline__ f3 sh_basis_vjp(f3 v, const float gc[NC]) {
    const float x = v.x, y = v.y, z = v.z;
    // degree <= 2 (matches kernel/render.cu:79-81 term by term)
    float gx = gc[1] + z * gc[4] + y * gc[5] - 2.0f * x * (gc[6] + gc[8]);
    float gy = gc[2] + z * gc[7] + x * gc[5] + 4.0f * y * gc[6];
    float gz = gc[3] + x * gc[4] + y * gc[7] + 2.0f * z * (gc[8] - gc[6]);
    if constexpr (NC > 9) {
        const float xx = x * x, yy = y * y, zz = z * z;
        gx += gc[9] * (3.0f * zz - 3.0f * xx) + gc[10] * (y * z) +
              gc[11] * (4.0f * yy - 3.0f * xx - zz) + gc[12] * (-6.0f * x * y) +
              gc[13] * (-2.0f * x * z) + gc[14] * (-2.0f * x * y) + gc[15] * (-6.0f * x * z);
        gy += gc[10] * (x * z) + gc[11] * (8.0f * x * y) +
              gc[12] * (6.0f * yy - 3.0f * xx - 3.0f * zz) + gc[13] * (8.0f * y * z) +
              gc[14] * (zz - xx);
        gz += gc[9] * (6.0f * x * z) + gc[10] * (x * y) + gc[11] * (-2.0f * x * z) +
              gc[12] * (-6.0f * y * z) + gc[13] * (4.0f * yy - xx - 3.0f * zz) +
              gc[14] * (2.0f * y * z) + gc[15] * (3.0f * zz - 3.0f * xx);
    }
    return f3{gx, gy, gz};
}

// Unit constants K_k used in gc (Y_k / basis polynomial).
template <int NC>
__device__ __forceinline__ float sh_const(int k) {
    switch (k) {
        case 0: return QSDVR_SH_K0;
        case 1: case 2: case 3: return QSDVR_SH_K1;
        case 4: case 5: case 7: return QSDVR_SH_K4;
        case 6: return QSDVR_SH_K6;
        case 8: return QSDVR_SH_K8;
        case 9: case 15: return QSDVR_SH_K9;
        case 10: return QSDVR_SH_K10;
        case 11: case 13: return QSDVR_SH_K11;
        case 12: return QSDVR_SH_K12;
        default: return QSDVR_SH_K14;
    }
}

struct ShadeOut {
    f3 rgb;
    f3 spec;
};

// rgb = ao * (diffuse * (1 - m) + L * m), L = SH(reflect(view, n)),
// reflect = v - n * (2 (v . n))   (render_torch.py:144-152)
template <int NC>
__device__ __forceinline__ ShadeOut shade(f3 n, f3 v, const float *d) {
    float vdotn = v.x * n.x + v.y * n.y + v.z * n.z;
    f3 r{v.x - n.x * (2.0f * vdotn), v.y - n.y * (2.0f * vdotn), v.z - n.z * (2.0f * vdotn)};
    float Y[NC];
    sh_basis<NC>(r, Y);
    f3 L{0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        L.x += Y[k] * d[3 * k + 0];
        L.y += Y[k] * d[3 * k + 1];
        L.z += Y[k] * d[3 * k + 2];
    }
    const float m = d[3 * NC + 3], ao = d[3 * NC + 4];
    const float om = 1.0f - m;
    ShadeOut o;
    o.spec = L;
    o.rgb = f3{(d[3 * NC + 0] * om + L.x * m) * ao, (d[3 * NC + 1] * om + L.y * m) * ao,
               (d[3 * NC + 2] * om + L.z * m) * ao};
    return o;
}

// Adjoint of shade: writes dd[3NC+5] (overwrite) and returns d rgb / d n.
template <int NC>
__device__ __forceinline__ f3 shade_vjp(f3 n, f3 v, const float *d, f3 L, f3 g, float *dd) {
    const float m = d[3 * NC + 3], ao = d[3 * NC + 4];
    const float om = 1.0f - m;
    f3 color{d[3 * NC + 0] * om + L.x * m, d[3 * NC + 1] * om + L.y * m,
             d[3 * NC + 2] * om + L.z * m};
    dd[3 * NC + 4] = g.x * color.x + g.y * color.y + g.z * color.z;  // ao
    f3 ga{g.x * ao, g.y * ao, g.z * ao};
    dd[3 * NC + 3] = ga.x * (L.x - d[3 * NC + 0]) + ga.y * (L.y - d[3 * NC + 1]) +
                     ga.z * (L.z - d[3 * NC + 2]);  // metallic
    dd[3 * NC + 0] = ga.x * om;
    dd[3 * NC + 1] = ga.y * om;
    dd[3 * NC + 2] = ga.z * om;
    f3 gl{ga.x * m, ga.y * m, ga.z * m};  // upstream of L
    float vdotn = v.x * n.x + v.y * n.y + v.z * n.z;
    f3 r{v.x - n.x * (2.0f * vdotn), v.y - n.y * (2.0f * vdotn), v.z - n.z * (2.0f * vdotn)};
    float Y[NC];
    sh_basis<NC>(r, Y);
    float gc[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        dd[3 * k + 0] = gl.x * Y[k];
        dd[3 * k + 1] = gl.y * Y[k];
        dd[3 * k + 2] = gl.z * Y[k];
        gc[k] = sh_const<NC>(k) * (d[3 * k] * gl.x + d[3 * k + 1] * gl.y + d[3 * k + 2] * gl.z);
    }
    f3 rb = sh_basis_vjp<NC>(r, gc);
    // r = v - 2 (v.n) n  =>  nbar = -2 [ (v.n) rbar + v (rbar . n) ]
    float rn = rb.x * n.x + rb.y * n.y + rb.z * n.z;
    return f3{-2.0f * (vdotn * rb.x + v.x * rn), -2.0f * (vdotn * rb.y + v.y * rn),
              -2.0f * (vdotn * rb.z + v.z * rn)};
}

// Trilinear weights of models/render_torch.py:128-143 (raw local point):
// w[dz*4 + dy*2 + dx] for the rows i + dx + dy R + dz R^2.
__device__ __forceinline__ void trilinear_weights(f3 p, float w[8]) {
    const float ux = 1.0f - p.x, uy = 1.0f - p.y, uz = 1.0f - p.z;
    w[0] = ux * uy * uz;
    w[1] = p.x * uy * uz;
    w[2] = ux * p.y * uz;
    w[3] = p.x * p.y * uz;
    w[4] = ux * uy * p.z;
    w[5] = p.x * uy * p.z;
    w[6] = ux * p.y * p.z;
    w[7] = p.x * p.y * p.z;
}

__device__ __forceinline__ float lerp_ref(float a, float b, float w) {
    // render_torch.py:170-172, with its one rounding of the sum fixed (an
    // fma of b w onto the rounded (1 - w) a) rather than left to contraction
    return __builtin_fmaf(b, w, (1.0f - w) * a);
}

}  // namespace qsdvr
