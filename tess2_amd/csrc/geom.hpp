// geom.hpp -- exact-order float32 geometry shared by the gfx950 kernels and
// the host helpers.  Every expression keeps the operand order and the
// float/double promotions of the reference C++ so that, built with
// -ffp-contract=off and correctly rounded division and square root, the
// results are bit-identical to the reference's x86 build.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>

#define T2_HD __host__ __device__ __forceinline__

namespace t2 {

// norm(): |x|^2 accumulated from 0, src/tet.cpp:122-128
T2_HD float nrm3(float a, float b, float c)
{
    float r = 0.0f;
    r += a * a;
    r += b * b;
    r += c * c;
    return r;
}

// circumcenter(): src/tet.cpp:37-66 (with norm, cross, determinant :122-143)
T2_HD void circumcenter(float ax, float ay, float az, float bx, float by, float bz,
                        float cx, float cy, float cz, float dx, float dy, float dz,
                        float &ox, float &oy, float &oz)
{
    const float t0 = ax - dx, t1 = ay - dy, t2 = az - dz;
    const float u0 = bx - dx, u1 = by - dy, u2 = bz - dz;
    const float v0 = cx - dx, v1 = cy - dy, v2 = cz - dz;
    const float nt = nrm3(t0, t1, t2);
    const float nu = nrm3(u0, u1, u2);
    const float nv = nrm3(v0, v1, v2);
    const float det = t0 * u1 * v2 + u0 * v1 * t2 + v0 * t1 * u2 - v0 * u1 * t2 -
                      u0 * t1 * v2 - t0 * v1 * u2;
    const float den = 2.0f * det;
    const float uv0 = u1 * v2 - u2 * v1, uv1 = u2 * v0 - u0 * v2, uv2 = u0 * v1 - u1 * v0;
    const float vt0 = v1 * t2 - v2 * t1, vt1 = v2 * t0 - v0 * t2, vt2 = v0 * t1 - v1 * t0;
    const float tu0 = t1 * u2 - t2 * u1, tu1 = t2 * u0 - t0 * u2, tu2 = t0 * u1 - t1 * u0;
    ox = dx + (nt * uv0 + nu * vt0 + nv * tu0) / den;
    oy = dy + (nt * uv1 + nu * vt1 + nv * tu1) / den;
    oz = dz + (nt * uv2 + nu * vt2 + nv * tu2) / den;
}

// A bound on |float circumcenter - exact circumcenter| (per component max):
// the relative perturbations of the float differences t, u, v and the
// rounding of every product and sum are bounded through the sums of the
// absolute values of the monomials of the numerator and of the determinant;
// 32 ulp on each with |c - d| times the determinant's, over |den|.
// Infinite when den is 0 or the inputs are not finite.
T2_HD float circumcenter_err(float ax, float ay, float az, float bx, float by, float bz,
                             float cx, float cy, float cz, float dx, float dy, float dz,
                             float ox, float oy, float oz)
{
    const float t0 = ax - dx, t1 = ay - dy, t2 = az - dz;
    const float u0 = bx - dx, u1 = by - dy, u2 = bz - dz;
    const float v0 = cx - dx, v1 = cy - dy, v2 = cz - dz;
    const float nt = nrm3(t0, t1, t2), nu = nrm3(u0, u1, u2), nv = nrm3(v0, v1, v2);
    const float a_uv0 = fabsf(u1 * v2) + fabsf(u2 * v1), a_uv1 = fabsf(u2 * v0) + fabsf(u0 * v2),
                a_uv2 = fabsf(u0 * v1) + fabsf(u1 * v0);
    const float a_vt0 = fabsf(v1 * t2) + fabsf(v2 * t1), a_vt1 = fabsf(v2 * t0) + fabsf(v0 * t2),
                a_vt2 = fabsf(v0 * t1) + fabsf(v1 * t0);
    const float a_tu0 = fabsf(t1 * u2) + fabsf(t2 * u1), a_tu1 = fabsf(t2 * u0) + fabsf(t0 * u2),
                a_tu2 = fabsf(t0 * u1) + fabsf(t1 * u0);
    const float sn0 = nt * a_uv0 + nu * a_vt0 + nv * a_tu0;
    const float sn1 = nt * a_uv1 + nu * a_vt1 + nv * a_tu1;
    const float sn2 = nt * a_uv2 + nu * a_vt2 + nv * a_tu2;
    const float sn = fmaxf(sn0, fmaxf(sn1, sn2));
    const float sd = 2.0f * (fabsf(t0 * u1 * v2) + fabsf(u0 * v1 * t2) + fabsf(v0 * t1 * u2) +
                             fabsf(v0 * u1 * t2) + fabsf(u0 * t1 * v2) + fabsf(t0 * v1 * u2));
    const float det = t0 * u1 * v2 + u0 * v1 * t2 + v0 * t1 * u2 - v0 * u1 * t2 -
                      u0 * t1 * v2 - t0 * v1 * u2;
    const float den = fabsf(2.0f * det);
    const float R = fmaxf(fabsf(ox - dx), fmaxf(fabsf(oy - dy), fabsf(oz - dz)));
    const float cmax = fmaxf(fabsf(ox), fmaxf(fabsf(oy), fabsf(oz)));
    const float ulp = 5.9604645e-8f;
    const float e = 32.0f * ulp * (sn + R * sd) / den + 4.0f * ulp * cmax;
    return (den > 0.0f && e == e) ? e : INFINITY;
}

// same solve in double (FP64 circumcenters)
T2_HD void circumcenter_f64(double ax, double ay, double az, double bx, double by,
                            double bz, double cx, double cy, double cz, double dx,
                            double dy, double dz, double &ox, double &oy, double &oz)
{
    const double t0 = ax - dx, t1 = ay - dy, t2 = az - dz;
    const double u0 = bx - dx, u1 = by - dy, u2 = bz - dz;
    const double v0 = cx - dx, v1 = cy - dy, v2 = cz - dz;
    const double nt = t0 * t0 + t1 * t1 + t2 * t2;
    const double nu = u0 * u0 + u1 * u1 + u2 * u2;
    const double nv = v0 * v0 + v1 * v1 + v2 * v2;
    const double uv0 = u1 * v2 - u2 * v1, uv1 = u2 * v0 - u0 * v2, uv2 = u0 * v1 - u1 * v0;
    const double vt0 = v1 * t2 - v2 * t1, vt1 = v2 * t0 - v0 * t2, vt2 = v0 * t1 - v1 * t0;
    const double tu0 = t1 * u2 - t2 * u1, tu1 = t2 * u0 - t0 * u2, tu2 = t0 * u1 - t1 * u0;
    const double den = 2.0 * (t0 * uv0 + t1 * uv1 + t2 * uv2);
    ox = dx + (nt * uv0 + nu * vt0 + nv * tu0) / den;
    oy = dy + (nt * uv1 + nu * vt1 + nv * tu1) / den;
    oz = dz + (nt * uv2 + nu * vt2 + nv * tu2) / den;
}

// idx2phys(): src/dense.cpp:1098-1106
T2_HD float idx2phys1(int idx, float step, float gmin)
{
    return (float)idx * step + gmin;
}

// phys2idx(): truncating, src/dense.cpp:1117-1125
T2_HD int phys2idx1(float pos, float step, float gmin)
{
    return (int)((pos - gmin) / step);
}

}  // namespace t2
