// Vector and colour arithmetic shared by host and device code.
//
// Every expression mirrors the reference's operator definitions
// (src/yafraycore/vector3d.h:145-219, color.h:37-115) and is evaluated in
// FP32, left to right, with floating-point contraction disabled for the
// whole library (-ffp-contract=off), so device results are bit-identical to
// a strict-IEEE CPU evaluation of the same expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define YG_HD __host__ __device__ __forceinline__

struct f3 {
    float x, y, z;
};

YG_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
YG_HD f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }
YG_HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
YG_HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
YG_HD f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
// operator*(vector, PFLOAT) and (PFLOAT, vector) both compute f*b.x (vector3d.h:123-136)
YG_HD f3 operator*(f3 b, float f) { return mk3(f * b.x, f * b.y, f * b.z); }
YG_HD f3 operator/(f3 b, float f) { return mk3(b.x / f, b.y / f, b.z / f); }
YG_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
YG_HD f3 cross(f3 a, f3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
YG_HD bool isnull3(f3 a) { return a.x == 0 && a.y == 0 && a.z == 0; }

// vector3d_t::normalize (vector3d.h:209-219).  The reference divides 1.0 (double)
// by the float root; rounding that quotient to float equals the float quotient.
YG_HD f3 norm3(f3 a)
{
    float len = a.x * a.x + a.y * a.y + a.z * a.z;
    if (len != 0) {
        len = 1.0f / sqrtf(len);
        a.x *= len; a.y *= len; a.z *= len;
    }
    return a;
}
YG_HD float len3(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
// normLenSqr (vector3d.h:61-69)
YG_HD float normlensqr3(f3& a)
{
    float vl = a.x * a.x + a.y * a.y + a.z * a.z;
    if (vl != 0.0f) {
        const float d = 1.0f / sqrtf(vl);
        a.x *= d; a.y *= d; a.z *= d;
    }
    return vl;
}
// normLen (vector3d.h:50-59)
YG_HD float normlen3(f3& a)
{
    float vl = a.x * a.x + a.y * a.y + a.z * a.z;
    if (vl != 0.0f) {
        vl = sqrtf(vl);
        const float d = 1.0f / vl;
        a.x *= d; a.y *= d; a.z *= d;
    }
    return vl;
}
// FACE_FORWARD (shader.h:157)
YG_HD f3 face_forward(f3 ng, f3 n, f3 i) { return (dot(ng, i) < 0) ? -n : n; }
// createCS (vector3d.h:240-258)
YG_HD void create_cs(f3 n, f3& u, f3& v)
{
    if (n.x == 0 && n.y == 0) {
        u = (n.z < 0) ? mk3(-1, 0, 0) : mk3(1, 0, 0);
        v = mk3(0, 1, 0);
    } else {
        const float d = 1.0f / sqrtf(n.y * n.y + n.x * n.x);
        u = mk3(n.y * d, -n.x * d, 0);
        v = cross(n, u);
    }
}

// colour (color_t)
struct c3 {
    float r, g, b;
};
YG_HD c3 mkc(float r, float g, float b) { c3 c; c.r = r; c.g = g; c.b = b; return c; }
YG_HD c3 cadd(c3 a, c3 b) { return mkc(a.r + b.r, a.g + b.g, a.b + b.b); }
YG_HD c3 csub(c3 a, c3 b) { return mkc(a.r - b.r, a.g - b.g, a.b - b.b); }
YG_HD c3 cmul(c3 a, c3 b) { return mkc(a.r * b.r, a.g * b.g, a.b * b.b); }
YG_HD c3 cscale(c3 a, float f) { return mkc(f * a.r, f * a.g, f * a.b); }
YG_HD float cenergy(c3 c) { return (c.r + c.g + c.b) * 0.333333f; }
// mix (color.cc:103-109)
YG_HD c3 cmix(c3 a, c3 b, float point)
{
    if (point < 0.0f) return b;
    if (point > 1.0f) return a;
    return cadd(cscale(a, point), cscale(b, 1 - point));
}

// Reference float trigonometry (cosf/sinf) evaluated through double and rounded,
// identically on host and device.
YG_HD float dcos(float x) { return (float)cos((double)x); }
YG_HD float dsin(float x) { return (float)sin((double)x); }
YG_HD float dpow(float x, float y) { return (float)pow((double)x, (double)y); }
