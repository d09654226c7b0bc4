// Device-side ray queries over the BVH: the reference triangle test and the
// closest / any-hit acceptance rules of kdTree_t::Intersect / IntersectS.
#pragma once

#include "yg_internal.h"

// triangle_t::hit (triangle.h:48-58): flip the ray to the normal's side, then
// the three edge triple products must be >= 0.  Shared edges give exactly
// opposite products, so no ray slips between two triangles.
__device__ __forceinline__ bool tri_hit(f3 a, f3 b, f3 c, f3 n, f3 from, f3 ray)
{
    const f3 va = a - from, vb = b - from, vc = c - from;
    f3 r = (dot(ray, n) < 0) ? -ray : ray;
    if (dot(cross(va, vb), r) < 0) return false;
    if (dot(cross(vb, vc), r) < 0) return false;
    if (dot(cross(vc, va), r) < 0) return false;
    return true;
}

// triangle_t::intersect (triangle.h:83-86)
__device__ __forceinline__ float tri_t(f3 a, f3 n, f3 from, f3 ray) { return dot(n, a - from) / dot(n, ray); }

// getInterpolation (triangle.cc:74-136): fb, fc of the dominant-axis projection
__device__ __forceinline__ void tri_bary(f3 n, f3 a, f3 b, f3 c, f3 h, float& fb, float& fc)
{
    float uu0, vv0, uu1, vv1, uu2, vv2;
    float nx = fabsf(n.x), ny = fabsf(n.y), nz = fabsf(n.z);
    if (nx >= ny && nx >= nz) {
        uu0 = h.y - a.y; vv0 = h.z - a.z; uu1 = b.y - a.y; vv1 = b.z - a.z; uu2 = c.y - a.y; vv2 = c.z - a.z;
    } else if (ny >= nx && ny >= nz) {
        uu0 = h.x - a.x; vv0 = h.z - a.z; uu1 = b.x - a.x; vv1 = b.z - a.z; uu2 = c.x - a.x; vv2 = c.z - a.z;
    } else {
        uu0 = h.x - a.x; vv0 = h.y - a.y; uu1 = b.x - a.x; vv1 = b.y - a.y; uu2 = c.x - a.x; vv2 = c.y - a.y;
    }
    float alpha_numer = uu0 * vv2 - vv0 * uu2;
    float beta_numer = uu1 * vv0 - vv1 * uu0;
    float denum = uu1 * vv2 - vv1 * uu2;
    fb = alpha_numer / denum;
    fc = beta_numer / denum;
}

// Loads from scene arrays in device memory.  The arrays reach the kernels
// through structs in memory, so the compiler cannot infer their address
// space and would emit flat loads, which share the LDS counter: every wait
// for a node would also drain the stack's LDS traffic.
#define YG_GLOBAL __attribute__((address_space(1)))
typedef float yg_v4f __attribute__((ext_vector_type(4)));
typedef int yg_v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int gload(const int32_t* p) { return *(const YG_GLOBAL int32_t*)p; }

__device__ __forceinline__ float4 gload(const float4* p)
{
    const yg_v4f v = *(const YG_GLOBAL yg_v4f*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ GNode4 gload(const GNode4* p)
{
    const YG_GLOBAL yg_v4f* q = (const YG_GLOBAL yg_v4f*)p;
    const yg_v4f a = q[0], b = q[1], c = q[2], d = q[3], e = q[4], f = q[5];
    const yg_v4i ch = ((const YG_GLOBAL yg_v4i*)p)[6];
    GNode4 n;
    n.lox = make_float4(a.x, a.y, a.z, a.w);
    n.hix = make_float4(b.x, b.y, b.z, b.w);
    n.loy = make_float4(c.x, c.y, c.z, c.w);
    n.hiy = make_float4(d.x, d.y, d.z, d.w);
    n.loz = make_float4(e.x, e.y, e.z, e.w);
    n.hiz = make_float4(f.x, f.y, f.z, f.w);
    n.child = make_int4(ch.x, ch.y, ch.z, ch.w);
    return n;
}

__device__ __forceinline__ GNode8 gload(const GNode8* p)
{
    const YG_GLOBAL yg_v4f* q = (const YG_GLOBAL yg_v4f*)p;
    GNode8 n;
#pragma unroll
    for (int r = 0; r < 6; r++) {
        const yg_v4f a = q[2 * r], b = q[2 * r + 1];
        n.box[r][0] = a.x; n.box[r][1] = a.y; n.box[r][2] = a.z; n.box[r][3] = a.w;
        n.box[r][4] = b.x; n.box[r][5] = b.y; n.box[r][6] = b.z; n.box[r][7] = b.w;
    }
    const yg_v4i c0 = ((const YG_GLOBAL yg_v4i*)p)[12], c1 = ((const YG_GLOBAL yg_v4i*)p)[13];
    n.child[0] = c0.x; n.child[1] = c0.y; n.child[2] = c0.z; n.child[3] = c0.w;
    n.child[4] = c1.x; n.child[5] = c1.y; n.child[6] = c1.z; n.child[7] = c1.w;
    return n;
}

__device__ __forceinline__ GNode gload(const GNode* p)
{
    const YG_GLOBAL yg_v4f* q = (const YG_GLOBAL yg_v4f*)p;
    const yg_v4f a = q[0], b = q[1], c = q[2];
    const yg_v4i d = ((const YG_GLOBAL yg_v4i*)p)[3];
    GNode n;
    n.a = make_float4(a.x, a.y, a.z, a.w);
    n.b = make_float4(b.x, b.y, b.z, b.w);
    n.c = make_float4(c.x, c.y, c.z, c.w);
    n.d = make_int4(d.x, d.y, d.z, d.w);
    return n;
}

__device__ __forceinline__ void load_tri(const float4* tris, int ti, f3& a, f3& b, f3& c, f3& n)
{
    const float4 A = gload(tris + 3 * ti), B = gload(tris + 3 * ti + 1), C = gload(tris + 3 * ti + 2);
    a = mk3(A.x, A.y, A.z);
    b = mk3(B.x, B.y, B.z);
    c = mk3(C.x, C.y, C.z);
    n = mk3(A.w, B.w, C.w);
}

// Per-lane traversal stack in LDS: entry k of lane t at lds[k*nlanes + t]
// (conflict free).  Launches size the dynamic LDS for the deepest tree they
// traverse, so no entry ever goes to private (scratch) memory.
// A lane's traversal stack: the first `lcap` entries in LDS ([entry][lane],
// conflict free), deeper ones (rare) in global memory ([entry][thread], one
// slot per thread of the launch).  A short LDS part lets more waves share a
// compute unit; lcap = INT_MAX keeps the whole stack in LDS.
struct LaneStack {
    int* lds;
    int lane, nlanes, sp;
    int lcap = 0x7fffffff;
    int* ovf = nullptr;
    uint32_t ostride = 0;
    __device__ void init(int* l, int ln, int nl)
    {
        lds = l; lane = ln; nlanes = nl; sp = 0;
    }
    // entries from cap on go to ovf + (entry - cap) * stride (cap 0: none)
    __device__ void spill_to(int cap, int* o, uint32_t stride)
    {
        if (cap > 0 && o) { lcap = cap; ovf = o; ostride = stride; }
    }
    __device__ void push(int v)
    {
        if (sp < lcap) lds[sp * nlanes + lane] = v;
        else ovf[(size_t)(sp - lcap) * ostride] = v;
        sp++;
    }
    __device__ int pop()
    {
        sp--;
        return (sp < lcap) ? lds[sp * nlanes + lane] : ovf[(size_t)(sp - lcap) * ostride];
    }
};

struct RayBox {
    float ix, iy, iz, ox, oy, oz;  // 1/d and o/d for the slab test
};

__device__ __forceinline__ RayBox make_raybox(f3 o, f3 d)
{
    RayBox r;
    // a zero component gets a tiny same-signed stand-in: that slab then
    // accepts everything on the origin's side, like bound_t::cross skipping the axis
    float dx = (fabsf(d.x) > 1e-30f) ? d.x : copysignf(1e-30f, d.x);
    float dy = (fabsf(d.y) > 1e-30f) ? d.y : copysignf(1e-30f, d.y);
    float dz = (fabsf(d.z) > 1e-30f) ? d.z : copysignf(1e-30f, d.z);
    r.ix = 1.0f / dx; r.iy = 1.0f / dy; r.iz = 1.0f / dz;
    r.ox = o.x * r.ix; r.oy = o.y * r.iy; r.oz = o.z * r.iz;
    return r;
}

// Relative slack of the slab interval: along grazing rays the float plane
// distance of an accepted triangle can fall a little outside its box's
// [near, far] interval, or just beyond the current closest distance.
#define YG_TEPS 6.103515625e-05f  // 2^-14
// Relative slack of the distance bound a box must start before (the current
// closest hit, or a shadow ray's length): near a shared vertex a grazing ray
// can pass the float edge test of a triangle at a plane distance in front of
// where it enters that triangle's box.
#define YG_HEPS 9.765625e-04f  // 2^-10
__device__ __forceinline__ float cull_bound(float t) { return __builtin_fmaf(fabsf(t), YG_HEPS, t); }

// Slab test of both children of a node.  Boxes are conservatively inflated at
// build time and the interval is widened by YG_TEPS, so fused multiply-adds
// here cannot drop a hit.
__device__ __forceinline__ void node_test(const GNode& n, const RayBox& r, float lo, float hi,
                                          float& t0, float& t1, bool& h0, bool& h1)
{
    float a0 = __builtin_fmaf(n.a.x, r.ix, -r.ox), a1 = __builtin_fmaf(n.a.y, r.ix, -r.ox);
    float a2 = __builtin_fmaf(n.a.z, r.iy, -r.oy), a3 = __builtin_fmaf(n.a.w, r.iy, -r.oy);
    float a4 = __builtin_fmaf(n.c.x, r.iz, -r.oz), a5 = __builtin_fmaf(n.c.y, r.iz, -r.oz);
    float b0 = __builtin_fmaf(n.b.x, r.ix, -r.ox), b1 = __builtin_fmaf(n.b.y, r.ix, -r.ox);
    float b2 = __builtin_fmaf(n.b.z, r.iy, -r.oy), b3 = __builtin_fmaf(n.b.w, r.iy, -r.oy);
    float b4 = __builtin_fmaf(n.c.z, r.iz, -r.oz), b5 = __builtin_fmaf(n.c.w, r.iz, -r.oz);
    float n0 = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), lo));
    float f0 = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fminf(fmaxf(a4, a5), hi));
    float n1 = fmaxf(fmaxf(fminf(b0, b1), fminf(b2, b3)), fmaxf(fminf(b4, b5), lo));
    float f1 = fminf(fminf(fmaxf(b0, b1), fmaxf(b2, b3)), fminf(fmaxf(b4, b5), hi));
    h0 = __builtin_fmaf(-fabsf(n0), YG_TEPS, n0) <= __builtin_fmaf(fabsf(f0), YG_TEPS, f0);
    h1 = __builtin_fmaf(-fabsf(n1), YG_TEPS, n1) <= __builtin_fmaf(fabsf(f1), YG_TEPS, f1);
    t0 = n0;
    t1 = n1;
}

// Slab test of the four children of a width-4 node: t[k] = entry distance of
// child k, or +inf when the ray misses it (also for the +inf empty slots,
// whose widened test evaluates to NaN and fails).
__device__ __forceinline__ void node_test4(const GNode4& n, const RayBox& r, float lo, float hi, float t[4])
{
    const float* lx = &n.lox.x; const float* hx = &n.hix.x;
    const float* ly = &n.loy.x; const float* hy = &n.hiy.x;
    const float* lz = &n.loz.x; const float* hz = &n.hiz.x;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float a0 = __builtin_fmaf(lx[k], r.ix, -r.ox), a1 = __builtin_fmaf(hx[k], r.ix, -r.ox);
        const float a2 = __builtin_fmaf(ly[k], r.iy, -r.oy), a3 = __builtin_fmaf(hy[k], r.iy, -r.oy);
        const float a4 = __builtin_fmaf(lz[k], r.iz, -r.oz), a5 = __builtin_fmaf(hz[k], r.iz, -r.oz);
        const float nr = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), lo));
        const float fr = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fminf(fmaxf(a4, a5), hi));
        const bool h = __builtin_fmaf(-fabsf(nr), YG_TEPS, nr) <= __builtin_fmaf(fabsf(fr), YG_TEPS, fr);
        t[k] = h ? nr : INFINITY;
    }
}

// Slab test of the eight children of a width-8 node (as node_test4).
__device__ __forceinline__ void node_test8(const GNode8& n, const RayBox& r, float lo, float hi, float t[8])
{
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float a0 = __builtin_fmaf(n.box[0][k], r.ix, -r.ox), a1 = __builtin_fmaf(n.box[1][k], r.ix, -r.ox);
        const float a2 = __builtin_fmaf(n.box[2][k], r.iy, -r.oy), a3 = __builtin_fmaf(n.box[3][k], r.iy, -r.oy);
        const float a4 = __builtin_fmaf(n.box[4][k], r.iz, -r.oz), a5 = __builtin_fmaf(n.box[5][k], r.iz, -r.oz);
        const float nr = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), lo));
        const float fr = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fminf(fmaxf(a4, a5), hi));
        const bool h = __builtin_fmaf(-fabsf(nr), YG_TEPS, nr) <= __builtin_fmaf(fabsf(fr), YG_TEPS, fr);
        t[k] = h ? nr : INFINITY;
    }
}

#define YG_POP 0x7fffffff  // "no child to descend into": pop the stack

__device__ __forceinline__ void cswap(float& ta, int& ca, float& tb, int& cb)
{
    const bool s = tb < ta;
    const float t = s ? tb : ta;
    const int c = s ? cb : ca;
    tb = s ? ta : tb;
    cb = s ? ca : cb;
    ta = t;
    ca = c;
}

// One inner-node visit: the hit children nearest first, the nearest
// returned, the others pushed far-to-near (so the next nearest pops first).
__device__ __forceinline__ int visit2(const BVHView& bvh, int node, const RayBox& rb, float lo, float hi,
                                      LaneStack& st)
{
    const GNode n = gload(bvh.nodes + node);
    float t0, t1;
    bool h0, h1;
    node_test(n, rb, lo, hi, t0, t1, h0, h1);
    if (h0 && h1) {
        const bool first0 = t0 <= t1;
        st.push(first0 ? n.d.y : n.d.x);
        return first0 ? n.d.x : n.d.y;
    }
    if (h0) return n.d.x;
    if (h1) return n.d.y;
    return YG_POP;
}

__device__ __forceinline__ int visit4(const BVHView& bvh, int node, const RayBox& rb, float lo, float hi,
                                      LaneStack& st)
{
    const GNode4 n = gload(bvh.nodes4 + node);
    float t[4];
    node_test4(n, rb, lo, hi, t);
    float t0 = t[0], t1 = t[1], t2 = t[2], t3 = t[3];
    int c0 = n.child.x, c1 = n.child.y, c2 = n.child.z, c3 = n.child.w;
    cswap(t0, c0, t1, c1);
    cswap(t2, c2, t3, c3);
    cswap(t0, c0, t2, c2);
    cswap(t1, c1, t3, c3);
    cswap(t1, c1, t2, c2);
    if (t0 == INFINITY) return YG_POP;
    if (t3 != INFINITY) st.push(c3);
    if (t2 != INFINITY) st.push(c2);
    if (t1 != INFINITY) st.push(c1);
    return c0;
}

// Eight children: sorted nearest first by a 19-comparator network (depth 6),
// the nearest returned, the other hit children pushed far-to-near.
__device__ __forceinline__ int visit8(const BVHView& bvh, int node, const RayBox& rb, float lo, float hi,
                                      LaneStack& st)
{
    const GNode8 n = gload(bvh.nodes8 + node);
    float t[8];
    node_test8(n, rb, lo, hi, t);
    int c[8];
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = n.child[k];
#define YG_CS(i, j) cswap(t[i], c[i], t[j], c[j])
    YG_CS(0, 2); YG_CS(1, 3); YG_CS(4, 6); YG_CS(5, 7);
    YG_CS(0, 4); YG_CS(1, 5); YG_CS(2, 6); YG_CS(3, 7);
    YG_CS(0, 1); YG_CS(2, 3); YG_CS(4, 5); YG_CS(6, 7);
    YG_CS(2, 4); YG_CS(3, 5);
    YG_CS(1, 4); YG_CS(3, 6);
    YG_CS(1, 2); YG_CS(3, 4); YG_CS(5, 6);
#undef YG_CS
    if (t[0] == INFINITY) return YG_POP;
#pragma unroll
    for (int k = 7; k >= 1; k--)
        if (t[k] != INFINITY) st.push(c[k]);
    return c[0];
}

// Kernels are instantiated per BVH width W (2, 4 or 8); a launch traverses
// trees of one width only.  Keeping the two node formats in separate kernels
// keeps each traversal loop small (one binary with both loops ran a third
// slower on either).
template <int W>
__device__ __forceinline__ int visit(const BVHView& bvh, int node, const RayBox& rb, float lo, float hi,
                                     LaneStack& st)
{
    if constexpr ((W & 15) == 8) return visit8(bvh, node, rb, lo, hi, st);
    else if constexpr ((W & 15) == 4) return visit4(bvh, node, rb, lo, hi, st);
    else return visit2(bvh, node, rb, lo, hi, st);
}

struct TraceStats {
    uint32_t nodes, tris;
};

// The traversal kernels' W template parameter is the BVH width (2, 4 or 8),
// plus 16 for launches with objects that cast no shadows: only those
// instantiations test object flags in their shadow queries, so the leaf loops
// of every other launch stay as they are.
#define YG_W_NOCAST 16
template <int W>
constexpr bool yg_nocast_w() { return (W & YG_W_NOCAST) != 0; }

// Whether a shadow query in this wave must look at object flags (lanes of
// scenes whose objects all cast shadows read flags that never skip).
template <int W>
__device__ __forceinline__ bool wave_nocast(const BVHView& bvh, bool shadow)
{
    if constexpr (!yg_nocast_w<W>()) return false;
    else return shadow && __ballot(bvh.nocast != 0) != 0ull;
}

// Whether a scene shadow query (scene_t::isShadowed, and firstHit with its
// shadow flag, scene.cc:698) skips triangle ti: its object casts no shadows.
__device__ __forceinline__ bool skips_shadow(const BVHView& bvh, int ti)
{
    const int o = (bvh.nobj == 1) ? 0 : gload(bvh.tobj + ti);
    return !(gload(bvh.oflags + o) & 1);
}

// Closest hit: accepts tmin <= t < tbest (kdtree.cc:690-694); exact ties go to
// the lowest primitive id so the result does not depend on the tree.
// Returns the leaf-ordered triangle index or -1; tbest holds the distance.
// shadow: a scene shadow query (objects that cast no shadows are skipped).
template <bool COUNT, int W>
__device__ int trace_closest(const BVHView& bvh, f3 o, f3 d, float tmin, float& tbest,
                             LaneStack& st, TraceStats* cs = nullptr, bool shadow = false)
{
    const bool skip = wave_nocast<W>(bvh, shadow);
    const RayBox rb = make_raybox(o, d);
    const float lo = tmin - 1e-4f * fabsf(tmin) - 1e-30f;
    int best = -1;
    int bestprim = 0x7fffffff;
    float tcull = cull_bound(tbest);
    int node = 0;
    st.sp = 0;
    for (;;) {
        if (node >= 0) {
            if (COUNT) cs->nodes++;
            node = visit<W>(bvh, node, rb, lo, tcull, st);
            if (node != YG_POP) continue;
        } else {
            const int code = ~node;
            const int first = code >> 4, cnt = (code & 15) + 1;
            for (int k = 0; k < cnt; k++) {
                const int ti = first + k;
                if (skip && skips_shadow(bvh, ti)) continue;
                f3 a, b, c, nn;
                load_tri(bvh.tris, ti, a, b, c, nn);
                if (COUNT) cs->tris++;
                if (!tri_hit(a, b, c, nn, o, d)) continue;
                const float t = tri_t(a, nn, o, d);
                if (!(t >= tmin)) continue;
                if (t < tbest) {
                    tbest = t;
                    tcull = cull_bound(t);
                    best = ti;
                    bestprim = gload(bvh.tprim + ti);
                } else if (t == tbest && best >= 0) {
                    const int p = gload(bvh.tprim + ti);
                    if (p < bestprim) { best = ti; bestprim = p; }
                }
            }
        }
        if (st.sp == 0) break;
        node = st.pop();
    }
    return best;
}

// Any hit: accepts tmin < t < tmax (kdtree.cc:835-841).  Triangles of object
// self_obj are tested from oself instead of o (scene_t::isShadowed's
// self_bias rule, scene.cc:115-141); [blo, bhi] is the box interval in o's
// parameter and must cover both origins' segments.
template <bool COUNT, int W>
__device__ bool trace_any(const BVHView& bvh, f3 o, f3 d, float tmin, float tmax, int self_obj,
                          f3 oself, float blo, float bhi, LaneStack& st, TraceStats* cs = nullptr,
                          bool shadow = false)
{
    const bool skip = wave_nocast<W>(bvh, shadow);
    const RayBox rb = make_raybox(o, d);
    const float hi = cull_bound(bhi);
    int node = 0;
    st.sp = 0;
    for (;;) {
        if (node >= 0) {
            if (COUNT) cs->nodes++;
            node = visit<W>(bvh, node, rb, blo, hi, st);
            if (node != YG_POP) continue;
        } else {
            const int code = ~node;
            const int first = code >> 4, cnt = (code & 15) + 1;
            for (int k = 0; k < cnt; k++) {
                const int ti = first + k;
                if (skip && skips_shadow(bvh, ti)) continue;
                f3 a, b, c, nn;
                load_tri(bvh.tris, ti, a, b, c, nn);
                if (COUNT) cs->tris++;
                // (one-object scenes need no per-triangle object id)
                const bool same = self_obj >= 0 && (bvh.nobj == 1 ? self_obj == 0 : gload(bvh.tobj + ti) == self_obj);
                const f3 org = same ? oself : o;
                if (!tri_hit(a, b, c, nn, org, d)) continue;
                const float t = tri_t(a, nn, org, d);
                if (t > tmin && t < tmax) return true;
            }
        }
        if (st.sp == 0) break;
        node = st.pop();
    }
    return false;
}

// scene_t::firstHit (scene.cc:688-721): origin from + ray*min_raydis, closest
// hit with Z > 0 (a zero-distance closest hit is rejected by the scene).
// shadow: firstHit's shadow flag (objects that cast no shadows are skipped).
template <int W>
__device__ __forceinline__ int first_hit(const BVHView& bvh, f3 from, f3 ray, float min_raydis,
                                         float& t, LaneStack& st, bool shadow = false)
{
    const f3 f = from + ray * min_raydis;
    t = INFINITY;
    int ti = trace_closest<false, W>(bvh, f, ray, 0.0f, t, st, nullptr, shadow);
    if (ti >= 0 && !(t > 0.0f)) ti = -1;
    return ti;
}

// scene_t::isShadowed, point form (scene.cc:107-146) or direction form (:148-185)
template <int W>
__device__ __forceinline__ bool is_shadowed(const BVHView& bvh, f3 p, f3 l, int self_obj,
                                            bool to_point, float min_raydis, float self_bias,
                                            LaneStack& st)
{
    f3 ray;
    float dist;
    if (to_point) {
        ray = l - p;
        dist = len3(ray);
    } else {
        ray = l;
        dist = INFINITY;
    }
    ray = norm3(ray);
    const f3 self = p + ray * self_bias;
    const f3 q = p + ray * min_raydis;
    const float delta = self_bias - min_raydis;
    const float slack = 1e-4f * (fabsf(self_bias) + fabsf(min_raydis)) + 1e-30f;
    float blo = fminf(0.0f, delta) - slack;
    float bhi = dist + fmaxf(0.0f, delta) + slack;
    return trace_any<false, W>(bvh, q, ray, 0.0f, dist, self_obj, self, blo, bhi, st, nullptr, true);
}
