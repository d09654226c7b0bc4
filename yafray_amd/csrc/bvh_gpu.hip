// Device BVH builder for libyafgpu: the GPU counterpart of the host binned-SAH
// builder (bvh_build.cpp), in the role of kdTree_t::kdTree_t / buildTree
// (kdtree.cc:57-582).  It writes the same node formats (widths 2, 4 and 8),
// the same leaf references and the same leaf-ordered triangle records, so the
// traversal kernels cannot tell the builders apart; the tree shape changes
// only speed, never results (boxes are widened as the host builder widens
// them, inflate_axis).
//
// Build (one scene, all on the device):
//  1. triangle boxes, the bounds of their centres, a 63-bit Morton code of
//     each centre, radix sort of the triangles by code;
//  2. Parallel Locally-Ordered Clustering (Meister and Bittner, "Parallel
//     Locally-Ordered Clustering for Bounding Volume Hierarchy Construction",
//     IEEE TVCG 2018): every cluster finds, among the PLOC_R clusters on
//     either side of it in curve order, the one whose union box has the least
//     surface area; mutual nearest neighbours merge into a new node; the
//     survivors are compacted in curve order; repeat until one cluster is
//     left.  Distance ties are broken by a hash of the pair, then by the
//     pair's positions, so the globally closest pair is always mutual and
//     every round merges;
//  3. bottom-up, per merge round: triangle counts and the SAH cost with unit
//     traversal and intersection costs, a subtree of at most max_leaf
//     triangles becoming a leaf where that is cheaper than its split;
//  4. top-down, rounds in reverse: each node's first triangle in depth-first
//     leaf order and its depth-first index among the inner nodes;
//  5. export: triangle records and ids scattered to leaf order, two-wide
//     nodes in depth-first order, and for widths 4 and 8 the level-by-level
//     collapse of the host builder (the largest-area inner child opened until
//     the node is full), nodes laid out level by level.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "yg_internal.h"

namespace {

#ifndef YG_PLOC_R
#define YG_PLOC_R 16
#endif
constexpr int PLOC_R = YG_PLOC_R;  // nearest-neighbour search radius in curve order
constexpr int BLK = 256;

template <class T>
struct DBuf {
    T* p = nullptr;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(n, 1)); }
    T* release()
    {
        T* q = p;
        p = nullptr;
        return q;
    }
};

#define YG_ALLOC(buf, n)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (buf).alloc(n);                                                            \
        if (e_ != hipSuccess) return yg_fail(YG_ERR_NOMEM, std::string("device BVH build: hipMalloc: ") + \
                                              hipGetErrorString(e_));                              \
    } while (0)

// ------------------------------------------------------------- kernels --

__global__ void k_prim_boxes(const float* __restrict__ verts, const uint32_t* __restrict__ idx, int n,
                             float4* __restrict__ tlo, float4* __restrict__ thi)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int k = 0; k < 3; k++) {
        const float* v = verts + 3 * (size_t)idx[3 * (size_t)t + k];
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(lo[a], v[a]);
            hi[a] = fmaxf(hi[a], v[a]);
        }
    }
    tlo[t] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    thi[t] = make_float4(hi[0], hi[1], hi[2], 0.0f);
}

__device__ __forceinline__ float4 centre(float4 lo, float4 hi)
{
    return make_float4(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z), 0.0f);
}

// bounds of the triangle centres: per-block partials, then one block
__global__ __launch_bounds__(BLK) void k_cbounds_partial(const float4* __restrict__ tlo,
                                                         const float4* __restrict__ thi, int n,
                                                         float* __restrict__ part)
{
    __shared__ float sm[6][BLK];
    float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int t = blockIdx.x * BLK + threadIdx.x; t < n; t += gridDim.x * BLK) {
        const float4 c = centre(tlo[t], thi[t]);
        v[0] = fminf(v[0], c.x); v[1] = fminf(v[1], c.y); v[2] = fminf(v[2], c.z);
        v[3] = fmaxf(v[3], c.x); v[4] = fmaxf(v[4], c.y); v[5] = fmaxf(v[5], c.z);
    }
    for (int a = 0; a < 6; a++) sm[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int s = BLK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 6; a++)
                sm[a][threadIdx.x] = a < 3 ? fminf(sm[a][threadIdx.x], sm[a][threadIdx.x + s])
                                           : fmaxf(sm[a][threadIdx.x], sm[a][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = sm[threadIdx.x][0];
}

__global__ __launch_bounds__(BLK) void k_cbounds_final(const float* __restrict__ part, int nb, float* __restrict__ out)
{
    __shared__ float sm[6][BLK];
    float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int b = threadIdx.x; b < nb; b += BLK)
        for (int a = 0; a < 6; a++) v[a] = a < 3 ? fminf(v[a], part[6 * b + a]) : fmaxf(v[a], part[6 * b + a]);
    for (int a = 0; a < 6; a++) sm[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int s = BLK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 6; a++)
                sm[a][threadIdx.x] = a < 3 ? fminf(sm[a][threadIdx.x], sm[a][threadIdx.x + s])
                                           : fmaxf(sm[a][threadIdx.x], sm[a][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = sm[threadIdx.x][0];
}

// 21 bits spread to every third bit of a 64-bit word
__device__ __forceinline__ uint64_t spread21(uint32_t x)
{
    uint64_t v = x & 0x1fffffu;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

__global__ void k_morton(const float4* __restrict__ tlo, const float4* __restrict__ thi, int n,
                         const float* __restrict__ cb, uint64_t* __restrict__ keys, int32_t* __restrict__ vals)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float4 c = centre(tlo[t], thi[t]);
    const float cc[3] = {c.x, c.y, c.z};
    // one scale for all axes (the bounds' largest extent): the curve keeps
    // the scene's proportions, so neighbours along it are neighbours in space
    const float ext = fmaxf(fmaxf(cb[3] - cb[0], cb[4] - cb[1]), cb[5] - cb[2]);
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        float u = ext > 0.0f ? (cc[a] - cb[a]) / ext : 0.0f;
        u = fminf(fmaxf(u * 2097151.0f, 0.0f), 2097151.0f);  // NaN -> 0
        q[a] = (uint32_t)u;
    }
    keys[t] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    vals[t] = t;
}

// leaves 0..n-1: the sorted triangles, one per leaf
__global__ void k_leaf_init(const int32_t* __restrict__ sp, int n, const float4* __restrict__ tlo,
                            const float4* __restrict__ thi, float4* __restrict__ nlo, float4* __restrict__ nhi,
                            int32_t* __restrict__ cl, int32_t* __restrict__ cnt, float* __restrict__ cost,
                            uint8_t* __restrict__ leaf, int32_t* __restrict__ ninner)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = sp[i];
    nlo[i] = tlo[p];
    nhi[i] = thi[p];
    cl[i] = i;
    cnt[i] = 1;
    cost[i] = 1.0f;
    leaf[i] = 1;
    ninner[i] = 0;
}

// half the surface area of the union of two boxes
__device__ __forceinline__ float union_area(float4 alo, float4 ahi, float4 blo, float4 bhi)
{
    const float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
    const float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
    const float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
    const float a = dx * dy + dy * dz + dz * dx;
    return a == a ? a : INFINITY;
}

// Symmetric hash of a pair of positions (lo < hi): breaks distance ties.
__device__ __forceinline__ uint32_t pair_hash(uint32_t lo, uint32_t hi)
{
    uint32_t h = (lo * 0x9E3779B1u) ^ (hi * 0x85EBCA77u);
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return h;
}

// Nearest neighbour of every cluster within PLOC_R positions, by the total
// order (distance, pair hash, pair positions): candidates are scanned in
// increasing position and replace the best only when strictly smaller in
// (distance, hash), so equal keys keep the lowest pair.  The globally
// smallest pair is then mutual (every round merges), and the hash spreads
// the merges of equal distances (a row of equal boxes, coincident
// triangles) over the sequence instead of merging one pair per round.
__global__ __launch_bounds__(BLK) void k_nn(const int32_t* __restrict__ cl, int m, const float4* __restrict__ nlo,
                                            const float4* __restrict__ nhi, int32_t* __restrict__ nn)
{
    __shared__ float4 slo[BLK + 2 * PLOC_R], shi[BLK + 2 * PLOC_R];
    const int b0 = (int)blockIdx.x * BLK - PLOC_R;
    for (int q = threadIdx.x; q < BLK + 2 * PLOC_R; q += BLK) {
        const int j = b0 + q;
        if (j >= 0 && j < m) {
            const int c = cl[j];
            slo[q] = nlo[c];
            shi[q] = nhi[c];
        }
    }
    __syncthreads();
    const int i = (int)blockIdx.x * BLK + threadIdx.x;
    if (i >= m) return;
    const float4 alo = slo[i - b0], ahi = shi[i - b0];
    const int j0 = max(0, i - PLOC_R), j1 = min(m - 1, i + PLOC_R);
    float best = INFINITY;
    uint32_t bh = 0;
    int bj = -1;
    for (int j = j0; j <= j1; j++) {
        if (j == i) continue;
        const float d = union_area(alo, ahi, slo[j - b0], shi[j - b0]);
        const uint32_t h = pair_hash((uint32_t)min(i, j), (uint32_t)max(i, j));
        if (bj < 0 || d < best || (d == best && h < bh)) {
            best = d;
            bh = h;
            bj = j;
        }
    }
    nn[i] = bj;
}

__global__ void k_merge_flags(const int32_t* __restrict__ nn, int m, uint32_t* __restrict__ f)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    f[i] = (j > i && nn[j] == i) ? 1u : 0u;
}

// fincl: inclusive scan of the merge flags.  A merging pair (i < j) writes its
// new node at next + fincl[i] - 1 into position i; position j is dropped.
__global__ void k_merge(const int32_t* __restrict__ cl, const int32_t* __restrict__ nn,
                        const uint32_t* __restrict__ fincl, int m, int next, float4* __restrict__ nlo,
                        float4* __restrict__ nhi, int32_t* __restrict__ left, int32_t* __restrict__ right,
                        int32_t* __restrict__ cl2, uint32_t* __restrict__ keep)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = nn[j] == i;
    if (mutual && i < j) {
        const int id = next + (int)fincl[i] - 1;
        const int a = cl[i], b = cl[j];
        const float4 alo = nlo[a], ahi = nhi[a], blo = nlo[b], bhi = nhi[b];
        nlo[id] = make_float4(fminf(alo.x, blo.x), fminf(alo.y, blo.y), fminf(alo.z, blo.z), 0.0f);
        nhi[id] = make_float4(fmaxf(ahi.x, bhi.x), fmaxf(ahi.y, bhi.y), fmaxf(ahi.z, bhi.z), 0.0f);
        left[id] = a;
        right[id] = b;
        cl2[i] = id;
        keep[i] = 1;
    } else if (mutual) {
        cl2[i] = -1;
        keep[i] = 0;
    } else {
        cl2[i] = cl[i];
        keep[i] = 1;
    }
}

__global__ void k_compact(const int32_t* __restrict__ cl2, const uint32_t* __restrict__ keep,
                          const uint32_t* __restrict__ kincl, int m, int32_t* __restrict__ cl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    if (keep[i]) cl[kincl[i] - 1] = cl2[i];
}

__device__ __forceinline__ float half_area(float4 lo, float4 hi)
{
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

// One merge round's nodes [a, b): their children were made in earlier rounds.
__global__ void k_bottom_up(int a, int b, const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                            const float4* __restrict__ nlo, const float4* __restrict__ nhi, int max_leaf,
                            int32_t* __restrict__ cnt, float* __restrict__ cost, uint8_t* __restrict__ leaf,
                            int32_t* __restrict__ ninner)
{
    const int k = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b) return;
    const int l = left[k], r = right[k];
    const int c = cnt[l] + cnt[r];
    const float A = half_area(nlo[k], nhi[k]);
    const float sub = half_area(nlo[l], nhi[l]) * cost[l] + half_area(nlo[r], nhi[r]) * cost[r];
    const float split = A > 0.0f ? 1.0f + sub / A : INFINITY;
    const bool lf = c <= max_leaf && (float)c <= split;
    cnt[k] = c;
    leaf[k] = lf ? 1 : 0;
    cost[k] = lf ? (float)c : (A > 0.0f ? split : 1.0f + cost[l] + cost[r]);
    ninner[k] = lf ? 0 : 1 + ninner[l] + ninner[r];
}

__global__ void k_root_init(int root, int32_t* off, int32_t* depth, uint8_t* under, int32_t* didx)
{
    off[root] = 0;
    depth[root] = 0;
    under[root] = 0;
    didx[root] = 0;
}

// One merge round's nodes [a, b), rounds taken last to first: a node's
// values are final before its children's are derived from them.
__global__ void k_top_down(int a, int b, const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                           const int32_t* __restrict__ cnt, const uint8_t* __restrict__ leaf,
                           const int32_t* __restrict__ ninner, int32_t* __restrict__ off, int32_t* __restrict__ depth,
                           uint8_t* __restrict__ under, int32_t* __restrict__ didx, int32_t* __restrict__ maxdepth)
{
    const int k = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b) return;
    const int l = left[k], r = right[k];
    const int o = off[k], d = depth[k];
    const uint8_t u = (uint8_t)(under[k] | leaf[k]);
    off[l] = o;
    off[r] = o + cnt[l];
    depth[l] = d + 1;
    depth[r] = d + 1;
    under[l] = u;
    under[r] = u;
    if (!u) {  // an inner node of the exported tree: its children are exported nodes
        didx[l] = didx[k] + 1;
        didx[r] = didx[k] + 1 + ninner[l];
        atomicMax(maxdepth, d + 1);
    }
}

__global__ void k_scatter_tris(int n, const int32_t* __restrict__ sp, const int32_t* __restrict__ off,
                               const float* __restrict__ verts, const uint32_t* __restrict__ idx,
                               const int32_t* __restrict__ obj_in, const int32_t* __restrict__ mat_in,
                               const float* __restrict__ vn_in, const uint8_t* __restrict__ smooth_in,
                               float4* __restrict__ tris, int32_t* __restrict__ tprim, int32_t* __restrict__ tobj,
                               int32_t* __restrict__ tmat, float4* __restrict__ tvn)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = sp[i];
    const size_t pos = off ? (size_t)off[i] : (size_t)i;
    const f3 a = ld3(verts + 3 * (size_t)idx[3 * (size_t)p]);
    const f3 b = ld3(verts + 3 * (size_t)idx[3 * (size_t)p + 1]);
    const f3 c = ld3(verts + 3 * (size_t)idx[3 * (size_t)p + 2]);
    const f3 nr = tri_normal(a, b, c);
    tris[3 * pos] = make_float4(a.x, a.y, a.z, nr.x);
    tris[3 * pos + 1] = make_float4(b.x, b.y, b.z, nr.y);
    tris[3 * pos + 2] = make_float4(c.x, c.y, c.z, nr.z);
    tprim[pos] = p;
    tobj[pos] = obj_in[p];
    tmat[pos] = mat_in[p];
    if (tvn) {
        const float* v = vn_in + 9 * (size_t)p;
        tvn[3 * pos] = make_float4(v[0], v[1], v[2], smooth_in[p] ? 1.0f : 0.0f);
        tvn[3 * pos + 1] = make_float4(v[3], v[4], v[5], 0.0f);
        tvn[3 * pos + 2] = make_float4(v[6], v[7], v[8], 0.0f);
    }
}

// the host builder's box widening (bvh_build.cpp inflate), per axis
__device__ __forceinline__ void inflate_axis(float lo, float hi, float G, float& olo, float& ohi)
{
    const float ext = hi - lo;
    const float e = (fabsf(lo) + fabsf(hi) + ext + G) * 3.814697265625e-06f + 1e-30f;
    olo = lo - e;
    ohi = hi + e;
}

__device__ __forceinline__ int32_t leaf_ref(int32_t first, int32_t count)
{
    return ~(first * YG_LEAF_MAX + (count - 1));
}

__device__ __forceinline__ int32_t child_ref(int c, const uint8_t* leaf, const int32_t* off, const int32_t* cnt,
                                             const int32_t* didx)
{
    return leaf[c] ? leaf_ref(off[c], cnt[c]) : didx[c];
}

__global__ void k_export2(int n, int total, const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                          const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                          const int32_t* __restrict__ cnt, const int32_t* __restrict__ off,
                          const int32_t* __restrict__ didx, const uint8_t* __restrict__ leaf,
                          const uint8_t* __restrict__ under, float G, GNode* __restrict__ out)
{
    const int k = n + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total || (under && under[k]) || leaf[k]) return;
    const int ch[2] = {left[k], right[k]};
    float lo[2][3], hi[2][3];
    int32_t ref[2];
    for (int q = 0; q < 2; q++) {
        const float4 a = nlo[ch[q]], b = nhi[ch[q]];
        inflate_axis(a.x, b.x, G, lo[q][0], hi[q][0]);
        inflate_axis(a.y, b.y, G, lo[q][1], hi[q][1]);
        inflate_axis(a.z, b.z, G, lo[q][2], hi[q][2]);
        ref[q] = child_ref(ch[q], leaf, off, cnt, didx);
    }
    GNode g;
    g.a = make_float4(lo[0][0], hi[0][0], lo[0][1], hi[0][1]);
    g.b = make_float4(lo[1][0], hi[1][0], lo[1][1], hi[1][1]);
    g.c = make_float4(lo[0][2], hi[0][2], lo[1][2], hi[1][2]);
    g.d = make_int4(ref[0], ref[1], 0, 0);
    out[didx[k]] = g;
}

// Children of a wide node: the binary node's two, then the largest-area
// inner child replaced by its own two (left-to-right order kept) until there
// are N or only leaves remain (bvh_build.cpp export_wide).
template <int N>
__global__ void k_wide_children(int nw, const int32_t* __restrict__ wb, const int32_t* __restrict__ left,
                                const int32_t* __restrict__ right, const uint8_t* __restrict__ leaf,
                                const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                int32_t* __restrict__ ch, int32_t* __restrict__ nch, uint32_t* __restrict__ ninn)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const int b = wb[w];
    int c[N];
    int n = 2;
    c[0] = left[b];
    c[1] = right[b];
    while (n < N) {
        int best = -1;
        double ba = -1.0;
        for (int k = 0; k < n; k++) {
            if (leaf[c[k]]) continue;
            const float4 lo = nlo[c[k]], hi = nhi[c[k]];
            const double dx = (double)hi.x - lo.x, dy = (double)hi.y - lo.y, dz = (double)hi.z - lo.z;
            const double ar = (dx < 0 || dy < 0 || dz < 0) ? 0.0 : 2.0 * (dx * dy + dy * dz + dz * dx);
            if (ar > ba) {
                ba = ar;
                best = k;
            }
        }
        if (best < 0) break;
        const int x = c[best];
        for (int k = n; k > best + 1; k--) c[k] = c[k - 1];
        c[best] = left[x];
        c[best + 1] = right[x];
        n++;
    }
    uint32_t inner = 0;
    for (int k = 0; k < N; k++) {
        ch[(size_t)w * N + k] = k < n ? c[k] : -1;
        if (k < n && !leaf[c[k]]) inner++;
    }
    nch[w] = n;
    ninn[w] = inner;
}

// Writes the wide nodes of one level and lists the next level's.  Node
// layout: float rows[6][N] (lo.x hi.x lo.y hi.y lo.z hi.z), int32 child[N],
// then padding (GNode4 and GNode8 alike).
template <int N, class Node>
__global__ void k_wide_write(int nw, const int32_t* __restrict__ wi, const int32_t* __restrict__ wp,
                             const int32_t* __restrict__ ch, const int32_t* __restrict__ nch,
                             const uint32_t* __restrict__ ninn, const uint32_t* __restrict__ iincl, int base,
                             const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                             const uint8_t* __restrict__ leaf, const int32_t* __restrict__ off,
                             const int32_t* __restrict__ cnt, float G, Node* __restrict__ out,
                             int32_t* __restrict__ nb, int32_t* __restrict__ ni, int32_t* __restrict__ np,
                             int32_t* __restrict__ stackmax)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const int n = nch[w];
    const int pend = wp[w] + (n - 1);
    uint32_t slot = iincl[w] - ninn[w];
    float* rows = reinterpret_cast<float*>(&out[wi[w]]);
    int32_t* child = reinterpret_cast<int32_t*>(rows + 6 * N);
    for (int k = 0; k < N; k++) {
        if (k >= n) {
            for (int r = 0; r < 6; r++) rows[r * N + k] = INFINITY;
            child[k] = 0;
            continue;
        }
        const int c = ch[(size_t)w * N + k];
        const float4 a = nlo[c], b = nhi[c];
        float lo[3], hi[3];
        inflate_axis(a.x, b.x, G, lo[0], hi[0]);
        inflate_axis(a.y, b.y, G, lo[1], hi[1]);
        inflate_axis(a.z, b.z, G, lo[2], hi[2]);
        for (int ax = 0; ax < 3; ax++) {
            rows[(2 * ax) * N + k] = lo[ax];
            rows[(2 * ax + 1) * N + k] = hi[ax];
        }
        if (leaf[c]) child[k] = leaf_ref(off[c], cnt[c]);
        else {
            const int ref = base + (int)slot;
            child[k] = ref;
            nb[slot] = c;
            ni[slot] = ref;
            np[slot] = pend;
            slot++;
        }
    }
    int32_t* pad = child + N;
    const int npad = (int)(sizeof(Node) / 4) - 7 * N;
    for (int k = 0; k < npad; k++) pad[k] = 0;
    atomicMax(stackmax, pend);
}

inline unsigned blocks(int64_t n) { return (unsigned)std::max<int64_t>((n + BLK - 1) / BLK, 1); }

// inclusive prefix sum of n uint32 values (temp storage grown on demand)
struct Scanner {
    DBuf<uint8_t> tmp;
    size_t bytes = 0;
    int run(const uint32_t* in, uint32_t* out, int n)
    {
        size_t need = 0;
        YG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, need, in, out, n, (hipStream_t)0));
        if (need > bytes) {
            if (tmp.p) (void)hipFree(tmp.p);
            tmp.p = nullptr;
            YG_ALLOC(tmp, need);
            bytes = need;
        }
        YG_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, need, in, out, n, (hipStream_t)0));
        return YG_OK;
    }
};

template <int N, class Node>
int export_wide_dev(int root, int ninner, const int32_t* left, const int32_t* right, const uint8_t* leaf,
                    const float4* nlo, const float4* nhi, const int32_t* off, const int32_t* cnt, float G,
                    Scanner& scan, Node** out_nodes, int64_t& n_out, int32_t& stack)
{
    const size_t cap = (size_t)ninner + 1;
    DBuf<Node> nodes;
    DBuf<int32_t> wb[2], wi[2], wp[2], ch, nch, smax;
    DBuf<uint32_t> ninn, iincl;
    YG_ALLOC(nodes, cap);
    for (int q = 0; q < 2; q++) {
        YG_ALLOC(wb[q], cap);
        YG_ALLOC(wi[q], cap);
        YG_ALLOC(wp[q], cap);
    }
    YG_ALLOC(ch, cap * N);
    YG_ALLOC(nch, cap);
    YG_ALLOC(ninn, cap);
    YG_ALLOC(iincl, cap);
    YG_ALLOC(smax, 1);
    const int32_t zero = 0;
    YG_HIP(hipMemcpy(wb[0].p, &root, sizeof(int32_t), hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(wi[0].p, &zero, sizeof(int32_t), hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(wp[0].p, &zero, sizeof(int32_t), hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(smax.p, &zero, sizeof(int32_t), hipMemcpyHostToDevice));
    int cur = 0, nw = 1;
    int64_t total = 1;
    while (nw > 0) {
        k_wide_children<N><<<blocks(nw), BLK>>>(nw, wb[cur].p, left, right, leaf, nlo, nhi, ch.p, nch.p, ninn.p);
        YG_HIP(hipGetLastError());
        int rc = scan.run(ninn.p, iincl.p, nw);
        if (rc) return rc;
        uint32_t next = 0;
        YG_HIP(hipMemcpy(&next, iincl.p + (nw - 1), sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (total + (int64_t)next > (int64_t)cap) return yg_fail(YG_ERR_STATE, "device BVH build: wide node overflow");
        k_wide_write<N, Node><<<blocks(nw), BLK>>>(nw, wi[cur].p, wp[cur].p, ch.p, nch.p, ninn.p, iincl.p, (int)total,
                                                   nlo, nhi, leaf, off, cnt, G, nodes.p, wb[1 - cur].p,
                                                   wi[1 - cur].p, wp[1 - cur].p, smax.p);
        YG_HIP(hipGetLastError());
        total += next;
        nw = (int)next;
        cur = 1 - cur;
    }
    YG_HIP(hipMemcpy(&stack, smax.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    // exact-size copy
    DBuf<Node> exact;
    YG_ALLOC(exact, (size_t)total);
    YG_HIP(hipMemcpy(exact.p, nodes.p, sizeof(Node) * (size_t)total, hipMemcpyDeviceToDevice));
    *out_nodes = exact.release();
    n_out = total;
    return YG_OK;
}


// ------------------------------------------------- binned SAH, top down --
//
// The host builder's algorithm (bvh_build.cpp Builder::build) run level by
// level: every node of a level is binned (centres into `bins` equal bins of
// the node's centre bounds per axis, double-precision bin index), the split
// of least SAH cost is chosen by the same sweep and the same leaf rules, and
// the node's triangles are partitioned stably in place, so the node sets are
// the host's.  Where the host falls back to an object median (no valid bin
// split, or depth >= 60) the device splits the node's triangles at half
// their count in their current order.  Nodes of more than SAH_SMALL
// triangles are binned by all threads of the level (block-private LDS bins
// flushed with atomics); smaller ones by one thread each, from their
// triangles directly.

constexpr int SAH_MAXB = 64;
constexpr int SAH_SMALL = 256;
constexpr int SAH_PER = 8;  // positions per thread of the binning passes

// float <-> unsigned with the same order (atomic min / max on floats)
__device__ __forceinline__ uint32_t fenc(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fdec(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
#define ENC_PINF 0xff800000u  // fenc(+inf)
#define ENC_NINF 0x007fffffu  // fenc(-inf)

// accumulator of 12 words per big node: box lo xyz, hi xyz, centre lo xyz, hi xyz
__global__ void k_acc_init(uint32_t* __restrict__ acc, int nbig)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 12 * nbig) return;
    acc[i] = ((i % 12) % 6) < 3 ? ENC_PINF : ENC_NINF;
}

// bin accumulators: count, then box lo xyz hi xyz (6 words) per (node, axis, bin)
__global__ void k_bins_init(uint32_t* __restrict__ bcnt, uint32_t* __restrict__ bbox, int64_t nb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    bcnt[i] = 0;
    for (int c = 0; c < 6; c++) bbox[6 * i + c] = c < 3 ? ENC_PINF : ENC_NINF;
}

__device__ __forceinline__ int sah_bin(float c, float lo, double scale, int bins)
{
    int b = (int)(((double)c - lo) * scale);
    return min(max(b, 0), bins - 1);
}

// big-node slot of position i's node, or -1 (done, small, or another level)
__device__ __forceinline__ int big_slot_of(int node, int lb, int le, const int32_t* __restrict__ bslot)
{
    return (node >= lb && node < le) ? bslot[node - lb] : -1;
}

__global__ __launch_bounds__(BLK) void k_sah_bounds(int n, int lb, int le, const int32_t* __restrict__ seg,
                                                    const int32_t* __restrict__ bslot, const int32_t* __restrict__ ord,
                                                    const float4* __restrict__ tlo, const float4* __restrict__ thi,
                                                    uint32_t* __restrict__ acc)
{
    __shared__ uint32_t sacc[12];
    const int r0 = (int)blockIdx.x * BLK * SAH_PER;
    const int r1 = min(n, r0 + BLK * SAH_PER);
    const int b0 = big_slot_of(seg[r0], lb, le, bslot);
    const bool uni = b0 >= 0 && seg[r1 - 1] == seg[r0];
    if (threadIdx.x < 12) sacc[threadIdx.x] = (threadIdx.x % 6) < 3 ? ENC_PINF : ENC_NINF;
    __syncthreads();
    float v[12] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY,
                   INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    bool any = false;
    for (int i = r0 + threadIdx.x; i < r1; i += BLK) {
        const int bs = big_slot_of(seg[i], lb, le, bslot);
        if (bs < 0) continue;
        const int p = ord[i];
        const float4 lo = tlo[p], hi = thi[p];
        const float4 c = centre(lo, hi);
        if (uni) {
            v[0] = fminf(v[0], lo.x); v[1] = fminf(v[1], lo.y); v[2] = fminf(v[2], lo.z);
            v[3] = fmaxf(v[3], hi.x); v[4] = fmaxf(v[4], hi.y); v[5] = fmaxf(v[5], hi.z);
            v[6] = fminf(v[6], c.x); v[7] = fminf(v[7], c.y); v[8] = fminf(v[8], c.z);
            v[9] = fmaxf(v[9], c.x); v[10] = fmaxf(v[10], c.y); v[11] = fmaxf(v[11], c.z);
            any = true;
        } else {
            uint32_t* a = acc + 12 * (size_t)bs;
            atomicMin(a + 0, fenc(lo.x)); atomicMin(a + 1, fenc(lo.y)); atomicMin(a + 2, fenc(lo.z));
            atomicMax(a + 3, fenc(hi.x)); atomicMax(a + 4, fenc(hi.y)); atomicMax(a + 5, fenc(hi.z));
            atomicMin(a + 6, fenc(c.x)); atomicMin(a + 7, fenc(c.y)); atomicMin(a + 8, fenc(c.z));
            atomicMax(a + 9, fenc(c.x)); atomicMax(a + 10, fenc(c.y)); atomicMax(a + 11, fenc(c.z));
        }
    }
    if (uni && any)
        for (int q = 0; q < 12; q++) {
            if ((q % 6) < 3) atomicMin(&sacc[q], fenc(v[q]));
            else atomicMax(&sacc[q], fenc(v[q]));
        }
    __syncthreads();
    if (uni && threadIdx.x < 12) {
        uint32_t* a = acc + 12 * (size_t)b0;
        if ((threadIdx.x % 6) < 3) atomicMin(a + threadIdx.x, sacc[threadIdx.x]);
        else atomicMax(a + threadIdx.x, sacc[threadIdx.x]);
    }
}

__device__ __forceinline__ void node_cb(const uint32_t* __restrict__ a, float cl[3], float ch[3])
{
    for (int k = 0; k < 3; k++) {
        cl[k] = fdec(a[6 + k]);
        ch[k] = fdec(a[9 + k]);
    }
}

__global__ __launch_bounds__(BLK) void k_sah_bin(int n, int lb, int le, int bins, const int32_t* __restrict__ seg,
                                                 const int32_t* __restrict__ bslot, const int32_t* __restrict__ ord,
                                                 const float4* __restrict__ tlo, const float4* __restrict__ thi,
                                                 const uint32_t* __restrict__ acc, uint32_t* __restrict__ bcnt,
                                                 uint32_t* __restrict__ bbox)
{
    __shared__ uint32_t scnt[3 * SAH_MAXB];
    __shared__ uint32_t sbox[3 * SAH_MAXB * 6];
    const int r0 = (int)blockIdx.x * BLK * SAH_PER;
    const int r1 = min(n, r0 + BLK * SAH_PER);
    const int b0 = big_slot_of(seg[r0], lb, le, bslot);
    const bool uni = b0 >= 0 && seg[r1 - 1] == seg[r0];
    for (int q = threadIdx.x; q < 3 * bins; q += BLK) {
        scnt[q] = 0;
        for (int c = 0; c < 6; c++) sbox[6 * q + c] = c < 3 ? ENC_PINF : ENC_NINF;
    }
    __syncthreads();
    for (int i = r0 + threadIdx.x; i < r1; i += BLK) {
        const int bs = big_slot_of(seg[i], lb, le, bslot);
        if (bs < 0) continue;
        float cl[3], ch[3];
        node_cb(acc + 12 * (size_t)bs, cl, ch);
        const int p = ord[i];
        const float4 lo = tlo[p], hi = thi[p];
        const float4 c = centre(lo, hi);
        const float cc[3] = {c.x, c.y, c.z};
        for (int ax = 0; ax < 3; ax++) {
            const double ext = (double)ch[ax] - cl[ax];
            if (!(ext > 0)) continue;
            const int b = sah_bin(cc[ax], cl[ax], bins / ext, bins);
            const int q = ax * bins + b;
            uint32_t* cn = uni ? &scnt[q] : &bcnt[((size_t)bs * 3) * bins + q];
            uint32_t* bx = uni ? &sbox[6 * q] : &bbox[6 * (((size_t)bs * 3) * bins + q)];
            atomicAdd(cn, 1u);
            atomicMin(bx + 0, fenc(lo.x)); atomicMin(bx + 1, fenc(lo.y)); atomicMin(bx + 2, fenc(lo.z));
            atomicMax(bx + 3, fenc(hi.x)); atomicMax(bx + 4, fenc(hi.y)); atomicMax(bx + 5, fenc(hi.z));
        }
    }
    __syncthreads();
    if (uni)
        for (int q = threadIdx.x; q < 3 * bins; q += BLK) {
            if (scnt[q] == 0) continue;
            const size_t g = ((size_t)b0 * 3) * bins + q;
            atomicAdd(&bcnt[g], scnt[q]);
            for (int c = 0; c < 6; c++) {
                if (c < 3) atomicMin(&bbox[6 * g + c], sbox[6 * q + c]);
                else atomicMax(&bbox[6 * g + c], sbox[6 * q + c]);
            }
        }
}

struct SahBox {
    float lo[3], hi[3];
    __device__ void empty()
    {
        for (int k = 0; k < 3; k++) { lo[k] = INFINITY; hi[k] = -INFINITY; }
    }
    __device__ void grow(const SahBox& o)
    {
        for (int k = 0; k < 3; k++) { lo[k] = fminf(lo[k], o.lo[k]); hi[k] = fmaxf(hi[k], o.hi[k]); }
    }
    // bvh_build.cpp BBox::area
    __device__ double area() const
    {
        const double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

// The split sweep of one axis over its bins (bvh_build.cpp Builder::build):
// best (cost, split bin) updated in place.
__device__ void sah_sweep(const uint32_t* cnt, const SahBox* box, int bins, int ax, double& best_cost,
                          int& best_axis, int& best_split)
{
    double rarea[SAH_MAXB];
    uint32_t rcnt[SAH_MAXB];
    SahBox acc;
    acc.empty();
    uint32_t c = 0;
    for (int b = bins - 1; b > 0; b--) {
        acc.grow(box[b]);
        c += cnt[b];
        rarea[b] = acc.area();
        rcnt[b] = c;
    }
    acc.empty();
    c = 0;
    for (int b = 0; b < bins - 1; b++) {
        acc.grow(box[b]);
        c += cnt[b];
        if (c == 0 || rcnt[b + 1] == 0) continue;
        const double cost = acc.area() * (double)c + rarea[b + 1] * (double)rcnt[b + 1];
        if (cost < best_cost) {
            best_cost = cost;
            best_axis = ax;
            best_split = b;
        }
    }
}

// One thread per node of the level: node box and centre bounds, split
// choice, leaf rules.  sp_axis: split axis, or -1 for a split at half the
// count in the current order; sp_left: triangles going left.
__global__ void k_sah_split(int lb, int le, int depth, int bins, int max_leaf, const int32_t* __restrict__ bslot,
                            const uint32_t* __restrict__ acc, const uint32_t* __restrict__ bcnt,
                            const uint32_t* __restrict__ bbox, const int32_t* __restrict__ ord,
                            const float4* __restrict__ tlo, const float4* __restrict__ thi,
                            const int32_t* __restrict__ nstart, const int32_t* __restrict__ ncount,
                            float4* __restrict__ nlo, float4* __restrict__ nhi, float4* __restrict__ ncbl,
                            float4* __restrict__ ncbh, uint8_t* __restrict__ leaf, int32_t* __restrict__ off,
                            int32_t* __restrict__ cnt, int32_t* __restrict__ sp_axis, int32_t* __restrict__ sp_bin,
                            int32_t* __restrict__ sp_left, uint32_t* __restrict__ nchild)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= le - lb) return;
    const int k = lb + s;
    const int start = nstart[k], count = ncount[k];
    const int bs = bslot[s];
    SahBox nb;
    float cl[3], ch[3];
    if (bs >= 0) {
        const uint32_t* a = acc + 12 * (size_t)bs;
        for (int q = 0; q < 3; q++) {
            nb.lo[q] = fdec(a[q]);
            nb.hi[q] = fdec(a[3 + q]);
        }
        node_cb(a, cl, ch);
    } else {
        nb.empty();
        for (int q = 0; q < 3; q++) { cl[q] = INFINITY; ch[q] = -INFINITY; }
        for (int i = start; i < start + count; i++) {
            const int p = ord[i];
            const float4 lo = tlo[p], hi = thi[p];
            const float4 c = centre(lo, hi);
            const float lv[3] = {lo.x, lo.y, lo.z}, hv[3] = {hi.x, hi.y, hi.z}, cv[3] = {c.x, c.y, c.z};
            for (int q = 0; q < 3; q++) {
                nb.lo[q] = fminf(nb.lo[q], lv[q]);
                nb.hi[q] = fmaxf(nb.hi[q], hv[q]);
                cl[q] = fminf(cl[q], cv[q]);
                ch[q] = fmaxf(ch[q], cv[q]);
            }
        }
    }
    nlo[k] = make_float4(nb.lo[0], nb.lo[1], nb.lo[2], 0.0f);
    nhi[k] = make_float4(nb.hi[0], nb.hi[1], nb.hi[2], 0.0f);
    ncbl[k] = make_float4(cl[0], cl[1], cl[2], 0.0f);
    ncbh[k] = make_float4(ch[0], ch[1], ch[2], 0.0f);
    off[k] = start;
    cnt[k] = count;
    leaf[k] = 1;
    nchild[s] = 0;
    if (count <= 1) return;
    int best_axis = -1, best_split = -1;
    double best_cost = DBL_MAX;
    const bool median = depth >= 60;  // keep the depth bounded for the traversal stack
    if (!median) {
        uint32_t c_[SAH_MAXB];
        SahBox b_[SAH_MAXB];
        for (int ax = 0; ax < 3; ax++) {
            const double ext = (double)ch[ax] - cl[ax];
            if (!(ext > 0)) continue;
            if (bs >= 0) {
                for (int b = 0; b < bins; b++) {
                    const size_t g = ((size_t)bs * 3 + ax) * bins + b;
                    c_[b] = bcnt[g];
                    for (int q = 0; q < 3; q++) {
                        b_[b].lo[q] = fdec(bbox[6 * g + q]);
                        b_[b].hi[q] = fdec(bbox[6 * g + 3 + q]);
                    }
                }
            } else {
                for (int b = 0; b < bins; b++) {
                    c_[b] = 0;
                    b_[b].empty();
                }
                const double scale = bins / ext;
                for (int i = start; i < start + count; i++) {
                    const int p = ord[i];
                    const float4 lo = tlo[p], hi = thi[p];
                    const float4 c = centre(lo, hi);
                    const float cv[3] = {c.x, c.y, c.z};
                    const int b = sah_bin(cv[ax], cl[ax], scale, bins);
                    c_[b]++;
                    SahBox tb;
                    tb.lo[0] = lo.x; tb.lo[1] = lo.y; tb.lo[2] = lo.z;
                    tb.hi[0] = hi.x; tb.hi[1] = hi.y; tb.hi[2] = hi.z;
                    b_[b].grow(tb);
                }
            }
            sah_sweep(c_, b_, bins, ax, best_cost, best_axis, best_split);
        }
        const double pa = nb.area();
        // SAH with unit traversal and intersection costs
        const double split_cost = (pa > 0) ? 1.0 + best_cost / pa : DBL_MAX;
        if (count <= max_leaf && (best_axis < 0 || split_cost >= (double)count)) return;
        if (count <= YG_LEAF_MAX && best_axis >= 0 && split_cost >= (double)count && count <= max_leaf * 2) return;
    }
    int nl = count / 2;
    if (best_axis >= 0) {
        // triangles in the bins up to the split on the chosen axis
        const double ext = (double)ch[best_axis] - cl[best_axis];
        const double scale = bins / ext;
        nl = 0;
        if (bs >= 0) {
            for (int b = 0; b <= best_split; b++) nl += (int)bcnt[((size_t)bs * 3 + best_axis) * bins + b];
        } else {
            for (int i = start; i < start + count; i++) {
                const int p = ord[i];
                const float4 c = centre(tlo[p], thi[p]);
                const float cv[3] = {c.x, c.y, c.z};
                if (sah_bin(cv[best_axis], cl[best_axis], scale, bins) <= best_split) nl++;
            }
        }
        if (nl == 0 || nl == count) {
            best_axis = -1;
            nl = count / 2;
        }
    }
    leaf[k] = 0;
    sp_axis[s] = best_axis;
    sp_bin[s] = best_split;
    sp_left[s] = nl;
    nchild[s] = 2;
}

// children of the level's split nodes, numbered from `le` in node order
__global__ void k_sah_children(int lb, int le, const uint32_t* __restrict__ nchild, const uint32_t* __restrict__ cincl,
                               const int32_t* nstart, const int32_t* ncount, const int32_t* __restrict__ sp_left,
                               int32_t* __restrict__ left, int32_t* __restrict__ right, int32_t* cstart,
                               int32_t* ccount)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= le - lb || nchild[s] == 0) return;
    const int k = lb + s;
    const int c0 = le + (int)cincl[s] - 2;
    left[k] = c0;
    right[k] = c0 + 1;
    cstart[c0] = nstart[k];
    ccount[c0] = sp_left[s];
    cstart[c0 + 1] = nstart[k] + sp_left[s];
    ccount[c0 + 1] = ncount[k] - sp_left[s];
}

__global__ void k_sah_flags(int n, int lb, int le, int bins, const int32_t* __restrict__ seg,
                            const int32_t* __restrict__ ord, const float4* __restrict__ tlo,
                            const float4* __restrict__ thi, const uint8_t* __restrict__ leaf,
                            const int32_t* __restrict__ nstart, const float4* __restrict__ ncbl,
                            const float4* __restrict__ ncbh, const int32_t* __restrict__ sp_axis,
                            const int32_t* __restrict__ sp_bin, const int32_t* __restrict__ sp_left,
                            uint32_t* __restrict__ flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int node = seg[i];
    uint32_t f = 0;
    if (node >= lb && node < le && !leaf[node]) {
        const int s = node - lb;
        const int ax = sp_axis[s];
        if (ax < 0) f = (i - nstart[node]) < sp_left[s] ? 1u : 0u;
        else {
            const float4 cl = ncbl[node], ch = ncbh[node];
            const float l3[3] = {cl.x, cl.y, cl.z}, h3[3] = {ch.x, ch.y, ch.z};
            const int p = ord[i];
            const float4 c = centre(tlo[p], thi[p]);
            const float cv[3] = {c.x, c.y, c.z};
            const double ext = (double)h3[ax] - l3[ax];
            f = sah_bin(cv[ax], l3[ax], bins / ext, bins) <= sp_bin[s] ? 1u : 0u;
        }
    }
    flag[i] = f;
}

// stable partition of every split node's triangles; triangles of leaves
// (and of nodes finished earlier) keep their positions and leave the build
__global__ void k_sah_scatter(int n, int lb, int le, const int32_t* __restrict__ seg, const int32_t* __restrict__ ord,
                              const uint8_t* __restrict__ leaf, const int32_t* __restrict__ nstart,
                              const int32_t* __restrict__ sp_left, const int32_t* __restrict__ left,
                              const int32_t* __restrict__ right, const uint32_t* __restrict__ flag,
                              const uint32_t* __restrict__ lincl, int32_t* __restrict__ ord2,
                              int32_t* __restrict__ seg2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int node = seg[i];
    if (node < lb || node >= le || leaf[node]) {
        ord2[i] = ord[i];
        seg2[i] = -1;
        return;
    }
    const int s = node - lb;
    const int start = nstart[node];
    const uint32_t l0 = start > 0 ? lincl[start - 1] : 0u;
    const int rank = (int)(lincl[i] - l0);  // lefts in [start, i]
    int pos;
    if (flag[i]) pos = start + rank - 1;
    else pos = start + sp_left[s] + (i - start + 1 - rank) - 1;
    ord2[pos] = ord[i];
    seg2[pos] = flag[i] ? left[node] : right[node];
}

__global__ void k_big_flags(int lb, int le, const int32_t* __restrict__ ncount, uint32_t* __restrict__ big)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= le - lb) return;
    big[s] = ncount[lb + s] > SAH_SMALL ? 1u : 0u;
}

__global__ void k_big_slots(int na, const uint32_t* __restrict__ big, const uint32_t* __restrict__ bincl,
                            int32_t* __restrict__ bslot)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= na) return;
    bslot[s] = big[s] ? (int)bincl[s] - 1 : -1;
}

__global__ void k_sah_init(int n, int32_t* __restrict__ ord, int32_t* __restrict__ seg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ord[i] = i;
    seg[i] = 0;
}

__global__ void k_set_root(int n, int32_t* __restrict__ nstart, int32_t* __restrict__ ncount)
{
    nstart[0] = 0;
    ncount[0] = n;
}

// inner nodes per subtree, one level at a time from the deepest
__global__ void k_ninner(int a, int b, const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                         const uint8_t* __restrict__ leaf, int32_t* __restrict__ ninner)
{
    const int k = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b) return;
    ninner[k] = leaf[k] ? 0 : 1 + ninner[left[k]] + ninner[right[k]];
}

// depth-first indices of the inner nodes, one level at a time from the root
__global__ void k_didx(int a, int b, const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                       const uint8_t* __restrict__ leaf, const int32_t* __restrict__ ninner,
                       int32_t* __restrict__ didx)
{
    const int k = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b || leaf[k]) return;
    const int l = left[k];
    didx[l] = didx[k] + 1;
    didx[right[k]] = didx[k] + 1 + ninner[l];
}

// The exported tree, whichever builder made it.
struct TreeView {
    int n = 0;                       // triangles
    int first_node = 0, total_nodes = 0;  // node ids [first_node, total_nodes) may be inner nodes
    int root = 0;
    const float4 *nlo = nullptr, *nhi = nullptr;
    const int32_t *left = nullptr, *right = nullptr, *cnt = nullptr, *off = nullptr, *didx = nullptr;
    const uint8_t* leaf = nullptr;
    const uint8_t* under = nullptr;  // inside a collapsed leaf (null: no node is)
    const int32_t* tri_order = nullptr;  // triangle id at sorted position i
    const int32_t* tri_pos = nullptr;    // leaf-order position of sorted position i (null: i itself)
    int32_t n_inner = 0;             // inner nodes of the two-wide tree
    int32_t max_depth = 0;           // its depth, root = 1
};

// Leaf-ordered triangle records, two-wide nodes and the wide nodes of the
// asked width, handed to the scene.
int finish_export(yg_scene* s, const TreeView& t, const float* verts, const uint32_t* idx, const int32_t* obj_in,
                  const int32_t* mat_in, int width, Scanner& scan, DeviceBVHInfo& info)
{
    const int n = t.n;
    float4 rlo, rhi;
    YG_HIP(hipMemcpy(&rlo, t.nlo + t.root, sizeof(float4), hipMemcpyDeviceToHost));
    YG_HIP(hipMemcpy(&rhi, t.nhi + t.root, sizeof(float4), hipMemcpyDeviceToHost));
    const float G = std::max({fabsf(rlo.x), fabsf(rlo.y), fabsf(rlo.z), fabsf(rhi.x), fabsf(rhi.y), fabsf(rhi.z)});

    DBuf<float4> tris, tvn;
    DBuf<int32_t> tprim, tobj, tmat;
    DBuf<float> vn_in;
    DBuf<uint8_t> smooth_in;
    YG_ALLOC(tris, 3 * (size_t)n);
    YG_ALLOC(tprim, (size_t)n);
    YG_ALLOC(tobj, (size_t)n);
    YG_ALLOC(tmat, (size_t)n);
    const bool smooth = !s->tri_smooth.empty();
    if (smooth) {
        YG_ALLOC(tvn, 3 * (size_t)n);
        YG_ALLOC(vn_in, 9 * (size_t)n);
        YG_ALLOC(smooth_in, (size_t)n);
        YG_HIP(hipMemcpy(vn_in.p, s->tri_vn.data(), sizeof(float) * 9 * (size_t)n, hipMemcpyHostToDevice));
        YG_HIP(hipMemcpy(smooth_in.p, s->tri_smooth.data(), (size_t)n, hipMemcpyHostToDevice));
    }
    k_scatter_tris<<<blocks(n), BLK>>>(n, t.tri_order, t.tri_pos, verts, idx, obj_in, mat_in, vn_in.p, smooth_in.p,
                                       tris.p, tprim.p, tobj.p, tmat.p, smooth ? tvn.p : nullptr);
    YG_HIP(hipGetLastError());

    // two-wide nodes, depth-first
    DBuf<GNode> nodes2;
    YG_ALLOC(nodes2, (size_t)t.n_inner);
    k_export2<<<blocks(t.total_nodes - t.first_node), BLK>>>(t.first_node, t.total_nodes, t.left, t.right, t.nlo,
                                                            t.nhi, t.cnt, t.off, t.didx, t.leaf, t.under, G,
                                                            nodes2.p);
    YG_HIP(hipGetLastError());

    info.n_nodes = t.n_inner;
    info.max_depth = t.max_depth;
    info.stack = info.max_depth;
    GNode4* n4 = nullptr;
    GNode8* n8 = nullptr;
    if (width == 4 || width == 8) {
        int64_t nw = 0;
        int32_t stack = 0;
        int rc = (width == 4) ? export_wide_dev<4, GNode4>(t.root, t.n_inner, t.left, t.right, t.leaf, t.nlo, t.nhi,
                                                           t.off, t.cnt, G, scan, &n4, nw, stack)
                              : export_wide_dev<8, GNode8>(t.root, t.n_inner, t.left, t.right, t.leaf, t.nlo, t.nhi,
                                                           t.off, t.cnt, G, scan, &n8, nw, stack);
        if (rc) return rc;
        info.n_nodes = nw;
        info.stack = stack;
    }
    YG_HIP(hipDeviceSynchronize());

    s->dev_bytes = 0;
    s->d_nodes = nodes2.release();
    s->dev_bytes += sizeof(GNode) * (size_t)t.n_inner;
    if (n4) { s->d_nodes4 = n4; s->dev_bytes += sizeof(GNode4) * (size_t)info.n_nodes; }
    if (n8) { s->d_nodes8 = n8; s->dev_bytes += sizeof(GNode8) * (size_t)info.n_nodes; }
    s->d_tris = tris.release();
    s->d_tprim = tprim.release();
    s->d_tobj = tobj.release();
    s->d_tmat = tmat.release();
    s->dev_bytes += (sizeof(float4) * 3 + 3 * sizeof(int32_t)) * (size_t)n;
    if (smooth) {
        s->d_tvn = tvn.release();
        s->dev_bytes += sizeof(float4) * 3 * (size_t)n;
    }
    return YG_OK;
}

}  // namespace

static int build_ploc(yg_scene* s, int max_leaf, int width, DeviceBVHInfo& info)
{
    const int n = (int)s->tri_mat.size();
    const int total_nodes = 2 * n - 1;

    // inputs
    DBuf<float> verts;
    DBuf<uint32_t> idx;
    DBuf<int32_t> obj_in, mat_in;
    YG_ALLOC(verts, s->verts.size());
    YG_ALLOC(idx, s->idx.size());
    YG_ALLOC(obj_in, (size_t)n);
    YG_ALLOC(mat_in, (size_t)n);
    YG_HIP(hipMemcpy(verts.p, s->verts.data(), sizeof(float) * s->verts.size(), hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(idx.p, s->idx.data(), sizeof(uint32_t) * s->idx.size(), hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(obj_in.p, s->tri_obj.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(mat_in.p, s->tri_mat.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));

    // 1. boxes, centre bounds, Morton order
    DBuf<float4> tlo, thi;
    YG_ALLOC(tlo, (size_t)n);
    YG_ALLOC(thi, (size_t)n);
    k_prim_boxes<<<blocks(n), BLK>>>(verts.p, idx.p, n, tlo.p, thi.p);
    YG_HIP(hipGetLastError());
    const int nb = (int)std::min<int64_t>(blocks(n), 1024);
    DBuf<float> part, cb;
    YG_ALLOC(part, 6 * (size_t)nb);
    YG_ALLOC(cb, 6);
    k_cbounds_partial<<<nb, BLK>>>(tlo.p, thi.p, n, part.p);
    YG_HIP(hipGetLastError());
    k_cbounds_final<<<1, BLK>>>(part.p, nb, cb.p);
    YG_HIP(hipGetLastError());
    DBuf<uint64_t> keys, keys2;
    DBuf<int32_t> vals, sp;
    YG_ALLOC(keys, (size_t)n);
    YG_ALLOC(keys2, (size_t)n);
    YG_ALLOC(vals, (size_t)n);
    YG_ALLOC(sp, (size_t)n);
    k_morton<<<blocks(n), BLK>>>(tlo.p, thi.p, n, cb.p, keys.p, vals.p);
    YG_HIP(hipGetLastError());
    {
        size_t bytes = 0;
        YG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys.p, keys2.p, vals.p, sp.p, n, 0, 63,
                                                  (hipStream_t)0));
        DBuf<uint8_t> tmp;
        YG_ALLOC(tmp, bytes);
        YG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, keys.p, keys2.p, vals.p, sp.p, n, 0, 63,
                                                  (hipStream_t)0));
    }

    // 2. clustering
    DBuf<float4> nlo, nhi;
    DBuf<int32_t> left, right, cnt, ninner, cl, cl2, nn;
    DBuf<float> cost;
    DBuf<uint8_t> leaf;
    DBuf<uint32_t> f, fincl, keep, kincl;
    YG_ALLOC(nlo, (size_t)total_nodes);
    YG_ALLOC(nhi, (size_t)total_nodes);
    YG_ALLOC(left, (size_t)total_nodes);
    YG_ALLOC(right, (size_t)total_nodes);
    YG_ALLOC(cnt, (size_t)total_nodes);
    YG_ALLOC(ninner, (size_t)total_nodes);
    YG_ALLOC(cost, (size_t)total_nodes);
    YG_ALLOC(leaf, (size_t)total_nodes);
    YG_ALLOC(cl, (size_t)n);
    YG_ALLOC(cl2, (size_t)n);
    YG_ALLOC(nn, (size_t)n);
    YG_ALLOC(f, (size_t)n);
    YG_ALLOC(fincl, (size_t)n);
    YG_ALLOC(keep, (size_t)n);
    YG_ALLOC(kincl, (size_t)n);
    k_leaf_init<<<blocks(n), BLK>>>(sp.p, n, tlo.p, thi.p, nlo.p, nhi.p, cl.p, cnt.p, cost.p, leaf.p, ninner.p);
    YG_HIP(hipGetLastError());
    Scanner scan;
    std::vector<int> rounds;  // first node of each merge round, then the end
    int m = n, next = n;
    rounds.push_back(next);
    while (m > 1) {
        k_nn<<<blocks(m), BLK>>>(cl.p, m, nlo.p, nhi.p, nn.p);
        YG_HIP(hipGetLastError());
        k_merge_flags<<<blocks(m), BLK>>>(nn.p, m, f.p);
        YG_HIP(hipGetLastError());
        int rc = scan.run(f.p, fincl.p, m);
        if (rc) return rc;
        uint32_t merges = 0;
        YG_HIP(hipMemcpy(&merges, fincl.p + (m - 1), sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (merges == 0) return yg_fail(YG_ERR_STATE, "device BVH build: a clustering round merged nothing");
        k_merge<<<blocks(m), BLK>>>(cl.p, nn.p, fincl.p, m, next, nlo.p, nhi.p, left.p, right.p, cl2.p, keep.p);
        YG_HIP(hipGetLastError());
        if ((rc = scan.run(keep.p, kincl.p, m))) return rc;
        k_compact<<<blocks(m), BLK>>>(cl2.p, keep.p, kincl.p, m, cl.p);
        YG_HIP(hipGetLastError());
        next += (int)merges;
        m -= (int)merges;
        rounds.push_back(next);
    }
    if (next != total_nodes) return yg_fail(YG_ERR_STATE, "device BVH build: node count mismatch");
    const int root = total_nodes - 1;

    // 3. counts, SAH collapse into leaves
    for (size_t q = 0; q + 1 < rounds.size(); q++) {
        const int a = rounds[q], b = rounds[q + 1];
        k_bottom_up<<<blocks(b - a), BLK>>>(a, b, left.p, right.p, nlo.p, nhi.p, max_leaf, cnt.p, cost.p, leaf.p,
                                           ninner.p);
        YG_HIP(hipGetLastError());
    }
    uint8_t root_leaf = 0;
    YG_HIP(hipMemcpy(&root_leaf, leaf.p + root, 1, hipMemcpyDeviceToHost));
    if (root_leaf) return yg_fail(YG_ERR_ARG, "device BVH build: the whole scene fits one leaf");

    // 4. depth-first offsets and inner-node indices
    DBuf<int32_t> off, depth, didx, maxd;
    DBuf<uint8_t> under;
    YG_ALLOC(off, (size_t)total_nodes);
    YG_ALLOC(depth, (size_t)total_nodes);
    YG_ALLOC(didx, (size_t)total_nodes);
    YG_ALLOC(under, (size_t)total_nodes);
    YG_ALLOC(maxd, 1);
    const int32_t zero = 0;
    YG_HIP(hipMemcpy(maxd.p, &zero, sizeof(int32_t), hipMemcpyHostToDevice));
    k_root_init<<<1, 1>>>(root, off.p, depth.p, under.p, didx.p);
    YG_HIP(hipGetLastError());
    for (size_t q = rounds.size() - 1; q > 0; q--) {
        const int a = rounds[q - 1], b = rounds[q];
        k_top_down<<<blocks(b - a), BLK>>>(a, b, left.p, right.p, cnt.p, leaf.p, ninner.p, off.p, depth.p, under.p,
                                          didx.p, maxd.p);
        YG_HIP(hipGetLastError());
    }
    int32_t n_inner = 0, max_depth = 0;
    YG_HIP(hipMemcpy(&n_inner, ninner.p + root, sizeof(int32_t), hipMemcpyDeviceToHost));
    YG_HIP(hipMemcpy(&max_depth, maxd.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (n_inner < 1 || n_inner > n - 1) return yg_fail(YG_ERR_STATE, "device BVH build: inconsistent tree");
    TreeView tv;
    tv.n = n;
    tv.first_node = n;
    tv.total_nodes = total_nodes;
    tv.root = root;
    tv.nlo = nlo.p; tv.nhi = nhi.p;
    tv.left = left.p; tv.right = right.p; tv.cnt = cnt.p; tv.off = off.p; tv.didx = didx.p;
    tv.leaf = leaf.p; tv.under = under.p;
    tv.tri_order = sp.p;
    tv.tri_pos = off.p;
    tv.n_inner = n_inner;
    tv.max_depth = max_depth + 1;
    return finish_export(s, tv, verts.p, idx.p, obj_in.p, mat_in.p, width, scan, info);
}

static int build_sah(yg_scene* s, int max_leaf, int bins, int width, DeviceBVHInfo& info)
{
    const int n = (int)s->tri_mat.size();
    const int cap_nodes = 2 * n - 1;
    DBuf<float> verts;
    DBuf<uint32_t> idx;
    DBuf<int32_t> obj_in, mat_in;
    YG_ALLOC(verts, s->verts.size());
    YG_ALLOC(idx, s->idx.size());
    YG_ALLOC(obj_in, (size_t)n);
    YG_ALLOC(mat_in, (size_t)n);
    YG_HIP(hipMemcpy(verts.p, s->verts.data(), sizeof(float) * s->verts.size(), hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(idx.p, s->idx.data(), sizeof(uint32_t) * s->idx.size(), hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(obj_in.p, s->tri_obj.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(mat_in.p, s->tri_mat.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
    DBuf<float4> tlo, thi;
    YG_ALLOC(tlo, (size_t)n);
    YG_ALLOC(thi, (size_t)n);
    k_prim_boxes<<<blocks(n), BLK>>>(verts.p, idx.p, n, tlo.p, thi.p);
    YG_HIP(hipGetLastError());

    // node records
    DBuf<float4> nlo, nhi, ncbl, ncbh;
    DBuf<int32_t> nstart, ncount, left, right, off, cnt, ninner, didx;
    DBuf<uint8_t> leaf;
    YG_ALLOC(nlo, (size_t)cap_nodes);
    YG_ALLOC(nhi, (size_t)cap_nodes);
    YG_ALLOC(ncbl, (size_t)cap_nodes);
    YG_ALLOC(ncbh, (size_t)cap_nodes);
    YG_ALLOC(nstart, (size_t)cap_nodes);
    YG_ALLOC(ncount, (size_t)cap_nodes);
    YG_ALLOC(left, (size_t)cap_nodes);
    YG_ALLOC(right, (size_t)cap_nodes);
    YG_ALLOC(off, (size_t)cap_nodes);
    YG_ALLOC(cnt, (size_t)cap_nodes);
    YG_ALLOC(ninner, (size_t)cap_nodes);
    YG_ALLOC(didx, (size_t)cap_nodes);
    YG_ALLOC(leaf, (size_t)cap_nodes);
    // per position
    DBuf<int32_t> ord, ord2, seg, seg2;
    DBuf<uint32_t> flag, lincl;
    YG_ALLOC(ord, (size_t)n);
    YG_ALLOC(ord2, (size_t)n);
    YG_ALLOC(seg, (size_t)n);
    YG_ALLOC(seg2, (size_t)n);
    YG_ALLOC(flag, (size_t)n);
    YG_ALLOC(lincl, (size_t)n);
    // per node of a level (at most n), per big node of a level (at most n / SAH_SMALL)
    const int cap_big = n / SAH_SMALL + 1;
    DBuf<int32_t> bslot, sp_axis, sp_bin, sp_left;
    DBuf<uint32_t> big, bincl, nchild, cincl, acc, bcnt, bbox;
    YG_ALLOC(bslot, (size_t)n);
    YG_ALLOC(sp_axis, (size_t)n);
    YG_ALLOC(sp_bin, (size_t)n);
    YG_ALLOC(sp_left, (size_t)n);
    YG_ALLOC(big, (size_t)n);
    YG_ALLOC(bincl, (size_t)n);
    YG_ALLOC(nchild, (size_t)n);
    YG_ALLOC(cincl, (size_t)n);
    YG_ALLOC(acc, 12 * (size_t)cap_big);
    YG_ALLOC(bcnt, (size_t)cap_big * 3 * bins);
    YG_ALLOC(bbox, (size_t)cap_big * 3 * bins * 6);

    k_sah_init<<<blocks(n), BLK>>>(n, ord.p, seg.p);
    YG_HIP(hipGetLastError());
    k_set_root<<<1, 1>>>(n, nstart.p, ncount.p);
    YG_HIP(hipGetLastError());
    Scanner scan;
    std::vector<int> levels = {0};  // first node of each level, then the end
    int lb = 0, le = 1, depth = 0;
    const unsigned pos_blocks = (unsigned)((n + BLK * SAH_PER - 1) / (BLK * SAH_PER));
    while (le > lb) {
        const int na = le - lb;
        if (depth >= YG_MAX_STACK) return yg_fail(YG_ERR_STATE, "device BVH build: tree deeper than the traversal stack");
        // big nodes of the level and their accumulator slots
        k_big_flags<<<blocks(na), BLK>>>(lb, le, ncount.p, big.p);
        YG_HIP(hipGetLastError());
        int rc = scan.run(big.p, bincl.p, na);
        if (rc) return rc;
        uint32_t nbig = 0;
        YG_HIP(hipMemcpy(&nbig, bincl.p + (na - 1), sizeof(uint32_t), hipMemcpyDeviceToHost));
        if ((int)nbig > cap_big) return yg_fail(YG_ERR_STATE, "device BVH build: big node overflow");
        k_big_slots<<<blocks(na), BLK>>>(na, big.p, bincl.p, bslot.p);
        YG_HIP(hipGetLastError());
        if (nbig > 0) {
            k_acc_init<<<blocks(12 * (int64_t)nbig), BLK>>>(acc.p, (int)nbig);
            YG_HIP(hipGetLastError());
            const int64_t nbins = (int64_t)nbig * 3 * bins;
            k_bins_init<<<blocks(nbins), BLK>>>(bcnt.p, bbox.p, nbins);
            YG_HIP(hipGetLastError());
            k_sah_bounds<<<pos_blocks, BLK>>>(n, lb, le, seg.p, bslot.p, ord.p, tlo.p, thi.p, acc.p);
            YG_HIP(hipGetLastError());
            k_sah_bin<<<pos_blocks, BLK>>>(n, lb, le, bins, seg.p, bslot.p, ord.p, tlo.p, thi.p, acc.p, bcnt.p, bbox.p);
            YG_HIP(hipGetLastError());
        }
        k_sah_split<<<blocks(na), BLK>>>(lb, le, depth, bins, max_leaf, bslot.p, acc.p, bcnt.p, bbox.p, ord.p, tlo.p,
                                        thi.p, nstart.p, ncount.p, nlo.p, nhi.p, ncbl.p, ncbh.p, leaf.p, off.p, cnt.p,
                                        sp_axis.p, sp_bin.p, sp_left.p, nchild.p);
        YG_HIP(hipGetLastError());
        if ((rc = scan.run(nchild.p, cincl.p, na))) return rc;
        uint32_t nnew = 0;
        YG_HIP(hipMemcpy(&nnew, cincl.p + (na - 1), sizeof(uint32_t), hipMemcpyDeviceToHost));
        if ((int64_t)le + nnew > (int64_t)cap_nodes) return yg_fail(YG_ERR_STATE, "device BVH build: node overflow");
        if (nnew > 0) {
            k_sah_children<<<blocks(na), BLK>>>(lb, le, nchild.p, cincl.p, nstart.p, ncount.p, sp_left.p, left.p,
                                                right.p, nstart.p, ncount.p);
            YG_HIP(hipGetLastError());
        }
        k_sah_flags<<<blocks(n), BLK>>>(n, lb, le, bins, seg.p, ord.p, tlo.p, thi.p, leaf.p, nstart.p, ncbl.p, ncbh.p,
                                        sp_axis.p, sp_bin.p, sp_left.p, flag.p);
        YG_HIP(hipGetLastError());
        if ((rc = scan.run(flag.p, lincl.p, n))) return rc;
        k_sah_scatter<<<blocks(n), BLK>>>(n, lb, le, seg.p, ord.p, leaf.p, nstart.p, sp_left.p, left.p, right.p,
                                          flag.p, lincl.p, ord2.p, seg2.p);
        YG_HIP(hipGetLastError());
        std::swap(ord.p, ord2.p);
        std::swap(seg.p, seg2.p);
        levels.push_back(le);  // end of the level just split = first node of the next
        lb = le;
        le += (int)nnew;
        depth++;
    }
    const int total = le;
    // inner-node counts from the deepest level up, depth-first indices from the root down
    for (size_t q = levels.size() - 1; q > 0; q--) {
        const int a = levels[q - 1], b = levels[q];
        if (b > a) {
            k_ninner<<<blocks(b - a), BLK>>>(a, b, left.p, right.p, leaf.p, ninner.p);
            YG_HIP(hipGetLastError());
        }
    }
    uint8_t root_leaf = 0;
    YG_HIP(hipMemcpy(&root_leaf, leaf.p, 1, hipMemcpyDeviceToHost));
    if (root_leaf) return yg_fail(YG_ERR_ARG, "device BVH build: the whole scene fits one leaf");
    const int32_t zero = 0;
    YG_HIP(hipMemcpy(didx.p, &zero, sizeof(int32_t), hipMemcpyHostToDevice));
    for (size_t q = 0; q + 1 < levels.size(); q++) {
        const int a = levels[q], b = levels[q + 1];
        if (b > a) {
            k_didx<<<blocks(b - a), BLK>>>(a, b, left.p, right.p, leaf.p, ninner.p, didx.p);
            YG_HIP(hipGetLastError());
        }
    }
    int32_t n_inner = 0;
    YG_HIP(hipMemcpy(&n_inner, ninner.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    // a two-wide tree over n triangles with more than one leaf: 1 .. n-1 inner nodes, 2 n_inner + 1 nodes
    if (n_inner < 1 || n_inner > n - 1 || 2 * n_inner + 1 != total)
        return yg_fail(YG_ERR_STATE, "device BVH build: inconsistent tree");
    TreeView tv;
    tv.n = n;
    tv.first_node = 0;
    tv.total_nodes = total;
    tv.root = 0;
    tv.nlo = nlo.p; tv.nhi = nhi.p;
    tv.left = left.p; tv.right = right.p; tv.cnt = cnt.p; tv.off = off.p; tv.didx = didx.p;
    tv.leaf = leaf.p; tv.under = nullptr;
    tv.tri_order = ord.p;
    tv.tri_pos = nullptr;
    tv.n_inner = n_inner;
    tv.max_depth = depth;  // levels: the deepest node's depth + 1
    return finish_export(s, tv, verts.p, idx.p, obj_in.p, mat_in.p, width, scan, info);
}

int build_bvh_device(yg_scene* s, int kind, int max_leaf, int bins, int width, DeviceBVHInfo& info)
{
    const int64_t nt = (int64_t)s->tri_mat.size();
    if (nt < 2) return yg_fail(YG_ERR_ARG, "device BVH build needs at least two triangles");
    if (nt > (int64_t)(INT32_MAX / YG_LEAF_MAX) || 2 * nt > (int64_t)INT32_MAX)
        return yg_fail(YG_ERR_ARG, "device BVH build: too many triangles");
    if (max_leaf < 1) max_leaf = 4;
    if (max_leaf > YG_LEAF_MAX) max_leaf = YG_LEAF_MAX;
    if (bins < 4) bins = 32;
    if (bins > SAH_MAXB) bins = SAH_MAXB;
    if (kind == YG_BUILDER_DEVICE_PLOC) return build_ploc(s, max_leaf, width, info);
    return build_sah(s, max_leaf, bins, width, info);
}
