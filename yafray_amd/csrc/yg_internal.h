// Internal types of libyafgpu: scene handle, BVH layout, error plumbing.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/yafgpu.h"
#include "dev_math.h"

// ----------------------------------------------------------------- BVH --
// Two-wide BVH, one 64-byte node per inner node holding both child boxes
// (one cache line per visit).  Children >= 0 are inner nodes; a leaf child is
// ~(first*16 + count-1): `count` (1..16) triangles from position `first` of
// the leaf-ordered triangle array.
struct alignas(64) GNode {
    float4 a;  // c0.lo.x c0.hi.x c0.lo.y c0.hi.y
    float4 b;  // c1.lo.x c1.hi.x c1.lo.y c1.hi.y
    float4 c;  // c0.lo.z c0.hi.z c1.lo.z c1.hi.z
    int4 d;    // child0 child1 - -
};

// Four-wide BVH node, 128 bytes (one L2 line): the boxes of up to four
// children as structure-of-arrays (component k = child k) and their
// references, encoded as for GNode.  Unused child slots have +inf boxes,
// which every slab test rejects.
struct alignas(128) GNode4 {
    float4 lox, hix, loy, hiy, loz, hiz;
    int4 child;
    int4 pad;
};

// Eight-wide BVH node, 256 bytes (two L2 lines): six box rows (lo.x hi.x
// lo.y hi.y lo.z hi.z) of eight children each, then the eight child
// references encoded as for GNode; unused slots have +inf boxes.
struct alignas(128) GNode8 {
    float box[6][8];
    int32_t child[8];
    int32_t pad[8];
};

#define YG_LEAF_MAX 16
#ifndef YG_DEFAULT_WIDTH
#define YG_DEFAULT_WIDTH 4
#endif
#ifndef YG_DEFAULT_STACK_LDS
#define YG_DEFAULT_STACK_LDS 16  // render launches: stack entries per lane in LDS
#endif
// Deepest traversal stack the kernels accept.  A launch sizes each lane's
// stack for the deepest its trees can need (one entry per level of a
// two-wide tree, up to three per level of a four-wide one): in LDS, or with
// the render launches the first stack_lds entries in LDS and the rest in
// device memory.
#define YG_MAX_STACK 200

// Dynamic LDS bytes of a traversal launch with `lanes` threads per block and
// `entries` stack entries per lane.
#ifndef YG_STACK_EXTRA
#define YG_STACK_EXTRA 0  // extra entries per lane (occupancy tuning builds)
#endif
static inline size_t stack_lds_bytes(int entries, int lanes)
{
    return (size_t)(entries + YG_STACK_EXTRA) * (size_t)lanes * sizeof(int);
}

// Leaf-ordered triangle record, 48 B: (a.xyz, n.x) (b.xyz, n.y) (c.xyz, n.z),
// n = the unit face normal of triangle_t::recNormal (triangle.cc:45-49).
struct BVHView {
    const GNode* nodes;     // width 2
    const GNode4* nodes4;   // width 4
    const GNode8* nodes8;   // width 8
    int32_t width;
    const float4* tris;     // 3 per triangle
    const int32_t* tprim;   // original primitive id per leaf-ordered triangle
    const int32_t* tobj;    // object id per leaf-ordered triangle
    const int32_t* tmat;    // material id per leaf-ordered triangle
    int32_t ntris;
    int32_t nobj;           // objects in the scene (1: every triangle is object 0)
    const int32_t* oflags;  // per object: bit0 cast shadows, bit1 receive, bit2 use for radiosity
    int32_t nocast;         // some object casts no shadows (scene shadow queries skip it)
    const float4* tvn;      // vertex normals per leaf-ordered triangle (3 float4), or null
    const float4* ptex;     // texture-coordinate data per primitive (4 float4), or null
    const float* oorco;     // per object: rows 0..2 of the orco matrix (12 floats)
};

struct HostBVH {
    std::vector<GNode> nodes;
    std::vector<GNode4> nodes4;  // the same tree collapsed to width 4 (when asked)
    int32_t stack4 = 0;          // deepest traversal stack the width-4 tree can need
    std::vector<GNode8> nodes8;  // ... or to width 8
    int32_t stack8 = 0;
    std::vector<float4> tris;
    std::vector<int32_t> tprim, tobj, tmat;
    int32_t max_depth = 0;
};

// Builds the BVH over world-space triangles (binned SAH).  prim ids are
// positions in idx/3.
void build_bvh(const float* verts, const uint32_t* idx, int64_t nt, const int32_t* tri_obj,
               const int32_t* tri_mat, int max_leaf, int bins, int threads, int width, HostBVH& out);

struct yg_scene;

// What the device builder reports: node count of the exported width, the
// two-wide tree's depth (root = 1) and, for widths 4 and 8, the deepest
// traversal stack the tree can need (entries - 1).
struct DeviceBVHInfo {
    int64_t n_nodes = 0;
    int32_t max_depth = 0;
    int32_t stack = 0;
};

// Builds the scene's BVH on the device (bvh_gpu.hip), kind
// YG_BUILDER_DEVICE (binned SAH) or YG_BUILDER_DEVICE_PLOC (clustering):
// writes s->d_nodes, d_nodes4 or d_nodes8 (width 4 or 8), d_tris, d_tprim,
// d_tobj, d_tmat and, for smoothed scenes, d_tvn; sets s->dev_bytes.  Needs
// at least two triangles and a scene that is not a single leaf.
int build_bvh_device(yg_scene* s, int kind, int max_leaf, int bins, int width, DeviceBVHInfo& info);

// Face normal exactly as triangle_t::recNormal computes it.
YG_HD f3 tri_normal(f3 a, f3 b, f3 c) { return norm3(cross(b - a, c - a)); }

// ---------------------------------------------------------- photon map --
struct PhotonMapDev {
    float* d_photons = nullptr;     // 9 floats each (dir, pos, colour) in cell order
    int64_t n = 0;
    int32_t* d_gstart = nullptr;    // uniform grid over photons: ncells+1
    int32_t* d_gitems = nullptr;
    float gmin[3] = {0, 0, 0};
    float gcell = 1;
    int gdim[3] = {1, 1, 1};
    int np_emit = 0;                // photonLight_t::Np
    int light = -1;
};

// ------------------------------------------------- global photon map --
#include "gphoton_dev.h"

// globalPhotonLight_t's pre-gathered boxes on the device (gphoton.hip)
struct GPhotonDev {
    int light = -1;                 // light index it was built for, -1: none
    uint64_t gen = 0;               // geometry generation it was built on
    int64_t n = 0;                  // boxes
    float cell = 1, radius = 1;
    int bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    int sy = 0, sz = 0;
    float gmin[3] = {0, 0, 0};
    int gdim[3] = {1, 1, 1};
    unsigned long long* d_key = nullptr;
    f3* d_N = nullptr;
    c3* d_irr = nullptr;
    f3* d_pos = nullptr;
    f3* d_qN = nullptr;
    c3* d_qirr = nullptr;
    int32_t* d_gstart = nullptr;
    int32_t* d_gitems = nullptr;
    float* d_read = nullptr;        // 15 floats per box (yg_gphoton_read)
    // raw photons in store order: decoded position / direction / colour,
    // the 5-word records of yg_gphoton_read_photons, grid (cell = radius)
    int64_t nrec = 0;
    f3* d_rpos = nullptr;
    f3* d_rdir = nullptr;
    c3* d_rcol = nullptr;
    uint32_t* d_rraw = nullptr;
    int32_t* d_rgstart = nullptr;
    int32_t* d_rgitems = nullptr;
    float rgmin[3] = {0, 0, 0};
    int rgdim[3] = {1, 1, 1};
    GPhotonView view() const
    {
        GPhotonView v;
        v.key = d_key; v.N = d_N; v.irr = d_irr; v.pos = d_pos; v.qN = d_qN; v.qirr = d_qirr;
        v.gstart = d_gstart; v.gitems = d_gitems;
        for (int a = 0; a < 3; a++) {
            v.gmin[a] = gmin[a]; v.gdim[a] = gdim[a]; v.bmin[a] = bmin[a]; v.bmax[a] = bmax[a];
        }
        v.cell = cell; v.radius = radius; v.sy = sy; v.sz = sz; v.n = n;
        v.rpos = d_rpos; v.rdir = d_rdir; v.rcol = d_rcol; v.rgstart = d_rgstart; v.rgitems = d_rgitems;
        for (int a = 0; a < 3; a++) { v.rgmin[a] = rgmin[a]; v.rgdim[a] = rgdim[a]; }
        v.nrec = nrec;
        return v;
    }
};
void gphoton_free(GPhotonDev& g);

// --------------------------------------------------------- scene handle --
struct yg_scene {
    std::vector<float> verts;
    std::vector<uint32_t> idx;
    std::vector<int32_t> tri_obj, tri_mat, obj_flags;
    int32_t nobj = 0;
    // autoSmooth vertex normals: 9 floats per triangle (corners a, b, c) in
    // mesh-add order, and whether the triangle has them (empty: none has)
    std::vector<float> tri_vn;
    std::vector<uint8_t> tri_smooth;
    float4* d_tvn = nullptr;  // leaf order: (na, 1), (nb, 0), (nc, 0); w of the first = has normals
    // yg_mesh_texcoords: per triangle in add order 16 floats (u_a v_a u_b v_b
    // | u_c v_c oa.x oa.y | oa.z ob.xyz | oc.xyz flags; flags bit0 uv, bit1
    // vertex colours, bit2 orco points) once any object has such data (empty:
    // none has); per object its first vertex, vertex count, orco matrix
    // (rows 0..2 of backOrco) and whether the caller set that matrix
    std::vector<float> tri_tex;
    std::vector<uint32_t> obj_vbase, obj_nv;
    std::vector<float> obj_orco;
    std::vector<uint8_t> obj_orco_set;
    float4* d_ptex = nullptr;  // tri_tex on the device
    float* d_oorco = nullptr;  // obj_orco on the device (defaults filled in by yg_build)

    bool built = false;
    int32_t max_depth = 0;
    int32_t width = 2;
    int32_t stack_entries = 1;  // traversal stack entries per lane the tree can need
    int32_t stack_lds = YG_DEFAULT_STACK_LDS;  // of those, kept in LDS by render launches (0: all)
    int32_t rays_per_lane = 1;   // GI level launches: rays each lane traces
    float compact_below = 1.0f;  // GI levels: compact the survivors below this fraction
    uint64_t n_nodes = 0;
    GNode* d_nodes = nullptr;
    GNode4* d_nodes4 = nullptr;
    GNode8* d_nodes8 = nullptr;
    float4* d_tris = nullptr;
    int32_t* d_tprim = nullptr;
    int32_t* d_tobj = nullptr;
    int32_t* d_tmat = nullptr;
    int32_t* d_obj_flags = nullptr;
    uint64_t dev_bytes = 0;

    bool has_cam = false;
    yg_camera cam{};
    std::vector<yg_material> mats;
    std::vector<yg_modulator> mods;
    std::vector<yg_texture> texs;
    std::vector<yg_light> lights;
    yg_material* d_mats = nullptr;
    yg_modulator* d_mods = nullptr;
    yg_texture* d_texs = nullptr;

    PhotonMapDev pm;
    GPhotonDev gp;
    std::vector<float*> ph_halton;  // yg_photon_build: device Halton tables per dimension
    int ph_halton_np = -1;          //   for this emitted-photon count
    int* d_any_ovf = nullptr;  // yg_trace_any: traversal stack entries beyond the LDS part
    size_t any_ovf_ints = 0;
    uint64_t gen = 0;  // bumped by every change of geometry, shading, lights or camera
    uint64_t geo_gen = 0;          // bumped by geometry changes only
    uint64_t checked_geo = ~0ull;  // geo_gen / material count at which the
    size_t checked_nmat = 0;       //   triangle material ids were last validated
    uint64_t uid = 0;              // process-unique scene number (never reused)

    // per-area Halton tables of the last render (dimension-major, hstride
    // entries per dimension), kept on the device while ndim/hstride repeat
    float* d_halton = nullptr;
    int halton_ndim = 0;
    int halton_stride = 0;

    BVHView view() const
    {
        BVHView v;
        v.nodes = d_nodes; v.nodes4 = d_nodes4; v.nodes8 = d_nodes8; v.width = width; v.tris = d_tris; v.tprim = d_tprim; v.tobj = d_tobj; v.tmat = d_tmat;
        v.ntris = (int32_t)tri_mat.size();
        v.nobj = nobj;
        v.oflags = d_obj_flags;
        v.tvn = d_tvn;
        v.ptex = d_ptex;
        v.oorco = d_oorco;
        v.nocast = 0;
        for (int32_t f : obj_flags)
            if (!(f & 1)) v.nocast = 1;
        return v;
    }
};

// ---------------------------------------------------------------- errors --
void yg_set_error(const std::string& msg);
int yg_fail(int code, const std::string& msg);

#define YG_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return yg_fail(YG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
static inline int dev_upload(T** dst, const T* src, size_t n, uint64_t* acct = nullptr)
{
    if (*dst) { (void)hipFree(*dst); *dst = nullptr; }
    if (n == 0) return YG_OK;
    hipError_t e = hipMalloc((void**)dst, sizeof(T) * n);
    if (e != hipSuccess) return yg_fail(YG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return yg_fail(YG_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    if (acct) *acct += sizeof(T) * n;
    return YG_OK;
}

static inline unsigned grid_for(uint64_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

// Launches the instantiation of a traversal kernel for the BVH width in use
// (K8 / K4 / K2: the kernel template-ids for widths 8, 4 and 2, parenthesised).
#define YG_LAUNCH_W(width, K8, K4, K2, grid, block, shm, stream, ...)                            \
    do {                                                                                        \
        if ((width) == 8) hipLaunchKernelGGL(K8, grid, block, shm, stream, __VA_ARGS__);        \
        else if ((width) == 4) hipLaunchKernelGGL(K4, grid, block, shm, stream, __VA_ARGS__);   \
        else hipLaunchKernelGGL(K2, grid, block, shm, stream, __VA_ARGS__);                     \
    } while (0)
// The same for kernels with shadow queries: K24 / K20 / K18 are the width-8 /
// 4 / 2 instantiations that test object flags (W + 16), used when nocast is set.
#define YG_LAUNCH_WS(width, nocast, K8, K4, K2, K24, K20, K18, grid, block, shm, stream, ...)     \
    do {                                                                                        \
        if (nocast) YG_LAUNCH_W(width, K24, K20, K18, grid, block, shm, stream, __VA_ARGS__);   \
        else YG_LAUNCH_W(width, K8, K4, K2, grid, block, shm, stream, __VA_ARGS__);             \
    } while (0)
