// Ray-query entry points of libyafgpu: closest / any hit over ray batches,
// scene_t::firstHit / isShadowed semantics, noise evaluation and heightfield
// generation on the device.
#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "noise_dev.h"
#include "traverse.h"

#define TB 64

// --------------------------------------------------------------- kernels --

template <bool COUNT, int W>
__global__ __launch_bounds__(TB) void k_trace_closest(BVHView bvh, const yg_ray* __restrict__ rays,
                                                       uint32_t n, yg_hit* __restrict__ hits,
                                                       unsigned long long* counters)
{
    extern __shared__ int s_stk[];
    const uint32_t i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    LaneStack st;
    st.init(s_stk, threadIdx.x, TB);
    const yg_ray r = rays[i];
    const f3 o = mk3(r.ox, r.oy, r.oz), d = mk3(r.dx, r.dy, r.dz);
    float t = r.tmax;
    TraceStats cs = {0, 0};
    const int ti = trace_closest<COUNT, W>(bvh, o, d, r.tmin, t, st, &cs);
    yg_hit h;
    if (ti >= 0) {
        f3 a, b, c, nn;
        load_tri(bvh.tris, ti, a, b, c, nn);
        const f3 hp = o + d * t;  // mesh.cc:570
        tri_bary(nn, a, b, c, hp, h.u, h.v);
        h.prim = bvh.tprim[ti];
        h.t = t;
    } else {
        h.prim = -1; h.t = r.tmax; h.u = 0; h.v = 0;
    }
    hits[i] = h;
    if (COUNT) {
        atomicAdd(&counters[0], (unsigned long long)cs.nodes);
        atomicAdd(&counters[1], (unsigned long long)cs.tris);
    }
}

// Any-hit batches: each lane keeps the first `lcap` entries of its traversal
// stack in LDS and deeper ones (only rays under deep subtrees reach them) in
// its own column of an overflow buffer, ovf[(entry - lcap) * nthreads +
// thread].  The short LDS part lets more waves share a compute unit than a
// whole stack per lane would (terrain, 4M rays: 4.7 vs 4.0 Grays/s).  One ray
// per thread while the overflow buffer stays within its cap; beyond that the
// grid shrinks and each thread loops over several rays.
#ifndef YG_ANY_LCAP
#define YG_ANY_LCAP 16
#endif
#define YG_ANY_OVF_MAX ((size_t)1 << 28)  // overflow ints of one launch (1 GiB)

template <int W>
__global__ __launch_bounds__(TB) void k_trace_any(BVHView bvh, const yg_ray* __restrict__ rays,
                                                  uint32_t n, uint8_t* __restrict__ occ, int lcap,
                                                  int* __restrict__ ovf)
{
    extern __shared__ int s_stk[];
    const uint32_t t = blockIdx.x * TB + threadIdx.x, nt = gridDim.x * TB;
    LaneStack st;
    st.init(s_stk, threadIdx.x, TB);
    if (ovf) st.spill_to(lcap, ovf + t, nt);
    for (uint32_t i = t; i < n; i += nt) {
        const yg_ray r = rays[i];
        const f3 o = mk3(r.ox, r.oy, r.oz), d = mk3(r.dx, r.dy, r.dz);
        const float blo = r.tmin - 1e-4f * fabsf(r.tmin) - 1e-30f;
        occ[i] = trace_any<false, W>(bvh, o, d, r.tmin, r.tmax, -1, o, blo, r.tmax, st) ? 1 : 0;
    }
}

// Launch shape of an any-hit batch: grid, LDS stack entries per lane, and the
// overflow entries per lane (0: the whole stack fits in LDS).  max_ovf_ints
// caps the overflow buffer: the grid shrinks until it fits.
struct AnyShape {
    unsigned grid = 0;
    int lcap = 0;
    int ovf_entries = 0;
    size_t ovf_ints() const { return (size_t)ovf_entries * grid * TB; }
};

static AnyShape any_shape(const yg_scene* s, uint32_t n, size_t max_ovf_ints)
{
    AnyShape a;
    a.grid = grid_for(n, TB);
    a.lcap = std::min(s->stack_entries, YG_ANY_LCAP);
    a.ovf_entries = s->stack_entries - a.lcap;
    if (a.ovf_entries > 0) {
        const size_t fit = max_ovf_ints / ((size_t)a.ovf_entries * TB);
        if (fit < a.grid) a.grid = (unsigned)fit;
    }
    return a;
}

static int launch_trace_any(yg_scene* s, const yg_ray* d_rays, uint32_t n, uint8_t* d_occ, const AnyShape& a,
                            int* ovf, hipStream_t st)
{
    YG_LAUNCH_W(s->width, (k_trace_any<8>), (k_trace_any<4>), (k_trace_any<2>), dim3(a.grid), dim3(TB), stack_lds_bytes(a.lcap, TB),
                st, s->view(), d_rays, n, d_occ, a.lcap, a.ovf_entries > 0 ? ovf : nullptr);
    YG_HIP(hipGetLastError());
    return YG_OK;
}

template <int W>
__global__ __launch_bounds__(TB) void k_first_hit(BVHView bvh, const float* __restrict__ from,
                                                  const float* __restrict__ dir, uint32_t n,
                                                  float min_raydis, yg_hit* __restrict__ hits)
{
    extern __shared__ int s_stk[];
    const uint32_t i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    LaneStack st;
    st.init(s_stk, threadIdx.x, TB);
    const f3 fr = ld3(from + 3 * (size_t)i), d = ld3(dir + 3 * (size_t)i);
    float t;
    const int ti = first_hit<W>(bvh, fr, d, min_raydis, t, st);
    yg_hit h;
    h.prim = (ti >= 0) ? bvh.tprim[ti] : -1;
    h.t = t;
    h.u = 0; h.v = 0;
    hits[i] = h;
}

template <int W>
__global__ __launch_bounds__(TB) void k_is_shadowed(BVHView bvh, const float* __restrict__ p,
                                                    const float* __restrict__ l,
                                                    const int32_t* __restrict__ self_obj,
                                                    uint32_t n, int to_point, float min_raydis,
                                                    float self_bias, uint8_t* __restrict__ out)
{
    extern __shared__ int s_stk[];
    const uint32_t i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    LaneStack st;
    st.init(s_stk, threadIdx.x, TB);
    const int so = self_obj ? self_obj[i] : -1;
    out[i] = is_shadowed<W>(bvh, ld3(p + 3 * (size_t)i), ld3(l + 3 * (size_t)i), so, to_point != 0,
                         min_raydis, self_bias, st) ? 1 : 0;
}

__global__ void k_noise_eval(int what, yg_texture tex, float pwHL, const float* __restrict__ pts, uint32_t n,
                             float* __restrict__ out)
{
    __shared__ unsigned char perm[512];
    __shared__ float grad[768];
    if (tex.noise == YG_NOISE_NEWPERLIN) perm_to_lds(perm);  // (tex is a kernel argument: uniform)
    else noise_tables_to_lds(perm, grad);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 p = ld3(pts + 3 * (size_t)i);
    const int kind = tex.noise;
    switch (what) {
    case YG_EVAL_NOISE: out[i] = noise3(kind, p.x, p.y, p.z, perm, grad); break;
    case YG_EVAL_FBM: out[i] = fbm_w(p.x, p.y, p.z, pwHL, tex.lacunarity, tex.octaves, perm, kind, grad); break;
    case YG_EVAL_TURBULENCE:
        out[i] = turbulence(p.x, p.y, p.z, tex.depth, tex.size, tex.hard, perm, kind, grad);
        break;
    case YG_EVAL_TEX_FLOAT: out[i] = tex_float(tex, p, perm, grad); break;
    case YG_EVAL_TEX_COLOR: {
        const c3 c = tex_color(tex, p, perm, grad);
        out[3 * (size_t)i] = c.r; out[3 * (size_t)i + 1] = c.g; out[3 * (size_t)i + 2] = c.b;
        break;
    }
    }
}

__global__ void k_heightfield(int nx, int ny, float ex, float ey, float amp, float freq, float pwHL,
                              float lac, float oct, float* __restrict__ xyz)
{
    __shared__ unsigned char perm[512];
    perm_to_lds(perm);
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * ny) return;
    const int i = (int)(k % nx), j = (int)(k / nx);
    const float x = ex * (float)i / (float)(nx - 1);
    const float y = ey * (float)j / (float)(ny - 1);
    const float z = amp * fbm_w(freq * x, freq * y, 0.0f, pwHL, lac, oct, perm);
    xyz[3 * k] = x; xyz[3 * k + 1] = y; xyz[3 * k + 2] = z;
}

// ------------------------------------------------------------- C entry --

static int check_scene(yg_scene* s)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    if (!s->built) return yg_fail(YG_ERR_STATE, "scene not built (call yg_build)");
    return YG_OK;
}

extern "C" int yg_trace_closest(yg_scene* s, const yg_ray* d_rays, uint32_t n, yg_hit* d_hits,
                                void* stream)
{
    int rc = check_scene(s);
    if (rc) return rc;
    if (n == 0) return YG_OK;
    if (!d_rays || !d_hits) return yg_fail(YG_ERR_ARG, "null buffer");
    YG_LAUNCH_W(s->width, (k_trace_closest<false, 8>), (k_trace_closest<false, 4>), (k_trace_closest<false, 2>), dim3(grid_for(n, TB)), dim3(TB),
                stack_lds_bytes(s->stack_entries, TB), (hipStream_t)stream, s->view(), d_rays, n, d_hits, nullptr);
    YG_HIP(hipGetLastError());
    return YG_OK;
}

// Closest-hit with per-query node-visit and triangle-test totals (counters[2]).
extern "C" int yg_trace_closest_counted(yg_scene* s, const yg_ray* d_rays, uint32_t n,
                                        yg_hit* d_hits, unsigned long long* d_counters,
                                        void* stream)
{
    int rc = check_scene(s);
    if (rc) return rc;
    if (n == 0) return YG_OK;
    YG_LAUNCH_W(s->width, (k_trace_closest<true, 8>), (k_trace_closest<true, 4>), (k_trace_closest<true, 2>), dim3(grid_for(n, TB)), dim3(TB),
                stack_lds_bytes(s->stack_entries, TB), (hipStream_t)stream, s->view(), d_rays, n, d_hits, d_counters);
    YG_HIP(hipGetLastError());
    return YG_OK;
}

extern "C" int yg_trace_any(yg_scene* s, const yg_ray* d_rays, uint32_t n, uint8_t* d_occluded,
                            void* stream)
{
    int rc = check_scene(s);
    if (rc) return rc;
    if (n == 0) return YG_OK;
    if (!d_rays || !d_occluded) return yg_fail(YG_ERR_ARG, "null buffer");
    const AnyShape a = any_shape(s, n, YG_ANY_OVF_MAX);
    if (a.ovf_ints() > s->any_ovf_ints) {
        // grow-only per scene; hipFree waits for work still using the old buffer
        if (s->d_any_ovf) (void)hipFree(s->d_any_ovf);
        s->d_any_ovf = nullptr;
        s->any_ovf_ints = 0;
        hipError_t e = hipMalloc((void**)&s->d_any_ovf, sizeof(int) * a.ovf_ints());
        if (e != hipSuccess) return yg_fail(YG_ERR_NOMEM, std::string("any-hit stack buffer: ") + hipGetErrorString(e));
        s->any_ovf_ints = a.ovf_ints();
    }
    return launch_trace_any(s, d_rays, n, d_occluded, a, s->d_any_ovf, (hipStream_t)stream);
}

// ------------------------------------------- allocation-free any-hit --
// yg_trace_any_noalloc keeps the stack overflow in this file-scope array
// instead of a buffer the library allocates.  Each GPU holds its own copy
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (the array's name is a
// host-side stand-in, not a device address).
#define YG_SCRATCH_INTS ((size_t)16 << 20)  // 64 MiB
#define YG_SCRATCH_DEVICES 64
static __device__ __align__(256) int scratch[YG_SCRATCH_INTS];
static int* g_scratch_dev[YG_SCRATCH_DEVICES];
static std::mutex g_scratch_mu;

// scratch's device address on the current device, looked up once per device
static int scratch_on_device(int** out)
{
    int dev = 0;
    YG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= YG_SCRATCH_DEVICES) return yg_fail(YG_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return yg_fail(YG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = (int*)p;
    }
    *out = g_scratch_dev[dev];
    return YG_OK;
}

extern "C" int yg_trace_any_noalloc(yg_scene* s, const yg_ray* d_rays, uint32_t n, uint8_t* d_occluded,
                                    void* stream)
{
    int rc = check_scene(s);
    if (rc) return rc;
    if (n == 0) return YG_OK;
    if (!d_rays || !d_occluded) return yg_fail(YG_ERR_ARG, "null buffer");
    int* scr = nullptr;
    if ((rc = scratch_on_device(&scr))) return rc;
    // the grid shrinks until every thread's overflow column fits the scratch
    const AnyShape a = any_shape(s, n, YG_SCRATCH_INTS);
    if (a.grid == 0 || a.ovf_ints() > YG_SCRATCH_INTS)
        return yg_fail(YG_ERR_UNSUPPORTED, "traversal stack of " + std::to_string(s->stack_entries) +
                                               " entries per lane does not fit the allocation-free scratch (" +
                                               std::to_string(YG_SCRATCH_INTS) + " ints)");
    return launch_trace_any(s, d_rays, n, d_occluded, a, scr, (hipStream_t)stream);
}

extern "C" int yg_first_hit(yg_scene* s, const float* d_from, const float* d_dir, uint32_t n,
                            float min_raydis, yg_hit* d_hits, void* stream)
{
    int rc = check_scene(s);
    if (rc) return rc;
    if (n == 0) return YG_OK;
    if (!d_from || !d_dir || !d_hits) return yg_fail(YG_ERR_ARG, "null buffer");
    YG_LAUNCH_W(s->width, (k_first_hit<8>), (k_first_hit<4>), (k_first_hit<2>), dim3(grid_for(n, TB)), dim3(TB),
                stack_lds_bytes(s->stack_entries, TB), (hipStream_t)stream, s->view(), d_from, d_dir, n, min_raydis,
                d_hits);
    YG_HIP(hipGetLastError());
    return YG_OK;
}

extern "C" int yg_is_shadowed(yg_scene* s, const float* d_p, const float* d_l,
                              const int32_t* d_self_obj, uint32_t n, int32_t to_point,
                              float min_raydis, float self_bias, uint8_t* d_out, void* stream)
{
    int rc = check_scene(s);
    if (rc) return rc;
    if (n == 0) return YG_OK;
    if (!d_p || !d_l || !d_out) return yg_fail(YG_ERR_ARG, "null buffer");
    YG_LAUNCH_WS(s->width, s->view().nocast != 0, (k_is_shadowed<8>), (k_is_shadowed<4>), (k_is_shadowed<2>), (k_is_shadowed<24>), (k_is_shadowed<20>), (k_is_shadowed<18>), dim3(grid_for(n, TB)), dim3(TB),
                stack_lds_bytes(s->stack_entries, TB), (hipStream_t)stream, s->view(), d_p, d_l, d_self_obj, n, to_point,
                min_raydis, self_bias, d_out);
    YG_HIP(hipGetLastError());
    return YG_OK;
}

extern "C" int yg_noise_eval(int32_t what, const yg_texture* tex, const float* d_pts, uint32_t n,
                             float* d_out, void* stream)
{
    if (what < YG_EVAL_NOISE || what > YG_EVAL_TEX_COLOR) return yg_fail(YG_ERR_ARG, "bad eval kind");
    yg_texture t{};
    if (tex) t = *tex;
    if ((what >= YG_EVAL_TEX_FLOAT) && !tex) return yg_fail(YG_ERR_ARG, "texture required");
    if (t.noise < YG_NOISE_NEWPERLIN || t.noise > YG_NOISE_BLENDER) return yg_fail(YG_ERR_ARG, "bad noise generator");
    if (what >= YG_EVAL_TEX_FLOAT && (t.type < YG_TEX_CLOUDS || t.type > YG_TEX_MUSGRAVE))
        return yg_fail(YG_ERR_ARG, "bad texture type");
    int rc = perm_upload();
    if (rc) return rc;
    if (n == 0) return YG_OK;
    hipLaunchKernelGGL(k_noise_eval, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       what, t, dpow(t.lacunarity, -t.H), d_pts, n, d_out);
    YG_HIP(hipGetLastError());
    return YG_OK;
}

extern "C" int yg_gen_heightfield(int32_t nx, int32_t ny, float ex, float ey, float amp,
                                  float freq, float H, float lacunarity, float octaves,
                                  float* d_xyz, void* stream)
{
    if (nx < 2 || ny < 2 || !d_xyz) return yg_fail(YG_ERR_ARG, "bad heightfield arguments");
    int rc = perm_upload();
    if (rc) return rc;
    const uint64_t n = (uint64_t)nx * ny;
    hipLaunchKernelGGL(k_heightfield, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       nx, ny, ex, ey, amp, freq, dpow(lacunarity, -H), lacunarity, octaves, d_xyz);
    YG_HIP(hipGetLastError());
    return YG_OK;
}
