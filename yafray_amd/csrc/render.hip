// Wavefront renderer: scene_t::render(colorOutput_t&) in mono mode
// (scene.cc:538-600, 724-775) for sun, area, path and photon lights.
//
// Work items are the pixels of the render areas (renderArea_t, 64x64 plus a
// one-pixel overlap) laid out area by area in the reference's processing
// order.  Sequential per-area state of the reference -- the Halton samplers a
// pathlight keeps per area (pathlight.cc:725-738) and the global ourRandom()
// stream the area light draws from (arealight.cc:139) -- is turned into
// indices by exclusive prefix sums over the work items and path segments, so
// every path reads exactly the sample values the CPU renderer would.
#include <array>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "photon_dev.h"
#include "qmc.h"
#include "render_dev.h"

#ifndef RB
#define RB 64
#endif

__constant__ uint64_t YG_LCG_POW[64];

// ------------------------------------------------------------- helpers --

// A launch's area array (nareas entries, ascending item_start) is followed in
// the same buffer by its bucket index: entry b = the area holding item
// b << YG_AREA_LUT_SHIFT.  The area of an item is then one index load plus a
// forward scan over the few areas that start inside its bucket.
#define YG_AREA_LUT_SHIFT 8
__device__ __forceinline__ int find_area(const Area* areas, int nareas, uint32_t item)
{
    const uint32_t* lut = reinterpret_cast<const uint32_t*>(areas + nareas);
    int k = (int)lut[item >> YG_AREA_LUT_SHIFT];
    while (k + 1 < nareas && areas[k + 1].item_start <= item) k++;
    return k;
}

// camera_t::shootRay perspective (camera.cc:184-189)
__device__ __forceinline__ f3 cam_ray(const RenderScene& R, float px, float py)
{
    return norm3((R.vright * px + R.vup * py) + R.vto);
}

__device__ __forceinline__ void item_pixel(const Area& a, uint32_t item, int& i, int& j)
{
    const uint32_t li = item - a.item_start;
    i = a.y + (int)(li / (uint32_t)a.w);
    j = a.x + (int)(li % (uint32_t)a.w);
}

// Pixel (i, j) and camera-plane point (px, py) of a work item
// (scene_t::render(renderArea_t&), scene.cc:740-796).  First-pass items are
// the area's pixels in row-major order, shot through the pixel centre or,
// with AA_jitterfirst, through (RI_vdC(n), RI_S(n)) for the area's n-th
// pixel.  An item of AA pass p is sample ms of a marked pixel, cursam =
// p * AA_minsamples + ms, at
//   fx = 0.5 + AA_pixelwidth * (RI_LP(cursam + pixelNumber) - 0.5),
//   fy = 0.5 + AA_pixelwidth * (cursam / (AA_minsamples * AA_passes) - 0.5),
// formed in double as the reference's double literals make it.
__device__ __forceinline__ void item_camera(const RenderScene& R, const Area& a, uint32_t item, int& i, int& j,
                                            float& px, float& py)
{
    float fx = 0.5f, fy = 0.5f;
    if (R.aa_pass < 0) {
        item_pixel(a, item, i, j);
        if (R.aa_jitter) {
            const uint32_t n = item - a.item_start + 1;
            fx = ri_vdc(n);
            fy = ri_s(n);
        }
    } else {
        const uint2 src = R.aa_items[item];
        const uint32_t li = src.x - a.src_start;
        i = a.y + (int)(li / (uint32_t)a.w);
        j = a.x + (int)(li % (uint32_t)a.w);
        const uint32_t cursam = (uint32_t)R.aa_pass * (uint32_t)R.aa_mins + src.y;
        const uint32_t pixnum = (uint32_t)(j + i * R.resx);
        fx = (float)(0.5 + (double)R.aa_pw * ((double)ri_lp(cursam + pixnum) - 0.5));
        fy = (float)(0.5 + (double)R.aa_pw * ((double)((float)cursam * R.aa_totsamdiv) - 0.5));
    }
    px = (float)j + fx;
    py = (float)i + fy;
}

struct Counters {
    unsigned long long ext, shadow, extra, bshadow, draws;
};

__device__ __forceinline__ void block_add(unsigned long long* dst, uint32_t v)
{
    // one atomic per wavefront
    uint64_t s = v;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(dst, (unsigned long long)s);
}

// ------------------------------------------------------------- kernels --

// Minimum waves per SIMD the compiler must leave room for (register budget)
// in the traversal kernels; 1 = no constraint.
// k_primary fits 8 waves (64 VGPRs) and the shadow-ray kernels 6 (79) without
// spilling; tighter budgets spill to scratch.
#ifndef YG_PRIMARY_WPE
#define YG_PRIMARY_WPE 8
#endif
#ifndef YG_DIRECT_WPE
#define YG_DIRECT_WPE 1
#endif
#ifndef YG_SHADOW_WPE
#define YG_SHADOW_WPE 6
#endif

// Camera rays and their closest hits (scene_t::raytrace, scene.cc:200-235).
template <int W>
__global__ __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(YG_PRIMARY_WPE))) void k_primary(
                                                const RenderScene* __restrict__ scenes,
                                                const Area* __restrict__ areas, int nareas,
                                                uint32_t nitems, int32_t* __restrict__ item_ti,
                                                float* __restrict__ item_t, Counters* __restrict__ cnt,
                                                int zero_counters, int lds_cap, int* __restrict__ stk_ovf)
{
    extern __shared__ int s_stk[];
    const uint32_t item = blockIdx.x * RB + threadIdx.x;
    // the frame's ray counters start here (kernels after this one add to
    // them); antialiasing passes keep counting
    if (item == 0 && zero_counters) *cnt = Counters{};
    if (item >= nitems) return;
    LaneStack st;
    st.init(s_stk, threadIdx.x, RB);
    st.spill_to(lds_cap, stk_ovf + item, gridDim.x * RB);
    const Area a = areas[find_area(areas, nareas, item)];
    const RenderScene& R = scenes[a.scene];
    int i, j;
    float px, py;
    item_camera(R, a, item, i, j, px, py);
    const f3 ray = cam_ray(R, px, py);
    float t = INFINITY;
    int ti = -1;
    if (R.raydepth > 0) ti = first_hit<W>(R.bvh, R.cam_pos, ray, R.min_raydis, t, st);
    item_ti[item] = ti;
    item_t[item] = t;
}

__device__ __forceinline__ Surf primary_surf(const RenderScene& R, const Area& a, uint32_t item, int ti,
                                             float t, f3& eye)
{
    int i, j;
    float px, py;
    item_camera(R, a, item, i, j, px, py);
    const f3 ray = cam_ray(R, px, py);
    const f3 f = R.cam_pos + ray * R.min_raydis;
    Surf sp = make_surf(R.bvh, ti, f, ray, t);
    eye = R.cam_pos - sp.P;
    // state.screenpos (scene.cc:749-750)
    sp.sx = (float)(2.0 * (double)(px / (float)R.resx) - 1.0);
    sp.sy = (float)(1.0 - 2.0 * (double)(py / (float)R.resy));
    return sp;
}

// guessLight (arealight.cc:176-200) needs, for each area light with
// psamples > 0 at each camera-ray hit in front of it, whether each of its
// sample points faces the hit (L.N >= 0) and is unshadowed: two bitsets per
// light (words woff[li].., facing then lit), traced here for every point so
// that the sequential draw scan (k_guess_scan) only reads bits.  app: bit li
// set when light li guesses at this item.
template <int W>
__global__ __launch_bounds__(RB) void k_guess_vis(const RenderScene* __restrict__ scenes,
                                                  const Area* __restrict__ areas, int nareas, uint32_t nitems,
                                                  const int32_t* __restrict__ item_ti,
                                                  const float* __restrict__ item_t, const int32_t* __restrict__ woff,
                                                  int nw, uint32_t* __restrict__ vis, uint32_t* __restrict__ app,
                                                  int lds_cap, int* __restrict__ stk_ovf)
{
    extern __shared__ int s_stk[];
    const uint32_t item = blockIdx.x * RB + threadIdx.x;
    if (item >= nitems) return;
    uint32_t mask = 0;
    const int ti = item_ti[item];
    if (ti >= 0) {
        LaneStack st;
        st.init(s_stk, threadIdx.x, RB);
        st.spill_to(lds_cap, stk_ovf + item, gridDim.x * RB);
        const Area a = areas[find_area(areas, nareas, item)];
        const RenderScene& R = scenes[a.scene];
        f3 eye;
        const Surf sp = primary_surf(R, a, item, ti, item_t[item], eye);
        const f3 N = face_forward(sp.Ng, sp.N, norm3(eye));
        for (int li = 0; li < R.nl; li++) {
            const DevLight& L = R.lights[li];
            if (!L.use_in_render || L.type != YG_LIGHT_AREA || L.fsamples <= 0 || L.samples <= 1) continue;
            if (!(dot(sp.P - L.from, L.dir) > 0)) continue;
            mask |= 1u << li;
            uint32_t* fw = vis + (size_t)item * nw + woff[li];
            const int nwl = (L.samples + 31) / 32;
            for (int w = 0; w < nwl; w++) {
                uint32_t fb = 0, lb = 0;
                for (int b = 0; b < 32 && w * 32 + b < L.samples; b++) {
                    const f3 p = R.apts[L.pts + w * 32 + b];
                    if (dot(p - sp.P, N) < 0) continue;
                    fb |= 1u << b;
                    if (!is_shadowed<W>(R.bvh, sp.P, p, sp.obj, true, R.min_raydis, R.self_bias, st)) lb |= 1u << b;
                }
                fw[w] = fb;
                fw[nwl + w] = lb;
            }
        }
    }
    app[item] = mask;
}

// guessLight's draws in render order (one lane: the reference's glibc
// rand() stream is consumed point after point, psamples draws at most, and
// how many a point takes depends on what its draws saw).  rnd: the stream
// after the block shuffle's draws.  guess: 2-bit verdicts per light.  Counts
// the shadow rays the reference traces for its test points.
__global__ void k_guess_scan(const RenderScene* __restrict__ scenes, uint32_t nitems, const uint32_t* __restrict__ app,
                             const int32_t* __restrict__ woff, int nw, const uint32_t* __restrict__ vis,
                             const int32_t* __restrict__ rnd, uint64_t nrnd, uint32_t* __restrict__ guess,
                             Counters* __restrict__ cnt, unsigned long long* __restrict__ err)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const RenderScene& R = scenes[0];
    uint64_t pos = 0;
    unsigned long long rays = 0;
    for (uint32_t item = 0; item < nitems; item++) {
        const uint32_t m = app[item];
        uint32_t g = 0;
        for (int li = 0; m >> li; li++) {
            if (!((m >> li) & 1u)) continue;
            const DevLight& L = R.lights[li];
            const uint32_t* fw = vis + (size_t)item * nw + woff[li];
            const int nwl = (L.samples + 31) / 32;
            bool luz = false, sombra = false;
            for (int i = 0; i < L.fsamples; ++i) {
                if (luz && sombra) break;
                if (pos >= nrnd) {
                    *err = 1;
                    return;
                }
                const int rs = rnd[pos++] % L.samples;
                const bool facing = (fw[rs >> 5] >> (rs & 31)) & 1u;
                const bool lit = (fw[nwl + (rs >> 5)] >> (rs & 31)) & 1u;
                if (!facing) sombra = true;
                else {
                    rays++;
                    if (!lit) sombra = true;
                    else luz = true;
                }
            }
            const uint32_t v = (luz && sombra) ? YG_GUESS_PENUMBRA : (luz ? YG_GUESS_LIGHT : YG_GUESS_SHADOW);
            g |= v << (2 * li);
        }
        guess[item] = g;
    }
    atomicAdd(&cnt->shadow, rays);
}

// Per item: ourRandom() draws of the area lights and whether the pathlight
// samples at all (pathlight.cc:398: diffuse energy >= 0.05).
__global__ __launch_bounds__(RB) void k_prep(const RenderScene* __restrict__ scenes,
                                             const Area* __restrict__ areas, int nareas,
                                             uint32_t nitems, const int32_t* __restrict__ item_ti,
                                             const float* __restrict__ item_t,
                                             uint32_t* __restrict__ draws, uint32_t* __restrict__ qual,
                                             uint32_t* __restrict__ pre,
                                             const uint32_t* __restrict__ guess = nullptr)
{
    const uint32_t item = blockIdx.x * RB + threadIdx.x;
    if (item >= nitems) return;
    const int ti = item_ti[item];
    uint32_t dr = 0, q = 0, drp = 0;
    if (ti >= 0) {
        const Area a = areas[find_area(areas, nareas, item)];
        const RenderScene& R = scenes[a.scene];
        f3 eye;
        const Surf sp = primary_surf(R, a, item, ti, item_t[item], eye);
        for (int li = 0; li < R.nl; li++) {
            const DevLight& L = R.lights[li];
            if (!L.use_in_render || L.type != YG_LIGHT_AREA) continue;
            // guessLight's LIGHT and SHADOW verdicts draw nothing
            if (guess && L.samples > 1 && ((guess[item] >> (2 * li)) & 3u) != YG_GUESS_PENUMBRA) continue;
            if (dot(sp.P - L.from, L.dir) > 0) {
                dr += 2u * (uint32_t)L.samples;
                if (li < R.path_light) drp += 2u * (uint32_t)L.samples;
            }
        }
        if (R.path_light >= 0 && R.lights[R.path_light].use_in_render) {
            const c3 dif = get_diffuse(R, sp, eye);
            q = (cenergy(dif) < 0.05f) ? 0u : 1u;
        }
    }
    draws[item] = dr;
    qual[item] = q;
    if (pre) pre[item] = drp;  // draws of the area lights before the pathlight
}

// Sun and area lights at the primary hit, one colour per light.
// Sun shadows of the camera rays' hits for k_direct<W, false>: bit li of
// sun_vis[item] is set when sun light li (casting shadows, used in render)
// is blocked (scene_t::isShadowed, direction form, scene.cc:148-185).
template <int W>
__global__ __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(YG_SHADOW_WPE))) void k_sun_vis(
    const RenderScene* __restrict__ scenes, const Area* __restrict__ areas, int nareas, uint32_t nitems,
    const int32_t* __restrict__ item_ti, const float* __restrict__ item_t, uint32_t* __restrict__ sun_vis,
    Counters* __restrict__ cnt, int lds_cap, int* __restrict__ stk_ovf)
{
    extern __shared__ int s_stk[];
    const uint32_t item = blockIdx.x * RB + threadIdx.x;
    uint32_t nshadow = 0, mask = 0;
    if (item < nitems && item_ti[item] >= 0) {
        LaneStack st;
        st.init(s_stk, threadIdx.x, RB);
        st.spill_to(lds_cap, stk_ovf + item, gridDim.x * RB);
        const Area a = areas[find_area(areas, nareas, item)];
        const RenderScene& R = scenes[a.scene];
        f3 eye;
        const Surf sp = primary_surf(R, a, item, item_ti[item], item_t[item], eye);
        for (int li = 0; li < R.nl && li < 32; li++) {
            const DevLight& L = R.lights[li];
            if (!L.use_in_render || L.type != YG_LIGHT_SUN || !L.cast_shadows) continue;
            nshadow++;
            if (is_shadowed<W>(R.bvh, sp.P, L.dir, sp.obj, false, R.min_raydis, R.self_bias, st)) mask |= 1u << li;
        }
    }
    if (item < nitems) sun_vis[item] = mask;
    block_add(&cnt->shadow, nshadow);
}

// Light sum of the camera rays' hits (scene_t::light, scene.cc:259-277).
// AREA: the launch has area lights, whose samples take ourRandom() draws
// here and trace their shadow rays here, as do the sun lights.  Without area
// lights the sun shadows come from k_sun_vis and this kernel only shades
// (no traversal: far fewer registers).
template <int W, bool AREA>
__global__ __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(YG_DIRECT_WPE))) void k_direct(
                                               const RenderScene* __restrict__ scenes,
                                               const Area* __restrict__ areas, int nareas,
                                               uint32_t nitems, const int32_t* __restrict__ item_ti,
                                               const float* __restrict__ item_t,
                                               const uint32_t* __restrict__ lcg_off,
                                               c3* __restrict__ lightc, int nl_max,
                                               Counters* __restrict__ cnt, int lds_cap, int* __restrict__ stk_ovf,
                                               const uint32_t* __restrict__ sun_vis,
                                               const uint32_t* __restrict__ pathd = nullptr,
                                               const uint32_t* __restrict__ guess = nullptr)
{
    extern __shared__ int s_stk[];
    const uint32_t item = blockIdx.x * RB + threadIdx.x;
    uint32_t nshadow = 0, ndraws = 0;
    if (item < nitems) {
        const int ti = item_ti[item];
        if (ti >= 0) {
            LaneStack st;
            if (AREA) {
                st.init(s_stk, threadIdx.x, RB);
                st.spill_to(lds_cap, stk_ovf + item, gridDim.x * RB);
            }
            const Area a = areas[find_area(areas, nareas, item)];
            const RenderScene& R = scenes[a.scene];
            f3 eye;
            const Surf sp = primary_surf(R, a, item, ti, item_t[item], eye);
            int32_t seed = 0;
            bool seeded = false;
            uint64_t skip = 0;  // draws of the pathlight's vertices passed before the first area light
            for (int li = 0; li < R.nl; li++) {
                const DevLight& L = R.lights[li];
                c3 c = mkc(0, 0, 0);
                if (AREA && pathd && li == R.path_light) {
                    // the pathlight's GI paths draw here in the stream (area lights at their vertices)
                    if (seeded) seed = lcg_jump((uint32_t)seed, pathd[item], YG_LCG_POW);
                    else skip += pathd[item];
                }
                if (L.use_in_render) {
                    if (L.type == YG_LIGHT_SUN) {
                        if (AREA) c = sun_illuminate<W>(R, L, sp, eye, st, nshadow);
                        else if (!(L.cast_shadows && li < 32 && ((sun_vis[item] >> li) & 1u)))
                            c = from_light(R, sp, L.dir, L.col, eye);  // sunLight_t::illuminate (sunlight.cc:35-57)
                    } else if (AREA && L.type == YG_LIGHT_AREA) {
                        if (!seeded) {
                            seed = lcg_jump(R.lcg_seed, (uint64_t)(lcg_off[item] - lcg_off[R.item_base]) + skip,
                                            YG_LCG_POW);
                            seeded = true;
                        }
                        const uint32_t gs =
                            (guess && L.samples > 1) ? ((guess[item] >> (2 * li)) & 3u) : YG_GUESS_PENUMBRA;
                        if (gs == YG_GUESS_PENUMBRA && dot(sp.P - L.from, L.dir) > 0)
                            ndraws += 2u * (uint32_t)L.samples;
                        c = area_illuminate<W>(R, L, sp, eye, seed, st, nshadow, gs);
                    }
                }
                lightc[(size_t)item * nl_max + li] = c;
            }
        }
    }
    block_add(&cnt->shadow, nshadow);
    block_add(&cnt->draws, ndraws);
}

// ------------------------------------------------ adjoint accumulators --
// Gradient buffers (double) and their block-private LDS copies; used by the
// GI levels' shading kernel in replay mode and by the per-pixel adjoint.

#define ADJ_MAXD 8
#define ADJ_LDS_MAT 256
#define ADJ_LDS_MOD 16
#define ADJ_LDS_LIGHT 8
#define ADJ_GRID 1024u  // blocks of a gradient-accumulating launch (about 4 per CU)

struct AdjAcc {
    double* mat;   // nmat*3
    double* mod;   // nmod*2 (d/dcolour, d/dlambda)
    double* light; // nl
    int nmat, nmod, nl;
};

// block-private accumulators in LDS, flushed once per block
struct AdjLds {
    double mat[ADJ_LDS_MAT * 3];
    double mod[ADJ_LDS_MOD * 2];
    double light[ADJ_LDS_LIGHT];
};

__device__ __forceinline__ void adj_add_vertex(AdjLds& L, const AdjAcc& G, bool lds_ok, int mat, const RhoD& r,
                                               c3 gamma)
{
    const c3 ga = cmul(gamma, r.d_alb);
    double* dst = lds_ok ? L.mat : G.mat;
    atomicAdd(&dst[3 * mat], (double)ga.r);
    atomicAdd(&dst[3 * mat + 1], (double)ga.g);
    atomicAdd(&dst[3 * mat + 2], (double)ga.b);
    if (r.mod >= 0) {
        double* dm = lds_ok ? L.mod : G.mod;
        const c3 a = cmul(gamma, r.d_amp), l = cmul(gamma, r.d_lam);
        atomicAdd(&dm[2 * r.mod], (double)(a.r + a.g + a.b));
        atomicAdd(&dm[2 * r.mod + 1], (double)(l.r + l.g + l.b));
    }
}

__device__ __forceinline__ void adj_lds_init(AdjLds& L)
{
    for (int k = threadIdx.x; k < ADJ_LDS_MAT * 3; k += blockDim.x) L.mat[k] = 0.0;
    for (int k = threadIdx.x; k < ADJ_LDS_MOD * 2; k += blockDim.x) L.mod[k] = 0.0;
    for (int k = threadIdx.x; k < ADJ_LDS_LIGHT; k += blockDim.x) L.light[k] = 0.0;
    __syncthreads();
}

__device__ __forceinline__ void adj_lds_flush(AdjLds& L, const AdjAcc& G, bool lds_ok)
{
    __syncthreads();
    if (!lds_ok) return;
    for (int k = threadIdx.x; k < G.nmat * 3; k += blockDim.x)
        if (L.mat[k] != 0.0) atomicAdd(&G.mat[k], L.mat[k]);
    for (int k = threadIdx.x; k < G.nmod * 2; k += blockDim.x)
        if (L.mod[k] != 0.0) atomicAdd(&G.mod[k], L.mod[k]);
    for (int k = threadIdx.x; k < G.nl; k += blockDim.x)
        if (L.light[k] != 0.0) atomicAdd(&G.light[k], L.light[k]);
}

__device__ __forceinline__ bool item_output(const Area& a, uint32_t item, int& i, int& j)
{
    item_pixel(a, item, i, j);
    return j >= a.rx && j < a.rx + a.rw && i >= a.ry && i < a.ry + a.rh;
}

// One bounce of the pathlight's full-GI loop (pathlight.cc:305-347) for the
// active path segments, in three launches so that each traversal kernel
// carries only the registers it needs:
//   k_ext     hemisphere direction from the Halton tables, closest hit;
//             queues the paths whose ray hit something
//   k_shadow  the sun shadow rays of the queued hit points (isShadowed)
//   k_shade   scene_t::light at the hit from the visibility bits, the
//             path's contribution and its next-vertex state
// Path p = item*S + sample.  Every value passed between them is a float the
// fused evaluation would hold in a register, so results are bit-identical.

// Per-level path queue of hit paths (order irrelevant: paths are independent).
__device__ __forceinline__ void queue_push(bool want, uint32_t p, uint32_t* qcount, uint32_t* queue)
{
    const uint64_t m = __ballot(want);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    uint32_t base = 0;
    const int leader = __ffsll((unsigned long long)m) - 1;
    if (lane == leader) base = atomicAdd(qcount, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (want) queue[base + __popcll(m & ((1ull << lane) - 1))] = p;
}

#ifndef YG_BOUNCE_WPE
#define YG_BOUNCE_WPE 1
#endif
// ADJ: replay for the adjoint (contri/nseg of the forward are kept); RPL: rays
// each lane traces, one after another (a compile-time count: a runtime loop
// costs the one-ray kernel 20 registers)
template <bool ADJ, int W, int RPL>
__global__ __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(YG_BOUNCE_WPE))) void k_ext(
    const RenderScene* __restrict__ scenes, const Area* __restrict__ areas, int nareas, int j,
    const uint32_t* __restrict__ active, const uint32_t* __restrict__ nact, uint32_t nactive_cap, int S, int D,
    const int32_t* __restrict__ item_ti, const float* __restrict__ item_t, const uint32_t* __restrict__ qoff,
    const uint32_t* __restrict__ hidx, const f3* __restrict__ st_where, const f3* __restrict__ st_hn,
    const int32_t* __restrict__ st_ti, c3* __restrict__ st_col, float* __restrict__ st_s,
    c3* __restrict__ contri, uint8_t* __restrict__ nseg, f3* __restrict__ h_p, f3* __restrict__ h_ray,
    int32_t* __restrict__ h_ti, uint32_t* __restrict__ qcount, uint32_t* __restrict__ queue,
    uint32_t* __restrict__ flag, const uint32_t* __restrict__ hbase, Counters* __restrict__ cnt,
    int lds_cap, int* __restrict__ stk_ovf, const uint32_t* __restrict__ dense)
{
    extern __shared__ int s_stk[];
    const uint32_t tid = blockIdx.x * RB + threadIdx.x;
    LaneStack st;
    st.init(s_stk, threadIdx.x, RB);
    st.spill_to(lds_cap, stk_ovf + tid, gridDim.x * RB);
    // *dense: every listed path is alive (k_path_init, or k_compact compacted
    // the list); else the list is the previous level's and only paths whose
    // flag k_shade set continue
    const bool all_alive = *dense != 0;
    // each lane traces RPL rays of the block's consecutive slice
    for (int r = 0; r < RPL; r++) {
    const uint32_t idx = ((uint32_t)blockIdx.x * (uint32_t)RPL + (uint32_t)r) * RB + threadIdx.x;
    uint32_t next = 0;
    bool hit = false;
    uint32_t p = 0;
    // *nact = this level's listed paths (written on the device by k_path_init
    // or k_compact); the grid covers every path
    if (idx < *nact && (all_alive || flag[active[idx]])) {
        p = active[idx];
        flag[p] = 0;  // set again by k_shade if the path continues
        const uint32_t item = p / (uint32_t)S;
        const int smp = (int)(p % (uint32_t)S);
        const int ai = find_area(areas, nareas, item);
        const Area a = areas[ai];
        const RenderScene& R = scenes[a.scene];
        f3 where, N, Ru, Rv;
        c3 raycolor;
        uint32_t k;
        if (j == 0) {
            f3 eye;
            const Surf sp = primary_surf(R, a, item, item_ti[item], item_t[item], eye);
            where = sp.P;
            N = face_forward(sp.Ng, sp.N, eye);
            Ru = sp.NU;
            Rv = sp.NV;
            raycolor = mkc(1, 1, 1);
            k = (qoff[item] - qoff[a.item_start]) * (uint32_t)S + (uint32_t)smp;
            if (hbase) k += hbase[(size_t)ai * D];  // level-0 draws of the area's earlier passes
        } else {
            where = st_where[p];
            N = st_hn[p];
            const int pti = st_ti[p];
            const float4 A = gload(R.bvh.tris + 3 * pti), B = gload(R.bvh.tris + 3 * pti + 1),
                         C = gload(R.bvh.tris + 3 * pti + 2);
            // sp.NU, sp.NV of the previous vertex: around its shading normal
            create_cs(surf_normal(R.bvh, pti, where, mk3(A.w, B.w, C.w)), Ru, Rv);
            raycolor = st_col[p];
            k = hidx[(size_t)j * nactive_cap + p];
        }
        f3 ray;
        if (R.ps_on) {
            // photonSampler_t::nextDirection (pathtools.cc:238-261): the area's
            // own Halton values; level 0 inside the walk's sector (the item's
            // plan), the ray weighed by weight x 2 z1
            const float* hv = R.ps_h + ((size_t)(ai - (int)R.area_base) * R.ps_hstride + k) * R.ps_nd;
            if (j == 0) {
                const int nsec = R.ps_par * R.ps_mer;
                const int32_t* hits = R.ps_hits + (size_t)item * nsec;
                int sec = 0, acc = hits[0];
                while (acc <= smp && sec + 1 < nsec) acc += hits[++sec];
                const int c0 = sec / R.ps_mer, c1 = sec % R.ps_mer;
                const float z1 = ((float)c0 + hv[0]) * R.ps_pdiv;
                const float z2 = ((float)c1 + hv[1]) * R.ps_mdiv;
                const float wv = R.ps_w[(size_t)item * nsec + sec] * 2.0f * z1;
                raycolor = mkc(wv, wv, wv);
                if (!ADJ) st_col[p] = raycolor;
                ray = ps_direction(z1, z2, N, Ru, Rv);
            } else {
                const float z1 = hv[2 * j];
                const float z2 = (float)((double)hv[2 * j + 1] * 2.0 * M_PI);
                ray = ps_direction(z1, z2, N, Ru, Rv);
            }
        } else {
            const float z1 = *(const YG_GLOBAL float*)(R.halton + (size_t)(2 * j) * R.hstride + k);
            const float z2 = *(const YG_GLOBAL float*)(R.halton + (size_t)(2 * j + 1) * R.hstride + k);
            ray = hemi_direction(z1, z2, N, Ru, Rv);
        }
        float s;
        if (j == 0) {
            s = dot(ray, N);
            st_s[p] = s;
        } else s = st_s[p];
        if (cenergy(raycolor) < 0.05f) {
            if (!ADJ) nseg[p] = (uint8_t)j;
        } else {
            next = 1;
            float t;
            const f3 f = where + ray * R.min_raydis;
            // pathlight.cc:317: firstHit with the shadow flag
            const int ti = first_hit<W>(R.bvh, where, ray, R.min_raydis, t, st, true);
            if (ti < 0) {
                if (!ADJ) {
                    contri[(size_t)p * D + j] = cmul(cscale(raycolor, s), R.bg);
                    nseg[p] = (uint8_t)(j + 1);
                }
            } else {
                h_p[p] = f + ray * t;  // make_surf's P
                h_ray[p] = ray;
                h_ti[p] = ti;
                hit = true;
            }
        }
    }
    queue_push(hit, p, qcount, queue);
    if (!ADJ) block_add(&cnt->ext, next);
    }
}

template <bool ADJ, int W, int RPL>
__global__ __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(YG_SHADOW_WPE))) void k_shadow(
                                               const RenderScene* __restrict__ scenes,
                                               const Area* __restrict__ areas, int nareas, int S,
                                               const uint32_t* __restrict__ qcount,
                                               const uint32_t* __restrict__ queue,
                                               const f3* __restrict__ h_p, const int32_t* __restrict__ h_ti,
                                               uint32_t* __restrict__ vis, Counters* __restrict__ cnt,
                                               int lds_cap, int* __restrict__ stk_ovf)
{
    extern __shared__ int s_stk[];
    const uint32_t tid = blockIdx.x * RB + threadIdx.x;
    LaneStack st;
    st.init(s_stk, threadIdx.x, RB);
    st.spill_to(lds_cap, stk_ovf + tid, gridDim.x * RB);
    for (int r = 0; r < RPL; r++) {
    const uint32_t idx = ((uint32_t)blockIdx.x * (uint32_t)RPL + (uint32_t)r) * RB + threadIdx.x;
    uint32_t nshadow = 0;
    if (idx < *qcount) {
        const uint32_t p = queue[idx];
        const Area& a = areas[find_area(areas, nareas, p / (uint32_t)S)];
        const RenderScene& R = scenes[a.scene];
        const f3 P = h_p[p];
        const int obj = gload(R.bvh.tobj + h_ti[p]);
        uint32_t mask = 0;
        for (int li = 0; li < R.nl && li < 32; li++) {
            const DevLight& L = R.lights[li];
            if (!L.use_in_indirect || L.type != YG_LIGHT_SUN || !L.cast_shadows) continue;
            nshadow++;
            if (is_shadowed<W>(R.bvh, P, L.dir, obj, false, R.min_raydis, R.self_bias, st)) mask |= 1u << li;
        }
        vis[p] = mask;
    }
    if (!ADJ) {
        block_add(&cnt->shadow, nshadow);
        block_add(&cnt->bshadow, nshadow);
    }
    }
}

#define SB 256  // shading block size (no traversal stack)

// Hit vertices of the GI paths whose area lights are evaluated after the
// levels (pathlight + area lights used in indirect light): entry p*D + j.
// Null pointers: the launch has no such lights.
struct VertexRec {
    f3* p;             // hit point
    f3* ray;           // direction the path arrived along
    int32_t* ti;       // leaf-ordered triangle
    uint32_t* vis;     // sun shadow bits of the vertex (k_shadow)
    c3* rs;            // raycolor * (N.w0) of the path at this vertex
    uint32_t* draws;   // ourRandom() draws of its area lights (0: no vertex or none lit)
};

// Forward: the level's light sum, contribution and next-vertex state.
// ADJ (replay): the same state update, plus the gradient terms of this path
// vertex.  With W = dLoss/dpixel * rho0 * mult * power (per item) and the
// estimator's pixel term rho0 * mult * power * sum_b contri_b,
// contri_b = s * R_b * X_b (X_b = lf_b * rho_b at a hit, bg at a miss),
//   d pixel / d rho_j = W * (s * R_j * lf_j + (sum_{b>j} contri_b) / rho_j)
//   d pixel / d power = W * sum_j s * R_j * lfp_j * rho_j
// where the suffix sum comes from the forward's contri of this path.  A
// vertex whose albedo has a zero channel (the suffix cannot be divided) is
// counted in *degenerate and left to the per-path replay.
template <bool ADJ>
__global__ __launch_bounds__(SB) void k_shade(const RenderScene* __restrict__ scenes,
                                              const Area* __restrict__ areas, int nareas, int j, int S, int D,
                                              const uint32_t* __restrict__ qcount,
                                              const uint32_t* __restrict__ queue,
                                              const f3* __restrict__ h_p, const f3* __restrict__ h_ray,
                                              const int32_t* __restrict__ h_ti, const uint32_t* __restrict__ vis,
                                              f3* __restrict__ st_where, f3* __restrict__ st_hn,
                                              int32_t* __restrict__ st_ti, c3* __restrict__ st_col,
                                              const float* __restrict__ st_s, c3* __restrict__ contri,
                                              uint8_t* __restrict__ nseg, uint32_t* __restrict__ flag,
                                              AdjAcc G, const c3* __restrict__ w_item,
                                              unsigned long long* __restrict__ degenerate, VertexRec vx)
{
    extern __shared__ double s_adj[];
    AdjLds& Ls = *reinterpret_cast<AdjLds*>(s_adj);
    const bool lds_ok = ADJ && G.nmat <= ADJ_LDS_MAT && G.nmod <= ADJ_LDS_MOD && G.nl <= ADJ_LDS_LIGHT;
    if (ADJ) adj_lds_init(Ls);
    uint32_t ndeg = 0;
    const uint32_t nq = *qcount;
    // grid-stride: the replay launches a bounded grid so that each block's
    // LDS accumulators gather many vertices before their one flush
    for (uint32_t idx = blockIdx.x * SB + threadIdx.x; idx < nq; idx += gridDim.x * SB) {
        const uint32_t p = queue[idx];
        const Area& a = areas[find_area(areas, nareas, p / (uint32_t)S)];
        const RenderScene& R = scenes[a.scene];
        const int ti = h_ti[p];
        const f3 ray = h_ray[p];
        Surf sp;
        sp.ti = ti;
        sp.P = h_p[p];
        const float4 A = R.bvh.tris[3 * ti], B = R.bvh.tris[3 * ti + 1], C = R.bvh.tris[3 * ti + 2];
        sp.Ng = mk3(A.w, B.w, C.w);
        sp.N = surf_normal(R.bvh, ti, sp.P, sp.Ng);
        create_cs(sp.N, sp.NU, sp.NV);
        sp.obj = R.bvh.tobj[ti];
        sp.mat = R.bvh.tmat[ti];
        // scene_t::light with indirect = true (sun lights only inside GI paths)
        const uint32_t mask = vis[p];
        const f3 eye = (sp.P + (-ray)) - sp.P;
        const c3 raycolor = (j == 0 && !R.ps_on) ? mkc(1, 1, 1) : st_col[p];
        const float s = st_s[p];
        c3 rho;
        if (!ADJ) {
            c3 fl = mkc(0, 0, 0);
            for (int li = 0; li < R.nl; li++) {
                const DevLight& L = R.lights[li];
                if (!L.use_in_indirect || L.type != YG_LIGHT_SUN) continue;
                if (L.cast_shadows && (mask >> li & 1u)) fl = cadd(fl, mkc(0, 0, 0));
                else fl = cadd(fl, from_light(R, sp, L.dir, L.col, eye));
            }
            // raycolor's factor: getDiffuse(state, tempsp, -ray) (pathlight.cc:339)
            rho = get_diffuse(R, sp, -ray);
            if (R.gp_on) {
                // getLight with the irradiance map (pathlight.cc:140-185):
                // irradiance x getDiffuse(state, sp, N), then the lights
                const f3 Nv = face_forward(sp.Ng, sp.N, -ray);
                const c3 rn = diffuse_reads_eye(R.mats, R.mods, sp.mat) ? get_diffuse(R, sp, Nv) : rho;
                fl = cadd(cmul(gp_irradiance(R.gp, sp.P, Nv), rn), fl);
            }
            contri[(size_t)p * D + j] = cmul(cscale(raycolor, s), fl);
            nseg[p] = (uint8_t)(j + 1);
            if (vx.draws) {
                // area lights used in indirect light draw 2 x samples at a vertex
                // in front of their plane (arealight.cc:93-153); k_vertex_light
                // redoes this vertex's light sum once their stream positions are known
                uint32_t dr = 0;
                for (int li = 0; li < R.nl; li++) {
                    const DevLight& L = R.lights[li];
                    if (L.use_in_indirect && L.type == YG_LIGHT_AREA && dot(sp.P - L.from, L.dir) > 0)
                        dr += 2u * (uint32_t)L.samples;
                }
                const size_t v = (size_t)p * D + j;
                vx.draws[v] = dr;
                if (dr) {
                    vx.p[v] = sp.P;
                    vx.ray[v] = ray;
                    vx.ti[v] = ti;
                    vx.vis[v] = mask;
                    vx.rs[v] = cscale(raycolor, s);
                }
            }
        } else {
            // diffuse_d's rho is get_diffuse's value (one modulator at most)
            const RhoD r = diffuse_d(R.mats, R.mods, R.texs, sp.mat, sp.P);
            rho = r.rho;
            const c3 W = w_item[p / (uint32_t)S];
            if (W.r != 0 || W.g != 0 || W.b != 0) {
                c3 lf = mkc(0, 0, 0), lfp = mkc(0, 0, 0);
                int sun = -1;
                for (int li = 0; li < R.nl; li++) {
                    const DevLight& L = R.lights[li];
                    if (!L.use_in_indirect || L.type != YG_LIGHT_SUN) continue;
                    if (sun < 0) sun = li;
                    if (L.cast_shadows && (mask >> li & 1u)) continue;
                    const f3 N = face_forward(sp.Ng, sp.N, norm3(eye));
                    const float inte = dot(N, L.dir);
                    if (inte < 0) continue;
                    lf = cadd(lf, cscale(L.col, inte));
                    lfp = cadd(lfp, cscale(L.color, inte));
                }
                c3 La = mkc(0, 0, 0);
                const int ns = nseg[p];
                for (int b = j + 1; b < ns; b++) La = cadd(La, contri[(size_t)p * D + b]);
                if ((W.r != 0 && rho.r == 0) || (W.g != 0 && rho.g == 0) || (W.b != 0 && rho.b == 0)) {
                    ndeg = 1;
                } else {
                    const c3 sR = cscale(raycolor, s);
                    const c3 div = mkc(rho.r != 0 ? La.r / rho.r : 0.0f, rho.g != 0 ? La.g / rho.g : 0.0f,
                                       rho.b != 0 ? La.b / rho.b : 0.0f);
                    adj_add_vertex(Ls, G, lds_ok, sp.mat, r, cmul(W, cadd(cmul(sR, lf), div)));
                    if (sun >= 0) {
                        const c3 q = cmul(W, cmul(cmul(sR, lfp), rho));
                        atomicAdd(lds_ok ? &Ls.light[sun] : &G.light[sun], (double)(q.r + q.g + q.b));
                    }
                }
            }
        }
        st_hn[p] = face_forward(sp.Ng, sp.N, -ray);
        c3 next_col = cmul(raycolor, rho);
        uint32_t cont = 1;
        if (!ADJ && R.rr_from > 0 && j + 1 >= R.rr_from) {
            // Russian roulette before the path's next level
            const float q = fminf(1.0f, cenergy(next_col));
            if (rr_unit(p, j + 1, R.rr_seed) < q) next_col = cscale(next_col, 1.0f / q);
            else cont = 0;
        }
        st_col[p] = next_col;
        st_where[p] = sp.P;
        st_ti[p] = ti;
        flag[p] = cont;
    }
    if (ADJ) {
        adj_lds_flush(Ls, G, lds_ok);
        block_add(degenerate, ndeg);
    }
}

// Scatter flagged paths into the next active list and give each its Halton
// call index within its render area (hidx_level: npaths entries of this level).
// The next level's Halton indices (every surviving path's rank in its area)
// and its path list: the survivors compacted when fewer than `compact_below`
// of this level's listed paths survive, else this level's list again (its
// ended paths' lanes exit at once; *dense_next tells the level which).
__global__ void k_compact(const Area* __restrict__ areas, int nareas, uint32_t npaths, int S,
                          const uint32_t* __restrict__ flag, const uint32_t* __restrict__ ex,
                          const uint32_t* __restrict__ active_cur, const uint32_t* __restrict__ nact_cur,
                          uint32_t* __restrict__ active, uint32_t* __restrict__ hidx_level,
                          uint32_t* __restrict__ nact_next, uint32_t* __restrict__ dense_next, float compact_below,
                          const uint32_t* __restrict__ hbase_level, int D)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t listed = *nact_cur, alive = ex[npaths - 1] + flag[npaths - 1];
    const bool compact = (float)alive < compact_below * (float)listed;
    if (p == npaths - 1) {
        *nact_next = compact ? alive : listed;
        *dense_next = compact ? 1u : 0u;
    }
    if (!compact && p < listed) active[p] = active_cur[p];
    if (p >= npaths || !flag[p]) return;
    const uint32_t item = p / (uint32_t)S;
    const int ai = find_area(areas, nareas, item);
    const uint32_t e = ex[p];
    if (compact) active[e] = p;
    // the path's rank among the area's paths at this level (plus, with
    // antialiasing, the level's draws in the area's earlier passes)
    hidx_level[p] = e - ex[(size_t)areas[ai].item_start * S] + (hbase_level ? hbase_level[(size_t)ai * D] : 0u);
}

// Clears the vertex draw counts of a frame (graph-safe: no memset node).
__global__ void k_fill_u32(uint32_t* __restrict__ a, size_t n, uint32_t v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// An item's ourRandom() draws: its primary-hit area lights plus every draw
// its GI paths' vertices take (vx_off: exclusive scan of the vertex draws
// in (item, sample, level) order, the order takeSample visits them).
__global__ void k_item_pathd(uint32_t nitems, int S, int D, const uint32_t* __restrict__ vx_draws,
                             const uint32_t* __restrict__ vx_off, const uint32_t* __restrict__ draws,
                             uint32_t* __restrict__ pathd, uint32_t* __restrict__ tdraws)
{
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    const size_t b = (size_t)item * S * D, e = b + (size_t)S * D - 1;
    const uint32_t pd = vx_off[e] + vx_draws[e] - vx_off[b];
    pathd[item] = pd;
    tdraws[item] = draws[item] + pd;
}

// 1 for a vertex that draws (the lit list's flags).
__global__ void k_vx_flag(size_t n, const uint32_t* __restrict__ vx_draws, uint32_t* __restrict__ flag)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = vx_draws[i] > 0 ? 1u : 0u;
}

// The drawing vertices in vertex order, and their count (pos: exclusive
// scan of the flags), so that k_vertex_light's waves hold lit vertices only.
__global__ void k_vx_list(size_t n, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                          uint32_t* __restrict__ lit, uint32_t* __restrict__ nlit)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) lit[pos[i]] = (uint32_t)i;
    if (i == n - 1) *nlit = pos[i] + flag[i];
}

// scene_t::light(indirect) at the GI vertices that light area lights
// (pathlight.cc:186-191): the sun terms from k_shadow's bits and the area
// lights drawing from the stream position of the vertex -- after the item's
// draws, its area lights before the pathlight and the earlier vertices of its
// paths -- summed in light order like k_shade's sum, which this replaces.
template <int W>
__global__ __launch_bounds__(RB) void k_vertex_light(const RenderScene* __restrict__ scenes,
                                                    const Area* __restrict__ areas, int nareas,
                                                    const uint32_t* __restrict__ lit,
                                                    const uint32_t* __restrict__ nlit,
                                                    int S, int D, VertexRec vx, const uint32_t* __restrict__ vx_off,
                                                    const uint32_t* __restrict__ lcg_off,
                                                    const uint32_t* __restrict__ pre, c3* __restrict__ contri,
                                                    Counters* __restrict__ cnt, int lds_cap,
                                                    int* __restrict__ stk_ovf)
{
    extern __shared__ int s_stk[];
    const uint32_t idx = blockIdx.x * RB + threadIdx.x;
    uint32_t nshadow = 0, ndraws = 0;
    if (idx < *nlit) {
        const uint32_t v = lit[idx];
        LaneStack st;
        st.init(s_stk, threadIdx.x, RB);
        st.spill_to(lds_cap, stk_ovf + idx, gridDim.x * RB);
        const uint32_t p = v / (uint32_t)D, item = p / (uint32_t)S;
        const Area& a = areas[find_area(areas, nareas, item)];
        const RenderScene& R = scenes[a.scene];
        const int ti = vx.ti[v];
        const f3 ray = vx.ray[v];
        Surf sp;
        sp.ti = ti;
        sp.P = vx.p[v];
        const float4 A = R.bvh.tris[3 * ti], B = R.bvh.tris[3 * ti + 1], C = R.bvh.tris[3 * ti + 2];
        sp.Ng = mk3(A.w, B.w, C.w);
        sp.N = surf_normal(R.bvh, ti, sp.P, sp.Ng);
        create_cs(sp.N, sp.NU, sp.NV);
        sp.obj = R.bvh.tobj[ti];
        sp.mat = R.bvh.tmat[ti];
        const f3 eye = (sp.P + (-ray)) - sp.P;
        const uint64_t pos = (uint64_t)(lcg_off[item] - lcg_off[R.item_base]) + pre[item] +
                             (vx_off[v] - vx_off[(size_t)item * S * D]);
        int32_t seed = lcg_jump(R.lcg_seed, pos, YG_LCG_POW);
        const uint32_t mask = vx.vis[v];
        c3 fl = mkc(0, 0, 0);
        for (int li = 0; li < R.nl; li++) {
            const DevLight& L = R.lights[li];
            if (!L.use_in_indirect) continue;
            if (L.type == YG_LIGHT_SUN) {
                if (L.cast_shadows && (mask >> li & 1u)) fl = cadd(fl, mkc(0, 0, 0));
                else fl = cadd(fl, from_light(R, sp, L.dir, L.col, eye));
            } else if (L.type == YG_LIGHT_AREA) {
                fl = cadd(fl, area_illuminate<W>(R, L, sp, eye, seed, st, nshadow));
            }
        }
        contri[v] = cmul(vx.rs[v], fl);
        ndraws = vx.draws[v];
    }
    block_add(&cnt->shadow, nshadow);
    block_add(&cnt->bshadow, nshadow);
    block_add(&cnt->draws, ndraws);
}

// Antialiasing passes: adds this pass's draws of one GI level to each area's
// Halton counter (level 0: every sample of every qualifying item; level L:
// the paths the compaction kept), so that the next pass continues the area's
// sequences where this one stopped.
__global__ void k_area_draws(const Area* __restrict__ areas, int nareas, uint32_t nitems, int S, int D,
                             int level, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ ex,
                             uint32_t* __restrict__ hbase)
{
    const int ai = blockIdx.x * blockDim.x + threadIdx.x;
    if (ai >= nareas) return;
    const uint32_t s = areas[ai].item_start, e = (ai + 1 < nareas) ? areas[ai + 1].item_start : nitems;
    if (e <= s) return;
    const uint32_t k = (level == 0) ? 1u : (uint32_t)S;  // level 0 counts items, levels > 0 paths
    const size_t lo = (size_t)s * k, hi = (size_t)e * k - 1;
    uint32_t n = ex[hi] + flag[hi] - ex[lo];
    if (level == 0) n *= (uint32_t)S;
    hbase[(size_t)ai * D + level] += n;
}

// Level-0 active list (every sample of every qualifying item) and its
// length nact[0]; continuation flags and per-level queue lengths start at 0.
// ------------------------------------------------ photonSampler_t --
//
// The sampler of a render area is made at its first takeSample
// (getSampler, pathlight.cc:725-738); its constructor starts every Halton
// dimension at an ourRandomI() draw (pathtools.cc:80-92), so the areas that
// make one draw 2 (depth + 1) values each, in area order.

// per area: does a pixel reach takeSample (one wave per area)
__global__ __launch_bounds__(64) void k_ps_flag(const RenderScene* __restrict__ scenes,
                                                const Area* __restrict__ areas, int nareas,
                                                const uint32_t* __restrict__ qual)
{
    const int a = blockIdx.x;
    if (a >= nareas) return;
    const Area A = areas[a];
    const uint32_t n = (uint32_t)(A.w * A.h);
    int any = 0;
    for (uint32_t k = threadIdx.x; k < n; k += 64) any |= qual[A.item_start + k] ? 1 : 0;
    any = __any(any);
    if (threadIdx.x == 0) scenes[A.scene].ps_flag[a - (int)scenes[A.scene].area_base] = any ? 1u : 0u;
}

// nextPrime (mcqmc.h:84-97)
__device__ __forceinline__ int next_prime_dev(int last)
{
    int np = last + (last & 1) + 1;
    for (;;) {
        int dv = 3;
        bool ispr = true;
        while (ispr && dv * dv <= np) {
            ispr = (np % dv) != 0;
            dv += 2;
        }
        if (ispr) return np;
        np += 2;
    }
}

// one thread per (area, dimension): Halton::setStart of its ourRandomI()
// draw (mcqmc.h:24-33), then getNext (mcqmc.h:35-49) for every draw of the
// area's qualifying samples
__global__ void k_ps_halton(const RenderScene* __restrict__ scenes, const Area* __restrict__ areas, int nareas,
                            const uint32_t* __restrict__ qual, const uint32_t* __restrict__ qoff, int S)
{
    const RenderScene& R = scenes[0];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = t / R.ps_nd, d = t % R.ps_nd;
    if (a >= nareas || !R.ps_flag[a]) return;
    uint32_t rank = 0;
    for (int b = 0; b < a; b++) rank += R.ps_flag[b];
    uint32_t i = (uint32_t)lcg_jump(R.lcg_seed, (uint64_t)rank * (uint64_t)R.ps_nd + (uint64_t)d + 1, YG_LCG_POW);
    int base = 2;
    for (int k = 0; k < d; k++) base = next_prime_dev(base);
    const double inv = 1.0 / (double)base;
    double value = 0.0, f = inv, factor = inv;
    while (i > 0) {
        value += (double)(i % (uint32_t)base) * factor;
        i /= (uint32_t)base;
        factor *= f;
    }
    const Area A = areas[a];
    const uint32_t last = A.item_start + (uint32_t)(A.w * A.h) - 1u;
    uint32_t cnt = (qoff[last] + qual[last] - qoff[A.item_start]) * (uint32_t)S;
    if (cnt > R.ps_hstride) cnt = R.ps_hstride;
    // [area][draw][dimension]: a wave's stores fill a few lines
    float* out = R.ps_h + (size_t)a * R.ps_hstride * R.ps_nd + d;
    for (uint32_t k = 0; k < cnt; k++) {
        const double r = 1.0 - value - 0.0000000001;
        if (inv < r) value += inv;
        else {
            double hh, h = inv;
            do {
                hh = h;
                h *= inv;
            } while (h >= r);
            value += hh + h - 1.0;
        }
        out[(size_t)k * R.ps_nd] = (float)value;
    }
}

// photonSampler_t::getCoords (pathtools.cc:118-145); a NaN sine (|z| above 1
// by rounding, undefined in the reference's int cast) is taken as 0
__device__ __forceinline__ int ps_sector(const RenderScene& R, f3 v, f3 N, f3 Ru, f3 Rv)
{
    float x = dot(v, Ru), y = dot(v, Rv), z = dot(v, N);
    float fix = (float)sqrt(1.0 - (double)(z * z));
    if (fix != fix) fix = 0.0f;
    int first = (int)(fix / R.ps_pdiv);
    if (first >= R.ps_par) first--;
    if (fix > 1.0f) fix = 1.0f;
    if (fix != 0.0f) {
        x /= fix;
        y /= fix;
    }
    if (x > 1.0f) x = 1.0f;
    if (x < -1.0f) x = -1.0f;
    float an = (float)acos((double)x);
    if (y < 0) an = (float)(2.0 * M_PI - (double)an);
    int second = (int)(an / R.ps_mdiv);
    if (second >= R.ps_mer) second--;
    return first * R.ps_mer + second;
}

__device__ __forceinline__ float max_abs_diff(c3 a, c3 b)
{
    return fmaxf(fabsf(a.r - b.r), fmaxf(fabsf(a.g - b.g), fabsf(a.b - b.b)));
}

// photonSampler_t::samplingFrom (pathtools.cc:172-236) at every qualifying
// item, one wave per item: the lanes gather the 3 x sectors nearest raw
// photons within the map's radius that face N (sorted by distance, then
// store index, with a bitonic network in LDS, cut whenever the buffer
// fills); lane 0 sums their colours per sector, spreads the samples by the
// sectors' energy differences (giveMaxDiff) and writes the plan
#define PS_BUF 1024
__global__ __launch_bounds__(64) void k_ps_plan(const RenderScene* __restrict__ scenes,
                                                const Area* __restrict__ areas, int nareas,
                                                const int32_t* __restrict__ item_ti,
                                                const float* __restrict__ item_t,
                                                const uint32_t* __restrict__ qual)
{
    __shared__ float s_d[PS_BUF];
    __shared__ int s_i[PS_BUF];
    __shared__ c3 s_e[64];
    __shared__ float s_w[64];
    __shared__ int s_h[64];
    const uint32_t item = blockIdx.x;
    const int lane = threadIdx.x;
    if (!qual[item]) return;
    const Area a = areas[find_area(areas, nareas, item)];
    const RenderScene& R = scenes[a.scene];
    f3 eye;
    const Surf sp = primary_surf(R, a, item, item_ti[item], item_t[item], eye);
    const f3 P = sp.P, N = face_forward(sp.Ng, sp.N, eye), Ru = sp.NU, Rv = sp.NV;
    const GPhotonView& G = R.gp;
    const int nsec = R.ps_par * R.ps_mer, K = 3 * nsec;
    const float radius = G.radius;
    auto sort_buf = [&](int n) {
        int n2 = 1;
        while (n2 < n) n2 <<= 1;
        for (int k = n + lane; k < n2; k += 64) {
            s_d[k] = INFINITY;
            s_i[k] = 0x7fffffff;
        }
        __syncthreads();
        for (int size = 2; size <= n2; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int k = lane; k < n2; k += 64) {
                    const int o = k ^ stride;
                    if (o > k) {
                        const bool up = (k & size) == 0;
                        const bool gt = (s_d[o] < s_d[k]) || (s_d[o] == s_d[k] && s_i[o] < s_i[k]);
                        if (gt == up) {
                            const float td = s_d[k]; s_d[k] = s_d[o]; s_d[o] = td;
                            const int ti = s_i[k]; s_i[k] = s_i[o]; s_i[o] = ti;
                        }
                    }
                }
                __syncthreads();
            }
    };
    int n = 0;
    // after the first cut only candidates below the K-th kept (distance, then
    // index) can still be among the K nearest
    float thr_d = INFINITY;
    int thr_i = 0x7fffffff;
    if (G.nrec > 0) {
        const float pp[3] = {P.x, P.y, P.z};
        int g0[3], g1[3];
        for (int q = 0; q < 3; q++) {
            long lo = (long)floorf((pp[q] - radius - G.rgmin[q]) / radius);
            long hh = (long)floorf((pp[q] + radius - G.rgmin[q]) / radius);
            if (lo < 0) lo = 0;
            if (hh >= G.rgdim[q]) hh = G.rgdim[q] - 1;
            g0[q] = (int)lo;
            g1[q] = (int)hh;
        }
        for (int z = g0[2]; z <= g1[2]; z++)
            for (int y = g0[1]; y <= g1[1]; y++)
                for (int x = g0[0]; x <= g1[0]; x++) {
                    const int c = (z * G.rgdim[1] + y) * G.rgdim[0] + x;
                    const int q0 = G.rgstart[c], q1 = G.rgstart[c + 1];
                    for (int q00 = q0; q00 < q1; q00 += 64) {
                        const int q = q00 + lane;
                        bool keep = false;
                        float dis = 0.f;
                        int idx = 0;
                        if (q < q1) {
                            idx = G.rgitems[q];
                            dis = len3(G.rpos[idx] - P);
                            keep = !(dis > radius) && ((dis < thr_d) || (dis == thr_d && idx < thr_i)) &&
                                   dot(G.rdir[idx], N) > 0.0f;
                        }
                        const uint64_t m = __ballot(keep);
                        if (!m) continue;
                        if (n + 64 > PS_BUF || (n >= 2 * K && thr_i == 0x7fffffff)) {
                            sort_buf(n);
                            if (n > K) n = K;
                            if (n == K) {
                                thr_d = s_d[K - 1];
                                thr_i = s_i[K - 1];
                            }
                        }
                        if (keep) s_d[n + (int)__popcll(m & ((1ull << lane) - 1ull))] = dis;
                        if (keep) s_i[n + (int)__popcll(m & ((1ull << lane) - 1ull))] = idx;
                        n += (int)__popcll(m);
                        __syncthreads();
                    }
                }
        if (n > 0) sort_buf(n);
        if (n > K) n = K;
    }
    if (lane != 0) return;
    const int par = R.ps_par, mer = R.ps_mer;
    for (int k = 0; k < nsec; k++) s_e[k] = mkc(0, 0, 0);
    for (int k = 0; k < n; k++) {
        const int idx = s_i[k];
        const f3 d = G.rdir[idx];
        if (dot(d, N) > 0.0f) {
            const int sec = ps_sector(R, d, N, Ru, Rv);
            s_e[sec] = cadd(s_e[sec], G.rcol[idx]);
        }
    }
    float totaldiff = 0;
    for (int i = 0; i < par; ++i)
        for (int j = 0; j < mer; ++j) {
            const c3 e = s_e[i * mer + j];
            float mx = 0, t;
            const int il = i - 1, ih = i + 1, jl = j - 1, jh = j + 1;
            if (il >= 0) {
                if (jl >= 0) { t = max_abs_diff(s_e[il * mer + jl], e); if (t > mx) mx = t; }
                t = max_abs_diff(s_e[il * mer + j], e); if (t > mx) mx = t;
                if (jh < mer) { t = max_abs_diff(s_e[il * mer + jh], e); if (t > mx) mx = t; }
            }
            if (jl >= 0) { t = max_abs_diff(s_e[i * mer + jl], e); if (t > mx) mx = t; }
            if (jh < mer) { t = max_abs_diff(s_e[i * mer + jh], e); if (t > mx) mx = t; }
            if (ih < par) {
                if (jl >= 0) { t = max_abs_diff(s_e[ih * mer + jl], e); if (t > mx) mx = t; }
                t = max_abs_diff(s_e[ih * mer + j], e); if (t > mx) mx = t;
                if (jh < mer) { t = max_abs_diff(s_e[ih * mer + jh], e); if (t > mx) mx = t; }
            }
            s_w[i * mer + j] = mx;
            totaldiff += mx;
        }
    int assign = R.samples - nsec;
    const float sph = (totaldiff != 0.0f) ? (float)assign / totaldiff : 0;
    for (int k = 0; k < nsec; k++) {
        int h = (int)((double)(sph * s_w[k]) + 0.5);
        if (h > assign) h = assign;
        if (h < 0) h = 0;  // only for a non-finite energy (the reference's cast is undefined there)
        assign -= h;
        s_h[k] = h + 1;
    }
    bool noway = false;
    while (assign > 0) {
        for (int k = 0; k < nsec && assign; k++)
            if ((s_h[k] > 1) || noway) {
                s_h[k]++;
                assign--;
            }
        noway = true;
    }
    int mx = 1;
    for (int k = 0; k < nsec; k++)
        if (s_h[k] > mx) mx = s_h[k];
    const float maxf = (float)mx;
    for (int k = 0; k < nsec; k++) {
        R.ps_hits[(size_t)item * nsec + k] = s_h[k];
        R.ps_w[(size_t)item * nsec + k] = maxf / (float)s_h[k];
    }
    R.ps_multi[item] = (float)(1.0 / (double)(maxf * (float)nsec));
}

__global__ void k_path_init(uint32_t nitems, int S, int D, const uint32_t* __restrict__ qual,
                            const uint32_t* __restrict__ qoff, uint32_t* __restrict__ active,
                            uint32_t* __restrict__ flag, uint32_t* __restrict__ nact,
                            uint32_t* __restrict__ qcnt, uint32_t* __restrict__ dense)
{
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item == 0) {
        for (int j = 0; j < D; j++) qcnt[j] = 0;
        dense[0] = 1;
    }
    if (item >= nitems) return;
    if (item == nitems - 1) nact[0] = (qoff[item] + qual[item]) * (uint32_t)S;
    for (int i = 0; i < S; i++) flag[(size_t)item * S + i] = 0;
    if (!qual[item]) return;
    for (int i = 0; i < S; i++) active[(size_t)qoff[item] * S + i] = item * (uint32_t)S + i;
}

// ------------------------------------------------------ antialiasing --
//
// scene_t::render(renderArea_t&) passes (scene.cc:777-816) over the
// first-pass items: every area keeps its own colours (overlap border
// included) in item_col.

// abscol2bri of a colour difference (color.h:71): the difference in float
// (colorA_t operator-), the weighted sum in double (color.h's unqualified
// fabs is the C library's double fabs), returned as float.
__device__ __forceinline__ float abscol2bri(c3 a, c3 b)
{
    return (float)(0.299f * fabs((double)(a.r - b.r)) + 0.587f * fabs((double)(a.g - b.g)) +
                   0.114f * fabs((double)(a.b - b.b)));
}

// renderArea_t::checkResample (renderblock.cc:22-69): an item is resampled
// when one of its 8 neighbours in the area (clamped at the area's edge)
// differs by abscol2bri >= threshold; count = AA_minsamples items then, else 0.
__global__ void k_aa_mask(const Area* __restrict__ areas, int nareas, uint32_t nitems, const c3* __restrict__ item_col,
                          float threshold, uint32_t mins, uint32_t* __restrict__ count)
{
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    const Area& a = areas[find_area(areas, nareas, item)];
    const int l = (int)(item - a.item_start), i = l / a.w, j = l % a.w;
    const int im = i > 0 ? i - 1 : 0, ip = i + 1 < a.h ? i + 1 : a.h - 1;
    const int jm = j > 0 ? j - 1 : 0, jp = j + 1 < a.w ? j + 1 : a.w - 1;
    const c3* A = item_col + a.item_start;
    const c3 c = A[l];
    const int nb[8] = {im * a.w + jm, im * a.w + j, im * a.w + jp, i * a.w + jm,
                       i * a.w + jp, ip * a.w + jm, ip * a.w + j, ip * a.w + jp};
    bool need = false;
    for (int q = 0; q < 8 && !need; q++) need = abscol2bri(c, A[nb[q]]) >= threshold;
    count[item] = need ? mins : 0u;
}

// The pass's items: for each marked first-pass item its samples 0..mins-1,
// in item order (areas, then the area's pixels row-major, then samples).
__global__ void k_aa_items(uint32_t nitems, const uint32_t* __restrict__ count, const uint32_t* __restrict__ off,
                           uint2* __restrict__ aa_items)
{
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    for (uint32_t ms = 0; ms < count[item]; ms++) aa_items[off[item] + ms] = make_uint2(item, ms);
}

// The running average of pass `pass` (scene.cc:809-814):
// colour = (mf * colour + totcol) / (mf + n), mf = pass * n + 1, where the
// colour division multiplies by the float reciprocal (color.h:282-286).
__global__ void k_aa_combine(uint32_t nitems, const uint32_t* __restrict__ count, const uint32_t* __restrict__ off,
                             const c3* __restrict__ aa_col, int pass, c3* __restrict__ item_col)
{
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    const uint32_t n = count[item];
    if (!n) return;
    c3 tot = mkc(0, 0, 0);
    for (uint32_t ms = 0; ms < n; ms++) tot = cadd(tot, aa_col[off[item] + ms]);
    const float mf = (float)(pass * (int)n + 1);
    float den = mf + (float)n;
    if (den != 0) den = (float)(1.0 / (double)den);
    const c3 c = item_col[item];
    item_col[item] = mkc((mf * c.r + tot.r) * den, (mf * c.g + tot.g) * den, (mf * c.b + tot.b) * den);
}

// First item of each area in the pass (starts[na] = the pass's item count).
__global__ void k_aa_starts(const Area* __restrict__ areas, int nareas, uint32_t nitems,
                            const uint32_t* __restrict__ count, const uint32_t* __restrict__ off,
                            uint32_t* __restrict__ starts)
{
    const int ai = blockIdx.x * blockDim.x + threadIdx.x;
    if (ai < nareas) starts[ai] = off[areas[ai].item_start];
    else if (ai == nareas) starts[ai] = off[nitems - 1] + count[nitems - 1];
}

// renderArea_t::out (renderblock.cc:8-20): the real rectangle of each area.
__global__ void k_aa_output(const RenderScene* __restrict__ scenes, const Area* __restrict__ areas, int nareas,
                            uint32_t nitems, const c3* __restrict__ item_col)
{
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    const Area& a = areas[find_area(areas, nareas, item)];
    int i, j;
    item_pixel(a, item, i, j);
    if (j < a.rx || j >= a.rx + a.rw || i < a.ry || i >= a.ry + a.rh) return;
    const RenderScene& R = scenes[a.scene];
    const c3 c = item_col[item];
    float* o = R.out_rgb + 3 * ((size_t)i * R.resx + j);
    o[0] = c.r; o[1] = c.g; o[2] = c.b;
}

// Ordered light sum, pathlight estimate (pathlight.cc:349-403) and pixel output.
__global__ __launch_bounds__(RB) void k_finalize(const RenderScene* __restrict__ scenes,
                                                 const Area* __restrict__ areas, int nareas,
                                                 uint32_t nitems, int S, int D,
                                                 const int32_t* __restrict__ item_ti,
                                                 const float* __restrict__ item_t,
                                                 const uint32_t* __restrict__ qual,
                                                 const c3* __restrict__ lightc, int nl_max,
                                                 const c3* __restrict__ contri,
                                                 const uint8_t* __restrict__ nseg,
                                                 Counters* __restrict__ cnt, c3* __restrict__ item_col,
                                                 const c3* __restrict__ cache_pathc)
{
    const uint32_t item = blockIdx.x * RB + threadIdx.x;
    uint32_t extra_n = 0;
    if (item < nitems) {
        const Area a = areas[find_area(areas, nareas, item)];
        const RenderScene& R = scenes[a.scene];
        int i, j;
        item_pixel(a, item, i, j);
        const int ti = item_ti[item];
        c3 col;
        if (ti < 0) col = R.bg;
        else {
            c3 pathc = mkc(0, 0, 0);
            if (cache_pathc) {
                // the irradiance cache's interpolation (k_cache_interp)
                if (R.path_light >= 0 && qual[item]) pathc = cache_pathc[item];
            } else if (R.path_light >= 0 && qual[item]) {
                c3 total = mkc(0, 0, 0), sub[4];
                for (int q = 0; q < 4; q++) sub[q] = mkc(0, 0, 0);
                for (int smp = 0; smp < S; smp++) {
                    const size_t p = (size_t)item * S + smp;
                    const int ns = nseg[p];
                    for (int b = 0; b < ns; b++) {
                        const c3 c = contri[p * D + b];
                        total = cadd(total, c);
                        sub[smp % 4] = cadd(sub[smp % 4], c);
                    }
                }
                const float mult = R.ps_on ? R.ps_multi[item] : (float)(1.0 / (double)(float)S);
                // adaptive repetition test (pathlight.cc:349-374): only counted
                float minR = 1000, minG = 1000, minB = 1000, maxR = 0, maxG = 0, maxB = 0;
                for (int q = 0; q < 4; q++) {
                    const c3 v = cscale(sub[q], mult);
                    maxR = fmaxf(maxR, v.r); maxG = fmaxf(maxG, v.g); maxB = fmaxf(maxB, v.b);
                    minR = fminf(minR, v.r); minG = fminf(minG, v.g); minB = fminf(minB, v.b);
                }
                const float ene = fmaxf(fabsf(minR), fmaxf(fabsf(minG), fabsf(minB)));
                const float dmax = fmaxf(fabsf(maxR - minR), fmaxf(fabsf(maxG - minG), fabsf(maxB - minB)));
                float diff = (float)((double)dmax / ((double)((ene < 1.0f) ? 1.0f : ene) * 2.0 * 2.0));
                diff /= R.threshold;
                if ((int)(diff * diff) - 1 > 0) extra_n = 1;
                total = cscale(total, mult / 1.0f);
                f3 eye;
                const Surf sp = primary_surf(R, a, item, ti, item_t[item], eye);
                const c3 dif = get_diffuse(R, sp, eye);
                pathc = cscale(cmul(dif, total), R.path_power);
            }
            col = mkc(0, 0, 0);
            for (int li = 0; li < R.nl; li++) {
                if (!R.lights[li].use_in_render) continue;
                col = cadd(col, (li == R.path_light) ? pathc : lightc[(size_t)item * nl_max + li]);
            }
        }
        if (R.tm) col = expgam_adjust(col, R.tm_e, R.tm_g, R.tm_clamp != 0);
        if (item_col) item_col[item] = col;  // antialiasing: the area keeps every colour
        else if (j >= a.rx && j < a.rx + a.rw && i >= a.ry && i < a.ry + a.rh) {
            float* o = R.out_rgb + 3 * ((size_t)i * R.resx + j);
            o[0] = col.r; o[1] = col.g; o[2] = col.b;
        }
    }
    block_add(&cnt->extra, extra_n);
}

#include "lcache_dev.h"

// ------------------------------------------------------------------ host --

struct HostCam {
    f3 pos, vright, vup, vto;
};

// camera_t::camera_t perspective part (camera.cc:27-87)
static HostCam camera_setup(const yg_camera& c)
{
    const f3 pos = ld3(c.pos), look = ld3(c.look), up = ld3(c.up);
    f3 vup = up - pos, vto = look - pos;
    f3 vright = cross(vup, vto);
    vup = cross(vright, vto);
    vup = norm3(vup);
    vright = norm3(vright);
    vright = vright * -1.0f;
    (void)normlen3(vto);
    vup = vup * (c.aspect * (float)c.resy / (float)c.resx);
    vto = vto * c.focal - (vup + vright) * 0.5f;
    vup = vup / (float)c.resy;
    vright = vright / (float)c.resx;
    HostCam h;
    h.pos = pos; h.vright = vright; h.vup = vup; h.vto = vto;
    return h;
}

// blockSpliter_t (renderblock.cc:71-101): areas in processing order (the
// mono loop takes regions.back() first, scene.cc:585-588)
static void block_regions(int width, int height, int b, uint32_t seed, std::vector<Area>& out)
{
    int bw = width / b, bh = height / b;
    if (width % b) bw++;
    if (height % b) bh++;
    const int n = bw * bh;
    std::vector<int> scram(n);
    std::vector<int32_t> rnd;
    glibc_rand(seed, n, rnd);
    for (int i = 0; i < n; ++i) scram[i] = i;
    for (int i = 0; i < n; ++i) std::swap(scram[i], scram[rnd[i] % n]);
    std::vector<Area> regions(n);
    int sec = 0;
    for (int i = 0; i < bh; ++i)
        for (int j = 0; j < bw; ++j) {
            Area r{};
            r.x = r.rx = j * b;
            r.y = r.ry = i * b;
            r.w = r.rw = width - r.rx;
            r.h = r.rh = height - r.ry;
            if (r.w > b) r.w = r.rw = b;
            if (r.h > b) r.h = r.rh = b;
            if (r.x > 0) { r.x--; r.w++; }
            if (r.y > 0) { r.y--; r.h++; }
            if ((r.x + r.w) < (width - 1)) r.w++;
            if ((r.y + r.h) < (height - 1)) r.h++;
            regions[scram[sec]] = r;
            sec++;
        }
    out.clear();
    for (int k = n - 1; k >= 0; k--) out.push_back(regions[k]);
}

// areaLight_t::fillQuad (arealight.cc:31-73) and constructor (:75-91)
static void area_setup(const yg_light& Lg, DevLight& L, std::vector<f3>& pts, std::vector<f3>& jh,
                       std::vector<f3>& jv)
{
    const f3 a = ld3(Lg.a), b = ld3(Lg.b), c = ld3(Lg.c), d = ld3(Lg.d);
    const int samp = Lg.samples;
    const float ab_len = len3(b - a), bc_len = len3(c - b), cd_len = len3(d - c), da_len = len3(a - d);
    const float max_v = (bc_len > da_len) ? bc_len : da_len;
    const float max_h = (ab_len > cd_len) ? ab_len : cd_len;
    const int v_samp = (int)((double)(max_v / (max_v + max_h)) * 2.0 * (double)sqrtf((float)samp));
    const int h_samp = (int)((double)(max_h / (max_v + max_h)) * 2.0 * (double)sqrtf((float)samp));
    L.pts = (int)pts.size();
    std::vector<f3> P(samp > 0 ? samp : 1, mk3(0, 0, 0)), H(P), V(P);
    int index = 0;
    if (samp == 1) index = 1;
    else {
        const f3 inc_v_a = (d - a) / (float)v_samp;
        const f3 inc_v_b = (c - b) / (float)v_samp;
        f3 line_a = a + inc_v_a * 0.5f;
        f3 line_b = b + inc_v_b * 0.5f;
        for (int i = 0; i < v_samp; ++i) {
            const f3 inc_h = (line_b - line_a) / (float)h_samp;
            f3 point = line_a + inc_h * 0.5f;
            for (int j = 0; j < h_samp; ++j) {
                const float vf = ((float)j) / ((float)h_samp);
                const f3 inc_v = inc_v_a * (float)(1.0 - (double)vf) + inc_v_b * vf;
                if (index < samp) { P[index] = point; H[index] = inc_h; V[index] = inc_v; }
                index++;
                point = point + inc_h;
            }
            line_a = line_a + inc_v_a;
            line_b = line_b + inc_v_b;
        }
    }
    L.samples = std::min(index, samp);
    pts.insert(pts.end(), P.begin(), P.end());
    jh.insert(jh.end(), H.begin(), H.end());
    jv.insert(jv.end(), V.begin(), V.end());
    L.dir = norm3(cross(b - a, d - a));
    L.col = cscale(mkc(Lg.color[0], Lg.color[1], Lg.color[2]), Lg.power);
    L.from = (((a + b) + c) + d) * 0.25f;
    L.fsamples = Lg.psamples;
    L.corner = a;
    L.toX = b - a;
    L.toY = d - a;
}

// Counts render-buffer allocations: a changed value means some buffer may
// have moved (a freed block can come back at the same address, so pointers
// alone do not tell).
static uint64_t g_alloc_serial = 0;

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    uint64_t serial = 0;  // g_alloc_serial of this allocation
    int ensure(size_t n)
    {
        if (n <= cap && p) return YG_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(n, 1);
        hipError_t e = hipMalloc((void**)&p, sizeof(T) * want);
        if (e != hipSuccess) return yg_fail(YG_ERR_NOMEM, std::string("render buffer: ") + hipGetErrorString(e));
        cap = want;
        serial = ++g_alloc_serial;
        return YG_OK;
    }
};

// Host bytes last copied into one device buffer.
struct UploadCache {
    uint64_t serial = 0;
    std::vector<uint8_t> bytes;
};

// Copies h into d unless d's allocation already holds exactly these bytes
// from the previous call: a repeated frame uploads nothing.
template <typename T>
static int upload_cached(DevBuf<T>& d, const std::vector<T>& h, UploadCache& c, hipStream_t st)
{
    const size_t n = sizeof(T) * h.size();
    if (n == 0) return YG_OK;
    const uint8_t* b = reinterpret_cast<const uint8_t*>(h.data());
    if (c.serial == d.serial && c.bytes.size() == n && memcmp(c.bytes.data(), b, n) == 0) return YG_OK;
    YG_HIP(hipMemcpyAsync(d.p, h.data(), n, hipMemcpyHostToDevice, st));
    c.bytes.assign(b, b + n);
    c.serial = d.serial;
    return YG_OK;
}

// Per-process render workspace (grow-only, reused across calls).
struct RenderWS {
    DevBuf<RenderScene> scenes;
    DevBuf<Area> areas;
    DevBuf<DevLight> lights;
    DevBuf<f3> apts, ajh, ajv;
    DevBuf<int32_t> item_ti;
    DevBuf<float> item_t;
    DevBuf<uint32_t> draws, lcg_off, qual, qoff;
    DevBuf<uint32_t> sun_vis;  // camera-ray hits: blocked sun lights (bit per light)
    DevBuf<c3> lightc;
    DevBuf<uint32_t> active0, active1, flag, ex, hidx;
    DevBuf<f3> st_where, st_hn;
    DevBuf<int32_t> st_ti;
    DevBuf<c3> st_col;
    DevBuf<float> st_s;
    DevBuf<c3> contri;
    DevBuf<uint8_t> nseg;
    DevBuf<f3> h_p, h_ray;        // hit point and direction of a level's extension ray
    DevBuf<int32_t> h_ti;
    DevBuf<uint32_t> vis, queue;  // shadow visibility bits; queue of hit paths
    DevBuf<uint32_t> qcnt;        // queue length per level
    DevBuf<uint32_t> nact;        // listed paths per level
    DevBuf<uint32_t> dense;       // per level: 1 = every listed path is alive
    int rays_lane = 1;            // rays each lane of k_ext / k_shadow traces
    float compact_below = 1.0f;   // survivor fraction below which a level is compacted
    DevBuf<Counters> cnt;
    DevBuf<uint8_t> scan_tmp;
    DevBuf<uint8_t> akey0, akey1;
    DevBuf<c3> w_item;                   // adjoint: per-item path weight
    DevBuf<unsigned long long> adeg;     // adjoint: vertices left to the per-path replay
    bool pow_ready = false;
    // shape of the last launch (the adjoint replays it)
    uint32_t nitems = 0, npaths = 0;
    int S = 0, D = 0, na = 0, nl_max = 0;
    bool any_path = false;
    size_t stk_bytes = 0;      // dynamic LDS of the traversal launches
    size_t stk_bytes_lvl = 0;  // ... of the frame launches (camera, direct, GI levels)
    int stk_entries = 1;       // stack bound of the launch's trees
    int stk_lds = 0;           // entries the frame launches keep in LDS (0: all)
    int lds_cap = 0;           // lds_cap argument of those launches (0: no overflow)
    DevBuf<int> stk_ovf;       // their stack entries beyond the LDS part
    int width = 2;         // BVH width of the scenes in the launch
    bool nocast = false;   // some scene of the launch has objects that cast no shadows
    // a pathlight with area lights used in indirect light: their draws at GI
    // vertices are placed in the stream after the levels (k_vertex_light)
    bool area_path = false;
    DevBuf<f3> vx_p, vx_ray;
    DevBuf<int32_t> vx_ti;
    DevBuf<uint32_t> vx_vis, vx_draws, vx_off, pre, pathd, tdraws;
    DevBuf<uint32_t> vx_flag, vx_pos, vx_lit, vx_nlit;  // the drawing vertices, compacted
    DevBuf<c3> vx_rs;
    bool any_area = false, any_photon = false;
    int photon_kmax = 0;  // largest K of the launch's photon lights (heap entries per lane)
    // antialiasing passes: first-pass areas and colours, the pass's samples
    DevBuf<Area> areas0;
    DevBuf<c3> item_col, aa_col;
    DevBuf<uint32_t> aa_cnt, aa_off, aa_starts, hbase;
    DevBuf<uint2> aa_items;
    int aa_pass = -1;             // pass of the launch being enqueued (-1: first pass)
    bool hbase_on = false;        // per-area Halton counters carried across passes
    c3* finalize_to = nullptr;    // per-item colours instead of image pixels
    // irradiance cache (pathlight cache = on): samples in item slots, the fake
    // pass's area levels and candidate lists, the render pass's grid
    DevBuf<f3> lc_pP, lc_P, lc_N;
    DevBuf<float> lc_adist, lc_M, lc_prec;
    DevBuf<c3> lc_col, lc_pathc;
    DevBuf<int32_t> lc_rank, lc_row, lc_lvl, lc_adjs, lc_adj, lc_gstart;
    DevBuf<uint32_t> lc_in, lc_list, lc_cnt, lc_nb, lc_qual0, lc_gitem;
    DevBuf<unsigned long long> lc_err;
    // guessLight (area lights with psamples > 0): verdicts, test bits, rand() stream
    bool any_guess = false;
    bool ps_on = false;  // photonSampler_t plans and Halton tables before the levels
    int ps_nd = 0;
    int g_nw = 0;
    DevBuf<uint32_t> g_status, g_app, g_vis;
    DevBuf<int32_t> g_woff, g_rnd;
    uint64_t g_nrnd = 0;
    DevBuf<unsigned long long> g_err;
    UploadCache up_grnd, up_gwoff;
    std::vector<int32_t> g_host_rnd;
    uint32_t g_rnd_seed = 0;
    uint64_t g_rnd_skip = 0;
    DevBuf<f3> lv_P, lv_N, lv_U, lv_V, lv_pP;
    DevBuf<float> lv_Z;
    DevBuf<c3> lv_dif;
    std::vector<uint32_t> lc_ins, lc_pix;  // the last cache frame's samples (items, pixels) in insertion order
    uint64_t lc_uid = 0;                    // its scene
    // the single-scene render the workspace holds (for YG_RENDER_REUSE)
    uint64_t last_uid = 0;  // yg_scene::uid, 0 = none
    uint64_t last_gen = 0;
    yg_render_opts last_opts{};
    // descriptor uploads of the last frame
    UploadCache up_scenes, up_areas, up_lights, up_apts, up_ajh, up_ajv;
    Counters* h_cnt = nullptr;  // pinned read-back of the frame's counters
    // captured frame graph: replayed for frames of the same launch shape
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t frame_exec = nullptr;
    std::array<uint64_t, 13> frame_key{}, seen_key{};
    bool seen_valid = false;
    bool graphs_off = false;  // a capture failed: launch kernel by kernel
};

static bool same_opts(const yg_render_opts& a, const yg_render_opts& b)
{
    yg_render_opts x = a, y = b;
    x.flags &= YG_RENDER_ROULETTE;  // the other flags do not change the image
    y.flags &= YG_RENDER_ROULETTE;
    return memcmp(&x, &y, sizeof(x)) == 0;
}

static RenderWS g_ws;

// temporary bytes of an exclusive scan of n elements
static int scan_bytes(size_t n, size_t& bytes, hipStream_t st)
{
    bytes = 0;
    YG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, st));
    return YG_OK;
}

static int exclusive_scan(const uint32_t* in, uint32_t* out, size_t n, hipStream_t st)
{
    size_t bytes = 0;
    YG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, st));
    int rc = g_ws.scan_tmp.ensure(bytes);
    if (rc) return rc;
    YG_HIP(hipcub::DeviceScan::ExclusiveSum(g_ws.scan_tmp.p, bytes, in, out, (int)n, st));
    return YG_OK;
}


// Prepared host description of one scene for a render launch.
struct ScenePlan {
    yg_scene* s;
    RenderScene R;
    std::vector<Area> areas;
    std::vector<DevLight> lights;
    std::vector<f3> apts, ajh, ajv;
    uint32_t nitems = 0;
    int max_area_items = 0;
    int aa_passes = 0;
    float aa_threshold = 0;
    bool area_lcg = false;  // antialiasing with area lights: areas in turn
    bool guess = false;     // an area light with psamples > 0 (guessLight)
    uint32_t rand_seed = 1; // srand() seed: the block shuffle's and guessLight's rand() stream
    // irradiance cache of the pathlight (cache = on)
    bool cache = false;
    float cache_size = 0, shadow_threshold = 0;
    int search = 0;
};

// photonSampler_t buffers of the frame being rendered (one scene at a time)
static struct {
    DevBuf<float> h;
    DevBuf<uint32_t> flag;
    DevBuf<int32_t> hits;
    DevBuf<float> w, multi;
} g_ps;

static int plan_scene(yg_scene* s, const yg_render_opts* o, float* d_rgb, ScenePlan& P)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    if (!s->built) return yg_fail(YG_ERR_STATE, "scene not built");
    if (!s->has_cam) return yg_fail(YG_ERR_STATE, "no camera (yg_set_camera)");
    if (s->mats.empty()) return yg_fail(YG_ERR_STATE, "no materials (yg_set_shading)");
    if (!o || !d_rgb) return yg_fail(YG_ERR_ARG, "null argument");
    if (s->checked_geo != s->geo_gen || s->checked_nmat != s->mats.size()) {
        for (size_t t = 0; t < s->tri_mat.size(); t++)
            if (s->tri_mat[t] < 0 || (size_t)s->tri_mat[t] >= s->mats.size())
                return yg_fail(YG_ERR_ARG, "triangle material id out of range");
        s->checked_geo = s->geo_gen;
        s->checked_nmat = s->mats.size();
    }
    P.s = s;
    RenderScene& R = P.R;
    memset(&R, 0, sizeof(R));
    R.bvh = s->view();
    R.mats = s->d_mats;
    R.mods = s->d_mods;
    R.texs = s->d_texs;
    R.obj_flags = s->d_obj_flags;
    const HostCam hc = camera_setup(s->cam);
    R.cam_pos = hc.pos; R.vright = hc.vright; R.vup = hc.vup; R.vto = hc.vto;
    R.resx = s->cam.resx;
    R.resy = s->cam.resy;
    R.min_raydis = o->min_raydis;
    R.self_bias = o->self_bias;
    R.raydepth = o->raydepth;
    R.bg = o->has_background ? mkc(o->background[0], o->background[1], o->background[2]) : mkc(0, 0, 0);
    R.lcg_seed = o->lcg_seed;
    R.out_rgb = d_rgb;
    R.path_light = -1;
    R.photon_light = -1;
    R.gp_on = 0;
    R.ps_on = 0;
    R.ps_h = nullptr;
    R.ps_flag = nullptr;
    R.ps_hits = nullptr;
    R.ps_w = nullptr;
    R.ps_multi = nullptr;
    R.nl = (int)s->lights.size();
    for (int li = 0; li < R.nl; li++) {
        const yg_light& Lg = s->lights[li];
        DevLight L;
        memset(&L, 0, sizeof(L));
        L.type = Lg.type;
        L.use_in_render = Lg.use_in_render;
        L.use_in_indirect = Lg.use_in_indirect;
        L.cast_shadows = Lg.cast_shadows;
        L.power = Lg.power;
        L.color = mkc(Lg.color[0], Lg.color[1], Lg.color[2]);
        if (Lg.type == YG_LIGHT_SUN) {
            // sunLight_t ctor (sunlight.cc:27-33)
            L.dir = norm3(ld3(Lg.dir));
            L.col = cscale(mkc(Lg.color[0], Lg.color[1], Lg.color[2]), Lg.power);
        } else if (Lg.type == YG_LIGHT_AREA && Lg.dummy) {
            // a dummy area light only shoots global photons (arealight.cc:98):
            // no light, no draws
            L.type = 0;
        } else if (Lg.type == YG_LIGHT_AREA) {
            area_setup(Lg, L, P.apts, P.ajh, P.ajv);
            if (Lg.psamples < 0) return yg_fail(YG_ERR_ARG, "area light psamples must be >= 0");
            if (Lg.psamples > 0 && L.samples > 1 && Lg.use_in_render) {
                if (li >= 16) return yg_fail(YG_ERR_UNSUPPORTED, "area lights with psamples > 0 must be among the first 16 lights");
                if (Lg.use_in_indirect)
                    for (const auto& Lp : s->lights)
                        if (Lp.type == YG_LIGHT_PATH)
                            return yg_fail(YG_ERR_UNSUPPORTED,
                                           "area light psamples > 0 inside GI paths is not on the GPU path");
                P.guess = true;
            }
        } else if (Lg.type == YG_LIGHT_GLOBALPHOTON) {
            if (R.gp_on) return yg_fail(YG_ERR_UNSUPPORTED, "more than one global photon light");
            if (s->gp.light != li || s->gp.gen != s->geo_gen)
                return yg_fail(YG_ERR_STATE, "global photon map not built (yg_gphoton_build)");
            R.gp_on = 1;
            R.gp = s->gp.view();
            L.type = 0;
        } else if (Lg.type == YG_LIGHT_PATH) {
            if (R.path_light >= 0) return yg_fail(YG_ERR_UNSUPPORTED, "more than one pathlight");
            R.path_light = li;
            R.samples = Lg.samples;
            R.depth = Lg.depth;
            R.path_power = Lg.power;
            R.threshold = Lg.threshold;
            if (Lg.cache) {
                // pathLight_t with cache (pathlight.cc:405-723): the GPU path
                // renders it without refinement passes and antialiasing
                if (Lg.threshold < 1.0f)
                    return yg_fail(YG_ERR_UNSUPPORTED,
                                   "pathlight cache with threshold < 1 (refinement passes) is not on the GPU path");
                if (!(Lg.cache_size > 0) || !(Lg.shadow_threshold > 0))
                    return yg_fail(YG_ERR_ARG, "pathlight cache_size and shadow_threshold must be > 0");
                if (Lg.search < 3 || Lg.search > LC_KMAX)
                    return yg_fail(YG_ERR_UNSUPPORTED, "pathlight cache search must be 3..1000");
                // the fake pass forms the screen position in double
                // (scene.cc:559-583), the render pass in float: a "window"
                // texture seen by both passes is not on the GPU path
                for (const yg_modulator& md : s->mods)
                    if (md.texco == YG_TEXCO_WIN)
                        return yg_fail(YG_ERR_UNSUPPORTED,
                                       "pathlight cache with a texco window modulator is not on the GPU path");
                P.cache = true;
                P.cache_size = Lg.cache_size;
                P.shadow_threshold = Lg.shadow_threshold;
                P.search = Lg.search;
            }
        } else if (Lg.type == YG_LIGHT_PHOTON) {
            if (R.photon_light >= 0) return yg_fail(YG_ERR_UNSUPPORTED, "more than one photonlight");
            if (s->pm.light != li) return yg_fail(YG_ERR_STATE, "photon map not built (yg_photon_build)");
            if (Lg.search < 1 || Lg.search > YG_KNN_MAX) return yg_fail(YG_ERR_UNSUPPORTED, "photon search must be 1..128");
            R.photon_light = li;
            L.K = Lg.search;
            L.fixedradius = Lg.fixedradius;
        }
        P.lights.push_back(L);
    }
    if (R.gp_on) {
        for (const auto& Lg : s->lights)
            if (Lg.type == YG_LIGHT_AREA && !Lg.dummy && Lg.use_in_indirect)
                return yg_fail(YG_ERR_UNSUPPORTED,
                               "area lights inside GI paths together with a global photon map are not on the GPU path");
        // with the global map published, a pathlight of more than 96 samples
        // samples by photonSampler_t (pathlight.cc:731): its Halton starts are
        // ourRandomI() draws in area order, so no other light may draw
        if (R.path_light >= 0 && R.samples > 96) {
            if (o->aa_passes > 0)
                return yg_fail(YG_ERR_UNSUPPORTED, "photonSampler_t (pathlight samples > 96) with antialiasing passes");
            for (const auto& Lg : s->lights)
                if (Lg.type == YG_LIGHT_AREA && !Lg.dummy && Lg.use_in_render)
                    return yg_fail(YG_ERR_UNSUPPORTED,
                                   "photonSampler_t (pathlight samples > 96) with area lights drawing from ourRandom()");
            R.ps_on = 1;
        }
    }
    if (P.cache) {
        if (o->aa_passes > 0)
            return yg_fail(YG_ERR_UNSUPPORTED, "pathlight cache with antialiasing passes is not on the GPU path");
        if (R.gp_on)
            return yg_fail(YG_ERR_UNSUPPORTED, "pathlight cache with a global photon map is not on the GPU path");
        for (const auto& Lg : s->lights)
            if (Lg.type == YG_LIGHT_AREA && !Lg.dummy && Lg.use_in_render)
                return yg_fail(YG_ERR_UNSUPPORTED,
                               "pathlight cache with area lights (their fake-pass draws) is not on the GPU path");
    }
    if (R.path_light >= 0) {
        if (R.depth > 255) return yg_fail(YG_ERR_UNSUPPORTED, "pathlight depth");
        if (R.nl > 32) return yg_fail(YG_ERR_UNSUPPORTED, "more than 32 lights with a pathlight");
    }
    if (o->aa_passes < 0) return yg_fail(YG_ERR_ARG, "aa_passes must be >= 0");
    if (P.guess && o->aa_passes > 0)
        return yg_fail(YG_ERR_UNSUPPORTED, "area light psamples > 0 with antialiasing passes is not on the GPU path");
    if (o->aa_passes > 0) {
        if (o->aa_minsamples < 1) return yg_fail(YG_ERR_ARG, "aa_minsamples must be >= 1");
        if (!(o->aa_pixelwidth >= 1.0f && o->aa_pixelwidth <= 2.0f))
            return yg_fail(YG_ERR_ARG, "aa_pixelwidth must be in [1, 2]");
        for (const auto& L : P.lights)
            if (L.type == YG_LIGHT_AREA) P.area_lcg = true;
    }
    R.aa_pass = -1;
    R.aa_mins = o->aa_passes > 0 ? o->aa_minsamples : 0;
    R.aa_jitter = (o->aa_jitterfirst && o->aa_passes != 0) ? 1 : 0;
    R.aa_pw = o->aa_pixelwidth;
    float tsd = (float)(o->aa_minsamples * o->aa_passes);
    if (tsd != 0) tsd = (float)(1.0 / (double)tsd);
    R.aa_totsamdiv = tsd;
    R.tm = o->tonemap ? 1 : 0;
    R.tm_clamp = o->clamp_rgb ? 1 : 0;
    R.tm_e = -o->exposure;
    R.tm_g = (o->gamma != 0.0f) ? (float)(1.0 / (double)o->gamma) : 0.0f;
    R.rr_from = (o->flags & YG_RENDER_ROULETTE) ? 2 : 0;
    R.rr_seed = o->lcg_seed;
    P.aa_passes = o->aa_passes;
    P.aa_threshold = o->aa_threshold;
    block_regions(R.resx, R.resy, o->block > 0 ? o->block : 64, o->rand_seed, P.areas);
    P.rand_seed = o->rand_seed;
    uint32_t acc = 0;
    for (auto& a : P.areas) {
        a.item_start = a.src_start = acc;
        acc += (uint32_t)(a.w * a.h);
        P.max_area_items = std::max(P.max_area_items, a.w * a.h);
    }
    P.nitems = acc;
    if (R.path_light >= 0) {
        const int ndim = 2 * (R.depth + 1);
        // an area's draws per level: its first pass, then up to AA_minsamples
        // per pixel in each antialiasing pass
        const int64_t hs = (int64_t)P.max_area_items * R.samples * (1 + (int64_t)P.aa_passes * std::max(R.aa_mins, 0)) + 1;
        if (hs > INT32_MAX) return yg_fail(YG_ERR_UNSUPPORTED, "too many sampler draws per render area");
        R.hstride = (int)hs;
        if (!s->d_halton || s->halton_ndim != ndim || s->halton_stride != R.hstride) {
            std::vector<float> all((size_t)ndim * R.hstride), tab;
            int base = 2;
            for (int d = 0; d < ndim; d++) {
                halton_incremental(base, (uint64_t)R.hstride, tab);
                memcpy(&all[(size_t)d * R.hstride], tab.data(), sizeof(float) * R.hstride);
                base = next_prime(base);
            }
            s->halton_ndim = 0;
            int rc = dev_upload(&s->d_halton, all.data(), all.size());
            if (rc) return rc;
            s->halton_ndim = ndim;
            s->halton_stride = R.hstride;
        }
        R.halton = s->d_halton;
        if (R.ps_on) {
            // photonSampler_t ctor, grid 36 (pathtools.cc:100-106)
            const float base = (float)sqrt((double)(float)36 / 2.0);
            R.ps_par = (int)(base + 0.5);
            R.ps_mer = 2 * R.ps_par;
            R.ps_nd = ndim;
            R.ps_pdiv = (float)(1.0 / (double)(float)R.ps_par);
            R.ps_mdiv = (float)(2.0 * M_PI / (double)(float)R.ps_mer);
            const int nsec = R.ps_par * R.ps_mer;
            if (R.samples < nsec) return yg_fail(YG_ERR_UNSUPPORTED, "photonSampler_t needs samples >= sectors");
            const uint64_t hs = (uint64_t)P.max_area_items * (uint64_t)R.samples;
            if (hs > UINT32_MAX) return yg_fail(YG_ERR_UNSUPPORTED, "too many sampler draws per render area");
            R.ps_hstride = (uint32_t)hs;
            int rc;
            if ((rc = g_ps.flag.ensure(P.areas.size())) || (rc = g_ps.h.ensure(P.areas.size() * (size_t)ndim * hs)) ||
                (rc = g_ps.hits.ensure((size_t)P.nitems * nsec)) || (rc = g_ps.w.ensure((size_t)P.nitems * nsec)) ||
                (rc = g_ps.multi.ensure(P.nitems)))
                return rc;
            R.ps_h = g_ps.h.p;
            R.ps_flag = g_ps.flag.p;
            R.ps_hits = g_ps.hits.p;
            R.ps_w = g_ps.w.p;
            R.ps_multi = g_ps.multi.p;
        }
    }
    if (R.photon_light >= 0) {
        R.pm.photons = s->pm.d_photons;
        R.pm.gstart = s->pm.d_gstart;
        R.pm.gitems = s->pm.d_gitems;
        for (int k = 0; k < 3; k++) { R.pm.gmin[k] = s->pm.gmin[k]; R.pm.gdim[k] = s->pm.gdim[k]; }
        R.pm.gcell = s->pm.gcell;
        R.pm.np_emit = s->pm.np_emit;
        R.pm.n = s->pm.n;
    }
    return YG_OK;
}

__global__ __launch_bounds__(64) void k_photon_direct(const RenderScene* __restrict__ scenes,
                                                      const Area* __restrict__ areas, int nareas,
                                                      uint32_t nitems, const int32_t* __restrict__ item_ti,
                                                      const float* __restrict__ item_t,
                                                      c3* __restrict__ lightc, int nl_max);

// The GI levels of the launch the workspace holds: per level the extension
// rays, the shadow rays of their hits, the shading, then the stable
// compaction that numbers the next level's Halton calls per render area.
// Level lengths stay on the device (W.nact): every launch covers all paths
// and its threads beyond the level's length return at once, so the host
// never waits inside a frame.
// ADJ replays the forward (same Halton indices, same hits) and accumulates
// gradients in the shading kernel; the forward's contri/nseg stay untouched.
template <bool ADJ>
static int trace_levels(hipStream_t st, bool ktiming, std::vector<hipEvent_t>& kev, const AdjAcc& G,
                        const c3* w_item, unsigned long long* degenerate)
{
    RenderWS& W = g_ws;
    const int S = W.S, D = W.D, na = W.na;
    const uint32_t npaths = W.npaths, nitems = W.nitems;
    int rc;
    if (W.ps_on) {
        hipLaunchKernelGGL(k_ps_flag, dim3(na), dim3(64), 0, st, W.scenes.p, W.areas.p, na, W.qual.p);
        YG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ps_halton, dim3(grid_for((size_t)na * W.ps_nd, 64)), dim3(64), 0, st, W.scenes.p,
                           W.areas.p, na, W.qual.p, W.qoff.p, S);
        YG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ps_plan, dim3(nitems), dim3(64), 0, st, W.scenes.p, W.areas.p, na, W.item_ti.p,
                           W.item_t.p, W.qual.p);
        YG_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_path_init, dim3(grid_for(nitems, 256)), dim3(256), 0, st, nitems, S, D, W.qual.p,
                       W.qoff.p, W.active0.p, W.flag.p, W.nact.p, W.qcnt.p, W.dense.p);
    YG_HIP(hipGetLastError());
    uint32_t* act = W.active0.p;
    uint32_t* nxt = W.active1.p;
    const unsigned grid = grid_for(npaths, RB * (ADJ ? 1 : W.rays_lane));
    for (int j = 0; j < D; j++) {
        hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
        if (ktiming) {
            for (auto& e : ev) {
                YG_HIP(hipEventCreate(&e));
                kev.push_back(e);
            }
            YG_HIP(hipEventRecord(ev[0], st));
        }
        uint32_t* qc = W.qcnt.p + j;
        const int rl = ADJ ? 1 : W.rays_lane;
#define YG_EXT(RL)                                                                                          \
    YG_LAUNCH_WS(W.width, W.nocast, (k_ext<ADJ, 8, RL>), (k_ext<ADJ, 4, RL>), (k_ext<ADJ, 2, RL>),           \
                 (k_ext<ADJ, 24, RL>), (k_ext<ADJ, 20, RL>), (k_ext<ADJ, 18, RL>), dim3(grid), dim3(RB),      \
                 W.stk_bytes_lvl, st, W.scenes.p, W.areas.p, na, j, act, W.nact.p + j, npaths, S, D,        \
                 W.item_ti.p, W.item_t.p, W.qoff.p, W.hidx.p, W.st_where.p, W.st_hn.p, W.st_ti.p, W.st_col.p, \
                 W.st_s.p, W.contri.p, W.nseg.p, W.h_p.p, W.h_ray.p, W.h_ti.p, qc, W.queue.p, W.flag.p,       \
                 W.hbase_on ? W.hbase.p : nullptr, W.cnt.p, W.lds_cap, W.stk_ovf.p, W.dense.p + j)
#define YG_SHADOW(RL)                                                                                       \
    YG_LAUNCH_WS(W.width, W.nocast, (k_shadow<ADJ, 8, RL>), (k_shadow<ADJ, 4, RL>), (k_shadow<ADJ, 2, RL>),  \
                 (k_shadow<ADJ, 24, RL>), (k_shadow<ADJ, 20, RL>), (k_shadow<ADJ, 18, RL>), dim3(grid),      \
                 dim3(RB), W.stk_bytes_lvl, st, W.scenes.p, W.areas.p, na, S, qc, W.queue.p, W.h_p.p,        \
                 W.h_ti.p, W.vis.p, W.cnt.p, W.lds_cap, W.stk_ovf.p)
        if constexpr (ADJ) YG_EXT(1);
        else if (rl == 4) YG_EXT(4);
        else if (rl == 2) YG_EXT(2);
        else YG_EXT(1);
        YG_HIP(hipGetLastError());
        if (ktiming) YG_HIP(hipEventRecord(ev[1], st));
        if constexpr (ADJ) YG_SHADOW(1);
        else if (rl == 4) YG_SHADOW(4);
        else if (rl == 2) YG_SHADOW(2);
        else YG_SHADOW(1);
#undef YG_EXT
#undef YG_SHADOW
        YG_HIP(hipGetLastError());
        if (ktiming) YG_HIP(hipEventRecord(ev[2], st));
        const unsigned sgrid = ADJ ? std::min(grid_for(npaths, SB), ADJ_GRID) : grid_for(npaths, SB);
        hipLaunchKernelGGL(k_shade<ADJ>, dim3(sgrid), dim3(SB), ADJ ? sizeof(AdjLds) : 0, st,
                           W.scenes.p, W.areas.p, na, j, S, D, qc, W.queue.p, W.h_p.p, W.h_ray.p, W.h_ti.p,
                           W.vis.p, W.st_where.p, W.st_hn.p, W.st_ti.p, W.st_col.p, W.st_s.p, W.contri.p,
                           W.nseg.p, W.flag.p, G, w_item, degenerate,
                           (!ADJ && W.area_path) ? VertexRec{W.vx_p.p, W.vx_ray.p, W.vx_ti.p, W.vx_vis.p, W.vx_rs.p,
                                                             W.vx_draws.p}
                                                 : VertexRec{});
        YG_HIP(hipGetLastError());
        if (ktiming) YG_HIP(hipEventRecord(ev[3], st));
        if (j + 1 >= D) break;
        if ((rc = exclusive_scan(W.flag.p, W.ex.p, npaths, st))) return rc;
        hipLaunchKernelGGL(k_compact, dim3(grid_for(npaths, 256)), dim3(256), 0, st, W.areas.p, na, npaths, S,
                           W.flag.p, W.ex.p, act, W.nact.p + j, nxt, W.hidx.p + (size_t)(j + 1) * npaths,
                           W.nact.p + j + 1, W.dense.p + j + 1, W.compact_below,
                           W.hbase_on ? W.hbase.p + j + 1 : nullptr, D);
        YG_HIP(hipGetLastError());
        if (W.hbase_on) {
            hipLaunchKernelGGL(k_area_draws, dim3(grid_for(na, 256)), dim3(256), 0, st, W.areas.p, na, nitems, S, D,
                               j + 1, W.flag.p, W.ex.p, W.hbase.p);
            YG_HIP(hipGetLastError());
        }
        std::swap(act, nxt);
    }
    if (W.hbase_on) {
        hipLaunchKernelGGL(k_area_draws, dim3(grid_for(na, 256)), dim3(256), 0, st, W.areas.p, na, nitems, S, D, 0,
                           W.qual.p, W.qoff.p, W.hbase.p);
        YG_HIP(hipGetLastError());
    }
    return YG_OK;
}

// Appends the bucket index of `areas` (ascending item_start, nitems items in
// all) behind them, as find_area expects it.
static void append_area_index(std::vector<Area>& areas, uint32_t nitems)
{
    const size_t na = areas.size();
    const uint32_t nb = (nitems >> YG_AREA_LUT_SHIFT) + 1;
    std::vector<uint32_t> lut(nb);
    size_t k = 0;
    for (uint32_t b = 0; b < nb; b++) {
        const uint64_t first = (uint64_t)b << YG_AREA_LUT_SHIFT;
        while (k + 1 < na && areas[k + 1].item_start <= first) k++;
        lut[b] = (uint32_t)k;
    }
    areas.resize(na + (sizeof(uint32_t) * nb + sizeof(Area) - 1) / sizeof(Area));
    memcpy(areas.data() + na, lut.data(), sizeof(uint32_t) * nb);
}

// Per-item and per-path buffers of a launch of nitems work items.
static int ensure_frame_buffers(uint32_t nitems, uint32_t npaths, int nl_max, bool any_path, int D, hipStream_t st)
{
    RenderWS& W = g_ws;
    int rc;
    if ((rc = W.item_ti.ensure(nitems)) || (rc = W.item_t.ensure(nitems)) || (rc = W.draws.ensure(nitems)) ||
        (rc = W.sun_vis.ensure(nitems)) ||
        (rc = W.lcg_off.ensure(nitems)) || (rc = W.qual.ensure(nitems)) || (rc = W.qoff.ensure(nitems)) ||
        (rc = W.lightc.ensure((size_t)nitems * nl_max)) || (rc = W.cnt.ensure(1)))
        return rc;
    if (any_path) {
        if ((rc = W.active0.ensure(npaths)) || (rc = W.active1.ensure(npaths)) || (rc = W.flag.ensure(npaths)) ||
            (rc = W.ex.ensure(npaths)) || (rc = W.hidx.ensure((size_t)npaths * D)) ||
            (rc = W.st_where.ensure(npaths)) || (rc = W.st_hn.ensure(npaths)) || (rc = W.st_ti.ensure(npaths)) ||
            (rc = W.st_col.ensure(npaths)) || (rc = W.st_s.ensure(npaths)) ||
            (rc = W.contri.ensure((size_t)npaths * D)) || (rc = W.nseg.ensure(npaths)) ||
            (rc = W.h_p.ensure(npaths)) || (rc = W.h_ray.ensure(npaths)) || (rc = W.h_ti.ensure(npaths)) ||
            (rc = W.vis.ensure(npaths)) || (rc = W.queue.ensure(npaths)) || (rc = W.qcnt.ensure(D)) ||
            (rc = W.nact.ensure(D)) || (rc = W.dense.ensure(D)))
            return rc;
    }
    // the frame launches keep W.stk_lds entries per lane in LDS and the rest
    // in global memory, one slot per thread of the largest of them
    W.stk_bytes_lvl = stack_lds_bytes(W.stk_entries, RB);
    W.lds_cap = 0;
    if (W.stk_lds > 0 && W.stk_entries > W.stk_lds) {
        const size_t threads = (size_t)grid_for(std::max<uint64_t>(nitems, any_path ? npaths : 0), RB) * RB;
        if ((rc = W.stk_ovf.ensure((size_t)(W.stk_entries - W.stk_lds) * threads))) return rc;
        W.stk_bytes_lvl = stack_lds_bytes(W.stk_lds, RB);
        W.lds_cap = W.stk_lds;
    }
    if ((rc = W.stk_ovf.ensure(1))) return rc;  // a valid pointer for launches that do not spill
    const size_t nv = W.area_path && any_path ? (size_t)npaths * D : 0;
    if (nv) {
        if ((rc = W.vx_p.ensure(nv)) || (rc = W.vx_ray.ensure(nv)) || (rc = W.vx_ti.ensure(nv)) ||
            (rc = W.vx_vis.ensure(nv)) || (rc = W.vx_draws.ensure(nv)) || (rc = W.vx_off.ensure(nv)) ||
            (rc = W.vx_rs.ensure(nv)) || (rc = W.pre.ensure(nitems)) || (rc = W.pathd.ensure(nitems)) ||
            (rc = W.tdraws.ensure(nitems)) || (rc = W.vx_flag.ensure(nv)) || (rc = W.vx_pos.ensure(nv)) ||
            (rc = W.vx_lit.ensure(nv)) || (rc = W.vx_nlit.ensure(1)))
            return rc;
        if (nv >= 0x7fffffffull) return yg_fail(YG_ERR_UNSUPPORTED, "too many GI vertices in one launch");
    }
    // hipcub scan storage for every scan of the launch
    size_t b1 = 0, b2 = 0, b3 = 0;
    if ((rc = scan_bytes(nitems, b1, st)) || (rc = scan_bytes(any_path ? npaths : 0, b2, st)) ||
        (rc = scan_bytes(nv, b3, st)) || (rc = W.scan_tmp.ensure(std::max(b1, std::max(b2, b3)))))
        return rc;
    return YG_OK;
}

// One frame of the launch the workspace holds, camera rays to pixel pass,
// enqueued on st.  Nothing in it waits for the host or reads host memory,
// so the same sequence can be captured as a graph.  before_final (if set) is
// recorded ahead of the pixel pass.
static int launch_frame(hipStream_t st, bool ktiming, std::vector<hipEvent_t>& kev, hipEvent_t before_final)
{
    RenderWS& W = g_ws;
    const uint32_t nitems = W.nitems;
    const int na = W.na, nl_max = W.nl_max;
    int rc;
    YG_LAUNCH_W(W.width, (k_primary<8>), (k_primary<4>), (k_primary<2>), dim3(grid_for(nitems, RB)), dim3(RB), W.stk_bytes_lvl, st,
                W.scenes.p, W.areas.p, na, nitems, W.item_ti.p, W.item_t.p, W.cnt.p, W.aa_pass < 0 ? 1 : 0,
                W.lds_cap, W.stk_ovf.p);
    YG_HIP(hipGetLastError());
    const bool area_path = W.area_path && W.any_path;
    if (W.any_guess) {
        YG_LAUNCH_WS(W.width, W.nocast, (k_guess_vis<8>), (k_guess_vis<4>), (k_guess_vis<2>), (k_guess_vis<24>),
                     (k_guess_vis<20>), (k_guess_vis<18>), dim3(grid_for(nitems, RB)), dim3(RB), W.stk_bytes_lvl, st,
                     W.scenes.p, W.areas.p, na, nitems, W.item_ti.p, W.item_t.p, W.g_woff.p, W.g_nw, W.g_vis.p,
                     W.g_app.p, W.lds_cap, W.stk_ovf.p);
        YG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_guess_scan, dim3(1), dim3(64), 0, st, W.scenes.p, nitems, W.g_app.p, W.g_woff.p, W.g_nw,
                           W.g_vis.p, W.g_rnd.p, W.g_nrnd, W.g_status.p, W.cnt.p, W.g_err.p);
        YG_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_prep, dim3(grid_for(nitems, RB)), dim3(RB), 0, st, W.scenes.p, W.areas.p, na,
                       nitems, W.item_ti.p, W.item_t.p, W.draws.p, W.qual.p, area_path ? W.pre.p : nullptr,
                       W.any_guess ? (const uint32_t*)W.g_status.p : nullptr);
    YG_HIP(hipGetLastError());
    if (area_path) {
        // the GI levels first: the items' ourRandom() positions depend on their paths' vertices
        if ((rc = exclusive_scan(W.qual.p, W.qoff.p, nitems, st))) return rc;
        const size_t nv = (size_t)W.npaths * W.D;
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nv, 256)), dim3(256), 0, st, W.vx_draws.p, nv, 0u);
        YG_HIP(hipGetLastError());
        if ((rc = trace_levels<false>(st, ktiming, kev, AdjAcc{}, nullptr, nullptr))) return rc;
        if ((rc = exclusive_scan(W.vx_draws.p, W.vx_off.p, nv, st))) return rc;
        hipLaunchKernelGGL(k_item_pathd, dim3(grid_for(nitems, 256)), dim3(256), 0, st, nitems, W.S, W.D,
                           W.vx_draws.p, W.vx_off.p, W.draws.p, W.pathd.p, W.tdraws.p);
        YG_HIP(hipGetLastError());
        if ((rc = exclusive_scan(W.tdraws.p, W.lcg_off.p, nitems, st))) return rc;
        YG_LAUNCH_WS(W.width, W.nocast, (k_direct<8, true>), (k_direct<4, true>), (k_direct<2, true>), (k_direct<24, true>), (k_direct<20, true>), (k_direct<18, true>), dim3(grid_for(nitems, RB)), dim3(RB), W.stk_bytes, st, W.scenes.p,
                     W.areas.p, na, nitems, W.item_ti.p, W.item_t.p, W.lcg_off.p, W.lightc.p, nl_max, W.cnt.p, 0,
                     W.stk_ovf.p, (const uint32_t*)nullptr, (const uint32_t*)W.pathd.p);
        YG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_vx_flag, dim3(grid_for(nv, 256)), dim3(256), 0, st, nv, W.vx_draws.p, W.vx_flag.p);
        YG_HIP(hipGetLastError());
        if ((rc = exclusive_scan(W.vx_flag.p, W.vx_pos.p, nv, st))) return rc;
        hipLaunchKernelGGL(k_vx_list, dim3(grid_for(nv, 256)), dim3(256), 0, st, nv, W.vx_flag.p, W.vx_pos.p,
                           W.vx_lit.p, W.vx_nlit.p);
        YG_HIP(hipGetLastError());
        const VertexRec vx{W.vx_p.p, W.vx_ray.p, W.vx_ti.p, W.vx_vis.p, W.vx_rs.p, W.vx_draws.p};
        YG_LAUNCH_WS(W.width, W.nocast, (k_vertex_light<8>), (k_vertex_light<4>), (k_vertex_light<2>), (k_vertex_light<24>), (k_vertex_light<20>), (k_vertex_light<18>), dim3(grid_for(nv, RB)), dim3(RB), W.stk_bytes, st, W.scenes.p, W.areas.p,
                     na, W.vx_lit.p, W.vx_nlit.p, W.S, W.D, vx, W.vx_off.p, W.lcg_off.p, W.pre.p, W.contri.p,
                     W.cnt.p, 0, W.stk_ovf.p);
        YG_HIP(hipGetLastError());
        if (W.any_photon) {
            hipLaunchKernelGGL(k_photon_direct, dim3(grid_for(nitems, 64)), dim3(64), knn_lds_bytes(W.photon_kmax), st, W.scenes.p,
                               W.areas.p, na, nitems, W.item_ti.p, W.item_t.p, W.lightc.p, nl_max);
            YG_HIP(hipGetLastError());
        }
        if (before_final) YG_HIP(hipEventRecord(before_final, st));
        hipLaunchKernelGGL(k_finalize, dim3(grid_for(nitems, RB)), dim3(RB), 0, st, W.scenes.p, W.areas.p, na,
                           nitems, W.S, W.D, W.item_ti.p, W.item_t.p, W.qual.p, W.lightc.p, nl_max, W.contri.p,
                           W.nseg.p, W.cnt.p, W.finalize_to, (const c3*)nullptr);
        YG_HIP(hipGetLastError());
        return YG_OK;
    }
    if (W.any_area && (rc = exclusive_scan(W.draws.p, W.lcg_off.p, nitems, st))) return rc;
    if (W.any_area || nl_max > 32) {
        // registers hold this kernel to 2 waves per SIMD, so LDS has room for
        // the whole stack
        YG_LAUNCH_WS(W.width, W.nocast, (k_direct<8, true>), (k_direct<4, true>), (k_direct<2, true>), (k_direct<24, true>), (k_direct<20, true>), (k_direct<18, true>), dim3(grid_for(nitems, RB)), dim3(RB),
                    W.stk_bytes, st, W.scenes.p, W.areas.p, na, nitems, W.item_ti.p, W.item_t.p, W.lcg_off.p,
                    W.lightc.p, nl_max, W.cnt.p, 0, W.stk_ovf.p, (const uint32_t*)nullptr, (const uint32_t*)nullptr,
                    W.any_guess ? (const uint32_t*)W.g_status.p : nullptr);
    } else {
        YG_LAUNCH_WS(W.width, W.nocast, (k_sun_vis<8>), (k_sun_vis<4>), (k_sun_vis<2>), (k_sun_vis<24>), (k_sun_vis<20>), (k_sun_vis<18>), dim3(grid_for(nitems, RB)), dim3(RB), W.stk_bytes_lvl,
                    st, W.scenes.p, W.areas.p, na, nitems, W.item_ti.p, W.item_t.p, W.sun_vis.p, W.cnt.p,
                    W.lds_cap, W.stk_ovf.p);
        YG_HIP(hipGetLastError());
        YG_LAUNCH_W(W.width, (k_direct<8, false>), (k_direct<4, false>), (k_direct<2, false>), dim3(grid_for(nitems, RB)), dim3(RB), 0,
                    st, W.scenes.p, W.areas.p, na, nitems, W.item_ti.p, W.item_t.p, W.lcg_off.p, W.lightc.p,
                    nl_max, W.cnt.p, 0, W.stk_ovf.p, (const uint32_t*)W.sun_vis.p);
    }
    YG_HIP(hipGetLastError());
    if (W.any_photon) {
        hipLaunchKernelGGL(k_photon_direct, dim3(grid_for(nitems, 64)), dim3(64), knn_lds_bytes(W.photon_kmax), st, W.scenes.p,
                           W.areas.p, na, nitems, W.item_ti.p, W.item_t.p, W.lightc.p, nl_max);
        YG_HIP(hipGetLastError());
    }
    if (W.any_path) {
        if ((rc = exclusive_scan(W.qual.p, W.qoff.p, nitems, st))) return rc;
        if ((rc = trace_levels<false>(st, ktiming, kev, AdjAcc{}, nullptr, nullptr))) return rc;
    }
    if (before_final) YG_HIP(hipEventRecord(before_final, st));
    hipLaunchKernelGGL(k_finalize, dim3(grid_for(nitems, RB)), dim3(RB), 0, st, W.scenes.p, W.areas.p, na,
                       nitems, W.S, W.D, W.item_ti.p, W.item_t.p, W.qual.p, W.lightc.p, nl_max, W.contri.p,
                       W.nseg.p, W.cnt.p, W.finalize_to, (const c3*)nullptr);
    YG_HIP(hipGetLastError());
    return YG_OK;
}

// ------------------------------------------------- irradiance cache (host) --

// hash3d_t::getBox (hash3d.h:224-234)
static void lc_getbox(float cell, f3 p, int32_t& nx, int32_t& ny, int32_t& nz)
{
    nx = (int32_t)(p.x / cell);
    ny = (int32_t)(p.y / cell);
    nz = (int32_t)(p.z / cell);
    if (p.x < 0.0f) nx--;
    if (p.y < 0.0f) ny--;
    if (p.z < 0.0f) nz--;
}

// Depth-first leaf order of buildGenericTree (bound.h:252-340) over v (leaf
// size 1): split at the mean of the longest axis of the bound, the points on
// the cut in both halves' bounds going to a third part; an axis whose split
// separates nothing is skipped.  Appends to out.
static void lc_tree_order(const std::vector<f3>& pp, std::vector<uint32_t>& v, bool sx, bool sy, bool sz,
                          std::vector<uint32_t>& out)
{
    const size_t n = v.size();
    if (n <= 1 || (sx && sy && sz)) {
        out.insert(out.end(), v.begin(), v.end());
        return;
    }
    float lo[3], hi[3];
    {
        const f3 p0 = pp[v[0]];
        lo[0] = hi[0] = p0.x; lo[1] = hi[1] = p0.y; lo[2] = hi[2] = p0.z;
        for (uint32_t id : v) {
            const float q[3] = {pp[id].x, pp[id].y, pp[id].z};
            for (int k = 0; k < 3; k++) {
                if (q[k] > hi[k]) hi[k] = q[k];
                if (q[k] < lo[k]) lo[k] = q[k];
            }
        }
    }
    const float lx = hi[0] - lo[0], ly = hi[1] - lo[1], lz = hi[2] - lo[2];
    int axis;
    if (((lx >= ly) || sy) && ((lx >= lz) || sz) && !sx) axis = 0;
    else if (((ly >= lx) || sx) && ((ly >= lz) || sz) && !sy) axis = 1;
    else axis = 2;
    auto coord = [&](uint32_t id) { return axis == 0 ? pp[id].x : (axis == 1 ? pp[id].y : pp[id].z); };
    float media = 0;
    for (uint32_t id : v) media += coord(id);
    media /= (float)n;
    auto inside = [&](uint32_t id, const float* a, const float* g) {
        const float q[3] = {pp[id].x, pp[id].y, pp[id].z};
        return q[0] >= a[0] && q[0] <= g[0] && q[1] >= a[1] && q[1] <= g[1] && q[2] >= a[2] && q[2] <= g[2];
    };
    float la[3] = {lo[0], lo[1], lo[2]}, lg[3] = {hi[0], hi[1], hi[2]};
    float ra[3] = {lo[0], lo[1], lo[2]}, rg[3] = {hi[0], hi[1], hi[2]};
    lg[axis] = media;
    ra[axis] = media;
    std::vector<uint32_t> vl, vr, vm;
    for (uint32_t id : v) {
        if (inside(id, la, lg)) {
            if (inside(id, ra, rg)) vm.push_back(id);
            else vl.push_back(id);
        } else
            vr.push_back(id);
    }
    if (vl.size() == n || vr.size() == n || vm.size() == n) {
        lc_tree_order(pp, v, sx || axis == 0, sy || axis == 1, sz || axis == 2, out);
        return;
    }
    std::vector<uint32_t>().swap(v);
    if (vl.empty()) {
        lc_tree_order(pp, vr, sx, sy, sz, out);
        lc_tree_order(pp, vm, sx, sy, sz, out);
    } else if (vr.empty()) {
        lc_tree_order(pp, vl, sx, sy, sz, out);
        lc_tree_order(pp, vm, sx, sy, sz, out);
    } else {
        lc_tree_order(pp, vl, sx, sy, sz, out);
        lc_tree_order(pp, vr, sx, sy, sz, out);
        if (!vm.empty()) lc_tree_order(pp, vm, sx, sy, sz, out);
    }
}

// One frame of a pathlight with the irradiance cache, single scene
// (scene_t::render with repeatFirst, scene.cc:559-603): camera rays; the
// fake pass, area level by area level (k_cache_fill); the samples'
// estimates through the GI levels; startUse's tree order and the candidate
// grid (host); the render pass: the other lights, the interpolation
// (k_cache_interp) and the pixel sums.
static int cache_frame(const ScenePlan& P, hipStream_t st, bool ktiming, std::vector<hipEvent_t>& kev,
                       hipEvent_t before_final)
{
    RenderWS& W = g_ws;
    const uint32_t nitems = W.nitems;
    const int na = W.na, nl_max = W.nl_max;
    const RenderScene& R = P.R;
    int rc;
    YG_LAUNCH_W(W.width, (k_primary<8>), (k_primary<4>), (k_primary<2>), dim3(grid_for(nitems, RB)), dim3(RB),
                W.stk_bytes_lvl, st, W.scenes.p, W.areas.p, na, nitems, W.item_ti.p, W.item_t.p, W.cnt.p, 1,
                W.lds_cap, W.stk_ovf.p);
    YG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_prep, dim3(grid_for(nitems, RB)), dim3(RB), 0, st, W.scenes.p, W.areas.p, na, nitems,
                       W.item_ti.p, W.item_t.p, W.draws.p, W.qual.p, (uint32_t*)nullptr);
    YG_HIP(hipGetLastError());
    if ((rc = W.lc_pP.ensure(nitems)) || (rc = W.lc_P.ensure(nitems)) || (rc = W.lc_N.ensure(nitems)) ||
        (rc = W.lc_adist.ensure(nitems)) || (rc = W.lc_M.ensure(nitems)) || (rc = W.lc_prec.ensure(nitems)) ||
        (rc = W.lc_col.ensure(nitems)) || (rc = W.lc_pathc.ensure(nitems)) || (rc = W.lc_rank.ensure(nitems)) ||
        (rc = W.lc_row.ensure(nitems)) || (rc = W.lc_in.ensure(nitems)) || (rc = W.lc_list.ensure(nitems)) ||
        (rc = W.lc_cnt.ensure((size_t)na)) || (rc = W.lc_nb.ensure(nitems)) || (rc = W.lc_qual0.ensure(nitems)) ||
        (rc = W.lc_err.ensure(2)) || (rc = W.lv_P.ensure(nitems)) || (rc = W.lv_N.ensure(nitems)) ||
        (rc = W.lv_U.ensure(nitems)) || (rc = W.lv_V.ensure(nitems)) || (rc = W.lv_pP.ensure(nitems)) ||
        (rc = W.lv_Z.ensure(nitems)) || (rc = W.lv_dif.ensure(nitems)))
        return rc;
    YG_HIP(hipMemsetAsync(W.lc_in.p, 0, sizeof(uint32_t) * nitems, st));
    YG_HIP(hipMemsetAsync(W.lc_cnt.p, 0, sizeof(uint32_t) * na, st));
    YG_HIP(hipMemsetAsync(W.lc_err.p, 0, sizeof(unsigned long long) * 2, st));
    CacheView cv;
    cv.pP = W.lc_pP.p; cv.P = W.lc_P.p; cv.N = W.lc_N.p; cv.adist = W.lc_adist.p; cv.M = W.lc_M.p;
    cv.prec = W.lc_prec.p; cv.col = W.lc_col.p; cv.rank = W.lc_rank.p; cv.in = W.lc_in.p; cv.list = W.lc_list.p;
    cv.cnt = W.lc_cnt.p; cv.row = W.lc_row.p; cv.nb = W.lc_nb.p; cv.err = W.lc_err.p;
    CacheParams cp{};
    cp.cs = P.cache_size;
    cp.ycorr = (float)(1.0 / (double)((float)R.resx / (float)R.resy));
    cp.worldres = (float)((1.0 / (double)(float)R.resx) / (double)P.s->cam.focal);
    const float desired = (float)(1.0 / (double)P.shadow_threshold);
    cp.wfill = desired * (float)(1.0 / (double)(0 + 1));
    cp.wlimit = (float)(0.8 * (double)desired);
    cp.sradius = (float)((double)P.cache_size * 2.0);
    cp.K = P.search;
    // pixels within cache_size of a polar position: one pixel is 2 / resx in
    // both screen coordinates (y is scaled by 1 / aspect)
    cp.reach = (int)std::ceil((double)P.cache_size * (double)R.resx * 0.5) + 1;
    // area levels of the fake pass: an area waits for the earlier areas
    // whose samples can reach its pixels
    const std::vector<Area>& A = P.areas;
    std::vector<int> level(na, 0);
    std::vector<int32_t> adjs(na + 1, 0), adj;
    for (int k = 0; k < na; k++) {
        for (int j = 0; j < k; j++) {
            const int gx = std::max(A[j].x - (A[k].x + A[k].w - 1), A[k].x - (A[j].x + A[j].w - 1));
            const int gy = std::max(A[j].y - (A[k].y + A[k].h - 1), A[k].y - (A[j].y + A[j].h - 1));
            if (gx <= cp.reach && gy <= cp.reach) {
                adj.push_back(j);
                level[k] = std::max(level[k], level[j] + 1);
            }
        }
        adjs[k + 1] = (int32_t)adj.size();
    }
    const int nlev = na ? *std::max_element(level.begin(), level.end()) + 1 : 0;
    std::vector<int32_t> lvl;
    std::vector<size_t> loff(nlev + 1, 0);
    for (int L = 0; L < nlev; L++) {
        for (int k = 0; k < na; k++)
            if (level[k] == L) lvl.push_back(k);
        loff[L + 1] = lvl.size();
    }
    if ((rc = W.lc_lvl.ensure(std::max<size_t>(lvl.size(), 1))) || (rc = W.lc_adjs.ensure(adjs.size())) ||
        (rc = W.lc_adj.ensure(std::max<size_t>(adj.size(), 1))))
        return rc;
    YG_HIP(hipMemcpyAsync(W.lc_lvl.p, lvl.data(), sizeof(int32_t) * lvl.size(), hipMemcpyHostToDevice, st));
    YG_HIP(hipMemcpyAsync(W.lc_adjs.p, adjs.data(), sizeof(int32_t) * adjs.size(), hipMemcpyHostToDevice, st));
    if (!adj.empty())
        YG_HIP(hipMemcpyAsync(W.lc_adj.p, adj.data(), sizeof(int32_t) * adj.size(), hipMemcpyHostToDevice, st));
    VisitView vv{W.lv_P.p, W.lv_N.p, W.lv_U.p, W.lv_V.p, W.lv_pP.p, W.lv_Z.p, W.lv_dif.p};
    hipLaunchKernelGGL(k_cache_prep, dim3(grid_for(nitems, RB)), dim3(RB), 0, st, W.scenes.p, W.areas.p, na, nitems,
                       W.item_ti.p, W.item_t.p, W.qual.p, vv, cp);
    YG_HIP(hipGetLastError());
    const size_t fill_lds = stack_lds_bytes(W.stk_entries, LC_WAVE);
    for (int L = 0; L < nlev; L++) {
        const unsigned nb = (unsigned)(loff[L + 1] - loff[L]);
        YG_LAUNCH_WS(W.width, W.nocast, (k_cache_fill<8>), (k_cache_fill<4>), (k_cache_fill<2>), (k_cache_fill<24>),
                     (k_cache_fill<20>), (k_cache_fill<18>), dim3(nb), dim3(LC_WAVE), fill_lds, st, W.scenes.p,
                     W.areas.p, na, W.lc_lvl.p + loff[L], W.lc_adjs.p, W.lc_adj.p, W.qual.p, vv, cv, cp);
        YG_HIP(hipGetLastError());
    }
    // the samples' estimates: takeSample over the sample items, through the GI levels
    YG_HIP(hipMemcpyAsync(W.lc_qual0.p, W.qual.p, sizeof(uint32_t) * nitems, hipMemcpyDeviceToDevice, st));
    YG_HIP(hipMemcpyAsync(W.qual.p, W.lc_in.p, sizeof(uint32_t) * nitems, hipMemcpyDeviceToDevice, st));
    if ((rc = exclusive_scan(W.qual.p, W.qoff.p, nitems, st))) return rc;
    if ((rc = trace_levels<false>(st, ktiming, kev, AdjAcc{}, nullptr, nullptr))) return rc;
    hipLaunchKernelGGL(k_cache_color, dim3(grid_for(nitems, RB)), dim3(RB), 0, st, W.scenes.p, W.areas.p, na, nitems,
                       W.S, W.D, W.contri.p, W.nseg.p, cv, W.cnt.p);
    YG_HIP(hipGetLastError());
    YG_HIP(hipMemcpyAsync(W.qual.p, W.lc_qual0.p, sizeof(uint32_t) * nitems, hipMemcpyDeviceToDevice, st));
    // startUse (lightcache.cc:48-57): the samples in the cache's iteration
    // order -- boxes by (x, y, z) as std::map orders them, each box's list
    // newest first -- give the generic tree, whose leaf order ranks them
    std::vector<uint32_t> cnt(na), list(nitems);
    std::vector<f3> pP(nitems);
    unsigned long long err[2];
    YG_HIP(hipMemcpyAsync(cnt.data(), W.lc_cnt.p, sizeof(uint32_t) * na, hipMemcpyDeviceToHost, st));
    YG_HIP(hipMemcpyAsync(list.data(), W.lc_list.p, sizeof(uint32_t) * nitems, hipMemcpyDeviceToHost, st));
    YG_HIP(hipMemcpyAsync(pP.data(), W.lc_pP.p, sizeof(f3) * nitems, hipMemcpyDeviceToHost, st));
    YG_HIP(hipMemcpyAsync(err, W.lc_err.p, sizeof(err), hipMemcpyDeviceToHost, st));
    YG_HIP(hipStreamSynchronize(st));
    if (err[0]) return yg_fail(YG_ERR_UNSUPPORTED, "pathlight cache: more samples within reach of a render area than it has pixels");
    std::vector<uint32_t> ins;  // insertion order: areas in processing order, each area's visits in order
    for (int k = 0; k < na; k++)
        for (uint32_t m = 0; m < cnt[k]; m++) ins.push_back(list[A[k].item_start + m]);
    struct BoxKey {
        int32_t x, y, z;
        uint32_t seq;  // insertion position
    };
    std::vector<BoxKey> keys(ins.size());
    const float cs = P.cache_size;
    for (size_t q = 0; q < ins.size(); q++) {
        int32_t nx, ny, nz;
        lc_getbox(cs, pP[ins[q]], nx, ny, nz);
        // findBox looks the box up again from its centre (hash3d.h:41-42, 195-205)
        f3 c = mk3(cs * (float)nx, cs * (float)ny, cs * (float)nz);
        c.x = (float)((double)c.x + 0.5 * (double)cs);
        c.y = (float)((double)c.y + 0.5 * (double)cs);
        c.z = (float)((double)c.z + 0.5 * (double)cs);
        lc_getbox(cs, c, nx, ny, nz);
        keys[q] = BoxKey{nx, ny, nz, (uint32_t)q};
    }
    std::sort(keys.begin(), keys.end(), [](const BoxKey& a, const BoxKey& b) {
        if (a.x != b.x) return a.x < b.x;
        if (a.y != b.y) return a.y < b.y;
        if (a.z != b.z) return a.z < b.z;
        return a.seq > b.seq;  // push_front: newest first
    });
    std::vector<uint32_t> v(keys.size()), order;
    for (size_t q = 0; q < keys.size(); q++) v[q] = ins[keys[q].seq];
    order.reserve(v.size());
    lc_tree_order(pP, v, false, false, false, order);
    std::vector<int32_t> rank(nitems, 0);
    for (size_t q = 0; q < order.size(); q++) rank[order[q]] = (int32_t)q;
    // candidate grid over (pP.x, pP.y): cells a little over the search
    // radius, entries in rank order
    float lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (uint32_t id : ins) {
        lo[0] = std::min(lo[0], pP[id].x); hi[0] = std::max(hi[0], pP[id].x);
        lo[1] = std::min(lo[1], pP[id].y); hi[1] = std::max(hi[1], pP[id].y);
    }
    if (ins.empty()) lo[0] = lo[1] = hi[0] = hi[1] = 0;
    cp.gcell = cp.sradius * (1.0f + 1.0f / 1024.0f);
    cp.gx0 = lo[0];
    cp.gy0 = lo[1];
    cp.gnx = std::min(4096, (int)((hi[0] - lo[0]) / cp.gcell) + 1);
    cp.gny = std::min(4096, (int)((hi[1] - lo[1]) / cp.gcell) + 1);
    const size_t ncell = (size_t)cp.gnx * cp.gny;
    std::vector<int32_t> gstart(ncell + 1, 0);
    std::vector<uint32_t> gitem(order.size());
    auto cell_of = [&](uint32_t id) {
        int cx = (int)((pP[id].x - cp.gx0) / cp.gcell), cy = (int)((pP[id].y - cp.gy0) / cp.gcell);
        cx = std::min(std::max(cx, 0), cp.gnx - 1);
        cy = std::min(std::max(cy, 0), cp.gny - 1);
        return (size_t)cy * cp.gnx + cx;
    };
    for (uint32_t id : order) gstart[cell_of(id) + 1]++;
    for (size_t c = 0; c < ncell; c++) gstart[c + 1] += gstart[c];
    {
        std::vector<int32_t> fillp(gstart.begin(), gstart.end() - 1);
        for (uint32_t id : order) gitem[fillp[cell_of(id)]++] = id;  // order: ascending rank per cell
    }
    if ((rc = W.lc_gstart.ensure(gstart.size())) || (rc = W.lc_gitem.ensure(std::max<size_t>(gitem.size(), 1))))
        return rc;
    YG_HIP(hipMemcpyAsync(W.lc_gstart.p, gstart.data(), sizeof(int32_t) * gstart.size(), hipMemcpyHostToDevice, st));
    if (!gitem.empty())
        YG_HIP(hipMemcpyAsync(W.lc_gitem.p, gitem.data(), sizeof(uint32_t) * gitem.size(), hipMemcpyHostToDevice, st));
    YG_HIP(hipMemcpyAsync(W.lc_rank.p, rank.data(), sizeof(int32_t) * nitems, hipMemcpyHostToDevice, st));
    cp.gstart = W.lc_gstart.p;
    cp.gitem = W.lc_gitem.p;
    // render pass: the other lights at the camera hits, then the interpolation
    if (nl_max > 32) {
        YG_LAUNCH_WS(W.width, W.nocast, (k_direct<8, true>), (k_direct<4, true>), (k_direct<2, true>),
                     (k_direct<24, true>), (k_direct<20, true>), (k_direct<18, true>), dim3(grid_for(nitems, RB)),
                     dim3(RB), W.stk_bytes, st, W.scenes.p, W.areas.p, na, nitems, W.item_ti.p, W.item_t.p,
                     W.lcg_off.p, W.lightc.p, nl_max, W.cnt.p, 0, W.stk_ovf.p, (const uint32_t*)nullptr);
    } else {
        YG_LAUNCH_WS(W.width, W.nocast, (k_sun_vis<8>), (k_sun_vis<4>), (k_sun_vis<2>), (k_sun_vis<24>),
                     (k_sun_vis<20>), (k_sun_vis<18>), dim3(grid_for(nitems, RB)), dim3(RB), W.stk_bytes_lvl, st,
                     W.scenes.p, W.areas.p, na, nitems, W.item_ti.p, W.item_t.p, W.sun_vis.p, W.cnt.p, W.lds_cap,
                     W.stk_ovf.p);
        YG_HIP(hipGetLastError());
        YG_LAUNCH_W(W.width, (k_direct<8, false>), (k_direct<4, false>), (k_direct<2, false>),
                    dim3(grid_for(nitems, RB)), dim3(RB), 0, st, W.scenes.p, W.areas.p, na, nitems, W.item_ti.p,
                    W.item_t.p, W.lcg_off.p, W.lightc.p, nl_max, W.cnt.p, 0, W.stk_ovf.p,
                    (const uint32_t*)W.sun_vis.p);
    }
    YG_HIP(hipGetLastError());
    if (W.any_photon) {
        hipLaunchKernelGGL(k_photon_direct, dim3(grid_for(nitems, 64)), dim3(64), knn_lds_bytes(W.photon_kmax), st,
                           W.scenes.p, W.areas.p, na, nitems, W.item_ti.p, W.item_t.p, W.lightc.p, nl_max);
        YG_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_cache_interp, dim3(na), dim3(LC_WAVE), 0, st, W.scenes.p, W.areas.p, na, W.qual.p, vv,
                       cv, cp, W.lc_pathc.p);
    YG_HIP(hipGetLastError());
    if (before_final) YG_HIP(hipEventRecord(before_final, st));
    hipLaunchKernelGGL(k_finalize, dim3(grid_for(nitems, RB)), dim3(RB), 0, st, W.scenes.p, W.areas.p, na, nitems,
                       W.S, W.D, W.item_ti.p, W.item_t.p, W.qual.p, W.lightc.p, nl_max, W.contri.p, W.nseg.p, W.cnt.p,
                       W.finalize_to, (const c3*)W.lc_pathc.p);
    YG_HIP(hipGetLastError());
    YG_HIP(hipMemcpyAsync(err, W.lc_err.p, sizeof(err), hipMemcpyDeviceToHost, st));
    YG_HIP(hipStreamSynchronize(st));
    if (err[0]) return yg_fail(YG_ERR_UNSUPPORTED, "pathlight cache: more than 1024 samples within a gather radius");
    if (err[1]) return yg_fail(YG_ERR_STATE, "pathlight cache: a render-pass gather found no sample");
    W.lc_ins = ins;
    W.lc_pix.resize(ins.size());
    for (int k = 0, q = 0; k < na; k++)
        for (uint32_t m = 0; m < cnt[k]; m++, q++) {
            const uint32_t li = list[A[k].item_start + m] - A[k].item_start;
            const int i = A[k].y + (int)(li / (uint32_t)A[k].w), j = A[k].x + (int)(li % (uint32_t)A[k].w);
            W.lc_pix[q] = (uint32_t)(j + i * R.resx);
        }
    W.lc_uid = P.s->uid;
    return YG_OK;
}

// Captures launch_frame on a private stream into W.frame_exec for the launch
// shape `key`.  A capture that fails leaves graphs off for the process: the
// frames are then launched kernel by kernel, as with YG_RENDER_NO_GRAPH.
static int capture_frame(const std::array<uint64_t, 13>& key)
{
    RenderWS& W = g_ws;
    if (!W.cap_stream) YG_HIP(hipStreamCreateWithFlags(&W.cap_stream, hipStreamNonBlocking));
    const uint64_t serial = g_alloc_serial;
    YG_HIP(hipStreamBeginCapture(W.cap_stream, hipStreamCaptureModeThreadLocal));
    std::vector<hipEvent_t> none;
    const int rc = launch_frame(W.cap_stream, false, none, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(W.cap_stream, &g);
    hipGraphExec_t x = nullptr;
    hipError_t ie = hipErrorUnknown;
    // a buffer (re)allocated while capturing would leave stale pointers in the graph
    if (rc == YG_OK && ce == hipSuccess && g && serial == g_alloc_serial)
        ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        (void)hipGetLastError();
        W.graphs_off = true;
        return YG_OK;
    }
    if (W.frame_exec) (void)hipGraphExecDestroy(W.frame_exec);
    W.frame_exec = x;
    W.frame_key = key;
    return YG_OK;
}

// ourRandom()'s state after n more draws: seed * 16807^n mod (2^31 - 1)
// (vector3d.h:264-298), by the same power table the kernels jump with.
static uint32_t lcg_advance(uint32_t seed, uint64_t n)
{
    static uint64_t pw[64];
    static bool ready = false;
    if (!ready) {
        lcg_pow_table(pw);
        ready = true;
    }
    uint64_t acc = seed;
    for (int k = 0; k < 64 && (n >> k); k++)
        if ((n >> k) & 1) acc = (acc * pw[k]) % 2147483647ull;
    return (uint32_t)acc;
}

// Launch the whole pipeline over a set of planned scenes (one launch per stage).
static int render_plans(std::vector<ScenePlan>& plans, hipStream_t st, yg_render_stats* stats,
                        bool timing, bool ktiming, bool use_graph)
{
    int rc;
    g_ws.last_uid = 0;  // the workspace is about to hold this launch
    if ((rc = perm_upload())) return rc;
    if (!g_ws.pow_ready) {
        uint64_t pw[64];
        lcg_pow_table(pw);
        YG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(YG_LCG_POW), pw, sizeof(pw)));
        g_ws.pow_ready = true;
    }
    bool cache = false;
    for (auto& P : plans) cache = cache || P.cache;
    if (cache && plans.size() > 1)
        return yg_fail(YG_ERR_UNSUPPORTED, "a pathlight cache in a batched render is not on the GPU path");
    // concatenate scenes
    std::vector<RenderScene> scenes;
    std::vector<Area> areas;
    std::vector<DevLight> lights;
    std::vector<f3> apts, ajh, ajv;
    uint32_t nitems = 0;
    int S = 0, D = 0, nl_max = 1;
    bool any_path = false, any_photon = false, any_area = false;
    int photon_kmax = 1;
    for (auto& P : plans) {
        RenderScene R = P.R;
        R.item_base = nitems;
        R.area_base = (uint32_t)areas.size();
        for (auto L : P.lights) {
            if (L.type == YG_LIGHT_AREA) { L.pts += (int)apts.size(); any_area = true; }
            lights.push_back(L);
        }
        R.lights = (const DevLight*)(uintptr_t)(lights.size() - P.lights.size());  // offset, fixed below
        R.apts = nullptr;
        apts.insert(apts.end(), P.apts.begin(), P.apts.end());
        ajh.insert(ajh.end(), P.ajh.begin(), P.ajh.end());
        ajv.insert(ajv.end(), P.ajv.begin(), P.ajv.end());
        for (auto a : P.areas) {
            a.scene = (int)scenes.size();
            a.item_start += nitems;
            a.src_start += nitems;
            areas.push_back(a);
        }
        nitems += P.nitems;
        nl_max = std::max(nl_max, R.nl);
        if (R.path_light >= 0) {
            if (any_path && (R.samples != S || R.depth != D))
                return yg_fail(YG_ERR_UNSUPPORTED, "batched pathlights need equal samples and depth");
            any_path = true;
            S = R.samples;
            D = R.depth;
        }
        if (R.photon_light >= 0) {
            any_photon = true;
            photon_kmax = std::max(photon_kmax, P.lights[R.photon_light].K);
        }
        scenes.push_back(R);
    }
    if (!any_path) { S = 1; D = 1; }
    if (D == 0) D = 1;
    RenderWS& W = g_ws;
    if ((rc = W.lights.ensure(lights.size()))) return rc;
    if ((rc = W.apts.ensure(apts.size()))) return rc;
    if ((rc = W.ajh.ensure(ajh.size()))) return rc;
    if ((rc = W.ajv.ensure(ajv.size()))) return rc;
    for (auto& R : scenes) {
        R.lights = W.lights.p + (uintptr_t)R.lights;
        R.apts = W.apts.p;
        R.ajh = W.ajh.p;
        R.ajv = W.ajv.p;
    }
    const uint64_t npaths64 = (uint64_t)nitems * S;
    if (npaths64 >= 0xffffffffull) return yg_fail(YG_ERR_UNSUPPORTED, "too many path samples in one launch");
    const uint32_t npaths = (uint32_t)npaths64;
    const int na_real = (int)areas.size();
    int aa_passes = 0;
    float aa_threshold = 0;
    for (auto& P : plans)
        if (P.aa_passes > 0) {
            if (plans.size() > 1) return yg_fail(YG_ERR_UNSUPPORTED, "antialiasing passes in a batched render");
            if (P.area_lcg && P.areas.size() > 1)
                return yg_fail(YG_ERR_STATE, "antialiasing with area lights renders one area per launch");
            aa_passes = P.aa_passes;
            aa_threshold = P.aa_threshold;
        }
    int max_depth = 1;  // stack entries
    for (auto& P : plans) {
        max_depth = std::max(max_depth, P.s->stack_entries);
        if (P.s->width != plans[0].s->width)
            return yg_fail(YG_ERR_UNSUPPORTED, "batched scenes need BVHs of one width");
    }
    W.stk_entries = max_depth;
    W.stk_lds = 0;
    for (auto& P : plans) W.stk_lds = std::max(W.stk_lds, P.s->stack_lds);
    // launch settings of the GI levels: the first scene's (a batch shares them)
    W.rays_lane = plans[0].s->rays_per_lane;
    W.compact_below = plans[0].s->compact_below;
    W.nocast = false;
    for (auto& P : plans)
        for (int32_t f : P.s->obj_flags)
            if (!(f & 1)) W.nocast = true;
    W.area_path = false;
    for (auto& P : plans)
        if (P.R.path_light >= 0)
            for (const auto& L : P.s->lights)
                if (L.use_in_indirect && L.type == YG_LIGHT_AREA && !L.dummy) W.area_path = true;
    if (W.area_path && aa_passes > 0)
        return yg_fail(YG_ERR_UNSUPPORTED, "antialiasing with area lights inside GI paths");
    if (W.area_path) {
        // stream offsets are 32-bit scans: bound the launch's draws (every
        // primary hit and every path vertex in front of every area light)
        uint64_t prim = 0, vert = 0;
        for (auto& P : plans) {
            uint64_t pp = 0, pv = 0;
            for (const auto& L : P.s->lights)
                if (L.type == YG_LIGHT_AREA && !L.dummy) {
                    if (L.use_in_render) pp += 2ull * (uint64_t)L.samples;
                    if (L.use_in_indirect) pv += 2ull * (uint64_t)L.samples;
                }
            prim = std::max(prim, pp);
            vert = std::max(vert, pv);
        }
        if ((double)nitems * (double)prim + (double)npaths * (double)D * (double)vert >= 4294967295.0)
            return yg_fail(YG_ERR_UNSUPPORTED, "too many area-light draws in one launch with GI paths");
    } else {
        // the primary hits' ourRandom() offsets are 32-bit scans too: a scene's
        // draws in one pass (all its items, or an AA pass's samples) must stay
        // below 2^32 (differences between the scenes of a batch survive a wrap)
        for (auto& P : plans) {
            uint64_t pp = 0;
            for (const auto& L : P.s->lights)
                if (L.type == YG_LIGHT_AREA && !L.dummy && L.use_in_render) pp += 2ull * (uint64_t)L.samples;
            const double per_pass = (double)P.nitems * (double)std::max(1, P.aa_passes > 0 ? P.R.aa_mins : 1);
            if (per_pass * (double)pp >= 4294967295.0)
                return yg_fail(YG_ERR_UNSUPPORTED, "too many area-light draws in one scene's pass (2^32 or more)");
        }
    }
    const std::vector<Area> areas0(areas);  // the first pass's areas
    append_area_index(areas, nitems);
    if ((rc = W.scenes.ensure(scenes.size())) || (rc = W.areas.ensure(areas.size()))) return rc;
    if ((rc = ensure_frame_buffers(nitems, npaths, nl_max, any_path, D, st))) return rc;
    if ((rc = upload_cached(W.scenes, scenes, W.up_scenes, st)) || (rc = upload_cached(W.areas, areas, W.up_areas, st)) ||
        (rc = upload_cached(W.lights, lights, W.up_lights, st)) || (rc = upload_cached(W.apts, apts, W.up_apts, st)) ||
        (rc = upload_cached(W.ajh, ajh, W.up_ajh, st)) || (rc = upload_cached(W.ajv, ajv, W.up_ajv, st)))
        return rc;

    W.stk_bytes = stack_lds_bytes(max_depth, RB);
    W.width = plans[0].s->width;
    W.nitems = nitems;
    W.npaths = npaths;
    W.S = S;
    W.D = D;
    W.na = na_real;
    W.nl_max = nl_max;
    W.any_path = any_path;
    W.any_area = any_area;
    W.any_photon = any_photon;
    W.photon_kmax = photon_kmax;
    // guessLight: test bits per light, and glibc rand() after the block
    // shuffle's draws (one per area), psamples draws per light at most per item
    W.any_guess = false;
    for (auto& P : plans) W.any_guess = W.any_guess || P.guess;
    W.ps_on = plans.size() == 1 && plans[0].R.ps_on;
    W.ps_nd = plans[0].R.ps_nd;
    if (W.any_guess) {
        if (plans.size() > 1)
            return yg_fail(YG_ERR_UNSUPPORTED, "area light psamples > 0 in a batched render is not on the GPU path");
        std::vector<int32_t> woff(std::max<size_t>(plans[0].lights.size(), 1), 0);
        int nw = 0;
        uint64_t per = 0;
        for (size_t li = 0; li < plans[0].lights.size(); li++) {
            const DevLight& L = plans[0].lights[li];
            if (L.type == YG_LIGHT_AREA && L.use_in_render && L.fsamples > 0 && L.samples > 1) {
                woff[li] = nw;
                nw += 2 * ((L.samples + 31) / 32);
                per += (uint64_t)L.fsamples;
            }
        }
        W.g_nw = nw;
        const uint64_t nrnd = (uint64_t)nitems * per;
        if ((rc = W.g_status.ensure(nitems)) || (rc = W.g_app.ensure(nitems)) ||
            (rc = W.g_vis.ensure((size_t)nitems * nw)) || (rc = W.g_woff.ensure(woff.size())) ||
            (rc = W.g_rnd.ensure(nrnd)) || (rc = W.g_err.ensure(1)))
            return rc;
        const uint32_t seed = plans[0].rand_seed;
        if (W.g_host_rnd.size() != nrnd || W.g_rnd_seed != seed || W.g_rnd_skip != (uint64_t)na_real) {
            std::vector<int32_t> all;
            glibc_rand(seed, (int64_t)(na_real + nrnd), all);
            W.g_host_rnd.assign(all.begin() + na_real, all.end());
            W.g_rnd_seed = seed;
            W.g_rnd_skip = (uint64_t)na_real;
        }
        W.g_nrnd = nrnd;
        if ((rc = upload_cached(W.g_rnd, W.g_host_rnd, W.up_grnd, st)) || (rc = upload_cached(W.g_woff, woff, W.up_gwoff, st)))
            return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    std::vector<hipEvent_t> kev;
    if (ktiming) timing = true;
    if (timing) {
        YG_HIP(hipEventCreate(&e0));
        YG_HIP(hipEventCreate(&e1));
        YG_HIP(hipEventCreate(&e2));
        YG_HIP(hipEventRecord(e0, st));
    }
    // (every buffer of the frame was allocated above, before its shape is taken)
    // The launch shape: a graph captured for it replays any frame of the same
    // shape (scenes, areas, lights and output pointers are read from the
    // uploaded descriptors, not baked into the launches).
    const std::array<uint64_t, 13> key = {g_alloc_serial, nitems, npaths, (uint64_t)S, (uint64_t)D,
                                          (uint64_t)W.na, (uint64_t)nl_max,
                                          (uint64_t)W.width | ((uint64_t)W.rays_lane << 8) |
                                              ((uint64_t)(uint32_t)__builtin_bit_cast(uint32_t, W.compact_below) << 16),
                                          W.stk_bytes,
                                          W.scan_tmp.cap, any_area, any_photon,
                                          (uint64_t)any_path | ((uint64_t)W.nocast << 1) |
                                              ((uint64_t)W.area_path << 2) | ((uint64_t)W.any_guess << 3)};
    int graph = 0;
    W.aa_pass = -1;
    W.hbase_on = aa_passes > 0 && any_path;
    W.finalize_to = nullptr;
    if (aa_passes > 0) {
        if ((rc = W.areas0.ensure(areas.size())) || (rc = W.item_col.ensure(nitems)) ||
            (rc = W.aa_cnt.ensure(nitems)) || (rc = W.aa_off.ensure(nitems)) ||
            (rc = W.aa_starts.ensure((size_t)na_real + 1)) || (rc = W.hbase.ensure((size_t)na_real * D)))
            return rc;
        YG_HIP(hipMemcpyAsync(W.areas0.p, areas.data(), sizeof(Area) * areas.size(), hipMemcpyHostToDevice, st));
        YG_HIP(hipMemsetAsync(W.hbase.p, 0, sizeof(uint32_t) * na_real * D, st));
        W.finalize_to = W.item_col.p;
    }
    if (use_graph && !timing && !W.graphs_off && aa_passes == 0 && !cache) {
        // capture on the second consecutive frame of a shape, replay from then on
        bool captured = false;
        if (!(W.frame_exec && W.frame_key == key) && W.seen_valid && W.seen_key == key) {
            if ((rc = capture_frame(key))) return rc;
            captured = true;
        }
        if (W.frame_exec && W.frame_key == key) {
            graph = captured ? 2 : 1;
            YG_HIP(hipGraphLaunch(W.frame_exec, st));
        }
    }
    W.seen_key = key;
    W.seen_valid = true;
    if (!graph && (rc = cache ? cache_frame(plans[0], st, ktiming, kev, timing ? e1 : nullptr)
                              : launch_frame(st, ktiming, kev, timing ? e1 : nullptr)))
        return rc;
    uint64_t aa_rays = 0;
    if (aa_passes > 0) {
        // scene.cc:777-816: each pass resamples the pixels the contrast mask
        // marks (per area, from the area's colours) with AA_minsamples camera
        // rays each, through the whole pipeline again, then folds them in
        const uint32_t n0 = nitems;
        const uint32_t mins = (uint32_t)plans[0].R.aa_mins;
        std::vector<uint32_t> starts((size_t)na_real + 1);
        std::vector<hipEvent_t> none;
        for (int pass = 0; pass < aa_passes; pass++) {
            hipLaunchKernelGGL(k_aa_mask, dim3(grid_for(n0, 256)), dim3(256), 0, st, W.areas0.p, na_real, n0,
                               W.item_col.p, aa_threshold, mins, W.aa_cnt.p);
            YG_HIP(hipGetLastError());
            if ((rc = exclusive_scan(W.aa_cnt.p, W.aa_off.p, n0, st))) return rc;
            hipLaunchKernelGGL(k_aa_starts, dim3(grid_for((uint64_t)na_real + 1, 256)), dim3(256), 0, st, W.areas0.p,
                               na_real, n0, W.aa_cnt.p, W.aa_off.p, W.aa_starts.p);
            YG_HIP(hipGetLastError());
            YG_HIP(hipMemcpyAsync(starts.data(), W.aa_starts.p, sizeof(uint32_t) * starts.size(),
                                  hipMemcpyDeviceToHost, st));
            if (any_area) {
                if (!W.h_cnt) YG_HIP(hipHostMalloc((void**)&W.h_cnt, sizeof(Counters), hipHostMallocDefault));
                YG_HIP(hipMemcpyAsync(W.h_cnt, W.cnt.p, sizeof(Counters), hipMemcpyDeviceToHost, st));
            }
            YG_HIP(hipStreamSynchronize(st));
            const uint32_t total = starts[na_real];
            // (one area per launch with area lights: the pass continues the
            // area's ourRandom() stream after the draws taken so far)
            if (any_area) scenes[0].lcg_seed = lcg_advance(plans[0].R.lcg_seed, W.h_cnt->draws);
            if (total == 0) break;  // no colour changes, so no later pass resamples either
            const uint64_t np64 = (uint64_t)total * S;
            if (np64 >= 0xffffffffull) return yg_fail(YG_ERR_UNSUPPORTED, "too many path samples in one launch");
            std::vector<Area> aa_areas(areas0);
            for (int a = 0; a < na_real; a++) aa_areas[a].item_start = starts[a];
            append_area_index(aa_areas, total);
            if ((rc = W.areas.ensure(aa_areas.size())) || (rc = W.aa_items.ensure(total)) ||
                (rc = W.aa_col.ensure(total)) ||
                (rc = ensure_frame_buffers(total, (uint32_t)np64, nl_max, any_path, D, st)))
                return rc;
            scenes[0].aa_pass = pass;
            scenes[0].aa_items = W.aa_items.p;
            if ((rc = upload_cached(W.scenes, scenes, W.up_scenes, st)) ||
                (rc = upload_cached(W.areas, aa_areas, W.up_areas, st)))
                return rc;
            hipLaunchKernelGGL(k_aa_items, dim3(grid_for(n0, 256)), dim3(256), 0, st, n0, W.aa_cnt.p, W.aa_off.p,
                               W.aa_items.p);
            YG_HIP(hipGetLastError());
            W.nitems = total;
            W.npaths = (uint32_t)np64;
            W.aa_pass = pass;
            W.finalize_to = W.aa_col.p;
            if ((rc = launch_frame(st, false, none, nullptr))) return rc;
            hipLaunchKernelGGL(k_aa_combine, dim3(grid_for(n0, 256)), dim3(256), 0, st, n0, W.aa_cnt.p, W.aa_off.p,
                               W.aa_col.p, pass, W.item_col.p);
            YG_HIP(hipGetLastError());
            aa_rays += total;
        }
        hipLaunchKernelGGL(k_aa_output, dim3(grid_for(n0, 256)), dim3(256), 0, st, W.scenes.p, W.areas0.p, na_real,
                           n0, W.item_col.p);
        YG_HIP(hipGetLastError());
    }
    if (stats) {
        if (!W.h_cnt) YG_HIP(hipHostMalloc((void**)&W.h_cnt, sizeof(Counters), hipHostMallocDefault));
        YG_HIP(hipMemcpyAsync(W.h_cnt, W.cnt.p, sizeof(Counters), hipMemcpyDeviceToHost, st));
        if (timing) YG_HIP(hipEventRecord(e2, st));
        YG_HIP(hipStreamSynchronize(st));
        const Counters c = *W.h_cnt;
        memset(stats, 0, sizeof(*stats));
        stats->rays_camera = nitems + aa_rays;
        stats->rays_extension = c.ext;
        stats->rays_shadow = c.shadow;
        stats->adaptive_extra = c.extra;
        stats->lcg_draws = c.draws;
        stats->rays_bounce = c.ext + c.bshadow;
        stats->graph = graph;
        if (timing) {
            float ms = 0, ms2 = 0;
            YG_HIP(hipEventElapsedTime(&ms, e0, e2));
            YG_HIP(hipEventElapsedTime(&ms2, e0, e1));
            stats->ms = ms;
            stats->ms_trace = ms2;
        }
        for (size_t q = 0; q + 3 < kev.size(); q += 4) {
            float m_ext = 0, m_sh = 0, m_all = 0;
            YG_HIP(hipEventElapsedTime(&m_ext, kev[q], kev[q + 1]));
            YG_HIP(hipEventElapsedTime(&m_sh, kev[q + 1], kev[q + 2]));
            YG_HIP(hipEventElapsedTime(&m_all, kev[q], kev[q + 3]));
            stats->ms_bounce += m_all;
            stats->ms_ext += m_ext;
            stats->ms_shadow += m_sh;
            stats->bounce_launches++;
        }
    }
    for (auto e : kev) (void)hipEventDestroy(e);
    if (timing) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
    }
    return YG_OK;
}

// scene_t::render in mono mode renders the areas one after another.  With
// anti-aliasing and area lights, an area's passes take draws from the global
// ourRandom() stream before the next area starts, and which pixels the passes
// resample depends on the colours those draws give: each area is then its
// own launch (first pass and passes), its stream starting where the previous
// areas' draws left it.
static int render_areas_in_turn(const ScenePlan& P, hipStream_t st, yg_render_stats* stats, bool timing)
{
    yg_render_stats tot;
    memset(&tot, 0, sizeof(tot));
    uint64_t draws = 0;
    for (const Area& a : P.areas) {
        std::vector<ScenePlan> one(1, P);
        Area b = a;
        b.item_start = b.src_start = 0;
        one[0].areas.assign(1, b);
        one[0].nitems = (uint32_t)(a.w * a.h);
        one[0].R.lcg_seed = lcg_advance(P.R.lcg_seed, draws);
        yg_render_stats s;
        int rc = render_plans(one, st, &s, timing, false, false);
        if (rc) return rc;
        draws += s.lcg_draws;
        tot.rays_camera += s.rays_camera;
        tot.rays_extension += s.rays_extension;
        tot.rays_shadow += s.rays_shadow;
        tot.adaptive_extra += s.adaptive_extra;
        tot.rays_bounce += s.rays_bounce;
        tot.ms += s.ms;
        tot.ms_trace += s.ms_trace;
    }
    tot.lcg_draws = draws;
    if (stats) *stats = tot;
    return YG_OK;
}

extern "C" int yg_render(yg_scene* s, const yg_render_opts* opts, float* d_rgb, yg_render_stats* stats,
                         void* stream)
{
    std::vector<ScenePlan> plans(1);
    int rc = plan_scene(s, opts, d_rgb, plans[0]);
    if (rc) return rc;
    if (plans[0].aa_passes > 0 && plans[0].area_lcg) {
        g_ws.last_uid = 0;
        return render_areas_in_turn(plans[0], (hipStream_t)stream, stats, opts->flags & YG_RENDER_TIMING);
    }
    g_ws.last_uid = 0;
    rc = render_plans(plans, (hipStream_t)stream, stats, opts->flags & YG_RENDER_TIMING,
                      opts->flags & YG_RENDER_KTIMING, !(opts->flags & YG_RENDER_NO_GRAPH));
    if (rc == YG_OK) {
        g_ws.last_uid = s->uid;
        g_ws.last_gen = s->gen;
        g_ws.last_opts = *opts;
    }
    return rc;
}

// The samples of the scene's last cache frame, 18 floats each in insertion
// order (the fake pass's visit order): P, N, pP, colour, adist, M,
// precision, rank, pixel number, devaluated (1).
extern "C" int yg_cache_read(yg_scene* s, float* d_out, uint64_t capacity, uint64_t* count, void* stream)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    RenderWS& W = g_ws;
    if (W.lc_uid != s->uid) return yg_fail(YG_ERR_STATE, "no cache frame of this scene (yg_render with cache on)");
    const size_t n = W.lc_ins.size();
    if (count) *count = n;
    if (!d_out || !n) return YG_OK;
    hipStream_t st = (hipStream_t)stream;
    const size_t ni = W.nitems;
    std::vector<f3> P(ni), N(ni), pP(ni);
    std::vector<c3> col(ni);
    std::vector<float> ad(ni), M(ni), pr(ni);
    std::vector<int32_t> rk(ni);
    YG_HIP(hipMemcpyAsync(P.data(), W.lc_P.p, sizeof(f3) * ni, hipMemcpyDeviceToHost, st));
    YG_HIP(hipMemcpyAsync(N.data(), W.lc_N.p, sizeof(f3) * ni, hipMemcpyDeviceToHost, st));
    YG_HIP(hipMemcpyAsync(pP.data(), W.lc_pP.p, sizeof(f3) * ni, hipMemcpyDeviceToHost, st));
    YG_HIP(hipMemcpyAsync(col.data(), W.lc_col.p, sizeof(c3) * ni, hipMemcpyDeviceToHost, st));
    YG_HIP(hipMemcpyAsync(ad.data(), W.lc_adist.p, sizeof(float) * ni, hipMemcpyDeviceToHost, st));
    YG_HIP(hipMemcpyAsync(M.data(), W.lc_M.p, sizeof(float) * ni, hipMemcpyDeviceToHost, st));
    YG_HIP(hipMemcpyAsync(pr.data(), W.lc_prec.p, sizeof(float) * ni, hipMemcpyDeviceToHost, st));
    YG_HIP(hipMemcpyAsync(rk.data(), W.lc_rank.p, sizeof(int32_t) * ni, hipMemcpyDeviceToHost, st));
    YG_HIP(hipStreamSynchronize(st));
    const size_t m = std::min<size_t>(n, capacity);
    std::vector<float> out(18 * m);
    for (size_t q = 0; q < m; q++) {
        const uint32_t id = W.lc_ins[q];
        float* o = &out[18 * q];
        o[0] = P[id].x; o[1] = P[id].y; o[2] = P[id].z;
        o[3] = N[id].x; o[4] = N[id].y; o[5] = N[id].z;
        o[6] = pP[id].x; o[7] = pP[id].y; o[8] = pP[id].z;
        o[9] = col[id].r; o[10] = col[id].g; o[11] = col[id].b;
        o[12] = ad[id]; o[13] = M[id]; o[14] = pr[id];
        o[15] = (float)rk[id]; o[16] = (float)W.lc_pix[q]; o[17] = 1.0f;
    }
    YG_HIP(hipMemcpyAsync(d_out, out.data(), sizeof(float) * out.size(), hipMemcpyHostToDevice, st));
    YG_HIP(hipStreamSynchronize(st));
    return YG_OK;
}

// Batched render: several scenes in one set of launches (yg_render per scene
// otherwise).  d_rgb[k] receives scene k's image.
extern "C" int yg_render_batch(yg_scene** scenes, uint32_t n, const yg_render_opts* opts, float** d_rgb,
                               yg_render_stats* stats, void* stream)
{
    if (!scenes || !opts || !d_rgb || n == 0) return yg_fail(YG_ERR_ARG, "null argument");
    std::vector<ScenePlan> plans(n);
    for (uint32_t k = 0; k < n; k++) {
        int rc = plan_scene(scenes[k], &opts[k], d_rgb[k], plans[k]);
        if (rc) return rc;
        if (plans[k].R.ps_on && n > 1)
            return yg_fail(YG_ERR_UNSUPPORTED, "photonSampler_t (pathlight samples > 96) in a batch");
    }
    bool timing = false, ktiming = false, graph = true;
    for (uint32_t k = 0; k < n; k++) {
        timing = timing || (opts[k].flags & YG_RENDER_TIMING);
        ktiming = ktiming || (opts[k].flags & YG_RENDER_KTIMING);
        graph = graph && !(opts[k].flags & YG_RENDER_NO_GRAPH);
    }
    return render_plans(plans, (hipStream_t)stream, stats, timing, ktiming, graph);
}

// photonLight_t::illuminate at every primary hit (photonlight.cc:297-395)
__global__ __launch_bounds__(64) void k_photon_direct(const RenderScene* __restrict__ scenes,
                                                      const Area* __restrict__ areas, int nareas,
                                                      uint32_t nitems, const int32_t* __restrict__ item_ti,
                                                      const float* __restrict__ item_t,
                                                      c3* __restrict__ lightc, int nl_max)
{
    // K-entry heaps of the block's 64 lanes: distances, then indices (entry k
    // of lane t at [k * 64 + t]), sized by the launch's largest K
    extern __shared__ float s_knn[];
    const uint32_t item = blockIdx.x * 64 + threadIdx.x;
    if (item >= nitems) return;
    const int ti = item_ti[item];
    if (ti < 0) return;
    const Area a = areas[find_area(areas, nareas, item)];
    const RenderScene& R = scenes[a.scene];
    if (R.photon_light < 0) return;
    const DevLight& L = R.lights[R.photon_light];
    if (!L.use_in_render) return;
    f3 eye;
    const Surf sp = primary_surf(R, a, item, ti, item_t[item], eye);
    const bool receive = (R.obj_flags[sp.obj] & 2) != 0;
    lightc[(size_t)item * nl_max + R.photon_light] =
        photon_estimate(R, R.pm, L.K, L.fixedradius, sp, eye, receive, s_knn + threadIdx.x,
                        (int*)(s_knn + (size_t)L.K * 64) + threadIdx.x, 64);
}

// ------------------------------------------------------------ adjoint --
//
// Gradient of a scalar loss through the image, given dLoss/dpixel, with
// visibility held fixed: path replay re-traces every pathlight sample with
// the Halton indices of the forward pass (kept per level), recomputes the
// vertex albedos with their partial derivatives and applies the product rule
// to the estimator of pathlight.cc:305-387 (contri_b = s * R_b * X_b, pixel
// term = rho0 * (mult * sum_b contri_b) * power).  Sun power enters every
// light term linearly (sunlight.cc:35-57).  Supported scenes: one pathlight
// plus sun lights, materials with at most one modulator.

// Per output pixel: the direct sun term at the primary hit,
// Dir = vis * inte * (colour*power) * rho0.  With path_terms, also the
// pathlight's dependence on rho0 through its pixel term
// rho0 * mult * power * sum_{samples, b} contri_b (from the forward's contri),
// and the per-item weight W = dLoss/dpixel * rho0 * mult * power that the
// replay's shading kernel applies at every path vertex (0 where the pixel
// has no pathlight term).
template <int W>
__global__ __launch_bounds__(RB) void k_adjoint_direct(const RenderScene* __restrict__ scenes,
                                                       const Area* __restrict__ areas, int nareas, uint32_t nitems,
                                                       const int32_t* __restrict__ item_ti,
                                                       const float* __restrict__ item_t,
                                                       const float* __restrict__ grad_rgb, AdjAcc G,
                                                       int path_terms, int S, int D,
                                                       const uint32_t* __restrict__ qual,
                                                       const c3* __restrict__ contri,
                                                       const uint8_t* __restrict__ nseg, c3* __restrict__ w_item)
{
    extern __shared__ int s_stk[];
    __shared__ AdjLds L;
    const bool lds_ok = G.nmat <= ADJ_LDS_MAT && G.nmod <= ADJ_LDS_MOD && G.nl <= ADJ_LDS_LIGHT;
    adj_lds_init(L);
    for (uint32_t item = blockIdx.x * RB + threadIdx.x; item < nitems; item += gridDim.x * RB) {
        c3 w = mkc(0, 0, 0);
        if (item_ti[item] >= 0) {
            const Area a = areas[find_area(areas, nareas, item)];
            int i, j;
            if (item_output(a, item, i, j)) {
                const RenderScene& R = scenes[a.scene];
                const float* gp = grad_rgb + 3 * ((size_t)i * R.resx + j);
                const c3 g = mkc(gp[0], gp[1], gp[2]);
                LaneStack st;
                st.init(s_stk, threadIdx.x, RB);
                f3 eye;
                const Surf sp = primary_surf(R, a, item, item_ti[item], item_t[item], eye);
                const RhoD r0 = diffuse_d(R.mats, R.mods, R.texs, sp.mat, sp.P);
                for (int li = 0; li < R.nl; li++) {
                    const DevLight& Lt = R.lights[li];
                    if (!Lt.use_in_render || Lt.type != YG_LIGHT_SUN) continue;
                    if (Lt.cast_shadows &&
                        is_shadowed<W>(R.bvh, sp.P, Lt.dir, sp.obj, false, R.min_raydis, R.self_bias, st))
                        continue;
                    const f3 N = face_forward(sp.Ng, sp.N, norm3(eye));
                    const float inte = dot(N, Lt.dir);
                    if (inte < 0) continue;
                    const c3 fac = cscale(Lt.col, inte);  // d Dir / d rho0
                    adj_add_vertex(L, G, lds_ok, sp.mat, r0, cmul(g, fac));
                    const c3 dp = cmul(cscale(Lt.color, inte), r0.rho);  // d Dir / d power
                    atomicAdd(lds_ok ? &L.light[li] : &G.light[li], (double)(g.r * dp.r + g.g * dp.g + g.b * dp.b));
                }
                if (path_terms && R.path_light >= 0 && R.lights[R.path_light].use_in_render && qual[item]) {
                    const float pm = (float)(1.0 / (double)(float)S) * R.path_power;
                    c3 T = mkc(0, 0, 0);
                    for (int smp = 0; smp < S; smp++) {
                        const size_t p = (size_t)item * S + smp;
                        const int ns = nseg[p];
                        for (int b = 0; b < ns; b++) T = cadd(T, contri[p * D + b]);
                    }
                    adj_add_vertex(L, G, lds_ok, sp.mat, r0, cmul(g, cscale(T, pm)));
                    w = cmul(g, cscale(r0.rho, pm));
                }
            }
        }
        if (path_terms) w_item[item] = w;
    }
    adj_lds_flush(L, G, lds_ok);
}

// Pathlight sample replay (one thread per qualifying path).
template <int W>
__global__ __launch_bounds__(RB) void k_adjoint_path(const RenderScene* __restrict__ scenes,
                                                     const Area* __restrict__ areas, int nareas,
                                                     const uint32_t* __restrict__ active, uint32_t nactive,
                                                     uint32_t npaths, int S, int D,
                                                     const int32_t* __restrict__ item_ti,
                                                     const float* __restrict__ item_t,
                                                     const uint32_t* __restrict__ qoff,
                                                     const uint32_t* __restrict__ hidx,
                                                     const float* __restrict__ grad_rgb, AdjAcc G)
{
    extern __shared__ int s_stk[];
    __shared__ AdjLds L;
    const bool lds_ok = G.nmat <= ADJ_LDS_MAT && G.nmod <= ADJ_LDS_MOD && G.nl <= ADJ_LDS_LIGHT;
    adj_lds_init(L);
    const uint32_t idx = blockIdx.x * RB + threadIdx.x;
    if (idx < nactive) {
        const uint32_t p = active[idx];
        const uint32_t item = p / (uint32_t)S;
        const int smp = (int)(p % (uint32_t)S);
        const Area a = areas[find_area(areas, nareas, item)];
        int pi, pj;
        if (item_output(a, item, pi, pj)) {
            const RenderScene& R = scenes[a.scene];
            const float* gp = grad_rgb + 3 * ((size_t)pi * R.resx + pj);
            const c3 g = mkc(gp[0], gp[1], gp[2]);
            LaneStack st;
            st.init(s_stk, threadIdx.x, RB);
            f3 eye;
            const Surf sp0 = primary_surf(R, a, item, item_ti[item], item_t[item], eye);
            const RhoD r0 = diffuse_d(R.mats, R.mods, R.texs, sp0.mat, sp0.P);
            f3 where = sp0.P, N = face_forward(sp0.Ng, sp0.N, eye), Ru = sp0.NU, Rv = sp0.NV;
            // replayed path: per bounce the hit vertex (albedo + derivatives), the
            // light factor of its sun term without / with power, or a miss
            RhoD rv[ADJ_MAXD];
            int mv[ADJ_MAXD];
            c3 lf[ADJ_MAXD], lfp[ADJ_MAXD];
            bool hit[ADJ_MAXD];
            int nb = 0;
            c3 raycolor = mkc(1, 1, 1);
            float s = 0.0f;
            for (int b = 0; b < D && b < ADJ_MAXD; b++) {
                const uint32_t k = (b == 0) ? (qoff[item] - qoff[a.item_start]) * (uint32_t)S + (uint32_t)smp
                                            : hidx[(size_t)b * npaths + p];
                const float z1 = R.halton[(size_t)(2 * b) * R.hstride + k];
                const float z2 = R.halton[(size_t)(2 * b + 1) * R.hstride + k];
                const f3 ray = hemi_direction(z1, z2, N, Ru, Rv);
                if (b == 0) s = dot(ray, N);
                if (cenergy(raycolor) < 0.05f) break;
                float t;
                const int ti = first_hit<W>(R.bvh, where, ray, R.min_raydis, t, st, true);
                nb = b + 1;
                if (ti < 0) {
                    hit[b] = false;
                    break;
                }
                hit[b] = true;
                const Surf sp = make_surf(R.bvh, ti, where + ray * R.min_raydis, ray, t);
                rv[b] = diffuse_d(R.mats, R.mods, R.texs, sp.mat, sp.P);
                mv[b] = sp.mat;
                // sun terms of getLight (scene_t::light indirect), without the albedo
                c3 fac = mkc(0, 0, 0), facp = mkc(0, 0, 0);
                const f3 from = sp.P + (-ray);
                const f3 eye2 = from - sp.P;
                for (int li = 0; li < R.nl; li++) {
                    const DevLight& Lt = R.lights[li];
                    if (!Lt.use_in_indirect || Lt.type != YG_LIGHT_SUN) continue;
                    if (Lt.cast_shadows &&
                        is_shadowed<W>(R.bvh, sp.P, Lt.dir, sp.obj, false, R.min_raydis, R.self_bias, st))
                        continue;
                    const f3 Nl = face_forward(sp.Ng, sp.N, norm3(eye2));
                    const float inte = dot(Nl, Lt.dir);
                    if (inte < 0) continue;
                    fac = cadd(fac, cscale(Lt.col, inte));
                    facp = cadd(facp, cscale(Lt.color, inte));
                }
                lf[b] = fac;
                lfp[b] = facp;
                N = face_forward(sp.Ng, sp.N, -ray);
                create_cs(sp.N, Ru, Rv);
                raycolor = cmul(raycolor, rv[b].rho);
                where = sp.P;
            }
            const float mult = (float)(1.0 / (double)(float)S);
            const float pm = mult * R.path_power;
            // T = sum_b s R_b X_b; d pixel / d rho0 = T*pm; d pixel / d T = rho0*pm
            c3 T = mkc(0, 0, 0), Rb = mkc(1, 1, 1);
            c3 dpow = mkc(0, 0, 0);
            for (int b = 0; b < nb; b++) {
                const c3 X = hit[b] ? cmul(lf[b], rv[b].rho) : R.bg;
                T = cadd(T, cscale(cmul(Rb, X), s));
                if (hit[b]) {
                    dpow = cadd(dpow, cscale(cmul(Rb, cmul(lfp[b], rv[b].rho)), s));
                    Rb = cmul(Rb, rv[b].rho);
                }
            }
            adj_add_vertex(L, G, lds_ok, sp0.mat, r0, cmul(g, cscale(T, pm)));
            const c3 Wp = cmul(g, cscale(r0.rho, pm));
            // d T / d rho(H_k): bounces after k through R_b, bounce k through X_k
            for (int k = 0; k < nb; k++) {
                if (!hit[k]) continue;
                c3 dk = mkc(0, 0, 0);
                c3 Rex = mkc(1, 1, 1);  // product of rho(H_k') for k' < b, k' != k
                for (int b = 0; b < nb; b++) {
                    if (b == k) dk = cadd(dk, cscale(cmul(Rex, lf[b]), s));
                    else if (b > k) {
                        const c3 X = hit[b] ? cmul(lf[b], rv[b].rho) : R.bg;
                        dk = cadd(dk, cscale(cmul(Rex, X), s));
                    }
                    if (hit[b] && b != k) Rex = cmul(Rex, rv[b].rho);
                }
                adj_add_vertex(L, G, lds_ok, mv[k], rv[k], cmul(Wp, dk));
            }
            // sun power: every hit term is linear in it (one sun light assumed)
            for (int li = 0; li < R.nl; li++) {
                const DevLight& Lt = R.lights[li];
                if (Lt.use_in_indirect && Lt.type == YG_LIGHT_SUN) {
                    const c3 q = cmul(Wp, dpow);
                    atomicAdd(lds_ok ? &L.light[li] : &G.light[li], (double)(q.r + q.g + q.b));
                    break;
                }
            }
        }
    }
    adj_lds_flush(L, G, lds_ok);
}

// sort key of a replayed path: the number of segments the forward traced
__global__ void k_adj_keys(const uint32_t* __restrict__ active, uint32_t n, const uint8_t* __restrict__ nseg,
                           uint8_t* __restrict__ key)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key[i] = nseg[active[i]];
}

// Forward render (image into d_rgb) followed by the replay.  Gradient
// buffers are overwritten: d_grad_mat[nmat*3] (albedo), d_grad_mod[nmod*2]
// (modulator colour, modulator size scale lambda), d_grad_light[nl] (power).
extern "C" int yg_render_adjoint(yg_scene* s, const yg_render_opts* opts, const float* d_grad_rgb, float* d_rgb,
                                 double* d_grad_mat, double* d_grad_mod, double* d_grad_light,
                                 yg_render_stats* stats, void* stream)
{
    if (!s || !opts || !d_grad_rgb || !d_rgb || !d_grad_mat || !d_grad_light)
        return yg_fail(YG_ERR_ARG, "null argument");
    if (opts->aa_passes > 0) return yg_fail(YG_ERR_UNSUPPORTED, "the adjoint does not differentiate antialiasing passes");
    if (opts->flags & YG_RENDER_ROULETTE)
        return yg_fail(YG_ERR_UNSUPPORTED, "the adjoint does not differentiate Russian roulette");
    int nsun = 0, npath = 0;
    for (const auto& L : s->lights) {
        if (L.type == YG_LIGHT_SUN) nsun++;
        else if (L.type == YG_LIGHT_PATH) npath++;
        else return yg_fail(YG_ERR_UNSUPPORTED, "adjoint supports pathlight + sun lights only");
    }
    if (nsun > 1) return yg_fail(YG_ERR_UNSUPPORTED, "adjoint supports one sun light");
    for (const auto& m : s->mats)
        if (m.mod_count > 1) return yg_fail(YG_ERR_UNSUPPORTED, "adjoint supports one modulator per material");
    for (const auto& t : s->texs) {
        if (t.noise != YG_NOISE_NEWPERLIN)
            return yg_fail(YG_ERR_UNSUPPORTED, "adjoint differentiates newPerlin_t textures only");
        if (t.type == YG_TEX_MUSGRAVE && t.mtype != YG_MUSGRAVE_FBM)
            return yg_fail(YG_ERR_UNSUPPORTED, "adjoint differentiates Musgrave fBm only, not the multifractals");
    }
    if (!s->mods.empty() && !d_grad_mod) return yg_fail(YG_ERR_ARG, "d_grad_mod required");
    for (const yg_modulator& md : s->mods)
        if (md.texco != YG_TEXCO_GLOBAL)
            return yg_fail(YG_ERR_UNSUPPORTED, "adjoint differentiates modulators with texco global only");
    hipStream_t st = (hipStream_t)stream;
    std::vector<ScenePlan> plans(1);
    int rc = plan_scene(s, opts, d_rgb, plans[0]);
    if (rc) return rc;
    if (plans[0].R.depth > ADJ_MAXD) return yg_fail(YG_ERR_UNSUPPORTED, "adjoint supports depth <= 8");
    if (plans[0].cache) return yg_fail(YG_ERR_UNSUPPORTED, "adjoint of a pathlight with the irradiance cache");
    if (plans[0].R.ps_on) return yg_fail(YG_ERR_UNSUPPORTED, "adjoint with photonSampler_t (pathlight samples > 96)");
    const bool timing = opts->flags & YG_RENDER_TIMING;
    RenderWS& W = g_ws;
    const bool reuse = (opts->flags & YG_RENDER_REUSE) && W.last_uid == s->uid && W.last_gen == s->gen &&
                       same_opts(W.last_opts, *opts);
    if (!reuse) {
        W.last_uid = 0;
        if ((rc = render_plans(plans, st, stats, timing, false, !(opts->flags & YG_RENDER_NO_GRAPH)))) return rc;
        W.last_uid = s->uid;
        W.last_gen = s->gen;
        W.last_opts = *opts;
    }
    AdjAcc G;
    G.mat = d_grad_mat;
    G.mod = d_grad_mod;
    G.light = d_grad_light;
    G.nmat = (int)s->mats.size();
    G.nmod = (int)s->mods.size();
    G.nl = (int)s->lights.size();
    YG_HIP(hipMemsetAsync(d_grad_mat, 0, sizeof(double) * 3 * G.nmat, st));
    if (G.nmod) YG_HIP(hipMemsetAsync(d_grad_mod, 0, sizeof(double) * 2 * G.nmod, st));
    YG_HIP(hipMemsetAsync(d_grad_light, 0, sizeof(double) * G.nl, st));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timing) {
        YG_HIP(hipEventCreate(&e0));
        YG_HIP(hipEventCreate(&e1));
        YG_HIP(hipEventRecord(e0, st));
    }
    // replay: the forward's GI levels are traced again and every path vertex
    // adds its gradient terms; vertices with a zero albedo channel fall back
    // to the per-path replay below
    bool fused = !W.any_path;
    uint32_t n0 = 0;
    if (W.any_path) {
        uint32_t last_q = 0, last_off = 0;
        YG_HIP(hipMemcpyAsync(&last_q, W.qual.p + W.nitems - 1, 4, hipMemcpyDeviceToHost, st));
        YG_HIP(hipMemcpyAsync(&last_off, W.qoff.p + W.nitems - 1, 4, hipMemcpyDeviceToHost, st));
        YG_HIP(hipStreamSynchronize(st));
        n0 = (last_q + last_off) * (uint32_t)W.S;
    }
    if (!fused) {
        if ((rc = W.w_item.ensure(W.nitems)) || (rc = W.adeg.ensure(1))) return rc;
        YG_HIP(hipMemsetAsync(W.adeg.p, 0, sizeof(unsigned long long), st));
        YG_LAUNCH_WS(W.width, W.nocast, (k_adjoint_direct<8>), (k_adjoint_direct<4>), (k_adjoint_direct<2>), (k_adjoint_direct<24>), (k_adjoint_direct<20>), (k_adjoint_direct<18>), dim3(std::min(grid_for(W.nitems, RB), 4 * ADJ_GRID)), dim3(RB), W.stk_bytes, st, W.scenes.p,
                           W.areas.p, W.na, W.nitems, W.item_ti.p, W.item_t.p, d_grad_rgb, G, 1, W.S, W.D,
                           W.qual.p, W.contri.p, W.nseg.p, W.w_item.p);
        YG_HIP(hipGetLastError());
        std::vector<hipEvent_t> none;
        if (n0 > 0 && (rc = trace_levels<true>(st, false, none, G, W.w_item.p, W.adeg.p))) return rc;
        unsigned long long ndeg = 0;
        YG_HIP(hipMemcpyAsync(&ndeg, W.adeg.p, sizeof(ndeg), hipMemcpyDeviceToHost, st));
        YG_HIP(hipStreamSynchronize(st));
        if (ndeg > 0) {
            fused = true;
            YG_HIP(hipMemsetAsync(d_grad_mat, 0, sizeof(double) * 3 * G.nmat, st));
            if (G.nmod) YG_HIP(hipMemsetAsync(d_grad_mod, 0, sizeof(double) * 2 * G.nmod, st));
            YG_HIP(hipMemsetAsync(d_grad_light, 0, sizeof(double) * G.nl, st));
        }
    }
    if (fused) {
        YG_LAUNCH_WS(W.width, W.nocast, (k_adjoint_direct<8>), (k_adjoint_direct<4>), (k_adjoint_direct<2>), (k_adjoint_direct<24>), (k_adjoint_direct<20>), (k_adjoint_direct<18>), dim3(std::min(grid_for(W.nitems, RB), 4 * ADJ_GRID)), dim3(RB), W.stk_bytes, st, W.scenes.p,
                           W.areas.p, W.na, W.nitems, W.item_ti.p, W.item_t.p, d_grad_rgb, G, 0, W.S, W.D,
                           W.qual.p, W.contri.p, W.nseg.p, (c3*)nullptr);
        YG_HIP(hipGetLastError());
    }
    if (fused && W.any_path) {
        hipLaunchKernelGGL(k_path_init, dim3(grid_for(W.nitems, 256)), dim3(256), 0, st, W.nitems, W.S, W.D,
                           W.qual.p, W.qoff.p, W.active0.p, W.flag.p, W.nact.p, W.qcnt.p, W.dense.p);
        YG_HIP(hipGetLastError());
        if (n0 > 0) {
            // Replay order: longest paths first (stable, so pixel order is kept
            // within a length), so that a wavefront's lanes run the same number
            // of bounces instead of idling behind the longest path.
            if ((rc = W.akey0.ensure(n0)) || (rc = W.akey1.ensure(n0))) return rc;
            hipLaunchKernelGGL(k_adj_keys, dim3(grid_for(n0, 256)), dim3(256), 0, st, W.active0.p, n0, W.nseg.p,
                               W.akey0.p);
            YG_HIP(hipGetLastError());
            size_t bytes = 0;
            YG_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, W.akey0.p, W.akey1.p, W.active0.p,
                                                                W.active1.p, (int)n0, 0, 4, st));
            if ((rc = W.scan_tmp.ensure(bytes))) return rc;
            YG_HIP(hipcub::DeviceRadixSort::SortPairsDescending(W.scan_tmp.p, bytes, W.akey0.p, W.akey1.p,
                                                                W.active0.p, W.active1.p, (int)n0, 0, 4, st));
            YG_LAUNCH_WS(W.width, W.nocast, (k_adjoint_path<8>), (k_adjoint_path<4>), (k_adjoint_path<2>), (k_adjoint_path<24>), (k_adjoint_path<20>), (k_adjoint_path<18>), dim3(grid_for(n0, RB)), dim3(RB), W.stk_bytes, st, W.scenes.p, W.areas.p, W.na,
                               W.active1.p, n0, W.npaths, W.S, W.D, W.item_ti.p, W.item_t.p, W.qoff.p, W.hidx.p,
                               d_grad_rgb, G);
            YG_HIP(hipGetLastError());
        }
    }
    if (timing) {
        YG_HIP(hipEventRecord(e1, st));
        YG_HIP(hipEventSynchronize(e1));
        float ms = 0;
        YG_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (stats) stats->ms_trace = ms;  // adjoint replay time
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return YG_OK;
}
