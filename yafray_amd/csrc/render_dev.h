// Device-side shading for the render pipeline: surface points, the generic
// diffuse shader, sun and area lights, and the pathlight hemisphere sampler.
#pragma once

#include "noise_dev.h"
#include "gphoton_dev.h"
#include "traverse.h"

// One light of the render's light table (see yg_light); derived quantities
// are precomputed on the host with the reference's float operations.
struct DevLight {
    int type, use_in_render, use_in_indirect, cast_shadows;
    f3 dir;       // sun: normalised direction; area: plane normal (b-a)^(d-a)
    c3 col;       // sun: colour*power; area: pow*colour
    f3 from;      // area centre
    int samples;  // area sample count after fillQuad
    int pts;      // area: first entry of the sample point / jitter arrays
    int K;        // photon: search count
    float fixedradius;
    float power;  // the light's "power" parameter
    c3 color;     // the light's "color" parameter
    int fsamples; // area "psamples": guessLight's test draws (0: always the penumbra loop)
    f3 corner, toX, toY;  // area: a, b - a, d - a (the one-sample light)
};

// guessLight's verdict for one area light at one shading point
// (arealight.cc:176-200): 2 bits per light index in an item's status word
#define YG_GUESS_PENUMBRA 0u
#define YG_GUESS_LIGHT 1u
#define YG_GUESS_SHADOW 2u

// Photon map view for the K-nearest gather.
struct PhotonView {
    const float* photons;   // 9 floats per photon: dir, pos, colour
    const int32_t* gstart;  // uniform grid cell ranges
    const int32_t* gitems;
    float gmin[3];
    float gcell;
    int gdim[3];
    int np_emit;
    int64_t n;
};

// Everything a render kernel needs about one scene.
struct RenderScene {
    BVHView bvh;
    const yg_material* mats;
    const yg_modulator* mods;
    const yg_texture* texs;
    const DevLight* lights;
    const f3* apts;          // area-light sample points
    const f3* ajh;           // area-light horizontal jitter vectors
    const f3* ajv;           // area-light vertical jitter vectors
    const int32_t* obj_flags;
    int nl;
    f3 cam_pos, vright, vup, vto;
    int resx, resy;
    float min_raydis, self_bias;
    int raydepth;
    c3 bg;
    int path_light;          // index of the pathlight or -1
    int samples, depth;
    float path_power, threshold;
    const float* halton;     // [2*(depth+1)][hstride] getNext() tables
    int hstride;
    uint32_t lcg_seed;
    int photon_light;        // index of the photon light or -1
    PhotonView pm;
    int gp_on;               // a global photon light's irradiance map feeds the path vertices
    GPhotonView gp;
    float* out_rgb;
    uint32_t item_base;      // first item of this scene
    uint32_t area_base;      // first area of this scene
    // antialiasing (scene.cc:740-816): the launch's pass, -1 = the first pass
    int aa_pass, aa_mins, aa_jitter;
    float aa_pw, aa_totsamdiv;
    const uint2* aa_items;   // AA-pass items: (first-pass item, sample number)
    // tone mapping (scene_t::setExposure / setGamma, scene.h:161-164)
    int tm, tm_clamp;
    float tm_e, tm_g;        // -exposure, 1/gamma (0 for gamma 0)
    // Russian roulette (YG_RENDER_ROULETTE): first GI level it decides, 0 = off
    int rr_from;
    uint32_t rr_seed;
    // photonSampler_t (pathtools.cc:80-266): the pathlight's sampler when the
    // global photon map is published and samples > 96 (pathlight.cc:731)
    int ps_on;
    int ps_par, ps_mer, ps_nd;   // hemigrid paralels x meridians, Halton dimensions
    float ps_pdiv, ps_mdiv;
    float* ps_h;                 // [area][ps_hstride][dim] Halton values from the area's ourRandomI() starts
    uint32_t ps_hstride;
    uint32_t* ps_flag;           // [area] the area's renderState_t creates the sampler
    int32_t* ps_hits;            // [item][sector] samples of each sector (samplingFrom)
    float* ps_w;                 // [item][sector] weight = max hits / hits
    float* ps_multi;             // [item] multiplier() = 1 / (max hits x sectors)
};

// Uniform [0, 1) of a path's roulette at one level: an integer hash of
// (path, level, seed) (Wellons' lowbias32 finaliser), so a frame is
// reproducible and launch order does not matter.
__device__ __forceinline__ float rr_unit(uint32_t p, int level, uint32_t seed)
{
    uint32_t h = p * 0x9E3779B1u ^ (uint32_t)(level + 1) * 0x85EBCA77u ^ seed * 0xC2B2AE3Du;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// color_t::expgam_Adjust (color.h:175-195).  color.h includes only <cmath>, so
// its unqualified exp and pow are the C library's double functions: each
// channel is formed in double and stored as float.
__device__ __forceinline__ float clamp0(float v) { return (v < 0) ? 0.f : v; }
__device__ __forceinline__ float clamp01(float v) { return (v < 0) ? 0.f : ((v > 1) ? 1.f : v); }

__device__ __forceinline__ c3 expgam_adjust(c3 c, float e, float g, bool clamp_rgb)
{
    if (e == 0.f && g == 1.f) {
        if (clamp_rgb) c = mkc(clamp01(c.r), clamp01(c.g), clamp01(c.b));
        return c;
    }
    if (e != 0.f) {
        c = mkc(clamp0(c.r), clamp0(c.g), clamp0(c.b));
        c = mkc((float)(1.0 - exp((double)(c.r * e))), (float)(1.0 - exp((double)(c.g * e))),
                (float)(1.0 - exp((double)(c.b * e))));
    }
    if (g != 1.f) {
        c = mkc(clamp0(c.r), clamp0(c.g), clamp0(c.b));
        c = mkc((float)pow((double)c.r, (double)g), (float)pow((double)c.g, (double)g),
                (float)pow((double)c.b, (double)g));
    }
    return c;
}

// A render area (renderArea_t): the extended rectangle x,y,w,h is rendered,
// the real rectangle rx,ry,rw,rh is written out (renderblock.cc:8-20, 71-101).
struct Area {
    int scene;
    int x, y, w, h, rx, ry, rw, rh;
    uint32_t item_start;  // first item of the area in this launch
    uint32_t src_start;   // first item of the area in the first pass
};

// RI_vdC / RI_S / RI_LP (mcqmc.h:59-81): radical inverses by bit reversal and
// the Sobol' and Larcher-Pillichshammer generator matrices, as float.
__device__ __forceinline__ float ri_vdc(uint32_t bits)
{
    return (float)((double)__builtin_bitreverse32(bits) * (1.0 / 4294967296.0));
}

__device__ __forceinline__ float ri_s(uint32_t i)
{
    uint32_t r = 0;
    for (uint32_t v = 1u << 31; i; i >>= 1, v ^= v >> 1)
        if (i & 1) r ^= v;
    return (float)((double)r * (1.0 / 4294967296.0));
}

__device__ __forceinline__ float ri_lp(uint32_t i)
{
    uint32_t r = 0;
    for (uint32_t v = 1u << 31; i; i >>= 1, v |= v >> 1)
        if (i & 1) r ^= v;
    return (float)((double)r * (1.0 / 4294967296.0));
}

// surfacePoint_t subset for a hit (mesh.cc:517-577, triangle.cc:138-160):
// P = from + Z*ray; Ng = face normal; N = Ng or the interpolated vertex
// normal (surf_normal); NU, NV = createCS(N).
struct Surf {
    f3 P, N, Ng, NU, NV;
    int obj, mat, ti;
    float sx = 0.0f, sy = 0.0f;  // screen position (camera-ray hits; 0 elsewhere, scene.cc:234)
};

// The shading normal of triangle_t::getSurface (triangle.cc:142-160) at the
// hit point P of leaf-ordered triangle ti with face normal Ng: Ng itself, or
// for a triangle of an autoSmooth'ed mesh the vertex normals interpolated
// with getInterpolation's weights (fa = 1 - fb - fc) and normalised.
__device__ __forceinline__ f3 surf_normal(const BVHView& bvh, int ti, f3 P, f3 Ng)
{
    if (!bvh.tvn) return Ng;
    const float4 A = bvh.tvn[3 * ti];
    if (A.w == 0.0f) return Ng;
    const float4 B = bvh.tvn[3 * ti + 1], C = bvh.tvn[3 * ti + 2];
    const float4 TA = bvh.tris[3 * ti], TB = bvh.tris[3 * ti + 1], TC = bvh.tris[3 * ti + 2];
    float fb, fc;
    tri_bary(Ng, mk3(TA.x, TA.y, TA.z), mk3(TB.x, TB.y, TB.z), mk3(TC.x, TC.y, TC.z), P, fb, fc);
    const float fa = 1.0f - fb - fc;
    const f3 nn = mk3(A.x, A.y, A.z) * fa + mk3(B.x, B.y, B.z) * fb + mk3(C.x, C.y, C.z) * fc;
    return norm3(nn);
}

__device__ __forceinline__ Surf make_surf(const BVHView& bvh, int ti, f3 f, f3 ray, float Z)
{
    Surf s;
    s.ti = ti;
    s.P = f + ray * Z;
    const float4 A = bvh.tris[3 * ti], B = bvh.tris[3 * ti + 1], C = bvh.tris[3 * ti + 2];
    s.Ng = mk3(A.w, B.w, C.w);
    s.N = surf_normal(bvh, ti, s.P, s.Ng);
    create_cs(s.N, s.NU, s.NV);
    s.obj = bvh.tobj[ti];
    s.mat = bvh.tmat[ti];
    s.sx = 0.0f;
    s.sy = 0.0f;
    return s;
}

// reflect (vector3d.h:221-226)
__device__ __forceinline__ f3 reflect3(f3 n, f3 v)
{
    const float vn = dot(v, n);
    if (vn < 0) return -v;
    return n * (2.0f * vn) - v;
}

// The texture coordinates of modulator_t::doMapping (texture.cc:125-165) for
// texco at hit sp, seen along eye (the eye vector of the shader call); vec is
// set for a direction ("normal", "reflect").  uv and orco points come from
// triangle_t::getSurface (triangle.cc:138-167): a triangle with no uv, vertex
// colours, vertex normals or orco points has u = v = -1 and no orco point;
// otherwise getInterpolation's weights interpolate them (u = v = 0 without uv).
__device__ __forceinline__ f3 tex_coords(const BVHView& bvh, int texco, const Surf& sp, f3 eye, bool& vec)
{
    vec = false;
    if (texco == YG_TEXCO_UV || texco == YG_TEXCO_ORCO) {
        const float4* r = bvh.ptex ? bvh.ptex + 4 * (size_t)bvh.tprim[sp.ti] : nullptr;
        const float4 r0 = r ? r[0] : make_float4(0, 0, 0, 0), r1 = r ? r[1] : make_float4(0, 0, 0, 0);
        const float4 r2 = r ? r[2] : make_float4(0, 0, 0, 0), r3 = r ? r[3] : make_float4(0, 0, 0, 0);
        const int fl = (int)r3.w;
        if (texco == YG_TEXCO_ORCO && !(fl & 4)) {
            // meshObject_t::toObjectOrco: backOrco * P
            const float* M = bvh.oorco + 12 * (size_t)sp.obj;
            const f3 P = sp.P;
            return mk3(M[0] * P.x + M[1] * P.y + M[2] * P.z + M[3], M[4] * P.x + M[5] * P.y + M[6] * P.z + M[7],
                       M[8] * P.x + M[9] * P.y + M[10] * P.z + M[11]);
        }
        const bool smooth = bvh.tvn && bvh.tvn[3 * sp.ti].w != 0.0f;
        float u = -1.0f, v = -1.0f;
        if (fl || smooth) {
            const float4 TA = bvh.tris[3 * sp.ti], TB = bvh.tris[3 * sp.ti + 1], TC = bvh.tris[3 * sp.ti + 2];
            float fb, fc;
            tri_bary(sp.Ng, mk3(TA.x, TA.y, TA.z), mk3(TB.x, TB.y, TB.z), mk3(TC.x, TC.y, TC.z), sp.P, fb, fc);
            const float fa = 1.0f - fb - fc;
            if (texco == YG_TEXCO_ORCO)
                return mk3(r1.z, r1.w, r2.x) * fa + mk3(r2.y, r2.z, r2.w) * fb + mk3(r3.x, r3.y, r3.z) * fc;
            u = (fl & 1) ? r0.x * fa + r0.z * fb + r1.x * fc : 0.0f;
            v = (fl & 1) ? r0.y * fa + r0.w * fb + r1.y * fc : 0.0f;
        }
        return mk3((float)((double)u * 2.0 - 1.0), (float)(1.0 - (double)v * 2.0), 0.0f);
    }
    if (texco == YG_TEXCO_WIN) return mk3(sp.sx, sp.sy, 0.0f);
    if (texco == YG_TEXCO_NOR) {
        vec = true;
        return sp.N;
    }
    if (texco == YG_TEXCO_REFL) {
        vec = true;
        const f3 ne = norm3(eye);
        return reflect3(face_forward(sp.Ng, sp.N, ne), ne);
    }
    return sp.P;
}

// Whether material mat has a modulator whose coordinates depend on the eye
// vector (texco "reflect")
__device__ __forceinline__ bool diffuse_reads_eye(const yg_material* mats, const yg_modulator* mods, int mat)
{
    const yg_material& m = mats[mat];
    for (int k = 0; k < m.mod_count; k++)
        if (mods[m.mod_first + k].texco == YG_TEXCO_REFL) return true;
    return false;
}

// genericShader_t::getDiffuse (basicshaders.cc:191-204) at hit sp, eye the
// call's eye vector
__device__ __forceinline__ c3 get_diffuse(const BVHView& bvh, const yg_material* mats, const yg_modulator* mods,
                                          const yg_texture* texs, const Surf& sp, f3 eye)
{
    const yg_material& m = mats[sp.mat];
    c3 c = mkc(m.color[0], m.color[1], m.color[2]);
    for (int k = 0; k < m.mod_count; k++) {
        const yg_modulator& md = mods[m.mod_first + k];
        bool vec;
        const f3 T = tex_coords(bvh, md.texco, sp, eye, vec);
        c = modulate(md, texs, c, T, vec);
    }
    return c;
}

__device__ __forceinline__ c3 get_diffuse(const RenderScene& R, const Surf& sp, f3 eye)
{
    return get_diffuse(R.bvh, R.mats, R.mods, R.texs, sp, eye);
}

// genericShader_t::fromLight, diffuse part (basicshaders.cc:46-70)
__device__ __forceinline__ c3 from_light(const RenderScene& R, const Surf& sp, f3 ldir, c3 lcol, f3 eye)
{
    const f3 edir = norm3(eye);
    const f3 N = face_forward(sp.Ng, sp.N, edir);
    const float inte = dot(N, ldir);
    if (inte < 0) return mkc(0, 0, 0);
    const c3 lkc = get_diffuse(R, sp, eye);
    return cmul(cscale(lcol, inte), lkc);
}

// sunLight_t::illuminate (sunlight.cc:35-57)
template <int W>
__device__ __forceinline__ c3 sun_illuminate(const RenderScene& R, const DevLight& L, const Surf& sp,
                                             f3 eye, LaneStack& st, uint32_t& nshadow)
{
    if (L.cast_shadows) {
        nshadow++;
        if (is_shadowed<W>(R.bvh, sp.P, L.dir, sp.obj, false, R.min_raydis, R.self_bias, st))
            return mkc(0, 0, 0);
    }
    return from_light(R, sp, L.dir, L.col, eye);
}

// ourRandom() (vector3d.h:288-297)
__device__ __forceinline__ float our_random(int32_t& seed)
{
    const int a = 16807, m = 2147483647, q = m / a, r = m % a;
    int32_t s = seed;
    s = a * (s % q) - r * (s / q);
    if (s < 0) s += m;
    seed = s;
    return (float)s / (float)m;
}

// state of ourRandom() after n further draws: seed * 16807^n mod (2^31-1)
__device__ __forceinline__ int32_t lcg_jump(uint32_t seed, uint64_t n, const uint64_t* pow2)
{
    const uint64_t m = 2147483647u;
    uint64_t acc = seed;
    for (int k = 0; n; k++, n >>= 1)
        if (n & 1) acc = (acc * pow2[k]) % m;
    return (int32_t)acc;
}

// areaLight_t::illuminate (arealight.cc:93-174), draws taken from the
// ourRandom() stream starting at *seed: the one-sample light (a random point
// of the quad), else by guessLight's verdict `status` (YG_GUESS_*; the
// penumbra loop when the light has no psamples)
template <int W>
__device__ __forceinline__ c3 area_illuminate(const RenderScene& R, const DevLight& L, const Surf& sp,
                                              f3 eye, int32_t& seed, LaneStack& st, uint32_t& nshadow,
                                              uint32_t status = YG_GUESS_PENUMBRA)
{
    c3 resul = mkc(0, 0, 0);
    const f3 edir = norm3(eye);
    const f3 N = face_forward(sp.Ng, sp.N, edir);
    const float plane_at = dot(sp.P - L.from, L.dir);
    if (plane_at <= 0) return mkc(0, 0, 0);
    if (L.samples == 1) {
        // corner + toX*ourRandom() + toY*ourRandom(): g++ evaluates the right
        // operand of + first, so toY takes the first draw
        const float r2 = our_random(seed), r1 = our_random(seed);
        const f3 sampleP = (L.corner + L.toX * r1) + L.toY * r2;
        const f3 Lv = sampleP - sp.P;
        if (dot(Lv, N) < 0) return mkc(0, 0, 0);
        nshadow++;
        if (!is_shadowed<W>(R.bvh, sp.P, sampleP, sp.obj, true, R.min_raydis, R.self_bias, st)) {
            f3 dir = Lv;
            float LD2 = normlensqr3(dir);
            LD2 = (LD2 != 0) ? (1.0f / LD2) : 1.0f;
            resul = cadd(resul, from_light(R, sp, dir, cscale(L.col, LD2), eye));
        }
        return cscale(resul, plane_at);
    }
    if (status == YG_GUESS_SHADOW) return mkc(0, 0, 0);  // fromLight of a zero colour
    if (status == YG_GUESS_LIGHT) {
        // every point unjittered, no shadow ray
        for (int i = 0; i < L.samples; ++i) {
            f3 dir = R.apts[L.pts + i] - sp.P;
            float LD2 = normlensqr3(dir);
            LD2 = (LD2 != 0) ? (1.0f / LD2) : 1.0f;
            resul = cadd(resul, from_light(R, sp, dir, cscale(L.col, LD2), eye));
        }
        const c3 r = cscale(resul, plane_at);
        const float fs = (float)L.samples;
        return mkc(r.r / fs, r.g / fs, r.b / fs);
    }
    for (int i = 0; i < L.samples; ++i) {
        const float dish = (float)((double)our_random(seed) - 0.5);
        const float disv = (float)((double)our_random(seed) - 0.5);
        const f3 sampleP = (R.apts[L.pts + i] + R.ajh[L.pts + i] * dish) + R.ajv[L.pts + i] * disv;
        const f3 Lv = sampleP - sp.P;
        if (dot(Lv, N) < 0) continue;
        nshadow++;
        if (!is_shadowed<W>(R.bvh, sp.P, sampleP, sp.obj, true, R.min_raydis, R.self_bias, st)) {
            f3 dir = Lv;
            float LD2 = normlensqr3(dir);
            LD2 = (LD2 != 0) ? (1.0f / LD2) : 1.0f;
            resul = cadd(resul, from_light(R, sp, dir, cscale(L.col, LD2), eye));
        }
    }
    const c3 r = cscale(resul, plane_at);
    const float fs = (float)L.samples;
    return mkc(r.r / fs, r.g / fs, r.b / fs);
}

// haltonSampler_t::nextDirection (pathtools.cc:32-43) with the two Halton
// values already fetched
__device__ __forceinline__ f3 hemi_direction(float z1, float z2, f3 N, f3 Ru, f3 Rv)
{
    if (z1 > 1.0f) z1 = 1.0f;
    z2 = (float)((double)z2 * (2.0 * M_PI));
    const f3 a = Ru * dcos(z2) + Rv * dsin(z2);
    return a * (float)sqrt(1.0 - (double)z1) + N * sqrtf(z1);
}

// photonSampler_t::nextDirection's vector (pathtools.cc:257-260): z1 is the
// sine of the angle to N
__device__ __forceinline__ f3 ps_direction(float z1, float z2, f3 N, f3 Ru, f3 Rv)
{
    if (z1 > 1.0f) z1 = 1.0f;
    const f3 a = Ru * dcos(z2) + Rv * dsin(z2);
    return a * z1 + N * (float)sqrt(1.0 - (double)(z1 * z1));
}

// scene_t::light with indirect = true restricted to sun lights (the only
// lights the GPU path supports inside GI paths): the getLight of
// pathlight.cc:186-191 at a path vertex, eye = (P + (-ray)) - P.
template <int W>
__device__ __forceinline__ c3 indirect_light(const RenderScene& R, const Surf& sp, f3 ray, LaneStack& st,
                                             uint32_t& nshadow)
{
    const f3 from = sp.P + (-ray);
    const f3 eye = from - sp.P;
    c3 fl = mkc(0, 0, 0);
    for (int li = 0; li < R.nl; li++) {
        const DevLight& L = R.lights[li];
        if (!L.use_in_indirect) continue;
        if (L.type == YG_LIGHT_SUN) fl = cadd(fl, sun_illuminate<W>(R, L, sp, eye, st, nshadow));
    }
    return fl;
}
