// Device-side procedural noise and textures: newPerlin_t, fBm_t, turbulence,
// the clouds / marble / wood / Musgrave textures and modulator_t's colour blend.
#pragma once

#include "yg_internal.h"

// The 256-entry permutation of noise.cc:118-127 (the reference table is this
// permutation twice); values index it with X+1 <= 256 and A+1 <= 511.
// Each translation unit holds its own copy (no relocatable device code);
// perm_upload() fills the copy of the unit it is called from.
static __constant__ unsigned char YG_PERM[512];
void perm_fill(unsigned char out[512]);

static int perm_upload()
{
    static bool done = false;
    if (done) return YG_OK;
    unsigned char h[512];
    perm_fill(h);
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(YG_PERM), h, 512);
    if (e != hipSuccess) return yg_fail(YG_ERR_HIP, std::string("perm upload: ") + hipGetErrorString(e));
    done = true;
    return YG_OK;
}

// newPerlin_t::fade / grad (noise.h:33-40)
__device__ __forceinline__ float pfade(float t) { return t * t * t * (t * (t * 6 - 15) + 10); }
__device__ __forceinline__ float pgrad(int hash, float x, float y, float z)
{
    int h = hash & 15;
    float u = h < 8 ? x : y, v = h < 4 ? y : h == 12 || h == 14 ? x : z;
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}
__device__ __forceinline__ float plerp(float t, float a, float b) { return a + t * (b - a); }

// newPerlin_t::operator() (noise.cc:161-183), value in [0,1]
__device__ __forceinline__ float new_perlin(float x, float y, float z)
{
    const unsigned char* H = YG_PERM;
    float u = floorf(x), v = floorf(y), w = floorf(z);
    int X = ((int)u) & 255, Y = ((int)v) & 255, Z = ((int)w) & 255;
    x -= u; y -= v; z -= w;
    u = pfade(x); v = pfade(y); w = pfade(z);
    int A = H[X] + Y, AA = H[A] + Z, AB = H[A + 1] + Z;
    int B = H[X + 1] + Y, BA = H[B] + Z, BB = H[B + 1] + Z;
    float nv = plerp(w,
                     plerp(v, plerp(u, pgrad(H[AA], x, y, z), pgrad(H[BA], x - 1, y, z)),
                           plerp(u, pgrad(H[AB], x, y - 1, z), pgrad(H[BB], x - 1, y - 1, z))),
                     plerp(v, plerp(u, pgrad(H[AA + 1], x, y, z - 1), pgrad(H[BA + 1], x - 1, y, z - 1)),
                           plerp(u, pgrad(H[AB + 1], x, y - 1, z - 1), pgrad(H[BB + 1], x - 1, y - 1, z - 1))));
    return 0.5f + 0.5f * nv;
}

__device__ __forceinline__ float signed_noise(float x, float y, float z) { return 2.0f * new_perlin(x, y, z) - 1.0f; }

// fBm_t::operator() (noise.cc:638-650)
__device__ __forceinline__ float fbm(float x, float y, float z, float H, float lac, float oct)
{
    float value = 0, pwr = 1, pwHL = dpow(lac, -H);
    for (int i = 0; i < (int)oct; i++) {
        value += signed_noise(x, y, z) * pwr;
        pwr *= pwHL;
        x *= lac; y *= lac; z *= lac;
    }
    float rmd = oct - floorf(oct);
    if (rmd != 0.f) value += rmd * signed_noise(x, y, z) * pwr;
    return value;
}

// turbulence (noise.cc:808-819); newPerlin_t adds no coordinate offset
__device__ __forceinline__ float turbulence(float x, float y, float z, int oct, float size, int hard)
{
    float val, amp = 1, sum = 0;
    x *= size; y *= size; z *= size;
    for (int i = 0; i <= oct; i++, amp *= 0.5f, x *= 2.0f, y *= 2.0f, z *= 2.0f) {
        val = new_perlin(x, y, z);
        if (hard) val = fabsf(2.0f * val - 1.0f);
        sum += amp * val;
    }
    return sum * ((float)(1 << oct) / (float)((1 << (oct + 1)) - 1));
}

// texture getFloat (basictex.cc:82-90, 133-153, 194-216, 677-680)
__device__ __forceinline__ float tex_float(const yg_texture& t, f3 p)
{
    switch (t.type) {
    case YG_TEX_CLOUDS: {
        float v = turbulence(p.x, p.y, p.z, t.depth, t.size, t.hard);
        if (t.bias) {
            v *= v;
            if (t.bias == 1) return -v;
        }
        return v;
    }
    case YG_TEX_MARBLE:
    case YG_TEX_WOOD: {
        float w;
        if (t.type == YG_TEX_MARBLE) {
            w = (float)((double)(p.x + p.y + p.z) * 5.0 +
                        ((t.turb == 0.0f) ? 0.0 : (double)(t.turb * turbulence(p.x, p.y, p.z, t.depth, t.size, t.hard))));
        } else {
            if (t.rings) w = (float)((double)sqrtf(p.x * p.x + p.y * p.y + p.z * p.z) * 20.0);
            else w = (float)((double)(p.x + p.y + p.z) * 10.0);
            w = (float)((double)w + ((t.turb == 0.0f) ? 0.0 : (double)(t.turb * turbulence(p.x, p.y, p.z, t.depth, t.size, t.hard))));
        }
        switch (t.shape) {
        case YG_SHAPE_SAW:
            w *= (float)(0.5 * M_1_PI);
            w -= floorf(w);
            break;
        case YG_SHAPE_TRI:
            w *= (float)(0.5 * M_1_PI);
            w = fabsf(2.0f * (w - floorf(w)) - 1.0f);
            break;
        default:
            w = 0.5f + 0.5f * dsin(w);
        }
        if (t.type == YG_TEX_MARBLE) return dpow(w, t.sharpness);
        return w;
    }
    case YG_TEX_MUSGRAVE:
        return t.iscale * fbm(p.x * t.size, p.y * t.size, p.z * t.size, t.H, t.lacunarity, t.octaves);
    }
    return 0.0f;
}

// getColor = color1 + getFloat(p)*(color2 - color1)
__device__ __forceinline__ c3 tex_color(const yg_texture& t, f3 p)
{
    float f = tex_float(t, p);
    c3 c1 = mkc(t.color1[0], t.color1[1], t.color1[2]);
    c3 c2 = mkc(t.color2[0], t.color2[1], t.color2[2]);
    return cadd(c1, cscale(csub(c2, c1), f));
}

// modulator_t::modulate colour part (texture.cc:499-534) with procedural
// global mapping texpt = size*(P + ofs) (texture.cc:311-317)
__device__ __forceinline__ c3 modulate(const yg_modulator& m, const yg_texture* texs, c3 C, f3 P)
{
    const yg_texture& t = texs[m.tex];
    f3 tp = mk3(m.size[0] * (P.x + m.ofs[0]), m.size[1] * (P.y + m.ofs[1]), m.size[2] * (P.z + m.ofs[2]));
    c3 tc = tex_color(t, tp);
    switch (m.mode) {
    case YG_MOD_MIX:
        if (m.color > 0) C = cmix(tc, C, m.color);
        break;
    case YG_MOD_MUL:
        if (m.color > 0) C = cmul(cmix(tc, mkc(1, 1, 1), m.color), C);
        break;
    case YG_MOD_ADD:
        if (m.color > 0) C = cadd(cscale(tc, m.color), C);
        break;
    case YG_MOD_SUB:
        if (m.color > 0) C = cadd(cscale(tc, -m.color), C);
        break;
    }
    return C;
}

// genericShader_t::getDiffuse (basicshaders.cc:191-204)
__device__ __forceinline__ c3 get_diffuse(const yg_material* mats, const yg_modulator* mods,
                                          const yg_texture* texs, int mat, f3 P)
{
    const yg_material& m = mats[mat];
    c3 c = mkc(m.color[0], m.color[1], m.color[2]);
    for (int k = 0; k < m.mod_count; k++) c = modulate(mods[m.mod_first + k], texs, c, P);
    return c;
}
