// Device-side procedural noise and textures: newPerlin_t, fBm_t, turbulence,
// the clouds / marble / wood / Musgrave textures and modulator_t's colour blend.
#pragma once

#include "yg_internal.h"

// The 256-entry permutation of noise.cc:118-127 (the reference table is this
// permutation twice); values index it with X+1 <= 256 and A+1 <= 511.
// Each translation unit holds its own copy (no relocatable device code);
// perm_upload() fills the copy of the unit it is called from.
static __constant__ unsigned char YG_PERM[512];
// The unit gradient vectors of stdPerlin_t and blenderNoise_t, 3 per entry.
static __constant__ float YG_GRAD[768];
void perm_fill(unsigned char out[512]);
void grad_fill(float out[768]);

static int perm_upload()
{
    static bool done = false;
    if (done) return YG_OK;
    unsigned char h[512];
    perm_fill(h);
    float g[768];
    grad_fill(g);
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(YG_PERM), h, 512);
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(YG_GRAD), g, sizeof(g));
    if (e != hipSuccess) return yg_fail(YG_ERR_HIP, std::string("noise table upload: ") + hipGetErrorString(e));
    done = true;
    return YG_OK;
}

// Copies the permutation into a block's LDS (all threads of the block call
// it, before any returns): the hash lookups then read LDS instead of three
// dependent levels of global loads per octave.
__device__ __forceinline__ void perm_to_lds(unsigned char* s)
{
    for (int i = threadIdx.x; i < 512; i += blockDim.x) s[i] = YG_PERM[i];
    __syncthreads();
}

// Both tables into LDS (kernels that may evaluate any noise generator).
__device__ __forceinline__ void noise_tables_to_lds(unsigned char* perm, float* grad)
{
    for (int i = threadIdx.x; i < 768; i += blockDim.x) {
        if (i < 512) perm[i] = YG_PERM[i];
        grad[i] = YG_GRAD[i];
    }
    __syncthreads();
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

// newPerlin_t::operator() (noise.cc:161-183), value in [0,1].  H: the
// permutation, YG_PERM or a block's LDS copy of it (perm_to_lds); the noise
// functions below pass it on.
__device__ __forceinline__ float new_perlin(float x, float y, float z, const unsigned char* H = YG_PERM)
{
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

// stdPerlin_t::operator() (noise.cc:372-435): Perlin's original noise over
// the permutation and the unit gradients, coordinates shifted by 10000 in
// float (which quantises them), the fade and the result formed in double.
__device__ __forceinline__ float std_perlin(float x, float y, float z, const unsigned char* H, const float* G)
{
    int b0[3], b1[3];
    float r0[3], r1[3];
    const float vec[3] = {x, y, z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float t = (float)((double)vec[k] + 10000.0);
        b0[k] = ((int)t) & 255;
        b1[k] = (b0[k] + 1) & 255;
        r0[k] = t - (float)(int)t;
        r1[k] = (float)((double)r0[k] - 1.0);
    }
    const int i = H[b0[0]], j = H[b1[0]];
    const int b00 = H[i + b0[1]], b10 = H[j + b0[1]], b01 = H[i + b1[1]], b11 = H[j + b1[1]];
    auto surve = [](float t) { return (float)((double)(t * t) * (3.0 - 2.0 * (double)t)); };
    const float sx = surve(r0[0]), sy = surve(r0[1]), sz = surve(r0[2]);
    // stdp_g has 514 rows (256 twice + 2): row k is gradient k & 255
    auto at = [&](int row, float rx, float ry, float rz) {
        const float* q = G + 3 * (row & 255);
        return rx * q[0] + ry * q[1] + rz * q[2];
    };
    float u = at(b00 + b0[2], r0[0], r0[1], r0[2]), v = at(b10 + b0[2], r1[0], r0[1], r0[2]);
    float a = plerp(sx, u, v);
    u = at(b01 + b0[2], r0[0], r1[1], r0[2]);
    v = at(b11 + b0[2], r1[0], r1[1], r0[2]);
    float b = plerp(sx, u, v);
    const float c = plerp(sy, a, b);
    u = at(b00 + b1[2], r0[0], r0[1], r1[2]);
    v = at(b10 + b1[2], r1[0], r0[1], r1[2]);
    a = plerp(sx, u, v);
    u = at(b01 + b1[2], r0[0], r1[1], r1[2]);
    v = at(b11 + b1[2], r1[0], r1[1], r1[2]);
    b = plerp(sx, u, v);
    const float d = plerp(sy, a, b);
    return (float)(0.5 + 0.75 * (double)plerp(sz, c, d));
}

// blenderNoise_t::operator() (noise.cc:439-509): corner weights
// 1 - 3t^2 + 2|t|^3 formed in double, gradients from the same unit vectors
// hashed through the permutation, clamped to [0, 1].
__device__ __forceinline__ float blender_noise(float x, float y, float z, const unsigned char* H, const float* G)
{
    const int ix = (int)floorf(x), iy = (int)floorf(y), iz = (int)floorf(z);
    const float ox = x - (float)ix, oy = y - (float)iy, oz = z - (float)iz;
    const float jx = ox - 1.0f, jy = oy - 1.0f, jz = oz - 1.0f;
    auto wp = [](float o) { const float c = o * o; return (float)(1.0 - 3.0 * (double)c + 2.0 * (double)c * (double)o); };
    auto wn = [](float j) { const float c = j * j; return (float)(1.0 - 3.0 * (double)c - 2.0 * (double)c * (double)j); };
    const float cn1 = wp(ox), cn2 = wp(oy), cn3 = wp(oz), cn4 = wn(jx), cn5 = wn(jy), cn6 = wn(jz);
    const int b00 = H[H[ix & 255] + (iy & 255)], b10 = H[H[(ix + 1) & 255] + (iy & 255)];
    const int b01 = H[H[ix & 255] + ((iy + 1) & 255)], b11 = H[H[(ix + 1) & 255] + ((iy + 1) & 255)];
    const int b20 = iz & 255, b21 = (iz + 1) & 255;
    auto term = [&](float i, int hidx, float px, float py, float pz) {
        const float* h = G + 3 * H[hidx];
        return i * (h[0] * px + h[1] * py + h[2] * pz);
    };
    float n = 0.5f;
    n += term(cn1 * cn2 * cn3, b20 + b00, ox, oy, oz);
    n += term(cn1 * cn2 * cn6, b21 + b00, ox, oy, jz);
    n += term(cn1 * cn5 * cn3, b20 + b01, ox, jy, oz);
    n += term(cn1 * cn5 * cn6, b21 + b01, ox, jy, jz);
    n += term(cn4 * cn2 * cn3, b20 + b10, jx, oy, oz);
    n += term(cn4 * cn2 * cn6, b21 + b10, jx, oy, jz);
    n += term(cn4 * cn5 * cn3, b20 + b11, jx, jy, oz);
    n += term(cn4 * cn5 * cn6, b21 + b11, jx, jy, jz);
    return (n < 0.0f) ? 0.0f : ((n > 1.0f) ? 1.0f : n);
}

// The noise generator of a texture (newNoise, basictex.cc:38-63)
__device__ __forceinline__ float noise3(int kind, float x, float y, float z, const unsigned char* perm,
                                        const float* grad)
{
    if (kind == YG_NOISE_STDPERLIN) return std_perlin(x, y, z, perm, grad);
    if (kind == YG_NOISE_BLENDER) return blender_noise(x, y, z, perm, grad);
    return new_perlin(x, y, z, perm);
}

// getSignedNoise (noise.h:245-248)
__device__ __forceinline__ float signed_noise(float x, float y, float z, const unsigned char* perm = YG_PERM,
                                              int kind = YG_NOISE_NEWPERLIN, const float* grad = YG_GRAD)
{
    return 2.0f * noise3(kind, x, y, z, perm, grad) - 1.0f;
}

// fBm_t::operator() (noise.cc:638-650) with pwHL = pow(lacunarity, -H)
// given: it is the same for every point of a texture, so bulk kernels take
// it from the host (dpow there equals the oracle's).
__device__ __forceinline__ float fbm_w(float x, float y, float z, float pwHL, float lac, float oct,
                                       const unsigned char* perm = YG_PERM, int kind = YG_NOISE_NEWPERLIN,
                                       const float* grad = YG_GRAD)
{
    float value = 0, pwr = 1;
    for (int i = 0; i < (int)oct; i++) {
        value += signed_noise(x, y, z, perm, kind, grad) * pwr;
        pwr *= pwHL;
        x *= lac; y *= lac; z *= lac;
    }
    float rmd = oct - floorf(oct);
    if (rmd != 0.f) value += rmd * signed_noise(x, y, z, perm, kind, grad) * pwr;
    return value;
}

__device__ __forceinline__ float fbm(float x, float y, float z, float H, float lac, float oct,
                                     const unsigned char* perm = YG_PERM, int kind = YG_NOISE_NEWPERLIN,
                                     const float* grad = YG_GRAD)
{
    return fbm_w(x, y, z, dpow(lac, -H), lac, oct, perm, kind, grad);
}

// The Musgrave multifractals over the texture's noise generator, each with
// pwHL = pow(lacunarity, -H) computed as fBm computes it.
// mFractal_t::operator() (noise.cc:668-681)
__device__ __forceinline__ float mfractal(float x, float y, float z, float H, float lac, float oct,
                                          const unsigned char* perm, int kind, const float* grad)
{
    float value = 1, pwr = 1;
    const float pwHL = dpow(lac, -H);
    for (int i = 0; i < (int)oct; i++) {
        value *= (pwr * signed_noise(x, y, z, perm, kind, grad) + 1.0f);
        pwr *= pwHL;
        x *= lac; y *= lac; z *= lac;
    }
    const float rmd = oct - floorf(oct);
    if (rmd != 0.0f) value *= (rmd * signed_noise(x, y, z, perm, kind, grad) * pwr + 1.0f);
    return value;
}

// heteroTerrain_t::operator() (noise.cc:695-719)
__device__ __forceinline__ float hetero_terrain(float x, float y, float z, float H, float lac, float oct, float offset,
                                                const unsigned char* perm, int kind, const float* grad)
{
    const float pwHL = dpow(lac, -H);
    float pwr = pwHL;
    float value = offset + signed_noise(x, y, z, perm, kind, grad);
    x *= lac; y *= lac; z *= lac;
    for (int i = 1; i < (int)oct; i++) {
        const float increment = (signed_noise(x, y, z, perm, kind, grad) + offset) * pwr * value;
        value += increment;
        pwr *= pwHL;
        x *= lac; y *= lac; z *= lac;
    }
    const float rmd = oct - floorf(oct);
    if (rmd != 0.0f) {
        const float increment = (signed_noise(x, y, z, perm, kind, grad) + offset) * pwr * value;
        value += rmd * increment;
    }
    return value;
}

// hybridMFractal_t::operator() (noise.cc:731-754)
__device__ __forceinline__ float hybrid_mfractal(float x, float y, float z, float H, float lac, float oct,
                                                 float offset, float gain, const unsigned char* perm, int kind,
                                                 const float* grad)
{
    const float pwHL = dpow(lac, -H);
    float pwr = pwHL;
    float result = signed_noise(x, y, z, perm, kind, grad) + offset;
    float weight = gain * result;
    x *= lac; y *= lac; z *= lac;
    for (int i = 1; (weight > 0.001f) && (i < (int)oct); i++) {
        if (weight > 1.0f) weight = 1.0f;
        const float signal = (signed_noise(x, y, z, perm, kind, grad) + offset) * pwr;
        pwr *= pwHL;
        result += weight * signal;
        weight *= gain * signal;
        x *= lac; y *= lac; z *= lac;
    }
    const float rmd = oct - floorf(oct);
    if (rmd != 0.0f) result += rmd * ((signed_noise(x, y, z, perm, kind, grad) + offset) * pwr);
    return result;
}

// ridgedMFractal_t::operator() (noise.cc:768-791)
__device__ __forceinline__ float ridged_mfractal(float x, float y, float z, float H, float lac, float oct,
                                                 float offset, float gain, const unsigned char* perm, int kind,
                                                 const float* grad)
{
    const float pwHL = dpow(lac, -H);
    float pwr = pwHL;
    float signal = offset - fabsf(signed_noise(x, y, z, perm, kind, grad));
    signal *= signal;
    float result = signal;
    float weight = 1.0f;
    for (int i = 1; i < (int)oct; i++) {
        x *= lac; y *= lac; z *= lac;
        weight = signal * gain;
        if (weight > 1.0f) weight = 1.0f;
        else if (weight < 0.0f) weight = 0.0f;
        signal = offset - fabsf(signed_noise(x, y, z, perm, kind, grad));
        signal *= signal;
        signal *= weight;
        result += signal * pwr;
        pwr *= pwHL;
    }
    return result;
}

// the generator textureMusgrave_t's "type" selects (basictex.cc:657-672)
__device__ __forceinline__ float musgrave(const yg_texture& t, float x, float y, float z, const unsigned char* perm,
                                          const float* grad)
{
    switch (t.mtype) {
    case YG_MUSGRAVE_MULTIFRACTAL:
        return mfractal(x, y, z, t.H, t.lacunarity, t.octaves, perm, t.noise, grad);
    case YG_MUSGRAVE_HETEROTERRAIN:
        return hetero_terrain(x, y, z, t.H, t.lacunarity, t.octaves, t.offset, perm, t.noise, grad);
    case YG_MUSGRAVE_HYBRIDMF:
        return hybrid_mfractal(x, y, z, t.H, t.lacunarity, t.octaves, t.offset, t.gain, perm, t.noise, grad);
    case YG_MUSGRAVE_RIDGEDMF:
        return ridged_mfractal(x, y, z, t.H, t.lacunarity, t.octaves, t.offset, t.gain, perm, t.noise, grad);
    default:
        return fbm(x, y, z, t.H, t.lacunarity, t.octaves, perm, t.noise, grad);
    }
}

// turbulence (noise.cc:808-819) over the texture's noise generator
__device__ __forceinline__ float turbulence(float x, float y, float z, int oct, float size, int hard,
                                            const unsigned char* perm = YG_PERM, int kind = YG_NOISE_NEWPERLIN,
                                            const float* grad = YG_GRAD)
{
    float val, amp = 1, sum = 0;
    if (kind == YG_NOISE_BLENDER) {  // only blenderNoise_t offsets the point (noise.h:60-61)
        x += 1.0f; y += 1.0f; z += 1.0f;
    }
    x *= size; y *= size; z *= size;
    for (int i = 0; i <= oct; i++, amp *= 0.5f, x *= 2.0f, y *= 2.0f, z *= 2.0f) {
        val = noise3(kind, x, y, z, perm, grad);
        if (hard) val = fabsf(2.0f * val - 1.0f);
        sum += amp * val;
    }
    return sum * ((float)(1 << oct) / (float)((1 << (oct + 1)) - 1));
}

// texture getFloat (basictex.cc:82-90, 133-153, 194-216, 677-680)
__device__ __forceinline__ float tex_float(const yg_texture& t, f3 p, const unsigned char* perm = YG_PERM,
                                           const float* grad = YG_GRAD)
{
    switch (t.type) {
    case YG_TEX_CLOUDS: {
        float v = turbulence(p.x, p.y, p.z, t.depth, t.size, t.hard, perm, t.noise, grad);
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
                        ((t.turb == 0.0f) ? 0.0 : (double)(t.turb * turbulence(p.x, p.y, p.z, t.depth, t.size, t.hard, perm, t.noise, grad))));
        } else {
            if (t.rings) w = (float)((double)sqrtf(p.x * p.x + p.y * p.y + p.z * p.z) * 20.0);
            else w = (float)((double)(p.x + p.y + p.z) * 10.0);
            w = (float)((double)w + ((t.turb == 0.0f) ? 0.0 : (double)(t.turb * turbulence(p.x, p.y, p.z, t.depth, t.size, t.hard, perm, t.noise, grad))));
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
        // textureMusgrave_t::getFloat (basictex.cc:677-680): iscale * mGen(p * size)
        return t.iscale * musgrave(t, p.x * t.size, p.y * t.size, p.z * t.size, perm, grad);
    }
    return 0.0f;
}

// getColor = color1 + getFloat(p)*(color2 - color1)
__device__ __forceinline__ c3 tex_color(const yg_texture& t, f3 p, const unsigned char* perm = YG_PERM,
                                        const float* grad = YG_GRAD)
{
    float f = tex_float(t, p, perm, grad);
    c3 c1 = mkc(t.color1[0], t.color1[1], t.color1[2]);
    c3 c2 = mkc(t.color2[0], t.color2[1], t.color2[2]);
    return cadd(c1, cscale(csub(c2, c1), f));
}

// modulator_t::doMapping for a procedural texture (texture.cc:167-180 and
// 311-317) from the texture coordinates P of the modulator's texco: the
// texture matrix (matrix4x4_t * point3d_t, matrix4.h:86-91; for a direction,
// texco "normal" or "reflect", matrix4x4_t * vector3d_t without the
// translation, matrix4.h:79-84), then per axis size * (T[projected axis] +
// ofs), or size * ofs for "n"
__device__ __forceinline__ f3 mod_texpt(const yg_modulator& m, f3 P, bool vec = false)
{
    float T[3] = {P.x, P.y, P.z};
    if (m.has_mtx) {
        const float* M = m.mtx;
        if (vec) {
            T[0] = M[0] * P.x + M[1] * P.y + M[2] * P.z;
            T[1] = M[4] * P.x + M[5] * P.y + M[6] * P.z;
            T[2] = M[8] * P.x + M[9] * P.y + M[10] * P.z;
        } else {
            T[0] = M[0] * P.x + M[1] * P.y + M[2] * P.z + M[3];
            T[1] = M[4] * P.x + M[5] * P.y + M[6] * P.z + M[7];
            T[2] = M[8] * P.x + M[9] * P.y + M[10] * P.z + M[11];
        }
    }
    float o[3];
    for (int a = 0; a < 3; a++) {
        const int pj = m.proj[a];
        o[a] = (pj < 0) ? m.size[a] * m.ofs[a] : m.size[a] * (T[pj == 0 ? a : pj - 1] + m.ofs[a]);
    }
    return mk3(o[0], o[1], o[2]);
}

// modulator_t::modulate colour part (texture.cc:499-534) at texture
// coordinates P (vec: a direction)
__device__ __forceinline__ c3 modulate(const yg_modulator& m, const yg_texture* texs, c3 C, f3 P, bool vec = false)
{
    const yg_texture& t = texs[m.tex];
    const f3 tp = mod_texpt(m, P, vec);
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

// ---- derivatives for the adjoint (gradient of the noise value in space) --

// gradient vector of grad(hash, .) (noise.h:35-40): grad is linear in (x,y,z)
__device__ __forceinline__ f3 pgrad_vec(int hash)
{
    const int h = hash & 15;
    f3 g = mk3(0, 0, 0);
    const float su = (h & 1) ? -1.0f : 1.0f, sv = (h & 2) ? -1.0f : 1.0f;
    if (h < 8) g.x += su; else g.y += su;
    if (h < 4) g.y += sv;
    else if (h == 12 || h == 14) g.x += sv;
    else g.z += sv;
    return g;
}

__device__ __forceinline__ float pfade_d(float t) { return 30.0f * t * t * (t - 1.0f) * (t - 1.0f); }

// d newPerlin_t / d(x,y,z)
__device__ __forceinline__ f3 new_perlin_grad(float x, float y, float z)
{
    const unsigned char* H = YG_PERM;
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    int X = ((int)fx) & 255, Y = ((int)fy) & 255, Z = ((int)fz) & 255;
    x -= fx; y -= fy; z -= fz;
    const float u = pfade(x), v = pfade(y), w = pfade(z);
    const float du = pfade_d(x), dv = pfade_d(y), dw = pfade_d(z);
    int A = H[X] + Y, AA = H[A] + Z, AB = H[A + 1] + Z;
    int B = H[X + 1] + Y, BA = H[B] + Z, BB = H[B + 1] + Z;
    const int hc[8] = {H[AA], H[BA], H[AB], H[BB], H[AA + 1], H[BA + 1], H[AB + 1], H[BB + 1]};
    float g[8];
    f3 G[8];
    for (int k = 0; k < 8; k++) {
        const float cx = x - (float)(k & 1), cy = y - (float)((k >> 1) & 1), cz = z - (float)((k >> 2) & 1);
        g[k] = pgrad(hc[k], cx, cy, cz);
        G[k] = pgrad_vec(hc[k]);
    }
    const float k1 = g[1] - g[0], k2 = g[2] - g[0], k3 = g[4] - g[0];
    const float k4 = g[0] - g[1] - g[2] + g[3], k5 = g[0] - g[2] - g[4] + g[6];
    const float k6 = g[0] - g[1] - g[4] + g[5];
    const float k7 = -g[0] + g[1] + g[2] - g[3] + g[4] - g[5] - g[6] + g[7];
    f3 d = mk3(du * (k1 + k4 * v + k6 * w + k7 * v * w), dv * (k2 + k4 * u + k5 * w + k7 * u * w),
               dw * (k3 + k5 * v + k6 * u + k7 * u * v));
    for (int k = 0; k < 8; k++) {
        const float wu = (k & 1) ? u : 1.0f - u, wv = ((k >> 1) & 1) ? v : 1.0f - v, ww = ((k >> 2) & 1) ? w : 1.0f - w;
        d = d + G[k] * (wu * wv * ww);
    }
    return d * 0.5f;
}

// d getFloat / dp for the textures with a derivative (Musgrave fBm, clouds)
__device__ __forceinline__ f3 tex_float_grad(const yg_texture& t, f3 p)
{
    if (t.type == YG_TEX_MUSGRAVE) {
        const float s = t.size;
        float x = p.x * s, y = p.y * s, z = p.z * s, pwr = 1.0f, scale = 1.0f;
        const float pwHL = dpow(t.lacunarity, -t.H);
        f3 g = mk3(0, 0, 0);
        for (int i = 0; i < (int)t.octaves; i++) {
            g = g + new_perlin_grad(x, y, z) * (2.0f * pwr * scale);
            pwr *= pwHL;
            x *= t.lacunarity; y *= t.lacunarity; z *= t.lacunarity;
            scale *= t.lacunarity;
        }
        const float rmd = t.octaves - floorf(t.octaves);
        if (rmd != 0.f) g = g + new_perlin_grad(x, y, z) * (2.0f * rmd * pwr * scale);
        return g * (t.iscale * s);
    }
    if (t.type == YG_TEX_CLOUDS) {
        float x = p.x * t.size, y = p.y * t.size, z = p.z * t.size, amp = 1.0f, scale = 1.0f;
        f3 g = mk3(0, 0, 0);
        float v = 0.0f;
        for (int i = 0; i <= t.depth; i++, amp *= 0.5f, scale *= 2.0f, x *= 2.0f, y *= 2.0f, z *= 2.0f) {
            const float n = new_perlin(x, y, z);
            f3 dn = new_perlin_grad(x, y, z) * scale;
            float val = n;
            if (t.hard) {
                val = fabsf(2.0f * n - 1.0f);
                dn = dn * ((2.0f * n - 1.0f) >= 0.0f ? 2.0f : -2.0f);
            }
            v += amp * val;
            g = g + dn * amp;
        }
        const float norm = (float)(1 << t.depth) / (float)((1 << (t.depth + 1)) - 1);
        v *= norm;
        g = g * (norm * t.size);
        if (t.bias) g = g * (t.bias == 1 ? -2.0f * v : 2.0f * v);
        return g;
    }
    return mk3(0, 0, 0);
}

// Albedo of a material with its partial derivatives: rho = modulated colour
// (at most one modulator), d rho / d albedo, d rho / d(modulator colour),
// d rho / d(lambda) where every modulator size is scaled by lambda (at 1).
struct RhoD {
    c3 rho, d_alb, d_amp, d_lam;
    int mod;  // modulator index or -1
};

__device__ __forceinline__ RhoD diffuse_d(const yg_material* mats, const yg_modulator* mods, const yg_texture* texs,
                                          int mat, f3 P)
{
    const yg_material& m = mats[mat];
    RhoD r;
    const c3 a = mkc(m.color[0], m.color[1], m.color[2]);
    r.rho = a;
    r.d_alb = mkc(1, 1, 1);
    r.d_amp = mkc(0, 0, 0);
    r.d_lam = mkc(0, 0, 0);
    r.mod = -1;
    if (m.mod_count == 0) return r;
    const yg_modulator& md = mods[m.mod_first];
    r.mod = m.mod_first;
    r.rho = modulate(md, texs, a, P);
    const float A = md.color;
    if (!(A > 0)) return r;
    const yg_texture& t = texs[md.tex];
    const f3 tp = mod_texpt(md, P);
    const c3 tc = tex_color(t, tp);
    const c3 c21 = mkc(t.color2[0] - t.color1[0], t.color2[1] - t.color1[1], t.color2[2] - t.color1[2]);
    const float dF = dot(tex_float_grad(t, tp), tp);  // d getFloat / d lambda
    const c3 dtex = cscale(c21, dF);
    const bool inside = A <= 1.0f;
    switch (md.mode) {
    case YG_MOD_MUL:
        if (inside) {
            r.d_alb = cadd(cscale(tc, A), mkc(1 - A, 1 - A, 1 - A));
            r.d_amp = cmul(csub(tc, mkc(1, 1, 1)), a);
            r.d_lam = cmul(cscale(dtex, A), a);
        } else {
            r.d_alb = tc;
            r.d_lam = cmul(dtex, a);
        }
        break;
    case YG_MOD_MIX:
        if (inside) {
            r.d_alb = mkc(1 - A, 1 - A, 1 - A);
            r.d_amp = csub(tc, a);
            r.d_lam = cscale(dtex, A);
        } else {
            r.d_alb = mkc(0, 0, 0);
            r.d_lam = dtex;
        }
        break;
    case YG_MOD_ADD:
        r.d_amp = tc;
        r.d_lam = cscale(dtex, A);
        break;
    case YG_MOD_SUB:
        r.d_amp = cscale(tc, -1.0f);
        r.d_lam = cscale(dtex, -A);
        break;
    }
    return r;
}

