/*
 * CPU oracle: improved Perlin noise, fBm, turbulence and the procedural
 * textures built on them.  TEST INFRASTRUCTURE ONLY (see yaf_oracle.h).
 *
 * Transcendentals whose reference call is a float libm function (sinf, powf)
 * are evaluated as (float)f((double)x) here and on the device, so both sides
 * agree bit for bit; they differ from libm's float functions by at most one
 * unit in the last place.
 */
#include <stdlib.h>
#include <string.h>
#include "yo_internal.h"

/* The 256-entry permutation of noise.cc:118-127 (the 512-entry table is this
 * permutation twice).  Kept as data: the noise values depend on it exactly. */
static const char* PERM_HEX =
    "a2a0193bf8ebaaeef31c67281ded00de952edc3f3a82354d6cba36d0f60c7932"
    "d159f4088b63892fb8b49783f28f18c751146587482042a880b54013b2227e57"
    "bc7f6b9d864cc8db7cd5254e5a557450cdb37abbc3cbb6e2e4ecfd980b96d39e"
    "5ca164f18161e1c42472498c904b843438ab78ca1f01d79311c158a931f9446d"
    "bf339c5f0994a38506c69a1e7b461530272b1b713c5bd66f62ac4fc2c00eb123"
    "a7df47b0776905e9e6e776730ffe6e9b56ef12a537fcaed9038edd10b9cec98d"
    "da2abd68179fbed40accd2e8433d70b7027d99d80d608a042c3e92e5af5307e0"
    "29a6c5e3f5f74a41266a165e522d21adf091ffea54fa661a4539cf75a488fb5d";

static unsigned char HASH[512];
static int hash_ready = 0;

static int hexv(char c) { return (c <= '9') ? c - '0' : c - 'a' + 10; }

static void hash_init(void)
{
    if (hash_ready) return;
    for (int i = 0; i < 256; i++) {
        unsigned char v = (unsigned char)(hexv(PERM_HEX[2 * i]) * 16 + hexv(PERM_HEX[2 * i + 1]));
        HASH[i] = v;
        HASH[i + 256] = v;
    }
    hash_ready = 1;
}

void yo_perm(unsigned char* out)
{
    hash_init();
    memcpy(out, HASH, 512);
}

static inline float dsin(float x) { return (float)sin((double)x); }
static inline float dpow(float x, float y) { return (float)pow((double)x, (double)y); }

/* newPerlin_t::fade / grad, noise.h:33-40 */
static inline float fade(float t) { return t * t * t * (t * (t * 6 - 15) + 10); }
static inline float grad(int hash, float x, float y, float z)
{
    int h = hash & 15;
    float u = h < 8 ? x : y, v = h < 4 ? y : h == 12 || h == 14 ? x : z;
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}
static inline float lerpf(float t, float a, float b) { return a + t * (b - a); }

/* newPerlin_t::operator(), noise.cc:161-183 */
float yo_newperlin(float x, float y, float z)
{
    hash_init();
    float u = floorf(x), v = floorf(y), w = floorf(z);
    int X = ((int)u) & 255, Y = ((int)v) & 255, Z = ((int)w) & 255;
    x -= u; y -= v; z -= w;
    u = fade(x); v = fade(y); w = fade(z);
    int A = HASH[X] + Y, AA = HASH[A] + Z, AB = HASH[A + 1] + Z;
    int B = HASH[X + 1] + Y, BA = HASH[B] + Z, BB = HASH[B + 1] + Z;
    float nv = lerpf(w,
                     lerpf(v, lerpf(u, grad(HASH[AA], x, y, z), grad(HASH[BA], x - 1, y, z)),
                           lerpf(u, grad(HASH[AB], x, y - 1, z), grad(HASH[BB], x - 1, y - 1, z))),
                     lerpf(v, lerpf(u, grad(HASH[AA + 1], x, y, z - 1), grad(HASH[BA + 1], x - 1, y, z - 1)),
                           lerpf(u, grad(HASH[AB + 1], x, y - 1, z - 1),
                                 grad(HASH[BB + 1], x - 1, y - 1, z - 1))));
    return (float)(0.5 + 0.5 * (double)nv);
}

/* getSignedNoise, noise.h:245-248 */
static inline float signed_noise(float x, float y, float z) { return 2.0f * yo_newperlin(x, y, z) - 1.0f; }

/* fBm_t::operator(), noise.cc:638-650 */
float yo_fbm(float x, float y, float z, float H, float lac, float oct)
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

/* turbulence, noise.cc:808-819 (newPerlin_t adds no offset) */
float yo_turbulence(float x, float y, float z, int oct, float size, int hard)
{
    float val, amp = 1, sum = 0;
    x *= size; y *= size; z *= size;
    for (int i = 0; i <= oct; i++, amp *= 0.5f, x *= 2.0f, y *= 2.0f, z *= 2.0f) {
        val = yo_newperlin(x, y, z);
        if (hard) val = (float)fabs(2.0 * (double)val - 1.0);
        sum += amp * val;
    }
    return sum * ((float)(1 << oct) / (float)((1 << (oct + 1)) - 1));
}

/* texture getFloat: clouds basictex.cc:82-90, marble :133-153, wood :194-216,
 * Musgrave :677-680 */
static float tex_float(const yg_texture* t, yo_v3 p)
{
    switch (t->type) {
    case YG_TEX_CLOUDS: {
        float v = yo_turbulence(p.x, p.y, p.z, t->depth, t->size, t->hard);
        if (t->bias) {
            v *= v;
            if (t->bias == 1) return -v;
        }
        return v;
    }
    case YG_TEX_MARBLE:
    case YG_TEX_WOOD: {
        float w;
        if (t->type == YG_TEX_MARBLE) {
            w = (float)((double)(p.x + p.y + p.z) * 5.0 +
                        ((t->turb == 0.0f) ? 0.0 : (double)(t->turb * yo_turbulence(p.x, p.y, p.z, t->depth, t->size, t->hard))));
        } else {
            if (t->rings) w = (float)((double)sqrtf(p.x * p.x + p.y * p.y + p.z * p.z) * 20.0);
            else w = (float)((double)(p.x + p.y + p.z) * 10.0);
            w = (float)((double)w + ((t->turb == 0.0f) ? 0.0 : (double)(t->turb * yo_turbulence(p.x, p.y, p.z, t->depth, t->size, t->hard))));
        }
        switch (t->shape) {
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
        if (t->type == YG_TEX_MARBLE) return dpow(w, t->sharpness);
        return w;
    }
    case YG_TEX_MUSGRAVE:
        return t->iscale * yo_fbm(p.x * t->size, p.y * t->size, p.z * t->size, t->H, t->lacunarity,
                                  t->octaves);
    }
    return 0.0f;
}

/* getColor: color1 + getFloat(p)*(color2 - color1) (basictex.cc:92-95 etc.) */
static yo_col tex_color(const yg_texture* t, yo_v3 p)
{
    float f = tex_float(t, p);
    yo_col c1 = col(t->color1[0], t->color1[1], t->color1[2]);
    yo_col c2 = col(t->color2[0], t->color2[1], t->color2[2]);
    return col_add(c1, col_scale(col_sub(c2, c1), f));
}

/* modulator_t::modulate colour part (texture.cc:499-534) with doMapping's
 * procedural branch for global coordinates, projections x,y,z (:311-317). */
yo_col yo_modulate(const yg_modulator* m, const yg_texture* texs, yo_col C, yo_v3 P)
{
    const yg_texture* t = &texs[m->tex];
    yo_v3 tp = v3(m->size[0] * (P.x + m->ofs[0]), m->size[1] * (P.y + m->ofs[1]),
                  m->size[2] * (P.z + m->ofs[2]));
    yo_col tc = tex_color(t, tp);
    switch (m->mode) {
    case YG_MOD_MIX:
        if (m->color > 0) C = col_mix(tc, C, m->color);
        break;
    case YG_MOD_MUL:
        if (m->color > 0) C = col_mul(col_mix(tc, col(1, 1, 1), m->color), C);
        break;
    case YG_MOD_ADD:
        if (m->color > 0) C = col_add(col_scale(tc, m->color), C);
        break;
    case YG_MOD_SUB:
        if (m->color > 0) C = col_add(col_scale(tc, -m->color), C);
        break;
    }
    return C;
}

int yo_noise_eval(int what, const yg_texture* tex, const float* pts, int64_t n, float* out)
{
    for (int64_t i = 0; i < n; i++) {
        yo_v3 p = v3_ld(pts + 3 * i);
        switch (what) {
        case YG_EVAL_NOISE: out[i] = yo_newperlin(p.x, p.y, p.z); break;
        case YG_EVAL_FBM: out[i] = yo_fbm(p.x, p.y, p.z, tex->H, tex->lacunarity, tex->octaves); break;
        case YG_EVAL_TURBULENCE: out[i] = yo_turbulence(p.x, p.y, p.z, tex->depth, tex->size, tex->hard); break;
        case YG_EVAL_TEX_FLOAT: out[i] = tex_float(tex, p); break;
        case YG_EVAL_TEX_COLOR: {
            yo_col c = tex_color(tex, p);
            out[3 * i] = c.r; out[3 * i + 1] = c.g; out[3 * i + 2] = c.b;
            break;
        }
        default: return -1;
        }
    }
    return 0;
}

/* Heightfield: x = ex*i/(nx-1), y = ey*j/(ny-1), z = amp*fBm(freq*(x,y,0)) */
void yo_gen_heightfield(int nx, int ny, float ex, float ey, float amp, float freq, float H,
                        float lac, float oct, float* xyz)
{
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            float x = ex * (float)i / (float)(nx - 1);
            float y = ey * (float)j / (float)(ny - 1);
            float z = amp * yo_fbm(freq * x, freq * y, 0.0f, H, lac, oct);
            int64_t k = (int64_t)j * nx + i;
            xyz[3 * k] = x; xyz[3 * k + 1] = y; xyz[3 * k + 2] = z;
        }
}

/* ---------------------------------------------- FP64 texture derivatives -- */

static double fade_d(double t) { return 30.0 * t * t * (t - 1.0) * (t - 1.0); }
static double fade_dd(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }

static void grad_vec(int hash, double g[3])
{
    int h = hash & 15;
    double su = (h & 1) ? -1.0 : 1.0, sv = (h & 2) ? -1.0 : 1.0;
    g[0] = g[1] = g[2] = 0;
    if (h < 8) g[0] += su; else g[1] += su;
    if (h < 4) g[1] += sv;
    else if (h == 12 || h == 14) g[0] += sv;
    else g[2] += sv;
}

/* d newPerlin_t / d(x,y,z) in double: the noise is a trilinear blend (fade
 * weights) of the 8 corner gradient ramps, differentiated analytically */
static void perlin_grad_d(double x, double y, double z, double out[3])
{
    hash_init();
    double fx = floor(x), fy = floor(y), fz = floor(z);
    int X = ((int)fx) & 255, Y = ((int)fy) & 255, Z = ((int)fz) & 255;
    x -= fx; y -= fy; z -= fz;
    double u = fade_dd(x), v = fade_dd(y), w = fade_dd(z);
    double du = fade_d(x), dv = fade_d(y), dw = fade_d(z);
    int A = HASH[X] + Y, AA = HASH[A] + Z, AB = HASH[A + 1] + Z;
    int B = HASH[X + 1] + Y, BA = HASH[B] + Z, BB = HASH[B + 1] + Z;
    int hc[8] = {HASH[AA], HASH[BA], HASH[AB], HASH[BB], HASH[AA + 1], HASH[BA + 1], HASH[AB + 1], HASH[BB + 1]};
    double g[8], G[8][3];
    for (int k = 0; k < 8; k++) {
        double c[3] = {x - (k & 1), y - ((k >> 1) & 1), z - ((k >> 2) & 1)};
        grad_vec(hc[k], G[k]);
        g[k] = G[k][0] * c[0] + G[k][1] * c[1] + G[k][2] * c[2];
    }
    double k1 = g[1] - g[0], k2 = g[2] - g[0], k3 = g[4] - g[0];
    double k4 = g[0] - g[1] - g[2] + g[3], k5 = g[0] - g[2] - g[4] + g[6], k6 = g[0] - g[1] - g[4] + g[5];
    double k7 = -g[0] + g[1] + g[2] - g[3] + g[4] - g[5] - g[6] + g[7];
    double d[3] = {du * (k1 + k4 * v + k6 * w + k7 * v * w), dv * (k2 + k4 * u + k5 * w + k7 * u * w),
                   dw * (k3 + k5 * v + k6 * u + k7 * u * v)};
    for (int k = 0; k < 8; k++) {
        double wu = (k & 1) ? u : 1 - u, wv = ((k >> 1) & 1) ? v : 1 - v, ww = ((k >> 2) & 1) ? w : 1 - w;
        for (int a = 0; a < 3; a++) d[a] += G[k][a] * wu * wv * ww;
    }
    for (int a = 0; a < 3; a++) out[a] = 0.5 * d[a];
}

/* d getFloat(p) / d lambda for p scaled by lambda at 1 (= grad . p), for
 * Musgrave fBm and clouds turbulence */
double yo_tex_dF_dlambda(const yg_texture* t, const float* pp)
{
    double p[3] = {pp[0], pp[1], pp[2]}, g[3] = {0, 0, 0}, q[3];
    if (t->type == YG_TEX_MUSGRAVE) {
        double s = t->size, pwr = 1, scale = 1, pwHL = pow((double)t->lacunarity, -(double)t->H);
        for (int i = 0; i < (int)t->octaves; i++) {
            perlin_grad_d(p[0] * s * scale, p[1] * s * scale, p[2] * s * scale, q);
            for (int a = 0; a < 3; a++) g[a] += 2.0 * pwr * scale * q[a];
            pwr *= pwHL;
            scale *= t->lacunarity;
        }
        double rmd = t->octaves - floor(t->octaves);
        if (rmd != 0) {
            perlin_grad_d(p[0] * s * scale, p[1] * s * scale, p[2] * s * scale, q);
            for (int a = 0; a < 3; a++) g[a] += 2.0 * rmd * pwr * scale * q[a];
        }
        for (int a = 0; a < 3; a++) g[a] *= t->iscale * s;
    } else if (t->type == YG_TEX_CLOUDS) {
        double s = t->size, amp = 1, scale = 1, v = 0;
        for (int i = 0; i <= t->depth; i++, amp *= 0.5, scale *= 2) {
            double x = p[0] * s * scale, y = p[1] * s * scale, z = p[2] * s * scale;
            double n = yo_newperlin((float)x, (float)y, (float)z);
            perlin_grad_d(x, y, z, q);
            double f = scale;
            double val = n;
            if (t->hard) {
                val = fabs(2 * n - 1);
                f *= (2 * n - 1) >= 0 ? 2.0 : -2.0;
            }
            v += amp * val;
            for (int a = 0; a < 3; a++) g[a] += amp * f * q[a];
        }
        double norm = (double)(1 << t->depth) / (double)((1 << (t->depth + 1)) - 1);
        v *= norm;
        for (int a = 0; a < 3; a++) g[a] *= norm * s;
        if (t->bias) for (int a = 0; a < 3; a++) g[a] *= (t->bias == 1) ? -2.0 * v : 2.0 * v;
    }
    return g[0] * p[0] + g[1] * p[1] + g[2] * p[2];
}

void yo_tex_color_f(const yg_texture* t, const float* p, float* out)
{
    yo_col c = tex_color(t, v3_ld(p));
    out[0] = c.r; out[1] = c.g; out[2] = c.b;
}
