/*
 * CPU oracle: camera, generic diffuse shader, sun / area / path / photon
 * lights and the mono render loop.  TEST INFRASTRUCTURE ONLY (see yaf_oracle.h).
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "yo_internal.h"

yo_col yo_modulate(const yg_modulator* m, const yg_texture* texs, yo_col C, yo_v3 P);
yo_col yo_modulate_v(const yg_modulator* m, const yg_texture* texs, yo_col C, yo_v3 P, int vec);
void yo_mod_texpt(const yg_modulator* m, yo_v3 P, float out[3]);
yo_v3 yo_shading_normal(const yo_scene* s, int32_t prim, yo_v3 h);
int yo_first_hit_one(const yo_scene* s, yo_v3 from, yo_v3 ray, float min_raydis, int32_t* prim,
                     float* t);
int yo_is_shadowed_one(const yo_scene* s, yo_v3 p, yo_v3 l, int self_obj, int to_point,
                       float min_raydis, float self_bias);

static inline float dcos(float x) { return (float)cos((double)x); }
static inline float dsin(float x) { return (float)sin((double)x); }

/* surfacePoint_t subset (surface.h:64-90) */
typedef struct {
    yo_v3 P, N, Ng, NU, NV;
    float Z;
    int32_t prim, obj, mat;
    float sx, sy;  /* screenpos: set at camera-ray hits (scene.cc:234), else 0 */
} yo_sp;

/* area light precomputation, arealight.cc:31-91 */
typedef struct {
    int samples;
    yo_v3* points;
    yo_v3* jit_h;
    yo_v3* jit_v;
    yo_v3 direction, from, corner, toX, toY;
    yo_col color;
    float pow;
} yo_area;

/* photonSampler_t (pathtools.h:64-126) of a render area: the hemigrid of
 * paralels x meridians sectors, the current shading point's plan, the walk */
#define YO_PS_MAXSEC 64
typedef struct {
    int made;              /* created by this area's first takeSample */
    float radius;          /* its gather radius state */
    int par, mer, sectors, search;
    float pdiv, mdiv, multi;
    int hits[YO_PS_MAXSEC];
    float weight[YO_PS_MAXSEC];
    int cur[3];
} yo_psamp;

typedef struct {
    const yo_scene* s;
    const yg_material* mats;
    int nmat;
    const yg_modulator* mods;
    const yg_texture* texs;
    const yg_light* lights;
    int nl;
    const int32_t* obj_flags;
    int nobj;
    yg_render_opts opts;
    yo_area* areas;        /* per light (only for area lights) */
    yo_v3* sun_dir;        /* per light */
    yo_col* sun_col;
    yo_halton** hseq;      /* per light: path sampler state of the current render area */
    int32_t lcg;           /* global myseed */
    uint64_t stats[5];
    /* adjoint (FP64 reference gradient) */
    int adj;
    double gpx[3];         /* dLoss/dpixel of the pixel being rendered, or 0 */
    int gpx_on;
    int primary;           /* inside the camera ray's scene_t::light */
    double* gmat;          /* nmat*3 */
    double* gmod;          /* nmod*2 */
    double* glight;        /* nl */
    struct adj_path_s* ap; /* the current pixel's recorded path samples */
    int nap, apcap;
    float gm_radius;       /* photonData_t::radius of this render thread (pathlight.cc:131) */
    int gp;                /* the scene has a global photon light whose map is built */
    /* irradiance cache of a pathlight with cache = on (pathlight.cc:405-723) */
    struct yo_lcache_s* lc;
    int lc_phase;          /* 1: fake pass (cached), 2: render pass (interpolate) */
    float scr_x, scr_y;    /* renderState_t::screenpos of the pixel being rendered */
    int64_t pixnum;        /* renderState_t::pixelNumber */
    float traveled;        /* renderState_t::traveled at the camera ray's hit (scene.cc:239-240) */
    float worldres;        /* scene_t::world_resolution, (1.0 / resx) / focal (scene.cc:103) */
    /* cacheProxy_t of the render area (cacheproxy.cc:10-15, 160-172) */
    float px_radius;
    int64_t px_id;
    yo_psamp ps;           /* photonSampler_t of the current area (pathlight samples > 96 with a global map) */
    struct yo_ls_s* px_created;
    int px_n, px_cap;
    uint64_t lc_new;       /* interpolate() found no sample and took one */
    int32_t* lc_cand;      /* gather candidates (one per cache sample) */
    struct yo_grand_s* grand;  /* glibc rand() state of the render (after the block shuffle's draws) */
} rctx;

static void adj_path_sample_begin(rctx* R);
static void adj_path_vertex(rctx* R, int hit, const yo_sp* sp, yo_col lf, yo_col lfp);
static void adj_path_sample_end(rctx* R, float s);
static void adj_path_pixel(rctx* R, int li, const yo_sp* sp0, float mult);
static void adj_direct_sun(rctx* R, int li, const yo_sp* sp, float inte);

static int obj_flag(const rctx* R, int obj, int bit)
{
    if (!R->obj_flags || obj < 0 || obj >= R->nobj) return 1;
    return (R->obj_flags[obj] >> bit) & 1;
}

/* meshObject_t::shoot + triangle_t::getSurface (mesh.cc:517-577,
 * triangle.cc:138-160): P = from + Z*ray, Ng = face normal, N = Ng or the
 * interpolated vertex normal of an autoSmooth'ed mesh, NU/NV from createCS. */
static void make_sp(const yo_scene* s, int32_t prim, yo_v3 f, yo_v3 ray, float Z, yo_sp* sp)
{
    sp->prim = prim;
    sp->Z = Z;
    sp->P = v3_add(f, v3_mul(ray, Z));
    sp->Ng = v3_ld(s->tri_n + 3 * (int64_t)prim);
    sp->N = yo_shading_normal(s, prim, sp->P);
    create_cs(sp->N, &sp->NU, &sp->NV);
    sp->obj = s->tri_obj[prim];
    sp->mat = s->tri_mat[prim];
    sp->sx = sp->sy = 0.0f;
}

/* scene_t::firstHit (scene.cc:688-721); shadow: its shadow flag, which
 * skips objects that cast no shadows (the pathlight's rays, pathlight.cc:317) */
static int first_hit(rctx* R, yo_v3 from, yo_v3 ray, yo_sp* sp, int shadow)
{
    int32_t p; float t;
    if (!yo_first_hit_one_s(R->s, from, ray, R->opts.min_raydis, shadow, &p, &t)) return 0;
    yo_v3 f = v3_add(from, v3_mul(ray, R->opts.min_raydis));
    make_sp(R->s, p, f, ray, t, sp);
    return 1;
}

/* modulator_t::doMapping's coordinates (texture.cc:125-165) for texco at sp,
 * eye the shader call's eye vector; *vec = 1 for a direction.  uv and orco
 * come from triangle_t::getSurface (triangle.cc:138-167): a triangle with no
 * uv, vertex colours, vertex normals or orco points has u = v = -1 and
 * hasOrco() false; else getInterpolation's weights (fa = 1 - fb - fc) apply,
 * with u = v = 0 on a face without uv. */
static yo_v3 tex_coords(const rctx* R, int texco, const yo_sp* sp, yo_v3 eye, int* vec)
{
    const yo_scene* S = R->s;
    *vec = 0;
    switch (texco) {
    case YG_TEXCO_UV:
    case YG_TEXCO_ORCO: {
        const float* r = S->tri_tex ? S->tri_tex + 16 * (int64_t)sp->prim : NULL;
        const int fl = r ? (int)r[15] : 0;
        if (texco == YG_TEXCO_ORCO && !(fl & 4)) {
            /* toObjectOrco: backOrco * P (matrix4.h:86-91) */
            const float* M = S->obj_orco + 12 * (int64_t)sp->obj;
            const yo_v3 P = sp->P;
            return v3(M[0] * P.x + M[1] * P.y + M[2] * P.z + M[3], M[4] * P.x + M[5] * P.y + M[6] * P.z + M[7],
                      M[8] * P.x + M[9] * P.y + M[10] * P.z + M[11]);
        }
        const int smooth = S->tri_smooth && S->tri_smooth[sp->prim];
        float u = -1.0f, v = -1.0f;
        if (fl || smooth) {
            float fb, fc;
            const yo_v3 A = tri_vert(S, sp->prim, 0), B = tri_vert(S, sp->prim, 1), C = tri_vert(S, sp->prim, 2);
            const float a[3] = {A.x, A.y, A.z}, b[3] = {B.x, B.y, B.z}, c[3] = {C.x, C.y, C.z};
            const float n[3] = {sp->Ng.x, sp->Ng.y, sp->Ng.z}, h[3] = {sp->P.x, sp->P.y, sp->P.z};
            yo_tri_bary(a, b, c, n, h, &fb, &fc);
            const float fa = 1.f - fb - fc;
            if (texco == YG_TEXCO_ORCO)
                return v3_add(v3_add(v3_mul(v3(r[6], r[7], r[8]), fa), v3_mul(v3(r[9], r[10], r[11]), fb)),
                              v3_mul(v3(r[12], r[13], r[14]), fc));
            if (fl & 1) {
                u = r[0] * fa + r[2] * fb + r[4] * fc;
                v = r[1] * fa + r[3] * fb + r[5] * fc;
            } else
                u = v = 0.0f;
        }
        /* texpt.x = u*2.0 - 1.0, texpt.y = 1.0 - v*2.0 (double literals) */
        return v3((float)((double)u * 2.0 - 1.0), (float)(1.0 - (double)v * 2.0), 0.0f);
    }
    case YG_TEXCO_WIN:
        return v3(sp->sx, sp->sy, 0.0f);
    case YG_TEXCO_NOR:
        *vec = 1;
        return sp->N;
    case YG_TEXCO_REFL: {
        /* N = FACE_FORWARD(Ng, N, neye); reflect(N, neye) (vector3d.h:221-226) */
        *vec = 1;
        const yo_v3 ne = v3_norm(eye);
        const yo_v3 n = face_forward(sp->Ng, sp->N, ne);
        const float vn = v3_dot(ne, n);
        if (vn < 0) return v3_neg(ne);
        return v3_sub(v3_mul(n, 2.0f * vn), ne);
    }
    default:
        return sp->P;
    }
}

/* genericShader_t::getDiffuse, basicshaders.cc:191-204, with the call's eye */
static yo_col get_diffuse(const rctx* R, const yo_sp* sp, yo_v3 eye)
{
    const yg_material* m = &R->mats[sp->mat];
    yo_col c = col(m->color[0], m->color[1], m->color[2]);
    for (int k = 0; k < m->mod_count; k++) {
        const yg_modulator* md = &R->mods[m->mod_first + k];
        int vec;
        const yo_v3 T = tex_coords(R, md->texco, sp, eye, &vec);
        c = yo_modulate_v(md, R->texs, c, T, vec);
    }
    return c;
}

/* genericShader_t::fromLight, basicshaders.cc:46-70, diffuse only (ks = 0, edif = 1) */
static yo_col from_light(const rctx* R, const yo_sp* sp, yo_v3 edir_dir, yo_col ecol, yo_v3 eye)
{
    yo_v3 edir = v3_norm(eye);
    yo_v3 N = face_forward(sp->Ng, sp->N, edir);
    float inte = v3_dot(N, edir_dir);
    if (inte < 0) return col(0, 0, 0);
    yo_col lkc = get_diffuse(R, sp, eye);
    return col_mul(col_scale(ecol, inte), lkc);
}

/* genericShader_t::fromRadiosity, basicshaders.cc:30-44 */
static yo_col from_radiosity(const rctx* R, const yo_sp* sp, yo_v3 edir_dir, yo_col ecol, yo_v3 eye)
{
    if (v3_dot(face_forward(sp->Ng, sp->N, eye), edir_dir) < 0) return col(0, 0, 0);
    yo_col lkc = get_diffuse(R, sp, eye);
    return col_mul(ecol, lkc);
}

static int is_shadowed(rctx* R, const yo_sp* sp, yo_v3 l, int to_point)
{
    R->stats[2]++;
    return yo_is_shadowed_one(R->s, sp->P, l, sp->obj, to_point, R->opts.min_raydis,
                              R->opts.self_bias);
}

/* sunLight_t::illuminate, sunlight.cc:35-57 */
static yo_col sun_illuminate(rctx* R, int li, const yo_sp* sp, yo_v3 eye)
{
    const yg_light* L = &R->lights[li];
    if (L->cast_shadows && is_shadowed(R, sp, R->sun_dir[li], 0)) return col(0, 0, 0);
    if (R->adj && R->gpx_on && R->primary) {
        const float inte = v3_dot(face_forward(sp->Ng, sp->N, v3_norm(eye)), R->sun_dir[li]);
        if (inte >= 0) adj_direct_sun(R, li, sp, inte);
    }
    return from_light(R, sp, R->sun_dir[li], R->sun_col[li], eye);
}

/* areaLight_t::fillQuad, arealight.cc:31-73 */
static void area_setup(const yg_light* L, yo_area* A)
{
    yo_v3 a = v3_ld(L->a), b = v3_ld(L->b), c = v3_ld(L->c), d = v3_ld(L->d);
    int samp = L->samples;
    A->points = (yo_v3*)calloc((size_t)(samp > 0 ? samp : 1), sizeof(yo_v3));
    A->jit_h = (yo_v3*)calloc((size_t)(samp > 0 ? samp : 1), sizeof(yo_v3));
    A->jit_v = (yo_v3*)calloc((size_t)(samp > 0 ? samp : 1), sizeof(yo_v3));
    float ab_len = v3_len(v3_sub(b, a)), bc_len = v3_len(v3_sub(c, b));
    float cd_len = v3_len(v3_sub(d, c)), da_len = v3_len(v3_sub(a, d));
    float max_v = (bc_len > da_len) ? bc_len : da_len;
    float max_h = (ab_len > cd_len) ? ab_len : cd_len;
    int v_samp = (int)((double)(max_v / (max_v + max_h)) * 2.0 * (double)sqrtf((float)samp));
    int h_samp = (int)((double)(max_h / (max_v + max_h)) * 2.0 * (double)sqrtf((float)samp));
    int index = 0;
    if (samp == 1) index = 1;
    else {
        yo_v3 inc_v_a = v3_div(v3_sub(d, a), (float)v_samp);
        yo_v3 inc_v_b = v3_div(v3_sub(c, b), (float)v_samp);
        yo_v3 line_a = v3_add(a, v3_mul(inc_v_a, 0.5f));
        yo_v3 line_b = v3_add(b, v3_mul(inc_v_b, 0.5f));
        for (int i = 0; i < v_samp; ++i) {
            yo_v3 inc_h = v3_div(v3_sub(line_b, line_a), (float)h_samp);
            yo_v3 point = v3_add(line_a, v3_mul(inc_h, 0.5f));
            for (int j = 0; j < h_samp; ++j) {
                float vf = ((float)j) / ((float)h_samp);
                yo_v3 inc_v = v3_add(v3_mul(inc_v_a, (float)(1.0 - (double)vf)), v3_mul(inc_v_b, vf));
                if (index < samp) {
                    A->points[index] = point;
                    A->jit_h[index] = inc_h;
                    A->jit_v[index] = inc_v;
                }
                index++;
                point = v3_add(point, inc_h);
            }
            line_a = v3_add(line_a, inc_v_a);
            line_b = v3_add(line_b, inc_v_b);
        }
    }
    A->samples = index;
    A->direction = v3_norm(v3_cross(v3_sub(b, a), v3_sub(d, a)));
    A->color = col(L->color[0], L->color[1], L->color[2]);
    A->pow = L->power;
    A->from = v3_mul(v3_add(v3_add(v3_add(a, b), c), d), 0.25f);
    A->corner = a;
    A->toX = v3_sub(b, a);
    A->toY = v3_sub(d, a);
}

/* glibc random_r TYPE_3 one output at a time (the recurrence of
 * yo_glibc_rand: r[i] = r[i-31] + r[i-3], outputs r[i] >> 1 from i = 344) */
typedef struct yo_grand_s { uint32_t r[34]; int64_t i; } yo_grand;

static int32_t grand_next(yo_grand* g)
{
    const uint32_t v = g->r[(g->i - 31) % 34] + g->r[(g->i - 3) % 34];
    g->r[g->i % 34] = v;
    g->i++;
    return (int32_t)(v >> 1);
}

/* the generator after srand(seed) and `skip` rand() calls */
static void grand_init(yo_grand* g, uint32_t seed, int64_t skip)
{
    if (seed == 0) seed = 1;
    g->r[0] = seed;
    for (int i = 1; i < 31; i++) {
        int64_t v = (16807LL * (int32_t)g->r[i - 1]) % 2147483647LL;
        if (v < 0) v += 2147483647LL;
        g->r[i] = (uint32_t)v;
    }
    for (int i = 31; i < 34; i++) g->r[i] = g->r[i - 31];
    g->i = 34;
    for (int64_t k = 34; k < 344 + skip; k++) (void)grand_next(g);
}

/* Test hook: n outputs of rand() after srand(seed) and `skip` calls */
void yo_grand_stream(uint32_t seed, int64_t skip, int64_t n, int32_t* out)
{
    yo_grand g;
    grand_init(&g, seed, skip);
    for (int64_t k = 0; k < n; k++) out[k] = grand_next(&g);
}

/* areaLight_t::illuminate, arealight.cc:93-200 */
static yo_col area_illuminate(rctx* R, int li, const yo_sp* sp, yo_v3 eye)
{
    const yo_area* A = &R->areas[li];
    yo_col resul = col(0, 0, 0);
    yo_v3 edir = v3_norm(eye);
    yo_v3 N = face_forward(sp->Ng, sp->N, edir);
    float plane_at = v3_dot(v3_sub(sp->P, A->from), A->direction);
    if (plane_at <= 0) return col(0, 0, 0); /* fromLight of a zero colour */
    yo_col pc = col_scale(A->color, A->pow);
    if (A->samples == 1) {
        /* corner + toX*ourRandom() + toY*ourRandom(): g++ evaluates the right
         * operand of + first, so toY takes the first draw (tests/evalorder_check.cpp) */
        float r2 = our_random(&R->lcg);
        float r1 = our_random(&R->lcg);
        R->stats[4] += 2;
        yo_v3 sampleP = v3_add(v3_add(A->corner, v3_mul(A->toX, r1)), v3_mul(A->toY, r2));
        yo_v3 L = v3_sub(sampleP, sp->P);
        if (v3_dot(L, N) < 0) return col(0, 0, 0);
        if (!is_shadowed(R, sp, sampleP, 1)) {
            yo_v3 dir = L;
            float LD2 = v3_normlensqr(&dir);
            LD2 = (LD2 != 0) ? (float)(1.0 / (double)LD2) : 1.0f;
            resul = col_add(resul, from_light(R, sp, dir, col_scale(pc, LD2), eye));
        }
        return col_scale(resul, plane_at);
    }
    /* guessLight (arealight.cc:176-200): psamples draws of glibc rand() pick
     * sample points; lit and shadowed ones both seen: PENUMBRA, only lit:
     * LIGHT (every point, unjittered, no shadow ray), none lit: SHADOW */
    const int fsam = R->lights[li].psamples;
    if (fsam > 0) {
        int luz = 0, sombra = 0, status = 0;  /* 0 PENUMBRA, 1 LIGHT, 2 SHADOW */
        for (int i = 0; i < fsam; ++i) {
            if (luz && sombra) break;
            const int rs = grand_next(R->grand) % A->samples;
            const yo_v3 L = v3_sub(A->points[rs], sp->P);
            if (v3_dot(L, N) < 0) sombra = 1;
            else if (is_shadowed(R, sp, A->points[rs], 1)) sombra = 1;
            else luz = 1;
        }
        if (luz && sombra) status = 0;
        else if (luz) status = 1;
        else status = 2;
        if (status == 2) return col(0, 0, 0);  /* fromLight of a zero colour */
        if (status == 1) {
            for (int i = 0; i < A->samples; ++i) {
                yo_v3 dir = v3_sub(A->points[i], sp->P);
                float LD2 = v3_normlensqr(&dir);
                LD2 = (LD2 != 0) ? (float)(1.0 / (double)LD2) : 1.0f;
                resul = col_add(resul, from_light(R, sp, dir, col_scale(pc, LD2), eye));
            }
            yo_col r = col_scale(resul, plane_at);
            const float fsm = (float)A->samples;
            return col(r.r / fsm, r.g / fsm, r.b / fsm);
        }
    }
    for (int i = 0; i < A->samples; ++i) {
        float dish = (float)((double)our_random(&R->lcg) - 0.5);
        float disv = (float)((double)our_random(&R->lcg) - 0.5);
        R->stats[4] += 2;
        yo_v3 sampleP = v3_add(v3_add(A->points[i], v3_mul(A->jit_h[i], dish)), v3_mul(A->jit_v[i], disv));
        yo_v3 L = v3_sub(sampleP, sp->P);
        if (v3_dot(L, N) < 0) continue;
        if (!is_shadowed(R, sp, sampleP, 1)) {
            yo_v3 dir = L;
            float LD2 = v3_normlensqr(&dir);
            LD2 = (LD2 != 0) ? (float)(1.0 / (double)LD2) : 1.0f;
            resul = col_add(resul, from_light(R, sp, dir, col_scale(pc, LD2), eye));
        }
    }
    yo_col r = col_scale(resul, plane_at);
    float fs = (float)A->samples;
    return col(r.r / fs, r.g / fs, r.b / fs);
}

static yo_col photon_illuminate(rctx* R, int li, const yo_sp* sp, yo_v3 eye);
static yo_col path_illuminate(rctx* R, int li, const yo_sp* sp, yo_v3 eye);
static yo_col gm_irradiance(rctx* R, const yo_sp* sp, yo_v3 N);

/* scene_t::light, scene.cc:260-277 (fromWorld of a diffuse shader is 0) */
static yo_col scene_light(rctx* R, const yo_sp* sp, yo_v3 from, int indirect)
{
    yo_col fl = col(0, 0, 0);
    yo_v3 eye = v3_sub(from, sp->P);
    for (int li = 0; li < R->nl; li++) {
        const yg_light* L = &R->lights[li];
        if (!indirect && !L->use_in_render) continue;
        if (indirect && !L->use_in_indirect) continue;
        switch (L->type) {
        case YG_LIGHT_SUN: fl = col_add(fl, sun_illuminate(R, li, sp, eye)); break;
        case YG_LIGHT_AREA:
            /* a dummy area light only shoots global photons (arealight.cc:98) */
            if (!L->dummy) fl = col_add(fl, area_illuminate(R, li, sp, eye));
            break;
        case YG_LIGHT_PATH: fl = col_add(fl, path_illuminate(R, li, sp, eye)); break;
        case YG_LIGHT_PHOTON: fl = col_add(fl, photon_illuminate(R, li, sp, eye)); break;
        }
    }
    return fl;
}

/* haltonSampler_t::nextDirection, pathtools.cc:32-43 */
static yo_v3 next_direction(yo_halton* hs, int* taken, yo_v3 N, yo_v3 Ru, yo_v3 Rv, int cursam,
                            int curlev)
{
    if (cursam > *taken) *taken = cursam;
    curlev <<= 1;
    float z1 = halton_next(&hs[curlev]);
    float z2 = halton_next(&hs[curlev + 1]);
    if (z1 > 1.0f) z1 = 1.0f;
    z2 = (float)((double)z2 * (2.0 * M_PI));
    yo_v3 a = v3_add(v3_mul(Ru, dcos(z2)), v3_mul(Rv, dsin(z2)));
    return v3_add(v3_mul(a, (float)sqrt(1.0 - (double)z1)), v3_mul(N, sqrtf(z1)));
}

static yo_col background(const rctx* R)
{
    if (!R->opts.has_background) return col(0, 0, 0);
    return col(R->opts.background[0], R->opts.background[1], R->opts.background[2]);
}

static void ps_create(rctx* R, int li);
static void ps_sampling_from(rctx* R, yo_psamp* S, yo_v3 P, yo_v3 N, yo_v3 Ru, yo_v3 Rv, int samples);
static yo_v3 ps_next_direction(yo_psamp* S, yo_halton* hs, yo_v3 N, yo_v3 Ru, yo_v3 Rv, int curlev,
                               yo_col* raycolor);

/* pathLight_t::takeSample full-GI branch, pathlight.cc:223-387 (no cache, no
 * caustics, direct = false); getSampler (pathlight.cc:725-738) gives the
 * photonSampler_t when a global photon map is published and samples > 96 */
static yo_col path_take_sample(rctx* R, int li, yo_v3 N, const yo_sp* sp, float* avgD, float* minD)
{
    const yg_light* L = &R->lights[li];
    yo_halton* hs = R->hseq[li];
    int samples = L->samples, maxdepth = L->depth;
    const int caching = avgD != NULL;
    /* maxdist = 1000000 * world resolution * sp.Z (pathlight.cc:227), in float */
    const float maxdist = (float)1000000 * R->worldres * sp->Z;
    float HNUM = 0, HD = 0, M = 0;
    yo_col total = col(0, 0, 0), subtotal[4];
    for (int k = 0; k < 4; k++) subtotal[k] = col(0, 0, 0);
    int taken = 0, first = 1;
    float repetitions = 0.0f;
    yo_psamp* PS = NULL;
    if (R->gp && R->s->gm.rec && samples > 96) {
        PS = &R->ps;
        if (!PS->made) ps_create(R, li);
        ps_sampling_from(R, PS, sp->P, N, sp->NU, sp->NV, samples);
    }
    for (int ite = 1; ite > 0; --ite, repetitions += 1.0f) {
        taken = 0;
        if (PS) PS->cur[0] = PS->cur[1] = PS->cur[2] = 0;  /* photonSampler_t::reset */
        for (int i = 0; i < samples; ++i) {
            yo_col raycolor = col(1, 1, 1);
            yo_v3 ray = PS ? ps_next_direction(PS, hs, N, sp->NU, sp->NV, 0, &raycolor)
                           : next_direction(hs, &taken, N, sp->NU, sp->NV, i, 0);
            yo_v3 startray = ray;
            yo_v3 where = sp->P;
            if (caching) HNUM += 1;
            const int rec = R->adj && R->gpx_on && first;
            if (rec) adj_path_sample_begin(R);
            for (int j = 0; j < maxdepth; ++j) {
                if (col_energy(raycolor) < 0.05f) break;
                yo_sp t;
                R->stats[1]++;
                if (!first_hit(R, where, ray, &t, 1)) {
                    yo_col contri = col_mul(col_scale(raycolor, v3_dot(startray, N)), background(R));
                    total = col_add(total, contri);
                    if (first) subtotal[i % 4] = col_add(subtotal[i % 4], contri);
                    if (rec) adj_path_vertex(R, 0, NULL, col(0, 0, 0), col(0, 0, 0));
                    break;
                }
                if (caching) {
                    /* harmonic mean and minimum of the first hits' distances (pathlight.cc:323-327) */
                    if (j == 0 && t.Z > 0) HD = (float)((double)HD + 1.0 / (double)t.Z);
                    if (j == 0 && ((t.Z < M) || (M == 0))) M = t.Z;
                }
                /* getLight, pathlight.cc:186-191: sc.light(indirect) at P + eye */
                yo_v3 eye = v3_neg(ray);
                if (rec) {
                    /* sun factors of this vertex without the albedo (with / without power) */
                    yo_col lf = col(0, 0, 0), lfp = col(0, 0, 0);
                    yo_v3 e2 = v3_sub(v3_add(t.P, eye), t.P);
                    for (int l2 = 0; l2 < R->nl; l2++) {
                        const yg_light* Ls = &R->lights[l2];
                        if (Ls->type != YG_LIGHT_SUN || !Ls->use_in_indirect) continue;
                        if (Ls->cast_shadows &&
                            yo_is_shadowed_one(R->s, t.P, R->sun_dir[l2], t.obj, 0, R->opts.min_raydis,
                                               R->opts.self_bias))
                            continue;
                        float inte = v3_dot(face_forward(t.Ng, t.N, v3_norm(e2)), R->sun_dir[l2]);
                        if (inte < 0) continue;
                        lf = col_add(lf, col_scale(R->sun_col[l2], inte));
                        lfp = col_add(lfp, col_scale(col(Ls->color[0], Ls->color[1], Ls->color[2]), inte));
                    }
                    adj_path_vertex(R, 1, &t, lf, lfp);
                }
                yo_col light = scene_light(R, &t, v3_add(t.P, eye), 1);
                if (R->gp) {
                    /* getLight with the irradiance map (pathlight.cc:140-185):
                     * irradiance x diffuse, then the lights */
                    const yo_v3 Nv = face_forward(t.Ng, t.N, eye);
                    light = col_add(col_mul(gm_irradiance(R, &t, Nv), get_diffuse(R, &t, Nv)), light);
                }
                yo_col contri = col_mul(col_scale(raycolor, v3_dot(startray, N)), light);
                total = col_add(total, contri);
                if (first) subtotal[i % 4] = col_add(subtotal[i % 4], contri);
                yo_v3 HN = face_forward(t.Ng, t.N, v3_neg(ray));
                raycolor = col_mul(raycolor, get_diffuse(R, &t, v3_neg(ray)));
                ray = PS ? ps_next_direction(PS, hs, HN, t.NU, t.NV, j + 1, &raycolor)
                         : next_direction(hs, &taken, HN, t.NU, t.NV, i, j + 1);
                where = t.P;
            }
            if (rec) adj_path_sample_end(R, v3_dot(startray, N));
        }
        if (first) {
            first = 0;
            float mult = PS ? PS->multi : (float)(1.0 / (double)(float)(taken + 1));
            float minR = 1000, minG = 1000, minB = 1000, maxR = 0, maxG = 0, maxB = 0;
            for (int i = 0; i < 4; ++i) {
                subtotal[i] = col_scale(subtotal[i], mult);
                if (subtotal[i].r > maxR) maxR = subtotal[i].r;
                if (subtotal[i].g > maxG) maxG = subtotal[i].g;
                if (subtotal[i].b > maxB) maxB = subtotal[i].b;
                if (subtotal[i].r < minR) minR = subtotal[i].r;
                if (subtotal[i].g < minG) minG = subtotal[i].g;
                if (subtotal[i].b < minB) minB = subtotal[i].b;
            }
            float ene = fmaxf(fabsf(minR), fmaxf(fabsf(minG), fabsf(minB)));
            float dmax = fmaxf(fabsf(maxR - minR), fmaxf(fabsf(maxG - minG), fabsf(maxB - minB)));
            float diff = (float)((double)dmax / ((double)((ene < 1.0f) ? 1.0f : ene) * 2.0 * 2.0));
            diff /= L->threshold;
            int extra = (int)(diff * diff) - 1;
            if (extra > 3) extra = 3;
            if (extra < 0) extra = 0;
            if (extra > 0) R->stats[3]++;
            ite += extra;
        }
    }
    if (caching) {
        float H;
        if (HD > 0) H = HNUM / HD; else H = maxdist;
        if (M <= 0) M = maxdist;
        *avgD = H;
        *minD = M;
    }
    float mult = PS ? PS->multi : (float)(1.0 / (double)(float)(taken + 1));
    return col_scale(total, mult / repetitions);
}

static yo_col lc_cached(rctx* R, int li, const yo_sp* sp, yo_v3 eye);
static yo_col lc_interpolate(rctx* R, int li, const yo_sp* sp, yo_v3 eye);

/* pathLight_t::illuminate (pathlight.cc:107-119): the cache's fake pass
 * (cached), its render pass (interpolate), else normalSample
 * (pathlight.cc:389-403) */
static yo_col path_illuminate(rctx* R, int li, const yo_sp* sp, yo_v3 eye)
{
    const yg_light* L = &R->lights[li];
    if (L->cache && R->lc) return (R->lc_phase == 1) ? lc_cached(R, li, sp, eye) : lc_interpolate(R, li, sp, eye);
    yo_v3 N = face_forward(sp->Ng, sp->N, eye);
    yo_col dif = get_diffuse(R, sp, eye);
    if (col_energy(dif) < 0.05f) return col(0, 0, 0);
    const int prim = R->primary;
    R->primary = 0;  /* lights evaluated along the paths are not the pixel's direct terms */
    yo_col total = path_take_sample(R, li, N, sp, NULL, NULL);
    R->primary = prim;
    if (R->adj && R->gpx_on) adj_path_pixel(R, li, sp, (float)(1.0 / (double)(float)L->samples));
    return col_scale(col_mul(dif, total), L->power);
}

/* --------------------------------------------- FP64 reference gradient -- */

double yo_tex_dF_dlambda(const yg_texture* t, const float* p);
void yo_tex_color_f(const yg_texture* t, const float* p, float* out);

#define ADJ_MAXD 8
typedef struct adj_path_s {
    int n, hit[ADJ_MAXD], mat[ADJ_MAXD];
    yo_v3 P[ADJ_MAXD];
    yo_col lf[ADJ_MAXD], lfp[ADJ_MAXD];
    float s;
} adj_path;

/* albedo with its partial derivatives, in double (one modulator at most):
 * rho, d rho/d albedo, d rho/d modulator colour, d rho/d lambda */
static int diffuse_d(const rctx* R, int mat, yo_v3 P, double rho[3], double dalb[3], double damp[3],
                     double dlam[3])
{
    const yg_material* m = &R->mats[mat];
    double a[3] = {m->color[0], m->color[1], m->color[2]};
    for (int c = 0; c < 3; c++) { rho[c] = a[c]; dalb[c] = 1; damp[c] = 0; dlam[c] = 0; }
    if (m->mod_count == 0) return -1;
    const yg_modulator* md = &R->mods[m->mod_first];
    const yg_texture* t = &R->texs[md->tex];
    float tp[3];
    yo_mod_texpt(md, P, tp);
    float tcf[3];
    yo_tex_color_f(t, tp, tcf);
    yo_col rc = yo_modulate(md, R->texs, col(m->color[0], m->color[1], m->color[2]), P);
    rho[0] = rc.r; rho[1] = rc.g; rho[2] = rc.b;
    double A = md->color;
    if (!(A > 0)) return m->mod_first;
    double dF = yo_tex_dF_dlambda(t, tp);
    double c21[3] = {t->color2[0] - t->color1[0], t->color2[1] - t->color1[1], t->color2[2] - t->color1[2]};
    int inside = A <= 1.0;
    for (int c = 0; c < 3; c++) {
        double tc = tcf[c], dtex = c21[c] * dF;
        switch (md->mode) {
        case YG_MOD_MUL:
            if (inside) { dalb[c] = tc * A + (1 - A); damp[c] = (tc - 1) * a[c]; dlam[c] = A * dtex * a[c]; }
            else { dalb[c] = tc; dlam[c] = dtex * a[c]; }
            break;
        case YG_MOD_MIX:
            if (inside) { dalb[c] = 1 - A; damp[c] = tc - a[c]; dlam[c] = A * dtex; }
            else { dalb[c] = 0; dlam[c] = dtex; }
            break;
        case YG_MOD_ADD: damp[c] = tc; dlam[c] = A * dtex; break;
        case YG_MOD_SUB: damp[c] = -tc; dlam[c] = -A * dtex; break;
        }
    }
    return m->mod_first;
}

static void adj_vertex_add(rctx* R, int mat, yo_v3 P, const double gamma[3])
{
    double rho[3], dalb[3], damp[3], dlam[3];
    int mod = diffuse_d(R, mat, P, rho, dalb, damp, dlam);
    for (int c = 0; c < 3; c++) R->gmat[3 * mat + c] += gamma[c] * dalb[c];
    if (mod >= 0) {
        for (int c = 0; c < 3; c++) {
            R->gmod[2 * mod] += gamma[c] * damp[c];
            R->gmod[2 * mod + 1] += gamma[c] * dlam[c];
        }
    }
}

static void adj_path_sample_begin(rctx* R)
{
    if (R->nap == R->apcap) {
        R->apcap = R->apcap ? 2 * R->apcap : 64;
        R->ap = (adj_path*)realloc(R->ap, sizeof(adj_path) * (size_t)R->apcap);
    }
    R->ap[R->nap].n = 0;
}

static void adj_path_vertex(rctx* R, int hit, const yo_sp* sp, yo_col lf, yo_col lfp)
{
    adj_path* a = &R->ap[R->nap];
    if (a->n >= ADJ_MAXD) return;
    a->hit[a->n] = hit;
    if (hit) { a->mat[a->n] = sp->mat; a->P[a->n] = sp->P; a->lf[a->n] = lf; a->lfp[a->n] = lfp; }
    a->n++;
}

static void adj_path_sample_end(rctx* R, float s)
{
    R->ap[R->nap].s = s;
    R->nap++;
}

/* product rule over the recorded samples of one pixel (see yg_render_adjoint) */
static void adj_path_pixel(rctx* R, int li, const yo_sp* sp0, float mult)
{
    const yg_light* L = &R->lights[li];
    double pm = (double)mult * (double)L->power;
    double rho0[3], d0a[3], d0m[3], d0l[3];
    diffuse_d(R, sp0->mat, sp0->P, rho0, d0a, d0m, d0l);
    double bg[3] = {0, 0, 0};
    if (R->opts.has_background) for (int c = 0; c < 3; c++) bg[c] = R->opts.background[c];
    double T[3] = {0, 0, 0}, dpow[3] = {0, 0, 0};
    double W[3];
    for (int c = 0; c < 3; c++) W[c] = R->gpx[c] * rho0[c] * pm;
    int sun = -1;
    for (int l2 = 0; l2 < R->nl; l2++)
        if (R->lights[l2].type == YG_LIGHT_SUN && R->lights[l2].use_in_indirect) { sun = l2; break; }
    for (int q = 0; q < R->nap; q++) {
        const adj_path* a = &R->ap[q];
        double rv[ADJ_MAXD][3];
        for (int b = 0; b < a->n; b++)
            if (a->hit[b]) {
                double da[3], dm[3], dl[3];
                diffuse_d(R, a->mat[b], a->P[b], rv[b], da, dm, dl);
            }
        double Rb[3] = {1, 1, 1};
        for (int b = 0; b < a->n; b++) {
            for (int c = 0; c < 3; c++) {
                double X = a->hit[b] ? (double)(&a->lf[b].r)[c] * rv[b][c] : bg[c];
                T[c] += a->s * Rb[c] * X;
                if (a->hit[b]) dpow[c] += a->s * Rb[c] * (double)(&a->lfp[b].r)[c] * rv[b][c];
            }
            if (a->hit[b]) for (int c = 0; c < 3; c++) Rb[c] *= rv[b][c];
        }
        for (int k = 0; k < a->n; k++) {
            if (!a->hit[k]) continue;
            double dk[3] = {0, 0, 0}, Rex[3] = {1, 1, 1};
            for (int b = 0; b < a->n; b++) {
                for (int c = 0; c < 3; c++) {
                    if (b == k) dk[c] += a->s * Rex[c] * (double)(&a->lf[b].r)[c];
                    else if (b > k) {
                        double X = a->hit[b] ? (double)(&a->lf[b].r)[c] * rv[b][c] : bg[c];
                        dk[c] += a->s * Rex[c] * X;
                    }
                }
                if (a->hit[b] && b != k) for (int c = 0; c < 3; c++) Rex[c] *= rv[b][c];
            }
            double gam[3] = {W[0] * dk[0], W[1] * dk[1], W[2] * dk[2]};
            adj_vertex_add(R, a->mat[k], a->P[k], gam);
        }
    }
    double g0[3];
    for (int c = 0; c < 3; c++) g0[c] = R->gpx[c] * T[c] * pm;
    adj_vertex_add(R, sp0->mat, sp0->P, g0);
    if (sun >= 0) R->glight[sun] += W[0] * dpow[0] + W[1] * dpow[1] + W[2] * dpow[2];
    R->nap = 0;
}

static void adj_direct_sun(rctx* R, int li, const yo_sp* sp, float inte)
{
    double rho[3], da[3], dm[3], dl[3];
    diffuse_d(R, sp->mat, sp->P, rho, da, dm, dl);
    const yg_light* L = &R->lights[li];
    double gam[3], dp = 0;
    for (int c = 0; c < 3; c++) {
        double colp = (double)(&R->sun_col[li].r)[c];
        gam[c] = R->gpx[c] * (double)(inte * colp);
        dp += R->gpx[c] * (double)inte * (double)L->color[c] * rho[c];
    }
    adj_vertex_add(R, sp->mat, sp->P, gam);
    R->glight[li] += dp;
}

/* ------------------------------------------------------------- photons -- */

/* randomVectorCone, vector3d.cc:162-168 */
static yo_v3 random_vector_cone(yo_v3 D, yo_v3 U, yo_v3 V, float cosang, float z1, float z2)
{
    float t1 = (float)(2.0 * M_PI * (double)z1);
    float t2 = (float)(1.0 - (1.0 - (double)cosang) * (double)z2);
    yo_v3 a = v3_add(v3_mul(U, dcos(t1)), v3_mul(V, dsin(t1)));
    return v3_add(v3_mul(a, (float)sqrt(1.0 - (double)(t2 * t2))), v3_mul(D, t2));
}

typedef struct {
    int32_t cx, cy, cz;
    int64_t seq;
    yo_v3 dir, pos;
    yo_col c;
} yo_mark;

typedef struct {
    yo_mark* v;
    int64_t n, cap;
} markvec;

static void mark_push(markvec* m, yo_mark x)
{
    if (m->n == m->cap) {
        m->cap = m->cap ? 2 * m->cap : 1024;
        m->v = (yo_mark*)realloc(m->v, sizeof(yo_mark) * (size_t)m->cap);
    }
    m->v[m->n++] = x;
}

/* hash3d_t::getBox (hash3d.h:218-228): truncation toward zero, minus one below 0 */
static void hash_cell(float cell, yo_v3 p, int32_t* nx, int32_t* ny, int32_t* nz)
{
    *nx = (int32_t)(p.x / cell);
    *ny = (int32_t)(p.y / cell);
    *nz = (int32_t)(p.z / cell);
    if (p.x < 0.0f) (*nx)--;
    if (p.y < 0.0f) (*ny)--;
    if (p.z < 0.0f) (*nz)--;
}

/* findBox = findCreateBox(getBox(p)) (hash3d.h:41-42): the cell of the cell centre */
static void find_box(float cell, yo_v3 p, int32_t* bx, int32_t* by, int32_t* bz)
{
    int32_t nx, ny, nz;
    hash_cell(cell, p, &nx, &ny, &nz);
    yo_v3 c = v3((float)((double)(cell * (float)nx) + 0.5 * (double)cell),
                 (float)((double)(cell * (float)ny) + 0.5 * (double)cell),
                 (float)((double)(cell * (float)nz) + 0.5 * (double)cell));
    hash_cell(cell, c, bx, by, bz);
}

typedef struct {
    rctx* R;
    const yg_light* L;
    yo_halton* hs;
    int depth;
    markvec* marks;
    int64_t photon_idx;
    uint64_t stored;
} pctx;

typedef struct { yo_v3 pos, lastpos; yo_col c; } yo_photon;

/* photonLight_t::shoot_photon_diffuse, photonlight.cc:203-244 */
static void shoot_diffuse(pctx* P, yo_photon* ph, yo_v3 dir)
{
    rctx* R = P->R;
    P->depth++;
    yo_sp sp;
    R->stats[1]++;
    if (!first_hit(R, ph->pos, dir, &sp, 0)) { P->depth--; return; }
    /* photon_t::position, photonlight.cc:39-47 */
    float ad = v3_len(v3_sub(ph->pos, sp.P));
    if (ad > P->L->bias) ph->lastpos = ph->pos;
    ph->pos = sp.P;
    yo_v3 edir = v3_norm(v3_sub(ph->lastpos, ph->pos));
    yo_v3 Ng = face_forward(sp.Ng, sp.Ng, edir);
    int canreceive = (P->depth > P->L->mindepth) && obj_flag(R, sp.obj, 1);
    if (canreceive) {
        yo_mark mk;
        mk.pos = ph->pos;
        mk.c = ph->c;
        mk.dir = v3_norm(v3_sub(ph->lastpos, ph->pos));
        find_box(P->L->cluster, mk.pos, &mk.cx, &mk.cy, &mk.cz);
        mk.seq = (int64_t)P->marks->n;
        mark_push(P->marks, mk);
        P->stored++;
    }
    if (obj_flag(R, sp.obj, 2) && P->depth <= P->L->depth) {
        edir = v3_norm(edir);
        int d2 = P->depth << 1;
        float r1 = halton_next(&P->hs[d2]);
        float r2 = halton_next(&P->hs[d2 + 1]);
        yo_v3 refDir = random_vector_cone(Ng, sp.NU, sp.NV, 0.05f, r1, r2);
        ph->c = from_radiosity(R, &sp, edir, ph->c, refDir);
        shoot_diffuse(P, ph, refDir);
    }
    P->depth--;
}

static int mark_cmp(const void* a, const void* b)
{
    const yo_mark* x = (const yo_mark*)a;
    const yo_mark* y = (const yo_mark*)b;
    if (x->cx != y->cx) return x->cx < y->cx ? -1 : 1;
    if (x->cy != y->cy) return x->cy < y->cy ? -1 : 1;
    if (x->cz != y->cz) return x->cz < y->cz ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

static int photon_np(const yg_light* L)
{
    int np = L->photons / (L->depth + 1 - L->mindepth);
    int r = (int)sqrtf((float)np);
    return r * r;
}

static void rctx_init(rctx* R, yo_scene* s, const yg_material* mats, int nmat,
                      const yg_modulator* mods, const yg_texture* texs, const yg_light* lights,
                      int nl, const int32_t* obj_flags, int nobj, const yg_render_opts* opts)
{
    memset(R, 0, sizeof(*R));
    R->s = s; R->mats = mats; R->nmat = nmat; R->mods = mods; R->texs = texs;
    s->oflags = obj_flags;  /* the scene's shadow queries see the render's object flags */
    s->noflags = nobj;
    R->lights = lights; R->nl = nl; R->obj_flags = obj_flags; R->nobj = nobj;
    R->opts = *opts;
    R->lcg = (int32_t)opts->lcg_seed;
    /* the global photon pass ran in init, before the render: the render's
     * ourRandom() draws continue after the photons' (when the pass ran here) */
    for (int li = 0; li < nl; li++)
        if (lights[li].type == YG_LIGHT_GLOBALPHOTON && s->gm.built && s->gm.own_stream) R->lcg = s->gm.lcg_after;
    R->gm_radius = s->gm.radius;
    for (int li = 0; li < nl; li++)
        if (lights[li].type == YG_LIGHT_GLOBALPHOTON && s->gm.built) R->gp = 1;
    R->areas = (yo_area*)calloc((size_t)(nl > 0 ? nl : 1), sizeof(yo_area));
    R->sun_dir = (yo_v3*)calloc((size_t)(nl > 0 ? nl : 1), sizeof(yo_v3));
    R->sun_col = (yo_col*)calloc((size_t)(nl > 0 ? nl : 1), sizeof(yo_col));
    R->hseq = (yo_halton**)calloc((size_t)(nl > 0 ? nl : 1), sizeof(yo_halton*));
    for (int li = 0; li < nl; li++) {
        const yg_light* L = &lights[li];
        if (L->type == YG_LIGHT_SUN) {
            /* sunLight_t ctor, sunlight.cc:27-33 */
            R->sun_dir[li] = v3_norm(v3_ld(L->dir));
            R->sun_col[li] = col_scale(col(L->color[0], L->color[1], L->color[2]), L->power);
        } else if (L->type == YG_LIGHT_AREA) {
            area_setup(L, &R->areas[li]);
        } else if (L->type == YG_LIGHT_PATH) {
            int nd = (L->depth + 1) * 2;
            R->hseq[li] = (yo_halton*)calloc((size_t)nd, sizeof(yo_halton));
        }
    }
}

static void rctx_free(rctx* R)
{
    for (int li = 0; li < R->nl; li++) {
        free(R->areas[li].points); free(R->areas[li].jit_h); free(R->areas[li].jit_v);
        free(R->hseq[li]);
    }
    free(R->areas); free(R->sun_dir); free(R->sun_col); free(R->hseq);
    free(R->ap);
    free(R->px_created);
    free(R->lc_cand);
}

static void grid_build(yo_photonmap* pm, float radius)
{
    free(pm->gstart); free(pm->gitems);
    pm->gstart = NULL; pm->gitems = NULL;
    if (pm->n == 0) return;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < pm->n; i++)
        for (int k = 0; k < 3; k++) {
            float v = pm->data[9 * i + 3 + k];
            if (v < lo[k]) lo[k] = v;
            if (v > hi[k]) hi[k] = v;
        }
    pm->gcell = radius > 0 ? radius : 1.0f;
    int64_t nc = 1;
    for (int k = 0; k < 3; k++) {
        pm->gmin[k] = lo[k];
        int64_t d = (int64_t)((hi[k] - lo[k]) / pm->gcell) + 1;
        if (d > 1024) d = 1024;
        pm->gdim[k] = (int)d;
        nc *= d;
    }
    pm->gstart = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
    pm->gitems = (int64_t*)malloc(sizeof(int64_t) * (size_t)pm->n);
    int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * (size_t)pm->n);
    for (int64_t i = 0; i < pm->n; i++) {
        int64_t c = 0;
        for (int k = 2; k >= 0; k--) {
            int64_t g = (int64_t)((pm->data[9 * i + 3 + k] - pm->gmin[k]) / pm->gcell);
            if (g < 0) g = 0;
            if (g >= pm->gdim[k]) g = pm->gdim[k] - 1;
            c = c * pm->gdim[k] + g;
        }
        cell[i] = c;
        pm->gstart[c + 1]++;
    }
    for (int64_t c = 0; c < nc; c++) pm->gstart[c + 1] += pm->gstart[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)nc);
    memcpy(fill, pm->gstart, sizeof(int64_t) * (size_t)nc);
    for (int64_t i = 0; i < pm->n; i++) pm->gitems[fill[cell[i]]++] = i;
    free(fill); free(cell);
}

/* photonLight_t::init, photonlight.cc:453-501 (+ preGathering :397-411) */
int yo_photon_build(yo_scene* s, const yg_material* mats, int nmat, const yg_modulator* mods,
                    int nmod, const yg_texture* texs, int ntex, const yg_light* lights, int nl,
                    int li, const int32_t* obj_flags, int nobj, const yg_render_opts* opts,
                    uint64_t* stats)
{
    (void)nmod; (void)ntex;
    if (li < 0 || li >= nl || lights[li].type != YG_LIGHT_PHOTON) return -1;
    rctx R;
    rctx_init(&R, s, mats, nmat, mods, texs, lights, nl, obj_flags, nobj, opts);
    const yg_light* L = &lights[li];
    int Np = photon_np(L);
    float angle_cos = (float)cos(((double)L->angle / 180.0) * M_PI);
    float pw = (float)((double)L->power * 2.0 * M_PI * (1.0 - (double)angle_cos));
    yo_col pcol = col_scale(col(L->color[0], L->color[1], L->color[2]), pw);
    int nd = (L->depth + 1) * 2;
    yo_halton* hs = (yo_halton*)calloc((size_t)nd, sizeof(yo_halton));
    int base = 2;
    for (int i = 0; i < nd; i++) { halton_init(&hs[i], base); base = yo_next_prime(base); }
    yo_v3 from = v3_ld(L->from), to = v3_ld(L->to);
    yo_v3 light_dir = v3_norm(v3_sub(to, from));
    yo_v3 LU, LV;
    create_cs(light_dir, &LU, &LV);
    markvec marks = {0};
    pctx P;
    P.R = &R; P.L = L; P.hs = hs; P.depth = 0; P.marks = &marks; P.stored = 0;
    int64_t emitted = 0;
    for (int64_t i = 0; emitted < Np; ++i) {
        yo_photon ph;
        ph.c = pcol; ph.pos = from; ph.lastpos = v3(0, 0, 0);
        float r1 = halton_next(&hs[0]);
        float r2 = halton_next(&hs[1]);
        yo_v3 dir = random_vector_cone(light_dir, LU, LV, angle_cos, r1, r2);
        if (v3_null(dir)) continue;
        P.photon_idx = emitted;
        shoot_diffuse(&P, &ph, dir);
        emitted++;
    }
    /* hash3d_t cells iterate in (x, y, z) order; sums follow insertion order */
    qsort(marks.v, (size_t)marks.n, sizeof(yo_mark), mark_cmp);
    yo_photonmap* pm = &s->pm;
    free(pm->data);
    pm->data = (float*)malloc(sizeof(float) * 9 * (size_t)(marks.n > 0 ? marks.n : 1));
    pm->n = 0;
    pm->cell = L->cluster;
    pm->np_emit = Np;
    for (int64_t a = 0; a < marks.n;) {
        int64_t b = a;
        yo_v3 dsum = v3(0, 0, 0), psum = v3(0, 0, 0);
        yo_col csum = col(0, 0, 0);
        float count = 0;
        while (b < marks.n && marks.v[b].cx == marks.v[a].cx && marks.v[b].cy == marks.v[a].cy &&
               marks.v[b].cz == marks.v[a].cz) {
            dsum = v3_add(dsum, marks.v[b].dir);
            psum = v3_add(psum, marks.v[b].pos);
            csum = col_add(csum, marks.v[b].c);
            count += 1;
            b++;
        }
        if (!v3_null(dsum)) {
            yo_v3 d = v3_norm(dsum);
            yo_v3 p = v3_div(psum, count);
            float* o = pm->data + 9 * pm->n;
            o[0] = d.x; o[1] = d.y; o[2] = d.z; o[3] = p.x; o[4] = p.y; o[5] = p.z;
            o[6] = csum.r; o[7] = csum.g; o[8] = csum.b;
            pm->n++;
        }
        a = b;
    }
    grid_build(pm, L->fixedradius);
    if (stats) {
        stats[0] = (uint64_t)emitted; stats[1] = P.stored; stats[2] = (uint64_t)pm->n;
        stats[3] = R.stats[1];
    }
    free(marks.v);
    free(hs);
    rctx_free(&R);
    return 0;
}

int yo_photon_load(yo_scene* s, const float* data, int64_t n, const yg_light* L)
{
    if (!L || L->type != YG_LIGHT_PHOTON || n < 0) return -1;
    yo_photonmap* pm = &s->pm;
    free(pm->data);
    pm->data = (float*)malloc(sizeof(float) * 9 * (size_t)(n > 0 ? n : 1));
    if (n > 0) memcpy(pm->data, data, sizeof(float) * 9 * (size_t)n);
    pm->n = n;
    pm->cell = L->cluster;
    pm->np_emit = photon_np(L);
    grid_build(pm, L->fixedradius);
    return 0;
}

int64_t yo_photon_read(yo_scene* s, float* out, int64_t cap)
{
    int64_t n = s->pm.n < cap ? s->pm.n : cap;
    if (out && n > 0) memcpy(out, s->pm.data, sizeof(float) * 9 * (size_t)n);
    return s->pm.n;
}

typedef struct { float dis; int64_t idx; } yo_found;

static int found_cmp(const void* a, const void* b)
{
    const yo_found* x = (const yo_found*)a;
    const yo_found* y = (const yo_found*)b;
    if (x->dis != y->dis) return x->dis < y->dis ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* photonLight_t::illuminate, photonlight.cc:297-395.  The K nearest photons
 * (dir.N > 0, dis <= fixedRadius) are summed in increasing (dis, index) order;
 * the reference sums in heap order, which changes only float rounding. */
static yo_col photon_estimate(const rctx* R, const yg_light* L, const yo_sp* sp, yo_v3 eye)
{
    const yo_photonmap* pm = &R->s->pm;
    if (!obj_flag(R, sp->obj, 1)) return col(0, 0, 0);
    yo_v3 N = face_forward(sp->Ng, sp->N, eye);
    int K = L->search;
    yo_found* cand = NULL;
    int64_t nc = 0, cap = 0;
    if (pm->n > 0 && pm->gstart) {
        int g0[3], g1[3];
        float pp[3] = {sp->P.x, sp->P.y, sp->P.z};
        for (int k = 0; k < 3; k++) {
            int64_t a = (int64_t)floorf((pp[k] - L->fixedradius - pm->gmin[k]) / pm->gcell);
            int64_t b = (int64_t)floorf((pp[k] + L->fixedradius - pm->gmin[k]) / pm->gcell);
            if (a < 0) a = 0;
            if (b >= pm->gdim[k]) b = pm->gdim[k] - 1;
            g0[k] = (int)a; g1[k] = (int)b;
        }
        for (int z = g0[2]; z <= g1[2]; z++)
            for (int y = g0[1]; y <= g1[1]; y++)
                for (int x = g0[0]; x <= g1[0]; x++) {
                    int64_t c = ((int64_t)z * pm->gdim[1] + y) * pm->gdim[0] + x;
                    for (int64_t q = pm->gstart[c]; q < pm->gstart[c + 1]; q++) {
                        int64_t i = pm->gitems[q];
                        const float* ph = pm->data + 9 * i;
                        if (v3_dot(v3(ph[0], ph[1], ph[2]), N) <= 0.0f) continue;
                        float dis = v3_len(v3_sub(sp->P, v3(ph[3], ph[4], ph[5])));
                        if (dis <= L->fixedradius) {
                            if (nc == cap) {
                                cap = cap ? 2 * cap : 256;
                                cand = (yo_found*)realloc(cand, sizeof(yo_found) * (size_t)cap);
                            }
                            cand[nc].dis = dis; cand[nc].idx = i; nc++;
                        }
                    }
                }
    }
    qsort(cand, (size_t)nc, sizeof(yo_found), found_cmp);
    if (nc > K) nc = K;
    if (nc < 2) { free(cand); return col(0, 0, 0); }
    const float radius = cand[nc - 1].dis;
    yo_col total = col(0, 0, 0);
    yo_v3 VX, VY;
    create_cs(N, &VX, &VY);
    float avgX = 0, avgY = 0, den = 0;
    float* sx = (float*)malloc(sizeof(float) * (size_t)nc);
    float* sy = (float*)malloc(sizeof(float) * (size_t)nc);
    float* sw = (float*)malloc(sizeof(float) * (size_t)nc);
    for (int64_t k = 0; k < nc; k++) {
        const float* ph = pm->data + 9 * cand[k].idx;
        const float dis = cand[k].dis;
        /* filterCone, photonlight.h:205-210 */
        float filter = (radius == 0) ? 1.0f : (float)((double)(float)(1.0 - (double)(dis / radius)) * 2.0);
        yo_v3 pdir = v3(ph[0], ph[1], ph[2]);
        total = col_add(total, from_light(R, sp, pdir, col_scale(col(ph[6], ph[7], ph[8]), filter), eye));
        yo_v3 sub = v3_div(v3_sub(v3(ph[3], ph[4], ph[5]), sp->P), radius);
        sx[k] = v3_dot(sub, VX);
        sy[k] = v3_dot(sub, VY);
        sw[k] = (float)((double)(v3_dot(pdir, N) * filter) * 0.5);
    }
    float area = (float)(M_PI * (double)radius * (double)radius);
    /* fillArea, photonlight.cc:259-290 */
    float fa = 1.0f;
    {
        const float SQRT2 = (float)sqrt(2.0);
        const float offset = (float)1.414213562373095;
        for (int64_t k = 0; k < nc; k++) { avgX += sx[k] * sw[k]; avgY += sy[k] * sw[k]; den += sw[k]; }
        if (den != 0.0f) {
            float avgR = 0;
            avgX /= den; avgY /= den;
            for (int64_t k = 0; k < nc; k++) {
                float dx = sx[k] - avgX, dy = sy[k] - avgY;
                avgR += sqrtf(dx * dx + dy * dy) * sw[k];
            }
            avgR = (avgR / den) * SQRT2 * offset;
            avgR *= avgR;
            if (avgR < 0.25f) avgR = 0.25f;
            fa = avgR;
        }
    }
    area *= fa;
    const float minarea = (float)(4.0 * M_PI / (double)(float)pm->np_emit);
    if (area < minarea) area = minarea;
    const float factor = (float)(4.0 * M_PI / (double)((float)pm->np_emit * area));
    free(sx); free(sy); free(sw); free(cand);
    return col_scale(total, factor);
}

static yo_col photon_illuminate(rctx* R, int li, const yo_sp* sp, yo_v3 eye)
{
    return photon_estimate(R, &R->lights[li], sp, eye);
}

void yo_photon_gather(yo_scene* s, const yg_material* mats, const yg_modulator* mods,
                      const yg_texture* texs, const yg_light* light, const float* p,
                      const float* n, const float* ng, const float* eye, const int32_t* mat,
                      int64_t cnt, float* rgb)
{
    rctx R;
    yg_render_opts o;
    memset(&o, 0, sizeof(o));
    rctx_init(&R, s, mats, 1 << 30, mods, texs, light, 1, NULL, 0, &o);
    for (int64_t i = 0; i < cnt; i++) {
        yo_sp sp;
        sp.P = v3_ld(p + 3 * i);
        sp.N = v3_ld(n + 3 * i);
        sp.Ng = v3_ld(ng + 3 * i);
        create_cs(sp.N, &sp.NU, &sp.NV);
        sp.Z = 0; sp.prim = -1; sp.obj = -1; sp.mat = mat[i];
        yo_col c = photon_estimate(&R, light, &sp, v3_ld(eye + 3 * i));
        rgb[3 * i] = c.r; rgb[3 * i + 1] = c.g; rgb[3 * i + 2] = c.b;
    }
    rctx_free(&R);
}

/* ------------------------------------------------------------- render -- */


/* ------------------------------------------------- global photon map -- */

/* dirConverter_t (photon.cc:11-25, photon.h:14-37): direction <-> (theta,
 * phi) bytes.  Table entries: the float angle's cosine and sine computed in
 * double and rounded to float. */
static float g_costh[255], g_sinth[255], g_cosph[256], g_sinph[256];
static pthread_once_t g_dirconv_once = PTHREAD_ONCE_INIT;
static void dirconv_fill(void)
{
    for (int i = 0; i < 255; ++i) {
        const float angle = (float)((double)(float)i * (1.0 / 255.0) * M_PI);
        g_costh[i] = (float)cos((double)angle);
        g_sinth[i] = (float)sin((double)angle);
    }
    for (int i = 0; i < 256; ++i) {
        const float angle = (float)((double)(float)i * (1.0 / 256.0) * 2.0 * M_PI);
        g_cosph[i] = (float)cos((double)angle);
        g_sinph[i] = (float)sin((double)angle);
    }
}
void yo_dirconv_tables(float* out)  /* costheta[255] sintheta[255] cosphi[256] sinphi[256] */
{
    pthread_once(&g_dirconv_once, dirconv_fill);
    memcpy(out, g_costh, sizeof(g_costh));
    memcpy(out + 255, g_sinth, sizeof(g_sinth));
    memcpy(out + 510, g_cosph, sizeof(g_cosph));
    memcpy(out + 766, g_sinph, sizeof(g_sinph));
}
/* storedPhoton_t::direction(dir) (photon.h:108-117) with convert (:24-34) */
static void dir_enc(yo_v3 d, uint8_t* th, uint8_t* ph)
{
    if (d.x == 0 && d.y == 0 && d.z == 0) { *th = 255; *ph = 0; return; }
    int t = (int)(acos((double)d.z) * (255.0 / M_PI));
    int p = (int)(atan2((double)d.y, (double)d.x) * (256.0 / (2.0 * M_PI)));
    if (t > 254) t = 254;
    else if (t < 0) t = 0;
    if (p > 255) p = 255;
    else if (p < 0) p += 256;
    *th = (uint8_t)t;
    *ph = (uint8_t)p;
}
static yo_v3 dir_dec(uint8_t th, uint8_t ph)
{
    pthread_once(&g_dirconv_once, dirconv_fill);
    if (th == 255) return v3(0, 0, 0);
    return v3(g_sinth[th] * g_cosph[ph], g_sinth[th] * g_sinph[ph], g_costh[th]);
}
/* rgbe_t (color.cc:140-155, color.h:154-170) */
static void rgbe_enc(yo_col c, uint8_t* e4)
{
    float v = c.r;
    if (c.g > v) v = c.g;
    if (c.b > v) v = c.b;
    if (v < 1e-32f) { e4[0] = e4[1] = e4[2] = e4[3] = 0; return; }
    int e;
    v = (float)((double)frexpf(v, &e) * 256.0 / (double)v);
    e4[0] = (uint8_t)(c.r * v);
    e4[1] = (uint8_t)(c.g * v);
    e4[2] = (uint8_t)(c.b * v);
    e4[3] = (uint8_t)(e + 128);
}
static yo_col rgbe_dec(const uint8_t* e4)
{
    if (!e4[3]) return col(0, 0, 0);
    const float f = (float)ldexp(1.0, (int)e4[3] - (int)(128 + 8));
    return col((float)e4[0] * f, (float)e4[1] * f, (float)e4[2] * f);
}
void yo_rgbe_roundtrip(const float* in, float* out, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        uint8_t e[4];
        rgbe_enc(col(in[3 * i], in[3 * i + 1], in[3 * i + 2]), e);
        const yo_col c = rgbe_dec(e);
        out[3 * i] = c.r; out[3 * i + 1] = c.g; out[3 * i + 2] = c.b;
    }
}
void yo_dir_roundtrip(const float* in, float* out, int32_t* code, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        uint8_t th, ph;
        dir_enc(v3(in[3 * i], in[3 * i + 1], in[3 * i + 2]), &th, &ph);
        const yo_v3 d = dir_dec(th, ph);
        out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
        if (code) { code[2 * i] = th; code[2 * i + 1] = ph; }
    }
}

/* HemiVec_CONE (globalphotonlight.cc:10-16, arealight.cc:219-225) */
static yo_v3 hemivec_cone(yo_v3 nrm, yo_v3 Ru, yo_v3 Rv, float cosang, float z1, float z2)
{
    const float t1 = (float)(2.0 * M_PI * (double)z1);
    const float t2 = (float)(1.0 - (1.0 - (double)cosang) * (double)z2);
    const yo_v3 a = v3_add(v3_mul(Ru, dcos(t1)), v3_mul(Rv, dsin(t1)));
    return v3_add(v3_mul(a, (float)sqrt(1.0 - (double)(t2 * t2))), v3_mul(nrm, t2));
}

typedef struct {
    yo_v3 corner, toX, toY, direction, NU, NV;
    yo_col scolor;
} yo_qemit;

/* areaLight_t::getEmitter + quadEmitter_t ctor / numSamples (arealight.h:90-95,
 * arealight.cc:200-217): colour = colour * power * |toX x toY| * 0.5 / n */
static void qemit_setup(const yg_light* L, int n, yo_qemit* q)
{
    yo_area A;
    memset(&A, 0, sizeof(A));
    yg_light L1 = *L;
    L1.samples = 1;  /* only the frame of the quad is needed */
    area_setup(&L1, &A);
    free(A.points); free(A.jit_h); free(A.jit_v);
    q->corner = A.corner;
    q->toX = A.toX;
    q->toY = A.toY;
    q->direction = A.direction;
    q->NU = v3_norm(A.toX);
    q->NV = v3_norm(A.toY);
    const float area = v3_len(v3_cross(A.toX, A.toY));
    const yo_col c = col_scale(col_scale(col(L->color[0], L->color[1], L->color[2]), L->power), area);
    const yo_col c2 = col_scale(c, 0.5f);
    q->scolor = col(c2.r / (float)n, c2.g / (float)n, c2.b / (float)n);
}

typedef struct {
    int32_t bx, by, bz;
    int64_t seq;
    yo_v3 pos, N;
    uint8_t th, ph, rgbe[4];
} yo_grec;

typedef struct { yo_grec* v; int64_t n, cap; } grecvec;

static void grec_push(grecvec* m, yo_grec x)
{
    if (m->n == m->cap) {
        m->cap = m->cap ? 2 * m->cap : 4096;
        m->v = (yo_grec*)realloc(m->v, sizeof(yo_grec) * (size_t)m->cap);
    }
    m->v[m->n++] = x;
}

typedef struct {
    rctx* R;
    int maxdepth;
    float cell;
    grecvec* recs;
    int64_t stored;
} gpctx;

/* globalPhotonLight_t::shoot (globalphotonlight.cc:19-95), no caustic objects */
static void gshoot(gpctx* G, yo_photon* ph, yo_v3 dir, int depth)
{
    rctx* R = G->R;
    if (depth > G->maxdepth) return;
    yo_sp sp;
    R->stats[1]++;
    if (!first_hit(R, ph->pos, dir, &sp, 0)) return;
    /* runningPhoton_t::position(p, MIN_RAYDIST), photon.cc:27-35 */
    if (v3_len(v3_sub(ph->pos, sp.P)) > R->opts.min_raydis) ph->lastpos = ph->pos;
    ph->pos = sp.P;
    const yo_v3 edir = v3_norm(v3_sub(ph->lastpos, ph->pos));
    const yo_v3 N = face_forward(sp.Ng, sp.N, edir);
    const yo_v3 Ng = face_forward(sp.Ng, sp.Ng, edir);
    /* storeFirst: quad emitters store the first hit (arealight.h:43) */
    if (obj_flag(R, sp.obj, 1)) {
        yo_grec r;
        r.pos = ph->pos;
        r.N = N;
        rgbe_enc(ph->c, r.rgbe);
        dir_enc(v3_norm(v3_sub(ph->lastpos, ph->pos)), &r.th, &r.ph);
        find_box(G->cell, r.pos, &r.bx, &r.by, &r.bz);
        r.seq = G->recs->n;
        grec_push(G->recs, r);
        G->stored++;
    }
    float diffuse = 0.0f;
    yo_col diffcolor = col(0, 0, 0);
    if (obj_flag(R, sp.obj, 2)) {
        diffcolor = get_diffuse(R, &sp, edir);
        diffuse = col_energy(diffcolor);
    }
    float sum = diffuse;  /* trans = 0 */
    if (sum > 0.0f) sum = (float)(1.0 / (double)sum);
    diffuse *= sum;
    if (sum > 0.0f) {
        R->stats[4]++;
        (void)our_random(&R->lcg);  /* ourRandom() < trans (= 0): never taken */
        diffcolor = col_scale(diffcolor, (float)(1.0 / (double)diffuse));
        const float r1 = our_random(&R->lcg), r2 = our_random(&R->lcg);
        R->stats[4] += 2;
        const yo_v3 refDir = hemivec_cone(Ng, sp.NU, sp.NV, 0.05f, r1, r2);
        ph->c = col_mul(ph->c, diffcolor);
        gshoot(G, ph, refDir, depth + 1);
    }
}

static int grec_cmp(const void* a, const void* b)
{
    const yo_grec* x = (const yo_grec*)a;
    const yo_grec* y = (const yo_grec*)b;
    if (x->bx != y->bx) return x->bx < y->bx ? -1 : 1;
    if (x->by != y->by) return x->by < y->by ? -1 : 1;
    if (x->bz != y->bz) return x->bz < y->bz ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

/* storeInHash (globalphotonlight.cc:177-195): the records of one box, in
 * the order they were stored */
static void gbox_fold(const yo_grec* r, int64_t n, yo_gbox* B)
{
    B->bx = r[0].bx; B->by = r[0].by; B->bz = r[0].bz;
    B->th = 255; B->ph = 0;
    memset(B->rgbe, 0, 4);
    memset(B->irr, 0, sizeof(B->irr));
    for (int64_t k = 0; k < n; k++) {
        const yo_grec* q = &r[k];
        if (B->th == 255) {
            B->pos[0] = q->pos.x; B->pos[1] = q->pos.y; B->pos[2] = q->pos.z;
            B->th = q->th; B->ph = q->ph;
            memcpy(B->rgbe, q->rgbe, 4);
            B->N[0] = q->N.x; B->N[1] = q->N.y; B->N[2] = q->N.z;
        } else {
            const yo_v3 nd = dir_dec(q->th, q->ph);
            if (v3_dot(nd, v3_ld(B->N)) > 0.0f) {
                const yo_col ac = rgbe_dec(B->rgbe), nc = rgbe_dec(q->rgbe);
                yo_v3 mixedd = v3_add(v3_mul(dir_dec(B->th, B->ph), col_energy(ac)), v3_mul(nd, col_energy(nc)));
                mixedd = v3_norm(mixedd);
                dir_enc(mixedd, &B->th, &B->ph);
                rgbe_enc(col_add(ac, nc), B->rgbe);
            }
        }
    }
}

typedef struct { float dis; int64_t idx; } yo_gfound;

static int gfound_cmp(const void* a, const void* b)
{
    const yo_gfound* x = (const yo_gfound*)a;
    const yo_gfound* y = (const yo_gfound*)b;
    if (x->dis != y->dis) return x->dis < y->dis ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* uniform grid over the box positions (cell = radius >= any gather radius) */
static void gm_grid(yo_gmap* M)
{
    free(M->gstart); free(M->gitems);
    M->gstart = NULL; M->gitems = NULL;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < M->n; i++)
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(lo[a], M->box[i].pos[a]);
            hi[a] = fmaxf(hi[a], M->box[i].pos[a]);
        }
    int64_t nc = 1;
    for (int a = 0; a < 3; a++) {
        M->gmin[a] = M->n ? lo[a] : 0.0f;
        int d = M->n ? (int)floorf((hi[a] - lo[a]) / M->radius) + 1 : 1;
        if (d < 1) d = 1;
        if (d > 256) d = 256;  /* items beyond the last cell stay in it (clamped) */
        M->gdim[a] = d;
        nc *= d;
    }
    M->gstart = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
    M->gitems = (int64_t*)malloc(sizeof(int64_t) * (size_t)(M->n > 0 ? M->n : 1));
    int64_t* cellof = (int64_t*)malloc(sizeof(int64_t) * (size_t)(M->n > 0 ? M->n : 1));
    for (int64_t i = 0; i < M->n; i++) {
        int c[3];
        for (int a = 0; a < 3; a++) {
            int k = (int)floorf((M->box[i].pos[a] - M->gmin[a]) / M->radius);
            if (k < 0) k = 0;
            if (k >= M->gdim[a]) k = M->gdim[a] - 1;
            c[a] = k;
        }
        cellof[i] = ((int64_t)c[2] * M->gdim[1] + c[1]) * M->gdim[0] + c[0];
        M->gstart[cellof[i] + 1]++;
    }
    for (int64_t c = 0; c < nc; c++) M->gstart[c + 1] += M->gstart[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)nc);
    memcpy(fill, M->gstart, sizeof(int64_t) * (size_t)nc);
    for (int64_t i = 0; i < M->n; i++) M->gitems[fill[cellof[i]]++] = i;
    free(fill); free(cellof);
}

/* uniform grid (cell = radius) over the raw map's photons */
static void gr_grid(yo_gmap* M)
{
    free(M->rgstart); free(M->rgitems);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < M->nrec; i++)
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(lo[a], M->rec[i].pos[a]);
            hi[a] = fmaxf(hi[a], M->rec[i].pos[a]);
        }
    int64_t nc = 1;
    for (int a = 0; a < 3; a++) {
        M->rgmin[a] = M->nrec ? lo[a] : 0.0f;
        int d = M->nrec ? (int)floorf((hi[a] - lo[a]) / M->radius) + 1 : 1;
        if (d < 1) d = 1;
        if (d > 256) d = 256;
        M->rgdim[a] = d;
        nc *= d;
    }
    M->rgstart = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
    M->rgitems = (int64_t*)malloc(sizeof(int64_t) * (size_t)(M->nrec > 0 ? M->nrec : 1));
    int64_t* cof = (int64_t*)malloc(sizeof(int64_t) * (size_t)(M->nrec > 0 ? M->nrec : 1));
    for (int64_t i = 0; i < M->nrec; i++) {
        int c[3];
        for (int a = 0; a < 3; a++) {
            int k = (int)floorf((M->rec[i].pos[a] - M->rgmin[a]) / M->radius);
            if (k < 0) k = 0;
            if (k >= M->rgdim[a]) k = M->rgdim[a] - 1;
            c[a] = k;
        }
        cof[i] = ((int64_t)c[2] * M->rgdim[1] + c[1]) * M->rgdim[0] + c[0];
        M->rgstart[cof[i] + 1]++;
    }
    for (int64_t c = 0; c < nc; c++) M->rgstart[c + 1] += M->rgstart[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc + 1));
    memcpy(fill, M->rgstart, sizeof(int64_t) * (size_t)nc);
    for (int64_t i = 0; i < M->nrec; i++) M->rgitems[fill[cof[i]]++] = i;
    free(fill); free(cof);
}

/* globalPhotonMap_t::gather (photon.cc:167-212) over the raw map, mincos 0:
 * the K nearest photons within *radius whose direction faces N, sorted by
 * (distance, store index); the radius doubles while fewer than K are reached
 * and shrinks by 0.95 when K is under 0.49 of those reached.  (The reference
 * sums its found photons in heap order; the sorted order here changes only
 * float rounding.) */
static int gr_gather(const yo_gmap* M, yo_v3 P, yo_v3 N, int K, float* radius, yo_gfound* found)
{
    int reached = 0, nf = 0;
    while ((reached < K) && (*radius <= M->radius)) {
        reached = 0;
        nf = 0;
        const float r = *radius;
        const float pp[3] = {P.x, P.y, P.z};
        int g0[3], g1[3];
        for (int a = 0; a < 3; a++) {
            long lo = (long)floorf((pp[a] - r - M->rgmin[a]) / M->radius);
            long hi = (long)floorf((pp[a] + r - M->rgmin[a]) / M->radius);
            if (lo < 0) lo = 0;
            if (hi >= M->rgdim[a]) hi = M->rgdim[a] - 1;
            g0[a] = (int)lo; g1[a] = (int)hi;
        }
        for (int z = g0[2]; z <= g1[2]; z++)
            for (int y = g0[1]; y <= g1[1]; y++)
                for (int x = g0[0]; x <= g1[0]; x++) {
                    const int64_t c = ((int64_t)z * M->rgdim[1] + y) * M->rgdim[0] + x;
                    for (int64_t q = M->rgstart[c]; q < M->rgstart[c + 1]; q++) {
                        const int64_t i = M->rgitems[q];
                        const yo_gphot* g = &M->rec[i];
                        const float D = v3_len(v3_sub(v3_ld(g->pos), P));
                        if ((D > r) || (v3_dot(dir_dec(g->th, g->ph), N) <= 0.0f)) continue;
                        reached++;
                        yo_gfound f = {D, i};
                        if (nf < K) { found[nf++] = f; continue; }
                        int far = 0;
                        for (int k = 1; k < nf; k++)
                            if (gfound_cmp(&found[k], &found[far]) > 0) far = k;
                        if (gfound_cmp(&f, &found[far]) < 0) found[far] = f;
                    }
                }
        if (reached < K) *radius *= 2;
    }
    if (reached > K) {
        const float f = (float)K / (float)reached;
        if (f < (float)(0.7 * 0.7)) *radius *= 0.95f;
    }
    if (*radius > M->radius) *radius = M->radius;
    qsort(found, (size_t)nf, sizeof(yo_gfound), gfound_cmp);
    return nf;
}

/* ------------------------------------------------ photonSampler_t -- */

/* ourRandomI (vector3d.h:264-274): the next Park-Miller state itself */
static int32_t our_random_i(int32_t* seed)
{
    const int a = 16807, m = 2147483647, q = m / a, r = m % a;
    int32_t v = *seed;
    v = a * (v % q) - r * (v / q);
    if (v < 0) v += m;
    *seed = v;
    return v;
}

/* Halton::setStart (mcqmc.h:24-33) */
static void halton_start(yo_halton* h, unsigned int i)
{
    double value = 0.0, f, factor;
    f = factor = h->inv_base;
    while (i > 0) {
        value += (double)(i % h->base) * factor;
        i /= h->base;
        factor *= f;
    }
    h->value = value;
}

static void ps_grid(yo_psamp* S, int grid)
{
    const float base = (float)sqrt((double)(float)grid / 2.0);
    S->par = (int)(base + 0.5);
    S->mer = S->par * 2;
    S->pdiv = (float)(1.0 / (double)(float)S->par);
    S->mdiv = (float)(2.0 * M_PI / (double)(float)S->mer);
    S->sectors = S->par * S->mer;
    S->search = 3 * S->sectors;
}

/* photonSampler_t ctor (pathtools.cc:80-115) at the area's first takeSample:
 * every Halton dimension starts at an ourRandomI() draw, grid 36 */
static void ps_create(rctx* R, int li)
{
    const yg_light* L = &R->lights[li];
    const int nd = (L->depth + 1) * 2;
    for (int d = 0; d < nd; d++) halton_start(&R->hseq[li][d], (unsigned int)our_random_i(&R->lcg));
    ps_grid(&R->ps, 36);
    R->ps.radius = R->s->gm.radius;
    R->ps.made = 1;
}

/* photonSampler_t::getCoords (pathtools.cc:118-145).  sqrt(1 - z z) is NaN
 * when |z| exceeds 1 by rounding (undefined in the reference's int cast):
 * taken as 0 here. */
static void ps_coords(const yo_psamp* S, yo_v3 v, yo_v3 N, yo_v3 Ru, yo_v3 Rv, int* c0, int* c1)
{
    float x = v3_dot(v, Ru), y = v3_dot(v, Rv), z = v3_dot(v, N);
    float fix = (float)sqrt(1.0 - (double)(z * z));
    if (fix != fix) fix = 0.0f;
    int first = (int)(fix / S->pdiv);
    if (first >= S->par) first--;
    if (fix > 1.0f) fix = 1.0f;
    if (fix != 0.0f) { x /= fix; y /= fix; }
    if (x > 1.0f) x = 1.0f;
    if (x < -1.0f) x = -1.0f;
    float an = (float)acos((double)x);
    if (y < 0) an = (float)(2.0 * M_PI - (double)an);
    int second = (int)(an / S->mdiv);
    if (second >= S->mer) second--;
    *c0 = first;
    *c1 = second;
}

static float max_abs_diff(yo_col a, yo_col b)
{
    return fmaxf(fabsf(a.r - b.r), fmaxf(fabsf(a.g - b.g), fabsf(a.b - b.b)));
}

/* photonSampler_t::samplingFrom (pathtools.cc:172-236) with giveMaxDiff
 * (:147-170): energy of the gathered photons per sector, hits spread by the
 * energy differences to the 8 neighbours, weight = max hits / hits */
static void ps_plan(const yo_gmap* M, yo_psamp* S, float* radius, yo_v3 P, yo_v3 N, yo_v3 Ru, yo_v3 Rv,
                    int samples, int* nfound)
{
    yo_gfound* found = (yo_gfound*)malloc(sizeof(yo_gfound) * (size_t)(S->search + 1));
    const int nf = gr_gather(M, P, N, S->search, radius, found);
    if (nfound) *nfound = nf;
    yo_col energy[YO_PS_MAXSEC];
    const int par = S->par, mer = S->mer;
    for (int k = 0; k < S->sectors; k++) { energy[k] = col(0, 0, 0); S->hits[k] = 0; }
    for (int k = 0; k < nf; k++) {
        const yo_gphot* g = &M->rec[found[k].idx];
        const yo_v3 d = dir_dec(g->th, g->ph);
        if (v3_dot(d, N) > 0.0f) {
            int c0, c1;
            ps_coords(S, d, N, Ru, Rv, &c0, &c1);
            energy[c0 * mer + c1] = col_add(energy[c0 * mer + c1], rgbe_dec(g->rgbe));
        }
    }
    free(found);
    float totaldiff = 0;
    for (int i = 0; i < par; ++i)
        for (int j = 0; j < mer; ++j) {
            const yo_col e = energy[i * mer + j];
            float mx = 0, t;
            const int il = i - 1, ih = i + 1, jl = j - 1, jh = j + 1;
            if (il >= 0) {
                if (jl >= 0) { t = max_abs_diff(energy[il * mer + jl], e); if (t > mx) mx = t; }
                t = max_abs_diff(energy[il * mer + j], e); if (t > mx) mx = t;
                if (jh < mer) { t = max_abs_diff(energy[il * mer + jh], e); if (t > mx) mx = t; }
            }
            if (jl >= 0) { t = max_abs_diff(energy[i * mer + jl], e); if (t > mx) mx = t; }
            if (jh < mer) { t = max_abs_diff(energy[i * mer + jh], e); if (t > mx) mx = t; }
            if (ih < par) {
                if (jl >= 0) { t = max_abs_diff(energy[ih * mer + jl], e); if (t > mx) mx = t; }
                t = max_abs_diff(energy[ih * mer + j], e); if (t > mx) mx = t;
                if (jh < mer) { t = max_abs_diff(energy[ih * mer + jh], e); if (t > mx) mx = t; }
            }
            S->weight[i * mer + j] = mx;
            totaldiff += mx;
        }
    int assign = samples - S->sectors;
    const float sph = (totaldiff != 0.0f) ? (float)assign / totaldiff : 0;
    for (int k = 0; k < S->sectors; k++) {
        int h = (int)((double)(sph * S->weight[k]) + 0.5);
        if (h > assign) h = assign;
        if (h < 0) h = 0;  /* only for a non-finite energy (the reference's cast is undefined there) */
        assign -= h;
        S->hits[k] = h + 1;
    }
    int noway = 0;
    while (assign) {
        for (int k = 0; k < S->sectors && assign; k++)
            if ((S->hits[k] > 1) || noway) { S->hits[k]++; assign--; }
        noway = 1;
    }
    int mx = 1;
    for (int k = 0; k < S->sectors; k++)
        if (S->hits[k] > mx) mx = S->hits[k];
    const float maxf = (float)mx;
    for (int k = 0; k < S->sectors; k++) S->weight[k] = maxf / (float)S->hits[k];
    S->multi = (float)(1.0 / (double)(maxf * (float)S->sectors));
    S->cur[0] = S->cur[1] = S->cur[2] = 0;
}

static void ps_sampling_from(rctx* R, yo_psamp* S, yo_v3 P, yo_v3 N, yo_v3 Ru, yo_v3 Rv, int samples)
{
    const yo_gmap* M = &R->s->gm;
    if (M->stateless) S->radius = M->radius;
    ps_plan(M, S, &S->radius, P, N, Ru, Rv, samples, NULL);
}

/* photonSampler_t::nextDirection (pathtools.cc:238-261) and nextSample
 * (pathtools.h:92-106): level 0 draws inside the walk's current sector and
 * weighs the ray by weight x 2 z1; deeper levels use z1 as the sine */
static yo_v3 ps_next_direction(yo_psamp* S, yo_halton* hs, yo_v3 N, yo_v3 Ru, yo_v3 Rv, int curlev,
                               yo_col* raycolor)
{
    float z1, z2;
    if (curlev == 0) {
        z1 = ((float)S->cur[0] + halton_next(&hs[0])) * S->pdiv;
        z2 = ((float)S->cur[1] + halton_next(&hs[1])) * S->mdiv;
        *raycolor = col_scale(*raycolor, S->weight[S->cur[0] * S->mer + S->cur[1]] * 2.0f * z1);
        if (++S->cur[2] == S->hits[S->cur[0] * S->mer + S->cur[1]]) {
            S->cur[2] = 0;
            if (++S->cur[1] == S->mer) {
                S->cur[1] = 0;
                if (++S->cur[0] == S->par) S->cur[0] = 0;
            }
        }
    } else {
        curlev <<= 1;
        z1 = halton_next(&hs[curlev]);
        z2 = (float)((double)halton_next(&hs[curlev + 1]) * 2.0 * M_PI);
    }
    if (z1 > 1.0f) z1 = 1.0f;
    yo_v3 a = v3_add(v3_mul(Ru, dcos(z2)), v3_mul(Rv, dsin(z2)));
    return v3_add(v3_mul(a, z1), v3_mul(N, (float)sqrt(1.0 - (double)(z1 * z1))));
}

int yo_psamp_plan(yo_scene* s, const float* P, const float* N, const float* Ru, const float* Rv, int samples,
                  int grid, int32_t* hits, float* weight, float* multi, int32_t* nfound)
{
    const yo_gmap* M = &s->gm;
    if (!M->rec) return -1;
    yo_psamp S;
    memset(&S, 0, sizeof(S));
    ps_grid(&S, grid);
    if (S.sectors > YO_PS_MAXSEC || samples < S.sectors) return -2;
    float radius = M->radius;
    int nf = 0;
    ps_plan(M, &S, &radius, v3_ld(P), v3_ld(N), v3_ld(Ru), v3_ld(Rv), samples, &nf);
    memcpy(hits, S.hits, sizeof(int32_t) * (size_t)S.sectors);
    memcpy(weight, S.weight, sizeof(float) * (size_t)S.sectors);
    *multi = S.multi;
    if (nfound) *nfound = nf;
    return S.sectors;
}

int64_t yo_gphoton_read_photons(yo_scene* s, uint32_t* out, int64_t cap)
{
    const yo_gmap* M = &s->gm;
    if (!M->rec) return -1;
    for (int64_t i = 0; i < M->nrec && i < cap; i++) {
        const yo_gphot* g = &M->rec[i];
        uint32_t* o = out + 5 * i;
        memcpy(o, g->pos, 12);
        o[3] = (uint32_t)g->th | ((uint32_t)g->ph << 8);
        o[4] = (uint32_t)g->rgbe[0] | ((uint32_t)g->rgbe[1] << 8) | ((uint32_t)g->rgbe[2] << 16) |
               ((uint32_t)g->rgbe[3] << 24);
    }
    return M->nrec;
}

int yo_gphoton_load_photons(yo_scene* s, const uint32_t* data, int64_t n)
{
    yo_gmap* M = &s->gm;
    if (!M->built) return -1;
    free(M->rec);
    M->rec = (yo_gphot*)calloc((size_t)(n > 0 ? n : 1), sizeof(yo_gphot));
    for (int64_t i = 0; i < n; i++) {
        const uint32_t* o = data + 5 * i;
        yo_gphot* g = &M->rec[i];
        memcpy(g->pos, o, 12);
        g->th = (uint8_t)(o[3] & 255u);
        g->ph = (uint8_t)((o[3] >> 8) & 255u);
        for (int k = 0; k < 4; k++) g->rgbe[k] = (uint8_t)(o[4] >> (8 * k));
    }
    M->nrec = n;
    gr_grid(M);
    return 0;
}

/* globalPhotonMap_t::gather (photon.cc:167-210) over the boxes' photons:
 * direction and colour of box i as `which` says (0: the merged photon, only
 * boxes with a direction; 1: the irradiance photon, direction N, colour
 * RGBE(irr)).  *radius: the caller's radius state.  Returns the number
 * found, sorted by (distance, box). */
static int gm_gather(const yo_gmap* M, int which, yo_v3 P, yo_v3 N, int K, float* radius, float mincos,
                     yo_gfound* found, int cap)
{
    int reached = 0, nf = 0;
    int first = 1;
    while (first || ((reached < K) && (*radius <= M->radius))) {
        first = 0;
        if (!(*radius <= M->radius)) break;
        reached = 0;
        nf = 0;
        const float r = *radius;
        int g0[3], g1[3];
        const float pp[3] = {P.x, P.y, P.z};
        for (int a = 0; a < 3; a++) {
            long lo = (long)floorf((pp[a] - r - M->gmin[a]) / M->radius);
            long hi = (long)floorf((pp[a] + r - M->gmin[a]) / M->radius);
            if (lo < 0) lo = 0;
            if (hi >= M->gdim[a]) hi = M->gdim[a] - 1;
            g0[a] = (int)lo; g1[a] = (int)hi;
        }
        for (int z = g0[2]; z <= g1[2]; z++)
            for (int y = g0[1]; y <= g1[1]; y++)
                for (int x = g0[0]; x <= g1[0]; x++) {
                    const int64_t c = ((int64_t)z * M->gdim[1] + y) * M->gdim[0] + x;
                    for (int64_t q = M->gstart[c]; q < M->gstart[c + 1]; q++) {
                        const int64_t i = M->gitems[q];
                        const yo_gbox* B = &M->box[i];
                        yo_v3 d;
                        if (which == 0) {
                            if (B->th == 255) continue;
                            d = dir_dec(B->th, B->ph);
                        } else {
                            uint8_t th, ph;
                            dir_enc(v3_ld(B->N), &th, &ph);
                            d = dir_dec(th, ph);
                        }
                        const float D = v3_len(v3_sub(v3_ld(B->pos), P));
                        if ((D > r) || (v3_dot(d, N) <= mincos)) continue;
                        reached++;
                        yo_gfound f = {D, i};
                        if (nf < K && nf < cap) { found[nf++] = f; continue; }
                        /* keep the K nearest: replace the farthest when nearer */
                        int far = 0;
                        for (int k = 1; k < nf; k++)
                            if (gfound_cmp(&found[k], &found[far]) > 0) far = k;
                        if (gfound_cmp(&f, &found[far]) < 0) found[far] = f;
                    }
                }
        if (reached < K) *radius *= 2;
        if (M->stateless) break;
    }
    if (reached > K) {
        const float f = (float)K / (float)reached;
        if (f < (float)(0.7 * 0.7)) *radius *= 0.95f;
    }
    if (*radius > M->radius) *radius = M->radius;
    qsort(found, (size_t)nf, sizeof(yo_gfound), gfound_cmp);
    return nf;
}

static yo_col gbox_irr_colour(const yo_gbox* B)
{
    uint8_t e[4];
    rgbe_enc(col(B->irr[0], B->irr[1], B->irr[2]), e);
    return rgbe_dec(e);
}

/* setIrradiance (globalphotonlight.cc:127-175) of every box, boxes in hash
 * order (x, y, z), the gather radius carried from box to box */
static void gm_irradiances(yo_gmap* M, int search)
{
    yo_gfound* found = (yo_gfound*)malloc(sizeof(yo_gfound) * (size_t)(search + 1));
    float radius = M->radius;
    for (int64_t i = 0; i < M->n; i++) {
        yo_gbox* B = &M->box[i];
        if (M->stateless) radius = M->radius;
        const yo_v3 N = v3_ld(B->N);
        const int nf = gm_gather(M, 0, v3_ld(B->pos), N, search, &radius, 0.0f, found, search + 1);
        yo_col total = col(0, 0, 0);
        if (nf == 0) {
            B->irr[0] = B->irr[1] = B->irr[2] = 0;
            continue;
        }
        const yo_gbox* F0 = &M->box[found[nf - 1].idx];  /* the heap's front: the farthest */
        if (nf == 1 || found[nf - 1].dis == 0) {
            float factor = v3_dot(dir_dec(F0->th, F0->ph), N);
            if (factor < 0.0f) factor = 0.0f;
            const yo_col c = col_scale(rgbe_dec(F0->rgbe), factor);
            B->irr[0] = c.r; B->irr[1] = c.g; B->irr[2] = c.b;
            continue;
        }
        const float farest = found[nf - 1].dis;
        for (int k = 0; k < nf; k++) {
            const yo_gbox* F = &M->box[found[k].idx];
            const float factor = (1.0f - found[k].dis / farest) * v3_dot(dir_dec(F->th, F->ph), N);
            if (factor > 0) total = col_add(total, col_scale(rgbe_dec(F->rgbe), factor));
        }
        float area = (float)(M_PI * (double)farest * (double)farest);
        if (area < 0.00005f) area = 0.00005f;
        const yo_col c = col_scale(total, (float)(4.0 * M_PI / (double)area));
        B->irr[0] = c.r; B->irr[1] = c.g; B->irr[2] = c.b;
    }
    free(found);
}

/* getLight's irradiance part (pathlight.cc:147-176): the box at P if its
 * normal agrees (N.N > 0.7), else the weighted mean of the 5 nearest
 * irradiance photons within the thread's radius (directions . N > 0.25) */
static yo_col gm_irradiance(rctx* R, const yo_sp* sp, yo_v3 N)
{
    const yo_gmap* M = &R->s->gm;
    int32_t bx, by, bz;
    hash_cell(M->cell, sp->P, &bx, &by, &bz);
    int64_t lo = 0, hi = M->n;
    while (lo < hi) {  /* boxes sorted by (bx, by, bz) */
        const int64_t mid = (lo + hi) / 2;
        const yo_gbox* B = &M->box[mid];
        const int less = B->bx != bx ? B->bx < bx : (B->by != by ? B->by < by : B->bz < bz);
        if (less) lo = mid + 1;
        else hi = mid;
    }
    if (lo < M->n && M->box[lo].bx == bx && M->box[lo].by == by && M->box[lo].bz == bz) {
        const yo_gbox* B = &M->box[lo];
        const float factor = v3_dot(v3_ld(B->N), N);
        if (factor > 0.7f) return col_scale(col(B->irr[0], B->irr[1], B->irr[2]), factor);
    }
    yo_gfound found[6];
    if (M->stateless) R->gm_radius = M->radius;
    const int nf = gm_gather(M, 1, sp->P, N, 5, &R->gm_radius, 0.25f, found, 6);
    yo_col total = col(0, 0, 0);
    if (!nf) return total;
    float farest = (nf == 1) ? R->gm_radius : found[nf - 1].dis;
    if (farest == 0.0f) farest = 1.0f;
    float div = 0.0f;
    for (int k = 0; k < nf; k++) {
        const yo_gbox* B = &M->box[found[k].idx];
        uint8_t th, ph;
        dir_enc(v3_ld(B->N), &th, &ph);
        const float factor = v3_dot(dir_dec(th, ph), N) * (1.0f - found[k].dis / farest);
        if (factor > 0) {
            total = col_add(total, col_scale(gbox_irr_colour(B), factor));
            div += factor;
        }
    }
    if (div > 0.0f) total = col_scale(total, (float)(1.0 / (double)div));
    return total;
}

/* globalPhotonLight_t::init (globalphotonlight.cc:234-278) for light li:
 * stats = {emitted, stored, boxes, ray count, ourRandom() draws} */
int yo_gphoton_build(yo_scene* s, const yg_material* mats, int nmat, const yg_modulator* mods,
                     const yg_texture* texs, const yg_light* lights, int nl, const int32_t* obj_flags,
                     int nobj, const yg_render_opts* opts, int li, int stateless, uint64_t* stats)
{
    if (li < 0 || li >= nl || lights[li].type != YG_LIGHT_GLOBALPHOTON) return -1;
    const yg_light* GL = &lights[li];
    if (GL->search < 1 || GL->photons < 1 || !(GL->fixedradius > 0)) return -2;
    rctx R;
    rctx_init(&R, s, mats, nmat, mods, texs, lights, nl, obj_flags, nobj, opts);
    R.lcg = (int32_t)opts->lcg_seed;
    yo_gmap* M = &s->gm;
    free(M->box); free(M->gstart); free(M->gitems);
    free(M->rec); free(M->rgstart); free(M->rgitems);
    memset(M, 0, sizeof(*M));
    M->light = li;
    M->radius = GL->fixedradius;
    M->cell = GL->fixedradius / sqrtf((float)GL->search);
    M->stateless = stateless;
    int numemitters = 0;
    for (int k = 0; k < nl; k++)
        if (lights[k].type == YG_LIGHT_AREA && lights[k].dummy) numemitters++;
    grecvec recs = {NULL, 0, 0};
    gpctx G = {&R, GL->depth, M->cell, &recs, 0};
    int64_t emitted = 0;
    if (numemitters) {
        const int ppl = GL->photons / numemitters;
        for (int k = 0; k < nl; k++) {
            if (lights[k].type != YG_LIGHT_AREA || !lights[k].dummy) continue;
            yo_qemit q;
            qemit_setup(&lights[k], ppl, &q);
            for (int j = 0; j < ppl; ++j) {
                /* quadEmitter_t::getDirection (arealight.cc:227-232); the
                 * argument and operand draws in GCC's order (right to left) */
                const float z2 = our_random(&R.lcg), z1 = our_random(&R.lcg);
                const yo_v3 dir = hemivec_cone(q.direction, q.NU, q.NV, 0.0001f, z1, z2);
                const float ry = our_random(&R.lcg), rx = our_random(&R.lcg);
                R.stats[4] += 4;
                const yo_v3 from = v3_add(v3_add(q.corner, v3_mul(q.toX, rx)), v3_mul(q.toY, ry));
                yo_photon ph;
                ph.pos = from;
                ph.lastpos = from;
                ph.c = col_scale(q.scolor, v3_dot(q.direction, dir));
                gshoot(&G, &ph, dir, 0);
                emitted++;
            }
        }
    }
    /* the raw map keeps the records in store order (photonMap->store) */
    M->nrec = recs.n;
    M->rec = (yo_gphot*)calloc((size_t)(recs.n > 0 ? recs.n : 1), sizeof(yo_gphot));
    for (int64_t i = 0; i < recs.n; i++) {
        yo_gphot* g = &M->rec[i];
        g->pos[0] = recs.v[i].pos.x; g->pos[1] = recs.v[i].pos.y; g->pos[2] = recs.v[i].pos.z;
        g->th = recs.v[i].th; g->ph = recs.v[i].ph;
        memcpy(g->rgbe, recs.v[i].rgbe, 4);
    }
    gr_grid(M);
    /* the hash boxes in (x, y, z) order, records of a box in store order */
    qsort(recs.v, (size_t)recs.n, sizeof(yo_grec), grec_cmp);
    int64_t nbox = 0;
    for (int64_t i = 0; i < recs.n; i++)
        if (i == 0 || recs.v[i].bx != recs.v[i - 1].bx || recs.v[i].by != recs.v[i - 1].by ||
            recs.v[i].bz != recs.v[i - 1].bz)
            nbox++;
    M->box = (yo_gbox*)calloc((size_t)(nbox > 0 ? nbox : 1), sizeof(yo_gbox));
    int64_t b = 0;
    for (int64_t i = 0; i < recs.n;) {
        int64_t j = i + 1;
        while (j < recs.n && recs.v[j].bx == recs.v[i].bx && recs.v[j].by == recs.v[i].by &&
               recs.v[j].bz == recs.v[i].bz)
            j++;
        gbox_fold(&recs.v[i], j - i, &M->box[b++]);
        i = j;
    }
    M->n = nbox;
    free(recs.v);
    gm_grid(M);
    gm_irradiances(M, GL->search);
    M->stored = G.stored;
    M->emitted = emitted;
    M->lcg_after = R.lcg;
    M->own_stream = 1;
    M->built = 1;
    if (stats) {
        stats[0] = (uint64_t)emitted; stats[1] = (uint64_t)G.stored; stats[2] = (uint64_t)nbox;
        stats[3] = R.stats[1]; stats[4] = R.stats[4];
    }
    rctx_free(&R);
    return 0;
}

/* the boxes as 15 floats each (position, N, irradiance, merged direction,
 * merged colour), like yg_gphoton_read */
int64_t yo_gphoton_read(yo_scene* s, float* out, int64_t cap)
{
    const yo_gmap* M = &s->gm;
    if (!M->built) return -1;
    for (int64_t i = 0; i < M->n && i < cap; i++) {
        const yo_gbox* B = &M->box[i];
        const yo_v3 d = dir_dec(B->th, B->ph);
        const yo_col c = rgbe_dec(B->rgbe);
        float* o = out + 15 * i;
        memcpy(o, B->pos, 12); memcpy(o + 3, B->N, 12); memcpy(o + 6, B->irr, 12);
        o[9] = d.x; o[10] = d.y; o[11] = d.z;
        o[12] = c.r; o[13] = c.g; o[14] = c.b;
    }
    return M->n;
}

/* the irradiance boxes of another implementation (15 floats each, in box
 * order; box indices recomputed from the positions with the light's box
 * size): renders then read them instead of a pass made here */
int yo_gphoton_load(yo_scene* s, const float* data, int64_t n, const yg_light* GL, int li, int stateless,
                    int recompute)
{
    yo_gmap* M = &s->gm;
    free(M->box); free(M->gstart); free(M->gitems);
    free(M->rec); free(M->rgstart); free(M->rgitems);
    memset(M, 0, sizeof(*M));
    M->light = li;
    M->radius = GL->fixedradius;
    M->cell = GL->fixedradius / sqrtf((float)GL->search);
    M->stateless = stateless;
    M->box = (yo_gbox*)calloc((size_t)(n > 0 ? n : 1), sizeof(yo_gbox));
    for (int64_t i = 0; i < n; i++) {
        const float* o = data + 15 * i;
        yo_gbox* B = &M->box[i];
        memcpy(B->pos, o, 12); memcpy(B->N, o + 3, 12); memcpy(B->irr, o + 6, 12);
        dir_enc(v3(o[9], o[10], o[11]), &B->th, &B->ph);
        rgbe_enc(col(o[12], o[13], o[14]), B->rgbe);
        find_box(M->cell, v3_ld(B->pos), &B->bx, &B->by, &B->bz);
    }
    M->n = n;
    gm_grid(M);
    /* recompute: pre-gather the loaded boxes' irradiance here (setIrradiance) */
    if (recompute) gm_irradiances(M, GL->search);
    M->built = 1;
    M->own_stream = 0;
    return 0;
}

typedef struct { yo_v3 pos, vright, vup, vto; } yo_cam;

/* camera_t::camera_t perspective part, camera.cc:27-87 */
static void cam_setup(const yg_camera* c, yo_cam* k)
{
    yo_v3 pos = v3_ld(c->pos), look = v3_ld(c->look), up = v3_ld(c->up);
    yo_v3 vup = v3_sub(up, pos), vto = v3_sub(look, pos);
    yo_v3 vright = v3_cross(vup, vto);
    vup = v3_cross(vright, vto);
    vup = v3_norm(vup);
    vright = v3_norm(vright);
    vright = v3_mul(vright, -1.0f);
    (void)v3_normlen(&vto);
    vup = v3_mul(vup, c->aspect * (float)c->resy / (float)c->resx);
    vto = v3_sub(v3_mul(vto, c->focal), v3_mul(v3_add(vup, vright), 0.5f));
    vup = v3_div(vup, (float)c->resy);
    vright = v3_div(vright, (float)c->resx);
    k->pos = pos; k->vright = vright; k->vup = vup; k->vto = vto;
}

/* camera_t::shootRay perspective, camera.cc:184-189 */
static yo_v3 cam_ray(const yo_cam* k, float px, float py)
{
    return v3_norm(v3_add(v3_add(v3_mul(k->vright, px), v3_mul(k->vup, py)), k->vto));
}

/* scene_t::raytrace, scene.cc:200-258 (raylevel 0, no fog, no displacement) */
static yo_col raytrace(rctx* R, yo_v3 from, yo_v3 ray)
{
    yo_sp sp;
    R->stats[0]++;
    if (first_hit(R, from, ray, &sp, 0)) {
        sp.sx = R->scr_x;  /* sp.setScreenPos(state.screenpos) (scene.cc:234) */
        sp.sy = R->scr_y;
        R->traveled = sp.Z;  /* state.traveled += sp.Z() from 0 at the camera ray (scene.cc:239-240) */
        R->primary = 1;
        yo_col c = scene_light(R, &sp, from, 0);
        R->primary = 0;
        return c;
    }
    return background(R);
}

typedef struct { int x, y, w, h, rx, ry, rw, rh; } yo_region;
typedef struct yo_lcache_s* yo_lcache_ptr;

/* blockSpliter_t::blockSpliter_t, renderblock.cc:71-101 */
int yo_block_regions(int width, int height, int b, uint32_t seed, yo_region** out)
{
    int bw = width / b, bh = height / b;
    if (width % b) bw++;
    if (height % b) bh++;
    int n = bw * bh;
    int* scram = (int*)malloc(sizeof(int) * (size_t)n);
    int32_t* rnd = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    yo_glibc_rand(seed, n, rnd);
    for (int i = 0; i < n; ++i) scram[i] = i;
    for (int i = 0; i < n; ++i) {
        int j = rnd[i] % n;
        int t = scram[i]; scram[i] = scram[j]; scram[j] = t;
    }
    yo_region* regions = (yo_region*)calloc((size_t)n, sizeof(yo_region));
    int sec = 0;
    for (int i = 0; i < bh; ++i)
        for (int j = 0; j < bw; ++j) {
            yo_region r;
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
    free(scram); free(rnd);
    *out = regions;
    return n;
}

/* the same regions into a caller buffer of cap x 8 ints (x y w h rx ry rw rh);
 * returns the region count (regions beyond cap are not written) */
int yo_block_regions_into(int width, int height, int b, uint32_t seed, int32_t* out, int cap)
{
    yo_region* r = NULL;
    const int n = yo_block_regions(width, height, b, seed, &r);
    for (int q = 0; q < n && q < cap; q++) {
        const int32_t v[8] = {r[q].x, r[q].y, r[q].w, r[q].h, r[q].rx, r[q].ry, r[q].rw, r[q].rh};
        memcpy(out + 8 * q, v, sizeof(v));
    }
    free(r);
    return n;
}

typedef struct {
    yo_scene* s;
    const yg_camera* cam;
    const yg_material* mats; int nmat;
    const yg_modulator* mods; const yg_texture* texs;
    const yg_light* lights; int nl;
    const int32_t* obj_flags; int nobj;
    const yg_render_opts* opts;
    const yo_cam* k;
    const yo_region* regions;
    int nreg;
    float* rgb;
    volatile int next;
    pthread_mutex_t mu;
    uint64_t stats[5];
    rctx* shared_R;   /* sequential mode only */
    const float* grad_rgb;  /* adjoint: dLoss/dpixel */
    double *gmat, *gmod, *glight;
    int nmod;
    yo_lcache_ptr lc;    /* irradiance cache of a cache = on pathlight, filled by the fake pass */
    float worldres;
    uint64_t lc_new;
    struct yo_grand_s* grand;  /* glibc rand() state shared by the render (areas in turn with area lights) */
} render_job;

/* color_t::expgam_Adjust (color.h:175-195) with scene_t's stored values
 * exposure = -e and gamma_R = 1/g (0 for g = 0) (scene.h:161-164); exp and
 * pow are the C library's double functions there. */
static float cl0(float v) { return v < 0 ? 0.f : v; }

static yo_col tonemap(yo_col c, const yg_render_opts* o)
{
    if (!o->tonemap) return c;
    const float e = -o->exposure, g = (o->gamma != 0.0f) ? (float)(1.0 / (double)o->gamma) : 0.0f;
    if (e == 0.f && g == 1.f) {
        if (o->clamp_rgb) {
            c.r = c.r < 0 ? 0.f : (c.r > 1 ? 1.f : c.r);
            c.g = c.g < 0 ? 0.f : (c.g > 1 ? 1.f : c.g);
            c.b = c.b < 0 ? 0.f : (c.b > 1 ? 1.f : c.b);
        }
        return c;
    }
    if (e != 0.f) {
        c.r = (float)(1.0 - exp((double)(cl0(c.r) * e)));
        c.g = (float)(1.0 - exp((double)(cl0(c.g) * e)));
        c.b = (float)(1.0 - exp((double)(cl0(c.b) * e)));
    }
    if (g != 1.f) {
        c.r = (float)pow((double)cl0(c.r), (double)g);
        c.g = (float)pow((double)cl0(c.g), (double)g);
        c.b = (float)pow((double)cl0(c.b), (double)g);
    }
    return c;
}

/* renderArea_t::checkResample, renderblock.cc:22-69: a pixel of the area's
 * image (W x H, the overlap border included) is resampled when any of its 8
 * neighbours (clamped to the area) differs by abscol2bri >= threshold. */
static float abscol2bri(yo_col a, yo_col b)
{
    /* color.h:71 on the difference (colorA_t operator-, color.h:293-296, in
     * float).  color.h includes only <cmath>, so its unqualified fabs is the C
     * library's fabs(double): the weighted sum is formed in double and
     * returned as float. */
    return (float)(0.299f * fabs((double)(a.r - b.r)) + 0.587f * fabs((double)(a.g - b.g)) +
                   0.114f * fabs((double)(a.b - b.b)));
}

static void check_resample(const yo_col* img, int W, int H, float threshold, unsigned char* mask)
{
    for (int i = 0; i < H; ++i) {
        const int im = i > 0 ? i - 1 : 0, ip = i + 1 < H ? i + 1 : H - 1;
        for (int j = 0; j < W; ++j) {
            const int jm = j > 0 ? j - 1 : 0, jp = j + 1 < W ? j + 1 : W - 1;
            const yo_col c = img[i * W + j];
            const int nb[8][2] = {{im, jm}, {im, j}, {im, jp}, {i, jm}, {i, jp}, {ip, jm}, {ip, j}, {ip, jp}};
            int need = 0;
            for (int q = 0; q < 8 && !need; q++) need = abscol2bri(c, img[nb[q][0] * W + nb[q][1]]) >= threshold;
            mask[i * W + j] = (unsigned char)need;
        }
    }
}

/* scene_t::render(renderArea_t&), scene.cc:724-816: the first pass (pixel
 * centres, or with AA_jitterfirst the area's n-th pixel at (RI_vdC(n),
 * RI_S(n))), then AA_passes passes that resample the pixels the contrast mask
 * marks with AA_minsamples camera rays each and fold them into a running
 * average.  The area keeps its own image, overlap border included; only the
 * real rectangle is written out. */
/* ====================================================== irradiance cache ==
 * pathLight_t with cache = on: lightSample_t and lightCache_t
 * (lightcache.h:14-176, lightcache.cc:48-239), cacheProxy_t
 * (cacheproxy.cc:10-184), pathLight_t::cached / interpolate and the weight
 * functions (pathlight.cc:409-723), and the fake pass of scene_t::render
 * (scene.cc:559-583, fakeRender scene.cc:825-854).  Restated one sample at a
 * time in the reference's order, with its list orders (move-to-front), its
 * generic-tree visiting order and libstdc++'s heap operations, since which
 * samples are found, and in which order ties fall, depends on them.
 * testRefinement (threshold < 1) is not restated: renders with it are refused. */

typedef struct yo_ls_s {
    yo_v3 N;
    yo_col color;
    float adist, M, precision;
    yo_v3 P, pP;
    float devaluated;
    int32_t next;       /* next sample of its box's list (front first), -1 at the end */
    int32_t visit;      /* fake-pass visit that took it */
} yo_ls;

typedef struct { int32_t x, y, z, head; } yo_lbox;

typedef struct yo_lcache_s {
    float cache_size, ycorr;
    yo_ls* s; int64_t n, cap;
    yo_lbox* b; int64_t nb, bcap;
    int64_t* ht; int64_t htcap;   /* (x, y, z) -> box index + 1 */
    int32_t* rank;                /* startUse: position in the generic tree's leaf order */
    /* candidates of a gather: grid over (pP.x, pP.y), cell gcell, entries in rank order */
    float gx0, gy0, gcell;
    int gnx, gny;
    int64_t* gstart;
    int32_t* gitem;
} yo_lcache;

/* foundSample_t (lightcache.h:38-43) */
typedef struct { int32_t S; float dis, weight; } yo_fs;

static uint64_t lc_hkey(int32_t x, int32_t y, int32_t z)
{
    uint64_t h = (uint64_t)(uint32_t)x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)y * 0xC2B2AE3D27D4EB4Full + (h >> 29);
    h ^= (uint64_t)(uint32_t)z * 0x165667B19E3779F9ull + (h >> 31);
    return h ^ (h >> 33);
}

static int64_t lc_find(const yo_lcache* C, int32_t x, int32_t y, int32_t z)
{
    if (!C->htcap) return -1;
    for (uint64_t q = lc_hkey(x, y, z) & (uint64_t)(C->htcap - 1);; q = (q + 1) & (uint64_t)(C->htcap - 1)) {
        const int64_t v = C->ht[q];
        if (!v) return -1;
        const yo_lbox* B = &C->b[v - 1];
        if (B->x == x && B->y == y && B->z == z) return v - 1;
    }
}

static int64_t lc_find_create(yo_lcache* C, int32_t x, int32_t y, int32_t z)
{
    int64_t f = lc_find(C, x, y, z);
    if (f >= 0) return f;
    if (2 * (C->nb + 1) > C->htcap) {
        const int64_t nc = C->htcap ? 2 * C->htcap : 1024;
        free(C->ht);
        C->ht = (int64_t*)calloc((size_t)nc, sizeof(int64_t));
        C->htcap = nc;
        for (int64_t i = 0; i < C->nb; i++) {
            uint64_t q = lc_hkey(C->b[i].x, C->b[i].y, C->b[i].z) & (uint64_t)(nc - 1);
            while (C->ht[q]) q = (q + 1) & (uint64_t)(nc - 1);
            C->ht[q] = i + 1;
        }
    }
    if (C->nb == C->bcap) {
        C->bcap = C->bcap ? 2 * C->bcap : 1024;
        C->b = (yo_lbox*)realloc(C->b, sizeof(yo_lbox) * (size_t)C->bcap);
    }
    yo_lbox* B = &C->b[C->nb];
    B->x = x; B->y = y; B->z = z; B->head = -1;
    uint64_t q = lc_hkey(x, y, z) & (uint64_t)(C->htcap - 1);
    while (C->ht[q]) q = (q + 1) & (uint64_t)(C->htcap - 1);
    C->ht[q] = ++C->nb;
    return C->nb - 1;
}

/* hash3d_t::getBox (hash3d.h:224-234): truncation, then one down for negative */
static void lc_getbox(float cell, yo_v3 p, int32_t* nx, int32_t* ny, int32_t* nz)
{
    *nx = (int32_t)(p.x / cell);
    *ny = (int32_t)(p.y / cell);
    *nz = (int32_t)(p.z / cell);
    if (p.x < 0.0f) (*nx)--;
    if (p.y < 0.0f) (*ny)--;
    if (p.z < 0.0f) (*nz)--;
}

/* lightCache_t::toPolar (lightcache.h:107-111): the pixel's screen position,
 * y scaled by 1/aspect, and the log of the distance travelled; log is the
 * C library's double log of the float distance */
static yo_v3 lc_to_polar(const rctx* R)
{
    return v3(R->scr_x, R->scr_y * R->lc->ycorr, (float)log((double)R->traveled));
}

/* pathLight_t::weight / weightNoPrec (pathlight.cc:409-465): prec = 1
 * subtracts twice the sample's precision from the distance (clamped at 0) */
static float lc_weight(const yo_ls* S, yo_v3 P, yo_v3 N, float maxweight, int prec)
{
    yo_v3 PP = v3_sub(P, S->P);
    const float avgOclu = S->adist;
    if (avgOclu == 0.0f) return 0;
    float D;
    if (prec) {
        D = (float)((double)v3_normlen(&PP) - (double)S->precision * 2.0);
        if (D < 0) D = 0;
    } else
        D = v3_normlen(&PP);
    const float a = D / avgOclu;
    const float cinv = (float)(1.000001 - (double)v3_dot(S->N, N));
    const float b = sqrtf(cinv);
    const float c = fabsf(v3_dot(N, PP));
    float f = S->devaluated * (a + ((b > c) ? b : c));
    if (f == 0) return maxweight;
    f = (float)(1.0 / (double)f);
    return (f > maxweight) ? maxweight : f;
}

/* lightCache_t::enoughFor (lightcache.cc:182-222): boxes cx, cx-1, cx+1 (and
 * so on per axis), each list from its front; the first sample within
 * cache_size whose weightNoPrec reaches wlimit moves to its list's front */
static int lc_enough_for(yo_lcache* C, yo_v3 P, yo_v3 N, yo_v3 pP, float wlimit)
{
    int32_t cx, cy, cz;
    lc_getbox(C->cache_size, pP, &cx, &cy, &cz);
    const float maxw = (float)((double)wlimit * 2.0);
    for (int i = cx; i <= cx + 1; i += (i == cx) ? -1 : ((i < cx) ? 2 : 1))
        for (int j = cy; j <= cy + 1; j += (j == cy) ? -1 : ((j < cy) ? 2 : 1))
            for (int k = cz; k <= cz + 1; k += (k == cz) ? -1 : ((k < cz) ? 2 : 1)) {
                const int64_t bi = lc_find(C, i, j, k);
                if (bi < 0) continue;
                yo_lbox* B = &C->b[bi];
                for (int32_t prev = -1, l = B->head; l >= 0; prev = l, l = C->s[l].next) {
                    const float pD = v3_len(v3_sub(pP, C->s[l].pP));
                    if (pD > C->cache_size) continue;
                    if (lc_weight(&C->s[l], P, N, maxw, 0) < wlimit) continue;
                    if (prev >= 0) {
                        C->s[prev].next = C->s[l].next;
                        C->s[l].next = B->head;
                        B->head = l;
                    }
                    return 1;
                }
            }
    return 0;
}

/* lightCache_t::insert (lightcache.cc:224-237): push_front into the box of
 * findBox(pP), which looks the box up again from its centre */
static void lc_insert(yo_lcache* C, yo_ls smp)
{
    int32_t nx, ny, nz;
    lc_getbox(C->cache_size, smp.pP, &nx, &ny, &nz);
    const float cs = C->cache_size;
    yo_v3 box = v3(cs * (float)nx, cs * (float)ny, cs * (float)nz);
    box.x = (float)((double)box.x + 0.5 * (double)cs);
    box.y = (float)((double)box.y + 0.5 * (double)cs);
    box.z = (float)((double)box.z + 0.5 * (double)cs);
    lc_getbox(cs, box, &nx, &ny, &nz);
    const int64_t bi = lc_find_create(C, nx, ny, nz);
    if (C->n == C->cap) {
        C->cap = C->cap ? 2 * C->cap : 4096;
        C->s = (yo_ls*)realloc(C->s, sizeof(yo_ls) * (size_t)C->cap);
    }
    smp.next = C->b[bi].head;
    C->s[C->n] = smp;
    C->b[bi].head = (int32_t)C->n;
    C->n++;
}

/* ---- startUse: buildGenericTree (bound.h:252-340) over the samples in the
 * cache's iteration order (boxes by x, y, z as std::map orders them; each
 * list from its front), leaf size 1; rank = position in the depth-first leaf
 * order, which is the order gObjectIterator_t hands candidates out */
static int lc_box_cmp(const void* a, const void* b)
{
    const yo_lbox *p = (const yo_lbox*)a, *q = (const yo_lbox*)b;
    if (p->x != q->x) return p->x < q->x ? -1 : 1;
    if (p->y != q->y) return p->y < q->y ? -1 : 1;
    if (p->z != q->z) return p->z < q->z ? -1 : 1;
    return 0;
}

typedef struct { float a[3], g[3]; } yo_gb;

/* path_calc_bound (lightcache.cc:11-33) */
static yo_gb lc_calc_bound(const yo_lcache* C, const int32_t* v, int64_t n)
{
    yo_gb b;
    if (n == 0) {
        for (int k = 0; k < 3; k++) b.a[k] = b.g[k] = 0;
        return b;
    }
    const yo_v3 p0 = C->s[v[0]].pP;
    b.a[0] = b.g[0] = p0.x; b.a[1] = b.g[1] = p0.y; b.a[2] = b.g[2] = p0.z;
    for (int64_t i = 0; i < n; i++) {
        const yo_v3 p = C->s[v[i]].pP;
        const float q[3] = {p.x, p.y, p.z};
        for (int k = 0; k < 3; k++) {
            if (q[k] > b.g[k]) b.g[k] = q[k];
            if (q[k] < b.a[k]) b.a[k] = q[k];
        }
    }
    return b;
}

static int lc_in_bound(const yo_v3 p, const yo_gb* b)
{
    return p.x >= b->a[0] && p.x <= b->g[0] && p.y >= b->a[1] && p.y <= b->g[1] && p.z >= b->a[2] &&
           p.z <= b->g[2];
}

/* appends the depth-first leaf order of the tree buildGenericTree makes of
 * v[0..n) to out */
static void lc_tree_order(const yo_lcache* C, const int32_t* v, int64_t n, int sx, int sy, int sz, int32_t* out,
                          int64_t* nout)
{
    if (n <= 1 || (sx && sy && sz)) {
        for (int64_t i = 0; i < n; i++) out[(*nout)++] = v[i];
        return;
    }
    const yo_gb bound = lc_calc_bound(C, v, n);
    const float lx = bound.g[0] - bound.a[0], ly = bound.g[1] - bound.a[1], lz = bound.g[2] - bound.a[2];
    int axis;
    if (((lx >= ly) || sy) && ((lx >= lz) || sz) && !sx) axis = 0;
    else if (((ly >= lx) || sx) && ((ly >= lz) || sz) && !sy) axis = 1;
    else axis = 2;
    float media = 0;
    for (int64_t i = 0; i < n; i++) {
        const yo_v3 p = C->s[v[i]].pP;
        media += (axis == 0) ? p.x : (axis == 1) ? p.y : p.z;
    }
    media /= (float)n;
    yo_gb bl = bound, br = bound;
    bl.g[axis] = media;
    br.a[axis] = media;
    int32_t* vl = (int32_t*)malloc(sizeof(int32_t) * (size_t)n * 3);
    int32_t *vr = vl + n, *vm = vl + 2 * n;
    int64_t nl = 0, nr = 0, nm = 0;
    for (int64_t i = 0; i < n; i++) {
        const yo_v3 p = C->s[v[i]].pP;
        if (lc_in_bound(p, &bl)) {
            if (lc_in_bound(p, &br)) vm[nm++] = v[i];
            else vl[nl++] = v[i];
        } else
            vr[nr++] = v[i];
    }
    if (nl == n || nr == n || nm == n) {
        lc_tree_order(C, v, n, sx || axis == 0, sy || axis == 1, sz || axis == 2, out, nout);
    } else if (nl == 0) {
        lc_tree_order(C, vr, nr, sx, sy, sz, out, nout);
        lc_tree_order(C, vm, nm, sx, sy, sz, out, nout);
    } else if (nr == 0) {
        lc_tree_order(C, vl, nl, sx, sy, sz, out, nout);
        lc_tree_order(C, vm, nm, sx, sy, sz, out, nout);
    } else {
        lc_tree_order(C, vl, nl, sx, sy, sz, out, nout);
        lc_tree_order(C, vr, nr, sx, sy, sz, out, nout);
        if (nm) lc_tree_order(C, vm, nm, sx, sy, sz, out, nout);
    }
    free(vl);
}

static const yo_lcache* g_lc_sort;
static int lc_item_cmp(const void* a, const void* b)
{
    const int32_t p = *(const int32_t*)a, q = *(const int32_t*)b;
    return (g_lc_sort->rank[p] < g_lc_sort->rank[q]) ? -1 : (g_lc_sort->rank[p] > g_lc_sort->rank[q]);
}

/* lightCache_t::startUse (lightcache.cc:48-57), plus the candidate grid */
static void lc_start_use(yo_lcache* C, float maxradius)
{
    yo_lbox* boxes = (yo_lbox*)malloc(sizeof(yo_lbox) * (size_t)(C->nb ? C->nb : 1));
    memcpy(boxes, C->b, sizeof(yo_lbox) * (size_t)C->nb);
    qsort(boxes, (size_t)C->nb, sizeof(yo_lbox), lc_box_cmp);
    int32_t* v = (int32_t*)malloc(sizeof(int32_t) * (size_t)(C->n ? C->n : 1));
    int64_t nv = 0;
    for (int64_t i = 0; i < C->nb; i++)
        for (int32_t l = boxes[i].head; l >= 0; l = C->s[l].next) v[nv++] = l;
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(C->n ? C->n : 1));
    int64_t no = 0;
    lc_tree_order(C, v, nv, 0, 0, 0, order, &no);
    free(C->rank);
    C->rank = (int32_t*)malloc(sizeof(int32_t) * (size_t)(C->n ? C->n : 1));
    for (int64_t i = 0; i < no; i++) C->rank[order[i]] = (int32_t)i;
    free(order);
    free(v);
    free(boxes);
    /* candidate grid */
    float lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (int64_t i = 0; i < C->n; i++) {
        const float q[2] = {C->s[i].pP.x, C->s[i].pP.y};
        for (int k = 0; k < 2; k++) {
            if (q[k] < lo[k]) lo[k] = q[k];
            if (q[k] > hi[k]) hi[k] = q[k];
        }
    }
    if (C->n == 0) lo[0] = lo[1] = hi[0] = hi[1] = 0;
    C->gcell = maxradius;
    C->gx0 = lo[0];
    C->gy0 = lo[1];
    C->gnx = (int)((hi[0] - lo[0]) / maxradius) + 1;
    C->gny = (int)((hi[1] - lo[1]) / maxradius) + 1;
    free(C->gstart);
    free(C->gitem);
    C->gstart = (int64_t*)calloc((size_t)C->gnx * C->gny + 1, sizeof(int64_t));
    C->gitem = (int32_t*)malloc(sizeof(int32_t) * (size_t)(C->n ? C->n : 1));
    int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * (size_t)(C->n ? C->n : 1));
    for (int64_t i = 0; i < C->n; i++) {
        int cx = (int)((C->s[i].pP.x - C->gx0) / C->gcell), cy = (int)((C->s[i].pP.y - C->gy0) / C->gcell);
        if (cx >= C->gnx) cx = C->gnx - 1;
        if (cy >= C->gny) cy = C->gny - 1;
        cell[i] = cy * C->gnx + cx;
        C->gstart[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < (int64_t)C->gnx * C->gny; c++) C->gstart[c + 1] += C->gstart[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * ((size_t)C->gnx * C->gny + 1));
    memcpy(fill, C->gstart, sizeof(int64_t) * ((size_t)C->gnx * C->gny + 1));
    for (int64_t i = 0; i < C->n; i++) C->gitem[fill[cell[i]]++] = (int32_t)i;
    g_lc_sort = C;
    for (int64_t c = 0; c < (int64_t)C->gnx * C->gny; c++)
        qsort(C->gitem + C->gstart[c], (size_t)(C->gstart[c + 1] - C->gstart[c]), sizeof(int32_t), lc_item_cmp);
    free(fill);
    free(cell);
}

/* std::push_heap / std::pop_heap of libstdc++ (the reference is built with
 * g++) with compareFound_f (a.weight > b.weight): the front is the smallest
 * weight; which of equal weights stays depends on these exact steps */
static int fs_comp(const yo_fs* a, const yo_fs* b) { return a->weight > b->weight; }

static void fs_sift_up(yo_fs* f, int64_t hole, int64_t top, yo_fs value)
{
    int64_t parent = (hole - 1) / 2;
    while (hole > top && fs_comp(&f[parent], &value)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = value;
}

static void fs_push_heap(yo_fs* f, int64_t len) { fs_sift_up(f, len - 1, 0, f[len - 1]); }

static void fs_pop_heap(yo_fs* f, int64_t len)
{
    if (len <= 1) return;
    const int64_t last = len - 1;
    yo_fs value = f[last];
    f[last] = f[0];
    /* __adjust_heap(first, 0, last, value) */
    const int64_t n = last;
    int64_t hole = 0, child = 0;
    while (child < (n - 1) / 2) {
        child = 2 * (child + 1);
        if (fs_comp(&f[child], &f[child - 1])) child--;
        f[hole] = f[child];
        hole = child;
    }
    if ((n & 1) == 0 && child == (n - 2) / 2) {
        child = 2 * (child + 1);
        f[hole] = f[child - 1];
        hole = child - 1;
    }
    fs_sift_up(f, hole, 0, value);
}

/* one candidate into the found heap (lightcache.cc:147-157, cacheproxy.cc:70-85) */
static void fs_offer(yo_fs* found, int64_t* nf, yo_fs t, unsigned limit)
{
    if (((uint64_t)*nf >= limit) && (t.weight < found[0].weight)) return;
    found[(*nf)++] = t;
    fs_push_heap(found, *nf);
    if ((uint64_t)*nf > limit) {
        fs_pop_heap(found, *nf);
        (*nf)--;
    }
}

/* lightCache_t::gatherSamples (lightcache.cc:111-180) with weight(); K
 * candidates at most in found (capacity K + 1) */
static float lc_gather(const yo_lcache* C, yo_v3 P, yo_v3 pP, yo_v3 N, yo_fs* found, int64_t* nf, unsigned K,
                       float* radius, float maxradius, unsigned minimun, float wlimit, int32_t* cand)
{
    float best = 0, adist = 0;
    float maxdist = 0, lastradius = 0;
    const float maxweight = (float)((double)wlimit * 2.5), minweight = (float)((double)wlimit * 0.6);
    unsigned reached = 0;
    int repeat = 1;
    *nf = 0;
    while (repeat) {
        /* the tree's candidates within the circle, in its visiting order */
        int64_t nc = 0;
        const int cx = (int)floorf((pP.x - C->gx0) / C->gcell), cy = (int)floorf((pP.y - C->gy0) / C->gcell);
        for (int y = cy - 1; y <= cy + 1; y++)
            for (int x = cx - 1; x <= cx + 1; x++) {
                if (x < 0 || y < 0 || x >= C->gnx || y >= C->gny) continue;
                const int64_t c = (int64_t)y * C->gnx + x;
                for (int64_t q = C->gstart[c]; q < C->gstart[c + 1]; q++) cand[nc++] = C->gitem[q];
            }
        g_lc_sort = C;
        qsort(cand, (size_t)nc, sizeof(int32_t), lc_item_cmp);
        for (int64_t q = 0; q < nc; q++) {
            const yo_ls* S = &C->s[cand[q]];
            const float pD = v3_len(v3_sub(pP, S->pP));
            if (pD >= *radius) continue;
            if (pD < lastradius) continue;
            reached++;
            yo_fs t;
            t.S = cand[q];
            t.dis = pD;
            t.weight = lc_weight(S, P, N, maxweight, 1);
            if (t.weight > best) {
                best = t.weight;
                adist = S->adist;
            }
            unsigned limit;
            if (t.weight <= wlimit) limit = minimun;
            else {
                if (pD > maxdist) maxdist = pD;
                limit = K;
            }
            if (!limit) continue;
            fs_offer(found, nf, t, limit);
        }
        float rrad = (*nf == 0 || found[0].dis == 0)
                         ? (float)(0.0001 * (double)adist)
                         : *radius * v3_len(v3_sub(C->s[found[0].S].P, P)) / found[0].dis;
        if (rrad == 0.0f) rrad = 1.0f;
        repeat = ((*nf == 0) || (((adist / rrad) > wlimit) && (reached < K)) || (best <= minweight)) &&
                 (*radius < maxradius);
        if (repeat) {
            lastradius = *radius;
            *radius *= 2;
        }
        if (*radius > maxradius) *radius = maxradius;
    }
    if ((maxdist > 0) && ((double)(maxdist / *radius) < (1.0 / sqrt(2.0)))) *radius = (float)((double)*radius * 0.9);
    return *nf ? found[0].weight : 0.0f;
}

static void lc_free(yo_lcache* C)
{
    if (!C) return;
    free(C->s); free(C->b); free(C->ht); free(C->rank); free(C->gstart); free(C->gitem);
    free(C);
}

/* pathLight_t::cached (pathlight.cc:676-700) at a camera ray's hit */
static yo_col lc_cached(rctx* R, int li, const yo_sp* sp, yo_v3 eye)
{
    const yg_light* L = &R->lights[li];
    if (col_energy(get_diffuse(R, sp, eye)) < 0.05f) return col(0, 0, 0);
    const yo_v3 N = face_forward(sp->Ng, sp->N, eye);
    const float desired = (float)(1.0 / (double)L->shadow_threshold);
    const float rq = (float)(1.0 / (double)(0 + 1));
    yo_col total = col(0, 0, 0);
    const yo_v3 pP = lc_to_polar(R);
    if (!lc_enough_for(R->lc, sp->P, N, pP, desired * rq)) {
        float H, M;
        const int prim = R->primary;
        R->primary = 0;
        total = path_take_sample(R, li, N, sp, &H, &M);
        R->primary = prim;
        yo_ls smp;
        memset(&smp, 0, sizeof(smp));
        smp.N = N; smp.color = total; smp.adist = H; smp.M = M;
        smp.precision = R->traveled * R->worldres;
        smp.P = sp->P; smp.pP = pP; smp.devaluated = 1.0f;
        smp.visit = (int32_t)R->pixnum;
        lc_insert(R->lc, smp);
        total = col(1, 1, 1);
    }
    return total;
}

/* pathLight_t::interpolate (pathlight.cc:543-599) with the area's
 * cacheProxy_t (cacheproxy.cc:17-87) */
static yo_col lc_interpolate(rctx* R, int li, const yo_sp* sp, yo_v3 eye)
{
    const yg_light* L = &R->lights[li];
    const yo_lcache* C = R->lc;
    const yo_col dif = get_diffuse(R, sp, eye);
    if (col_energy(dif) < 0.05f) return col(0, 0, 0);
    const yo_v3 N = face_forward(sp->Ng, sp->N, eye);
    const yo_v3 pP = lc_to_polar(R);
    const float desired = (float)(1.0 / (double)L->shadow_threshold);
    const float wlimit = (float)(0.8 * (double)desired);
    const float search_radius = (float)((double)L->cache_size * 2.0);
    const unsigned K = (unsigned)L->search;
    /* cacheProxy_t::gatherSamples: a new pixel resets the proxy */
    if (R->pixnum != R->px_id) {
        if (R->px_n > 600) R->px_n = 0;
        R->px_id = R->pixnum;
    }
    yo_fs* found = (yo_fs*)malloc(sizeof(yo_fs) * (K + 2));
    int64_t nf = 0;
    lc_gather(C, sp->P, pP, N, found, &nf, K, &R->px_radius, search_radius, 0, wlimit, R->lc_cand);
    const yo_ls* pool = C->s;
    if (nf == 0) {
        /* cacheProxy_t::newSearch: the proxy's own new samples */
        const float maxweight = (float)((double)wlimit * 2.5);
        for (int i = 0; i < R->px_n; i++) {
            yo_fs t;
            t.dis = v3_len(v3_sub(pP, R->px_created[i].pP));
            if (t.dis > search_radius) continue;
            t.S = i;
            t.weight = lc_weight(&R->px_created[i], sp->P, N, maxweight, 1);
            const unsigned limit = (t.weight <= wlimit) ? 0u : K;
            if (!limit) continue;
            fs_offer(found, &nf, t, limit);
        }
        pool = R->px_created;
    }
    float farest = 0;
    if (nf) farest = found[0].weight;
    if (nf == 1) farest = 0;
    else if (farest > wlimit) farest = wlimit;
    for (int64_t i = 0; i < nf; i++)
        found[i].weight = (float)((double)(found[i].weight - farest) * (1.0 - (double)(found[i].dis / search_radius)));
    yo_col total = col(0, 0, 0);
    float amount = 0;
    for (int64_t i = 0; i < nf; i++) {
        total = col_add(total, col_scale(pool[found[i].S].color, found[i].weight));
        amount += found[i].weight;
    }
    free(found);
    if (amount != 0) {
        amount = (float)(1.0 / (double)amount);
        total = col_scale(total, amount);
        return col_scale(col_mul(dif, total), L->power);
    }
    /* no sample: take one (pathlight.cc:585-597) */
    float H, M;
    const int prim = R->primary;
    R->primary = 0;
    const yo_col ncol = path_take_sample(R, li, N, sp, &H, &M);
    R->primary = prim;
    R->lc_new++;
    if (R->px_n == R->px_cap) {
        R->px_cap = R->px_cap ? 2 * R->px_cap : 64;
        R->px_created = (yo_ls*)realloc(R->px_created, sizeof(yo_ls) * (size_t)R->px_cap);
    }
    yo_ls smp;
    memset(&smp, 0, sizeof(smp));
    smp.N = N; smp.color = ncol; smp.adist = H; smp.M = M;
    smp.precision = R->traveled * R->worldres;
    smp.P = sp->P; smp.pP = pP; smp.devaluated = 1.0f;
    smp.next = -1; smp.visit = -1;
    R->px_created[R->px_n++] = smp;
    return col_scale(col_mul(dif, ncol), L->power);
}

static void render_area(rctx* R, const render_job* J, int ri)
{
    const yo_region* a = &J->regions[ri];
    const yg_render_opts* o = J->opts;
    /* a fresh renderState_t (and sampler context) per area, scene.cc:726 */
    R->ps.made = 0;
    for (int li = 0; li < J->nl; li++)
        if (J->lights[li].type == YG_LIGHT_PATH) {
            int nd = (J->lights[li].depth + 1) * 2, base = 2;
            for (int d = 0; d < nd; d++) { halton_init(&R->hseq[li][d], base); base = yo_next_prime(base); }
        }
    const int W = a->w, H = a->h, resx = J->cam->resx;
    const int passes = o->aa_passes > 0 ? o->aa_passes : 0;
    /* a fresh cacheProxy_t per area (cacheproxy.cc:10-15): radius = maxradius */
    for (int li = 0; li < J->nl; li++)
        if (J->lights[li].type == YG_LIGHT_PATH && J->lights[li].cache)
            R->px_radius = (float)((double)J->lights[li].cache_size * 2.0);
    R->px_id = -1;
    R->px_n = 0;
    yo_col* img = (yo_col*)malloc(sizeof(yo_col) * (size_t)W * H);
    unsigned int sc = 0;
    for (int i = a->y; i < a->y + a->h; ++i)
        for (int j = a->x; j < a->x + a->w; ++j) {
            float fx = 0.5f, fy = 0.5f;
            if (o->aa_jitterfirst && passes != 0) {
                ++sc;
                fx = yo_ri_vdc(sc, 0);
                fy = yo_ri_s(sc, 0);
            }
            yo_v3 ray = cam_ray(J->k, (float)j + fx, (float)i + fy);
            /* state.screenpos and pixelNumber (scene.cc:749-756) */
            R->scr_x = (float)(2.0 * (double)(((float)j + fx) / (float)resx) - 1.0);
            R->scr_y = (float)(1.0 - 2.0 * (double)(((float)i + fy) / (float)J->cam->resy));
            R->pixnum = (int64_t)j + (int64_t)i * resx;
            const int out = j >= a->rx && j < a->rx + a->rw && i >= a->ry && i < a->ry + a->rh;
            R->gpx_on = 0;
            if (R->adj && out) {
                const float* g = J->grad_rgb + 3 * ((int64_t)i * resx + j);
                R->gpx[0] = g[0]; R->gpx[1] = g[1]; R->gpx[2] = g[2];
                R->gpx_on = (g[0] != 0 || g[1] != 0 || g[2] != 0);
            }
            img[(i - a->y) * W + (j - a->x)] = tonemap(raytrace(R, J->k->pos, ray), o);
        }
    if (passes > 0) {
        unsigned char* mask = (unsigned char*)malloc((size_t)W * H);
        const int mins = o->aa_minsamples;
        float totsamdiv = (float)(mins * passes);
        if (totsamdiv != 0) totsamdiv = (float)(1.0 / (double)totsamdiv);
        for (int pass = 0; pass < passes; pass++) {
            check_resample(img, W, H, o->aa_threshold, mask);
            for (int i = a->y; i < a->y + a->h; ++i)
                for (int j = a->x; j < a->x + a->w; ++j) {
                    yo_col* px = &img[(i - a->y) * W + (j - a->x)];
                    if (!mask[(i - a->y) * W + (j - a->x)]) continue;
                    const unsigned int pixnum = (unsigned int)(j + i * resx);
                    yo_col tot = col(0, 0, 0);
                    int n = 0;
                    for (int ms = 0; ms < mins; ms++) {
                        const unsigned int cursam = (unsigned int)(pass * mins + ms);
                        /* 0.5 is a double literal: the offsets are formed in double */
                        const float fx = (float)(0.5 + (double)o->aa_pixelwidth *
                                                           ((double)yo_ri_lp(cursam + pixnum, 0) - 0.5));
                        const float fy = (float)(0.5 + (double)o->aa_pixelwidth *
                                                           ((double)((float)cursam * totsamdiv) - 0.5));
                        /* state.screenpos of the sample (scene.cc:798-799) */
                        R->scr_x = (float)(2.0 * (double)(((float)j + fx) / (float)resx) - 1.0);
                        R->scr_y = (float)(1.0 - 2.0 * (double)(((float)i + fy) / (float)J->cam->resy));
                        const yo_col c =
                            tonemap(raytrace(R, J->k->pos, cam_ray(J->k, (float)j + fx, (float)i + fy)), o);
                        tot = col_add(tot, c);
                        n++;
                    }
                    /* (mf*image + totcol) / (mf + n), colorA_t / f = * (1/f) (color.h:282-286) */
                    const float mf = (float)(pass * n + 1);
                    float den = mf + (float)n;
                    if (den != 0) den = (float)(1.0 / (double)den);
                    px->r = (mf * px->r + tot.r) * den;
                    px->g = (mf * px->g + tot.g) * den;
                    px->b = (mf * px->b + tot.b) * den;
                }
        }
        free(mask);
    }
    for (int i = a->ry; i < a->ry + a->rh; ++i)
        for (int j = a->rx; j < a->rx + a->rw; ++j) {
            const yo_col c = img[(i - a->y) * W + (j - a->x)];
            const int64_t px = (int64_t)i * resx + j;
            J->rgb[3 * px] = c.r; J->rgb[3 * px + 1] = c.g; J->rgb[3 * px + 2] = c.b;
        }
    free(img);
}

static void* render_worker(void* arg)
{
    render_job* J = (render_job*)arg;
    rctx R;
    rctx_init(&R, J->s, J->mats, J->nmat, J->mods, J->texs, J->lights, J->nl, J->obj_flags, J->nobj, J->opts);
    R.grand = J->grand;
    if (J->lc) {
        R.lc = J->lc;
        R.lc_phase = 2;
        R.worldres = J->worldres;
        R.lc_cand = (int32_t*)malloc(sizeof(int32_t) * (size_t)(J->lc->n + 1));
    }
    double* priv = NULL;
    const size_t ng = 3 * (size_t)J->nmat + 2 * (size_t)J->nmod + (size_t)J->nl;
    if (J->grad_rgb) {
        /* per-thread gradient sums, added to the caller's under the lock */
        priv = (double*)calloc(ng, sizeof(double));
        R.adj = 1;
        R.gmat = priv;
        R.gmod = priv + 3 * (size_t)J->nmat;
        R.glight = priv + 3 * (size_t)J->nmat + 2 * (size_t)J->nmod;
    }
    for (;;) {
        pthread_mutex_lock(&J->mu);
        int q = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (q >= J->nreg) break;
        render_area(&R, J, J->nreg - 1 - q);
    }
    pthread_mutex_lock(&J->mu);
    for (int k = 0; k < 5; k++) J->stats[k] += R.stats[k];
    J->lc_new += R.lc_new;
    if (priv) {
        for (int q = 0; q < 3 * J->nmat; q++) J->gmat[q] += priv[q];
        if (J->gmod)
            for (int q = 0; q < 2 * J->nmod; q++) J->gmod[q] += priv[3 * (size_t)J->nmat + q];
        for (int q = 0; q < J->nl; q++) J->glight[q] += priv[3 * (size_t)J->nmat + 2 * (size_t)J->nmod + q];
    }
    pthread_mutex_unlock(&J->mu);
    free(priv);
    rctx_free(&R);
    return NULL;
}

/* The cache's fake pass (scene.cc:559-583, fakeRender scene.cc:825-854):
 * every area in processing order, a fresh renderState_t (samplers) per area,
 * every pixel (overlap border included) through its centre; the pathlight's
 * cached() at the camera ray's hit.  The other lights' terms of the fake
 * pass are discarded by the reference and have no side effect (area lights,
 * whose draws would, are refused), so they are not evaluated; the fake
 * pass's camera rays are not counted (the render pass traces the same ones). */
static void lc_fake_pass(rctx* R, const render_job* J, int li)
{
    const int resx = J->cam->resx, resy = J->cam->resy;
    R->lc_phase = 1;
    for (int q = 0; q < J->nreg; q++) {
        const yo_region* a = &J->regions[J->nreg - 1 - q];
        for (int l = 0; l < J->nl; l++)
            if (J->lights[l].type == YG_LIGHT_PATH) {
                int nd = (J->lights[l].depth + 1) * 2, base = 2;
                for (int d = 0; d < nd; d++) { halton_init(&R->hseq[l][d], base); base = yo_next_prime(base); }
            }
        for (int i = a->y; i < a->y + a->h; ++i)
            for (int j = a->x; j < a->x + a->w; ++j) {
                /* fakeRender forms the screen position with double literals */
                R->scr_x = (float)(2.0 * (((double)j + 0.5) / (double)(float)resx) - 1.0);
                R->scr_y = (float)(1.0 - 2.0 * (((double)i + 0.5) / (double)(float)resy));
                R->pixnum = (int64_t)j + (int64_t)i * resx;
                const yo_v3 ray = cam_ray(J->k, (float)j + 0.5f, (float)i + 0.5f);
                yo_sp sp;
                const int hit = first_hit(R, J->k->pos, ray, &sp, 0);
                if (hit) R->traveled = sp.Z;
                yo_scene* S = (yo_scene*)R->s;
                if (S->lc_rec) {
                    if (S->lc_nvis == S->lc_viscap) {
                        S->lc_viscap = S->lc_viscap ? 2 * S->lc_viscap : 65536;
                        S->lc_visits = (float*)realloc(S->lc_visits, sizeof(float) * 13 * (size_t)S->lc_viscap);
                    }
                    float* o = S->lc_visits + 13 * S->lc_nvis++;
                    memset(o, 0, sizeof(float) * 13);
                    o[0] = (float)R->pixnum;
                    o[1] = (float)q;
                    if (hit) {
                        const yo_v3 N = face_forward(sp.Ng, sp.N, v3_sub(J->k->pos, sp.P));
                        const yo_v3 pP = lc_to_polar(R);
                        o[2] = col_energy(get_diffuse(R, &sp, v3_sub(J->k->pos, sp.P))) < 0.05f ? 0.0f : 1.0f;
                        o[3] = sp.P.x; o[4] = sp.P.y; o[5] = sp.P.z;
                        o[6] = N.x; o[7] = N.y; o[8] = N.z;
                        o[9] = pP.x; o[10] = pP.y; o[11] = pP.z;
                        o[12] = (float)sp.mat;
                    }
                }
                if (!hit) continue;
                (void)lc_cached(R, li, &sp, v3_sub(J->k->pos, sp.P));
            }
    }
    R->lc_phase = 0;
}

/* scene_t::render(colorOutput_t&) mono loop, scene.cc:538-600; per area
 * scene_t::render(renderArea_t&) first pass, scene.cc:724-775.  Areas are
 * independent unless an area light draws from the global ourRandom() stream,
 * so without area lights they may be rendered by several threads. */
int yo_render_mt(yo_scene* s, const yg_camera* cam, const yg_material* mats, int nmat,
                 const yg_modulator* mods, int nmod, const yg_texture* texs, int ntex,
                 const yg_light* lights, int nl, const int32_t* obj_flags, int nobj,
                 const yg_render_opts* opts, float* rgb, uint64_t* stats, int nthreads)
{
    (void)nmod; (void)ntex;
    for (int li = 0; li < nl; li++)
        if (lights[li].type == YG_LIGHT_GLOBALPHOTON) {
            if (!s->gm.built || s->gm.light != li) return -3;  /* map not built */
            /* one thread: the gathers carry a radius from call to call, and
             * photonSampler_t (pathlight samples > 96) draws its Halton starts
             * from ourRandomI() in area order */
            nthreads = 1;
        }
    int sequential = 0;
    for (int li = 0; li < nl; li++)
        if (lights[li].type == YG_LIGHT_AREA) sequential = 1;
    /* the irradiance cache: no refinement (threshold >= 1), no AA passes, no
     * lights whose fake-pass terms would draw from ourRandom(), no global map */
    int cli = -1;
    for (int li = 0; li < nl; li++)
        if (lights[li].type == YG_LIGHT_PATH && lights[li].cache) {
            if (cli >= 0) return -7;  /* "Several cached pathlights" (pathlight.cc:41-45) */
            cli = li;
        }
    if (cli >= 0) {
        if (lights[cli].threshold < 1.0f || opts->aa_passes > 0) return -7;
        for (int li = 0; li < nl; li++)
            if ((lights[li].type == YG_LIGHT_AREA && lights[li].use_in_render && !lights[li].dummy) ||
                lights[li].type == YG_LIGHT_GLOBALPHOTON)
                return -7;
        if (!(lights[cli].cache_size > 0) || !(lights[cli].shadow_threshold > 0) || lights[cli].search < 3) return -2;
        /* the fake pass's screen position is formed in double, the render
         * pass's in float: "window" textures stay off with the cache */
        for (int m = 0; m < nmat; m++)
            for (int k = 0; k < mats[m].mod_count; k++)
                if (mods[mats[m].mod_first + k].texco == YG_TEXCO_WIN) return -7;
    }
    if (nthreads < 1 || sequential) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    yo_cam k;
    cam_setup(cam, &k);
    yo_region* regions = NULL;
    int nreg = yo_block_regions(cam->resx, cam->resy, opts->block > 0 ? opts->block : 64,
                                opts->rand_seed, &regions);
    render_job J;
    memset(&J, 0, sizeof(J));
    J.s = s; J.cam = cam; J.mats = mats; J.nmat = nmat; J.mods = mods; J.texs = texs;
    J.lights = lights; J.nl = nl; J.obj_flags = obj_flags; J.nobj = nobj; J.opts = opts;
    J.k = &k; J.regions = regions; J.nreg = nreg; J.rgb = rgb; J.next = 0;
    pthread_mutex_init(&J.mu, NULL);
    /* glibc rand() continues after the block shuffle's draws: guessLight
     * (area lights with psamples > 0) takes its draws from there */
    yo_grand grand;
    grand_init(&grand, opts->rand_seed, nreg);
    J.grand = &grand;
    if (cli >= 0) {
        /* pathLight_t::init and the fake pass, then postInit's startUse */
        yo_lcache* C = (yo_lcache*)calloc(1, sizeof(yo_lcache));
        C->cache_size = lights[cli].cache_size;
        const float aspect = (float)cam->resx / (float)cam->resy;
        C->ycorr = (float)(1.0 / (double)aspect);
        J.lc = C;
        J.worldres = (float)((1.0 / (double)(float)cam->resx) / (double)cam->focal);
        rctx F;
        rctx_init(&F, s, mats, nmat, mods, texs, lights, nl, obj_flags, nobj, opts);
        F.lc = C;
        F.worldres = J.worldres;
        lc_fake_pass(&F, &J, cli);
        for (int q = 0; q < 5; q++) J.stats[q] += F.stats[q];
        rctx_free(&F);
        lc_start_use(C, (float)((double)lights[cli].cache_size * 2.0));
    }
    pthread_t th[64];
    if (nthreads == 1) render_worker(&J);
    else {
        for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, render_worker, &J);
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    }
    pthread_mutex_destroy(&J.mu);
    if (stats) for (int q = 0; q < 5; q++) stats[q] = J.stats[q];
    free(regions);
    if (J.lc) {
        lc_free(s->lc_last);
        s->lc_last = J.lc;
        s->lc_new = J.lc_new;
    }
    return 0;
}

void yo_cache_release(yo_scene* s)
{
    lc_free(s->lc_last);
    s->lc_last = NULL;
    free(s->lc_visits);
    s->lc_visits = NULL;
    s->lc_nvis = s->lc_viscap = 0;
}

/* Test hook: record the next cache render's fake-pass visits, 13 floats
 * each: pixel number, area (processing order), qualifies (hit, diffuse
 * energy >= 0.05), P, face-forwarded N, pP, material */
void yo_cache_record(yo_scene* s, int on)
{
    s->lc_rec = on;
    s->lc_nvis = 0;
}

int64_t yo_cache_visits(yo_scene* s, float* out, int64_t cap)
{
    if (out) memcpy(out, s->lc_visits, sizeof(float) * 13 * (size_t)(cap < s->lc_nvis ? cap : s->lc_nvis));
    return s->lc_nvis;
}

/* Test hook: gatherSamples' offer of candidates with weights w[0..n) (ids
 * 0..n-1, in that order) into a found heap of at most `limit`; out: the
 * final heap array (id, weight) in its array order; returns its size */
int64_t yo_heap_offer_trace(const float* w, int64_t n, unsigned limit, float* out)
{
    yo_fs* f = (yo_fs*)malloc(sizeof(yo_fs) * ((size_t)limit + 2));
    int64_t nf = 0;
    for (int64_t i = 0; i < n; i++) {
        yo_fs t;
        t.S = (int32_t)i;
        t.dis = 0;
        t.weight = w[i];
        fs_offer(f, &nf, t, limit);
    }
    for (int64_t i = 0; i < nf; i++) {
        out[2 * i] = (float)f[i].S;
        out[2 * i + 1] = f[i].weight;
    }
    free(f);
    return nf;
}

/* The samples of the last cache render: 18 floats each, in insertion order:
 * P, N, pP, colour, adist, M, precision, generic-tree rank, fake-pass visit
 * (pixel number), devaluated.  Returns the count; *new_samples: samples the
 * render pass took (interpolate found none). */
int64_t yo_cache_read(yo_scene* s, float* out, int64_t cap, uint64_t* new_samples)
{
    const yo_lcache* C = s->lc_last;
    if (new_samples) *new_samples = s->lc_new;
    if (!C) return 0;
    for (int64_t i = 0; i < C->n && i < cap; i++) {
        const yo_ls* S = &C->s[i];
        float* o = out + 18 * i;
        o[0] = S->P.x; o[1] = S->P.y; o[2] = S->P.z;
        o[3] = S->N.x; o[4] = S->N.y; o[5] = S->N.z;
        o[6] = S->pP.x; o[7] = S->pP.y; o[8] = S->pP.z;
        o[9] = S->color.r; o[10] = S->color.g; o[11] = S->color.b;
        o[12] = S->adist; o[13] = S->M; o[14] = S->precision;
        o[15] = (float)C->rank[i]; o[16] = (float)S->visit; o[17] = S->devaluated;
    }
    return C->n;
}

/* scene_t::raytrace of camera rays through camera-plane points (px[q], py[q])
 * (camera_t::shootRay coordinates), one after another.  Test hook for scenes
 * whose lights keep no per-area state (sun lights, photon lights whose map
 * yo_photon_build made, background): returns -6 for path and area lights. */
int yo_trace_camera(yo_scene* s, const yg_camera* cam, const yg_material* mats, int nmat,
                    const yg_modulator* mods, const yg_texture* texs, const yg_light* lights, int nl,
                    const int32_t* obj_flags, int nobj, const yg_render_opts* opts, const float* px,
                    const float* py, int64_t n, float* rgb)
{
    for (int li = 0; li < nl; li++)
        if (lights[li].type != YG_LIGHT_SUN && lights[li].type != YG_LIGHT_PHOTON) return -6;
    for (int li = 0; li < nl; li++)
        if (lights[li].type == YG_LIGHT_PHOTON && !s->pm.data) return -3;  /* photon map not built */
    rctx R;
    rctx_init(&R, s, mats, nmat, mods, texs, lights, nl, obj_flags, nobj, opts);
    yo_cam k;
    cam_setup(cam, &k);
    for (int64_t q = 0; q < n; q++) {
        const yo_col c = raytrace(&R, k.pos, cam_ray(&k, px[q], py[q]));
        rgb[3 * q] = c.r; rgb[3 * q + 1] = c.g; rgb[3 * q + 2] = c.b;
    }
    rctx_free(&R);
    return 0;
}

int yo_render(yo_scene* s, const yg_camera* cam, const yg_material* mats, int nmat,
              const yg_modulator* mods, int nmod, const yg_texture* texs, int ntex,
              const yg_light* lights, int nl, const int32_t* obj_flags, int nobj,
              const yg_render_opts* opts, float* rgb, uint64_t* stats)
{
    return yo_render_mt(s, cam, mats, nmat, mods, nmod, texs, ntex, lights, nl, obj_flags, nobj,
                        opts, rgb, stats, 1);
}

/* FP64 reference gradient of a loss through the image (see yg_render_adjoint):
 * renders like yo_render and accumulates, per output pixel,
 * dLoss/dparams = sum over the estimator's terms of grad_rgb * d(term)/d(param)
 * in double precision.  gmat: nmat*3, gmod: nmod*2, glight: nl (zeroed here).
 * nthreads > 1 renders areas on several threads (each summing its own
 * gradient; the sums are added at the end, so the last bits depend on the
 * areas' order); scenes with area lights run on one. */
int yo_render_adjoint(yo_scene* s, const yg_camera* cam, const yg_material* mats, int nmat,
                      const yg_modulator* mods, int nmod, const yg_texture* texs, int ntex,
                      const yg_light* lights, int nl, const int32_t* obj_flags, int nobj,
                      const yg_render_opts* opts, const float* grad_rgb, float* rgb, double* gmat,
                      double* gmod, double* glight, int nthreads)
{
    if (opts->aa_passes > 0) return -5; /* antialiasing passes are not differentiated */
    for (int k = 0; k < nmod; k++)
        if (mods[k].texco != YG_TEXCO_GLOBAL) return -7; /* hit-point coordinates only */
    for (int k = 0; k < ntex; k++)
        if (texs[k].type == YG_TEX_MUSGRAVE && texs[k].mtype != YG_MUSGRAVE_FBM) return -7; /* fBm only */
    for (int li = 0; li < nl; li++)
        if (lights[li].type == YG_LIGHT_GLOBALPHOTON) return -8; /* no irradiance-map derivatives */
    for (int li = 0; li < nl; li++)
        if (lights[li].type == YG_LIGHT_AREA) nthreads = 1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    memset(gmat, 0, sizeof(double) * 3 * (size_t)nmat);
    if (gmod) memset(gmod, 0, sizeof(double) * 2 * (size_t)nmod);
    memset(glight, 0, sizeof(double) * (size_t)nl);
    yo_cam k;
    cam_setup(cam, &k);
    yo_region* regions = NULL;
    int nreg = yo_block_regions(cam->resx, cam->resy, opts->block > 0 ? opts->block : 64,
                                opts->rand_seed, &regions);
    render_job J;
    memset(&J, 0, sizeof(J));
    J.s = s; J.cam = cam; J.mats = mats; J.nmat = nmat; J.mods = mods; J.texs = texs;
    J.lights = lights; J.nl = nl; J.obj_flags = obj_flags; J.nobj = nobj; J.opts = opts;
    J.k = &k; J.regions = regions; J.nreg = nreg; J.rgb = rgb; J.next = 0;
    J.grad_rgb = grad_rgb; J.gmat = gmat; J.gmod = gmod; J.glight = glight; J.nmod = gmod ? nmod : 0;
    pthread_mutex_init(&J.mu, NULL);
    pthread_t th[64];
    if (nthreads == 1) render_worker(&J);
    else {
        for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, render_worker, &J);
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    }
    pthread_mutex_destroy(&J.mu);
    free(regions);
    return 0;
}
