/*
 * FP64 oracle: scene_t::raytrace of camera rays (scene.cc:200-258) with the
 * same lights and estimators as yo_render.c -- sun and area lights, and the
 * pathlight's full-GI estimator with its energy cut-off and adaptive
 * repetitions (pathlight.cc:223-403) -- evaluated in double precision.
 * TEST INFRASTRUCTURE ONLY (see yaf_oracle.h).
 *
 * Two sample sources:
 *   YO_FP64_REFERENCE  the reference's own streams over the whole image in
 *                      render-area order: per-area Halton sequences for the
 *                      pathlight (a fresh sampler per area, scene.cc:726) and
 *                      the global ourRandom() stream for area-light jitter.
 *                      The result differs from the FP32 restatement by float
 *                      rounding only.
 *   YO_FP64_CONVERGED  per listed pixel, `samples` samples of every sampled
 *                      light: pathlight sample i of a pixel is the Halton
 *                      point of index offset + i + 1 (bases 2, 3, 5, ... per
 *                      dimension, offset hashed from the pixel), and area
 *                      lights take a fillQuad grid of `samples` points
 *                      jittered from a per-pixel splitmix64 stream.  Same
 *                      estimator, more samples, no float rounding: the
 *                      converged image a rendered image is measured against.
 *
 * Geometry: vertices are the scene's floats widened to double, face normals
 * are recomputed in double, the BVH is the oracle's canonical FP64 tree.
 * Textured albedos come from the FP32 modulator restatement at the hit point
 * rounded to float (the noise tables are single precision); untextured
 * albedos are the material colour.
 */
#include <float.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include "yo_internal.h"

yo_col yo_modulate(const yg_modulator* m, const yg_texture* texs, yo_col C, yo_v3 P);

/* ------------------------------------------------------- double vectors -- */
typedef struct { double x, y, z; } d3;
typedef struct { double r, g, b; } dc;

static inline d3 D3(double x, double y, double z) { d3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline d3 dld(const float* p) { return D3(p[0], p[1], p[2]); }
static inline d3 dldd(const double* p) { return D3(p[0], p[1], p[2]); }
static inline d3 dadd(d3 a, d3 b) { return D3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 dsub(d3 a, d3 b) { return D3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 dneg(d3 a) { return D3(-a.x, -a.y, -a.z); }
static inline d3 dmul(d3 a, double f) { return D3(f * a.x, f * a.y, f * a.z); }
static inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 dcross(d3 a, d3 b)
{
    return D3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double dlen(d3 a) { return sqrt(ddot(a, a)); }
static inline d3 dnorm(d3 a)
{
    const double l = ddot(a, a);
    return (l != 0) ? dmul(a, 1.0 / sqrt(l)) : a;
}
/* FACE_FORWARD, shader.h:157 */
static inline d3 dff(d3 ng, d3 n, d3 i) { return (ddot(ng, i) < 0) ? dneg(n) : n; }
/* createCS, vector3d.h:240-258 */
static void dcreate_cs(d3 n, d3* u, d3* v)
{
    if (n.x == 0 && n.y == 0) {
        *u = (n.z < 0) ? D3(-1, 0, 0) : D3(1, 0, 0);
        *v = D3(0, 1, 0);
    } else {
        const double d = 1.0 / sqrt(n.y * n.y + n.x * n.x);
        *u = D3(n.y * d, -n.x * d, 0);
        *v = dcross(n, *u);
    }
}

static inline dc DC(double r, double g, double b) { dc c; c.r = r; c.g = g; c.b = b; return c; }
static inline dc cadd(dc a, dc b) { return DC(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline dc cmul(dc a, dc b) { return DC(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline dc cscale(dc a, double f) { return DC(f * a.r, f * a.g, f * a.b); }
static inline double cenergy(dc c) { return (c.r + c.g + c.b) * 0.333333; }

/* ------------------------------------------------------------ samplers -- */
typedef struct { unsigned base; double inv_base, value; } dhalton;

/* Halton::getNext, mcqmc.h:35-49, without the float rounding of the result */
static double dhalton_next(dhalton* h)
{
    const double r = 1.0 - h->value - 0.0000000001;
    if (h->inv_base < r)
        h->value += h->inv_base;
    else {
        double hh, hv = h->inv_base;
        do { hh = hv; hv *= h->inv_base; } while (hv >= r);
        h->value += hh + hv - 1.0;
    }
    return h->value;
}

/* radical inverse of index k in base b, in double */
static double radinv(uint64_t k, unsigned b)
{
    const double inv = 1.0 / (double)b;
    double f = inv, r = 0;
    while (k) {
        r += (double)(k % b) * f;
        k /= b;
        f *= inv;
    }
    return r;
}

static uint64_t splitmix(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

#define MAXDIM 32

/* per-thread sample state */
typedef struct {
    int32_t* lcg;            /* reference: the shared ourRandom() state */
    dhalton hs[MAXDIM];      /* reference: the path light's per-area sequences */
    uint64_t rng;            /* converged: jitter stream of the pixel */
    uint64_t hoff, hcount;   /* converged: Halton index offset, samples taken */
    uint64_t k;              /* converged: Halton index of the current path */
} fsamp;

/* ------------------------------------------------------------- context -- */
typedef struct {
    int samples;
    d3* points; d3* jit_h; d3* jit_v;
    d3 direction, from, corner, toX, toY;
    dc pc;                   /* colour * power */
} darea;

typedef struct {
    const yo_scene* s;
    double* tn;              /* face normals in double, 3 per triangle */
    const yg_material* mats;
    const yg_modulator* mods;
    const yg_texture* texs;
    const yg_light* lights;
    int nl;
    const int32_t* oflags;   /* per object, bit0 cast shadows (NULL: all cast) */
    int nobj;
    yg_render_opts o;
    int mode, samples;
    unsigned bases[MAXDIM];
    d3* sun_dir; dc* sun_col; darea* areas;
} fctx;

/* ------------------------------------------------------------ geometry -- */
static inline d3 vert(const yo_scene* s, int64_t tri, int k)
{
    return dld(s->verts + 3 * (int64_t)s->idx[3 * tri + k]);
}

static int dbox(const yo_node* nd, const double* o, const double* d, double t0, double t1,
                double* tnear)
{
    for (int k = 0; k < 3; k++) {
        if (d[k] == 0.0) {
            if (o[k] < nd->lo[k] || o[k] > nd->hi[k]) return 0;
            continue;
        }
        double ta = (nd->lo[k] - o[k]) / d[k], tb = (nd->hi[k] - o[k]) / d[k];
        if (ta > tb) { const double x = ta; ta = tb; tb = x; }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
    }
    *tnear = t0;
    return t0 - 1e-9 * fabs(t0) - 1e-12 <= t1 + 1e-9 * fabs(t1) + 1e-12;
}

/* triangle_t::hit (triangle.h:48-58) and intersect (triangle.h:83-86) in
 * double; mode 0 closest (tmin <= t < best, ties to the lower id), 1 any
 * (tmin < t < tmax); the object filter of yo_query */
static int dquery(const fctx* F, d3 o, d3 d, double tmin, double tmax, int mode, int filter_obj,
                  int include, int shadow, int32_t* prim, double* t)
{
    const yo_scene* s = F->s;
    int32_t best = -1;
    double bt = tmax;
    *prim = -1; *t = tmax;
    if (s->nt == 0) return 0;
    const double od[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const yo_node* nd = &s->nodes[stack[--sp]];
        const double hi = isinf(bt) ? DBL_MAX : bt;
        double tn, tl, tr;
        if (!dbox(nd, od, dd, tmin, hi, &tn)) continue;
        if (nd->left >= 0) {
            /* near child on top */
            const int hl = dbox(&s->nodes[nd->left], od, dd, tmin, hi, &tl);
            const int hr = dbox(&s->nodes[nd->right], od, dd, tmin, hi, &tr);
            if (hl && hr) {
                if (tl <= tr) { stack[sp++] = nd->right; stack[sp++] = nd->left; }
                else { stack[sp++] = nd->left; stack[sp++] = nd->right; }
            } else if (hl) stack[sp++] = nd->left;
            else if (hr) stack[sp++] = nd->right;
            continue;
        }
        for (int32_t i = nd->start; i < nd->start + nd->count; i++) {
            const int32_t tri = s->order[i];
            if (shadow && F->oflags) {  /* shadow queries skip objects that cast no shadows */
                const int32_t ob = s->tri_obj[tri];
                if (ob >= 0 && ob < F->nobj && !(F->oflags[ob] & 1)) continue;
            }
            if (filter_obj >= 0) {
                const int same = s->tri_obj[tri] == filter_obj;
                if (include ? !same : same) continue;
            }
            const d3 a = vert(s, tri, 0), b = vert(s, tri, 1), c = vert(s, tri, 2);
            const d3 n = dldd(F->tn + 3 * (int64_t)tri);
            const d3 va = dsub(a, o), vb = dsub(b, o), vc = dsub(c, o);
            const d3 r = (ddot(d, n) < 0) ? dneg(d) : d;
            if (ddot(dcross(va, vb), r) < 0) continue;
            if (ddot(dcross(vb, vc), r) < 0) continue;
            if (ddot(dcross(vc, va), r) < 0) continue;
            const double rt = ddot(n, dsub(a, o)) / ddot(n, d);
            if (mode == 0) {
                /* coplanar faces tie to the lower id, as exact float ties do
                 * in yo_query; in double their distances differ by rounding */
                const double tol = 1e-12 * fabs(bt);
                if (rt >= tmin &&
                    (best < 0 ? rt < bt : (rt < bt - tol || (rt <= bt + tol && tri < best)))) {
                    bt = rt;
                    best = tri;
                }
            } else if (rt > tmin && rt < tmax) {
                *prim = tri;
                *t = rt;
                return 1;
            }
        }
    }
    *prim = best;
    *t = bt;
    return best >= 0;
}

/* ----------------------------------------------------- surface + queries -- */
typedef struct {
    d3 P, N, Ng, NU, NV;
    int32_t prim, obj, mat;
} dsp;

/* meshObject_t::shoot + triangle_t::getSurface (mesh.cc:517-577) */
static void make_sp(const fctx* F, int32_t prim, d3 f, d3 ray, double Z, dsp* sp)
{
    sp->prim = prim;
    sp->P = dadd(f, dmul(ray, Z));
    sp->Ng = dldd(F->tn + 3 * (int64_t)prim);
    sp->N = sp->Ng;
    const yo_scene* s = F->s;
    if (s->tri_smooth && s->tri_smooth[prim]) {
        /* getSurface's interpolated shading normal (triangle.cc:142-160) in double:
         * barycentric weights of the hit point (dominant-axis projection) */
        d3 a = {0, 0, 0}, b = a, c = a;
        for (int k = 0; k < 3; k++) {
            (&a.x)[k] = s->verts[3 * (int64_t)s->idx[3 * (int64_t)prim] + k];
            (&b.x)[k] = s->verts[3 * (int64_t)s->idx[3 * (int64_t)prim + 1] + k];
            (&c.x)[k] = s->verts[3 * (int64_t)s->idx[3 * (int64_t)prim + 2] + k];
        }
        const d3 n = sp->Ng, h = sp->P;
        const double nx = fabs(n.x), ny = fabs(n.y), nz = fabs(n.z);
        int u = 0, v = 1;
        if (nx >= ny && nx >= nz) { u = 1; v = 2; }
        else if (ny >= nx && ny >= nz) { u = 0; v = 2; }
        const double uu0 = (&h.x)[u] - (&a.x)[u], vv0 = (&h.x)[v] - (&a.x)[v];
        const double uu1 = (&b.x)[u] - (&a.x)[u], vv1 = (&b.x)[v] - (&a.x)[v];
        const double uu2 = (&c.x)[u] - (&a.x)[u], vv2 = (&c.x)[v] - (&a.x)[v];
        const double den = uu1 * vv2 - vv1 * uu2;
        const double fb = (uu0 * vv2 - vv0 * uu2) / den, fc = (uu1 * vv0 - vv1 * uu0) / den, fa = 1 - fb - fc;
        const float* vn = s->tri_vn + 9 * (int64_t)prim;
        d3 nn = {vn[0] * fa + vn[3] * fb + vn[6] * fc, vn[1] * fa + vn[4] * fb + vn[7] * fc,
                 vn[2] * fa + vn[5] * fb + vn[8] * fc};
        sp->N = dnorm(nn);
    }
    dcreate_cs(sp->N, &sp->NU, &sp->NV);
    sp->obj = F->s->tri_obj[prim];
    sp->mat = F->s->tri_mat[prim];
}

/* scene_t::firstHit (scene.cc:688-721): closest hit from from + ray*min_raydis,
 * Z > 0; shadow: its shadow flag (objects that cast no shadows are skipped) */
static int first_hit(const fctx* F, d3 from, d3 ray, dsp* sp, int shadow)
{
    const d3 f = dadd(from, dmul(ray, (double)F->o.min_raydis));
    int32_t p;
    double t;
    dquery(F, f, ray, 0.0, INFINITY, 0, -1, 0, shadow, &p, &t);
    if (p < 0 || !(t > 0)) return 0;
    make_sp(F, p, f, ray, t, sp);
    return 1;
}

/* scene_t::isShadowed (scene.cc:107-185): own object from p + ray*self_bias,
 * the others from p + ray*min_raydis, any hit with 0 < t < dist occludes */
static int is_shadowed(const fctx* F, d3 p, d3 l, int self_obj, int to_point)
{
    d3 ray = to_point ? dsub(l, p) : l;
    const double dist = to_point ? dlen(ray) : INFINITY;
    ray = dnorm(ray);
    const d3 self = dadd(p, dmul(ray, (double)F->o.self_bias));
    const d3 q = dadd(p, dmul(ray, (double)F->o.min_raydis));
    int32_t pr;
    double tt;
    if (self_obj >= 0) {
        if (dquery(F, self, ray, 0.0, dist, 1, self_obj, 1, 1, &pr, &tt)) return 1;
        return dquery(F, q, ray, 0.0, dist, 1, self_obj, 0, 1, &pr, &tt);
    }
    return dquery(F, q, ray, 0.0, dist, 1, -1, 0, 1, &pr, &tt);
}

/* ------------------------------------------------------------- shading -- */
/* genericShader_t::getDiffuse (basicshaders.cc:191-204) */
static dc albedo(const fctx* F, const dsp* sp)
{
    const yg_material* m = &F->mats[sp->mat];
    if (m->mod_count == 0) return DC(m->color[0], m->color[1], m->color[2]);
    const yo_v3 P = v3((float)sp->P.x, (float)sp->P.y, (float)sp->P.z);
    yo_col c = col(m->color[0], m->color[1], m->color[2]);
    for (int k = 0; k < m->mod_count; k++) c = yo_modulate(&F->mods[m->mod_first + k], F->texs, c, P);
    return DC(c.r, c.g, c.b);
}

/* genericShader_t::fromLight, diffuse part (basicshaders.cc:46-70) */
static dc from_light(const fctx* F, const dsp* sp, d3 ldir, dc lcol, d3 eye)
{
    const d3 N = dff(sp->Ng, sp->N, dnorm(eye));
    const double inte = ddot(N, ldir);
    if (inte < 0) return DC(0, 0, 0);
    return cmul(cscale(lcol, inte), albedo(F, sp));
}

static double uniform(const fctx* F, fsamp* S)
{
    if (F->mode == YO_FP64_REFERENCE) {
        /* ourRandom(), vector3d.h:288-297, the quotient in double */
        const int a = 16807, m = 2147483647, q = m / a, r = m % a;
        int32_t s = *S->lcg;
        s = a * (s % q) - r * (s / q);
        if (s < 0) s += m;
        *S->lcg = s;
        return (double)s / (double)m;
    }
    return (double)(splitmix(&S->rng) >> 11) * (1.0 / 9007199254740992.0);
}

/* sunLight_t::illuminate (sunlight.cc:35-57) */
static dc sun_illuminate(const fctx* F, int li, const dsp* sp, d3 eye)
{
    if (F->lights[li].cast_shadows && is_shadowed(F, sp->P, F->sun_dir[li], sp->obj, 0))
        return DC(0, 0, 0);
    return from_light(F, sp, F->sun_dir[li], F->sun_col[li], eye);
}

/* areaLight_t::fillQuad (arealight.cc:31-73) */
static void area_setup(const yg_light* L, int samp, darea* A)
{
    const d3 a = dld(L->a), b = dld(L->b), c = dld(L->c), d = dld(L->d);
    const size_t cap = (size_t)(samp > 0 ? samp : 1);
    A->points = (d3*)calloc(cap, sizeof(d3));
    A->jit_h = (d3*)calloc(cap, sizeof(d3));
    A->jit_v = (d3*)calloc(cap, sizeof(d3));
    const double ab = dlen(dsub(b, a)), bc = dlen(dsub(c, b)), cd = dlen(dsub(d, c)), da = dlen(dsub(a, d));
    const double max_v = (bc > da) ? bc : da, max_h = (ab > cd) ? ab : cd;
    const int v_samp = (int)(max_v / (max_v + max_h) * 2.0 * sqrt((double)samp));
    const int h_samp = (int)(max_h / (max_v + max_h) * 2.0 * sqrt((double)samp));
    int index = 0;
    if (samp == 1) index = 1;
    else {
        const d3 inc_v_a = dmul(dsub(d, a), 1.0 / v_samp), inc_v_b = dmul(dsub(c, b), 1.0 / v_samp);
        d3 line_a = dadd(a, dmul(inc_v_a, 0.5)), line_b = dadd(b, dmul(inc_v_b, 0.5));
        for (int i = 0; i < v_samp; ++i) {
            const d3 inc_h = dmul(dsub(line_b, line_a), 1.0 / h_samp);
            d3 point = dadd(line_a, dmul(inc_h, 0.5));
            for (int j = 0; j < h_samp; ++j) {
                const double vf = (double)j / (double)h_samp;
                if (index < samp) {
                    A->points[index] = point;
                    A->jit_h[index] = inc_h;
                    A->jit_v[index] = dadd(dmul(inc_v_a, 1.0 - vf), dmul(inc_v_b, vf));
                }
                index++;
                point = dadd(point, inc_h);
            }
            line_a = dadd(line_a, inc_v_a);
            line_b = dadd(line_b, inc_v_b);
        }
    }
    A->samples = index;
    A->direction = dnorm(dcross(dsub(b, a), dsub(d, a)));
    A->pc = cscale(DC(L->color[0], L->color[1], L->color[2]), L->power);
    A->from = dmul(dadd(dadd(dadd(a, b), c), d), 0.25);
    A->corner = a;
    A->toX = dsub(b, a);
    A->toY = dsub(d, a);
}

/* areaLight_t::illuminate, PENUMBRA branch (arealight.cc:93-174) */
static dc area_illuminate(const fctx* F, fsamp* S, int li, const dsp* sp, d3 eye)
{
    const darea* A = &F->areas[li];
    const d3 N = dff(sp->Ng, sp->N, dnorm(eye));
    const double plane_at = ddot(dsub(sp->P, A->from), A->direction);
    if (plane_at <= 0) return DC(0, 0, 0);
    dc resul = DC(0, 0, 0);
    for (int i = 0; i < A->samples; ++i) {
        d3 sampleP;
        if (A->samples == 1) {
            /* toY takes the first draw (g++ evaluates + right operand first) */
            const double r2 = uniform(F, S), r1 = uniform(F, S);
            sampleP = dadd(dadd(A->corner, dmul(A->toX, r1)), dmul(A->toY, r2));
        } else {
            const double dish = uniform(F, S) - 0.5, disv = uniform(F, S) - 0.5;
            sampleP = dadd(dadd(A->points[i], dmul(A->jit_h[i], dish)), dmul(A->jit_v[i], disv));
        }
        const d3 L = dsub(sampleP, sp->P);
        if (ddot(L, N) < 0) continue;
        if (is_shadowed(F, sp->P, sampleP, sp->obj, 1)) continue;
        const double l2 = ddot(L, L);
        const double LD2 = (l2 != 0) ? 1.0 / l2 : 1.0;
        resul = cadd(resul, from_light(F, sp, dnorm(L), cscale(A->pc, LD2), eye));
    }
    if (A->samples == 1) return cscale(resul, plane_at);
    return cscale(resul, plane_at / (double)A->samples);
}

static dc path_illuminate(const fctx* F, fsamp* S, int li, const dsp* sp, d3 eye);

/* scene_t::light (scene.cc:260-277) */
static dc scene_light(const fctx* F, fsamp* S, const dsp* sp, d3 from, int indirect)
{
    dc fl = DC(0, 0, 0);
    const d3 eye = dsub(from, sp->P);
    for (int li = 0; li < F->nl; li++) {
        const yg_light* L = &F->lights[li];
        if (indirect ? !L->use_in_indirect : !L->use_in_render) continue;
        switch (L->type) {
        case YG_LIGHT_SUN: fl = cadd(fl, sun_illuminate(F, li, sp, eye)); break;
        case YG_LIGHT_AREA: fl = cadd(fl, area_illuminate(F, S, li, sp, eye)); break;
        case YG_LIGHT_PATH: fl = cadd(fl, path_illuminate(F, S, li, sp, eye)); break;
        default: break;
        }
    }
    return fl;
}

static dc background(const fctx* F)
{
    if (!F->o.has_background) return DC(0, 0, 0);
    return DC(F->o.background[0], F->o.background[1], F->o.background[2]);
}

/* haltonSampler_t::nextDirection (pathtools.cc:32-43) */
static d3 next_direction(const fctx* F, fsamp* S, d3 N, d3 Ru, d3 Rv, int curlev)
{
    double z1, z2;
    if (F->mode == YO_FP64_REFERENCE) {
        z1 = dhalton_next(&S->hs[2 * curlev]);
        z2 = dhalton_next(&S->hs[2 * curlev + 1]);
    } else {
        z1 = radinv(S->k, F->bases[2 * curlev]);
        z2 = radinv(S->k, F->bases[2 * curlev + 1]);
    }
    if (z1 > 1.0) z1 = 1.0;
    z2 *= 2.0 * M_PI;
    const d3 a = dadd(dmul(Ru, cos(z2)), dmul(Rv, sin(z2)));
    return dadd(dmul(a, sqrt(1.0 - z1)), dmul(N, sqrt(z1)));
}

/* pathLight_t::takeSample, full-GI branch (pathlight.cc:223-387) */
static dc path_take_sample(const fctx* F, fsamp* S, int li, d3 N, const dsp* sp)
{
    const yg_light* L = &F->lights[li];
    const int samples = (F->mode == YO_FP64_CONVERGED && F->samples > 0) ? F->samples : L->samples;
    dc total = DC(0, 0, 0), subtotal[4];
    for (int k = 0; k < 4; k++) subtotal[k] = DC(0, 0, 0);
    int taken = 0, first = 1;
    double repetitions = 0;
    for (int ite = 1; ite > 0; --ite, repetitions += 1) {
        taken = 0;
        for (int i = 0; i < samples; ++i) {
            S->k = S->hoff + (++S->hcount);
            dc raycolor = DC(1, 1, 1);
            if (i > taken) taken = i;
            d3 ray = next_direction(F, S, N, sp->NU, sp->NV, 0);
            const d3 startray = ray;
            const double cos0 = ddot(startray, N);
            d3 where = sp->P;
            for (int j = 0; j < L->depth; ++j) {
                if (cenergy(raycolor) < 0.05) break;
                dsp t;
                if (!first_hit(F, where, ray, &t, 1)) {
                    const dc contri = cmul(cscale(raycolor, cos0), background(F));
                    total = cadd(total, contri);
                    if (first) subtotal[i % 4] = cadd(subtotal[i % 4], contri);
                    break;
                }
                const d3 eye = dneg(ray);
                const dc light = scene_light(F, S, &t, dadd(t.P, eye), 1);
                const dc contri = cmul(cscale(raycolor, cos0), light);
                total = cadd(total, contri);
                if (first) subtotal[i % 4] = cadd(subtotal[i % 4], contri);
                const d3 HN = dff(t.Ng, t.N, dneg(ray));
                raycolor = cmul(raycolor, albedo(F, &t));
                ray = next_direction(F, S, HN, t.NU, t.NV, j + 1);
                where = t.P;
            }
        }
        if (first) {
            first = 0;
            const double mult = 1.0 / (double)(taken + 1);
            double minR = 1000, minG = 1000, minB = 1000, maxR = 0, maxG = 0, maxB = 0;
            for (int i = 0; i < 4; ++i) {
                const dc st = cscale(subtotal[i], mult);
                if (st.r > maxR) maxR = st.r;
                if (st.g > maxG) maxG = st.g;
                if (st.b > maxB) maxB = st.b;
                if (st.r < minR) minR = st.r;
                if (st.g < minG) minG = st.g;
                if (st.b < minB) minB = st.b;
            }
            const double ene = fmax(fabs(minR), fmax(fabs(minG), fabs(minB)));
            const double dmax = fmax(fabs(maxR - minR), fmax(fabs(maxG - minG), fabs(maxB - minB)));
            const double diff = dmax / (((ene < 1.0) ? 1.0 : ene) * 2.0 * 2.0) / (double)L->threshold;
            int extra = (int)(diff * diff) - 1;
            if (extra > 3) extra = 3;
            if (extra < 0) extra = 0;
            ite += extra;
        }
    }
    return cscale(total, 1.0 / (double)(taken + 1) / repetitions);
}

/* pathLight_t::normalSample (pathlight.cc:389-403) */
static dc path_illuminate(const fctx* F, fsamp* S, int li, const dsp* sp, d3 eye)
{
    const d3 N = dff(sp->Ng, sp->N, eye);
    const dc dif = albedo(F, sp);
    if (cenergy(dif) < 0.05) return DC(0, 0, 0);
    const dc total = path_take_sample(F, S, li, N, sp);
    return cscale(cmul(dif, total), (double)F->lights[li].power);
}

/* ------------------------------------------------------------- camera -- */
typedef struct { d3 pos, vright, vup, vto; } dcam;

/* camera_t::camera_t perspective part (camera.cc:27-87) */
static void cam_setup(const yg_camera* c, dcam* k)
{
    const d3 pos = dld(c->pos), look = dld(c->look), up = dld(c->up);
    d3 vup = dsub(up, pos), vto = dsub(look, pos);
    d3 vright = dcross(vup, vto);
    vup = dnorm(dcross(vright, vto));
    vright = dmul(dnorm(vright), -1.0);
    vto = dnorm(vto);
    vup = dmul(vup, (double)c->aspect * (double)c->resy / (double)c->resx);
    vto = dsub(dmul(vto, (double)c->focal), dmul(dadd(vup, vright), 0.5));
    k->vup = dmul(vup, 1.0 / (double)c->resy);
    k->vright = dmul(vright, 1.0 / (double)c->resx);
    k->vto = vto;
    k->pos = pos;
}

/* scene_t::raytrace (scene.cc:200-258) of the ray through the centre of pixel (x, y) */
static dc raytrace_pixel(const fctx* F, fsamp* S, const dcam* k, int x, int y)
{
    const d3 ray = dnorm(dadd(dadd(dmul(k->vright, x + 0.5), dmul(k->vup, y + 0.5)), k->vto));
    dsp sp;
    if (first_hit(F, k->pos, ray, &sp, 0)) return scene_light(F, S, &sp, k->pos, 0);
    return background(F);
}

/* ------------------------------------------------------------- drivers -- */
typedef struct {
    const fctx* F;
    const dcam* k;
    const yg_camera* cam;
    /* reference mode */
    const int32_t* regions;  /* 8 ints per area */
    int nreg;
    int32_t* lcg;
    /* converged mode */
    const int32_t *px, *py;
    int64_t n;
    uint64_t seed;
    double* rgb;
    int64_t next;
    pthread_mutex_t mu;
} fjob;

static int64_t take(fjob* J, int64_t step)
{
    pthread_mutex_lock(&J->mu);
    const int64_t q = J->next;
    J->next += step;
    pthread_mutex_unlock(&J->mu);
    return q;
}

static void* ref_worker(void* arg)
{
    fjob* J = (fjob*)arg;
    const fctx* F = J->F;
    fsamp S;
    memset(&S, 0, sizeof(S));
    S.lcg = J->lcg;
    const int resx = J->cam->resx;
    for (;;) {
        const int64_t q = take(J, 1);
        if (q >= J->nreg) break;
        /* areas in the oracle renderer's order, scene.cc:538-600 */
        const int32_t* a = J->regions + 8 * (J->nreg - 1 - q);
        unsigned base = 2;
        for (int d = 0; d < MAXDIM; d++) {
            S.hs[d].base = base;
            S.hs[d].inv_base = 1.0 / (double)base;
            S.hs[d].value = 0;
            base = (unsigned)yo_next_prime((int)base);
        }
        for (int i = a[1]; i < a[1] + a[3]; ++i)
            for (int j = a[0]; j < a[0] + a[2]; ++j) {
                const dc c = raytrace_pixel(F, &S, J->k, j, i);
                if (j >= a[4] && j < a[4] + a[6] && i >= a[5] && i < a[5] + a[7]) {
                    double* o = J->rgb + 3 * ((int64_t)i * resx + j);
                    o[0] = c.r; o[1] = c.g; o[2] = c.b;
                }
            }
    }
    return NULL;
}

static void* conv_worker(void* arg)
{
    fjob* J = (fjob*)arg;
    fsamp S;
    memset(&S, 0, sizeof(S));
    for (;;) {
        const int64_t q0 = take(J, 16);
        if (q0 >= J->n) break;
        for (int64_t q = q0; q < q0 + 16 && q < J->n; q++) {
            uint64_t h = J->seed ^ ((uint64_t)(uint32_t)J->px[q] << 32) ^ (uint64_t)(uint32_t)J->py[q];
            S.hoff = splitmix(&h) & 0xFFFFFFull;
            S.rng = splitmix(&h);
            S.hcount = 0;
            const dc c = raytrace_pixel(J->F, &S, J->k, J->px[q], J->py[q]);
            J->rgb[3 * q] = c.r; J->rgb[3 * q + 1] = c.g; J->rgb[3 * q + 2] = c.b;
        }
    }
    return NULL;
}

int yo_render_fp64(yo_scene* s, const yg_camera* cam, const yg_material* mats, int nmat,
                   const yg_modulator* mods, const yg_texture* texs, const yg_light* lights, int nl,
                   const int32_t* obj_flags, int nobj, const yg_render_opts* opts, int mode,
                   int samples, const int32_t* px, const int32_t* py, int64_t n, uint64_t seed,
                   double* rgb, int nthreads)
{
    (void)nmat;
    if (!s || !cam || !opts || !rgb) return -2;
    if (mode != YO_FP64_REFERENCE && mode != YO_FP64_CONVERGED) return -2;
    if (mode == YO_FP64_CONVERGED && (samples < 1 || (n > 0 && (!px || !py)))) return -2;
    if (opts->aa_passes > 0 || opts->tonemap) return -5;
    int npath = 0, sequential = 0;
    for (int li = 0; li < nl; li++) {
        const yg_light* L = &lights[li];
        if (L->type == YG_LIGHT_PHOTON) return -6;
        if (L->type == YG_LIGHT_PATH) {
            if (L->use_in_indirect || 2 * (L->depth + 1) > MAXDIM) return -6;
            npath++;
        }
        if (L->type == YG_LIGHT_AREA) sequential = 1;
        /* guessLight's rand() stream is not restated here: the converged mode
         * samples the penumbra everywhere, which is the integral guessLight
         * approximates */
        if (L->type == YG_LIGHT_AREA && L->psamples > 0 && mode == YO_FP64_REFERENCE) return -6;
    }
    if (npath > 1) return -7;
    /* texture coordinates other than the hit point are restated only in the
     * float oracle */
    for (int m = 0; m < nmat; m++)
        for (int k = 0; k < mats[m].mod_count; k++)
            if (mods[mats[m].mod_first + k].texco != YG_TEXCO_GLOBAL) return -6;

    fctx F;
    memset(&F, 0, sizeof(F));
    F.s = s; F.mats = mats; F.mods = mods; F.texs = texs; F.lights = lights; F.nl = nl;
    F.oflags = obj_flags;
    F.nobj = nobj;
    F.o = *opts;
    F.mode = mode;
    F.samples = samples;
    unsigned base = 2;
    for (int d = 0; d < MAXDIM; d++) { F.bases[d] = base; base = (unsigned)yo_next_prime((int)base); }
    F.tn = (double*)malloc(sizeof(double) * 3 * (size_t)(s->nt > 0 ? s->nt : 1));
    for (int64_t t = 0; t < s->nt; t++) {
        const d3 a = vert(s, t, 0), b = vert(s, t, 1), c = vert(s, t, 2);
        const d3 nn = dnorm(dcross(dsub(b, a), dsub(c, a)));
        F.tn[3 * t] = nn.x; F.tn[3 * t + 1] = nn.y; F.tn[3 * t + 2] = nn.z;
    }
    const size_t nls = (size_t)(nl > 0 ? nl : 1);
    F.sun_dir = (d3*)calloc(nls, sizeof(d3));
    F.sun_col = (dc*)calloc(nls, sizeof(dc));
    F.areas = (darea*)calloc(nls, sizeof(darea));
    for (int li = 0; li < nl; li++) {
        const yg_light* L = &lights[li];
        if (L->type == YG_LIGHT_SUN) {
            /* sunLight_t ctor, sunlight.cc:27-33 */
            F.sun_dir[li] = dnorm(dld(L->dir));
            F.sun_col[li] = cscale(DC(L->color[0], L->color[1], L->color[2]), L->power);
        } else if (L->type == YG_LIGHT_AREA) {
            area_setup(L, mode == YO_FP64_CONVERGED ? samples : L->samples, &F.areas[li]);
        }
    }
    dcam k;
    cam_setup(cam, &k);

    fjob J;
    memset(&J, 0, sizeof(J));
    J.F = &F; J.k = &k; J.cam = cam; J.rgb = rgb; J.seed = seed;
    J.px = px; J.py = py; J.n = n;
    int32_t lcg = (int32_t)opts->lcg_seed;
    int32_t* regions = NULL;
    if (mode == YO_FP64_REFERENCE) {
        const int b = opts->block > 0 ? opts->block : 64;
        const int cap = ((cam->resx + b - 1) / b) * ((cam->resy + b - 1) / b);
        regions = (int32_t*)malloc(sizeof(int32_t) * 8 * (size_t)(cap > 0 ? cap : 1));
        J.nreg = yo_block_regions_into(cam->resx, cam->resy, b, opts->rand_seed, regions, cap);
        J.regions = regions;
        J.lcg = &lcg;
        memset(rgb, 0, sizeof(double) * 3 * (size_t)cam->resx * (size_t)cam->resy);
        if (sequential) nthreads = 1;
    }
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_mutex_init(&J.mu, NULL);
    void* (*fn)(void*) = (mode == YO_FP64_REFERENCE) ? ref_worker : conv_worker;
    if (nthreads == 1) fn(&J);
    else {
        pthread_t th[64];
        for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, fn, &J);
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    }
    pthread_mutex_destroy(&J.mu);
    for (int li = 0; li < nl; li++) {
        free(F.areas[li].points); free(F.areas[li].jit_h); free(F.areas[li].jit_v);
    }
    free(F.areas); free(F.sun_dir); free(F.sun_col); free(F.tn); free(regions);
    return 0;
}
