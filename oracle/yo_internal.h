/* Internal helpers of the CPU oracle (test infrastructure only). */
#ifndef YO_INTERNAL_H
#define YO_INTERNAL_H

#include <math.h>
#include <stdint.h>
#include "yaf_oracle.h"

/* vector3d_t operators, vector3d.h:145-219.  Evaluated in float, left to right. */
static inline yo_v3 v3(float x, float y, float z) { yo_v3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline yo_v3 v3_ld(const float* p) { return v3(p[0], p[1], p[2]); }
static inline yo_v3 v3_add(yo_v3 a, yo_v3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline yo_v3 v3_sub(yo_v3 a, yo_v3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline yo_v3 v3_neg(yo_v3 a) { return v3(-a.x, -a.y, -a.z); }
/* operator*(vector, PFLOAT) and operator*(PFLOAT, vector) both compute f*b.x */
static inline yo_v3 v3_mul(yo_v3 b, float f) { return v3(f * b.x, f * b.y, f * b.z); }
static inline yo_v3 v3_div(yo_v3 b, float f) { return v3(b.x / f, b.y / f, b.z / f); }
static inline float v3_dot(yo_v3 a, yo_v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline yo_v3 v3_cross(yo_v3 a, yo_v3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* vector3d_t::normalize, vector3d.h:209-219: len = 1.0/sqrt(len) is a double
 * division of a float square root, whose float rounding equals 1.0f/sqrtf. */
static inline yo_v3 v3_norm(yo_v3 a)
{
    float len = a.x * a.x + a.y * a.y + a.z * a.z;
    if (len != 0) {
        len = (float)(1.0 / (double)sqrtf(len));
        a.x *= len; a.y *= len; a.z *= len;
    }
    return a;
}
/* vector3d_t::length, vector3d.h:199-202 */
static inline float v3_len(yo_v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
/* normLenSqr, vector3d.h:61-69 */
static inline float v3_normlensqr(yo_v3* a)
{
    float vl = a->x * a->x + a->y * a->y + a->z * a->z;
    if (vl != 0.0f) {
        const float d = (float)(1.0 / (double)sqrtf(vl));
        a->x *= d; a->y *= d; a->z *= d;
    }
    return vl;
}
/* normLen, vector3d.h:50-59 */
static inline float v3_normlen(yo_v3* a)
{
    float vl = a->x * a->x + a->y * a->y + a->z * a->z;
    if (vl != 0.0f) {
        vl = sqrtf(vl);
        const float d = (float)(1.0 / (double)vl);
        a->x *= d; a->y *= d; a->z *= d;
    }
    return vl;
}
static inline int v3_null(yo_v3 a) { return a.x == 0 && a.y == 0 && a.z == 0; }
/* FACE_FORWARD(Ng,N,I), shader.h:157 */
static inline yo_v3 face_forward(yo_v3 ng, yo_v3 n, yo_v3 i) { return (v3_dot(ng, i) < 0) ? v3_neg(n) : n; }
/* createCS, vector3d.h:240-258 */
static inline void create_cs(yo_v3 n, yo_v3* u, yo_v3* v)
{
    if (n.x == 0 && n.y == 0) {
        *u = (n.z < 0) ? v3(-1, 0, 0) : v3(1, 0, 0);
        *v = v3(0, 1, 0);
    } else {
        const float d = (float)(1.0 / (double)sqrtf(n.y * n.y + n.x * n.x));
        *u = v3(n.y * d, -n.x * d, 0);
        *v = v3_cross(n, *u);
    }
}

/* color_t (color.h:37-115): three floats */
typedef struct { float r, g, b; } yo_col;
static inline yo_col col(float r, float g, float b) { yo_col c; c.r = r; c.g = g; c.b = b; return c; }
static inline yo_col col_add(yo_col a, yo_col b) { return col(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline yo_col col_sub(yo_col a, yo_col b) { return col(a.r - b.r, a.g - b.g, a.b - b.b); }
static inline yo_col col_mul(yo_col a, yo_col b) { return col(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline yo_col col_scale(yo_col a, float f) { return col(f * a.r, f * a.g, f * a.b); }
static inline float col_energy(yo_col c) { return (c.r + c.g + c.b) * 0.333333f; }
/* mix, color.cc:103-109 */
static inline yo_col col_mix(yo_col a, yo_col b, float point)
{
    if (point < 0.0f) return b;
    if (point > 1.0f) return a;
    return col_add(col_scale(a, point), col_scale(b, 1 - point));
}

/* Halton incremental generator, mcqmc.h:11-54 */
typedef struct { unsigned int base; double inv_base; double value; } yo_halton;
static inline void halton_init(yo_halton* h, int base)
{
    h->base = (unsigned)base; h->inv_base = 1.0 / (double)base; h->value = 0;
}
static inline float halton_next(yo_halton* h)
{
    double r = 1.0 - h->value - 0.0000000001;
    if (h->inv_base < r)
        h->value += h->inv_base;
    else {
        double hh, hv = h->inv_base;
        do { hh = hv; hv *= h->inv_base; } while (hv >= r);
        h->value += hh + hv - 1.0;
    }
    return (float)h->value;
}

/* ourRandom(), vector3d.h:288-297 (Park-Miller minimal standard, Schrage) */
static inline float our_random(int32_t* seed)
{
    const int a = 16807, m = 2147483647, q = m / a, r = m % a;
    int32_t s = *seed;
    s = a * (s % q) - r * (s / q);
    if (s < 0) s += m;
    *seed = s;
    return (float)s / (float)m;
}

/* scene internals */
typedef struct {
    double lo[3], hi[3];
    int32_t left, right;   /* children, -1 for leaf */
    int32_t start, count;  /* leaf triangle range in order[] */
} yo_node;

typedef struct {
    float cell;            /* hash cell size */
    int64_t n;
    float* data;           /* 9 floats per photon */
    /* uniform grid for the K-nearest query */
    float gmin[3];
    float gcell;
    int gdim[3];
    int64_t* gstart;       /* ncells+1 */
    int64_t* gitems;
    int np_emit;           /* Np of the light */
} yo_photonmap;

/* globalPhotonLight_t's hash boxes after pre-gathering (globalphotonlight.cc
 * :177-228): one record per box, boxes sorted by (bx, by, bz) */
typedef struct {
    int32_t bx, by, bz;
    float pos[3];          /* the box photon's position (the first stored) */
    uint8_t th, ph;        /* its quantised direction (th 255: none) */
    uint8_t rgbe[4];       /* its RGBE colour */
    float N[3];            /* normal of the first stored photon */
    float irr[3];          /* irradiance (setIrradiance) */
} yo_gbox;

typedef struct {
    int built;
    int light;             /* light index of the global photon light */
    float cell;            /* hash box size: radius / sqrt(search) */
    float radius;          /* maxradius of the photon and irradiance maps */
    int64_t n;             /* boxes */
    yo_gbox* box;
    int64_t stored, emitted;
    int stateless;         /* gathers start at radius each time (no per-call radius state) */
    int32_t lcg_after;     /* ourRandom() state after the photon pass (render continues it) */
    int own_stream;        /* the pass ran here (1) or the boxes were loaded (0) */
    /* uniform grid (cell = radius) over the box positions for the gathers */
    float gmin[3];
    int gdim[3];
    int64_t* gstart;
    int64_t* gitems;
    float irr_radius;      /* pathLight_t's photonData_t radius state (pathlight.cc:131) */
    /* the raw photon map (globalPhotonMap_t photonMap, globalphotonlight.cc:38) in
     * store order, with a uniform grid (cell = radius) for photonSampler_t's gathers */
    int64_t nrec;
    struct yo_gphot_s* rec;
    float rgmin[3];
    int rgdim[3];
    int64_t* rgstart;
    int64_t* rgitems;
} yo_gmap;

/* storedPhoton_t of the raw map: position, 8-bit direction, RGBE colour */
typedef struct yo_gphot_s { float pos[3]; uint8_t th, ph; uint8_t rgbe[4]; } yo_gphot;

struct yo_scene {
    int64_t nv, nt;
    float* verts;
    uint32_t* idx;
    int32_t* tri_obj;
    int32_t* tri_mat;
    float* tri_n;          /* cached unit normals, triangle.cc:45-49 */
    yo_node* nodes;
    int64_t nnodes;
    int32_t* order;
    yo_photonmap pm;
    yo_gmap gm;
    float* tri_vn;         /* autoSmooth vertex normals, 9 per triangle (NULL: none) */
    uint8_t* tri_smooth;   /* per triangle: has vertex normals */
    /* texture coordinates: per triangle u_a v_a u_b v_b u_c v_c, the orco
     * points of corners a, b, c, flags (bit0 uv, bit1 vertex colours, bit2
     * orco points) = 16 floats (NULL: no triangle has any); per object the
     * orco matrix, rows 0..2 of backOrco (mesh.cc:488-491) */
    float* tri_tex;
    int32_t nobj;
    float* obj_orco;
    struct yo_lcache_s* lc_last;  /* irradiance cache of the last cache = on render */
    uint64_t lc_new;
    int lc_rec;                   /* record the fake pass's visits (test hook) */
    float* lc_visits;             /* 13 floats per visit */
    int64_t lc_nvis, lc_viscap;
    /* object flags of the render in progress (bit0 cast shadows), or NULL */
    const int32_t* oflags;
    int64_t noflags;
};

/* scene shadow queries (isShadowed; firstHit with its shadow flag,
 * scene.cc:698) skip objects that cast no shadows */
static inline int yo_skips_shadow(const yo_scene* s, int32_t obj)
{
    return s->oflags && obj >= 0 && obj < s->noflags && !(s->oflags[obj] & 1);
}

/* closest query with object filter: only triangles with (filter_obj < 0) or
 * (include ? obj == filter_obj : obj != filter_obj).  mode 0 closest, 1 any. */
int yo_query(const yo_scene* s, yo_v3 o, yo_v3 d, float tmin, float tmax, int mode,
             int filter_obj, int include, int32_t* prim, float* t, uint64_t* cn, uint64_t* ct);
int yo_first_hit_one_s(const yo_scene* s, yo_v3 from, yo_v3 ray, float min_raydis, int shadow,
                       int32_t* prim, float* t);
/* the same, skipping objects that cast no shadows when shadow != 0 */
int yo_query_s(const yo_scene* s, yo_v3 o, yo_v3 d, float tmin, float tmax, int mode,
               int filter_obj, int include, int shadow, int32_t* prim, float* t, uint64_t* cn,
               uint64_t* ct);

static inline yo_v3 tri_vert(const yo_scene* s, int64_t tri, int k)
{
    return v3_ld(s->verts + 3 * (int64_t)s->idx[3 * tri + k]);
}

#endif
