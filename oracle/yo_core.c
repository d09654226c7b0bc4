/*
 * CPU oracle: triangle primitive, canonical BVH2, scene queries, QMC and RNG.
 * TEST INFRASTRUCTURE ONLY (see yaf_oracle.h).
 */
#include <float.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include "yo_internal.h"

/* ------------------------------------------------------------ triangle -- */

/* triangle_t::recNormal, triangle.cc:45-49: normal = (b-a)^(c-a), normalised */
static yo_v3 tri_normal_v(yo_v3 a, yo_v3 b, yo_v3 c)
{
    return v3_norm(v3_cross(v3_sub(b, a), v3_sub(c, a)));
}

void yo_tri_normal(const float* a, const float* b, const float* c, float* n)
{
    yo_v3 r = tri_normal_v(v3_ld(a), v3_ld(b), v3_ld(c));
    n[0] = r.x; n[1] = r.y; n[2] = r.z;
}

/* triangle_t::hit, triangle.h:48-58: the ray is flipped to the normal's side,
 * then the three edge triple products (va^vb).r, (vb^vc).r, (vc^va).r must all
 * be >= 0, with va = a-from etc. */
static int tri_hit_v(yo_v3 a, yo_v3 b, yo_v3 c, yo_v3 n, yo_v3 from, yo_v3 ray)
{
    const yo_v3 va = v3_sub(a, from), vb = v3_sub(b, from), vc = v3_sub(c, from);
    yo_v3 r;
    if (v3_dot(ray, n) < 0) r = v3_neg(ray);
    else r = ray;
    if (v3_dot(v3_cross(va, vb), r) < 0) return 0;
    if (v3_dot(v3_cross(vb, vc), r) < 0) return 0;
    if (v3_dot(v3_cross(vc, va), r) < 0) return 0;
    return 1;
}

/* triangle_t::intersect, triangle.h:83-86: t = (N.(a-from)) / (N.ray) */
static float tri_intersect_v(yo_v3 a, yo_v3 n, yo_v3 from, yo_v3 ray)
{
    return v3_dot(n, v3_sub(a, from)) / v3_dot(n, ray);
}

int yo_tri_hit(const float* a, const float* b, const float* c, const float* n,
               const float* from, const float* ray)
{
    return tri_hit_v(v3_ld(a), v3_ld(b), v3_ld(c), v3_ld(n), v3_ld(from), v3_ld(ray));
}

float yo_tri_intersect(const float* a, const float* n, const float* from, const float* ray)
{
    return tri_intersect_v(v3_ld(a), v3_ld(n), v3_ld(from), v3_ld(ray));
}

/* getInterpolation, triangle.cc:74-136: 2-D barycentrics in the plane that
 * drops the normal's dominant axis. */
static void tri_bary_v(yo_v3 n, yo_v3 a, yo_v3 b, yo_v3 c, yo_v3 h, float* fb, float* fc)
{
    float uu0, vv0, uu1, vv1, uu2, vv2;
    float nx = fabsf(n.x), ny = fabsf(n.y), nz = fabsf(n.z);
    if (nx >= ny && nx >= nz) {
        uu0 = h.y - a.y; vv0 = h.z - a.z; uu1 = b.y - a.y; vv1 = b.z - a.z; uu2 = c.y - a.y; vv2 = c.z - a.z;
    } else if (ny >= nx && ny >= nz) {
        uu0 = h.x - a.x; vv0 = h.z - a.z; uu1 = b.x - a.x; vv1 = b.z - a.z; uu2 = c.x - a.x; vv2 = c.z - a.z;
    } else {
        uu0 = h.x - a.x; vv0 = h.y - a.y; uu1 = b.x - a.x; vv1 = b.y - a.y; uu2 = c.x - a.x; vv2 = c.y - a.y;
    }
    float alpha_numer = uu0 * vv2 - vv0 * uu2;
    float beta_numer = uu1 * vv0 - vv1 * uu0;
    float denum = uu1 * vv2 - vv1 * uu2;
    *fb = alpha_numer / denum;
    *fc = beta_numer / denum;
}

void yo_tri_bary(const float* a, const float* b, const float* c, const float* n,
                 const float* h, float* fb, float* fc)
{
    tri_bary_v(v3_ld(n), v3_ld(a), v3_ld(b), v3_ld(c), v3_ld(h), fb, fc);
}

/* --------------------------------------------------- canonical BVH2 build -- */
/* Binned SAH, 32 bins, leaf <= 4 triangles, FP64 boxes.  The tree only prunes:
 * results equal the brute-force scan except for numerically degenerate
 * (grazing) candidates, so it is result-neutral like kdTree_t (kdtree.cc:57-582). */

#define YO_BINS 32
#define YO_LEAF 4
#define YO_TEPS 6.103515625e-05  /* 2^-14 */
#define YO_HEPS 7.8125e-03       /* 2^-7 */

typedef struct { double lo[3], hi[3]; } box_d;

static void box_empty(box_d* b)
{
    for (int k = 0; k < 3; k++) { b->lo[k] = DBL_MAX; b->hi[k] = -DBL_MAX; }
}
static void box_grow_pt(box_d* b, const double* p)
{
    for (int k = 0; k < 3; k++) {
        if (p[k] < b->lo[k]) b->lo[k] = p[k];
        if (p[k] > b->hi[k]) b->hi[k] = p[k];
    }
}
static void box_grow(box_d* b, const box_d* o)
{
    for (int k = 0; k < 3; k++) {
        if (o->lo[k] < b->lo[k]) b->lo[k] = o->lo[k];
        if (o->hi[k] > b->hi[k]) b->hi[k] = o->hi[k];
    }
}
static double box_area(const box_d* b)
{
    double dx = b->hi[0] - b->lo[0], dy = b->hi[1] - b->lo[1], dz = b->hi[2] - b->lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

typedef struct {
    yo_scene* s;
    box_d* tb;         /* per-triangle boxes */
    double* cen;       /* per-triangle centroids */
    int64_t cap;
    double G;          /* largest vertex coordinate magnitude */
} build_ctx;

static int32_t node_alloc(build_ctx* bc)
{
    yo_scene* s = bc->s;
    if (s->nnodes == bc->cap) {
        bc->cap = bc->cap ? bc->cap * 2 : 1024;
        s->nodes = (yo_node*)realloc(s->nodes, sizeof(yo_node) * (size_t)bc->cap);
    }
    return (int32_t)(s->nnodes++);
}

static void build_rec(build_ctx* bc, int32_t ni, int64_t start, int64_t count)
{
    yo_scene* s = bc->s;
    int32_t* ord = s->order;
    box_d nb, cb;
    box_empty(&nb); box_empty(&cb);
    for (int64_t i = start; i < start + count; i++) {
        box_grow(&nb, &bc->tb[ord[i]]);
        box_grow_pt(&cb, &bc->cen[3 * (int64_t)ord[i]]);
    }
    /* inflate so that every triangle the float edge test accepts is reached:
     * that test's slack scales with the ray origin's distance and the
     * scene's magnitude G (far rays over small triangles pass it up to ~1e-5
     * outside the triangle), not with this box's own coordinates, so boxes
     * grow by 5% of their extent plus 2^-14 G (a kd-tree cell of the
     * reference is likewise much larger than that slack) */
    for (int k = 0; k < 3; k++) {
        double e = 0.05 * (nb.hi[k] - nb.lo[k]) + 1e-7 * (fabs(nb.lo[k]) + fabs(nb.hi[k]) + 1e-30) +
                   bc->G * 6.103515625e-05;
        s->nodes[ni].lo[k] = nb.lo[k] - e;
        s->nodes[ni].hi[k] = nb.hi[k] + e;
    }
    s->nodes[ni].left = s->nodes[ni].right = -1;
    s->nodes[ni].start = (int32_t)start;
    s->nodes[ni].count = (int32_t)count;
    if (count <= YO_LEAF) return;

    int best_axis = -1, best_split = -1;
    double best_cost = DBL_MAX;
    for (int ax = 0; ax < 3; ax++) {
        double ext = cb.hi[ax] - cb.lo[ax];
        if (!(ext > 0)) continue;
        box_d bb[YO_BINS];
        int64_t bn[YO_BINS];
        for (int b = 0; b < YO_BINS; b++) { box_empty(&bb[b]); bn[b] = 0; }
        double scale = YO_BINS / ext;
        for (int64_t i = start; i < start + count; i++) {
            int b = (int)((bc->cen[3 * (int64_t)ord[i] + ax] - cb.lo[ax]) * scale);
            if (b >= YO_BINS) b = YO_BINS - 1;
            if (b < 0) b = 0;
            bn[b]++;
            box_grow(&bb[b], &bc->tb[ord[i]]);
        }
        double rarea[YO_BINS];
        int64_t rcnt[YO_BINS];
        box_d acc; box_empty(&acc);
        int64_t c = 0;
        for (int b = YO_BINS - 1; b > 0; b--) {
            box_grow(&acc, &bb[b]); c += bn[b];
            rarea[b] = box_area(&acc); rcnt[b] = c;
        }
        box_empty(&acc); c = 0;
        for (int b = 0; b < YO_BINS - 1; b++) {
            box_grow(&acc, &bb[b]); c += bn[b];
            if (c == 0 || rcnt[b + 1] == 0) continue;
            double cost = box_area(&acc) * (double)c + rarea[b + 1] * (double)rcnt[b + 1];
            if (cost < best_cost) { best_cost = cost; best_axis = ax; best_split = b; }
        }
    }
    int64_t mid;
    if (best_axis < 0) {
        mid = start + count / 2;  /* all centroids coincide: split by index */
    } else {
        double ext = cb.hi[best_axis] - cb.lo[best_axis];
        double scale = YO_BINS / ext;
        int64_t i = start, j = start + count - 1;
        while (i <= j) {
            int b = (int)((bc->cen[3 * (int64_t)ord[i] + best_axis] - cb.lo[best_axis]) * scale);
            if (b >= YO_BINS) b = YO_BINS - 1;
            if (b < 0) b = 0;
            if (b <= best_split) i++;
            else { int32_t tmp = ord[i]; ord[i] = ord[j]; ord[j] = tmp; j--; }
        }
        mid = i;
        if (mid == start || mid == start + count) mid = start + count / 2;
    }
    int32_t l = node_alloc(bc);
    int32_t r = node_alloc(bc);
    s->nodes[ni].left = l;
    s->nodes[ni].right = r;
    s->nodes[ni].count = 0;
    build_rec(bc, l, start, mid - start);
    build_rec(bc, r, mid, start + count - mid);
}

yo_scene* yo_scene_new(const float* verts, int64_t nv, const uint32_t* idx, int64_t nt,
                       const int32_t* tri_obj, const int32_t* tri_mat)
{
    yo_scene* s = (yo_scene*)calloc(1, sizeof(yo_scene));
    s->nv = nv; s->nt = nt;
    s->verts = (float*)malloc(sizeof(float) * 3 * (size_t)(nv ? nv : 1));
    memcpy(s->verts, verts, sizeof(float) * 3 * (size_t)nv);
    s->idx = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (size_t)(nt ? nt : 1));
    memcpy(s->idx, idx, sizeof(uint32_t) * 3 * (size_t)nt);
    s->tri_obj = (int32_t*)calloc((size_t)(nt ? nt : 1), sizeof(int32_t));
    s->tri_mat = (int32_t*)calloc((size_t)(nt ? nt : 1), sizeof(int32_t));
    if (tri_obj) memcpy(s->tri_obj, tri_obj, sizeof(int32_t) * (size_t)nt);
    if (tri_mat) memcpy(s->tri_mat, tri_mat, sizeof(int32_t) * (size_t)nt);
    s->tri_n = (float*)malloc(sizeof(float) * 3 * (size_t)(nt ? nt : 1));
    s->order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nt ? nt : 1));
    build_ctx bc;
    bc.s = s; bc.cap = 0; bc.G = 0.0;
    bc.tb = (box_d*)malloc(sizeof(box_d) * (size_t)(nt ? nt : 1));
    bc.cen = (double*)malloc(sizeof(double) * 3 * (size_t)(nt ? nt : 1));
    for (int64_t i = 0; i < nt; i++) {
        yo_v3 a = tri_vert(s, i, 0), b = tri_vert(s, i, 1), c = tri_vert(s, i, 2);
        yo_v3 n = tri_normal_v(a, b, c);
        s->tri_n[3 * i] = n.x; s->tri_n[3 * i + 1] = n.y; s->tri_n[3 * i + 2] = n.z;
        box_empty(&bc.tb[i]);
        double pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z}, pc[3] = {c.x, c.y, c.z};
        box_grow_pt(&bc.tb[i], pa); box_grow_pt(&bc.tb[i], pb); box_grow_pt(&bc.tb[i], pc);
        for (int k = 0; k < 3; k++) {
            bc.cen[3 * i + k] = 0.5 * (bc.tb[i].lo[k] + bc.tb[i].hi[k]);
            if (fabs(bc.tb[i].lo[k]) > bc.G) bc.G = fabs(bc.tb[i].lo[k]);
            if (fabs(bc.tb[i].hi[k]) > bc.G) bc.G = fabs(bc.tb[i].hi[k]);
        }
        s->order[i] = (int32_t)i;
    }
    int32_t root = node_alloc(&bc);
    if (nt > 0) build_rec(&bc, root, 0, nt);
    else {
        for (int k = 0; k < 3; k++) { s->nodes[0].lo[k] = 1; s->nodes[0].hi[k] = -1; }
        s->nodes[0].left = s->nodes[0].right = -1; s->nodes[0].start = 0; s->nodes[0].count = 0;
    }
    free(bc.tb); free(bc.cen);
    /* every object's orco matrix for the identity transform */
    s->nobj = 0;
    for (int64_t i = 0; i < nt; i++)
        if (s->tri_obj[i] + 1 > s->nobj) s->nobj = s->tri_obj[i] + 1;
    s->obj_orco = (float*)calloc((size_t)(12 * (s->nobj ? s->nobj : 1)), sizeof(float));
    for (int64_t t0 = 0; t0 < nt;) {
        int64_t t1 = t0;
        while (t1 < nt && s->tri_obj[t1] == s->tri_obj[t0]) t1++;
        yo_orco_matrix(NULL, s->verts, nv, s->idx + 3 * t0, t1 - t0, s->obj_orco + 12 * (int64_t)s->tri_obj[t0]);
        t0 = t1;
    }
    return s;
}

/* matrix4x4_t::inverse (matrix4.cc:78-112): Gauss-Jordan on the rows with the
 * largest-magnitude pivot, float arithmetic; 0 when singular */
static int inv4(float m[4][4])
{
    float e[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    for (int c = 0; c < 4; c++) {
        int piv = 0;
        float best = 0;
        for (int r = c; r < 4; r++) {
            const float a = fabsf(m[r][c]);
            if (a > best) { best = a; piv = r; }
        }
        if (best == 0) return 0;
        for (int k = 0; k < 4; k++) {
            float t = m[c][k]; m[c][k] = m[piv][k]; m[piv][k] = t;
            t = e[c][k]; e[c][k] = e[piv][k]; e[piv][k] = t;
        }
        const float d = m[c][c];
        for (int k = 0; k < 4; k++) m[c][k] = m[c][k] / d;
        for (int k = 0; k < 4; k++) e[c][k] = e[c][k] / d;
        for (int r = 0; r < 4; r++) {
            if (r == c) continue;
            const float f = m[r][c];
            for (int k = 0; k < 4; k++) m[r][k] = m[r][k] - m[c][k] * f;
            for (int k = 0; k < 4; k++) e[r][k] = e[r][k] - e[c][k] * f;
        }
    }
    memcpy(m, e, sizeof(e));
    return 1;
}

/* recalcBound (mesh.cc:343-380) over the triangles' corners, +-BOX_MIN (1e-5) */
static void corner_bound(const float* xyz, const uint32_t* idx, int64_t nt, float lo[3], float hi[3])
{
    for (int a = 0; a < 3; a++) lo[a] = hi[a] = xyz[3 * (int64_t)idx[0] + a];
    for (int64_t k = 0; k < 3 * nt; k++)
        for (int a = 0; a < 3; a++) {
            const float v = xyz[3 * (int64_t)idx[k] + a];
            if (v < lo[a]) lo[a] = v;
            if (v > hi[a]) hi[a] = v;
        }
    for (int a = 0; a < 3; a++) {
        lo[a] = (float)((double)lo[a] - 1e-5);
        hi[a] = (float)((double)hi[a] + 1e-5);
    }
}

int yo_orco_matrix(const float* m16, const float* xyz, int64_t nv, const uint32_t* idx, int64_t nt,
                   float* out12)
{
    if (nt <= 0) return -1;
    float M[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    if (m16) memcpy(M, m16, sizeof(M));
    float lo[3], hi[3];
    corner_bound(xyz, idx, nt, lo, hi);
    float B[4][4];
    memcpy(B, M, sizeof(B));
    /* matrix4x4_t::scale (matrix4.cc:180-185): columns 0..2 of rows 0..2 */
    for (int c = 0; c < 3; c++) {
        const float ext = hi[c] - lo[c];
        const float half = (float)(0.5 * (double)ext);
        for (int r = 0; r < 3; r++) B[r][c] = B[r][c] * half;
    }
    /* the vertices transformed by M, each row ((m0 x + m1 y) + m2 z) + m3 */
    float* w = (float*)malloc(sizeof(float) * 3 * (size_t)(nv ? nv : 1));
    for (int64_t v = 0; v < nv; v++)
        for (int r = 0; r < 3; r++)
            w[3 * v + r] = M[r][0] * xyz[3 * v] + M[r][1] * xyz[3 * v + 1] + M[r][2] * xyz[3 * v + 2] + M[r][3];
    corner_bound(w, idx, nt, lo, hi);
    free(w);
    for (int r = 0; r < 3; r++) {
        const float sum = lo[r] + hi[r];
        B[r][3] = 0.5f * sum;
    }
    if (!inv4(B)) return -2;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) out12[4 * r + c] = B[r][c];
    return 0;
}

int yo_mesh_texcoords(yo_scene* s, int obj, const float* face_uv, const uint8_t* face_flags,
                      const float* corner_orco, const float* orco_mtx)
{
    if (obj < 0 || obj >= s->nobj) return -1;
    int64_t t0 = 0;
    while (t0 < s->nt && s->tri_obj[t0] != obj) t0++;
    int64_t t1 = t0;
    while (t1 < s->nt && s->tri_obj[t1] == obj) t1++;
    if (orco_mtx) memcpy(s->obj_orco + 12 * (int64_t)obj, orco_mtx, 12 * sizeof(float));
    if (!(face_uv || face_flags || corner_orco)) return 0;
    if (!s->tri_tex) s->tri_tex = (float*)calloc((size_t)(16 * (s->nt ? s->nt : 1)), sizeof(float));
    for (int64_t t = t0; t < t1; t++) {
        float* r = s->tri_tex + 16 * t;
        const int64_t k = t - t0;
        int fl = face_flags ? (face_flags[k] & 3) : (face_uv ? 1 : 0);
        if (!face_uv) fl &= ~1;
        for (int c = 0; c < 6; c++) r[c] = (fl & 1) ? face_uv[6 * k + c] : 0.0f;
        if (corner_orco) {
            for (int c = 0; c < 9; c++) r[6 + c] = corner_orco[9 * k + c];
            fl |= 4;
        }
        r[15] = (float)fl;
    }
    return 0;
}

/* meshObject_t::autoSmooth (mesh.cc:199-282) for the triangles of object
 * obj: no smoothing below 1 degree; at 180 or more one normal per vertex, the
 * face normals of its triangles summed in triangle order and normalised;
 * otherwise per triangle corner the normalised sum, in triangle order, of the
 * face normals of the vertex's triangles whose dot with this triangle's face
 * normal exceeds cos(angle). */
int yo_mesh_autosmooth(yo_scene* s, int obj, float angle)
{
    if (angle < 1) return 0;
    if (!s->tri_vn) {
        s->tri_vn = (float*)calloc((size_t)9 * (size_t)(s->nt ? s->nt : 1), sizeof(float));
        s->tri_smooth = (uint8_t*)calloc((size_t)(s->nt ? s->nt : 1), 1);
    }
    int64_t* vcount = (int64_t*)calloc((size_t)(s->nv ? s->nv : 1), sizeof(int64_t));
    for (int64_t t = 0; t < s->nt; t++)
        if (s->tri_obj[t] == obj)
            for (int k = 0; k < 3; k++) vcount[s->idx[3 * t + k]]++;
    if (angle >= 180) {
        yo_v3* vn = (yo_v3*)calloc((size_t)(s->nv ? s->nv : 1), sizeof(yo_v3));
        for (int64_t t = 0; t < s->nt; t++) {
            if (s->tri_obj[t] != obj) continue;
            yo_v3 n = v3_ld(s->tri_n + 3 * t);
            for (int k = 0; k < 3; k++) {
                yo_v3* a = &vn[s->idx[3 * t + k]];
                *a = v3_add(*a, n);
            }
        }
        for (int64_t v = 0; v < s->nv; v++)
            if (vcount[v]) vn[v] = v3_norm(vn[v]);
        for (int64_t t = 0; t < s->nt; t++) {
            if (s->tri_obj[t] != obj) continue;
            for (int k = 0; k < 3; k++) {
                yo_v3 n = vn[s->idx[3 * t + k]];
                s->tri_vn[9 * t + 3 * k] = n.x; s->tri_vn[9 * t + 3 * k + 1] = n.y; s->tri_vn[9 * t + 3 * k + 2] = n.z;
            }
            s->tri_smooth[t] = 1;
        }
        free(vn);
    } else {
        float cosa = (float)cos((double)angle * M_PI / 180.0);
        /* the triangles of each vertex, in triangle order */
        int64_t* start = (int64_t*)calloc((size_t)s->nv + 1, sizeof(int64_t));
        for (int64_t v = 0; v < s->nv; v++) start[v + 1] = start[v] + vcount[v];
        int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->nv ? s->nv : 1));
        memcpy(fill, start, sizeof(int64_t) * (size_t)s->nv);
        int64_t* tris = (int64_t*)malloc(sizeof(int64_t) * (size_t)(start[s->nv] ? start[s->nv] : 1));
        for (int64_t t = 0; t < s->nt; t++)
            if (s->tri_obj[t] == obj)
                for (int k = 0; k < 3; k++) tris[fill[s->idx[3 * t + k]]++] = t;
        for (int64_t t = 0; t < s->nt; t++) {
            if (s->tri_obj[t] != obj) continue;
            yo_v3 N = v3_ld(s->tri_n + 3 * t);
            for (int k = 0; k < 3; k++) {
                uint32_t v = s->idx[3 * t + k];
                yo_v3 sn = v3(0, 0, 0);
                for (int64_t q = start[v]; q < start[v + 1]; q++) {
                    yo_v3 o = v3_ld(s->tri_n + 3 * tris[q]);
                    if (v3_dot(o, N) > cosa) sn = v3_add(sn, o);
                }
                sn = v3_norm(sn);
                s->tri_vn[9 * t + 3 * k] = sn.x; s->tri_vn[9 * t + 3 * k + 1] = sn.y; s->tri_vn[9 * t + 3 * k + 2] = sn.z;
            }
            s->tri_smooth[t] = 1;
        }
        free(start); free(fill); free(tris);
    }
    free(vcount);
    return 0;
}

/* triangle_t::getSurface's shading normal (triangle.cc:142-160) of prim at
 * hit point h: the face normal, or the vertex normals interpolated with
 * getInterpolation's weights and normalised. */
yo_v3 yo_shading_normal(const yo_scene* s, int32_t prim, yo_v3 h)
{
    yo_v3 ng = v3_ld(s->tri_n + 3 * (int64_t)prim);
    if (!s->tri_smooth || !s->tri_smooth[prim]) return ng;
    float fb, fc;
    tri_bary_v(ng, tri_vert(s, prim, 0), tri_vert(s, prim, 1), tri_vert(s, prim, 2), h, &fb, &fc);
    float fa = 1.f - fb - fc;
    const float* n = s->tri_vn + 9 * (int64_t)prim;
    yo_v3 nn = v3_add(v3_add(v3_mul(v3(n[0], n[1], n[2]), fa), v3_mul(v3(n[3], n[4], n[5]), fb)),
                      v3_mul(v3(n[6], n[7], n[8]), fc));
    return v3_norm(nn);
}

void yo_shading_normals(yo_scene* s, const int32_t* prim, const float* h, int64_t n, float* out)
{
    for (int64_t i = 0; i < n; i++) {
        yo_v3 r = yo_shading_normal(s, prim[i], v3_ld(h + 3 * i));
        out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
    }
}

void yo_scene_free(yo_scene* s)
{
    if (!s) return;
    free(s->tri_vn); free(s->tri_smooth);
    free(s->tri_tex); free(s->obj_orco);
    yo_cache_release(s);
    free(s->verts); free(s->idx); free(s->tri_obj); free(s->tri_mat); free(s->tri_n);
    free(s->nodes); free(s->order);
    free(s->pm.data); free(s->pm.gstart); free(s->pm.gitems);
    free(s->gm.box); free(s->gm.gstart); free(s->gm.gitems);
    free(s->gm.rec); free(s->gm.rgstart); free(s->gm.rgitems);
    free(s);
}

int64_t yo_scene_nodes(yo_scene* s) { return s->nnodes; }

/* ------------------------------------------------------------ traversal -- */

/* slab test in double; axes with a zero direction component are range checks
 * (bound_t::cross skips them the same way, bound.cc:340-437) */
static int box_hit(const yo_node* nd, const double* o, const double* d, double t0, double t1,
                   double* tnear)
{
    for (int k = 0; k < 3; k++) {
        if (d[k] == 0.0) {
            if (o[k] < nd->lo[k] || o[k] > nd->hi[k]) return 0;
            continue;
        }
        double ta = (nd->lo[k] - o[k]) / d[k], tb = (nd->hi[k] - o[k]) / d[k];
        if (ta > tb) { double x = ta; ta = tb; tb = x; }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
    }
    /* relative slack: along grazing rays the float plane distance of an
     * accepted triangle can fall a little outside its box's interval */
    if (t0 - YO_TEPS * fabs(t0) > t1 + YO_TEPS * fabs(t1)) return 0;
    *tnear = t0;
    return 1;
}

/* Accept rules: closest tmin <= t < tmax (kdtree.cc:690-694), any tmin < t < tmax
 * (kdtree.cc:835-841).  Exact ties of t go to the lowest primitive id. */
int yo_query(const yo_scene* s, yo_v3 o, yo_v3 d, float tmin, float tmax, int mode,
             int filter_obj, int include, int32_t* prim, float* t, uint64_t* cn, uint64_t* ct)
{
    return yo_query_s(s, o, d, tmin, tmax, mode, filter_obj, include, 0, prim, t, cn, ct);
}

int yo_query_s(const yo_scene* s, yo_v3 o, yo_v3 d, float tmin, float tmax, int mode,
               int filter_obj, int include, int shadow, int32_t* prim, float* t, uint64_t* cn,
               uint64_t* ct)
{
    int32_t best = -1;
    float bt = tmax;
    if (s->nt == 0) { *prim = -1; *t = tmax; return 0; }
    double od[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    /* slack so that float plane distances near a box face are never pruned */
    double span_lo = (double)tmin - 1e-6 * (fabs((double)tmin) + 1.0);
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const yo_node* nd = &s->nodes[stack[--sp]];
        double tn;
        double hi = (double)bt;
        /* near a shared vertex a grazing ray can pass the float edge test of
         * a triangle at a plane distance in front of its box: boxes are kept
         * until they start 2^-7 (relative) beyond the bound */
        if (isinf(hi)) hi = DBL_MAX; else hi = hi + YO_HEPS * fabs(hi) + 1e-6;
        if (cn) (*cn)++;
        if (!box_hit(nd, od, dd, span_lo, hi, &tn)) continue;
        if (nd->left < 0) {
            for (int32_t i = nd->start; i < nd->start + nd->count; i++) {
                int32_t tri = s->order[i];
                if (shadow && yo_skips_shadow(s, s->tri_obj[tri])) continue;
                if (filter_obj >= 0) {
                    int same = s->tri_obj[tri] == filter_obj;
                    if (include ? !same : same) continue;
                }
                if (ct) (*ct)++;
                yo_v3 a = tri_vert(s, tri, 0), b = tri_vert(s, tri, 1), c = tri_vert(s, tri, 2);
                yo_v3 n = v3_ld(s->tri_n + 3 * (int64_t)tri);
                if (!tri_hit_v(a, b, c, n, o, d)) continue;
                float rt = tri_intersect_v(a, n, o, d);
                if (mode == 0) {
                    if (rt >= tmin && (rt < bt || (rt == bt && best >= 0 && tri < best))) {
                        bt = rt; best = tri;
                    }
                } else {
                    if (rt > tmin && rt < tmax) { *prim = tri; *t = rt; return 1; }
                }
            }
        } else {
            const yo_node* l = &s->nodes[nd->left];
            const yo_node* r = &s->nodes[nd->right];
            double tl = 0, tr = 0;
            /* near child first: push the far one underneath */
            int hl = box_hit(l, od, dd, span_lo, hi, &tl);
            int hr = box_hit(r, od, dd, span_lo, hi, &tr);
            if (hl && hr) {
                if (tl <= tr) { stack[sp++] = nd->right; stack[sp++] = nd->left; }
                else { stack[sp++] = nd->left; stack[sp++] = nd->right; }
            } else if (hl) stack[sp++] = nd->left;
            else if (hr) stack[sp++] = nd->right;
            if (cn) (*cn) += 0;
        }
    }
    *prim = best;
    *t = bt;
    return best >= 0;
}

static void hit_uv(const yo_scene* s, int32_t tri, yo_v3 o, yo_v3 d, float t, float* u, float* v)
{
    /* h = from + Z*ray (mesh.cc:570) */
    yo_v3 h = v3_add(o, v3_mul(d, t));
    yo_v3 a = tri_vert(s, tri, 0), b = tri_vert(s, tri, 1), c = tri_vert(s, tri, 2);
    tri_bary_v(v3_ld(s->tri_n + 3 * (int64_t)tri), a, b, c, h, u, v);
}

void yo_trace_brute(yo_scene* s, const float* rays, int64_t n, int mode, int32_t* prim,
                    float* t, float* u, float* v)
{
    for (int64_t r = 0; r < n; r++) {
        const float* ry = rays + 8 * r;
        yo_v3 o = v3(ry[0], ry[1], ry[2]), d = v3(ry[4], ry[5], ry[6]);
        float tmin = ry[3], tmax = ry[7];
        int32_t best = -1;
        float bt = tmax;
        for (int64_t tri = 0; tri < s->nt; tri++) {
            yo_v3 a = tri_vert(s, tri, 0), b = tri_vert(s, tri, 1), c = tri_vert(s, tri, 2);
            yo_v3 nn = v3_ld(s->tri_n + 3 * tri);
            if (!tri_hit_v(a, b, c, nn, o, d)) continue;
            float rt = tri_intersect_v(a, nn, o, d);
            if (mode == 0) {
                if (rt >= tmin && rt < bt) { bt = rt; best = (int32_t)tri; }
            } else if (rt > tmin && rt < tmax) { best = (int32_t)tri; bt = rt; break; }
        }
        prim[r] = best;
        if (t) t[r] = (best >= 0) ? bt : tmax;
        if (u && v) {
            if (best >= 0 && mode == 0) hit_uv(s, best, o, d, bt, &u[r], &v[r]);
            else { u[r] = 0; v[r] = 0; }
        }
    }
}

typedef struct {
    yo_scene* s; const float* rays; int64_t lo, hi; int mode;
    int32_t* prim; float *t, *u, *v; uint64_t nodes, tris;
} trace_job;

static void* trace_worker(void* arg)
{
    trace_job* j = (trace_job*)arg;
    for (int64_t r = j->lo; r < j->hi; r++) {
        const float* ry = j->rays + 8 * r;
        yo_v3 o = v3(ry[0], ry[1], ry[2]), d = v3(ry[4], ry[5], ry[6]);
        int32_t p; float tt;
        yo_query(j->s, o, d, ry[3], ry[7], j->mode, -1, 0, &p, &tt, &j->nodes, &j->tris);
        j->prim[r] = p;
        if (j->t) j->t[r] = tt;
        if (j->u && j->v) {
            if (p >= 0 && j->mode == 0) hit_uv(j->s, p, o, d, tt, &j->u[r], &j->v[r]);
            else { j->u[r] = 0; j->v[r] = 0; }
        }
    }
    return NULL;
}

void yo_trace_bvh(yo_scene* s, const float* rays, int64_t n, int mode, int32_t* prim,
                  float* t, float* u, float* v, uint64_t* nodes, uint64_t* tris, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    trace_job jobs[64];
    pthread_t th[64];
    for (int i = 0; i < nthreads; i++) {
        jobs[i].s = s; jobs[i].rays = rays; jobs[i].mode = mode;
        jobs[i].lo = n * i / nthreads; jobs[i].hi = n * (i + 1) / nthreads;
        jobs[i].prim = prim; jobs[i].t = t; jobs[i].u = u; jobs[i].v = v;
        jobs[i].nodes = 0; jobs[i].tris = 0;
    }
    if (nthreads == 1) trace_worker(&jobs[0]);
    else {
        for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, trace_worker, &jobs[i]);
        for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    }
    uint64_t sn = 0, st = 0;
    for (int i = 0; i < nthreads; i++) { sn += jobs[i].nodes; st += jobs[i].tris; }
    if (nodes) *nodes = sn;
    if (tris) *tris = st;
}

/* Query log: while enabled, every yo_first_hit_one / yo_is_shadowed_one call
 * appends its inputs and its answer (8 floats per record), so that a full
 * render's queries can be replayed elsewhere and compared one by one.
 *   first hit: from xyz, ray xyz, prim (int32 bits), t
 *   shadow:    p xyz, l xyz, self_obj (int32 bits), to_point*2 + result (int32 bits) */
static struct {
    float *ext, *sh;
    int64_t ext_cap, sh_cap, ext_n, sh_n;
    int on;
} g_log;

void yo_log_begin(float* ext, int64_t ext_cap, float* sh, int64_t sh_cap)
{
    g_log.ext = ext; g_log.ext_cap = ext ? ext_cap : 0;
    g_log.sh = sh; g_log.sh_cap = sh ? sh_cap : 0;
    g_log.ext_n = 0; g_log.sh_n = 0;
    g_log.on = 1;
}

void yo_log_end(int64_t* ext_n, int64_t* sh_n)
{
    g_log.on = 0;
    if (ext_n) *ext_n = g_log.ext_n;
    if (sh_n) *sh_n = g_log.sh_n;
}

static void log_put(float* buf, int64_t cap, int64_t* n, const float* rec)
{
    int64_t k = __atomic_fetch_add(n, 1, __ATOMIC_RELAXED);
    if (k < cap) memcpy(buf + 8 * k, rec, 8 * sizeof(float));
}

/* scene_t::firstHit, scene.cc:688-721: f = from + ray*min_raydis; closest hit
 * with Z > 0 (a zero-distance closest hit counts as a miss, as the per-object
 * kd result would be rejected by the Z > 0 test). */
int yo_first_hit_one(const yo_scene* s, yo_v3 from, yo_v3 ray, float min_raydis, int32_t* prim,
                     float* t)
{
    return yo_first_hit_one_s(s, from, ray, min_raydis, 0, prim, t);
}

/* shadow: firstHit's shadow flag (scene.cc:698, objects that cast no shadows
 * are skipped) */
int yo_first_hit_one_s(const yo_scene* s, yo_v3 from, yo_v3 ray, float min_raydis, int shadow,
                       int32_t* prim, float* t)
{
    yo_v3 f = v3_add(from, v3_mul(ray, min_raydis));
    int32_t p; float tt;
    yo_query_s(s, f, ray, 0.0f, INFINITY, 0, -1, 0, shadow, &p, &tt, NULL, NULL);
    if (p >= 0 && !(tt > 0.0f)) p = -1;
    *prim = p;
    *t = tt;
    if (g_log.on) {
        float rec[8] = {from.x, from.y, from.z, ray.x, ray.y, ray.z, 0, tt};
        memcpy(&rec[6], &p, 4);
        log_put(g_log.ext, g_log.ext_cap, &g_log.ext_n, rec);
    }
    return p >= 0;
}

void yo_first_hit(yo_scene* s, const float* from, const float* dir, int64_t n, float min_raydis,
                  int32_t* prim, float* t)
{
    for (int64_t i = 0; i < n; i++)
        yo_first_hit_one(s, v3_ld(from + 3 * i), v3_ld(dir + 3 * i), min_raydis, &prim[i], &t[i]);
}

/* scene_t::isShadowed, scene.cc:107-146 (point) and 148-185 (direction): the
 * direction is normalised; the shading point's own object is tested from
 * p + ray*self_bias, every other object from p + ray*min_raydis; any hit with
 * 0 < t < dist occludes (dist = |l-p| or infinity). */
static int is_shadowed_q(const yo_scene* s, yo_v3 p, yo_v3 l, int self_obj, int to_point,
                         float min_raydis, float self_bias);

int yo_is_shadowed_one(const yo_scene* s, yo_v3 p, yo_v3 l, int self_obj, int to_point,
                       float min_raydis, float self_bias)
{
    const int r = is_shadowed_q(s, p, l, self_obj, to_point, min_raydis, self_bias);
    if (g_log.on) {
        float rec[8] = {p.x, p.y, p.z, l.x, l.y, l.z, 0, 0};
        const int32_t code = to_point * 2 + r;
        memcpy(&rec[6], &self_obj, 4);
        memcpy(&rec[7], &code, 4);
        log_put(g_log.sh, g_log.sh_cap, &g_log.sh_n, rec);
    }
    return r;
}

static int is_shadowed_q(const yo_scene* s, yo_v3 p, yo_v3 l, int self_obj, int to_point,
                         float min_raydis, float self_bias)
{
    yo_v3 ray;
    float dist;
    if (to_point) {
        ray = v3_sub(l, p);
        dist = v3_len(ray);
    } else {
        ray = l;
        dist = INFINITY;
    }
    ray = v3_norm(ray);
    yo_v3 self = v3_add(p, v3_mul(ray, self_bias));
    yo_v3 q = v3_add(p, v3_mul(ray, min_raydis));
    int32_t pr; float tt;
    if (self_obj >= 0) {
        if (yo_query_s(s, self, ray, 0.0f, dist, 1, self_obj, 1, 1, &pr, &tt, NULL, NULL)) return 1;
        return yo_query_s(s, q, ray, 0.0f, dist, 1, self_obj, 0, 1, &pr, &tt, NULL, NULL);
    }
    return yo_query_s(s, q, ray, 0.0f, dist, 1, -1, 0, 1, &pr, &tt, NULL, NULL);
}

void yo_is_shadowed(yo_scene* s, const float* p, const float* l, const int32_t* self_obj,
                    int64_t n, int to_point, float min_raydis, float self_bias, uint8_t* out)
{
    for (int64_t i = 0; i < n; i++)
        out[i] = (uint8_t)yo_is_shadowed_one(s, v3_ld(p + 3 * i), v3_ld(l + 3 * i),
                                             self_obj ? self_obj[i] : -1, to_point, min_raydis,
                                             self_bias);
}

/* ------------------------------------------------------------------ QMC -- */

void yo_halton_seq(int base, int64_t n, float* out)
{
    yo_halton h;
    halton_init(&h, base);
    for (int64_t i = 0; i < n; i++) out[i] = halton_next(&h);
}

/* Halton::setStart(i) digit loop, mcqmc.h:25-33, rounded like getNext's return */
float yo_halton_at(int base, uint32_t i)
{
    double inv = 1.0 / (double)base;
    double value = 0.0, f = inv, factor = inv;
    while (i > 0) {
        value += (double)(i % (uint32_t)base) * factor;
        i /= (uint32_t)base;
        factor *= f;
    }
    return (float)value;
}

/* RI_vdC, mcqmc.h:59-67: bit reversal */
float yo_ri_vdc(uint32_t bits, uint32_t r)
{
    bits = (bits << 16) | (bits >> 16);
    bits = ((bits & 0x00ff00ffu) << 8) | ((bits & 0xff00ff00u) >> 8);
    bits = ((bits & 0x0f0f0f0fu) << 4) | ((bits & 0xf0f0f0f0u) >> 4);
    bits = ((bits & 0x33333333u) << 2) | ((bits & 0xccccccccu) >> 2);
    bits = ((bits & 0x55555555u) << 1) | ((bits & 0xaaaaaaaau) >> 1);
    return (float)((double)(bits ^ r) / 4294967296.0);
}

/* RI_S, mcqmc.h:69-74: Sobol' generator matrix */
float yo_ri_s(uint32_t i, uint32_t r)
{
    for (uint32_t v = 1u << 31; i; i >>= 1, v ^= v >> 1)
        if (i & 1) r ^= v;
    return (float)((double)r / 4294967296.0);
}

/* RI_LP, mcqmc.h:76-81: Larcher-Pillichshammer generator matrix */
float yo_ri_lp(uint32_t i, uint32_t r)
{
    for (uint32_t v = 1u << 31; i; i >>= 1, v |= v >> 1)
        if (i & 1) r ^= v;
    return (float)((double)r / 4294967296.0);
}

/* nextPrime, mcqmc.h:84-97 */
int yo_next_prime(int last)
{
    int np = last + (last & 1) + 1;
    for (;;) {
        int dv = 3, ispr = 1;
        while (ispr && dv * dv <= np) { ispr = (np % dv) != 0; dv += 2; }
        if (ispr) break;
        np += 2;
    }
    return np;
}

float yo_our_random(int32_t* seed) { return our_random(seed); }

/* seed * 16807^n mod (2^31-1): the state ourRandom() reaches after n draws */
int32_t yo_lcg_jump(int32_t seed, uint64_t n)
{
    const uint64_t m = 2147483647u;
    uint64_t acc = 1, base = 16807;
    while (n) {
        if (n & 1) acc = (acc * base) % m;
        base = (base * base) % m;
        n >>= 1;
    }
    return (int32_t)(((uint64_t)seed * acc) % m);
}

/* glibc random_r TYPE_3 (the generator behind rand()): r[0] = seed,
 * r[i] = 16807 r[i-1] mod (2^31-1) for i < 31, r[i] = r[i-31] for i = 31..33,
 * r[i] = r[i-31] + r[i-3] (mod 2^32) after; outputs r[i] >> 1 from i = 344.
 * Drives blockSpliter_t's shuffle (renderblock.cc:82). */
void yo_glibc_rand(uint32_t seed, int64_t n, int32_t* out)
{
    int64_t total = 344 + n;
    uint32_t* r = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)total);
    if (seed == 0) seed = 1;
    r[0] = seed;
    for (int i = 1; i < 31; i++) {
        int64_t v = (16807LL * (int32_t)r[i - 1]) % 2147483647LL;
        if (v < 0) v += 2147483647LL;
        r[i] = (uint32_t)v;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (int64_t i = 34; i < total; i++) r[i] = r[i - 31] + r[i - 3];
    for (int64_t k = 0; k < n; k++) out[k] = (int32_t)(r[344 + k] >> 1);
    free(r);
}
