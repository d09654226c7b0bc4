/* Canonical BVH2 visit counts for the frozen algorithmic byte model of the
 * roofline (TEST INFRASTRUCTURE: used by tests/golden/make_byte_model.py only).
 *
 * The byte model prices a ray query as
 *     32 B ray + 16 B hit (4 B any hit) + 32 B x nodes visited + 36 B x triangles tested
 * with the counts of a canonical two-wide BVH: binned SAH over 32 centroid
 * bins, at most 4 triangles per leaf, boxes exactly the FP64 bounds of their
 * triangles (no inflation), an exact FP64 slab test, near child first, and
 * closest-hit distance shrinking.  A node is visited when it is taken from the
 * stack (the root, and every child whose box the ray enters within the
 * current distance bound): n_node counts visits, one 32-byte node each.  The
 * child boxes a visit tests to order its children are counted separately
 * (*boxes_tested); a layout with both child boxes in the parent fetches those.
 * Independent of the parity tree of yo_core.c,
 * whose boxes are grown for bit-exact agreement with the reference's float
 * edge test and would overstate the work.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "yaf_oracle.h"

#define CB_BINS 32
#define CB_LEAF 4

typedef struct { double lo[3], hi[3]; } cb_box;
typedef struct { cb_box b; int32_t left, right, start, count; } cb_node;

typedef struct {
    const float* v;
    const uint32_t* idx;
    cb_box* tb;
    double* cen;
    int32_t* order;
    cb_node* nodes;
    int64_t nn, cap;
} cb_tree;

static void cb_empty(cb_box* b)
{
    for (int k = 0; k < 3; k++) { b->lo[k] = DBL_MAX; b->hi[k] = -DBL_MAX; }
}
static void cb_grow(cb_box* b, const cb_box* o)
{
    for (int k = 0; k < 3; k++) {
        if (o->lo[k] < b->lo[k]) b->lo[k] = o->lo[k];
        if (o->hi[k] > b->hi[k]) b->hi[k] = o->hi[k];
    }
}
static double cb_area(const cb_box* b)
{
    double dx = b->hi[0] - b->lo[0], dy = b->hi[1] - b->lo[1], dz = b->hi[2] - b->lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

static int32_t cb_alloc(cb_tree* t)
{
    if (t->nn == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->nodes = (cb_node*)realloc(t->nodes, sizeof(cb_node) * (size_t)t->cap);
    }
    return (int32_t)t->nn++;
}

static int cb_bin(double c, double lo, double scale)
{
    int b = (int)((c - lo) * scale);
    return b < 0 ? 0 : (b >= CB_BINS ? CB_BINS - 1 : b);
}

static void cb_build(cb_tree* t, int32_t ni, int64_t start, int64_t count)
{
    cb_box nb, cbx;
    cb_empty(&nb); cb_empty(&cbx);
    for (int64_t i = start; i < start + count; i++) {
        const int32_t p = t->order[i];
        cb_grow(&nb, &t->tb[p]);
        for (int k = 0; k < 3; k++) {
            const double c = t->cen[3 * (int64_t)p + k];
            if (c < cbx.lo[k]) cbx.lo[k] = c;
            if (c > cbx.hi[k]) cbx.hi[k] = c;
        }
    }
    t->nodes[ni].b = nb;
    t->nodes[ni].left = t->nodes[ni].right = -1;
    t->nodes[ni].start = (int32_t)start;
    t->nodes[ni].count = (int32_t)count;
    if (count <= CB_LEAF) return;
    int bax = -1, bsplit = -1;
    double bcost = DBL_MAX;
    for (int ax = 0; ax < 3; ax++) {
        const double ext = cbx.hi[ax] - cbx.lo[ax];
        if (!(ext > 0)) continue;
        const double scale = CB_BINS / ext;
        cb_box bb[CB_BINS];
        int64_t bn[CB_BINS];
        for (int b = 0; b < CB_BINS; b++) { cb_empty(&bb[b]); bn[b] = 0; }
        for (int64_t i = start; i < start + count; i++) {
            const int32_t p = t->order[i];
            const int b = cb_bin(t->cen[3 * (int64_t)p + ax], cbx.lo[ax], scale);
            bn[b]++;
            cb_grow(&bb[b], &t->tb[p]);
        }
        double ra[CB_BINS];
        int64_t rc[CB_BINS];
        cb_box acc;
        cb_empty(&acc);
        int64_t c = 0;
        for (int b = CB_BINS - 1; b > 0; b--) {
            cb_grow(&acc, &bb[b]); c += bn[b];
            ra[b] = cb_area(&acc); rc[b] = c;
        }
        cb_empty(&acc); c = 0;
        for (int b = 0; b < CB_BINS - 1; b++) {
            cb_grow(&acc, &bb[b]); c += bn[b];
            if (c == 0 || rc[b + 1] == 0) continue;
            const double cost = cb_area(&acc) * (double)c + ra[b + 1] * (double)rc[b + 1];
            if (cost < bcost) { bcost = cost; bax = ax; bsplit = b; }
        }
    }
    int64_t mid = start + count / 2;
    if (bax >= 0) {
        const double scale = CB_BINS / (cbx.hi[bax] - cbx.lo[bax]);
        int64_t i = start, j = start + count - 1;
        while (i <= j) {
            if (cb_bin(t->cen[3 * (int64_t)t->order[i] + bax], cbx.lo[bax], scale) <= bsplit) i++;
            else { const int32_t x = t->order[i]; t->order[i] = t->order[j]; t->order[j] = x; j--; }
        }
        mid = (i == start || i == start + count) ? start + count / 2 : i;
    }
    const int32_t l = cb_alloc(t), r = cb_alloc(t);
    t->nodes[ni].left = l;
    t->nodes[ni].right = r;
    t->nodes[ni].count = 0;
    cb_build(t, l, start, mid - start);
    cb_build(t, r, mid, start + count - mid);
}

/* exact FP64 slab test: entry distance of [t0, t1] against the box, or -1 */
static int cb_slab(const cb_box* b, const double* o, const double* inv, double t0, double t1, double* tn)
{
    for (int k = 0; k < 3; k++) {
        double ta = (b->lo[k] - o[k]) * inv[k], tb = (b->hi[k] - o[k]) * inv[k];
        if (isnan(ta) || isnan(tb)) {          /* zero direction on this axis: a range check */
            if (o[k] < b->lo[k] || o[k] > b->hi[k]) return 0;
            continue;
        }
        if (ta > tb) { const double x = ta; ta = tb; tb = x; }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
    }
    if (t0 > t1) return 0;
    *tn = t0;
    return 1;
}

/* FP64 ray / triangle distance (Moller-Trumbore); -1 on a miss */
static double cb_tri(const cb_tree* t, int32_t p, const double* o, const double* d)
{
    double a[3], b[3], c[3];
    for (int k = 0; k < 3; k++) {
        a[k] = t->v[3 * (int64_t)t->idx[3 * (int64_t)p] + k];
        b[k] = t->v[3 * (int64_t)t->idx[3 * (int64_t)p + 1] + k];
        c[k] = t->v[3 * (int64_t)t->idx[3 * (int64_t)p + 2] + k];
    }
    const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det == 0) return -1;
    const double inv = 1.0 / det;
    const double tv[3] = {o[0] - a[0], o[1] - a[1], o[2] - a[2]};
    const double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
    if (u < 0 || u > 1) return -1;
    const double q[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
    const double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    if (v < 0 || u + v > 1) return -1;
    return (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
}

static void cb_trace(const cb_tree* t, const float* ry, int mode, uint64_t* visits, uint64_t* tris,
                     uint64_t* boxes)
{
    const double o[3] = {ry[0], ry[1], ry[2]}, d[3] = {ry[4], ry[5], ry[6]};
    const double tmin = ry[3];
    double tmax = isinf(ry[7]) ? DBL_MAX : (double)ry[7];
    const double inv[3] = {1.0 / d[0], 1.0 / d[1], 1.0 / d[2]};
    int32_t stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const cb_node* n = &t->nodes[stack[--sp]];
        double tn;
        (*visits)++;
        (*boxes)++;
        /* the node's box against the current bound (a child pushed before
         * the bound shrank may now lie beyond it) */
        if (!cb_slab(&n->b, o, inv, tmin, tmax, &tn)) continue;
        if (n->left < 0) {
            for (int32_t i = n->start; i < n->start + n->count; i++) {
                (*tris)++;
                const double h = cb_tri(t, t->order[i], o, d);
                if (h > tmin && h < tmax) {
                    if (mode == 1) return;       /* any hit */
                    tmax = h;                    /* closest hit: shrink */
                }
            }
            continue;
        }
        double tl, tr;
        *boxes += 2;
        const int hl = cb_slab(&t->nodes[n->left].b, o, inv, tmin, tmax, &tl);
        const int hr = cb_slab(&t->nodes[n->right].b, o, inv, tmin, tmax, &tr);
        if (hl && hr) {
            if (tl <= tr) { stack[sp++] = n->right; stack[sp++] = n->left; }
            else { stack[sp++] = n->left; stack[sp++] = n->right; }
        } else if (hl) stack[sp++] = n->left;
        else if (hr) stack[sp++] = n->right;
    }
}

typedef struct {
    const cb_tree* t; const float* rays; int64_t lo, hi; int mode; uint64_t visits, tris, boxes;
} cb_job;

static void* cb_worker(void* a)
{
    cb_job* j = (cb_job*)a;
    for (int64_t r = j->lo; r < j->hi; r++)
        cb_trace(j->t, j->rays + 8 * r, j->mode, &j->visits, &j->tris, &j->boxes);
    return NULL;
}

int64_t yo_canonical_counts(const float* verts, int64_t nv, const uint32_t* idx, int64_t nt, const float* rays,
                            int64_t n, int mode, int threads, uint64_t* visits, uint64_t* tris,
                            uint64_t* boxes_tested)
{
    (void)nv;
    cb_tree t;
    memset(&t, 0, sizeof(t));
    t.v = verts;
    t.idx = idx;
    t.tb = (cb_box*)malloc(sizeof(cb_box) * (size_t)(nt ? nt : 1));
    t.cen = (double*)malloc(sizeof(double) * 3 * (size_t)(nt ? nt : 1));
    t.order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nt ? nt : 1));
    for (int64_t i = 0; i < nt; i++) {
        cb_empty(&t.tb[i]);
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) {
                const double x = verts[3 * (int64_t)idx[3 * i + c] + k];
                if (x < t.tb[i].lo[k]) t.tb[i].lo[k] = x;
                if (x > t.tb[i].hi[k]) t.tb[i].hi[k] = x;
            }
        for (int k = 0; k < 3; k++) t.cen[3 * i + k] = 0.5 * (t.tb[i].lo[k] + t.tb[i].hi[k]);
        t.order[i] = (int32_t)i;
    }
    const int32_t root = cb_alloc(&t);
    cb_build(&t, root, 0, nt);
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    cb_job jobs[64];
    for (int k = 0; k < threads; k++) {
        jobs[k].t = &t; jobs[k].rays = rays; jobs[k].mode = mode;
        jobs[k].visits = 0; jobs[k].tris = 0; jobs[k].boxes = 0;
        jobs[k].lo = n * k / threads; jobs[k].hi = n * (k + 1) / threads;
        pthread_create(&th[k], NULL, cb_worker, &jobs[k]);
    }
    *visits = 0;
    *tris = 0;
    *boxes_tested = 0;
    for (int k = 0; k < threads; k++) {
        pthread_join(th[k], NULL);
        *visits += jobs[k].visits;
        *tris += jobs[k].tris;
        *boxes_tested += jobs[k].boxes;
    }
    const int64_t nodes = t.nn;
    free(t.tb); free(t.cen); free(t.order); free(t.nodes);
    return nodes;
}
