// Host-side binned-SAH BVH builder for libyafgpu.
//
// Replaces kdTree_t::kdTree_t / buildTree / pigeonMinCost (kdtree.cc:57-582)
// as the acceleration structure over each scene.  The tree shape never
// changes results (the triangle acceptance rules do), only speed; it is a
// two-wide BVH because 64-lane wavefronts traverse it with one 64-byte node
// fetch per step, which a kd-tree's split-plane walk cannot match.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>

#include "yg_internal.h"

namespace {

struct BBox {
    float lo[3], hi[3];
    void empty()
    {
        for (int k = 0; k < 3; k++) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
    }
    void grow(const BBox& o)
    {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], o.lo[k]);
            hi[k] = std::max(hi[k], o.hi[k]);
        }
    }
    void grow(const float* p)
    {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    double area() const
    {
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode {
    BBox box;
    int32_t left, right;   // -1 for leaves
    int32_t start, count;
    int32_t depth;
};

struct Builder {
    const BBox* tb;
    const float* cen;      // 3 per triangle
    int32_t* ord;
    BNode* nodes;
    std::atomic<int64_t> next{0};
    std::atomic<int> active{1};
    int max_threads = 1;
    int max_leaf = 4;
    int bins = 32;
    std::atomic<int> max_depth{0};

    int64_t alloc2() { return next.fetch_add(2); }

    void build(int64_t ni, int64_t start, int64_t count, int depth)
    {
        BNode& n = nodes[ni];
        BBox nb, cb;
        nb.empty();
        cb.empty();
        for (int64_t i = start; i < start + count; i++) {
            nb.grow(tb[ord[i]]);
            cb.grow(&cen[3 * (int64_t)ord[i]]);
        }
        n.box = nb;
        n.left = n.right = -1;
        n.start = (int32_t)start;
        n.count = (int32_t)count;
        n.depth = depth;
        int md = max_depth.load();
        while (depth > md && !max_depth.compare_exchange_weak(md, depth)) {}
        if (count <= 1) return;

        int best_axis = -1, best_split = -1;
        double best_cost = DBL_MAX;
        const bool median = depth >= 60;  // keep the depth bounded for the traversal stack
        if (!median) {
            std::vector<BBox> bb(bins);
            std::vector<int64_t> bn(bins);
            std::vector<double> rarea(bins);
            std::vector<int64_t> rcnt(bins);
            for (int ax = 0; ax < 3; ax++) {
                double ext = (double)cb.hi[ax] - cb.lo[ax];
                if (!(ext > 0)) continue;
                for (int b = 0; b < bins; b++) { bb[b].empty(); bn[b] = 0; }
                double scale = bins / ext;
                for (int64_t i = start; i < start + count; i++) {
                    int b = (int)(((double)cen[3 * (int64_t)ord[i] + ax] - cb.lo[ax]) * scale);
                    b = std::min(std::max(b, 0), bins - 1);
                    bn[b]++;
                    bb[b].grow(tb[ord[i]]);
                }
                BBox acc;
                acc.empty();
                int64_t c = 0;
                for (int b = bins - 1; b > 0; b--) {
                    acc.grow(bb[b]);
                    c += bn[b];
                    rarea[b] = acc.area();
                    rcnt[b] = c;
                }
                acc.empty();
                c = 0;
                for (int b = 0; b < bins - 1; b++) {
                    acc.grow(bb[b]);
                    c += bn[b];
                    if (c == 0 || rcnt[b + 1] == 0) continue;
                    double cost = acc.area() * (double)c + rarea[b + 1] * (double)rcnt[b + 1];
                    if (cost < best_cost) { best_cost = cost; best_axis = ax; best_split = b; }
                }
            }
            double pa = nb.area();
            // SAH with unit traversal and intersection costs
            double split_cost = (pa > 0) ? 1.0 + best_cost / pa : DBL_MAX;
            if (count <= max_leaf && (best_axis < 0 || split_cost >= (double)count)) return;
            if (count <= YG_LEAF_MAX && best_axis >= 0 && split_cost >= (double)count * 1.0 &&
                count <= max_leaf * 2)
                return;
        }
        int64_t mid;
        if (best_axis < 0) {
            // all centroids coincide (or depth cap): object median on the widest axis
            int ax = 0;
            float w = -1;
            for (int k = 0; k < 3; k++)
                if (cb.hi[k] - cb.lo[k] > w) { w = cb.hi[k] - cb.lo[k]; ax = k; }
            mid = start + count / 2;
            std::nth_element(ord + start, ord + mid, ord + start + count,
                             [&](int32_t x, int32_t y) { return cen[3 * (int64_t)x + ax] < cen[3 * (int64_t)y + ax]; });
        } else {
            double ext = (double)cb.hi[best_axis] - cb.lo[best_axis];
            double scale = bins / ext;
            int32_t* p = std::partition(ord + start, ord + start + count, [&](int32_t t) {
                int b = (int)(((double)cen[3 * (int64_t)t + best_axis] - cb.lo[best_axis]) * scale);
                b = std::min(std::max(b, 0), bins - 1);
                return b <= best_split;
            });
            mid = p - ord;
            if (mid == start || mid == start + count) mid = start + count / 2;
        }
        int64_t l = alloc2();
        int64_t r = l + 1;
        nodes[ni].left = (int32_t)l;
        nodes[ni].right = (int32_t)r;
        int64_t nl = mid - start, nr = start + count - mid;
        bool spawn = false;
        if (std::min(nl, nr) > 20000) {
            int a = active.load();
            while (a < max_threads && !(spawn = active.compare_exchange_weak(a, a + 1))) {}
        }
        if (spawn) {
            std::thread th([this, l, start, nl, depth]() {
                build(l, start, nl, depth + 1);
                active.fetch_sub(1);
            });
            build(r, mid, nr, depth + 1);
            th.join();
        } else {
            build(l, start, nl, depth + 1);
            build(r, mid, nr, depth + 1);
        }
    }
};

// Boxes are widened outward by 2^-18 relative to their own coordinates and to
// the scene's largest coordinate G.  The float edge test of a triangle accepts
// rays that pass up to a few ulps of the operands' magnitude outside it; the
// G term keeps such rays inside the box of a triangle whose own coordinates
// are small (a terrain cell near z = 0 has almost no relative slack in z).
static void inflate(const BBox& b, float G, float lo[3], float hi[3])
{
    for (int k = 0; k < 3; k++) {
        float ext = b.hi[k] - b.lo[k];
        float e = (fabsf(b.lo[k]) + fabsf(b.hi[k]) + ext + G) * 3.814697265625e-06f + 1e-30f;
        lo[k] = b.lo[k] - e;
        hi[k] = b.hi[k] + e;
    }
}

}  // namespace

// Box rows of a wide node (0..5: lo.x hi.x lo.y hi.y lo.z hi.z, one float
// per child) and its child references.
static float* node_row(GNode4& g, int r) { return &(&g.lox)[r].x; }
static int32_t* node_child(GNode4& g) { return &g.child.x; }
static float* node_row(GNode8& g, int r) { return g.box[r]; }
static int32_t* node_child(GNode8& g) { return g.child; }

// Collapses the binary tree into N-wide nodes (N = 4 or 8): a node's children
// start as the binary node's two and the largest-area inner child is replaced
// by its own two children until there are N or only leaves remain.  Boxes and
// leaf references are those of the binary tree, so the width changes the
// traversal order only, never which triangles a ray is tested against.
// `stack` gets the deepest traversal stack the tree can need: a visit pushes
// at most (children - 1) entries on top of what its ancestors left.
template <int N, class Node>
static void export_wide(const BNode* bn, float G, std::vector<Node>& nodes, int32_t& stack,
                        int32_t (*enc_leaf)(const BNode&), void (*inflate_fn)(const BBox&, float, float*, float*))
{
    nodes.clear();
    stack = 0;
    auto set_empty = [](Node& g, int k) {
        float* f[6] = {node_row(g, 0), node_row(g, 1), node_row(g, 2), node_row(g, 3), node_row(g, 4), node_row(g, 5)};
        for (int c = 0; c < 6; c++) f[c][k] = INFINITY;
        node_child(g)[k] = 0;
    };
    auto set_child = [&](Node& g, int k, const BNode& c, int32_t ref) {
        float lo[3], hi[3];
        inflate_fn(c.box, G, lo, hi);
        for (int a = 0; a < 3; a++) {
            node_row(g, 2 * a)[k] = lo[a];
            node_row(g, 2 * a + 1)[k] = hi[a];
        }
        node_child(g)[k] = ref;
    };
    auto children = [&](int64_t bi, int64_t* ch) -> int {
        int n = 0;
        ch[n++] = bn[bi].left;
        ch[n++] = bn[bi].right;
        while (n < N) {
            int best = -1;
            double ba = -1.0;
            for (int k = 0; k < n; k++)
                if (bn[ch[k]].left >= 0 && bn[ch[k]].box.area() > ba) { ba = bn[ch[k]].box.area(); best = k; }
            if (best < 0) break;
            const int64_t x = ch[best];
            for (int k = n; k > best + 1; k--) ch[k] = ch[k - 1];  // keep left-to-right order
            ch[best] = bn[x].left;
            ch[best + 1] = bn[x].right;
            n++;
        }
        return n;
    };
    const BNode& root = bn[0];
    if (root.left < 0) {
        Node g;
        memset(&g, 0, sizeof(g));
        for (int k = 0; k < N; k++) set_empty(g, k);
        if (root.count > 0) set_child(g, 0, root, enc_leaf(root));
        nodes.push_back(g);
        stack = 0;
        return;
    }
    // (binary node, slot, stack entries pending above this node)
    struct Work { int64_t bi; int64_t gi; int32_t pend; };
    std::vector<Work> work;
    nodes.emplace_back();
    work.push_back({0, 0, 0});
    while (!work.empty()) {
        const Work w = work.back();
        work.pop_back();
        int64_t ch[N];
        const int n = children(w.bi, ch);
        Node g;
        memset(&g, 0, sizeof(g));
        int32_t refs[N];
        for (int k = 0; k < n; k++) {
            if (bn[ch[k]].left < 0) refs[k] = enc_leaf(bn[ch[k]]);
            else {
                refs[k] = (int32_t)nodes.size();
                nodes.emplace_back();
            }
        }
        for (int k = 0; k < N; k++) {
            if (k < n) set_child(g, k, bn[ch[k]], refs[k]);
            else set_empty(g, k);
        }
        nodes[w.gi] = g;
        const int32_t pend = w.pend + (n - 1);
        stack = std::max(stack, pend);
        for (int k = n - 1; k >= 0; k--)
            if (bn[ch[k]].left >= 0) work.push_back({ch[k], refs[k], pend});
    }
}

static void export_width(int width, const BNode* bn, float G, HostBVH& out)
{
    auto enc = [](const BNode& n) -> int32_t { return ~(int32_t)(n.start * YG_LEAF_MAX + (n.count - 1)); };
    if (width == 4) export_wide<4>(bn, G, out.nodes4, out.stack4, enc, inflate);
    if (width == 8) export_wide<8>(bn, G, out.nodes8, out.stack8, enc, inflate);
}

void build_bvh(const float* verts, const uint32_t* idx, int64_t nt, const int32_t* tri_obj,
               const int32_t* tri_mat, int max_leaf, int bins, int threads, int width, HostBVH& out)
{
    out.nodes.clear();
    out.tris.clear();
    out.tprim.clear();
    out.tobj.clear();
    out.tmat.clear();
    if (max_leaf < 1) max_leaf = 4;
    if (max_leaf > YG_LEAF_MAX) max_leaf = YG_LEAF_MAX;
    if (bins < 4) bins = 32;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<BBox> tb((size_t)std::max<int64_t>(nt, 1));
    std::vector<float> cen((size_t)std::max<int64_t>(3 * nt, 3));
    std::vector<int32_t> ord((size_t)std::max<int64_t>(nt, 1));
    {
        std::vector<std::thread> ths;
        for (int t = 0; t < threads; t++)
            ths.emplace_back([&, t]() {
                for (int64_t i = nt * t / threads; i < nt * (t + 1) / threads; i++) {
                    BBox b;
                    b.empty();
                    for (int k = 0; k < 3; k++) b.grow(&verts[3 * (int64_t)idx[3 * i + k]]);
                    tb[i] = b;
                    for (int k = 0; k < 3; k++) cen[3 * i + k] = 0.5f * (b.lo[k] + b.hi[k]);
                    ord[i] = (int32_t)i;
                }
            });
        for (auto& th : ths) th.join();
    }
    // uninitialised storage: only the nodes the build touches are written
    std::unique_ptr<BNode[]> bn(new BNode[(size_t)std::max<int64_t>(2 * nt + 2, 2)]);
    Builder B;
    B.tb = tb.data();
    B.cen = cen.data();
    B.ord = ord.data();
    B.nodes = bn.get();
    B.max_threads = threads;
    B.max_leaf = max_leaf;
    B.bins = bins;
    B.next = 1;
    if (nt > 0) B.build(0, 0, nt, 0);
    else {
        bn[0].box.empty();
        bn[0].left = bn[0].right = -1;
        bn[0].start = 0;
        bn[0].count = 0;
    }
    out.max_depth = B.max_depth.load() + 1;

    // leaf-ordered triangles
    out.tris.resize((size_t)nt * 3);
    out.tprim.resize((size_t)nt);
    out.tobj.resize((size_t)nt);
    out.tmat.resize((size_t)nt);
    for (int64_t i = 0; i < nt; i++) {
        int32_t p = ord[i];
        f3 a = ld3(&verts[3 * (int64_t)idx[3 * p]]);
        f3 b = ld3(&verts[3 * (int64_t)idx[3 * p + 1]]);
        f3 c = ld3(&verts[3 * (int64_t)idx[3 * p + 2]]);
        f3 n = tri_normal(a, b, c);
        out.tris[3 * i] = make_float4(a.x, a.y, a.z, n.x);
        out.tris[3 * i + 1] = make_float4(b.x, b.y, b.z, n.y);
        out.tris[3 * i + 2] = make_float4(c.x, c.y, c.z, n.z);
        out.tprim[i] = p;
        out.tobj[i] = tri_obj ? tri_obj[p] : 0;
        out.tmat[i] = tri_mat ? tri_mat[p] : 0;
    }

    // scene magnitude for the box inflation
    float G = 0.0f;
    if (nt > 0)
        for (int k = 0; k < 3; k++) G = std::max(G, std::max(fabsf(bn[0].box.lo[k]), fabsf(bn[0].box.hi[k])));

    // inner nodes in depth-first order, both child boxes per node
    auto enc_leaf = [&](const BNode& n) -> int32_t {
        return ~(int32_t)(n.start * YG_LEAF_MAX + (n.count - 1));
    };
    std::vector<int64_t> stack;
    std::vector<std::pair<int64_t, int64_t>> work;  // (builder node, gnode slot)
    auto set_child = [&](GNode& g, int which, const BNode& c, int32_t ref) {
        float lo[3], hi[3];
        if (c.count == 0 && c.left < 0) {
            lo[0] = lo[1] = lo[2] = 1.0f;
            hi[0] = hi[1] = hi[2] = -1.0f;
        } else inflate(c.box, G, lo, hi);
        if (which == 0) {
            g.a.x = lo[0]; g.a.y = hi[0]; g.a.z = lo[1]; g.a.w = hi[1];
            g.c.x = lo[2]; g.c.y = hi[2];
            g.d.x = ref;
        } else {
            g.b.x = lo[0]; g.b.y = hi[0]; g.b.z = lo[1]; g.b.w = hi[1];
            g.c.z = lo[2]; g.c.w = hi[2];
            g.d.y = ref;
        }
    };
    out.nodes.reserve((size_t)(B.next.load() / 2 + 1));
    const BNode& root = bn[0];
    if (root.left < 0) {
        // a single leaf: wrap it in a root with an empty second child
        GNode g;
        memset(&g, 0, sizeof(g));
        set_child(g, 0, root, root.count > 0 ? enc_leaf(root) : ~0);
        BNode emptyn;
        emptyn.box.empty();
        emptyn.left = emptyn.right = -1;
        emptyn.count = 0;
        emptyn.start = 0;
        set_child(g, 1, emptyn, ~0);
        if (root.count == 0) set_child(g, 0, emptyn, ~0);
        out.nodes.push_back(g);
        export_width(width, bn.get(), G, out);
        return;
    }
    out.nodes.emplace_back();
    work.push_back({0, 0});
    while (!work.empty()) {
        auto [bi, gi] = work.back();
        work.pop_back();
        const BNode& n = bn[bi];
        GNode g;
        memset(&g, 0, sizeof(g));
        const BNode* ch[2] = {&bn[n.left], &bn[n.right]};
        int32_t refs[2];
        for (int k = 0; k < 2; k++) {
            if (ch[k]->left < 0) refs[k] = enc_leaf(*ch[k]);
            else {
                refs[k] = (int32_t)out.nodes.size();
                out.nodes.emplace_back();
            }
        }
        for (int k = 0; k < 2; k++) set_child(g, k, *ch[k], refs[k]);
        out.nodes[gi] = g;
        // right pushed first so the left subtree is laid out next (depth-first)
        for (int k = 1; k >= 0; k--)
            if (ch[k]->left >= 0) work.push_back({(int64_t)(k == 0 ? n.left : n.right), refs[k]});
    }
    export_width(width, bn.get(), G, out);
}
