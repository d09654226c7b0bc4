// Autotuner over the BVH build (node width 2 / 4 / 8, leaf size, SAH bins)
// and three launch settings of the GI levels, timed on the caller's own
// frame: how many traversal stack entries stay in LDS (fewer: more waves
// per compute unit, more device-memory entries for deep rays), how many rays
// each lane of the extension and shadow launches traces (64 per wave and
// ray, fewer waves for more), and the survivor fraction below which a
// level's paths are compacted for the next level (above it the next level
// reuses the list and ended paths' lanes exit at once).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <vector>

#include "yg_internal.h"

namespace {

std::vector<int> bit_values(int mask, int lo, int hi)
{
    std::vector<int> v;
    for (int n = lo; n <= hi; n++)
        if (mask & (1 << n)) v.push_back(n);
    return v;
}

// The launch settings of build options b on the built scene s (yg_build's
// rules for their defaults)
void apply_launch(yg_scene* s, const yg_build_opts& b)
{
    s->stack_lds = b.stack_lds == 0 ? YG_DEFAULT_STACK_LDS : (b.stack_lds < 0 ? 0 : b.stack_lds);
    s->rays_per_lane = b.rays_per_lane ? b.rays_per_lane : 1;
    s->compact_below = b.compact_below == 0.0f ? 1.0f : (b.compact_below == -1.0f ? 0.0f : b.compact_below);
}

}  // namespace

extern "C" int yg_autotune(yg_scene* s, const yg_render_opts* opts, float* d_rgb, const yg_tune_opts* t,
                           yg_tune_result* r, void* stream)
{
    if (!s || !opts || !d_rgb || !r) return yg_fail(YG_ERR_ARG, "null argument");
    const int wmask = (t && t->widths) ? t->widths : (1 << 2) | (1 << 4) | (1 << 8);
    const int lmask = (t && t->leaves) ? t->leaves : (1 << 2) | (1 << 4) | (1 << 8);
    const int bmask = (t && t->bins) ? t->bins : (1 << 5);
    const int repeats = (t && t->repeats > 0) ? t->repeats : 2;
    const int threads = t ? t->threads : 0;
    const int smask = (t && t->stacks) ? t->stacks : 0;
    if (smask & ~0x79) return yg_fail(YG_ERR_ARG, "tune stacks: bits 1<<0 and 1<<3 .. 1<<6 only");
    const int rmask = (t && t->rays) ? t->rays : 1;
    const int cmask = t ? t->compacts : 0;
    if (rmask & ~0x7) return yg_fail(YG_ERR_ARG, "tune rays: bits 1<<0 .. 1<<2 only");
    if (cmask & ~0x1f) return yg_fail(YG_ERR_ARG, "tune compacts: bits 1<<0 .. 1<<4 only");
    std::vector<int> rays;  // rays_per_lane values
    for (int k : bit_values(rmask, 0, 2)) rays.push_back(1 << k);
    std::vector<float> compacts;  // compact_below values (-1 = never)
    if (!cmask) compacts.push_back(1.0f);
    if (cmask & 1) compacts.push_back(-1.0f);
    for (int k : bit_values(cmask, 1, 4)) compacts.push_back(0.25f * (float)k);
    if (wmask & ~((1 << 2) | (1 << 4) | (1 << 8))) return yg_fail(YG_ERR_ARG, "tune widths: bits 1<<2, 1<<4 and 1<<8 only");
    if (lmask & ~0x1fe) return yg_fail(YG_ERR_ARG, "tune leaves: bits 1<<1 .. 1<<8 only");
    if (bmask & ~0x78) return yg_fail(YG_ERR_ARG, "tune bins: bits 1<<3 .. 1<<6 only");
    const std::vector<int> widths = bit_values(wmask, 2, 8), leaves = bit_values(lmask, 1, 8),
                           bins = bit_values(bmask, 3, 6);
    std::vector<int> stacks;  // stack_lds values (0 = the build's default)
    if (!smask) stacks.push_back(0);
    if (smask & 1) stacks.push_back(-1);
    for (int k : bit_values(smask, 3, 6)) stacks.push_back(1 << k);
    const size_t n = widths.size() * leaves.size() * bins.size() * stacks.size() * rays.size() * compacts.size();
    if (n == 0 || n > YG_TUNE_MAX) return yg_fail(YG_ERR_ARG, "tune: 1..YG_TUNE_MAX candidates");
    memset(r, 0, sizeof(*r));
    yg_render_opts o = *opts;
    o.flags = (o.flags | YG_RENDER_TIMING) & ~YG_RENDER_KTIMING;
    int best = -1;
    for (int w : widths)
        for (int lf : leaves)
            for (int bk : bins) {
                yg_build_opts b{};
                b.width = w;
                b.max_leaf = lf;
                b.sah_bins = 1 << bk;
                b.threads = threads;
                const auto t0 = std::chrono::steady_clock::now();
                int rc = yg_build(s, &b);
                if (rc) return rc;
                const double build_ms =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                for (int sl : stacks)
                for (int rl : rays)
                for (float cb : compacts) {
                    // the LDS part, rays per lane and compaction threshold are
                    // launch settings of the built tree: no rebuild
                    b.stack_lds = sl;
                    b.rays_per_lane = rl;
                    b.compact_below = cb;
                    apply_launch(s, b);
                    float ms = 0;
                    for (int k = 0; k <= repeats; k++) {
                        yg_render_stats st;
                        if ((rc = yg_render(s, &o, d_rgb, &st, stream))) return rc;
                        if (k > 0 && (k == 1 || st.ms < ms)) ms = st.ms;  // the first render warms up
                    }
                    const int c = r->ncand++;
                    r->cand[c] = b;
                    r->cand_ms[c] = ms;
                    r->cand_build_ms[c] = (float)build_ms;
                    if (best < 0 || ms < r->cand_ms[best]) best = c;
                }
            }
    r->best = r->cand[best];
    r->best_ms = r->cand_ms[best];
    const yg_build_opts& last = r->cand[r->ncand - 1];
    if (r->best.width != last.width || r->best.max_leaf != last.max_leaf || r->best.sah_bins != last.sah_bins)
        return yg_build(s, &r->best);
    apply_launch(s, r->best);
    return YG_OK;
}
