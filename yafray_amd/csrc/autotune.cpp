// Autotuner over the BVH build (node width, leaf size, SAH bins) and the
// traversal stack's LDS part, timed on the caller's own frame.  The stack's
// depth bound is exact per tree; what is chosen is how many of its entries
// stay in LDS (fewer: more waves per compute unit, more device-memory
// entries for deep rays).  The other launch constants are not tuned: every
// lane traces one ray (64 per wave), and a level's compaction is a scan plus
// one scatter (well under 1% of a frame), so skipping it would save nothing.
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

}  // namespace

extern "C" int yg_autotune(yg_scene* s, const yg_render_opts* opts, float* d_rgb, const yg_tune_opts* t,
                           yg_tune_result* r, void* stream)
{
    if (!s || !opts || !d_rgb || !r) return yg_fail(YG_ERR_ARG, "null argument");
    const int wmask = (t && t->widths) ? t->widths : (1 << 2) | (1 << 4);
    const int lmask = (t && t->leaves) ? t->leaves : (1 << 2) | (1 << 4) | (1 << 8);
    const int bmask = (t && t->bins) ? t->bins : (1 << 5);
    const int repeats = (t && t->repeats > 0) ? t->repeats : 2;
    const int threads = t ? t->threads : 0;
    const int smask = (t && t->stacks) ? t->stacks : 0;
    if (smask & ~0x79) return yg_fail(YG_ERR_ARG, "tune stacks: bits 1<<0 and 1<<3 .. 1<<6 only");
    if (wmask & ~((1 << 2) | (1 << 4))) return yg_fail(YG_ERR_ARG, "tune widths: bits 1<<2 and 1<<4 only");
    if (lmask & ~0x1fe) return yg_fail(YG_ERR_ARG, "tune leaves: bits 1<<1 .. 1<<8 only");
    if (bmask & ~0x78) return yg_fail(YG_ERR_ARG, "tune bins: bits 1<<3 .. 1<<6 only");
    const std::vector<int> widths = bit_values(wmask, 2, 4), leaves = bit_values(lmask, 1, 8),
                           bins = bit_values(bmask, 3, 6);
    std::vector<int> stacks;  // stack_lds values (0 = the build's default)
    if (!smask) stacks.push_back(0);
    if (smask & 1) stacks.push_back(-1);
    for (int k : bit_values(smask, 3, 6)) stacks.push_back(1 << k);
    const size_t n = widths.size() * leaves.size() * bins.size() * stacks.size();
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
                for (int sl : stacks) {
                    // the LDS part is a launch setting of the built tree: no rebuild
                    b.stack_lds = sl;
                    s->stack_lds = sl == 0 ? YG_DEFAULT_STACK_LDS : (sl < 0 ? 0 : sl);
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
    s->stack_lds = r->best.stack_lds == 0 ? YG_DEFAULT_STACK_LDS : (r->best.stack_lds < 0 ? 0 : r->best.stack_lds);
    return YG_OK;
}
