// Irradiance cache of the pathlight (cache = on): the fake pass that places
// the samples (pathLight_t::cached, pathlight.cc:676-700, with
// lightCache_t::enoughFor / insert, lightcache.cc:182-237) and the render
// pass that interpolates them (pathLight_t::interpolate, pathlight.cc:543-599,
// through the area's cacheProxy_t and lightCache_t::gatherSamples,
// cacheproxy.cc:17-87, lightcache.cc:111-180).  Included by render.hip.
//
// Work layout.  Both passes are sequential within a render area in the
// reference: a visit takes a sample only if no earlier sample covers it, and
// a gather starts from the radius the area's previous gather left.  One wave
// walks one area's pixels in the reference's order; the lanes share each
// step (the coverage scan, the sample's first-hit rays, the candidate scan).
// Areas run side by side: in the fake pass an area waits only for the
// earlier areas within reach of its samples (the host groups areas into
// levels), in the render pass areas are independent (the cache is final).
// A sample lives in the slot of the item whose visit took it.
#pragma once

struct CacheView {
    f3* pP;            // (screen x, screen y / aspect, log distance)
    f3* P;
    f3* N;             // face-forwarded shading normal
    float* adist;      // harmonic mean distance of the first hits (H)
    float* M;          // smallest first-hit distance
    float* prec;       // distance x world resolution
    c3* col;           // takeSample's estimate
    int32_t* rank;     // position in the generic tree's leaf order (render pass)
    uint32_t* in;      // 1: the item's visit took a sample
    uint32_t* list;    // area a's samples in order at [a.item_start, a.item_start + cnt[a])
    uint32_t* cnt;     // samples per area
    int32_t* row;      // image row of the sample's pixel
    uint32_t* nb;      // fake pass: an area's candidates from earlier areas (item ids), same slots
    unsigned long long* err;  // [0]: candidate overflow, [1]: empty gathers
};

struct CacheParams {
    float cs;          // cache_size
    float ycorr;       // 1 / aspect
    float worldres;    // (1 / resx) / focal
    float wfill;       // desiredWeight * rq, rq = 1 at camera hits
    float wlimit;      // weightLimit = 0.8 desiredWeight
    float sradius;     // searchRadius = 2 cache_size
    int K;             // search
    int reach;         // pixel rows within cache_size of a pixel's polar position
    // render pass candidate grid over (pP.x, pP.y), cell = sradius
    const int32_t* gstart;
    const uint32_t* gitem;
    float gx0, gy0, gcell;
    int gnx, gny;
};

// hash3d_t-free distance of two polar positions, polarDist (lightcache.h:112)
__device__ __forceinline__ float lc_pdist(f3 a, f3 b) { return len3(a - b); }

// pathLight_t::weight (prec = true) / weightNoPrec (pathlight.cc:409-465);
// samples are never devaluated here (refinement is off)
__device__ __forceinline__ float lc_weight(f3 sP, f3 sN, float adist, float prec_v, f3 P, f3 N, float maxweight,
                                           bool prec)
{
    f3 PP = P - sP;
    if (adist == 0.0f) return 0.0f;
    float D;
    if (prec) {
        D = (float)((double)normlen3(PP) - (double)prec_v * 2.0);
        if (D < 0) D = 0;
    } else
        D = normlen3(PP);
    const float a = D / adist;
    const float cinv = (float)(1.000001 - (double)dot(sN, N));
    const float b = sqrtf(cinv);
    const float c = fabsf(dot(N, PP));
    float f = a + ((b > c) ? b : c);
    if (f == 0) return maxweight;
    f = (float)(1.0 / (double)f);
    return (f > maxweight) ? maxweight : f;
}

// lightCache_t::toPolar (lightcache.h:107-111) for pixel (i, j): the fake
// pass forms the screen position with double literals (scene.cc:841-842),
// the render pass in float (scene.cc:749-750); log is the C library's double
// log of the float distance
__device__ __forceinline__ f3 lc_polar(int i, int j, int resx, int resy, float ycorr, float Z, bool fake)
{
    float sx, sy;
    if (fake) {
        sx = (float)(2.0 * (((double)j + 0.5) / (double)(float)resx) - 1.0);
        sy = (float)(1.0 - 2.0 * (((double)i + 0.5) / (double)(float)resy));
    } else {
        sx = (float)(2.0 * (double)(((float)j + 0.5f) / (float)resx) - 1.0);
        sy = (float)(1.0 - 2.0 * (double)(((float)i + 0.5f) / (float)resy));
    }
    return mk3(sx, sy * ycorr, (float)log((double)Z));
}

#define LC_WAVE 64
#define LC_OWN 1024  // ring of an area's most recent samples kept in LDS by the fake pass
#define LC_NBL 1024  // candidates from earlier areas kept in LDS

// Per-item surface data of the camera rays' hits, computed once for all
// items (k_cache_prep) so that the area walks read it in coalesced chunks.
struct VisitView {
    f3* P;
    f3* N;      // face-forwarded shading normal
    f3* U;      // sp.NU, sp.NV
    f3* V;
    f3* pP;     // the fake pass's polar position
    float* Z;
    c3* dif;    // getDiffuse(state, sp, eye) at the camera hit (pathlight.cc:536,582)
};

__global__ __launch_bounds__(RB) void k_cache_prep(const RenderScene* __restrict__ scenes,
                                                   const Area* __restrict__ areas, int nareas, uint32_t nitems,
                                                   const int32_t* __restrict__ item_ti,
                                                   const float* __restrict__ item_t,
                                                   const uint32_t* __restrict__ qual, VisitView vv, CacheParams cp)
{
    const uint32_t item = blockIdx.x * RB + threadIdx.x;
    if (item >= nitems || !qual[item]) return;
    const Area a = areas[find_area(areas, nareas, item)];
    const RenderScene& R = scenes[a.scene];
    const float Z = item_t[item];
    f3 eye;
    const Surf sp = primary_surf(R, a, item, item_ti[item], Z, eye);
    int i, j;
    item_pixel(a, item, i, j);
    vv.P[item] = sp.P;
    vv.N[item] = face_forward(sp.Ng, sp.N, eye);
    vv.U[item] = sp.NU;
    vv.V[item] = sp.NV;
    vv.pP[item] = lc_polar(i, j, R.resx, R.resy, cp.ycorr, Z, true);
    vv.Z[item] = Z;
    vv.dif[item] = get_diffuse(R, sp, eye);
}

// Fake pass of the areas of one level: one wave per area.  Candidates of a
// visit: the area's own samples so far (the most recent LC_OWN in an LDS
// ring) and the samples of earlier areas within reach of the area (gathered
// into cv.nb at the start, sorted by row, the first LC_NBL copied to LDS),
// both scanned only over the rows within reach; own samples newest first.
template <int W>
__global__ __launch_bounds__(LC_WAVE) void k_cache_fill(const RenderScene* __restrict__ scenes,
                                                        const Area* __restrict__ areas, int nareas,
                                                        const int32_t* __restrict__ level_areas,
                                                        const int32_t* __restrict__ adj_start,
                                                        const int32_t* __restrict__ adj,
                                                        const uint32_t* __restrict__ qual, VisitView vv,
                                                        CacheView cv, CacheParams cp)
{
    extern __shared__ int s_stk[];
    __shared__ float s_z[LC_WAVE];
    __shared__ f3 o_pP[LC_OWN], o_P[LC_OWN], o_N[LC_OWN];
    __shared__ float o_H[LC_OWN];
    __shared__ int o_row[LC_OWN];
    __shared__ f3 n_pP[LC_NBL], n_P[LC_NBL], n_N[LC_NBL];
    __shared__ float n_H[LC_NBL];
    __shared__ int n_row[LC_NBL];
    __shared__ f3 v_P[LC_WAVE], v_N[LC_WAVE], v_U[LC_WAVE], v_V[LC_WAVE], v_pP[LC_WAVE];
    __shared__ float v_Z[LC_WAVE];
    __shared__ uint32_t v_q[LC_WAVE];
    const int lane = threadIdx.x;
    const int ai = level_areas[blockIdx.x];
    const Area a = areas[ai];
    const RenderScene& R = scenes[a.scene];
    LaneStack st;
    st.init(s_stk, lane, LC_WAVE);
    const int S = R.samples;
    const uint32_t base = a.item_start;
    const uint32_t n_area = (uint32_t)a.w * (uint32_t)a.h;
    // candidates of the earlier areas in reach: their samples within `reach`
    // pixels of this area's rectangle
    uint32_t nnb = 0;
    const int rlo = a.y - cp.reach, rhi = a.y + a.h - 1 + cp.reach;
    const int clo = a.x - cp.reach, chi = a.x + a.w - 1 + cp.reach;
    for (int q = adj_start[ai]; q < adj_start[ai + 1]; q++) {
        const int bi = adj[q];
        const Area b = areas[bi];
        const uint32_t nb = cv.cnt[bi];
        for (uint32_t k0 = 0; k0 < nb; k0 += LC_WAVE) {
            const uint32_t k = k0 + lane;
            bool keep = false;
            uint32_t id = 0;
            if (k < nb) {
                id = cv.list[b.item_start + k];
                int pi, pj;
                item_pixel(b, id, pi, pj);
                keep = pi >= rlo && pi <= rhi && pj >= clo && pj <= chi;
            }
            const uint64_t m = __ballot(keep);
            if (keep) {
                const uint32_t pos = nnb + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                if (pos < n_area) cv.nb[base + pos] = id;
            }
            nnb += (uint32_t)__popcll(m);
        }
    }
    if (nnb > n_area) {
        if (lane == 0) atomicAdd(&cv.err[0], 1ull);
        return;
    }
    // by image row (lane 0: a few hundred entries along the area's border)
    if (lane == 0) {
        for (uint32_t k = 1; k < nnb; k++) {
            const uint32_t id = cv.nb[base + k];
            const int pi = cv.row[id];
            uint32_t m = k;
            while (m > 0 && cv.row[cv.nb[base + m - 1]] > pi) {
                cv.nb[base + m] = cv.nb[base + m - 1];
                m--;
            }
            cv.nb[base + m] = id;
        }
    }
    __syncthreads();
    for (uint32_t k = lane; k < nnb && k < LC_NBL; k += LC_WAVE) {
        const uint32_t id = cv.nb[base + k];
        n_pP[k] = cv.pP[id];
        n_P[k] = cv.P[id];
        n_N[k] = cv.N[id];
        n_H[k] = cv.adist[id];
        n_row[k] = cv.row[id];
    }
    __syncthreads();
    uint32_t nin = 0, own_lo = 0, nb_lo = 0, nb_hi = 0;
    const float maxw = (float)((double)cp.wfill * 2.0);
    for (uint32_t c0 = 0; c0 < n_area; c0 += LC_WAVE) {
        const uint32_t nc = (n_area - c0 < LC_WAVE) ? n_area - c0 : LC_WAVE;
        {
            const uint32_t item = base + c0 + lane;
            const uint32_t q = (lane < (int)nc) ? qual[item] : 0u;
            v_q[lane] = q;
            if (q) {
                v_P[lane] = vv.P[item];
                v_N[lane] = vv.N[item];
                v_U[lane] = vv.U[item];
                v_V[lane] = vv.V[item];
                v_pP[lane] = vv.pP[item];
                v_Z[lane] = vv.Z[item];
            }
        }
        __syncthreads();
        for (uint32_t u = 0; u < nc; u++) {
            if (!v_q[u]) continue;
            const uint32_t li = c0 + u, item = base + li;
            const int pi = a.y + (int)(li / (uint32_t)a.w);
            const f3 P = v_P[u], N = v_N[u], pP = v_pP[u];
            // rows within reach: advance the windows
            while (own_lo < nin) {
                const int r = (nin - own_lo <= LC_OWN) ? o_row[own_lo % LC_OWN] : cv.row[cv.list[base + own_lo]];
                if (r >= pi - cp.reach) break;
                own_lo++;
            }
            while (nb_lo < nnb) {
                const int r = (nb_lo < LC_NBL) ? n_row[nb_lo] : cv.row[cv.nb[base + nb_lo]];
                if (r >= pi - cp.reach) break;
                nb_lo++;
            }
            if (nb_hi < nb_lo) nb_hi = nb_lo;
            while (nb_hi < nnb) {
                const int r = (nb_hi < LC_NBL) ? n_row[nb_hi] : cv.row[cv.nb[base + nb_hi]];
                if (r > pi + cp.reach) break;
                nb_hi++;
            }
            // enoughFor: a sample within cache_size whose weightNoPrec reaches the limit
            bool cov = false;
            const uint32_t n_own = nin - own_lo;
            for (uint32_t k0 = 0; k0 < n_own && !cov; k0 += LC_WAVE) {
                bool c = false;
                if (k0 + lane < n_own) {
                    const uint32_t k = nin - 1 - (k0 + lane);  // newest first
                    f3 q, sP, sN;
                    float H;
                    if (nin - k <= LC_OWN) {
                        const uint32_t r = k % LC_OWN;
                        q = o_pP[r]; sP = o_P[r]; sN = o_N[r]; H = o_H[r];
                    } else {
                        const uint32_t id = cv.list[base + k];
                        q = cv.pP[id]; sP = cv.P[id]; sN = cv.N[id]; H = cv.adist[id];
                    }
                    if (lc_pdist(pP, q) <= cp.cs) c = lc_weight(sP, sN, H, 0.f, P, N, maxw, false) >= cp.wfill;
                }
                cov = __ballot(c) != 0;
            }
            for (uint32_t k0 = nb_lo; k0 < nb_hi && !cov; k0 += LC_WAVE) {
                const uint32_t k = k0 + lane;
                bool c = false;
                if (k < nb_hi) {
                    f3 q, sP, sN;
                    float H;
                    if (k < LC_NBL) {
                        q = n_pP[k]; sP = n_P[k]; sN = n_N[k]; H = n_H[k];
                    } else {
                        const uint32_t id = cv.nb[base + k];
                        q = cv.pP[id]; sP = cv.P[id]; sN = cv.N[id]; H = cv.adist[id];
                    }
                    if (lc_pdist(pP, q) <= cp.cs) c = lc_weight(sP, sN, H, 0.f, P, N, maxw, false) >= cp.wfill;
                }
                cov = __ballot(c) != 0;
            }
            if (cov) continue;
            // takeSample's first hits (pathlight.cc:305-327): the level-0 rays of
            // this sample's paths, Halton calls nin*S + r of the area's sampler
            const f3 U = v_U[u], V = v_V[u];
            const float Z = v_Z[u];
            float HD = 0.f, Mv = 0.f;
            for (int r0 = 0; r0 < S; r0 += LC_WAVE) {
                const int r = r0 + lane;
                float z = -1.f;
                if (r < S) {
                    const uint32_t k = nin * (uint32_t)S + (uint32_t)r;
                    const float z1 = *(const YG_GLOBAL float*)(R.halton + k);
                    const float z2 = *(const YG_GLOBAL float*)(R.halton + (size_t)R.hstride + k);
                    const f3 ray = hemi_direction(z1, z2, N, U, V);
                    float t;
                    if (first_hit<W>(R.bvh, P, ray, R.min_raydis, t, st, true) >= 0) z = t;
                }
                s_z[lane] = z;
                __syncthreads();
                if (lane == 0) {
                    const int n = (S - r0 < LC_WAVE) ? S - r0 : LC_WAVE;
                    for (int q = 0; q < n; q++) {
                        const float t = s_z[q];
                        if (t < 0.f) continue;
                        if (t > 0) HD = (float)((double)HD + 1.0 / (double)t);
                        if ((t < Mv) || (Mv == 0)) Mv = t;
                    }
                }
                __syncthreads();
            }
            if (lane == 0) {
                const float maxdist = (float)1000000 * cp.worldres * Z;
                const float HNUM = (float)S;
                const float H = (HD > 0) ? HNUM / HD : maxdist;
                if (Mv <= 0) Mv = maxdist;
                cv.pP[item] = pP;
                cv.P[item] = P;
                cv.N[item] = N;
                cv.adist[item] = H;
                cv.M[item] = Mv;
                cv.prec[item] = Z * cp.worldres;
                cv.in[item] = 1u;
                cv.row[item] = pi;
                cv.list[base + nin] = item;
                const uint32_t r = nin % LC_OWN;
                o_pP[r] = pP;
                o_P[r] = P;
                o_N[r] = N;
                o_H[r] = H;
                o_row[r] = pi;
            }
            __syncthreads();
            nin++;
        }
        __syncthreads();
    }
    if (lane == 0) cv.cnt[ai] = nin;
}

// takeSample's estimate of every sample: its paths' contributions (traced
// by the GI levels over the sample items) times the sampler's multiplier;
// also counts the samples whose spread asks for adaptive repetitions.
__global__ __launch_bounds__(RB) void k_cache_color(const RenderScene* __restrict__ scenes,
                                                    const Area* __restrict__ areas, int nareas, uint32_t nitems,
                                                    int S, int D, const c3* __restrict__ contri,
                                                    const uint8_t* __restrict__ nseg, CacheView cv,
                                                    Counters* __restrict__ cnt)
{
    const uint32_t item = blockIdx.x * RB + threadIdx.x;
    uint32_t extra_n = 0;
    if (item < nitems && cv.in[item]) {
        const Area a = areas[find_area(areas, nareas, item)];
        const RenderScene& R = scenes[a.scene];
        c3 total = mkc(0, 0, 0), sub[4];
        for (int q = 0; q < 4; q++) sub[q] = mkc(0, 0, 0);
        for (int smp = 0; smp < S; smp++) {
            const size_t p = (size_t)item * S + smp;
            const int ns = nseg[p];
            for (int b = 0; b < ns; b++) {
                const c3 c = contri[p * D + b];
                total = cadd(total, c);
                sub[smp % 4] = cadd(sub[smp % 4], c);
            }
        }
        const float mult = (float)(1.0 / (double)(float)S);
        float minR = 1000, minG = 1000, minB = 1000, maxR = 0, maxG = 0, maxB = 0;
        for (int q = 0; q < 4; q++) {
            const c3 v = cscale(sub[q], mult);
            maxR = fmaxf(maxR, v.r); maxG = fmaxf(maxG, v.g); maxB = fmaxf(maxB, v.b);
            minR = fminf(minR, v.r); minG = fminf(minG, v.g); minB = fminf(minB, v.b);
        }
        const float ene = fmaxf(fabsf(minR), fmaxf(fabsf(minG), fabsf(minB)));
        const float dmax = fmaxf(fabsf(maxR - minR), fmaxf(fabsf(maxG - minG), fabsf(maxB - minB)));
        float diff = (float)((double)dmax / ((double)((ene < 1.0f) ? 1.0f : ene) * 2.0 * 2.0));
        diff /= R.threshold;
        if ((int)(diff * diff) - 1 > 0) extra_n = 1;
        cv.col[item] = cscale(total, mult / 1.0f);
    }
    block_add(&cnt->extra, extra_n);
}

// std::push_heap / std::pop_heap (libstdc++) with compareFound_f
// (a.weight > b.weight) over the found arrays in LDS: equal weights fall
// where the reference's heap puts them
struct LcFound {
    float w, dis;
    uint32_t id;
};

__device__ __forceinline__ void lc_sift_up(LcFound* f, int hole, int top, LcFound v)
{
    int parent = (hole - 1) / 2;
    while (hole > top && f[parent].w > v.w) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = v;
}

__device__ __forceinline__ void lc_pop_heap(LcFound* f, int len)
{
    if (len <= 1) return;
    const int n = len - 1;
    LcFound v = f[n];
    f[n] = f[0];
    int hole = 0, child = 0;
    while (child < (n - 1) / 2) {
        child = 2 * (child + 1);
        if (f[child].w > f[child - 1].w) child--;
        f[hole] = f[child];
        hole = child;
    }
    if ((n & 1) == 0 && child == (n - 2) / 2) {
        child = 2 * (child + 1);
        f[hole] = f[child - 1];
        hole = child - 1;
    }
    lc_sift_up(f, hole, 0, v);
}

#define LC_CAND 1024   // candidates of one gather ring (LDS)
#define LC_KMAX 1000   // largest "search" the render pass keeps in LDS

// Render pass: one wave per area walks its pixels in order; a qualifying
// pixel's gather (lightCache_t::gatherSamples with weight(), minimun 0)
// starts from the radius the area's previous gather left.  Each ring's
// candidates are collected by the lanes, ordered by their rank (the order
// the generic tree hands them out), and offered to the heap in that order
// by lane 0.  Writes the pathlight's term dif x interpolation x power.
__global__ __launch_bounds__(LC_WAVE) void k_cache_interp(const RenderScene* __restrict__ scenes,
                                                          const Area* __restrict__ areas, int nareas,
                                                          const uint32_t* __restrict__ qual, VisitView vv,
                                                          CacheView cv, CacheParams cp, c3* __restrict__ pathc)
{
    __shared__ int32_t c_rank[LC_CAND];
    __shared__ uint32_t c_id[LC_CAND];
    __shared__ float c_dis[LC_CAND], c_w[LC_CAND], c_ad[LC_CAND];
    __shared__ LcFound found[LC_KMAX + 1];
    __shared__ c3 f_col[LC_KMAX + 1];
    __shared__ int s_nf;
    __shared__ uint32_t s_nc;
    __shared__ int s_repeat;
    __shared__ float s_radius, s_last;
    const int lane = threadIdx.x;
    const int ai = blockIdx.x;
    if (ai >= nareas) return;
    const Area a = areas[ai];
    const RenderScene& R = scenes[a.scene];
    const uint32_t base = a.item_start, n_area = (uint32_t)a.w * (uint32_t)a.h;
    const float maxweight = (float)((double)cp.wlimit * 2.5), minweight = (float)((double)cp.wlimit * 0.6);
    const unsigned K = (unsigned)cp.K;
    // a fresh cacheProxy_t per area: radius = maxradius
    if (lane == 0) s_radius = cp.sradius;
    __syncthreads();
    for (uint32_t li = 0; li < n_area; li++) {
        const uint32_t item = base + li;
        if (!qual[item]) continue;
        const float Z = vv.Z[item];
        const f3 P = vv.P[item], N = vv.N[item];
        const int pi = a.y + (int)(li / (uint32_t)a.w), pj = a.x + (int)(li % (uint32_t)a.w);
        const f3 pP = lc_polar(pi, pj, R.resx, R.resy, cp.ycorr, Z, false);
        const int gx = (int)floorf((pP.x - cp.gx0) / cp.gcell), gy = (int)floorf((pP.y - cp.gy0) / cp.gcell);
        // gatherSamples state (lane 0)
        float best = 0, adist = 0, maxdist = 0;
        unsigned reached = 0;
        int nf = 0;
        if (lane == 0) {
            s_last = 0.f;
            s_repeat = 1;
        }
        __syncthreads();
        while (s_repeat) {
            const float radius = s_radius, lastradius = s_last;
            if (lane == 0) s_nc = 0;
            __syncthreads();
            for (int cy = gy - 1; cy <= gy + 1; cy++) {
                if (cy < 0 || cy >= cp.gny) continue;
                for (int cx = gx - 1; cx <= gx + 1; cx++) {
                    if (cx < 0 || cx >= cp.gnx) continue;
                    const int c = cy * cp.gnx + cx;
                    const int q0 = cp.gstart[c], q1 = cp.gstart[c + 1];
                    for (int q = q0 + lane; q < q1; q += LC_WAVE) {
                        const uint32_t id = cp.gitem[q];
                        const float pD = lc_pdist(pP, cv.pP[id]);
                        if (pD >= radius || pD < lastradius) continue;
                        const uint32_t pos = atomicAdd(&s_nc, 1u);
                        if (pos < LC_CAND) {
                            c_rank[pos] = cv.rank[id];
                            c_id[pos] = id;
                            c_dis[pos] = pD;
                            c_ad[pos] = cv.adist[id];
                            c_w[pos] = lc_weight(cv.P[id], cv.N[id], cv.adist[id], cv.prec[id], P, N, maxweight,
                                                 true);
                        }
                    }
                }
            }
            __syncthreads();
            const uint32_t nc = s_nc;
            if (nc > LC_CAND) {
                if (lane == 0) {
                    atomicAdd(&cv.err[0], 1ull);
                    s_repeat = 0;
                }
                __syncthreads();
                break;
            }
            // order the ring's candidates by rank (bitonic sort, padded)
            uint32_t n2 = 1;
            while (n2 < nc) n2 <<= 1;
            for (uint32_t k = nc + lane; k < n2; k += LC_WAVE) c_rank[k] = 0x7fffffff;
            __syncthreads();
            for (uint32_t size = 2; size <= n2 && nc > 1; size <<= 1)
                for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                    for (uint32_t k = lane; k < n2; k += LC_WAVE) {
                        const uint32_t o = k ^ stride;
                        if (o > k) {
                            const bool up = (k & size) == 0;
                            if ((c_rank[k] > c_rank[o]) == up) {
                                int32_t tr = c_rank[k]; c_rank[k] = c_rank[o]; c_rank[o] = tr;
                                uint32_t ti2 = c_id[k]; c_id[k] = c_id[o]; c_id[o] = ti2;
                                float td = c_dis[k]; c_dis[k] = c_dis[o]; c_dis[o] = td;
                                float tw = c_w[k]; c_w[k] = c_w[o]; c_w[o] = tw;
                                float ta = c_ad[k]; c_ad[k] = c_ad[o]; c_ad[o] = ta;
                            }
                        }
                    }
                    __syncthreads();
                }
            if (lane == 0) {
                for (uint32_t k = 0; k < nc; k++) {
                    reached++;
                    const float w = c_w[k], pD = c_dis[k];
                    const uint32_t id = c_id[k];
                    if (w > best) {
                        best = w;
                        adist = c_ad[k];
                    }
                    unsigned limit;
                    if (w <= cp.wlimit) limit = 0u;
                    else {
                        if (pD > maxdist) maxdist = pD;
                        limit = K;
                    }
                    if (!limit) continue;
                    if (((unsigned)nf >= limit) && (w < found[0].w)) continue;
                    LcFound t;
                    t.w = w;
                    t.dis = pD;
                    t.id = id;
                    found[nf++] = t;
                    lc_sift_up(found, nf - 1, 0, found[nf - 1]);
                    if ((unsigned)nf > limit) {
                        lc_pop_heap(found, nf);
                        nf--;
                    }
                }
                float rrad = (nf == 0 || found[0].dis == 0)
                                 ? (float)(0.0001 * (double)adist)
                                 : radius * len3(cv.P[found[0].id] - P) / found[0].dis;
                if (rrad == 0.0f) rrad = 1.0f;
                const int rep = ((nf == 0) || (((adist / rrad) > cp.wlimit) && (reached < K)) || (best <= minweight)) &&
                                (radius < cp.sradius);
                float nr = radius;
                if (rep) {
                    s_last = radius;
                    nr = radius * 2;
                }
                if (nr > cp.sradius) nr = cp.sradius;
                s_radius = nr;
                s_repeat = rep;
            }
            __syncthreads();
        }
        // the found samples' colours, loaded by all lanes
        if (lane == 0) s_nf = nf;
        __syncthreads();
        for (int k = lane; k < s_nf; k += LC_WAVE) f_col[k] = cv.col[found[k].id];
        __syncthreads();
        if (lane == 0) {
            float radius = s_radius;
            if ((maxdist > 0) && ((double)(maxdist / radius) < (1.0 / sqrt(2.0))))
                radius = (float)((double)radius * 0.9);
            s_radius = radius;
            // interpolate (pathlight.cc:567-583)
            float farest = 0;
            if (nf) farest = found[0].w;
            if (nf == 1) farest = 0;
            else if (farest > cp.wlimit) farest = cp.wlimit;
            c3 total = mkc(0, 0, 0);
            float amount = 0;
            for (int k = 0; k < nf; k++) {
                const float w = (float)((double)(found[k].w - farest) *
                                        (1.0 - (double)(found[k].dis / cp.sradius)));
                total = cadd(total, cscale(f_col[k], w));
                amount += w;
            }
            c3 out = mkc(0, 0, 0);
            if (amount != 0) {
                amount = (float)(1.0 / (double)amount);
                total = cscale(total, amount);
                out = cscale(cmul(vv.dif[item], total), R.path_power);
            } else
                atomicAdd(&cv.err[1], 1ull);
            pathc[item] = out;
        }
        __syncthreads();
    }
}
