// Global photon map on the device: globalPhotonLight_t::init
// (globalphotonlight.cc:234-278) -- photons from every dummy area light,
// diffuse bounces, storeInHash's box merging and computeIrradiances'
// pre-gathering -- leaving the irradiance boxes the renderer reads at path
// vertices (gphoton_dev.h).
//
// The reference shoots photons one after another from the ourRandom()
// stream (4 draws to emit, 3 per diffuse bounce), so where photon k's draws
// start depends on every earlier photon's path.  Here photon k draws from
// its own segment of the same Park-Miller sequence, starting at draw
// k * (4 + 3 * (depth + 1)): the same generator and distributions, other
// numbers; the maps agree statistically.  Everything after the tracing is
// the reference's arithmetic: records keep their (photon, bounce) order
// through a stable sort by box, each box folds its records in that order
// (one thread per box), and every box gathers its "search" nearest box
// photons.  The gathers start at the light's radius on each call (the
// reference carries a radius from call to call, globalPhotonMap_t::gather,
// photon.cc:167-210, which only matters where fewer than "search" photons
// lie within the radius).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "gphoton_dev.h"
#include "photon_dev.h"
#include "qmc.h"

__constant__ uint64_t GP_LCG_POW[64];

namespace {

struct GpEmit {
    f3 corner, toX, toY, direction, NU, NV;
    c3 scolor;
};

// one stored photon: position, face-forwarded normal, quantised direction
// (theta | phi << 8), RGBE colour, hash box
struct GRec {
    f3 pos, N;
    uint32_t code;
    uint32_t rgbe;
    int bx, by, bz;
};

// dirConverter_t tables (photon.cc:11-25): costheta[255] sintheta[255]
// cosphi[256] sinphi[256], float angles, cosine and sine in double
static void dirconv_tables(float* t)
{
    for (int i = 0; i < 255; ++i) {
        const float angle = (float)((double)(float)i * (1.0 / 255.0) * M_PI);
        t[i] = (float)cos((double)angle);
        t[255 + i] = (float)sin((double)angle);
    }
    for (int i = 0; i < 256; ++i) {
        const float angle = (float)((double)(float)i * (1.0 / 256.0) * 2.0 * M_PI);
        t[510 + i] = (float)cos((double)angle);
        t[766 + i] = (float)sin((double)angle);
    }
}

// dirConverter_t::convert (photon.h:24-34); 255: no direction
__device__ __forceinline__ uint32_t dir_enc(f3 d)
{
    if (d.x == 0 && d.y == 0 && d.z == 0) return 255u;
    int t = (int)(acos((double)d.z) * (255.0 / M_PI));
    int p = (int)(atan2((double)d.y, (double)d.x) * (256.0 / (2.0 * M_PI)));
    if (t > 254) t = 254;
    else if (t < 0) t = 0;
    if (p > 255) p = 255;
    else if (p < 0) p += 256;
    return (uint32_t)t | ((uint32_t)p << 8);
}

__device__ __forceinline__ f3 dir_dec(uint32_t code, const float* __restrict__ tab)
{
    const uint32_t th = code & 255u, ph = (code >> 8) & 255u;
    if (th == 255u) return mk3(0, 0, 0);
    return mk3(tab[255 + th] * tab[510 + ph], tab[255 + th] * tab[766 + ph], tab[th]);
}

// rgbe_t (color.cc:140-155, color.h:159-170)
__device__ __forceinline__ uint32_t rgbe_enc(c3 c)
{
    float v = c.r;
    if (c.g > v) v = c.g;
    if (c.b > v) v = c.b;
    if (v < 1e-32f) return 0u;
    int e;
    v = (float)((double)frexpf(v, &e) * 256.0 / (double)v);
    const uint32_t r = (uint32_t)(unsigned char)(c.r * v), g = (uint32_t)(unsigned char)(c.g * v),
                   b = (uint32_t)(unsigned char)(c.b * v);
    return r | (g << 8) | (b << 16) | ((uint32_t)(unsigned char)(e + 128) << 24);
}

__device__ __forceinline__ c3 rgbe_dec(uint32_t q)
{
    const int e = (int)(q >> 24);
    if (!e) return mkc(0, 0, 0);
    const float f = (float)ldexp(1.0, e - (128 + 8));
    return mkc((float)(q & 255u) * f, (float)((q >> 8) & 255u) * f, (float)((q >> 16) & 255u) * f);
}

__device__ __forceinline__ float energy(c3 c) { return (c.r + c.g + c.b) * 0.333333f; }

// HemiVec_CONE (globalphotonlight.cc:10-16, arealight.cc:219-225)
__device__ __forceinline__ f3 hemivec_cone(f3 nrm, f3 Ru, f3 Rv, float cosang, float z1, float z2)
{
    const float t1 = (float)(2.0 * M_PI * (double)z1);
    const float t2 = (float)(1.0 - (1.0 - (double)cosang) * (double)z2);
    const f3 a = Ru * dcos(t1) + Rv * dsin(t1);
    return a * (float)sqrt(1.0 - (double)(t2 * t2)) + nrm * t2;
}

// findBox: the box of the box centre (hash3d.h:41-42)
__device__ __forceinline__ void gp_find_box(float cell, f3 p, int& bx, int& by, int& bz)
{
    int nx, ny, nz;
    gp_hash_cell(cell, p, nx, ny, nz);
    const f3 c = mk3((float)((double)(cell * (float)nx) + 0.5 * (double)cell),
                     (float)((double)(cell * (float)ny) + 0.5 * (double)cell),
                     (float)((double)(cell * (float)nz) + 0.5 * (double)cell));
    gp_hash_cell(cell, c, bx, by, bz);
}

#define GB 64
// One photon per thread: quadEmitter_t::getDirection (arealight.cc:227-232)
// then globalPhotonLight_t::shoot (globalphotonlight.cc:19-95) without
// caustic objects; record slot k * (depth + 1) + bounce.
template <int W>
__global__ __launch_bounds__(GB) void k_gp_trace(const GpEmit* __restrict__ em, int ppl, uint32_t nph, int maxdepth,
                                                  uint32_t seed0, uint32_t budget, BVHView bvh,
                                                  const yg_material* __restrict__ mats,
                                                  const yg_modulator* __restrict__ mods,
                                                  const yg_texture* __restrict__ texs,
                                                  const int32_t* __restrict__ obj_flags, float min_raydis,
                                                  float cell, GRec* __restrict__ rec, uint32_t* __restrict__ valid,
                                                  unsigned long long* __restrict__ nrays)
{
    extern __shared__ int s_stk[];
    const uint32_t k = blockIdx.x * GB + threadIdx.x;
    if (k >= nph) return;
    LaneStack st;
    st.init(s_stk, threadIdx.x, GB);
    const GpEmit& E = em[k / (uint32_t)ppl];
    int32_t seed = lcg_jump(seed0, (uint64_t)k * budget, GP_LCG_POW);
    // argument and operand draws in GCC's order (right to left)
    const float z2 = our_random(seed), z1 = our_random(seed);
    f3 dir = hemivec_cone(E.direction, E.NU, E.NV, 0.0001f, z1, z2);
    const float ry = our_random(seed), rx = our_random(seed);
    f3 pos = (E.corner + E.toX * rx) + E.toY * ry;
    f3 lastpos = pos;
    c3 col = cscale(E.scolor, dot(E.direction, dir));
    unsigned long long rays = 0;
    for (int depth = 0; depth <= maxdepth; depth++) {
        rays++;
        float t;
        const int ti = first_hit<W>(bvh, pos, dir, min_raydis, t, st);
        if (ti < 0) break;
        const f3 hp = (pos + dir * min_raydis) + dir * t;
        // runningPhoton_t::position(p, MIN_RAYDIST) (photon.cc:27-35)
        if (len3(pos - hp) > min_raydis) lastpos = pos;
        pos = hp;
        const f3 edir = norm3(lastpos - pos);
        const float4 A = bvh.tris[3 * ti], B = bvh.tris[3 * ti + 1], C = bvh.tris[3 * ti + 2];
        const f3 Ng = mk3(A.w, B.w, C.w);
        const f3 Nsh = surf_normal(bvh, ti, pos, Ng);
        const f3 N = face_forward(Ng, Nsh, edir);
        const f3 Ngf = face_forward(Ng, Ng, edir);
        const int fl = obj_flags[bvh.tobj[ti]];
        if (fl & 2) {  // storeFirst (arealight.h:43) and reciveRadiosity
            GRec r;
            r.pos = pos;
            r.N = N;
            r.code = dir_enc(norm3(lastpos - pos));
            r.rgbe = rgbe_enc(col);
            gp_find_box(cell, pos, r.bx, r.by, r.bz);
            const size_t slot = (size_t)k * (maxdepth + 1) + depth;
            rec[slot] = r;
            valid[slot] = 1;
        }
        if (!(fl & 4)) break;
        // getDiffuse(nullstate, sp, edir) (globalphotonlight.cc:54,93)
        Surf hs;
        hs.P = pos;
        hs.N = Nsh;
        hs.Ng = Ng;
        hs.ti = ti;
        hs.obj = bvh.tobj[ti];
        hs.mat = bvh.tmat[ti];
        c3 dif = get_diffuse(bvh, mats, mods, texs, hs, edir);
        float diffuse = energy(dif);
        float sum = diffuse;
        if (sum > 0.0f) sum = (float)(1.0 / (double)sum);
        diffuse *= sum;
        if (!(sum > 0.0f)) break;
        (void)our_random(seed);  // ourRandom() < trans (= 0): never taken
        dif = cscale(dif, (float)(1.0 / (double)diffuse));
        const float r1 = our_random(seed), r2 = our_random(seed);
        f3 NU, NV;
        create_cs(Nsh, NU, NV);
        dir = hemivec_cone(Ngf, NU, NV, 0.05f, r1, r2);
        col = cmul(col, dif);
    }
    atomicAdd(nrays, rays);
}

__device__ __forceinline__ int gp_wave_min(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int gp_wave_max(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// valid records to their (photon, bounce)-ordered positions; box bounds per wave
__global__ void k_gp_collect(uint32_t nslots, const uint32_t* __restrict__ valid, const uint32_t* __restrict__ vpos,
                             const GRec* __restrict__ slots, GRec* __restrict__ out, int* __restrict__ bounds)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    if (s < nslots && valid[s]) {
        const GRec r = slots[s];
        out[vpos[s]] = r;
        const int b[3] = {r.bx, r.by, r.bz};
        for (int a = 0; a < 3; a++) { lo[a] = b[a]; hi[a] = b[a]; }
    }
    for (int a = 0; a < 3; a++) {  // every lane reaches the reductions
        lo[a] = gp_wave_min(lo[a]);
        hi[a] = gp_wave_max(hi[a]);
    }
    if ((threadIdx.x & 63) == 0 && lo[0] != INT32_MAX)
        for (int a = 0; a < 3; a++) {
            atomicMin(&bounds[2 * a], lo[a]);
            atomicMax(&bounds[2 * a + 1], hi[a]);
        }
}

// the raw photon map: records in store order decoded for the gathers of
// photonSampler_t, and as 5-word records (position bits, direction code, RGBE)
__global__ void k_gp_raw(uint32_t n, const GRec* __restrict__ rec, const float* __restrict__ tab, f3* __restrict__ rpos,
                         f3* __restrict__ rdir, c3* __restrict__ rcol, uint32_t* __restrict__ raw)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GRec r = rec[i];
    rpos[i] = r.pos;
    rdir[i] = dir_dec(r.code, tab);
    rcol[i] = rgbe_dec(r.rgbe);
    raw[5 * (size_t)i + 0] = __float_as_uint(r.pos.x);
    raw[5 * (size_t)i + 1] = __float_as_uint(r.pos.y);
    raw[5 * (size_t)i + 2] = __float_as_uint(r.pos.z);
    raw[5 * (size_t)i + 3] = r.code;
    raw[5 * (size_t)i + 4] = r.rgbe;
}

__global__ void k_gp_keys(uint32_t n, const GRec* __restrict__ rec, int bx0, int by0, int bz0, int sy, int sz,
                          unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int bmin[3] = {bx0, by0, bz0};
    keys[i] = gp_box_key(rec[i].bx, rec[i].by, rec[i].bz, bmin, sy, sz);
    vals[i] = i;
}

__global__ void k_gp_heads(uint32_t n, const unsigned long long* __restrict__ keys, uint32_t* __restrict__ head)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_gp_starts(uint32_t n, const uint32_t* __restrict__ head, const uint32_t* __restrict__ hpos,
                            uint32_t* __restrict__ starts)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    starts[hpos[i]] = i;
}

// storeInHash (globalphotonlight.cc:177-195): one box's records in store order
__global__ void k_gp_fold(uint32_t nbox, uint32_t n, const uint32_t* __restrict__ starts,
                          const uint32_t* __restrict__ order, const unsigned long long* __restrict__ skeys,
                          const GRec* __restrict__ rec, const float* __restrict__ tab, f3* __restrict__ bpos,
                          uint32_t* __restrict__ bcode, uint32_t* __restrict__ brgbe, f3* __restrict__ bN,
                          unsigned long long* __restrict__ bkey)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const uint32_t a = starts[b], e = (b + 1 < nbox) ? starts[b + 1] : n;
    uint32_t code = 255u, q = 0u;
    f3 P = mk3(0, 0, 0), N = mk3(0, 0, 0);
    for (uint32_t k = a; k < e; k++) {
        const GRec& r = rec[order[k]];
        if ((code & 255u) == 255u) {
            P = r.pos;
            code = r.code;
            q = r.rgbe;
            N = r.N;
        } else {
            const f3 nd = dir_dec(r.code, tab);
            if (dot(nd, N) > 0.0f) {
                const c3 ac = rgbe_dec(q), nc = rgbe_dec(r.rgbe);
                f3 mixedd = dir_dec(code, tab) * energy(ac) + nd * energy(nc);
                mixedd = norm3(mixedd);
                code = dir_enc(mixedd);
                q = rgbe_enc(cadd(ac, nc));
            }
        }
    }
    bpos[b] = P;
    bcode[b] = code;
    brgbe[b] = q;
    bN[b] = N;
    bkey[b] = skeys[a];
}

// setIrradiance (globalphotonlight.cc:127-175) of every box: the K nearest
// box photons with a direction, within the radius, facing N (dot > 0);
// (dis, box) max-heaps in global memory, [entry][box]
__global__ void k_gp_irr(uint32_t nbox, int K, float radius, const f3* __restrict__ bpos,
                         const uint32_t* __restrict__ bcode, const uint32_t* __restrict__ brgbe,
                         const f3* __restrict__ bN, const int32_t* __restrict__ gstart,
                         const int32_t* __restrict__ gitems, float gx, float gy, float gz, int dx, int dy, int dz,
                         const float* __restrict__ tab, float* __restrict__ hd, int* __restrict__ hi,
                         c3* __restrict__ irr)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const size_t S = nbox;
    const f3 P = bpos[b], N = bN[b];
    const float gmin[3] = {gx, gy, gz};
    const int gdim[3] = {dx, dy, dz};
    const float pp[3] = {P.x, P.y, P.z};
    int g0[3], g1[3];
    for (int a = 0; a < 3; a++) {
        long lo = (long)floorf((pp[a] - radius - gmin[a]) / radius);
        long hh = (long)floorf((pp[a] + radius - gmin[a]) / radius);
        if (lo < 0) lo = 0;
        if (hh >= gdim[a]) hh = gdim[a] - 1;
        g0[a] = (int)lo;
        g1[a] = (int)hh;
    }
    int n = 0;
    for (int z = g0[2]; z <= g1[2]; z++)
        for (int y = g0[1]; y <= g1[1]; y++)
            for (int x = g0[0]; x <= g1[0]; x++) {
                const int c = (z * gdim[1] + y) * gdim[0] + x;
                for (int q = gstart[c]; q < gstart[c + 1]; q++) {
                    const int idx = gitems[q];
                    if ((bcode[idx] & 255u) == 255u) continue;
                    const float dis = len3(bpos[idx] - P);
                    if ((dis > radius) || (dot(dir_dec(bcode[idx], tab), N) <= 0.0f)) continue;
                    if (n < K) {
                        int pos = n++;
                        while (pos > 0) {
                            const int par = (pos - 1) >> 1;
                            if (!knn_less(hd[par * S + b], hi[par * S + b], dis, idx)) break;
                            hd[pos * S + b] = hd[par * S + b];
                            hi[pos * S + b] = hi[par * S + b];
                            pos = par;
                        }
                        hd[pos * S + b] = dis;
                        hi[pos * S + b] = idx;
                    } else if (knn_less(dis, idx, hd[b], hi[b])) {
                        int pos = 0;
                        for (;;) {
                            const int l = 2 * pos + 1, r = l + 1;
                            int big = -1;
                            float bd = dis;
                            int bi = idx;
                            if (l < n && knn_less(bd, bi, hd[l * S + b], hi[l * S + b])) {
                                big = l; bd = hd[l * S + b]; bi = hi[l * S + b];
                            }
                            if (r < n && knn_less(bd, bi, hd[r * S + b], hi[r * S + b])) big = r;
                            if (big < 0) break;
                            hd[pos * S + b] = hd[big * S + b];
                            hi[pos * S + b] = hi[big * S + b];
                            pos = big;
                        }
                        hd[pos * S + b] = dis;
                        hi[pos * S + b] = idx;
                    }
                }
            }
    if (n == 0) {
        irr[b] = mkc(0, 0, 0);
        return;
    }
    // heap -> ascending order
    for (int end = n - 1; end > 0; end--) {
        const float td = hd[end * S + b];
        const int tix = hi[end * S + b];
        hd[end * S + b] = hd[b];
        hi[end * S + b] = hi[b];
        int pos = 0;
        for (;;) {
            const int l = 2 * pos + 1, r = l + 1;
            int big = -1;
            float bd = td;
            int bi = tix;
            if (l < end && knn_less(bd, bi, hd[l * S + b], hi[l * S + b])) {
                big = l; bd = hd[l * S + b]; bi = hi[l * S + b];
            }
            if (r < end && knn_less(bd, bi, hd[r * S + b], hi[r * S + b])) big = r;
            if (big < 0) break;
            hd[pos * S + b] = hd[big * S + b];
            hi[pos * S + b] = hi[big * S + b];
            pos = big;
        }
        hd[pos * S + b] = td;
        hi[pos * S + b] = tix;
    }
    const float farest = hd[(n - 1) * S + b];
    if (n == 1 || farest == 0.0f) {
        const int f0 = hi[(n - 1) * S + b];
        float factor = dot(dir_dec(bcode[f0], tab), N);
        if (factor < 0.0f) factor = 0.0f;
        irr[b] = cscale(rgbe_dec(brgbe[f0]), factor);
        return;
    }
    c3 total = mkc(0, 0, 0);
    for (int k = 0; k < n; k++) {
        const int idx = hi[k * S + b];
        const float factor = (1.0f - hd[k * S + b] / farest) * dot(dir_dec(bcode[idx], tab), N);
        if (factor > 0) total = cadd(total, cscale(rgbe_dec(brgbe[idx]), factor));
    }
    float area = (float)(M_PI * (double)farest * (double)farest);
    if (area < 0.00005f) area = 0.00005f;
    irr[b] = cscale(total, (float)(4.0 * M_PI / (double)area));
}

// The same pre-gathering with one wave per box: the lanes scan the box's
// grid cells and collect the box photons within the radius that face N into
// LDS; whenever the buffer fills, it is sorted by (dis, box) and cut to the
// K nearest; the final K are summed by lane 0 in that order.  (dis, box) is
// a total order, so the K nearest and their order are the heap's.
#define GP_IRR_BUF 2048
__global__ __launch_bounds__(64) void k_gp_irr_wave(uint32_t nbox, int K, float radius, const f3* __restrict__ bpos,
                                                    const uint32_t* __restrict__ bcode,
                                                    const uint32_t* __restrict__ brgbe, const f3* __restrict__ bN,
                                                    const int32_t* __restrict__ gstart,
                                                    const int32_t* __restrict__ gitems, float gx, float gy, float gz,
                                                    int dx, int dy, int dz, const float* __restrict__ tab,
                                                    c3* __restrict__ irr)
{
    __shared__ float s_d[GP_IRR_BUF];
    __shared__ int s_i[GP_IRR_BUF];
    const uint32_t b = blockIdx.x;
    const int lane = threadIdx.x;
    if (b >= nbox) return;
    const f3 P = bpos[b], N = bN[b];
    const float gmin[3] = {gx, gy, gz};
    const int gdim[3] = {dx, dy, dz};
    const float pp[3] = {P.x, P.y, P.z};
    int g0[3], g1[3];
    for (int a = 0; a < 3; a++) {
        long lo = (long)floorf((pp[a] - radius - gmin[a]) / radius);
        long hh = (long)floorf((pp[a] + radius - gmin[a]) / radius);
        if (lo < 0) lo = 0;
        if (hh >= gdim[a]) hh = gdim[a] - 1;
        g0[a] = (int)lo;
        g1[a] = (int)hh;
    }
    // sort s_d/s_i[0..n) by (dis, box), padding to a power of two
    auto sort_buf = [&](int n) {
        int n2 = 1;
        while (n2 < n) n2 <<= 1;
        for (int k = n + lane; k < n2; k += 64) {
            s_d[k] = INFINITY;
            s_i[k] = 0x7fffffff;
        }
        __syncthreads();
        for (int size = 2; size <= n2; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int k = lane; k < n2; k += 64) {
                    const int o = k ^ stride;
                    if (o > k) {
                        const bool up = (k & size) == 0;
                        const bool gt = knn_less(s_d[o], s_i[o], s_d[k], s_i[k]);
                        if (gt == up) {
                            const float td = s_d[k]; s_d[k] = s_d[o]; s_d[o] = td;
                            const int ti = s_i[k]; s_i[k] = s_i[o]; s_i[o] = ti;
                        }
                    }
                }
                __syncthreads();
            }
    };
    int n = 0;
    for (int z = g0[2]; z <= g1[2]; z++)
        for (int y = g0[1]; y <= g1[1]; y++)
            for (int x = g0[0]; x <= g1[0]; x++) {
                const int c = (z * gdim[1] + y) * gdim[0] + x;
                const int q0 = gstart[c], q1 = gstart[c + 1];
                for (int q00 = q0; q00 < q1; q00 += 64) {
                    const int q = q00 + lane;
                    bool keep = false;
                    float dis = 0.f;
                    int idx = 0;
                    if (q < q1) {
                        idx = gitems[q];
                        const uint32_t code = bcode[idx];
                        if ((code & 255u) != 255u) {
                            dis = len3(bpos[idx] - P);
                            keep = !(dis > radius) && dot(dir_dec(code, tab), N) > 0.0f;
                        }
                    }
                    const uint64_t m = __ballot(keep);
                    if (!m) continue;
                    if (n + 64 > GP_IRR_BUF) {
                        sort_buf(n);
                        if (n > K) n = K;
                    }
                    if (keep) {
                        const int pos = n + (int)__popcll(m & ((1ull << lane) - 1ull));
                        s_d[pos] = dis;
                        s_i[pos] = idx;
                    }
                    n += (int)__popcll(m);
                    __syncthreads();
                }
            }
    if (n == 0) {
        if (lane == 0) irr[b] = mkc(0, 0, 0);
        return;
    }
    sort_buf(n);
    if (n > K) n = K;
    if (lane != 0) return;
    const float farest = s_d[n - 1];
    if (n == 1 || farest == 0.0f) {
        const int f0 = s_i[n - 1];
        float factor = dot(dir_dec(bcode[f0], tab), N);
        if (factor < 0.0f) factor = 0.0f;
        irr[b] = cscale(rgbe_dec(brgbe[f0]), factor);
        return;
    }
    c3 total = mkc(0, 0, 0);
    for (int k = 0; k < n; k++) {
        const int idx = s_i[k];
        const float factor = (1.0f - s_d[k] / farest) * dot(dir_dec(bcode[idx], tab), N);
        if (factor > 0) total = cadd(total, cscale(rgbe_dec(brgbe[idx]), factor));
    }
    float area = (float)(M_PI * (double)farest * (double)farest);
    if (area < 0.00005f) area = 0.00005f;
    irr[b] = cscale(total, (float)(4.0 * M_PI / (double)area));
}

// irradiance photons (direction N quantised, colour irr through RGBE) and
// the 15-float read-out records
__global__ void k_gp_finish(uint32_t nbox, const f3* __restrict__ bpos, const uint32_t* __restrict__ bcode,
                            const uint32_t* __restrict__ brgbe, const f3* __restrict__ bN, const c3* __restrict__ irr,
                            const float* __restrict__ tab, f3* __restrict__ qN, c3* __restrict__ qirr,
                            float* __restrict__ out)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    qN[b] = dir_dec(dir_enc(bN[b]), tab);
    qirr[b] = rgbe_dec(rgbe_enc(irr[b]));
    const f3 d = dir_dec(bcode[b], tab);
    const c3 c = rgbe_dec(brgbe[b]);
    float* o = out + 15 * (size_t)b;
    o[0] = bpos[b].x; o[1] = bpos[b].y; o[2] = bpos[b].z;
    o[3] = bN[b].x; o[4] = bN[b].y; o[5] = bN[b].z;
    o[6] = irr[b].r; o[7] = irr[b].g; o[8] = irr[b].b;
    o[9] = d.x; o[10] = d.y; o[11] = d.z;
    o[12] = c.r; o[13] = c.g; o[14] = c.b;
}

struct GFreer {
    std::vector<void*> ptrs;
    ~GFreer()
    {
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
    template <class T>
    int alloc(T** p, size_t n)
    {
        hipError_t e = hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1));
        if (e != hipSuccess) return yg_fail(YG_ERR_NOMEM, std::string("global photon buffer: ") + hipGetErrorString(e));
        ptrs.push_back(*p);
        return YG_OK;
    }
};

template <class T>
int dmalloc_keep(T** p, size_t n)
{
    hipError_t e = hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1));
    if (e != hipSuccess) return yg_fail(YG_ERR_NOMEM, std::string("global photon map: ") + hipGetErrorString(e));
    return YG_OK;
}

int gp_scan(GFreer& F, const uint32_t* in, uint32_t* out, size_t n, hipStream_t st)
{
    size_t bytes = 0;
    YG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, st));
    uint8_t* tmp = nullptr;
    int rc = F.alloc(&tmp, bytes);
    if (rc) return rc;
    YG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, st));
    return YG_OK;
}

uint32_t gp_last_sum(const uint32_t* flag, const uint32_t* ex, size_t n, hipStream_t st)
{
    uint32_t a = 0, b = 0;
    (void)hipMemcpyAsync(&a, flag + n - 1, 4, hipMemcpyDeviceToHost, st);
    (void)hipMemcpyAsync(&b, ex + n - 1, 4, hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    return a + b;
}

// uniform grid (cell edge `cell`, at most 256 cells per axis; points beyond
// the last cell stay in it) over points, built on the host
static void host_grid(const std::vector<f3>& pts, float cell, float* gmin, int* gdim, std::vector<int32_t>& gs,
                      std::vector<int32_t>& gi)
{
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const f3& p : pts) {
        const float v[3] = {p.x, p.y, p.z};
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], v[a]);
            hi[a] = std::max(hi[a], v[a]);
        }
    }
    int64_t ncell = 1;
    for (int a = 0; a < 3; a++) {
        gmin[a] = pts.empty() ? 0.f : lo[a];
        int dd = pts.empty() ? 1 : (int)floorf((hi[a] - lo[a]) / cell) + 1;
        if (dd < 1) dd = 1;
        if (dd > 256) dd = 256;
        gdim[a] = dd;
        ncell *= dd;
    }
    gs.assign((size_t)ncell + 1, 0);
    gi.assign(pts.size(), 0);
    std::vector<int32_t> cof(pts.size());
    for (size_t i = 0; i < pts.size(); i++) {
        const float v[3] = {pts[i].x, pts[i].y, pts[i].z};
        int c[3];
        for (int a = 0; a < 3; a++) {
            int kk = (int)floorf((v[a] - gmin[a]) / cell);
            if (kk < 0) kk = 0;
            if (kk >= gdim[a]) kk = gdim[a] - 1;
            c[a] = kk;
        }
        cof[i] = (c[2] * gdim[1] + c[1]) * gdim[0] + c[0];
        gs[(size_t)cof[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) gs[c + 1] += gs[c];
    std::vector<int32_t> fill(gs.begin(), gs.end() - 1);
    for (size_t i = 0; i < pts.size(); i++) gi[fill[cof[i]]++] = (int32_t)i;
}

}  // namespace

void gphoton_free(GPhotonDev& g)
{
    void* ptrs[] = {g.d_key,  g.d_N,    g.d_irr,  g.d_pos,     g.d_qN,     g.d_qirr,   g.d_gstart,
                    g.d_gitems, g.d_read, g.d_rpos, g.d_rdir, g.d_rcol, g.d_rraw, g.d_rgstart, g.d_rgitems};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    g = GPhotonDev();
}

extern "C" int yg_gphoton_build(yg_scene* s, uint32_t li, const yg_render_opts* opts, yg_photon_stats* stats,
                                void* stream)
{
    if (!s || !opts) return yg_fail(YG_ERR_ARG, "null argument");
    if (!s->built) return yg_fail(YG_ERR_STATE, "scene not built");
    if (li >= s->lights.size() || s->lights[li].type != YG_LIGHT_GLOBALPHOTON)
        return yg_fail(YG_ERR_ARG, "light index is not a global photon light");
    if (s->mats.empty()) return yg_fail(YG_ERR_STATE, "no materials (yg_set_shading)");
    const yg_light& GL = s->lights[li];
    if (GL.depth < 0 || GL.depth > 50) return yg_fail(YG_ERR_ARG, "global photon depth must be 0..50");
    if (GL.search < 1 || GL.search > 1024) return yg_fail(YG_ERR_UNSUPPORTED, "global photon search must be 1..1024");
    if (!(GL.fixedradius > 0)) return yg_fail(YG_ERR_ARG, "global photon radius must be > 0");
    if (GL.photons < 1) return yg_fail(YG_ERR_ARG, "global photon count must be >= 1");
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = perm_upload())) return rc;
    auto t0 = std::chrono::steady_clock::now();
    {
        uint64_t pw[64];
        lcg_pow_table(pw);
        YG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(GP_LCG_POW), pw, sizeof(pw)));
    }
    // emitters: the dummy area lights, in light order (globalphotonlight.cc:241-257)
    std::vector<GpEmit> em;
    for (const auto& L : s->lights)
        if (L.type == YG_LIGHT_AREA && L.dummy) em.emplace_back();
    GPhotonDev& G = s->gp;
    gphoton_free(G);
    // YG_RENDER_KTIMING: events around the tracing and pre-gathering launches
    const bool ktime = (opts->flags & YG_RENDER_KTIMING) != 0;
    struct Events {
        hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } evs;
    hipEvent_t* ev = evs.e;
    if (ktime)
        for (int i = 0; i < 4; i++) YG_HIP(hipEventCreate(&ev[i]));
    bool irr_timed = false;
    uint64_t nph = 0, nrec = 0, nbox = 0;
    unsigned long long nrays = 0;
    GFreer F;
    float* tab;
    if ((rc = F.alloc(&tab, 1022))) return rc;
    {
        float t[1022];
        dirconv_tables(t);
        YG_HIP(hipMemcpyAsync(tab, t, sizeof(t), hipMemcpyHostToDevice, st));
    }
    G.cell = GL.fixedradius / sqrtf((float)GL.search);
    G.radius = GL.fixedradius;
    if (!em.empty()) {
        const int ppl = GL.photons / (int)em.size();
        size_t k = 0;
        for (const auto& L : s->lights) {
            if (L.type != YG_LIGHT_AREA || !L.dummy) continue;
            // areaLight_t::getEmitter, quadEmitter_t ctor and numSamples (arealight.h:90-95,
            // arealight.cc:200-217)
            const f3 a = ld3(L.a), b = ld3(L.b), d = ld3(L.d);
            GpEmit& E = em[k++];
            E.corner = a;
            E.toX = b - a;
            E.toY = d - a;
            E.direction = norm3(cross(E.toX, E.toY));
            E.NU = norm3(E.toX);
            E.NV = norm3(E.toY);
            const float area = len3(cross(E.toX, E.toY));
            const c3 c = cscale(cscale(cscale(mkc(L.color[0], L.color[1], L.color[2]), L.power), area), 0.5f);
            E.scolor = mkc(c.r / (float)ppl, c.g / (float)ppl, c.b / (float)ppl);
        }
        nph = (uint64_t)ppl * em.size();
        const int slots_per = GL.depth + 1;
        const uint64_t nslots64 = nph * (uint64_t)slots_per;
        if (nslots64 >= (1ull << 31)) return yg_fail(YG_ERR_UNSUPPORTED, "too many global photon records");
        const uint32_t nslots = (uint32_t)nslots64;
        const uint32_t budget = 4u + 3u * (uint32_t)slots_per;
        GpEmit* d_em;
        GRec *slots, *recs;
        uint32_t *valid, *vpos;
        unsigned long long* d_rays;
        int* bounds;
        if ((rc = F.alloc(&d_em, em.size())) || (rc = F.alloc(&slots, nslots)) || (rc = F.alloc(&valid, nslots)) ||
            (rc = F.alloc(&vpos, nslots)) || (rc = F.alloc(&d_rays, 1)) || (rc = F.alloc(&bounds, 6)))
            return rc;
        YG_HIP(hipMemcpyAsync(d_em, em.data(), sizeof(GpEmit) * em.size(), hipMemcpyHostToDevice, st));
        YG_HIP(hipMemsetAsync(valid, 0, sizeof(uint32_t) * nslots, st));
        YG_HIP(hipMemsetAsync(d_rays, 0, sizeof(unsigned long long), st));
        if (ktime) YG_HIP(hipEventRecord(ev[0], st));
        YG_LAUNCH_W(s->width, (k_gp_trace<8>), (k_gp_trace<4>), (k_gp_trace<2>), dim3(grid_for(nph, GB)), dim3(GB),
                    stack_lds_bytes(s->stack_entries, GB), st, d_em, ppl, (uint32_t)nph, GL.depth, opts->lcg_seed,
                    budget, s->view(), s->d_mats, s->d_mods, s->d_texs, s->d_obj_flags, opts->min_raydis, G.cell,
                    slots, valid, d_rays);
        YG_HIP(hipGetLastError());
        if (ktime) YG_HIP(hipEventRecord(ev[1], st));
        YG_HIP(hipMemcpyAsync(&nrays, d_rays, sizeof(nrays), hipMemcpyDeviceToHost, st));
        if ((rc = gp_scan(F, valid, vpos, nslots, st))) return rc;
        nrec = gp_last_sum(valid, vpos, nslots, st);
        if ((rc = F.alloc(&recs, nrec))) return rc;
        const int binit[6] = {INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
        YG_HIP(hipMemcpyAsync(bounds, binit, sizeof(binit), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_gp_collect, dim3(grid_for(nslots, 256)), dim3(256), 0, st, nslots, valid, vpos, slots,
                           recs, bounds);
        YG_HIP(hipGetLastError());
        int hb[6];
        YG_HIP(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, st));
        YG_HIP(hipStreamSynchronize(st));
        if (nrec > 0) {
            // the raw map for photonSampler_t
            if ((rc = dmalloc_keep(&G.d_rpos, nrec)) || (rc = dmalloc_keep(&G.d_rdir, nrec)) ||
                (rc = dmalloc_keep(&G.d_rcol, nrec)) || (rc = dmalloc_keep(&G.d_rraw, 5 * nrec)))
                return rc;
            hipLaunchKernelGGL(k_gp_raw, dim3(grid_for(nrec, 256)), dim3(256), 0, st, (uint32_t)nrec, recs, tab,
                               G.d_rpos, G.d_rdir, G.d_rcol, G.d_rraw);
            YG_HIP(hipGetLastError());
            std::vector<f3> rp(nrec);
            YG_HIP(hipMemcpyAsync(rp.data(), G.d_rpos, sizeof(f3) * nrec, hipMemcpyDeviceToHost, st));
            YG_HIP(hipStreamSynchronize(st));
            std::vector<int32_t> rgs, rgi;
            host_grid(rp, G.radius, G.rgmin, G.rgdim, rgs, rgi);
            if ((rc = dev_upload(&G.d_rgstart, rgs.data(), rgs.size())) ||
                (rc = dev_upload(&G.d_rgitems, rgi.data(), rgi.size())))
                return rc;
            G.nrec = (int64_t)nrec;
            auto bits = [](long span) {
                int b = 0;
                while (b < 62 && (1L << b) <= span) b++;
                return b;
            };
            const int sx = bits((long)hb[1] - hb[0]), sy = bits((long)hb[3] - hb[2]), sz = bits((long)hb[5] - hb[4]);
            if (sx + sy + sz > 64) return yg_fail(YG_ERR_UNSUPPORTED, "global photon boxes span too far");
            for (int a = 0; a < 3; a++) {
                G.bmin[a] = hb[2 * a];
                G.bmax[a] = hb[2 * a + 1];
            }
            G.sy = sy;
            G.sz = sz;
            unsigned long long *keys, *skeys;
            uint32_t *vals, *order, *head, *hpos, *starts;
            if ((rc = F.alloc(&keys, nrec)) || (rc = F.alloc(&skeys, nrec)) || (rc = F.alloc(&vals, nrec)) ||
                (rc = F.alloc(&order, nrec)) || (rc = F.alloc(&head, nrec)) || (rc = F.alloc(&hpos, nrec)))
                return rc;
            hipLaunchKernelGGL(k_gp_keys, dim3(grid_for(nrec, 256)), dim3(256), 0, st, (uint32_t)nrec, recs, hb[0],
                               hb[2], hb[4], sy, sz, keys, vals);
            YG_HIP(hipGetLastError());
            size_t bytes = 0;
            YG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys, skeys, vals, order, (int)nrec, 0,
                                                      std::max(sx + sy + sz, 1), st));
            uint8_t* tmp;
            if ((rc = F.alloc(&tmp, bytes))) return rc;
            YG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys, skeys, vals, order, (int)nrec, 0,
                                                      std::max(sx + sy + sz, 1), st));
            hipLaunchKernelGGL(k_gp_heads, dim3(grid_for(nrec, 256)), dim3(256), 0, st, (uint32_t)nrec, skeys, head);
            YG_HIP(hipGetLastError());
            if ((rc = gp_scan(F, head, hpos, nrec, st))) return rc;
            nbox = gp_last_sum(head, hpos, nrec, st);
            if ((rc = F.alloc(&starts, nbox))) return rc;
            hipLaunchKernelGGL(k_gp_starts, dim3(grid_for(nrec, 256)), dim3(256), 0, st, (uint32_t)nrec, head, hpos,
                               starts);
            YG_HIP(hipGetLastError());
            uint32_t *bcode, *brgbe;
            if ((rc = F.alloc(&bcode, nbox)) || (rc = F.alloc(&brgbe, nbox))) return rc;
            if ((rc = dmalloc_keep(&G.d_pos, nbox)) || (rc = dmalloc_keep(&G.d_N, nbox)) ||
                (rc = dmalloc_keep(&G.d_key, nbox)) || (rc = dmalloc_keep(&G.d_irr, nbox)) ||
                (rc = dmalloc_keep(&G.d_qN, nbox)) || (rc = dmalloc_keep(&G.d_qirr, nbox)) ||
                (rc = dmalloc_keep(&G.d_read, 15 * nbox)))
                return rc;
            hipLaunchKernelGGL(k_gp_fold, dim3(grid_for(nbox, 256)), dim3(256), 0, st, (uint32_t)nbox, (uint32_t)nrec,
                               starts, order, skeys, recs, tab, G.d_pos, bcode, brgbe, G.d_N, G.d_key);
            YG_HIP(hipGetLastError());
            // uniform grid (cell = radius) over the box positions, built on the host
            std::vector<f3> bp(nbox);
            YG_HIP(hipMemcpyAsync(bp.data(), G.d_pos, sizeof(f3) * nbox, hipMemcpyDeviceToHost, st));
            YG_HIP(hipStreamSynchronize(st));
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (const f3& p : bp) {
                const float v[3] = {p.x, p.y, p.z};
                for (int a = 0; a < 3; a++) {
                    lo[a] = std::min(lo[a], v[a]);
                    hi[a] = std::max(hi[a], v[a]);
                }
            }
            int64_t ncell = 1;
            for (int a = 0; a < 3; a++) {
                G.gmin[a] = lo[a];
                int dd = (int)floorf((hi[a] - lo[a]) / G.radius) + 1;
                if (dd < 1) dd = 1;
                if (dd > 256) dd = 256;  // items beyond the last cell stay in it (clamped)
                G.gdim[a] = dd;
                ncell *= dd;
            }
            std::vector<int32_t> gs((size_t)ncell + 1, 0), gi(nbox), cof(nbox);
            for (uint64_t i = 0; i < nbox; i++) {
                const float v[3] = {bp[i].x, bp[i].y, bp[i].z};
                int c[3];
                for (int a = 0; a < 3; a++) {
                    int kk = (int)floorf((v[a] - G.gmin[a]) / G.radius);
                    if (kk < 0) kk = 0;
                    if (kk >= G.gdim[a]) kk = G.gdim[a] - 1;
                    c[a] = kk;
                }
                cof[i] = (c[2] * G.gdim[1] + c[1]) * G.gdim[0] + c[0];
                gs[(size_t)cof[i] + 1]++;
            }
            for (int64_t c = 0; c < ncell; c++) gs[c + 1] += gs[c];
            std::vector<int32_t> fill(gs.begin(), gs.end() - 1);
            for (uint64_t i = 0; i < nbox; i++) gi[fill[cof[i]]++] = (int32_t)i;
            if ((rc = dev_upload(&G.d_gstart, gs.data(), gs.size())) || (rc = dev_upload(&G.d_gitems, gi.data(), gi.size())))
                return rc;
            // pre-gathering
            if ((uint64_t)GL.search * nbox > (1ull << 31)) return yg_fail(YG_ERR_UNSUPPORTED, "global photon search x boxes too large");
            float* hd;
            int* hidx;
            if ((rc = F.alloc(&hd, (size_t)GL.search * nbox)) || (rc = F.alloc(&hidx, (size_t)GL.search * nbox)))
                return rc;
            if (ktime) YG_HIP(hipEventRecord(ev[2], st));
            if (GL.search <= GP_IRR_BUF - 64)
                hipLaunchKernelGGL(k_gp_irr_wave, dim3((unsigned)nbox), dim3(64), 0, st, (uint32_t)nbox, GL.search,
                                   G.radius, G.d_pos, bcode, brgbe, G.d_N, G.d_gstart, G.d_gitems, G.gmin[0], G.gmin[1],
                                   G.gmin[2], G.gdim[0], G.gdim[1], G.gdim[2], tab, G.d_irr);
            else
                hipLaunchKernelGGL(k_gp_irr, dim3(grid_for(nbox, 64)), dim3(64), 0, st, (uint32_t)nbox, GL.search,
                                   G.radius, G.d_pos, bcode, brgbe, G.d_N, G.d_gstart, G.d_gitems, G.gmin[0], G.gmin[1],
                                   G.gmin[2], G.gdim[0], G.gdim[1], G.gdim[2], tab, hd, hidx, G.d_irr);
            YG_HIP(hipGetLastError());
            if (ktime) YG_HIP(hipEventRecord(ev[3], st));
            irr_timed = ktime;
            hipLaunchKernelGGL(k_gp_finish, dim3(grid_for(nbox, 256)), dim3(256), 0, st, (uint32_t)nbox, G.d_pos, bcode,
                               brgbe, G.d_N, G.d_irr, tab, G.d_qN, G.d_qirr, G.d_read);
            YG_HIP(hipGetLastError());
        }
    }
    YG_HIP(hipStreamSynchronize(st));
    G.n = (int64_t)nbox;
    G.light = (int)li;
    G.gen = s->geo_gen;
    if (stats) {
        stats->emitted = nph;
        stats->stored = nrec;
        stats->clustered = nbox;
        stats->ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->rays = nrays;
        stats->ms_trace = 0.f;
        stats->ms_irr = 0.f;
        if (ktime && nph > 0) YG_HIP(hipEventElapsedTime(&stats->ms_trace, ev[0], ev[1]));
        if (irr_timed) YG_HIP(hipEventElapsedTime(&stats->ms_irr, ev[2], ev[3]));
    }
    return YG_OK;
}

extern "C" int yg_gphoton_read(yg_scene* s, float* d_out, uint64_t capacity, uint64_t* count, void* stream)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    if (s->gp.light < 0) return yg_fail(YG_ERR_STATE, "global photon map not built (yg_gphoton_build)");
    const uint64_t n = (uint64_t)s->gp.n;
    if (count) *count = n;
    if (d_out && n) {
        const uint64_t m = std::min(n, capacity);
        YG_HIP(hipMemcpyAsync(d_out, s->gp.d_read, sizeof(float) * 15 * m, hipMemcpyDeviceToDevice,
                              (hipStream_t)stream));
        YG_HIP(hipStreamSynchronize((hipStream_t)stream));
    }
    return YG_OK;
}

extern "C" int yg_gphoton_read_photons(yg_scene* s, uint32_t* d_out, uint64_t capacity, uint64_t* count,
                                       void* stream)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    if (s->gp.light < 0) return yg_fail(YG_ERR_STATE, "global photon map not built (yg_gphoton_build)");
    const uint64_t n = (uint64_t)s->gp.nrec;
    if (count) *count = n;
    if (d_out && n) {
        const uint64_t m = std::min(n, capacity);
        YG_HIP(hipMemcpyAsync(d_out, s->gp.d_rraw, sizeof(uint32_t) * 5 * m, hipMemcpyDeviceToDevice,
                              (hipStream_t)stream));
        YG_HIP(hipStreamSynchronize((hipStream_t)stream));
    }
    return YG_OK;
}
