// Device-side irradiance lookup of the global photon map: the irradiance
// part of pathLight_t::getLight (pathlight.cc:140-176) at a path vertex.
#pragma once

#include "dev_math.h"

// The pre-gathered hash boxes of globalPhotonLight_t (globalphotonlight.cc
// :197-228), sorted by box index (x, y, z).
struct GPhotonView {
    const unsigned long long* key;  // box index packed as in gp_box_key
    const f3* N;                    // compPhoton_t::N (normal of the first stored photon)
    const c3* irr;                  // compPhoton_t::irr
    const f3* pos;                  // box photon position
    const f3* qN;                   // irradiance photon direction: N quantised and decoded
    const c3* qirr;                 // irradiance photon colour: irr through RGBE
    const int32_t* gstart;          // uniform grid (cell = radius) over the box positions
    const int32_t* gitems;
    float gmin[3];
    int gdim[3];
    float cell;                     // hash box size
    float radius;                   // maxradius of the irradiance map
    int bmin[3], bmax[3];           // box index bounds
    int sy, sz;                     // key shifts
    int64_t n;
    // the raw photon map (photonMap, globalphotonlight.cc:38) in store order,
    // decoded, with a uniform grid (cell = radius): photonSampler_t's gathers
    const f3* rpos;
    const f3* rdir;
    const c3* rcol;
    const int32_t* rgstart;
    const int32_t* rgitems;
    float rgmin[3];
    int rgdim[3];
    int64_t nrec;
};

// hash3d_t::getBox (hash3d.h:218-228)
__device__ __host__ __forceinline__ void gp_hash_cell(float cell, f3 p, int& nx, int& ny, int& nz)
{
    nx = (int)(p.x / cell);
    ny = (int)(p.y / cell);
    nz = (int)(p.z / cell);
    if (p.x < 0.0f) nx--;
    if (p.y < 0.0f) ny--;
    if (p.z < 0.0f) nz--;
}

__device__ __host__ __forceinline__ unsigned long long gp_box_key(int bx, int by, int bz, const int* bmin, int sy,
                                                                   int sz)
{
    return ((unsigned long long)(bx - bmin[0]) << (sy + sz)) | ((unsigned long long)(by - bmin[1]) << sz) |
           (unsigned long long)(bz - bmin[2]);
}

// the 5 nearest of (dis, idx) in ascending order, kept in registers
struct GpNear5 {
    float d[5];
    int i[5];
    int n;
    __device__ void init() { n = 0; }
    __device__ static bool less(float da, int ia, float db, int ib) { return da < db || (da == db && ia < ib); }
    __device__ void offer(float dis, int idx)
    {
        if (n == 5 && !less(dis, idx, d[4], i[4])) return;
        int k = (n < 5) ? n++ : 4;
        while (k > 0 && less(dis, idx, d[k - 1], i[k - 1])) {
            d[k] = d[k - 1];
            i[k] = i[k - 1];
            k--;
        }
        d[k] = dis;
        i[k] = idx;
    }
};

// getLight's irradiance (pathlight.cc:147-176): the box holding P when its
// normal agrees with N (N.N > 0.7: irr x factor), otherwise the mean of the
// 5 nearest irradiance photons within the radius whose directions face N
// (dot > 0.25), weighted by dot x (1 - dis / farthest).  The gather starts at
// the map's radius on every call (the reference carries a per-thread radius
// from call to call, globalPhotonMap_t::gather, photon.cc:167-210).
__device__ __forceinline__ c3 gp_irradiance(const GPhotonView& G, f3 P, f3 N)
{
    if (G.n == 0) return mkc(0, 0, 0);
    int bx, by, bz;
    gp_hash_cell(G.cell, P, bx, by, bz);
    if (bx >= G.bmin[0] && bx <= G.bmax[0] && by >= G.bmin[1] && by <= G.bmax[1] && bz >= G.bmin[2] &&
        bz <= G.bmax[2]) {
        const unsigned long long k = gp_box_key(bx, by, bz, G.bmin, G.sy, G.sz);
        int64_t lo = 0, hi = G.n;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (G.key[mid] < k) lo = mid + 1;
            else hi = mid;
        }
        if (lo < G.n && G.key[lo] == k) {
            const float factor = dot(G.N[lo], N);
            if (factor > 0.7f) return cscale(G.irr[lo], factor);
        }
    }
    GpNear5 f;
    f.init();
    const float r = G.radius;
    const float pp[3] = {P.x, P.y, P.z};
    int g0[3], g1[3];
    for (int a = 0; a < 3; a++) {
        long lo = (long)floorf((pp[a] - r - G.gmin[a]) / r);
        long hi = (long)floorf((pp[a] + r - G.gmin[a]) / r);
        if (lo < 0) lo = 0;
        if (hi >= G.gdim[a]) hi = G.gdim[a] - 1;
        g0[a] = (int)lo;
        g1[a] = (int)hi;
    }
    for (int z = g0[2]; z <= g1[2]; z++)
        for (int y = g0[1]; y <= g1[1]; y++)
            for (int x = g0[0]; x <= g1[0]; x++) {
                const int c = (z * G.gdim[1] + y) * G.gdim[0] + x;
                for (int q = G.gstart[c]; q < G.gstart[c + 1]; q++) {
                    const int i = G.gitems[q];
                    const float D = len3(G.pos[i] - P);
                    if ((D > r) || (dot(G.qN[i], N) <= 0.25f)) continue;
                    f.offer(D, i);
                }
            }
    c3 total = mkc(0, 0, 0);
    if (!f.n) return total;
    float farest = (f.n == 1) ? r : f.d[f.n - 1];
    if (farest == 0.0f) farest = 1.0f;
    float div = 0.0f;
    for (int k = 0; k < f.n; k++) {
        const float factor = dot(G.qN[f.i[k]], N) * (1.0f - f.d[k] / farest);
        if (factor > 0) {
            total = cadd(total, cscale(G.qirr[f.i[k]], factor));
            div += factor;
        }
    }
    if (div > 0.0f) total = cscale(total, (float)(1.0 / (double)div));
    return total;
}
