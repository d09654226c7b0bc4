// Device-side photon-map radiance estimate: photonLight_t::illuminate
// (photonlight.cc:297-395) with filterCone (photonlight.h:205-210) and
// fillArea (photonlight.cc:259-290).
#pragma once

#include "render_dev.h"

#define YG_KNN_MAX 128  // photon "search" (K) at most: two K-entry rows per lane, 64 KB per 64-lane block

// Dynamic LDS of a 64-lane K-nearest launch: a K-entry distance heap and a
// K-entry index heap per lane ([k * 64 + lane]), sized for the launch's largest K
static inline size_t knn_lds_bytes(int kmax) { return (size_t)2 * (size_t)kmax * 64 * sizeof(float); }

// randomVectorCone (vector3d.cc:162-168)
__device__ __host__ __forceinline__ f3 random_vector_cone(f3 D, f3 U, f3 V, float cosang, float z1, float z2)
{
    const float t1 = (float)(2.0 * M_PI * (double)z1);
    const float t2 = (float)(1.0 - (1.0 - (double)cosang) * (double)z2);
    const f3 a = U * dcos(t1) + V * dsin(t1);
    return a * (float)sqrt(1.0 - (double)(t2 * t2)) + D * t2;
}

// (dis, idx) ordering of candidate photons: nearest first, index breaks ties
__device__ __forceinline__ bool knn_less(float da, int ia, float db, int ib)
{
    return da < db || (da == db && ia < ib);
}

// K nearest photons with dir.N > 0 and dis <= fixedRadius, then the cone-
// filtered estimate.  The K photons are summed in increasing (dis, index)
// order.  heap_d / heap_i: this lane's K-entry scratch (stride `stride`).
__device__ __forceinline__ c3 photon_estimate(const RenderScene& R, const PhotonView& pm, int K, float fixedradius,
                                              const Surf& sp, f3 eye, bool receive, float* heap_d, int* heap_i,
                                              int stride)
{
    if (!receive) return mkc(0, 0, 0);
    const f3 N = face_forward(sp.Ng, sp.N, eye);
    int n = 0;
    const float pp[3] = {sp.P.x, sp.P.y, sp.P.z};
    int g0[3], g1[3];
    for (int k = 0; k < 3; k++) {
        long a = (long)floorf((pp[k] - fixedradius - pm.gmin[k]) / pm.gcell);
        long b = (long)floorf((pp[k] + fixedradius - pm.gmin[k]) / pm.gcell);
        if (a < 0) a = 0;
        if (b >= pm.gdim[k]) b = pm.gdim[k] - 1;
        g0[k] = (int)a;
        g1[k] = (int)b;
    }
    for (int z = g0[2]; z <= g1[2]; z++)
        for (int y = g0[1]; y <= g1[1]; y++)
            for (int x = g0[0]; x <= g1[0]; x++) {
                const int c = (z * pm.gdim[1] + y) * pm.gdim[0] + x;
                for (int q = pm.gstart[c]; q < pm.gstart[c + 1]; q++) {
                    const int idx = pm.gitems[q];
                    const float* ph = pm.photons + 9 * (size_t)idx;
                    if (dot(mk3(ph[0], ph[1], ph[2]), N) <= 0.0f) continue;
                    const float dis = len3(sp.P - mk3(ph[3], ph[4], ph[5]));
                    if (!(dis <= fixedradius)) continue;
                    if (n < K) {
                        // sift up into a max-heap on (dis, idx)
                        int pos = n++;
                        while (pos > 0) {
                            const int par = (pos - 1) >> 1;
                            if (!knn_less(heap_d[par * stride], heap_i[par * stride], dis, idx)) break;
                            heap_d[pos * stride] = heap_d[par * stride];
                            heap_i[pos * stride] = heap_i[par * stride];
                            pos = par;
                        }
                        heap_d[pos * stride] = dis;
                        heap_i[pos * stride] = idx;
                    } else if (knn_less(dis, idx, heap_d[0], heap_i[0])) {
                        // replace the farthest and sift down
                        int pos = 0;
                        for (;;) {
                            const int l = 2 * pos + 1, r = l + 1;
                            int big = -1;
                            float bd = dis;
                            int bi = idx;
                            if (l < n && knn_less(bd, bi, heap_d[l * stride], heap_i[l * stride])) {
                                big = l; bd = heap_d[l * stride]; bi = heap_i[l * stride];
                            }
                            if (r < n && knn_less(bd, bi, heap_d[r * stride], heap_i[r * stride])) big = r;
                            if (big < 0) break;
                            heap_d[pos * stride] = heap_d[big * stride];
                            heap_i[pos * stride] = heap_i[big * stride];
                            pos = big;
                        }
                        heap_d[pos * stride] = dis;
                        heap_i[pos * stride] = idx;
                    }
                }
            }
    if (n < 2) return mkc(0, 0, 0);
    // heap -> ascending order (in-place heap sort)
    for (int end = n - 1; end > 0; end--) {
        const float td = heap_d[end * stride];
        const int tix = heap_i[end * stride];
        heap_d[end * stride] = heap_d[0];
        heap_i[end * stride] = heap_i[0];
        int pos = 0;
        for (;;) {
            const int l = 2 * pos + 1, r = l + 1;
            int big = -1;
            float bd = td;
            int bi = tix;
            if (l < end && knn_less(bd, bi, heap_d[l * stride], heap_i[l * stride])) {
                big = l; bd = heap_d[l * stride]; bi = heap_i[l * stride];
            }
            if (r < end && knn_less(bd, bi, heap_d[r * stride], heap_i[r * stride])) big = r;
            if (big < 0) break;
            heap_d[pos * stride] = heap_d[big * stride];
            heap_i[pos * stride] = heap_i[big * stride];
            pos = big;
        }
        heap_d[pos * stride] = td;
        heap_i[pos * stride] = tix;
    }
    const float radius = heap_d[(n - 1) * stride];
    c3 total = mkc(0, 0, 0);
    f3 VX, VY;
    create_cs(N, VX, VY);
    float avgX = 0, avgY = 0, den = 0;
    // first pass: radiance and the weighted centroid of the sample disc
    for (int k = 0; k < n; k++) {
        const float* ph = pm.photons + 9 * (size_t)heap_i[k * stride];
        const float dis = heap_d[k * stride];
        const float filter = (radius == 0) ? 1.0f : (float)((double)(float)(1.0 - (double)(dis / radius)) * 2.0);
        const f3 pdir = mk3(ph[0], ph[1], ph[2]);
        total = cadd(total, from_light(R, sp, pdir, cscale(mkc(ph[6], ph[7], ph[8]), filter), eye));
        const f3 sub = (mk3(ph[3], ph[4], ph[5]) - sp.P) / radius;
        const float sx = dot(sub, VX), sy = dot(sub, VY);
        const float sw = (float)((double)(dot(pdir, N) * filter) * 0.5);
        avgX += sx * sw;
        avgY += sy * sw;
        den += sw;
    }
    float area = (float)(M_PI * (double)radius * (double)radius);
    float fa = 1.0f;
    if (den != 0.0f) {
        const float SQRT2 = (float)sqrt(2.0);
        const float offset = (float)1.414213562373095;
        float avgR = 0;
        avgX /= den;
        avgY /= den;
        for (int k = 0; k < n; k++) {
            const float* ph = pm.photons + 9 * (size_t)heap_i[k * stride];
            const float dis = heap_d[k * stride];
            const float filter = (radius == 0) ? 1.0f : (float)((double)(float)(1.0 - (double)(dis / radius)) * 2.0);
            const f3 pdir = mk3(ph[0], ph[1], ph[2]);
            const f3 sub = (mk3(ph[3], ph[4], ph[5]) - sp.P) / radius;
            const float dx = dot(sub, VX) - avgX, dy = dot(sub, VY) - avgY;
            const float sw = (float)((double)(dot(pdir, N) * filter) * 0.5);
            avgR += sqrtf(dx * dx + dy * dy) * sw;
        }
        avgR = (avgR / den) * SQRT2 * offset;
        avgR *= avgR;
        if (avgR < 0.25f) avgR = 0.25f;
        fa = avgR;
    }
    area *= fa;
    const float minarea = (float)(4.0 * M_PI / (double)(float)pm.np_emit);
    if (area < minarea) area = minarea;
    const float factor = (float)(4.0 * M_PI / (double)((float)pm.np_emit * area));
    return cscale(total, factor);
}
