// Photon map on the device: photonLight_t::init (photonlight.cc:453-501) --
// emission from the spot cone, diffuse photon tracing, hash-cell clustering
// and pre-gathering -- plus the gather index used by the renderer.
//
// The reference traces photons one after another and feeds a std::map of
// cells; here every photon is a thread.  The sequential parts are recovered
// exactly: Halton call indices of the bounce dimensions come from an
// exclusive scan over the photons that bounced, stored photon marks keep
// their (photon, depth) insertion order through a stable radix sort by cell,
// and each cell is summed in that order by one thread.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "photon_dev.h"
#include "qmc.h"


struct PhotonSrc {
    f3 from, light_dir, LU, LV;
    float angle_cos;
    c3 pcol;
    float bias, cluster;
    int mindepth, maxdepth;
};

struct PhotonState {
    f3 pos, lastpos;
    c3 col;
    f3 edir, ng2;
    int ti;
};

struct Mark {
    f3 dir, pos;
    c3 col;
};

// hash3d_t::getBox (hash3d.h:218-228)
__device__ __forceinline__ void hash_cell(float cell, f3 p, int& nx, int& ny, int& nz)
{
    nx = (int)(p.x / cell);
    ny = (int)(p.y / cell);
    nz = (int)(p.z / cell);
    if (p.x < 0.0f) nx--;
    if (p.y < 0.0f) ny--;
    if (p.z < 0.0f) nz--;
}

// findBox = findCreateBox(getBox(p)): the cell of the cell centre (hash3d.h:41-42)
__device__ __forceinline__ void find_box(float cell, f3 p, int& bx, int& by, int& bz)
{
    int nx, ny, nz;
    hash_cell(cell, p, nx, ny, nz);
    const f3 c = mk3((float)((double)(cell * (float)nx) + 0.5 * (double)cell),
                     (float)((double)(cell * (float)ny) + 0.5 * (double)cell),
                     (float)((double)(cell * (float)nz) + 0.5 * (double)cell));
    hash_cell(cell, c, bx, by, bz);
}

__global__ void k_emit_valid(PhotonSrc src, const float* __restrict__ h0, const float* __restrict__ h1,
                             uint32_t natt, uint32_t* __restrict__ valid)
{
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natt) return;
    const f3 d = random_vector_cone(src.light_dir, src.LU, src.LV, src.angle_cos, h0[a], h1[a]);
    valid[a] = isnull3(d) ? 0u : 1u;
}

__global__ void k_emit_pick(uint32_t natt, uint32_t np, const uint32_t* __restrict__ valid,
                            const uint32_t* __restrict__ vpos, uint32_t* __restrict__ att)
{
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natt || !valid[a] || vpos[a] >= np) return;
    att[vpos[a]] = a;
}

// One photon segment at depth d (shoot_photon_diffuse, photonlight.cc:203-244)
#define PB 64
template <int W>
__global__ __launch_bounds__(PB) void k_photon_seg(PhotonSrc src, BVHView bvh, const yg_material* __restrict__ mats,
                                                    const yg_modulator* __restrict__ mods,
                                                    const yg_texture* __restrict__ texs,
                                                    const int32_t* __restrict__ obj_flags, float min_raydis,
                                                    int d, const uint32_t* __restrict__ active, uint32_t nactive,
                                                    const uint32_t* __restrict__ att, const float* __restrict__ h0,
                                                    const float* __restrict__ h1, const uint32_t* __restrict__ bidx,
                                                    PhotonState* __restrict__ pst, Mark* __restrict__ marks,
                                                    uint32_t* __restrict__ mvalid, uint32_t* __restrict__ bflag,
                                                    unsigned long long* __restrict__ nrays)
{
    extern __shared__ int s_stk[];
    const uint32_t idx = blockIdx.x * PB + threadIdx.x;
    if (idx >= nactive) return;
    LaneStack st;
    st.init(s_stk, threadIdx.x, PB);
    const uint32_t p = active[idx];
    f3 pos, lastpos, dir;
    c3 col;
    if (d == 1) {
        const uint32_t a = att[p];
        dir = random_vector_cone(src.light_dir, src.LU, src.LV, src.angle_cos, h0[a], h1[a]);
        pos = src.from;
        lastpos = mk3(0, 0, 0);
        col = src.pcol;
    } else {
        const PhotonState s = pst[p];
        const uint32_t k = bidx[p];
        f3 NU, NV;
        const float4 A = bvh.tris[3 * s.ti], B = bvh.tris[3 * s.ti + 1], C = bvh.tris[3 * s.ti + 2];
        const f3 nrm = mk3(A.w, B.w, C.w);
        // sp.N, sp.NU, sp.NV of the previous hit (shading normal: interpolated on
        // autoSmooth'ed meshes)
        const f3 nsh = surf_normal(bvh, s.ti, s.pos, nrm);
        create_cs(nsh, NU, NV);
        dir = random_vector_cone(s.ng2, NU, NV, 0.05f, h0[k], h1[k]);
        // fromRadiosity (basicshaders.cc:30-44) at the previous hit
        const f3 prevP = s.pos;
        if (dot(face_forward(nrm, nsh, dir), s.edir) < 0) col = mkc(0, 0, 0);
        else {
            // the modulators see eye = refDir (photonlight.cc:237)
            Surf hs;
            hs.P = prevP;
            hs.N = nsh;
            hs.Ng = nrm;
            hs.ti = s.ti;
            hs.obj = bvh.tobj[s.ti];
            hs.mat = bvh.tmat[s.ti];
            col = cmul(s.col, get_diffuse(bvh, mats, mods, texs, hs, dir));
        }
        pos = s.pos;
        lastpos = s.lastpos;
    }
    atomicAdd(nrays, 1ull);
    float t;
    const int ti = first_hit<W>(bvh, pos, dir, min_raydis, t, st);
    if (ti < 0) return;
    const f3 f = pos + dir * min_raydis;
    const f3 hp = f + dir * t;
    // photon_t::position (photonlight.cc:39-47)
    const float ad = len3(pos - hp);
    if (ad > src.bias) lastpos = pos;
    pos = hp;
    f3 edir = norm3(lastpos - pos);
    const float4 A = bvh.tris[3 * ti], B = bvh.tris[3 * ti + 1], C = bvh.tris[3 * ti + 2];
    const f3 Ng = mk3(A.w, B.w, C.w);
    const f3 ng2 = face_forward(Ng, Ng, edir);
    const int fl = obj_flags[bvh.tobj[ti]];
    if (d > src.mindepth && (fl & 2)) {
        Mark m;
        m.pos = pos;
        m.col = col;
        m.dir = norm3(lastpos - pos);
        marks[(size_t)p * (src.maxdepth + 1) + (d - 1)] = m;
        mvalid[(size_t)p * (src.maxdepth + 1) + (d - 1)] = 1;
    }
    if ((fl & 4) && d <= src.maxdepth) {
        PhotonState s;
        s.pos = pos;
        s.lastpos = lastpos;
        s.col = col;
        s.edir = norm3(edir);
        s.ng2 = ng2;
        s.ti = ti;
        pst[p] = s;
        bflag[p] = 1;
    }
}

__global__ void k_photon_compact(uint32_t n, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ ex,
                                 uint32_t* __restrict__ active, uint32_t* __restrict__ bidx)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !flag[p]) return;
    active[ex[p]] = p;
    bidx[p] = ex[p];
}

__device__ __forceinline__ int wave_min(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Stored marks to their insertion-ordered positions, with their hash cells;
// the cell range of all marks is reduced per wave (one atomic per wave and
// bound instead of one per mark)
__global__ void k_mark_gather(uint32_t nslots, const uint32_t* __restrict__ mvalid, const uint32_t* __restrict__ mpos,
                              const Mark* __restrict__ slots, Mark* __restrict__ out, float cell,
                              int* __restrict__ cells, int* __restrict__ bounds)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    if (s < nslots && mvalid[s]) {
        const uint32_t o = mpos[s];
        const Mark m = slots[s];
        out[o] = m;
        int b[3];
        find_box(cell, m.pos, b[0], b[1], b[2]);
        for (int k = 0; k < 3; k++) {
            cells[3 * (size_t)o + k] = b[k];
            lo[k] = b[k];
            hi[k] = b[k];
        }
    }
    // every lane of the wave reaches the reductions (no early return above)
    for (int k = 0; k < 3; k++) {
        lo[k] = wave_min(lo[k]);
        hi[k] = wave_max(hi[k]);
    }
    if ((threadIdx.x & 63) == 0 && lo[0] != INT32_MAX) {
        for (int k = 0; k < 3; k++) {
            atomicMin(&bounds[2 * k], lo[k]);
            atomicMax(&bounds[2 * k + 1], hi[k]);
        }
    }
}

__global__ void k_mark_keys(uint32_t n, const int* __restrict__ cells, int bx0, int by0, int bz0, int sy, int sz,
                            unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long kx = (unsigned long long)(cells[3 * (size_t)i] - bx0);
    const unsigned long long ky = (unsigned long long)(cells[3 * (size_t)i + 1] - by0);
    const unsigned long long kz = (unsigned long long)(cells[3 * (size_t)i + 2] - bz0);
    keys[i] = (kx << (sy + sz)) | (ky << sz) | kz;
    vals[i] = i;
}

__global__ void k_seg_heads(uint32_t n, const unsigned long long* __restrict__ keys, uint32_t* __restrict__ head)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_seg_starts(uint32_t n, const uint32_t* __restrict__ head, const uint32_t* __restrict__ hpos,
                             uint32_t* __restrict__ starts)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    starts[hpos[i]] = i;
}

// preGathering (photonlight.cc:397-411): per cell, sums in insertion order
__global__ void k_seg_reduce(uint32_t nseg, uint32_t n, const uint32_t* __restrict__ starts,
                             const uint32_t* __restrict__ order, const Mark* __restrict__ marks,
                             float* __restrict__ rec, uint32_t* __restrict__ keep)
{
    const uint32_t sg = blockIdx.x * blockDim.x + threadIdx.x;
    if (sg >= nseg) return;
    const uint32_t a = starts[sg], b = (sg + 1 < nseg) ? starts[sg + 1] : n;
    f3 ds = mk3(0, 0, 0), ps = mk3(0, 0, 0);
    c3 cs = mkc(0, 0, 0);
    float count = 0;
    for (uint32_t k = a; k < b; k++) {
        const Mark m = marks[order[k]];
        ds = ds + m.dir;
        ps = ps + m.pos;
        cs = cadd(cs, m.col);
        count += 1;
    }
    keep[sg] = isnull3(ds) ? 0u : 1u;
    const f3 d = norm3(ds);
    const f3 p = ps / count;
    float* o = rec + 9 * (size_t)sg;
    o[0] = d.x; o[1] = d.y; o[2] = d.z; o[3] = p.x; o[4] = p.y; o[5] = p.z;
    o[6] = cs.r; o[7] = cs.g; o[8] = cs.b;
}

__global__ void k_seg_keep(uint32_t nseg, const uint32_t* __restrict__ keep, const uint32_t* __restrict__ kpos,
                           const float* __restrict__ rec, float* __restrict__ out)
{
    const uint32_t sg = blockIdx.x * blockDim.x + threadIdx.x;
    if (sg >= nseg || !keep[sg]) return;
    for (int k = 0; k < 9; k++) out[9 * (size_t)kpos[sg] + k] = rec[9 * (size_t)sg + k];
}

__global__ __launch_bounds__(64) void k_gather_points(RenderScene R, PhotonView pm, int K, float fr,
                                                      const float* __restrict__ P, const float* __restrict__ N,
                                                      const float* __restrict__ Ng, const float* __restrict__ E,
                                                      const int32_t* __restrict__ mat, uint32_t n,
                                                      float* __restrict__ rgb)
{
    extern __shared__ float s_knn[];  // K distances, then K indices per lane
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    Surf sp;
    sp.P = ld3(P + 3 * (size_t)i);
    sp.N = ld3(N + 3 * (size_t)i);
    sp.Ng = ld3(Ng + 3 * (size_t)i);
    create_cs(sp.N, sp.NU, sp.NV);
    sp.mat = mat[i];
    sp.obj = -1;
    sp.ti = -1;
    const c3 c = photon_estimate(R, pm, K, fr, sp, ld3(E + 3 * (size_t)i), true, s_knn + threadIdx.x,
                                 (int*)(s_knn + (size_t)K * 64) + threadIdx.x, 64);
    rgb[3 * (size_t)i] = c.r;
    rgb[3 * (size_t)i + 1] = c.g;
    rgb[3 * (size_t)i + 2] = c.b;
}

// ------------------------------------------------------------------ host --

template <class T>
static int dmalloc(T** p, size_t n)
{
    hipError_t e = hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1));
    if (e != hipSuccess) return yg_fail(YG_ERR_NOMEM, std::string("photon buffer: ") + hipGetErrorString(e));
    return YG_OK;
}

struct Freer {
    std::vector<void*> ptrs;
    ~Freer()
    {
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
    template <class T>
    int alloc(T** p, size_t n)
    {
        int rc = dmalloc(p, n);
        if (!rc) ptrs.push_back(*p);
        return rc;
    }
};

static int scan_u32(Freer& F, const uint32_t* in, uint32_t* out, size_t n, hipStream_t st)
{
    size_t bytes = 0;
    YG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, st));
    uint8_t* tmp = nullptr;
    int rc = F.alloc(&tmp, bytes);
    if (rc) return rc;
    YG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, st));
    return YG_OK;
}

static uint32_t read_last_sum(const uint32_t* flag, const uint32_t* ex, size_t n, hipStream_t st)
{
    uint32_t a = 0, b = 0;
    (void)hipMemcpyAsync(&a, flag + n - 1, 4, hipMemcpyDeviceToHost, st);
    (void)hipMemcpyAsync(&b, ex + n - 1, 4, hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    return a + b;
}

extern "C" int yg_photon_build(yg_scene* s, uint32_t li, const yg_render_opts* opts, yg_photon_stats* stats,
                               void* stream)
{
    if (!s || !opts) return yg_fail(YG_ERR_ARG, "null argument");
    if (!s->built) return yg_fail(YG_ERR_STATE, "scene not built");
    if (li >= s->lights.size() || s->lights[li].type != YG_LIGHT_PHOTON)
        return yg_fail(YG_ERR_ARG, "light index is not a photonlight");
    if (s->mats.empty()) return yg_fail(YG_ERR_STATE, "no materials (yg_set_shading)");
    const yg_light& L = s->lights[li];
    if (L.depth < 0 || L.depth + 1 - L.mindepth <= 0) return yg_fail(YG_ERR_ARG, "photon depth/mindepth");
    if (L.search < 1 || L.search > YG_KNN_MAX) return yg_fail(YG_ERR_UNSUPPORTED, "photon search must be 1..128");
    if (!(L.cluster > 0) || !(L.fixedradius > 0)) return yg_fail(YG_ERR_ARG, "cluster and fixedradius must be > 0");
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = perm_upload())) return rc;
    auto t0 = std::chrono::steady_clock::now();
    // photonLight_t ctor (photonlight.cc:63-103) and init (:453-470)
    int np = L.photons / (L.depth + 1 - L.mindepth);
    int rt = (int)sqrtf((float)np);
    np = rt * rt;
    PhotonSrc src;
    src.angle_cos = (float)cos(((double)L.angle / 180.0) * M_PI);
    const float pw = (float)((double)L.power * 2.0 * M_PI * (1.0 - (double)src.angle_cos));
    src.pcol = cscale(mkc(L.color[0], L.color[1], L.color[2]), pw);
    src.from = ld3(L.from);
    src.light_dir = norm3(ld3(L.to) - src.from);
    create_cs(src.light_dir, src.LU, src.LV);
    src.bias = L.bias;
    src.cluster = L.cluster;
    src.mindepth = L.mindepth;
    src.maxdepth = L.depth;
    const uint32_t natt = (uint32_t)np + 64;
    const int ndim = 2 * (L.depth + 1);
    Freer F;
    // Halton tables of the emission (dimensions 0, 1) and bounce dimensions,
    // built on the host with the reference's incremental rule and kept on the
    // device while the photon count and depth repeat
    if (s->ph_halton_np != np || (int)s->ph_halton.size() != ndim) {
        for (float* p : s->ph_halton)
            if (p) (void)hipFree(p);
        s->ph_halton.assign(ndim, nullptr);
        s->ph_halton_np = -1;
        std::vector<float> tab;
        int base = 2;
        for (int dd = 0; dd < ndim; dd++) {
            const uint64_t len = (dd < 2) ? natt : (uint64_t)np + 1;
            halton_incremental(base, len, tab);
            if ((rc = dmalloc(&s->ph_halton[dd], len))) return rc;
            YG_HIP(hipMemcpy(s->ph_halton[dd], tab.data(), sizeof(float) * len, hipMemcpyHostToDevice));
            base = next_prime(base);
        }
        s->ph_halton_np = np;
    }
    const std::vector<float*>& hdim = s->ph_halton;
    uint32_t *valid, *vpos, *att;
    if ((rc = F.alloc(&valid, natt)) || (rc = F.alloc(&vpos, natt)) || (rc = F.alloc(&att, np))) return rc;
    hipLaunchKernelGGL(k_emit_valid, dim3(grid_for(natt, 256)), dim3(256), 0, st, src, hdim[0], hdim[1], natt, valid);
    YG_HIP(hipGetLastError());
    if ((rc = scan_u32(F, valid, vpos, natt, st))) return rc;
    if (read_last_sum(valid, vpos, natt, st) < (uint32_t)np)
        return yg_fail(YG_ERR_STATE, "too many null emission directions");
    hipLaunchKernelGGL(k_emit_pick, dim3(grid_for(natt, 256)), dim3(256), 0, st, natt, (uint32_t)np, valid, vpos, att);
    YG_HIP(hipGetLastError());

    const int slots_per = L.depth + 1;
    const size_t nslots = (size_t)np * slots_per;
    PhotonState* pst;
    Mark *slots, *marks;
    uint32_t *mvalid, *mpos, *bflag, *bex, *bidx, *act0, *act1;
    unsigned long long* nrays;
    if ((rc = F.alloc(&pst, np)) || (rc = F.alloc(&slots, nslots)) || (rc = F.alloc(&mvalid, nslots)) ||
        (rc = F.alloc(&mpos, nslots)) || (rc = F.alloc(&bflag, np)) || (rc = F.alloc(&bex, np)) ||
        (rc = F.alloc(&bidx, np)) || (rc = F.alloc(&act0, np)) || (rc = F.alloc(&act1, np)) ||
        (rc = F.alloc(&nrays, 1)))
        return rc;
    YG_HIP(hipMemsetAsync(mvalid, 0, sizeof(uint32_t) * nslots, st));
    YG_HIP(hipMemsetAsync(nrays, 0, sizeof(unsigned long long), st));
    {
        std::vector<uint32_t> ident(np);
        for (int p = 0; p < np; p++) ident[p] = (uint32_t)p;
        YG_HIP(hipMemcpyAsync(act0, ident.data(), sizeof(uint32_t) * np, hipMemcpyHostToDevice, st));
        YG_HIP(hipStreamSynchronize(st));
    }
    uint32_t nactive = (uint32_t)np;
    uint32_t* act = act0;
    uint32_t* nxt = act1;
    for (int d = 1; d <= L.depth + 1 && nactive > 0; d++) {
        YG_HIP(hipMemsetAsync(bflag, 0, sizeof(uint32_t) * np, st));
        const float* h0 = (d == 1) ? hdim[0] : hdim[2 * (d - 1)];
        const float* h1 = (d == 1) ? hdim[1] : hdim[2 * (d - 1) + 1];
        YG_LAUNCH_W(s->width, (k_photon_seg<8>), (k_photon_seg<4>), (k_photon_seg<2>), dim3(grid_for(nactive, PB)), dim3(PB),
                    stack_lds_bytes(s->stack_entries, PB), st, src, s->view(), s->d_mats, s->d_mods, s->d_texs,
                    s->d_obj_flags, opts->min_raydis, d, act, nactive, att, h0, h1, bidx, pst, slots, mvalid, bflag,
                    nrays);
        YG_HIP(hipGetLastError());
        if (d > L.depth) break;
        if ((rc = scan_u32(F, bflag, bex, np, st))) return rc;
        hipLaunchKernelGGL(k_photon_compact, dim3(grid_for(np, 256)), dim3(256), 0, st, (uint32_t)np, bflag, bex,
                           nxt, bidx);
        YG_HIP(hipGetLastError());
        nactive = read_last_sum(bflag, bex, np, st);
        std::swap(act, nxt);
    }
    // compact marks in (photon, depth) order
    if ((rc = scan_u32(F, mvalid, mpos, nslots, st))) return rc;
    const uint32_t nmarks = read_last_sum(mvalid, mpos, nslots, st);
    int* cells;
    int* bounds;
    if ((rc = F.alloc(&marks, nmarks)) || (rc = F.alloc(&cells, 3 * (size_t)nmarks)) || (rc = F.alloc(&bounds, 6)))
        return rc;
    const int binit[6] = {INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
    YG_HIP(hipMemcpyAsync(bounds, binit, sizeof(binit), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_mark_gather, dim3(grid_for(nslots, 256)), dim3(256), 0, st, (uint32_t)nslots, mvalid, mpos,
                       slots, marks, src.cluster, cells, bounds);
    YG_HIP(hipGetLastError());
    int hb[6];
    YG_HIP(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, st));
    YG_HIP(hipStreamSynchronize(st));
    PhotonMapDev& pm = s->pm;
    if (pm.d_photons) (void)hipFree(pm.d_photons);
    if (pm.d_gstart) (void)hipFree(pm.d_gstart);
    if (pm.d_gitems) (void)hipFree(pm.d_gitems);
    pm = PhotonMapDev();
    uint32_t nphot = 0;
    std::vector<float> hphot;
    if (nmarks > 0) {
        auto bits = [](int64_t range) {
            int b = 0;
            while ((1ll << b) < range + 1) b++;
            return b;
        };
        const int sx = bits((int64_t)hb[1] - hb[0]), sy = bits((int64_t)hb[3] - hb[2]), sz = bits((int64_t)hb[5] - hb[4]);
        if (sx + sy + sz > 64) return yg_fail(YG_ERR_UNSUPPORTED, "photon cell range too large");
        unsigned long long *k0, *k1;
        uint32_t *v0, *v1, *head, *hpos, *starts;
        if ((rc = F.alloc(&k0, nmarks)) || (rc = F.alloc(&k1, nmarks)) || (rc = F.alloc(&v0, nmarks)) ||
            (rc = F.alloc(&v1, nmarks)) || (rc = F.alloc(&head, nmarks)) || (rc = F.alloc(&hpos, nmarks)))
            return rc;
        hipLaunchKernelGGL(k_mark_keys, dim3(grid_for(nmarks, 256)), dim3(256), 0, st, nmarks, cells, hb[0], hb[2],
                           hb[4], sy, sz, k0, v0);
        YG_HIP(hipGetLastError());
        size_t bytes = 0;
        const int endbit = std::max(1, sx + sy + sz);
        YG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k0, k1, v0, v1, (int)nmarks, 0, endbit, st));
        uint8_t* tmp;
        if ((rc = F.alloc(&tmp, bytes))) return rc;
        YG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, k0, k1, v0, v1, (int)nmarks, 0, endbit, st));
        hipLaunchKernelGGL(k_seg_heads, dim3(grid_for(nmarks, 256)), dim3(256), 0, st, nmarks, k1, head);
        YG_HIP(hipGetLastError());
        if ((rc = scan_u32(F, head, hpos, nmarks, st))) return rc;
        const uint32_t nseg = read_last_sum(head, hpos, nmarks, st);
        float* rec;
        uint32_t *keep, *kpos;
        if ((rc = F.alloc(&starts, nseg)) || (rc = F.alloc(&rec, 9 * (size_t)nseg)) || (rc = F.alloc(&keep, nseg)) ||
            (rc = F.alloc(&kpos, nseg)))
            return rc;
        hipLaunchKernelGGL(k_seg_starts, dim3(grid_for(nmarks, 256)), dim3(256), 0, st, nmarks, head, hpos, starts);
        hipLaunchKernelGGL(k_seg_reduce, dim3(grid_for(nseg, 128)), dim3(128), 0, st, nseg, nmarks, starts, v1, marks,
                           rec, keep);
        YG_HIP(hipGetLastError());
        if ((rc = scan_u32(F, keep, kpos, nseg, st))) return rc;
        nphot = read_last_sum(keep, kpos, nseg, st);
        if ((rc = dmalloc(&pm.d_photons, 9 * (size_t)nphot))) return rc;
        hipLaunchKernelGGL(k_seg_keep, dim3(grid_for(nseg, 256)), dim3(256), 0, st, nseg, keep, kpos, rec,
                           pm.d_photons);
        YG_HIP(hipGetLastError());
        hphot.resize(9 * (size_t)nphot);
        YG_HIP(hipMemcpyAsync(hphot.data(), pm.d_photons, sizeof(float) * hphot.size(), hipMemcpyDeviceToHost, st));
        YG_HIP(hipStreamSynchronize(st));
    }
    // uniform grid (cell = fixedradius) over the clustered photons for the K-nearest query
    pm.n = nphot;
    pm.gcell = L.fixedradius;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < nphot; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], hphot[9 * (size_t)i + 3 + k]);
            hi[k] = std::max(hi[k], hphot[9 * (size_t)i + 3 + k]);
        }
    int64_t ncell = 1;
    for (int k = 0; k < 3; k++) {
        if (nphot == 0) { lo[k] = 0; hi[k] = 0; }
        pm.gmin[k] = lo[k];
        int64_t dd = (int64_t)((hi[k] - lo[k]) / pm.gcell) + 1;
        if (dd > 1024) dd = 1024;
        pm.gdim[k] = (int)dd;
        ncell *= dd;
    }
    std::vector<int32_t> gstart((size_t)ncell + 1, 0), gitems(std::max<uint32_t>(nphot, 1));
    std::vector<int64_t> cellof(nphot);
    for (uint32_t i = 0; i < nphot; i++) {
        int64_t c = 0;
        for (int k = 2; k >= 0; k--) {
            int64_t g = (int64_t)((hphot[9 * (size_t)i + 3 + k] - pm.gmin[k]) / pm.gcell);
            g = std::min<int64_t>(std::max<int64_t>(g, 0), pm.gdim[k] - 1);
            c = c * pm.gdim[k] + g;
        }
        cellof[i] = c;
        gstart[c + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) gstart[c + 1] += gstart[c];
    {
        std::vector<int32_t> fill(gstart.begin(), gstart.end() - 1);
        for (uint32_t i = 0; i < nphot; i++) gitems[fill[cellof[i]]++] = (int32_t)i;
    }
    if ((rc = dev_upload(&pm.d_gstart, gstart.data(), gstart.size()))) return rc;
    if ((rc = dev_upload(&pm.d_gitems, gitems.data(), gitems.size()))) return rc;
    if (!pm.d_photons && (rc = dmalloc(&pm.d_photons, 9))) return rc;
    pm.np_emit = np;
    pm.light = (int)li;
    if (stats) {
        unsigned long long nr = 0;
        YG_HIP(hipMemcpy(&nr, nrays, sizeof(nr), hipMemcpyDeviceToHost));
        stats->emitted = (uint64_t)np;
        stats->stored = nmarks;
        stats->clustered = nphot;
        stats->ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return YG_OK;
}

extern "C" int yg_photon_read(yg_scene* s, float* d_out, uint64_t capacity, uint64_t* count, void* stream)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    if (s->pm.light < 0) return yg_fail(YG_ERR_STATE, "photon map not built");
    if (count) *count = (uint64_t)s->pm.n;
    const uint64_t n = std::min<uint64_t>(capacity, (uint64_t)s->pm.n);
    if (n && d_out)
        YG_HIP(hipMemcpyAsync(d_out, s->pm.d_photons, sizeof(float) * 9 * n, hipMemcpyDeviceToDevice,
                              (hipStream_t)stream));
    return YG_OK;
}

extern "C" int yg_photon_gather(yg_scene* s, const float* d_p, const float* d_n, const float* d_ng,
                                const float* d_eye, const int32_t* d_mat, uint32_t n, float* d_rgb, void* stream)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    if (s->pm.light < 0) return yg_fail(YG_ERR_STATE, "photon map not built");
    if (s->mats.empty()) return yg_fail(YG_ERR_STATE, "no materials");
    if (n == 0) return YG_OK;
    if (!d_p || !d_n || !d_ng || !d_eye || !d_mat || !d_rgb) return yg_fail(YG_ERR_ARG, "null buffer");
    int rc = perm_upload();
    if (rc) return rc;
    const yg_light& L = s->lights[s->pm.light];
    RenderScene R;
    memset(&R, 0, sizeof(R));
    R.bvh = s->view();
    R.mats = s->d_mats;
    R.mods = s->d_mods;
    R.texs = s->d_texs;
    PhotonView pv;
    pv.photons = s->pm.d_photons;
    pv.gstart = s->pm.d_gstart;
    pv.gitems = s->pm.d_gitems;
    for (int k = 0; k < 3; k++) { pv.gmin[k] = s->pm.gmin[k]; pv.gdim[k] = s->pm.gdim[k]; }
    pv.gcell = s->pm.gcell;
    pv.np_emit = s->pm.np_emit;
    pv.n = s->pm.n;
    hipLaunchKernelGGL(k_gather_points, dim3(grid_for(n, 64)), dim3(64), knn_lds_bytes(L.search), (hipStream_t)stream, R, pv, L.search,
                       L.fixedradius, d_p, d_n, d_ng, d_eye, d_mat, n, d_rgb);
    YG_HIP(hipGetLastError());
    return YG_OK;
}
