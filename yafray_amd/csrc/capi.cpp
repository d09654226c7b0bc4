// C ABI of libyafgpu: errors, scene lifecycle, geometry ingestion, BVH build,
// shading descriptors, QMC tables.
#include <atomic>
#include <cmath>
#include <chrono>
#include <cstring>
#include <mutex>

#include "qmc.h"
#include "yg_internal.h"

static thread_local std::string g_err;

void yg_set_error(const std::string& msg) { g_err = msg; }

int yg_fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

extern "C" int yg_abi_version(void) { return YG_ABI_VERSION; }

extern "C" int yg_last_error(char* buf, size_t len)
{
    if (!buf || len == 0) return YG_ERR_ARG;
    size_t n = std::min(len - 1, g_err.size());
    memcpy(buf, g_err.data(), n);
    buf[n] = 0;
    return YG_OK;
}

extern "C" int yg_device_count(int* n)
{
    if (!n) return yg_fail(YG_ERR_ARG, "null");
    YG_HIP(hipGetDeviceCount(n));
    return YG_OK;
}

extern "C" int yg_scene_create(yg_scene** out)
{
    if (!out) return yg_fail(YG_ERR_ARG, "null out");
    *out = new (std::nothrow) yg_scene();
    if (!*out) return yg_fail(YG_ERR_NOMEM, "scene allocation");
    static std::atomic<uint64_t> next_uid{1};
    (*out)->uid = next_uid++;
    return YG_OK;
}

static void free_dev(yg_scene* s)
{
    void* ptrs[] = {s->d_nodes, s->d_nodes4, s->d_nodes8, s->d_tvn, s->d_ptex, s->d_oorco, s->d_tris, s->d_tprim, s->d_tobj, s->d_tmat, s->d_obj_flags,
                    s->d_mats, s->d_mods, s->d_texs, s->pm.d_photons, s->pm.d_gstart,
                    s->pm.d_gitems, s->d_halton, s->d_any_ovf};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    s->d_nodes = nullptr; s->d_nodes4 = nullptr; s->d_nodes8 = nullptr; s->d_tvn = nullptr; s->d_tris = nullptr;
    s->d_ptex = nullptr; s->d_oorco = nullptr;
    s->d_tprim = nullptr; s->d_tobj = nullptr; s->d_tmat = nullptr; s->d_obj_flags = nullptr; s->d_mats = nullptr;
    s->d_mods = nullptr;
    s->d_texs = nullptr; s->pm = PhotonMapDev();
    gphoton_free(s->gp);
    s->d_halton = nullptr; s->halton_ndim = 0; s->halton_stride = 0;
    s->d_any_ovf = nullptr; s->any_ovf_ints = 0;
    for (float* p : s->ph_halton)
        if (p) (void)hipFree(p);
    s->ph_halton.clear();
    s->ph_halton_np = -1;
}

extern "C" int yg_scene_destroy(yg_scene* s)
{
    if (!s) return YG_OK;
    free_dev(s);
    delete s;
    return YG_OK;
}

extern "C" int yg_mesh_add(yg_scene* s, const float* xyz, uint32_t nv, const uint32_t* idx,
                           uint32_t nt, const int32_t* mat, int32_t flags, int32_t* obj_id)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    if ((nv && !xyz) || (nt && !idx)) return yg_fail(YG_ERR_ARG, "null geometry");
    for (uint64_t i = 0; i < 3ull * nt; i++)
        if (idx[i] >= nv) return yg_fail(YG_ERR_ARG, "vertex index out of range");
    const uint32_t vbase = (uint32_t)(s->verts.size() / 3);
    if ((uint64_t)vbase + nv > 0xffffffffull) return yg_fail(YG_ERR_ARG, "too many vertices");
    if (s->tri_mat.size() + nt > (size_t)(0x7fffffff / YG_LEAF_MAX))
        return yg_fail(YG_ERR_ARG, "too many triangles");
    s->verts.insert(s->verts.end(), xyz, xyz + 3ull * nv);
    for (uint64_t i = 0; i < 3ull * nt; i++) s->idx.push_back(idx[i] + vbase);
    const int32_t oid = s->nobj++;
    s->tri_obj.insert(s->tri_obj.end(), nt, oid);
    if (mat) s->tri_mat.insert(s->tri_mat.end(), mat, mat + nt);
    else s->tri_mat.insert(s->tri_mat.end(), nt, 0);
    s->obj_flags.push_back(flags);
    s->obj_vbase.push_back(vbase);
    s->obj_nv.push_back(nv);
    s->obj_orco.insert(s->obj_orco.end(), 12, 0.0f);
    s->obj_orco_set.push_back(0);
    if (!s->tri_smooth.empty()) {
        s->tri_smooth.resize(s->tri_mat.size(), 0);
        s->tri_vn.resize(9 * s->tri_mat.size(), 0.0f);
    }
    if (!s->tri_tex.empty()) s->tri_tex.resize(16 * s->tri_mat.size(), 0.0f);
    s->built = false;
    s->gen++;
    s->geo_gen++;
    if (obj_id) *obj_id = oid;
    return YG_OK;
}

extern "C" int yg_mesh_autosmooth(yg_scene* s, int32_t obj_id, float angle)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    if (obj_id < 0 || obj_id >= s->nobj) return yg_fail(YG_ERR_ARG, "object id out of range");
    if (!(angle >= 1.0f)) return YG_OK;  // mesh.cc:204: no smoothing below 1 degree
    const size_t nt = s->tri_mat.size();
    // the object's triangles (contiguous in add order) and its vertex range
    size_t t0 = 0;
    while (t0 < nt && s->tri_obj[t0] != obj_id) t0++;
    size_t t1 = t0;
    while (t1 < nt && s->tri_obj[t1] == obj_id) t1++;
    if (t0 == t1) return YG_OK;
    if (s->tri_smooth.empty()) {
        s->tri_smooth.assign(nt, 0);
        s->tri_vn.assign(9 * nt, 0.0f);
    }
    uint32_t v0 = UINT32_MAX, v1 = 0;
    for (size_t t = t0; t < t1; t++)
        for (int k = 0; k < 3; k++) {
            v0 = std::min(v0, s->idx[3 * t + k]);
            v1 = std::max(v1, s->idx[3 * t + k] + 1);
        }
    auto vert = [&](uint32_t i) { return ld3(&s->verts[3 * (size_t)i]); };
    // triangle_t::N(): the normal of triangle_t::recNormal (triangle.cc:45-49)
    std::vector<f3> tn(t1 - t0);
    for (size_t t = t0; t < t1; t++)
        tn[t - t0] = tri_normal(vert(s->idx[3 * t]), vert(s->idx[3 * t + 1]), vert(s->idx[3 * t + 2]));
    const size_t nv = v1 - v0;
    if (angle >= 180.0f) {
        // one normal per vertex: the face normals summed in triangle order, normalised
        std::vector<f3> vn(nv, mk3(0, 0, 0));
        for (size_t t = t0; t < t1; t++)
            for (int k = 0; k < 3; k++) {
                f3& n = vn[s->idx[3 * t + k] - v0];
                n = n + tn[t - t0];
            }
        for (auto& n : vn) n = norm3(n);
        for (size_t t = t0; t < t1; t++)
            for (int k = 0; k < 3; k++) {
                const f3 n = vn[s->idx[3 * t + k] - v0];
                s->tri_vn[9 * t + 3 * k] = n.x; s->tri_vn[9 * t + 3 * k + 1] = n.y; s->tri_vn[9 * t + 3 * k + 2] = n.z;
            }
    } else {
        // per corner: the face normals of the vertex's triangles (in triangle
        // order) whose dot with this face's normal exceeds cos(angle)
        const float cosa = (float)cos((double)angle * M_PI / 180.0);
        std::vector<std::vector<uint32_t>> cnx(nv);
        for (size_t t = t0; t < t1; t++)
            for (int k = 0; k < 3; k++) cnx[s->idx[3 * t + k] - v0].push_back((uint32_t)(t - t0));
        for (size_t t = t0; t < t1; t++) {
            const f3 N = tn[t - t0];
            for (int k = 0; k < 3; k++) {
                f3 sn = mk3(0, 0, 0);
                for (uint32_t o : cnx[s->idx[3 * t + k] - v0])
                    if (dot(tn[o], N) > cosa) sn = sn + tn[o];
                sn = norm3(sn);
                s->tri_vn[9 * t + 3 * k] = sn.x; s->tri_vn[9 * t + 3 * k + 1] = sn.y; s->tri_vn[9 * t + 3 * k + 2] = sn.z;
            }
        }
    }
    for (size_t t = t0; t < t1; t++) s->tri_smooth[t] = 1;
    s->built = false;
    s->gen++;
    s->geo_gen++;
    return YG_OK;
}

// ------------------------------------------------- texture coordinates --

// matrix4x4_t::inverse (matrix4.cc:78-112): Gauss-Jordan elimination with
// partial pivoting on the rows, in float, the identity receiving the same row
// operations.  False for a singular matrix (where the reference exits).
static bool inverse4(float m[4][4])
{
    float id[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    for (int i = 0; i < 4; ++i) {
        float mx = 0;
        int ci = 0;
        for (int k = i; k < 4; ++k)
            if (std::fabs(m[k][i]) > mx) {
                mx = std::fabs(m[k][i]);
                ci = k;
            }
        if (mx == 0) return false;
        for (int j = 0; j < 4; ++j) {
            std::swap(m[i][j], m[ci][j]);
            std::swap(id[i][j], id[ci][j]);
        }
        float f = m[i][i];
        for (int j = 0; j < 4; ++j) m[i][j] /= f;
        for (int j = 0; j < 4; ++j) id[i][j] /= f;
        for (int k = 0; k < 4; ++k) {
            if (k == i) continue;
            f = m[k][i];
            for (int j = 0; j < 4; ++j) {
                const float t = m[i][j] * f;
                m[k][j] -= t;
            }
            for (int j = 0; j < 4; ++j) {
                const float t = id[i][j] * f;
                id[k][j] -= t;
            }
        }
    }
    memcpy(m, id, sizeof(id));
    return true;
}

// meshObject_t::recalcBound (mesh.cc:343-380): the corners of the triangles,
// widened by BOX_MIN (1e-5, a double, so each bound is rounded once)
static void tri_bound(const float* xyz, const uint32_t* idx, uint32_t nt, float lo[3], float hi[3])
{
    for (int a = 0; a < 3; a++) lo[a] = hi[a] = xyz[3 * (size_t)idx[0] + a];
    for (uint64_t k = 0; k < 3ull * nt; k++)
        for (int a = 0; a < 3; a++) {
            const float v = xyz[3 * (size_t)idx[k] + a];
            if (v > hi[a]) hi[a] = v;
            if (v < lo[a]) lo[a] = v;
        }
    for (int a = 0; a < 3; a++) {
        lo[a] = (float)((double)lo[a] - 0.00001);
        hi[a] = (float)((double)hi[a] + 0.00001);
    }
}

extern "C" int yg_orco_matrix(const float* m16, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                              float* out12)
{
    if (!xyz || !idx || !out12 || nt == 0) return yg_fail(YG_ERR_ARG, "null or empty mesh");
    for (uint64_t i = 0; i < 3ull * nt; i++)
        if (idx[i] >= nv) return yg_fail(YG_ERR_ARG, "vertex index out of range");
    float M[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    if (m16) memcpy(M, m16, sizeof(M));
    // backOrco = M scaled by half the untransformed bound's extent (columns 0..2)
    float lo[3], hi[3];
    tri_bound(xyz, idx, nt, lo, hi);
    float B[4][4];
    memcpy(B, M, sizeof(B));
    for (int c = 0; c < 3; c++) {
        const float sc = (float)(0.5 * (double)(hi[c] - lo[c]));
        for (int r = 0; r < 3; r++) B[r][c] *= sc;
    }
    // the transformed vertices (matrix4x4_t * point3d_t, matrix4.h:86-91)
    std::vector<float> w(3 * (size_t)nv);
    for (uint32_t v = 0; v < nv; v++) {
        const float x = xyz[3 * (size_t)v], y = xyz[3 * (size_t)v + 1], z = xyz[3 * (size_t)v + 2];
        for (int r = 0; r < 3; r++) {
            float acc = M[r][0] * x;
            const float ty = M[r][1] * y;
            acc += ty;
            const float tz = M[r][2] * z;
            acc += tz;
            acc += M[r][3];
            w[3 * (size_t)v + r] = acc;
        }
    }
    // translation = centre of the transformed bound, then the inverse
    tri_bound(w.data(), idx, nt, lo, hi);
    for (int r = 0; r < 3; r++) {
        const float sum = lo[r] + hi[r];
        B[r][3] = 0.5f * sum;
    }
    if (!inverse4(B)) return yg_fail(YG_ERR_ARG, "singular orco matrix");
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) out12[4 * r + c] = B[r][c];
    return YG_OK;
}

extern "C" int yg_mesh_texcoords(yg_scene* s, int32_t obj_id, const float* face_uv, const uint8_t* face_flags,
                                 const float* orco, const float* orco_mtx)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    if (obj_id < 0 || obj_id >= s->nobj) return yg_fail(YG_ERR_ARG, "object id out of range");
    const size_t nt = s->tri_mat.size();
    size_t t0 = 0;
    while (t0 < nt && s->tri_obj[t0] != obj_id) t0++;
    size_t t1 = t0;
    while (t1 < nt && s->tri_obj[t1] == obj_id) t1++;
    if (orco_mtx) {
        memcpy(&s->obj_orco[12 * (size_t)obj_id], orco_mtx, 12 * sizeof(float));
        s->obj_orco_set[obj_id] = 1;
    }
    if ((face_uv || face_flags || orco) && t1 > t0) {
        if (s->tri_tex.empty()) s->tri_tex.assign(16 * nt, 0.0f);
        const uint32_t vb = s->obj_vbase[obj_id];
        for (size_t t = t0; t < t1; t++) {
            float* r = &s->tri_tex[16 * t];
            const size_t k = t - t0;
            int fl = face_flags ? (face_flags[k] & 3) : (face_uv ? 1 : 0);
            if (face_uv && (fl & 1))
                for (int c = 0; c < 6; c++) r[c] = face_uv[6 * k + c];
            else
                for (int c = 0; c < 6; c++) r[c] = 0.0f;
            if (!face_uv) fl &= ~1;
            if (orco) {
                for (int c = 0; c < 3; c++) {
                    const size_t v = s->idx[3 * t + c] - vb;
                    for (int a = 0; a < 3; a++) r[6 + 3 * c + a] = orco[3 * v + a];
                }
                fl |= 4;
            }
            r[15] = (float)fl;
        }
    }
    s->built = false;
    s->gen++;
    s->geo_gen++;
    return YG_OK;
}

// Device copies of the texture-coordinate data; the orco matrix of every
// object the caller did not give one is its identity-transform matrix.
static int upload_texdata(yg_scene* s)
{
    int rc;
    const size_t nt = s->tri_mat.size();
    for (int32_t o = 0; o < s->nobj; o++) {
        if (s->obj_orco_set[o]) continue;
        size_t t0 = 0;
        while (t0 < nt && s->tri_obj[t0] != o) t0++;
        size_t t1 = t0;
        while (t1 < nt && s->tri_obj[t1] == o) t1++;
        float* out = &s->obj_orco[12 * (size_t)o];
        if (t1 == t0) continue;
        // BOX_MIN keeps the scaled matrix invertible for any mesh
        if ((rc = yg_orco_matrix(nullptr, s->verts.data(), (uint32_t)(s->verts.size() / 3), &s->idx[3 * t0],
                                 (uint32_t)(t1 - t0), out)))
            return rc;
    }
    if (s->nobj > 0 && (rc = dev_upload(&s->d_oorco, s->obj_orco.data(), s->obj_orco.size(), &s->dev_bytes)))
        return rc;
    if (!s->tri_tex.empty())
        if ((rc = dev_upload(&s->d_ptex, reinterpret_cast<const float4*>(s->tri_tex.data()), s->tri_tex.size() / 4,
                             &s->dev_bytes)))
            return rc;
    return YG_OK;
}

extern "C" int yg_build(yg_scene* s, const yg_build_opts* opts)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    int max_leaf = opts ? opts->max_leaf : 4;
    int bins = opts ? opts->sah_bins : 32;
    int threads = opts ? opts->threads : 0;
    int width = (opts && opts->width) ? opts->width : YG_DEFAULT_WIDTH;
    if (width != 2 && width != 4 && width != 8) return yg_fail(YG_ERR_ARG, "BVH width must be 2, 4 or 8");
    const int stack_lds = (opts && opts->stack_lds) ? opts->stack_lds : YG_DEFAULT_STACK_LDS;
    if (stack_lds < -1 || stack_lds > YG_MAX_STACK) return yg_fail(YG_ERR_ARG, "stack_lds must be -1 or 0..YG_MAX_STACK");
    const int rays_lane = (opts && opts->rays_per_lane) ? opts->rays_per_lane : 1;
    if (rays_lane != 1 && rays_lane != 2 && rays_lane != 4) return yg_fail(YG_ERR_ARG, "rays_per_lane must be 1, 2 or 4");
    float compact_below = (opts && opts->compact_below != 0.0f) ? opts->compact_below : 1.0f;
    if (compact_below == -1.0f) compact_below = 0.0f;  // never compact (0 < 0 is never true)
    else if (!(compact_below > 0.0f && compact_below <= 1.0f))
        return yg_fail(YG_ERR_ARG, "compact_below must be in (0, 1] or -1");
    const int builder = opts ? opts->builder : YG_BUILDER_DEVICE;
    if (builder != YG_BUILDER_HOST && builder != YG_BUILDER_DEVICE && builder != YG_BUILDER_DEVICE_PLOC)
        return yg_fail(YG_ERR_ARG, "builder must be YG_BUILDER_HOST, YG_BUILDER_DEVICE or YG_BUILDER_DEVICE_PLOC");
    const int64_t nt = (int64_t)s->tri_mat.size();
    // the device copy replaces the previous one
    auto free_old = [s]() {
        void* old[] = {s->d_nodes, s->d_nodes4, s->d_nodes8, s->d_tvn, s->d_tris, s->d_tprim, s->d_tobj,
                       s->d_tmat, s->d_obj_flags, s->d_ptex, s->d_oorco};
        for (void* p : old)
            if (p) (void)hipFree(p);
        s->d_ptex = nullptr; s->d_oorco = nullptr;
        s->d_nodes = nullptr; s->d_nodes4 = nullptr; s->d_nodes8 = nullptr; s->d_tvn = nullptr;
        s->d_tris = nullptr; s->d_tprim = nullptr; s->d_tobj = nullptr; s->d_tmat = nullptr;
        s->d_obj_flags = nullptr;
        s->dev_bytes = 0;
    };
    int rc;
    if (builder != YG_BUILDER_HOST && nt > 64) {
        free_old();
        DeviceBVHInfo di;
        if ((rc = build_bvh_device(s, builder, max_leaf, bins, width, di))) {
            free_old();
            s->built = false;
            return rc;
        }
        const int entries = (width == 2) ? di.max_depth + 1 : di.stack + 1;
        if (entries > YG_MAX_STACK) {
            free_old();
            s->built = false;
            return yg_fail(YG_ERR_STATE, "BVH deeper than the traversal stack");
        }
        if ((rc = dev_upload(&s->d_obj_flags, s->obj_flags.data(), s->obj_flags.size(), &s->dev_bytes))) return rc;
        if ((rc = upload_texdata(s))) return rc;
        s->n_nodes = di.n_nodes;
        s->max_depth = di.max_depth;
        s->width = width;
        s->stack_entries = entries;
        s->stack_lds = stack_lds < 0 ? 0 : stack_lds;
        s->rays_per_lane = rays_lane;
        s->compact_below = compact_below;
        s->built = true;
        s->gen++;
        return YG_OK;
    }
    HostBVH h;
    build_bvh(s->verts.data(), s->idx.data(), nt, s->tri_obj.data(), s->tri_mat.data(), max_leaf,
              bins, threads, width, h);
    const int entries = (width == 8) ? h.stack8 + 1 : (width == 4) ? h.stack4 + 1 : h.max_depth + 1;
    if (entries > YG_MAX_STACK) return yg_fail(YG_ERR_STATE, "BVH deeper than the traversal stack");
    free_old();
    if ((rc = dev_upload(&s->d_nodes, h.nodes.data(), h.nodes.size(), &s->dev_bytes))) return rc;
    if (width == 4 && (rc = dev_upload(&s->d_nodes4, h.nodes4.data(), h.nodes4.size(), &s->dev_bytes))) return rc;
    if (width == 8 && (rc = dev_upload(&s->d_nodes8, h.nodes8.data(), h.nodes8.size(), &s->dev_bytes))) return rc;
    if ((rc = dev_upload(&s->d_tris, h.tris.data(), h.tris.size(), &s->dev_bytes))) return rc;
    if ((rc = dev_upload(&s->d_tprim, h.tprim.data(), h.tprim.size(), &s->dev_bytes))) return rc;
    if ((rc = dev_upload(&s->d_tobj, h.tobj.data(), h.tobj.size(), &s->dev_bytes))) return rc;
    if ((rc = dev_upload(&s->d_tmat, h.tmat.data(), h.tmat.size(), &s->dev_bytes))) return rc;
    if ((rc = dev_upload(&s->d_obj_flags, s->obj_flags.data(), s->obj_flags.size(), &s->dev_bytes))) return rc;
    if ((rc = upload_texdata(s))) return rc;
    if (!s->tri_smooth.empty()) {
        // vertex normals in leaf order, with a has-normals flag per triangle
        std::vector<float4> tvn(3 * (size_t)nt);
        for (int64_t i = 0; i < nt; i++) {
            const int32_t p = h.tprim[i];
            const float* n = &s->tri_vn[9 * (size_t)p];
            tvn[3 * i] = make_float4(n[0], n[1], n[2], s->tri_smooth[p] ? 1.0f : 0.0f);
            tvn[3 * i + 1] = make_float4(n[3], n[4], n[5], 0.0f);
            tvn[3 * i + 2] = make_float4(n[6], n[7], n[8], 0.0f);
        }
        if ((rc = dev_upload(&s->d_tvn, tvn.data(), tvn.size(), &s->dev_bytes))) return rc;
    }
    if (nt == 0) {
        // keep valid (dummy) triangle arrays so kernels never see null pointers
        float4 z[3] = {};
        int32_t zi = 0;
        if ((rc = dev_upload(&s->d_tris, z, 3))) return rc;
        if ((rc = dev_upload(&s->d_tprim, &zi, 1))) return rc;
        if ((rc = dev_upload(&s->d_tobj, &zi, 1))) return rc;
        if ((rc = dev_upload(&s->d_tmat, &zi, 1))) return rc;
    }
    s->n_nodes = (width == 8) ? h.nodes8.size() : (width == 4) ? h.nodes4.size() : h.nodes.size();
    s->max_depth = h.max_depth;
    s->width = width;
    s->stack_entries = entries;
    s->stack_lds = stack_lds < 0 ? 0 : stack_lds;
    s->rays_per_lane = rays_lane;
    s->compact_below = compact_below;
    s->built = true;
    s->gen++;
    return YG_OK;
}

extern "C" int yg_scene_info(yg_scene* s, uint64_t* n_tris, uint64_t* n_nodes, uint64_t* dev_bytes,
                             int32_t* max_depth)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    if (n_tris) *n_tris = s->tri_mat.size();
    if (n_nodes) *n_nodes = s->n_nodes;
    if (dev_bytes) *dev_bytes = s->dev_bytes;
    if (max_depth) *max_depth = s->max_depth;
    return YG_OK;
}

// ------------------------------------------------------------- shading --

extern "C" int yg_set_camera(yg_scene* s, const yg_camera* cam)
{
    if (!s || !cam) return yg_fail(YG_ERR_ARG, "null argument");
    if (cam->resx <= 0 || cam->resy <= 0) return yg_fail(YG_ERR_ARG, "bad resolution");
    s->cam = *cam;
    s->has_cam = true;
    s->gen++;
    return YG_OK;
}

extern "C" int yg_set_shading(yg_scene* s, const yg_material* mats, uint32_t nmat,
                              const yg_modulator* mods, uint32_t nmod, const yg_texture* texs,
                              uint32_t ntex)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    if (nmat == 0 || !mats) return yg_fail(YG_ERR_ARG, "at least one material is required");
    for (uint32_t m = 0; m < nmat; m++)
        if (mats[m].mod_count < 0 || mats[m].mod_first < 0 ||
            (uint64_t)mats[m].mod_first + mats[m].mod_count > nmod)
            return yg_fail(YG_ERR_ARG, "material modulator range out of bounds");
    for (uint32_t k = 0; k < nmod; k++) {
        if (mods[k].tex < 0 || (uint32_t)mods[k].tex >= ntex)
            return yg_fail(YG_ERR_ARG, "modulator texture index out of bounds");
        for (int a = 0; a < 3; a++)
            if (mods[k].proj[a] < -1 || mods[k].proj[a] > 3) return yg_fail(YG_ERR_ARG, "modulator projection out of range");
        if (mods[k].texco < YG_TEXCO_GLOBAL || mods[k].texco > YG_TEXCO_REFL)
            return yg_fail(YG_ERR_ARG, "modulator texco out of range");
    }
    for (uint32_t k = 0; k < ntex; k++) {
        if (texs[k].noise < YG_NOISE_NEWPERLIN || texs[k].noise > YG_NOISE_BLENDER)
            return yg_fail(YG_ERR_ARG, "bad noise generator");
        if (texs[k].type < YG_TEX_CLOUDS || texs[k].type > YG_TEX_MUSGRAVE)
            return yg_fail(YG_ERR_ARG, "bad texture type");
        if (texs[k].type == YG_TEX_MUSGRAVE && (texs[k].mtype < YG_MUSGRAVE_FBM || texs[k].mtype > YG_MUSGRAVE_RIDGEDMF))
            return yg_fail(YG_ERR_ARG, "bad Musgrave type");
    }
    // same table sizes: overwrite the device copies in place (parameter
    // updates of a differentiable render do this every step)
    const bool same = s->d_mats && s->d_mods && s->d_texs && s->mats.size() == nmat &&
                      s->mods.size() == nmod && s->texs.size() == ntex;
    s->mats.assign(mats, mats + nmat);
    s->mods.assign(mods, mods + nmod);
    s->texs.assign(texs, texs + ntex);
    s->gen++;
    yg_modulator dm{};
    yg_texture dt{};
    const yg_modulator* mp = nmod ? mods : &dm;
    const yg_texture* tp = ntex ? texs : &dt;
    if (same) {
        // renders still in flight on any stream may read the old tables
        YG_HIP(hipDeviceSynchronize());
        YG_HIP(hipMemcpy(s->d_mats, mats, sizeof(yg_material) * nmat, hipMemcpyHostToDevice));
        YG_HIP(hipMemcpy(s->d_mods, mp, sizeof(yg_modulator) * (nmod ? nmod : 1), hipMemcpyHostToDevice));
        YG_HIP(hipMemcpy(s->d_texs, tp, sizeof(yg_texture) * (ntex ? ntex : 1), hipMemcpyHostToDevice));
        return YG_OK;
    }
    int rc;
    if ((rc = dev_upload(&s->d_mats, mats, nmat))) return rc;
    if ((rc = dev_upload(&s->d_mods, mp, nmod ? nmod : 1))) return rc;
    if ((rc = dev_upload(&s->d_texs, tp, ntex ? ntex : 1))) return rc;
    return YG_OK;
}

extern "C" int yg_set_lights(yg_scene* s, const yg_light* lights, uint32_t n)
{
    if (!s) return yg_fail(YG_ERR_ARG, "null scene");
    for (uint32_t i = 0; i < n; i++) {
        const yg_light& L = lights[i];
        if (L.type < YG_LIGHT_SUN || L.type > YG_LIGHT_GLOBALPHOTON) return yg_fail(YG_ERR_ARG, "bad light type");
        if (L.type == YG_LIGHT_AREA && L.psamples < 0) return yg_fail(YG_ERR_ARG, "area light psamples must be >= 0");
        if (L.type == YG_LIGHT_PATH && !L.use_qmc)
            return yg_fail(YG_ERR_UNSUPPORTED, "pathlight needs use_QMC=on (the LCG sampler is not implemented)");
        if (L.type == YG_LIGHT_PHOTON && !L.use_qmc)
            return yg_fail(YG_ERR_UNSUPPORTED, "photonlight needs use_QMC=on");
        if ((L.type == YG_LIGHT_PATH || L.type == YG_LIGHT_AREA) && L.samples < 1)
            return yg_fail(YG_ERR_ARG, "samples must be >= 1");
        if (L.type == YG_LIGHT_PATH && (L.depth < 0 || L.depth > 8))
            return yg_fail(YG_ERR_UNSUPPORTED, "pathlight depth must be in 0..8");
    }
    s->lights.assign(lights, lights + n);
    s->pm.light = -1;
    s->gp.light = -1;
    s->gen++;
    return YG_OK;
}

// ----------------------------------------------------------------- QMC --

extern "C" int yg_halton(int32_t base, uint32_t i0, uint32_t n, float* d_out, void* stream)
{
    if (base < 2 || (!d_out && n)) return yg_fail(YG_ERR_ARG, "bad halton arguments");
    if (n == 0) return YG_OK;
    std::vector<float> h;
    halton_incremental(base, (uint64_t)i0 + n, h);
    YG_HIP(hipMemcpyAsync(d_out, h.data() + i0, sizeof(float) * n, hipMemcpyHostToDevice,
                          (hipStream_t)stream));
    YG_HIP(hipStreamSynchronize((hipStream_t)stream));
    return YG_OK;
}
