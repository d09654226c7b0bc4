/*
 * yaf_oracle.h -- CPU restatement of YafRay 0.0.9's ray-intersection / GI path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for libyafgpu.so:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product path never calls it.
 *
 * Every function restates a reference function (file:line cited at the
 * definition) in plain C99, FP32 where the reference is FP32 and double where
 * the reference computes in double.  Built with -ffp-contract=off and without
 * -ffast-math so each expression is evaluated exactly as written, left to right.
 */
#ifndef YAF_ORACLE_H
#define YAF_ORACLE_H

#include <stdint.h>
#include "../include/yafgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z; } yo_v3;

/* ------------------------------------------------------------ triangle -- */
void yo_tri_normal(const float* a, const float* b, const float* c, float* n);
int yo_tri_hit(const float* a, const float* b, const float* c, const float* n,
               const float* from, const float* ray);
float yo_tri_intersect(const float* a, const float* n, const float* from, const float* ray);
void yo_tri_bary(const float* a, const float* b, const float* c, const float* n,
                 const float* h, float* fb, float* fc);

/* --------------------------------------------------------------- scene -- */
typedef struct yo_scene yo_scene;

/* verts: nv*3, idx: nt*3, tri_obj/tri_mat: nt (may be NULL -> 0). */
yo_scene* yo_scene_new(const float* verts, int64_t nv, const uint32_t* idx, int64_t nt,
                       const int32_t* tri_obj, const int32_t* tri_mat);
void yo_scene_free(yo_scene* s);
/* meshObject_t::autoSmooth (mesh.cc:199-282) of object obj's triangles */
int yo_mesh_autosmooth(yo_scene* s, int obj, float angle);
/* meshObject_t::transform's orco matrix (mesh.cc:422-493): M (4x4 row-major,
 * NULL = identity) scaled by half the untransformed triangle bound's extent,
 * translation = centre of the transformed bound, inverted; rows 0..2 */
int yo_orco_matrix(const float* m16, const float* xyz, int64_t nv, const uint32_t* idx, int64_t nt,
                   float* out12);
/* texture-coordinate data of object obj's triangles (in scene order):
 * face_uv 6 per triangle or NULL, face_flags per triangle (bit0 uv, bit1
 * vertex colours; NULL: all have uv when face_uv is given), corner_orco 9
 * per triangle (the orco points of a, b, c) or NULL, orco_mtx 12 or NULL
 * (keep the identity-transform matrix yo_scene_new computed) */
int yo_mesh_texcoords(yo_scene* s, int obj, const float* face_uv, const uint8_t* face_flags,
                      const float* corner_orco, const float* orco_mtx);
/* getSurface's shading normal of n hits (prim, hit point) */
void yo_shading_normals(yo_scene* s, const int32_t* prim, const float* h, int64_t n, float* out);
/* canonical BVH2 node count (binned SAH, 32 bins, leaf <= 4, FP64 boxes) */
int64_t yo_scene_nodes(yo_scene* s);

/* brute force over every triangle; mode 0 closest (tmin <= t < tmax),
 * 1 any (tmin < t < tmax).  rays: n*8 floats in yg_ray layout. */
void yo_trace_brute(yo_scene* s, const float* rays, int64_t n, int mode,
                    int32_t* prim, float* t, float* u, float* v);
/* same results through the canonical BVH2; counts node visits / tri tests */
void yo_trace_bvh(yo_scene* s, const float* rays, int64_t n, int mode, int32_t* prim,
                  float* t, float* u, float* v, uint64_t* nodes, uint64_t* tris,
                  int nthreads);

/* scene_t::firstHit (closest, Z > 0, origin from+ray*min_raydis) */
void yo_first_hit(yo_scene* s, const float* from, const float* dir, int64_t n,
                  float min_raydis, int32_t* prim, float* t);
/* scene_t::isShadowed, point form (to_point=1) or direction form */
void yo_is_shadowed(yo_scene* s, const float* p, const float* l, const int32_t* self_obj,
                    int64_t n, int to_point, float min_raydis, float self_bias,
                    uint8_t* out);

/* ------------------------------------------------------------------ QMC -- */
/* n successive Halton::getNext() values of a fresh Halton(base) (mcqmc.h:35-49) */
/* query log of yo_first_hit_one / yo_is_shadowed_one (8 floats per record) */
void yo_log_begin(float* ext, int64_t ext_cap, float* sh, int64_t sh_cap);
void yo_log_end(int64_t* ext_n, int64_t* sh_n);

void yo_halton_seq(int base, int64_t n, float* out);
/* index-addressed value via Halton::setStart's digit loop (mcqmc.h:25-33) */
float yo_halton_at(int base, uint32_t i);
float yo_ri_vdc(uint32_t bits, uint32_t r);
float yo_ri_s(uint32_t i, uint32_t r);
float yo_ri_lp(uint32_t i, uint32_t r);
int yo_next_prime(int p);
/* ourRandom(): advances *seed, returns the float (vector3d.h:288-297) */
float yo_our_random(int32_t* seed);
/* seed after n steps via a^n mod m */
int32_t yo_lcg_jump(int32_t seed, uint64_t n);
/* glibc TYPE_3 rand(): n outputs after srand(seed) */
void yo_glibc_rand(uint32_t seed, int64_t n, int32_t* out);

/* ---------------------------------------------------------------- noise -- */
void yo_perm(unsigned char* out512);
float yo_newperlin(float x, float y, float z);
float yo_stdperlin(float x, float y, float z);
float yo_blendernoise(float x, float y, float z);
/* the 256 unit gradients (768 floats) of stdPerlin_t / blenderNoise_t */
void yo_grad_table(float* out768);
float yo_fbm(float x, float y, float z, float H, float lac, float oct);
float yo_turbulence(float x, float y, float z, int oct, float size, int hard);
/* what: YG_EVAL_* ; out gets 1 float (3 for TEX_COLOR) per point */
int yo_noise_eval(int what, const yg_texture* tex, const float* pts, int64_t n, float* out);
void yo_gen_heightfield(int nx, int ny, float ex, float ey, float amp, float freq, float H,
                        float lac, float oct, float* xyz);
/* genericShader_t::getDiffuse (basicshaders.cc:191-204) of material `mat` at
 * n points: its colour passed through its modulators in order. out: 3 per point */
void yo_diffuse_at(const yg_material* mats, const yg_modulator* mods, const yg_texture* texs, int mat,
                   const float* pts, int64_t n, float* out);

/* ------------------------------------------------------------ rendering -- */
/* scene_t::render(colorOutput_t&) mono mode, AA_passes 0 (scene.cc:538-600,
 * 724-775).  rgb: resx*resy*3.  stats: [camera, extension, shadow, extra,
 * lcg_draws] ray counts.  Photon lights need yo_photon_build first. */
int yo_render(yo_scene* s, const yg_camera* cam, const yg_material* mats, int nmat,
              const yg_modulator* mods, int nmod, const yg_texture* texs, int ntex,
              const yg_light* lights, int nl, const int32_t* obj_flags, int nobj,
              const yg_render_opts* opts, float* rgb, uint64_t* stats);

/* same, areas spread over nthreads (only without area lights) */
int yo_render_mt(yo_scene* s, const yg_camera* cam, const yg_material* mats, int nmat,
                 const yg_modulator* mods, int nmod, const yg_texture* texs, int ntex,
                 const yg_light* lights, int nl, const int32_t* obj_flags, int nobj,
                 const yg_render_opts* opts, float* rgb, uint64_t* stats, int nthreads);

/* blockSpliter_t regions (renderblock.cc:71-101) in render order, 8 ints each */
int yo_block_regions_into(int width, int height, int b, uint32_t seed, int32_t* out, int cap);

/* raytrace of camera rays through camera-plane points (px, py); sun lights only */
int yo_trace_camera(yo_scene* s, const yg_camera* cam, const yg_material* mats, int nmat,
                    const yg_modulator* mods, const yg_texture* texs, const yg_light* lights, int nl,
                    const int32_t* obj_flags, int nobj, const yg_render_opts* opts, const float* px,
                    const float* py, int64_t n, float* rgb);

/* FP64 reference gradient of a loss through the image (single thread):
 * gmat nmat*3, gmod nmod*2 (colour, size scale), glight nl (power) */
int yo_render_adjoint(yo_scene* s, const yg_camera* cam, const yg_material* mats, int nmat,
                      const yg_modulator* mods, int nmod, const yg_texture* texs, int ntex,
                      const yg_light* lights, int nl, const int32_t* obj_flags, int nobj,
                      const yg_render_opts* opts, const float* grad_rgb, float* rgb, double* gmat,
                      double* gmod, double* glight, int nthreads);

/* FP64 render of camera rays through pixel centres (yo_fp64.c): sun, area
 * and path lights in double precision.  mode YO_FP64_REFERENCE: the whole
 * image with the reference's sample streams (rgb: resx*resy*3; px, py, n,
 * samples, seed unused); YO_FP64_CONVERGED: the n pixels (px[q], py[q]) with
 * `samples` samples of every sampled light (rgb: n*3).  Returns 0, -2 bad
 * arguments, -5 AA or tone mapping, -6 photon
 * lights or a path light used in indirect lighting, -7 several path lights. */
#define YO_FP64_REFERENCE 0
#define YO_FP64_CONVERGED 1
int yo_render_fp64(yo_scene* s, const yg_camera* cam, const yg_material* mats, int nmat,
                   const yg_modulator* mods, const yg_texture* texs, const yg_light* lights, int nl,
                   const int32_t* obj_flags, int nobj, const yg_render_opts* opts, int mode,
                   int samples, const int32_t* px, const int32_t* py, int64_t n, uint64_t seed,
                   double* rgb, int nthreads);

/* photonLight_t::init for light li; keeps the clustered photons in s */
int yo_photon_build(yo_scene* s, const yg_material* mats, int nmat, const yg_modulator* mods,
                    int nmod, const yg_texture* texs, int ntex, const yg_light* lights,
                    int nl, int li, const int32_t* obj_flags, int nobj,
                    const yg_render_opts* opts, uint64_t* stats);
/* copy clustered photons (9 floats each: dir, pos, colour) */
int64_t yo_photon_read(yo_scene* s, float* out, int64_t cap);
/* Installs a clustered photon map (n marks of 9 floats, e.g. read back from
 * the GPU) as photon light L's map, with its gather grid, so that renders
 * and gathers use it instead of a map yo_photon_build made. */
int yo_photon_load(yo_scene* s, const float* data, int64_t n, const yg_light* L);

/* globalPhotonLight_t::init (globalphotonlight.cc:234-278) for light li, one
 * photon after another from the ourRandom() stream at opts->lcg_seed; a
 * later yo_render continues that stream.  stateless: every gather starts at
 * the light's radius (no radius carried between calls).  stats: emitted,
 * stored, boxes, rays, draws. */
int yo_gphoton_build(yo_scene* s, const yg_material* mats, int nmat, const yg_modulator* mods,
                     const yg_texture* texs, const yg_light* lights, int nl, const int32_t* obj_flags,
                     int nobj, const yg_render_opts* opts, int li, int stateless, uint64_t* stats);
/* boxes: 15 floats each (position, N, irradiance, merged direction, merged colour) */
int64_t yo_gphoton_read(yo_scene* s, float* out, int64_t cap);
/* the raw photon map in store order, 5 words per photon: position (3 floats),
 * direction code (theta | phi << 8; theta 255: none), RGBE (r | g << 8 | b << 16 | e << 24) */
int64_t yo_gphoton_read_photons(yo_scene* s, uint32_t* out, int64_t cap);
/* replace the raw photon map by another implementation's (same 5-word format) */
int yo_gphoton_load_photons(yo_scene* s, const uint32_t* data, int64_t n);
/* photonSampler_t's sector plan (pathtools.cc:172-236) for one shading point:
 * hits[sectors], weight[sectors], multi; the gather is stateless (radius = maxradius).
 * Returns the number of sectors, or -1 without a raw map. */
int yo_psamp_plan(yo_scene* s, const float* P, const float* N, const float* Ru, const float* Rv, int samples,
                  int grid, int32_t* hits, float* weight, float* multi, int32_t* nfound);

/* Irradiance cache (pathlight cache = on) of the last yo_render*: its
 * samples, 18 floats each (P, N, pP, colour, adist, M, precision, rank,
 * visit, devaluated); *new_samples: samples its render pass had to take.
 * yo_cache_release frees it (yo_scene_free does). */
int64_t yo_cache_read(yo_scene* s, float* out, int64_t cap, uint64_t* new_samples);
void yo_cache_release(yo_scene* s);
void yo_cache_record(yo_scene* s, int on);
int64_t yo_cache_visits(yo_scene* s, float* out, int64_t cap);
int64_t yo_heap_offer_trace(const float* w, int64_t n, unsigned limit, float* out);
void yo_grand_stream(uint32_t seed, int64_t skip, int64_t n, int32_t* out);
int yo_gphoton_load(yo_scene* s, const float* data, int64_t n, const yg_light* GL, int li, int stateless,
                    int recompute);
void yo_rgbe_roundtrip(const float* in, float* out, int64_t n);
void yo_dir_roundtrip(const float* in, float* out, int32_t* code, int64_t n);
void yo_dirconv_tables(float* out);
/* photonLight_t::illuminate at explicit shading points (see yg_photon_gather) */
void yo_photon_gather(yo_scene* s, const yg_material* mats, const yg_modulator* mods,
                      const yg_texture* texs, const yg_light* light, const float* p,
                      const float* n, const float* ng, const float* eye, const int32_t* mat,
                      int64_t cnt, float* rgb);

#ifdef __cplusplus
}
#endif
/* Canonical BVH2 visit counts of the roofline byte model (yo_bytemodel.c):
 * builds a tight FP64 binned-SAH BVH2 (32 bins, leaf <= 4) over the
 * triangles, traces n rays (8 floats: o, tmin, d, tmax) closest (mode 0) or
 * any hit (mode 1), and returns the node count; *visits / *tris /
 * *boxes_tested get the summed node visits, triangle tests and box tests. */
int64_t yo_canonical_counts(const float* verts, int64_t nv, const uint32_t* idx, int64_t nt, const float* rays,
                            int64_t n, int mode, int threads, uint64_t* visits, uint64_t* tris,
                            uint64_t* boxes_tested);

#endif
