/*
 * yafgpu.h -- C ABI of libyafgpu.so, the MI355X (gfx950) ray-intersection and
 * global-illumination engine that drops in for YafRay 0.0.9's per-ray CPU calls.
 *
 * Conventions (mirroring src/interface/interface_impl.cc:141-148 and
 * src/yafraycore/scene.cc:69-78 of the reference):
 *   - every entry point returns YG_OK (0) or a negative status; no exceptions
 *     cross this boundary and nothing calls exit();
 *   - yg_last_error() returns the text of the last failure on the calling thread;
 *   - the caller owns host buffers, a scene handle owns its device buffers;
 *   - buffers named d_* are device (HBM) pointers, all device work is ordered
 *     on the HIP stream passed in (NULL = the null stream);
 *   - a handle must not be used from two threads at once without external locking.
 *
 * Arithmetic is FP32 like the reference (PFLOAT = float, configio.py:135-140)
 * except where the reference itself computes in double (Halton state,
 * mcqmc.h:35-49; sample-direction trigonometry, pathtools.cc:32-43).
 */
#ifndef YAFGPU_H
#define YAFGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YG_OK 0
#define YG_ERR_ARG (-1)
#define YG_ERR_HIP (-2)
#define YG_ERR_STATE (-3)
#define YG_ERR_UNSUPPORTED (-4)
#define YG_ERR_NOMEM (-5)

#define YG_ABI_VERSION 14

/* ---------------------------------------------------------------- rays -- */

/* One traversal query, 32 B.  Closest-hit accepts tmin <= t < tmax
 * (kdTree_t::Intersect, kdtree.cc:690-694 with tmin = 0, tmax = dist);
 * any-hit accepts tmin < t < tmax (kdTree_t::IntersectS, kdtree.cc:835-841). */
typedef struct yg_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} yg_ray;

/* Closest-hit result, 16 B.  prim = index of the triangle in mesh-add order
 * (-1 on miss); t = plane distance (triangle_t::intersect, triangle.h:83-86);
 * u, v = barycentric weights of vertices b and c from the dominant-axis
 * projection of triangle.cc:74-136 (fb, fc). */
typedef struct yg_hit {
    int32_t prim;
    float t;
    float u;
    float v;
} yg_hit;

/* ---------------------------------------------------------- descriptors -- */

/* Perspective pinhole camera; camera_t::camera_t (camera.cc:27-87). */
typedef struct yg_camera {
    float pos[3];
    float look[3];
    float up[3];     /* a point, like the reference ("up" - pos is the up vector) */
    int32_t resx, resy;
    float aspect;    /* render param "aspect_ratio" (1 by default) */
    float focal;     /* render param "focal" (df) */
} yg_camera;

enum {
    YG_NOISE_NEWPERLIN = 0, /* newPerlin_t, noise.cc:161-183 */
    YG_NOISE_STDPERLIN = 1, /* stdPerlin_t, noise.cc:372-435 */
    YG_NOISE_BLENDER = 2    /* blenderNoise_t, noise.cc:439-509 (turbulence offsets its point by 1) */
};

enum {
    YG_TEX_CLOUDS = 1,   /* textureClouds_t, basictex.cc:70-115 */
    YG_TEX_MARBLE = 2,   /* textureMarble_t, basictex.cc:121-177 */
    YG_TEX_WOOD = 3,     /* textureWood_t, basictex.cc:183-230 */
    YG_TEX_MUSGRAVE = 4  /* textureMusgrave_t with fBm_t, basictex.cc:657-709 */
};

enum { YG_SHAPE_SIN = 0, YG_SHAPE_SAW = 1, YG_SHAPE_TRI = 2 };

/* Procedural texture; fields not used by a type are ignored. */
typedef struct yg_texture {
    int32_t type;
    int32_t noise;       /* YG_NOISE_* */
    float color1[3];
    float color2[3];
    int32_t depth;       /* clouds "depth", marble/wood "depth" (octaves) */
    float size;          /* turbulence "size" (clouds/marble/wood), Musgrave "size" */
    int32_t hard;
    int32_t bias;        /* clouds: 0 none, 1 positive, 2 negative */
    float turb;          /* marble/wood "turbulence" */
    float sharpness;     /* marble: 1/shp if shp > 1 else 1 (basictex.cc:125-126) */
    int32_t shape;       /* YG_SHAPE_* */
    int32_t rings;       /* wood "type" == rings */
    float H, lacunarity, octaves, iscale; /* Musgrave parameters */
    int32_t mtype;       /* Musgrave "type": YG_MUSGRAVE_* (textureMusgrave_t ctor, basictex.cc:657-672) */
    float offset, gain;  /* Musgrave "offset", "gain" (heteroterrain, hybridmf, ridgedmf) */
} yg_texture;

enum {
    YG_MUSGRAVE_FBM = 0,          /* fBm_t (noise.cc:638-650) */
    YG_MUSGRAVE_MULTIFRACTAL = 1, /* mFractal_t (noise.cc:668-681) */
    YG_MUSGRAVE_HETEROTERRAIN = 2,/* heteroTerrain_t (noise.cc:695-719) */
    YG_MUSGRAVE_HYBRIDMF = 3,     /* hybridMFractal_t (noise.cc:731-754) */
    YG_MUSGRAVE_RIDGEDMF = 4      /* ridgedMFractal_t (noise.cc:768-791) */
};

enum { YG_MOD_MIX = 0, YG_MOD_MUL = 1, YG_MOD_ADD = 2, YG_MOD_SUB = 3 };

/* Texture coordinates of a modulator (modulator_t::doMapping,
 * texture.cc:125-165; names as string2texcotype, texture.h:135-149). */
enum {
    YG_TEXCO_GLOBAL = 0, /* "global": the hit point P */
    YG_TEXCO_ORCO = 1,   /* "orco" (the factory default): the mesh's interpolated
                            orco points, or the object's orco matrix times P */
    YG_TEXCO_UV = 2,     /* "uv" (any other name): (2u - 1, 1 - 2v, 0) */
    YG_TEXCO_WIN = 3,    /* "window": the camera sample's screen position */
    YG_TEXCO_NOR = 4,    /* "normal": the shading normal */
    YG_TEXCO_REFL = 5    /* "reflect": the eye vector reflected about the
                            face-forwarded shading normal */
};

/* modulator_t on a procedural texture (modulator_t::doMapping,
 * texture.cc:122-317, procedural branch): the texture coordinates T of
 * `texco`, times the texture matrix when there is one (texpt = M * T; no
 * translation for "normal" and "reflect"), then per output axis a:
 * texpt[a] = size[a] * (T[proj[a]] + ofs[a]); colour blend by mode and
 * factor "color" (texture.cc:499-534). */
typedef struct yg_modulator {
    int32_t tex;         /* index into the texture table */
    int32_t mode;        /* YG_MOD_* */
    float color;         /* blend factor ("color" parameter) */
    float size[3];       /* sizex, sizey, sizez (each times "size") */
    float ofs[3];
    int32_t proj[3];     /* proj_x/y/z: 0 = the same axis, 1..3 = x, y, z, -1 = none ("n") */
    int32_t has_mtx;     /* 0: no texture matrix (identity) */
    float mtx[12];       /* rows 0..2 of the texture matrix m00..m23, row-major */
    int32_t texco;       /* YG_TEXCO_* */
} yg_modulator;

/* genericShader_t in its diffuse form (basicshaders.cc:30-70,191-204):
 * scolor = albedo, edif = 1, ks = 0, no reflection/refraction. */
typedef struct yg_material {
    float color[3];
    int32_t mod_first;
    int32_t mod_count;
} yg_material;

enum {
    YG_LIGHT_SUN = 1,    /* sunLight_t, sunlight.cc:27-57 */
    YG_LIGHT_AREA = 2,   /* areaLight_t, arealight.cc:31-174 */
    YG_LIGHT_PATH = 3,   /* pathLight_t full-GI, non-cached (pathlight.cc:107-403) */
    YG_LIGHT_PHOTON = 4, /* photonLight_t, diffuse mode (photonlight.cc:33-501) */
    YG_LIGHT_GLOBALPHOTON = 5 /* globalPhotonLight_t (globalphotonlight.cc:19-287): builds the
                                 irradiance map the pathlight's path vertices read
                                 (pathlight.cc:140-185); photons from "dummy" area lights */
};

/* One light; lights are summed in the order given, which must be the
 * reference's order (alphabetical by light name, interface_impl.cc:571-576). */
typedef struct yg_light {
    int32_t type;
    int32_t use_in_render;   /* light_t::useInRender */
    int32_t use_in_indirect; /* light_t::useInIndirect (pathlight sets false) */
    float color[3];
    float power;
    /* sun */
    float dir[3];            /* "from", interpreted as a direction */
    int32_t cast_shadows;
    /* area */
    float a[3], b[3], c[3], d[3];
    int32_t samples;         /* area "samples", path "samples" */
    int32_t psamples;        /* area "psamples": guessLight's test draws of glibc rand() (0: every
                                lit point samples the penumbra; > 0 needs the light among the
                                first 16, outside GI paths, no antialiasing, no batching) */
    /* path */
    int32_t depth;           /* path "depth" / photon "depth" */
    int32_t use_qmc;         /* path/photon "use_QMC" */
    float threshold;         /* path "threshold" */
    /* photon */
    float from[3], to[3];
    float angle;
    int32_t photons;
    int32_t search;          /* K */
    int32_t mindepth;
    float bias;
    float fixedradius;       /* photon "fixedradius"; global photon "radius" */
    float cluster;
    int32_t dummy;           /* area "dummy": shoots global photons only, no direct light */
    /* pathlight irradiance cache (pathlight.cc:405-723, lightcache.cc, cacheproxy.cc);
       "search" above is the cache's search (at least 3, pathlight.cc:795) */
    int32_t cache;           /* path "cache" */
    float cache_size;        /* path "cache_size": box edge in (screen x, screen y, log
                                distance) space; samples cover within it, gathers reach twice it */
    float shadow_threshold;  /* path "shadow_threshold": a sample covers a point while its
                                weight is at least 1 / shadow_threshold */
} yg_light;

/* Render parameters (interface_impl.cc:480-644 / scene.cc:50-60). */
typedef struct yg_render_opts {
    float min_raydis;        /* MIN_RAYDIST, 5e-5 (configio.py:95-100) */
    float self_bias;         /* render param "bias" */
    int32_t raydepth;        /* maxraylevel */
    float background[3];     /* constBackground_t colour */
    int32_t has_background;
    int32_t block;           /* render area edge, 64 (scene.cc:541) */
    uint32_t rand_seed;      /* srand() seed before the block shuffle (glibc default 1) */
    uint32_t lcg_seed;       /* ourRandom() state at render start (vector3d.cc:160: 123212) */
    int32_t flags;           /* YG_RENDER_* */
    /* adaptive antialiasing of scene_t::render(renderArea_t&) (scene.cc:740-816),
     * with the ranges interface_impl.cc:501-533 enforces; aa_passes = 0: off */
    int32_t aa_passes;       /* AA_passes */
    int32_t aa_minsamples;   /* AA_minsamples, >= 1 when aa_passes > 0 */
    float aa_pixelwidth;     /* AA_pixelwidth, 1..2 */
    float aa_threshold;      /* AA_threshold */
    int32_t aa_jitterfirst;  /* AA_jitterfirst */
    /* tone mapping of every raytraced colour (scene.cc:763 with color.h:175-195);
     * the reference turns it on for 8-bit (targa) output, off for hdr */
    int32_t tonemap;
    float exposure;          /* render param "exposure", 0 = off */
    float gamma;             /* render param "gamma", 1 = off */
    int32_t clamp_rgb;       /* render param "clamp_rgb" */
} yg_render_opts;

#define YG_RENDER_TIMING 1   /* fill yg_render_stats.ms with device time */
#define YG_RENDER_KTIMING 2  /* also time each GI bounce launch (ms_bounce) */
#define YG_RENDER_REUSE 4    /* yg_render_adjoint: replay the last yg_render of this scene
                                if nothing changed since, instead of rendering again */
#define YG_RENDER_NO_GRAPH 8 /* launch every kernel of the frame from the host instead of
                                replaying the captured frame graph (see yg_render) */
#define YG_RENDER_ROULETTE 16 /* Russian roulette from the third bounce on: a path continues
                                 with probability q = min(1, energy of its throughput) and is
                                 divided by q.  Not in the reference (pathlight.cc:313-315 ends
                                 paths only at energy < 0.05 or the depth): images differ from
                                 it; rejected by yg_render_adjoint */

typedef struct yg_render_stats {
    uint64_t rays_camera;
    uint64_t rays_extension;
    uint64_t rays_shadow;
    uint64_t adaptive_extra;  /* pixels whose pathlight asked for extra repetitions */
    uint64_t lcg_draws;       /* ourRandom() draws consumed (area lights) */
    float ms;                 /* whole render, device time */
    float ms_trace;           /* up to the final pixel pass */
    float ms_bounce;          /* sum of the GI bounce launches (YG_RENDER_KTIMING) */
    uint32_t bounce_launches; /* GI levels traced */
    uint64_t rays_bounce;     /* extension + shadow rays traced inside the bounce launches */
    float ms_ext;             /* sum of the extension-ray (closest hit) launches */
    float ms_shadow;          /* sum of the bounce shadow-ray (any hit) launches */
    int32_t graph;            /* 1: the frame ran as a replay of its captured hipGraph,
                                 2: it was captured by this call, 0: launched kernel by kernel */
} yg_render_stats;

typedef struct yg_build_opts {
    int32_t max_leaf;      /* triangles per leaf, 1..8 */
    int32_t sah_bins;      /* 8..64 */
    int32_t threads;       /* host build threads, 0 = all */
    int32_t width;         /* BVH node width: 2, 4 or 8 (0 = default, 4) */
    int32_t stack_lds;     /* traversal stack entries per lane kept in LDS by the render
                              launches, deeper ones in device memory (0 = default 16,
                              -1 = the whole stack in LDS) */
    int32_t rays_per_lane; /* rays each lane of the GI level launches (extension and shadow
                              rays) traces one after another: 64 x this per wave (0 = 1; 1, 2, 4) */
    float compact_below;   /* a GI level's surviving paths are compacted into a dense list
                              for the next level when fewer than this fraction of the
                              level's listed paths survive; otherwise the next level reuses
                              the list and its ended paths' lanes exit at once
                              (0 = default 1: always compact; 0 < x <= 1, or -1 = never) */
    int32_t builder;       /* YG_BUILDER_DEVICE (0, default): binned SAH on the GPU, level by
                              level (sah_bins at most 64), the host builder's tree;
                              YG_BUILDER_HOST: binned SAH on host threads;
                              YG_BUILDER_DEVICE_PLOC: locally-ordered clustering on the GPU
                              (fastest build, trees that trace more slowly);
                              scenes of at most 64 triangles are built on the host */
} yg_build_opts;

#define YG_BUILDER_DEVICE 0
#define YG_BUILDER_HOST 1
#define YG_BUILDER_DEVICE_PLOC 2

typedef struct yg_scene yg_scene;

/* ------------------------------------------------------------ lifecycle -- */

int yg_abi_version(void);
int yg_last_error(char* buf, size_t len);
int yg_device_count(int* n);

/* interface_impl.cc:51 (interfaceImpl_t ctor) */
int yg_scene_create(yg_scene** out);
/* interface_impl.cc:58-91 (~interfaceImpl_t / clear) */
int yg_scene_destroy(yg_scene* s);

/* addObject_trimesh (interface_impl.cc:193-265): world-space float3 vertices,
 * uint32 index triples, optional per-face material ids (NULL: all 0).
 * Primitive ids continue across calls in add order; *obj_id receives the
 * object index (objects are numbered in add order, which must be the
 * reference's alphabetical object order, render.cc:1158-1164).
 * flags: bit0 = cast shadows, bit1 = receive radiosity, bit2 = use for radiosity
 * (object3d_t flags; 7 = reference defaults). */
int yg_mesh_add(yg_scene* s, const float* xyz, uint32_t nv, const uint32_t* idx,
                uint32_t nt, const int32_t* mat, int32_t flags, int32_t* obj_id);

/* meshObject_t::autoSmooth (mesh.cc:199-282), called by the scene loader for a
 * mesh with autosmooth="angle" (render.cc:764): vertex normals for object
 * obj_id from the face normals of its triangles that share the vertex and
 * deviate by less than `angle` degrees (at 180 or more: all of them; below 1:
 * nothing).  Hits on the object then get the interpolated, normalised
 * shading normal of triangle_t::getSurface (triangle.cc:142-160) and the
 * tangent frame around it.  Call after yg_mesh_add, before yg_build. */
int yg_mesh_autosmooth(yg_scene* s, int32_t obj_id, float angle);

/* Texture-coordinate data of mesh obj_id (faces as join_face / join_lface read
 * them, msin_object.cc:116-230, and render.cc:743-764; addObject_trimesh,
 * interface_impl.cc:229-238), read by modulators with texco "uv" and "orco":
 *   face_uv     6 floats per triangle of the object (u_a v_a u_b v_b u_c v_c),
 *               or NULL;
 *   face_flags  per triangle: bit0 = the face has uv, bit1 = it has vertex
 *               colours (NULL: every face has uv when face_uv is given, none
 *               has vertex colours).  A triangle with neither, no autoSmooth
 *               normals and no orco points has u = v = -1 (triangle.cc:140),
 *               else u = v = 0 unless it has uv;
 *   orco        3 floats per vertex of the object (has_orco="on": the orco
 *               point stored after each vertex), or NULL;
 *   orco_mtx    rows 0..2 of meshObject_t's backOrco (toObjectOrco, used when
 *               orco is NULL), from yg_orco_matrix; NULL: the matrix of an
 *               object with the identity transform, computed from its
 *               vertices.
 * Vertex colours are not stored: the generic shader does not read them.
 * Call after yg_mesh_add (and yg_mesh_autosmooth), before yg_build. */
int yg_mesh_texcoords(yg_scene* s, int32_t obj_id, const float* face_uv, const uint8_t* face_flags,
                      const float* orco, const float* orco_mtx);

/* meshObject_t::transform's orco matrix (mesh.cc:422-493): M (4x4 row-major
 * object transform, render_t::M; NULL = identity) scaled by half the extent of
 * the untransformed bound of the triangles' vertices xyz (recalcBound with
 * BOX_MIN, mesh.cc:343-380), its translation replaced by the centre of the
 * transformed bound, inverted (matrix4x4_t::inverse, matrix4.cc:78-112); rows
 * 0..2 into out12.  Vertices in float as the reference: M * p. */
int yg_orco_matrix(const float* m16, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                   float* out12);

/* kdTree_t::kdTree_t (kdtree.cc:57-138) + buildObjectTree (object3d.cc:30-53):
 * builds the SAH BVH over all meshes and uploads it. */
int yg_build(yg_scene* s, const yg_build_opts* opts);

/* Node / triangle counts and device bytes of the built tree. */
int yg_scene_info(yg_scene* s, uint64_t* n_tris, uint64_t* n_nodes, uint64_t* dev_bytes,
                  int32_t* max_depth);

/* -------------------------------------------------------- ray queries -- */

/* kdTree_t::Intersect over every object (kdtree.cc:591-727): closest hit. */
int yg_trace_closest(yg_scene* s, const yg_ray* d_rays, uint32_t n, yg_hit* d_hits,
                     void* stream);
/* kdTree_t::IntersectS (kdtree.cc:730-866): any hit; d_occluded gets 0/1 per ray. */
int yg_trace_any(yg_scene* s, const yg_ray* d_rays, uint32_t n, uint8_t* d_occluded,
                 void* stream);

/* yg_trace_any without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid it).  No counterpart in the reference.  Same results as
 * yg_trace_any; the traversal stack entries beyond the LDS part go to a
 * 64 MiB scratch array that the library holds on each GPU until the process
 * ends, instead of a buffer it allocates.  The launch shrinks its grid (each
 * thread then traces several rays) until the scratch holds every thread's
 * entries, so no ray count is refused; a tree whose stack does not fit the
 * scratch even for one block of 64 lanes returns YG_ERR_UNSUPPORTED (deeper
 * than any tree yg_build accepts).  The scratch keeps the last call's
 * contents and every call on one GPU shares it: calls that overlap on one
 * GPU (other streams or threads) race and must be serialised by the caller.
 * The first call on a device looks up the scratch's address there. */
int yg_trace_any_noalloc(yg_scene* s, const yg_ray* d_rays, uint32_t n, uint8_t* d_occluded,
                         void* stream);

/* scene_t::firstHit (scene.cc:688-721): origin from + ray*min_raydis, closest
 * with Z > 0.  d_from, d_dir: n float3.  Writes prim/t/u/v per query. */
int yg_first_hit(yg_scene* s, const float* d_from, const float* d_dir, uint32_t n,
                 float min_raydis, yg_hit* d_hits, void* stream);

/* scene_t::isShadowed (scene.cc:107-146 point form when to_point != 0,
 * scene.cc:148-185 direction form otherwise): d_p = shading points,
 * d_l = light point or direction, d_self_obj = object of the shading point
 * (or -1).  Writes 0/1. */
int yg_is_shadowed(yg_scene* s, const float* d_p, const float* d_l,
                   const int32_t* d_self_obj, uint32_t n, int32_t to_point,
                   float min_raydis, float self_bias, uint8_t* d_out, void* stream);

/* ------------------------------------------------------------- noise -- */

enum {
    YG_EVAL_NOISE = 0,      /* newPerlin_t::operator() in [0,1] */
    YG_EVAL_FBM = 1,        /* fBm_t::operator() (noise.cc:638-650) */
    YG_EVAL_TURBULENCE = 2, /* turbulence() (noise.cc:808-819) */
    YG_EVAL_TEX_FLOAT = 3,  /* texture getFloat */
    YG_EVAL_TEX_COLOR = 4   /* texture getColor (3 floats per point) */
};

/* Evaluate noise/texture functions at n float3 points on the device.
 * For NOISE/FBM/TURBULENCE the parameters come from tex (H, lacunarity,
 * octaves for FBM; depth, size, hard for TURBULENCE). */
int yg_noise_eval(int32_t what, const yg_texture* tex, const float* d_pts, uint32_t n,
                  float* d_out, void* stream);

/* Heightfield vertices on the device: nx*ny vertices over [0,ex]x[0,ey],
 * z = amp * fBm(freq * (x, y, 0)) with fBm(H, lacunarity, octaves) over
 * newPerlin_t.  d_xyz receives nx*ny float3, row-major in y then x. */
int yg_gen_heightfield(int32_t nx, int32_t ny, float ex, float ey, float amp, float freq,
                       float H, float lacunarity, float octaves, float* d_xyz,
                       void* stream);

/* ------------------------------------------------------- QMC / RNG -- */

/* Halton radical inverse of index i (first getNext() of a fresh Halton(base)
 * returns index 1; mcqmc.h:35-49) for n consecutive indices starting at i0. */
int yg_halton(int32_t base, uint32_t i0, uint32_t n, float* d_out, void* stream);

/* -------------------------------------------------------- rendering -- */

int yg_set_camera(yg_scene* s, const yg_camera* cam);
int yg_set_shading(yg_scene* s, const yg_material* mats, uint32_t nmat,
                   const yg_modulator* mods, uint32_t nmod,
                   const yg_texture* texs, uint32_t ntex);
int yg_set_lights(yg_scene* s, const yg_light* lights, uint32_t n);

/* scene_t::render(colorOutput_t&) in mono mode (scene.cc:538-600), each area
 * rendered as scene_t::render(renderArea_t&) (scene.cc:724-816) including the
 * antialiasing passes of opts->aa_* (with area lights, whose ourRandom()
 * stream runs through each area's passes in turn, the areas are then launched
 * one after another): d_rgb receives resx*resy*3 floats (row-major, top row
 * first).
 * Lights supported: sun, area (psamples >= 0), path (use_QMC on; with cache on:
 * the fake pass, the samples and the interpolating pass in this one call),
 * photon (diffuse mode; yg_photon_build must have run), global photon
 * (yg_gphoton_build must have run). */
int yg_render(yg_scene* s, const yg_render_opts* opts, float* d_rgb,
              yg_render_stats* stats, void* stream);

/* Several scenes rendered by one set of launches (batched mode): the work
 * items of all scenes share the kernels, so small scenes no longer leave the
 * GPU idle.  d_rgb[k] receives scene k's image; opts[k] are per scene.
 * Every scene's pathlight must use the same samples and depth.  No reference
 * counterpart: the reference renders one scene per process. */
int yg_render_batch(yg_scene** scenes, uint32_t n, const yg_render_opts* opts, float** d_rgb,
                    yg_render_stats* stats, void* stream);

/* --------------------------------------------------------- autotuner -- */

typedef struct yg_tune_opts {
    int32_t widths;   /* BVH node widths to try, as bits: 1<<2 (two-wide), 1<<4 (four-wide), 1<<8 (eight-wide);
                         0 = all three */
    int32_t leaves;   /* max_leaf values to try, as bits 1<<n, n in 1..8; 0 = 2, 4 and 8 */
    int32_t bins;     /* SAH bin counts to try, as bits 1<<k for 2^k bins, k in 3..6; 0 = 32 only */
    int32_t repeats;  /* timed renders per candidate after one untimed render; 0 = 2 */
    int32_t threads;  /* host build threads; 0 = all */
    int32_t stacks;   /* stack_lds values to try, as bits: 1<<0 = whole stack in LDS,
                         1<<k = 2^k entries, k in 3..6; 0 = the build's default only */
    int32_t rays;     /* rays_per_lane values to try, as bits 1<<k for 2^k rays, k in 0..2;
                         0 = 1 only */
    int32_t compacts; /* compact_below values to try, as bits: 1<<k = k/4 for k in 1..4,
                         1<<0 = never; 0 = always compact only */
} yg_tune_opts;

#define YG_TUNE_MAX 32
typedef struct yg_tune_result {
    yg_build_opts best;               /* the build the scene holds on return */
    float best_ms;                    /* its frame time */
    int32_t ncand;                    /* candidates tried, in the order below */
    yg_build_opts cand[YG_TUNE_MAX];
    float cand_ms[YG_TUNE_MAX];       /* fastest frame of each candidate (device time) */
    float cand_build_ms[YG_TUNE_MAX]; /* build + upload of each candidate (host time) */
} yg_tune_result;

/* Autotuner (no reference counterpart: the reference's kd-tree has fixed
 * build constants, kdtree.cc:57-138).  Builds the scene with every candidate
 * (node width x leaf size x SAH bins, and for each build the launch settings
 * LDS stack entries x rays per lane x compaction threshold), times yg_render of the caller's frame
 * (opts, d_rgb: the image of the last render is left there) and leaves the
 * scene built with the fastest.  Images do not depend on the build (exact
 * distance ties go to the lowest primitive id).  At most YG_TUNE_MAX
 * candidates. */
int yg_autotune(yg_scene* s, const yg_render_opts* opts, float* d_rgb, const yg_tune_opts* t,
                yg_tune_result* r, void* stream);

/* Gradient of a loss through the image (no reference counterpart: the
 * reference has no derivatives).  Renders the image into d_rgb, then replays
 * every pathlight path with the same Halton samples and visibility held
 * fixed, applying the chain rule to the estimator of pathlight.cc:305-403 and
 * sunlight.cc:35-57 with d_grad_rgb = dLoss/dpixel (resx*resy*3).  Outputs
 * (overwritten, FP64): d_grad_mat[nmat*3] = dLoss/d albedo, d_grad_mod[nmod*2]
 * = dLoss/d modulator colour and dLoss/d lambda where every modulator size is
 * scaled by lambda (at lambda = 1), d_grad_light[nl] = dLoss/d power.
 * Scenes: one pathlight + at most one sun light, at most one modulator per
 * material (Musgrave fBm or clouds texture). */
int yg_render_adjoint(yg_scene* s, const yg_render_opts* opts, const float* d_grad_rgb, float* d_rgb,
                      double* d_grad_mat, double* d_grad_mod, double* d_grad_light,
                      yg_render_stats* stats, void* stream);

/* ----------------------------------------------------------- photons -- */

typedef struct yg_photon_stats {
    uint64_t emitted;
    uint64_t stored;
    uint64_t clustered;   /* photons kept after preGathering */
    float ms;
    uint64_t rays;        /* yg_gphoton_build: photon ray segments traced */
    float ms_trace;       /* yg_gphoton_build with YG_RENDER_KTIMING: device time of the
                             photon tracing launch */
    float ms_irr;         /* ... and of the pre-gathering launch (irradiance of every box) */
} yg_photon_stats;

/* photonLight_t::init (photonlight.cc:453-501) for light index li of the
 * light table: emission, tracing, cell clustering, pre-gathering and the
 * gather index.  Results stay on the device for yg_render / yg_photon_gather. */
int yg_photon_build(yg_scene* s, uint32_t li, const yg_render_opts* opts,
                    yg_photon_stats* stats, void* stream);

/* Copy the clustered photons (photonMark_t: dir, pos, colour = 9 floats each,
 * in the reference's std::map cell order) to d_out (capacity in photons). */
int yg_photon_read(yg_scene* s, float* d_out, uint64_t capacity, uint64_t* count,
                   void* stream);

/* photonLight_t::illuminate (photonlight.cc:297-395) at n shading points:
 * d_p = position, d_n = shading normal (already FACE_FORWARDed), d_ng =
 * geometric normal, d_eye = eye vector, d_mat = material id; d_rgb gets the
 * radiance estimate. */
int yg_photon_gather(yg_scene* s, const float* d_p, const float* d_n, const float* d_ng,
                     const float* d_eye, const int32_t* d_mat, uint32_t n, float* d_rgb,
                     void* stream);

/* globalPhotonLight_t::init (globalphotonlight.cc:234-287) for light index li:
 * photons from every dummy area light (quadEmitter_t, arealight.cc:200-232),
 * traced with diffuse bounces up to "depth" (globalphotonlight.cc:19-95),
 * merged per hash box (storeInHash, :177-195), then the irradiance of every
 * box from its "search" nearest box photons within "radius"
 * (computeIrradiances / setIrradiance, :127-175 and :197-228).  The photon
 * random numbers come from disjoint segments of the ourRandom() sequence,
 * one per photon (the reference draws them one photon after another).
 * stats: emitted photons, stored photons, boxes. */
int yg_gphoton_build(yg_scene* s, uint32_t li, const yg_render_opts* opts, yg_photon_stats* stats,
                     void* stream);

/* The irradiance cache of the scene's last yg_render whose pathlight has
 * cache on (pathlight.cc:405-723; no counterpart call in the reference, which
 * keeps it inside the light): its samples in insertion order, 18 floats
 * each -- P, N, polar position, colour, harmonic-mean distance, minimum
 * distance, precision, rank in the cache's tree order, pixel number, 1 --
 * into d_out (capacity in samples); *count = the number of samples. */
int yg_cache_read(yg_scene* s, float* d_out, uint64_t capacity, uint64_t* count, void* stream);

/* Copy the irradiance boxes (15 floats each: position, normal N, irradiance,
 * merged photon direction, merged photon colour; in box order x, y, z) to
 * d_out (capacity in boxes). */
int yg_gphoton_read(yg_scene* s, float* d_out, uint64_t capacity, uint64_t* count, void* stream);

/* The raw photon map (globalPhotonMap_t photonMap, globalphotonlight.cc:38,
 * the map photonSampler_t gathers from, pathtools.cc:172-174) in store
 * order: 5 words per photon -- position (3 float bit patterns), direction
 * code theta | phi << 8 (dirConverter_t bytes, photon.h:14-37; theta 255:
 * none), RGBE colour r | g << 8 | b << 16 | e << 24 (color.cc:140-155).
 * d_out: device memory of capacity records, or NULL to query *count. */
int yg_gphoton_read_photons(yg_scene* s, uint32_t* d_out, uint64_t capacity, uint64_t* count, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* YAFGPU_H */
