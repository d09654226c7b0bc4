// Host-side mirror of the reference plugin factories (see yafgpu_host.h).
#include "yafgpu_host.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <iostream>

namespace yafgpu {

static std::string g_host_err;

static int fail(int code, const std::string& msg)
{
    g_host_err = msg;
    std::cerr << "[Error] " << msg << "\n";  // interface_impl.cc:44-46 prefixes
    return code;
}

const std::string& lastError() { return g_host_err; }
int hostFail(int code, const std::string& msg) { return fail(code, msg); }

bool paramMap_t::getParam(const std::string& k, float& out) const
{
    auto it = f_.find(k);
    if (it != f_.end()) { out = it->second; return true; }
    auto ii = i_.find(k);
    if (ii != i_.end()) { out = (float)ii->second; return true; }
    return false;
}
bool paramMap_t::getParam(const std::string& k, int& out) const
{
    auto it = i_.find(k);
    if (it == i_.end()) return false;
    out = it->second;
    return true;
}
bool paramMap_t::getParam(const std::string& k, bool& out) const
{
    auto it = b_.find(k);
    if (it == b_.end()) return false;
    out = it->second;
    return true;
}
bool paramMap_t::getParam(const std::string& k, std::string& out) const
{
    auto it = s_.find(k);
    if (it == s_.end()) return false;
    out = it->second;
    return true;
}
bool paramMap_t::getParam(const std::string& k, float out[3]) const
{
    auto it = p_.find(k);
    if (it == p_.end()) return false;
    for (int c = 0; c < 3; c++) out[c] = it->second[c];
    return true;
}

static void set3(float* d, float a, float b, float c) { d[0] = a; d[1] = b; d[2] = c; }

int lightFactory(const std::string& type, const paramMap_t& p, yg_light& L)
{
    memset(&L, 0, sizeof(L));
    L.use_in_render = 1;
    L.use_in_indirect = 1;  // light_t ctor (light.h:61)
    if (type == "sunlight") {
        // sunLight_t::factory (sunlight.cc:59-73)
        L.type = YG_LIGHT_SUN;
        set3(L.color, 1, 1, 1);
        L.power = 1.0f;
        bool shadows = true;
        p.getParam("from", L.dir);
        p.getParam("color", L.color);
        p.getParam("power", L.power);
        p.getParam("cast_shadows", shadows);
        L.cast_shadows = shadows;
        return YG_OK;
    }
    if (type == "arealight") {
        // areaLight_t::factory (arealight.cc:257-282)
        L.type = YG_LIGHT_AREA;
        L.power = 1.0f;
        L.samples = 50;
        L.psamples = 0;
        bool dummy = false;
        p.getParam("a", L.a);
        p.getParam("b", L.b);
        p.getParam("c", L.c);
        p.getParam("d", L.d);
        p.getParam("color", L.color);
        p.getParam("power", L.power);
        p.getParam("samples", L.samples);
        p.getParam("psamples", L.psamples);
        p.getParam("dummy", dummy);
        L.dummy = dummy ? 1 : 0;  // emits global photons only (arealight.cc:98)
        if (L.psamples < 0) L.psamples = 0;
        return YG_OK;
    }
    if (type == "globalphotonlight") {
        // globalPhotonLight_t::factory (globalphotonlight.cc:280-294)
        L.type = YG_LIGHT_GLOBALPHOTON;
        L.fixedradius = 1.0f;
        L.depth = 2;
        L.photons = 50000;
        L.search = 200;
        p.getParam("radius", L.fixedradius);
        p.getParam("depth", L.depth);
        p.getParam("photons", L.photons);
        p.getParam("search", L.search);
        int cdepth = 4;
        p.getParam("caus_depth", cdepth);  // caustic objects are refused by the loader
        L.use_in_indirect = 0;
        return YG_OK;
    }
    if (type == "pathlight") {
        // pathLight_t::factory (pathlight.cc:753-807)
        L.type = YG_LIGHT_PATH;
        L.power = 1.0f;
        L.threshold = 0.1f;
        L.samples = 16;
        L.depth = 3;
        bool qmc = false, cache = false, direct = false;
        std::string mode;
        p.getParam("power", L.power);
        p.getParam("depth", L.depth);
        p.getParam("samples", L.samples);
        p.getParam("use_QMC", qmc);
        p.getParam("cache", cache);
        p.getParam("direct", direct);
        p.getParam("threshold", L.threshold);
        p.getParam("mode", mode);
        if (L.samples < 1) {
            std::cerr << "[Warning] Samples value too low, minimum is one\n";
            L.samples = 1;
        }
        L.cache = cache ? 1 : 0;
        L.cache_size = 0.01f;
        L.shadow_threshold = 0.3f;
        L.search = 50;
        if (cache) {
            bool show = false, grad = false, ignorms = false;
            p.getParam("cache_size", L.cache_size);
            p.getParam("shadow_threshold", L.shadow_threshold);
            p.getParam("search", L.search);
            p.getParam("show_samples", show);
            p.getParam("gradient", grad);
            p.getParam("ignore_bumpnormals", ignorms);
            if (L.search < 3) L.search = 3;
            if (show) return fail(YG_ERR_UNSUPPORTED, "pathlight: show_samples is not on the GPU path");
            if (ignorms) return fail(YG_ERR_UNSUPPORTED, "pathlight: ignore_bumpnormals is not on the GPU path");
            (void)grad;  // read and unused by the reference (pathlight.cc:790)
        }
        // photonSampler_t's hemigrid (pathlight.cc:774, pathtools.cc:100-106):
        // the device plans 4 x 8 sectors, the grid of the default 36
        int grid = 36;
        p.getParam("grid", grid);
        if (grid != 36) return fail(YG_ERR_UNSUPPORTED, "pathlight: grid other than 36 is not on the GPU path");
        if (direct) return fail(YG_ERR_UNSUPPORTED, "pathlight: direct=on is not on the GPU path");
        if (mode == "occlusion") return fail(YG_ERR_UNSUPPORTED, "pathlight: occlusion mode is not on the GPU path");
        if (!qmc) return fail(YG_ERR_UNSUPPORTED, "pathlight: use_QMC=off (LCG sampler) is not on the GPU path");
        L.use_qmc = 1;
        L.use_in_indirect = 0;  // pathLight_t::init (pathlight.cc:100)
        return YG_OK;
    }
    if (type == "photonlight") {
        // photonLight_t::factory (photonlight.cc:692-732)
        L.type = YG_LIGHT_PHOTON;
        set3(L.from, 0, 0, 1);
        set3(L.to, 0, 0, 0);
        set3(L.color, 1, 1, 1);
        L.power = 1.0f;
        L.angle = 45.0f;
        L.bias = 0.001f;
        L.fixedradius = 1.0f;
        L.photons = 5000;
        L.search = 50;
        L.depth = 3;
        L.mindepth = 1;
        bool qmc = false;
        std::string mode;
        p.getParam("from", L.from);
        p.getParam("to", L.to);
        p.getParam("color", L.color);
        p.getParam("photons", L.photons);
        p.getParam("search", L.search);
        p.getParam("power", L.power);
        p.getParam("angle", L.angle);
        p.getParam("depth", L.depth);
        p.getParam("mindepth", L.mindepth);
        p.getParam("bias", L.bias);
        p.getParam("use_QMC", qmc);
        p.getParam("mode", mode);
        if (!p.getParam("fixedradius", L.fixedradius))
            std::cerr << "[Warning] Missing fixedradius, using default won't work.\n";
        if (!p.getParam("cluster", L.cluster)) L.cluster = L.fixedradius / 4;
        if (mode != "diffuse") return fail(YG_ERR_UNSUPPORTED, "photonlight: caustic mode is not on the GPU path");
        if (!qmc) return fail(YG_ERR_UNSUPPORTED, "photonlight: use_QMC=off is not on the GPU path");
        L.use_qmc = 1;
        L.use_in_indirect = 0;  // photonLight_t ctor (photonlight.cc:102)
        return YG_OK;
    }
    return fail(YG_ERR_UNSUPPORTED, "light type '" + type + "' is not on the GPU path");
}

int textureFactory(const std::string& type, const paramMap_t& p, yg_texture& T)
{
    memset(&T, 0, sizeof(T));
    set3(T.color1, 0, 0, 0);
    set3(T.color2, 1, 1, 1);
    // newNoise (basictex.cc:38-63): any other name is newPerlin_t
    std::string ntype;
    p.getParam("noise_type", ntype);
    if (ntype.find("voronoi") != std::string::npos || ntype == "cellnoise")
        return fail(YG_ERR_UNSUPPORTED, "noise_type '" + ntype + "' is not on the GPU path");
    T.noise = (ntype == "blender") ? YG_NOISE_BLENDER : (ntype == "stdperlin") ? YG_NOISE_STDPERLIN : YG_NOISE_NEWPERLIN;
    p.getParam("color1", T.color1);
    p.getParam("color2", T.color2);
    std::string shape;
    p.getParam("shape", shape);
    T.shape = (shape == "saw") ? YG_SHAPE_SAW : (shape == "tri") ? YG_SHAPE_TRI : YG_SHAPE_SIN;
    if (type == "clouds") {
        // textureClouds_t::factory (basictex.cc:98-115)
        T.type = YG_TEX_CLOUDS;
        T.depth = 2;
        T.size = 1;
        bool hard = false;
        std::string bias;
        p.getParam("depth", T.depth);
        p.getParam("size", T.size);
        p.getParam("hard", hard);
        p.getParam("bias", bias);
        T.hard = hard;
        T.bias = (bias == "positive") ? 1 : (bias == "negative") ? 2 : 0;
        return YG_OK;
    }
    if (type == "marble" || type == "wood") {
        // textureMarble_t / textureWood_t factories (basictex.cc:159-177, 218-230)
        T.type = (type == "marble") ? YG_TEX_MARBLE : YG_TEX_WOOD;
        T.depth = 2;
        T.turb = 1.0f;
        T.size = 1.0f;
        float shp = 1.0f;
        bool hard = false;
        std::string wtype;
        p.getParam("depth", T.depth);
        p.getParam("turbulence", T.turb);
        p.getParam("sharpness", shp);
        p.getParam("size", T.size);
        p.getParam("hard", hard);
        p.getParam("type", wtype);
        T.hard = hard;
        T.sharpness = (shp > 1) ? 1.0f / shp : 1.0f;
        T.rings = (wtype == "rings");
        return YG_OK;
    }
    if (type == "musgrave") {
        // textureMusgrave_t::factory (basictex.cc:687-709) and its ctor's
        // choice of generator (:657-672; any other name is fBm)
        T.type = YG_TEX_MUSGRAVE;
        T.H = 1;
        T.lacunarity = 2;
        T.octaves = 2;
        T.size = 1;
        T.iscale = 1;
        T.offset = 1;
        T.gain = 1;
        std::string mtype;
        p.getParam("type", mtype);
        if (mtype == "multifractal") T.mtype = YG_MUSGRAVE_MULTIFRACTAL;
        else if (mtype == "heteroterrain") T.mtype = YG_MUSGRAVE_HETEROTERRAIN;
        else if (mtype == "hybridmf") T.mtype = YG_MUSGRAVE_HYBRIDMF;
        else if (mtype == "ridgedmf") T.mtype = YG_MUSGRAVE_RIDGEDMF;
        else T.mtype = YG_MUSGRAVE_FBM;
        p.getParam("H", T.H);
        p.getParam("lacunarity", T.lacunarity);
        p.getParam("octaves", T.octaves);
        p.getParam("offset", T.offset);
        p.getParam("gain", T.gain);
        p.getParam("size", T.size);
        p.getParam("intensity", T.iscale);
        return YG_OK;
    }
    return fail(YG_ERR_UNSUPPORTED, "texture type '" + type + "' is not on the GPU path");
}

int shaderFactory(const std::string& type, const paramMap_t& p, const std::vector<paramMap_t>& mods,
                  const std::map<std::string, int>& textures, yg_material& M, std::vector<yg_modulator>& out)
{
    if (type != "generic") return fail(YG_ERR_UNSUPPORTED, "shader type '" + type + "' is not on the GPU path");
    memset(&M, 0, sizeof(M));
    p.getParam("color", M.color);
    float spec[3] = {0, 0, 0}, refl[3] = {0, 0, 0}, trans[3] = {0, 0, 0};
    p.getParam("specular", spec);
    p.getParam("reflected", refl);
    p.getParam("transmitted", trans);
    for (int c = 0; c < 3; c++)
        if (spec[c] != 0 || refl[c] != 0 || trans[c] != 0)
            return fail(YG_ERR_UNSUPPORTED, "generic shader: specular/reflection/refraction are not on the GPU path");
    M.mod_first = (int)out.size();
    M.mod_count = 0;
    for (const paramMap_t& mp : mods) {
        // modulator params (basicshaders.cc:262-357)
        std::string texname, mode = "mix", texco = "orco";
        mp.getParam("texname", texname);
        auto it = textures.find(texname);
        if (it == textures.end()) {
            std::cerr << "Undefined texture : " << texname << "\n";  // skipped like the reference
            continue;
        }
        yg_modulator m;
        memset(&m, 0, sizeof(m));
        m.tex = it->second;
        float size = 1, sx = 1, sy = 1, sz = 1, col = 0;
        mp.getParam("size", size);
        mp.getParam("sizex", sx);
        mp.getParam("sizey", sy);
        mp.getParam("sizez", sz);
        mp.getParam("color", col);
        mp.getParam("ofsx", m.ofs[0]);
        mp.getParam("ofsy", m.ofs[1]);
        mp.getParam("ofsz", m.ofs[2]);
        mp.getParam("mode", mode);
        mp.getParam("texco", texco);
        // string2texcotype (texture.h:135-149): an unknown name is "uv"
        m.texco = YG_TEXCO_UV;
        if (texco == "global") m.texco = YG_TEXCO_GLOBAL;
        else if (texco == "orco") m.texco = YG_TEXCO_ORCO;
        else if (texco == "window") m.texco = YG_TEXCO_WIN;
        else if (texco == "normal") m.texco = YG_TEXCO_NOR;
        else if (texco == "reflect") m.texco = YG_TEXCO_REFL;
        // string2texprojection (texture.h:166-173): "nxyz".find(axis), 'n' or
        // an unknown name = no axis
        const char* defaults[3] = {"x", "y", "z"};
        const char* keys[3] = {"proj_x", "proj_y", "proj_z"};
        for (int a = 0; a < 3; a++) {
            std::string ax = defaults[a];
            mp.getParam(keys[a], ax);
            const size_t k = std::string("nxyz").find(ax);
            m.proj[a] = (k == std::string::npos || k == 0) ? -1 : (int32_t)k;
        }
        // the texture matrix m00..m23 (basicshaders.cc:317-333; row 3 does
        // not enter matrix4x4_t * point3d_t)
        float mtx[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++) mp.getParam("m" + std::to_string(r) + std::to_string(c), mtx[4 * r + c]);
        const float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        if (memcmp(mtx, ident, sizeof(mtx)) != 0) {
            m.has_mtx = 1;
            memcpy(m.mtx, mtx, sizeof(mtx));
        }
        set3(m.size, sx, sy, sz);
        if (size != 1.0f) set3(m.size, size, size, size);
        m.color = col;
        if (mode == "mix") m.mode = YG_MOD_MIX;
        else if (mode == "mul") m.mode = YG_MOD_MUL;
        else if (mode == "add") m.mode = YG_MOD_ADD;
        else if (mode == "sub") m.mode = YG_MOD_SUB;
        else return fail(YG_ERR_UNSUPPORTED, "modulator mode '" + mode + "'");
        out.push_back(m);
        M.mod_count++;
    }
    return YG_OK;
}

// ------------------------------------------------------------ interface --

gpuInterface_t::gpuInterface_t() {}
gpuInterface_t::~gpuInterface_t() {}

int gpuInterface_t::addTexture(const std::string& name, const std::string& type, const paramMap_t& p)
{
    yg_texture t;
    int rc = textureFactory(type, p, t);
    if (rc) return rc;
    textures_[name] = t;  // registerFactory-style overwrite by name
    return YG_OK;
}

int gpuInterface_t::addShader(const std::string& name, const std::string& type, const paramMap_t& p,
                              const std::vector<paramMap_t>& mods)
{
    if (type != "generic") return fail(YG_ERR_UNSUPPORTED, "shader type '" + type + "' is not on the GPU path");
    shaders_[name] = std::make_pair(p, mods);
    return YG_OK;
}

int gpuInterface_t::addLight(const std::string& name, const std::string& type, const paramMap_t& p)
{
    yg_light L;
    int rc = lightFactory(type, p, L);
    if (rc) return rc;
    // render_t::light (render.cc:781-815) applies use_in_render / use_in_indirect
    // after the factory; path and photon lights leave indirect light again in
    // their init (pathlight.cc:100, photonlight.cc:102), which runs later
    bool in_render = true, in_indirect = true;
    p.getParam("use_in_render", in_render);
    p.getParam("use_in_indirect", in_indirect);
    L.use_in_render = in_render ? 1 : 0;
    if (L.type != YG_LIGHT_PATH && L.type != YG_LIGHT_PHOTON) L.use_in_indirect = in_indirect ? 1 : 0;
    lights_[name] = L;
    return YG_OK;
}

int gpuInterface_t::addBackground(const std::string& name, const paramMap_t& p)
{
    // constBackground_t::factory (sunsky.cc:12-17)
    std::vector<float> c = {1, 1, 1};
    float col[3] = {1, 1, 1};
    p.getParam("color", col);
    c.assign(col, col + 3);
    backgrounds_[name] = c;
    return YG_OK;
}

bool gpuInterface_t::cameraResolution(const std::string& name, int& resx, int& resy) const
{
    auto it = cameras_.find(name);
    if (it == cameras_.end()) return false;
    resx = it->second.resx;
    resy = it->second.resy;
    return true;
}

int gpuInterface_t::addCamera(const std::string& name, const paramMap_t& p)
{
    yg_camera c;
    memset(&c, 0, sizeof(c));
    c.resx = 320;
    c.resy = 200;
    c.aspect = 1.0f;
    c.focal = 1.0f;
    set3(c.up, 0, 1, 0);
    p.getParam("from", c.pos);
    p.getParam("to", c.look);
    p.getParam("up", c.up);
    p.getParam("resx", c.resx);
    p.getParam("resy", c.resy);
    p.getParam("focal", c.focal);
    p.getParam("aspect_ratio", c.aspect);
    if (c.resx <= 0 || c.resy <= 0) return fail(YG_ERR_ARG, "camera resolution");
    cameras_[name] = c;
    return YG_OK;
}

int gpuInterface_t::addObject_trimesh(const std::string& name, const std::vector<float>& verts,
                                      const std::vector<uint32_t>& faces, const std::vector<std::string>& shaders,
                                      bool cast_shadows, bool recv_rad, bool use_rad, float smooth_angle)
{
    if (faces.size() % 3 || verts.size() % 3) return fail(YG_ERR_ARG, "addObject_trimesh: bad array sizes");
    if (shaders.size() != 1 && shaders.size() != faces.size() / 3)
        return fail(YG_ERR_ARG, "addObject_trimesh: one shader name or one per face");
    Obj o;
    o.v = verts;
    o.f = faces;
    o.sh = shaders;
    o.flags = (cast_shadows ? 1 : 0) | (recv_rad ? 2 : 0) | (use_rad ? 4 : 0);
    o.smooth = smooth_angle;
    objects_[name] = o;
    return YG_OK;
}

int gpuInterface_t::setObjectTexcoords(const std::string& name, const std::vector<float>& face_uv,
                                       const std::vector<uint8_t>& face_flags, const std::vector<float>& orco,
                                       const std::vector<float>& orco_mtx)
{
    auto it = objects_.find(name);
    if (it == objects_.end()) return fail(YG_ERR_ARG, "setObjectTexcoords: no object " + name);
    Obj& o = it->second;
    const size_t nf = o.f.size() / 3;
    if ((!face_uv.empty() && face_uv.size() != 6 * nf) || (!face_flags.empty() && face_flags.size() != nf) ||
        (!orco.empty() && orco.size() != o.v.size()) || (!orco_mtx.empty() && orco_mtx.size() != 12))
        return fail(YG_ERR_ARG, "setObjectTexcoords: bad array sizes");
    o.uv = face_uv;
    o.fflags = face_flags;
    o.orco = orco;
    o.mtx = orco_mtx;
    return YG_OK;
}

int gpuInterface_t::describe(std::vector<yg_light>& lights, std::vector<yg_material>& mats,
                             std::vector<yg_modulator>& mods, std::vector<yg_texture>& texs)
{
    lights.clear();
    mats.clear();
    mods.clear();
    texs.clear();
    std::map<std::string, int> tindex;
    for (auto& kv : textures_) {
        tindex[kv.first] = (int)texs.size();
        texs.push_back(kv.second);
    }
    for (auto& kv : shaders_) {
        yg_material m;
        int rc = shaderFactory("generic", kv.second.first, kv.second.second, tindex, m, mods);
        if (rc) return rc;
        mats.push_back(m);
    }
    for (auto& kv : lights_) lights.push_back(kv.second);  // alphabetical (interface_impl.cc:571-576)
    return YG_OK;
}

yg_render_opts gpuInterface_t::render_opts(const paramMap_t& params)
{
    yg_render_opts o;
    memset(&o, 0, sizeof(o));
    o.min_raydis = 0.00005f;  // MIN_RAYDIST (configio.py:95-100)
    o.self_bias = 0.1f;       // scene_t ctor (scene.cc:56)
    o.raydepth = 3;           // maxraylevel
    o.block = 64;
    o.rand_seed = 1;
    o.lcg_seed = 123212;      // myseed (vector3d.cc:160)
    params.getParam("bias", o.self_bias);
    params.getParam("raydepth", o.raydepth);
    // antialiasing parameters and their ranges (interface_impl.cc:501-533)
    int aa_passes = 0, aa_min = 1;
    float aa_pw = 1.5f, aa_thr = 0.05f;
    bool aa_jit = false;
    if (params.getParam("samples", aa_passes)) aa_passes -= 1;  // obsolete name, counts the first pass
    else params.getParam("AA_passes", aa_passes);
    if (aa_passes < 0) aa_passes = 1;
    params.getParam("AA_minsamples", aa_min);
    if (aa_passes && aa_min < 1) aa_min = 1;
    params.getParam("AA_pixelwidth", aa_pw);
    if (aa_pw < 1.0f) aa_pw = 1.0f;
    else if (aa_pw > 2.0f) aa_pw = 2.0f;
    if (!params.getParam("tolerance", aa_thr)) params.getParam("AA_threshold", aa_thr);
    params.getParam("AA_jitterfirst", aa_jit);
    o.aa_passes = aa_passes;
    o.aa_minsamples = aa_min;
    o.aa_pixelwidth = aa_pw;
    o.aa_threshold = aa_thr;
    o.aa_jitterfirst = aa_jit ? 1 : 0;
    // exposure and gamma (interface_impl.cc:545-548); the reference tone-maps
    // only for 8-bit output (render.cc:1233-1257)
    bool clamp = false;
    std::string out_type = "tga";
    o.exposure = 0.0f;
    o.gamma = 1.0f;
    params.getParam("exposure", o.exposure);
    params.getParam("gamma", o.gamma);
    params.getParam("clamp_rgb", clamp);
    params.getParam("output_type", out_type);
    o.clamp_rgb = clamp ? 1 : 0;
    o.tonemap = (out_type == "hdr" || out_type.find("exr") != std::string::npos) ? 0 : 1;
    return o;
}

int gpuInterface_t::render(const paramMap_t& params, std::vector<float>& rgb, yg_render_stats* stats)
{
    std::vector<yg_light> lights;
    std::vector<yg_material> mats;
    std::vector<yg_modulator> mods;
    std::vector<yg_texture> texs;
    int rc = describe(lights, mats, mods, texs);
    if (rc) return rc;
    std::string camname, bgname;
    params.getParam("camera_name", camname);
    params.getParam("background_name", bgname);
    auto cit = cameras_.find(camname);
    if (cit == cameras_.end()) return fail(YG_ERR_ARG, "camera '" + camname + "' not found");
    std::map<std::string, int> shindex;
    int k = 0;
    for (auto& kv : shaders_) shindex[kv.first] = k++;
    yg_scene* s = nullptr;
    if ((rc = yg_scene_create(&s))) return rc;
    struct Guard {
        yg_scene* s;
        float* d = nullptr;
        ~Guard()
        {
            if (d) (void)hipFree(d);
            yg_scene_destroy(s);
        }
    } guard{s};
    for (auto& kv : objects_) {  // objects in name order (render.cc:1158-1164)
        const Obj& o = kv.second;
        std::vector<int32_t> fm(o.f.size() / 3);
        for (size_t t = 0; t < fm.size(); t++) {
            const std::string& sn = o.sh.size() == 1 ? o.sh[0] : o.sh[t];
            auto it = shindex.find(sn);
            if (it == shindex.end()) return fail(YG_ERR_ARG, "shader '" + sn + "' not found");
            fm[t] = it->second;
        }
        int32_t oid = -1;
        if ((rc = yg_mesh_add(s, o.v.data(), (uint32_t)(o.v.size() / 3), o.f.data(), (uint32_t)fm.size(),
                              fm.data(), o.flags, &oid)))
            return rc;
        // render.cc:764: obj->autoSmooth(angle) after the mesh is built
        if (o.smooth >= 1.0f && (rc = yg_mesh_autosmooth(s, oid, o.smooth))) return rc;
        if (!o.uv.empty() || !o.fflags.empty() || !o.orco.empty() || !o.mtx.empty())
            if ((rc = yg_mesh_texcoords(s, oid, o.uv.empty() ? nullptr : o.uv.data(),
                                        o.fflags.empty() ? nullptr : o.fflags.data(),
                                        o.orco.empty() ? nullptr : o.orco.data(),
                                        o.mtx.empty() ? nullptr : o.mtx.data())))
                return rc;
    }
    if ((rc = yg_build(s, nullptr))) return rc;
    if ((rc = yg_set_camera(s, &cit->second))) return rc;
    if ((rc = yg_set_shading(s, mats.data(), (uint32_t)mats.size(), mods.data(), (uint32_t)mods.size(),
                             texs.data(), (uint32_t)texs.size())))
        return rc;
    if ((rc = yg_set_lights(s, lights.data(), (uint32_t)lights.size()))) return rc;
    yg_render_opts o = render_opts(params);
    auto bit = backgrounds_.find(bgname);
    if (bit != backgrounds_.end()) {
        o.has_background = 1;
        for (int c = 0; c < 3; c++) o.background[c] = bit->second[c];
    }
    // the photon lights' init (scene_t::setupLights, scene.cc:279-295) before the render
    for (uint32_t li = 0; li < lights.size(); li++) {
        if (lights[li].type == YG_LIGHT_PHOTON && (rc = yg_photon_build(s, li, &o, nullptr, nullptr))) return rc;
        if (lights[li].type == YG_LIGHT_GLOBALPHOTON && (rc = yg_gphoton_build(s, li, &o, nullptr, nullptr)))
            return rc;
    }
    const yg_camera& cam = cit->second;
    const size_t n = (size_t)cam.resx * cam.resy * 3;
    if (hipMalloc((void**)&guard.d, n * sizeof(float)) != hipSuccess) return fail(YG_ERR_NOMEM, "output buffer");
    if ((rc = yg_render(s, &o, guard.d, stats, nullptr))) return rc;
    rgb.resize(n);
    if (hipMemcpy(rgb.data(), guard.d, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(YG_ERR_HIP, "output copy");
    return YG_OK;
}

std::string gpuInterface_t::renderWarnings(const yg_render_stats& st)
{
    if (st.adaptive_extra == 0) return std::string();
    return "[Warning]: " + std::to_string(st.adaptive_extra) +
           " pixels asked the pathlight for adaptive extra repetitions (its \"threshold\" is too low for "
           "them to stay off); the GPU path renders the first round only, so these pixels differ from the "
           "reference's. Set the pathlight threshold to 1000 or more for the reference's image.";
}

}  // namespace yafgpu

// C entry for callers without the C++ types: the factory defaults, so a
// binding can check its parameter mapping (no device work).
extern "C" int yg_host_light_defaults(const char* type, yg_light* out)
{
    yafgpu::paramMap_t p;
    if (std::string(type) == "pathlight") p.setBool("use_QMC", true);
    if (std::string(type) == "photonlight") {
        p.setBool("use_QMC", true);
        p.setString("mode", "diffuse");
        p.setFloat("fixedradius", 1.0f);
    }
    return yafgpu::lightFactory(type, p, *out);
}
