// Host-side mirror of the reference's plugin interface for the GPU path.
//
// The reference builds lights, shaders, textures and backgrounds through
// factories registered by name that read a paramMap_t
// (src/interface/external.h:112-117, params.h:54-73; e.g. pathLight_t::factory
// pathlight.cc:753-807) and ingests meshes through addObject_trimesh
// (interface_impl.cc:193-265).  This header offers the same names, parameter
// keys, defaults and error behaviour (factories return an error instead of an
// object, never exit), but each factory produces the POD descriptor of
// include/yafgpu.h, and render() runs the frame through libyafgpu.so.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "../../include/yafgpu.h"

namespace yafgpu {

// paramMap_t (params.h:54-73): string keys to typed values.
class paramMap_t {
public:
    void setFloat(const std::string& k, float v) { f_[k] = v; }
    void setInt(const std::string& k, int v) { i_[k] = v; }
    void setBool(const std::string& k, bool v) { b_[k] = v; }
    void setString(const std::string& k, const std::string& v) { s_[k] = v; }
    void setPoint(const std::string& k, float x, float y, float z) { p_[k] = {x, y, z}; }
    void setColor(const std::string& k, float r, float g, float b) { p_[k] = {r, g, b}; }

    // getParam: leaves `out` untouched and returns false when the key is absent
    bool getParam(const std::string& k, float& out) const;
    bool getParam(const std::string& k, int& out) const;
    bool getParam(const std::string& k, bool& out) const;
    bool getParam(const std::string& k, std::string& out) const;
    bool getParam(const std::string& k, float out[3]) const;

private:
    std::map<std::string, float> f_;
    std::map<std::string, int> i_;
    std::map<std::string, bool> b_;
    std::map<std::string, std::string> s_;
    std::map<std::string, std::vector<float>> p_;
};

// Factories by plugin name, with the reference defaults.  Each returns YG_OK
// and fills the descriptor, or a negative status (message in lastError()).
int lightFactory(const std::string& type, const paramMap_t& p, yg_light& out);
int textureFactory(const std::string& type, const paramMap_t& p, yg_texture& out);
// genericShader_t::factory (basicshaders.cc:238-367): base params + one
// paramMap_t per modulator; texture names resolve through `textures`.
int shaderFactory(const std::string& type, const paramMap_t& p, const std::vector<paramMap_t>& mods,
                  const std::map<std::string, int>& textures, yg_material& out,
                  std::vector<yg_modulator>& mods_out);
const std::string& lastError();
// records an error (lastError, cerr with the reference's prefix) and returns code
int hostFail(int code, const std::string& msg);

// renderEnvironment / yafrayInterface_t subset for the GPU path.
class gpuInterface_t {
public:
    gpuInterface_t();
    ~gpuInterface_t();
    int addTexture(const std::string& name, const std::string& type, const paramMap_t& p);
    int addShader(const std::string& name, const std::string& type, const paramMap_t& p,
                  const std::vector<paramMap_t>& mods);
    int addLight(const std::string& name, const std::string& type, const paramMap_t& p);
    int addBackground(const std::string& name, const paramMap_t& p);
    int addCamera(const std::string& name, const paramMap_t& p);
    // addObject_trimesh: world-space vertices, index triples, one shader name
    // per face (or a single name for all faces); smooth_angle: the mesh's
    // autosmooth angle in degrees (meshObject_t::autoSmooth, below 1: none)
    int addObject_trimesh(const std::string& name, const std::vector<float>& verts,
                          const std::vector<uint32_t>& faces, const std::vector<std::string>& shaders,
                          bool cast_shadows = true, bool recv_rad = true, bool use_rad = true,
                          float smooth_angle = 0.0f);
    // texture-coordinate data of object name (render.cc:743-764,
    // interface_impl.cc:229-238): face_uv 6 per face (u_a v_a u_b v_b u_c
    // v_c) or empty, face_flags per face (bit0 uv, bit1 vertex colours) or
    // empty, orco 3 per vertex (has_orco) or empty, orco_mtx the 12-float
    // orco matrix (yg_orco_matrix of the object's transform) or empty
    int setObjectTexcoords(const std::string& name, const std::vector<float>& face_uv,
                           const std::vector<uint8_t>& face_flags, const std::vector<float>& orco,
                           const std::vector<float>& orco_mtx);
    // render params: camera_name, background_name, bias, raydepth, AA_*, exposure,
    // gamma, clamp_rgb, output_type; rgb gets resx*resy*3 floats, tone-mapped
    // unless output_type is hdr or exr.  Photon lights are built first.
    int render(const paramMap_t& params, std::vector<float>& rgb, yg_render_stats* stats = nullptr);

    // the render options render() passes to yg_render for these params
    // (interfaceImpl_t::render's defaults and ranges, interface_impl.cc:480-533)
    static yg_render_opts render_opts(const paramMap_t& params);

    // What a finished render did differently from the reference, as warning
    // text ("" when nothing): pixels whose pathlight asked for adaptive extra
    // repetitions (pathlight.cc:349-374; the GPU path counts them in
    // stats.adaptive_extra but renders the first round only).
    static std::string renderWarnings(const yg_render_stats& st);

    // resolution of a defined camera; false when there is none of that name
    bool cameraResolution(const std::string& name, int& resx, int& resy) const;

    // the descriptors the next render would use (object order, light order)
    int describe(std::vector<yg_light>& lights, std::vector<yg_material>& mats,
                 std::vector<yg_modulator>& mods, std::vector<yg_texture>& texs);

private:
    struct Obj {
        std::vector<float> v;
        std::vector<uint32_t> f;
        std::vector<std::string> sh;
        int flags;
        float smooth = 0.0f;
        std::vector<float> uv, orco, mtx;
        std::vector<uint8_t> fflags;
    };
    std::map<std::string, yg_texture> textures_;
    std::map<std::string, std::pair<paramMap_t, std::vector<paramMap_t>>> shaders_;
    std::map<std::string, yg_light> lights_;
    std::map<std::string, std::vector<float>> backgrounds_;
    std::map<std::string, yg_camera> cameras_;
    std::map<std::string, Obj> objects_;
};

}  // namespace yafgpu
