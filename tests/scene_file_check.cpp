// Scene-file checks without a GPU: `describe <scene.xml>` prints what the
// loader built, `images <out.tga> <out.hdr>` writes a synthetic image,
// `errors` loads malformed documents and prints the status of each.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

#include "yafgpu_xml.h"

using namespace yafgpu;

int main(int argc, char** argv)
{
    if (argc >= 3 && !strcmp(argv[1], "describe")) {
        std::ifstream in(argv[2]);
        std::stringstream ss;
        ss << in.rdbuf();
        gpuInterface_t I;
        std::vector<paramMap_t> renders;
        int rc = loadScene(ss.str(), I, renders);
        printf("load %d\n", rc);
        if (rc) return 0;
        std::vector<yg_light> lights;
        std::vector<yg_material> mats;
        std::vector<yg_modulator> mods;
        std::vector<yg_texture> texs;
        if (I.describe(lights, mats, mods, texs)) return 3;
        for (auto& L : lights)
            printf("light %d %d %d %.9g %.9g %.9g %.9g %.9g %.9g %.9g %d %d %d %d %.9g\n", L.type, L.use_in_render,
                   L.use_in_indirect, L.power, L.color[0], L.color[1], L.color[2], L.dir[0], L.dir[1], L.dir[2],
                   L.cast_shadows, L.samples, L.depth, L.use_qmc, L.threshold);
        for (auto& m : mats) printf("mat %.9g %.9g %.9g\n", m.color[0], m.color[1], m.color[2]);
        for (auto& p : renders) {
            std::string cam, out, type = "tga";
            p.getParam("camera_name", cam);
            p.getParam("outfile", out);
            p.getParam("output_type", type);
            int w = 0, h = 0;
            I.cameraResolution(cam, w, h);
            const yg_render_opts o = gpuInterface_t::render_opts(p);
            printf("render %s %d %d %s %s %d %.9g %d %d %d %.9g %.9g\n", cam.c_str(), w, h, out.c_str(), type.c_str(),
                   o.raydepth, o.self_bias, o.aa_passes, o.aa_minsamples, o.tonemap, o.exposure, o.gamma);
        }
        return 0;
    }
    if (argc >= 4 && !strcmp(argv[1], "images")) {
        const int w = 37, h = 5;
        std::vector<float> rgb((size_t)w * h * 3);
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++)
                for (int c = 0; c < 3; c++) {
                    // runs, literal spans, values above 1, tiny and zero pixels
                    float v = (x < 12) ? 0.25f + 0.125f * c : (float)(x * (c + 1) + y) / 23.0f;
                    if (x == 20) v = 1e-35f;
                    if (x == 21 && c == 1) v = 0.0f;
                    if (y == 4) v = std::ldexp(1.0f + 0.1f * c, x - 10);
                    rgb[3 * ((size_t)y * w + x) + c] = v;
                }
        printf("tga %d\n", writeTGA(argv[2], w, h, rgb.data()));
        printf("hdr %d\n", writeHDR(argv[3], w, h, rgb.data()));
        for (float v : rgb) printf("%.9g\n", v);
        return 0;
    }
    if (argc >= 2 && !strcmp(argv[1], "errors")) {
        const char* docs[] = {
            "<scene><shader name=\"a\" type=\"generic\"></scene>",  // mismatched end tag
            "<scene>\n<object name=\"x\">\n<mesh><points><p x=1/></points></mesh></object></scene>",  // unquoted
            "<scene><light type=\"sunlight\" name=\"s\"><from x=\"0\" y=\"0\" z=\"1\"/></light>",  // unclosed
            "<world/>",  // wrong root
            "<scene><teapot name=\"t\"/></scene>",  // unknown element
            "<scene><shader type=\"phong\" name=\"p\"/></scene>",  // not on the GPU path
            "<?xml version=\"1.0\"?><!-- c --><scene><light type=\"sunlight\" name=\"s\" power=\"2\"/></scene>",
        };
        for (const char* d : docs) {
            gpuInterface_t I;
            std::vector<paramMap_t> renders;
            const int rc = loadScene(d, I, renders);
            printf("error %d %s\n", rc, rc ? lastError().c_str() : "-");
        }
        return 0;
    }
    return 2;
}
