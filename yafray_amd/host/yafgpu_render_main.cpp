// yafgpu-render: renders a YafRay XML scene file on the GPU, the counterpart
// of the reference's command-line renderer (loader/loader.cc with
// render_t::render, render.cc:1052-1263): every <render> element of the file
// is rendered and written to its "outfile" (default salida.tga) as targa or,
// with output_type "hdr", as Radiance RGBE (exr falls back to hdr, as a
// reference build without OpenEXR does).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "yafgpu_xml.h"

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::cerr << "usage: " << argv[0] << " scene.xml\n";
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    if (!in) {
        std::cerr << "[Error] cannot open " << argv[1] << "\n";
        return 1;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    yafgpu::gpuInterface_t I;
    std::vector<yafgpu::paramMap_t> renders;
    if (yafgpu::loadScene(ss.str(), I, renders)) return 1;
    if (renders.empty()) std::cerr << "[Warning]: no <render> element\n";
    for (const auto& p : renders) {
        std::string cam, outfile = "salida.tga", type = "tga";
        bool alpha = false;
        p.getParam("camera_name", cam);
        p.getParam("outfile", outfile);
        p.getParam("output_type", type);
        p.getParam("save_alpha", alpha);
        int w = 0, h = 0;
        if (!I.cameraResolution(cam, w, h)) {
            std::cerr << "[Error] Camera " << cam << " does not exist\n";  // render.cc:1131-1140
            return 1;
        }
        std::vector<float> rgb;
        yg_render_stats st;
        if (I.render(p, rgb, &st)) return 1;
        std::cout << "Render: " << st.rays_camera << " camera, " << st.rays_extension << " extension, "
                  << st.rays_shadow << " shadow rays\n";
        const std::string warn = yafgpu::gpuInterface_t::renderWarnings(st);
        if (!warn.empty()) std::cerr << warn << "\n";
        int rc;
        if (type == "hdr" || type.find("exr") != std::string::npos) {
            if (type != "hdr") std::cout << "Built without OpenEXR: image saved in hdr format instead.\n";
            std::cout << "Saving HDR file as " << outfile << "\n";
            rc = yafgpu::writeHDR(outfile, w, h, rgb.data());
        } else {
            if (alpha) std::cerr << "[Warning]: save_alpha is not on the GPU path, the targa has no alpha\n";
            std::cout << "Saving Targa file as " << outfile << "\n";
            rc = yafgpu::writeTGA(outfile, w, h, rgb.data());
        }
        if (rc) return 1;
        std::cout << "OK\n";
    }
    return 0;
}
