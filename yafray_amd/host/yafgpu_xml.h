// YafRay XML scene files for the GPU path: the element grammar of
// src/gram.yafray and the element handlers of render_t (loader/render.cc),
// feeding gpuInterface_t.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "yafgpu_host.h"

namespace yafgpu {

// One XML element: name, attributes in file order, child elements (text
// content is not part of the scene grammar and is dropped).
struct xmlNode_t {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<xmlNode_t> kids;
};

// Parses a document; returns YG_OK or YG_ERR_ARG with the line in lastError().
int parseXML(const std::string& text, xmlNode_t& root);

// Loads <scene> into `I` in file order (shaders, textures, objects, lights,
// cameras, backgrounds, transforms) and returns the parameters of each
// <render> element, in order.
int loadScene(const std::string& text, gpuInterface_t& I, std::vector<paramMap_t>& renders);

// Image files of render_t::render (render.cc:1225-1257): 24/32-bit targa,
// top-left origin (targaIO.cc:83-131), and Radiance RGBE with run-length
// scanlines (HDR_io.cc:76-160).  rgb: w*h*3 floats, top row first.
int writeTGA(const std::string& path, int w, int h, const float* rgb);
int writeHDR(const std::string& path, int w, int h, const float* rgb);

}  // namespace yafgpu
