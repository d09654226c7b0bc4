// YafRay XML scene files (src/gram.yafray; render_t in loader/render.cc) and
// the image files render_t writes, for the GPU path.
#include "yafgpu_xml.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

namespace yafgpu {

// ------------------------------------------------------------- parser --

namespace {

struct Reader {
    const std::string& s;
    size_t i = 0;
    int line = 1;
    explicit Reader(const std::string& t) : s(t) {}
    bool eof() const { return i >= s.size(); }
    char peek() const { return eof() ? '\0' : s[i]; }
    void adv(size_t n = 1)
    {
        for (size_t k = 0; k < n && i < s.size(); k++, i++)
            if (s[i] == '\n') line++;
    }
    bool starts(const char* p) const { return s.compare(i, strlen(p), p) == 0; }
    void skip_ws()
    {
        while (!eof() && isspace((unsigned char)s[i])) adv();
    }
    // skips to just after `end`; false at end of input
    bool skip_past(const char* end)
    {
        const size_t k = s.find(end, i);
        if (k == std::string::npos) return false;
        adv(k + strlen(end) - i);
        return true;
    }
};

int xml_error(const Reader& r, const std::string& what)
{
    return hostFail(YG_ERR_ARG, "scene file, line " + std::to_string(r.line) + ": " + what);
}

std::string decode_entities(const std::string& v)
{
    std::string o;
    for (size_t k = 0; k < v.size(); k++) {
        if (v[k] != '&') { o += v[k]; continue; }
        static const char* names[] = {"&amp;", "&lt;", "&gt;", "&quot;", "&apos;"};
        static const char chars[] = {'&', '<', '>', '"', '\''};
        bool hit = false;
        for (int q = 0; q < 5 && !hit; q++)
            if (v.compare(k, strlen(names[q]), names[q]) == 0) {
                o += chars[q];
                k += strlen(names[q]) - 1;
                hit = true;
            }
        if (!hit) o += '&';
    }
    return o;
}

bool name_char(char c) { return isalnum((unsigned char)c) || c == '_' || c == '-' || c == ':' || c == '.'; }

// comments, processing instructions and declarations between elements
int skip_misc(Reader& r)
{
    for (;;) {
        r.skip_ws();
        if (r.starts("<!--")) {
            if (!r.skip_past("-->")) return xml_error(r, "unterminated comment");
        } else if (r.starts("<?")) {
            if (!r.skip_past("?>")) return xml_error(r, "unterminated processing instruction");
        } else if (r.starts("<!")) {
            if (!r.skip_past(">")) return xml_error(r, "unterminated declaration");
        } else if (!r.eof() && r.peek() != '<') {
            r.adv();  // character data: not part of the scene grammar
        } else {
            return YG_OK;
        }
    }
}

int parse_element(Reader& r, xmlNode_t& n, int depth)
{
    if (depth > 256) return xml_error(r, "elements nested too deeply");
    if (r.peek() != '<') return xml_error(r, "expected an element");
    r.adv();
    const size_t b = r.i;
    while (!r.eof() && name_char(r.peek())) r.adv();
    if (r.i == b) return xml_error(r, "element without a name");
    n.name = r.s.substr(b, r.i - b);
    for (;;) {
        r.skip_ws();
        if (r.eof()) return xml_error(r, "unterminated element <" + n.name + ">");
        if (r.starts("/>")) {
            r.adv(2);
            return YG_OK;
        }
        if (r.peek() == '>') {
            r.adv();
            break;
        }
        const size_t ab = r.i;
        while (!r.eof() && name_char(r.peek())) r.adv();
        if (r.i == ab) return xml_error(r, "bad attribute in <" + n.name + ">");
        const std::string key = r.s.substr(ab, r.i - ab);
        r.skip_ws();
        if (r.peek() != '=') return xml_error(r, "attribute " + key + " without a value");
        r.adv();
        r.skip_ws();
        const char q = r.peek();
        if (q != '"' && q != '\'') return xml_error(r, "attribute " + key + ": value not quoted");
        r.adv();
        const size_t vb = r.i;
        while (!r.eof() && r.peek() != q) r.adv();
        if (r.eof()) return xml_error(r, "attribute " + key + ": unterminated value");
        n.attrs.emplace_back(key, decode_entities(r.s.substr(vb, r.i - vb)));
        r.adv();
    }
    for (;;) {
        int rc = skip_misc(r);
        if (rc) return rc;
        if (r.eof()) return xml_error(r, "missing </" + n.name + ">");
        if (r.starts("</")) {
            r.adv(2);
            const size_t eb = r.i;
            while (!r.eof() && name_char(r.peek())) r.adv();
            if (r.s.compare(eb, r.i - eb, n.name) != 0 || r.i - eb != n.name.size())
                return xml_error(r, "</" + r.s.substr(eb, r.i - eb) + "> closes <" + n.name + ">");
            r.skip_ws();
            if (r.peek() != '>') return xml_error(r, "bad end tag of <" + n.name + ">");
            r.adv();
            return YG_OK;
        }
        n.kids.emplace_back();
        if ((rc = parse_element(r, n.kids.back(), depth + 1))) return rc;
    }
}

}  // namespace

int parseXML(const std::string& text, xmlNode_t& root)
{
    Reader r(text);
    int rc = skip_misc(r);
    if (rc) return rc;
    if (r.eof()) return xml_error(r, "no root element");
    if ((rc = parse_element(r, root, 0))) return rc;
    if ((rc = skip_misc(r))) return rc;
    if (!r.eof()) return xml_error(r, "content after the root element");
    return YG_OK;
}

// ------------------------------------------------------------- loader --

namespace {

bool as_float(const std::string& v, float& f)
{
    if (v.empty()) return false;
    char* e = nullptr;
    f = strtof(v.c_str(), &e);
    return e && *e == '\0';
}

// an attribute as render_t's parser types it: T_FLOAT values are numbers
// (also readable as integers when whole); "on"/"off" read as booleans
void set_value(paramMap_t& p, const std::string& k, const std::string& v)
{
    float f;
    if (as_float(v, f)) {
        p.setFloat(k, f);
        if (std::floor(f) == f && std::fabs(f) < 2147483520.0f) p.setInt(k, (int)f);
        return;
    }
    p.setString(k, v);
    if (v == "on" || v == "true") p.setBool(k, true);
    else if (v == "off" || v == "false") p.setBool(k, false);
}

const std::string* attr(const xmlNode_t& n, const char* k)
{
    for (const auto& a : n.attrs)
        if (a.first == k) return &a.second;
    return nullptr;
}

// A child label <name .../> (msin_general.cc): a colour when its attributes
// are 3 or 4 numbers named r, g, b (a); a point for exactly x, y, z; a value
// for one attribute "value"; other labels carry no parameter.
void add_label(paramMap_t& p, const xmlNode_t& l)
{
    bool rgb = (l.attrs.size() == 3 || l.attrs.size() == 4), xyz = l.attrs.size() == 3;
    float c[4] = {0, 0, 0, 1}, x[3] = {0, 0, 0};
    for (const auto& a : l.attrs) {
        float f;
        const bool num = as_float(a.second, f) && a.first.size() == 1;
        const char ch = a.first.empty() ? 0 : a.first[0];
        if (!num || !strchr("rgba", ch)) rgb = false;
        else c[ch == 'r' ? 0 : ch == 'g' ? 1 : ch == 'b' ? 2 : 3] = f;
        if (!num || !strchr("xyz", ch)) xyz = false;
        else x[ch - 'x'] = f;
    }
    if (rgb) p.setColor(l.name, c[0], c[1], c[2]);
    else if (xyz) p.setPoint(l.name, x[0], x[1], x[2]);
    else if (l.attrs.size() == 1 && l.attrs[0].first == "value") set_value(p, l.name, l.attrs[0].second);
}

static bool has_descendant(const xmlNode_t& n, const char* name)
{
    for (const auto& k : n.kids)
        if (k.name == name || (k.name != "mesh" && has_descendant(k, name))) return true;
    return false;
}

paramMap_t generic_params(const xmlNode_t& n)
{
    paramMap_t p;
    for (const auto& a : n.attrs) set_value(p, a.first, a.second);
    for (const auto& k : n.kids)
        if (k.kids.empty()) add_label(p, k);
    return p;
}

struct Loader {
    gpuInterface_t& I;
    std::vector<paramMap_t>& renders;
    float M[4][4];  // current transform (render_t::M, a float matrix4x4_t)

    int item(const xmlNode_t& n);

    int transform(const xmlNode_t& n)
    {
        // render_t::transform (render.cc:114-140): M = old * T for the enclosed
        // items, the float product of matrix4.h:65-77 (sums from 0 over j)
        float T[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) {
                const std::string key = "m" + std::to_string(r) + std::to_string(c);
                float f;
                const std::string* v = attr(n, key.c_str());
                if (v && as_float(*v, f)) T[r][c] = f;
            }
        float old[4][4];
        memcpy(old, M, sizeof(M));
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) {
                float s = 0;
                for (int k = 0; k < 4; k++) {
                    const float p = old[r][k] * T[k][c];
                    s += p;
                }
                M[r][c] = s;
            }
        int rc = YG_OK;
        for (const auto& k : n.kids)
            if ((rc = item(k))) break;
        memcpy(M, old, sizeof(M));
        return rc;
    }

    int object(const xmlNode_t& n)
    {
        // render_t::object / render_t::mesh (render.cc:647-768)
        const std::string* name = attr(n, "name");
        const std::string* shader = attr(n, "shader_name");
        if (!name || name->empty()) return hostFail(YG_ERR_ARG, "object without a name");
        if (attr(n, "original")) return hostFail(YG_ERR_UNSUPPORTED, "object references (original=) are not on the GPU path");
        const xmlNode_t* mesh = nullptr;
        bool sphere = false;
        for (const auto& k : n.kids)
            if (k.name == "mesh") mesh = &k;
            else if (k.name == "sphere") sphere = true;
        if (!mesh && sphere) {
            // render_t::sphere returns NULL (render.cc:768-776): the reference
            // adds no object for a sphere
            std::cerr << "[Warning]: sphere object " << *name << " ignored (the scene loader makes no spheres)\n";
            return YG_OK;
        }
        if (!mesh) return hostFail(YG_ERR_UNSUPPORTED, "object " + *name + ": only meshes are on the GPU path");
        // Settings that change what the reference renders for this object and
        // that the GPU path does not implement are refused, not ignored:
        // caustic settings turn on obj->caustics (msin_object.cc:448-476),
        // which makes the pathlight follow caustic bounces (pathlight.cc:194-222).
        // Face uv, vertex colours and orco points are read: the uv tangents
        // they give (meshObject_t::tangentsFromUV, mesh.cc:284-340) set only
        // surfacePoint_t's TU/TV, which the generic shader and the lights do
        // not read (the pathlight samples around NU/NV, createCS of N), and
        // the generic shader does not read vertex colours.
        if (attr(n, "caus_IOR") || has_descendant(n, "caus_rcolor") || has_descendant(n, "caus_tcolor"))
            return hostFail(YG_ERR_UNSUPPORTED, "object " + *name + ": caustic settings (caus_IOR, caus_rcolor, "
                                                "caus_tcolor) are not on the GPU path");
        bool has_orco = false;
        if (const std::string* orco = attr(*mesh, "has_orco")) has_orco = (*orco == "on");
        float smooth_angle = 0;
        if (const std::string* as = attr(*mesh, "autosmooth")) {
            // render.cc:764 with msin_object.cc:271-280: a float angle; below 1
            // degree meshObject_t::autoSmooth does nothing (mesh.cc:204)
            if (!as_float(*as, smooth_angle)) {
                std::cerr << "[Warning]: Only float value accepted for autosmooth\n";
                smooth_angle = 0;
            }
        }
        bool shadow = true, emit = true, recv = true;
        paramMap_t op = generic_params(n);
        op.getParam("shadow", shadow);
        op.getParam("emit_rad", emit);
        op.getParam("recv_rad", recv);
        std::vector<float> raw, v;
        std::vector<uint32_t> f;
        std::vector<std::string> sh;
        std::vector<float> uv;
        std::vector<uint8_t> fflags;
        bool any_tex = false;
        for (const auto& part : mesh->kids) {
            if (part.name == "points") {
                for (const auto& pt : part.kids) {
                    float x[3] = {0, 0, 0};
                    for (const auto& a : pt.attrs) {
                        float val;
                        if (a.first.size() == 1 && strchr("xyz", a.first[0]) && as_float(a.second, val))
                            x[a.first[0] - 'x'] = val;
                    }
                    raw.insert(raw.end(), x, x + 3);
                }
            } else if (part.name == "faces") {
                for (const auto& fc : part.kids) {
                    float a = 0, b = 0, c = 0;
                    const std::string *sa = attr(fc, "a"), *sb = attr(fc, "b"), *sc = attr(fc, "c");
                    if (!sa || !sb || !sc || !as_float(*sa, a) || !as_float(*sb, b) || !as_float(*sc, c))
                        return hostFail(YG_ERR_ARG, "object " + *name + ": face without a, b, c");
                    f.push_back((uint32_t)a);
                    f.push_back((uint32_t)b);
                    f.push_back((uint32_t)c);
                    // join_face (msin_object.cc:116-213): u_a..v_c into uv[0..5]
                    // (u at even, v at odd places), vcol_* set has_vcol; a face
                    // that names some of them keeps 0 for the rest
                    float fuv[6] = {0, 0, 0, 0, 0, 0};
                    uint8_t fl = 0;
                    static const char* uvk[6] = {"u_a", "v_a", "u_b", "v_b", "u_c", "v_c"};
                    for (int k = 0; k < 6; k++)
                        if (const std::string* t = attr(fc, uvk[k])) {
                            float val;
                            if (as_float(*t, val)) {
                                fuv[k] = val;
                                fl |= 1;
                            }
                        }
                    for (const auto& at : fc.attrs)
                        if (at.first.compare(0, 5, "vcol_") == 0) fl |= 2;
                    uv.insert(uv.end(), fuv, fuv + 6);
                    fflags.push_back(fl);
                    if (fl) any_tex = true;
                    const std::string* fs = attr(fc, "shader_name");
                    sh.push_back(fs ? *fs : (shader ? *shader : std::string()));
                }
            }
        }
        const uint32_t np = (uint32_t)(raw.size() / 3);
        for (uint32_t& k : f)
            if (k >= np) {
                std::cerr << "[Warning]: Point " << k << " out of bounds in object\n";  // render.cc:719-731
                k = 0;
            }
        // meshObject_t::transform (mesh.cc:422-469): vertices = M * p in float;
        // with has_orco every second point is the orco point of the one
        // before it and keeps its untransformed value
        v.resize(raw.size());
        std::vector<float> orco;
        if (has_orco) orco.assign(raw.size(), 0.0f);
        for (uint32_t i = 0; i < np; i++) {
            const float* x = &raw[3 * (size_t)i];
            const bool is_orco = has_orco && (i & 1);
            for (int r = 0; r < 3; r++) {
                float acc = M[r][0] * x[0];
                const float t1 = M[r][1] * x[1];
                acc += t1;
                const float t2 = M[r][2] * x[2];
                acc += t2;
                acc += M[r][3];
                v[3 * (size_t)i + r] = is_orco ? x[r] : acc;
            }
            if (has_orco && i + 1 < np)
                for (int r = 0; r < 3; r++) orco[3 * (size_t)i + r] = raw[3 * (size_t)(i + 1) + r];
        }
        int rc = I.addObject_trimesh(*name, v, f, sh, shadow, recv, emit, smooth_angle);
        if (rc || f.empty()) return rc;
        std::vector<float> mtx(12);
        if ((rc = yg_orco_matrix(&M[0][0], raw.data(), np, f.data(), (uint32_t)(f.size() / 3), mtx.data()))) {
            char msg[256];
            yg_last_error(msg, sizeof(msg));
            return hostFail(rc, "object " + *name + ": " + msg);
        }
        if (!any_tex) {
            uv.clear();
            fflags.clear();
        }
        return I.setObjectTexcoords(*name, uv, fflags, orco, mtx);
    }

    int shader(const xmlNode_t& n)
    {
        // render_t::shader (render.cc:231-273): base params from the element
        // and its <attributes>, one parameter map per <modulator>
        const std::string* name = attr(n, "name");
        const std::string* type = attr(n, "type");
        if (!name || !type) return hostFail(YG_ERR_ARG, "shader without name or type");
        paramMap_t p;
        for (const auto& a : n.attrs) set_value(p, a.first, a.second);
        std::vector<paramMap_t> mods;
        for (const auto& k : n.kids) {
            if (k.name == "attributes")
                for (const auto& l : k.kids) add_label(p, l);
            else if (k.name == "modulator")
                mods.push_back(generic_params(k));
        }
        return I.addShader(*name, *type, p, mods);
    }
};

int Loader::item(const xmlNode_t& n)
{
    const std::string* name = attr(n, "name");
    const std::string* type = attr(n, "type");
    if (n.name == "shader") return shader(n);
    if (n.name == "object") return object(n);
    if (n.name == "transform") return transform(n);
    if (n.name == "texture") {
        if (!name || !type) return hostFail(YG_ERR_ARG, "texture without name or type");
        return I.addTexture(*name, *type, generic_params(n));
    }
    if (n.name == "light") {
        if (!name || !type) return hostFail(YG_ERR_ARG, "light without name or type");
        return I.addLight(*name, *type, generic_params(n));
    }
    if (n.name == "camera") {
        if (!name) return hostFail(YG_ERR_ARG, "camera without a name");
        const std::string* ct = attr(n, "type");
        if (ct && *ct != "perspective")
            return hostFail(YG_ERR_UNSUPPORTED, "camera type '" + *ct + "' is not on the GPU path");
        return I.addCamera(*name, generic_params(n));
    }
    if (n.name == "background") {
        if (!name || !type) return hostFail(YG_ERR_ARG, "background without name or type");
        if (*type != "constant") return hostFail(YG_ERR_UNSUPPORTED, "background type '" + *type + "' is not on the GPU path");
        return I.addBackground(*name, generic_params(n));
    }
    if (n.name == "render") {
        renders.push_back(generic_params(n));
        return YG_OK;
    }
    if (n.name == "filter") {
        std::cerr << "[Warning]: filters are not on the GPU path, ignored\n";
        return YG_OK;
    }
    return hostFail(YG_ERR_ARG, "unknown scene element <" + n.name + ">");
}

}  // namespace

int loadScene(const std::string& text, gpuInterface_t& I, std::vector<paramMap_t>& renders)
{
    xmlNode_t root;
    int rc = parseXML(text, root);
    if (rc) return rc;
    if (root.name != "scene") return hostFail(YG_ERR_ARG, "the root element is <" + root.name + ">, not <scene>");
    Loader L{I, renders, {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}}};
    for (const auto& k : root.kids)
        if ((rc = L.item(k))) return rc;
    return YG_OK;
}

// -------------------------------------------------------- image files --

int writeTGA(const std::string& path, int w, int h, const float* rgb)
{
    if (w <= 0 || h <= 0 || w > 65535 || h > 65535) return hostFail(YG_ERR_ARG, "targa size");
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp) return hostFail(YG_ERR_ARG, "cannot write " + path);
    // uncompressed true-colour header, 24 bits, top-left origin (targaIO.cc:14, 92-112)
    const unsigned char head[18] = {0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, (unsigned char)(w & 255),
                                    (unsigned char)(w >> 8), (unsigned char)(h & 255), (unsigned char)(h >> 8),
                                    0x18, 0x20};
    bool ok = fwrite(head, 1, 18, fp) == 18;
    std::vector<unsigned char> line((size_t)w * 3);
    // 8-bit channels as color.cc:35-40 converts them, written B, G, R
    auto byte = [](float c) -> unsigned char { return (c < 0.f) ? 0 : ((c >= 1.f) ? 255 : (unsigned char)(255.f * c)); };
    for (int y = 0; y < h && ok; y++) {
        for (int x = 0; x < w; x++) {
            const float* p = rgb + 3 * ((size_t)y * w + x);
            line[3 * x] = byte(p[2]);
            line[3 * x + 1] = byte(p[1]);
            line[3 * x + 2] = byte(p[0]);
        }
        ok = fwrite(line.data(), 1, line.size(), fp) == line.size();
    }
    ok = (fclose(fp) == 0) && ok;
    return ok ? YG_OK : hostFail(YG_ERR_ARG, "write error on " + path);
}

int writeHDR(const std::string& path, int w, int h, const float* rgb)
{
    if (w <= 0 || h <= 0) return hostFail(YG_ERR_ARG, "hdr size");
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp) return hostFail(YG_ERR_ARG, "cannot write " + path);
    fprintf(fp, "#?RADIANCE\n# Created with YafRay\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=%25.13f\n\n-Y %d +X %d\n", 1.0, h, w);
    std::vector<unsigned char> px((size_t)w * 4), out;
    bool ok = true;
    for (int y = 0; y < h && ok; y++) {
        // shared-exponent pixels (HDR_io.cc:76-94): mantissas scaled by
        // 256 / 2^e where max channel = m 2^e, m in [0.5, 1)
        for (int x = 0; x < w; x++) {
            const float* p = rgb + 3 * ((size_t)y * w + x);
            const float d = std::max(p[0], std::max(p[1], p[2]));
            unsigned char* q = &px[4 * (size_t)x];
            if (d <= 1e-32f) {
                q[0] = q[1] = q[2] = q[3] = 0;
            } else {
                int e;
                const float s = (float)(frexp(d, &e) * 256.f / d);
                q[0] = (unsigned char)(p[0] * s);
                q[1] = (unsigned char)(p[1] * s);
                q[2] = (unsigned char)(p[2] * s);
                q[3] = (unsigned char)(e + 128);
            }
        }
        out.clear();
        if (w < 8 || w > 0x7fff) {
            out.assign(px.begin(), px.end());  // flat scanline
        } else {
            // run-length scanline: marker, then each channel as runs (count
            // 128 + n, one byte) and literal spans (count n, n bytes)
            out.push_back(2);
            out.push_back(2);
            out.push_back((unsigned char)(w >> 8));
            out.push_back((unsigned char)(w & 255));
            for (int c = 0; c < 4; c++) {
                int x = 0;
                while (x < w) {
                    int run = 1;
                    while (x + run < w && run < 127 && px[4 * (x + run) + c] == px[4 * x + c]) run++;
                    if (run >= 4) {
                        out.push_back((unsigned char)(128 + run));
                        out.push_back(px[4 * x + c]);
                        x += run;
                        continue;
                    }
                    // literal span up to the next run of 4 or 128 bytes
                    int n = 0;
                    while (x + n < w && n < 128) {
                        int r = 1;
                        while (x + n + r < w && r < 4 && px[4 * (x + n + r) + c] == px[4 * (x + n) + c]) r++;
                        if (r >= 4) break;
                        n++;
                    }
                    out.push_back((unsigned char)n);
                    for (int k = 0; k < n; k++) out.push_back(px[4 * (x + k) + c]);
                    x += n;
                }
            }
        }
        ok = fwrite(out.data(), 1, out.size(), fp) == out.size();
    }
    ok = (fclose(fp) == 0) && ok;
    return ok ? YG_OK : hostFail(YG_ERR_ARG, "write error on " + path);
}

}  // namespace yafgpu
