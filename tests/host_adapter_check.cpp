// Drives the C++ host adapter (factories by plugin name) without a GPU and
// prints the descriptors it would hand to libyafgpu.so, one per line.
#include <cstdio>
#include <cstring>

#include "yafgpu_host.h"

using yafgpu::paramMap_t;

static void pl(const yg_light& L)
{
    printf("light %d %d %d %g %g %g %g %g %g %g %d %d %d %d %g %g %g %g %d %d %d %g %g %g\n", L.type,
           L.use_in_render, L.use_in_indirect, L.color[0], L.color[1], L.color[2], L.power, L.dir[0], L.dir[1],
           L.dir[2], L.cast_shadows, L.samples, L.depth, L.use_qmc, L.threshold, L.from[2], L.to[2], L.angle,
           L.photons, L.search, L.mindepth, L.bias, L.fixedradius, L.cluster);
}

int main()
{
    yafgpu::gpuInterface_t I;
    paramMap_t sun;
    sun.setPoint("from", 1, 0.5f, 1.5f);
    sun.setFloat("power", 3.0f);
    paramMap_t path;
    path.setInt("samples", 4);
    path.setBool("use_QMC", true);
    path.setFloat("threshold", 1000.0f);
    paramMap_t ph;
    ph.setBool("use_QMC", true);
    ph.setString("mode", "diffuse");
    ph.setFloat("fixedradius", 5.0f);
    ph.setInt("photons", 3998792);
    ph.setInt("depth", 1);
    ph.setInt("mindepth", 0);
    paramMap_t area;
    area.setPoint("a", 0, 0, 1);
    area.setInt("samples", 16);
    area.setFloat("power", 10.0f);
    area.setColor("color", 1, 1, 1);
    // names out of alphabetical order on purpose: the interface sorts them
    if (I.addLight("sun", "sunlight", sun) || I.addLight("path", "pathlight", path) ||
        I.addLight("photons", "photonlight", ph) || I.addLight("area", "arealight", area))
        return 1;
    paramMap_t mus;
    mus.setString("type", "fBm");
    mus.setFloat("octaves", 4.0f);
    mus.setFloat("intensity", 0.5f);
    if (I.addTexture("mus", "musgrave", mus)) return 2;
    paramMap_t sh;
    sh.setColor("color", 0.5f, 0.4f, 0.3f);
    paramMap_t mod;
    mod.setString("texname", "mus");
    mod.setString("mode", "mul");
    mod.setString("texco", "global");
    mod.setFloat("color", 0.6f);
    mod.setFloat("size", 3.0f);
    if (I.addShader("ground", "generic", sh, {mod})) return 3;
    std::vector<yg_light> lights;
    std::vector<yg_material> mats;
    std::vector<yg_modulator> mods;
    std::vector<yg_texture> texs;
    if (I.describe(lights, mats, mods, texs)) return 4;
    for (auto& L : lights) pl(L);
    for (auto& m : mats) printf("mat %g %g %g %d %d\n", m.color[0], m.color[1], m.color[2], m.mod_first, m.mod_count);
    for (auto& m : mods)
        printf("mod %d %d %g %g %g %g\n", m.tex, m.mode, m.color, m.size[0], m.size[1], m.size[2]);
    for (auto& t : texs) printf("tex %d %g %g %g %g\n", t.type, t.H, t.lacunarity, t.octaves, t.iscale);
    // render options: defaults, and the AA values clamped to their ranges
    paramMap_t r0, r1, r2;
    r1.setInt("AA_passes", 2);
    r1.setInt("AA_minsamples", 0);
    r1.setFloat("AA_pixelwidth", 2.5f);
    r1.setFloat("AA_threshold", 0.03f);
    r1.setBool("AA_jitterfirst", true);
    r1.setFloat("bias", 0.002f);
    r2.setInt("samples", 3);  // obsolete name: first pass included
    r2.setFloat("tolerance", 0.1f);
    r2.setFloat("AA_pixelwidth", 0.5f);
    for (const paramMap_t* p : {&r0, &r1, &r2}) {
        const yg_render_opts o = yafgpu::gpuInterface_t::render_opts(*p);
        printf("ropts %g %d %d %d %g %g %d\n", o.self_bias, o.raydepth, o.aa_passes, o.aa_minsamples,
               o.aa_pixelwidth, o.aa_threshold, o.aa_jitterfirst);
    }
    // noise_type names as newNoise (basictex.cc:38-63) reads them
    for (const char* nt : {"blender", "stdperlin", "newperlin", "", "anything", "voronoi_f1", "cellnoise"}) {
        paramMap_t tp;
        tp.setString("noise_type", nt);
        yg_texture T;
        const int rc = yafgpu::textureFactory("clouds", tp, T);
        printf("noise %s %d %d\n", *nt ? nt : "-", rc, rc ? -1 : T.noise);
    }
    // Musgrave type names (textureMusgrave_t ctor, basictex.cc:657-672), offset and gain
    for (const char* mt : {"fBm", "multifractal", "heteroterrain", "hybridmf", "ridgedmf", "", "other"}) {
        paramMap_t tp;
        tp.setString("type", mt);
        tp.setFloat("offset", 0.7f);
        tp.setFloat("gain", 2.0f);
        yg_texture T;
        const int rc = yafgpu::textureFactory("musgrave", tp, T);
        printf("mtype %s %d %d %g %g\n", *mt ? mt : "-", rc, T.mtype, T.offset, T.gain);
    }
    // modulator projection axes and texture matrix (basicshaders.cc:300-357)
    {
        paramMap_t sh2, m2;
        m2.setString("texname", "mus");
        m2.setString("texco", "global");
        m2.setString("proj_x", "z");
        m2.setString("proj_y", "n");
        m2.setString("proj_z", "x");
        m2.setFloat("m03", 0.25f);
        m2.setFloat("m11", 2.0f);
        std::vector<yg_modulator> out;
        yg_material M;
        const std::map<std::string, int> tix{{"mus", 0}};
        const int rc = yafgpu::shaderFactory("generic", sh2, {m2}, tix, M, out);
        printf("modmap %d %d %d %d %d %g %g %g\n", rc, out[0].proj[0], out[0].proj[1], out[0].proj[2], out[0].has_mtx,
               out[0].mtx[3], out[0].mtx[5], out[0].mtx[0]);
        paramMap_t m3;
        m3.setString("texname", "mus");
        m3.setString("texco", "global");
        out.clear();
        const int rc3 = yafgpu::shaderFactory("generic", sh2, {m3}, tix, M, out);
        printf("moddef %d %d %d %d %d\n", rc3, out[0].proj[0], out[0].proj[1], out[0].proj[2], out[0].has_mtx);
        // texco names (string2texcotype, texture.h:135-149; factory default "orco")
        const char* names[] = {"global", "orco", "uv", "window", "normal", "reflect", "sticky", "-"};
        for (const char* nm : names) {
            paramMap_t m4;
            m4.setString("texname", "mus");
            if (strcmp(nm, "-") != 0) m4.setString("texco", nm);
            out.clear();
            const int rc4 = yafgpu::shaderFactory("generic", sh2, {m4}, tix, M, out);
            printf("texco %s %d %d\n", nm, rc4, rc4 ? -1 : out[0].texco);
        }
    }
    // unsupported features are refused with a status, not an exit
    paramMap_t cache;
    cache.setBool("cache", true);
    cache.setBool("show_samples", true);
    printf("refused %d\n", I.addLight("cached", "pathlight", cache));
    // the irradiance cache's parameters (pathlight.cc:785-797): search at least 3
    {
        paramMap_t pc;
        pc.setBool("cache", true);
        pc.setBool("use_QMC", true);
        pc.setFloat("cache_size", 0.02f);
        pc.setFloat("shadow_threshold", 0.25f);
        pc.setInt("search", 1);
        yg_light L{};
        const int rc = yafgpu::lightFactory("pathlight", pc, L);
        printf("pcache %d %d %g %g %d\n", rc, L.cache, L.cache_size, L.shadow_threshold, L.search);
    }
    // warnings of a finished render: adaptive repetitions counted but not run
    yg_render_stats st{};
    printf("warn0 %d\n", (int)yafgpu::gpuInterface_t::renderWarnings(st).size());
    st.adaptive_extra = 17;
    const std::string w = yafgpu::gpuInterface_t::renderWarnings(st);
    printf("warn1 %d %d\n", (int)(w.find("[Warning]: 17 pixels") == 0), (int)(w.find("threshold") != std::string::npos));
    return 0;
}
