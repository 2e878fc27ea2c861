// Command-line renderer over the C ABI, with the reference's options
// (scene/command_line_option.h, main.cu): pbrtm-render [--spp N]
// [--integrator path|volpath|megakernelpath|ambientocclusion|surfacenormal]
// [--outfile out.png] scene.pbrt
// Parses the file, builds the scene on device 0, renders with the file's
// Integrator (or --integrator, which overrides it as in scene_builder.cu:366-371)
// and writes the PNG (the Film filename unless --outfile; a .pfm name writes the
// linear film as floats).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pbrtm.h"

int main(int argc, char **argv) {
    std::string input, outfile, integrator;
    int spp = -1;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a.rfind("--", 0) == 0) {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", a.c_str());
                return 1;
            }
            if (a == "--spp") spp = std::atoi(argv[++i]);
            else if (a == "--outfile") outfile = argv[++i];
            else if (a == "--integrator") integrator = argv[++i];
            else {
                std::fprintf(stderr, "unknown arg: `%s`\n", a.c_str());
                return 1;
            }
        } else {
            input = a;
        }
    }
    if (input.empty()) {
        std::fprintf(stderr, "please provide a PBRT input file:\n$ pbrtm-render example.pbrt\n");
        return 1;
    }
    uint32_t integrator_flags = PBRTM_INTEGRATOR_PATH;
    if (!integrator.empty() && pbrtm_integrator_from_name(integrator.c_str(), &integrator_flags) != PBRTM_OK) {
        std::fprintf(stderr, "integrator `%s` is not supported\n", integrator.c_str());
        return 1;
    }
    char err[512] = {0};
    pbrtm_pbrt_file *file = nullptr;
    if (pbrtm_pbrt_parse(input.c_str(), &file, err, sizeof(err)) != PBRTM_OK) {
        std::fprintf(stderr, "%s: %s\n", input.c_str(), err);
        return 1;
    }
    pbrtm_render_params params;
    pbrtm_pbrt_render_params(file, &params);
    if (!integrator.empty()) params.flags = (params.flags & ~PBRTM_INTEGRATOR_MASK) | integrator_flags;
    if (spp > 0) {
        int s = 1;
        while ((s + 1) * (s + 1) <= spp) ++s;
        params.spp = uint32_t(s * s);  // stratified: rounded down to a square
    }
    const pbrtm_scene_desc *desc = pbrtm_pbrt_desc(file);
    if (outfile.empty()) outfile = pbrtm_pbrt_output(file);
    pbrtm_ctx *ctx = nullptr;
    int rc = pbrtm_create(&ctx, 0);
    if (rc != PBRTM_OK) {
        std::fprintf(stderr, "no HIP device (%d)\n", rc);
        pbrtm_pbrt_free(file);
        return 1;
    }
    pbrtm_scene *scene = nullptr;
    rc = pbrtm_scene_create(ctx, desc, &scene);
    std::vector<float> rgb;
    pbrtm_render_stats st{};
    if (rc == PBRTM_OK) {
        rgb.resize(size_t(desc->width) * desc->height * 3);
        rc = pbrtm_render(ctx, scene, &params, rgb.data(), nullptr, &st);
    }
    if (rc != PBRTM_OK) std::fprintf(stderr, "error: %s\n", pbrtm_last_error(ctx));
    const bool pfm = outfile.size() > 4 && outfile.compare(outfile.size() - 4, 4, ".pfm") == 0;
    if (rc == PBRTM_OK)
        rc = pfm ? pbrtm_write_pfm(rgb.data(), desc->width, desc->height, outfile.c_str())
                 : pbrtm_write_png(rgb.data(), desc->width, desc->height, outfile.c_str());
    if (rc == PBRTM_OK)
        std::printf("%ux%u, %u spp, maxdepth %u: %.2f ms (%.1f Msamples/s), %llu rays, %llu shadow rays -> %s\n",
                    desc->width, desc->height, params.spp, params.max_depth, st.ms_total,
                    double(desc->width) * desc->height * params.spp / (st.ms_total * 1e3),
                    (unsigned long long)st.rays, (unsigned long long)st.shadow_rays, outfile.c_str());
    if (scene) pbrtm_scene_free(scene);
    pbrtm_destroy(ctx);
    pbrtm_pbrt_free(file);
    return rc == PBRTM_OK ? 0 : 1;
}
