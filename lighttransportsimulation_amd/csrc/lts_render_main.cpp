// lts_render: the reference executable's flow (src/LightTransportSimulation.cu:23-146) over the C ABI.
//
//   lts_render [projectDir] [--seed N] [--integrator mega|wavefront|cached|bdpt]
//
// Reads <projectDir>/app.params (positional: W H / spp / bounces / roulette start / format /
// output / scene / subject, core/FileIO.cuh:234-257), loads the scene and subject relative to
// projectDir, renders (the plain path tracer unless --integrator names another; an unknown name
// exits with status 2 before any device call) and writes <projectDir>/<output>.pfm (format 1)
// or a P3 PPM to <projectDir>/<output>.hdr (format 0), as the reference names it. The scene and
// render stages print the reference's "(1) Scene creation" / "(2) Ray tracing" timing lines.
// The reference seeds with time(0) (LightTransportSimulation.cu:104); here the seed defaults to 0
// so that a render is reproducible, and --seed overrides it.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lts.h"

int main(int argc, char** argv) {
    std::string dir = ".";
    unsigned seed = 0;
    int integrator = LTS_INTEGRATOR_MEGAKERNEL;
    const char* technique = "Path Tracing";
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) seed = (unsigned)std::strtoul(argv[++i], nullptr, 0);
        else if (!std::strcmp(argv[i], "--integrator") && i + 1 < argc) {
            const std::string v = argv[++i];
            if (v == "mega") integrator = LTS_INTEGRATOR_MEGAKERNEL;
            else if (v == "wavefront") integrator = LTS_INTEGRATOR_WAVEFRONT;
            else if (v == "cached") integrator = LTS_INTEGRATOR_CACHED;
            else if (v == "bdpt") integrator = LTS_INTEGRATOR_BDPT;
            else {
                std::fprintf(stderr, "unknown integrator '%s' (mega, wavefront, cached or bdpt)\n", v.c_str());
                return 2;
            }
            technique = integrator == LTS_INTEGRATOR_BDPT     ? "Bidirectional Path Tracing"
                        : integrator == LTS_INTEGRATOR_CACHED ? "Path Tracing with a learned radiance cache"
                                                              : "Path Tracing";
        } else dir = argv[i];
    }
    if (!dir.empty() && dir.back() != '/') dir += '/';
    int32_t W, H, spp, bounce, rr, fmt;
    char out[1024], scene[1024], subject[1024];
    if (lts_read_app_params((dir + "app.params").c_str(), &W, &H, &spp, &bounce, &rr, &fmt, out, scene, subject, 1024) != LTS_OK) {
        std::fprintf(stderr, "cannot read %sapp.params\n", dir.c_str());
        return 1;
    }
    std::printf("---%s---\n", lts_version());
    std::printf("===RENDERING PARAMETERS===\n\tTechnique: %s\n\tImage resolution: %dx%d\n"
                "\tSamples per pixel: %d\n\tMaximum bounces: %d\n", technique, W, H, spp, bounce);
    lts_ctx* ctx = nullptr;
    if (lts_create(&ctx, 0) != LTS_OK) {
        std::fprintf(stderr, "no HIP device\n");
        return 1;
    }
    auto t0 = std::chrono::steady_clock::now();
    if (lts_load_scene(ctx, dir.c_str(), scene, subject, W, H) != LTS_OK) {
        std::fprintf(stderr, "scene load failed: %s\n", lts_last_error(ctx));
        lts_destroy(ctx);
        return 1;
    }
    auto t1 = std::chrono::steady_clock::now();
    std::printf("(1) Scene creation finished in %.3f seconds.\n", std::chrono::duration<double>(t1 - t0).count());
    lts_render_params p{};
    p.spp = spp; p.max_bounce = bounce; p.rr_start = rr; p.seed = seed; p.integrator = integrator;
    std::vector<float> rgb((size_t)W * H * 3);
    lts_stats st{};
    t0 = std::chrono::steady_clock::now();
    if (lts_render(ctx, &p, rgb.data(), nullptr, &st) != LTS_OK) {
        std::fprintf(stderr, "render failed: %s\n", lts_last_error(ctx));
        lts_destroy(ctx);
        return 1;
    }
    t1 = std::chrono::steady_clock::now();
    std::printf("(2) Ray tracing finished in %.3f seconds. (%.1f Msamples/s)\n",
                std::chrono::duration<double>(t1 - t0).count(), st.msamples_per_s);
    int rc = lts_write_image((dir + out).c_str(), rgb.data(), W, H, fmt);
    if (rc != LTS_OK) std::fprintf(stderr, "cannot write %s%s\n", dir.c_str(), out);
    lts_destroy(ctx);
    return rc == LTS_OK ? 0 : 1;
}
