// Parses a flat YAML render config exactly as bin/path_tracer does and prints
// the resolved settings (one "key value" line each), or the error message and
// exit status 1 -- the host-side checks of path_tracer/tests/config_test.cpp
// without a GPU.
#include <cstdio>
#include <exception>

#include "config.h"
#include "flat_yaml.h"

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: config_check <config.yaml>\n");
        return 2;
    }
    try {
        const host::Settings s = host::parse_settings(host::FlatYaml::load_file(argv[1]));
        const rpd_render_config& r = s.render;
        std::printf("image_width %d\nimage_height %d\nsamples_per_pixel %d\nuse_statmc %d\n", r.image_width,
                    r.image_height, r.samples_per_pixel, r.use_statmc);
        std::printf("rpf_tile_size %d\ncolor_window_radius %d\ncolor_sigma_max %g\nvar_iterations %d\n",
                    r.rpf_tile_size, r.color_window_radius, r.color_sigma_max, r.var_iterations);
        std::printf("adaptive_spp %d\nadaptive_passes %d\nadaptive_importance_smoothing_radius %d\n", r.adaptive_spp,
                    r.adaptive_passes, r.adaptive_importance_smoothing_radius);
        std::printf("fov %g\nfocus_distance %g\naperture %g\n", s.fov, s.focus_distance, s.aperture);
        std::printf("camera_up %g %g %g\n", s.camera_up[0], s.camera_up[1], s.camera_up[2]);
        std::printf("tonemap %s\nscene %s\n", s.tonemap.c_str(), s.scene_name.value_or("").c_str());
    } catch (const std::exception& e) {
        std::printf("error %s\n", e.what());
        return 1;
    }
    return 0;
}
