// Render settings parsed from a flat YAML config (path_tracer/core/config.h:15-201).
#pragma once
#include <optional>
#include <string>

#include "../../include/rpd_engine.h"
#include "flat_yaml.h"

namespace host {

struct Settings {
    rpd_render_config render{};
    float fov = 45.0f;
    float focus_distance = 1.0f;
    float aperture = 0.0f;
    float camera_position[3] = {0, 0, 0};
    float camera_direction[3] = {0, 0, 1};
    float camera_up[3] = {0, 1, 0};
    std::optional<std::string> scene_name;
    std::optional<std::string> output_dir;
    std::string tonemap = "agx";
    bool debug_statmc_outputs = false;
};

// Strict parse: unknown keys, missing required keys and out-of-range values
// throw std::runtime_error with the reference's messages.
Settings parse_settings(const FlatYaml& y);

}  // namespace host
