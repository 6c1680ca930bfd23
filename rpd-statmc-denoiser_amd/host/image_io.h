// Output images: Radiance RGBE (.hdr) and 8-bit PNG with the reference's
// tone mapping presets (shared/image.h:29-137).
#pragma once
#include <string>
#include <vector>

namespace host {

enum class ToneMapping { AgxDefault, AgxGolden, AgxPunchy, Aces };

// case-insensitive preset names of main.cpp:257-264
bool tonemap_from_string(const std::string& name, ToneMapping& out);

// rgb: width*height*3 floats, row 0 at the top
bool write_hdr(const std::string& path, int width, int height, const float* rgb);
bool write_png(const std::string& path, int width, int height, const unsigned char* rgb8);
bool write_png_tonemapped(const std::string& path, int width, int height, const float* rgb, ToneMapping tm);
// scalar image written as grey RGB (core/buffers.h:170-180)
bool write_hdr_scalar(const std::string& path, int width, int height, const float* v);

// exposed for tests
void tonemap_pixel(const float in[3], ToneMapping tm, float out[3]);

}  // namespace host
