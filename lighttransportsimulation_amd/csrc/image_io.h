#pragma once
#include <string>
#include <vector>

namespace lts {

// 8-bit PNG or binary PPM -> linear RGB floats, bottom row first. False with `err` on failure.
bool load_image_linear(const std::string& path, int& w, int& h, std::vector<float>& rgb, std::string& err);
bool write_pfm(const std::string& path, const float* rgb, int w, int h);
bool write_ppm(const std::string& path, const float* rgb, int w, int h);
// Radiance RGBE (.hdr), flat scanlines, top row first: the format of the reference's unused
// outputFile (core/ImageO.cuh:80-92).
bool write_hdr(const std::string& path, const float* rgb, int w, int h);
// 8-bit RGB PNG with the PPM writer's tone map (toRGB, core/ImageO.cuh:52).
bool write_png(const std::string& path, const float* rgb, int w, int h);

}  // namespace lts
