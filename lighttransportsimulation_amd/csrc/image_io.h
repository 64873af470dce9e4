#pragma once
#include <string>
#include <vector>

namespace lts {

// 8-bit PNG or binary PPM -> linear RGB floats, bottom row first. False with `err` on failure.
bool load_image_linear(const std::string& path, int& w, int& h, std::vector<float>& rgb, std::string& err);
bool write_pfm(const std::string& path, const float* rgb, int w, int h);
bool write_ppm(const std::string& path, const float* rgb, int w, int h);

}  // namespace lts
