#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lts {

// Largest image accepted: 32768 texels a side and 2^28 texels (a 16384^2 image), so that the
// decoded bytes and the pyramid stay far below the sizes where allocation sizes could overflow.
constexpr uint32_t kMaxImageSide = 1u << 15;
constexpr uint64_t kMaxImageTexels = 1ull << 28;

// Baseline (sequential, Huffman-coded, 8-bit) JPEG with one or three components -> RGB bytes,
// top row first, the bytes stb_image 2.27 returns for the file (jpeg.cpp). False with `err` on
// failure.
bool jpeg_to_rgb8(const std::vector<uint8_t>& bytes, int& w, int& h, std::vector<uint8_t>& rgb, std::string& err);

// 8-bit PNG, binary PPM or baseline JPEG -> linear RGB floats, bottom row first. False with `err`
// on failure.
bool load_image_linear(const std::string& path, int& w, int& h, std::vector<float>& rgb, std::string& err);
bool write_pfm(const std::string& path, const float* rgb, int w, int h);
bool write_ppm(const std::string& path, const float* rgb, int w, int h);
// Radiance RGBE (.hdr), flat scanlines, top row first: the format of the reference's unused
// outputFile (core/ImageO.cuh:80-92).
bool write_hdr(const std::string& path, const float* rgb, int w, int h);
// 8-bit RGB PNG with the PPM writer's tone map (toRGB, core/ImageO.cuh:52).
bool write_png(const std::string& path, const float* rgb, int w, int h);

}  // namespace lts
