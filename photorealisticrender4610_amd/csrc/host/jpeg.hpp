// jpeg.hpp — JPEG decoder for glTF textures (jpeg.cpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace prr_host {

// Decodes a sequential or progressive Huffman JPEG (8-bit samples; grey,
// three components, or four: CMYK / YCCK) to RGBA8 (alpha 255). Throws
// std::runtime_error("JPEG: ...") on a malformed or unsupported file.
void decodeJpeg(const uint8_t *data, size_t n, int &width, int &height, std::vector<uint8_t> &rgba);

}  // namespace prr_host
