// image_io.hpp -- the image formats the scene API reads and writes, written
// for this library from the published formats:
//   * Radiance RGBE (.hdr): header lines, a "-Y H +X W" resolution line and
//     run-length-encoded or flat scanlines; a texel is (r, g, b) * 2^(e-136),
//     zero when e == 0 -- the decode stb_image applies (stb_image.h:6907-6930).
//   * PFM: "PF\n<w> <h>\n<scale>\n" then float RGB rows bottom to top,
//     little-endian when scale < 0.
//   * PPM (P6): "P6 <w> <h> <maxval>" then 8-bit RGB rows top to bottom.
//   * PNG (written only): 8-bit RGB, one IDAT of stored (uncompressed) deflate blocks, CRC-32
//     and Adler-32 as RFC 1950/1951/2083 define them.
// Environment maps come back in the layout pt_scene_set_env takes: texel
// (x, y) at rgb[(y*w + x)*3], y = 0 being the bottom row of the image (the
// direction -Y), as ImageTexture stores them (texture.cpp:21-45).
#pragma once

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace ptamd {
namespace image_io {

struct Image {
  uint32_t width = 0, height = 0;
  std::vector<float> rgb;  // width*height*3, row y = 0 at the bottom
};

inline std::vector<unsigned char> read_file(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  return std::vector<unsigned char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

// Radiance RGBE.  Rows in the file run top to bottom ("-Y"); the result is
// flipped so that row 0 is the bottom one.
inline Image read_hdr(const std::string &path) {
  const std::vector<unsigned char> b = read_file(path);
  size_t p = 0;
  auto line = [&]() {
    std::string s;
    while (p < b.size() && b[p] != '\n') s.push_back((char)b[p++]);
    if (p < b.size()) ++p;
    return s;
  };
  std::string magic = line();
  if (magic.rfind("#?", 0) != 0) throw std::runtime_error(path + ": not a Radiance HDR file");
  bool rgbe = true;
  for (;;) {
    if (p >= b.size()) throw std::runtime_error(path + ": truncated header");
    std::string s = line();
    if (s.empty()) break;
    if (s.rfind("FORMAT=", 0) == 0 && s != "FORMAT=32-bit_rle_rgbe") rgbe = false;
  }
  if (!rgbe) throw std::runtime_error(path + ": only 32-bit_rle_rgbe is supported");
  std::istringstream res(line());
  std::string ya, xa;
  int h = 0, w = 0;
  res >> ya >> h >> xa >> w;
  if (ya != "-Y" || xa != "+X" || w <= 0 || h <= 0)
    throw std::runtime_error(path + ": unsupported resolution line (want -Y H +X W)");
  Image img;
  img.width = (uint32_t)w;
  img.height = (uint32_t)h;
  img.rgb.assign((size_t)w * h * 3, 0.f);
  std::vector<unsigned char> scan((size_t)w * 4);
  auto need = [&](size_t n) {
    if (p + n > b.size()) throw std::runtime_error(path + ": truncated pixel data");
  };
  for (int row = 0; row < h; ++row) {
    need(4);
    const bool rle = w >= 8 && w < 32768 && b[p] == 2 && b[p + 1] == 2 && ((b[p + 2] << 8) | b[p + 3]) == w &&
                     !(b[p + 2] & 0x80);
    if (rle) {
      p += 4;
      for (int c = 0; c < 4; ++c) {  // channel planes r, g, b, e
        int x = 0;
        while (x < w) {
          need(1);
          int n = b[p++];
          if (n > 128) {  // run
            n -= 128;
            need(1);
            unsigned char v = b[p++];
            if (x + n > w) throw std::runtime_error(path + ": bad run length");
            for (int k = 0; k < n; ++k) scan[4 * (size_t)(x++) + c] = v;
          } else {  // literal
            if (n == 0 || x + n > w) throw std::runtime_error(path + ": bad literal length");
            need((size_t)n);
            for (int k = 0; k < n; ++k) scan[4 * (size_t)(x++) + c] = b[p++];
          }
        }
      }
    } else {
      need((size_t)w * 4);
      std::memcpy(scan.data(), &b[p], (size_t)w * 4);
      p += (size_t)w * 4;
    }
    const int y = h - 1 - row;
    for (int x = 0; x < w; ++x) {
      const unsigned char *q = &scan[4 * (size_t)x];
      float *o = &img.rgb[3 * ((size_t)y * w + x)];
      if (q[3] == 0) {
        o[0] = o[1] = o[2] = 0.f;
      } else {
        const float f = std::ldexp(1.0f, (int)q[3] - 136);
        o[0] = q[0] * f;
        o[1] = q[1] * f;
        o[2] = q[2] * f;
      }
    }
  }
  return img;
}

inline Image read_pfm(const std::string &path) {
  const std::vector<unsigned char> b = read_file(path);
  size_t p = 0;
  auto token = [&]() {
    while (p < b.size() && std::isspace(b[p])) ++p;
    std::string s;
    while (p < b.size() && !std::isspace(b[p])) s.push_back((char)b[p++]);
    return s;
  };
  std::string magic = token();
  if (magic != "PF") throw std::runtime_error(path + ": only colour PFM (PF) is supported");
  Image img;
  img.width = (uint32_t)std::stoul(token());
  img.height = (uint32_t)std::stoul(token());
  const float scale = std::stof(token());
  ++p;  // the single whitespace byte after the scale
  const size_t n = (size_t)img.width * img.height * 3;
  if (p + 4 * n > b.size()) throw std::runtime_error(path + ": truncated pixel data");
  img.rgb.resize(n);
  const bool little = scale < 0.f;
  for (size_t i = 0; i < n; ++i) {
    unsigned char q[4];
    std::memcpy(q, &b[p + 4 * i], 4);
    uint32_t u = little ? (uint32_t)q[0] | (uint32_t)q[1] << 8 | (uint32_t)q[2] << 16 | (uint32_t)q[3] << 24
                        : (uint32_t)q[3] | (uint32_t)q[2] << 8 | (uint32_t)q[1] << 16 | (uint32_t)q[0] << 24;
    float v;
    std::memcpy(&v, &u, 4);
    img.rgb[i] = v;
  }
  return img;
}

// Environment map by extension (.hdr, .pfm).  8-bit formats need a decoder
// this library does not carry: convert them to .hdr or .pfm first.
inline Image read_env(const std::string &path) {
  auto ends = [&](const char *e) {
    size_t n = std::strlen(e);
    if (path.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
      if (std::tolower((unsigned char)path[path.size() - n + i]) != e[i]) return false;
    return true;
  };
  if (ends(".hdr") || ends(".pic")) return read_hdr(path);
  if (ends(".pfm")) return read_pfm(path);
  throw std::runtime_error(path + ": environment maps are read from .hdr or .pfm files");
}

// Binary PPM (P6, maxval <= 255): rows top to bottom in the file, flipped so
// that row 0 is the bottom one; channel values are returned as v / maxval.
inline Image read_ppm(const std::string &path) {
  const std::vector<unsigned char> b = read_file(path);
  size_t p = 0;
  auto token = [&]() {
    for (;;) {
      while (p < b.size() && std::isspace(b[p])) ++p;
      if (p < b.size() && b[p] == '#') {
        while (p < b.size() && b[p] != '\n') ++p;
        continue;
      }
      break;
    }
    std::string s;
    while (p < b.size() && !std::isspace(b[p])) s.push_back((char)b[p++]);
    return s;
  };
  if (token() != "P6") throw std::runtime_error(path + ": only binary PPM (P6) is supported");
  const long w = std::stol(token()), h = std::stol(token()), maxv = std::stol(token());
  if (w <= 0 || h <= 0 || maxv <= 0 || maxv > 255) throw std::runtime_error(path + ": unsupported PPM header");
  ++p;  // the single whitespace byte after maxval
  const size_t n = (size_t)w * h * 3;
  if (p + n > b.size()) throw std::runtime_error(path + ": truncated pixel data");
  Image img;
  img.width = (uint32_t)w;
  img.height = (uint32_t)h;
  img.rgb.resize(n);
  for (long row = 0; row < h; ++row)
    for (size_t i = 0; i < (size_t)w * 3; ++i)
      img.rgb[(size_t)(h - 1 - row) * w * 3 + i] = (float)b[p + (size_t)row * w * 3 + i] / (float)maxv;
  return img;
}

// ImageTexture(file, flip_y, isHDR) (texture.cpp:3-91): texels with row 0 at
// the bottom of the image unless flip_y, which keeps the file's top row
// first.  The reference reads through stb_image, which converts between its
// 8-bit and float forms: an HDR file read with isHDR false becomes
// (uint8)(x^(1/2.2) * 255 + 0.5) clamped to [0, 255], then / 255; an 8-bit
// file read with isHDR true becomes (v / 255)^2.2 (stb_image.h's
// stbi__hdr_to_ldr / stbi__ldr_to_hdr with their default gamma 2.2, scale 1).
inline Image read_texture(const std::string &path, bool flip_y, bool is_hdr) {
  auto ends = [&](const char *e) {
    size_t n = std::strlen(e);
    if (path.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
      if (std::tolower((unsigned char)path[path.size() - n + i]) != e[i]) return false;
    return true;
  };
  Image img;
  if (ends(".hdr") || ends(".pic") || ends(".pfm")) {
    img = ends(".pfm") ? read_pfm(path) : read_hdr(path);
    if (!is_hdr)
      for (float &x : img.rgb) {
        float z = (float)std::pow((double)x, (double)(1.0f / 2.2f)) * 255.f + 0.5f;
        z = z < 0.f ? 0.f : (z > 255.f ? 255.f : z);
        x = (float)(int)z / 255.0f;
      }
  } else if (ends(".ppm")) {
    img = read_ppm(path);
    if (is_hdr)
      for (float &x : img.rgb) x = (float)std::pow((double)x, (double)2.2f);
  } else {
    throw std::runtime_error(path + ": image textures are read from .hdr, .pfm or .ppm files");
  }
  if (flip_y)
    for (uint32_t y = 0; y < img.height / 2; ++y)
      std::swap_ranges(img.rgb.begin() + (size_t)y * img.width * 3, img.rgb.begin() + (size_t)(y + 1) * img.width * 3,
                       img.rgb.begin() + (size_t)(img.height - 1 - y) * img.width * 3);
  return img;
}

inline void write_pfm(const std::string &path, const float *rgb, uint32_t w, uint32_t h) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  f << "PF\n" << w << " " << h << "\n-1.0\n";
  for (size_t i = 0; i < (size_t)w * h * 3; ++i) {
    uint32_t u;
    std::memcpy(&u, &rgb[i], 4);
    unsigned char q[4] = {(unsigned char)u, (unsigned char)(u >> 8), (unsigned char)(u >> 16), (unsigned char)(u >> 24)};
    f.write((const char *)q, 4);
  }
}

inline uint32_t crc32(const unsigned char *d, size_t n, uint32_t c = 0xffffffffu) {
  for (size_t i = 0; i < n; ++i) {
    c ^= d[i];
    for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xedb88320u & (0u - (c & 1u)));
  }
  return c;
}

// FrameBuffer::writeToFile (framebuffer.cpp:27-58): per channel NaN -> 0,
// c^(1/gamma), clamp to [0,1], *255 truncated; rows written top to bottom
// from the film's last row (vertical flip).
inline void write_png(const std::string &path, const float *rgb, uint32_t w, uint32_t h, float gamma) {
  std::vector<unsigned char> raw;
  raw.reserve((size_t)(3 * w + 1) * h);
  for (uint32_t r = 0; r < h; ++r) {
    const uint32_t y = h - 1 - r;
    raw.push_back(0);  // filter: none
    for (uint32_t x = 0; x < 3 * w; ++x) {
      float c = rgb[(size_t)y * 3 * w + x];
      c = (c != c) ? 0.f : std::pow(c, 1.0f / gamma);
      c = (c > 0.f) ? (c > 1.f ? 1.f : c) : 0.f;  // also maps the NaN of a negative's root to 0
      raw.push_back((unsigned char)(c * 255));
    }
  }
  // zlib stream of stored deflate blocks (at most 65535 bytes each)
  std::vector<unsigned char> z = {0x78, 0x01};
  size_t off = 0;
  do {
    size_t n = std::min<size_t>(65535, raw.size() - off);
    z.push_back(off + n == raw.size() ? 1 : 0);
    z.push_back((unsigned char)n);
    z.push_back((unsigned char)(n >> 8));
    z.push_back((unsigned char)~n);
    z.push_back((unsigned char)(~n >> 8));
    z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
    off += n;
  } while (off < raw.size());
  uint32_t a = 1, bsum = 0;
  for (unsigned char c : raw) {
    a = (a + c) % 65521u;
    bsum = (bsum + a) % 65521u;
  }
  const uint32_t adler = (bsum << 16) | a;
  for (int s = 24; s >= 0; s -= 8) z.push_back((unsigned char)(adler >> s));

  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  auto be32 = [](std::vector<unsigned char> &v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back((unsigned char)(x >> s));
  };
  auto chunk = [&](const char *tag, const std::vector<unsigned char> &data) {
    std::vector<unsigned char> c;
    be32(c, (uint32_t)data.size());
    c.insert(c.end(), tag, tag + 4);
    c.insert(c.end(), data.begin(), data.end());
    be32(c, crc32(c.data() + 4, c.size() - 4) ^ 0xffffffffu);
    f.write((const char *)c.data(), (std::streamsize)c.size());
  };
  static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  f.write((const char *)sig, 8);
  std::vector<unsigned char> ihdr;
  be32(ihdr, w);
  be32(ihdr, h);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, deflate, no filter, no interlace
  chunk("IHDR", ihdr);
  chunk("IDAT", z);
  chunk("IEND", {});
}

}  // namespace image_io
}  // namespace ptamd
