// output.cpp -- the steps after the path, on the FP64 accumulation buffers (host).
//
//   * tpt_resolve_pixels: resolve_pixels (output.rs:96-102), acc / max(acc_w, 1).
//   * tpt_write_image: OutputFormat::from_path + write (output.rs:45-94): PPM (exposure,
//     tonemap, sRGB encode, 8 bit, ASCII P3), Radiance HDR (RGBE with the run-length scheme of
//     hdr.rs:149-195) and OpenEXR.  The reference writes EXR through a third-party encoder
//     (exr.rs:44-75: linear ACEScg float32 R, G, B, ACEScg chromaticities, a lossless codec);
//     this writer produces the same channels, values and chromaticities in an uncompressed
//     single-part scanline file of its own.
//   * tpt_checkpoint_save / load: the HYPCKPT v1 file of checkpoint.rs:48-138, byte for byte.
//
// Compiled with -ffp-contract=off: the colour arithmetic restates the reference's f64 steps.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "../../include/tinypt_gpu.h"
#include "scene_xml.h"

namespace tpt {
namespace {

double clamp(double x, double a, double b) { return std::fmin(std::fmax(x, a), b); }  // math.rs:173

double linear_to_srgb(double c) {  // math.rs:185-191
  return c <= 0.0031308 ? 12.92 * c : 1.055 * std::pow(c, 1.0 / 2.4) - 0.055;
}
uint8_t to_u8(double x) {  // output.rs:130-134
  double v = linear_to_srgb(clamp(x, 0.0, 1.0));
  return (uint8_t)(v * 255.0 + 0.5);
}
double aces_fitted(double x) {  // output.rs:137-148 (Narkowicz)
  x = std::fmax(x, 0.0);
  const double a = 2.51, b = 0.03, c = 2.43, d = 0.59, e = 0.14;
  return clamp((x * (a * x + b)) / (x * (c * x + d) + e), 0.0, 1.0);
}

// ---- colour matrices (aces.rs:13-90)
struct M3 {
  double m[3][3];
};
M3 mul(const M3& a, const M3& b) {
  M3 r{};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
  return r;
}
void mul_vec(const M3& a, const double* v, double* r) {
  for (int i = 0; i < 3; i++) r[i] = a.m[i][0] * v[0] + a.m[i][1] * v[1] + a.m[i][2] * v[2];
}
M3 invert(const M3& a) {  // Mat3::invert (math.rs:151-169), cofactors
  const auto& m = a.m;
  double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  double i = 1.0 / det;
  M3 r;
  r.m[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * i;
  r.m[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * i;
  r.m[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * i;
  r.m[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * i;
  r.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * i;
  r.m[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * i;
  r.m[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * i;
  r.m[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * i;
  r.m[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * i;
  return r;
}
M3 diag(const double* v) { return M3{{{v[0], 0, 0}, {0, v[1], 0}, {0, 0, v[2]}}}; }
void xy_to_xyz(double x, double y, double* out) {
  out[0] = x / y;
  out[1] = 1.0;
  out[2] = (1.0 - x - y) / y;
}
M3 rgb_to_xyz(const double pr[3][2], double wx, double wy) {
  double r[3], g[3], b[3], w[3], s[3];
  xy_to_xyz(pr[0][0], pr[0][1], r);
  xy_to_xyz(pr[1][0], pr[1][1], g);
  xy_to_xyz(pr[2][0], pr[2][1], b);
  M3 m{{{r[0], g[0], b[0]}, {r[1], g[1], b[1]}, {r[2], g[2], b[2]}}};
  xy_to_xyz(wx, wy, w);
  mul_vec(invert(m), w, s);
  return mul(m, diag(s));
}
M3 bradford(const double* src, const double* dst) {
  const M3 m{{{0.8951, 0.2664, -0.1614}, {-0.7502, 1.7135, 0.0367}, {0.0389, -0.0685, 1.0296}}};
  const M3 mi{{{0.9869929, -0.1470543, 0.1599627}, {0.4323053, 0.5183603, 0.0492912}, {-0.0085287, 0.0400428, 0.9684867}}};
  double s[3], d[3], k[3];
  mul_vec(m, src, s);
  mul_vec(m, dst, d);
  for (int i = 0; i < 3; i++) k[i] = d[i] / s[i];
  return mul(mul(mi, diag(k)), m);
}
M3 srgb_to_acescg() {
  static const M3 M = [] {
    const double rec709[3][2] = {{0.64, 0.33}, {0.30, 0.60}, {0.15, 0.06}};
    const double aces[3][2] = {{0.713, 0.293}, {0.165, 0.830}, {0.128, 0.044}};
    double d65[3], d60[3];
    xy_to_xyz(0.3127, 0.3290, d65);
    xy_to_xyz(0.32168, 0.33767, d60);
    M3 src = rgb_to_xyz(rec709, 0.3127, 0.3290), dst = rgb_to_xyz(aces, 0.32168, 0.33767);
    return mul(mul(invert(dst), bradford(d65, d60)), src);
  }();
  return M;
}

// ---- RGBE (hdr.rs:65-88, 149-195, 211-226)
void color_to_rgbe(const double* c, uint8_t* o) {
  double r = std::fmax(c[0], 0.0), g = std::fmax(c[1], 0.0), b = std::fmax(c[2], 0.0);
  double mx = std::fmax(std::fmax(r, g), b);
  if (mx < 1e-32) {
    o[0] = o[1] = o[2] = o[3] = 0;
    return;
  }
  int e = (int)std::floor(std::log2(mx)) + 1;
  double f = std::ldexp(1.0, e - 8);
  o[0] = (uint8_t)clamp(r / f, 0.0, 255.0);
  o[1] = (uint8_t)clamp(g / f, 0.0, 255.0);
  o[2] = (uint8_t)clamp(b / f, 0.0, 255.0);
  o[3] = (uint8_t)(e + 128);
}
void rgbe_scanline(std::string* out, const std::vector<uint8_t>& s, size_t w) {
  if (w < 8 || w > 0x7fff) {
    out->append((const char*)s.data(), s.size());
    return;
  }
  out->push_back(2);
  out->push_back(2);
  out->push_back((char)((w >> 8) & 0xff));
  out->push_back((char)(w & 0xff));
  for (size_t ch = 0; ch < 4; ch++) {
    size_t x = 0;
    while (x < w) {
      size_t run = 1;
      uint8_t v = s[x * 4 + ch];
      while (x + run < w && run < 127 && s[(x + run) * 4 + ch] == v) run++;
      if (run >= 4) {
        out->push_back((char)(128 + run));
        out->push_back((char)v);
        x += run;
      } else {
        size_t start = x, count = 0;
        while (x < w && count < 128) {
          uint8_t u = s[x * 4 + ch];
          size_t look = 1;
          while (x + look < w && look < 127 && s[(x + look) * 4 + ch] == u) look++;
          if (look >= 4) break;
          x++;
          count++;
        }
        out->push_back((char)count);
        for (size_t i = 0; i < count; i++) out->push_back((char)s[(start + i) * 4 + ch]);
      }
    }
  }
}

// ---- OpenEXR, uncompressed scanline, FLOAT channels B, G, R (alphabetical)
void put_u32(std::string* o, uint32_t v) { o->append((const char*)&v, 4); }
void put_i32(std::string* o, int32_t v) { o->append((const char*)&v, 4); }
void put_f32(std::string* o, float v) { o->append((const char*)&v, 4); }
void put_attr(std::string* o, const char* name, const char* type, const std::string& value) {
  o->append(name);
  o->push_back('\0');
  o->append(type);
  o->push_back('\0');
  put_i32(o, (int32_t)value.size());
  o->append(value);
}
std::string exr_bytes(int w, int h, const std::vector<float>& rgb) {
  static_assert(sizeof(float) == 4, "IEEE float");
  std::string o;
  put_u32(&o, 20000630u);  // magic
  put_u32(&o, 2u);         // version 2, single-part scanline
  std::string ch;
  for (const char* c : {"B", "G", "R"}) {
    ch.append(c);
    ch.push_back('\0');
    put_i32(&ch, 2);  // FLOAT
    ch.append(4, '\0');  // pLinear, reserved
    put_i32(&ch, 1);     // xSampling
    put_i32(&ch, 1);     // ySampling
  }
  ch.push_back('\0');
  put_attr(&o, "channels", "chlist", ch);
  std::string chroma;
  for (float v : {0.713f, 0.293f, 0.165f, 0.830f, 0.128f, 0.044f, 0.32168f, 0.33767f}) put_f32(&chroma, v);
  put_attr(&o, "chromaticities", "chromaticities", chroma);
  put_attr(&o, "compression", "compression", std::string(1, '\0'));
  std::string box;
  for (int v : {0, 0, w - 1, h - 1}) put_i32(&box, v);
  put_attr(&o, "dataWindow", "box2i", box);
  put_attr(&o, "displayWindow", "box2i", box);
  put_attr(&o, "lineOrder", "lineOrder", std::string(1, '\0'));
  put_attr(&o, "name", "string", "ACEScg");  // the reference's layer name
  std::string par, swc, sww;
  put_f32(&par, 1.0f);
  put_f32(&swc, 0.0f);
  put_f32(&swc, 0.0f);
  put_f32(&sww, 1.0f);
  put_attr(&o, "pixelAspectRatio", "float", par);
  put_attr(&o, "screenWindowCenter", "v2f", swc);
  put_attr(&o, "screenWindowWidth", "float", sww);
  o.push_back('\0');  // end of header
  const uint64_t table = o.size();
  const uint64_t line_bytes = (uint64_t)w * 3 * 4;
  const uint64_t first = table + 8 * (uint64_t)h;
  for (int y = 0; y < h; y++) {
    uint64_t off = first + (uint64_t)y * (8 + line_bytes);
    o.append((const char*)&off, 8);
  }
  for (int y = 0; y < h; y++) {
    put_i32(&o, y);
    put_i32(&o, (int32_t)line_bytes);
    for (int c : {2, 1, 0})  // B, G, R planes
      for (int x = 0; x < w; x++) put_f32(&o, rgb[3 * ((size_t)y * w + x) + c]);
  }
  return o;
}

bool write_file(const std::string& path, const std::string& bytes, std::string* err) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) {
    *err = "cannot create " + path;
    return false;
  }
  f.write(bytes.data(), (std::streamsize)bytes.size());
  if (!f) {
    *err = "write failed: " + path;
    return false;
  }
  return true;
}

}  // namespace
}  // namespace tpt

extern "C" {

int tpt_resolve_pixels(int64_t n, const double* acc_rgb, const double* acc_w, double* rgb_out) {
  if (n < 0 || (n > 0 && (!acc_rgb || !acc_w || !rgb_out))) return tpt::set_last_error(TPT_ERR_INVALID_ARGUMENT, "bad argument");
  for (int64_t i = 0; i < n; i++) {
    double d = std::fmax(acc_w[i], 1.0);
    for (int c = 0; c < 3; c++) rgb_out[3 * i + c] = acc_rgb[3 * i + c] / d;
  }
  return TPT_OK;
}

int tpt_write_image(const char* path, int32_t w, int32_t h, const double* acc_rgb, const double* acc_w, double exposure,
                    int32_t tonemap) {
  using namespace tpt;
  if (!path || w <= 0 || h <= 0 || !acc_rgb || !acc_w) return set_last_error(TPT_ERR_INVALID_ARGUMENT, "bad argument");
  if (tonemap != TPT_TONEMAP_NONE && tonemap != TPT_TONEMAP_ACES) return set_last_error(TPT_ERR_INVALID_ARGUMENT, "unknown tonemap");
  try {
    const size_t n = (size_t)w * (size_t)h;
    std::vector<double> px(3 * n);
    tpt_resolve_pixels((int64_t)n, acc_rgb, acc_w, px.data());
    std::string low = path;
    for (char& c : low) c = (char)std::tolower((unsigned char)c);
    auto ends = [&](const char* e) {
      size_t k = std::strlen(e);
      return low.size() >= k && low.compare(low.size() - k, k, e) == 0;
    };
    std::string bytes, err;
    if (ends(".exr")) {  // scene-referred linear, converted to ACEScg
      M3 m = srgb_to_acescg();
      std::vector<float> out(3 * n);
      for (size_t i = 0; i < n; i++) {
        double v[3];
        mul_vec(m, &px[3 * i], v);
        for (int c = 0; c < 3; c++) out[3 * i + c] = (float)v[c];
      }
      bytes = exr_bytes(w, h, out);
    } else if (ends(".hdr")) {  // scene-referred linear, RGBE
      bytes = "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " + std::to_string(h) + " +X " + std::to_string(w) + "\n";
      std::vector<uint8_t> line(4 * (size_t)w);
      for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) color_to_rgbe(&px[3 * ((size_t)y * w + x)], &line[4 * (size_t)x]);
        rgbe_scanline(&bytes, line, (size_t)w);
      }
    } else {  // PPM: exposure, tonemap, clamp, sRGB, 8 bit
      bytes = "P3\n" + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
      const double scale = std::pow(2.0, exposure);
      for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
          const double* p = &px[3 * ((size_t)y * w + x)];
          double c[3];
          for (int k = 0; k < 3; k++) {
            double v = p[k] * scale;
            if (tonemap == TPT_TONEMAP_ACES) v = aces_fitted(v);
            c[k] = std::fmin(std::fmax(v, 0.0), 1.0);  // Color::clamp01
          }
          bytes += std::to_string(to_u8(c[0])) + " " + std::to_string(to_u8(c[1])) + " " + std::to_string(to_u8(c[2])) + " ";
        }
        bytes += "\n";
      }
    }
    if (!write_file(path, bytes, &err)) return set_last_error(TPT_ERR_INVALID_ARGUMENT, err);
    return TPT_OK;
  } catch (const std::bad_alloc&) {
    return set_last_error(TPT_ERR_OUT_OF_MEMORY, "host allocation failed while writing the image");
  }
}

int tpt_checkpoint_save(const char* path, uint64_t scene_hash, int32_t w, int32_t h, uint64_t next_id,
                        const double* acc_rgb, const double* acc_w) {
  using namespace tpt;
  if (!path || w <= 0 || h <= 0 || !acc_rgb || !acc_w) return set_last_error(TPT_ERR_INVALID_ARGUMENT, "bad argument");
  const size_t n = (size_t)w * (size_t)h;
  std::string b("HYPCKPT\0", 8);
  auto u32 = [&](uint32_t v) { b.append((const char*)&v, 4); };
  auto u64 = [&](uint64_t v) { b.append((const char*)&v, 8); };
  u32(1u);
  u64(scene_hash);
  u32((uint32_t)w);
  u32((uint32_t)h);
  u64(next_id);
  b.append((const char*)acc_rgb, sizeof(double) * 3 * n);
  b.append((const char*)acc_w, sizeof(double) * n);
  std::string tmp = std::string(path) + ".tmp", err;
  if (!write_file(tmp, b, &err)) return set_last_error(TPT_ERR_INVALID_ARGUMENT, err);
  if (std::rename(tmp.c_str(), path) != 0) return set_last_error(TPT_ERR_INVALID_ARGUMENT, "cannot rename checkpoint");
  return TPT_OK;
}

int tpt_checkpoint_load(const char* path, uint64_t scene_hash, int32_t w, int32_t h, uint64_t* next_id, double* acc_rgb,
                        double* acc_w, int32_t* found) {
  using namespace tpt;
  if (!path || w <= 0 || h <= 0 || !next_id || !acc_rgb || !acc_w || !found)
    return set_last_error(TPT_ERR_INVALID_ARGUMENT, "bad argument");
  *found = 0;
  struct stat sb;
  if (::stat(path, &sb) != 0) return TPT_OK;  // Path::exists is false: no checkpoint
  std::ifstream f(path, std::ios::binary);
  if (!f) return set_last_error(TPT_ERR_INVALID_ARGUMENT, std::string("cannot open checkpoint ") + path);
  const size_t n = (size_t)w * (size_t)h;
  char magic[8];
  uint32_t ver = 0, fw = 0, fh = 0;
  uint64_t hash = 0, nid = 0;
  // A short read is an error (read_exact), a field that does not match is "no checkpoint".
  bool ok = true;
  auto rd = [&](void* p, size_t k) { return ok = ok && (bool)f.read((char*)p, (std::streamsize)k); };
  const int truncated = TPT_ERR_INVALID_ARGUMENT;
  if (!rd(magic, 8)) return set_last_error(truncated, "truncated checkpoint");
  if (std::memcmp(magic, "HYPCKPT\0", 8) != 0) return TPT_OK;
  if (!rd(&ver, 4)) return set_last_error(truncated, "truncated checkpoint");
  if (ver != 1u) return TPT_OK;
  if (!rd(&hash, 8)) return set_last_error(truncated, "truncated checkpoint");
  if (hash != scene_hash) return TPT_OK;
  if (!rd(&fw, 4) || !rd(&fh, 4)) return set_last_error(truncated, "truncated checkpoint");
  if (fw != (uint32_t)w || fh != (uint32_t)h) return TPT_OK;
  if (!rd(&nid, 8)) return set_last_error(truncated, "truncated checkpoint");
  std::vector<double> a(3 * n), wt(n);
  if (!rd(a.data(), sizeof(double) * 3 * n) || !rd(wt.data(), sizeof(double) * n))
    return set_last_error(truncated, "truncated checkpoint");
  std::memcpy(acc_rgb, a.data(), sizeof(double) * 3 * n);
  std::memcpy(acc_w, wt.data(), sizeof(double) * n);
  *next_id = nid;
  *found = 1;
  return TPT_OK;
}

}  // extern "C"
