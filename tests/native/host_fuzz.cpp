// host_fuzz.cpp -- the host code behind the C ABI (scene_xml.cpp, scene_host.cpp, output.cpp)
// driven over mutated inputs, built by tests/test_host_sanitizers.py with AddressSanitizer and
// UndefinedBehaviorSanitizer: any out-of-bounds access, use after free, leak or undefined
// operation ends the run with a report.  Test infrastructure; no device is used.
//
//   host_fuzz <scene dir> <scratch dir> <iterations>
//
// Loads every sample scene and builds its host scene (BVHs, four-wide trees, light table); then,
// from a fixed seed, loads mutated copies of the sample scenes (byte flips, truncations,
// deletions, inserted markup and extreme numbers), mutated OBJ files (through tpt_obj_load,
// count then copy) and random Radiance HDR files (flat and run-length scanlines, random bytes,
// stray numbers) as environment maps, and writes random images (PPM / HDR / EXR) and
// checkpoints, reading the checkpoints back whole, with the wrong size and truncated.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tinypt_gpu.h"
#include "scene_host.h"
#include "scene_xml.h"

namespace tpt {
// The C ABI's error slot lives in api.hip, which this host-only build leaves out.
int set_last_error(int code, const std::string&) { return code; }
}  // namespace tpt

namespace {

std::string read_all(const std::string& p) {
  std::ifstream f(p, std::ios::binary);
  std::stringstream s;
  s << f.rdbuf();
  return s.str();
}

void write_all(const std::string& p, const std::string& d) {
  std::ofstream f(p, std::ios::binary);
  f << d;
}

// Loads a scene file and, when it loads, builds its host scene; true when it loaded.
bool load_and_build(const std::string& path, bool* built) {
  tpt_render_config cfg;
  std::memset(&cfg, 0, sizeof cfg);
  cfg.width = 64;
  cfg.height = 36;
  cfg.spp = 4;
  cfg.max_bounces = 8;
  cfg.rr_start = 3;
  tpt_loaded_scene* L = nullptr;
  if (tpt_scene_load_xml(path.c_str(), &cfg, &L) != 0) return false;
  const tpt_scene_desc* d = nullptr;
  if (tpt_loaded_scene_desc(L, &d) == 0) {
    tpt::HostScene H;
    const bool ok = H.build(d);
    if (built) *built = ok;
  }
  tpt_loaded_scene_destroy(L);
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s scene_dir scratch_dir iterations\n", argv[0]);
    return 2;
  }
  std::setvbuf(stdout, nullptr, _IONBF, 0);
  const std::string dir = argv[1], tmp = argv[2];
  const int iters = std::atoi(argv[3]);
  const char* scenes[] = {"cornell.xml", "default.xml", "mesh.xml", "env_scene.xml"};
  for (const char* s : scenes) {
    bool built = false;
    if (!load_and_build(dir + "/" + s, &built) || !built) {
      std::printf("sample scene %s did not load and build\n", s);
      return 1;
    }
  }
  write_all(tmp + "/cube.obj", read_all(dir + "/cube.obj"));  // meshes the mutated scenes refer to
  std::mt19937_64 rng(7);
  auto pick = [&](uint64_t n) { return (size_t)(rng() % n); };

  int loaded = 0;
  for (int it = 0; it < iters; it++) {
    std::string src = read_all(dir + "/" + scenes[it % 4]);
    for (int m = 0, muts = 1 + (int)pick(6); m < muts && !src.empty(); m++) {
      const size_t pos = pick(src.size());
      switch (pick(5)) {
        case 0: src[pos] = (char)pick(256); break;
        case 1: src.resize(pos); break;
        case 2: src.insert(pos, std::string(1 + pick(4), "<>\"=/&;0-9e"[pick(11)])); break;
        case 3: src.erase(pos, 1 + pick(16)); break;
        default: {
          const char* nums[] = {"1e308", "-1e308", "nan", "inf", "99999999999999999999", "-1", "0", "4294967296"};
          src.insert(pos, nums[pick(8)]);
        }
      }
    }
    write_all(tmp + "/fuzz.xml", src);
    bool built = false;
    loaded += load_and_build(tmp + "/fuzz.xml", &built) ? 1 : 0;
  }
  std::printf("xml: %d of %d mutated scenes loaded\n", loaded, iters);

  const std::string obj = read_all(dir + "/cube.obj");
  int64_t tris = 0;
  for (int it = 0; it < iters; it++) {
    std::string src = obj;
    for (int m = 0, muts = 1 + (int)pick(6); m < muts && !src.empty(); m++) {
      const size_t pos = pick(src.size());
      switch (pick(4)) {
        case 0: src[pos] = (char)pick(256); break;
        case 1: src.resize(pos); break;
        case 2: {
          const char* lines[] = {"f 1 2 999999\n", "f -9999 1 2\n", "v nan 1 2\n", "f 1/2/3 4//5 6\n",
                                 "f 1\n",          "v 1e400 0 0\n", "f 0 0 0\n"};
          src.insert(pos, lines[pick(7)]);
          break;
        }
        default: src.erase(pos, 1 + pick(16));
      }
    }
    write_all(tmp + "/fuzz.obj", src);
    int64_t n = 0;
    if (tpt_obj_load((tmp + "/fuzz.obj").c_str(), &n, nullptr) == 0) {
      std::vector<double> v((size_t)std::max<int64_t>(n, 1) * 9);
      int64_t cap = std::max<int64_t>(n, 1);
      if (tpt_obj_load((tmp + "/fuzz.obj").c_str(), &cap, v.data()) == 0) tris += cap;
    }
  }
  std::printf("obj: %lld triangles read\n", (long long)tris);

  write_all(tmp + "/env.xml",
            "<scene><emitter type=\"envmap\"><string name=\"filename\" value=\"fuzz.hdr\"/></emitter></scene>");
  for (int it = 0; it < iters; it++) {
    const int w = 1 + (int)pick(40), h = 1 + (int)pick(20);
    std::string hdr = "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " + std::to_string(h) + " +X " + std::to_string(w) + "\n";
    for (size_t i = 0, body = pick((uint64_t)(w * h * 6 + 8)); i < body; i++) {
      if (pick(7) == 0 && w >= 8) {  // a run-length scanline header
        hdr += (char)2;
        hdr += (char)2;
        hdr += (char)(w >> 8);
        hdr += (char)(w & 255);
      } else {
        hdr += (char)pick(256);
      }
    }
    if (pick(5) == 0) hdr.insert(pick(hdr.size()), std::to_string(rng() % 100000000000ull));
    write_all(tmp + "/fuzz.hdr", hdr);
    load_and_build(tmp + "/env.xml", nullptr);
  }
  std::printf("hdr: done\n");

  for (int it = 0; it < iters / 10 + 1; it++) {
    const int w = 1 + (int)pick(50), h = 1 + (int)pick(30);
    std::vector<double> acc((size_t)w * h * 3), acc_w((size_t)w * h);
    std::uniform_real_distribution<double> u(-1.0, 100.0);
    for (double& x : acc) x = u(rng);
    for (double& x : acc_w) x = (double)pick(5);
    const char* ext[] = {".ppm", ".hdr", ".exr"};
    tpt_write_image((tmp + "/img" + ext[it % 3]).c_str(), w, h, acc.data(), acc_w.data(), 0.5, it % 2);
    const std::string ck = tmp + "/ck.bin";
    tpt_checkpoint_save(ck.c_str(), 1234, w, h, 7, acc.data(), acc_w.data());
    uint64_t next = 0;
    int32_t found = 0;
    const int w2 = (it % 4 == 0) ? w + 1 : w;
    std::vector<double> acc2((size_t)w2 * h * 3), acc_w2((size_t)w2 * h);
    tpt_checkpoint_load(ck.c_str(), 1234, w2, h, &next, acc2.data(), acc_w2.data(), &found);
    std::string bytes = read_all(ck);
    bytes.resize(pick(bytes.size() + 1));
    write_all(ck, bytes);
    tpt_checkpoint_load(ck.c_str(), 1234, w, h, &next, acc.data(), acc_w.data(), &found);
  }
  std::printf("writers: done\n");
  return 0;
}
