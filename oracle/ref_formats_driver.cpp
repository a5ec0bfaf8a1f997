// Driver that runs the reference's own file readers (checker only).
//
// Compiled together with the reference's Renderer/HDRloader.cpp and
// Renderer/npy.h, in place under the reference tree (see Makefile target
// `ref`); it only calls them and dumps what they return, so tests can hold
// this build's readers against the reference's.
//
//   ref_formats hdr  <file.hdr> <out.bin>  -> int32 w, h; float cols[h][w][3]
//                                             (file order, HDRLoader::load);
//                                             float px[h][w][3] via get(u, v)
//                                             (rows flipped, Renderer/HDRloader.h:21-25)
//   ref_formats npyf <file.npy> <out.bin>  -> int32 ndim, fortran; uint64 shape[ndim];
//                                             float data[] (LoadArrayFromNumpy<float>)
//   ref_formats npyi <file.npy> <out.bin>  -> same with int32 data
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "HDRloader.h"
#include "npy.h"

template <class T>
static int dump_npy(const char *in, const char *out) {
  std::vector<unsigned long> shape;
  bool fortran = false;
  std::vector<T> data;
  npy::LoadArrayFromNumpy(std::string(in), shape, fortran, data);
  std::ofstream f(out, std::ios::binary);
  const int32_t nd = (int32_t)shape.size(), fo = fortran ? 1 : 0;
  f.write((const char *)&nd, 4);
  f.write((const char *)&fo, 4);
  for (unsigned long s : shape) {
    const uint64_t v = s;
    f.write((const char *)&v, 8);
  }
  f.write((const char *)data.data(), (std::streamsize)(data.size() * sizeof(T)));
  return f ? 0 : 1;
}

int main(int argc, char **argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: ref_formats hdr|npyf|npyi <in> <out.bin>\n");
    return 2;
  }
  const std::string mode = argv[1];
  if (mode == "hdr") {
    nert_renderer::HDRLoaderResult res;
    res.cols = nullptr;
    if (!nert_renderer::HDRLoader::load(argv[2], res)) {
      std::fprintf(stderr, "HDRLoader::load failed\n");
      return 1;
    }
    std::ofstream f(argv[3], std::ios::binary);
    const int32_t w = res.width, h = res.height;
    f.write((const char *)&w, 4);
    f.write((const char *)&h, 4);
    f.write((const char *)res.cols, (std::streamsize)((size_t)w * h * 3 * sizeof(float)));
    for (int u = 0; u < h; ++u)
      for (int v = 0; v < w; ++v) {
        const gdt::vec3f px = res.get(u, v);
        const float t[3] = {px.x, px.y, px.z};
        f.write((const char *)t, sizeof(t));
      }
    return f ? 0 : 1;
  }
  try {
    if (mode == "npyf") return dump_npy<float>(argv[2], argv[3]);
    if (mode == "npyi") return dump_npy<int>(argv[2], argv[3]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "npy: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
  return 2;
}
