// nwprt_offline -- offline trajectory relight on the RenderEngine facade.
//
// Mirrors the wavelet half of the reference's offline mode
// (Renderer/offlineRenderer.cpp:12-32, 70-142; flags of Renderer/main.cpp:
// 64-89): read a trajectory file, and for each entry in [start, end) load
// the entry's environment map from the env folder, run the two TCNN passes
// and write predictedIndirect/<i>_<name>.hdr and predictedWithDirect/
// <i>_<name>.hdr, rows flipped as OutputWriter::writeRGBA does
// (Renderer/OutputWriter.cpp:24-40, 146-149).
//
// G-buffers: with --obj (or --mesh; a .ply file is read as PLY), frame i's
// G-buffer comes from the ray caster over the mesh, seen from the trajectory's camera (from = columns 0-2, at =
// columns 3-5, up +y, Renderer/offlineRenderer.cpp:118), positions
// normalised by the mesh's bounds.  The frame's environment
// (<env folder>/cube/<name>.hdr) is both the view and the lighting map, and
// the direct-light buffer gets samplesPerPixel x directSamplesPerLight
// estimates per pixel (trajectory keywords, defaults 1024 x 1 as
// renderConfig's, Renderer/RenderEngine.h:30-31) with random stream
// (--seed + i).  An OBJ's material maps (map_Kd, map_Ks, map_Ns of its
// MTL, relative to the OBJ's folder) become textures of the mesh when the
// files exist and the OBJ has texture coordinates.  Without --obj, frame i's G-buffer is the seeded synthetic
// one with seed (--seed + i) and the camera columns are parsed and ignored.
//
//   nwprt_offline <results_dir> -w <model_dir> -m <env_folder> -t <trajectory>
//                 [-s start] [-e end] [-l] [-k K] [--seed S] [--precision bf16|fp16|fp32]
//                 [--obj scene.obj]
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <vector>

#include "../../include/nwprt_render_engine.hpp"

namespace fs = std::filesystem;

namespace {

struct Trajectory {
  std::string envmapFolder;
  bool finalDenoiser = false;
  int samplesPerPixel = 1024, directSamplesPerLight = 1;
  std::vector<std::string> envmaps;
  std::vector<nwprt_camera> cameras;
};

// Renderer/offlineRenderer.cpp:70-123: keywords before "trajectoryBegin",
// then one "px py pz lx ly lz envname" line per frame.
Trajectory readTrajectory(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open trajectory file " + path);
  Trajectory t;
  bool started = false;
  std::string line;
  const char *keywords[] = {"IndirectSamplesPerDirect",
                            "finalDenoiser", "directIndirectDenoiser", "trainingMode",
                            "waveletSamplesPerPixel", "waveletSamplesPerLight",
                            "waveletDirectDenoiser"};
  while (std::getline(in, line)) {
    line.erase(std::remove_if(line.begin(), line.end(),
                              [](char c) { return c == '\t' || c == '\r' || c == '\n'; }),
               line.end());
    if (line.empty()) continue;
    std::istringstream ss(line);
    std::vector<std::string> tok;
    for (std::string w; std::getline(ss, w, ' ');)
      if (!w.empty()) tok.push_back(w);
    if (tok.empty()) continue;
    if (tok[0] == "trajectoryBegin") {
      started = true;
    } else if (tok[0] == "envmapFolder" && tok.size() > 1) {
      t.envmapFolder = tok[1];
    } else if (tok[0] == "samplesPerPixel" && tok.size() > 1) {
      t.samplesPerPixel = std::stoi(tok[1]);
    } else if (tok[0] == "directSamplesPerLight" && tok.size() > 1) {
      t.directSamplesPerLight = std::stoi(tok[1]);
    } else if (tok[0] == "waveletFinalDenoiser" && tok.size() > 1) {
      t.finalDenoiser = tok[1] == "true";
    } else if (std::find_if(std::begin(keywords), std::end(keywords),
                            [&](const char *k) { return tok[0] == k; }) != std::end(keywords)) {
      continue;
    } else {
      if (!started) throw std::runtime_error("Unknown things before trajectory start " + tok[0]);
      if (tok.size() < 7) throw std::runtime_error("trajectory line needs 7 fields: " + line);
      nwprt_camera cam{};
      for (int i = 0; i < 3; ++i) {
        cam.from[i] = std::stof(tok[(size_t)i]);
        cam.at[i] = std::stof(tok[(size_t)i + 3]);
      }
      cam.up[1] = 1.f;
      t.cameras.push_back(cam);
      t.envmaps.push_back(tok[6]);
    }
  }
  return t;
}

void writeFlipped(const fs::path &file, std::vector<float> rgba, int w, int h) {
  for (int y = 0; y < h / 2; ++y)
    std::swap_ranges(rgba.begin() + (size_t)y * w * 4, rgba.begin() + (size_t)(y + 1) * w * 4,
                     rgba.begin() + (size_t)(h - 1 - y) * w * 4);
  fs::create_directories(file.parent_path());
  if (nwprt_hdr_write(file.string().c_str(), rgba.data(), w, h, 4) != NWPRT_OK)
    throw std::runtime_error(std::string("writing ") + file.string() + ": " + nwprt_last_error());
}

// Texture ids per material for the kd / ks / roughness maps the MTL names,
// as textureManager::addTexture gives them (Renderer/textures.cpp:134-153,
// Model.cpp:133-136): a name that is no file, or a file that does not
// decode, leaves the field constant (-1); one texture per distinct absolute
// file name.
std::vector<int32_t> loadTextures(nwprt_ctx ctx, const char *maps, int64_t bytes, int32_t nm,
                                  std::map<std::string, int32_t> &known) {
  std::vector<std::string> names;
  for (int64_t i = 0; i < bytes;) {
    names.emplace_back(maps + i);
    i += (int64_t)names.back().size() + 1;
  }
  if ((int64_t)names.size() != 3 * (int64_t)nm)
    throw std::runtime_error("material map list does not match the materials");
  const int field[3] = {NWPRT_TEX_KD, NWPRT_TEX_KS, NWPRT_TEX_ROUGHNESS};
  std::vector<int32_t> ids((size_t)nm * NWPRT_TEX_FIELDS, -1);
  for (int32_t m = 0; m < nm; ++m)
    for (int slot = 0; slot < 3; ++slot) {
      const std::string &name = names[(size_t)m * 3 + (size_t)slot];
      std::error_code ec;
      if (name.empty() || !fs::is_regular_file(name, ec)) continue;
      const std::string key = fs::absolute(name, ec).string();
      auto it = known.find(key);
      if (it == known.end()) {
        uint8_t *rgba = nullptr;
        int32_t w = 0, h = 0, id = -1;
        if (nwprt_image_read_rgba8(key.c_str(), &rgba, &w, &h) != NWPRT_OK) {
          std::cerr << "nwprt_offline: texture " << key << " not read: " << nwprt_last_error()
                    << "\n";
          continue;
        }
        const nwprt_status st = nwprt_texture_create(ctx, rgba, w, h, &id);
        nwprt_free_host(rgba);
        if (st != NWPRT_OK)
          throw std::runtime_error("texture " + key + ": " + nwprt_last_error());
        it = known.emplace(key, id).first;
      }
      ids[(size_t)m * NWPRT_TEX_FIELDS + (size_t)field[slot]] = it->second;
    }
  return ids;
}

}  // namespace

int main(int argc, char **argv) {
  std::string results, model, envFolder, trajFile, precision = "bf16", obj;
  int start = 0, end = 2000, k = 64;
  bool large = false;
  uint64_t seed = 0;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error("missing value for " + a);
      return argv[++i];
    };
    try {
      if (a == "-w") model = next();
      else if (a == "-m") envFolder = next();
      else if (a == "-t") trajFile = next();
      else if (a == "-s") start = std::stoi(next());
      else if (a == "-e") end = std::stoi(next());
      else if (a == "-k") k = std::stoi(next());
      else if (a == "-l") large = true;
      else if (a == "--seed") seed = std::stoull(next());
      else if (a == "--precision") precision = next();
      else if (a == "--obj" || a == "--mesh") obj = next();
      else if (!a.empty() && a[0] != '-' && results.empty()) results = a;
      else throw std::runtime_error("unknown argument " + a);
    } catch (const std::exception &e) {
      std::cerr << "nwprt_offline: " << e.what() << "\n";
      return 2;
    }
  }
  if (results.empty() || model.empty() || trajFile.empty()) {
    std::cerr << "usage: nwprt_offline <results_dir> -w <model_dir> -m <env_folder> -t "
                 "<trajectory> [-s start] [-e end] [-l] [-k K] [--seed S] [--precision p] "
                 "[--obj scene.obj]\n";
    return 2;
  }
  try {
    nwprt::renderConfig cfg;
    cfg.precision = precision == "fp32"   ? NWPRT_PREC_FP32
                    : precision == "fp16" ? NWPRT_PREC_FP16
                                          : NWPRT_PREC_BF16;
    const Trajectory traj = readTrajectory(trajFile);
    if (envFolder.empty()) envFolder = traj.envmapFolder;
    cfg.finalDenoiser = cfg.indirectDenoiser = traj.finalDenoiser;
    const int w = large ? 800 : 512, h = large ? 600 : 512;  // Renderer/main.cpp:156
    nwprt::RenderEngine engine(0);
    engine.resize(w, h);
    engine.initTCNN(model);
    int32_t mesh = -1;
    std::map<std::string, int32_t> textures;  // absolute file name -> texture id
    if (!obj.empty()) {
      float *v = nullptr, *n = nullptr;
      int32_t *ix = nullptr, *mid = nullptr, nm = 0;
      nwprt_material *mats = nullptr;
      int64_t nv = 0, nt = 0;
      auto need = [](nwprt_status st, const char *what) {
        if (st != NWPRT_OK) throw std::runtime_error(std::string(what) + ": " + nwprt_last_error());
      };
      const bool ply = obj.size() > 4 && obj.compare(obj.size() - 4, 4, ".ply") == 0;
      float *tc = nullptr;
      char *maps = nullptr;
      int64_t map_bytes = 0;
      if (ply)
        need(nwprt_ply_read(obj.c_str(), &v, &nv, &n, &ix, &nt, &mid, &mats, &nm),
             "reading the mesh");
      else
        need(nwprt_obj_read_textured(obj.c_str(), &v, &nv, &n, &tc, &ix, &nt, &mid, &mats, &nm,
                                     &maps, &map_bytes),
             "reading the mesh");
      nwprt_status st = nwprt_mesh_create(engine.context(), v, nv, n, ix, nt, mid, mats, nm, &mesh);
      if (st == NWPRT_OK && tc) {
        const std::vector<int32_t> matTex =
            loadTextures(engine.context(), maps, map_bytes, nm, textures);
        st = nwprt_mesh_set_textures(engine.context(), mesh, tc, matTex.data());
      }
      for (void *p : {(void *)v, (void *)n, (void *)tc, (void *)ix, (void *)mid, (void *)mats,
                      (void *)maps})
        nwprt_free_host(p);
      need(st, "uploading the mesh");
      engine.setMesh(mesh);
    }
    float *envFaces = nullptr;  // device cube faces of the current frame's map
    int envN = 0;
    end = std::min<int>(end, (int)traj.envmaps.size());
    for (int i = start; i < end; ++i) {
      engine.buildLights({{envFolder, traj.envmaps[(size_t)i]}}, k);
      if (mesh >= 0) {
        const std::string cross =
            (fs::path(envFolder) / "cube" / (traj.envmaps[(size_t)i] + ".hdr")).string();
        float *rgb = nullptr;
        int32_t cw = 0, ch = 0;
        if (nwprt_hdr_read(cross.c_str(), &rgb, &cw, &ch) != NWPRT_OK)
          throw std::runtime_error(std::string("reading ") + cross + ": " + nwprt_last_error());
        const int n = cw / 4;
        std::vector<float> faces((size_t)6 * n * n * 3);
        const nwprt_status cs = nwprt_cubemap_cross_to_faces(rgb, cw, ch, faces.data());
        nwprt_free_host(rgb);
        if (cs != NWPRT_OK)
          throw std::runtime_error(std::string("cube cross ") + cross + ": " + nwprt_last_error());
        if (n != envN) {
          if (envFaces) (void)hipFree(envFaces);
          envFaces = nullptr;
          if (hipMalloc(&envFaces, faces.size() * sizeof(float)) != hipSuccess)
            throw std::runtime_error("device allocation of the environment faces failed");
          envN = n;
        }
        if (hipMemcpy(envFaces, faces.data(), faces.size() * sizeof(float),
                      hipMemcpyHostToDevice) != hipSuccess)
          throw std::runtime_error("uploading the environment faces failed");
        engine.setCamera(traj.cameras[(size_t)i]);
        engine.render(1, nullptr, envFaces, envN, envFaces, envN,
                      traj.samplesPerPixel * traj.directSamplesPerLight,
                      (uint32_t)(seed + (uint64_t)i));
      } else {
        engine.writeSyntheticGBuffer(1, seed + (uint64_t)i);
      }
      engine.TCNNPass_p1(1, nullptr, cfg);
      engine.TCNNPass_p2(0, 1, cfg);
      std::vector<float> ind, wd;
      engine.downloadPredictions(1, ind, wd);
      const std::string name = std::to_string(i) + "_" + traj.envmaps[(size_t)i] + ".hdr";
      writeFlipped(fs::path(results) / "predictedIndirect" / name, ind, w, h);
      writeFlipped(fs::path(results) / "predictedWithDirect" / name, wd, w, h);
      std::cout << "Trajectory: " << i << " Done. Total is: " << end << std::endl;
    }
    engine.destroyLights();
    if (mesh >= 0) nwprt_mesh_destroy(engine.context(), mesh);
    for (const auto &t : textures) nwprt_texture_destroy(engine.context(), t.second);
    if (envFaces) (void)hipFree(envFaces);
  } catch (const std::exception &e) {
    std::cerr << "nwprt_offline: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
