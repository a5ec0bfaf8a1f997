// Reference weight folder -> host scene arrays.
//
// Layout read by RenderEngine::initTCNN (Renderer/tinycudann/tcnn.cu:49-81):
//   config_comp.json  {"vtx_enc": grid, "out_prd": mlp, "cmp_enc": enc,
//                      "cmp_net": mlp}  (prt_network.h:38-61)
//   vtx_enc.npy       hash-grid parameters, flat
//   out_prd.npy       MLP hash-feature (.) wavelet-feature -> n_output_dims (64)
//   cmp_net.npy       MLP [out_prd output | aux encoding] -> 3
//   wve_enc.npy       wavelet feature table [32][6 N^2], flat (shape ignored)
// Every array is read as float and cast to the network precision, FP16
// (tcnn.cu:30-38); this loader rounds to FP16 the same way.
//
// Conventions this loader assumes for the absent tiny-cuda-nn library
// (none can be checked against its source here): each MLP layer is a
// row-major [out][in] matrix; the output layer is padded to 16 rows; the
// hash-grid table is level-major with the F features of an entry adjacent.
//
// The per-pair chain out_prd -> concat(aux) -> cmp_net is folded into this
// build's topology: out_prd's last layer is linear (output_activation None),
// so it is multiplied into the out_prd-facing columns of cmp_net's first
// layer; the aux-facing columns (13 encodings + 3 padding) stay a separate
// [H][16] matrix.
#include <cmath>
#include <fstream>
#include <sstream>

#include "nwprt_internal.hpp"

namespace nwprt {
namespace {

struct MlpCfg {
  int n_neurons, n_hidden_layers;
  std::string activation, output_activation;
};

MlpCfg mlp_cfg(const Json *j, const char *name) {
  require(j && j->type == Json::Object, NWPRT_ERR_FORMAT,
          std::string("config_comp.json: missing object ") + name);
  const std::string otype = j->string_or("otype", "FullyFusedMLP");
  require(otype == "FullyFusedMLP" || otype == "CutlassMLP" || otype == "MLP",
          NWPRT_ERR_UNSUPPORTED, std::string(name) + ": unsupported network " + otype);
  MlpCfg c;
  c.n_neurons = (int)j->number_or("n_neurons", 128);
  c.n_hidden_layers = (int)j->number_or("n_hidden_layers", 2);
  c.activation = j->string_or("activation", "ReLU");
  c.output_activation = j->string_or("output_activation", "None");
  require(c.activation == "ReLU", NWPRT_ERR_UNSUPPORTED,
          std::string(name) + ": hidden activation must be ReLU");
  require(c.n_hidden_layers >= 0 && c.n_neurons > 0, NWPRT_ERR_FORMAT,
          std::string(name) + ": bad layer counts");
  return c;
}

std::vector<float> round_f16(std::vector<float> v) {
  for (float &x : v) x = f16_bits_to_f32(f32_to_f16_bits(x));
  return v;
}

std::string read_text(const std::string &path) {
  std::ifstream f(path);
  if (!f) fail(NWPRT_ERR_IO, "cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

// Layer matrices of a tiny-cuda-nn MLP with `in` inputs and `out` outputs
// (output padded to a multiple of 16), as [out][in] blocks in order.
std::vector<std::vector<float>> split_mlp(const std::vector<float> &p, int in, int out,
                                          const MlpCfg &c, const char *name) {
  const int w = c.n_neurons, h = c.n_hidden_layers;
  const int out_pad = (out + 15) / 16 * 16;
  std::vector<std::pair<int, int>> dims;  // (rows, cols)
  if (h == 0) {
    dims.push_back({out_pad, in});
  } else {
    dims.push_back({w, in});
    for (int i = 1; i < h; ++i) dims.push_back({w, w});
    dims.push_back({out_pad, w});
  }
  size_t need = 0;
  for (auto &d : dims) need += (size_t)d.first * d.second;
  require(p.size() == need, NWPRT_ERR_FORMAT,
          std::string(name) + ".npy: expected " + std::to_string(need) + " parameters, found " +
              std::to_string(p.size()));
  std::vector<std::vector<float>> mats;
  size_t off = 0;
  for (auto &d : dims) {
    mats.emplace_back(p.begin() + off, p.begin() + off + (size_t)d.first * d.second);
    off += (size_t)d.first * d.second;
  }
  return mats;
}

}  // namespace

HostScene load_model_folder(const std::string &dir) {
  const Json cfg = json_parse(read_text(dir + "/config_comp.json"));
  require(cfg.type == Json::Object, NWPRT_ERR_FORMAT, "config_comp.json: not an object");

  HostScene hs;
  nwprt_scene_desc &d = hs.desc;
  std::memset(&d, 0, sizeof(d));

  // vtx_enc: hash grid over the 3 position inputs.
  const Json *vtx = cfg.get("vtx_enc");
  require(vtx && vtx->type == Json::Object, NWPRT_ERR_FORMAT, "config_comp.json: missing vtx_enc");
  const std::string votype = vtx->string_or("otype", "HashGrid");
  require(votype == "HashGrid" || (votype == "Grid" && vtx->string_or("type", "Hash") == "Hash"),
          NWPRT_ERR_UNSUPPORTED, "vtx_enc: only hash grids are supported, got " + votype);
  require(vtx->string_or("interpolation", "Linear") == "Linear", NWPRT_ERR_UNSUPPORTED,
          "vtx_enc: only linear interpolation is supported");
  d.grid.n_levels = (int)vtx->number_or("n_levels", 16);
  d.grid.n_features_per_level = (int)vtx->number_or("n_features_per_level", 2);
  d.grid.log2_hashmap_size = (int)vtx->number_or("log2_hashmap_size", 19);
  d.grid.base_resolution = (int)vtx->number_or("base_resolution", 16);
  d.grid.per_level_scale = (float)vtx->number_or("per_level_scale", 2.0);
  d.hash_dtype = NWPRT_DTYPE_F16;
  d.feature_dim = d.grid.n_levels * d.grid.n_features_per_level;
  d.has_projection = 0;

  // cmp_enc: encoding of the 13 aux inputs (prt_network.h:46, 124-131).
  const Json *cenc = cfg.get("cmp_enc");
  int aux_out = 0;
  if (cenc && cenc->type == Json::Object && !cenc->obj.empty()) {
    const std::string eo = cenc->string_or("otype", "Identity");
    require(eo == "Identity", NWPRT_ERR_UNSUPPORTED,
            "cmp_enc: only the Identity encoding is supported, got " + eo);
    aux_out = 13;
  }
  d.aux_dims = aux_out;
  d.aux_pad_value = 1.0f;

  const MlpCfg op = mlp_cfg(cfg.get("out_prd"), "out_prd");
  const MlpCfg cn = mlp_cfg(cfg.get("cmp_net"), "cmp_net");
  const int op_out = (int)cfg.get("out_prd")->number_or("n_output_dims", 64);
  require(op.output_activation == "None", NWPRT_ERR_UNSUPPORTED,
          "out_prd: output_activation must be None");
  require(cn.output_activation == "None" || cn.output_activation == "Softplus",
          NWPRT_ERR_UNSUPPORTED, "cmp_net: output_activation must be None or Softplus");
  d.output_softplus = cn.output_activation == "Softplus";

  // Parameter arrays.
  nwprt_hashgrid_levels lv;
  hashgrid_levels(d.grid, lv);
  const std::vector<float> vtx_p = npy_read(dir + "/vtx_enc.npy").as_f32();
  const size_t entries = lv.offset[lv.n_levels];
  require(vtx_p.size() == entries * (size_t)d.grid.n_features_per_level, NWPRT_ERR_FORMAT,
          "vtx_enc.npy: expected " + std::to_string(entries * d.grid.n_features_per_level) +
              " parameters, found " + std::to_string(vtx_p.size()));
  hs.hash_table.resize(vtx_p.size() * 2);
  for (size_t i = 0; i < vtx_p.size(); ++i) {
    const uint16_t h = f32_to_f16_bits(vtx_p[i]);
    std::memcpy(&hs.hash_table[2 * i], &h, 2);
  }

  hs.wavelet_features = round_f16(npy_read(dir + "/wve_enc.npy").as_f32());
  const size_t per_feat = hs.wavelet_features.size() / (size_t)d.feature_dim;
  require(per_feat * d.feature_dim == hs.wavelet_features.size() && per_feat % 6 == 0,
          NWPRT_ERR_FORMAT, "wve_enc.npy: size is not feature_dim * 6 * N^2");
  const int n = (int)std::lround(std::sqrt((double)(per_feat / 6)));
  require((size_t)6 * n * n == per_feat, NWPRT_ERR_FORMAT, "wve_enc.npy: 6 N^2 is not square");
  d.face_size = n;

  const int in_pad = (aux_out + op_out + 15) / 16 * 16;  // tcnn alignment 16
  const auto opm = split_mlp(round_f16(npy_read(dir + "/out_prd.npy").as_f32()), d.feature_dim,
                             op_out, op, "out_prd");
  const auto cnm = split_mlp(round_f16(npy_read(dir + "/cmp_net.npy").as_f32()), in_pad, 3, cn,
                             "cmp_net");

  // Fold: chain = out_prd layers [0, h_o) with ReLU, then
  // (cmp_net first layer, out_prd columns) x (out_prd last layer), then the
  // remaining cmp_net layers.
  const int H = cn.n_hidden_layers >= 1 ? cn.n_neurons : 3;
  require(cn.n_hidden_layers >= 1, NWPRT_ERR_UNSUPPORTED, "cmp_net: needs a hidden layer");
  require(op.n_hidden_layers == 0 || op.n_neurons == H, NWPRT_ERR_UNSUPPORTED,
          "out_prd and cmp_net widths must match");
  require(aux_out == 0 || op.n_hidden_layers == 1, NWPRT_ERR_UNSUPPORTED,
          "with an aux encoding, out_prd must have exactly one hidden layer "
          "(the aux inputs then join the second layer)");
  const std::vector<float> &c1 = cnm[0];  // [H][in_pad]
  const std::vector<float> &olast = opm.back();  // [op_out][prev]
  const int prev = op.n_hidden_layers == 0 ? d.feature_dim : op.n_neurons;
  std::vector<float> folded((size_t)H * prev);
  for (int r = 0; r < H; ++r)
    for (int c = 0; c < prev; ++c) {
      double acc = 0;
      for (int m = 0; m < op_out; ++m)
        acc += (double)c1[(size_t)r * in_pad + m] * olast[(size_t)m * prev + c];
      folded[(size_t)r * prev + c] = (float)acc;
    }
  std::vector<std::vector<float>> chain(opm.begin(), opm.end() - 1);
  chain.push_back(folded);
  for (size_t i = 1; i + 1 < cnm.size(); ++i) chain.push_back(cnm[i]);
  std::vector<float> out3(cnm.back().begin(), cnm.back().begin() + 3 * (size_t)H);
  d.hidden_width = H;
  d.n_hidden_layers = (int)chain.size();
  validate_scene_desc(d);  // rejects topologies the kernels do not implement
  for (auto &m : chain) hs.layers.insert(hs.layers.end(), m.begin(), m.end());
  hs.layers.insert(hs.layers.end(), out3.begin(), out3.end());
  if (aux_out) {
    for (int r = 0; r < H; ++r)
      for (int c = 0; c < 16; ++c) {
        const int col = op_out + c;
        hs.layers.push_back(col < in_pad ? c1[(size_t)r * in_pad + col] : 0.0f);
      }
  }
  require(hs.layers.size() == scene_layer_weight_count(d), NWPRT_ERR_FORMAT,
          "model folder: folded layer count mismatch");
  return hs;
}

}  // namespace nwprt
