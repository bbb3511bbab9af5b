// api_elementwise.cpp -- C entry points of the element-wise kernels.
#include "tmx_kernels.hpp"

using namespace tmx;

namespace {
struct GenCache {
  DevBuf blocks;
  int64_t nblocks = 0, max_block = 0;
};
// Heap singletons, never destroyed (see runtime.cpp).
std::mutex& gen_mu() {
  static std::mutex* m = new std::mutex();
  return *m;
}
std::map<uint64_t, std::unique_ptr<GenCache>>& gen_cache() {
  static auto* c = new std::map<uint64_t, std::unique_ptr<GenCache>>();
  return *c;
}

const Tensor& tensor_of(tmx_tensor t) {
  TMX_REQUIRE(t, "null tensor");
  return **reinterpret_cast<std::shared_ptr<Tensor>*>(t);
}

// Per-block global origins/extents of a tensor (uploaded once per tensor).
GenCache& gen_blocks(const Tensor& t) {
  std::lock_guard<std::mutex> g(gen_mu());
  auto& g_gen = gen_cache();
  auto it = g_gen.find(t.uid);
  if (it != g_gen.end()) return *it->second;
  std::vector<GenBlock> v;
  int64_t mx = 0;
  std::vector<int64_t> bid(t.nmodes());
  for (size_t b = 0; b < t.keys.size(); ++b) {
    t.key_to_bid(t.keys[b], bid.data());
    GenBlock gb;
    std::memset(&gb, 0, sizeof(gb));
    gb.off = t.offs[b];
    for (int d = 0; d < t.nmodes(); ++d) {
      gb.lo[d] = (int32_t)t.modes[d]->toff[bid[d]];
      gb.ext[d] = (int32_t)t.modes[d]->tsize(bid[d]);
    }
    mx = std::max(mx, t.sizes[b]);
    v.push_back(gb);
  }
  auto c = std::make_unique<GenCache>();
  c->nblocks = (int64_t)v.size();
  c->max_block = mx;
  c->blocks = upload(v);
  auto& ref = *c;
  g_gen[t.uid] = std::move(c);
  return ref;
}
}  // namespace

extern "C" {

int tmx_fill(double* data, int64_t n, double value, void* stream) {
  return guarded([&] {
    TMX_REQUIRE(data || n == 0, "null data");
    launch_fill(data, n, value, pick_stream(stream));
  });
}

int tmx_memset_zero(void* ptr, size_t bytes, void* stream) {
  return guarded([&] {
    TMX_REQUIRE(ptr || bytes == 0, "null data");
    launch_fill_bytes_zero(ptr, bytes, pick_stream(stream));
  });
}

int tmx_scale(double* data, int64_t n, double alpha, void* stream) {
  return guarded([&] { launch_scale(data, n, alpha, pick_stream(stream)); });
}

int tmx_axpy(int64_t n, double alpha, const double* x, double* y, void* stream) {
  return guarded([&] { launch_axpy(n, alpha, x, y, pick_stream(stream)); });
}

int tmx_fill_normal(tmx_tensor t, double* data, uint64_t seed, uint64_t tensor_id, double scale,
                    int symmetrize_pairs, void* stream) {
  return guarded([&] {
    hipStream_t s = pick_stream(stream);
    const Tensor& tt = tensor_of(t);
    TMX_REQUIRE(data || tt.total == 0, "null data");
    if (symmetrize_pairs) {
      TMX_REQUIRE(tt.nmodes() == 4 && tt.modes[0]->n == tt.modes[1]->n &&
                      tt.modes[2]->n == tt.modes[3]->n,
                  "pair symmetrisation needs a 4-mode tensor with n0 == n1 and n2 == n3");
    }
    GenCache& gc = gen_blocks(tt);
    std::vector<int64_t> mode_n;
    for (int d = 0; d < tt.nmodes(); ++d) mode_n.push_back(tt.modes[d]->n);
    launch_fill_normal(gc.blocks.as<GenBlock>(), gc.nblocks, tt.nmodes(), mode_n.data(), gc.max_block,
                       data, seed, tensor_id, scale, symmetrize_pairs, s);
  });
}

int tmx_norm(const double* data, int64_t n, double* result, void* stream) {
  return guarded([&] {
    TMX_REQUIRE(result, "null result");
    *result = std::sqrt(reduce_sumsq(data, n, pick_stream(stream)));
  });
}

int tmx_norm_noalloc(const double* data, int64_t n, double* result, void* stream) {
  return guarded([&] {
    TMX_REQUIRE(result, "null result");
    TMX_REQUIRE(data || n == 0, "null data");
    hipStream_t s = pick_stream(stream);
    *result = std::sqrt(reduce_sumsq_noalloc(data, n, s, runtime().device));
  });
}

int tmx_dot(const double* x, const double* y, int64_t n, double* result, void* stream) {
  return guarded([&] {
    TMX_REQUIRE(result, "null result");
    *result = reduce_dot(x, y, n, pick_stream(stream));
  });
}

int tmx_rel_diff(const double* x, const double* y, int64_t n, double* result, void* stream) {
  return guarded([&] {
    TMX_REQUIRE(result, "null result");
    hipStream_t s = pick_stream(stream);
    double d = reduce_diffsq(x, y, n, s), r = reduce_sumsq(y, n, s);
    *result = r > 0 ? std::sqrt(d / r) : std::sqrt(d);
  });
}

int tmx_energy_denominator(tmx_tensor R, double* data, const double* e_occ, const double* e_vir,
                           void* stream) {
  return guarded([&] {
    hipStream_t s = pick_stream(stream);
    const Tensor& tt = tensor_of(R);
    TMX_REQUIRE(tt.nmodes() == 4, "denominator update needs a (V,V,O,O) tensor");
    TMX_REQUIRE(e_occ && e_vir, "null orbital energies");
    GenCache& gc = gen_blocks(tt);
    TMX_REQUIRE(gc.max_block < (int64_t(1) << 31), "denominator update: a block of 2^31 or more elements");
    launch_denominator(gc.blocks.as<GenBlock>(), gc.nblocks, tt.total, data, e_occ, e_vir, s);
  });
}

}  // extern "C"
