// structure.cpp -- tiled index spaces, block sparsity and the block layout of tensors.
//
//   tiling:   TiledIndexSpace construction, src/tamm/tiled_index_space.hpp:1135-1194 (single
//             tile size: tiles end at every named/spin/spatial boundary and are shortened there)
//             and :1205-1260 (explicit tile vector; every boundary must be a tile start).
//             A tile's spin/irrep is that of its first index (tiled_index_space.hpp:1327).
//   sparsity: TensorBase::is_non_zero, src/tamm/tensor_base.hpp:243-263 (sum of spins over the
//             upper modes == sum over the lower modes), spin mask as TensorImpl's spin-size
//             constructor (src/tamm/tensor_impl.hpp:155-178).  Extension: with the irrep rule
//             the XOR of the tiles' C2v irreps over the non-ignored modes must be 0.
//   layout:   Distribution_NW (src/tamm/distribution.hpp:197-365): key = sum blockid[i] *
//             prod_{j>i} ntiles_j; nonzero blocks sorted by key and packed contiguously; each
//             block row-major in the tensor's mode order.
#include <atomic>
#include <numeric>

#include "tmx_internal.hpp"

namespace tmx {

static std::atomic<uint64_t> g_uid{1};

std::shared_ptr<Space> make_space(int64_t n, int use_vector, const int64_t* tiles, int64_t ntiles,
                                  int nspin, const int64_t* slo, const int64_t* shi, const int* sval,
                                  int nirrep, const int64_t* ilo, const int64_t* ihi, const int* ival,
                                  int nnamed, const int64_t* nlo, const int64_t* nhi) {
  TMX_REQUIRE(n >= 0, "negative index-space size");
  TMX_REQUIRE(tiles && ntiles >= 1, "tile sizes required");
  std::vector<int64_t> cuts;
  auto add_ranges = [&](int cnt, const int64_t* lo, const int64_t* hi) {
    for (int i = 0; i < cnt; ++i) {
      TMX_REQUIRE(lo[i] >= 0 && hi[i] <= n && lo[i] <= hi[i], "attribute range outside the space");
      cuts.push_back(lo[i]);
      cuts.push_back(hi[i]);
    }
  };
  add_ranges(nnamed, nlo, nhi);
  add_ranges(nspin, slo, shi);
  add_ranges(nirrep, ilo, ihi);
  std::sort(cuts.begin(), cuts.end());
  cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());

  auto sp = std::make_shared<Space>();
  sp->n = n;
  std::vector<int64_t>& t = sp->toff;
  if (n == 0) {
    t.push_back(0);
  } else if (!use_vector) {
    int64_t ts = tiles[0];
    TMX_REQUIRE(ts > 0, "tile size must be positive");
    int64_t pos = 0;
    size_t next_cut = 0;
    while (next_cut < cuts.size() && cuts[next_cut] <= 0) ++next_cut;
    while (pos < n) {
      t.push_back(pos);
      int64_t stop = pos + ts;
      if (next_cut < cuts.size() && stop >= cuts[next_cut]) stop = cuts[next_cut++];
      pos = stop;
    }
    t.push_back(n);
  } else {
    int64_t pos = 0;
    for (int64_t i = 0; i < ntiles; ++i) {
      TMX_REQUIRE(tiles[i] > 0, "tile sizes must be positive");
      t.push_back(pos);
      pos += tiles[i];
    }
    TMX_REQUIRE(pos == n, "tile vector does not sum to the index-space size");
    t.push_back(n);
    for (int64_t c : cuts)
      TMX_REQUIRE(std::binary_search(t.begin(), t.end(), c),
                  "a spin/irrep/subspace boundary is not a tile start");
  }
  auto attr = [](int cnt, const int64_t* lo, const int64_t* hi, const int* val, int64_t idx,
                 int dflt) {
    for (int i = 0; i < cnt; ++i)
      if (idx >= lo[i] && idx < hi[i]) return val[i];
    return dflt;
  };
  for (int64_t k = 0; k + 1 < (int64_t)t.size(); ++k) {
    sp->tspin.push_back(attr(nspin, slo, shi, sval, t[k], 0));
    sp->tirrep.push_back(attr(nirrep, ilo, ihi, ival, t[k], -1));
  }
  return sp;
}

bool Tensor::is_non_zero(const int64_t* bid) const {
  if (has_spin) {
    int up = 0, lo = 0;
    for (int i = 0; i < nmodes(); ++i) {
      int s = modes[i]->tspin[bid[i]];
      if (mask[i] == 0) up += s;
      else if (mask[i] == 1) lo += s;
    }
    if (up != lo) return false;
  }
  if (irrep_rule) {
    int x = 0;
    for (int i = 0; i < nmodes(); ++i)
      if (mask[i] != 2 || !has_spin) x ^= modes[i]->tirrep[bid[i]];
    if (x != 0) return false;
  }
  return true;
}

std::shared_ptr<Tensor> make_tensor(const std::vector<std::shared_ptr<const Space>>& modes,
                                    int nupper, int nlower, int irrep_rule) {
  TMX_REQUIRE((int)modes.size() <= MAXR, "at most 8 modes per tensor");
  auto t = std::make_shared<Tensor>();
  t->modes = modes;
  int r = (int)modes.size();
  if (nupper >= 0) {
    TMX_REQUIRE(nlower >= 0 && nupper + nlower <= r, "spin mask larger than the tensor rank");
    t->has_spin = true;
    for (int i = 0; i < r; ++i) t->mask.push_back(i < nupper ? 0 : (i < nupper + nlower ? 1 : 2));
  } else {
    t->mask.assign(r, 2);
  }
  t->irrep_rule = irrep_rule != 0;
  t->key_off.assign(r, 1);
  for (int i = r - 2; i >= 0; --i) t->key_off[i] = t->key_off[i + 1] * modes[i + 1]->ntiles();
  // Enumerate block ids in row-major order; keys come out ascending.
  std::vector<int64_t> bid(r, 0);
  bool any = true;
  for (int i = 0; i < r; ++i) if (modes[i]->ntiles() == 0) any = false;
  int64_t off = 0;
  while (any) {
    if (t->is_non_zero(bid.data())) {
      int64_t sz = 1;
      for (int i = 0; i < r; ++i) sz *= modes[i]->tsize(bid[i]);
      t->keys.push_back(t->key(bid.data()));
      t->offs.push_back(off);
      t->sizes.push_back(sz);
      off += sz;
    }
    int d = r - 1;
    for (; d >= 0; --d) {
      if (++bid[d] < modes[d]->ntiles()) break;
      bid[d] = 0;
    }
    if (d < 0) break;
  }
  t->total = off;
  t->uid = g_uid++;
  return t;
}

LOperand make_loperand(const tmx_operand* op) {
  TMX_REQUIRE(op && op->tensor, "null operand");
  LOperand o;
  o.t = *reinterpret_cast<const std::shared_ptr<Tensor>*>(op->tensor);
  int r = o.t->nmodes();
  for (int i = 0; i < r; ++i) {
    TMX_REQUIRE(op->labels || r == 0, "labels required");
    o.labels.push_back(op->labels[i]);
    int64_t off = op->tile_offset ? op->tile_offset[i] : 0;
    int64_t nt = op->label_ntiles ? op->label_ntiles[i] : o.t->modes[i]->ntiles() - off;
    TMX_REQUIRE(off >= 0 && nt >= 0 && off + nt <= o.t->modes[i]->ntiles(),
                "label tile range outside the mode's index space");
    o.toff.push_back(off);
    o.ntl.push_back(nt);
  }
  return o;
}

}  // namespace tmx

using namespace tmx;

// Handles are heap-allocated shared_ptr objects so plans can keep structures alive.
struct tmx_space_s {};
struct tmx_tensor_s {};

extern "C" {

int tmx_space_create(int64_t n, int use_tile_vector, const int64_t* tiles, int64_t ntiles,
                     int nspin, const int64_t* spin_lo, const int64_t* spin_hi, const int* spin_val,
                     int nirrep, const int64_t* irrep_lo, const int64_t* irrep_hi,
                     const int* irrep_val, int nnamed, const int64_t* named_lo,
                     const int64_t* named_hi, tmx_space* out) {
  return guarded([&] {
    TMX_REQUIRE(out, "null output");
    auto sp = make_space(n, use_tile_vector, tiles, ntiles, nspin, spin_lo, spin_hi, spin_val, nirrep,
                         irrep_lo, irrep_hi, irrep_val, nnamed, named_lo, named_hi);
    *out = reinterpret_cast<tmx_space>(new std::shared_ptr<const Space>(sp));
  });
}

int tmx_space_destroy(tmx_space s) {
  return guarded([&] { delete reinterpret_cast<std::shared_ptr<const Space>*>(s); });
}

int tmx_space_ntiles(tmx_space s, int64_t* nt) {
  return guarded([&] {
    TMX_REQUIRE(s && nt, "null argument");
    *nt = (*reinterpret_cast<std::shared_ptr<const Space>*>(s))->ntiles();
  });
}

int tmx_space_tiles(tmx_space s, int64_t* offsets, int* spin, int* irrep) {
  return guarded([&] {
    TMX_REQUIRE(s, "null space");
    const Space& sp = **reinterpret_cast<std::shared_ptr<const Space>*>(s);
    for (int64_t t = 0; t <= sp.ntiles(); ++t) if (offsets) offsets[t] = sp.toff[t];
    for (int64_t t = 0; t < sp.ntiles(); ++t) {
      if (spin) spin[t] = sp.tspin[t];
      if (irrep) irrep[t] = sp.tirrep[t];
    }
  });
}

int tmx_tensor_create(int nmodes, const tmx_space* spaces, int nupper, int nlower, int irrep_rule,
                      tmx_tensor* out) {
  return guarded([&] {
    TMX_REQUIRE(out && nmodes >= 0 && (nmodes == 0 || spaces), "bad tensor descriptor");
    std::vector<std::shared_ptr<const Space>> m;
    for (int i = 0; i < nmodes; ++i) {
      TMX_REQUIRE(spaces[i], "null space");
      m.push_back(*reinterpret_cast<std::shared_ptr<const Space>*>(spaces[i]));
    }
    auto t = make_tensor(m, nupper, nlower, irrep_rule);
    *out = reinterpret_cast<tmx_tensor>(new std::shared_ptr<Tensor>(t));
  });
}

int tmx_tensor_destroy(tmx_tensor t) {
  return guarded([&] { delete reinterpret_cast<std::shared_ptr<Tensor>*>(t); });
}

static const Tensor& T(tmx_tensor t) {
  TMX_REQUIRE(t, "null tensor");
  return **reinterpret_cast<std::shared_ptr<Tensor>*>(t);
}

int tmx_tensor_size(tmx_tensor t, int64_t* n) {
  return guarded([&] { *n = T(t).total; });
}
int tmx_tensor_nblocks(tmx_tensor t, int64_t* n) {
  return guarded([&] { *n = (int64_t)T(t).keys.size(); });
}
int tmx_tensor_blocks(tmx_tensor t, int64_t* keys, int64_t* offsets, int64_t* sizes) {
  return guarded([&] {
    const Tensor& tt = T(t);
    for (size_t i = 0; i < tt.keys.size(); ++i) {
      if (keys) keys[i] = tt.keys[i];
      if (offsets) offsets[i] = tt.offs[i];
      if (sizes) sizes[i] = tt.sizes[i];
    }
  });
}
int tmx_tensor_locate(tmx_tensor t, const int64_t* bid, int64_t* offset, int64_t* size) {
  return guarded([&] {
    const Tensor& tt = T(t);
    for (int i = 0; i < tt.nmodes(); ++i)
      TMX_REQUIRE(bid[i] >= 0 && bid[i] < tt.modes[i]->ntiles(), "block id out of range");
    int64_t k = tt.key(bid);
    auto it = std::lower_bound(tt.keys.begin(), tt.keys.end(), k);
    bool found = it != tt.keys.end() && *it == k;
    if (offset) *offset = found ? tt.offs[it - tt.keys.begin()] : -1;
    if (size) {
      int64_t s = 1;
      for (int i = 0; i < tt.nmodes(); ++i) s *= tt.modes[i]->tsize(bid[i]);
      *size = s;
    }
  });
}

}  // extern "C"
