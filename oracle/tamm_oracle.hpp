// tamm_oracle.hpp -- TEST INFRASTRUCTURE ONLY: data structures of the CPU restatement
// of TAMM's contraction path (see tamm_oracle.cpp for the file:line map).
#pragma once
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

extern thread_local std::string g_err;

struct Space {
  int64_t n = 0;
  std::vector<int64_t> toff;  // ntiles + 1 tile boundaries
  std::vector<int> tspin;     // spin value of each tile (0 when the space has none)
  std::vector<int> tirrep;    // irrep of each tile (-1 when the space has none)
  bool has_spin = false, has_irrep = false;
  int64_t ntiles() const { return (int64_t)toff.size() - 1; }
  int64_t tsize(int64_t t) const { return toff[t + 1] - toff[t]; }
};

struct AttrRange {
  int64_t lo, hi;
  int val;
};

static int attr_of(const std::vector<AttrRange>& rs, int64_t idx, int dflt) {
  for (const auto& r : rs)
    if (idx >= r.lo && idx < r.hi) return r.val;
  return dflt;
}

// tiled_index_space.hpp:1135-1194 (single tile size) and :1205-1260 (tile vector)
static Space make_space(int64_t n, int use_vector, const int64_t* tiles, int64_t ntiles_in,
                        const std::vector<AttrRange>& spin, const std::vector<AttrRange>& irrep,
                        const std::vector<std::pair<int64_t, int64_t>>& named) {
  Space s;
  s.n = n;
  std::vector<int64_t> bnd;
  for (auto& r : named) { bnd.push_back(r.first); bnd.push_back(r.second); }
  for (auto& r : spin) { bnd.push_back(r.lo); bnd.push_back(r.hi); }
  for (auto& r : irrep) { bnd.push_back(r.lo); bnd.push_back(r.hi); }
  std::sort(bnd.begin(), bnd.end());
  bnd.erase(std::unique(bnd.begin(), bnd.end()), bnd.end());
  std::vector<int64_t> ret;
  if (n == 0) {
    ret.push_back(0);
  } else if (!use_vector) {
    int64_t ts = tiles[0];
    if (ts <= 0) throw std::runtime_error("tile size must be positive");
    if (bnd.empty()) {
      for (int64_t i = 0; i < n; i += ts) ret.push_back(i);
      ret.push_back(n);
    } else {
      int64_t i = 0;
      size_t j = (i == bnd[0]) ? 1 : 0;
      while (i < n) {
        ret.push_back(i);
        if (j < bnd.size() && i + ts >= bnd[j]) i = bnd[j++];
        else i = i + ts;
      }
      ret.push_back(n);
    }
  } else {
    int64_t tot = 0;
    for (int64_t t = 0; t < ntiles_in; ++t) tot += tiles[t];
    if (tot != n) throw std::runtime_error("tile vector does not sum to the space size");
    int64_t i = 0;
    for (int64_t t = 0; i < n; i += tiles[t++]) ret.push_back(i);
    ret.push_back(n);
    for (auto b : bnd)
      if (!std::binary_search(ret.begin(), ret.end(), b))
        throw std::runtime_error("attribute/subspace boundary is not a tile start");
  }
  s.toff = ret;
  s.has_spin = !spin.empty();
  s.has_irrep = !irrep.empty();
  for (int64_t t = 0; t + 1 < (int64_t)ret.size(); ++t) {
    s.tspin.push_back(attr_of(spin, ret[t], 0));
    s.tirrep.push_back(attr_of(irrep, ret[t], -1));
  }
  return s;
}

struct Tensor {
  std::vector<const Space*> modes;
  bool has_spin = false;
  std::vector<int> mask;  // 0 upper, 1 lower, 2 ignore
  bool irrep_rule = false;
  std::vector<int64_t> key_off;
  std::vector<int64_t> keys, offs, sizes;  // nonzero blocks, ascending key
  int64_t total = 0;

  int nmodes() const { return (int)modes.size(); }

  // tensor_base.hpp:243-263 plus the irrep extension
  bool is_non_zero(const int64_t* bid) const {
    if (has_spin) {
      int up = 0, lo = 0;
      for (int i = 0; i < nmodes(); ++i) {
        int sp = modes[i]->tspin[bid[i]];
        if (mask[i] == 0) up += sp;
        else if (mask[i] == 1) lo += sp;
      }
      if (up != lo) return false;
    }
    if (irrep_rule) {
      int x = 0;
      for (int i = 0; i < nmodes(); ++i)
        if (!has_spin || mask[i] != 2) x ^= modes[i]->tirrep[bid[i]];
      if (x != 0) return false;
    }
    return true;
  }
  int64_t key(const int64_t* bid) const {
    int64_t k = 0;
    for (int i = 0; i < nmodes(); ++i) k += bid[i] * key_off[i];
    return k;
  }
  int64_t block_size(const int64_t* bid) const {
    int64_t s = 1;
    for (int i = 0; i < nmodes(); ++i) s *= modes[i]->tsize(bid[i]);
    return s;
  }
  // distribution.hpp:270-285 (single process: offset inside the one buffer)
  int64_t locate(const int64_t* bid) const {
    int64_t k = key(bid);
    auto it = std::lower_bound(keys.begin(), keys.end(), k);
    if (it == keys.end() || *it != k) throw std::runtime_error("locate: block is not stored");
    return offs[it - keys.begin()];
  }
};

}  // namespace orc
