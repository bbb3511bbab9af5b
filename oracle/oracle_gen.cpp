// oracle_gen.cpp -- TEST INFRASTRUCTURE ONLY: seeded synthetic inputs for the oracle.
// Compiled with -ffp-contract=off so that values match the device generator bit for bit
// (both evaluate include/tmx_rng.h).
#include <cstdint>
#include <vector>

#include "tamm_oracle.hpp"
#include "tmx_rng.h"

extern "C" {

// Decode block key -> block id (row-major over tile counts).
static void key_to_bid(const orc::Tensor* t, int64_t key, int64_t* bid) {
  for (int i = 0; i < t->nmodes(); ++i) { bid[i] = key / t->key_off[i]; key %= t->key_off[i]; }
}

// Fill every stored element with scale * N(0,1) from the counter-based generator,
// keyed by the element's row-major position in the full dense index space.
// symmetrize_pairs: v(a,b,c,d) = scale * 0.5 * (z(a,b,c,d) + z(b,a,d,c)) for 4-mode tensors.
void orc_fill_normal(void* tp, double* data, uint64_t seed, uint64_t tensor_id, double scale,
                     int symmetrize_pairs) {
  auto* t = (orc::Tensor*)tp;
  int r = t->nmodes();
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t b = 0; b < (int64_t)t->keys.size(); ++b) {
    std::vector<int64_t> bid(r), lo(r), ext(r), it(r, 0), n(r);
    key_to_bid(t, t->keys[b], bid.data());
    for (int i = 0; i < r; ++i) {
      lo[i] = t->modes[i]->toff[bid[i]];
      ext[i] = t->modes[i]->tsize(bid[i]);
      n[i] = t->modes[i]->n;
    }
    double* dst = data + t->offs[b];
    for (int64_t e = 0; e < t->sizes[b]; ++e) {
      int64_t g = 0;
      for (int i = 0; i < r; ++i) g = g * n[i] + lo[i] + it[i];
      double z = tmx_rng_normal(seed, tensor_id, (uint64_t)g);
      if (symmetrize_pairs && r == 4) {
        int64_t g2 = (((lo[1] + it[1]) * n[0] + lo[0] + it[0]) * n[3] + lo[3] + it[3]) * n[2] + lo[2] + it[2];
        double z2 = tmx_rng_normal(seed, tensor_id, (uint64_t)g2);
        z = 0.5 * (z + z2);
      }
      dst[e] = scale * z;
      for (int i = r - 1; i >= 0; --i) {
        if (++it[i] < ext[i]) break;
        it[i] = 0;
      }
    }
  }
}

}  // extern "C"
