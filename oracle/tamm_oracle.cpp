// tamm_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of TAMM's CPU contraction path, used as the checker for the
// MI355X engine.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may load this library; the product library never links it.
//
// What it restates (file:line in the TAMM tree):
//   * tile construction ........... src/tamm/tiled_index_space.hpp:1135-1194 (one tile
//                                   size, split at named/spin/spatial boundaries) and
//                                   :1205-1260 (explicit tile vector, boundaries must be
//                                   tile starts); spin of a tile = spin of its first
//                                   index (tiled_index_space.hpp:1327)
//   * spin block sparsity ......... src/tamm/tensor_base.hpp:243-263 (sum of upper spins
//                                   == sum of lower spins), mask built as in
//                                   src/tamm/tensor_impl.hpp:155-178
//   * irrep block sparsity ........ extension (absent from TAMM): XOR of the C2v irrep
//                                   labels of the non-ignored modes must be 0
//   * block layout ................ src/tamm/distribution.hpp:197-365 (key = sum blockid[i]
//                                   * prod_{j>i} ntiles_j, nonzero blocks sorted by key,
//                                   packed contiguously, row-major inside a block)
//   * contraction loop ............ src/tamm/multop.hpp:576-893 (execute_bufacc: output
//                                   blocks in row-major label order, reduction labels =
//                                   sorted unique RHS labels minus LHS labels, zero
//                                   blocks skipped, cbuf += per pair, then C.add(cbuf))
//   * block kernel ................ src/tamm/kernels/multiply.hpp:212-381 (label classes
//                                   batch/aouter/bouter/inner/areduce/breduce; inter orders
//                                   [AR,B,M,K], [BR,B,K,N], [B,M,N]; gemm loops :42-74;
//                                   row-major gemm as src/tamm/kernels/cpu_blas.cpp:5-10)
//   * block assign ................ src/tamm/kernels/assign.hpp:312-485 (dst(dlabels) =/+=
//                                   scale * src(slabels); label permutation, or the general
//                                   loop over the union of labels when they differ)
//   * norm ........................ src/tamm/tamm_utils.hpp:1197-1254 (sqrt of the sum of
//                                   squares over all stored blocks)
//
// Pinning (how each part is checked, see tests/test_oracle.py):
//   tiling    -> tile vectors quoted by TAMM's tests/survey (cfg shapes), boundary rules
//   sparsity  -> closed-form survival fraction (3/8 for 4-index spin tensors), brute force
//   layout    -> key/offset invariants, brute-force recomputation
//   contract  -> closed-form KATs of tests/tamm/Test_Ops.cpp and Test_Mult_Ops.cpp, and a
//                dense numpy.einsum over the fully materialised tensors (independent code)
//   assign    -> numpy.transpose round trips
//   norm      -> closed form (Test_Mult_Ops norm 2.625e8)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <dlfcn.h>
#endif

#include "tamm_oracle.hpp"

namespace orc {

thread_local std::string g_err;

// distribution.hpp:197-256, :335-356
static void build_layout(Tensor& t) {
  int r = t.nmodes();
  t.key_off.assign(r, 1);
  for (int i = r - 2; i >= 0; --i) t.key_off[i] = t.key_off[i + 1] * t.modes[i + 1]->ntiles();
  std::vector<int64_t> bid(r, 0);
  std::vector<std::pair<int64_t, int64_t>> kv;
  bool done = false;
  while (!done) {
    bool empty = false;
    for (int i = 0; i < r; ++i) if (t.modes[i]->ntiles() == 0) empty = true;
    if (empty) break;
    if (t.is_non_zero(bid.data())) kv.push_back({t.key(bid.data()), t.block_size(bid.data())});
    int d = r - 1;
    while (d >= 0) {
      if (++bid[d] < t.modes[d]->ntiles()) break;
      bid[d] = 0;
      --d;
    }
    if (d < 0) done = true;
  }
  std::sort(kv.begin(), kv.end());
  int64_t off = 0;
  for (auto& p : kv) {
    t.keys.push_back(p.first);
    t.offs.push_back(off);
    t.sizes.push_back(p.second);
    off += p.second;
  }
  t.total = off;
}

// ----------------------------------------------------------------------------
// kernels::assign restatement (assign.hpp:312-485): dst(dl) (+)= scale * src(sl)
// General form = loop over the union of labels (destination labels first), as in
// ip_gen_loop; a permutation is the special case where both label sets agree.
static void assign(double* dst, const std::vector<int64_t>& ddims, const std::vector<int>& dl,
                   double scale, const double* src, const std::vector<int64_t>& sdims,
                   const std::vector<int>& sl, bool is_assign) {
  size_t nd = ddims.size(), ns = sdims.size();
  std::vector<int64_t> sld(ns + 1, 1), dld(nd + 1, 1);
  for (int i = (int)ns - 1; i >= 0; --i) sld[i] = sld[i + 1] * sdims[i];
  for (int i = (int)nd - 1; i >= 0; --i) dld[i] = dld[i + 1] * ddims[i];
  std::vector<int> loop;
  for (int l : dl) if (std::find(loop.begin(), loop.end(), l) == loop.end()) loop.push_back(l);
  for (int l : sl) if (std::find(loop.begin(), loop.end(), l) == loop.end()) loop.push_back(l);
  size_t nl = loop.size();
  std::vector<int64_t> ext(nl, 1), sst(nl, 0), dst_st(nl, 0);
  for (size_t i = 0; i < nl; ++i) {
    for (size_t s = 0; s < ns; ++s)
      if (sl[s] == loop[i]) { sst[i] += sld[s + 1]; ext[i] = sdims[s]; }
    for (size_t d = 0; d < nd; ++d)
      if (dl[d] == loop[i]) { dst_st[i] += dld[d + 1]; ext[i] = ddims[d]; }
  }
  int64_t dtotal = dld[0];
  int64_t total = 1;
  for (auto e : ext) total *= e;
  // A permutation (every loop label is a destination label, and the loop covers the
  // destination once: no repeated labels) visits each destination element exactly once: an
  // assignment then stores instead of clearing and accumulating.
  bool perm = total == dtotal;
  for (size_t i = 0; i < nl; ++i) perm = perm && dst_st[i] != 0;
  const bool store = is_assign && perm;
  if (is_assign && !perm) std::fill(dst, dst + dtotal, 0.0);
  if (nl == 0) {  // scalars
    dst[0] = (store ? 0.0 : dst[0]) + scale * src[0];
    return;
  }
  if (total == 0) return;
  // Odometer over the loop labels in order (the last fastest), stepped once per run of the
  // last label, which is a strided loop of its own: the same visiting order, hence the same
  // accumulation order per destination element, as stepping element by element.
  const int64_t run = ext[nl - 1], s_run = sst[nl - 1], d_run = dst_st[nl - 1];
  std::vector<int64_t> it(nl, 0);
  for (int64_t c = 0; c < total; c += run) {
    int64_t so = 0, dof = 0;
    for (size_t i = 0; i + 1 < nl; ++i) { so += it[i] * sst[i]; dof += it[i] * dst_st[i]; }
    const double* sp = src + so;
    double* dp = dst + dof;
    if (store) {
      for (int64_t k = 0; k < run; ++k) dp[k * d_run] = scale * sp[k * s_run];
    } else if (s_run == 1 && d_run == 1) {
      for (int64_t k = 0; k < run; ++k) dp[k] += scale * sp[k];
    } else {
      for (int64_t k = 0; k < run; ++k) dp[k * d_run] += scale * sp[k * s_run];
    }
    for (int i = (int)nl - 2; i >= 0; --i) {
      if (++it[i] < ext[i]) break;
      it[i] = 0;
    }
  }
}

// Optional system BLAS for the GEMMs (TAMM's CPU path calls blas::gemm, cpu_blas.cpp:5-10):
// cblas_dgemm from a library named by orc_use_system_blas, else the built-in loops below.
// Parity checks use the built-in loops; the CPU baseline may time either and records which.
using cblas_dgemm_fn = void (*)(int, int, int, int, int, int, double, const double*, int,
                                const double*, int, double, double*, int);
static cblas_dgemm_fn g_dgemm = nullptr;
static std::string g_gemm_backend = "built-in blocked loops (OpenMP)";

// Row-major C[MxN] += alpha * A[MxK] * B[KxN] (cpu_blas.cpp:5-10 semantics with beta=1
// on a zeroed intermediate).  Blocked, OpenMP over row panels.
static void gemm_acc(int64_t M, int64_t N, int64_t K, double alpha, const double* A,
                     const double* B, double* C) {
  if (g_dgemm != nullptr && M < INT32_MAX && N < INT32_MAX && K < INT32_MAX) {
    constexpr int kRowMajor = 101, kNoTrans = 111;  // CBLAS enumerators
    g_dgemm(kRowMajor, kNoTrans, kNoTrans, (int)M, (int)N, (int)K, alpha, A, (int)K, B, (int)N, 1.0,
            C, (int)N);
    return;
  }
  const int64_t MB = 32, KB = 256;
  int64_t work = M * N * K;
#pragma omp parallel for schedule(dynamic, 1) if (work > (1 << 20))
  for (int64_t i0 = 0; i0 < M; i0 += MB) {
    int64_t i1 = std::min(M, i0 + MB);
    std::vector<double> acc((size_t)(i1 - i0) * N, 0.0);
    for (int64_t k0 = 0; k0 < K; k0 += KB) {
      int64_t k1 = std::min(K, k0 + KB);
      for (int64_t i = i0; i < i1; ++i) {
        double* ci = acc.data() + (i - i0) * N;
        const double* ai = A + i * K;
        for (int64_t k = k0; k < k1; ++k) {
          double a = ai[k];
          const double* bk = B + k * N;
          for (int64_t j = 0; j < N; ++j) ci[j] += a * bk[j];
        }
      }
    }
    for (int64_t i = i0; i < i1; ++i)
      for (int64_t j = 0; j < N; ++j) C[i * N + j] += alpha * acc[(i - i0) * N + j];
  }
}

// kernels::block_multiply CPU branch (multiply.hpp:212-381, real path :358-381):
// cbuf(cl) += alpha * sum over reduce labels of A(al) * B(bl).
static void block_multiply(double alpha, const double* abuf, const std::vector<int64_t>& adims,
                           const std::vector<int>& al, const double* bbuf,
                           const std::vector<int64_t>& bdims, const std::vector<int>& bl,
                           double* cbuf, const std::vector<int64_t>& cdims,
                           const std::vector<int>& cl) {
  auto in = [](const std::vector<int>& v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); };
  std::vector<int> batch, aouter, bouter, inner, ared, bred;
  std::vector<int64_t> batch_d, aouter_d, bouter_d, inner_d, ared_d, bred_d;
  int64_t Bn = 1, M = 1, N = 1, K = 1, AR = 1, BR = 1;
  for (size_t i = 0; i < cl.size(); ++i) {  // multiply.hpp:238-263
    bool ia = in(al, cl[i]), ib = in(bl, cl[i]);
    if (ia && ib) { batch.push_back(cl[i]); batch_d.push_back(cdims[i]); Bn *= cdims[i]; }
    else if (ia) { aouter.push_back(cl[i]); aouter_d.push_back(cdims[i]); M *= cdims[i]; }
    else if (ib) { bouter.push_back(cl[i]); bouter_d.push_back(cdims[i]); N *= cdims[i]; }
  }
  for (size_t i = 0; i < al.size(); ++i) {  // :265-283
    bool ib = in(bl, al[i]), ic = in(cl, al[i]);
    if (ib && !ic) { inner.push_back(al[i]); inner_d.push_back(adims[i]); K *= adims[i]; }
    else if (!ib && !ic) { ared.push_back(al[i]); ared_d.push_back(adims[i]); AR *= adims[i]; }
  }
  for (size_t i = 0; i < bl.size(); ++i) {  // :285-302
    bool ia = in(al, bl[i]), ic = in(cl, bl[i]);
    if (!ia && !ic) { bred.push_back(bl[i]); bred_d.push_back(bdims[i]); BR *= bdims[i]; }
  }
  auto cat = [](std::initializer_list<const std::vector<int>*> parts) {
    std::vector<int> r;
    for (auto p : parts) r.insert(r.end(), p->begin(), p->end());
    return r;
  };
  auto catd = [](std::initializer_list<const std::vector<int64_t>*> parts) {
    std::vector<int64_t> r;
    for (auto p : parts) r.insert(r.end(), p->begin(), p->end());
    return r;
  };
  // :304-330
  std::vector<int> ail = cat({&ared, &batch, &aouter, &inner});
  std::vector<int> bil = cat({&bred, &batch, &inner, &bouter});
  std::vector<int> cil = cat({&batch, &aouter, &bouter});
  std::vector<int64_t> aid = catd({&ared_d, &batch_d, &aouter_d, &inner_d});
  std::vector<int64_t> bid = catd({&bred_d, &batch_d, &inner_d, &bouter_d});
  std::vector<int64_t> cid = catd({&batch_d, &aouter_d, &bouter_d});
  int64_t asz = AR * Bn * M * K, bsz = BR * Bn * K * N, csz = Bn * M * N;
  // staging buffers: the assigns below write every element of ai and bi
  std::unique_ptr<double[]> ai(new double[(size_t)std::max<int64_t>(asz, 1)]);
  std::unique_ptr<double[]> bi(new double[(size_t)std::max<int64_t>(bsz, 1)]);
  std::vector<double> ci(csz, 0.0);
  assign(ai.get(), aid, ail, 1.0, abuf, adims, al, true);  // transpose_inputs :191-192
  assign(bi.get(), bid, bil, 1.0, bbuf, bdims, bl, true);
  for (int64_t ar = 0; ar < AR; ++ar)  // gemm_wrapper :42-74
    for (int64_t br = 0; br < BR; ++br)
      for (int64_t b = 0; b < Bn; ++b)
        gemm_acc(M, N, K, alpha, ai.get() + ar * Bn * M * K + b * M * K,
                 bi.get() + br * Bn * K * N + b * K * N, ci.data() + b * M * N);
  assign(cbuf, cdims, cl, 1.0, ci.data(), cid, cil, false);  // transpose_output :196-210
}

struct LabeledOperand {
  const Tensor* t;
  std::vector<int> labels;
  std::vector<int64_t> toff;  // tile offset of the label's space inside the mode's space
};

// execute_bufacc restatement (multop.hpp:576-893), one process.
// label_ntiles: number of tiles of each label's own tiled space, keyed by label id.
// sel_keys (optional): compute only these C blocks (keys of C's layout).
static void contract(const LabeledOperand& C, double* Cd, double alpha, const LabeledOperand& A,
                     const double* Ad, const LabeledOperand& B, const double* Bd,
                     const std::map<int, int64_t>& label_ntiles,
                     const std::vector<int64_t>* sel_keys, int64_t* npairs_out) {
  const std::vector<int>& cl = C.labels;
  std::vector<int> rhs = A.labels;
  rhs.insert(rhs.end(), B.labels.begin(), B.labels.end());
  std::vector<int> uniq = rhs;
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  std::vector<int> lsorted = cl;
  std::sort(lsorted.begin(), lsorted.end());
  std::vector<int> red;
  std::set_difference(uniq.begin(), uniq.end(), lsorted.begin(), lsorted.end(), std::back_inserter(red));
  auto map_of = [&](const std::vector<int>& labs, const std::vector<int>& target) {
    std::vector<int> m;
    for (int l : labs) {
      auto it = std::find(target.begin(), target.end(), l);
      m.push_back(it == target.end() ? -1 : (int)(it - target.begin()));
    }
    return m;
  };
  auto a_out = map_of(A.labels, cl), a_red = map_of(A.labels, red);
  auto b_out = map_of(B.labels, cl), b_red = map_of(B.labels, red);
  std::vector<int64_t> cext, rext;
  for (int l : cl) cext.push_back(label_ntiles.at(l));
  for (int l : red) rext.push_back(label_ntiles.at(l));

  int64_t npairs = 0;
  std::vector<int64_t> it(cl.size(), 0);
  auto next = [](std::vector<int64_t>& v, const std::vector<int64_t>& e) {
    for (int i = (int)v.size() - 1; i >= 0; --i) {
      if (++v[i] < e[i]) return true;
      v[i] = 0;
    }
    return false;
  };
  auto empty_nest = [](const std::vector<int64_t>& e) {
    for (auto x : e) if (x == 0) return true;
    return false;
  };
  if (empty_nest(cext)) return;
  do {
    std::vector<int64_t> cb(cl.size());
    for (size_t i = 0; i < cl.size(); ++i) cb[i] = C.toff[i] + it[i];
    if (!C.t->is_non_zero(cb.data())) continue;
    if (sel_keys) {
      int64_t k = C.t->key(cb.data());
      if (!std::binary_search(sel_keys->begin(), sel_keys->end(), k)) continue;
    }
    std::vector<int64_t> cdims(cl.size());
    for (size_t i = 0; i < cl.size(); ++i) cdims[i] = C.t->modes[i]->tsize(cb[i]);
    int64_t csize = C.t->block_size(cb.data());
    std::vector<double> cbuf(csize, 0.0);
    std::vector<int64_t> rit(red.size(), 0);
    if (!empty_nest(rext)) {
      do {
        std::vector<int64_t> ab(A.labels.size()), bb(B.labels.size());
        for (size_t i = 0; i < A.labels.size(); ++i)
          ab[i] = A.toff[i] + (a_out[i] != -1 ? it[a_out[i]] : rit[a_red[i]]);
        if (!A.t->is_non_zero(ab.data())) continue;
        for (size_t i = 0; i < B.labels.size(); ++i)
          bb[i] = B.toff[i] + (b_out[i] != -1 ? it[b_out[i]] : rit[b_red[i]]);
        if (!B.t->is_non_zero(bb.data())) continue;
        std::vector<int64_t> adims(ab.size()), bdims(bb.size());
        for (size_t i = 0; i < ab.size(); ++i) adims[i] = A.t->modes[i]->tsize(ab[i]);
        for (size_t i = 0; i < bb.size(); ++i) bdims[i] = B.t->modes[i]->tsize(bb[i]);
        block_multiply(alpha, Ad + A.t->locate(ab.data()), adims, A.labels,
                       Bd + B.t->locate(bb.data()), bdims, B.labels, cbuf.data(), cdims, cl);
        ++npairs;
      } while (next(rit, rext));
    }
    double* cdst = Cd + C.t->locate(cb.data());  // ctensor.add (multop.hpp:837)
    for (int64_t e = 0; e < csize; ++e) cdst[e] += cbuf[e];
  } while (next(it, cext));
  if (npairs_out) *npairs_out = npairs;
}

}  // namespace orc

// ============================================================================
// C interface (ctypes)
extern "C" {

const char* orc_last_error() { return orc::g_err.c_str(); }

// spin_*: nspin ranges [lo,hi) with values; irrep_*: nirrep ranges; named_*: nnamed ranges.
void* orc_space_new(int64_t n, int use_vector, const int64_t* tiles, int64_t ntiles, int nspin,
                    const int64_t* spin_lo, const int64_t* spin_hi, const int* spin_val, int nirrep,
                    const int64_t* ir_lo, const int64_t* ir_hi, const int* ir_val, int nnamed,
                    const int64_t* nm_lo, const int64_t* nm_hi) {
  try {
    std::vector<orc::AttrRange> sp, ir;
    std::vector<std::pair<int64_t, int64_t>> nm;
    for (int i = 0; i < nspin; ++i) sp.push_back({spin_lo[i], spin_hi[i], spin_val[i]});
    for (int i = 0; i < nirrep; ++i) ir.push_back({ir_lo[i], ir_hi[i], ir_val[i]});
    for (int i = 0; i < nnamed; ++i) nm.push_back({nm_lo[i], nm_hi[i]});
    return new orc::Space(orc::make_space(n, use_vector, tiles, ntiles, sp, ir, nm));
  } catch (std::exception& e) {
    orc::g_err = e.what();
    return nullptr;
  }
}
void orc_space_free(void* s) { delete (orc::Space*)s; }
int64_t orc_space_ntiles(void* s) { return ((orc::Space*)s)->ntiles(); }
void orc_space_tiles(void* s, int64_t* offsets, int* spin, int* irrep) {
  auto* sp = (orc::Space*)s;
  for (int64_t t = 0; t <= sp->ntiles(); ++t) offsets[t] = sp->toff[t];
  for (int64_t t = 0; t < sp->ntiles(); ++t) { spin[t] = sp->tspin[t]; irrep[t] = sp->tirrep[t]; }
}

// nupper < 0: no spin mask (TensorBase::has_spin() false).
void* orc_tensor_new(int nmodes, void** spaces, int nupper, int nlower, int irrep_rule) {
  auto* t = new orc::Tensor();
  for (int i = 0; i < nmodes; ++i) t->modes.push_back((const orc::Space*)spaces[i]);
  if (nupper >= 0) {
    t->has_spin = true;
    for (int i = 0; i < nmodes; ++i) t->mask.push_back(i < nupper ? 0 : (i < nupper + nlower ? 1 : 2));
  } else {
    t->mask.assign(nmodes, 2);
  }
  t->irrep_rule = irrep_rule != 0;
  orc::build_layout(*t);
  return t;
}
void orc_tensor_free(void* t) { delete (orc::Tensor*)t; }
int64_t orc_tensor_nblocks(void* t) { return (int64_t)((orc::Tensor*)t)->keys.size(); }
int64_t orc_tensor_size(void* t) { return ((orc::Tensor*)t)->total; }
void orc_tensor_blocks(void* t, int64_t* keys, int64_t* offs, int64_t* sizes) {
  auto* tt = (orc::Tensor*)t;
  for (size_t i = 0; i < tt->keys.size(); ++i) { keys[i] = tt->keys[i]; offs[i] = tt->offs[i]; sizes[i] = tt->sizes[i]; }
}
int orc_tensor_is_non_zero(void* t, const int64_t* bid) { return ((orc::Tensor*)t)->is_non_zero(bid); }

// labels/toff/ntl arrays per operand: int label ids, tile offset of the label's space inside
// the mode's space, number of tiles of the label's space.
int orc_contract(void* Ct, const int* cl, const int64_t* coff, const int64_t* cnt, double* Cd,
                 double alpha, void* At, const int* al, const int64_t* aoff, const int64_t* ant,
                 const double* Ad, void* Bt, const int* bl, const int64_t* boff, const int64_t* bnt,
                 const double* Bd, int64_t nsel, const int64_t* sel_keys, int64_t* npairs) {
  try {
    auto mk = [](void* t, const int* l, const int64_t* off) {
      orc::LabeledOperand o;
      o.t = (const orc::Tensor*)t;
      for (int i = 0; i < o.t->nmodes(); ++i) {
        o.labels.push_back(l[i]);
        o.toff.push_back(off ? off[i] : 0);
      }
      return o;
    };
    orc::LabeledOperand C = mk(Ct, cl, coff), A = mk(At, al, aoff), B = mk(Bt, bl, boff);
    std::map<int, int64_t> ntl;
    auto addn = [&](const orc::LabeledOperand& o, const int64_t* nt) {
      for (int i = 0; i < o.t->nmodes(); ++i) {
        int64_t v = nt ? nt[i] : o.t->modes[i]->ntiles();
        auto f = ntl.find(o.labels[i]);
        if (f != ntl.end() && f->second != v) throw std::runtime_error("label used with two different tile counts");
        ntl[o.labels[i]] = v;
      }
    };
    addn(C, cnt); addn(A, ant); addn(B, bnt);
    std::vector<int64_t> sel;
    if (nsel > 0) { sel.assign(sel_keys, sel_keys + nsel); std::sort(sel.begin(), sel.end()); }
    orc::contract(C, Cd, alpha, A, Ad, B, Bd, ntl, nsel > 0 ? &sel : nullptr, npairs);
    return 0;
  } catch (std::exception& e) {
    orc::g_err = e.what();
    return 1;
  }
}

// Block-level restatements (kernels::block_multiply CPU branch, kernels::assign).
void orc_block_multiply(double alpha, const double* A, int an, const int64_t* adims, const int* al,
                        const double* B, int bn, const int64_t* bdims, const int* bl, double* C,
                        int cn, const int64_t* cdims, const int* cl) {
  orc::block_multiply(alpha, A, std::vector<int64_t>(adims, adims + an), std::vector<int>(al, al + an),
                      B, std::vector<int64_t>(bdims, bdims + bn), std::vector<int>(bl, bl + bn), C,
                      std::vector<int64_t>(cdims, cdims + cn), std::vector<int>(cl, cl + cn));
}
void orc_assign(double* dst, int dn, const int64_t* ddims, const int* dl, double scale,
                const double* src, int sn, const int64_t* sdims, const int* sl, int is_assign) {
  orc::assign(dst, std::vector<int64_t>(ddims, ddims + dn), std::vector<int>(dl, dl + dn), scale, src,
              std::vector<int64_t>(sdims, sdims + sn), std::vector<int>(sl, sl + sn), is_assign != 0);
}

// tamm_utils.hpp:1197-1254
double orc_norm(const double* data, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += data[i] * data[i];
  return std::sqrt(s);
}

// Loads cblas_dgemm from the shared library at `path` for the GEMMs (1 on success).  An MKL
// runtime is switched to the GNU OpenMP threading layer, the oracle's own runtime.
int orc_use_system_blas(const char* path) {
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (h == nullptr) return 0;
  using layer_fn = int (*)(int);
  constexpr int kMklThreadingGnu = 3;
  if (auto set_layer = reinterpret_cast<layer_fn>(dlsym(h, "MKL_Set_Threading_Layer")))
    set_layer(kMklThreadingGnu);
  auto fn = reinterpret_cast<orc::cblas_dgemm_fn>(dlsym(h, "cblas_dgemm"));
  if (fn == nullptr) return 0;
  orc::g_dgemm = fn;
  orc::g_gemm_backend = std::string("cblas_dgemm from ") + path;
  return 1;
}
void orc_use_builtin_gemm() {
  orc::g_dgemm = nullptr;
  orc::g_gemm_backend = "built-in blocked loops (OpenMP)";
}
const char* orc_gemm_backend() { return orc::g_gemm_backend.c_str(); }

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

}  // extern "C"
