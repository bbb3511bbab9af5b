// plan_bp.cpp -- label analysis, execute_bufacc pair enumeration, the fused block-pair plan,
// the contraction object and the block-level entry points.
#include <atomic>
#include <cmath>
#include <numeric>

#include "plan.hpp"

namespace tmx {

std::vector<int64_t> strides_of(const std::vector<int64_t>& dims) {
  std::vector<int64_t> s(dims.size(), 1);
  for (int i = (int)dims.size() - 2; i >= 0; --i) s[i] = s[i + 1] * dims[i + 1];
  return s;
}

LabelInfo analyse_labels(const LOperand& C, const LOperand& A, const LOperand& B) {
  LabelInfo li;
  auto uniq = [](const std::vector<int>& v, const char* who) {
    std::vector<int> s = v;
    std::sort(s.begin(), s.end());
    TMX_REQUIRE(std::adjacent_find(s.begin(), s.end()) == s.end(),
                std::string("a label repeats within operand ") + who);
  };
  uniq(C.labels, "C");
  uniq(A.labels, "A");
  uniq(B.labels, "B");
  auto in = [](const std::vector<int>& v, int l) { return pos_of(v, l) >= 0; };
  for (int l : C.labels) {  // multiply.hpp:238-263
    bool ia = in(A.labels, l), ib = in(B.labels, l);
    TMX_REQUIRE(ia || ib, "an output label does not occur on the right-hand side");
    if (ia && ib) li.batch.push_back(l);
    else if (ia) li.aouter.push_back(l);
    else li.bouter.push_back(l);
  }
  for (int l : A.labels) {  // :265-283
    bool ib = in(B.labels, l), ic = in(C.labels, l);
    if (ib && !ic) li.inner.push_back(l);
    else if (!ib && !ic) li.ared.push_back(l);
  }
  for (int l : B.labels)  // :285-302
    if (!in(A.labels, l) && !in(C.labels, l)) li.bred.push_back(l);
  std::vector<int> rhs = A.labels;
  rhs.insert(rhs.end(), B.labels.begin(), B.labels.end());
  std::sort(rhs.begin(), rhs.end());
  rhs.erase(std::unique(rhs.begin(), rhs.end()), rhs.end());
  for (int l : rhs)
    if (!in(C.labels, l)) li.red.push_back(l);
  // Tile sizes of each label's space; every occurrence must agree (labels of one space).
  auto add = [&](const LOperand& o) {
    for (size_t i = 0; i < o.labels.size(); ++i) {
      std::vector<int64_t> ts;
      const Space& sp = *o.t->modes[i];
      for (int64_t t = 0; t < o.ntl[i]; ++t) ts.push_back(sp.tsize(o.toff[i] + t));
      auto it = li.tile_sizes.find(o.labels[i]);
      if (it == li.tile_sizes.end()) li.tile_sizes[o.labels[i]] = ts;
      else TMX_REQUIRE(it->second == ts, "a label is used over index spaces with different tilings");
    }
  };
  add(C);
  add(A);
  add(B);
  return li;
}

std::vector<int64_t> block_dims(const LOperand& op, const std::vector<int64_t>& label_tiles) {
  std::vector<int64_t> d;
  for (size_t i = 0; i < op.labels.size(); ++i)
    d.push_back(op.t->modes[i]->tsize(op.toff[i] + label_tiles[i]));
  return d;
}

std::vector<CBlockWork> enumerate_pairs(const LOperand& C, const LOperand& A, const LOperand& B,
                                        const LabelInfo& li) {
  std::vector<CBlockWork> out;
  const size_t nc = C.labels.size(), nr = li.red.size();
  std::vector<int64_t> cext(nc), rext(nr);
  for (size_t i = 0; i < nc; ++i) cext[i] = C.ntl[i];
  for (size_t i = 0; i < nr; ++i) rext[i] = (int64_t)li.tile_sizes.at(li.red[i]).size();
  for (auto e : cext) if (e == 0) return out;
  bool red_empty = false;
  for (auto e : rext) if (e == 0) red_empty = true;
  // operand label -> source (C label index or reduction label index)
  auto src_map = [&](const LOperand& o, std::vector<int>& from_c, std::vector<int>& from_r) {
    for (int l : o.labels) {
      from_c.push_back(pos_of(C.labels, l));
      from_r.push_back(pos_of(li.red, l));
    }
  };
  std::vector<int> a_c, a_r, b_c, b_r;
  src_map(A, a_c, a_r);
  src_map(B, b_c, b_r);
  auto next = [](std::vector<int64_t>& v, const std::vector<int64_t>& e) {
    for (int i = (int)v.size() - 1; i >= 0; --i) {
      if (++v[i] < e[i]) return true;
      v[i] = 0;
    }
    return false;
  };
  std::vector<int64_t> ct(nc, 0), cb(nc), ab(A.labels.size()), bb(B.labels.size());
  do {
    for (size_t i = 0; i < nc; ++i) cb[i] = C.toff[i] + ct[i];
    if (!C.t->is_non_zero(cb.data())) continue;
    CBlockWork w;
    w.c_off = C.t->find(cb.data());
    w.ct = ct;
    if (!red_empty) {
      std::vector<int64_t> rt(nr, 0);
      do {
        for (size_t i = 0; i < ab.size(); ++i)
          ab[i] = A.toff[i] + (a_c[i] >= 0 ? ct[a_c[i]] : rt[a_r[i]]);
        if (!A.t->is_non_zero(ab.data())) continue;
        for (size_t i = 0; i < bb.size(); ++i)
          bb[i] = B.toff[i] + (b_c[i] >= 0 ? ct[b_c[i]] : rt[b_r[i]]);
        if (!B.t->is_non_zero(bb.data())) continue;
        PairRef pr;
        pr.a_off = A.t->find(ab.data());
        pr.b_off = B.t->find(bb.data());
        pr.rt = rt;
        w.pairs.push_back(std::move(pr));
      } while (next(rt, rext));
    }
    out.push_back(std::move(w));
  } while (next(ct, cext));
  return out;
}

namespace {

Radix make_radix(const std::vector<int64_t>& ext, const std::vector<int64_t>& stride) {
  Radix r;
  std::memset(&r, 0, sizeof(r));
  TMX_REQUIRE(ext.size() <= (size_t)MAXR, "too many modes in one GEMM role");
  r.n = (int)ext.size();
  for (int i = 0; i < r.n; ++i) {
    TMX_REQUIRE(ext[i] < (1ll << 31) && stride[i] < (1ll << 31), "block too large");
    r.ext[i] = (int)ext[i];
    r.stride[i] = (int)stride[i];
  }
  return r;
}

template <class T>
std::vector<int64_t> key_of(const T& v) {
  std::vector<int64_t> k(sizeof(T) / sizeof(int32_t));
  const int32_t* p = reinterpret_cast<const int32_t*>(&v);
  for (size_t i = 0; i < k.size(); ++i) k[i] = p[i];
  return k;
}

template <class T>
int intern(std::vector<T>& table, std::map<std::vector<int64_t>, int>& index, const T& v) {
  auto k = key_of(v);
  auto it = index.find(k);
  if (it != index.end()) return it->second;
  int id = (int)table.size();
  table.push_back(v);
  index[k] = id;
  return id;
}

}  // namespace

static void build_bp_plan_as(BPPlan& plan, const LOperand& C, const LOperand& A, const LOperand& B,
                             const LabelInfo& li, const std::vector<CBlockWork>& work,
                             bool per_pair_launches, double* flops_out, int64_t* npairs_out,
                             int tile) {
  const int TM = bp_tile_rows(tile), TN = bp_tile_cols(tile);
  plan.tile = tile;
  std::vector<BPTask> tasks;
  std::vector<BPPair> pairs;
  std::vector<BPCShape> cshapes;
  std::vector<BPKShape> kshapes;
  std::map<std::vector<int64_t>, int> cidx, kidx;
  std::vector<BPTask> pair_tasks;
  double flops = 0;
  int64_t npairs = 0;
  std::vector<std::pair<int64_t, int>> cost_task;  // for LPT ordering

  // Label -> tile index lookup for one triple.
  auto tiles_for = [&](const LOperand& o, const std::vector<int64_t>& ct,
                       const std::vector<int64_t>& rt) {
    std::vector<int64_t> t;
    for (int l : o.labels) {
      int pc = pos_of(C.labels, l);
      t.push_back(pc >= 0 ? ct[pc] : rt[pos_of(li.red, l)]);
    }
    return t;
  };
  auto dims_of = [&](const LOperand& o, const std::vector<int64_t>& tiles, const std::vector<int>& ls,
                     std::vector<int64_t>& ext, std::vector<int64_t>& str) {
    std::vector<int64_t> d = block_dims(o, tiles), s = strides_of(d);
    ext.clear();
    str.clear();
    for (int l : ls) {
      int p = pos_of(o.labels, l);
      ext.push_back(p >= 0 ? d[p] : 0);
      str.push_back(p >= 0 ? s[p] : 0);
    }
  };
  std::vector<int> kl = li.inner;  // K modes: inner, A-reduce, B-reduce
  kl.insert(kl.end(), li.ared.begin(), li.ared.end());
  kl.insert(kl.end(), li.bred.begin(), li.bred.end());
  // Digit order of each GEMM role.  The M, N and K indices are enumerations private to this
  // plan (the A/B/C radices of a role all use one order), so each role's last, fastest digit
  // is chosen as the stride-1 mode of the operand the kernel reads along it: A's last mode
  // when it is an M mode, B's when it is an N mode, and for K the stride-1 K mode of the
  // larger operand (A's on a tie).  A wave's 16 or 64 consecutive indices are then one
  // contiguous run of that operand's block; in TAMM's label order they can be a stride of
  // thousands of elements (R(a,b,i,j) += W(k,b,c,j) T(a,c,i,k): K = (k,c) puts c, of stride
  // 256 in T, where the loads along T's k expect stride 1).
  std::vector<int> mo = li.aouter, no = li.bouter;
  auto to_back = [](std::vector<int>& v, int l) {
    auto it = std::find(v.begin(), v.end(), l);
    if (it == v.end()) return false;
    v.erase(it);
    v.push_back(l);
    return true;
  };
  if (!A.labels.empty()) to_back(mo, A.labels.back());
  if (!B.labels.empty()) to_back(no, B.labels.back());
  {
    const bool ak = !A.labels.empty() && pos_of(kl, A.labels.back()) >= 0;
    const bool bk = !B.labels.empty() && pos_of(kl, B.labels.back()) >= 0;
    if (ak && (!bk || A.t->total >= B.t->total)) to_back(kl, A.labels.back());
    else if (bk) to_back(kl, B.labels.back());
  }

  for (const CBlockWork& w : work) {
    if (w.pairs.empty()) continue;
    std::vector<int64_t> rt0(li.red.size(), 0);
    std::vector<int64_t> ctiles = tiles_for(C, w.ct, rt0);
    std::vector<int64_t> cext, cstr, nce, ncs, bte, bts;
    dims_of(C, ctiles, mo, cext, cstr);
    dims_of(C, ctiles, no, nce, ncs);
    dims_of(C, ctiles, li.batch, bte, bts);
    int64_t M = 1, N = 1, Bn = 1;
    for (auto e : cext) M *= e;
    for (auto e : nce) N *= e;
    for (auto e : bte) Bn *= e;
    BPCShape cs;
    std::memset(&cs, 0, sizeof(cs));
    cs.M = (int)M;
    cs.N = (int)N;
    cs.cm = make_radix(cext, cstr);
    cs.cn = make_radix(nce, ncs);
    int csid = intern(cshapes, cidx, cs);

    // Per pair: K shape and base offsets (batch element added below).
    struct PInfo { BPPair p; std::vector<int64_t> abt, bbt; int64_t K; };
    std::vector<PInfo> pinfo;
    for (const PairRef& pr : w.pairs) {
      std::vector<int64_t> at = tiles_for(A, w.ct, pr.rt), bt = tiles_for(B, w.ct, pr.rt);
      BPKShape ks;
      std::memset(&ks, 0, sizeof(ks));
      std::vector<int64_t> e, s, e2, s2;
      dims_of(A, at, mo, e, s);
      ks.am = make_radix(e, s);
      dims_of(B, bt, no, e, s);
      ks.bn = make_radix(e, s);
      // K radix: extents from whichever operand holds the label, strides 0 where absent.
      std::vector<int64_t> kext, ka, kb;
      {
        std::vector<int64_t> ad = block_dims(A, at), as = strides_of(ad);
        std::vector<int64_t> bd = block_dims(B, bt), bs = strides_of(bd);
        for (int l : kl) {
          int pa = pos_of(A.labels, l), pb = pos_of(B.labels, l);
          kext.push_back(pa >= 0 ? ad[pa] : bd[pb]);
          ka.push_back(pa >= 0 ? as[pa] : 0);
          kb.push_back(pb >= 0 ? bs[pb] : 0);
        }
        int64_t K = 1;
        for (auto x : kext) K *= x;
        TMX_REQUIRE(K < (1ll << 31), "reduction block too large");
        ks.K = (int)K;
        ks.ak = make_radix(kext, ka);
        ks.bk = make_radix(kext, kb);
        // Read along K when the operand's stride-1 mode is the fastest K digit (above), else
        // along its rows, whose fastest digit is then its stride-1 mode where that is an
        // outer mode.
        ks.a_kcontig = (!kl.empty() && !A.labels.empty() && A.labels.back() == kl.back()) ? 1 : 0;
        ks.b_kcontig = (!kl.empty() && !B.labels.empty() && B.labels.back() == kl.back()) ? 1 : 0;
        PInfo pi;
        pi.p.a_off = pr.a_off;
        pi.p.b_off = pr.b_off;
        pi.p.kshape = intern(kshapes, kidx, ks);
        pi.p.pad = 0;
        pi.abt = at;
        pi.bbt = bt;
        pi.K = K;
        pinfo.push_back(pi);
        flops += 2.0 * (double)M * (double)N * (double)K * (double)Bn;
        ++npairs;
      }
    }
    // Batch elements: offsets of each batch index in C, A (per pair), B (per pair).
    std::vector<int64_t> bidx(li.batch.size(), 0);
    for (int64_t be = 0; be < Bn; ++be) {
      int64_t coff = w.c_off;
      {
        int64_t rem = be;
        for (int i = (int)li.batch.size() - 1; i >= 0; --i) { bidx[i] = rem % bte[i]; rem /= bte[i]; }
        for (size_t i = 0; i < li.batch.size(); ++i) coff += bidx[i] * bts[i];
      }
      int pbeg = (int)pairs.size();
      int64_t kcost = 0;
      for (const PInfo& pi : pinfo) {
        BPPair p = pi.p;
        std::vector<int64_t> e, s;
        dims_of(A, pi.abt, li.batch, e, s);
        for (size_t i = 0; i < li.batch.size(); ++i) p.a_off += bidx[i] * s[i];
        dims_of(B, pi.bbt, li.batch, e, s);
        for (size_t i = 0; i < li.batch.size(); ++i) p.b_off += bidx[i] * s[i];
        pairs.push_back(p);
        kcost += pi.K;
      }
      int pend = (int)pairs.size();
      for (int64_t m0 = 0; m0 < M; m0 += TM)
        for (int64_t n0 = 0; n0 < N; n0 += TN) {
          BPTask t;
          t.c_off = coff;
          t.m0 = (int)m0;
          t.n0 = (int)n0;
          t.cshape = csid;
          t.pair_begin = pbeg;
          t.pair_end = pend;
          t.pad = 0;
          cost_task.push_back({kcost, (int)tasks.size()});
          tasks.push_back(t);
        }
    }
  }
  // Tasks stay in C-block order (the tiles of a block in a row, neighbouring blocks sharing
  // operand blocks) and are cut into 8 contiguous ranges of equal reduction work, one per XCD,
  // so each XCD's L2 serves the re-reads of its own blocks' operands.
  std::vector<BPTask> ordered = tasks;
  {
    double total = 0;
    for (auto& ct : cost_task) total += (double)ct.first;
    int x = 0;
    double acc = 0;
    plan.xcd.lo[0] = 0;
    for (size_t i = 0; i < cost_task.size(); ++i) {
      acc += (double)cost_task[i].first;
      while (x < 7 && acc >= total * (x + 1) / 8.0) {
        plan.xcd.hi[x] = (int)i + 1;
        plan.xcd.lo[++x] = (int)i + 1;
      }
    }
    for (; x < 7; ) { plan.xcd.hi[x] = (int)cost_task.size(); plan.xcd.lo[++x] = (int)cost_task.size(); }
    plan.xcd.hi[7] = (int)cost_task.size();
  }
  if (per_pair_launches) {
    // One launch per block pair, in execute_bufacc order: the tiles of the C block, one pair.
    std::vector<BPTask> byblock = tasks;  // original (block-major) order
    size_t i = 0;
    while (i < byblock.size()) {
      size_t j = i;
      while (j < byblock.size() && byblock[j].pair_begin == byblock[i].pair_begin &&
             byblock[j].c_off == byblock[i].c_off)
        ++j;
      for (int p = byblock[i].pair_begin; p < byblock[i].pair_end; ++p) {
        int beg = (int)pair_tasks.size();
        for (size_t k = i; k < j; ++k) {
          BPTask t = byblock[k];
          t.pair_begin = p;
          t.pair_end = p + 1;
          pair_tasks.push_back(t);
        }
        plan.per_pair.push_back({beg, (int)pair_tasks.size()});
        plan.per_pair_first.push_back(p == byblock[i].pair_begin ? 1 : 0);
      }
      i = j;
    }
    plan.pair_tasks = upload(pair_tasks);
  }
  {
    std::vector<int64_t> zoff, zpre{0};
    for (const CBlockWork& w : work) {
      if (!w.pairs.empty()) continue;
      int64_t n = 1;
      for (int64_t d : block_dims(C, w.ct)) n *= d;
      if (n == 0) continue;
      zoff.push_back(w.c_off);
      zpre.push_back(zpre.back() + n);
    }
    plan.nzero = (int)zoff.size();
    plan.zero_elems = zpre.back();
    if (plan.nzero > 0) {
      plan.zero_off = upload(zoff);
      plan.zero_prefix = upload(zpre);
    }
  }
  plan.ntasks = (int)ordered.size();
  plan.tasks = upload(ordered);
  plan.pairs = upload(pairs);
  plan.cshapes = upload(cshapes);
  plan.kshapes = upload(kshapes);
  if (flops_out) *flops_out = flops;
  if (npairs_out) *npairs_out = npairs;
}

void build_bp_plan(BPPlan& plan, const LOperand& C, const LOperand& A, const LOperand& B,
                   const LabelInfo& li, const std::vector<CBlockWork>& work,
                   bool per_pair_launches, double* flops_out, int64_t* npairs_out) {
  // Tile shape from the C blocks' row (A-outer) and column (B-outer) extents: the MFMA work
  // each shape pads to, weighted by the blocks' reduction lengths.  The 16x256 tile pays off
  // when one side is a single small tile (one-body terms, their adjoints); the plan then
  // puts that side on the rows, swapping A and B when it is the column side.
  double w64 = 0, w_rows = 0, w_cols = 0;
  auto pad = [](double x, double t) { return std::ceil(x / t) * t; };
  for (const CBlockWork& w : work) {
    if (w.pairs.empty()) continue;
    const std::vector<int64_t> cd = block_dims(C, w.ct);
    double M = 1, N = 1, Bn = 1, K = 0;
    for (size_t i = 0; i < C.labels.size(); ++i) {
      const int l = C.labels[i];
      if (pos_of(li.aouter, l) >= 0) M *= (double)cd[i];
      else if (pos_of(li.bouter, l) >= 0) N *= (double)cd[i];
      else Bn *= (double)cd[i];
    }
    for (const PairRef& pr : w.pairs) {
      double k = 1;
      for (size_t i = 0; i < li.red.size(); ++i) k *= (double)li.tile_sizes.at(li.red[i])[pr.rt[i]];
      K += k;
    }
    w64 += pad(M, 64) * pad(N, 64) * K * Bn;
    w_rows += pad(M, 16) * pad(N, 256) * K * Bn;
    w_cols += pad(N, 16) * pad(M, 256) * K * Bn;
  }
  // The 16x256 tile reads (16 + 256) / (16 * 256) operand elements per MFMA product against
  // 2 / 64 for 64x64: taken only when it saves a third of the padded work or more.
  if (std::min(w_rows, w_cols) * 1.5 < w64) {
    if (w_cols < w_rows) {
      const LabelInfo li2 = analyse_labels(C, B, A);
      const std::vector<CBlockWork> work2 = enumerate_pairs(C, B, A, li2);
      build_bp_plan_as(plan, C, B, A, li2, work2, per_pair_launches, flops_out, npairs_out,
                       BP_TILE_16x256);
      plan.swapped = true;
      return;
    }
    build_bp_plan_as(plan, C, A, B, li, work, per_pair_launches, flops_out, npairs_out,
                     BP_TILE_16x256);
    return;
  }
  build_bp_plan_as(plan, C, A, B, li, work, per_pair_launches, flops_out, npairs_out,
                   BP_TILE_64x64);
}

void launch_bp_plan(const BPPlan& bp, const BPTask* tasks, int ntasks, const BPXcd* xr,
                    const double* A, const double* B, double* C, double alpha, int is_assign,
                    hipStream_t s) {
  if (bp.swapped) std::swap(A, B);
  launch_bp_f64(tasks, ntasks, bp.pairs.as<BPPair>(), bp.cshapes.as<BPCShape>(),
                bp.kshapes.as<BPKShape>(), A, B, C, alpha, is_assign, s, xr, bp.tile);
}

}  // namespace tmx

using namespace tmx;

// ---------------------------------------------------------------------------
// Contraction object
struct tmx_contraction_s {
  int mode = TMX_EXEC_GROUPED_F64;
  double alpha = 1.0;
  LOperand C, A, B;
  double flops = 0;
  int64_t npairs = 0;
  BPPlan bp;
  SectorPlanPtr sec;
  // Inputs of the sector plan, kept so the autotuner can rebuild it with other split-K counts.
  // When C's labels list the B-outer labels before the A-outer ones, the sector plan is built
  // for C = B*A (swapped = true): a C block is then row-major over the plan's dense rows and
  // columns, which lets the split path add its results into C in one fused pass.
  LabelInfo li;
  std::vector<CBlockWork> work;
  int split_bits = 0;
  int parallel_split = 0;  // tmx_set_parallel_split at creation; 0 lets the autotuner measure split-K
  bool swapped = false;
  // tmx_contraction_set_constant: operands packed once and kept (call's A / B).
  bool const_a = false, const_b = false;
  // tmx_contraction_set_assign: C = alpha*A*B instead of C += alpha*A*B.
  bool assign = false;
  // Holds of live graphs that captured an execution of the contraction (GraphHold).
  std::atomic<int> graph_holds{0};
  const LOperand& sa() const { return swapped ? B : A; }
  const LOperand& sb() const { return swapped ? A : B; }
};

// Contraction and AddOp handles are owned through a registry of shared pointers: destroy drops
// the registry's reference, and a graph that captured the object holds another
// (capture_hold), so the object and its device memory live until both are gone.
template <class T>
struct HandleRegistry {
  std::mutex mu;
  std::map<const T*, std::shared_ptr<T>> live;
  T* add(std::shared_ptr<T> p) {
    std::lock_guard<std::mutex> g(mu);
    T* raw = p.get();
    live[raw] = std::move(p);
    return raw;
  }
  std::shared_ptr<T> get(const T* raw) {
    std::lock_guard<std::mutex> g(mu);
    auto it = live.find(raw);
    return it == live.end() ? nullptr : it->second;
  }
  // References beyond the registry's own (graphs holding the object).
  long held(const T* raw) {
    std::lock_guard<std::mutex> g(mu);
    auto it = live.find(raw);
    return it == live.end() ? 0 : it->second.use_count() - 1;
  }
  void drop(const T* raw) {
    std::shared_ptr<T> p;  // released outside the lock
    std::lock_guard<std::mutex> g(mu);
    auto it = live.find(raw);
    TMX_REQUIRE(it != live.end(), "unknown or already destroyed handle");
    p = std::move(it->second);
    live.erase(it);
  }
};
// Heap singletons, never destroyed (objects they still hold must not free device memory
// during static destruction).
static HandleRegistry<tmx_contraction_s>& contractions() {
  static auto* r = new HandleRegistry<tmx_contraction_s>();
  return *r;
}

// A graph's hold on a contraction, counted on the contraction, so that changes a replay must
// not see (which operands are constant, whether it assigns) are refused only while a graph
// holds it (a level's reference does not count).
struct GraphHold {
  std::shared_ptr<tmx_contraction_s> c;
  explicit GraphHold(std::shared_ptr<tmx_contraction_s> p) : c(std::move(p)) { ++c->graph_holds; }
  ~GraphHold() { --c->graph_holds; }
  GraphHold(const GraphHold&) = delete;
  GraphHold& operator=(const GraphHold&) = delete;
};
static void hold_for_graph(hipStream_t s, std::shared_ptr<tmx_contraction_s> c) {
  if (c && stream_capturing(s)) capture_hold(s, std::make_shared<GraphHold>(std::move(c)));
}

// Executes a contraction's sector plan with its operands in the plan's order.
// Executes a contraction's block-pair plan (GROUPED: one launch; BLOCKPAIR: one launch per
// block pair).  Assigning, the blocks without pairs are zeroed and each C tile is written by
// its first pair (GROUPED: its only task; BLOCKPAIR: the launch of the block's first pair).
static void run_bp(const tmx_contraction_s* c, const double* A, const double* B, double* C,
                   hipStream_t s) {
  if (c->assign)
    launch_zero_ranges(c->bp.zero_off.as<int64_t>(), c->bp.zero_prefix.as<int64_t>(), c->bp.nzero,
                       c->bp.zero_elems, C, s);
  if (c->mode == TMX_EXEC_GROUPED_F64) {
    launch_bp_plan(c->bp, c->bp.tasks.as<BPTask>(), c->bp.ntasks, &c->bp.xcd, A, B, C, c->alpha,
                   c->assign ? 1 : 0, s);
    return;
  }
  const BPTask* pt = c->bp.pair_tasks.as<BPTask>();
  for (size_t i = 0; i < c->bp.per_pair.size(); ++i) {
    const auto& r = c->bp.per_pair[i];
    launch_bp_plan(c->bp, pt + r.first, r.second - r.first, nullptr, A, B, C, c->alpha,
                   c->assign && c->bp.per_pair_first[i] ? 1 : 0, s);
  }
}

// pack_ev: optional two events around the main packing pass of the call's A operand.
static void run_sector(const tmx_contraction_s* c, SectorPlan& sp, const double* A, const double* B,
                       double* C, double alpha, hipStream_t s, hipEvent_t* ev = nullptr,
                       hipEvent_t* gemm_ev = nullptr, hipEvent_t* pack_ev = nullptr) {
  set_sector_constant(sp, c->swapped ? c->const_b : c->const_a, c->swapped ? c->const_a : c->const_b);
  set_sector_assign(sp, c->assign);
  if (c->swapped) std::swap(A, B);
  execute_sector_plan(sp, A, B, C, alpha, s, ev, gemm_ev, c->swapped ? nullptr : pack_ev,
                      c->swapped ? pack_ev : nullptr);
}

// Library-wide settings: set by the tmx_set_* calls, read when plans are built or executed,
// possibly from other threads.
static std::atomic<int> g_split_bits{35};
static std::atomic<int> g_parallel_split{0};  // 0: launch model

// Times a sector plan executing into a scratch C (average of 3 after one warm-up), in ms.
static float time_sector_plan(const tmx_contraction_s* c, SectorPlan& sp, const double* A,
                              const double* B, double* scratch, hipStream_t s, hipEvent_t e0,
                              hipEvent_t e1) {
  run_sector(c, sp, A, B, scratch, 1.0, s);
  TMX_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < 3; ++r) run_sector(c, sp, A, B, scratch, 1.0, s);
  TMX_HIP(hipEventRecord(e1, s));
  TMX_HIP(hipEventSynchronize(e1));
  float ms = 0;
  TMX_HIP(hipEventElapsedTime(&ms, e0, e1));
  return ms / 3;
}

// Split-K counts by measurement: the plan is rebuilt with uniform part counts 1, 2, 3 and 4
// (pairs whose reduction is too short keep fewer), each timed end to end (packing, GEMM,
// recombination, scatter) into a scratch copy of C, and the fastest kept if it beats the launch
// model's plan by 2%; a replacement then has its tile shapes tuned.  Plans whose execution takes
// under half a millisecond are left alone.  Returns 1 when the plan was replaced.
static int autotune_split_k(tmx_contraction_s* c, const double* A, const double* B, hipStream_t s) {
  hipEvent_t e0, e1;
  TMX_HIP(hipEventCreate(&e0));
  TMX_HIP(hipEventCreate(&e1));
  struct Cleanup {
    hipEvent_t a, b;
    ~Cleanup() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
  } cleanup{e0, e1};
  DevBuf scratch(sizeof(double) * (size_t)std::max<int64_t>(c->C.t->total, 1));
  launch_fill_bytes_zero(scratch.p, scratch.bytes, s);
  double* sc = scratch.as<double>();
  float best_ms = time_sector_plan(c, *c->sec, A, B, sc, s, e0, e1);
  if (best_ms < 0.5f) return 0;
  const bool split = c->mode == TMX_EXEC_SPLIT_I8;
  SectorPlanPtr best;
  int last_rounds = -1;
  for (int p = 1; p <= 4; ++p) {
    SectorPlanPtr cand = build_sector_plan(c->C, c->sa(), c->sb(), c->li, c->work, split, c->split_bits, p);
    int rounds = 0;
    sector_schedule_info(*cand, &rounds, nullptr, nullptr);
    if (rounds == last_rounds) continue;
    last_rounds = rounds;
    const float ms = time_sector_plan(c, *cand, A, B, sc, s, e0, e1);
    if (ms < best_ms * 0.98f) {
      best_ms = ms;
      best = std::move(cand);
    }
  }
  if (!best) return 0;
  autotune_sector_plan(*best, s);
  c->sec = std::move(best);
  return 1;
}

static bool sector_eligible(const LabelInfo& li) {
  return li.batch.empty() && li.ared.empty() && li.bred.empty() && !li.aouter.empty() &&
         !li.bouter.empty() && !li.inner.empty();
}

extern "C" {

int tmx_set_split_bits(int bits) {
  return guarded([&] {
    TMX_REQUIRE(bits >= 8 && bits <= 50, "split bits must be in [8, 50]");
    g_split_bits = bits;
  });
}

int tmx_set_parallel_split(int parts) {
  return guarded([&] {
    TMX_REQUIRE(parts >= 0 && parts <= 16, "parallel split must be in [0, 16]");
    g_parallel_split = parts;
  });
}

int tmx_set_scale_reuse(int on) {
  return guarded([&] {
    TMX_REQUIRE(on == 0 || on == 1, "scale reuse must be 0 or 1");
    set_scale_reuse(on != 0);
  });
}

int tmx_contraction_create(const tmx_operand* C, double alpha, const tmx_operand* A,
                           const tmx_operand* B, int exec_mode, tmx_contraction* out) {
  return guarded([&] {
    require_init();
    TMX_REQUIRE(out, "null output");
    TMX_REQUIRE(exec_mode >= 0 && exec_mode <= TMX_EXEC_AUTO_SPLIT, "unknown execution mode");
    auto c = std::make_shared<tmx_contraction_s>();
    c->alpha = alpha;
    c->C = make_loperand(C);
    c->A = make_loperand(A);
    c->B = make_loperand(B);
    TMX_REQUIRE(c->C.t.get() != c->A.t.get() && c->C.t.get() != c->B.t.get(),
                "output tensor also appears on the right-hand side (MultOp::validate)");
    LabelInfo li = analyse_labels(c->C, c->A, c->B);
    std::vector<CBlockWork> work = enumerate_pairs(c->C, c->A, c->B, li);
    int mode = exec_mode;
    if (mode == TMX_EXEC_AUTO || mode == TMX_EXEC_AUTO_SPLIT) {
      // Average reduction length per output element: below kMinSectorK the contraction is
      // memory-bound, and the sector path's operand repacking and dense intermediate cost more
      // than the fused block-pair kernel that reads the blocks in place.
      double fl = 0, celems = 0;
      for (const CBlockWork& w : work) {
        std::vector<int64_t> cd = block_dims(c->C, w.ct);
        double mn = 1;
        for (auto d : cd) mn *= (double)d;
        celems += mn;
        for (const PairRef& pr : w.pairs) {
          double k = 1;
          for (size_t i = 0; i < li.red.size(); ++i) k *= (double)li.tile_sizes.at(li.red[i])[pr.rt[i]];
          fl += 2.0 * mn * k;
        }
      }
      const double avg_k = celems > 0 ? fl / (2.0 * celems) : 0.0;
      constexpr double kMinSectorK = 128;
      const bool sector = sector_eligible(li) && avg_k >= kMinSectorK;
      mode = !sector ? TMX_EXEC_GROUPED_F64
                     : (exec_mode == TMX_EXEC_AUTO_SPLIT ? TMX_EXEC_SPLIT_I8 : TMX_EXEC_SECTOR_F64);
    }
    if (mode == TMX_EXEC_SECTOR_F64 || mode == TMX_EXEC_SPLIT_I8) {
      if (!sector_eligible(li))
        throw Error(TMX_ERR_UNSUPPORTED,
                    "sector execution needs aouter, bouter and inner labels and no batch or "
                    "reduce-only labels; use TMX_EXEC_GROUPED_F64");
      double fl = 0;
      int64_t np = 0;
      for (const CBlockWork& w : work) np += (int64_t)w.pairs.size();
      std::vector<int> ab(li.aouter), ba(li.bouter);
      ab.insert(ab.end(), li.bouter.begin(), li.bouter.end());
      ba.insert(ba.end(), li.aouter.begin(), li.aouter.end());
      c->swapped = c->C.labels != ab && c->C.labels == ba;
      c->li = c->swapped ? analyse_labels(c->C, c->B, c->A) : li;
      c->work = c->swapped ? enumerate_pairs(c->C, c->B, c->A, c->li) : work;
      // One read of each setting: the plan and the fields the autotuner rebuilds it from agree.
      c->split_bits = g_split_bits.load();
      c->parallel_split = g_parallel_split.load();
      // AUTO_SPLIT: split precision pays only where the FP64 GEMM is MFMA-bound.  A sector GEMM
      // reads (mrows + nrows) K operand elements for 2 mrows nrows K flops; below
      // kMinSplitMN (flops-weighted mrows nrows / (mrows + nrows) of the sectors, e.g. the one-O-
      // tile classes of Fo(k,i) += U(k,l,c,d) T(c,d,i,l)) both precisions stream their packed
      // operands at the same HBM rate and split packs 11 bytes per element against 8, so the
      // planner builds the plan in FP64 (before allocating anything).
      constexpr double kMinSplitMN = 32;
      c->sec = build_sector_plan(c->C, c->sa(), c->sb(), c->li, c->work, mode == TMX_EXEC_SPLIT_I8,
                                 c->split_bits, c->parallel_split,
                                 exec_mode == TMX_EXEC_AUTO_SPLIT ? kMinSplitMN : 0.0);
      if (mode == TMX_EXEC_SPLIT_I8 && !sector_is_split(*c->sec)) mode = TMX_EXEC_SECTOR_F64;
      // algorithmic flops: sum over triples of 2*M*N*K
      for (const CBlockWork& w : work) {
        std::vector<int64_t> rt0(li.red.size(), 0);
        std::vector<int64_t> cd = block_dims(c->C, w.ct);
        double mn = 1;
        for (auto d : cd) mn *= (double)d;
        for (const PairRef& pr : w.pairs) {
          double k = 1;
          for (size_t i = 0; i < li.red.size(); ++i) k *= (double)li.tile_sizes.at(li.red[i])[pr.rt[i]];
          fl += 2.0 * mn * k;
        }
      }
      c->flops = fl;
      c->npairs = np;
    } else {
      build_bp_plan(c->bp, c->C, c->A, c->B, li, work, mode == TMX_EXEC_BLOCKPAIR_F64, &c->flops,
                    &c->npairs);
    }
    c->mode = mode;
    *out = contractions().add(std::move(c));
  });
}

int tmx_contraction_execute(tmx_contraction c, const double* A, const double* B, double* C,
                            void* stream) {
  return guarded([&] {
    TMX_REQUIRE(c, "null contraction");
    hipStream_t s = pick_stream(stream);
    TMX_REQUIRE(A && B && C, "null data pointer");
    hold_for_graph(s, contractions().get(c));
    if (c->mode == TMX_EXEC_GROUPED_F64 || c->mode == TMX_EXEC_BLOCKPAIR_F64) run_bp(c, A, B, C, s);
    else run_sector(c, *c->sec, A, B, C, c->alpha, s);
  });
}

int tmx_contraction_autotune(tmx_contraction c, const double* A, const double* B, double* C,
                             void* stream, int* changed) {
  return guarded([&] {
    TMX_REQUIRE(c, "null contraction");
    TMX_REQUIRE(A && B && C, "null data pointer");
    hipStream_t s = pick_stream(stream);
    // A graph or a level (whose terms scatter from each other's products) holds the plan.
    if (contractions().held(c) > 0)
      throw Error(TMX_ERR_STATE,
                  "the contraction is held by a captured graph or a level (its plan would be "
                  "rebuilt under them): destroy them before autotuning");
    int n = 0;
    if (c->mode == TMX_EXEC_SECTOR_F64 || c->mode == TMX_EXEC_SPLIT_I8) {
      run_sector(c, *c->sec, A, B, C, c->alpha, s);
      n = autotune_sector_plan(*c->sec, s);
      if (c->parallel_split == 0) n += autotune_split_k(c, A, B, s);
    } else {
      const int rc = tmx_contraction_execute(c, A, B, C, stream);
      if (rc != TMX_OK) throw Error(rc, tmx_last_error());
    }
    TMX_HIP(hipStreamSynchronize(s));
    if (changed) *changed = n;
  });
}

// Executes once with events between the stages; ms[0] / ms[1] = packing of the call's A / B
// operand, ms[2] = GEMM, ms[3] = scatter (fused block-pair modes: their kernel as ms[2]),
// ms[4] = the longest single GEMM launch, ms[5] = the main packing pass of the call's A operand
// alone (0 when nothing was packed).
static void execute_stages(tmx_contraction c, const double* A, const double* B, double* C,
                           hipStream_t s, float ms[6]) {
  constexpr int kMaxLaunchEvents = 16, kEvents = 7 + kMaxLaunchEvents;
  hipEvent_t ev[kEvents];
  for (auto& e : ev) TMX_HIP(hipEventCreate(&e));
  struct Cleanup {
    hipEvent_t* ev;
    ~Cleanup() { for (int i = 0; i < kEvents; ++i) (void)hipEventDestroy(ev[i]); }
  } cleanup{ev};
  hipEvent_t* pack_ev = ev + 5 + kMaxLaunchEvents;
  // both recorded here first: a pass that does not run leaves them equal
  TMX_HIP(hipEventRecord(pack_ev[0], s));
  TMX_HIP(hipEventRecord(pack_ev[1], s));
  int launches = 1;
  if (c->mode == TMX_EXEC_SECTOR_F64 || c->mode == TMX_EXEC_SPLIT_I8) {
    int64_t t = 0;
    int r = 0;
    sector_schedule_info(*c->sec, &r, &launches, &t);
    const bool per_launch = launches > 1 && launches <= kMaxLaunchEvents + 1;
    run_sector(c, *c->sec, A, B, C, c->alpha, s, ev, per_launch ? ev + 5 : nullptr, pack_ev);
    if (!per_launch) launches = 1;
  } else {
    TMX_HIP(hipEventRecord(ev[0], s));
    TMX_HIP(hipEventRecord(ev[1], s));
    TMX_HIP(hipEventRecord(ev[2], s));
    run_bp(c, A, B, C, s);
    TMX_HIP(hipEventRecord(ev[3], s));
    TMX_HIP(hipEventRecord(ev[4], s));
  }
  TMX_HIP(hipEventSynchronize(ev[4]));
  for (int i = 0; i < 4; ++i) TMX_HIP(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
  if (c->swapped) std::swap(ms[0], ms[1]);  // the plan packed B first
  ms[4] = ms[2];
  if (launches > 1) {  // launch i runs from event (i - 1) (ev[2] for the first) to event i
    ms[4] = 0;
    for (int i = 0; i < launches; ++i) {
      hipEvent_t a = i == 0 ? ev[2] : ev[5 + i - 1], b = i + 1 == launches ? ev[3] : ev[5 + i];
      float x = 0;
      TMX_HIP(hipEventElapsedTime(&x, a, b));
      ms[4] = std::max(ms[4], x);
    }
  }
  TMX_HIP(hipEventElapsedTime(&ms[5], pack_ev[0], pack_ev[1]));
}

int tmx_contraction_execute_timed(tmx_contraction c, const double* A, const double* B, double* C,
                                  void* stream, float* stage_ms) {
  return guarded([&] {
    TMX_REQUIRE(c && stage_ms, "null argument");
    TMX_REQUIRE(A && B && C, "null data pointer");
    float ms[6];
    execute_stages(c, A, B, C, pick_stream(stream), ms);
    stage_ms[0] = ms[0] + ms[1];
    stage_ms[1] = ms[2];
    stage_ms[2] = ms[3];
  });
}

int tmx_contraction_execute_stages(tmx_contraction c, const double* A, const double* B, double* C,
                                   void* stream, float* stage_ms) {
  return tmx_contraction_execute_stages_n(c, A, B, C, stream, stage_ms, 5);
}

int tmx_contraction_execute_stages_n(tmx_contraction c, const double* A, const double* B, double* C,
                                     void* stream, float* stage_ms, int n) {
  return guarded([&] {
    TMX_REQUIRE(c && stage_ms, "null argument");
    TMX_REQUIRE(A && B && C, "null data pointer");
    TMX_REQUIRE(n >= 1 && n <= 6, "stage count must be in [1, 6]");
    float ms[6];
    execute_stages(c, A, B, C, pick_stream(stream), ms);
    for (int i = 0; i < n; ++i) stage_ms[i] = ms[i];
  });
}

int tmx_contraction_info(tmx_contraction c, double* flops, int64_t* npairs, int64_t* ws, int* mode) {
  return guarded([&] {
    TMX_REQUIRE(c, "null contraction");
    if (flops) *flops = c->flops;
    if (npairs) *npairs = c->npairs;
    if (ws) {
      int64_t b = (int64_t)(c->bp.tasks.bytes + c->bp.pairs.bytes + c->bp.cshapes.bytes +
                            c->bp.kshapes.bytes + c->bp.pair_tasks.bytes);
      if (c->sec) b += sector_workspace_bytes(*c->sec);
      *ws = b;
    }
    if (mode) *mode = c->mode;
  });
}

int tmx_contraction_schedule_info(tmx_contraction c, int* rounds, int* launches, int64_t* tasks) {
  return guarded([&] {
    TMX_REQUIRE(c, "null contraction");
    if (!c->sec) {
      if (rounds) *rounds = 1;
      if (launches) *launches = 1;
      if (tasks) *tasks = 0;
      return;
    }
    sector_schedule_info(*c->sec, rounds, launches, tasks);
  });
}

int tmx_contraction_split_info(tmx_contraction c, int* bits_a, int* bits_b, int* nmoduli) {
  return guarded([&] {
    TMX_REQUIRE(c, "null contraction");
    if (c->mode != TMX_EXEC_SPLIT_I8 || !c->sec) {
      if (bits_a) *bits_a = 53;
      if (bits_b) *bits_b = 53;
      if (nmoduli) *nmoduli = 0;
      return;
    }
    sector_split_info(*c->sec, bits_a, bits_b, nmoduli);
  });
}

int tmx_contraction_set_constant(tmx_contraction c, int a_constant, int b_constant) {
  return guarded([&] {
    TMX_REQUIRE(c, "null contraction");
    TMX_REQUIRE((a_constant == 0 || a_constant == 1) && (b_constant == 0 || b_constant == 1),
                "constant flags must be 0 or 1");
    if (c->graph_holds > 0)
      throw Error(TMX_ERR_STATE,
                  "the contraction is held by a captured graph: destroy the graph before changing "
                  "which operands are constant");
    c->const_a = a_constant != 0;
    c->const_b = b_constant != 0;
    if (c->sec) forget_sector_packing(*c->sec);  // the next execution packs afresh
  });
}

int tmx_contraction_set_assign(tmx_contraction c, int assign) {
  return guarded([&] {
    TMX_REQUIRE(c, "null contraction");
    TMX_REQUIRE(assign == 0 || assign == 1, "assign must be 0 or 1");
    if (c->graph_holds > 0)
      throw Error(TMX_ERR_STATE,
                  "the contraction is held by a captured graph: destroy the graph before changing "
                  "whether it assigns");
    c->assign = assign != 0;
    if (c->sec) set_sector_assign(*c->sec, c->assign);  // also for a level's scatter-only terms
  });
}

// ---------------------------------------------------------------------------
// Level: contractions executed in order as one unit (a Scheduler level), where terms that
// compute the same product share it.  A term's product is the packed-operand GEMM its sector
// plan runs; two plans have the same product when they pack the same tensors with the same
// jobs and have the same sector geometry (product_key), e.g. the four permutations of a CC
// ring term, R(a,b,i,j) += W(k,b,c,j) T(a,c,i,k) and its P(ab), P(ij), P(ab)P(ij) images,
// which all pack W and T alike and differ only in where (and with which sign) the product
// lands in R.  Such a term runs only its scatter stage, from the first term's product.
struct tmx_level_s {
  std::vector<std::shared_ptr<tmx_contraction_s>> terms;
  std::vector<int> leader;  // term whose product term i reuses; i itself when it computes its own
};

static bool sector_mode(const tmx_contraction_s& c) {
  return c.mode == TMX_EXEC_SECTOR_F64 || c.mode == TMX_EXEC_SPLIT_I8;
}

int tmx_level_create(const tmx_contraction* cs, int n, tmx_level* out) {
  return guarded([&] {
    TMX_REQUIRE(out && (cs || n == 0) && n >= 0, "bad level arguments");
    auto lv = std::make_unique<tmx_level_s>();
    for (int i = 0; i < n; ++i) {
      auto c = contractions().get(cs[i]);
      TMX_REQUIRE(c, "unknown contraction in level");
      int lead = i;
      if (sector_mode(*c) && c->sec) {
        for (int j = 0; j < i; ++j) {
          const tmx_contraction_s& d = *lv->terms[j];
          if (lv->leader[j] == j && sector_mode(d) && d.mode == c->mode &&
              sector_product_key(*d.sec) == sector_product_key(*c->sec)) {
            lead = j;
            break;
          }
        }
      }
      lv->terms.push_back(std::move(c));
      lv->leader.push_back(lead);
    }
    *out = lv.release();
  });
}

int tmx_level_products(tmx_level lv, int* nproducts) {
  return guarded([&] {
    TMX_REQUIRE(lv && nproducts, "null argument");
    int k = 0;
    for (size_t i = 0; i < lv->leader.size(); ++i) k += lv->leader[i] == (int)i;
    *nproducts = k;
  });
}

int tmx_level_execute(tmx_level lv, const double* const* A, const double* const* B, double* const* C,
                      void* stream) {
  return guarded([&] {
    TMX_REQUIRE(lv && A && B && C, "null argument");
    hipStream_t s = pick_stream(stream);
    const size_t n = lv->terms.size();
    for (size_t i = 0; i < n; ++i) TMX_REQUIRE(A[i] && B[i] && C[i], "null data pointer");
    for (size_t i = 0; i < n; ++i) {
      tmx_contraction_s* c = lv->terms[i].get();
      const size_t l = (size_t)lv->leader[i];
      // Reuse only when the leader saw the same operand data and no term since the leader
      // wrote into it (a level has no such dependence, but the call is checked).
      bool reuse = l != i && A[i] == A[l] && B[i] == B[l];
      for (size_t j = l; reuse && j < i; ++j) reuse = C[j] != A[l] && C[j] != B[l];
      if (reuse) {
        if (stream_capturing(s)) {
          hold_for_graph(s, lv->terms[i]);
          hold_for_graph(s, lv->terms[l]);
        }
        execute_sector_scatter_from(*c->sec, *lv->terms[l]->sec, C[i], c->alpha, s);
      } else {
        const int rc = tmx_contraction_execute(c, A[i], B[i], C[i], stream);
        if (rc != TMX_OK) throw Error(rc, tmx_last_error());
      }
    }
  });
}

int tmx_level_destroy(tmx_level lv) {
  return guarded([&] {
    if (!lv) return;
    quiesce();
    delete lv;
  });
}

int tmx_contraction_destroy(tmx_contraction c) {
  return guarded([&] {
    if (!c) return;
    quiesce();  // the plan's workspace may still be in use by queued kernels
    contractions().drop(c);  // a graph that captured it keeps it until tmx_graph_destroy
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Block-level entry points.  Plans are cached per (dims, labels) signature so a call
// allocates nothing once its shape has been seen.
namespace {

struct BlockMulPlan {
  BPPlan bp;
};
struct AssignPlan {
  DevBuf jobs, prefix, shapes;
  int njobs = 0;
  int64_t ntiles = 0;
};
// Heap singletons, never destroyed (their device buffers must not be freed during static
// destruction, see runtime.cpp).
std::mutex& cache_mu() {
  static std::mutex* m = new std::mutex();
  return *m;
}
std::map<std::vector<int64_t>, std::unique_ptr<BlockMulPlan>>& bm_cache() {
  static auto* c = new std::map<std::vector<int64_t>, std::unique_ptr<BlockMulPlan>>();
  return *c;
}
std::map<std::vector<int64_t>, std::unique_ptr<AssignPlan>>& as_cache() {
  static auto* c = new std::map<std::vector<int64_t>, std::unique_ptr<AssignPlan>>();
  return *c;
}

// Builds a one-block "tensor" whose single block has the given dims (one tile per mode).
LOperand single_block_operand(int n, const int64_t* dims, const int* labels) {
  std::vector<std::shared_ptr<const Space>> modes;
  for (int i = 0; i < n; ++i) {
    TMX_REQUIRE(dims[i] > 0, "block dims must be positive");
    int64_t t = dims[i];
    modes.push_back(make_space(dims[i], 0, &t, 1, 0, nullptr, nullptr, nullptr, 0, nullptr, nullptr,
                               nullptr, 0, nullptr, nullptr));
  }
  LOperand o;
  o.t = make_tensor(modes, -1, 0, 0);
  for (int i = 0; i < n; ++i) {
    o.labels.push_back(labels[i]);
    o.toff.push_back(0);
    o.ntl.push_back(1);
  }
  return o;
}

}  // namespace

namespace tmx {

// Assign plan between two operands whose labels are a permutation of each other.
std::unique_ptr<AssignPlan> build_assign_plan(const LOperand& D, const LOperand& S,
                                              std::vector<std::pair<int64_t, int64_t>>* zero_blocks,
                                              bool is_assign) {
  TMX_REQUIRE(D.labels.size() == S.labels.size(), "assign: label sets differ");
  for (int l : D.labels) TMX_REQUIRE(pos_of(S.labels, l) >= 0, "assign: labels are not a permutation");
  std::vector<AssignJob> jobs;
  std::vector<AssignShape> shapes;
  std::map<std::vector<int64_t>, int> sidx;
  std::vector<int64_t> prefix{0};
  const size_t r = D.labels.size();
  std::vector<int64_t> ext(r);
  for (size_t i = 0; i < r; ++i) ext[i] = D.ntl[i];
  std::vector<int64_t> it(r, 0), db(r), sb(r);
  bool empty = false;
  for (auto e : ext) if (e == 0) empty = true;
  while (!empty) {
    for (size_t i = 0; i < r; ++i) db[i] = D.toff[i] + it[i];
    if (D.t->is_non_zero(db.data())) {
      for (size_t j = 0; j < r; ++j) sb[j] = S.toff[j] + it[pos_of(D.labels, S.labels[j])];
      int64_t doff = D.t->find(db.data());
      int64_t soff = S.t->is_non_zero(sb.data()) ? S.t->find(sb.data()) : -1;
      std::vector<int64_t> dd(r), sd(r);
      for (size_t i = 0; i < r; ++i) dd[i] = D.t->modes[i]->tsize(db[i]);
      for (size_t i = 0; i < r; ++i) sd[i] = S.t->modes[i]->tsize(sb[i]);
      if (soff < 0) {
        if (is_assign && zero_blocks) {
          int64_t sz = 1;
          for (auto x : dd) sz *= x;
          zero_blocks->push_back({doff, sz});
        }
      } else {
        std::vector<int64_t> ds = strides_of(dd), ss = strides_of(sd);
        AssignShape sh;
        std::memset(&sh, 0, sizeof(sh));
        int ld = r ? D.labels.back() : 0, ls = r ? S.labels.back() : 0;
        std::vector<int64_t> oe, os, od;
        int64_t ed = r ? dd.back() : 1, es = r ? sd.back() : 1;
        sh.same = (r == 0 || ld == ls) ? 1 : 0;
        for (size_t i = 0; i < r; ++i) {
          int l = D.labels[i];
          if (l == ld) continue;
          if (!sh.same && l == ls) continue;
          oe.push_back(dd[i]);
          od.push_back(ds[i]);
          os.push_back(ss[pos_of(S.labels, l)]);
        }
        sh.ed = (int)ed;
        sh.es = (int)es;
        int64_t to = 1;
        for (auto x : oe) to *= x;
        sh.tiles_o = (int)to;
        if (sh.same) {
          sh.tiles_d = (int)((ed + 1023) / 1024);
          sh.tiles_s = 1;
        } else {
          sh.tiles_d = (int)((ed + 31) / 32);
          sh.tiles_s = (int)((es + 31) / 32);
          sh.src_stride_d = ss[pos_of(S.labels, ld)];
          sh.dst_stride_s = ds[pos_of(D.labels, ls)];
        }
        sh.src_outer = make_radix(oe, os);
        sh.dst_outer = make_radix(oe, od);
        AssignJob jb;
        jb.src_off = soff;
        jb.dst_off = doff;
        jb.shape = intern(shapes, sidx, sh);
        jb.pad = 0;
        jobs.push_back(jb);
        prefix.push_back(prefix.back() + (int64_t)sh.tiles_o * sh.tiles_d * sh.tiles_s);
      }
    }
    int d = (int)r - 1;
    for (; d >= 0; --d) {
      if (++it[d] < ext[d]) break;
      it[d] = 0;
    }
    if (d < 0) break;
  }
  auto p = std::make_unique<AssignPlan>();
  p->njobs = (int)jobs.size();
  p->ntiles = prefix.back();
  p->jobs = upload(jobs);
  p->prefix = upload(prefix);
  p->shapes = upload(shapes);
  return p;
}

}  // namespace tmx

struct tmx_addop_s {
  LOperand C, A;
  double alpha = 1;
  int is_assign = 0;
  std::unique_ptr<AssignPlan> plan;
  std::vector<std::pair<int64_t, int64_t>> zero_blocks;
  bool whole = false;  // identical labels and layouts over whole tensors: one streaming pass
};

// C(l) (=|+=) alpha * A(l) with the same label list, both operands spanning their whole
// tensors, and the two tensors laid out identically (same blocks, same offsets): no
// permutation and no block bookkeeping, so the op is one element-wise pass over the buffer.
static bool same_layout_whole(const LOperand& C, const LOperand& A) {
  if (C.labels != A.labels || C.t->nmodes() != A.t->nmodes()) return false;
  for (int i = 0; i < C.t->nmodes(); ++i) {
    if (C.toff[i] != 0 || A.toff[i] != 0) return false;
    if (C.ntl[i] != C.t->modes[i]->ntiles() || A.ntl[i] != A.t->modes[i]->ntiles()) return false;
    if (C.t->modes[i]->ntiles() != A.t->modes[i]->ntiles()) return false;
    for (int64_t k = 0; k < C.t->modes[i]->ntiles(); ++k)
      if (C.t->modes[i]->tsize(k) != A.t->modes[i]->tsize(k)) return false;
  }
  return C.t->total == A.t->total && C.t->keys == A.t->keys && C.t->offs == A.t->offs;
}

extern "C" {

int tmx_block_multiply(double alpha, const double* A, int andim, const int64_t* adims,
                       const int* alabels, const double* B, int bndim, const int64_t* bdims,
                       const int* blabels, double* C, int cndim, const int64_t* cdims,
                       const int* clabels, int is_assign, void* stream) {
  return guarded([&] {
    hipStream_t s = pick_stream(stream);
    TMX_REQUIRE(A && B && C, "null block pointer");
    TMX_REQUIRE(andim >= 0 && bndim >= 0 && cndim >= 0, "negative rank");
    std::vector<int64_t> key{andim, bndim, cndim};
    for (int i = 0; i < andim; ++i) { key.push_back(adims[i]); key.push_back(alabels[i]); }
    for (int i = 0; i < bndim; ++i) { key.push_back(bdims[i]); key.push_back(blabels[i]); }
    for (int i = 0; i < cndim; ++i) { key.push_back(cdims[i]); key.push_back(clabels[i]); }
    BlockMulPlan* plan;
    {
      std::lock_guard<std::mutex> g(cache_mu());
      auto& g_bm_cache = bm_cache();
      auto it = g_bm_cache.find(key);
      if (it == g_bm_cache.end()) {
        LOperand Co = single_block_operand(cndim, cdims, clabels);
        LOperand Ao = single_block_operand(andim, adims, alabels);
        LOperand Bo = single_block_operand(bndim, bdims, blabels);
        LabelInfo li = analyse_labels(Co, Ao, Bo);
        std::vector<CBlockWork> work = enumerate_pairs(Co, Ao, Bo, li);
        auto p = std::make_unique<BlockMulPlan>();
        build_bp_plan(p->bp, Co, Ao, Bo, li, work, false, nullptr, nullptr);
        it = g_bm_cache.emplace(key, std::move(p)).first;
      }
      plan = it->second.get();
    }
    launch_bp_plan(plan->bp, plan->bp.tasks.as<BPTask>(), plan->bp.ntasks, &plan->bp.xcd, A, B, C,
                   alpha, is_assign, s);
  });
}

int tmx_block_assign(double* dst, int ndim, const int64_t* ddims, const int* dlabels, double scale,
                     const double* src, const int64_t* sdims, const int* slabels, int is_assign,
                     void* stream) {
  return guarded([&] {
    hipStream_t s = pick_stream(stream);
    TMX_REQUIRE(dst && src, "null block pointer");
    std::vector<int64_t> key{ndim};
    for (int i = 0; i < ndim; ++i) { key.push_back(ddims[i]); key.push_back(dlabels[i]); }
    for (int i = 0; i < ndim; ++i) { key.push_back(sdims[i]); key.push_back(slabels[i]); }
    AssignPlan* plan;
    {
      std::lock_guard<std::mutex> g(cache_mu());
      auto& g_as_cache = as_cache();
      auto it = g_as_cache.find(key);
      if (it == g_as_cache.end()) {
        LOperand D = single_block_operand(ndim, ddims, dlabels);
        LOperand S = single_block_operand(ndim, sdims, slabels);
        for (int i = 0; i < ndim; ++i) {
          int p = pos_of(S.labels, D.labels[i]);
          TMX_REQUIRE(p >= 0 && sdims[p] == ddims[i], "assign: dims do not match the permutation");
        }
        it = g_as_cache.emplace(key, build_assign_plan(D, S, nullptr, false)).first;
      }
      plan = it->second.get();
    }
    launch_assign(plan->jobs.as<AssignJob>(), plan->njobs, plan->prefix.as<int64_t>(), plan->ntiles,
                  plan->shapes.as<AssignShape>(), src, dst, scale, is_assign, s);
  });
}

static HandleRegistry<tmx_addop_s>& addops() {
  static auto* r = new HandleRegistry<tmx_addop_s>();
  return *r;
}

int tmx_addop_create(const tmx_operand* C, double alpha, const tmx_operand* A, int is_assign,
                     tmx_addop* out) {
  return guarded([&] {
    require_init();
    auto op = std::make_shared<tmx_addop_s>();
    op->C = make_loperand(C);
    op->A = make_loperand(A);
    op->alpha = alpha;
    op->is_assign = is_assign;
    LabelInfo li = analyse_labels(op->C, op->A, op->A);  // validates tilings of shared labels
    (void)li;
    op->whole = same_layout_whole(op->C, op->A);
    if (!op->whole) op->plan = build_assign_plan(op->C, op->A, &op->zero_blocks, is_assign != 0);
    *out = addops().add(std::move(op));
  });
}

int tmx_addop_execute(tmx_addop op, const double* A, double* C, void* stream) {
  return guarded([&] {
    TMX_REQUIRE(op && A && C, "null argument");
    hipStream_t s = pick_stream(stream);
    if (stream_capturing(s)) capture_hold(s, addops().get(op));
    if (op->whole) {
      if (op->is_assign) launch_scale_copy(op->C.t->total, op->alpha, A, C, s);
      else launch_axpy(op->C.t->total, op->alpha, A, C, s);
      return;
    }
    for (auto& z : op->zero_blocks) launch_fill(C + z.first, z.second, 0.0, s);
    launch_assign(op->plan->jobs.as<AssignJob>(), op->plan->njobs, op->plan->prefix.as<int64_t>(),
                  op->plan->ntiles, op->plan->shapes.as<AssignShape>(), A, C, op->alpha,
                  op->is_assign, s);
  });
}

int tmx_addop_destroy(tmx_addop op) {
  return guarded([&] {
    if (!op) return;
    quiesce();
    addops().drop(op);
  });
}

}  // extern "C"
