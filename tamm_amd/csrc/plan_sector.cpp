// plan_sector.cpp -- sector plans: block-sparse contraction as a few dense GEMMs.
//
// A contraction C(M,N) += A(M,K) * B(N,K) over block-sparse tensors decomposes into
// sectors.  Row blocks of A (tuples of aouter tiles) with the same set of compatible K
// blocks form an A row class; likewise for B row blocks (bouter tiles).  For TAMM's spin
// rule (and the irrep extension) these classes are the conserved-quantity sectors, e.g. the
// ladder on spin orbitals has three: alpha-alpha, alpha-beta+beta-alpha, beta-beta.  Every
// (A class, B class) pair with a common K block is one dense GEMM whose K range is a short
// list of segments.  The plan gathers each operand's blocks into MFMA fragment-major
// class matrices (fusing the permutation into that gather), runs the dense GEMMs, and
// scatters the dense results back into C's blocks with C's own mode order.  The result
// equals TAMM's sum over nonzero block pairs (multop.hpp:576-893); only the grouping of
// the work differs.
//
// Split precision (TMX_EXEC_SPLIT_I8) replaces the FP64 GEMM by Ozaki scheme II (Ozaki,
// Uchino, Imamura 2025): each row of a class matrix is scaled so its largest element maps
// to 2^bits - 1 and rounded to an integer; the integers are reduced modulo pairwise coprime
// moduli m_l <= 256 into int8 residues; one int8 MFMA GEMM per modulus accumulates exact
// int32 sums; the residues of the product are recombined by the Chinese remainder theorem
// (a weighted sum of the residues held as exact FP64 limbs, reduced modulo prod(m_l)), then
// unscaled.  The product is exact for the rounded operands when
// prod(m_l) > 2 * K * 2^(bits_a + bits_b), so the only error is the operand rounding
// (~2^-bits relative per element).
#include <cmath>
#include <numeric>
#include <set>

#include "plan.hpp"
#include "sector.hpp"

namespace tmx {

namespace {

// Every K block is padded to a multiple of KGRAN columns (an int8 or FP64 MFMA fragment
// width); on the split path to KGRAN_SPLIT, so a GEMM stage of two 32-column k-steps never
// straddles two K segments.
constexpr int KGRAN = 32;
constexpr int KGRAN_SPLIT = 64;
// Longest split-K part of a split-precision sector: keeps the int32 residue sums exact
// (K * 128 * 128 < 2^31 needs K <= 131071) and the modulus count at 11 for 35-bit operands.
constexpr int64_t kSplitKPart = 32768;
// Class matrix rows are padded to a multiple of 32 on the split path (the int8 MFMA, CRT and
// fused-scatter tiles are 32x32) and of 16 on the FP64 path (its MFMA and fragments are 16
// rows): a class of 8 rows, e.g. one O tile, then streams 2x its bytes through the GEMM, not 4x.
constexpr int RGRAN_SPLIT = 32, RGRAN_F64 = 16;
// Shortest split-K part cut for parallelism (columns).
constexpr int64_t kMinPartK = 1024;
// Most split-K parts of one sector.
constexpr int64_t kMaxParts = 256;
// Cost model of an extra split-K part: rate of its scratch traffic.
constexpr double kScratchBytesPerSecond = 2.5e12;

int64_t roundup(int64_t x, int64_t g) { return (x + g - 1) / g * g; }

// FNV-1a over raw bytes (product fingerprints).
uint64_t fnv(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
  return h;
}

Radix radix_of(const std::vector<int64_t>& ext, const std::vector<int64_t>& str) {
  Radix r;
  std::memset(&r, 0, sizeof(r));
  TMX_REQUIRE(ext.size() <= (size_t)MAXR, "too many modes in one GEMM role");
  r.n = (int)ext.size();
  for (int i = 0; i < r.n; ++i) {
    r.ext[i] = (int)ext[i];
    r.stride[i] = (int)str[i];
  }
  return r;
}

// Enumerate tuples of tile indices for a label list.
std::vector<std::vector<int64_t>> tuples(const std::vector<int>& labels, const LabelInfo& li) {
  std::vector<int64_t> ext;
  for (int l : labels) ext.push_back((int64_t)li.tile_sizes.at(l).size());
  std::vector<std::vector<int64_t>> out;
  for (auto e : ext) if (e == 0) return out;
  std::vector<int64_t> t(labels.size(), 0);
  while (true) {
    out.push_back(t);
    int d = (int)t.size() - 1;
    for (; d >= 0; --d) {
      if (++t[d] < ext[d]) break;
      t[d] = 0;
    }
    if (d < 0) break;
  }
  return out;
}

int64_t tuple_size(const std::vector<int>& labels, const std::vector<int64_t>& t, const LabelInfo& li) {
  int64_t s = 1;
  for (size_t i = 0; i < labels.size(); ++i) s *= li.tile_sizes.at(labels[i])[t[i]];
  return s;
}

// Block id of operand `o` from the tiles of two label groups.
std::vector<int64_t> bid_of(const LOperand& o, const std::vector<int>& l1, const std::vector<int64_t>& t1,
                            const std::vector<int>& l2, const std::vector<int64_t>& t2) {
  std::vector<int64_t> b(o.labels.size());
  for (size_t i = 0; i < o.labels.size(); ++i) {
    int p = pos_of(l1, o.labels[i]);
    int64_t t = p >= 0 ? t1[p] : t2[pos_of(l2, o.labels[i])];
    b[i] = o.toff[i] + t;
  }
  return b;
}

// Row / column radices of one operand block: rows over `rl`, columns over `cl`.
void block_radix(const LOperand& o, const std::vector<int64_t>& bid, const std::vector<int>& rl,
                 const std::vector<int>& cl, Radix& rr, Radix& cr) {
  std::vector<int64_t> dims;
  for (size_t i = 0; i < o.labels.size(); ++i) dims.push_back(o.t->modes[i]->tsize(bid[i]));
  std::vector<int64_t> st = strides_of(dims);
  std::vector<int64_t> re, rs, ce, cs;
  for (int l : rl) { int p = pos_of(o.labels, l); re.push_back(dims[p]); rs.push_back(st[p]); }
  for (int l : cl) { int p = pos_of(o.labels, l); ce.push_back(dims[p]); cs.push_back(st[p]); }
  rr = radix_of(re, rs);
  cr = radix_of(ce, cs);
}

// Pairwise coprime moduli <= 256 (largest first); residues fit int8 symmetrically.
const int kModuli[kMaxModuli] = {256, 255, 253, 251, 247, 241, 239, 233,
                                 229, 227, 223, 217, 211, 199, 197, 193};

int inv_mod(int a, int m) {
  a %= m;
  if (a < 0) a += m;
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  throw Error(TMX_ERR_INVALID, "moduli are not coprime");
}

// Balanced base-2^42 limbs of a signed 128-bit integer: v = l2*2^84 + l1*2^42 + l0.
void limbs42(__int128 v, double out[3]) {
  const __int128 L = (__int128)1 << 42;
  for (int k = 0; k < 3; ++k) {
    __int128 d = v % L;
    if (d > L / 2) d -= L;
    else if (d < -L / 2) d += L;
    out[k] = (double)(long long)d;
    v = (v - d) / L;
  }
  TMX_REQUIRE(v == 0, "CRT constant exceeds three limbs");
}

}  // namespace

CrtConsts make_crt_consts(int n, const int* mod) {
  TMX_REQUIRE(n >= 1 && n <= kMaxModuli, "modulus count out of range");
  CrtConsts c;
  std::memset(&c, 0, sizeof(c));
  c.n = n;
  unsigned __int128 M = 1;
  for (int i = 0; i < n; ++i) M *= (unsigned)mod[i];
  for (int i = 0; i < n; ++i) {
    c.m[i] = mod[i];
    c.rcp[i] = 1.0 / (double)mod[i];
    const unsigned __int128 Mi = M / (unsigned)mod[i];
    const int inv = inv_mod((int)(Mi % (unsigned)mod[i]), mod[i]);
    const unsigned __int128 w = Mi * (unsigned)inv;  // < M
    const __int128 ws = w > M / 2 ? (__int128)w - (__int128)M : (__int128)w;
    limbs42(ws, c.w[i]);
  }
  limbs42((__int128)M, c.M);
  c.inv_m_total = 1.0 / (double)M;
  return c;
}

SectorPlanPtr build_sector_plan(const LOperand& C, const LOperand& A, const LOperand& B,
                                const LabelInfo& li, const std::vector<CBlockWork>& work,
                                bool split, int split_bits, int parallel_split,
                                double min_split_mn) {
  SectorPlanPtr sp(new SectorPlan());
  sp->split = split;
  const std::vector<int>& ML = li.aouter;
  const std::vector<int>& NL = li.bouter;
  // Order of the K labels (the class matrices' column order): that of an operand whose last,
  // stride-1 mode is a K label (the larger operand if both are), so its K columns come in
  // contiguous runs for the pack loads; else A's order.  E.g. for R(abij) += W(kbcj) T(acik),
  // K = (c,k) follows T; W's contiguous mode j is a row mode and is read along rows anyway.
  std::vector<int> KL = li.inner;
  {
    auto k_order = [&](const LOperand& o) {
      std::vector<int> v;
      for (int l : o.labels)
        if (pos_of(li.inner, l) >= 0) v.push_back(l);
      return v;
    };
    const bool a_k = !A.labels.empty() && pos_of(li.inner, A.labels.back()) >= 0;
    const bool b_k = !B.labels.empty() && pos_of(li.inner, B.labels.back()) >= 0;
    if (a_k && (!b_k || A.t->total >= B.t->total)) KL = k_order(A);
    else if (b_k) KL = k_order(B);
  }
  auto mt = tuples(ML, li), nt = tuples(NL, li), kt = tuples(KL, li);
  const size_t nK = kt.size();

  // Compatibility sets.
  auto sets = [&](const LOperand& o, const std::vector<int>& rl,
                  const std::vector<std::vector<int64_t>>& rt) {
    std::vector<std::vector<bool>> s(rt.size(), std::vector<bool>(nK, false));
    for (size_t r = 0; r < rt.size(); ++r)
      for (size_t k = 0; k < nK; ++k) {
        auto b = bid_of(o, rl, rt[r], KL, kt[k]);
        s[r][k] = o.t->is_non_zero(b.data());
      }
    return s;
  };
  auto SA = sets(A, ML, mt), SB = sets(B, NL, nt);
  auto classes = [](const std::vector<std::vector<bool>>& s, std::vector<int>& cls) {
    std::map<std::vector<bool>, int> idx;
    std::vector<std::vector<bool>> reps;
    cls.assign(s.size(), -1);
    for (size_t r = 0; r < s.size(); ++r) {
      bool any = false;
      for (bool x : s[r]) any |= x;
      if (!any) continue;
      auto it = idx.find(s[r]);
      if (it == idx.end()) { it = idx.emplace(s[r], (int)reps.size()).first; reps.push_back(s[r]); }
      cls[r] = it->second;
    }
    return reps;
  };
  std::vector<int> mcls, ncls;
  auto Arep = classes(SA, mcls), Brep = classes(SB, ncls);
  const int nMC = (int)Arep.size(), nNC = (int)Brep.size();

  // K classes: K blocks with the same (compatible A classes, compatible B classes).
  std::vector<std::pair<std::vector<bool>, std::vector<bool>>> ksig(nK);
  for (size_t k = 0; k < nK; ++k) {
    ksig[k].first.resize(nMC);
    ksig[k].second.resize(nNC);
    for (int c = 0; c < nMC; ++c) ksig[k].first[c] = Arep[c][k];
    for (int c = 0; c < nNC; ++c) ksig[k].second[c] = Brep[c][k];
  }
  std::map<std::pair<std::vector<bool>, std::vector<bool>>, int> kcidx;
  std::vector<int> kcls(nK);
  for (size_t k = 0; k < nK; ++k) {
    auto it = kcidx.find(ksig[k]);
    if (it == kcidx.end()) it = kcidx.emplace(ksig[k], (int)kcidx.size()).first;
    kcls[k] = it->second;
  }
  std::vector<size_t> korder(nK);
  std::iota(korder.begin(), korder.end(), 0);
  std::stable_sort(korder.begin(), korder.end(), [&](size_t a, size_t b) { return kcls[a] < kcls[b]; });
  // Geometry of the sector GEMMs before anything precision-specific is laid out: the
  // flops-weighted mrows * nrows / (mrows + nrows) over the (A class, B class) pairs, with real
  // rows and reduction lengths.  An FP64 GEMM reads (mrows + nrows) K elements for
  // 2 mrows nrows K flops, so below min_split_mn it streams its operands and split precision,
  // which packs 11 bytes per element against 8, cannot gain: the plan is then built in FP64.
  {
    std::vector<double> mr(nMC, 0.0), nr(nNC, 0.0), kl(nK, 0.0);
    for (size_t r = 0; r < mt.size(); ++r)
      if (mcls[r] >= 0) mr[mcls[r]] += (double)tuple_size(ML, mt[r], li);
    for (size_t r = 0; r < nt.size(); ++r)
      if (ncls[r] >= 0) nr[ncls[r]] += (double)tuple_size(NL, nt[r], li);
    for (size_t k = 0; k < nK; ++k) kl[k] = (double)tuple_size(KL, kt[k], li);
    double fl = 0, fl_mn = 0;
    for (int a = 0; a < nMC; ++a)
      for (int b = 0; b < nNC; ++b) {
        double klen = 0;
        for (size_t k = 0; k < nK; ++k)
          if (Arep[a][k] && Brep[b][k]) klen += kl[k];
        const double f = mr[a] * nr[b] * klen;
        fl += f;
        if (f > 0) fl_mn += f * (mr[a] * nr[b] / (mr[a] + nr[b]));
      }
    sp->mn_intensity = fl > 0 ? fl_mn / fl : 0.0;
  }
  if (split && min_split_mn > 0 && sp->mn_intensity < min_split_mn) {
    split = false;
    sp->split = false;
  }
  std::vector<int64_t> kpadlen(nK);
  const int64_t kgran = split ? KGRAN_SPLIT : KGRAN;
  for (size_t k = 0; k < nK; ++k) kpadlen[k] = roundup(tuple_size(KL, kt[k], li), kgran);

  // Class matrices (rows, padded columns, per-K-block column offsets).
  struct ClassMat {
    std::vector<int64_t> col;  // per K block, -1 when not compatible
    int64_t kpad = 0, rows = 0, rows_pad = 0, base = 0, growbase = 0;
  };
  auto build_side = [&](const std::vector<std::vector<bool>>& rep, const std::vector<int>& cls,
                        const std::vector<std::vector<int64_t>>& rt, const std::vector<int>& rl,
                        std::vector<ClassMat>& cm, std::vector<int64_t>& rowoff, int64_t& total,
                        int64_t& grows) {
    cm.assign(rep.size(), ClassMat());
    for (size_t c = 0; c < rep.size(); ++c) {
      cm[c].col.assign(nK, -1);
      int64_t off = 0;
      for (size_t k : korder)
        if (rep[c][k]) { cm[c].col[k] = off; off += kpadlen[k]; }
      cm[c].kpad = off;
    }
    rowoff.assign(rt.size(), -1);
    for (size_t r = 0; r < rt.size(); ++r) {
      if (cls[r] < 0) continue;
      ClassMat& m = cm[cls[r]];
      rowoff[r] = m.rows;
      m.rows += tuple_size(rl, rt[r], li);
    }
    total = 0;
    grows = 0;
    for (auto& m : cm) {
      m.rows_pad = roundup(std::max<int64_t>(m.rows, 1), split ? RGRAN_SPLIT : RGRAN_F64);
      m.base = total;
      m.growbase = grows;
      total += m.rows_pad * m.kpad;
      grows += m.rows_pad;
    }
  };
  std::vector<ClassMat> AM, BM;
  std::vector<int64_t> arow, brow;
  int64_t atotal, btotal, agrows, bgrows;
  build_side(Arep, mcls, mt, ML, AM, arow, atotal, agrows);
  build_side(Brep, ncls, nt, NL, BM, brow, btotal, bgrows);
  TMX_REQUIRE(agrows < (1ll << 31) && bgrows < (1ll << 31), "class matrices too large");

  // Pack jobs.
  uint64_t key = 0xcbf29ce484222325ull;  // product fingerprint (FNV-1a)
  auto pack_jobs = [&](const LOperand& o, const std::vector<int>& rl,
                       const std::vector<std::vector<int64_t>>& rt, const std::vector<int>& cls,
                       const std::vector<int64_t>& rowoff, const std::vector<ClassMat>& cm,
                       PackSide& side) {
    std::vector<PackJob> jobs;
    std::vector<TileUnit> units, rowunits, kfunits;
    const bool kcontig = !o.labels.empty() && pos_of(KL, o.labels.back()) >= 0;
    bool kfull = o.labels.size() >= KL.size();
    for (size_t i = 0; kfull && i < KL.size(); ++i)
      kfull = o.labels[o.labels.size() - KL.size() + i] == KL[i];
    for (size_t r = 0; r < rt.size(); ++r) {
      if (cls[r] < 0) continue;
      const ClassMat& m = cm[cls[r]];
      for (size_t k = 0; k < nK; ++k) {
        if (m.col[k] < 0) continue;
        auto b = bid_of(o, rl, rt[r], KL, kt[k]);
        int64_t off = o.t->find(b.data());
        if (off < 0) continue;
        PackJob j;
        std::memset(&j, 0, sizeof(j));
        j.src_off = off;
        j.dst_base = m.base;
        j.row0 = (int32_t)rowoff[r];
        j.col0 = (int32_t)m.col[k];
        j.rows = (int32_t)tuple_size(rl, rt[r], li);
        j.cols = (int32_t)tuple_size(KL, kt[k], li);
        j.kpad = (int32_t)m.kpad;
        j.growbase = (int32_t)m.growbase;
        j.kcontig = kcontig ? 1 : 0;
        j.kfull = kfull ? 1 : 0;
        block_radix(o, b, rl, KL, j.r, j.c);
        // Row maxima: k_rowmax streams whole rows (contiguous K) or 64 rows at a time along the
        // rows (row-contiguous blocks), for blocks with more than 32 rows; k_rowmax_t takes the
        // other blocks' pack tiles.  Blocks of more than 32 rows with contiguous K runs of at
        // least 128 columns pack through k_pack_kf's 32x128 tiles; all others through k_pack's
        // tiles of 4096 cells, 2^lt rows x 4096 / 2^lt columns (lt = 3..8), shaped to leave the
        // fewest padding cells (64x64 preferred on ties: its loads follow short digit runs).
        if (j.rows > 32 && (kfull || !kcontig))
          for (int32_t r0 = 0; r0 < j.rows; r0 += 64) rowunits.push_back({(int32_t)jobs.size(), r0, 0, 0});
        if (kfull && j.rows > 32 && j.cols >= 128) {
          for (int32_t r0 = 0; r0 < j.rows; r0 += 32)
            for (int32_t c0 = 0; c0 < j.cols; c0 += 128) kfunits.push_back({(int32_t)jobs.size(), r0, c0, 0});
        } else {
          int lt = 6;
          int64_t best = -1;
          for (int c : {6, 5, 7, 4, 8, 3}) {
            const int64_t cells = roundup(j.rows, (int64_t)1 << c) * roundup(j.cols, (int64_t)4096 >> c);
            if (best < 0 || cells < best) { best = cells; lt = c; }
          }
          const int32_t tr = 1 << lt, tc = 4096 >> lt;
          for (int32_t r0 = 0; r0 < j.rows; r0 += tr)
            for (int32_t c0 = 0; c0 < j.cols; c0 += tc) units.push_back({(int32_t)jobs.size(), r0, c0, lt});
        }
        jobs.push_back(j);
      }
    }
    side.njobs = (int)jobs.size();
    {
      const void* tid = o.t.get();
      key = fnv(key, &tid, sizeof(tid));
      if (!jobs.empty()) key = fnv(key, jobs.data(), jobs.size() * sizeof(PackJob));
    }
    side.nunits = (int64_t)units.size();
    side.nrowunits = (int64_t)rowunits.size();
    side.nkfunits = (int64_t)kfunits.size();
    side.jobs = upload(jobs);
    side.units = upload(units);
    side.rowunits = upload(rowunits);
    side.kfunits = upload(kfunits);
  };
  pack_jobs(A, ML, mt, mcls, arow, AM, sp->A);
  pack_jobs(B, NL, nt, ncls, brow, BM, sp->B);
  sp->A.total = atotal;
  sp->B.total = btotal;
  sp->A.nrows = agrows;
  sp->B.nrows = bgrows;

  // Sectors: (A class, B class) pairs with a common K block, each a dense GEMM whose K range is
  // a list of segments.  A sector is cut into split-K parts when its reduction is longer than
  // the split path's exact-accumulation budget, and into up to p parts (the parallel split
  // factor, chosen below) when that shortens the launch; every part has its own dense result
  // region (rows_pad(A) x rows_pad(B), a whole number of 32x32 tiles) and is scattered into C in
  // its own round.
  struct Sector { int mc, nc; int64_t c_base; int seg_b, seg_e; int64_t klen; int round; };
  struct Pair { int mc, nc; std::vector<KSeg> segs; int64_t klen; };
  std::vector<Pair> cpairs;
  for (int a = 0; a < nMC; ++a)
    for (int b = 0; b < nNC; ++b) {
      Pair pr{a, b, {}, 0};
      for (size_t k : korder) {
        if (AM[a].col[k] < 0 || BM[b].col[k] < 0) continue;
        KSeg g{(int32_t)AM[a].col[k], (int32_t)BM[b].col[k], (int32_t)kpadlen[k], 0};
        pr.klen += g.len;
        if (!pr.segs.empty()) {
          KSeg& last = pr.segs.back();
          if (last.a_col + last.len == g.a_col && last.b_col + last.len == g.b_col) {
            last.len += g.len;
            continue;
          }
        }
        pr.segs.push_back(g);
      }
      if (!pr.segs.empty()) cpairs.push_back(std::move(pr));
    }
  struct Parts {
    std::vector<Sector> secs;
    std::vector<KSeg> segs;
    int64_t cbase = 0, kmax = 0, extra_elems = 0;
    int nrounds = 0;
  };
  // Minimum part count of each pair (the split path's exact-accumulation budget) and the most
  // parts a pair may be cut into for parallelism.
  std::vector<int64_t> min_parts(cpairs.size()), max_parts(cpairs.size());
  for (size_t i = 0; i < cpairs.size(); ++i) {
    min_parts[i] = split ? (cpairs[i].klen + kSplitKPart - 1) / kSplitKPart : 1;
    max_parts[i] = std::max(min_parts[i], std::min<int64_t>(kMaxParts, cpairs[i].klen / kMinPartK));
  }
  auto make_parts = [&](const std::vector<int64_t>& nparts) {
    Parts P;
    for (size_t i = 0; i < cpairs.size(); ++i) {
      const Pair& pr = cpairs[i];
      const int64_t mr = AM[pr.mc].rows_pad, nr = BM[pr.nc].rows_pad;
      const int64_t parts = std::min(std::max(nparts[i], min_parts[i]), max_parts[i]);
      const int64_t target = roundup((pr.klen + parts - 1) / parts, kgran);
      int round = 0;
      Sector cur{pr.mc, pr.nc, 0, (int)P.segs.size(), 0, 0, 0};
      auto close = [&]() {
        cur.seg_e = (int)P.segs.size();
        cur.c_base = P.cbase;
        cur.round = round++;
        P.cbase += mr * nr;
        P.kmax = std::max(P.kmax, cur.klen);
        P.secs.push_back(cur);
        cur = Sector{pr.mc, pr.nc, 0, (int)P.segs.size(), 0, 0, 0};
      };
      for (const KSeg& g0 : pr.segs) {
        KSeg g = g0;
        while (g.len > 0) {
          const int32_t take = (int32_t)std::min<int64_t>(g.len, target - cur.klen);
          P.segs.push_back(KSeg{g.a_col, g.b_col, take, 0});
          cur.klen += take;
          g.a_col += take;
          g.b_col += take;
          g.len -= take;
          if (cur.klen >= target) close();
        }
      }
      if (cur.klen > 0) close();
      P.extra_elems += (int64_t)(round - 1) * mr * nr;
      P.nrounds = std::max(P.nrounds, round);
    }
    return P;
  };
  auto moduli_for = [&](int64_t kmax) {
    const double need = 2.0 * split_bits + std::log2((double)std::max<int64_t>(kmax, 1)) + 1.05;
    double have = 0;
    int n = 0;
    while (n < kMaxModuli && have < need) have += std::log2((double)kModuli[n++]);
    return have < need ? -1 : n;
  };
  auto make_geoms = [&](const Parts& P, int nmod) {
    std::vector<SectorGeom> geoms;
    for (const Sector& s2 : P.secs) {
      const ClassMat& am = AM[s2.mc];
      const ClassMat& bm = BM[s2.nc];
      SectorGeom g;
      std::memset(&g, 0, sizeof(g));
      g.a_base = am.base;
      g.b_base = bm.base;
      g.c_base = s2.c_base;
      g.seg_b = s2.seg_b;
      g.seg_e = s2.seg_e;
      g.a_kpad = (int32_t)am.kpad;
      g.b_kpad = (int32_t)bm.kpad;
      g.mrows_pad = (int32_t)am.rows_pad;
      g.nrows_pad = (int32_t)bm.rows_pad;
      g.mrows = (int32_t)am.rows;
      g.nrows = (int32_t)bm.rows;
      g.klen = s2.klen;
      choose_tile(split, nmod, g.mrows_pad, g.nrows_pad, g.klen, &g.tm, &g.tn);
      geoms.push_back(g);
    }
    return geoms;
  };
  // Parallel split: part counts per pair.  A fixed count applies to every pair; otherwise the
  // launch model decides -- the best uniform count first, then the pairs with the most work
  // (up to eight) doubled or stepped one by one while that helps -- charging each extra
  // part's partial result (written by the GEMM, read back by the scatter, and for the split
  // path reduced from its residue planes).  A change must gain 3% to be taken.
  auto estimate = [&](const std::vector<int64_t>& np) {
    Parts E = make_parts(np);
    const int nm = split ? moduli_for(E.kmax) : 0;
    if (split && nm < 0) return 1e30;
    const double bytes_per_elem = split ? 2.0 * nm + 16.0 : 16.0;
    return estimate_sector_seconds(split, nm, make_geoms(E, nm)) +
           (double)E.extra_elems * bytes_per_elem / kScratchBytesPerSecond;
  };
  std::vector<int64_t> nparts(cpairs.size(), 1);
  if (parallel_split > 0) {
    std::fill(nparts.begin(), nparts.end(), (int64_t)parallel_split);
  } else if (!cpairs.empty()) {
    double cur = estimate(nparts);
    std::vector<int64_t> best = nparts;
    for (int64_t pfac : {2, 3, 4, 6, 8, 16, 32, 64}) {
      std::vector<int64_t> trial(cpairs.size(), pfac);
      const double t = estimate(trial);
      if (t < cur * 0.97) { cur = t; best = trial; }
    }
    nparts = best;
    for (size_t i = 0; i < cpairs.size(); ++i)
      nparts[i] = std::min(std::max(nparts[i], min_parts[i]), max_parts[i]);
    std::vector<size_t> order(cpairs.size());
    std::iota(order.begin(), order.end(), (size_t)0);
    auto work = [&](size_t i) {
      return (double)AM[cpairs[i].mc].rows_pad * (double)BM[cpairs[i].nc].rows_pad * (double)cpairs[i].klen;
    };
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work(x) > work(y); });
    if (order.size() > 8) order.resize(8);
    for (int iter = 0; iter < 12; ++iter) {
      double best_t = cur;
      size_t best_i = cpairs.size();
      int64_t best_n = 0;
      for (size_t i : order)
        for (int step = 0; step < 2; ++step) {
          const int64_t n = step == 0 ? nparts[i] * 2 : nparts[i] + 1;
          if (n > max_parts[i] || (step == 1 && n == nparts[i] * 2)) continue;
          std::vector<int64_t> trial = nparts;
          trial[i] = n;
          const double t = estimate(trial);
          if (t < best_t) { best_t = t; best_i = i; best_n = n; }
        }
      if (best_i == cpairs.size() || best_t >= cur * 0.97) break;
      nparts[best_i] = best_n;
      cur = best_t;
    }
  }
  Parts P = make_parts(nparts);
  std::vector<Sector>& secs = P.secs;
  std::vector<KSeg>& segs = P.segs;
  const int64_t cbase = P.cbase, kmax = P.kmax;
  sp->c_total = cbase;
  sp->kmax = kmax;
  sp->split_factor = P.nrounds;
  sp->geoms = make_geoms(P, split ? std::max(moduli_for(kmax), 0) : 0);
  sp->ksegs = upload(segs);
  for (const SectorGeom& g : sp->geoms) {
    const int64_t f[] = {g.a_base, g.b_base, g.c_base, g.seg_b, g.seg_e, g.a_kpad, g.b_kpad,
                         g.mrows_pad, g.nrows_pad, g.mrows, g.nrows, g.klen};
    key = fnv(key, f, sizeof(f));
  }
  if (!segs.empty()) key = fnv(key, segs.data(), segs.size() * sizeof(KSeg));
  {
    const int64_t f[] = {split ? 1 : 0, split ? split_bits : 0, cbase, kmax, (int64_t)P.nrounds};
    key = fnv(key, f, sizeof(f));
  }
  sp->product_key = key;

  // Scatter jobs: dense sector results -> C blocks (C labels = aouter + bouter here), one
  // launch; a job sums the pair's split-K parts, whose dense results are consecutive regions.
  {
    std::map<std::pair<int, int>, std::pair<const Sector*, int>> parts_of;  // first part, count
    for (const Sector& s2 : secs) {
      auto it = parts_of.find({s2.mc, s2.nc});
      if (it == parts_of.end()) parts_of[{s2.mc, s2.nc}] = {&s2, 1};
      else ++it->second.second;
    }
    const bool ccontig = !C.labels.empty() && pos_of(NL, C.labels.back()) >= 0;
    std::vector<ScatterJob> jobs;
    std::vector<std::pair<size_t, size_t>> job_rc;  // (A tuple, B tuple) of each job
    std::vector<TileUnit> units;
    for (size_t r = 0; r < mt.size(); ++r) {
      if (mcls[r] < 0) continue;
      for (size_t c = 0; c < nt.size(); ++c) {
        if (ncls[c] < 0) continue;
        auto it = parts_of.find({mcls[r], ncls[c]});
        if (it == parts_of.end()) continue;
        auto cb = bid_of(C, ML, mt[r], NL, nt[c]);
        if (!C.t->is_non_zero(cb.data())) continue;
        ScatterJob j;
        std::memset(&j, 0, sizeof(j));
        j.c_off = C.t->find(cb.data());
        j.src_base = it->second.first->c_base;
        j.pstride = AM[mcls[r]].rows_pad * BM[ncls[c]].rows_pad;
        j.nparts = it->second.second;
        j.row0 = (int32_t)arow[r];
        j.col0 = (int32_t)brow[c];
        j.rows = (int32_t)tuple_size(ML, mt[r], li);
        j.cols = (int32_t)tuple_size(NL, nt[c], li);
        j.ldc = (int32_t)BM[ncls[c]].rows_pad;
        j.arow0 = (int32_t)(AM[mcls[r]].growbase + arow[r]);
        j.bcol0 = (int32_t)(BM[ncls[c]].growbase + brow[c]);
        j.ccontig = ccontig ? 1 : 0;
        block_radix(C, cb, ML, NL, j.r, j.c);
        for (int32_t r0 = 0; r0 < j.rows; r0 += 64)
          for (int32_t c0 = 0; c0 < j.cols; c0 += 64) units.push_back({(int32_t)jobs.size(), r0, c0, 0});
        jobs.push_back(j);
        job_rc.push_back({r, c});
      }
    }
    // Fused CRT + scatter tables (split path): every 32x32 tile of a pair's dense result with
    // valid rows and columns; row / column tuple maps and the (A tuple, B tuple) -> job table.
    // When C's labels are the A-outer labels then the B-outer labels, a block element sits at
    // row-within-tuple * block columns + column-within-tuple (layout 1).  Other orders (layouts
    // 2 and 3) add each row's and column's mixed-radix digits within its tuple, which needs at
    // most 4 modes per side and tile extents of at most 256.
    std::vector<int> mn(ML);
    mn.insert(mn.end(), NL.begin(), NL.end());
    auto fits_digits = [&](const std::vector<int>& labels) {
      if (labels.size() > 4) return false;
      for (int l : labels)
        for (int64_t e : li.tile_sizes.at(l))
          if (e > 256) return false;
      return true;
    };
    const int layout = C.labels == mn ? 1 : (fits_digits(ML) && fits_digits(NL)) ? (ccontig ? 2 : 3) : 0;
    if (split && layout > 0) {
      std::vector<int2> ainfo((size_t)agrows, int2{-1, 0}), binfo((size_t)bgrows, int2{-1, 0});
      for (size_t r = 0; r < mt.size(); ++r) {
        if (mcls[r] < 0) continue;
        const int64_t base = AM[mcls[r]].growbase + arow[r], n = tuple_size(ML, mt[r], li);
        for (int64_t i = 0; i < n; ++i) ainfo[(size_t)(base + i)] = int2{(int)r, (int)i};
      }
      for (size_t c = 0; c < nt.size(); ++c) {
        if (ncls[c] < 0) continue;
        const int64_t base = BM[ncls[c]].growbase + brow[c], n = tuple_size(NL, nt[c], li);
        for (int64_t i = 0; i < n; ++i) binfo[(size_t)(base + i)] = int2{(int)c, (int)i};
      }
      std::vector<int32_t> jt(mt.size() * nt.size(), -1);
      for (size_t jn = 0; jn < jobs.size(); ++jn) jt[job_rc[jn].first * nt.size() + job_rc[jn].second] = (int32_t)jn;
      // The job covering a whole 32x32 tile, or -1 when the tile straddles tuples or padding.
      auto tile_job = [&](int64_t ar0, int64_t br0, FusedTile& ft) {
        ft.job = -1;
        ft.rin = ft.cin = 0;
        if (ar0 + 32 > agrows || br0 + 32 > bgrows) return;
        const int2 a0 = ainfo[(size_t)ar0], a1 = ainfo[(size_t)(ar0 + 31)];
        const int2 b0 = binfo[(size_t)br0], b1 = binfo[(size_t)(br0 + 31)];
        if (a0.x < 0 || a0.x != a1.x || a1.y != a0.y + 31) return;
        if (b0.x < 0 || b0.x != b1.x || b1.y != b0.y + 31) return;
        ft.job = jt[(size_t)a0.x * nt.size() + (size_t)b0.x];
        ft.rin = a0.y;
        ft.cin = b0.y;
      };
      std::vector<CrtPair> pairs;
      std::vector<FusedTile> ftiles;
      for (const auto& kv : parts_of) {
        const int mc = kv.first.first, nc = kv.first.second;
        CrtPair cp;
        std::memset(&cp, 0, sizeof(cp));
        cp.c_base = kv.second.first->c_base;
        cp.pstride = AM[mc].rows_pad * BM[nc].rows_pad;
        cp.nparts = kv.second.second;
        cp.ntc = (int32_t)(BM[nc].rows_pad / 32);
        cp.agrow0 = (int32_t)AM[mc].growbase;
        cp.bgrow0 = (int32_t)BM[nc].growbase;
        const int32_t id = (int32_t)pairs.size();
        pairs.push_back(cp);
        for (int32_t tr = 0; tr * 32 < AM[mc].rows; ++tr)
          for (int32_t tc = 0; tc * 32 < BM[nc].rows; ++tc) {
            FusedTile ft;
            std::memset(&ft, 0, sizeof(ft));
            ft.pair = id;
            ft.tr = tr;
            ft.tc = tc;
            tile_job(AM[mc].growbase + 32 * tr, BM[nc].growbase + 32 * tc, ft);
            ftiles.push_back(ft);
          }
      }
      if (layout >= 2) {
        // digits of each row (column) within its tuple, the last mode fastest
        auto digits = [&](const std::vector<int>& labels, const std::vector<int64_t>& tup, int64_t i) {
          uint32_t packed = 0;
          for (int d = (int)labels.size() - 1; d >= 0; --d) {
            const int64_t e = li.tile_sizes.at(labels[d])[tup[d]];
            packed |= (uint32_t)(i % e) << (8 * d);
            i /= e;
          }
          return packed;
        };
        std::vector<uint32_t> adig((size_t)agrows, 0), bdig((size_t)bgrows, 0);
        for (size_t r = 0; r < mt.size(); ++r) {
          if (mcls[r] < 0) continue;
          const int64_t base = AM[mcls[r]].growbase + arow[r], n = tuple_size(ML, mt[r], li);
          for (int64_t i = 0; i < n; ++i) adig[(size_t)(base + i)] = digits(ML, mt[r], i);
        }
        for (size_t c = 0; c < nt.size(); ++c) {
          if (ncls[c] < 0) continue;
          const int64_t base = BM[ncls[c]].growbase + brow[c], n = tuple_size(NL, nt[c], li);
          for (int64_t i = 0; i < n; ++i) bdig[(size_t)(base + i)] = digits(NL, nt[c], i);
        }
        sp->arow_digits = upload(adig);
        sp->brow_digits = upload(bdig);
      }
      sp->fused_layout = layout;
      sp->nfused = (int64_t)ftiles.size();
      sp->fused_tiles = upload(ftiles);
      sp->crt_pairs = upload(pairs);
      sp->arow_info = upload(ainfo);
      sp->brow_info = upload(binfo);
      sp->jobtab = upload(jt);
      sp->jobtab_n = (int32_t)nt.size();
    }
    // The contraction's C blocks that no job writes (no compatible block pair), zeroed when the
    // plan assigns rather than accumulates.
    {
      std::set<int64_t> written;
      for (const ScatterJob& j : jobs) written.insert(j.c_off);
      std::vector<int64_t> zoff, zpre{0};
      for (const CBlockWork& w : work) {
        if (written.count(w.c_off)) continue;
        int64_t n = 1;
        for (int64_t d : block_dims(C, w.ct)) n *= d;
        if (n == 0) continue;
        zoff.push_back(w.c_off);
        zpre.push_back(zpre.back() + n);
      }
      sp->nzero = (int)zoff.size();
      sp->zero_elems = zpre.back();
      if (sp->nzero > 0) {
        sp->zero_off = upload(zoff);
        sp->zero_prefix = upload(zpre);
      }
    }
    SectorPlan::ScatterRound sr;
    sr.njobs = (int)jobs.size();
    sr.nunits = (int64_t)units.size();
    sr.jobs = upload(jobs);
    sr.units = upload(units);
    sp->scatter.push_back(std::move(sr));
  }

  // Precision parameters and buffers.
  if (split) {
    sp->bits_a = sp->bits_b = split_bits;
    TMX_REQUIRE(kmax <= 131071, "split precision: a sector's reduction length exceeds 131071 "
                                "(int32 accumulation bound); use FP64 execution");
    const int n = moduli_for(kmax);
    if (n < 0) throw Error(TMX_ERR_UNSUPPORTED, "split precision: too many bits requested for 16 moduli");
    sp->nmod = n;
    sp->crt = make_crt_consts(n, kModuli);
    sp->A.buf = DevBuf((size_t)atotal * n);
    sp->B.buf = DevBuf((size_t)btotal * n);
    sp->A.rowmax = DevBuf(sizeof(double) * agrows);
    sp->B.rowmax = DevBuf(sizeof(double) * bgrows);
    sp->A.invscale = DevBuf(2 * sizeof(double) * agrows);  // (f / top, e) per row
    sp->B.invscale = DevBuf(2 * sizeof(double) * bgrows);
    sp->A.rowmax_new = DevBuf(sizeof(double) * agrows);
    sp->B.rowmax_new = DevBuf(sizeof(double) * bgrows);
    sp->A.changed = DevBuf(16);
    sp->B.changed = DevBuf(16);
    sp->A.changed_host.alloc();
    sp->B.changed_host.alloc();
    sp->cres = DevBuf((size_t)cbase * n);
    // the fused CRT + scatter goes from the residue planes straight into C
    if (sp->nfused == 0) sp->cdense = DevBuf(sizeof(double) * (size_t)cbase);
    std::vector<CrtSector> cs;
    int64_t tiles = 0;
    for (const Sector& s2 : secs) {
      CrtSector x;
      std::memset(&x, 0, sizeof(x));
      x.c_base = s2.c_base;
      x.ntc = (int32_t)(BM[s2.nc].rows_pad / 32);
      x.ldc = (int32_t)BM[s2.nc].rows_pad;
      x.tiles = (AM[s2.mc].rows_pad / 32) * x.ntc;
      x.agrow0 = (int32_t)AM[s2.mc].growbase;
      x.bgrow0 = (int32_t)BM[s2.nc].growbase;
      tiles += x.tiles;
      cs.push_back(x);
    }
    TMX_REQUIRE(tiles * 1024 == cbase, "dense result is not a whole number of 32x32 tiles");
    sp->ncrt_secs = (int)cs.size();
    sp->crt_tiles = tiles;
    sp->crt_secs = upload(cs);
  } else {
    sp->A.buf = DevBuf(sizeof(double) * (size_t)atotal);
    sp->B.buf = DevBuf(sizeof(double) * (size_t)btotal);
    sp->cdense = DevBuf(sizeof(double) * (size_t)cbase);
  }
  build_sector_tasks(*sp);
  // Padding (rows beyond a class's rows, columns beyond a block's columns) must read as zero;
  // pack kernels only ever write real elements, so clear once here.
  hipStream_t s = runtime().stream;
  launch_fill_bytes_zero(sp->A.buf.p, sp->A.buf.bytes, s);
  launch_fill_bytes_zero(sp->B.buf.p, sp->B.buf.bytes, s);
  // Row scales of padding rows (never packed) read as zero in the CRT pass.
  if (sp->A.invscale.p) launch_fill_bytes_zero(sp->A.invscale.p, sp->A.invscale.bytes, s);
  if (sp->B.invscale.p) launch_fill_bytes_zero(sp->B.invscale.p, sp->B.invscale.bytes, s);
  // Recomputed row maxima accumulate by atomicMax from zero; each check clears them again.
  for (PackSide* side : {&sp->A, &sp->B})
    if (side->rowmax_new.p) {
      launch_fill_bytes_zero(side->rowmax_new.p, side->rowmax_new.bytes, s);
      side->fresh_clean = true;
    }
  TMX_HIP(hipStreamSynchronize(s));
  sp->ws = (int64_t)(sp->A.buf.bytes + sp->B.buf.bytes + sp->cdense.bytes + sp->cres.bytes + sp->A.jobs.bytes +
                     sp->B.jobs.bytes + sp->A.rowmax.bytes * 3 +
                     sp->B.rowmax.bytes * 3);
  return sp;
}

int64_t sector_workspace_bytes(const SectorPlan& sp) { return sp.ws; }
uint64_t sector_product_key(const SectorPlan& sp) { return sp.product_key; }
double sector_mn_intensity(const SectorPlan& sp) { return sp.mn_intensity; }
bool sector_is_split(const SectorPlan& sp) { return sp.split; }
void set_sector_assign(SectorPlan& sp, bool assign) { sp.assign = assign; }

void sector_split_info(const SectorPlan& sp, int* a, int* b, int* n) {
  if (a) *a = sp.bits_a;
  if (b) *b = sp.bits_b;
  if (n) *n = sp.nmod;
}

void destroy_sector_plan(SectorPlan* sp) { delete sp; }

}  // namespace tmx
