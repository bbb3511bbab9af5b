// sector.hpp -- sector (dense class-matrix) plan structures and kernel launchers.
#pragma once

#include "plan.hpp"

namespace tmx {

constexpr int kMaxModuli = 16;

// Moduli of the split-precision residue GEMMs and the constants of their CRT recombination.
// x = sum_l r_l * w_l (mod M) with M = prod m_l and w_l = (M/m_l) * ((M/m_l)^-1 mod m_l),
// reduced to (-M/2, M/2].  Each w_l and M is held as three balanced base-2^42 FP64 limbs
// (|limb| <= 2^41), so the limb sums over up to 16 residues |r_l| <= 128 stay below 2^52 and
// are exact in FP64; the quotient round(S/M) is recovered from a rounded FP64 evaluation.
struct CrtConsts {
  int n;
  int m[kMaxModuli];
  double rcp[kMaxModuli];   // 1/m_l (rounded)
  double w[kMaxModuli][3];  // limbs: w = w2*2^84 + w1*2^42 + w0
  double M[3];
  double inv_m_total;       // 1/M (rounded)
};

// Dense-result region of one sector, for the CRT pass over 32x32 tiles.
struct CrtSector {
  int64_t c_base;   // first element of the sector's dense result
  int64_t tiles;    // 32x32 tiles of the sector (mrows_pad/32 * nrows_pad/32)
  int32_t ntc;      // tiles per tile row (nrows_pad / 32)
  int32_t ldc;      // nrows_pad
  int32_t agrow0;   // global A row of dense row 0 (row scales)
  int32_t bgrow0;   // global B row of dense col 0 (column scales)
};

// Fused CRT + scatter (split path): one 32x32 tile of a pair's dense result, and the pair's
// residue regions (its split-K parts are consecutive regions).
struct FusedTile {
  int32_t pair;    // CrtPair index
  int32_t tr, tc;  // tile row / column in the pair's dense result
  int32_t job;     // ScatterJob holding the whole tile (all rows in one A tuple, all columns
                   // in one B tuple, block nonzero), else -1
  int32_t rin, cin;  // with job >= 0: the tile's first row / column within its tuples
  int32_t pad[2];
};
struct CrtPair {
  int64_t c_base;   // first element of part 0's dense result
  int64_t pstride;  // elements between consecutive parts
  int32_t nparts;
  int32_t ntc;      // tiles per tile row
  int32_t agrow0;   // global A row of dense row 0 (row scales)
  int32_t bgrow0;   // global B row of dense col 0 (column scales)
};

struct PackSide {
  DevBuf jobs;
  int njobs = 0;
  int64_t nitems = 0;  // (row, 16-column group) work items
  int64_t total = 0;   // elements (FP64) or bytes per residue plane (int8)
  int64_t nrows = 0;   // rows over all class matrices (row-scale arrays)
  DevBuf units;        // 64x64 tiles of the jobs
  int64_t nunits = 0;
  DevBuf rowunits;     // 64-row groups of the contiguous-row jobs (all columns), for k_rowmax
  int64_t nrowunits = 0;
  DevBuf kfunits;      // 32x128 tiles of the jobs with contiguous rows (k_pack_kf)
  int64_t nkfunits = 0;
  DevBuf buf;          // packed FP64 matrices, or nmod int8 residue planes
  DevBuf rowmax, invscale;
  // Row-scale reuse (split precision).  rowmax keeps the row maxima of the operand packed by
  // the last execution; a reusing pack scales with them, recomputes the maxima into
  // rowmax_new in the same pass (its first block also clears `changed`), and k_check_rowmax
  // copies them back, clears rowmax_new for the next execution and raises `changed` (and
  // changed_host) when any differs, which enables the gated second pack.
  DevBuf rowmax_new, changed;
  HostFlag changed_host;         // raised by k_check_rowmax in pinned memory, polled by the host
  bool fresh_clean = false;      // rowmax_new is known to be all zero
  const void* last_src = nullptr;
  bool have_max = false;
  int backoff = 0;               // executions left before reuse is tried again
  // Constant operand (tmx_contraction_set_constant): the caller promises the same values at
  // every execution from the same address, so `buf` (and the row scales) packed from
  // packed_src are reused without a pass over the operand.
  bool constant = false;
  const void* packed_src = nullptr;
};

// One sector GEMM: geometry of its dense product and its tile shape (autotunable).
struct SectorGeom {
  int64_t a_base, b_base, c_base;
  int32_t seg_b, seg_e;
  int32_t a_kpad, b_kpad;
  int32_t mrows_pad, nrows_pad;  // padded rows of the A / B class matrices
  int32_t mrows, nrows;          // rows that carry data
  int64_t klen;
  int tm, tn;
};

struct ShapeTasks {
  int tm = 0, tn = 0;
  DevBuf tasks;
  int ntasks = 0;  // laid-out list, padding included
  int nreal = 0;   // tasks with work
};

struct SectorPlan {
  bool split = false;
  int bits_a = 0, bits_b = 0, nmod = 0;
  CrtConsts crt;
  PackSide A, B;
  std::vector<SectorGeom> geoms;
  std::vector<ShapeTasks> shapes;
  DevBuf ksegs;
  DevBuf cdense;        // dense FP64 result, row-major per sector (ld = nrows_pad)
  DevBuf cres;          // split: nmod residue planes of the dense result, 32x32-tile native
  DevBuf crt_secs;      // split: CrtSector per sector
  int ncrt_secs = 0;
  int64_t crt_tiles = 0;
  // split: fused CRT + scatter (used when C's labels are the A-outer labels followed by the
  // B-outer labels, so a C block is row-major over (dense row, dense column)).  Each
  // global A / B row maps to (tuple index, row within the tuple), -1 for padding; jobtab maps
  // (A tuple, B tuple) to the scatter job of their C block in scatter[0], -1 if it is zero.
  DevBuf fused_tiles, crt_pairs, arow_info, brow_info, jobtab;
  int64_t nfused = 0;
  int32_t jobtab_n = 0;  // B tuples per jobtab row
  // Fused pass layout: 1 = C's labels are the A-outer then the B-outer labels; 2 / 3 = other
  // orders, element offsets from the rows' and columns' digits (arow_digits / brow_digits, 8
  // bits per mode, mode d at bits 8d) and the block's strides, with C contiguous along the
  // dense columns (2) or rows (3).
  int fused_layout = 0;
  DevBuf arow_digits, brow_digits;
  int64_t c_total = 0;  // dense result elements (= bytes per residue plane)
  int64_t kmax = 0;
  // Scatter launch: each job sums its C block's split-K parts and adds (or, assigning, writes)
  // the block once.
  struct ScatterRound {
    DevBuf jobs, units;
    int njobs = 0;
    int64_t nunits = 0;
  };
  std::vector<ScatterRound> scatter;
  int64_t ws = 0;
  double exec_flops = 0;  // MFMA work including padding (per modulus for split)
  int split_factor = 1;   // parallel split-K factor chosen by the launch model
  // Fingerprint of the product the plan computes: operand tensors, pack jobs, sector geometry
  // (tile shapes excluded), K segments and precision.  Plans with equal keys fed the same
  // operand data hold identical packed operands and dense results, laid out identically.
  uint64_t product_key = 0;
  // C = alpha*A*B instead of +=: the scatter writes, and the C blocks of the contraction that no
  // job writes (listed as element ranges, zero_prefix holding running lengths) are zeroed.
  bool assign = false;
  DevBuf zero_off, zero_prefix;
  int nzero = 0;
  int64_t zero_elems = 0;
  // Flops-weighted mean over the sectors of mrows * nrows / (mrows + nrows) (real class rows):
  // what an FP64 GEMM's intensity grows with, since it reads (mrows + nrows) K elements for
  // 2 mrows nrows K flops.
  double mn_intensity = 0;
};

// Tile shape (rows of A x rows of B per workgroup) for a sector GEMM: the autotuner's
// recorded choice for this geometry if there is one, else the cost model.
void choose_tile(bool split, int nmod, int64_t mrows, int64_t nrows, int64_t klen, int* tm, int* tn);
// Builds the GEMM task lists of all sectors from their geometry and tile shapes, one launch
// per shape, each laid out for the XCD-aware workgroup mapping (balanced XCD ranges, longest
// tasks first, padded with empty tasks), and the plan's executed-flop count.
void build_sector_tasks(SectorPlan& sp);
// Simulated duration of the GEMM launches of these sectors (list schedule of the same layout
// over the CU slots each kernel's occupancy allows), for choosing the split-K factor.
double estimate_sector_seconds(bool split, int nmod, const std::vector<SectorGeom>& geoms);
// Times every candidate tile shape of every sector on the plan's packed operands (which must
// be current: run after an execution) and keeps the fastest; records the choices in the tune
// cache.  Returns the number of sectors whose shape changed.
int autotune_sector_plan(SectorPlan& sp, hipStream_t s);
// CRT constants for the first n moduli (host, exact 128-bit integer arithmetic).
CrtConsts make_crt_consts(int n, const int* moduli);

}  // namespace tmx
