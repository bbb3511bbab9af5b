// k_sector.hip -- sector GEMM pipeline on gfx950:
//   pack (block layout -> MFMA fragment-major class matrices, permutation fused into the
//   gather; FP64 copy or int8 residues), dense GEMM (v_mfma_f64_16x16x4f64 or
//   v_mfma_i32_32x32x32_i8), scatter (dense result -> C blocks in C's mode order, FP64
//   accumulate, or Garner CRT reconstruction of the int8 residues into FP64).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <queue>
#include <tuple>

#include "sector.hpp"
#include "tmx_device.hpp"

namespace tmx {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));
typedef int i16v __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// Row scales of split precision.  Row maxima are kept as the IEEE bits of max|x| and combined
// as unsigned integers (atomicMax): for x >= 0 the bit order is the value order, with +Inf
// above every finite value and the quiet NaN 0x7ff8... above +Inf.  fmax carries an Inf
// through but skips NaN, so the loops also note whether any x was NaN and mark_nonfinite then
// turns the row's maximum into that one NaN bit pattern (whatever NaN the data held).
constexpr long long kQuietNaNBits = 0x7ff8000000000000ll;
__device__ __forceinline__ double mark_nonfinite(double m, bool nan_seen) {
  return nan_seen ? __longlong_as_double(kQuietNaNBits) : m;
}
__device__ __forceinline__ unsigned long long max_bits(unsigned long long a, unsigned long long b) {
  return a > b ? a : b;
}
__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}

// Scale of a row whose maximum |x| is rm: sigma = top / rm, applied as (x * 2^-e) * (top / f)
// with rm = f * 2^e, f in [1, 2).  The power-of-two step is exact, so sigma cannot overflow for
// tiny or subnormal rm, and the scaled integers equal rint(x * (top / rm)) bit for bit wherever
// top / rm is a normal number.  The inverse is kept as (f / top, e), e as the integer bits of
// the second double; the CRT applies ldexp(v * (f_a / top) * (f_b / top), e_a + e_b), one
// rounding of the exponent at the end.
// A row holding Inf or NaN (rm not finite) gets scale 0 and inverse (NaN, 0): whatever its
// residues hold (Inf * 0 is NaN, which packs as some residue), every element of its output row
// or column is NaN, and a row of A meets only its own output row (a row of B its column).  A
// zero row packs zeros with inverse (0, 0).
struct RowScale {
  double sig;
  int e;
};
__device__ __forceinline__ RowScale row_scale(double rm, double top) {
  RowScale r;
  const bool pos = rm > 0.0 && rm <= 1.7976931348623157e308;  // finite and nonzero
  r.e = pos ? ilogb(rm) : 0;
  r.sig = pos ? top / ldexp(rm, -r.e) : 0.0;
  return r;
}
__device__ __forceinline__ double2 row_inverse(double rm, const RowScale& r, double top) {
  const double m = !(rm <= 1.7976931348623157e308) ? __longlong_as_double(kQuietNaNBits)
                                                    : (rm > 0.0 ? ldexp(rm, -r.e) / top : 0.0);
  return make_double2(m, __longlong_as_double((long long)r.e));
}
__device__ __forceinline__ double scale_elem(double t, const RowScale& r) {
  return rint(ldexp(t, -r.e) * r.sig);
}
// v * (f_a / top) * (f_b / top) * 2^(e_a + e_b) for the inverse ia of a row and (bm, be) of a
// column (be = the exponent of ib, taken once per column).
__device__ __forceinline__ int inv_exp(double2 ib) { return (int)__double_as_longlong(ib.y); }
__device__ __forceinline__ double unscale(double v, double2 ia, double bm, int be) {
  return ldexp(v * ia.x * bm, inv_exp(ia) + be);
}

// Split packing of 16 consecutive cells of one class-matrix row (in LDS): each is scaled by
// the row's scale and rounded, x = rint(cell * sigma), and reduced modulo every modulus; plane
// l gets the 16 residue bytes at dst + l * plane.  Also takes the cells' max |x| into mrow,
// which turns NaN at a NaN cell and keeps it (an Inf cell makes it Inf).  sigma = sig * 2^-e
// is applied as one multiply unless it overflows (row maxima below about 2^-987), where the
// cells are scaled by 2^-e first.  (The max is written as a compare and select rather than
// fmax, which skips NaN: a separate NaN flag per cell costs the kernel its 128-VGPR budget.)
__device__ __forceinline__ void residues16(const double* cells, const RowScale& rs, const CrtConsts& gc,
                                           uint8_t* dst, int64_t plane, double& mrow) {
  double x[16];
  uint32_t xlo[16];
  const double sigma = ldexp(rs.sig, -rs.e);
  const bool tiny = !(sigma <= 1.7976931348623157e308);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const double t = cells[j];  // padding reads 0
    mrow = (fabs(t) > mrow || t != t) ? fabs(t) : mrow;
    x[j] = rint(t * sigma);
  }
  if (tiny) {
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = rint(ldexp(cells[j], -rs.e) * rs.sig);
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    // x + 1.5*2^52 is exact for |x| <= 2^50 and its low 32 bits are x in two's complement.
    xlo[j] = (uint32_t)__double_as_longlong(x[j] + 0x1.8p52);
  }
  for (int l = 0; l < gc.n; ++l) {
    // The quotient q = round(x / m): fma(x, 1/m, 1.5*2^52) rounds the exact product x * (1/m),
    // which for |x| <= 2^50 lies within 2^-3/m of x/m, inside the 1/(2m) rounding margin of odd
    // m (m = 256 is exact; a tie rounds to even), and its low 32 bits are q.  The residue
    // r = x - q m lies in [-(m-1)/2, (m-1)/2], or [-128, 128] for m = 256 (128 and -128 share
    // the byte 0x80); its low byte is that of x + q (2^24 - m) taken modulo 2^24, one 24-bit
    // multiply-add on the low words.
    const double im = gc.rcp[l];
    const uint32_t negm = (1u << 24) - (uint32_t)gc.m[l];
    uint32_t d[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const uint32_t q = (uint32_t)__double_as_longlong(fma(x[j], im, 0x1.8p52));
      d[j] = xlo[j] + __umul24(q, negm);
    }
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)  // low bytes of four words -> one word
      w[k] = __builtin_amdgcn_perm(d[4 * k + 1], d[4 * k], 0x0c0c0400u) |
             __builtin_amdgcn_perm(d[4 * k + 3], d[4 * k + 2], 0x04000c0cu);
    *reinterpret_cast<uint4*>(dst + (int64_t)l * plane) = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// ---------------------------------------------------------------------------
// Loads the 64x64 tile (rows nr, columns nc valid; the rest reads 0) of a pack job that is not
// contiguous along whole rows into LDS, lanes along the source's contiguous direction.  When
// the fastest column digit (kcontig) or row digit (otherwise) has an extent L dividing 64, one
// wave load covers 64/L rows x L columns (or L rows x 64/L columns): 64 consecutive doubles when
// the next-fastest source mode is the other side's fastest digit, e.g. T(a,c,i,k) with rows
// (a,i) and K = (c,k).  Otherwise lanes run along 64 columns or 64 rows.  Addresses always come
// from the row and column offset tables; the mapping only decides which lanes share a load.
__device__ __forceinline__ void load_tile64(double (*tile)[65], const PackJob& jb, const double* s,
                                            const int* roff, const int* coff, int nr, int nc,
                                            int tid) {
  const int lane = tid & 63, w = tid >> 6;
  const int L = jb.kcontig ? jb.c.ext[jb.c.n - 1] : jb.r.ext[jb.r.n - 1];
  if (L >= 2 && L < 64 && (L & (L - 1)) == 0) {
    const int lg = __builtin_ctz(L), sh = 6 - lg;  // 2^sh = 64 / L
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
      const int f = w + 4 * i;  // wave load 0..63: group f >> sh, sub-group f & (2^sh - 1)
      int rr, cc;
      if (jb.kcontig) {
        rr = (f >> sh) * (64 >> lg) + (lane >> lg);
        cc = (f & ((1 << sh) - 1)) * L + (lane & (L - 1));
      } else {
        rr = (f & ((1 << sh) - 1)) * L + (lane & (L - 1));
        cc = (f >> sh) * (64 >> lg) + (lane >> lg);
      }
      tile[rr][cc] = (rr < nr && cc < nc) ? s[roff[rr] + coff[cc]] : 0.0;
    }
  } else if (jb.kcontig) {
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
      const int rr = w + 4 * i, cc = lane;
      tile[rr][cc] = (rr < nr && cc < nc) ? s[roff[rr] + coff[cc]] : 0.0;
    }
  } else {
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
      const int rr = lane, cc = w + 4 * i;
      tile[rr][cc] = (rr < nr && cc < nc) ? s[roff[rr] + coff[cc]] : 0.0;
    }
  }
}

// Row stride in doubles of an LDS pack tile of 2^lt rows x 4096 / 2^lt columns (lt = 3..8):
// 65 for 64 rows (load_tile64's layout), columns + 1 for 128 and 256 rows, else columns + 2.
// Every shape fits kPackTile doubles (256 x 17 the largest).
constexpr int kPackTile = 256 * 17;
__device__ __forceinline__ int pack_ld(int lt) {
  return lt == 6 ? 65 : (4096 >> lt) + (lt > 6 ? 1 : 2);
}

// Loads a pack tile of TR = 2^lt rows x TC = 4096 / TR columns (rows nr, columns nc valid; the
// rest reads 0) into LDS with row stride pack_ld(lt).  64-row tiles use load_tile64; the other
// shapes put lanes along columns when the source is contiguous along K and along rows
// otherwise, e.g. L(a,b,i,j) with rows j of 8: 8 consecutive j of 8 consecutive (a,b,i)
// columns, one contiguous run per wave load; T(a,b,k,l) with K = (k,l) of 16: 256 x 16 tiles,
// 4 rows of 16 contiguous columns per wave load.
__device__ __forceinline__ void load_pack_tile(double* tile, int lt, const PackJob& jb, const double* s,
                                               const int* roff, const int* coff, int nr, int nc,
                                               int tid) {
  if (lt == 6) {
    load_tile64(reinterpret_cast<double(*)[65]>(tile), jb, s, roff, coff, nr, nc, tid);
    return;
  }
  const int ld = pack_ld(lt), ltc = 12 - lt;
#pragma unroll 4
  for (int i = 0; i < 16; ++i) {
    const int e = tid + 256 * i;
    const int rr = jb.kcontig ? e >> ltc : e & ((1 << lt) - 1);
    const int cc = jb.kcontig ? e & ((1 << ltc) - 1) : e >> lt;
    tile[rr * ld + cc] = (rr < nr && cc < nc) ? s[roff[rr] + coff[cc]] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Packing.  One workgroup iteration handles a tile of one source block: 2^lt rows x 4096 / 2^lt
// columns (64x64 for blocks with more than 32 rows).  The tile's row and column offsets go to
// LDS once, the source tile is read with lanes along the source's contiguous direction
// (coalesced) into an LDS tile, and then each thread emits one row's 16 consecutive columns in
// MFMA fragment-major order (int8: one 16-byte store per residue plane; FP64: 16-row runs of
// 128 bytes).
// Int8 residues scale rows by `rowmax` (from k_rowmax / k_rowmax_t, or kept from the previous
// execution); with `maxout` the pass also takes each row's maximum |x| over the tile (one
// atomicMax per row and tile), with `gate` the whole grid returns unless *gate != 0, and with
// `zero_flag` the first block clears that flag (for a check that runs after the pass).
template <int MODE>  // 0: FP64 copy, 2: int8 residues
__global__ __launch_bounds__(256) void k_pack(const PackJob* __restrict__ jobs,
                                              const TileUnit* __restrict__ units, int64_t nunits,
                                              const double* __restrict__ src,
                                              double* __restrict__ dst64,
                                              uint8_t* __restrict__ dst8, int64_t plane,
                                              const unsigned long long* __restrict__ rowmax,
                                              double2* __restrict__ invscale, CrtConsts gc,
                                              double scale_top,
                                              unsigned long long* __restrict__ maxout,
                                              const int* __restrict__ gate,
                                              int* __restrict__ zero_flag) {
  __shared__ double tile[kPackTile];
  __shared__ int roff[256], coff[512];
  __shared__ unsigned long long red[256];
  const int tid = threadIdx.x;
  if (gate != nullptr && *gate == 0) return;
  if (zero_flag != nullptr && blockIdx.x == 0 && tid == 0) *zero_flag = 0;
  for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const TileUnit un = units[u];
    const PackJob& jb = jobs[un.job];
    const int lt = un.lt, TR = 1 << lt, TC = 4096 >> lt, ld = pack_ld(lt);
    const int nr = min(TR, jb.rows - un.r0), nc = min(TC, jb.cols - un.c0);
    if (tid < TR) roff[tid] = tid < nr ? radix_off(jb.r, un.r0 + tid) : 0;
    for (int c = tid; c < TC; c += 256) coff[c] = c < nc ? radix_off(jb.c, un.c0 + c) : 0;
    __syncthreads();
    load_pack_tile(tile, lt, jb, src + jb.src_off, roff, coff, nr, nc, tid);
    __syncthreads();
    const int rr = tid & (TR - 1), q = tid >> lt;  // row, 16-column group
    const double* trow = tile + rr * ld + 16 * q;
    const int row = jb.row0 + un.r0 + rr;
    const int c0 = un.c0 + 16 * q;
    const bool active = rr < nr && c0 < jb.cols;
    double mrow = 0.0;  // max |x| of this thread's 16 cells (NaN once a NaN is seen)
    if (MODE == 0) {
      if (active) {
        const int ncol = min(16, jb.cols - c0);
        for (int j = 0; j < ncol; ++j)
          dst64[jb.dst_base + frag_f64(row, jb.col0 + c0 + j, jb.kpad)] = trow[j];
      }
    } else if (active) {
      const int grow = jb.growbase + row;
      const double rm = __longlong_as_double((long long)rowmax[grow]);
      const RowScale rs = row_scale(rm, scale_top);
      if (un.c0 == 0 && q == 0) invscale[grow] = row_inverse(rm, rs, scale_top);
      uint8_t* dst = dst8 + jb.dst_base + frag_i8(row, jb.col0 + c0, jb.kpad);
      residues16(trow, rs, gc, dst, plane, mrow);
    }
    // = red[q * TR + rr]; inactive threads hold padding
    if (MODE == 2 && maxout != nullptr) red[tid] = dbits(mark_nonfinite(mrow, mrow != mrow));
    __syncthreads();
    if (MODE == 2 && maxout != nullptr && tid < nr) {
      // red is rewritten only after two more barriers of the next iteration
      unsigned long long m = 0;
      for (int g = 0; g < (256 >> lt); ++g) m = max_bits(m, red[g * TR + tid]);
      if (m != 0) atomicMax(maxout + jb.growbase + jb.row0 + un.r0 + tid, m);
    }
  }
}

// ---------------------------------------------------------------------------
// Row maxima |x| (split precision) of the blocks k_rowmax leaves out: those whose last mode is
// a K mode but whose K columns are not one contiguous run per row (e.g. T(a,c,i,k) with
// K = (c,k)), and all blocks of at most 32 rows (e.g. L(a,b,i,j) with rows j): the same tiles
// and coalesced LDS-staged loads as k_pack, then 256 / 2^lt threads per row reduce its columns
// (16 each) and one atomicMax (on the IEEE bits, monotonic for x >= 0) per row and tile.
__global__ __launch_bounds__(256) void k_rowmax_t(const PackJob* __restrict__ jobs,
                                                  const TileUnit* __restrict__ units, int64_t nunits,
                                                  const double* __restrict__ src,
                                                  unsigned long long* __restrict__ rowmax) {
  __shared__ double tile[kPackTile];
  __shared__ int roff[256], coff[512];
  __shared__ unsigned long long red[256];
  const int tid = threadIdx.x;
  for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const TileUnit un = units[u];
    const PackJob& jb = jobs[un.job];
    const int lt = un.lt, TR = 1 << lt, TC = 4096 >> lt, ld = pack_ld(lt);
    if (jb.rows > 32 && (jb.kfull || !jb.kcontig)) continue;  // k_rowmax's rows; uniform
    const int nr = min(TR, jb.rows - un.r0), nc = min(TC, jb.cols - un.c0);
    if (tid < TR) roff[tid] = tid < nr ? radix_off(jb.r, un.r0 + tid) : 0;
    for (int c = tid; c < TC; c += 256) coff[c] = c < nc ? radix_off(jb.c, un.c0 + c) : 0;
    __syncthreads();
    load_pack_tile(tile, lt, jb, src + jb.src_off, roff, coff, nr, nc, tid);
    __syncthreads();
    const int rr = tid & (TR - 1), q = tid >> lt;
    double m = 0.0;
    bool nan_seen = false;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const double t = tile[rr * ld + 16 * q + j];
      m = fmax(m, fabs(t));
      nan_seen |= t != t;
    }
    red[tid] = dbits(mark_nonfinite(m, nan_seen));
    __syncthreads();
    if (tid < nr) {
      unsigned long long mb = 0;
      for (int g = 0; g < (256 >> lt); ++g) mb = max_bits(mb, red[g * TR + tid]);
      atomicMax(rowmax + jb.growbase + jb.row0 + un.r0 + tid, mb);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Packing of blocks with more than 32 rows whose columns (K) are one contiguous run per row
// (K labels trail the source block's modes in order, e.g. V(a,b,c,d) with K = (c,d); such
// blocks with fewer rows go through k_pack's wider tiles).  A workgroup iteration takes
// 32 rows x 128 columns: each wave reads 8 rows with 16-byte loads, 1 KB of one row per
// instruction (8-byte loads when a row run is not 16-byte aligned), into an LDS tile; then each
// thread emits one row's 16 consecutive columns in fragment order, as k_pack does.
// `rowmax`, `maxout`, `gate` and `zero_flag` as in k_pack.
template <int MODE>  // 0: FP64 copy, 2: int8 residues
__global__ __launch_bounds__(256, 4) void k_pack_kf(const PackJob* __restrict__ jobs,
                                                 const TileUnit* __restrict__ units, int64_t nunits,
                                                 const double* __restrict__ src,
                                                 double* __restrict__ dst64,
                                                 uint8_t* __restrict__ dst8, int64_t plane,
                                                 const unsigned long long* __restrict__ rowmax,
                                                 double2* __restrict__ invscale, CrtConsts gc,
                                                 double scale_top,
                                                 unsigned long long* __restrict__ maxout,
                                                 const int* __restrict__ gate,
                                                 int* __restrict__ zero_flag) {
  constexpr int TR = 32, TC = 128, LD = TC + 2;  // row stride 260 dwords: 4 banks apart
  __shared__ double tile[TR][LD];
  __shared__ int roff[TR];
  __shared__ unsigned long long red[4][TR];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (gate != nullptr && *gate == 0) return;
  if (zero_flag != nullptr && blockIdx.x == 0 && tid == 0) *zero_flag = 0;
  for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const TileUnit un = units[u];
    const PackJob& jb = jobs[un.job];
    const int nr = min(TR, jb.rows - un.r0), nc = min(TC, jb.cols - un.c0);
    if (tid < TR) roff[tid] = tid < nr ? radix_off(jb.r, un.r0 + tid) : 0;
    __syncthreads();
    const double* s = src + jb.src_off + un.c0;
    const int cc = 2 * lane;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int rr = w * 8 + i;
      double v0 = 0.0, v1 = 0.0;
      if (rr < nr) {
        const double* rp = s + roff[rr];
        if (cc + 1 < nc && ((reinterpret_cast<uintptr_t>(rp) & 15) == 0)) {
          const double2 v = *reinterpret_cast<const double2*>(rp + cc);
          v0 = v.x;
          v1 = v.y;
        } else {
          if (cc < nc) v0 = rp[cc];
          if (cc + 1 < nc) v1 = rp[cc + 1];
        }
      }
      tile[rr][cc] = v0;
      tile[rr][cc + 1] = v1;
    }
    __syncthreads();
    const int rr = tid & 31, q = tid >> 5;  // row, 16-column group
    const int row = jb.row0 + un.r0 + rr;
    double mrow = 0.0;  // max |x| of this thread's 16 cells (NaN once a NaN is seen)
    const int c0 = un.c0 + 16 * q;
    const bool active = rr < nr && c0 < jb.cols;
    if (MODE == 0) {
      if (active) {
        const int ncol = min(16, jb.cols - c0);
        for (int j = 0; j < ncol; ++j)
          dst64[jb.dst_base + frag_f64(row, jb.col0 + c0 + j, jb.kpad)] = tile[rr][16 * q + j];
      }
    } else if (active) {
      const int grow = jb.growbase + row;
      const double rm = __longlong_as_double((long long)rowmax[grow]);
      const RowScale rs = row_scale(rm, scale_top);
      if (un.c0 == 0 && q == 0) invscale[grow] = row_inverse(rm, rs, scale_top);
      uint8_t* dst = dst8 + jb.dst_base + frag_i8(row, jb.col0 + c0, jb.kpad);
      residues16(&tile[rr][16 * q], rs, gc, dst, plane, mrow);  // beyond the block: 0
    }
    if (MODE == 2 && maxout != nullptr) {
      // lanes l and l + 32 of a wave hold the same row (column groups 2w and 2w + 1);
      // inactive threads hold only padding
      const double mk = mark_nonfinite(mrow, mrow != mrow);
      const unsigned long long m = max_bits(dbits(mk), dbits(__shfl_xor(mk, 32)));
      if (lane < 32) red[w][rr] = m;
    }
    __syncthreads();
    if (MODE == 2 && maxout != nullptr && tid < nr) {
      // red is rewritten only after two more barriers of the next iteration
      const unsigned long long m = max_bits(max_bits(red[0][tid], red[1][tid]), max_bits(red[2][tid], red[3][tid]));
      if (m != 0) atomicMax(maxout + jb.growbase + jb.row0 + un.r0 + tid, m);
    }
  }
}

// Row-scale reuse check: rows whose recomputed maximum differs from the kept one take the
// new value and set *changed (the gate of the second packing pass) and, when given,
// *host_changed (pinned memory the host polls).  The recomputed maxima are cleared for the
// next reusing pass, which accumulates into them with atomicMax.
__global__ __launch_bounds__(256) void k_check_rowmax(unsigned long long* __restrict__ fresh,
                                                      unsigned long long* __restrict__ kept,
                                                      int64_t n, int* __restrict__ changed,
                                                      int* __restrict__ host_changed) {
  bool any = false;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const unsigned long long f = fresh[i];
    fresh[i] = 0;
    if (f != kept[i]) {
      kept[i] = f;
      any = true;
    }
  }
  // one write per block that saw a change (the host flag is in system memory)
  if (__syncthreads_or(any) && threadIdx.x == 0) {
    *changed = 1;
    if (host_changed != nullptr) *host_changed = 1;
  }
}

// ---------------------------------------------------------------------------
// Row maxima |x| of the class-matrix rows (split precision) for blocks whose K columns are one
// contiguous run per row (each wave streams whole rows) or whose last mode is a row mode
// (lanes along rows, eight loads in flight per lane).  A workgroup iteration takes 64 rows of
// one pack job across all of the job's columns, keeping running maxima in registers; then one
// LDS reduction per row and one atomicMax (on the IEEE bits, monotonic for x >= 0) per row
// and job.  Blocks whose last mode is a K mode but whose K columns are not one run per row go
// to k_rowmax_t.
__global__ __launch_bounds__(256) void k_rowmax(const PackJob* __restrict__ jobs,
                                                const TileUnit* __restrict__ units, int64_t nunits,
                                                const double* __restrict__ src,
                                                unsigned long long* __restrict__ rowmax) {
  __shared__ int roff[64];
  __shared__ int coff[256];
  __shared__ unsigned long long red[64][65];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const TileUnit un = units[u];
    const PackJob& jb = jobs[un.job];
    const int nr = min(64, jb.rows - un.r0);
    if (tid < 64) roff[tid] = tid < nr ? radix_off(jb.r, un.r0 + tid) : 0;
    const double* s = src + jb.src_off;
    if (jb.kfull) {
      // each row's columns are one contiguous run: wave w streams rows 16w .. 16w+15 whole
      __syncthreads();  // roff visible
      for (int i = 0; i < 16; ++i) {
        const int rr = w * 16 + i;
        if (rr >= nr) break;
        const double* rp = s + roff[rr];
        double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
        bool nan_seen = false;
        int c = lane;
        for (; c + 192 < jb.cols; c += 256) {
          const double t0 = rp[c], t1 = rp[c + 64], t2 = rp[c + 128], t3 = rp[c + 192];
          m0 = fmax(m0, fabs(t0));
          m1 = fmax(m1, fabs(t1));
          m2 = fmax(m2, fabs(t2));
          m3 = fmax(m3, fabs(t3));
          nan_seen |= (t0 != t0) | (t1 != t1) | (t2 != t2) | (t3 != t3);
        }
        for (; c < jb.cols; c += 64) {
          const double t = rp[c];
          m0 = fmax(m0, fabs(t));
          nan_seen |= t != t;
        }
        red[rr][lane] = dbits(mark_nonfinite(fmax(fmax(m0, m1), fmax(m2, m3)), nan_seen));
      }
      __syncthreads();
      if (tid < nr) {
        unsigned long long m = 0;
        for (int k = 0; k < 64; ++k) m = max_bits(m, red[tid][k]);
        atomicMax(rowmax + jb.growbase + jb.row0 + un.r0 + tid, m);
      }
      __syncthreads();
      continue;
    }
    double mx[8];
    bool nan_seen = false;
#pragma unroll
    for (int i = 0; i < 8; ++i) mx[i] = 0.0;
    for (int cb = 0; cb < jb.cols; cb += 256) {
      const int ncb = min(256, jb.cols - cb);
      __syncthreads();  // previous chunk's coff reads are done (and roff is visible)
      if (tid < ncb) coff[tid] = radix_off(jb.c, cb + tid);
      __syncthreads();
      if (lane < nr) {
        // lanes along rows; wave w owns columns w, w+4, ..., eight loads in flight per lane
        const int ro = roff[lane];
        int cc = w;
        for (; cc + 28 < ncb; cc += 32) {
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const double t = s[ro + coff[cc + 4 * i]];
            mx[i] = fmax(mx[i], fabs(t));
            nan_seen |= t != t;
          }
        }
        for (; cc < ncb; cc += 4) {
          const double t = s[ro + coff[cc]];
          mx[0] = fmax(mx[0], fabs(t));
          nan_seen |= t != t;
        }
      }
    }
    {
      double m = mx[0];
#pragma unroll
      for (int i = 1; i < 8; ++i) m = fmax(m, mx[i]);
      red[lane][w] = dbits(mark_nonfinite(m, nan_seen));
    }
    __syncthreads();
    if (tid < nr) {
      unsigned long long m = 0;
      for (int k = 0; k < 4; ++k) m = max_bits(m, red[tid][k]);
      atomicMax(rowmax + jb.growbase + jb.row0 + un.r0 + tid, m);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ int xcd_task(int ntasks) {
  // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range
  // of tasks so tiles sharing A rows (M-panel-major order) share that XCD's L2.
  const int b = blockIdx.x, per = (int)((gridDim.x + 7) >> 3);
  return (b & 7) * per + (b >> 3);
}

__device__ __forceinline__ void lds_dma16(const void* g, __attribute__((address_space(3))) void* l) {
  __builtin_amdgcn_global_load_lds(g, l, 16, 0, 0);
}

// Cursor over the K stream of a task in GEMM stages.  A stage is SCOLS columns inside one K
// segment (segment lengths are multiples of SCOLS); in 1 KB load units (a 32-column int8
// fragment, or two 4-column FP64 fragments = 8 columns) stage j of segment seg starts at unit
// a + j*UPS of A and b + j*UPS of B.  Advancing is branch-free: the description of the
// segment after the current one is reloaded on every step (clamped to the task's last
// segment) and only read on the next step, and past the end the cursor stays on the last
// stage, so loads issued beyond the stream stay in bounds (their data is never consumed).
// All fields are wave-uniform.
template <int SCOLS, int UCOLS>  // columns per stage, columns per 1 KB unit
struct StageCursor {
  static constexpr int UPS = SCOLS / UCOLS;
  const KSeg* segs;
  int seg, last, j, len, a, b, nlen, na, nb;
  __device__ __forceinline__ void init(const KSeg* s, int b0, int e) {
    segs = s;
    seg = b0;
    last = e - 1;
    j = 0;
    const KSeg k0 = s[b0];
    len = (int)((uint32_t)k0.len / SCOLS);
    a = (int)((uint32_t)k0.a_col / UCOLS);
    b = (int)((uint32_t)k0.b_col / UCOLS);
    const KSeg k1 = s[min(b0 + 1, last)];
    nlen = (int)((uint32_t)k1.len / SCOLS);
    na = (int)((uint32_t)k1.a_col / UCOLS);
    nb = (int)((uint32_t)k1.b_col / UCOLS);
  }
  __device__ __forceinline__ int unit_a() const { return a + j * UPS; }
  __device__ __forceinline__ int unit_b() const { return b + j * UPS; }
  __device__ __forceinline__ void advance() {
    const bool in_seg = j + 1 < len;
    const bool sw = !in_seg && seg < last;
    j = in_seg ? j + 1 : (sw ? 0 : j);
    seg = sw ? seg + 1 : seg;
    len = sw ? nlen : len;
    a = sw ? na : a;
    b = sw ? nb : b;
    const KSeg k = segs[min(seg + 1, last)];
    nlen = (int)((uint32_t)k.len / SCOLS);
    na = (int)((uint32_t)k.a_col / UCOLS);
    nb = (int)((uint32_t)k.b_col / UCOLS);
  }
};

// FP64 sector GEMM: workgroup tile TM x TN over 4 waves (2x2), each (TM/2) x (TN/2) =
// (TM/32) x (TN/32) v_mfma_f64_16x16x4f64 tiles (128x128 and 128x160 instantiated; the
// 16-column granularity lets 160 cover N = 800 exactly).  The fragment-major class matrices
// hold a 16-row tile's k-steps back to back (512 bytes each), so one global_load_lds_dwordx4
// wave instruction copies two k-steps of one row tile (1 KB).  A stage is those two k-steps
// for every row tile of A and B; stages go into a ring of NS LDS slots and stage st+NS is
// issued as soon as every wave has read stage st (one barrier per stage), so NS-1 stages of
// loads are in flight under the MFMA work.  Every wave issues the same number of loads per
// stage (spares read the zero fragment into a scratch slot), so the vmcnt thresholds are
// constants.  Fragment reads are double-buffered in registers: the next k-step's reads go out
// right after the first MFMA of the current one.  Two workgroups share a CU.
template <int TM, int TN, int NS>
__global__ __launch_bounds__(256, 2) void k_gemm_f64(const SecTask* __restrict__ tasks, int ntasks,
                                                     const KSeg* __restrict__ segs,
                                                     const double* __restrict__ Ap,
                                                     const double* __restrict__ Bp,
                                                     double* __restrict__ Cd) {
  constexpr int AF = TM / 16, BF = TN / 16, SF = AF + BF;  // 1 KB units per stage
  constexpr int FPW = (SF + 3) / 4;
  constexpr int SLOT = SF;
  constexpr int WTM = TM / 32, WTN = TN / 32;
  static_assert(WTM * 32 == TM && WTN * 32 == TN, "tile must split over 2x2 waves");
  __shared__ i4 lds[NS][SLOT * 64];
  const int task = xcd_task(ntasks);
  if (task >= ntasks) return;
  const SecTask tk = tasks[task];
  if (tk.mrows == 0) return;  // padding of the XCD layout
  const uint8_t* A = reinterpret_cast<const uint8_t*>(Ap + tk.a_base);
  const uint8_t* B = reinterpret_cast<const uint8_t*>(Bp + tk.b_base);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  const int wm = wave & 1, wn = wave >> 1;
  const int art = tk.mrows >> 4, brt = tk.nrows >> 4;
  const int64_t akt = tk.a_kpad >> 2, bkt = tk.b_kpad >> 2;  // k-steps per 16-row tile
  const int rt0 = tk.m0 >> 4, ct0 = tk.n0 >> 4;
  int nst = 0;
  for (int sg = tk.kseg_begin; sg < tk.kseg_end; ++sg) nst += (int)((uint32_t)segs[sg].len >> 3);
  StageCursor<8, 8> kc;
  kc.init(segs, tk.kseg_begin, tk.kseg_end);
  // This wave's loads: base address (row tile, unit 0), operand and LDS unit of each.  Spare
  // loads (when the units do not split evenly over the waves) repeat unit 0's load (same data,
  // same LDS place), so every wave issues the same count.
  const uint8_t* qb[FPW];
  bool qa[FPW];
  uint32_t ql[FPW];
#pragma unroll
  for (int q = 0; q < FPW; ++q) {
    const int f = wv * FPW + q;
    const bool spare = f >= SF;
    qa[q] = f < AF || spare;
    const int fl = spare ? 0 : (f < AF ? f : f - AF);
    const int rt = min((qa[q] ? rt0 : ct0) + fl, (qa[q] ? art : brt) - 1);
    qb[q] = qa[q] ? A + (int64_t)rt * akt * 512 : B + (int64_t)rt * bkt * 512;
    ql[q] = (uint32_t)(spare ? 0 : f) * 1024u;
  }
  const uint32_t lane16 = (uint32_t)lane * 16u;
  uint8_t* const lds0 = reinterpret_cast<uint8_t*>(&lds[0][0]);
  auto dma = [&](int q, uint32_t slot_off, int64_t ca, int64_t cb) {
    const uint8_t* base = qb[q] + (qa[q] ? ca : cb);
    uint32_t lo = lane16;
    asm volatile("" : "+v"(lo));  // per-lane offset stays a fresh VGPR (scalar-base addressing)
    lds_dma16(base + lo, (__attribute__((address_space(3))) void*)(lds0 + slot_off + ql[q]));
  };
  d4 acc[WTM][WTN];
#pragma unroll
  for (int i = 0; i < WTM; ++i)
#pragma unroll
    for (int j = 0; j < WTN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  double fa[2][WTM], fb[2][WTN];
  const int abase = wm * WTM * 128 + lane, bbase = (AF + wn * WTN) * 128 + lane;
#define TMX_RD(BUF, SLOT_, KK)                                                                     \
  {                                                                                               \
    const double* L = reinterpret_cast<const double*>(&lds[SLOT_][0]);                            \
    _Pragma("unroll") for (int i = 0; i < WTM; ++i) fa[BUF][i] = L[abase + i * 128 + (KK) * 64]; \
    _Pragma("unroll") for (int j = 0; j < WTN; ++j) fb[BUF][j] = L[bbase + j * 128 + (KK) * 64]; \
  }
#define TMX_MM1(BUF, I, J) \
  acc[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[BUF][I], fb[BUF][J], acc[I][J], 0, 0, 0);
#define TMX_MM_REST(BUF)                                                                    \
  _Pragma("unroll") for (int i = 0; i < WTM; ++i)                                           \
  _Pragma("unroll") for (int j = 0; j < WTN; ++j)                                           \
      if (i + j > 0) { TMX_MM1(BUF, i, j) }
#pragma unroll
  for (int s0 = 0; s0 < NS; ++s0) {
    const int64_t ca = (int64_t)kc.unit_a() * 1024, cb = (int64_t)kc.unit_b() * 1024;
#pragma unroll
    for (int q = 0; q < FPW; ++q) dma(q, (uint32_t)s0 * SLOT * 1024u, ca, cb);
    kc.advance();
  }
  asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"((NS - 1) * FPW) : "memory");
  TMX_RD(0, 0, 0)
  constexpr int NMM = WTM * WTN - 1;  // MFMAs of a k-step after its first
  constexpr int NIT = FPW + 2;        // interleaved items: offsets, fragment reads, loads
  constexpr int NSTEP = NMM > NIT ? NMM : NIT;
  for (int st = 0; st < nst; ++st) {
    const int slot = st % NS, nslot = (st + 1) % NS;
    // k-step 0 of stage st (fragments in buffer 0); k-step 1's reads follow its first MFMA
    TMX_MM1(0, 0, 0)
    __builtin_amdgcn_sched_barrier(0);
    TMX_RD(1, slot, 1)
    __builtin_amdgcn_sched_barrier(0);
    TMX_MM_REST(0)
    __builtin_amdgcn_sched_barrier(0);
    // k-step 1 (buffer 1); every read of stage st is issued, so after the barrier its slot
    // takes stage st+NS
    TMX_MM1(1, 0, 0)
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"((NS - 2) * FPW)
                 : "memory");
    __builtin_amdgcn_sched_barrier(0);
    // the rest of k-step 1, one MFMA before each of: the load offsets of stage st+NS, the
    // fragment reads of stage st+1's first k-step, each load of stage st+NS (the last one
    // followed by the cursor step)
    int64_t ca = 0, cb = 0;
#pragma unroll
    for (int t = 0; t < NSTEP; ++t) {
      if (t < NMM) { TMX_MM1(1, (t + 1) / WTN, (t + 1) % WTN) }
      __builtin_amdgcn_sched_barrier(0);
      if (t == 0) {
        ca = (int64_t)kc.unit_a() * 1024;
        cb = (int64_t)kc.unit_b() * 1024;
      } else if (t == 1) {
        TMX_RD(0, nslot, 0)
      } else if (t - 2 < FPW) {
        dma(t - 2, (uint32_t)slot * SLOT * 1024u, ca, cb);
        if (t - 2 == FPW - 1) kc.advance();
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#undef TMX_RD
#undef TMX_MM1
#undef TMX_MM_REST
  // No LDS-DMA may still be writing when the workgroup retires.
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#pragma unroll
  for (int i = 0; i < WTM; ++i)
#pragma unroll
    for (int j = 0; j < WTN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        // v_mfma_f64_16x16x4f64 result: register e of lane l holds row 4e + l/16, col l%16.
        const int row = tk.m0 + 16 * (wm * WTM + i) + 4 * e + (lane >> 4);
        const int col = tk.n0 + 16 * (wn * WTN + j) + (lane & 15);
        if (row < tk.mrows && col < tk.nrows) Cd[tk.c_base + (int64_t)row * tk.ldc + col] = acc[i][j][e];
      }
}

// ---------------------------------------------------------------------------
// int8 sector GEMM, one residue plane (modulus) per task.  Workgroup tile TM x TN over NW = WGM x WGN
// waves (wave tile (TM/WGM) x (TN/WGN) of 32x32 v_mfma_i32_32x32x32_i8 tiles).  The K range
// of a task (its list of K segments) is one stream of 32-column k-steps; a stage is two
// k-steps, i.e. (TM + TN) / 32 * 2 fragments of 1 KB, copied global -> LDS by
// global_load_lds_dwordx4 (one fragment per wave instruction) into a ring of NS stages.
// Stage st+NS is issued as soon as every wave has finished reading stage st (one barrier per
// stage), so NS-1 stages of loads are in flight under the MFMA work.  A missing second k-step
// at the end of the stream reads an all-zero fragment, which adds nothing.  Every wave issues
// the same number of loads per stage, so the vmcnt thresholds are constants.  Fragment reads
// from LDS are double-buffered in registers: the reads for the next k-step go out right after
// the first MFMA of the current one.  The epilogue reduces the exact int32 sums modulo m_l and
// stores each 32x32 tile in the MFMA-native order (lane l's 16 result registers as 16
// consecutive bytes), one 16-byte store per tile and lane.

template <int TM, int TN, int WGM, int WGN, int NS>
__global__ __launch_bounds__(64 * WGM * WGN, 2) void k_gemm_i8(
    const SecTask* __restrict__ tasks, int ntasks, const KSeg* __restrict__ segs,
    const uint8_t* __restrict__ Ap, int64_t aplane, const uint8_t* __restrict__ Bp, int64_t bplane,
    uint8_t* __restrict__ Cr, int64_t cplane, CrtConsts cc) {
  constexpr int NW = WGM * WGN;
  constexpr int AF = (TM / 32) * 2, BF = (TN / 32) * 2, SF = AF + BF;
  // Fragments per wave per stage; when SF does not split evenly the spare loads repeat the
  // load of fragment 0 (same data, same LDS place), so every wave issues the same count.
  constexpr int FPW = (SF + NW - 1) / NW;
  constexpr int SLOT = SF;
  constexpr int WTM = TM / 32 / WGM, WTN = TN / 32 / WGN;
  static_assert(WTM * 32 * WGM == TM && WTN * 32 * WGN == TN, "tile must split over the waves");
  __shared__ i4 lds[NS][SLOT * 64];
  const int task = xcd_task(ntasks);
  if (task >= ntasks) return;
  const SecTask tk = tasks[task];
  if (tk.mrows == 0) return;  // padding of the XCD layout
  const int l = tk.plane;
  const uint8_t* A = Ap + (int64_t)l * aplane + tk.a_base;
  const uint8_t* B = Bp + (int64_t)l * bplane + tk.b_base;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  const int wm = wave % WGM, wn = wave / WGM;
  const int art = tk.mrows >> 5, brt = tk.nrows >> 5;
  const int64_t akt = tk.a_kpad >> 5, bkt = tk.b_kpad >> 5;
  const int rt0 = tk.m0 >> 5, ct0 = tk.n0 >> 5;
  int nst = 0;
  for (int sg = tk.kseg_begin; sg < tk.kseg_end; ++sg) nst += (int)((uint32_t)segs[sg].len >> 6);
  StageCursor<64, 32> kc;
  kc.init(segs, tk.kseg_begin, tk.kseg_end);
  // This wave's loads: base address (row tile and k-step of the stage), operand and LDS
  // fragment of each.  Spare loads (when the fragments do not split evenly over the waves)
  // repeat fragment 0's load.
  const uint8_t* qb[FPW];
  bool qa[FPW];
  uint32_t ql[FPW];
#pragma unroll
  for (int q = 0; q < FPW; ++q) {
    const int f = wv * FPW + q;
    const bool spare = f >= SF;
    qa[q] = f < AF || spare;
    const int fl = spare ? 0 : (f < AF ? f : f - AF);
    const int rt = min((qa[q] ? rt0 : ct0) + (fl >> 1), (qa[q] ? art : brt) - 1);
    qb[q] = (qa[q] ? A + (int64_t)rt * akt * 1024 : B + (int64_t)rt * bkt * 1024) + (fl & 1) * 1024;
    ql[q] = (uint32_t)(spare ? 0 : f) * 1024u;
  }
  const uint32_t lane16 = (uint32_t)lane * 16u;
  uint8_t* const lds0 = reinterpret_cast<uint8_t*>(&lds[0][0]);
  auto dma = [&](int q, uint32_t slot_off, int64_t ca, int64_t cb) {
    const uint8_t* base = qb[q] + (qa[q] ? ca : cb);
    uint32_t lo = lane16;
    asm volatile("" : "+v"(lo));  // per-lane offset stays a fresh VGPR (scalar-base addressing)
    lds_dma16(base + lo, (__attribute__((address_space(3))) void*)(lds0 + slot_off + ql[q]));
  };
  i16v acc[WTM][WTN];
#pragma unroll
  for (int i = 0; i < WTM; ++i)
#pragma unroll
    for (int j = 0; j < WTN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0;
  i4 fa[2][WTM], fb[2][WTN];
  const int abase = (wm * WTM * 2) * 64 + lane, bbase = (AF + wn * WTN * 2) * 64 + lane;
#define TMX_RD(BUF, SLOT_, KK)                                                               \
  _Pragma("unroll") for (int i = 0; i < WTM; ++i) fa[BUF][i] = lds[SLOT_][abase + (i * 2 + (KK)) * 64]; \
  _Pragma("unroll") for (int j = 0; j < WTN; ++j) fb[BUF][j] = lds[SLOT_][bbase + (j * 2 + (KK)) * 64];
#define TMX_MM1(BUF, I, J) \
  acc[I][J] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[BUF][I], fb[BUF][J], acc[I][J], 0, 0, 0);
#define TMX_MM_REST(BUF)                                                                    \
  _Pragma("unroll") for (int i = 0; i < WTM; ++i)                                           \
  _Pragma("unroll") for (int j = 0; j < WTN; ++j)                                           \
      if (i + j > 0) { TMX_MM1(BUF, i, j) }
#pragma unroll
  for (int s0 = 0; s0 < NS; ++s0) {
    const int64_t ca = (int64_t)kc.unit_a() * 1024, cb = (int64_t)kc.unit_b() * 1024;
#pragma unroll
    for (int q = 0; q < FPW; ++q) dma(q, (uint32_t)s0 * SLOT * 1024u, ca, cb);
    kc.advance();
  }
  asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"((NS - 1) * FPW) : "memory");
  TMX_RD(0, 0, 0)
  constexpr int NMM = WTM * WTN - 1;  // MFMAs of a k-step after its first
  constexpr int NIT = FPW + 2;        // interleaved items: offsets, fragment reads, loads
  constexpr int NSTEP = NMM > NIT ? NMM : NIT;
  for (int st = 0; st < nst; ++st) {
    const int slot = st % NS, nslot = (st + 1) % NS;
    // k-step 0 of stage st (fragments in buffer 0); k-step 1's reads follow its first MFMA
    TMX_MM1(0, 0, 0)
    __builtin_amdgcn_sched_barrier(0);
    TMX_RD(1, slot, 1)
    __builtin_amdgcn_sched_barrier(0);
    TMX_MM_REST(0)
    __builtin_amdgcn_sched_barrier(0);
    // k-step 1 (buffer 1); every read of stage st is issued, so after the barrier its slot
    // takes stage st+NS
    TMX_MM1(1, 0, 0)
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"((NS - 2) * FPW)
                 : "memory");
    __builtin_amdgcn_sched_barrier(0);
    // the rest of k-step 1, one MFMA before each of: the load offsets of stage st+NS, the
    // fragment reads of stage st+1's first k-step, each load of stage st+NS (the last one
    // followed by the cursor step)
    int64_t ca = 0, cb = 0;
#pragma unroll
    for (int t = 0; t < NSTEP; ++t) {
      if (t < NMM) { TMX_MM1(1, (t + 1) / WTN, (t + 1) % WTN) }
      __builtin_amdgcn_sched_barrier(0);
      if (t == 0) {
        ca = (int64_t)kc.unit_a() * 1024;
        cb = (int64_t)kc.unit_b() * 1024;
      } else if (t == 1) {
        TMX_RD(0, nslot, 0)
      } else if (t - 2 < FPW) {
        dma(t - 2, (uint32_t)slot * SLOT * 1024u, ca, cb);
        if (t - 2 == FPW - 1) kc.advance();
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#undef TMX_RD
#undef TMX_MM1
#undef TMX_MM_REST
  // No LDS-DMA may still be writing when the workgroup retires.
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  const double md = (double)cc.m[l], rm = cc.rcp[l];
  uint8_t* Cp = Cr + (int64_t)l * cplane + tk.c_base;
  const int ntc = tk.ldc >> 5;
#pragma unroll
  for (int i = 0; i < WTM; ++i)
#pragma unroll
    for (int j = 0; j < WTN; ++j) {
      const int tr = rt0 + wm * WTM + i, tc = ct0 + wn * WTN + j;
      if (tr < art && tc < brt) {
        uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          // |acc| < 2^31: the FP64 quotient is exact to far below the rounding margin, so the
          // residue lies in [-(m-1)/2, (m-1)/2] (odd m) or [-128, 128] (m = 256; 128 and -128
          // share the byte 0x80 and are congruent).
          const double xd = (double)acc[i][j][e];
          const int r = (int)fma(-rint(xd * rm), md, xd);
          w[e >> 2] |= ((uint32_t)r & 0xffu) << (8 * (e & 3));
        }
        *reinterpret_cast<i4*>(Cp + ((int64_t)tr * ntc + tc) * 1024 + lane * 16) =
            i4{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
      }
    }
}

// ---------------------------------------------------------------------------
// CRT recombination: residue planes (32x32-tile native) -> dense FP64 result (row-major per
// sector, ld = nrows_pad), unscaled by the row scales of A and B.  A wave handles one tile;
// each lane reads its 16 bytes from every plane and reconstructs 16 elements.
__global__ __launch_bounds__(256) void k_crt(const CrtSector* __restrict__ secs, int nsec,
                                             int64_t ntiles, const int8_t* __restrict__ Cr,
                                             int64_t cplane, CrtConsts cc,
                                             const double2* __restrict__ invA,
                                             const double2* __restrict__ invB,
                                             double* __restrict__ Cd) {
  const int lane = threadIdx.x & 63;
  const double L1 = 0x1p42, L2 = 0x1p84, iL1 = 0x1p-42;
  for (int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < ntiles;
       t += (int64_t)gridDim.x * 4) {
    // sector of tile t (sectors are few; linear scan over cumulative tile counts)
    int s = 0;
    int64_t t0 = 0;
    while (s + 1 < nsec && t >= t0 + secs[s].tiles) { t0 += secs[s].tiles; ++s; }
    const CrtSector sc = secs[s];
    const int64_t tl = t - t0;
    const int tr = (int)(tl / sc.ntc), tc = (int)(tl - (int64_t)tr * sc.ntc);
    const int64_t boff = sc.c_base + tl * 1024 + lane * 16;
    double S0[16], S1[16], S2[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) S0[e] = S1[e] = S2[e] = 0.0;
    for (int k = 0; k < cc.n; ++k) {
      const i4 q = *reinterpret_cast<const i4*>(Cr + (int64_t)k * cplane + boff);
      const double w0 = cc.w[k][0], w1 = cc.w[k][1], w2 = cc.w[k][2];
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const double r = (double)(int)(int8_t)((uint32_t)q[e >> 2] >> (8 * (e & 3)));
        S0[e] = fma(r, w0, S0[e]);
        S1[e] = fma(r, w1, S1[e]);
        S2[e] = fma(r, w2, S2[e]);
      }
    }
    const int col = tc * 32 + (lane & 31);
    const double2 sb2 = invB[sc.bgrow0 + col];
    const double sb = sb2.x;
    const int sbe = inv_exp(sb2);
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      // S = S2*2^84 + S1*2^42 + S0 (exact limbs); q = round(S / M); x = S - q*M, exact per limb.
      const double D = fma(S2[e], L2, fma(S1[e], L1, S0[e]));
      const double qq = rint(D * cc.inv_m_total);
      double x0 = fma(-qq, cc.M[0], S0[e]);
      double x1 = fma(-qq, cc.M[1], S1[e]);
      double x2 = fma(-qq, cc.M[2], S2[e]);
      const double c0 = rint(x0 * iL1);
      x0 = fma(-c0, L1, x0);
      x1 += c0;
      const double c1 = rint(x1 * iL1);
      x1 = fma(-c1, L1, x1);
      x2 += c1;
      const double x = fma(x2, L2, fma(x1, L1, x0));
      const int row = tr * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
      Cd[sc.c_base + (int64_t)row * sc.ldc + col] = unscale(x, invA[sc.agrow0 + row], sb, sbe);
    }
  }
}

// ---------------------------------------------------------------------------
// Fused CRT + scatter: residue planes -> C blocks without the dense FP64 intermediate.  A wave
// takes one 32x32 tile of a pair's dense result: each lane reads its 16 bytes from every plane
// of every split-K part, reconstructs its 16 elements as k_crt does, unscales them and sums the
// parts (in part order); then alpha times the sum is added into C, as k_crt + k_scatter would.
// Lane l holds column l % 32 of the tile and rows 8g + 4(l / 32) + e (g, e = 0..3), element
// 4g + e.  Each element finds its C block from its row's and column's tuples (tiles may
// straddle blocks; padding rows and columns and zero blocks are skipped), looking the block up
// only when a tuple changes.  Per element: nmod bytes read per part + the C read-modify-write.
template <int NM>  // moduli count, or 0 for the count in cc (runtime loop)
__device__ __forceinline__ void crt_tile_values(const FusedTile& ft, const CrtPair& pr, int lane,
                                                const int8_t* __restrict__ Cr, int64_t cplane,
                                                const CrtConsts& cc, const double2* __restrict__ invA,
                                                double2 sb2, double v[16]) {
  const double sb = sb2.x;
  const int sbe = inv_exp(sb2);
  const double L1 = 0x1p42, L2 = 0x1p84, iL1 = 0x1p-42;
  const int64_t toff = ((int64_t)ft.tr * pr.ntc + ft.tc) * 1024 + lane * 16;
  // Reconstruction of one element from its limb sums over the planes: the exact reduction mod
  // M (as k_crt).
  auto finish = [&](double s0, double s1, double s2) {
    const double D = fma(s2, L2, fma(s1, L1, s0));
    const double qq = rint(D * cc.inv_m_total);
    double x0 = fma(-qq, cc.M[0], s0);
    double x1 = fma(-qq, cc.M[1], s1);
    double x2 = fma(-qq, cc.M[2], s2);
    const double c0 = rint(x0 * iL1);
    x0 = fma(-c0, L1, x0);
    x1 += c0;
    const double c1 = rint(x1 * iL1);
    x1 = fma(-c1, L1, x1);
    x2 += c1;
    return fma(x2, L2, fma(x1, L1, x0));
  };
#pragma unroll
  for (int e = 0; e < 16; ++e) v[e] = 0.0;
  if constexpr (NM > 0) {
    // moduli count known: a part's 16 bytes per plane are loaded once into registers
    for (int p = 0; p < pr.nparts; ++p) {
      const int8_t* rp = Cr + pr.c_base + p * pr.pstride + toff;
      i4 q[NM];
#pragma unroll
      for (int k = 0; k < NM; ++k) q[k] = *reinterpret_cast<const i4*>(rp + (int64_t)k * cplane);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        double S0[4], S1[4], S2[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) S0[e] = S1[e] = S2[e] = 0.0;
#pragma unroll
        for (int k = 0; k < NM; ++k) {
          const double w0 = cc.w[k][0], w1 = cc.w[k][1], w2 = cc.w[k][2];
          const uint32_t word = (uint32_t)q[k][g];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const double r = (double)(int)(int8_t)(word >> (8 * e));
            S0[e] = fma(r, w0, S0[e]);
            S1[e] = fma(r, w1, S1[e]);
            S2[e] = fma(r, w2, S2[e]);
          }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = ft.tr * 32 + 8 * g + 4 * (lane >> 5) + e;
          v[4 * g + e] += unscale(finish(S0[e], S1[e], S2[e]), invA[pr.agrow0 + row], sb, sbe);
        }
      }
    }
  } else {
    // runtime moduli count: group by group, one 4-byte residue word per plane
    for (int g = 0; g < 4; ++g)
      for (int p = 0; p < pr.nparts; ++p) {
        const int8_t* rp = Cr + pr.c_base + p * pr.pstride + toff + 4 * g;
        double S0[4], S1[4], S2[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) S0[e] = S1[e] = S2[e] = 0.0;
        for (int k = 0; k < cc.n; ++k) {
          const double w0 = cc.w[k][0], w1 = cc.w[k][1], w2 = cc.w[k][2];
          const uint32_t word = *reinterpret_cast<const uint32_t*>(rp + (int64_t)k * cplane);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const double r = (double)(int)(int8_t)(word >> (8 * e));
            S0[e] = fma(r, w0, S0[e]);
            S1[e] = fma(r, w1, S1[e]);
            S2[e] = fma(r, w2, S2[e]);
          }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = ft.tr * 32 + 8 * g + 4 * (lane >> 5) + e;
          v[4 * g + e] += unscale(finish(S0[e], S1[e], S2[e]), invA[pr.agrow0 + row], sb, sbe);
        }
      }
  }
}

// C += alpha * v (assign: C = alpha * v) at four addresses (negative: skipped): the four loads,
// then the four stores, so their latencies overlap.
__device__ __forceinline__ void add4(double* __restrict__ C, const int64_t* addr, const double* v,
                                     double alpha, bool assign) {
  double old[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) old[e] = addr[e] >= 0 && !assign ? C[addr[e]] : 0.0;
#pragma unroll
  for (int e = 0; e < 4; ++e)
    if (addr[e] >= 0) C[addr[e]] = old[e] + alpha * v[e];
}

// Natural layout: C's labels are the A-outer labels followed by the B-outer labels, so a C
// block is row-major over the dense rows and columns of its tuples and an element sits at
// row-within-tuple * block columns + column-within-tuple.  For each register, lanes 0-31 and
// 32-63 write 32 consecutive columns of two rows.
template <int NM>
__global__ __launch_bounds__(256) void k_crt_scatter(const FusedTile* __restrict__ tiles, int64_t ntiles,
                                                     const CrtPair* __restrict__ pairs,
                                                     const ScatterJob* __restrict__ jobs,
                                                     const int2* __restrict__ ainfo,
                                                     const int2* __restrict__ binfo,
                                                     const int32_t* __restrict__ jobtab, int jobtab_n,
                                                     const int8_t* __restrict__ Cr, int64_t cplane,
                                                     CrtConsts cc, const double2* __restrict__ invA,
                                                     const double2* __restrict__ invB,
                                                     double* __restrict__ C, double alpha, int assign) {
  const int lane = threadIdx.x & 63;
  for (int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < ntiles; t += (int64_t)gridDim.x * 4) {
    const FusedTile ft = tiles[t];
    const CrtPair pr = pairs[ft.pair];
    const int col = ft.tc * 32 + (lane & 31);  // dense column of this lane
    if (ft.job >= 0) {
      // The whole tile lies in one C block: addresses need no lookups, and the C loads go
      // out before the residue loads so both latencies overlap.
      const int ncols = jobs[ft.job].cols;
      const int64_t base = jobs[ft.job].c_off + (int64_t)ft.rin * ncols + ft.cin + (lane & 31);
      double old[16], v[16];
#pragma unroll
      for (int e = 0; e < 16; ++e)
        old[e] = assign ? 0.0 : C[base + (int64_t)(8 * (e >> 2) + 4 * (lane >> 5) + (e & 3)) * ncols];
      crt_tile_values<NM>(ft, pr, lane, Cr, cplane, cc, invA, invB[pr.bgrow0 + col], v);
#pragma unroll
      for (int e = 0; e < 16; ++e)
        C[base + (int64_t)(8 * (e >> 2) + 4 * (lane >> 5) + (e & 3)) * ncols] = old[e] + alpha * v[e];
      continue;
    }
    const int2 bi = binfo[pr.bgrow0 + col];     // (B tuple, column within it), -1 for padding
    double v[16];
    crt_tile_values<NM>(ft, pr, lane, Cr, cplane, cc, invA, invB[pr.bgrow0 + col], v);
    if (bi.x < 0) continue;
    int cur_rt = -1, jn = -1, ncols = 0;
    int64_t base = 0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      int64_t addr[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = ft.tr * 32 + 8 * g + 4 * (lane >> 5) + e;
        const int2 ai = ainfo[pr.agrow0 + row];
        addr[e] = -1;
        if (ai.x < 0) continue;
        if (ai.x != cur_rt) {
          cur_rt = ai.x;
          jn = jobtab[(int64_t)ai.x * jobtab_n + bi.x];
          if (jn >= 0) {
            base = jobs[jn].c_off + bi.y;
            ncols = jobs[jn].cols;
          }
        }
        if (jn >= 0) addr[e] = base + (int64_t)ai.y * ncols;
      }
      add4(C, addr, v + 4 * g, alpha, assign);
    }
  }
}

// Offset of a tuple element inside a C block: its digits (8 bits per mode, mode d at bits 8d)
// times the block's strides of those modes.
__device__ __forceinline__ int64_t digit_off(uint32_t dig, const Radix& r) {
  int64_t o = 0;
#pragma unroll
  for (int d = 0; d < 4; ++d)
    if (d < r.n) o += (int64_t)((dig >> (8 * d)) & 255u) * r.stride[d];
  return o;
}

// Other label orders (A-outer and B-outer labels interleaved in C, e.g. R(a,b,i,j) from
// W(k,b,c,j) T(a,c,i,k) with rows (b,j) and columns (a,i)): an element's offset in its block is
// digit_off of its row digits (adig) and column digits (bdig) with the block's row and column
// strides.  ROWC = false: C is contiguous along the dense columns, lanes keep their column.
// ROWC = true: C is contiguous along the dense rows; the tile goes through LDS so that lane l
// takes row l % 32 and columns 16(l / 32) .. +15, and 32 lanes write 32 consecutive rows.
// Every wave of a workgroup runs the same number of iterations (the transpose's barriers).
// Tiles inside one C block (FusedTile::job >= 0) take the block's strides directly.
template <int NM, bool ROWC>
__global__ __launch_bounds__(256, 3) void k_crt_scatter_g(const FusedTile* __restrict__ tiles, int64_t ntiles,
                                                       const CrtPair* __restrict__ pairs,
                                                       const ScatterJob* __restrict__ jobs,
                                                       const int2* __restrict__ ainfo,
                                                       const int2* __restrict__ binfo,
                                                       const uint32_t* __restrict__ adig,
                                                       const uint32_t* __restrict__ bdig,
                                                       const int32_t* __restrict__ jobtab, int jobtab_n,
                                                       const int8_t* __restrict__ Cr, int64_t cplane,
                                                       CrtConsts cc, const double2* __restrict__ invA,
                                                       const double2* __restrict__ invB,
                                                       double* __restrict__ C, double alpha, int assign) {
  __shared__ double xt[ROWC ? 4 : 1][32][33];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t t0 = (int64_t)blockIdx.x * 4; t0 < ntiles; t0 += (int64_t)gridDim.x * 4) {
    const int64_t t = t0 + w;
    const bool live = t < ntiles;
    FusedTile ft{};
    CrtPair pr{};
    double v[16];
    if (live) {
      ft = tiles[t];
      pr = pairs[ft.pair];
      const int col = ft.tc * 32 + (lane & 31);
      crt_tile_values<NM>(ft, pr, lane, Cr, cplane, cc, invA, invB[pr.bgrow0 + col], v);
    }
    if constexpr (!ROWC) {
      if (!live) continue;
      const int col = ft.tc * 32 + (lane & 31);
      if (ft.job >= 0) {
        // the whole tile lies in one C block: its strides serve every element
        const int64_t base = jobs[ft.job].c_off + digit_off(bdig[pr.bgrow0 + col], jobs[ft.job].c);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          int64_t addr[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = ft.tr * 32 + 8 * g + 4 * (lane >> 5) + e;
            addr[e] = base + digit_off(adig[pr.agrow0 + row], jobs[ft.job].r);
          }
          add4(C, addr, v + 4 * g, alpha, assign);
        }
        continue;
      }
      const int2 bi = binfo[pr.bgrow0 + col];
      if (bi.x < 0) continue;
      const uint32_t bd = bdig[pr.bgrow0 + col];
      int cur_rt = -1, jn = -1;
      int64_t base = 0;
      Radix rr{};
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        int64_t addr[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = ft.tr * 32 + 8 * g + 4 * (lane >> 5) + e;
          const int2 ai = ainfo[pr.agrow0 + row];
          addr[e] = -1;
          if (ai.x < 0) continue;
          if (ai.x != cur_rt) {
            cur_rt = ai.x;
            jn = jobtab[(int64_t)ai.x * jobtab_n + bi.x];
            if (jn >= 0) {
              rr = jobs[jn].r;
              base = jobs[jn].c_off + digit_off(bd, jobs[jn].c);
            }
          }
          if (jn >= 0) addr[e] = base + digit_off(adig[pr.agrow0 + row], rr);
        }
        add4(C, addr, v + 4 * g, alpha, assign);
      }
    } else {
      if (live) {
#pragma unroll
        for (int e = 0; e < 16; ++e) xt[w][8 * (e >> 2) + 4 * (lane >> 5) + (e & 3)][lane & 31] = v[e];
      }
      __syncthreads();
      if (live && ft.job >= 0) {
        // the whole tile lies in one C block: its strides serve every element
        const int rloc = lane & 31, c0 = 16 * (lane >> 5);
        const int64_t base =
            jobs[ft.job].c_off + digit_off(adig[pr.agrow0 + ft.tr * 32 + rloc], jobs[ft.job].r);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          int64_t addr[4];
          double x[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int cl = c0 + 4 * g + e;
            x[e] = xt[w][rloc][cl];
            addr[e] = base + digit_off(bdig[pr.bgrow0 + ft.tc * 32 + cl], jobs[ft.job].c);
          }
          add4(C, addr, x, alpha, assign);
        }
      } else if (live) {
        const int rloc = lane & 31, c0 = 16 * (lane >> 5);
        const int row = ft.tr * 32 + rloc;
        const int2 ai = ainfo[pr.agrow0 + row];
        if (ai.x >= 0) {
          const uint32_t ad = adig[pr.agrow0 + row];
          int cur_ct = -1, jn = -1;
          int64_t base = 0;
          Radix cr{};
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            int64_t addr[4];
            double x[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int cl = c0 + 4 * g + e;
              const int col = ft.tc * 32 + cl;
              const int2 bi = binfo[pr.bgrow0 + col];
              x[e] = xt[w][rloc][cl];
              addr[e] = -1;
              if (bi.x < 0) continue;
              if (bi.x != cur_ct) {
                cur_ct = bi.x;
                jn = jobtab[(int64_t)ai.x * jobtab_n + bi.x];
                if (jn >= 0) {
                  cr = jobs[jn].c;
                  base = jobs[jn].c_off + digit_off(ad, jobs[jn].r);
                }
              }
              if (jn >= 0) addr[e] = base + digit_off(bdig[pr.bgrow0 + col], cr);
            }
            add4(C, addr, x, alpha, assign);
          }
        }
      }
      __syncthreads();  // the next iteration rewrites xt
    }
  }
}

// ---------------------------------------------------------------------------
// Scatter: dense FP64 sector result -> C block (C's mode order), C += alpha * value, where the
// value is the sum of the sector's split-K parts (in part order).  A workgroup iteration
// handles a 64x64 tile: values are read with lanes along the dense columns (coalesced), staged
// in LDS, and added into C with lanes along C's contiguous direction.
__global__ __launch_bounds__(256) void k_scatter(const ScatterJob* __restrict__ jobs,
                                                 const TileUnit* __restrict__ units, int64_t nunits,
                                                 const double* __restrict__ Cd,
                                                 double* __restrict__ C, double alpha, int assign) {
  __shared__ double tile[64][65];
  __shared__ int roff[64], coff[64];
  const int tid = threadIdx.x;
  for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const TileUnit un = units[u];
    const ScatterJob& jb = jobs[un.job];
    const int nr = min(64, jb.rows - un.r0), nc = min(64, jb.cols - un.c0);
    if (tid < 64) roff[tid] = tid < nr ? radix_off(jb.r, un.r0 + tid) : 0;
    else if (tid < 128) coff[tid - 64] = (tid - 64) < nc ? radix_off(jb.c, un.c0 + tid - 64) : 0;
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
      const int rr = (tid >> 6) + 4 * i, cc = tid & 63;
      double v = 0.0;
      if (rr < nr && cc < nc) {
        const double* src = Cd + jb.src_base + (int64_t)(jb.row0 + un.r0 + rr) * jb.ldc + jb.col0 + un.c0 + cc;
        for (int p = 0; p < jb.nparts; ++p) v += src[(int64_t)p * jb.pstride];
      }
      tile[rr][cc] = alpha * v;
    }
    __syncthreads();
    double* cb = C + jb.c_off;
    if (jb.ccontig) {
      for (int i = 0; i < 16; ++i) {
        const int rr = (tid >> 6) + 4 * i, cc = tid & 63;
        if (rr < nr && cc < nc) cb[roff[rr] + coff[cc]] = (assign ? 0.0 : cb[roff[rr] + coff[cc]]) + tile[rr][cc];
      }
    } else {
      for (int i = 0; i < 16; ++i) {
        const int rr = tid & 63, cc = (tid >> 6) + 4 * i;
        if (rr < nr && cc < nc) cb[roff[rr] + coff[cc]] = (assign ? 0.0 : cb[roff[rr] + coff[cc]]) + tile[rr][cc];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Host side
// GEMM tile shapes and their relative MFMA efficiency (int8: measured on MI355X with the
// ladder and square shapes; wider tiles need fewer LDS and L2 bytes per MFMA).
// small > 0: offered only to sectors with at most `small` rows on one side.  64x64 serves
// products with few rows or columns over a long K (e.g. Fo(k,i) = U(k,l,c,d) T(c,d,i,l) with
// 60x60 outputs), where larger tiles would stream padding rows of A or B; on large sectors it
// can time faster alone yet slow the plan down as a launch of its own.  FP64 32x32 serves
// classes of one O tile (16 padded rows, e.g. that product's 8 x 8 sectors), which a 64-row
// tile would load four times over.
struct I8Shape { int tm, tn; double eff; int small; };
static const I8Shape kF64Shapes[] = {{128, 128, 1.0, 0}, {128, 160, 1.0, 0}, {64, 64, 0.5, 64}, {32, 32, 0.25, 32}};
static const I8Shape kI8Shapes[] = {{256, 256, 1.0, 0}, {512, 160, 1.0, 0}, {256, 160, 0.95, 0},
                                    {256, 128, 0.82, 0}, {128, 256, 0.82, 0}, {128, 128, 0.6, 0},
                                    {64, 64, 0.3, 64}};
static bool shape_allowed(const I8Shape& sh, int64_t mrows, int64_t nrows) {
  return sh.small == 0 || std::min(mrows, nrows) <= sh.small;
}

// Autotuner cache: sector geometry -> tile shape measured fastest.  Kept in memory and, when
// TAMM_MI355X_TUNE_FILE names a file, loaded from and saved to that file (one entry per line:
// "i8 nmod mrows_pad nrows_pad klen tm tn ms").
namespace {
struct TuneKey {
  int split, nmod;
  int64_t m, n, k;
  bool operator<(const TuneKey& o) const {
    return std::tie(split, nmod, m, n, k) < std::tie(o.split, o.nmod, o.m, o.n, o.k);
  }
};
struct TuneVal { int tm, tn; double ms; };
std::mutex& tune_mu() { static std::mutex* m = new std::mutex; return *m; }
std::map<TuneKey, TuneVal>& tune_map() {
  static auto* m = new std::map<TuneKey, TuneVal>;
  return *m;
}
const char* tune_file() { return std::getenv("TAMM_MI355X_TUNE_FILE"); }
void tune_load_locked() {
  static bool loaded = false;
  if (loaded) return;
  loaded = true;
  const char* f = tune_file();
  if (!f) return;
  std::ifstream in(f);
  std::string tag;
  TuneKey k;
  TuneVal v;
  while (in >> tag >> k.nmod >> k.m >> k.n >> k.k >> v.tm >> v.tn >> v.ms) {
    k.split = tag == "i8" ? 1 : 0;
    tune_map()[k] = v;
  }
}
void tune_save_locked() {
  const char* f = tune_file();
  if (!f) return;
  std::ofstream out(f, std::ios::trunc);
  for (auto& kv : tune_map())
    out << (kv.first.split ? "i8" : "f64") << ' ' << kv.first.nmod << ' ' << kv.first.m << ' '
        << kv.first.n << ' ' << kv.first.k << ' ' << kv.second.tm << ' ' << kv.second.tn << ' '
        << kv.second.ms << '\n';
}
}  // namespace

template <size_t N>
static void cost_model_tile(const I8Shape (&shapes)[N], int64_t mrows, int64_t nrows, int* tm, int* tn) {
  double best = 0;
  for (const I8Shape& sh : shapes) {
    if (!shape_allowed(sh, mrows, nrows)) continue;
    const double work = (double)((mrows + sh.tm - 1) / sh.tm * sh.tm) *
                        (double)((nrows + sh.tn - 1) / sh.tn * sh.tn);
    const double cost = work / sh.eff;
    if (best == 0 || cost < best) { best = cost; *tm = sh.tm; *tn = sh.tn; }
  }
}

void choose_tile(bool split, int nmod, int64_t mrows, int64_t nrows, int64_t klen, int* tm, int* tn) {
  {
    std::lock_guard<std::mutex> g(tune_mu());
    tune_load_locked();
    auto it = tune_map().find(TuneKey{split ? 1 : 0, nmod, mrows, nrows, klen});
    if (it != tune_map().end()) { *tm = it->second.tm; *tn = it->second.tn; return; }
  }
  if (split) cost_model_tile(kI8Shapes, mrows, nrows, tm, tn);
  else cost_model_tile(kF64Shapes, mrows, nrows, tm, tn);
}

// ---------------------------------------------------------------------------
// Launch scheduling.  A GEMM launch's workgroups go round-robin to the 8 XCDs and, on each
// XCD, start in id order as CU slots free up; xcd_task gives XCD x the task range
// [x*per, (x+1)*per).  A split-path task is one residue plane of an output tile.  Tasks are
// laid out so that each XCD's range holds an equal share of the work, longest tasks first,
// with the M panels (the tasks sharing a tile row of A, and for the split path one plane of
// it) kept whole on one XCD so they share its L2; short ranges are padded with empty tasks
// (mrows = 0).  The same list schedule, simulated
// on the host, estimates a launch's duration for choosing the split-K factor.

// int8 GEMM instantiations: (TM, TN, WGM, WGN, NS).  8-wave tiles run one workgroup per CU,
// 4-wave tiles two (the two waves sharing a SIMD then belong to different workgroups and
// reach their stage barriers independently).
#define TMX_I8_KERNELS(X) \
  X(256, 256, 2, 4, 4)    \
  X(512, 160, 8, 1, 3)    \
  X(256, 160, 4, 1, 3)    \
  X(256, 128, 2, 2, 3)    \
  X(128, 256, 2, 2, 3)    \
  X(128, 128, 2, 2, 4)    \
  X(64, 64, 2, 2, 4)

static int gemm_wg_per_cu(bool split, int tm, int tn) {
  static std::mutex mu;
  static std::map<int, int> cache;
  const int key = (split ? 1 : 0) * 1000000 + tm * 1000 + tn;
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const void* fn = nullptr;
  int threads = 256;
  if (!split) {
    if (tm == 128 && tn == 128) fn = reinterpret_cast<const void*>(&k_gemm_f64<128, 128, 4>);
    else if (tm == 128 && tn == 160) fn = reinterpret_cast<const void*>(&k_gemm_f64<128, 160, 4>);
    else if (tm == 64 && tn == 64) fn = reinterpret_cast<const void*>(&k_gemm_f64<64, 64, 4>);
    else if (tm == 32 && tn == 32) fn = reinterpret_cast<const void*>(&k_gemm_f64<32, 32, 8>);
  } else {
#define TMX_I8_FN(TM_, TN_, WGM_, WGN_, NS_)                                             \
  if (!fn && tm == TM_ && tn == TN_) {                                                  \
    fn = reinterpret_cast<const void*>(&k_gemm_i8<TM_, TN_, WGM_, WGN_, NS_>);          \
    threads = 64 * WGM_ * WGN_;                                                         \
  }
    TMX_I8_KERNELS(TMX_I8_FN)
#undef TMX_I8_FN
  }
  TMX_REQUIRE(fn, "no GEMM kernel for the requested tile shape");
  int nb = 0;
  TMX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, threads, 0));
  nb = std::max(nb, 1);
  cache[key] = nb;
  return nb;
}

static double shape_eff(bool split, int tm, int tn) {
  for (const I8Shape& sh : kI8Shapes)
    if (split && sh.tm == tm && sh.tn == tn) return sh.eff;
  for (const I8Shape& sh : kF64Shapes)
    if (!split && sh.tm == tm && sh.tn == tn) return sh.eff;
  return 1.0;
}

// Modelled duration of one task: its multiply-adds (plus a pipeline fill and epilogue worth a
// few k-steps) at one workgroup slot's share of the kernel's sustained rate.
static double task_seconds(bool split, int tm, int tn, int64_t klen, int wgpc) {
  const double macs = (double)tm * tn * (double)(klen + (split ? 256 : 128)) / shape_eff(split, tm, tn);
  const double chip = split ? 1.0e15 : 3.4e13;  // sustained multiply-adds per second
  return macs / (chip / (256.0 * wgpc));
}

// Lays out one launch's tasks (given in panel order) for the XCD mapping when `out` is non-null,
// and returns the simulated makespan in seconds.
static double layout_launch(bool split, int tm, int tn, const std::vector<SecTask>& tasks,
                            const std::vector<int64_t>& klens, std::vector<SecTask>* out) {
  constexpr int kXcds = 8, kCusPerXcd = 32;
  const int wgpc = gemm_wg_per_cu(split, tm, tn);
  struct Group { size_t b, e; double cost, each; };
  std::vector<Group> groups;
  for (size_t i = 0; i < tasks.size();) {
    size_t j = i;
    while (j < tasks.size() && tasks[j].c_base == tasks[i].c_base && tasks[j].m0 == tasks[i].m0 &&
           tasks[j].plane == tasks[i].plane)
      ++j;
    const double each = task_seconds(split, tm, tn, klens[i], wgpc);
    groups.push_back({i, j, each * (double)(j - i), each});
    i = j;
  }
  std::stable_sort(groups.begin(), groups.end(),
                   [](const Group& x, const Group& y) { return x.cost > y.cost; });
  std::vector<std::vector<const Group*>> bins(kXcds);
  std::vector<double> load(kXcds, 0.0);
  for (const Group& g : groups) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    bins[x].push_back(&g);
    load[x] += g.cost;
  }
  double makespan = 0;
  const int slots = kCusPerXcd * wgpc;
  for (int x = 0; x < kXcds; ++x) {
    std::priority_queue<double, std::vector<double>, std::greater<double>> free_at;
    for (int k = 0; k < slots; ++k) free_at.push(0.0);
    double end = 0;
    for (const Group* g : bins[x])
      for (size_t t = g->b; t < g->e; ++t) {
        const double f = free_at.top() + g->each;
        free_at.pop();
        free_at.push(f);
        end = std::max(end, f);
      }
    makespan = std::max(makespan, end);
  }
  if (out) {
    size_t per = 0;
    for (const auto& b : bins) {
      size_t n = 0;
      for (const Group* g : b) n += g->e - g->b;
      per = std::max(per, n);
    }
    SecTask empty;
    std::memset(&empty, 0, sizeof(empty));
    out->assign(per * kXcds, empty);
    for (int x = 0; x < kXcds; ++x) {
      size_t i = per * x;
      for (const Group* g : bins[x])
        for (size_t t = g->b; t < g->e; ++t) (*out)[i++] = tasks[t];
    }
  }
  return makespan;
}

// Tasks of every sector in panel order (split path: plane-major, one copy per residue plane),
// grouped by tile shape (shape id tm*1000 + tn).
static void tasks_by_shape(const std::vector<SectorGeom>& geoms, int planes,
                           std::map<int, std::pair<std::vector<SecTask>, std::vector<int64_t>>>& by,
                           double* exec_flops) {
  for (int pl = 0; pl < planes; ++pl)
  for (const SectorGeom& g : geoms) {
    auto& dst = by[g.tm * 1000 + g.tn];
    for (int64_t m0 = 0; m0 < g.mrows; m0 += g.tm)
      for (int64_t n0 = 0; n0 < g.nrows; n0 += g.tn) {
        SecTask t;
        std::memset(&t, 0, sizeof(t));
        t.a_base = g.a_base;
        t.b_base = g.b_base;
        t.c_base = g.c_base;
        t.m0 = (int32_t)m0;
        t.n0 = (int32_t)n0;
        t.kseg_begin = g.seg_b;
        t.kseg_end = g.seg_e;
        t.a_kpad = g.a_kpad;
        t.b_kpad = g.b_kpad;
        t.ldc = g.nrows_pad;
        t.mrows = g.mrows_pad;
        t.nrows = g.nrows_pad;
        t.plane = pl;
        dst.first.push_back(t);
        dst.second.push_back(g.klen);
        if (exec_flops && pl == 0) *exec_flops += 2.0 * g.tm * g.tn * (double)g.klen;
      }
  }
}

double estimate_sector_seconds(bool split, int nmod, const std::vector<SectorGeom>& geoms) {
  std::map<int, std::pair<std::vector<SecTask>, std::vector<int64_t>>> by;
  tasks_by_shape(geoms, split ? std::max(nmod, 1) : 1, by, nullptr);
  double t = 0;
  for (auto& kv : by)
    t += layout_launch(split, kv.first / 1000, kv.first % 1000, kv.second.first, kv.second.second, nullptr);
  return t;
}

// XCD-aware task lists of a plan, one per tile shape.
void build_sector_tasks(SectorPlan& sp) {
  std::map<int, std::pair<std::vector<SecTask>, std::vector<int64_t>>> by;
  double exec_flops = 0;
  tasks_by_shape(sp.geoms, sp.split ? std::max(sp.nmod, 1) : 1, by, &exec_flops);
  sp.exec_flops = exec_flops;
  sp.shapes.clear();
  for (auto& kv : by) {
    ShapeTasks st;
    st.tm = kv.first / 1000;
    st.tn = kv.first % 1000;
    std::vector<SecTask> laid;
    layout_launch(sp.split, st.tm, st.tn, kv.second.first, kv.second.second, &laid);
    st.ntasks = (int)laid.size();
    st.nreal = (int)kv.second.first.size() / (sp.split ? std::max(sp.nmod, 1) : 1);
    st.tasks = upload(laid);
    sp.shapes.push_back(std::move(st));
  }
}

void sector_schedule_info(const SectorPlan& sp, int* rounds, int* launches, int64_t* tasks) {
  int64_t n = 0;
  for (const ShapeTasks& st : sp.shapes) n += st.nreal;
  if (rounds) *rounds = sp.split_factor;
  if (launches) *launches = (int)sp.shapes.size();
  if (tasks) *tasks = n;
}

static int grid_units(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(n, 8192)); }

static std::atomic<bool> g_scale_reuse{true};  // tmx_set_scale_reuse, read at execution
constexpr int kScaleReuseBackoff = 8;

void set_scale_reuse(bool on) { g_scale_reuse = on; }


// mev: optional two events recorded around the main packing pass alone (without the row-maximum
// pass or the row-scale reuse bookkeeping around it).
static void pack_side(PackSide& side, const double* src, bool split, const SectorPlan& sp,
                      hipStream_t s, hipEvent_t* mev = nullptr) {
  if (side.nunits == 0 && side.nkfunits == 0) return;
  if (side.constant && side.packed_src == src) return;  // packed by an earlier execution
  side.packed_src = side.constant ? src : nullptr;
  auto mark = [&](int i) {
    if (mev) TMX_HIP(hipEventRecord(mev[i], s));
  };
  if (!split) {
    mark(0);
    if (side.nunits > 0)
      hipLaunchKernelGGL(k_pack<0>, dim3(grid_units(side.nunits)), dim3(256), 0, s,
                         side.jobs.as<PackJob>(), side.units.as<TileUnit>(), side.nunits, src,
                         side.buf.as<double>(), nullptr, (int64_t)0, nullptr, nullptr, sp.crt, 0.0,
                         nullptr, nullptr, nullptr);
    if (side.nkfunits > 0)
      hipLaunchKernelGGL(k_pack_kf<0>, dim3(grid_units(side.nkfunits)), dim3(256), 0, s,
                         side.jobs.as<PackJob>(), side.kfunits.as<TileUnit>(), side.nkfunits, src,
                         side.buf.as<double>(), nullptr, (int64_t)0, nullptr, nullptr, sp.crt, 0.0,
                         nullptr, nullptr, nullptr);
    TMX_HIP(hipGetLastError());
    mark(1);
    return;
  }
  const double top = std::ldexp(1.0, sp.bits_a) - 1.0;
  // maxout: recompute the row maxima into it; gate: the gated second pass, whose grid is
  // smaller (it returns at once unless enabled, and its blocks loop over the units);
  // zero_flag: cleared by the first block of the first launch.
  auto pack = [&](unsigned long long* maxout, const int* gate, int* zero_flag) {
    const int64_t cap = gate ? 1024 : 8192;
    if (side.nunits > 0) {
      hipLaunchKernelGGL(k_pack<2>, dim3(std::min<int64_t>(grid_units(side.nunits), cap)), dim3(256), 0, s,
                         side.jobs.as<PackJob>(), side.units.as<TileUnit>(), side.nunits, src, nullptr,
                         side.buf.as<uint8_t>(), side.total, side.rowmax.as<unsigned long long>(),
                         side.invscale.as<double2>(), sp.crt, top, maxout, gate, zero_flag);
      zero_flag = nullptr;
    }
    if (side.nkfunits > 0)
      hipLaunchKernelGGL(k_pack_kf<2>, dim3(std::min<int64_t>(grid_units(side.nkfunits), cap)), dim3(256), 0, s,
                         side.jobs.as<PackJob>(), side.kfunits.as<TileUnit>(), side.nkfunits, src,
                         nullptr, side.buf.as<uint8_t>(), side.total,
                         side.rowmax.as<unsigned long long>(), side.invscale.as<double2>(), sp.crt, top,
                         maxout, gate, zero_flag);
    TMX_HIP(hipGetLastError());
  };
  // Reuse policy: the host polls the flag a check raises in pinned memory, without blocking;
  // a changed operand pauses reuse for the next executions.  Under stream capture the choice
  // made now is what the graph replays (both paths give the same results).
  const bool capturing = stream_capturing(s);
  if (!capturing && side.changed_host.poll() > 0) side.backoff = kScaleReuseBackoff;
  const bool reuse = g_scale_reuse && side.have_max && side.last_src == src && side.backoff == 0;
  if (side.backoff > 0) --side.backoff;
  if (!reuse) {
    launch_fill_bytes_zero(side.rowmax.p, side.rowmax.bytes, s);
    if (side.nrowunits > 0)
      hipLaunchKernelGGL(k_rowmax, dim3(grid_units(side.nrowunits)), dim3(256), 0, s,
                         side.jobs.as<PackJob>(), side.rowunits.as<TileUnit>(), side.nrowunits, src,
                         side.rowmax.as<unsigned long long>());
    if (side.nunits > 0)
      hipLaunchKernelGGL(k_rowmax_t, dim3(grid_units(side.nunits)), dim3(256), 0, s,
                         side.jobs.as<PackJob>(), side.units.as<TileUnit>(), side.nunits, src,
                         side.rowmax.as<unsigned long long>());
    TMX_HIP(hipGetLastError());
    mark(0);
    pack(nullptr, nullptr, nullptr);
    mark(1);
  } else {
    // Three launches: the pass with the kept scales (recomputing the maxima into rowmax_new,
    // which the previous check left at zero, and clearing `changed`), the check, and the
    // gated second pass.
    if (!side.fresh_clean) launch_fill_bytes_zero(side.rowmax_new.p, side.rowmax_new.bytes, s);
    side.fresh_clean = false;
    mark(0);
    pack(side.rowmax_new.as<unsigned long long>(), nullptr, side.changed.as<int>());
    mark(1);
    const int g = grid_units((side.nrows + 255) / 256);
    hipLaunchKernelGGL(k_check_rowmax, dim3(g), dim3(256), 0, s,
                       side.rowmax_new.as<unsigned long long>(), side.rowmax.as<unsigned long long>(),
                       side.nrows, side.changed.as<int>(), side.changed_host.dev);
    TMX_HIP(hipGetLastError());
    side.fresh_clean = true;
    pack(nullptr, side.changed.as<int>(), nullptr);
  }
  side.have_max = true;
  side.last_src = src;
}

static void launch_gemm(SectorPlan& sp, const ShapeTasks& st, hipStream_t s) {
  const int grid = (st.ntasks + 7) / 8 * 8;
  if (!sp.split) {
#define TMX_F64_LAUNCH(TM_, TN_, NS_)                                                          \
  hipLaunchKernelGGL((k_gemm_f64<TM_, TN_, NS_>), dim3(grid), dim3(256), 0, s, st.tasks.as<SecTask>(), \
                     st.ntasks, sp.ksegs.as<KSeg>(), sp.A.buf.as<double>(), sp.B.buf.as<double>(), \
                     sp.cdense.as<double>())
    if (st.tm == 128 && st.tn == 128) TMX_F64_LAUNCH(128, 128, 4);
    else if (st.tm == 128 && st.tn == 160) TMX_F64_LAUNCH(128, 160, 4);
    else if (st.tm == 64 && st.tn == 64) TMX_F64_LAUNCH(64, 64, 4);
    else if (st.tm == 32 && st.tn == 32) TMX_F64_LAUNCH(32, 32, 8);
    else throw Error(TMX_ERR_INVALID, "no FP64 kernel for the requested tile shape");
#undef TMX_F64_LAUNCH
  } else {
#define TMX_I8_LAUNCH(TM_, TN_, WGM_, WGN_, NS_)                                                \
  hipLaunchKernelGGL((k_gemm_i8<TM_, TN_, WGM_, WGN_, NS_>), dim3(grid),                        \
                     dim3(64 * WGM_ * WGN_), 0, s, st.tasks.as<SecTask>(), st.ntasks,          \
                     sp.ksegs.as<KSeg>(), sp.A.buf.as<uint8_t>(), sp.A.total,                  \
                     sp.B.buf.as<uint8_t>(), sp.B.total, sp.cres.as<uint8_t>(), sp.c_total,    \
                     sp.crt)
    bool launched = false;
#define TMX_I8_CASE(TM_, TN_, WGM_, WGN_, NS_)                  \
  if (!launched && st.tm == TM_ && st.tn == TN_) {              \
    TMX_I8_LAUNCH(TM_, TN_, WGM_, WGN_, NS_);                   \
    launched = true;                                            \
  }
    TMX_I8_KERNELS(TMX_I8_CASE)
#undef TMX_I8_CASE
    if (!launched) throw Error(TMX_ERR_INVALID, "no int8 kernel for the requested tile shape");
#undef TMX_I8_LAUNCH
  }
  TMX_HIP(hipGetLastError());
}

int autotune_sector_plan(SectorPlan& sp, hipStream_t s) {
  hipEvent_t e0, e1;
  TMX_HIP(hipEventCreate(&e0));
  TMX_HIP(hipEventCreate(&e1));
  struct Cleanup {
    hipEvent_t a, b;
    ~Cleanup() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
  } cleanup{e0, e1};
  // The shortest of `reps` timed runs of fn: timings taken one by one, so that a run slowed by
  // a clock ramp or by other work on the device does not decide a choice.
  auto min_time_ms = [&](int reps, auto&& fn) {
    double best = 0;
    for (int r = 0; r < reps; ++r) {
      TMX_HIP(hipEventRecord(e0, s));
      fn();
      TMX_HIP(hipEventRecord(e1, s));
      TMX_HIP(hipEventSynchronize(e1));
      float ms = 0;
      TMX_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (r == 0 || ms < best) best = ms;
    }
    return best;
  };
  int changed = 0;
  for (SectorGeom& g : sp.geoms) {
    int best_tm = g.tm, best_tn = g.tn;
    double best_ms = 0;
    const I8Shape* shapes = sp.split ? kI8Shapes : kF64Shapes;
    const size_t nshapes = sp.split ? sizeof(kI8Shapes) / sizeof(I8Shape) : sizeof(kF64Shapes) / sizeof(I8Shape);
    // The plan's current shape is timed first, and another replaces the best so far only when
    // it is more than 1% faster: shapes within timing noise of each other do not trade places
    // from one run to the next.
    int cur = -1;
    for (size_t si = 0; si < nshapes; ++si)
      if (shapes[si].tm == g.tm && shapes[si].tn == g.tn && shape_allowed(shapes[si], g.mrows_pad, g.nrows_pad)) cur = (int)si;
    for (int oi = -1; oi < (int)nshapes; ++oi) {
      if (oi >= 0 && oi == cur) continue;
      if (oi < 0 && cur < 0) continue;
      const I8Shape& sh = shapes[oi < 0 ? cur : oi];
      if (!shape_allowed(sh, g.mrows_pad, g.nrows_pad)) continue;
      // this sector alone with this shape
      SectorGeom one = g;
      one.tm = sh.tm;
      one.tn = sh.tn;
      SectorPlan probe_tasks;  // only its task list is used
      probe_tasks.split = sp.split;
      probe_tasks.nmod = sp.nmod;
      probe_tasks.geoms.push_back(one);
      build_sector_tasks(probe_tasks);
      const ShapeTasks& st = probe_tasks.shapes.at(0);
      launch_gemm(sp, st, s);  // warm-up
      const double ms = min_time_ms(3, [&] { launch_gemm(sp, st, s); });
      if (best_ms == 0 || ms < best_ms * 0.99) { best_ms = ms; best_tm = sh.tm; best_tn = sh.tn; }
    }
    if (best_tm != g.tm || best_tn != g.tn) ++changed;
    g.tm = best_tm;
    g.tn = best_tn;
    std::lock_guard<std::mutex> lk(tune_mu());
    tune_load_locked();
    tune_map()[TuneKey{sp.split ? 1 : 0, sp.nmod, g.mrows_pad, g.nrows_pad, g.klen}] =
        TuneVal{best_tm, best_tn, best_ms};
  }
  // The shapes chosen sector by sector give one launch per distinct shape, run one after
  // another, and were timed with each sector alone.  Sectors side by side in one launch behave
  // differently (each may fill only part of the device alone), so the whole GEMM stage is also
  // timed with every sector on one common shape, which replaces the per-sector choice when it
  // is more than 1% faster.
  std::vector<std::pair<int, int>> per_sector;
  for (const SectorGeom& g : sp.geoms) per_sector.emplace_back(g.tm, g.tn);
  if (sp.geoms.size() > 1) {
    auto stage_ms = [&](const std::vector<std::pair<int, int>>& shp) {
      SectorPlan probe;  // only its task lists are used
      probe.split = sp.split;
      probe.nmod = sp.nmod;
      probe.geoms = sp.geoms;
      for (size_t i = 0; i < probe.geoms.size(); ++i) {
        probe.geoms[i].tm = shp[i].first;
        probe.geoms[i].tn = shp[i].second;
      }
      build_sector_tasks(probe);
      for (const ShapeTasks& st : probe.shapes) launch_gemm(sp, st, s);  // warm-up
      return min_time_ms(3, [&] {
        for (const ShapeTasks& st : probe.shapes) launch_gemm(sp, st, s);
      });
    };
    const I8Shape* shapes = sp.split ? kI8Shapes : kF64Shapes;
    const size_t nshapes = sp.split ? sizeof(kI8Shapes) / sizeof(I8Shape) : sizeof(kF64Shapes) / sizeof(I8Shape);
    std::vector<std::pair<int, int>> best = per_sector;
    double best_ms = stage_ms(per_sector);
    for (size_t si = 0; si < nshapes; ++si) {
      bool ok = true;
      for (const SectorGeom& g : sp.geoms) ok = ok && shape_allowed(shapes[si], g.mrows_pad, g.nrows_pad);
      if (!ok) continue;
      const std::vector<std::pair<int, int>> common(sp.geoms.size(), {shapes[si].tm, shapes[si].tn});
      const double ms = stage_ms(common);
      if (ms < best_ms * 0.99) {
        best_ms = ms;
        best = common;
      }
    }
    if (best != per_sector) {
      ++changed;
      std::lock_guard<std::mutex> lk(tune_mu());
      for (size_t i = 0; i < sp.geoms.size(); ++i) {
        SectorGeom& g = sp.geoms[i];
        g.tm = best[i].first;
        g.tn = best[i].second;
        tune_map()[TuneKey{sp.split ? 1 : 0, sp.nmod, g.mrows_pad, g.nrows_pad, g.klen}].tm = g.tm;
        tune_map()[TuneKey{sp.split ? 1 : 0, sp.nmod, g.mrows_pad, g.nrows_pad, g.klen}].tn = g.tn;
      }
    }
  }
  build_sector_tasks(sp);
  {
    std::lock_guard<std::mutex> lk(tune_mu());
    tune_save_locked();
  }
  return changed;
}

void set_sector_constant(SectorPlan& sp, bool a, bool b) {
  if (sp.A.constant != a) sp.A.packed_src = nullptr;
  if (sp.B.constant != b) sp.B.packed_src = nullptr;
  sp.A.constant = a;
  sp.B.constant = b;
}

void forget_sector_packing(SectorPlan& sp) {
  sp.A.packed_src = sp.B.packed_src = nullptr;
}

// Scatter stage: the product held by `src` (residue planes and row scales, or the dense FP64
// result) into C through the scatter tables of `sp`.  src is sp itself, or a plan whose packed
// operands and sector geometry are the same (same product_key), so its buffers have sp's
// layout.
static void scatter_stage(const SectorPlan& sp, const SectorPlan& src, double* C, double alpha,
                          hipStream_t s) {
  TMX_REQUIRE(!sp.assign || sp.scatter.size() <= 1, "assign needs one scatter launch");
  if (sp.assign)
    launch_zero_ranges(sp.zero_off.as<int64_t>(), sp.zero_prefix.as<int64_t>(), sp.nzero,
                       sp.zero_elems, C, s);
  if (sp.split && sp.nfused > 0) {
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((sp.nfused + 3) / 4, 8192));
#define TMX_CRT_SCATTER(NM_)                                                                  \
  hipLaunchKernelGGL(k_crt_scatter<NM_>, dim3(g), dim3(256), 0, s, sp.fused_tiles.as<FusedTile>(), \
                     sp.nfused, sp.crt_pairs.as<CrtPair>(), sp.scatter[0].jobs.as<ScatterJob>(),  \
                     sp.arow_info.as<int2>(), sp.brow_info.as<int2>(), sp.jobtab.as<int32_t>(),   \
                     sp.jobtab_n, src.cres.as<int8_t>(), sp.c_total, sp.crt,                       \
                     src.A.invscale.as<double2>(), src.B.invscale.as<double2>(), C, alpha,      \
                     (int)sp.assign)
#define TMX_CRT_SCATTER_G(NM_, ROWC_)                                                          \
  hipLaunchKernelGGL((k_crt_scatter_g<NM_, ROWC_>), dim3(g), dim3(256), 0, s,                   \
                     sp.fused_tiles.as<FusedTile>(), sp.nfused, sp.crt_pairs.as<CrtPair>(),      \
                     sp.scatter[0].jobs.as<ScatterJob>(), sp.arow_info.as<int2>(),              \
                     sp.brow_info.as<int2>(), sp.arow_digits.as<uint32_t>(),                   \
                     sp.brow_digits.as<uint32_t>(), sp.jobtab.as<int32_t>(), sp.jobtab_n,       \
                     src.cres.as<int8_t>(), sp.c_total, sp.crt, src.A.invscale.as<double2>(),    \
                     src.B.invscale.as<double2>(), C, alpha, (int)sp.assign)
    if (sp.fused_layout == 1) {
      if (sp.nmod == 11) TMX_CRT_SCATTER(11);
      else TMX_CRT_SCATTER(0);
    } else if (sp.fused_layout == 2) {
      if (sp.nmod == 11) TMX_CRT_SCATTER_G(11, false);
      else TMX_CRT_SCATTER_G(0, false);
    } else {
      if (sp.nmod == 11) TMX_CRT_SCATTER_G(11, true);
      else TMX_CRT_SCATTER_G(0, true);
    }
#undef TMX_CRT_SCATTER
#undef TMX_CRT_SCATTER_G
    TMX_HIP(hipGetLastError());
    return;
  }
  // The dense FP64 result: for split plans without the fused pass, src's own execution has
  // already reconstructed it (k_crt) into src.cdense.
  for (const SectorPlan::ScatterRound& sr : sp.scatter) {
    if (sr.nunits == 0) continue;
    const int g = grid_units(sr.nunits);
    hipLaunchKernelGGL(k_scatter, dim3(g), dim3(256), 0, s, sr.jobs.as<ScatterJob>(),
                       sr.units.as<TileUnit>(), sr.nunits, src.cdense.as<double>(), C, alpha,
                       (int)sp.assign);
    TMX_HIP(hipGetLastError());
  }
}

void execute_sector_plan(SectorPlan& sp, const double* A, const double* B, double* C, double alpha,
                         hipStream_t s, hipEvent_t* ev, hipEvent_t* gemm_ev, hipEvent_t* pack_ev_a,
                         hipEvent_t* pack_ev_b) {
  if (ev) TMX_HIP(hipEventRecord(ev[0], s));
  pack_side(sp.A, A, sp.split, sp, s, pack_ev_a);
  if (ev) TMX_HIP(hipEventRecord(ev[1], s));
  pack_side(sp.B, B, sp.split, sp, s, pack_ev_b);
  if (ev) TMX_HIP(hipEventRecord(ev[2], s));
  for (size_t i = 0; i < sp.shapes.size(); ++i) {
    launch_gemm(sp, sp.shapes[i], s);
    if (gemm_ev && i + 1 < sp.shapes.size()) TMX_HIP(hipEventRecord(gemm_ev[i], s));
  }
  if (ev) TMX_HIP(hipEventRecord(ev[3], s));
  if (sp.split && sp.nfused == 0 && sp.crt_tiles > 0) {
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((sp.crt_tiles + 3) / 4, 8192));
    hipLaunchKernelGGL(k_crt, dim3(g), dim3(256), 0, s, sp.crt_secs.as<CrtSector>(), sp.ncrt_secs,
                       sp.crt_tiles, sp.cres.as<int8_t>(), sp.c_total, sp.crt,
                       sp.A.invscale.as<double2>(), sp.B.invscale.as<double2>(),
                       sp.cdense.as<double>());
    TMX_HIP(hipGetLastError());
  }
  scatter_stage(sp, sp, C, alpha, s);
  if (ev) TMX_HIP(hipEventRecord(ev[4], s));
}

void execute_sector_scatter_from(SectorPlan& sp, const SectorPlan& src, double* C, double alpha,
                                 hipStream_t s) {
  TMX_REQUIRE(sp.product_key == src.product_key && sp.split == src.split && sp.nmod == src.nmod,
              "scatter from a plan with another product");
  scatter_stage(sp, src, C, alpha, s);
}

}  // namespace tmx
