// k_sector.hip -- sector GEMM pipeline on gfx950:
//   pack (block layout -> MFMA fragment-major class matrices, permutation fused into the
//   gather; FP64 copy or int8 residues), dense GEMM (v_mfma_f64_16x16x4f64 or
//   v_mfma_i32_32x32x32_i8), scatter (dense result -> C blocks in C's mode order, FP64
//   accumulate, or Garner CRT reconstruction of the int8 residues into FP64).
#include <cmath>

#include "sector.hpp"
#include "tmx_device.hpp"

namespace tmx {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));
typedef int i16v __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// Packing.  One workgroup iteration handles a 64x64 tile (rows x columns) of one source
// block: the row and column offsets of the tile go to LDS once, the source tile is read
// with lanes along the source's contiguous direction (coalesced) into an LDS tile, and then
// each thread emits one row's 16 consecutive columns in MFMA fragment-major order (int8:
// one 16-byte store per residue plane; FP64: 16-row runs of 128 bytes).
template <int MODE>  // 0: FP64 copy, 1: row max |x|, 2: int8 residues
__global__ __launch_bounds__(256) void k_pack(const PackJob* __restrict__ jobs,
                                              const TileUnit* __restrict__ units, int64_t nunits,
                                              const double* __restrict__ src,
                                              double* __restrict__ dst64,
                                              uint8_t* __restrict__ dst8, int64_t plane,
                                              unsigned long long* __restrict__ rowmax,
                                              double* __restrict__ invscale, GarnerConsts gc,
                                              double scale_top) {
  __shared__ double tile[64][65];
  __shared__ int roff[64], coff[64];
  __shared__ double pmax[4][64];
  const int tid = threadIdx.x;
  for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const TileUnit un = units[u];
    const PackJob& jb = jobs[un.job];
    const int nr = min(64, jb.rows - un.r0), nc = min(64, jb.cols - un.c0);
    if (tid < 64) roff[tid] = tid < nr ? radix_off(jb.r, un.r0 + tid) : 0;
    else if (tid < 128) coff[tid - 64] = (tid - 64) < nc ? radix_off(jb.c, un.c0 + tid - 64) : 0;
    __syncthreads();
    const double* s = src + jb.src_off;
    if (jb.kcontig) {
#pragma unroll 4
      for (int i = 0; i < 16; ++i) {
        const int rr = (tid >> 6) + 4 * i, cc = tid & 63;
        tile[rr][cc] = (rr < nr && cc < nc) ? s[roff[rr] + coff[cc]] : 0.0;
      }
    } else {
#pragma unroll 4
      for (int i = 0; i < 16; ++i) {
        const int rr = tid & 63, cc = (tid >> 6) + 4 * i;
        tile[rr][cc] = (rr < nr && cc < nc) ? s[roff[rr] + coff[cc]] : 0.0;
      }
    }
    __syncthreads();
    const int rr = tid & 63, q = tid >> 6;
    const int row = jb.row0 + un.r0 + rr;
    const int c0 = un.c0 + 16 * q;
    const bool active = rr < nr && c0 < jb.cols;
    if (MODE == 0) {
      if (active) {
        const int ncol = min(16, jb.cols - c0);
        for (int j = 0; j < ncol; ++j)
          dst64[jb.dst_base + frag_f64(row, jb.col0 + c0 + j, jb.kpad)] = tile[rr][16 * q + j];
      }
    } else if (MODE == 1) {
      double m = 0.0;
#pragma unroll
      for (int j = 0; j < 16; ++j) m = fmax(m, fabs(tile[rr][16 * q + j]));
      pmax[q][rr] = m;
      __syncthreads();
      if (q == 0 && rr < nr) {
        m = fmax(fmax(pmax[0][rr], pmax[1][rr]), fmax(pmax[2][rr], pmax[3][rr]));
        atomicMax(rowmax + jb.growbase + row, (unsigned long long)__double_as_longlong(m));
      }
    } else if (active) {
      const int grow = jb.growbase + row;
      const double rm = __longlong_as_double((long long)rowmax[grow]);
      const double sigma = rm > 0.0 ? scale_top / rm : 0.0;
      if (un.c0 == 0 && q == 0) invscale[grow] = rm / scale_top;
      double x[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) x[j] = rint(tile[rr][16 * q + j] * sigma);  // padding reads 0
      const int64_t off = jb.dst_base + frag_i8(row, jb.col0 + c0, jb.kpad);
      for (int l = 0; l < gc.n; ++l) {
        const double m = (double)gc.m[l];
        const double im = 1.0 / m;
        const int hi = (gc.m[l] - 1) / 2, lo = -(gc.m[l] / 2);
        uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const double qq = rint(x[j] * im);
          int r = (int)fma(-qq, m, x[j]);
          if (r > hi) r -= gc.m[l];
          else if (r < lo) r += gc.m[l];
          w[j >> 2] |= ((uint32_t)(r & 0xff)) << (8 * (j & 3));
        }
        *reinterpret_cast<uint4*>(dst8 + (int64_t)l * plane + off) = make_uint4(w[0], w[1], w[2], w[3]);
      }
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

// FP64 sector GEMM: workgroup tile 128x128, 4 waves (2x2) of 64x64 = 4x4 MFMA 16x16
// tiles.  A K stage is 16 columns = 4 fragments of 512 bytes per 16-row tile, contiguous in
// the fragment-major class matrix; each stage of A (8 row tiles) and B (8 row tiles) is
// copied with 16-byte loads into LDS, double buffered, so one stage of MFMA work
// (4 k-steps x 16 MFMAs per wave) covers the global latency of the next stage.
__global__ __launch_bounds__(256) void k_gemm_f64(const SecTask* __restrict__ tasks, int ntasks,
                                                  const KSeg* __restrict__ segs,
                                                  const double* __restrict__ Ap,
                                                  const double* __restrict__ Bp,
                                                  double* __restrict__ Cd) {
  constexpr int RT = 8;                 // 16-row tiles per operand per workgroup
  constexpr int CH = RT * 4 * 512 / 16;  // 16-byte chunks per operand stage (1024)
  __shared__ uint4 lds[2][2 * CH];
  const int task = xcd_task(ntasks);
  if (task >= ntasks) return;
  const SecTask tk = tasks[task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int art = tk.mrows >> 4, brt = tk.nrows >> 4;
  const int rt0 = tk.m0 >> 4, ct0 = tk.n0 >> 4;
  const int64_t akp = tk.a_kpad >> 2, bkp = tk.b_kpad >> 2;
  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  uint4 stage[8];
  for (int sg = tk.kseg_begin; sg < tk.kseg_end; ++sg) {
    const KSeg seg = segs[sg];
    const int nst = (seg.len + 15) >> 4;
    auto load = [&](int st) {
      const int ksa = (seg.a_col >> 2) + 4 * st, ksb = (seg.b_col >> 2) + 4 * st;
      const int kleft = (seg.len >> 2) - 4 * st;  // k-steps of this segment still ahead
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int id = q * 256 + tid;  // chunk id within the A (0..1023) or B (1024..2047) stage
        const int op = id >> 10, c = id & 1023;
        const int rtl = c >> 7, ks = (c >> 5) & 3, part = c & 31;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (op == 0) {
          const int rt = rt0 + rtl;
          if (rt < art && ks < kleft)
            v = *reinterpret_cast<const uint4*>(Ap + tk.a_base + ((int64_t)rt * akp + ksa + ks) * 64 + part * 2);
        } else {
          const int rt = ct0 + rtl;
          if (rt < brt && ks < kleft)
            v = *reinterpret_cast<const uint4*>(Bp + tk.b_base + ((int64_t)rt * bkp + ksb + ks) * 64 + part * 2);
        }
        stage[q] = v;
      }
    };
    auto store = [&](int buf) {
#pragma unroll
      for (int q = 0; q < 8; ++q) lds[buf][q * 256 + tid] = stage[q];
    };
    load(0);
    __syncthreads();
    store(0);
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
      const int buf = st & 1;
      if (st + 1 < nst) load(st + 1);
      const double* la = reinterpret_cast<const double*>(&lds[buf][0]);
      const double* lb = reinterpret_cast<const double*>(&lds[buf][CH]);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a[i] = la[((wm * 4 + i) * 4 + ks) * 64 + lane];
          b[i] = lb[((wn * 4 + i) * 4 + ks) * 64 + lane];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      }
      if (st + 1 < nst) store(buf ^ 1);
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        // v_mfma_f64_16x16x4f64 result: register e of lane l holds row 4e + l/16, col l%16.
        const int row = tk.m0 + 64 * wm + 16 * i + 4 * e + (lane >> 4);
        const int col = tk.n0 + 64 * wn + 16 * j + (lane & 15);
        if (row < tk.mrows && col < tk.nrows) Cd[tk.c_base + (int64_t)row * tk.ldc + col] = acc[i][j][e];
      }
}

// ---------------------------------------------------------------------------
// int8 sector GEMM, one modulus per blockIdx.y.  4 waves (2x2), workgroup tile TM x TN,
// wave tile (TM/2) x (TN/2) of 32x32 MFMA tiles.  A K stage is 64 columns: two 1 KB
// fragments per 32-row tile of A and B, copied with 16-byte loads (whole fragments) into a
// double-buffered LDS stage.  Loads run two stages ahead of their LDS store (two register
// sets), so two stages of MFMA work cover the global latency.  The int32 sums are exact;
// the epilogue reduces them mod m_l to symmetric int8 residues of the dense result.
template <int TM, int TN>
__global__ __launch_bounds__(256) void k_gemm_i8(const SecTask* __restrict__ tasks, int ntasks,
                                                 const KSeg* __restrict__ segs,
                                                 const uint8_t* __restrict__ Ap, int64_t aplane,
                                                 const uint8_t* __restrict__ Bp, int64_t bplane,
                                                 uint8_t* __restrict__ Cr, int64_t cplane,
                                                 GarnerConsts gc) {
  constexpr int AF = (TM / 32) * 2, BF = (TN / 32) * 2;  // fragments per stage
  constexpr int NCH = (AF + BF) * 64 / 256;               // 16-byte chunks per thread per stage
  static_assert((AF + BF) % 4 == 0, "stage must split evenly over 256 threads");
  constexpr int WTM = TM / 2 / 32, WTN = TN / 2 / 32;     // MFMA tiles per wave
  __shared__ i4 lds[2][(AF + BF) * 64];
  const int task = xcd_task(ntasks);
  if (task >= ntasks) return;
  const int l = blockIdx.y;
  const SecTask tk = tasks[task];
  const uint8_t* A = Ap + (int64_t)l * aplane + tk.a_base;
  const uint8_t* B = Bp + (int64_t)l * bplane + tk.b_base;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  const int wm = wave & 1, wn = wave >> 1;
  const int art = tk.mrows >> 5, brt = tk.nrows >> 5;
  const int64_t akt = tk.a_kpad >> 5, bkt = tk.b_kpad >> 5;
  const int rt0 = tk.m0 >> 5, ct0 = tk.n0 >> 5;

  i16v acc[WTM][WTN];
#pragma unroll
  for (int i = 0; i < WTM; ++i)
#pragma unroll
    for (int j = 0; j < WTN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0;

  i4 ra[NCH], rb[NCH];
  // Every load issues exactly NCH unconditional loads with addresses clamped into the class
  // matrices, so the count of loads in flight is path-independent and each LDS store waits
  // only for its own register set.  Clamped rows only feed output rows / columns beyond the
  // class, which are never stored; clamped K columns are skipped in the MFMA step.
  // Chunk q of a thread is fragment 4q + wave (wave-uniform), 16 bytes at lane*16.
#define TMX_I8_LOAD(R, ST)                                                                  \
  _Pragma("unroll") for (int q = 0; q < NCH; ++q) {                                         \
    const int f = 4 * q + wv;                                                               \
    const bool isa = f < AF;                                                                \
    const int fl = isa ? f : f - AF;                                                        \
    const int lim = isa ? art : brt;                                                        \
    const int rt = min((isa ? rt0 : ct0) + (fl >> 1), lim - 1);                             \
    const int kt = min(2 * (ST) + (fl & 1), nkt - 1);                                       \
    const uint8_t* g = isa ? A + ((int64_t)rt * akt + kt0a + kt) * 1024                     \
                           : B + ((int64_t)rt * bkt + kt0b + kt) * 1024;                    \
    R[q] = *reinterpret_cast<const i4*>(g + lane * 16);                                     \
  }
#define TMX_I8_STORE(R, BUF) \
  _Pragma("unroll") for (int q = 0; q < NCH; ++q) lds[BUF][(4 * q + wv) * 64 + lane] = R[q];
#define TMX_I8_COMPUTE(BUF, ST)                                                             \
  _Pragma("unroll") for (int kk = 0; kk < 2; ++kk) {                                        \
    if (kk < nkt - 2 * (ST)) {                                                              \
      i4 a[WTM], b[WTN];                                                                    \
      _Pragma("unroll") for (int i = 0; i < WTM; ++i)                                       \
          a[i] = lds[BUF][((wm * WTM + i) * 2 + kk) * 64 + lane];                           \
      _Pragma("unroll") for (int j = 0; j < WTN; ++j)                                       \
          b[j] = lds[BUF][(AF + (wn * WTN + j) * 2 + kk) * 64 + lane];                      \
      _Pragma("unroll") for (int i = 0; i < WTM; ++i)                                       \
      _Pragma("unroll") for (int j = 0; j < WTN; ++j)                                       \
          acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i], b[j], acc[i][j], 0, 0, 0); \
    }                                                                                       \
  }
  for (int sg = tk.kseg_begin; sg < tk.kseg_end; ++sg) {
    const KSeg seg = segs[sg];
    const int nst = (seg.len + 63) >> 6, nkt = seg.len >> 5;
    const int kt0a = seg.a_col >> 5, kt0b = seg.b_col >> 5;
    // Stage st lives in ra for even st and in rb for odd st; a trip handles two stages.
    TMX_I8_LOAD(ra, 0)
    TMX_I8_LOAD(rb, 1)
    __syncthreads();
    TMX_I8_STORE(ra, 0)
    __syncthreads();
    for (int st = 0; st < nst; st += 2) {
      TMX_I8_LOAD(ra, st + 2)
      TMX_I8_COMPUTE(0, st)
      TMX_I8_STORE(rb, 1)
      __syncthreads();
      TMX_I8_LOAD(rb, st + 3)
      TMX_I8_COMPUTE(1, st + 1)
      TMX_I8_STORE(ra, 0)
      __syncthreads();
    }
  }
#undef TMX_I8_LOAD
#undef TMX_I8_STORE
#undef TMX_I8_COMPUTE
  // Epilogue: exact int32 sums -> symmetric residues mod m_l.
  const int m = gc.m[l], hi = (m - 1) / 2, lo = -(m / 2);
  uint8_t* Cp = Cr + (int64_t)l * cplane + tk.c_base;
#pragma unroll
  for (int i = 0; i < WTM; ++i)
#pragma unroll
    for (int j = 0; j < WTN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        // v_mfma_i32_32x32x32_i8 result: register e of lane l holds row 8(e/4) + 4(l/32) + e%4,
        // column l%32.
        const int row = tk.m0 + (wm * WTM + i) * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        const int col = tk.n0 + (wn * WTN + j) * 32 + (lane & 31);
        if (row < tk.mrows && col < tk.nrows) {
          int r = acc[i][j][e] % m;
          if (r > hi) r -= m;
          else if (r < lo) r += m;
          Cp[(int64_t)row * tk.ldc + col] = (uint8_t)(r & 0xff);
        }
      }
}

// ---------------------------------------------------------------------------
// Scatter: dense sector result -> C block (C's mode order), C += alpha * value.  A workgroup
// iteration handles a 64x64 tile: values are produced with lanes along the dense columns
// (coalesced dense reads; for the split path each thread runs Garner's reconstruction for
// its element), staged in LDS, and added into C with lanes along C's contiguous direction.
__device__ __forceinline__ int mod_sym(int x, int m, float rm, int hi, int lo) {
  // |x| < 2^24: the float quotient is exact to well under 1/2, so one correction suffices.
  const int q = __float2int_rn((float)x * rm);
  int r = x - q * m;
  if (r > hi) r -= m;
  else if (r < lo) r += m;
  return r;
}

template <int SPLIT>
__global__ __launch_bounds__(256) void k_scatter(const ScatterJob* __restrict__ jobs,
                                                 const TileUnit* __restrict__ units, int64_t nunits,
                                                 const double* __restrict__ Cd,
                                                 const int8_t* __restrict__ Cr, int64_t cplane,
                                                 const double* __restrict__ invA,
                                                 const double* __restrict__ invB, GarnerConsts gc,
                                                 double* __restrict__ C, double alpha) {
  __shared__ double tile[64][65];
  __shared__ int roff[64], coff[64];
  const int tid = threadIdx.x;
  float rm[kMaxModuli];
  int hi[kMaxModuli], lo[kMaxModuli];
  if (SPLIT)
    for (int i = 0; i < gc.n; ++i) {
      rm[i] = 1.0f / (float)gc.m[i];
      hi[i] = (gc.m[i] - 1) / 2;
      lo[i] = -(gc.m[i] / 2);
    }
  for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const TileUnit un = units[u];
    const ScatterJob& jb = jobs[un.job];
    const int nr = min(64, jb.rows - un.r0), nc = min(64, jb.cols - un.c0);
    if (tid < 64) roff[tid] = tid < nr ? radix_off(jb.r, un.r0 + tid) : 0;
    else if (tid < 128) coff[tid - 64] = (tid - 64) < nc ? radix_off(jb.c, un.c0 + tid - 64) : 0;
#pragma unroll 1
    for (int i = 0; i < 16; ++i) {
      const int rr = (tid >> 6) + 4 * i, cc = tid & 63;
      double v = 0.0;
      if (rr < nr && cc < nc) {
        const int64_t di = jb.src_base + (int64_t)(jb.row0 + un.r0 + rr) * jb.ldc + jb.col0 + un.c0 + cc;
        if (!SPLIT) {
          v = Cd[di];
        } else {
          int dig[kMaxModuli];
          for (int k = 0; k < gc.n; ++k) {
            int t = Cr[(int64_t)k * cplane + di];
            for (int j = 0; j < k; ++j) t = mod_sym((t - dig[j]) * gc.inv[j][k], gc.m[k], rm[k], hi[k], lo[k]);
            dig[k] = t;
          }
          double x = (double)dig[gc.n - 1];
          for (int k = gc.n - 2; k >= 0; --k) x = fma(x, (double)gc.m[k], (double)dig[k]);
          v = x * invA[jb.arow0 + un.r0 + rr] * invB[jb.bcol0 + un.c0 + cc];
        }
      }
      tile[rr][cc] = alpha * v;
    }
    __syncthreads();
    double* cb = C + jb.c_off;
    if (jb.ccontig) {
      for (int i = 0; i < 16; ++i) {
        const int rr = (tid >> 6) + 4 * i, cc = tid & 63;
        if (rr < nr && cc < nc) cb[roff[rr] + coff[cc]] += tile[rr][cc];
      }
    } else {
      for (int i = 0; i < 16; ++i) {
        const int rr = tid & 63, cc = (tid >> 6) + 4 * i;
        if (rr < nr && cc < nc) cb[roff[rr] + coff[cc]] += tile[rr][cc];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Host side
void choose_tile(bool split, int64_t /*mrows*/, int64_t /*nrows*/, int* tm, int* tn) {
  if (split) { *tm = 256; *tn = 128; }
  else { *tm = 128; *tn = 128; }
}

void split_tasks_by_shape(SectorPlan& sp, const std::vector<SecTask>& tasks) {
  std::map<int, std::vector<SecTask>> by;
  for (const SecTask& t : tasks) by[t.pad].push_back(t);
  for (auto& kv : by) {
    ShapeTasks st;
    st.tm = kv.first / 1000;
    st.tn = kv.first % 1000;
    st.ntasks = (int)kv.second.size();
    st.tasks = upload(kv.second);
    sp.shapes.push_back(std::move(st));
  }
}

static int grid_units(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(n, 8192)); }

static void pack_side(const PackSide& side, const double* src, bool split, const SectorPlan& sp,
                      hipStream_t s) {
  if (side.nunits == 0) return;
  const int g = grid_units(side.nunits);
  if (!split) {
    hipLaunchKernelGGL(k_pack<0>, dim3(g), dim3(256), 0, s, side.jobs.as<PackJob>(),
                       side.units.as<TileUnit>(), side.nunits, src, side.buf.as<double>(), nullptr,
                       (int64_t)0, nullptr, nullptr, sp.garner, 0.0);
  } else {
    const double top = std::ldexp(1.0, sp.bits_a) - 1.0;
    hipLaunchKernelGGL(k_pack<1>, dim3(g), dim3(256), 0, s, side.jobs.as<PackJob>(),
                       side.units.as<TileUnit>(), side.nunits, src, nullptr, nullptr, (int64_t)0,
                       side.rowmax.as<unsigned long long>(), nullptr, sp.garner, top);
    TMX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pack<2>, dim3(g), dim3(256), 0, s, side.jobs.as<PackJob>(),
                       side.units.as<TileUnit>(), side.nunits, src, nullptr, side.buf.as<uint8_t>(),
                       side.total, side.rowmax.as<unsigned long long>(), side.invscale.as<double>(),
                       sp.garner, top);
  }
  TMX_HIP(hipGetLastError());
}

void execute_sector_plan(SectorPlan& sp, const double* A, const double* B, double* C, double alpha,
                         hipStream_t s, hipEvent_t* ev) {
  if (ev) TMX_HIP(hipEventRecord(ev[0], s));
  if (sp.split) {
    launch_fill_bytes_zero(sp.A.rowmax.p, sp.A.rowmax.bytes, s);
    launch_fill_bytes_zero(sp.B.rowmax.p, sp.B.rowmax.bytes, s);
  }
  pack_side(sp.A, A, sp.split, sp, s);
  pack_side(sp.B, B, sp.split, sp, s);
  if (ev) TMX_HIP(hipEventRecord(ev[1], s));
  for (const ShapeTasks& st : sp.shapes) {
    const int grid = (st.ntasks + 7) / 8 * 8;
    if (!sp.split) {
      hipLaunchKernelGGL(k_gemm_f64, dim3(grid), dim3(256), 0, s, st.tasks.as<SecTask>(), st.ntasks,
                         sp.ksegs.as<KSeg>(), sp.A.buf.as<double>(), sp.B.buf.as<double>(),
                         sp.cdense.as<double>());
    } else if (st.tm == 256 && st.tn == 128) {
      hipLaunchKernelGGL((k_gemm_i8<256, 128>), dim3(grid, sp.nmod), dim3(256), 0, s,
                         st.tasks.as<SecTask>(), st.ntasks, sp.ksegs.as<KSeg>(),
                         sp.A.buf.as<uint8_t>(), sp.A.total, sp.B.buf.as<uint8_t>(), sp.B.total,
                         sp.cdense.as<uint8_t>(), sp.c_total, sp.garner);
    } else {
      throw Error(TMX_ERR_INVALID, "no kernel for the requested tile shape");
    }
    TMX_HIP(hipGetLastError());
  }
  if (ev) TMX_HIP(hipEventRecord(ev[2], s));
  if (sp.nsunits > 0) {
    const int g = grid_units(sp.nsunits);
    if (!sp.split)
      hipLaunchKernelGGL(k_scatter<0>, dim3(g), dim3(256), 0, s, sp.sjobs.as<ScatterJob>(),
                         sp.sunits.as<TileUnit>(), sp.nsunits, sp.cdense.as<double>(), nullptr,
                         (int64_t)0, nullptr, nullptr, sp.garner, C, alpha);
    else
      hipLaunchKernelGGL(k_scatter<1>, dim3(g), dim3(256), 0, s, sp.sjobs.as<ScatterJob>(),
                         sp.sunits.as<TileUnit>(), sp.nsunits, nullptr, sp.cdense.as<int8_t>(),
                         sp.c_total, sp.A.invscale.as<double>(), sp.B.invscale.as<double>(),
                         sp.garner, C, alpha);
    TMX_HIP(hipGetLastError());
  }
  if (ev) TMX_HIP(hipEventRecord(ev[3], s));
}

}  // namespace tmx
