// k_blockpair.hip -- fused block-pair FP64 contraction and LDS-tiled block assign (gfx950).
//
// k_bp_f64 replaces TAMM's per-pair transpose -> dgemm -> transpose sequence
// (src/tamm/kernels/multiply.hpp:159-210 + gpu_blas.cpp:43-85): the permutation of A and B
// into GEMM order is folded into the global->LDS load (per-pair offset tables in LDS), the
// product runs on v_mfma_f64_16x16x4f64, and the epilogue accumulates straight into the C
// block in its own mode order.  One workgroup owns one C tile and loops over every
// reduction pair of that C block, so C is read and written once per tile and needs no
// atomics (TAMM zeroes, accumulates and copies a C staging buffer per block instead,
// multop.hpp:690-839).
//
// k_assign replaces kernels::assign / librett (assign.hpp:466-485, multiply.hpp:108-157):
// dst(dl) (+)= scale * src(sl) for a label permutation, tiled 32x32 through LDS so both the
// source read (along the source's contiguous mode) and the destination write (along the
// destination's contiguous mode) are coalesced.
#include "tmx_device.hpp"
#include "tmx_kernels.hpp"

namespace tmx {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Fused block-pair FP64 GEMM.  256 threads = 4 waves in a WM x (4/WM) grid over a TM x TN C
// tile; each wave owns (TM/WM) x (TN*WM/4) = FM x FN MFMA 16x16 tiles; BK k per LDS stage,
// double buffered (global -> registers -> LDS).  Tile shapes: 64x64 (2x2 waves of 32x32, BK 16)
// for general blocks; 16x256 (1x4 waves of 16x64, BK 8) for blocks with few rows, such as
// one-body terms R(a,b,i,j) += F(k,i) T(a,b,k,j) whose row extent is one O tile, where a
// 64-row tile would be mostly padding (the plan puts the small side on the rows).
//
// LDS layout: As[m][k] and Bs[n][k] with an odd row stride (BK + 1 doubles).  The MFMA reads of
// a 16x4 fragment (lane l -> row l%16, k l/16) and both global->LDS write patterns (k-fastest
// or row-fastest) then spread over the banks.
constexpr int BP_KC = 512;

// 64x64: four workgroups per CU (128 VGPRs, 40 KB of LDS; 7-9% faster than three, which 136
// registers gave); 16x256: three (45 KB of LDS).
template <int TM, int TN, int WM, int BK>
__global__ __launch_bounds__(256, TM == 64 ? 4 : 3) void k_bp_f64(const BPTask* __restrict__ tasks, int ntasks,
                                                const BPPair* __restrict__ pairs,
                                                const BPCShape* __restrict__ cshapes,
                                                const BPKShape* __restrict__ kshapes,
                                                const double* __restrict__ A,
                                                const double* __restrict__ B,
                                                double* __restrict__ C, double alpha,
                                                int is_assign, BPXcd xr) {
  constexpr int WN = 4 / WM;
  constexpr int WTM = TM / WM, WTN = TN / WN;  // wave tile
  constexpr int FM = WTM / 16, FN = WTN / 16;  // MFMA tiles per wave
  constexpr int LD = BK + 1;
  constexpr int NA = (TM * BK + 255) / 256, NB = (TN * BK + 255) / 256;  // loads per thread
  static_assert(WM * WN == 4 && FM >= 1 && FN >= 1 && WTM % 16 == 0 && WTN % 16 == 0, "wave grid");
  static_assert(BK % 4 == 0 && BP_KC % BK == 0, "k stage");
  __shared__ double As[2][TM * LD];
  __shared__ double Bs[2][TN * LD];
  __shared__ int offAm[TM], offBn[TN], offCm[TM], offCn[TN];
  __shared__ int offAk[BP_KC], offBk[BP_KC];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;

  const int xcd = blockIdx.x & 7, per_xcd = gridDim.x >> 3;
  for (int task_id = xr.lo[xcd] + (blockIdx.x >> 3); task_id < xr.hi[xcd]; task_id += per_xcd) {
    const BPTask tk = tasks[task_id];
    const BPCShape& cs = cshapes[tk.cshape];
    const int M = cs.M, N = cs.N;

    d4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    for (int p = tk.pair_begin; p < tk.pair_end; ++p) {
      const BPPair pr = pairs[p];
      const BPKShape& ks = kshapes[pr.kshape];
      const int K = ks.K;
      const double* Ablk = A + pr.a_off;
      const double* Bblk = B + pr.b_off;
      __syncthreads();  // previous users of the offset tables are done
      for (int t = tid; t < TM + TN; t += 256) {
        if (t < TM) {
          int m = tk.m0 + t;
          offAm[t] = m < M ? radix_off(ks.am, m) : -1;
        } else {
          int c = t - TM;
          int n = tk.n0 + c;
          offBn[c] = n < N ? radix_off(ks.bn, n) : -1;
        }
      }
      for (int kc = 0; kc < K; kc += BP_KC) {
        const int klen = min(BP_KC, K - kc);
        __syncthreads();
        for (int t = tid; t < BP_KC; t += 256) {
          int k = kc + t;
          bool ok = t < klen;
          offAk[t] = ok ? radix_off(ks.ak, k) : -1;
          offBk[t] = ok ? radix_off(ks.bk, k) : -1;
        }
        __syncthreads();
        const int nsteps = (klen + BK - 1) / BK;

        // Element q = tid + 256 e of a stage tile (rows x BK): along k when the operand is read
        // along K (its stride-1 mode is the fastest K digit), else along rows.
        double ra[NA], rb[NB];
        auto coord = [&](int q, int rows, bool kcontig, int& r, int& k) {
          if (kcontig) { k = q % BK; r = q / BK; }
          else { r = q % rows; k = q / rows; }
        };
        auto load_stage = [&](int step) {
          const int kb = step * BK;
#pragma unroll
          for (int e = 0; e < NA; ++e) {
            const int q = tid + 256 * e;
            int r, k;
            coord(q, TM, ks.a_kcontig, r, k);
            double v = 0.0;
            if (q < TM * BK) {
              int om = offAm[r], ok = offAk[kb + k];
              if (om >= 0 && ok >= 0) v = Ablk[om + ok];
            }
            ra[e] = v;
          }
#pragma unroll
          for (int e = 0; e < NB; ++e) {
            const int q = tid + 256 * e;
            int r, k;
            coord(q, TN, ks.b_kcontig, r, k);
            double v = 0.0;
            if (q < TN * BK) {
              int on = offBn[r], ok = offBk[kb + k];
              if (on >= 0 && ok >= 0) v = Bblk[on + ok];
            }
            rb[e] = v;
          }
        };
        auto store_stage = [&](int buf) {
#pragma unroll
          for (int e = 0; e < NA; ++e) {
            const int q = tid + 256 * e;
            int r, k;
            coord(q, TM, ks.a_kcontig, r, k);
            if (q < TM * BK) As[buf][r * LD + k] = ra[e];
          }
#pragma unroll
          for (int e = 0; e < NB; ++e) {
            const int q = tid + 256 * e;
            int r, k;
            coord(q, TN, ks.b_kcontig, r, k);
            if (q < TN * BK) Bs[buf][r * LD + k] = rb[e];
          }
        };

        load_stage(0);
        store_stage(0);
        __syncthreads();
        for (int s = 0; s < nsteps; ++s) {
          const int buf = s & 1;
          if (s + 1 < nsteps) load_stage(s + 1);
#pragma unroll
          for (int kk = 0; kk < BK / 4; ++kk) {
            const int kq = 4 * kk + (lane >> 4);
            double a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) a[i] = As[buf][(WTM * wm + 16 * i + (lane & 15)) * LD + kq];
#pragma unroll
            for (int j = 0; j < FN; ++j) b[j] = Bs[buf][(WTN * wn + 16 * j + (lane & 15)) * LD + kq];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
              for (int j = 0; j < FN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
          }
          if (s + 1 < nsteps) store_stage(buf ^ 1);
          __syncthreads();
        }
      }
    }

    // Epilogue: C(clabels) (=|+=) alpha * acc, straight into the block's own mode order.
    __syncthreads();
    for (int t = tid; t < TM + TN; t += 256) {
      if (t < TM) {
        int m = tk.m0 + t;
        offCm[t] = m < M ? radix_off(cs.cm, m) : -1;
      } else {
        int c = t - TM;
        int n = tk.n0 + c;
        offCn[c] = n < N ? radix_off(cs.cn, n) : -1;
      }
    }
    __syncthreads();
    double* Cblk = C + tk.c_off;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          // v_mfma_f64_16x16x4f64 result: register e of lane l holds row 4e + l/16, col l%16.
          int row = WTM * wm + 16 * i + 4 * e + (lane >> 4);
          int col = WTN * wn + 16 * j + (lane & 15);
          int om = offCm[row], on = offCn[col];
          if (om >= 0 && on >= 0) {
            double* dst = Cblk + om + on;
            double v = alpha * acc[i][j][e];
            *dst = is_assign ? v : *dst + v;
          }
        }
  }
}

int bp_tile_rows(int tile) { return tile == BP_TILE_16x256 ? 16 : 64; }
int bp_tile_cols(int tile) { return tile == BP_TILE_16x256 ? 256 : 64; }

void launch_bp_f64(const BPTask* tasks, int ntasks, const BPPair* pairs, const BPCShape* cshapes,
                   const BPKShape* kshapes, const double* A, const double* B, double* C,
                   double alpha, int is_assign, hipStream_t s, const BPXcd* xr, int tile) {
  if (ntasks <= 0) return;
  BPXcd r;
  if (xr) {
    r = *xr;
  } else {
    for (int x = 0; x < 8; ++x) {
      r.lo[x] = (int)((int64_t)ntasks * x / 8);
      r.hi[x] = (int)((int64_t)ntasks * (x + 1) / 8);
    }
  }
  int longest = 1;
  for (int x = 0; x < 8; ++x) {
    TMX_REQUIRE(0 <= r.lo[x] && r.lo[x] <= r.hi[x] && r.hi[x] <= ntasks, "bad XCD task ranges");
    longest = std::max(longest, r.hi[x] - r.lo[x]);
  }
  const int grid = 8 * std::min(longest, 65535);
  if (tile == BP_TILE_16x256)
    hipLaunchKernelGGL((k_bp_f64<16, 256, 1, 8>), dim3(grid), dim3(256), 0, s, tasks, ntasks, pairs,
                       cshapes, kshapes, A, B, C, alpha, is_assign, r);
  else
    hipLaunchKernelGGL((k_bp_f64<64, 64, 2, 16>), dim3(grid), dim3(256), 0, s, tasks, ntasks, pairs,
                       cshapes, kshapes, A, B, C, alpha, is_assign, r);
  TMX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Block assign / transpose-accumulate.
//
// Job j covers one (dst block, src block) pair and sh.tiles_o * sh.tiles_d * sh.tiles_s tiles.
// d = dst's contiguous label, s = src's contiguous label, outer = the remaining labels.
//   transpose case (d != s): a tile is 32 (along s) x 32 (along d); the source is read along s
//     and the destination written along d, the 32x32 block passing through LDS.
//   same case (d == s): both are contiguous along d; a tile is 1024 consecutive elements.
__global__ __launch_bounds__(256) void k_assign(const AssignJob* __restrict__ jobs, int njobs,
                                                const int64_t* __restrict__ tile_prefix,
                                                const AssignShape* __restrict__ shapes,
                                                const double* __restrict__ src,
                                                double* __restrict__ dst, double scale,
                                                int is_assign) {
  __shared__ double tile[32][33];
  const int64_t total = tile_prefix[njobs];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int64_t t = blockIdx.x; t < total; t += gridDim.x) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (tile_prefix[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const AssignJob jb = jobs[lo];
    const AssignShape& sh = shapes[jb.shape];
    int64_t local = t - tile_prefix[lo];
    const int ts = (int)(local % sh.tiles_s);
    local /= sh.tiles_s;
    const int td = (int)(local % sh.tiles_d);
    const int outer = (int)(local / sh.tiles_d);
    const double* sp = src + jb.src_off + radix_off(sh.src_outer, outer);
    double* dp = dst + jb.dst_off + radix_off(sh.dst_outer, outer);
    if (sh.same) {
      for (int e = threadIdx.x; e < 1024; e += 256) {
        int d = td * 1024 + e;
        if (d < sh.ed) {
          double v = scale * sp[d];
          dp[d] = is_assign ? v : dp[d] + v;
        }
      }
      continue;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
      int s_i = ts * 32 + tx, d_i = td * 32 + r;
      tile[r][tx] = (s_i < sh.es && d_i < sh.ed)
                        ? sp[(int64_t)s_i + (int64_t)d_i * sh.src_stride_d] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
      int d_i = td * 32 + tx, s_i = ts * 32 + r;
      if (s_i < sh.es && d_i < sh.ed) {
        double* q = dp + (int64_t)d_i + (int64_t)s_i * sh.dst_stride_s;
        double v = scale * tile[tx][r];
        *q = is_assign ? v : *q + v;
      }
    }
  }
}

void launch_assign(const AssignJob* jobs, int njobs, const int64_t* tile_prefix, int64_t ntiles,
                   const AssignShape* shapes, const double* src, double* dst, double scale,
                   int is_assign, hipStream_t s) {
  if (njobs <= 0 || ntiles <= 0) return;
  int grid = (int)std::min<int64_t>(ntiles, 65536);
  hipLaunchKernelGGL(k_assign, dim3(grid), dim3(256), 0, s, jobs, njobs, tile_prefix, shapes, src,
                     dst, scale, is_assign);
  TMX_HIP(hipGetLastError());
}

}  // namespace tmx
