// k_elementwise.hip -- fills, scaling, reductions, synthetic inputs, energy denominator.
//
// These cover TAMM's SetOp (C() = v), scale/axpy updates, norm (tamm_utils.hpp:1197-1254:
// sqrt of the sum of squares over stored blocks) and the element-wise update of a doubles
// residual by orbital-energy denominators.  All are HBM-bound streaming kernels with
// 16-byte accesses and a grid sized to a few waves per CU.
#include "tmx_kernels.hpp"
#include "tmx_rng.h"

namespace tmx {

namespace {
inline int grid_for(int64_t n, int per_block) {
  int64_t g = (n + per_block - 1) / per_block;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}
}  // namespace

__global__ __launch_bounds__(256) void k_fill(double* __restrict__ p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}
__global__ __launch_bounds__(256) void k_zero_bytes(uint8_t* __restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0;
}
__global__ __launch_bounds__(256) void k_zero_u4(uint4* __restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    p[i] = make_uint4(0, 0, 0, 0);
}
__global__ __launch_bounds__(256) void k_scale(double* __restrict__ p, int64_t n, double a) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] *= a;
}
__global__ __launch_bounds__(256) void k_axpy(int64_t n, double a, const double* __restrict__ x,
                                              double* __restrict__ y) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    y[i] += a * x[i];
}

void launch_fill(double* p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256 * 4)), dim3(256), 0, s, p, n, v);
  TMX_HIP(hipGetLastError());
}
// Zeroes the element ranges [off[i], off[i] + prefix[i+1] - prefix[i]) of C (prefix: running
// sums of the range lengths, prefix[0] = 0, prefix[n] = total): the blocks of an assigning
// contraction that no block pair writes.
__global__ __launch_bounds__(256) void k_zero_ranges(const int64_t* __restrict__ off,
                                                     const int64_t* __restrict__ prefix, int n,
                                                     double* __restrict__ C) {
  const int64_t total = prefix[n];
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= e) lo = mid;
      else hi = mid - 1;
    }
    C[off[lo] + (e - prefix[lo])] = 0.0;
  }
}
void launch_zero_ranges(const int64_t* off, const int64_t* prefix, int n, int64_t total, double* C,
                        hipStream_t s) {
  if (n <= 0 || total <= 0) return;
  hipLaunchKernelGGL(k_zero_ranges, dim3(grid_for(total, 256 * 4)), dim3(256), 0, s, off, prefix, n, C);
  TMX_HIP(hipGetLastError());
}
void launch_fill_bytes_zero(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  uintptr_t a = reinterpret_cast<uintptr_t>(p);
  if (a % 16 == 0 && bytes % 16 == 0) {
    hipLaunchKernelGGL(k_zero_u4, dim3(grid_for((int64_t)(bytes / 16), 1024)), dim3(256), 0, s,
                       static_cast<uint4*>(p), bytes / 16);
  } else {
    hipLaunchKernelGGL(k_zero_bytes, dim3(grid_for((int64_t)bytes, 1024)), dim3(256), 0, s,
                       static_cast<uint8_t*>(p), bytes);
  }
  TMX_HIP(hipGetLastError());
}
void launch_scale(double* p, int64_t n, double a, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n, 1024)), dim3(256), 0, s, p, n, a);
  TMX_HIP(hipGetLastError());
}
void launch_axpy(int64_t n, double a, const double* x, double* y, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n, 1024)), dim3(256), 0, s, n, a, x, y);
  TMX_HIP(hipGetLastError());
}

// y = a * x (an assigning AddOp over whole tensors: one read and one write per element).
__global__ __launch_bounds__(256) void k_scale_copy(int64_t n, double a, const double* __restrict__ x,
                                                    double* __restrict__ y) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    y[i] = a * x[i];
}

void launch_scale_copy(int64_t n, double a, const double* x, double* y, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale_copy, dim3(grid_for(n, 1024)), dim3(256), 0, s, n, a, x, y);
  TMX_HIP(hipGetLastError());
}

// ---- reductions: per-block partial sums in FP64, finished on the host in FP64 ----
template <int OP>
__global__ __launch_bounds__(256) void k_reduce(const double* __restrict__ x,
                                                const double* __restrict__ y, int64_t n,
                                                double* __restrict__ partial) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double a = x[i];
    if (OP == 0) s += a * a;
    else if (OP == 1) s += a * y[i];
    else { double d = a - y[i]; s += d * d; }
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

template <int OP>
static double reduce(const double* x, const double* y, int64_t n, hipStream_t s) {
  if (n <= 0) return 0.0;
  int g = std::min(grid_for(n, 1024), 2048);
  DevBuf part(sizeof(double) * g);
  hipLaunchKernelGGL(k_reduce<OP>, dim3(g), dim3(256), 0, s, x, y, n, part.as<double>());
  TMX_HIP(hipGetLastError());
  std::vector<double> h(g);
  TMX_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * g, hipMemcpyDeviceToHost, s));
  TMX_HIP(hipStreamSynchronize(s));  // `part` returns to the pool only after this
  double r = 0.0;
  for (double v : h) r += v;
  return r;
}
// ---- norm without allocation (tmx_norm_noalloc) ----
// The partial sums of k_reduce (at most kNoAllocPartials workgroups) and the finished sum go
// to a file-scope device array instead of pool memory.  Each GPU holds its own copy of the
// array until the process ends; its address on a device comes from hipGetSymbolAddress, taken
// once per device by noalloc_setup (called from tmx_init with that device current).  Calls
// that overlap on one GPU share the array and race.
constexpr int kNoAllocPartials = 2048;
__device__ __align__(256) double scratch[kNoAllocPartials + 1];  // partials, then the sum
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};

void noalloc_setup(int device) {
  TMX_REQUIRE(device >= 0 && device < kMaxDevices, "device index beyond the scratch table");
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  TMX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<double*>(p);
}

// Sums the partials in index order on one thread: the same additions, in the same order, as
// reduce()'s host loop, so both entry points return the same bits.
__global__ void k_reduce_finish(const double* __restrict__ partial, int g, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double r = 0.0;
    for (int i = 0; i < g; ++i) r += partial[i];
    *out = r;
  }
}

double reduce_sumsq_noalloc(const double* x, int64_t n, hipStream_t s, int device) {
  TMX_REQUIRE(device >= 0 && device < kMaxDevices && g_scratch_ptr[device],
              "no-allocation scratch not set up on this device (tmx_init)");
  if (n <= 0) return 0.0;
  double* sc = g_scratch_ptr[device];
  const int g = std::min(grid_for(n, 1024), kNoAllocPartials);
  hipLaunchKernelGGL(k_reduce<0>, dim3(g), dim3(256), 0, s, x, nullptr, n, sc);
  hipLaunchKernelGGL(k_reduce_finish, dim3(1), dim3(64), 0, s, sc, g, sc + kNoAllocPartials);
  TMX_HIP(hipGetLastError());
  double r = 0.0;
  TMX_HIP(hipMemcpyAsync(&r, sc + kNoAllocPartials, sizeof(double), hipMemcpyDeviceToHost, s));
  TMX_HIP(hipStreamSynchronize(s));
  return r;
}

double reduce_sumsq(const double* x, int64_t n, hipStream_t s) { return reduce<0>(x, nullptr, n, s); }
double reduce_dot(const double* x, const double* y, int64_t n, hipStream_t s) { return reduce<1>(x, y, n, s); }
double reduce_diffsq(const double* x, const double* y, int64_t n, hipStream_t s) { return reduce<2>(x, y, n, s); }

// ---- synthetic inputs: one workgroup per block, elements in block (row-major) order ----
struct GenDims {
  int nmodes;
  int64_t n[MAXR];
};

__global__ __launch_bounds__(256) void k_fill_normal(const GenBlock* __restrict__ blocks,
                                                     int64_t nblocks, GenDims gd,
                                                     double* __restrict__ data, uint64_t seed,
                                                     uint64_t tensor_id, double scale,
                                                     int symmetrize) {
  for (int64_t b = blockIdx.y * (int64_t)gridDim.x + blockIdx.x; b < nblocks;
       b += (int64_t)gridDim.x * gridDim.y) {
    const GenBlock gb = blocks[b];
    int64_t size = 1;
    for (int d = 0; d < gd.nmodes; ++d) size *= gb.ext[d];
    for (int64_t e = threadIdx.x; e < size; e += 256) {
      int64_t rem = e;
      int64_t idx[MAXR];
      for (int d = gd.nmodes - 1; d >= 0; --d) {
        idx[d] = gb.lo[d] + rem % gb.ext[d];
        rem /= gb.ext[d];
      }
      int64_t g = 0;
      for (int d = 0; d < gd.nmodes; ++d) g = g * gd.n[d] + idx[d];
      double z = tmx_rng_normal(seed, tensor_id, (uint64_t)g);
      if (symmetrize && gd.nmodes == 4) {
        int64_t g2 = ((idx[1] * gd.n[0] + idx[0]) * gd.n[3] + idx[3]) * gd.n[2] + idx[2];
        double z2 = tmx_rng_normal(seed, tensor_id, (uint64_t)g2);
        z = 0.5 * (z + z2);
      }
      data[gb.off + e] = scale * z;
    }
  }
}

void launch_fill_normal(const GenBlock* blocks, int64_t nblocks, int nmodes, const int64_t* mode_n,
                        int64_t /*max_block*/, double* data, uint64_t seed, uint64_t tensor_id,
                        double scale, int symmetrize, hipStream_t s) {
  if (nblocks <= 0) return;
  GenDims gd;
  std::memset(&gd, 0, sizeof(gd));
  gd.nmodes = nmodes;
  for (int i = 0; i < nmodes; ++i) gd.n[i] = mode_n[i];
  int gx = (int)std::min<int64_t>(nblocks, 65535);
  int gy = (int)std::min<int64_t>((nblocks + gx - 1) / gx, 64);
  hipLaunchKernelGGL(k_fill_normal, dim3(gx, gy), dim3(256), 0, s, blocks, nblocks, gd, data, seed,
                     tensor_id, scale, symmetrize);
  TMX_HIP(hipGetLastError());
}

// R(a,b,i,j) /= (e_i + e_j - e_a - e_b); blocks of a (V,V,O,O) tensor, lo = global indices.
// A workgroup takes a chunk of kDenomChunk consecutive elements of the tensor, wherever the
// block boundaries fall, so large blocks spread over many workgroups.  The chunk's first block
// comes from a binary search over the blocks' offsets (ascending); a thread steps to the next
// block when its element passes the current one's end.
constexpr int kDenomChunk = 4096;
__global__ __launch_bounds__(256) void k_denominator(const GenBlock* __restrict__ blocks,
                                                     int64_t nblocks, int64_t total,
                                                     double* __restrict__ data,
                                                     const double* __restrict__ eo,
                                                     const double* __restrict__ ev) {
  auto block_size = [](const GenBlock& gb) {
    return (int64_t)gb.ext[0] * gb.ext[1] * gb.ext[2] * gb.ext[3];
  };
  for (int64_t c0 = (int64_t)blockIdx.x * kDenomChunk; c0 < total; c0 += (int64_t)gridDim.x * kDenomChunk) {
    int64_t lo = 0, hi = nblocks - 1;  // the last block starting at or before c0
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) / 2;
      if (blocks[mid].off <= c0) lo = mid;
      else hi = mid - 1;
    }
    int64_t b = lo;
    GenBlock gb = blocks[b];
    int64_t end = gb.off + block_size(gb);
    for (int k = 0; k < kDenomChunk / 256; ++k) {
      const int64_t g = c0 + k * 256 + threadIdx.x;
      if (g >= total) break;
      while (g >= end && b + 1 < nblocks) {
        gb = blocks[++b];
        end = gb.off + block_size(gb);
      }
      if (g < gb.off || g >= end) continue;
      // a block holds fewer than 2^31 elements: its indices decompose in 32-bit arithmetic
      const uint32_t n1 = (uint32_t)gb.ext[1], n2 = (uint32_t)gb.ext[2], n3 = (uint32_t)gb.ext[3];
      uint32_t r = (uint32_t)(g - gb.off);
      const uint32_t j = r % n3;
      r /= n3;
      const uint32_t i = r % n2;
      r /= n2;
      const uint32_t bb = r % n1, a = r / n1;
      const double d = eo[gb.lo[2] + i] + eo[gb.lo[3] + j] - ev[gb.lo[0] + a] - ev[gb.lo[1] + bb];
      data[g] /= d;
    }
  }
}

void launch_denominator(const GenBlock* blocks, int64_t nblocks, int64_t total, double* data,
                        const double* e_occ, const double* e_vir, hipStream_t s) {
  if (nblocks <= 0 || total <= 0) return;
  const int g = (int)std::min<int64_t>((total + kDenomChunk - 1) / kDenomChunk, 16384);
  hipLaunchKernelGGL(k_denominator, dim3(g), dim3(256), 0, s, blocks, nblocks, total, data, e_occ, e_vir);
  TMX_HIP(hipGetLastError());
}

}  // namespace tmx
