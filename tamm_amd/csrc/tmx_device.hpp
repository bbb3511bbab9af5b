// tmx_device.hpp -- small device helpers shared by the kernels.
#pragma once

#include "tmx_internal.hpp"

namespace tmx {

// Mixed-radix offset of a row-major index.
__device__ __forceinline__ int radix_off(const Radix& r, int idx) {
  int off = 0;
  for (int d = r.n - 1; d >= 0; --d) {
    int e = r.ext[d];
    int q = idx / e;
    off += (idx - q * e) * r.stride[d];
    idx = q;
  }
  return off;
}

// Incremental mixed-radix counter over consecutive indices.
struct RadixCounter {
  int digit[MAXR];
  int off;
  __device__ __forceinline__ void init(const Radix& r, int idx) {
    off = 0;
    for (int d = r.n - 1; d >= 0; --d) {
      int e = r.ext[d];
      int q = idx / e;
      digit[d] = idx - q * e;
      off += digit[d] * r.stride[d];
      idx = q;
    }
  }
  __device__ __forceinline__ void next(const Radix& r) {
    int d = r.n - 1;
    if (d < 0) return;
    ++digit[d];
    off += r.stride[d];
    while (d > 0 && digit[d] == r.ext[d]) {
      off -= r.ext[d] * r.stride[d];
      digit[d] = 0;
      --d;
      ++digit[d];
      off += r.stride[d];
    }
  }
};

// Job containing work item t, scanning forward from a hint (jobs are item-prefix ordered).
__device__ __forceinline__ int job_of(const int64_t* __restrict__ prefix, int j, int64_t t) {
  while (prefix[j + 1] <= t) ++j;
  return j;
}

// Element (r, c) of an FP64 class matrix in v_mfma_f64_16x16x4f64 fragment-major order:
// 16x4 fragments of 64 doubles; lane l of a fragment holds row l%16, column l/16.
__device__ __forceinline__ int64_t frag_f64(int r, int c, int kpad) {
  return ((int64_t)(r >> 4) * (kpad >> 2) + (c >> 2)) * 64 + (c & 3) * 16 + (r & 15);
}

// Byte (r, c) of an int8 class matrix in v_mfma_i32_32x32x32_i8 fragment-major order:
// 32x32 fragments of 1024 bytes; lane l holds row l%32, columns 16*(l/32) .. +15.
__device__ __forceinline__ int64_t frag_i8(int r, int c, int kpad) {
  return ((int64_t)(r >> 5) * (kpad >> 5) + (c >> 5)) * 1024 + (((((c & 31) >> 4) << 5) + (r & 31)) << 4) +
         (c & 15);
}

}  // namespace tmx
