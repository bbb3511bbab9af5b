// tmx_kernels.hpp -- kernel descriptors and launcher declarations.
#pragma once

#include "tmx_internal.hpp"

namespace tmx {

// Block assign (permutation) descriptors.
struct AssignShape {
  int32_t same;       // 1 when dst and src share their contiguous label
  int32_t ed, es;     // extents of the dst-contiguous (d) and src-contiguous (s) labels
  int32_t tiles_o, tiles_d, tiles_s;
  int64_t src_stride_d;  // src stride of label d (transpose case)
  int64_t dst_stride_s;  // dst stride of label s (transpose case)
  Radix src_outer, dst_outer;  // remaining labels -> src / dst offsets
};
struct AssignJob {
  int64_t src_off, dst_off;
  int32_t shape, pad;
};

// xr: XCD task ranges (null: the tasks split into 8 equal contiguous ranges).
// C-tile shapes of the fused block-pair kernel (rows x columns of one task).
enum { BP_TILE_64x64 = 0, BP_TILE_16x256 = 1 };
int bp_tile_rows(int tile);
int bp_tile_cols(int tile);
void launch_bp_f64(const BPTask* tasks, int ntasks, const BPPair* pairs, const BPCShape* cshapes,
                   const BPKShape* kshapes, const double* A, const double* B, double* C,
                   double alpha, int is_assign, hipStream_t s, const BPXcd* xr = nullptr,
                   int tile = BP_TILE_64x64);
void launch_assign(const AssignJob* jobs, int njobs, const int64_t* tile_prefix, int64_t ntiles,
                   const AssignShape* shapes, const double* src, double* dst, double scale,
                   int is_assign, hipStream_t s);

// Element-wise kernels (k_elementwise.hip)
void launch_fill(double* p, int64_t n, double v, hipStream_t s);
void launch_fill_bytes_zero(void* p, size_t bytes, hipStream_t s);
// Zeroes n element ranges of C: range i starts at off[i] and has prefix[i+1] - prefix[i]
// elements (device arrays; total = prefix[n]).
void launch_zero_ranges(const int64_t* off, const int64_t* prefix, int n, int64_t total, double* C,
                        hipStream_t s);
void launch_scale(double* p, int64_t n, double a, hipStream_t s);
void launch_axpy(int64_t n, double a, const double* x, double* y, hipStream_t s);
void launch_scale_copy(int64_t n, double a, const double* x, double* y, hipStream_t s);
// Partial sums of squares / products / squared differences -> host result (synchronous).
double reduce_sumsq(const double* x, int64_t n, hipStream_t s);
double reduce_dot(const double* x, const double* y, int64_t n, hipStream_t s);
double reduce_diffsq(const double* x, const double* y, int64_t n, hipStream_t s);
// reduce_sumsq with its partial sums in a file-scope device array (no allocation); the
// array's address on `device` is taken by noalloc_setup, which tmx_init calls.
void noalloc_setup(int device);
double reduce_sumsq_noalloc(const double* x, int64_t n, hipStream_t s, int device);

struct GenBlock {  // one block of a generated tensor
  int64_t off;
  int32_t lo[MAXR];   // global index of the block's first element per mode
  int32_t ext[MAXR];  // block extent per mode
};
void launch_fill_normal(const GenBlock* blocks, int64_t nblocks, int nmodes, const int64_t* mode_n,
                        int64_t max_block, double* data, uint64_t seed, uint64_t tensor_id,
                        double scale, int symmetrize, hipStream_t s);
void launch_denominator(const GenBlock* blocks, int64_t nblocks, int64_t total, double* data,
                        const double* e_occ, const double* e_vir, hipStream_t s);

}  // namespace tmx
