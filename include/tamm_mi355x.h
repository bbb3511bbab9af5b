/*
 * tamm_mi355x.h -- C ABI of the MI355X block-sparse contraction engine.
 *
 * The engine executes TAMM's labeled contractions
 *     C(clabels) += alpha * A(alabels) * B(blabels)
 * on one MI355X (gfx950) with tensors resident in HBM in TAMM's block layout
 * (Distribution_NW: nonzero blocks sorted by block key, packed contiguously,
 * row-major inside a block; src/tamm/distribution.hpp:197-365).
 *
 * Conventions
 *   - Every function returns TMX_OK (0) or an error code; tmx_last_error()
 *     gives the message of the last failure on the calling thread.  No C++
 *     exception crosses this boundary.
 *   - `stream` arguments are hipStream_t values passed as void*; NULL selects
 *     the engine's own stream (created by tmx_init, like TAMM's one-stream
 *     GPUStreamPool, src/tamm/gpu_streams.hpp:353-428).
 *   - Data pointers are device pointers (HBM) unless stated otherwise.
 *   - Labels are ints; equal ints name the same index label (TAMM's
 *     fillin_int_labels convention, src/tamm/multop.hpp:955-970).
 */
#ifndef TAMM_MI355X_H
#define TAMM_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TMX_OK = 0,
  TMX_ERR_INVALID = 1,     /* bad argument or descriptor (TAMM: EXPECTS throws) */
  TMX_ERR_HIP = 2,         /* HIP runtime failure (TAMM: HIP_CHECK throws) */
  TMX_ERR_OOM = 3,         /* device pool exhausted (TAMM: tamm_terminate) */
  TMX_ERR_UNSUPPORTED = 4, /* valid request this engine does not execute */
  TMX_ERR_STATE = 5        /* engine not initialised / already initialised */
};

/* How a contraction is executed. */
enum {
  TMX_EXEC_BLOCKPAIR_F64 = 0, /* TAMM's loop structure: one fused kernel per block pair */
  TMX_EXEC_GROUPED_F64 = 1,   /* one grouped launch over all block pairs, fused permutes */
  TMX_EXEC_SECTOR_F64 = 2,    /* block pairs gathered into dense sector GEMMs, FP64 MFMA */
  TMX_EXEC_SPLIT_I8 = 3,      /* sector GEMMs on int8 residues (Ozaki scheme II), FP64 result */
  TMX_EXEC_AUTO = 4,          /* engine picks, FP64: SECTOR_F64 when eligible and the average
                                 reduction length per output element is >= 128, else GROUPED_F64 */
  TMX_EXEC_AUTO_SPLIT = 5     /* as AUTO, with SPLIT_I8 in place of SECTOR_F64 where the GEMMs
                                 are MFMA-bound: memory-bound contractions stay FP64, those
                                 with short reductions on the grouped path and those with thin
                                 products (flops-weighted M N / (M + N) of the sector GEMMs
                                 < 32, e.g. one-body intermediates whose classes are one O or
                                 V tile) on the sector path */
};

enum { TMX_H2D = 0, TMX_D2H = 1, TMX_D2D = 2 };

typedef struct tmx_space_s* tmx_space;
typedef struct tmx_tensor_s* tmx_tensor;
typedef struct tmx_contraction_s* tmx_contraction;
typedef struct tmx_addop_s* tmx_addop;
typedef struct tmx_level_s* tmx_level;

/* ---- runtime (replaces tamm::initialize's GPU part, src/tamm/tamm.cpp:36-52) ---- */
const char* tmx_last_error(void);
const char* tmx_version(void);
/* Select `device`, create the engine stream and one device pool of
 * pool_fraction x free HBM.  pool_fraction <= 0 reads TAMM_GPU_POOL (percent,
 * default 80) as RMMMemoryManager::initialize does (src/tamm/rmm_memory_pool.hpp:230-237). */
int tmx_init(int device, double pool_fraction);
int tmx_finalize(void);
int tmx_get_stream(void** stream);
int tmx_synchronize(void);
/* Device pool (replaces RMMMemoryManager::getDeviceMemoryPool, rmm_memory_pool.hpp:73-85). */
int tmx_malloc(size_t bytes, void** ptr);
int tmx_free(void* ptr);
int tmx_pool_info(size_t* capacity, size_t* in_use, size_t* peak);
int tmx_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream);
int tmx_memset_zero(void* ptr, size_t bytes, void* stream); /* fill kernel, graph safe */

/* ---- index spaces (TiledIndexSpace{IndexSpace, tile(s)}, tiled_index_space.hpp:31,49) ----
 * use_tile_vector = 0: tiles[0] is the single tile size (tiled_index_space.hpp:1135-1194);
 * use_tile_vector = 1: tiles[0..ntiles) is the explicit tile vector (:1205-1260).
 * Spin / irrep attribute ranges and named subspace ranges are [lo, hi). */
int tmx_space_create(int64_t n, int use_tile_vector, const int64_t* tiles, int64_t ntiles,
                     int nspin, const int64_t* spin_lo, const int64_t* spin_hi, const int* spin_val,
                     int nirrep, const int64_t* irrep_lo, const int64_t* irrep_hi,
                     const int* irrep_val, int nnamed, const int64_t* named_lo,
                     const int64_t* named_hi, tmx_space* out);
int tmx_space_destroy(tmx_space s);
int tmx_space_ntiles(tmx_space s, int64_t* ntiles);
int tmx_space_tiles(tmx_space s, int64_t* offsets /*ntiles+1*/, int* spin, int* irrep);

/* ---- tensors: block structure and layout (Tensor<T>{{tis...},{nupper,nlower}},
 * tensor.hpp:115-134; is_non_zero tensor_base.hpp:243-263; Distribution_NW).
 * nupper < 0: no spin mask.  irrep_rule != 0: also require XOR of the C2v irreps
 * of the non-ignored modes to be 0 (extension). ---- */
int tmx_tensor_create(int nmodes, const tmx_space* spaces, int nupper, int nlower,
                      int irrep_rule, tmx_tensor* out);
int tmx_tensor_destroy(tmx_tensor t);
int tmx_tensor_size(tmx_tensor t, int64_t* nelems);
int tmx_tensor_nblocks(tmx_tensor t, int64_t* nblocks);
int tmx_tensor_blocks(tmx_tensor t, int64_t* keys, int64_t* offsets, int64_t* sizes);
/* Offset/size of a block (Distribution_NW::locate, distribution.hpp:270-285); offset -1 for a
 * zero block. */
int tmx_tensor_locate(tmx_tensor t, const int64_t* blockid, int64_t* offset, int64_t* size);

/* ---- labeled contraction (MultOp::execute_bufacc, multop.hpp:576-893) ----
 * tile_offset[i]: first tile of the label's space inside mode i's space (label on a named
 * subspace, internal::translate_blockid, utils.hpp:263-286); NULL = 0.
 * label_ntiles[i]: tile count of the label's space; NULL = the mode's tile count. */
typedef struct {
  tmx_tensor tensor;
  const int* labels;
  const int64_t* tile_offset;
  const int64_t* label_ntiles;
} tmx_operand;

int tmx_contraction_create(const tmx_operand* C, double alpha, const tmx_operand* A,
                           const tmx_operand* B, int exec_mode, tmx_contraction* out);
/* C += alpha * A * B on `stream` (asynchronous; graph-capturable for GROUPED/SECTOR/SPLIT:
 * no allocation or memset happens inside). */
int tmx_contraction_execute(tmx_contraction c, const double* A, const double* B, double* C,
                            void* stream);
/* flops: algorithmic 2*M*N*K over nonzero block triples; npairs: nonzero block pairs;
 * workspace: device bytes held by the plan; mode: resolved execution mode. */
int tmx_contraction_info(tmx_contraction c, double* flops, int64_t* npairs,
                         int64_t* workspace_bytes, int* mode);
/* Split-precision controls (TMX_EXEC_SPLIT_I8): bits kept per operand element and the
 * number of int8 moduli (0 = smallest count that keeps the product exact). */
int tmx_contraction_split_info(tmx_contraction c, int* bits_a, int* bits_b, int* nmoduli);
int tmx_contraction_destroy(tmx_contraction c);
/* Executes once with HIP events between the stages on `stream` and returns their durations
 * (synchronous): stage_ms[0] = operand packing, [1] = GEMM, [2] = scatter into C.  Fused
 * block-pair modes report their single kernel as stage 1. */
int tmx_contraction_execute_timed(tmx_contraction c, const double* A, const double* B, double* C,
                                  void* stream, float* stage_ms);
/* As tmx_contraction_execute_timed with the packing split by operand: stage_ms[0] = packing of
 * A, [1] = packing of B, [2] = GEMM, [3] = scatter into C, [4] = the longest single GEMM launch
 * of the GEMM stage (five floats).  No counterpart in TAMM (its OpProfiler times whole ops,
 * op_profiler.hpp:5-38). */
int tmx_contraction_execute_stages(tmx_contraction c, const double* A, const double* B, double* C,
                                   void* stream, float* stage_ms);
/* The first n (1..6) of: the five stage times above, then [5] = the main packing pass of A
 * alone (the packing kernel, without the row-maximum pass or the row-scale reuse bookkeeping
 * launched around it; 0 when A was not packed, e.g. a constant operand already packed). */
int tmx_contraction_execute_stages_n(tmx_contraction c, const double* A, const double* B, double* C,
                                     void* stream, float* stage_ms, int n);
/* Executes once (C += alpha*A*B, as tmx_contraction_execute) and then autotunes the plan:
 * for each sector GEMM every tile shape is timed on the packed operands and the fastest kept,
 * the plan's current shape unless another is more than 1% faster; for a plan of several
 * sectors the whole GEMM stage is also timed with every sector on one common shape (one
 * launch), which replaces the per-sector choice when more than 1% faster.  Each candidate's
 * time is the shortest of three runs (split results do not change: the int32 sums are exact
 * in any tile order).  Unless a
 * split-K count is fixed (tmx_set_parallel_split), plans that take over 0.5 ms are then
 * rebuilt with 1..4 split-K parts per sector, each timed end to end into scratch memory, and
 * the fastest kept when it beats the current plan by 2% (partial sums are reassociated:
 * results stay within the mode's tolerance).  Choices are
 * remembered per sector geometry for later plans in the process and, when the environment
 * variable TAMM_MI355X_TUNE_FILE names a file, loaded from / saved to it.  *changed (may be
 * null) receives the number of sectors whose shape changed, plus one when the plan was
 * replaced by one with other split-K counts.  Synchronous.  FP64 sector plans choose among the
 * FP64 tile shapes the same way; grouped and per-pair block-pair plans have a single kernel
 * shape: they only execute.  Refused (TMX_ERR_STATE) while a captured graph or a level holds
 * the contraction.  No counterpart in TAMM
 * (its GPU path has no tile choices); it serves the batched path's tile/schedule selection. */
int tmx_contraction_autotune(tmx_contraction c, const double* A, const double* B, double* C,
                             void* stream, int* changed);
/* Constant operands (sector and split plans).  With a_constant (b_constant) = 1 the caller
 * promises that A (B) holds the same values at every execution that passes the same address,
 * as the integrals V, W, X of a coupled-cluster iteration do: the plan packs that operand once
 * (FP64 class matrices, or residue planes and row scales) at the next execution and reuses the
 * packed copy while the address stays the same.  Every call also drops the kept copies, so
 * calling it again after the operand's values changed makes the next execution pack afresh.
 * Refused (TMX_ERR_STATE) while a captured graph holds the contraction; a graph captured after
 * the packing replays without it.  Block-pair plans read operands in place and ignore the
 * flags.  No counterpart in TAMM, which re-reads every block at every execution. */
int tmx_contraction_set_constant(tmx_contraction c, int a_constant, int b_constant);
/* assign != 0: later executions compute C = alpha*A*B instead of C += alpha*A*B, TAMM's
 * `C(..) = alpha*A(..)*B(..)` (a MultOp with is_assign, multop.hpp:207-239).  TAMM runs it as a
 * zero fill of C's addressed blocks followed by `+=` (canonicalize, multop.hpp:255-269); the
 * engine fuses the fill: the scatter (or the block-pair epilogue) writes instead of adding, and
 * only the addressed blocks no block pair contributes to are zeroed.  Results equal the fill
 * followed by accumulation (up to the sign of zero).  Refused (TMX_ERR_STATE) while a captured
 * graph holds the contraction. */
int tmx_contraction_set_assign(tmx_contraction c, int assign);
/* ---- a Scheduler level of contractions (scheduler.hpp:230-264) ----
 * tmx_level_execute runs the terms in order, as n tmx_contraction_execute calls with the
 * given operand pointers would, except that a sector or split term whose product (packed
 * operands and GEMM) is that of an earlier term of the level runs only its scatter, from
 * that term's product: e.g. a CC ring term and its P(ab), P(ij) and P(ab)P(ij) images pack
 * and multiply W and T once and scatter four times into R.  Products are matched when the
 * level is created (same operand tensors, pack jobs, sector geometry and precision) and the
 * reuse is taken only when the operand pointers are the same and no term in between wrote
 * into them.  The level holds its contractions (destroying their handles is fine).  Results
 * are bit-identical to the separate executions.  No counterpart in TAMM, whose levels run
 * each op on its own (scheduler.hpp:230-264). */
int tmx_level_create(const tmx_contraction* terms, int n, tmx_level* out);
int tmx_level_products(tmx_level level, int* nproducts); /* distinct products (pack + GEMM) */
int tmx_level_execute(tmx_level level, const double* const* A, const double* const* B,
                      double* const* C, void* stream);
int tmx_level_destroy(tmx_level level);
/* Default split-precision bits per operand for contractions created afterwards. */
int tmx_set_split_bits(int bits);
/* Split-K parts per sector for sector plans created afterwards: 0 (default) lets the launch
 * model choose per sector, up to 256 parts (simulated XCD list schedule of the GEMM tasks
 * against the traffic of the extra partial results, which the scatter sums into C in one
 * pass); 1..16 fixes the count (still at least what the split path's
 * exact int32 accumulation needs, and parts of at least 1024 columns).  Results agree to the
 * mode's tolerance for every count (FP64 sums are reassociated).  A contraction keeps the
 * value it was created with; created with a fixed count, its autotuner keeps that count.
 * No counterpart in TAMM. */
int tmx_set_parallel_split(int parts);
/* Row-scale reuse of split-precision packing (default 1 = on).  Split packing scales each
 * class-matrix row by its maximum |x|, which normally takes a read of the operand of its own
 * before the packing pass.  With reuse on, an operand packed by the plan's previous execution
 * from the same address is packed with that execution's row maxima while the pass recomputes
 * them; when any maximum differs the operand is packed again with the new ones (the second
 * pass is gated on the device, so execution stays asynchronous and capturable).  Results are
 * bit-identical either way; an operand that changes between executions pays one extra pack,
 * after which reuse is paused for it for 8 executions.  No counterpart in TAMM. */
int tmx_set_scale_reuse(int on);
/* Schedule of a sector plan: split-K rounds (most parts of any sector), GEMM launches (one
 * per tile shape) and GEMM tasks; 1, 1, 0 for other plans. */
int tmx_contraction_schedule_info(tmx_contraction c, int* rounds, int* launches, int64_t* tasks);

/* ---- tensor I/O in TAMM's HDF5 format (write_to_disk / read_from_disk,
 * tamm_io.hpp:223-460, :735-980): a 1-D dataset "tensor" of doubles with the tensor's dense
 * extent, holding the nonzero blocks back to back in row-major block order from offset 0
 * (the Distribution_NW buffer).  `data` is the tensor's device buffer; the transfer is staged
 * through host memory and is synchronous.  libhdf5 is loaded on first use;
 * TMX_ERR_UNSUPPORTED if it is not available.  Reading checks the dense extent. ---- */
int tmx_tensor_write_hdf5(tmx_tensor t, const double* data, const char* path, void* stream);
int tmx_tensor_read_hdf5(tmx_tensor t, double* data, const char* path, void* stream);

/* ---- block kernels (kernels::block_multiply multiply.hpp:212-220, kernels::assign
 * assign.hpp:466-485) ---- */
/* C(clabels) (=|+=) alpha * sum over RHS-only labels of A(alabels) * B(blabels), one block
 * each, fused permute-load -> FP64 MFMA -> permute-store.  is_assign != 0 overwrites C. */
int tmx_block_multiply(double alpha, const double* A, int andim, const int64_t* adims,
                       const int* alabels, const double* B, int bndim, const int64_t* bdims,
                       const int* blabels, double* C, int cndim, const int64_t* cdims,
                       const int* clabels, int is_assign, void* stream);
/* dst(dlabels) (=|+=) scale * src(slabels); labels must be a permutation of each other. */
int tmx_block_assign(double* dst, int ndim, const int64_t* ddims, const int* dlabels,
                     double scale, const double* src, const int64_t* sdims, const int* slabels,
                     int is_assign, void* stream);

/* ---- tensor-level AddOp (addop.hpp:242-271): C(cl) (=|+=) alpha * A(al) ---- */
int tmx_addop_create(const tmx_operand* C, double alpha, const tmx_operand* A, int is_assign,
                     tmx_addop* out);
int tmx_addop_execute(tmx_addop op, const double* A, double* C, void* stream);
int tmx_addop_destroy(tmx_addop op);

/* ---- element-wise (SetOp, fill_tensor, norm, update_tensor; tamm_utils.hpp) ---- */
int tmx_fill(double* data, int64_t n, double value, void* stream);
int tmx_scale(double* data, int64_t n, double alpha, void* stream);
int tmx_axpy(int64_t n, double alpha, const double* x, double* y, void* stream);
/* scale * N(0,1) from include/tmx_rng.h keyed by the dense element index. */
int tmx_fill_normal(tmx_tensor t, double* data, uint64_t seed, uint64_t tensor_id, double scale,
                    int symmetrize_pairs, void* stream);
/* Synchronous reductions on `stream` (result copied to host). */
int tmx_norm(const double* data, int64_t n, double* result, void* stream);
int tmx_dot(const double* x, const double* y, int64_t n, double* result, void* stream);
/* tmx_norm without allocation: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under graph capture, tight serving loops, harnesses that forbid
 * it).  Its partial sums (at most 2048) and the finished sum go to a fixed 16 KB device array
 * of the library (one per GPU, held until the process ends) instead of pool memory; the result
 * is bit-identical to tmx_norm's.  Calls that overlap on one GPU share that array and race.
 * Synchronous, like tmx_norm. */
int tmx_norm_noalloc(const double* data, int64_t n, double* result, void* stream);
/* Relative Frobenius difference ||x - y|| / ||y|| (synchronous). */
int tmx_rel_diff(const double* x, const double* y, int64_t n, double* result, void* stream);
/* Doubles-residual update R(a,b,i,j) /= (e_i + e_j - e_a - e_b) for a tensor over (V,V,O,O);
 * e_occ / e_vir are device arrays indexed by the global orbital index of each space. */
int tmx_energy_denominator(tmx_tensor R, double* data, const double* e_occ,
                           const double* e_vir, void* stream);

/* ---- timing and graphs ---- */
int tmx_event_create(void** ev);
int tmx_event_record(void* ev, void* stream);
int tmx_event_elapsed_ms(void* start, void* stop, float* ms); /* synchronises on stop */
int tmx_event_destroy(void* ev);
int tmx_stream_create(void** stream); /* default priority */
int tmx_stream_destroy(void* stream);
int tmx_stream_synchronize(void* stream);
int tmx_graph_begin(void* stream);
int tmx_graph_end(void* stream, void** graph_exec);
int tmx_graph_launch(void* graph_exec, void* stream);
int tmx_graph_destroy(void* graph_exec);

#ifdef __cplusplus
}
#endif

#endif /* TAMM_MI355X_H */
