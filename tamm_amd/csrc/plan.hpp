// plan.hpp -- host-side contraction planning (label analysis, block-pair enumeration).
#pragma once

#include "tmx_kernels.hpp"

namespace tmx {

// Label classes of a contraction C(cl) += alpha * A(al) * B(bl), as kernels::block_multiply
// classifies them (src/tamm/kernels/multiply.hpp:238-302):
//   batch  = in C, A and B (C order)      aouter = in C and A only (C order)
//   bouter = in C and B only (C order)    inner  = in A and B only (A order)
//   ared   = in A only (A order)          bred   = in B only (B order)
struct LabelInfo {
  std::vector<int> batch, aouter, bouter, inner, ared, bred;
  std::vector<int> red;  // RHS labels not in C, sorted (execute_bufacc loop, multop.hpp:587-601)
  std::map<int, std::vector<int64_t>> tile_sizes;  // label -> tile sizes of its space
};

LabelInfo analyse_labels(const LOperand& C, const LOperand& A, const LOperand& B);

// Position of label `l` in an operand (-1 if absent).
inline int pos_of(const std::vector<int>& labels, int l) {
  for (size_t i = 0; i < labels.size(); ++i)
    if (labels[i] == l) return (int)i;
  return -1;
}

// One nonzero (C block, A block, B block) triple, in execute_bufacc order.
struct PairRef {
  int64_t a_off, b_off;
  std::vector<int64_t> rt;  // tile index of each reduction label
};
struct CBlockWork {
  int64_t c_off;
  std::vector<int64_t> ct;  // tile index of each C label
  std::vector<PairRef> pairs;
};

// Enumerates output blocks (row-major over C labels, skipping zero blocks) and, for each,
// the reduction blocks whose A and B blocks are both nonzero (multop.hpp:645-891).
std::vector<CBlockWork> enumerate_pairs(const LOperand& C, const LOperand& A, const LOperand& B,
                                        const LabelInfo& li);

// Block dims (tile sizes) of an operand given the tile index of each of its labels.
std::vector<int64_t> block_dims(const LOperand& op, const std::vector<int64_t>& label_tiles);
// Row-major strides for dims.
std::vector<int64_t> strides_of(const std::vector<int64_t>& dims);

// Block-pair (fused FP64) plan.
struct BPPlan {
  DevBuf tasks, pairs, cshapes, kshapes;
  int ntasks = 0;
  int tile = 0;          // BP_TILE_64x64 or BP_TILE_16x256
  bool swapped = false;  // built for C = B*A (the call's B is the kernel's A)
  BPXcd xcd{};  // cost-balanced XCD ranges of `tasks` (C blocks in key order, tiles in a row)
  std::vector<std::pair<int, int>> per_pair;  // task ranges, one launch per pair
  std::vector<char> per_pair_first;           // the launch holds its C block's first pair
  DevBuf pair_tasks;                           // tasks of the per-pair launches
  // C blocks with no block pair (element ranges; zero_prefix holds running lengths), zeroed
  // when the contraction assigns.
  DevBuf zero_off, zero_prefix;
  int nzero = 0;
  int64_t zero_elems = 0;
};

// Builds the fused plan from the enumerated work.  per_pair_launches: also build the one
// launch per block pair variant (TAMM's loop structure).
void build_bp_plan(BPPlan& plan, const LOperand& C, const LOperand& A, const LOperand& B,
                   const LabelInfo& li, const std::vector<CBlockWork>& work,
                   bool per_pair_launches, double* flops, int64_t* npairs);
// Launches tasks [0, ntasks) of `tasks` (the plan's own or its per-pair ranges) with the call's
// operands (swapped plans take them in the order they were built for).
void launch_bp_plan(const BPPlan& bp, const BPTask* tasks, int ntasks, const BPXcd* xr,
                    const double* A, const double* B, double* C, double alpha, int is_assign,
                    hipStream_t s);

// Sector plans (plan_sector.cpp)
struct SectorPlan;
void destroy_sector_plan(SectorPlan* sp);
// Split-K rounds (most parts of any sector), GEMM launches and GEMM tasks of a plan.
// Row-scale reuse of split packing for executions after this call (process-wide).
void set_scale_reuse(bool on);
void sector_schedule_info(const SectorPlan& sp, int* rounds, int* launches, int64_t* tasks);
struct SectorPlanDeleter {
  void operator()(SectorPlan* p) const { destroy_sector_plan(p); }
};
using SectorPlanPtr = std::unique_ptr<SectorPlan, SectorPlanDeleter>;
SectorPlanPtr build_sector_plan(const LOperand& C, const LOperand& A,
                                              const LOperand& B, const LabelInfo& li,
                                              const std::vector<CBlockWork>& work, bool split,
                                              int split_bits, int parallel_split,
                                              double min_split_mn = 0);
// ev: optional 5 events recorded before packing A, before packing B, before the GEMM, before
// the scatter and at the end; gemm_ev: optional events between consecutive GEMM launches (one
// fewer than the plan's launches); pack_ev_a / pack_ev_b: optional two events around the main
// packing pass of the plan's A / B operand alone (not recorded when the operand is constant and
// already packed).
void execute_sector_plan(SectorPlan& sp, const double* A, const double* B, double* C,
                         double alpha, hipStream_t s, hipEvent_t* ev = nullptr,
                         hipEvent_t* gemm_ev = nullptr, hipEvent_t* pack_ev_a = nullptr,
                         hipEvent_t* pack_ev_b = nullptr);
int64_t sector_workspace_bytes(const SectorPlan& sp);
// Scatter stage only: `src`'s product (same product_key: computed by src's last execution)
// into C through sp's scatter tables.
void execute_sector_scatter_from(SectorPlan& sp, const SectorPlan& src, double* C, double alpha,
                                 hipStream_t s);
uint64_t sector_product_key(const SectorPlan& sp);
double sector_mn_intensity(const SectorPlan& sp);
bool sector_is_split(const SectorPlan& sp);
// C = alpha*A*B (assign) instead of C += alpha*A*B for the plan's later executions and scatters.
void set_sector_assign(SectorPlan& sp, bool assign);
// Marks the plan's operands constant (packed once per source address and kept) or not.
void set_sector_constant(SectorPlan& sp, bool a, bool b);
// Drops what the plan remembers of earlier packings (the next execution packs both operands).
void forget_sector_packing(SectorPlan& sp);
// Per-sector tile-shape autotuning on the plan's current packed operands (sector.hpp).
int autotune_sector_plan(SectorPlan& sp, hipStream_t s);
void sector_split_info(const SectorPlan& sp, int* bits_a, int* bits_b, int* nmod);

}  // namespace tmx
