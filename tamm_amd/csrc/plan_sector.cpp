// plan_sector.cpp -- sector (dense block-class) GEMM plans: FP64 and int8-residue split.
#include "plan.hpp"

namespace tmx {

struct SectorPlan {
  int64_t ws = 0;
};

SectorPlanPtr build_sector_plan(const LOperand&, const LOperand&, const LOperand&, const LabelInfo&,
                                const std::vector<CBlockWork>&, bool, int) {
  throw Error(TMX_ERR_UNSUPPORTED, "sector execution is not available in this build");
}
void execute_sector_plan(SectorPlan&, const double*, const double*, double*, double, hipStream_t) {}
int64_t sector_workspace_bytes(const SectorPlan& sp) { return sp.ws; }
void sector_split_info(const SectorPlan&, int* a, int* b, int* n) {
  if (a) *a = 0;
  if (b) *b = 0;
  if (n) *n = 0;
}
void destroy_sector_plan(SectorPlan* sp) { delete sp; }

}  // namespace tmx
