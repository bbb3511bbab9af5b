// sector.hpp -- sector (dense class-matrix) plan structures and kernel launchers.
#pragma once

#include "plan.hpp"

namespace tmx {

constexpr int kMaxModuli = 16;

// Garner mixed-radix constants: inv[j][i] = m_j^{-1} mod m_i (j < i).
struct GarnerConsts {
  int n;
  int m[kMaxModuli];
  int inv[kMaxModuli][kMaxModuli];
};

struct PackSide {
  DevBuf jobs;
  int njobs = 0;
  int64_t nitems = 0;  // (row, 16-column group) work items
  int64_t total = 0;   // elements (FP64) or bytes per residue plane (int8)
  int64_t nrows = 0;   // rows over all class matrices (row-scale arrays)
  DevBuf units;        // 64x64 tiles of the jobs
  int64_t nunits = 0;
  DevBuf buf;          // packed FP64 matrices, or nmod int8 residue planes
  DevBuf rowmax, invscale;
};

struct ShapeTasks {
  int tm = 0, tn = 0;
  DevBuf tasks;
  int ntasks = 0;
};

struct SectorPlan {
  bool split = false;
  int bits_a = 0, bits_b = 0, nmod = 0;
  GarnerConsts garner;
  PackSide A, B;
  std::vector<ShapeTasks> shapes;
  DevBuf ksegs;
  DevBuf cdense;
  int64_t c_total = 0;  // dense result elements (FP64) or bytes per plane (int8)
  int64_t kmax = 0;
  DevBuf sjobs, sunits;
  int nsjobs = 0;
  int64_t nsunits = 0;
  int64_t ws = 0;
  double exec_flops = 0;  // MFMA work including padding (per modulus for split)
};

// Tile shape (rows of A x rows of B per workgroup) for a sector GEMM.
void choose_tile(bool split, int64_t mrows, int64_t nrows, int* tm, int* tn);
// Groups tasks by tile shape (SecTask::pad holds tm*1000+tn) and uploads them.
void split_tasks_by_shape(SectorPlan& sp, const std::vector<SecTask>& tasks);


}  // namespace tmx
