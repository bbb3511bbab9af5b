// tmx_internal.hpp -- shared host/device definitions of the MI355X contraction engine.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "tamm_mi355x.h"

namespace tmx {

// ---------------------------------------------------------------------------
// Errors: exceptions inside the library, codes at the C boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define TMX_REQUIRE(cond, msg)                                                        \
  do {                                                                                \
    if (!(cond)) throw ::tmx::Error(TMX_ERR_INVALID, std::string("invalid: ") + (msg)); \
  } while (0)

#define TMX_HIP(call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::tmx::Error(TMX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Runs `f`, converting exceptions to return codes.
template <class F>
int guarded(F&& f) {
  try {
    f();
    return TMX_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return TMX_ERR_OOM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return TMX_ERR_INVALID;
  }
}

// ---------------------------------------------------------------------------
// Runtime
struct Runtime {
  bool initialised = false;
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  int num_xcds = 8;
};
Runtime& runtime();
void require_init();
hipStream_t pick_stream(void* s);
void* pool_alloc(size_t bytes);
void pool_free(void* p);
// True while `s` is being captured into a graph.
bool stream_capturing(hipStream_t s);
// Graph lifetime.  While a stream is captured through tmx_graph_begin / tmx_graph_end, every
// plan object executed into it is recorded with that capture (capture_hold, a no-op when `s`
// is not captured that way), and the graph made from the capture keeps a reference to each
// until tmx_graph_destroy.  A plan whose handle is destroyed while a graph can still replay it
// therefore keeps its device memory until the graph goes.
void capture_hold(hipStream_t s, std::shared_ptr<void> obj);

// RAII device buffer from the pool.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) : p(b ? pool_alloc(b) : nullptr), bytes(b) {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { reset(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p) pool_free(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// A flag that kernels raise in pinned host memory (1: the event happened) and the host reads
// without blocking.  poll() returns 1 and clears the flag when it was raised, else 0; the
// atomic exchange loses no raise that lands while it runs.  The memory is freed only when no
// queued kernel can still write it (plans are destroyed after the device is idle).
struct HostFlag {
  int* host = nullptr;
  int* dev = nullptr;  // the same memory as the device addresses it
  HostFlag() = default;
  HostFlag(const HostFlag&) = delete;
  HostFlag& operator=(const HostFlag&) = delete;
  ~HostFlag() {
    if (host) (void)hipHostFree(host);
  }
  void alloc();
  int poll();
};

// Copies a host vector to a fresh device buffer (plan time only).  Ordered on the engine
// stream and completed before returning, so it never overtakes engine work still using
// memory the pool has handed out again.
void upload_bytes(void* dst, const void* src, size_t bytes);
template <class T>
DevBuf upload(const std::vector<T>& v) {
  DevBuf b(v.size() * sizeof(T));
  if (!v.empty()) upload_bytes(b.p, v.data(), v.size() * sizeof(T));
  return b;
}
// Waits for all engine work before plan memory returns to the pool.
void quiesce();

// ---------------------------------------------------------------------------
// Index spaces and tensors (host structure)
struct Space {
  int64_t n = 0;
  std::vector<int64_t> toff;  // ntiles + 1
  std::vector<int> tspin;     // per tile (0 = none)
  std::vector<int> tirrep;    // per tile (-1 = none)
  int64_t ntiles() const { return (int64_t)toff.size() - 1; }
  int64_t tsize(int64_t t) const { return toff[t + 1] - toff[t]; }
};

struct Tensor {
  std::vector<std::shared_ptr<const Space>> modes;
  bool has_spin = false;
  std::vector<int> mask;  // 0 upper, 1 lower, 2 ignore
  bool irrep_rule = false;
  std::vector<int64_t> key_off;
  std::vector<int64_t> keys, offs, sizes;
  int64_t total = 0;
  uint64_t uid = 0;

  int nmodes() const { return (int)modes.size(); }
  bool is_non_zero(const int64_t* bid) const;
  int64_t key(const int64_t* bid) const {
    int64_t k = 0;
    for (int i = 0; i < nmodes(); ++i) k += bid[i] * key_off[i];
    return k;
  }
  // Offset of a stored block or -1.
  int64_t find(const int64_t* bid) const {
    int64_t k = key(bid);
    auto it = std::lower_bound(keys.begin(), keys.end(), k);
    if (it == keys.end() || *it != k) return -1;
    return offs[it - keys.begin()];
  }
  void key_to_bid(int64_t key, int64_t* bid) const {
    for (int i = 0; i < nmodes(); ++i) { bid[i] = key / key_off[i]; key %= key_off[i]; }
  }
};

std::shared_ptr<Space> make_space(int64_t n, int use_vector, const int64_t* tiles, int64_t ntiles,
                                  int nspin, const int64_t* slo, const int64_t* shi, const int* sval,
                                  int nirrep, const int64_t* ilo, const int64_t* ihi, const int* ival,
                                  int nnamed, const int64_t* nlo, const int64_t* nhi);
std::shared_ptr<Tensor> make_tensor(const std::vector<std::shared_ptr<const Space>>& modes,
                                    int nupper, int nlower, int irrep_rule);

// Labeled operand after translation (labels + tile offsets + label tile counts).
struct LOperand {
  std::shared_ptr<Tensor> t;
  std::vector<int> labels;
  std::vector<int64_t> toff, ntl;
};
LOperand make_loperand(const tmx_operand* op);

// ---------------------------------------------------------------------------
// Device-side descriptors (plain data, shared by planner and kernels)
constexpr int MAXR = 8;  // max modes per GEMM role (TAMM tensors have at most 8 modes)

// Mixed-radix index map: idx in [0, prod(ext)) -> sum_d digit_d * stride_d (row-major digits).
struct Radix {
  int n;
  int ext[MAXR];
  int stride[MAXR];
};

// Fused block-pair GEMM (grouped FP64): one task = one C tile.
struct BPTask {
  int64_t c_off;      // element offset of the C block
  int32_t m0, n0;     // tile origin inside the block's GEMM view
  int32_t cshape;     // index into BPCShape table
  int32_t pair_begin, pair_end;
  int32_t pad;
};
struct BPCShape {
  int32_t M, N;
  Radix cm, cn;  // C strides of the M (aouter) and N (bouter) modes
};
struct BPPair {
  int64_t a_off, b_off;
  int32_t kshape;
  int32_t pad;
};
// Task ranges per XCD: workgroups are dispatched round-robin over the 8 XCDs, and workgroup b
// works through tasks [lo[b % 8], hi[b % 8]) so that neighbouring tiles, which read the same
// operand blocks, share that XCD's L2.
struct BPXcd {
  int32_t lo[8], hi[8];
};
struct BPKShape {
  int32_t K;
  int32_t a_kcontig;  // 1 when the A block is contiguous along K (loads coalesce along k)
  int32_t b_kcontig;  // 1 when the B block is contiguous along K
  int32_t pad;
  Radix am;  // A strides of the M modes
  Radix ak;  // A strides of the K modes (inner + A-reduce; 0 stride for B-reduce)
  Radix bk;  // B strides of the K modes (inner + B-reduce; 0 stride for A-reduce)
  Radix bn;  // B strides of the N modes
};

// Sector packing job: gathers one source block into a packed operand matrix.  A packed
// matrix holds the rows (M or N elements) of one row class against the padded columns
// (K elements) of the K blocks compatible with that class, in MFMA fragment-major order.
struct PackJob {
  int64_t src_off;   // element offset of the source block in the tensor
  int64_t dst_base;  // offset of the class matrix in the packed buffer (elements / bytes)
  int32_t row0;      // first row of this block inside the class matrix
  int32_t col0;      // first (padded) column of this block inside the class matrix
  int32_t rows, cols;  // block extent in rows / columns (columns before padding)
  int32_t kpad;      // padded column count of the class matrix
  int32_t growbase;  // global row index of the class matrix's row 0 (row scales)
  int32_t kcontig;   // 1 when the source block is contiguous along columns (K)
  int32_t kfull;     // 1 when a row's columns are one contiguous run (K labels trail in order)
  Radix r;           // row index -> source offset
  Radix c;           // column index -> source offset
};

// 64x64 tile of a pack or scatter job (one workgroup iteration).
struct TileUnit {
  int32_t job, r0, c0;
  int32_t lt;  // pack tiles (k_pack, k_rowmax_t): 2^lt rows x 4096 / 2^lt columns; else unused
};

// Dense sector GEMM tile: C_dense[m0.., n0..] = sum over K segments of A rows x B rows.
struct SecTask {
  int64_t a_base, b_base;  // offsets of the packed A / B class matrices
  int64_t c_base;          // offset of the dense sector result
  int32_t m0, n0;          // tile origin (rows of A / rows of B)
  int32_t kseg_begin, kseg_end;
  int32_t a_kpad, b_kpad;  // padded column counts of the A / B class matrices
  int32_t ldc;             // leading dimension of the dense result
  int32_t mrows, nrows;    // valid rows of the A / B class matrices (0: empty task)
  int32_t plane;           // residue plane (split path)
};
struct KSeg {
  int32_t a_col, b_col, len, pad;
};

// Scatter of a dense sector result into the C tensor's block layout.
struct ScatterJob {
  int64_t c_off;     // element offset of the destination C block
  int64_t src_base;  // offset of the dense sector result (its first split-K part)
  int64_t pstride;   // distance between the dense results of consecutive parts
  int32_t nparts;    // split-K parts summed into C
  int32_t pad;
  int32_t row0, col0;  // first dense row / col of this block
  int32_t rows, cols;
  int32_t ldc;         // leading dimension of the dense result
  int32_t arow0;       // global A-row index of dense row 0 (split: row scales)
  int32_t bcol0;       // global B-row index of dense col 0 (split: column scales)
  int32_t ccontig;     // 1 when the C block is contiguous along dense columns (N)
  Radix r, c;          // dense row / col -> C block offset
};

}  // namespace tmx
