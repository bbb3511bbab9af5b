// runtime.cpp -- device selection, engine stream, device memory pool, events, graphs.
//
// Pool policy follows RMMMemoryManager::initialize (src/tamm/rmm_memory_pool.hpp:230-237):
// one up-front hipMalloc of pool_fraction x free HBM (TAMM_GPU_POOL percent, default 80),
// sub-allocated best-fit from a coalescing free list (src/tamm/mr/coalescing_free_list.hpp),
// never grown.  Unlike TAMM's pool (not thread-safe, mr/stream_ordered_memory_resource.hpp:
// 91-118) this one is guarded by a mutex, and blocks are 256-byte aligned.
// On MI355X (288 GB HBM3E) the default pool is ~230 GB.
#include <cstdlib>
#include <map>
#include <mutex>
#include <set>

#include "tmx_kernels.hpp"

namespace tmx {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

static Runtime g_rt;
Runtime& runtime() { return g_rt; }

void require_init() {
  if (!g_rt.initialised) throw Error(TMX_ERR_STATE, "tmx_init has not been called");
}

hipStream_t pick_stream(void* s) {
  require_init();
  return s ? static_cast<hipStream_t>(s) : g_rt.stream;
}

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  TMX_HIP(hipStreamIsCapturing(s, &st));
  return st != hipStreamCaptureStatusNone;
}

namespace {
// Objects held by the captures in progress (tmx_graph_begin .. tmx_graph_end), per stream.
// Heap singletons, never destroyed (see the pool below).
std::mutex& capture_mu() {
  static std::mutex* m = new std::mutex();
  return *m;
}
std::map<hipStream_t, std::vector<std::shared_ptr<void>>>& captures() {
  static auto* c = new std::map<hipStream_t, std::vector<std::shared_ptr<void>>>();
  return *c;
}
// A graph handle: the executable graph and what its kernels point into.
struct GraphObj {
  hipGraphExec_t exec = nullptr;
  std::vector<std::shared_ptr<void>> keep;
};
}  // namespace

void capture_hold(hipStream_t s, std::shared_ptr<void> obj) {
  if (!obj || !stream_capturing(s)) return;
  std::lock_guard<std::mutex> g(capture_mu());
  auto it = captures().find(s);
  if (it == captures().end()) return;
  for (auto& o : it->second)
    if (o.get() == obj.get()) return;
  it->second.push_back(std::move(obj));
}

namespace {
constexpr size_t kAlign = 256;

struct Pool {
  std::mutex mu;
  char* base = nullptr;
  size_t capacity = 0, in_use = 0, peak = 0;
  std::map<size_t, size_t> free_by_off;             // offset -> size
  std::set<std::pair<size_t, size_t>> free_by_size;  // (size, offset)
  std::map<size_t, size_t> used;                     // offset -> size

  void add_free(size_t off, size_t sz) {
    // coalesce with neighbours
    auto nxt = free_by_off.lower_bound(off);
    if (nxt != free_by_off.end() && off + sz == nxt->first) {
      free_by_size.erase({nxt->second, nxt->first});
      sz += nxt->second;
      nxt = free_by_off.erase(nxt);
    }
    if (nxt != free_by_off.begin()) {
      auto prv = std::prev(nxt);
      if (prv->first + prv->second == off) {
        free_by_size.erase({prv->second, prv->first});
        off = prv->first;
        sz += prv->second;
        free_by_off.erase(prv);
      }
    }
    free_by_off[off] = sz;
    free_by_size.insert({sz, off});
  }
  void* alloc(size_t bytes) {
    std::lock_guard<std::mutex> g(mu);
    if (!base) throw Error(TMX_ERR_STATE, "device pool not initialised");
    size_t sz = (std::max<size_t>(bytes, 1) + kAlign - 1) / kAlign * kAlign;
    auto it = free_by_size.lower_bound({sz, 0});  // best fit
    if (it == free_by_size.end())
      throw Error(TMX_ERR_OOM, "device pool exhausted: requested " + std::to_string(bytes) +
                                   " bytes, in use " + std::to_string(in_use) + " of " +
                                   std::to_string(capacity));
    size_t off = it->second, fsz = it->first;
    free_by_size.erase(it);
    free_by_off.erase(off);
    if (fsz > sz) {
      free_by_off[off + sz] = fsz - sz;
      free_by_size.insert({fsz - sz, off + sz});
    }
    used[off] = sz;
    in_use += sz;
    peak = std::max(peak, in_use);
    return base + off;
  }
  void free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu);
    size_t off = static_cast<char*>(p) - base;
    auto it = used.find(off);
    if (it == used.end()) throw Error(TMX_ERR_INVALID, "tmx_free: pointer not from the pool");
    size_t sz = it->second;
    used.erase(it);
    in_use -= sz;
    add_free(off, sz);
  }
};
// Never destroyed: cached plans elsewhere free into the pool from static destructors, whose
// order across translation units is unspecified.
Pool& pool_singleton() {
  static Pool* p = new Pool();
  return *p;
}
#define g_pool (pool_singleton())
}  // namespace

void upload_bytes(void* dst, const void* src, size_t bytes) {
  require_init();
  TMX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_rt.stream));
  TMX_HIP(hipStreamSynchronize(g_rt.stream));
}

void HostFlag::alloc() {
  if (host) return;
  // coherent (fine-grained): a kernel's store reaches the host without a later release
  TMX_HIP(hipHostMalloc(reinterpret_cast<void**>(&host), sizeof(int),
                        hipHostMallocMapped | hipHostMallocCoherent));
  *host = 0;
  TMX_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&dev), host, 0));
}

int HostFlag::poll() {
  if (!host) return 0;
  return __atomic_exchange_n(host, 0, __ATOMIC_ACQ_REL) != 0 ? 1 : 0;
}

void quiesce() {
  if (g_rt.initialised) (void)hipDeviceSynchronize();
}

void* pool_alloc(size_t bytes) {
  require_init();
  return g_pool.alloc(bytes);
}
void pool_free(void* p) {
  // Called from destructors: never throws (a pointer the pool does not own is ignored).
  if (!g_rt.initialised) return;
  try {
    g_pool.free(p);
  } catch (...) {
  }
}

}  // namespace tmx

using namespace tmx;

extern "C" {

const char* tmx_last_error(void) { return g_last_error.c_str(); }
const char* tmx_version(void) { return "tamm-mi355x 0.1 (gfx950)"; }

int tmx_init(int device, double pool_fraction) {
  return guarded([&] {
    if (g_rt.initialised) throw Error(TMX_ERR_STATE, "tmx_init called twice");
    int ndev = 0;
    TMX_HIP(hipGetDeviceCount(&ndev));
    TMX_REQUIRE(device >= 0 && device < ndev, "device index out of range");
    TMX_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    TMX_HIP(hipGetDeviceProperties(&prop, device));
    std::string arch = prop.gcnArchName;
    if (arch.find("gfx950") == std::string::npos)
      throw Error(TMX_ERR_UNSUPPORTED, "this engine is built for gfx950 (MI355X); found " + arch);
    g_rt.num_cus = prop.multiProcessorCount;
    if (pool_fraction <= 0) {
      const char* env = std::getenv("TAMM_GPU_POOL");
      pool_fraction = env ? std::atof(env) / 100.0 : 0.8;
    }
    TMX_REQUIRE(pool_fraction > 0 && pool_fraction <= 0.95, "pool fraction must be in (0, 0.95]");
    size_t freeb = 0, totalb = 0;
    TMX_HIP(hipMemGetInfo(&freeb, &totalb));
    size_t cap = static_cast<size_t>(pool_fraction * static_cast<double>(freeb)) / kAlign * kAlign;
    void* base = nullptr;
    TMX_HIP(hipMalloc(&base, cap));
    TMX_HIP(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking));
    {
      std::lock_guard<std::mutex> g(g_pool.mu);
      g_pool.base = static_cast<char*>(base);
      g_pool.capacity = cap;
      g_pool.in_use = g_pool.peak = 0;
      g_pool.free_by_off.clear();
      g_pool.free_by_size.clear();
      g_pool.used.clear();
      g_pool.add_free(0, cap);
    }
    noalloc_setup(device);  // device-side scratch of tmx_norm_noalloc, with `device` current
    g_rt.device = device;
    g_rt.initialised = true;
  });
}

int tmx_finalize(void) {
  return guarded([&] {
    if (!g_rt.initialised) return;
    TMX_HIP(hipStreamSynchronize(g_rt.stream));
    TMX_HIP(hipStreamDestroy(g_rt.stream));
    {
      std::lock_guard<std::mutex> g(g_pool.mu);
      TMX_HIP(hipFree(g_pool.base));
      g_pool.base = nullptr;
      g_pool.capacity = 0;
      g_pool.free_by_off.clear();
      g_pool.free_by_size.clear();
      g_pool.used.clear();
    }
    g_rt.initialised = false;
    g_rt.stream = nullptr;
  });
}

int tmx_get_stream(void** stream) {
  return guarded([&] {
    require_init();
    TMX_REQUIRE(stream, "null output");
    *stream = g_rt.stream;
  });
}

int tmx_synchronize(void) {
  return guarded([&] {
    require_init();
    TMX_HIP(hipStreamSynchronize(g_rt.stream));
    TMX_HIP(hipDeviceSynchronize());
  });
}

int tmx_malloc(size_t bytes, void** ptr) {
  return guarded([&] {
    TMX_REQUIRE(ptr, "null output");
    *ptr = pool_alloc(bytes);
  });
}

int tmx_free(void* ptr) {
  return guarded([&] {
    require_init();
    // Memory may be handed out again immediately: let queued work that uses it finish.
    TMX_HIP(hipDeviceSynchronize());
    g_pool.free(ptr);
  });
}

int tmx_pool_info(size_t* capacity, size_t* in_use, size_t* peak) {
  return guarded([&] {
    require_init();
    std::lock_guard<std::mutex> g(g_pool.mu);
    if (capacity) *capacity = g_pool.capacity;
    if (in_use) *in_use = g_pool.in_use;
    if (peak) *peak = g_pool.peak;
  });
}

int tmx_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream) {
  return guarded([&] {
    hipStream_t s = pick_stream(stream);
    hipMemcpyKind k = kind == TMX_H2D ? hipMemcpyHostToDevice
                    : kind == TMX_D2H ? hipMemcpyDeviceToHost
                                      : hipMemcpyDeviceToDevice;
    TMX_REQUIRE(kind >= 0 && kind <= 2, "bad copy kind");
    if (bytes == 0) return;
    TMX_HIP(hipMemcpyAsync(dst, src, bytes, k, s));
    if (kind != TMX_D2D) TMX_HIP(hipStreamSynchronize(s));
  });
}

int tmx_event_create(void** ev) {
  return guarded([&] {
    require_init();
    hipEvent_t e;
    TMX_HIP(hipEventCreate(&e));
    *ev = e;
  });
}
int tmx_event_record(void* ev, void* stream) {
  return guarded([&] { TMX_HIP(hipEventRecord(static_cast<hipEvent_t>(ev), pick_stream(stream))); });
}
int tmx_event_elapsed_ms(void* start, void* stop, float* ms) {
  return guarded([&] {
    TMX_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    TMX_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
  });
}
int tmx_event_destroy(void* ev) {
  return guarded([&] { TMX_HIP(hipEventDestroy(static_cast<hipEvent_t>(ev))); });
}
int tmx_stream_create(void** stream) {
  return guarded([&] {
    require_init();
    hipStream_t s;
    TMX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
  });
}
int tmx_stream_destroy(void* stream) {
  return guarded([&] { TMX_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream))); });
}
int tmx_stream_synchronize(void* stream) {
  return guarded([&] { TMX_HIP(hipStreamSynchronize(pick_stream(stream))); });
}
int tmx_graph_begin(void* stream) {
  return guarded([&] {
    hipStream_t s = pick_stream(stream);
    TMX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    std::lock_guard<std::mutex> g(capture_mu());
    captures()[s].clear();
  });
}
int tmx_graph_end(void* stream, void** graph_exec) {
  return guarded([&] {
    TMX_REQUIRE(graph_exec, "null output");
    *graph_exec = nullptr;
    hipStream_t s = pick_stream(stream);
    auto go = std::make_unique<GraphObj>();
    {
      std::lock_guard<std::mutex> g(capture_mu());
      auto it = captures().find(s);
      if (it != captures().end()) {
        go->keep = std::move(it->second);
        captures().erase(it);
      }
    }
    hipGraph_t g;
    TMX_HIP(hipStreamEndCapture(s, &g));
    hipGraphExec_t ex;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    TMX_HIP(e);
    go->exec = ex;
    *graph_exec = go.release();
  });
}
int tmx_graph_launch(void* graph_exec, void* stream) {
  return guarded([&] {
    TMX_REQUIRE(graph_exec, "null graph");
    TMX_HIP(hipGraphLaunch(static_cast<GraphObj*>(graph_exec)->exec, pick_stream(stream)));
  });
}
int tmx_graph_destroy(void* graph_exec) {
  return guarded([&] {
    if (!graph_exec) return;
    std::unique_ptr<GraphObj> go(static_cast<GraphObj*>(graph_exec));
    quiesce();  // replays still running may use the held plans' memory
    TMX_HIP(hipGraphExecDestroy(go->exec));
  });
}

}  // extern "C"
