// Host-side helpers shared by the search, Clifford and T-count translation units: the thread's last error
// message, HIP error checking, grow-only device arrays, HIP-event timing of kernel families, launch sizing.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace chz {

inline std::string &err_slot() { static thread_local std::string s; return s; }
inline void set_err(const std::string &s) { err_slot() = s; }
inline const char *get_err() { return err_slot().c_str(); }

struct DeviceError {};

#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      ::chz::set_err(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x);  \
      throw ::chz::DeviceError();                                                      \
    }                                                                                  \
  } while (0)

// Process-wide lock between stream capture and the HIP calls that act on the whole device or allocate.  A thread
// capturing a level step holds it from hipStreamBeginCapture to the graph's instantiation; hipMalloc, hipFree,
// hipHostMalloc/hipHostFree, device-wide synchronisations and hipMemGetInfo take it too.  A device-wide
// synchronisation includes the stream under capture, and hipFree synchronises the device: issued by another
// thread (another engine of the request server) while a stream is being captured, they must wait for the capture
// to end.
inline std::recursive_mutex &capture_mu() { static std::recursive_mutex *m = new std::recursive_mutex(); return *m; }

// The stream the calling thread's grow-only device buffers belong to, set by each context's entry points for the
// duration of the call: a buffer's growth then waits for that stream only (its old block can be in use only by
// kernels queued there), not for other contexts' kernels on the same device.
inline hipStream_t &owner_stream() { static thread_local hipStream_t s = nullptr; return s; }

// The calling thread's stream-capture interaction mode, relaxed for the duration of a library call.  In HIP's
// default (global) mode a thread's synchronous or allocating call while any stream of the process is being
// captured invalidates that capture, even one in thread-local mode on another thread; the library's own calls never
// touch a stream under capture (device-wide synchronisations and allocations also wait on capture_mu), so its
// threads use the relaxed mode.
struct RelaxedCapture {
  hipStreamCaptureMode prev = hipStreamCaptureModeRelaxed;
  RelaxedCapture() { (void)hipThreadExchangeStreamCaptureMode(&prev); }
  ~RelaxedCapture() { (void)hipThreadExchangeStreamCaptureMode(&prev); }
};

struct StreamScope {
  hipStream_t prev;
  RelaxedCapture relaxed;
  explicit StreamScope(hipStream_t s) : prev(owner_stream()) { owner_stream() = s; }
  ~StreamScope() { owner_stream() = prev; }
};
// a synchronous copy on the context's own stream (a copy on the legacy default stream would depend on every
// blocking stream, which is not allowed while another thread captures one)
inline void copy_sync(void *dst, const void *src, size_t n, hipMemcpyKind kind, hipStream_t s) {
  HIPCHK(hipMemcpyAsync(dst, src, n, kind, s));
  HIPCHK(hipStreamSynchronize(s));
}

// no kernel of the calling context still uses its blocks once this returns
inline hipError_t sync_owner() {
  if (owner_stream()) return hipStreamSynchronize(owner_stream());
  std::lock_guard<std::recursive_mutex> g(capture_mu());
  return hipDeviceSynchronize();
}

// Process-wide cache of device blocks by (device, size class): contexts opened and closed one after another
// (one search per context, as the command line does) reuse memory instead of paying hipMalloc/hipFree each time.
// Size classes are 2^k rounded up to a quarter step (at most 25 % slack).  A block enters the cache only when no
// kernel can still use it: after a device synchronisation on growth, and from closed contexts whose stream has
// finished.  At most kMaxCached bytes stay cached; the rest is freed.
struct DevPool {
  static constexpr size_t kMaxCached = 16ull << 30;
  std::mutex mu;
  std::multimap<std::pair<int, size_t>, void *> free_blocks;
  size_t cached = 0;
  static size_t size_class(size_t b) {
    if (b <= 256) return 256;
    int lg = 63 - __builtin_clzll((unsigned long long)b);
    const size_t step = (size_t)1 << (lg - 2);
    return (b + step - 1) / step * step;
  }
  void *get(size_t bytes, size_t &got) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    got = size_class(bytes);
    {
      std::lock_guard<std::mutex> g(mu);
      auto it = free_blocks.find({dev, got});
      if (it != free_blocks.end()) {
        void *p = it->second;
        free_blocks.erase(it);
        cached -= got;
        return p;
      }
    }
    void *p = nullptr;
    std::lock_guard<std::recursive_mutex> cg(capture_mu());
    if (hipMalloc(&p, got) != hipSuccess) {
      (void)hipGetLastError();
      trim();                       // give the cache back to the device and try once more
      HIPCHK(hipMalloc(&p, got));
    }
    return p;
  }
  void put(void *p, size_t got) {
    if (!p) return;
    int dev = 0;
    std::lock_guard<std::recursive_mutex> cg(capture_mu());
    if (hipGetDevice(&dev) != hipSuccess) { (void)hipFree(p); return; }
    std::lock_guard<std::mutex> g(mu);
    if (cached + got > kMaxCached) { (void)hipFree(p); return; }
    free_blocks.insert({{dev, got}, p});
    cached += got;
  }
  void trim() {
    std::lock_guard<std::recursive_mutex> cg(capture_mu());
    std::lock_guard<std::mutex> g(mu);
    for (auto &kv : free_blocks) (void)hipFree(kv.second);
    free_blocks.clear();
    cached = 0;
  }
};
inline DevPool &dev_pool() { static DevPool *p = new DevPool(); return *p; }   // never destroyed: no frees at exit

// untyped blocks from the cache for code that frees by pointer; dev_free requires that no queued kernel still
// uses the block (synchronise first)
struct LiveBlocks {
  std::mutex mu;
  std::map<void *, size_t> sizes;
};
inline LiveBlocks &live_blocks() { static LiveBlocks *l = new LiveBlocks(); return *l; }
inline void *dev_alloc(size_t bytes) {
  size_t got = 0;
  void *p = dev_pool().get(std::max<size_t>(bytes, 1), got);
  std::lock_guard<std::mutex> g(live_blocks().mu);
  live_blocks().sizes[p] = got;
  return p;
}
inline void dev_free(void *p) {
  if (!p) return;
  size_t got = 0;
  {
    std::lock_guard<std::mutex> g(live_blocks().mu);
    auto it = live_blocks().sizes.find(p);
    if (it == live_blocks().sizes.end()) {
      std::lock_guard<std::recursive_mutex> cg(capture_mu());
      (void)hipFree(p);
      return;
    }
    got = it->second;
    live_blocks().sizes.erase(it);
  }
  dev_pool().put(p, got);
}

template <typename T> struct DBuf {     // grow-only device array on the block cache
  T *p = nullptr;
  size_t cap = 0, bytes = 0;
  void need(size_t n) {
    if (n <= cap) return;
    const size_t c = std::max<size_t>(std::max<size_t>(n, cap + cap / 2), 1);
    size_t got = 0;
    T *np = (T *)dev_pool().get(c * sizeof(T), got);
    if (p) {
      // kernels already queued may still read the old block
      HIPCHK(sync_owner());
      dev_pool().put(p, bytes);
    }
    p = np;
    bytes = got;
    cap = got / sizeof(T);
  }
  // the caller guarantees that no queued kernel still uses the block (its stream has synchronised)
  void release() { if (p) dev_pool().put(p, bytes); p = nullptr; cap = 0; bytes = 0; }
};

// Pinned host staging for a context's small per-level transfers (budgets, counters, flags): copies from and to
// pinned memory are asynchronous DMA, where pageable ones stall the host on a staging copy each.  The caller
// uses it only between stream synchronisations (no copy through it is in flight when it grows or is rewritten).
struct Pinned {
  uint8_t *p = nullptr;
  size_t cap = 0;
  uint8_t *need(size_t n) {
    if (n <= cap) return p;
    std::lock_guard<std::recursive_mutex> cg(capture_mu());
    if (p) HIPCHK(hipHostFree(p));
    p = nullptr;
    const size_t c = std::max<size_t>(n, std::max<size_t>(4096, cap * 2));
    HIPCHK(hipHostMalloc((void **)&p, c, hipHostMallocDefault));
    cap = c;
    return p;
  }
  void release() {
    std::lock_guard<std::recursive_mutex> cg(capture_mu());
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

inline double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// HIP-event timing of kernel families on one stream; pairs are read after the stream synchronised.  Only the
// families in `mask` record events (two records and an elapsed-time query per launch cost host time on every
// level); the others count launches only.
struct EventLog {
  unsigned mask = ~0u;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> open;
  hipEvent_t take() {
    if (pool.empty()) { hipEvent_t e; HIPCHK(hipEventCreate(&e)); return e; }
    hipEvent_t e = pool.back(); pool.pop_back(); return e;
  }
  void begin(int fam, hipStream_t s) {
    if (!((mask >> fam) & 1u)) { open.push_back({fam, {nullptr, nullptr}}); return; }
    hipEvent_t a = take(), b = take();
    HIPCHK(hipEventRecord(a, s));
    open.push_back({fam, {a, b}});
  }
  void end(hipStream_t s) {
    if (hipEvent_t b = open.back().second.second) HIPCHK(hipEventRecord(b, s));
  }
  void flush(double *ms, long *launches) {
    for (auto &o : open) {
      launches[o.first]++;
      if (!o.second.first) continue;
      float x = 0;
      HIPCHK(hipEventElapsedTime(&x, o.second.first, o.second.second));
      ms[o.first] += x;
      pool.push_back(o.second.first);
      pool.push_back(o.second.second);
    }
    open.clear();
  }
  void destroy() {
    for (auto &o : open)
      if (o.second.first) { pool.push_back(o.second.first); pool.push_back(o.second.second); }
    open.clear();
    for (auto e : pool) (void)hipEventDestroy(e);
    pool.clear();
  }
};

// blocks of 256 threads for a number of waves, capped
inline int grid_for(long long waves_needed, int max_blocks = 8192) {
  long long blocks = (waves_needed + 3) / 4;
  if (blocks < 1) blocks = 1;
  return (int)std::min<long long>(blocks, max_blocks);
}

}  // namespace chz
