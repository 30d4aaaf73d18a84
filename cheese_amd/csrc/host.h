// Host-side helpers shared by the search, Clifford and T-count translation units: the thread's last error
// message, HIP error checking, grow-only device arrays, HIP-event timing of kernel families, launch sizing.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
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

template <typename T> struct DBuf {     // grow-only device array
  T *p = nullptr;
  size_t cap = 0;
  void need(size_t n) {
    if (n <= cap) return;
    if (p) HIPCHK(hipFree(p));
    p = nullptr;
    const size_t c = std::max<size_t>(n, cap + cap / 2);
    HIPCHK(hipMalloc((void **)&p, std::max<size_t>(c, 1) * sizeof(T)));
    cap = c;
  }
  void release() { if (p) (void)hipFree(p); p = nullptr; cap = 0; }
};

inline double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// HIP-event timing of kernel families on one stream; pairs are read after the stream synchronised
struct EventLog {
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> open;
  hipEvent_t take() {
    if (pool.empty()) { hipEvent_t e; HIPCHK(hipEventCreate(&e)); return e; }
    hipEvent_t e = pool.back(); pool.pop_back(); return e;
  }
  void begin(int fam, hipStream_t s) {
    hipEvent_t a = take(), b = take();
    HIPCHK(hipEventRecord(a, s));
    open.push_back({fam, {a, b}});
  }
  void end(hipStream_t s) { HIPCHK(hipEventRecord(open.back().second.second, s)); }
  void flush(double *ms, long *launches) {
    for (auto &o : open) {
      float x = 0;
      HIPCHK(hipEventElapsedTime(&x, o.second.first, o.second.second));
      ms[o.first] += x;
      launches[o.first]++;
      pool.push_back(o.second.first);
      pool.push_back(o.second.second);
    }
    open.clear();
  }
  void destroy() {
    for (auto &o : open) { pool.push_back(o.second.first); pool.push_back(o.second.second); }
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
