// Common infrastructure of the engine: error propagation (no exceptions
// cross the C ABI), device memory pool, strided tensor views.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sisi_engine.h"

namespace sisi {

// Internal failures are C++ exceptions carrying a SISI_ERR_* code; the C ABI
// layer converts them into a status and the thread-local message.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define SISI_HIP(call)                                                      \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      throw ::sisi::Error(SISI_ERR_HIP, std::string(#call) + ": " +         \
                                            hipGetErrorString(e_) + " (" + \
                                            __FILE__ + ":" +                \
                                            std::to_string(__LINE__) + ")"); \
  } while (0)

#define SISI_REQUIRE(cond, code, msg)                                       \
  do {                                                                      \
    if (!(cond)) throw ::sisi::Error((code), (msg));                        \
  } while (0)

// Device memory pool: buffers are recycled by exact byte size, so the
// allocation pattern of one CCSD iteration is served without hipMalloc after
// the first iteration.  Not thread-safe (one engine, one host thread).
// Device-wide operations (hipDeviceSynchronize, hipFree, stream creation and
// destruction) are not allowed while another thread captures its stream into
// a hipGraph: they take this process-wide lock, and so does a capture.
inline std::recursive_mutex &device_wide_mutex() {
  static std::recursive_mutex m;
  return m;
}

class Pool {
 public:
  ~Pool() { release_all(); }
  void *get(size_t bytes) {
    if (bytes == 0) bytes = 8;
    bytes = (bytes + 255) & ~size_t(255);
    auto it = free_.find(bytes);
    if (it != free_.end() && !it->second.empty()) {
      void *p = it->second.back();
      it->second.pop_back();
      live_.insert(p);
      if (capturing_) captured_.push_back(p);
      return p;
    }
    void *p = nullptr;
    hipError_t err = hipMalloc(&p, bytes);
    if (err != hipSuccess) {
      (void)hipGetLastError();
      // Drop cached buffers and retry once.
      trim();
      err = hipMalloc(&p, bytes);
      if (err != hipSuccess) {
        (void)hipGetLastError();
        throw Error(SISI_ERR_ALLOC, "device allocation of " +
                                        std::to_string(bytes) + " bytes failed");
      }
    }
    size_[p] = bytes;
    live_.insert(p);
    if (capturing_) captured_.push_back(p);
    return p;
  }
  void put(void *p) {
    if (!p) return;
    auto it = size_.find(p);
    if (it == size_.end()) return;
    live_.erase(p);
    if (reserved_.count(p)) return;  // held by a captured graph
    free_[it->second].push_back(p);
  }
  // Graph capture: every buffer handed out between begin_capture and
  // end_capture is referenced by the captured graph, so end_capture takes
  // them out of circulation until release() (when the graph is destroyed);
  // a buffer that still has an owner at release() returns to the free list
  // when that owner gives it back.
  void begin_capture() {
    capturing_ = true;
    captured_.clear();
  }
  std::vector<void *> end_capture() {
    capturing_ = false;
    std::vector<void *> held;
    for (void *p : captured_) {
      if (reserved_.count(p)) continue;
      reserved_.insert(p);
      held.push_back(p);
      auto &fl = free_[size_[p]];
      for (size_t k = 0; k < fl.size(); ++k)
        if (fl[k] == p) {
          fl.erase(fl.begin() + k);
          break;
        }
    }
    captured_.clear();
    return held;
  }
  void release(const std::vector<void *> &held) {
    for (void *p : held) {
      reserved_.erase(p);
      auto it = size_.find(p);
      if (it != size_.end() && !live_.count(p)) free_[it->second].push_back(p);
    }
  }
  // Bytes held in the free lists (device memory the pool can hand out
  // again, or give back with trim()).
  size_t cached_bytes() const {
    size_t b = 0;
    for (const auto &kv : free_) b += kv.first * kv.second.size();
    return b;
  }
  void trim() {
    std::lock_guard<std::recursive_mutex> lock(device_wide_mutex());
    (void)hipDeviceSynchronize();
    for (auto &kv : free_)
      for (void *p : kv.second) {
        (void)hipFree(p);
        size_.erase(p);
      }
    free_.clear();
  }
  void release_all() {
    std::lock_guard<std::recursive_mutex> lock(device_wide_mutex());
    (void)hipDeviceSynchronize();
    for (auto &kv : size_) (void)hipFree(kv.first);
    size_.clear();
    free_.clear();
    reserved_.clear();
    live_.clear();
  }

 private:
  std::map<size_t, std::vector<void *>> free_;
  std::map<void *, size_t> size_;
  std::set<void *> reserved_;
  std::set<void *> live_;  // handed out and not yet given back
  std::vector<void *> captured_;
  bool capturing_ = false;
};

// Owning device buffer of doubles, returned to the pool on destruction.
class DBuf {
 public:
  DBuf() = default;
  DBuf(Pool *pool, size_t n) : pool_(pool), n_(n) {
    p_ = static_cast<double *>(pool->get(n * sizeof(double)));
  }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept { swap(o); }
  DBuf &operator=(DBuf &&o) noexcept {
    reset();
    swap(o);
    return *this;
  }
  ~DBuf() { reset(); }
  void reset() {
    if (pool_ && p_) pool_->put(p_);
    p_ = nullptr;
    n_ = 0;
  }
  double *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void swap(DBuf &o) {
    std::swap(pool_, o.pool_);
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
  }
  Pool *pool_ = nullptr;
  double *p_ = nullptr;
  size_t n_ = 0;
};

constexpr int kMaxOrder = 6;

// Strided view of a real tensor: element (x0..x_{order-1}) lives at
// p[sum_k x_k * stride[k]].  Dense tensors have stride[0] = 1 and
// stride[k] = stride[k-1] * len[k-1] (first index fastest).
struct View {
  double *p = nullptr;
  int order = 0;
  int64_t len[kMaxOrder] = {0};
  int64_t stride[kMaxOrder] = {0};

  int64_t size() const {
    int64_t n = 1;
    for (int k = 0; k < order; ++k) n *= len[k];
    return n;
  }
};

inline View dense(double *p, std::initializer_list<int64_t> lens) {
  View v;
  v.p = p;
  v.order = 0;
  int64_t s = 1;
  for (int64_t l : lens) {
    v.len[v.order] = l;
    v.stride[v.order] = s;
    s *= l;
    v.order++;
  }
  return v;
}

// View with explicit strides (element x at p[sum_k x_k * strides_k]).
inline View dense_strided(double *p, std::initializer_list<int64_t> lens,
                          std::initializer_list<int64_t> strides) {
  View v;
  v.p = p;
  v.order = 0;
  auto s = strides.begin();
  for (int64_t l : lens) {
    v.len[v.order] = l;
    v.stride[v.order] = *s++;
    v.order++;
  }
  return v;
}

// Sub-block [lo_k, hi_k) of a view (same order).
inline View subview(const View &v, std::initializer_list<int64_t> lo,
                    std::initializer_list<int64_t> hi) {
  View r = v;
  int k = 0;
  auto l = lo.begin();
  auto h = hi.begin();
  for (; k < v.order; ++k, ++l, ++h) {
    r.p += (*l) * v.stride[k];
    r.len[k] = (*h) - (*l);
  }
  return r;
}

}  // namespace sisi
