// Device context: one HIP stream, a caching device allocator and the last
// error string.  One context per host thread (the C ABI's contract).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace tk {

struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};
struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// A caller-provided output array is too small (TK_ERR_CAPACITY).
struct CapacityError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// DmrgError::Cancelled (tk-dmrg/src/error.rs:53): a cancel flag was raised.
struct CancelledError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define TK_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::tk::HipError(e_, std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)

// Size-class caching allocator: freed blocks are kept for reuse until the
// context is destroyed, so the sweep's per-step buffers cost no hipMalloc.
// A block freed on the host may be handed out again while kernels queued
// before the free still use it: every use goes through the context's one
// stream (launches and staged copies alike), so the new owner's work runs
// after theirs and no synchronisation is needed at a free.
class DevicePool {
 public:
  DevicePool() = default;
  ~DevicePool() { release_all(); }
  void* alloc(size_t bytes) {
    if (bytes == 0) bytes = 256;
    size_t cls = size_class(bytes);
    std::lock_guard<std::mutex> g(mu_);
    auto& fl = free_[cls];
    if (!fl.empty()) {
      void* p = fl.back();
      fl.pop_back();
      return p;
    }
    void* p = nullptr;
    TK_HIP(hipMalloc(&p, cls));
    size_of_[p] = cls;
    held_ += cls;
    return p;
  }
  void free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    auto it = size_of_.find(p);
    if (it == size_of_.end()) return;
    free_[it->second].push_back(p);
  }
  size_t held_bytes() const { return held_; }
  void release_all() {
    std::lock_guard<std::mutex> g(mu_);
    for (auto& [p, s] : size_of_) (void)hipFree(p);
    size_of_.clear();
    free_.clear();
    held_ = 0;
  }

 private:
  static size_t size_class(size_t b) {
    // 256-byte granularity below 1 MiB, then powers of two in quarter steps
    if (b <= (1u << 20)) return (b + 255) & ~size_t(255);
    size_t p = 1;
    while (p < b) p <<= 1;
    size_t q = p >> 2;
    size_t c = p >> 1;
    while (c < b) c += q;
    return c;
  }
  std::mutex mu_;
  std::map<size_t, std::vector<void*>> free_;
  std::map<void*, size_t> size_of_;
  size_t held_ = 0;
};

// Optional HIP-event timing of chosen launch sites on the context's stream
// (benchmark instrumentation; off unless enabled).  Each mark brackets one or
// more back-to-back launches of one kernel and carries their algorithmic
// bytes and flops.
class KernelTimer {
 public:
  struct Total {
    std::string name;
    int64_t launches = 0;
    double ms = 0, bytes = 0, flops = 0;
  };
  bool on() const { return on_; }
  void enable(bool on) {
    on_ = on;
    clear();
  }
  // -1 when timing is off
  int begin(hipStream_t s) {
    if (!on_) return -1;
    Mark m;
    m.a = take();
    m.b = take();
    TK_HIP(hipEventRecord(m.a, s));
    marks_.push_back(m);
    return int(marks_.size()) - 1;
  }
  void end(int tok, hipStream_t s, const char* name, double bytes, double flops, int64_t launches) {
    if (tok < 0) return;
    Mark& m = marks_[size_t(tok)];
    TK_HIP(hipEventRecord(m.b, s));
    m.name = name;
    m.bytes = bytes;
    m.flops = flops;
    m.launches = launches;
  }
  // Waits for the recorded events and returns the totals per kernel name.
  std::vector<Total> read() {
    std::vector<Total> out;
    for (Mark& m : marks_) {
      TK_HIP(hipEventSynchronize(m.b));
      float ms = 0;
      TK_HIP(hipEventElapsedTime(&ms, m.a, m.b));
      size_t i = 0;
      while (i < out.size() && out[i].name != m.name) ++i;
      if (i == out.size()) {
        out.push_back(Total{});
        out.back().name = m.name;
      }
      out[i].launches += m.launches;
      out[i].ms += ms;
      out[i].bytes += m.bytes;
      out[i].flops += m.flops;
    }
    clear();
    return out;
  }
  ~KernelTimer() {
    clear();
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
  }

 private:
  struct Mark {
    hipEvent_t a = nullptr, b = nullptr;
    std::string name;
    double bytes = 0, flops = 0;
    int64_t launches = 0;
  };
  hipEvent_t take() {
    if (!pool_.empty()) {
      hipEvent_t e = pool_.back();
      pool_.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    TK_HIP(hipEventCreate(&e));
    return e;
  }
  void clear() {
    for (Mark& m : marks_) {
      pool_.push_back(m.a);
      pool_.push_back(m.b);
    }
    marks_.clear();
  }
  bool on_ = false;
  std::vector<Mark> marks_;
  std::vector<hipEvent_t> pool_;
};

struct Context {
  int device = 0;
  // contraction tiling: 0 the default tiling, 1 measured per list signature
  // on first use and cached (tune.hpp)
  int tune_mode = 0;
  KernelTimer timer;
  hipStream_t stream = nullptr;
  DevicePool pool;
  std::string last_error;
  int n_cu = 256;
  // Stream-ordered scratch for reduction partials; reused across launches.
  void* scratch_p = nullptr;
  size_t scratch_n = 0;
  void* scratch(size_t bytes) {
    if (bytes > scratch_n) {
      if (scratch_p) pool.free(scratch_p);
      scratch_p = pool.alloc(bytes);
      scratch_n = bytes;
    }
    return scratch_p;
  }
  // Pinned staging for host-to-device uploads (a context is driven by one
  // host thread).  A copy out of pageable memory may hold the host until the
  // stream reaches it, leaving the GPU idle while the host prepares the next
  // launches; staged in pinned memory it is a DMA in stream order.  Slices
  // are recycled once the stream has drained (sync(), or when the arena is
  // full); an upload larger than the arena is copied directly.
  static constexpr size_t kStageBytes = size_t(32) << 20;
  char* stage_p = nullptr;
  size_t stage_off = 0;
  void upload(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    if (stage_p == nullptr || bytes > kStageBytes) {
      TK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
      return;
    }
    if (stage_off + bytes > kStageBytes) {
      TK_HIP(hipStreamSynchronize(stream));
      stage_off = 0;
    }
    std::memcpy(stage_p + stage_off, src, bytes);
    TK_HIP(hipMemcpyAsync(dst, stage_p + stage_off, bytes, hipMemcpyHostToDevice, stream));
    stage_off += (bytes + 255) & ~size_t(255);
  }
  explicit Context(int dev) : device(dev) {
    TK_HIP(hipSetDevice(dev));
    TK_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    TK_HIP(hipGetDeviceProperties(&prop, dev));
    n_cu = prop.multiProcessorCount;
    TK_HIP(hipHostMalloc(reinterpret_cast<void**>(&stage_p), kStageBytes, hipHostMallocDefault));
  }
  ~Context() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
    if (stage_p) (void)hipHostFree(stage_p);
    pool.release_all();
  }
  void sync() {
    TK_HIP(hipStreamSynchronize(stream));
    stage_off = 0;
  }
};

// RAII device buffer drawn from the context pool.
template <typename T>
class DBuf {
 public:
  DBuf() = default;
  DBuf(Context* c, size_t n) { reset(c, n); }
  ~DBuf() { release(); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept { *this = std::move(o); }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      ctx_ = o.ctx_; p_ = o.p_; n_ = o.n_;
      o.ctx_ = nullptr; o.p_ = nullptr; o.n_ = 0;
    }
    return *this;
  }
  void reset(Context* c, size_t n) {
    release();
    ctx_ = c;
    n_ = n;
    p_ = static_cast<T*>(c->pool.alloc(n * sizeof(T)));
  }
  void release() {
    if (p_ && ctx_) ctx_->pool.free(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  void upload(const T* h, size_t n) {
    if (n > n_) throw ArgError("upload larger than buffer");
    ctx_->upload(p_, h, n * sizeof(T));
  }
  void download(T* h, size_t n) const {
    if (n > n_) throw ArgError("download larger than buffer");
    if (n) TK_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, ctx_->stream));
  }

 private:
  Context* ctx_ = nullptr;
  T* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace tk
