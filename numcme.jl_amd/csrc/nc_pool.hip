// Caches of device blocks, pinned host blocks and streams (host code only).
//
// Device blocks: a freed block goes into a per-device size-ordered free list and serves
// any later request of at most its size and at least half of it (best fit). Requests are
// rounded up to 4 KiB, or 2 MiB above 1 MiB, so that the repeated rebuilds of an adaptive
// solve, whose sizes creep upwards, find their blocks again. The cache holds at most
// kCacheCap bytes; a block freed beyond that goes back to the driver. When hipMalloc
// fails, every cached block is returned and the allocation is retried once.
//
// The caches live until the process ends and are never torn down (the HIP runtime may
// already be gone when static destructors run).
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "nc_internal.h"

namespace nc {
namespace {

constexpr size_t kSmallGrain = 4096;
constexpr size_t kLargeGrain = size_t(2) << 20;
constexpr size_t kCacheCap = size_t(16) << 30;
// Streams kept for reuse per device. Every object holds a stream while it lives, so a user
// holding many objects at once (a batch built from hundreds of generators) makes as many;
// keeping them all afterwards would slow every later device-wide synchronisation, which
// visits each stream.
constexpr size_t kStreamCacheCap = 32;

size_t round_bytes(size_t b) {
  const size_t g = b > (size_t(1) << 20) ? kLargeGrain : kSmallGrain;
  return (b + g - 1) / g * g;
}

struct Cache {
  std::mutex mu;
  std::map<int, std::multimap<size_t, void*>> free_dev;  // device -> size -> block
  std::unordered_map<void*, std::pair<int, size_t>> live_dev;
  size_t cached_dev = 0;
  std::multimap<size_t, void*> free_host;
  std::unordered_map<void*, size_t> live_host;
  std::map<int, std::vector<hipStream_t>> streams;
  std::unordered_map<hipStream_t, int> stream_dev;
};

Cache& cache() {
  static Cache* c = new Cache();  // intentionally never destroyed
  return *c;
}

int current_device() {
  int d = 0;
  (void)hipGetDevice(&d);
  return d;
}

void trim_locked(Cache& c) {
  for (auto& [dev, lst] : c.free_dev) {
    if (lst.empty()) continue;
    int cur = current_device();
    (void)hipSetDevice(dev);
    for (auto& [sz, p] : lst) (void)hipFree(p);
    lst.clear();
    (void)hipSetDevice(cur);
  }
  c.cached_dev = 0;
}

}  // namespace

void* pool_alloc(size_t bytes) {
  Cache& c = cache();
  const size_t want = round_bytes(bytes);
  const int dev = current_device();
  std::lock_guard<std::mutex> lk(c.mu);
  auto& lst = c.free_dev[dev];
  auto it = lst.lower_bound(want);
  if (it != lst.end() && it->first <= 2 * want) {
    void* p = it->second;
    const size_t sz = it->first;
    lst.erase(it);
    c.cached_dev -= sz;
    c.live_dev[p] = {dev, sz};
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim_locked(c);
    e = hipMalloc(&p, want);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(NC_ERR_ALLOC, "hipMalloc of " + std::to_string(want) + " bytes failed: " + hipGetErrorString(e));
  }
  c.live_dev[p] = {dev, want};
  return p;
}

void pool_free(void* p) {
  if (!p) return;
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.live_dev.find(p);
  if (it == c.live_dev.end()) return;  // not ours
  const auto [dev, sz] = it->second;
  c.live_dev.erase(it);
  if (c.cached_dev + sz > kCacheCap) {
    (void)hipFree(p);
    return;
  }
  c.free_dev[dev].emplace(sz, p);
  c.cached_dev += sz;
}

void pool_trim() {
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  trim_locked(c);
}

void* pinned_alloc(size_t bytes) {
  Cache& c = cache();
  const size_t want = round_bytes(bytes);
  {
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.free_host.lower_bound(want);
    if (it != c.free_host.end() && it->first <= 2 * want) {
      void* p = it->second;
      c.live_host[p] = it->first;
      c.free_host.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  NC_CHECK_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
  std::lock_guard<std::mutex> lk(c.mu);
  c.live_host[p] = want;
  return p;
}

void pinned_free(void* p) {
  if (!p) return;
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.live_host.find(p);
  if (it == c.live_host.end()) return;
  c.free_host.emplace(it->second, p);
  c.live_host.erase(it);
}

// A thread's staging block goes back to the pinned cache when the thread ends (a worker of
// a batch of solves: its queued copies completed when its objects were freed).
struct StageBlock {
  unsigned char* p = nullptr;
  ~StageBlock() {
    if (p) pinned_free(p);
  }
};

void* host_stage(int slot) {
  thread_local StageBlock blk;
  if (!blk.p) blk.p = static_cast<unsigned char*>(pinned_alloc(16 * kHostStageSlot));
  if (slot < 0 || slot >= 16) throw Error(NC_ERR_INTERNAL, "bad host staging slot");
  return blk.p + (size_t)slot * kHostStageSlot;
}

// The streams this host thread holds (taken here and not yet released), and whether its
// buffer releases wait for these only (inside a PrivateStreams scope).
thread_local std::vector<hipStream_t> tl_streams;
thread_local int tl_private = 0;

hipStream_t stream_acquire() {
  Cache& c = cache();
  const int dev = current_device();
  hipStream_t s = nullptr;
  {
    std::lock_guard<std::mutex> lk(c.mu);
    auto& v = c.streams[dev];
    if (!v.empty()) {
      s = v.back();
      v.pop_back();
    }
  }
  if (!s) {
    NC_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::lock_guard<std::mutex> lk(c.mu);
    c.stream_dev[s] = dev;
  }
  if (std::find(tl_streams.begin(), tl_streams.end(), s) == tl_streams.end()) tl_streams.push_back(s);
  return s;
}

void stream_release(hipStream_t s) {
  if (!s) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    // a capture that failed part-way and was not ended: the stream cannot serve again, so
    // it is ended and destroyed instead of going back to the cache
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    (void)hipGetLastError();
    for (size_t k = 0; k < tl_streams.size(); ++k)
      if (tl_streams[k] == s) {
        tl_streams[k] = tl_streams.back();
        tl_streams.pop_back();
        break;
      }
    Cache& c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    c.stream_dev.erase(s);
    return;
  }
  (void)hipStreamSynchronize(s);
  for (size_t k = 0; k < tl_streams.size(); ++k)
    if (tl_streams[k] == s) {
      tl_streams[k] = tl_streams.back();
      tl_streams.pop_back();
      break;
    }
  Cache& c = cache();
  {
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.stream_dev.find(s);
    if (it == c.stream_dev.end()) return;
    auto& v = c.streams[it->second];
    if (v.size() < kStreamCacheCap) {
      v.push_back(s);
      return;
    }
    c.stream_dev.erase(it);
  }
  (void)hipStreamDestroy(s);  // beyond the cache's cap (its work is complete)
}

void release_sync() {
  if (!tl_private) {
    (void)hipDeviceSynchronize();
    return;
  }
  for (hipStream_t s : tl_streams) (void)hipStreamSynchronize(s);
}

PrivateStreams::PrivateStreams() { ++tl_private; }
PrivateStreams::~PrivateStreams() { --tl_private; }

void streams_settled() { tl_streams.clear(); }

// Device-to-host copy into pageable memory through two pinned chunks: the DMA of chunk
// c + 1 runs while the host copies chunk c out of its pinned buffer (a direct copy to
// pageable memory measured 6-9 GB/s here for snapshots of 8 MB).
namespace {

// Host copies of the read-back chunks split over the calling thread and kCopyHelpers
// helper threads. The destination of a download is usually fresh memory (a new numpy
// array), whose pages the kernel zeroes on first touch: one thread reaches about 10 GB/s
// that way, below what the DMA engine delivers. The helpers start with the first large
// download and live until the process ends; one download uses them at a time, and a
// download that finds them busy copies on its own thread.
class CopyHelpers {
 public:
  static CopyHelpers& get() {
    static CopyHelpers* p = new CopyHelpers();  // never destroyed: the helpers may outlive statics
    return *p;
  }
  // memcpy(dst, src, len) by the pool; false (nothing copied) when another download holds it
  bool run(unsigned char* dst, const unsigned char* src, size_t len) {
    std::unique_lock<std::mutex> owner(run_mu_, std::try_to_lock);
    if (!owner.owns_lock()) return false;
    {
      std::lock_guard<std::mutex> lk(mu_);
      dst_ = dst;
      src_ = src;
      len_ = len;
      pending_ = kCopyHelpers;
      ++gen_;
    }
    cv_.notify_all();
    slice(0, dst, src, len);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return pending_ == 0; });
    return true;
  }

 private:
  static constexpr int kCopyHelpers = 3;
  CopyHelpers() {
    for (int k = 1; k <= kCopyHelpers; ++k) std::thread([this, k] { loop(k); }).detach();
  }
  // part k of kCopyHelpers + 1 near-equal parts, cut at 64-byte boundaries
  static void slice(int k, unsigned char* dst, const unsigned char* src, size_t len) {
    const size_t per = (len / (kCopyHelpers + 1) + 63) / 64 * 64, off = (size_t)k * per;
    if (off < len) std::memcpy(dst + off, src + off, std::min(per, len - off));
  }
  void loop(int k) {
    uint64_t seen = 0;
    for (;;) {
      unsigned char* d;
      const unsigned char* s;
      size_t n;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        d = dst_;
        s = src_;
        n = len_;
      }
      slice(k, d, s, n);
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::mutex run_mu_, mu_;
  std::condition_variable cv_, done_;
  unsigned char* dst_ = nullptr;
  const unsigned char* src_ = nullptr;
  size_t len_ = 0;
  int pending_ = 0;
  uint64_t gen_ = 0;
};

constexpr size_t kParallelCopyMin = size_t(1) << 20;  // smaller chunks: one thread

void host_copy(unsigned char* dst, const unsigned char* src, size_t len) {
  if (len < kParallelCopyMin || !CopyHelpers::get().run(dst, src, len)) std::memcpy(dst, src, len);
}

}  // namespace

void copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
  constexpr size_t kChunk = size_t(4) << 20;
  if (bytes < (size_t(1) << 20)) {
    NC_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    return;
  }
  struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
  };
  thread_local Events ev;  // one pair per host thread, kept until the process ends
  for (auto& e : ev.e)
    if (!e) NC_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  unsigned char* buf = static_cast<unsigned char*>(pinned_alloc(2 * kChunk));
  try {
    const size_t n = (bytes + kChunk - 1) / kChunk;
    const auto* sp = static_cast<const unsigned char*>(src);
    auto* dp = static_cast<unsigned char*>(dst);
    auto issue = [&](size_t c) {
      const size_t off = c * kChunk, len = std::min(kChunk, bytes - off);
      NC_CHECK_HIP(hipMemcpyAsync(buf + (c & 1) * kChunk, sp + off, len, hipMemcpyDeviceToHost, s));
      NC_CHECK_HIP(hipEventRecord(ev.e[c & 1], s));
    };
    issue(0);
    for (size_t c = 0; c < n; ++c) {
      // buffer (c + 1) & 1 was emptied by the host copy of chunk c - 1 just before
      if (c + 1 < n) issue(c + 1);
      NC_CHECK_HIP(hipEventSynchronize(ev.e[c & 1]));
      const size_t off = c * kChunk, len = std::min(kChunk, bytes - off);
      host_copy(dp + off, buf + (c & 1) * kChunk, len);
    }
  } catch (...) {
    (void)hipStreamSynchronize(s);
    pinned_free(buf);
    throw;
  }
  pinned_free(buf);
}

}  // namespace nc
