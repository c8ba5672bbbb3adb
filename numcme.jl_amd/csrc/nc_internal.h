// Internal helpers shared by the numcme-hip library sources: error state, HIP
// checks, device buffers and small launch utilities. Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "numcme_hip.h"

namespace nc {

// ---- error handling --------------------------------------------------------------------
// Every C entry point catches nc::Error (and std::exception) and turns it into a
// status code plus a thread-local message retrievable with nc_last_error().
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define NC_CHECK_HIP(expr)                                                                 \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      throw ::nc::Error(NC_ERR_HIP, std::string(#expr " failed: ") + hipGetErrorString(_e) + \
                                        " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define NC_REQUIRE(cond, code, msg)              \
  do {                                           \
    if (!(cond)) throw ::nc::Error((code), (msg)); \
  } while (0)

template <class F>
int guard(F&& f) {
  try {
    f();
    return NC_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return NC_ERR_ALLOC;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return NC_ERR_INTERNAL;
  }
}

// ---- memory and stream caches ---------------------------------------------------------
// Adaptive solves rebuild their generator and integrator at every space adaptation, so
// device blocks, pinned host blocks and streams are cached for reuse instead of going back
// to the driver (hipMalloc / hipHostMalloc / hipStreamCreate cost tens to hundreds of
// microseconds each). A freed device block is reused for a request of at most its size and
// at least half of it; a failed allocation empties the cache and retries. Callers release
// a block only when no queued work uses it.
void* pool_alloc(size_t bytes);
void pool_free(void* p);
void pool_trim();  // return every cached device block to the driver
void* pinned_alloc(size_t bytes);
void pinned_free(void* p);
// Pinned staging for small device-to-host reads, per host thread: slot k (0..15) is its own
// 4 KiB region, so reads whose copy and use interleave take different slots. A copy into
// pageable memory blocks the host until it completes; into these it is asynchronous, and
// several reads share one synchronisation.
constexpr size_t kHostStageSlot = 4096;
void* host_stage(int slot);
hipStream_t stream_acquire();        // non-blocking stream, default priority
void stream_release(hipStream_t s);  // synchronises it, then keeps it for reuse
// What must be complete before a device block goes back to the cache: by default every
// stream's work (the whole device). Inside a PrivateStreams scope, only the work the calling
// thread has queued on the streams it holds.
void release_sync();
// For worker threads that create, use and free every object of their own (the threads of a
// batch of independent solves): their buffer releases then wait for their own streams only,
// so that the threads do not wait for each other whenever one of them frees a buffer.
struct PrivateStreams {
  PrivateStreams();
  ~PrivateStreams();
  PrivateStreams(const PrivateStreams&) = delete;
  PrivateStreams& operator=(const PrivateStreams&) = delete;
};
// For such a thread whose finished objects outlive their work (a builder that hands its
// objects on): once their streams have been synchronised, the thread stops tracking them,
// so that its later releases wait only for the streams of the work still in progress.
void streams_settled();
// Copy bytes from device memory to pageable host memory on stream s and wait for it
// (large copies go through pinned staging chunks with the transfers overlapped).
void copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t s);

// ---- device buffers -------------------------------------------------------------------
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    p = static_cast<T*>(pool_alloc(count * sizeof(T)));
    n = count;
  }
  // grow to at least `count` elements; contents are NOT preserved
  void reserve(size_t count) { if (count > n) alloc(count); }
  // grow to at least `count` elements, preserving the first `keep` elements
  void grow_keep(size_t count, size_t keep, hipStream_t s) {
    if (count <= n) return;
    T* q = static_cast<T*>(pool_alloc(count * sizeof(T)));
    if (keep && p) NC_CHECK_HIP(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    release();
    p = q;
    n = count;
  }
  void release() {
    if (p) {
      // queued work may still read the buffer
      release_sync();
      pool_free(p);
    }
    p = nullptr;
    n = 0;
  }
  T* get() const { return p; }
};

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return static_cast<unsigned>(g < 1 ? 1 : g);
}

// Exclusive prefix sum of `n` uint32 values (rocPRIM); the total is written to *total_host
// when non-null (synchronises `s`). Temporary storage lives in `tmp`.
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, DevBuf<unsigned char>& tmp,
                        hipStream_t s, uint32_t* total_host);

// Dense n x n row-major matrix exponential on the host (Pade 13 scaling and squaring),
// and its first column alone (exp(A) e1, what a Krylov step uses).
void expm_dense(int n, const double* a, double* out);
void expm_first_column(int n, const double* a, double* out);

// Stream with library-owned lifetime; default priority only.
struct Stream {
  hipStream_t s = nullptr;
  Stream() { s = stream_acquire(); }
  ~Stream() { if (s) stream_release(s); }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  // Give the stream back before the owner goes (its work is complete, and the owner will
  // only be read by work on other streams from here on); s becomes the null stream.
  void release_now() {
    if (s) stream_release(s);
    s = nullptr;
  }
};

}  // namespace nc
