// Device-resident sparse multi-indexed vectors (internal): the FSP snapshot container.
//
// Layout in HBM for n entries over ns species:
//   states  int32  [n][ns]   row-major (shared between snapshots of one state space)
//   values  double [n]
#pragma once

#include <memory>

#include "nc_internal.h"

// A vector owns no stream: a solve may hand out thousands of snapshots, and each operation
// takes a stream from the library's cache for its own duration.
struct nc_fspvec_s {
  int ns = 0;
  int64_t n = 0;
  std::shared_ptr<nc::DevBuf<int32_t>> states;  // may be shared with other vectors
  nc::DevBuf<double> values;
};

namespace nc {
// A vector over `states` (shared, n x ns on the device) with values copied from the
// device array `values` (n doubles), ordered after the work queued on `src`.
nc_fspvec_s* fspvec_from_device(int ns, int64_t n, std::shared_ptr<DevBuf<int32_t>> states, const double* values,
                                hipStream_t src);
// Copy of n x ns device states into a new shared buffer (on stream s).
std::shared_ptr<DevBuf<int32_t>> share_states(const int32_t* d_states, int64_t n, int ns, hipStream_t s);
}  // namespace nc
