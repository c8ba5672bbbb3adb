// Learned space adapter (internal): boundary-state features, network selection and the
// training records. No reference counterpart; the reactions to expand and the boundary
// states follow SelectiveRStepAdapter (rstepadapters.jl:86-110) and expand!
// (sparsestatespace.jl:150-193).
#pragma once

#include "nc_internal.h"
#include "nc_mlp.h"
#include "nc_model.h"
#include "nc_space.h"

struct nc_records_s {
  int64_t capacity = 0, per_adaptation = 0;
  double label_mass = 0.0;
  int64_t n_labelled = 0;         // rows [0, n_labelled) hold labels
  int64_t n_pending = 0;          // rows [n_labelled, n_labelled + n_pending) wait for theirs
  int ns = 0;
  nc::DevBuf<float> features;     // capacity x 8
  nc::DevBuf<float> labels;       // capacity
  nc::DevBuf<int32_t> pend_states;  // n_pending x ns
  nc::DevBuf<uint32_t> pend_mask;   // n_pending: sink reactions to expand at record time
  nc::DevBuf<int32_t> pend_stoich;  // ns x nr of the space that recorded them
  int pend_nr = 0;
};

namespace nc {

struct LearnWork {
  DevBuf<uint32_t> flags, ranks, select, counter;
  DevBuf<int32_t> E, xmax;
  DevBuf<float> feat, logits;
  DevBuf<unsigned char> scan_tmp;
};

// Boundary states of the space (ascending ids into w.E: states with a sink on a reaction
// of rmask) and their feature rows into w.feat, for probabilities p (device, sp->n) at
// time t. Returns the number of boundary states.
int64_t learned_features(nc_space_s* sp, const nc_model_s* m, const double* p, const double* dsinks, uint32_t rmask,
                         double t, double tend, LearnWork& w, hipStream_t s);

// Network selection over the boundary states of the last learned_features call: sets
// w.select (one flag per state) and returns how many boundary states are kept.
int64_t learned_select(nc_space_s* sp, nc_mlp_s* M, int64_t nE, double threshold, LearnWork& w, hipStream_t s);

// Training records: label the pending rows against the current space and probabilities p,
// then (records_add) queue up to per_adaptation of the last learned_features rows.
void records_label(nc_records_s* R, nc_space_s* sp, const double* p, hipStream_t s);
void records_add(nc_records_s* R, nc_space_s* sp, int64_t nE, uint32_t rmask, LearnWork& w, hipStream_t s);

}  // namespace nc
