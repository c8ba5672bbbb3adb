// Time integrators (internal): device-resident Krylov expv and Dormand-Prince 5(4).
#pragma once

#include <memory>

#include "nc_sens.h"

struct nc_solver_s {
  nc_matrix_s* A = nullptr;  // borrowed; must outlive the solver (the base generator for SA)
  nc_sensmatrix_s* SA = nullptr;  // sensitivity operator (borrowed), or null
  int blocks = 1;                 // P+1 for the sensitivity system
  int64_t Nb = 0, Npb = 0;        // true and padded rows per block
  nc_solve_opts opts{};
  int method = NC_METHOD_KRYLOV;
  int64_t N = 0;   // true vector length (blocks * Nb)
  int64_t Np = 0;  // padded length (blocks * Npb, Npb a multiple of 1024); padding stays zero
  size_t vs = 8;
  double t = 0.0;
  bool have_state = false;
  hipStream_t s = nullptr;  // the matrix's stream
  nc::DevBuf<double> u;     // FP64 master state [p; sinks]
  nc::DevBuf<double> u0;    // state uploaded by the last set_state (for restart)
  double t0 = 0.0, beta0 = 0.0;
  nc::DevBuf<double> snaps; // n_out x Np (padded layout)
  int32_t n_snaps = 0;
  // reduction scratch
  nc::DevBuf<double> partials;
  nc::DevBuf<double> fold;  // folded totals of a vector pass and the fold's arrival counter
  unsigned red_blocks = 0;
  int64_t dblk = 0;  // Krylov: 512-row projection blocks over the states (0: grid-stride passes)
  // Krylov
  int m = 30;
  nc::DevBuf<unsigned char> V;  // (m+1) x N, VT, unnormalised
  nc::DevBuf<double> q;         // FP64 mat-vec output
  nc::DevBuf<double> kbuf;      // H (m+2)^2 | h (m+2) | sc (m+2) | misc (8) | sink rows
  nc::DevBuf<double> gram;      // (m+1)^2 Gram matrix of the scaled basis
  nc::DevBuf<double> coef;      // combination coefficients (m+2)
  double* hbuf = nullptr;       // pinned host copy of kbuf
  double* hcoef = nullptr;      // pinned host staging of coef
  double* hpart = nullptr;      // pinned host copy of per-block partial sums
  int64_t hpart_cap = 0;        // partial slots of hpart before its sink-value region
  double* hpart_dev = nullptr;  // the same pinned block as a device address: the RK error
                                // pass writes its partials there directly
  double beta = 0.0, t_new = 0.0, anorm = 0.0;
  bool beta_pending = false;  // beta was updated on the device and not yet read back
  bool started = false;
  hipGraph_t arnoldi_graph = nullptr;  // captured Arnoldi process (replayed every step)
  hipGraphExec_t arnoldi_exec = nullptr;
  uint64_t arnoldi_gen = 0;  // the matrix's launch_gen when the graph was captured
  // RK
  nc::DevBuf<unsigned char> rkv;   // VT copy of the stage input
  nc::DevBuf<unsigned char> rkv2;  // the other stage-input buffer (epilogue-formed stages)
  nc::DevBuf<double> k;            // 7 x N stage derivatives (FP64)
  nc::DevBuf<double> ytmp;         // N
  double h = 0.0;
  int ksw = 0;  // 1: RK stages 0 and 6 have traded slots in k
  bool fsal_valid = false;
  nc_solve_stats st{};
  // state list shared by the snapshot vectors handed out (nc_solver_get_fspvec)
  std::shared_ptr<nc::DevBuf<int32_t>> snap_states;
  uint64_t snap_space_uid = 0, snap_version = 0;
  ~nc_solver_s() {
    if (arnoldi_exec) (void)hipGraphExecDestroy(arnoldi_exec);
    if (arnoldi_graph) (void)hipGraphDestroy(arnoldi_graph);
    if (s) (void)hipStreamSynchronize(s);  // copies into the pinned blocks have landed
    if (hbuf) nc::pinned_free(hbuf);
    if (hcoef) nc::pinned_free(hcoef);
    if (hpart) nc::pinned_free(hpart);
  }
};

namespace nc {
// FSP sink-mass budget of the adaptive solve: integration stops where
// g(t) = sum(sinks) - fsptol * t / tend crosses zero from below (fspsolve.jl:141-156).
struct EventSpec {
  double fsptol = 0.0, tend = 1.0;
  bool fired = false;
};
// Advance to t_target (or to the budget crossing when ev is given); with one_step the
// call returns after the first accepted step.
void krylov_advance(nc_solver_s* S, double t_target, EventSpec* ev = nullptr, bool one_step = false);
void rk_advance(nc_solver_s* S, double t_target, EventSpec* ev = nullptr, bool one_step = false);
// Internal entry points shared with the adaptive driver.
nc_solver_s* solver_create(nc_matrix_s* A, const nc_solve_opts* opts, nc_sensmatrix_s* SA = nullptr);
void solver_set_state_device(nc_solver_s* S, double t0, const double* u_dev);
// Step size the integrator would try next (RK: h, Krylov: tau), 0 before its first step.
double solver_step_hint(const nc_solver_s* S);
// Start the next step from h instead of the automatic first step (h <= 0: no change).
void solver_seed_step(nc_solver_s* S, double h);
void solver_reduce_setup(nc_solver_s* S, int nvals);
}  // namespace nc
