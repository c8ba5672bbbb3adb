// Model object: stoichiometry, parameters and propensity programs (CmeModel,
// src/cmemodel/cmemodel.jl:49-53; propensity kinds src/cmemodel/propensity.jl:50-153;
// gradients src/cmemodel/propensitygrad.jl:16-67).
#include <algorithm>

#include "nc_model.h"
#include "nc_space.h"

namespace nc {

void prog_validate(const nc_instr* code, int64_t len, int ns, int np, bool allow_x, bool allow_t, bool allow_empty,
                   const char* what) {
  const std::string w(what);
  if (len == 0) {
    NC_REQUIRE(allow_empty, NC_ERR_ARG, w + ": empty program");
    return;
  }
  int depth = 0;
  for (int64_t k = 0; k < len; ++k) {
    const nc_instr& in = code[k];
    int pop = 0, push = 1;
    switch (in.op) {
      case NC_OP_CONST: break;
      case NC_OP_PARAM:
        NC_REQUIRE(in.arg >= 0 && in.arg < np, NC_ERR_ARG, w + ": parameter index out of range");
        break;
      case NC_OP_SPECIES:
        NC_REQUIRE(allow_x, NC_ERR_ARG, w + ": program may not read the state");
        NC_REQUIRE(in.arg >= 0 && in.arg < ns, NC_ERR_ARG, w + ": species index out of range");
        break;
      case NC_OP_TIME:
        NC_REQUIRE(allow_t, NC_ERR_ARG, w + ": program may not read time");
        break;
      case NC_OP_NEG: case NC_OP_EXP: case NC_OP_LOG: case NC_OP_SIN: case NC_OP_COS: case NC_OP_SQRT: case NC_OP_ABS:
        pop = 1;
        break;
      case NC_OP_ADD: case NC_OP_SUB: case NC_OP_MUL: case NC_OP_DIV: case NC_OP_POW: case NC_OP_MAX: case NC_OP_MIN:
      case NC_OP_LT: case NC_OP_LE: case NC_OP_GT: case NC_OP_GE:
        pop = 2;
        break;
      case NC_OP_SELECT: pop = 3; break;
      default: throw Error(NC_ERR_ARG, w + ": unknown opcode " + std::to_string(in.op));
    }
    NC_REQUIRE(depth >= pop, NC_ERR_ARG, w + ": stack underflow");
    depth += push - pop;
    NC_REQUIRE(depth <= NC_PROG_STACK, NC_ERR_ARG, w + ": stack deeper than NC_PROG_STACK");
  }
  NC_REQUIRE(depth == 1, NC_ERR_ARG, w + ": program must leave exactly one value");
}

int prog_max_depth(const nc_instr* code, int64_t len) {
  int depth = 0, peak = 0;
  for (int64_t k = 0; k < len; ++k) {
    switch (code[k].op) {
      case NC_OP_CONST: case NC_OP_PARAM: case NC_OP_SPECIES: case NC_OP_TIME: ++depth; break;
      case NC_OP_NEG: case NC_OP_EXP: case NC_OP_LOG: case NC_OP_SIN: case NC_OP_COS: case NC_OP_SQRT: case NC_OP_ABS:
        break;
      case NC_OP_SELECT: depth -= 2; break;
      default: --depth; break;  // binary operations
    }
    peak = std::max(peak, depth);
  }
  return peak;
}

namespace {

__global__ void k_eval(const nc_instr* code, int64_t len, double t, const int32_t* states, int ns, int64_t n,
                       const double* theta, double* out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t x[kMaxSpecies];
  for (int s = 0; s < ns; ++s) x[s] = states[i * ns + s];
  out[i] = prog_eval(code, len, t, x, theta);
}

}  // namespace

}  // namespace nc

using namespace nc;

extern "C" {

int nc_model_create(int32_t ns, int32_t nr, const int32_t* stoich, int32_t np, const double* params,
                    const int32_t* kinds, const int64_t* offs, const nc_instr* code, nc_model* out) {
  return guard([&] {
    NC_REQUIRE(out && stoich && kinds && offs, NC_ERR_ARG, "null argument");
    NC_REQUIRE(ns >= 1 && ns <= kMaxSpecies, NC_ERR_UNSUPPORTED, "species count must be in 1..16");
    NC_REQUIRE(nr >= 1 && nr <= kMaxReactions, NC_ERR_UNSUPPORTED, "reaction count must be in 1..32");
    NC_REQUIRE(np >= 0 && (np == 0 || params), NC_ERR_ARG, "bad parameter vector");
    NC_REQUIRE(offs[0] == 0, NC_ERR_ARG, "program offsets must start at 0");
    for (int k = 0; k < 2 * nr; ++k) NC_REQUIRE(offs[k + 1] >= offs[k], NC_ERR_ARG, "program offsets must not decrease");
    const int64_t total = offs[2 * nr];
    NC_REQUIRE(total == 0 || code, NC_ERR_ARG, "null program code");
    for (int r = 0; r < nr; ++r) {
      const int kind = kinds[r];
      NC_REQUIRE(kind == NC_PROP_TIME_INVARIANT || kind == NC_PROP_SEPARABLE || kind == NC_PROP_JOINT, NC_ERR_ARG,
                 "unknown propensity kind");
      const std::string tag = "reaction " + std::to_string(r);
      prog_validate(code + offs[2 * r], offs[2 * r + 1] - offs[2 * r], ns, np, true, kind == NC_PROP_JOINT, false,
                    (tag + " propensity").c_str());
      prog_validate(code + offs[2 * r + 1], offs[2 * r + 2] - offs[2 * r + 1], ns, np, false, true,
                    kind != NC_PROP_SEPARABLE, (tag + " time factor").c_str());
      if (kind != NC_PROP_SEPARABLE)
        NC_REQUIRE(offs[2 * r + 2] == offs[2 * r + 1], NC_ERR_ARG, tag + ": only separable propensities have a time factor");
    }
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
      throw Error(NC_ERR_NO_DEVICE, "no HIP device available");
    auto m = new nc_model_s();
    try {
      m->ns = ns;
      m->nr = nr;
      m->np = np;
      m->stoich.assign(stoich, stoich + (size_t)ns * nr);
      m->params.assign(params, params + np);
      m->kinds.assign(kinds, kinds + nr);
      m->offs.assign(offs, offs + 2 * nr + 1);
      m->code.assign(code, code + total);
      for (int k = 0; k < 2 * nr; ++k)
        m->max_depth = std::max(m->max_depth, prog_max_depth(code + offs[k], offs[k + 1] - offs[k]));
      m->d_stoich.alloc((size_t)ns * nr);
      NC_CHECK_HIP(hipMemcpy(m->d_stoich.p, stoich, sizeof(int32_t) * ns * nr, hipMemcpyHostToDevice));
      m->d_code.alloc(std::max<int64_t>(total, 1));
      if (total) NC_CHECK_HIP(hipMemcpy(m->d_code.p, code, sizeof(nc_instr) * total, hipMemcpyHostToDevice));
      m->d_offs.alloc(2 * nr + 1);
      NC_CHECK_HIP(hipMemcpy(m->d_offs.p, offs, sizeof(int64_t) * (2 * nr + 1), hipMemcpyHostToDevice));
      m->d_params.alloc(std::max(np, 1));
      if (np) NC_CHECK_HIP(hipMemcpy(m->d_params.p, params, sizeof(double) * np, hipMemcpyHostToDevice));
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int nc_model_destroy(nc_model m) {
  return guard([&] { delete m; });
}

int nc_model_set_gradients(nc_model m, const uint8_t* pattern, const int64_t* doffs, const nc_instr* dcode) {
  return guard([&] {
    NC_REQUIRE(m && pattern && doffs, NC_ERR_ARG, "null argument");
    const int nr = m->nr, np = m->np;
    const int64_t npairs = (int64_t)nr * np;
    NC_REQUIRE(doffs[0] == 0, NC_ERR_ARG, "gradient offsets must start at 0");
    for (int64_t k = 0; k < 2 * npairs; ++k)
      NC_REQUIRE(doffs[k + 1] >= doffs[k], NC_ERR_ARG, "gradient offsets must not decrease");
    const int64_t total = doffs[2 * npairs];
    NC_REQUIRE(total == 0 || dcode, NC_ERR_ARG, "null gradient code");
    for (int r = 0; r < nr; ++r)
      for (int p = 0; p < np; ++p) {
        const int64_t k = (int64_t)r * np + p;
        const int kind = m->kinds[r];
        const std::string tag = "d a_" + std::to_string(r) + "/d theta_" + std::to_string(p);
        const bool dep = pattern[k] != 0;
        prog_validate(dcode + doffs[2 * k], doffs[2 * k + 1] - doffs[2 * k], m->ns, np, true, kind == NC_PROP_JOINT,
                      !dep || kind == NC_PROP_SEPARABLE, tag.c_str());
        prog_validate(dcode + doffs[2 * k + 1], doffs[2 * k + 2] - doffs[2 * k + 1], m->ns, np, false, true, true,
                      (tag + " (time factor)").c_str());
        if (kind != NC_PROP_SEPARABLE)
          NC_REQUIRE(doffs[2 * k + 2] == doffs[2 * k + 1], NC_ERR_ARG, tag + ": only separable propensities have a time-factor derivative");
        if (!dep)
          NC_REQUIRE(doffs[2 * k + 2] == doffs[2 * k], NC_ERR_ARG, tag + ": derivative given outside the sparsity pattern");
      }
    m->pattern.assign(pattern, pattern + npairs);
    m->doffs.assign(doffs, doffs + 2 * npairs + 1);
    m->dcode.assign(dcode, dcode + total);
    m->d_dcode.alloc(std::max<int64_t>(total, 1));
    if (total) NC_CHECK_HIP(hipMemcpy(m->d_dcode.p, dcode, sizeof(nc_instr) * total, hipMemcpyHostToDevice));
    m->d_doffs.alloc(2 * npairs + 1);
    NC_CHECK_HIP(hipMemcpy(m->d_doffs.p, doffs, sizeof(int64_t) * (2 * npairs + 1), hipMemcpyHostToDevice));
    m->has_grad = true;
  });
}

int nc_model_eval(nc_model m, int32_t r, int32_t which, double t, int64_t n, const int32_t* states, double* out) {
  return guard([&] {
    NC_REQUIRE(m && r >= 0 && r < m->nr && n >= 0 && (n == 0 || (states && out)), NC_ERR_ARG, "bad arguments");
    NC_REQUIRE(which == 0 || which == 1, NC_ERR_ARG, "which must be 0 (x program) or 1 (t program)");
    if (n == 0) return;
    const int64_t off = which == 0 ? m->offs[2 * r] : m->offs[2 * r + 1];
    const int64_t len = which == 0 ? m->xlen(r) : m->tlen(r);
    hipStream_t s = m->stream.s;
    DevBuf<int32_t> dx((size_t)n * m->ns);
    DevBuf<double> dout((size_t)n);
    NC_CHECK_HIP(hipMemcpyAsync(dx.p, states, sizeof(int32_t) * n * m->ns, hipMemcpyHostToDevice, s));
    k_eval<<<grid_for(n, 256), 256, 0, s>>>(m->d_code.p + off, len, t, dx.p, m->ns, n, m->d_params.p, dout.p);
    NC_CHECK_HIP(hipGetLastError());
    NC_CHECK_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
