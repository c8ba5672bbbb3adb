// Propensity programs and the model object (internal).
//
// A propensity is a postfix program over an FP64 stack (NumCME.jl's propensities are Julia
// closures, src/cmemodel/propensity.jl:50-153; the host shim lowers them to programs).
// The same evaluator runs on the host (time factors, once per stage time) and on the
// device (state factors, once per state during the matrix build).
#pragma once

#include <cmath>

#include "nc_internal.h"

namespace nc {

__host__ __device__ inline double prog_eval(const nc_instr* code, int64_t len, double t, const int32_t* x,
                                            const double* theta) {
  double st[NC_PROG_STACK];
  int sp = 0;
  for (int64_t k = 0; k < len; ++k) {
    const nc_instr in = code[k];
    switch (in.op) {
      case NC_OP_CONST: st[sp++] = in.c; break;
      case NC_OP_PARAM: st[sp++] = theta[in.arg]; break;
      case NC_OP_SPECIES: st[sp++] = (double)x[in.arg]; break;
      case NC_OP_TIME: st[sp++] = t; break;
      case NC_OP_ADD: --sp; st[sp - 1] = st[sp - 1] + st[sp]; break;
      case NC_OP_SUB: --sp; st[sp - 1] = st[sp - 1] - st[sp]; break;
      case NC_OP_MUL: --sp; st[sp - 1] = st[sp - 1] * st[sp]; break;
      case NC_OP_DIV: --sp; st[sp - 1] = st[sp - 1] / st[sp]; break;
      case NC_OP_POW: --sp; st[sp - 1] = pow(st[sp - 1], st[sp]); break;
      case NC_OP_NEG: st[sp - 1] = -st[sp - 1]; break;
      case NC_OP_EXP: st[sp - 1] = exp(st[sp - 1]); break;
      case NC_OP_LOG: st[sp - 1] = log(st[sp - 1]); break;
      case NC_OP_SIN: st[sp - 1] = sin(st[sp - 1]); break;
      case NC_OP_COS: st[sp - 1] = cos(st[sp - 1]); break;
      case NC_OP_SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
      case NC_OP_ABS: st[sp - 1] = fabs(st[sp - 1]); break;
      case NC_OP_MAX: { --sp; double a = st[sp - 1], b = st[sp]; st[sp - 1] = (b < a) ? a : b; } break;
      case NC_OP_MIN: { --sp; double a = st[sp - 1], b = st[sp]; st[sp - 1] = (b < a) ? b : a; } break;
      case NC_OP_LT: --sp; st[sp - 1] = (st[sp - 1] < st[sp]) ? 1.0 : 0.0; break;
      case NC_OP_LE: --sp; st[sp - 1] = (st[sp - 1] <= st[sp]) ? 1.0 : 0.0; break;
      case NC_OP_GT: --sp; st[sp - 1] = (st[sp - 1] > st[sp]) ? 1.0 : 0.0; break;
      case NC_OP_GE: --sp; st[sp - 1] = (st[sp - 1] >= st[sp]) ? 1.0 : 0.0; break;
      case NC_OP_SELECT: { sp -= 2; double c = st[sp - 1], a = st[sp], b = st[sp + 1]; st[sp - 1] = (c != 0.0) ? a : b; } break;
      default: break;
    }
  }
  return sp > 0 ? st[sp - 1] : 0.0;
}

// Register-resident evaluator for programs whose stack stays within kRegStack entries (the
// model records its deepest program). A program's instructions are the same for every lane
// of a wave, so the stack pointer is wave-uniform: the stack slots are eight named
// registers picked by scalar branches, and nothing goes to per-lane scratch memory (the
// array stack of prog_eval does, once per inlined call). Species values come from xs(index).
// The same operations in the same order as prog_eval, so the same results.
constexpr int kRegStack = 8;
template <class XS>
__device__ inline double prog_eval_reg(const nc_instr* code, int64_t len, double t, XS xs, const double* theta) {
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0, r4 = 0.0, r5 = 0.0, r6 = 0.0, r7 = 0.0;
  auto get = [&](int k) -> double {
    switch (k) {
      case 0: return r0;
      case 1: return r1;
      case 2: return r2;
      case 3: return r3;
      case 4: return r4;
      case 5: return r5;
      case 6: return r6;
      default: return r7;
    }
  };
  auto put = [&](int k, double v) {
    switch (k) {
      case 0: r0 = v; break;
      case 1: r1 = v; break;
      case 2: r2 = v; break;
      case 3: r3 = v; break;
      case 4: r4 = v; break;
      case 5: r5 = v; break;
      case 6: r6 = v; break;
      default: r7 = v; break;
    }
  };
  int sp = 0;
  for (int64_t k = 0; k < len; ++k) {
    const nc_instr in = code[k];
    switch (in.op) {
      case NC_OP_CONST: put(sp++, in.c); break;
      case NC_OP_PARAM: put(sp++, theta[in.arg]); break;
      case NC_OP_SPECIES: put(sp++, (double)xs(in.arg)); break;
      case NC_OP_TIME: put(sp++, t); break;
      case NC_OP_ADD: { const double b = get(--sp); put(sp - 1, get(sp - 1) + b); } break;
      case NC_OP_SUB: { const double b = get(--sp); put(sp - 1, get(sp - 1) - b); } break;
      case NC_OP_MUL: { const double b = get(--sp); put(sp - 1, get(sp - 1) * b); } break;
      case NC_OP_DIV: { const double b = get(--sp); put(sp - 1, get(sp - 1) / b); } break;
      case NC_OP_POW: { const double b = get(--sp); put(sp - 1, pow(get(sp - 1), b)); } break;
      case NC_OP_NEG: put(sp - 1, -get(sp - 1)); break;
      case NC_OP_EXP: put(sp - 1, exp(get(sp - 1))); break;
      case NC_OP_LOG: put(sp - 1, log(get(sp - 1))); break;
      case NC_OP_SIN: put(sp - 1, sin(get(sp - 1))); break;
      case NC_OP_COS: put(sp - 1, cos(get(sp - 1))); break;
      case NC_OP_SQRT: put(sp - 1, sqrt(get(sp - 1))); break;
      case NC_OP_ABS: put(sp - 1, fabs(get(sp - 1))); break;
      case NC_OP_MAX: { const double b = get(--sp), a = get(sp - 1); put(sp - 1, (b < a) ? a : b); } break;
      case NC_OP_MIN: { const double b = get(--sp), a = get(sp - 1); put(sp - 1, (b < a) ? b : a); } break;
      case NC_OP_LT: { const double b = get(--sp); put(sp - 1, (get(sp - 1) < b) ? 1.0 : 0.0); } break;
      case NC_OP_LE: { const double b = get(--sp); put(sp - 1, (get(sp - 1) <= b) ? 1.0 : 0.0); } break;
      case NC_OP_GT: { const double b = get(--sp); put(sp - 1, (get(sp - 1) > b) ? 1.0 : 0.0); } break;
      case NC_OP_GE: { const double b = get(--sp); put(sp - 1, (get(sp - 1) >= b) ? 1.0 : 0.0); } break;
      case NC_OP_SELECT: {
        sp -= 2;
        const double c = get(sp - 1), a = get(sp), b = get(sp + 1);
        put(sp - 1, (c != 0.0) ? a : b);
      } break;
      default: break;
    }
  }
  return sp > 0 ? get(sp - 1) : 0.0;
}

// Deepest stack a (valid) program reaches.
int prog_max_depth(const nc_instr* code, int64_t len);

// Host-side validation: stack discipline, operand indices. `allow_x` / `allow_t` say
// whether the program may read species / time. Empty programs are allowed only when
// `allow_empty`.
void prog_validate(const nc_instr* code, int64_t len, int ns, int np, bool allow_x, bool allow_t, bool allow_empty,
                   const char* what);

}  // namespace nc

struct nc_model_s {
  int ns = 0, nr = 0, np = 0;
  std::vector<int32_t> stoich;  // ns x nr column-major
  std::vector<double> params;
  std::vector<int32_t> kinds;
  std::vector<int64_t> offs;  // 2*nr+1
  std::vector<nc_instr> code;
  nc::DevBuf<int32_t> d_stoich;
  nc::DevBuf<nc_instr> d_code;
  nc::DevBuf<int64_t> d_offs;
  nc::DevBuf<double> d_params;
  // gradients (CmeModelWithSensitivity)
  int max_depth = 0;  // deepest stack of the propensity programs (prog_eval_reg up to kRegStack)
  bool has_grad = false;
  std::vector<uint8_t> pattern;  // nr x np
  std::vector<int64_t> doffs;    // 2*nr*np+1
  std::vector<nc_instr> dcode;
  nc::DevBuf<nc_instr> d_dcode;
  nc::DevBuf<int64_t> d_doffs;
  nc::Stream stream;

  const nc_instr* xprog(int r) const { return code.data() + offs[2 * r]; }
  int64_t xlen(int r) const { return offs[2 * r + 1] - offs[2 * r]; }
  const nc_instr* tprog(int r) const { return code.data() + offs[2 * r + 1]; }
  int64_t tlen(int r) const { return offs[2 * r + 2] - offs[2 * r + 1]; }
  double tfactor(int r, double t) const { return nc::prog_eval(tprog(r), tlen(r), t, nullptr, params.data()); }
  // derivative programs of pair (r, p)
  const nc_instr* dxprog(int r, int p) const { return dcode.data() + doffs[2 * (r * np + p)]; }
  int64_t dxlen(int r, int p) const { return doffs[2 * (r * np + p) + 1] - doffs[2 * (r * np + p)]; }
  const nc_instr* dtprog(int r, int p) const { return dcode.data() + doffs[2 * (r * np + p) + 1]; }
  int64_t dtlen(int r, int p) const { return doffs[2 * (r * np + p) + 2] - doffs[2 * (r * np + p) + 1]; }
};
