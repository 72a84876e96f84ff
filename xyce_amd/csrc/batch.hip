// Batched Monte Carlo engine: thousands of independent circuits of uneven
// size, each evaluated, stamped, factored and solved by one workgroup, one
// launch per Newton iteration of every instance, the launches captured in a
// hipGraph.  Each instance runs its own DC operating point and its own
// OneStep (trapezoid) step control on the device, so no host round trip
// happens between launches.
//
// Per launch and instance (reference call stack, paths relative to the Xyce
// source root):
//   1. device load at x       DeviceMgr::updateState / loadDAEVectors
//                             (src/DeviceModelPKG/Core/N_DEV_DeviceMgr.C:3857,4156)
//   2. residual               OneStep::obtainResidual (src/TimeIntegrationPKG/N_TIA_OneStep.C:219)
//   3. convergence of the previous update
//                             DampedNewton::converged_ (src/NonlinearSolverPKG/N_NLS_DampedNewton.C:1191)
//   4a. not converged: Jacobian (loadDAEMatrices + obtainJacobian :469), LU
//       refactor + solve (AmesosSolver::doSolve), x += dx (updateX_ :808)
//   4b. converged: error estimate and completeStep / rejectStep
//       (N_TIA_OneStep.C:2124,1946), updateHistory (:1390), next predictor (:135)
//
// Instances with identical structure share one copy of their pattern, gather
// maps and LU schedules ("structure classes": all rings with the same stage
// count share one).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <vector>

#include "mos1_device.h"
#include "xg_internal.h"

#define HIPCHK(x)                          \
  do {                                     \
    hipError_t e_ = (x);                   \
    if (e_ != hipSuccess) return XG_E_HIP; \
  } while (0)

namespace xg {

// Every multiply-add in the engine is rounded twice, as in a build of the
// reference for baseline x86-64 and in the oracle: no fused multiply-adds in
// the residual and Jacobian assembly (EpetraMatrix::linearCombo,
// OneStep::obtainResidual), the LU and solves, the Newton update and norms,
// the predictor and the step control.
#pragma clang fp contract(off)

// ---- per structure class (offsets into the concatenated int arrays)
struct ClassDesc {
  int n, nnz, nlu, n_mos, n_src, nops, nterm;
  // term array of an instance: tsize = kterm * n_mos doubles, the term of
  // device d with rank r among its nonzero planes at r * n_mos + d
  int kterm, tsize;
  int nflev, nbs;
  // int32 arrays (global): device node indices, per-device plane masks
  // (low words, then high words), the term array index of every slot-ordered
  // term (o_tix, nterm entries), term ranges, sources
  int o_lid, o_tmask, o_tix, o_vptr, o_mptr, o_bgp, o_bge;
  // groups of >= kLongSeg terms, summed by a whole wavefront: vector groups
  // (4*row + vector) then matrix groups (2*slot + F/Q), at o_long
  int n_vlong, n_mlong, o_long;
  // uint16 program blob (staged into LDS each launch); offsets in uint16 units
  int blob_off, blob_len, blob_lds;   // blob_lds: the leading part staged into LDS (the rest is read in HBM)
  int b_flp, b_fmeta, b_fops, b_lurp, b_luci, b_diag, b_bsp, b_bsr, b_prow, b_pcol,
      b_rp, b_cc, b_a2lu;
  // long LU targets / solve rows (>= kLongOps terms, one wavefront each) come
  // last in their level: first long index per level
  int b_fls, b_bss;
  int b_vrow;        // n flags: 1 = node-voltage row (gmin stepping adds its conductance there)
  // term ranges, linear entries (CSR slots with a nonzero linear value, per
  // row) and per-row source lists, all uint16
  int b_vptr, b_mptr, b_lrp, b_lk, b_bgp, b_bge;
};

// ---- per instance state (one 512-byte record in HBM)
struct InstState {
  int mode;          // 0 DC, 1 transient, 2 done, -1 DC failed, -2 step failure
  int newton_iter;
  int order;
  int nef;
  int beginning;
  int bpi;
  int stall_count;
  int last_status;
  long long n_iters, n_loads, n_steps, n_rejects, n_rec, n_dense;
  double t, h, psi0, psi1, psi2, last_h, h_sum, saved_h, est, ck, beta1, beta2, errwt;
  double norm_init, norm_old, tmp_rate, wt_dx;
  double steps_taken;
  // learned predictor: step sizes of the last 4 accepted steps (newest
  // first), accepted points in the history, flags and counters
  double hP[4];
  int n_acc, feat_valid, pred_skip, used_learned;
  long long n_pairs, n_learned, n_learned_rejects;
  // Newton iterations per step started from each predictor (decaying
  // averages): Newton convergence decides between the network and the
  // polynomial predictor
  double it_learned, st_learned, it_poly, st_poly;
  long long n_choice;
  // Gear12: predictor weight of x_n and the corrector weights of Q_{n+1},
  // Q_n, Q_{n-1} (N_TIA_Gear12.C:1050-1111)
  double beta0, a0, a1, a2;
  // DC operating point strategy (Interface::spiceStrategy): stage 0 Newton,
  // 1 gmin stepping, 2 source stepping; the continuation parameter, its last
  // converged value and step, the conductance and source scale in force,
  // whether the next solve starts from initJct, and the steps taken
  double dc_stage, cont_par, cont_ok, cont_step, gcond, src_scale, cont_first, cont_k;
  // which of the two store/state buffers holds the last accepted values (the
  // other takes the iterate's)
  int sbuf, pad_i;
  double pad[9];
};

static_assert(sizeof(InstState) == 512, "InstState is one 512-byte record");

struct InstDesc {
  int cls, dev_off, vec_off, nnz_off, src_off, rec_off;
  long long dense_off;   // n*n scratch for the partial-pivoting fallback
  long long term_off;    // this instance's term array (device-major, kterm x n_mos)
  int bp_off, n_bp;      // this instance's breakpoints (sorted, ending at tstop)
  int win_off;           // first predictor window (rows of 256 unknowns)
  long long pair_off;    // first training-pair record (capacity * windows)
};

struct BatchArgs {
  int n_inst;
  const int* order;
  const InstDesc* inst;
  InstState* st;
  const ClassDesc* cls;
  const int* ci;               // concatenated class int arrays
  // devices
  int nt_mos;
  const double* kcls; const int* mcls; const double* tvbi; const double* beta; const double* tvto;
  const unsigned long long* mask;
  // store / state values, two buffers each: an instance's accepted values are
  // in buffer sbuf, its iterate's in the other (accepting a step flips sbuf)
  double* sto_cur; double* sto_nxt; double* sta_cur; double* sta_nxt;
  double* terms;               // per-instance term arrays (device-major)
  const unsigned short* blob;  // per-class uint16 programs
  // per-instance vectors (offset vec_off, length n)
  double* x; double* xn0; double* xH0; double* xH1; double* xH2; double* qH0; double* qH1; double* qH2;
  double* F; double* Q; double* B; double* F0; double* B0; double* wt;
  // linear part (offset nnz_off) and sources (offset src_off)
  const double* linF; const double* linQ;
  const int* src_wave; const int* wave_type; const int* wave_off; const double* wave_data;
  // breakpoints: per instance, concatenated (InstDesc::bp_off / n_bp)
  const double* bp;
  // transient parameters
  double tstop, tstep, hmax;
  int method;                  // 0 = OneStep trapezoid, 1 = Gear12 (BDF1/BDF2)
  // recording
  int rec_max; double* rec_t; double* rec_x;
  int max_n, max_nlu, max_blob;   // over all instances (max_n: stride of the recorded points)
  int lds_n, lds_nlu;              // LDS layout of this launch's instances (their largest class)
  int lds_terms, lds_blob;         // fused form: the term array in LDS after the program (lds_blob uint16)
  int n_kcls;                  // entries of the MOSFET class table
  int iters;                   // Newton iterations per launch
  unsigned long long* prof;    // per-phase wall-clock ticks of one instance (null = off)
  double* dense;           // per-instance n*n scratch
  // separate device-evaluation kernel: owning instance of every MOSFET and
  // the per-instance "some device limited its voltages" flag it raises
  const int* dev_inst;
  int* lim;
  // learned predictor (mode 0 off, 1 collect training pairs, 2 serve)
  int pred_mode;
  int pred_base;           // the network corrects 0: the polynomial predictor, 1: the cubic extrapolation
  double pred_dead;        // corrections |y| <= pred_dead (Newton weights) are not applied
  int pair_cap;            // training pairs per instance
  double* xP;              // last 4 accepted solutions and x_cubic, planes of total_n
  long long total_n;
  unsigned short* feat;    // [windows][XG_PRED_FEAT] bf16 features of the current step
  const float* pred_out;   // [windows][XG_PRED_WIN] network output
  unsigned short* pair_x;  // [records][XG_PRED_FEAT]
  float* pair_t;           // [records][XG_PRED_WIN] target (x_new - x_poly) / scale
  float* pair_m;           // [records][XG_PRED_WIN] 1 for rows that exist
};

constexpr double kRelTol = 1e-3, kAbsTol = 1e-6;                       // TIA defaults
constexpr double kNlRel = 1e-2, kNlAbs = 1e-6, kDxTol = 0.33, kRhsTol = 1e-2;  // transient Newton
// learned predictor: training targets clipped to +-kPredClip Newton weights
// before compression; proposals with an output beyond g(512) = 9 are not used
constexpr double kPredClip = 256.0;
constexpr float kPredTrustLog = 9.0f;
constexpr int kMaxNewton = 20;
constexpr double kDcRel = 1e-3, kDcAbs = 1e-12, kDcDxTol = 1.0, kDcRhsTol = 1e-6;
constexpr int kDcMaxNewton = 200;
constexpr int kContMaxNewton = 20;       // per continuation step (LOCA "Max Nonlinear Iterations")
constexpr double kHmin = 1e-18;
constexpr int kLongSeg = 32;     // term groups this long are reduced by a wavefront
constexpr int kMaxLong = 16;     // per class
constexpr int kLongOps = 16;     // LU targets / solve rows this long are reduced by a wavefront
constexpr int kMaxKclsLds = 4;   // MOSFET class records the fused kernel stages in LDS
constexpr int kMaxSrc = 64;      // sources per instance (their values are staged in LDS)

// Sum over rows 0..n-1 whose value does not depend on the workgroup size:
// each 64-row chunk is summed by one wavefront with a fixed shuffle tree
// (lane = row mod 64), the chunk sums are added in chunk order.  Callers loop
// `for (base = 64 * wave; base < n; base += blockDim.x)` (row = base + lane,
// the same trip count for the whole wavefront), sum each chunk with
// wave_tree_sum, store it in chs[base / 64] from lane 0 and then call
// chunk_total.
constexpr int kMaxChunks = 128;          // rows per instance <= 64 * kMaxChunks
// An instance's accepted (cur) and iterate (nxt) store / state planes.
struct DevState {
  double *sto_cur, *sto_nxt, *sta_cur, *sta_nxt;
};
__device__ __forceinline__ DevState dev_state(const BatchArgs& A, int sbuf) {
  DevState d;
  d.sto_cur = sbuf ? A.sto_nxt : A.sto_cur; d.sto_nxt = sbuf ? A.sto_cur : A.sto_nxt;
  d.sta_cur = sbuf ? A.sta_nxt : A.sta_cur; d.sta_nxt = sbuf ? A.sta_cur : A.sta_nxt;
  return d;
}

__device__ __forceinline__ double wave_tree_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// Every thread adds the chunk sums itself, in the same order, so one barrier
// suffices; chs is written again only after later phases' barriers.
__device__ __forceinline__ double chunk_total(const double* chs, int n) {
  __syncthreads();
  double r = 0.0;
  for (int c = 0; c < (n + 63) / 64; ++c) r += chs[c];
  return r;
}
// Block maximum: 64-lane wave shuffles, then one LDS exchange across waves
// (a maximum does not depend on the order).
__device__ __forceinline__ double block_max(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = red[0];
  for (int i = 1; i < nw; ++i) r = fmax(r, red[i]);
  __syncthreads();
  return r;
}

// J = a*dQdx + b*dFdx (EpetraMatrix::linearCombo rounding)
__device__ __forceinline__ double jac_combine(double a, double dq, double b, double df) { return a * dq + b * df; }

// A device's nonzero record planes are a 48-bit mask (lo: planes 0-31, hi:
// planes 32-47); plane p's term sits at its rank among them.  With p a
// constant at every call site both reduce to a mask test and a bit count.
__device__ __forceinline__ bool term_on(unsigned lo, unsigned hi, int p) {
  return p < 32 ? (lo >> p) & 1u : (hi >> (p - 32)) & 1u;
}
__device__ __forceinline__ int term_rank(unsigned lo, unsigned hi, int p) {
  return p < 32 ? __popc(lo & ((1u << p) - 1u)) : __popc(lo) + __popc(hi & ((1u << (p - 32)) - 1u));
}

// Sums the term groups [b[0], b[1]), ..., [b[G-1], b[G]) (consecutive in slot
// order; tix maps a slot to its place in the term array) into out[0..G-1],
// each in slot order from 0.0.  The span's indices are loaded 8 at a time
// before any of their values, so a span costs one round trip per 8 terms.
template <int G>
__device__ __forceinline__ void gather_groups(const double* T, const int* tix, const int (&b)[G + 1], double (&out)[G]) {
#pragma unroll
  for (int v = 0; v < G; ++v) out[v] = 0.0;
  for (int c = b[0]; c < b[G]; c += 8) {
    int ix[8];
    double tv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) ix[q] = c + q < b[G] ? tix[c + q] : 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) tv[q] = c + q < b[G] ? T[ix[q]] : 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = c + q;
#pragma unroll
      for (int v = 0; v < G; ++v)
        if (e >= b[v] && e < b[v + 1]) out[v] += tv[q];
    }
  }
}
// residual combine (OneStep::obtainResidual / NoTimeIntegration::obtainResidual rounding)
__device__ __forceinline__ double residual_combine(bool dc, int order, double h, double f, double q, double b,
                                                   double fl, double ql, double q0, double q2) {
  double res;
  if (dc) {
    res = 1.0 * f + -1.0 * b;
    res *= -1.0;
    res = res + 1.0 * fl;
  } else {
    res = 1.0 * q + -1.0 * q0;
    if (order == 2) {
      res = (1.0 / h) * res + 0.5 * f + -0.5 * b;
      res = res + 0.5 * q2;
    } else {
      res = (1.0 / h) * res + 1.0 * f + -1.0 * b;
    }
    res *= -1.0;
    res = res + 1.0 * ((1.0 / h) * ql);
    res = res + (order == 2 ? 0.5 : 1.0) * fl;
  }
  return res;
}
// Gear12::obtainResidual (N_TIA_Gear12.C:208-263) rounding: q0 = Q_n, q1 = Q_{n-1}
__device__ __forceinline__ double residual_gear(int order, double h, double a0, double a1, double a2, double f,
                                                double q, double b, double fl, double ql, double q0, double q1) {
  double res = a0 * q + a1 * q0;
  if (order == 2) res = res + a2 * q1;
  res = (1.0 / h) * res + 1.0 * f + -1.0 * b;
  res *= -1.0;
  res = res + 1.0 * ((a0 / h) * ql);
  res = res + 1.0 * fl;
  return res;
}

__device__ double next_bp(const BatchArgs& A, const InstDesc& I, InstState& S) {
  const double* bp = A.bp + I.bp_off;
  while (S.bpi < I.n_bp && bp[S.bpi] <= S.t * (1 + 1e-13) + 1e-30) ++S.bpi;
  return S.bpi < I.n_bp ? bp[S.bpi] : A.tstop;
}

__device__ __forceinline__ unsigned short d_f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// Learned-predictor features of the step about to be taken, per window of
// XG_PRED_WIN unknowns, in units of each unknown's Newton weight
//   w = 1e-2 max(|x_n|, |x_poly|) + 1e-6
// (the transient Newton weight of the first iteration, DampedNewton's
// updateWeights_): with x_n..x_{n-3} the last accepted solutions, x_poly the
// polynomial predictor and x_cubic the cubic through the 4 points at t_{n+1},
// the planes
//   g((x_n - x_{n-1}) / w), g((x_{n-1} - x_{n-2}) / w), g((x_{n-2} - x_{n-3}) / w),
//   g((x_poly - x_n) / w), g((x_cubic - x_poly) / w), x_n,
// g(v) = sign(v) log2(1 + |v|) (BF16 keeps 3 digits of every magnitude,
// so nodes near 0 V are resolved as well as the rails), then the step-size
// ratios.  The network corrects a base extrapolation computed here in FP64,
// the polynomial predictor or the cubic (pred_base), in the same compressed
// units: its output g gives y = sign(g)(2^|g| - 1) and x = x_base + w y
// (g = 0 proposes the base itself), so that a correction of a fraction of a
// weight is learned as precisely, relative to its size, as one of hundreds; a
// proposal within 1/3 of a weight of the solution on every unknown converges
// in one Newton iteration.  x_base is kept in the fifth history plane.
__device__ __forceinline__ double pred_weight(double xn, double xp) { return kNlRel * fmax(fabs(xn), fabs(xp)) + kNlAbs; }
__device__ __forceinline__ float pred_slog(double v) { return (float)copysign(log2(1.0 + fabs(v)), v); }
__device__ __forceinline__ double pred_sexp(double g) { return copysign(exp2(fabs(g)) - 1.0, g); }

__device__ void pred_features(const BatchArgs& A, const InstDesc& I, const ClassDesc& C, InstState& S) {
  const int n = C.n, off = I.vec_off, tid = threadIdx.x, nth = blockDim.x;
  const long long TN = A.total_n;
  const int nw = (n + XG_PRED_WIN - 1) / XG_PRED_WIN;
  // Lagrange weights of the cubic at t_{n+1} (times relative to t_n)
  const double h1 = S.h, ha = S.hP[0], hb = S.hP[1], hc = S.hP[2];
  const double tau[4] = {0.0, -ha, -ha - hb, -ha - hb - hc};
  double L[4];
  for (int j = 0; j < 4; ++j) {
    double v = 1.0;
    for (int k = 0; k < 4; ++k)
      if (k != j) v *= (h1 - tau[k]) / (tau[j] - tau[k]);
    L[j] = v;
  }
  for (int w = 0; w < nw; ++w) {
    const int r0 = w * XG_PRED_WIN, r1 = min(n, r0 + XG_PRED_WIN);
    unsigned short* f = A.feat + (size_t)(I.win_off + w) * XG_PRED_FEAT;
    for (int j = tid; j < XG_PRED_WIN; j += nth) {
      const int r = r0 + j;
      float v[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (r < r1) {
        const double x0 = A.xP[off + r], x1 = A.xP[TN + off + r], x2 = A.xP[2 * TN + off + r],
                     x3 = A.xP[3 * TN + off + r];
        const double xp = A.xn0[off + r];
        const double xc = L[0] * x0 + L[1] * x1 + L[2] * x2 + L[3] * x3;
        const double wr = pred_weight(x0, xp);
        A.xP[4 * TN + off + r] = A.pred_base ? xc : xp;
        v[0] = pred_slog((x0 - x1) / wr); v[1] = pred_slog((x1 - x2) / wr); v[2] = pred_slog((x2 - x3) / wr);
        v[3] = pred_slog((xp - x0) / wr); v[4] = pred_slog((xc - xp) / wr); v[5] = (float)x0;
      }
      for (int q = 0; q < 6; ++q) f[q * XG_PRED_WIN + j] = d_f2bf(v[q]);
    }
    for (int j = 6 * XG_PRED_WIN + tid; j < XG_PRED_FEAT; j += nth) {
      const int q = j - 6 * XG_PRED_WIN;
      float v = 0.f;
      if (q == 0) v = (float)(h1 / ha);
      else if (q == 1) v = (float)(ha / hb);
      else if (q == 2) v = (float)(hb / hc);
      else if (q == 3) v = (float)S.order;
      else if (q == 4) v = (float)((log10(h1) + 12.0) / 4.0);
      else if (q == 5) v = 1.0f;
      f[j] = d_f2bf(v);
    }
  }
}

// Per-instance predictor selector: decaying averages of the Newton
// iterations of steps started from the network's proposal and from the
// polynomial predictor.
__device__ __forceinline__ void pred_score(InstState& S, bool learned, int iters) {
  const double k = 0.97;
  if (learned) { S.it_learned = k * S.it_learned + iters; S.st_learned = k * S.st_learned + 1.0; }
  else { S.it_poly = k * S.it_poly + iters; S.st_poly = k * S.st_poly + 1.0; }
}
// Use the network for this step?  Both predictors are tried while their
// averages are young and every 16th choice goes to the one not preferred;
// otherwise the one with fewer iterations per step wins.
__device__ __forceinline__ bool pred_choose(InstState& S) {
  const long long c = S.n_choice++;
  bool pick;
  if (S.st_learned < 4.0 || S.st_poly < 4.0) pick = (c & 1) == 0;
  else pick = S.it_learned / S.st_learned <= S.it_poly / S.st_poly;
  if (c % 16 == 15) pick = !pick;
  return pick;
}

// Accepted step: the training pair of this step (collect mode) and the
// shift of the solution history.
__device__ void pred_accept(const BatchArgs& A, const InstDesc& I, const ClassDesc& C, InstState& S,
                            const double* xs) {
  const int n = C.n, off = I.vec_off, tid = threadIdx.x, nth = blockDim.x;
  const long long TN = A.total_n;
  const int nw = (n + XG_PRED_WIN - 1) / XG_PRED_WIN;
  if (A.pred_mode == 1 && S.feat_valid && S.n_pairs < A.pair_cap) {
    for (int w = 0; w < nw; ++w) {
      const size_t rec = (size_t)I.pair_off + (size_t)S.n_pairs * nw + w;
      const unsigned short* f = A.feat + (size_t)(I.win_off + w) * XG_PRED_FEAT;
      for (int j = tid; j < XG_PRED_FEAT; j += nth) A.pair_x[rec * XG_PRED_FEAT + j] = f[j];
      for (int j = tid; j < XG_PRED_WIN; j += nth) {
        const int r = w * XG_PRED_WIN + j;
        const bool ok = r < n;
        double tv = 0.0;
        if (ok) {
          const double xp = A.xn0[off + r];
          tv = (xs[r] - A.xP[4 * TN + off + r]) / pred_weight(A.xP[off + r], xp);
          tv = pred_slog(fmin(fmax(tv, -kPredClip), kPredClip));
        }
        A.pair_t[rec * XG_PRED_WIN + j] = (float)tv;
        A.pair_m[rec * XG_PRED_WIN + j] = ok ? 1.f : 0.f;
      }
    }
  }
  for (int r = tid; r < n; r += nth) {
    A.xP[3 * TN + off + r] = A.xP[2 * TN + off + r];
    A.xP[2 * TN + off + r] = A.xP[TN + off + r];
    A.xP[TN + off + r] = A.xP[off + r];
    A.xP[off + r] = xs[r];
  }
  __syncthreads();
  if (tid == 0) {
    if (A.pred_mode == 2 && S.feat_valid) pred_score(S, S.used_learned != 0, S.newton_iter);
    if (A.pred_mode == 1 && S.feat_valid && S.n_pairs < A.pair_cap) S.n_pairs++;
    S.hP[3] = S.hP[2]; S.hP[2] = S.hP[1]; S.hP[1] = S.hP[0]; S.hP[0] = S.h;
    S.n_acc++;
    S.feat_valid = 0;
    S.pred_skip = 0;
    S.used_learned = 0;
  }
}

// Gear12::updateCoeffs (N_TIA_Gear12.C:1050-1111): variable-step BDF2
// predictor and corrector weights, or BDF1 (psi already shifted; out of line
// to keep the iteration kernel's registers)
__device__ __noinline__ void gear_coeffs(InstState& S) {
  const double t1 = S.h;
  if (S.order == 2) {
    S.beta2 = t1 / S.psi2 * (t1 + S.psi1) / (S.psi1 + S.psi2);
    S.beta1 = -t1 / S.psi1 - S.beta2 * (S.psi1 + S.psi2) / S.psi1;
    S.beta0 = 1.0 - S.beta2 - S.beta1;
    double a2 = -t1 / S.psi1 * t1 / (2 * t1 + S.psi1);
    double a1 = 1 - a2;
    double a0 = -a1 - a2 * (1 + S.psi1 / t1);
    S.a2 = a2 / a0; S.a1 = a1 / a0; S.a0 = -1 / a0;
    S.ck = S.h / (t1 + S.psi1 + S.psi2);
  } else {
    S.beta0 = 1.0 + t1 / S.psi1;
    S.beta1 = -t1 / S.psi1;
    S.beta2 = 0.0;
    S.ck = S.h / (t1 + S.psi1);
  }
}

// Top of a time step: clamp to the next breakpoint, OneStep::updateCoeffs
// (N_TIA_OneStep.C:1584), predictor (:135), error weights (NEWLTE=1).
__device__ void begin_step(const BatchArgs& A, InstState& S, const ClassDesc& C, const InstDesc& I, double* red) {
  __shared__ double sh[4];
  if (threadIdx.x == 0) {
    const double stop = next_bp(A, I, S);
    if (S.t + S.h > stop) { S.saved_h = S.h; S.h = stop - S.t; S.h_sum = S.h + S.last_h; }
    const double t1 = S.h;
    if (S.order == 2) S.psi2 = S.psi1;
    S.psi1 = S.psi0; S.psi0 = t1;
    S.beta0 = 1.0; S.a0 = 1.0; S.a1 = -1.0; S.a2 = 0.0;
    if (A.method == 1) {
      gear_coeffs(S);
    } else if (S.order == 2) {
      const double t2 = S.psi1;
      S.beta1 = t1 / t2 + (t1 / t2) * (t1 / t2) / 2;
      S.beta2 = -1.0 * t1 * t1 / t2 / S.psi2 / 2;
      S.ck = (S.h / S.h_sum) / 3.0;
    } else {
      S.beta1 = t1 / S.psi1;
      S.beta2 = 0.0;
      S.ck = S.h / S.h_sum;
    }
    S.newton_iter = 0;
    sh[0] = S.beta1; sh[1] = S.beta2; sh[2] = (double)S.order; sh[3] = S.beta0;
  }
  __syncthreads();
  const double b1 = sh[0], b2 = sh[1], b0 = sh[3];
  const int order = (int)sh[2];
  const bool gear = A.method == 1;
  double mx = 0.0;
  const int off = I.vec_off;
  for (int i = threadIdx.x; i < C.n; i += blockDim.x) {
    double v;
    if (gear) {
      // Gear12::obtainPredictor: sum of beta_i x_{n-i}
      v = b0 * A.xH0[off + i];
      v = b1 * A.xH1[off + i] + v;
      if (order == 2) v = b2 * A.xH2[off + i] + v;
    } else {
      v = A.xH0[off + i];
      v += b1 * A.xH1[off + i];
      if (order == 2) v += b2 * A.xH2[off + i];
    }
    A.xn0[off + i] = v;
    A.x[off + i] = v;
    mx = fmax(mx, fabs(A.xH0[off + i]));
  }
  mx = block_max(mx, red);
  if (threadIdx.x == 0) S.errwt = kRelTol * mx + kAbsTol;
  __syncthreads();
  if (A.pred_mode != 0 && S.n_acc >= 4) {
    pred_features(A, I, C, S);
    __syncthreads();
    if (threadIdx.x == 0) {
      S.feat_valid = 1;
      // serve: wait for the network's proposal (the apply kernel resumes the
      // instance); after a Newton failure with it, retry with the polynomial one
      if (A.pred_mode == 2 && !S.pred_skip && pred_choose(S)) S.mode = 3;
    }
    __syncthreads();
  }
}

// Dense LU with partial pivoting of one instance's Jacobian, in the
// instance's n*n HBM scratch, solving in place: rhs (LDS) becomes dx.  Used
// for DC iterations (ring oscillators at their DC point have gmin-sized
// diagonals: static pivots would grow by 1e8) and when the sparse static
// pivots fail KLU's threshold test, i.e. where KLU's re-pivoting refactor
// (N_LAS_AmesosSolver.C:318) would pick different pivots.  Returns nonzero if
// a zero pivot is met (singular).
__device__ int dense_pivot_solve(const BatchArgs& A, const ClassDesc& C, const InstDesc& I, const int* ci,
                                 const unsigned short* prog, const double* T, double ja, double jb, double gdiag,
                                 double* rhs, double* red, int* sh_i) {
  const int n = C.n, tid = threadIdx.x, nth = blockDim.x;
  double* D = A.dense + I.dense_off;
  const unsigned short* gprog = A.blob + C.blob_off;
  const unsigned short* rp = gprog + C.b_rp; const unsigned short* cc = gprog + C.b_cc;
  const unsigned short* mptr = gprog + C.b_mptr;
  const double* linF = A.linF + I.nnz_off; const double* linQ = A.linQ + I.nnz_off;
  const int* tix = ci + C.o_tix;
  for (int e = tid; e < n * n; e += nth) D[e] = 0.0;
  __syncthreads();
  for (int r = tid; r < n; r += nth)
    for (int s = rp[r]; s < rp[r + 1]; ++s) {
      double df = 0.0, dq = 0.0;
      for (int e = mptr[2 * s]; e < mptr[2 * s + 1]; ++e) df += T[tix[e]];
      for (int e = mptr[2 * s + 1]; e < mptr[2 * s + 2]; ++e) dq += T[tix[e]];
      if (linF[s] != 0.0) df += linF[s];
      if (linQ[s] != 0.0) dq += linQ[s];
      D[(size_t)r * n + cc[s]] = jac_combine(ja, dq, jb, df);
    }
  __syncthreads();
  if (gdiag != 0.0) {
    // gmin stepping: the continuation conductance on every node-voltage row
    // (GStepping::augmentJacobian)
    const unsigned short* vr = gprog + C.b_vrow;
    for (int r = tid; r < n; r += nth)
      if (vr[r]) D[(size_t)r * n + r] += gdiag;
    __syncthreads();
  }
  int singular = 0;
  for (int k = 0; k < n; ++k) {
    // pivot: largest |D(i,k)|, i >= k (lowest index on ties)
    double best = -1.0; int bi = k;
    for (int i = k + tid; i < n; i += nth) {
      const double v = fabs(D[(size_t)i * n + k]);
      if (v > best) { best = v; bi = i; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { red[tid >> 6] = best; sh_i[tid >> 6] = bi; }
    __syncthreads();
    double pmax = red[0];
    int p = sh_i[0];
    for (int w = 1; w < (nth >> 6); ++w)
      if (red[w] > pmax || (red[w] == pmax && sh_i[w] < p)) { pmax = red[w]; p = sh_i[w]; }
    __syncthreads();
    if (!(pmax > 0.0) || !isfinite(pmax)) { singular = 1; break; }
    if (p != k) {
      for (int j = tid; j < n; j += nth) {
        const double t = D[(size_t)k * n + j]; D[(size_t)k * n + j] = D[(size_t)p * n + j]; D[(size_t)p * n + j] = t;
      }
      if (tid == 0) { const double t = rhs[k]; rhs[k] = rhs[p]; rhs[p] = t; }
      __syncthreads();
    }
    const double piv = D[(size_t)k * n + k];
    const int m = n - k - 1;
    for (int i = k + 1 + tid; i < n; i += nth) D[(size_t)i * n + k] /= piv;
    __syncthreads();
    for (int e = tid; e < m * m; e += nth) {
      const int i = k + 1 + e / m, j = k + 1 + e % m;
      D[(size_t)i * n + j] -= D[(size_t)i * n + k] * D[(size_t)k * n + j];
    }
    if (tid == 0)
      for (int i = k + 1; i < n; ++i) rhs[i] -= D[(size_t)i * n + k] * rhs[k];
    __syncthreads();
  }
  if (singular) return 1;
  for (int k = n - 1; k >= 0; --k) {
    if (tid == 0) rhs[k] = rhs[k] / D[(size_t)k * n + k];
    __syncthreads();
    const double xk = rhs[k];
    for (int i = tid; i < k; i += nth) rhs[i] -= D[(size_t)i * n + k] * xk;
    __syncthreads();
  }
  return 0;
}

// One Newton iteration of one instance (the body of every launch; a launch
// runs A.iters of them back to back with the instance state kept in LDS).
// One decision of the DC strategy after a Newton solve ended with `status`
// (scalar code, kept out of line so that the iteration kernel's registers are
// not spent on it).  Returns 0 done (to transient), 1 next continuation step
// from the converged point, 2 failed, 3 a strategy starts over from zero,
// 4 retry from the last converged point.
__device__ __noinline__ int dc_strategy_step(InstState& S, int status) {
  int act = 0;       // 0 done (to transient), 1 new solve, 2 failed
  const int stage = (int)S.dc_stage;
  if (status > 0) {
    if (stage == 0) act = 0;
    else {
      S.cont_first = 0.0;
      S.cont_ok = S.cont_par;
      const double p1 = stage == 1 ? -4.0 : 1.0;
      if (S.cont_par == p1) act = 0;
      else {
        const double r = S.newton_iter / 20.0;
        S.cont_step *= 1.0 + (stage == 2 ? 1.0 : 0.01) * (1.0 - r * r);   // LOCA aggressiveness
        if (stage == 2 && fabs(S.cont_step) > 0.2) S.cont_step = 0.2;
        act = 1;
      }
    }
  } else if (stage == 0) {
    S.dc_stage = 1; S.cont_par = 4.0; S.cont_ok = 4.0; S.cont_step = -2.0; S.cont_first = 1.0; S.cont_k = 0.0;
    act = 3;         // restart from zero
  } else {
    S.cont_step *= 0.5;
    const double min_step = stage == 1 ? 1e-6 : 1e-4;
    if (fabs(S.cont_step) < min_step || S.cont_k >= 400.0) {
      if (stage == 1) {
        // source stepping disables initJct (setDisableInitJctFlags)
        S.dc_stage = 2; S.cont_par = 0.0; S.cont_ok = 0.0; S.cont_step = 0.2; S.cont_first = 0.0; S.cont_k = 0.0;
        act = 3;
      } else act = 2;
    } else act = 4;  // retry from the last converged point
  }
  if (act == 1 || act == 4) {
    const double p0 = stage == 1 ? 4.0 : 0.0, p1 = stage == 1 ? -4.0 : 1.0;
    S.cont_par = S.cont_ok + S.cont_step;
    if ((p1 - S.cont_par) * (p1 - p0) <= 0.0) S.cont_par = p1;
    S.cont_k += 1.0;
  }
  if (act == 1 || act == 3 || act == 4) {
    const int st2 = (int)S.dc_stage;
    S.gcond = st2 == 1 ? pow(10.0, S.cont_par) - pow(10.0, -4.0) : 0.0;
    S.src_scale = st2 == 2 ? S.cont_par : 1.0;
    S.newton_iter = 0;
  }
  if (act == 0) { S.gcond = 0.0; S.src_scale = 1.0; }
  if (act == 2) S.mode = -1;
  return act;
}

template <bool FUSED>
__device__ void batch_iteration(const BatchArgs& A, int inst, const InstDesc& I, const ClassDesc& C, InstState& S,
                                double* red, int* red_i, int* sh_int, double* sh_d, double* xs, double* rhs,
                                double* lu, const unsigned short* prog, const double* kcls,
                                double* srcv, double* T) {
  const int tid = threadIdx.x, nth = blockDim.x;
  const int n = C.n, off = I.vec_off;
  const int* ci = A.ci;
  const unsigned short* gprog = A.blob + C.blob_off;   // the program's part read in HBM
  const int NT = A.nt_mos;
  const bool dc = S.mode == 0;
  const DevState DS = dev_state(A, S.sbuf);
  // T: the instance's device-major term array (HBM, or LDS in the fused form)
  // phase timers (one instance, thread 0, after each phase's barrier); the
  // launch groups may run side by side on their own streams, so the sums are
  // atomic
  const bool prof_on = A.prof != nullptr && blockIdx.x == 0 && tid == 0;
  unsigned long long prof_t = prof_on ? wall_clock64() : 0;
#define XG_PHASE(k)                                              \
  do {                                                           \
    if (prof_on) {                                               \
      const unsigned long long t_ = wall_clock64();              \
      atomicAdd(&A.prof[k], t_ - prof_t);                        \
      prof_t = t_;                                               \
    }                                                            \
  } while (0)
  if (prof_on) atomicAdd(&A.prof[15], 1ull);
  // (the loop's barrier after the previous iteration has retired every
  // reader of these; their first use in this one follows the next barrier)
  if (tid < 8) { sh_int[tid] = 0; sh_d[tid] = 0.0; }
  for (int i = tid; i < n; i += nth) xs[i] = A.x[off + i];
  if (tid == 0) xs[n] = 0.0;
  if (!FUSED && tid == 0) {
    // the evaluation kernel ran just before this launch
    sh_int[0] = A.lim[inst];
    A.lim[inst] = 0;
  }
  __syncthreads();
  XG_PHASE(0);
  // ---- 2. device evaluation (one thread per MOSFET); each nonzero term goes
  //         to the device's column of the instance's term array, at its rank
  //         among the device's nonzero planes (an index computed in registers:
  //         no load stands between two term stores)
  if (FUSED) {
    MosFlags fl;
    fl.dcop = dc; fl.tran = !dc; fl.init_jct = dc && S.newton_iter == 0 && S.cont_first != 0.0;
    fl.tranop = dc;   // every batch runs a transient: its DC point is a transient's operating point
    fl.newton_iter = S.newton_iter;
    const int* lid = ci + C.o_lid;
    const int* tmask = ci + C.o_tmask;
    int lim_any = 0;
    for (int d = tid; d < C.n_mos; d += nth) {
      // one device at a time: without this fence the compiler starts the
      // next device's loads inside the current one and holds both devices'
      // values at once (a hundred more registers)
      asm volatile("" ::: "memory");
      const int g = I.dev_off + d;
      const double* k = kcls + (size_t)A.mcls[g] * KSTRIDE;
      double volt[6], sc[6], sn[6], ac[8];
      for (int j = 0; j < 6; ++j) volt[j] = xs[lid[j * C.n_mos + d]];
      for (int j = 0; j < 6; ++j) { sc[j] = DS.sto_cur[(size_t)j * NT + g]; sn[j] = DS.sto_nxt[(size_t)j * NT + g]; }
      for (int j = 0; j < 8; ++j) ac[j] = DS.sta_cur[(size_t)j * NT + g];
      // class arrays are plane-major ([plane][device]): a wavefront's loads are coalesced
      const unsigned mlo = (unsigned)tmask[d], mhi = (unsigned)tmask[C.n_mos + d];
      // mos1_eval emits the planes in increasing order, so a plane's rank
      // among the device's nonzero planes is the count of stores before it:
      // the term pointer advances by one device column after each store
      double* Tq = T + d;
      const int nm = C.n_mos;
      auto emit = [&](int plane, double v) {
        if (term_on(mlo, mhi, plane)) { *Tq = v; Tq += nm; }
      };
      auto keep = [&](int j, double v) {
        if (j < 6) DS.sto_nxt[(size_t)j * NT + g] = v;
        else DS.sta_nxt[(size_t)(j - 6) * NT + g] = v;
      };
      lim_any |= mos1_eval(k, A.tvbi[g], A.beta[g], A.tvto[g], volt, sc, sn, ac, fl, keep, emit);
    }
    if (lim_any) atomicOr(&sh_int[0], 1);
  }
  {
    // source values at this iteration's time, staged in LDS for the vector
    // gather; the last threads (one evaluation pass at most) compute them
    const double ts = dc ? 0.0 : S.t + S.h;
    const double sscale = dc ? S.src_scale : 1.0;     // source stepping (VSRCSCALE)
    for (int k = nth - 1 - tid; k < C.n_src; k += nth) {
      if (k < 0) break;
      const int w = A.src_wave[I.src_off + k];
      const int wo = A.wave_off[w];
      srcv[k] = wave_value(A.wave_type[w], A.wave_data + wo, A.wave_off[w + 1] - wo, ts, A.tstep, A.tstop);
      if (sscale != 1.0) srcv[k] *= sscale;
    }
  }
  __syncthreads();
  XG_PHASE(1);
  const int limited = sh_int[0];
  // long term groups (e.g. a supply node shared by every PMOS): one wavefront
  // each, lane-strided loads and a fixed shuffle tree (deterministic)
  __shared__ double lsum[kMaxLong];
  __shared__ double chs[kMaxChunks];     // per-64-row chunk sums (workgroup-size independent totals)
  {
    const int nl = C.n_vlong + C.n_mlong;
    const int* lg = ci + C.o_long;
    const unsigned short* vptr = gprog + C.b_vptr;
    const unsigned short* mptr = gprog + C.b_mptr;
    const int* tix = ci + C.o_tix;
    const int wave = tid >> 6, lane = tid & 63, nwaves = nth >> 6;
    for (int j = wave; j < nl; j += nwaves) {
      const int g = lg[j];
      const int beg = j < C.n_vlong ? vptr[g] : mptr[g];
      const int end = j < C.n_vlong ? vptr[g + 1] : mptr[g + 1];
      double sacc = 0.0;
      // four strides at a time: indices, then values, then the sums in order
      for (int e0 = beg + lane; e0 < end; e0 += 256) {
        int ix[4];
        double tv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) ix[q] = e0 + 64 * q < end ? tix[e0 + 64 * q] : 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) tv[q] = e0 + 64 * q < end ? T[ix[q]] : 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (e0 + 64 * q < end) sacc += tv[q];
      }
      for (int o = 32; o > 0; o >>= 1) sacc += __shfl_xor(sacc, o, 64);
      if (lane == 0) lsum[j] = sacc;
    }
    if (nl) __syncthreads();
  }
  const unsigned short* cc = gprog + C.b_cc;
  const unsigned short* vrow = gprog + C.b_vrow;
  // ---- 3. vectors per row (contiguous term ranges) and the residual
  const double h = S.h;
  double n2 = 0.0, ninf = 0.0;
  {
    const unsigned short* vptr = gprog + C.b_vptr;
    const unsigned short* bgp = gprog + C.b_bgp; const unsigned short* bge = gprog + C.b_bge;
    const unsigned short* lrp = gprog + C.b_lrp; const unsigned short* lk = gprog + C.b_lk;
    const double* linF = A.linF + I.nnz_off; const double* linQ = A.linQ + I.nnz_off;
    const int* tix = ci + C.o_tix;
    for (int base = tid & ~63; base < n; base += nth) {
      const int r = base + (tid & 63);
      double res2 = 0.0;
      if (r < n) {
      // loads that do not wait for the term gather go out first
      const int lk0 = lrp[r], lk1 = lrp[r + 1], bg0 = bgp[r], bg1 = bgp[r + 1];
      const double qh0 = A.qH0[off + r];
      const double qh12 = (!dc && A.method == 1) ? A.qH1[off + r] : A.qH2[off + r];
      double acc[4];
      const int gb[5] = {vptr[4 * r], vptr[4 * r + 1], vptr[4 * r + 2], vptr[4 * r + 3], vptr[4 * r + 4]};
      bool lng = false;
      for (int v = 0; v < 4; ++v) lng |= gb[v + 1] - gb[v] >= kLongSeg;
      if (!lng) {
        gather_groups<4>(T, tix, gb, acc);
      } else {
        for (int v = 0; v < 4; ++v) {
          double s = 0.0;
          if (gb[v + 1] - gb[v] >= kLongSeg) {
            const int* lg = ci + C.o_long;
            for (int j = 0; j < C.n_vlong; ++j) if (lg[j] == 4 * r + v) s = lsum[j];
          } else {
            for (int e = gb[v]; e < gb[v + 1]; ++e) s += T[tix[e]];
          }
          acc[v] = s;
        }
      }
      double f = acc[0], q = acc[1];
      const double fl = acc[2], ql = acc[3];
      double sf = 0.0, sq = 0.0;
      for (int k = lk0; k < lk1; ++k) {
        const int sl = lk[k];
        const double xv = xs[cc[sl]];
        sf += linF[sl] * xv;
        sq += linQ[sl] * xv;
      }
      f += sf; q += sq;
      double b = 0.0;
      for (int e = bg0; e < bg1; ++e) {
        const int ent = bge[e];
        const double v = srcv[ent >> 1];
        if (ent & 1) b -= v; else b += v;
      }
      A.F[off + r] = f; A.Q[off + r] = q; A.B[off + r] = b;
      const double res = (!dc && A.method == 1)
                             ? residual_gear(S.order, h, S.a0, S.a1, S.a2, f, q, b, fl, ql, qh0, qh12)
                             : residual_combine(dc, S.order, h, f, q, b, fl, ql, qh0, qh12);
      rhs[r] = (dc && S.gcond != 0.0 && vrow[r]) ? res - S.gcond * xs[r] : res;
      res2 = res * res;
      ninf = fmax(ninf, fabs(res));
      }
      const double cs = wave_tree_sum(res2);
      if ((tid & 63) == 0) chs[base >> 6] = cs;
    }
  }
  XG_PHASE(2);
  n2 = chunk_total(chs, n);
  ninf = block_max(ninf, red);
  XG_PHASE(3);
  // ---- 4. Newton bookkeeping (DampedNewton::solve / converged_)
  if (tid == 0) {
    S.n_loads++;
    const double normRHS = sqrt(n2);
    int status = 0;
    if (S.newton_iter == 0) {
      S.norm_init = normRHS; S.norm_old = normRHS; S.stall_count = 0; S.tmp_rate = 1.0;
      status = isfinite(normRHS) ? 0 : -6;
    } else {
      const int maxstep = dc ? (S.dc_stage != 0.0 ? kContMaxNewton : kDcMaxNewton) : kMaxNewton;
      const double rhstol = dc ? kDcRhsTol : kRhsTol, dxtol = dc ? kDcDxTol : kDxTol;
      const int nl = S.newton_iter;
      if (S.last_status < 0) status = S.last_status;          // linear failure / NaN from the solve
      else if (limited && nl < maxstep) status = 0;
      else if (limited) status = -1;
      else if (!isfinite(normRHS)) status = -6;
      else if (normRHS < 2.220446049250313e-16) status = 1;
      else {
        const double rel = normRHS / S.norm_init, rate = normRHS / S.norm_old;
        S.norm_old = normRHS;
        const double upd = S.wt_dx;
        if (ninf <= rhstol && upd <= dxtol) status = 2;
        else if (!dc && nl >= maxstep && rel <= 0.9 && rate <= 1.0) status = -3;
        else if (upd <= 1e-6) status = 4;
        else if (nl >= maxstep) status = -1;
        else if (rate > 0.5 * 1.7976931348623157e308) status = -2;
        else if (!dc) {
          if (fabs(rate - 1.0) <= 1.0e-3) {
            if (S.stall_count == 0 || rate < S.tmp_rate) S.tmp_rate = rate;
            S.stall_count++;
          } else S.stall_count = 0;
          if (S.stall_count == 5) status = -3;
        }
      }
    }
    sh_int[1] = status;
  }
  __syncthreads();
  const int status = sh_int[1];
  XG_PHASE(4);

  if (status == 0) {
    // ---- 5a. Jacobian, LU refactor, solve, update
    const int nlu = C.nlu;
    if (S.newton_iter == 0 && !dc) {
      // transient weights from the predictor and the last accepted x (updateWeights_)
      for (int i = tid; i < n; i += nth)
        A.wt[off + i] = kNlRel * fmax(fabs(A.xH0[off + i]), fabs(xs[i])) + kNlAbs;
    }
    const double ja = dc ? 1.0e-20 : (A.method == 1 ? S.a0 / h : 1.0 / h);
    const double jb = dc ? 1.0 : (A.method == 1 ? 1.0 : (S.order == 2 ? 0.5 : 1.0));
    bool use_dense = dc;
    if (!dc) {
      for (int p = tid; p < nlu; p += nth) lu[p] = 0.0;
      // the forward solve runs inside the refactor: its n entries after the
      // factors start from the permuted right-hand side
      {
        const unsigned short* prow = gprog + C.b_prow;
        for (int i = tid; i < n; i += nth) lu[nlu + i] = rhs[prow[i]];
      }
      __syncthreads();
      {
        const unsigned short* mptr = gprog + C.b_mptr;
        const unsigned short* a2lu = gprog + C.b_a2lu;
        const double* linF = A.linF + I.nnz_off; const double* linQ = A.linQ + I.nnz_off;
        const int* lg = ci + C.o_long;
        const int* tix = ci + C.o_tix;
        for (int s = tid; s < C.nnz; s += nth) {
          // loads that do not wait for the term gather go out first
          const double lf = linF[s], lq = linQ[s];
          const int dst = a2lu[s];
          double part[2];
          const int b[3] = {mptr[2 * s], mptr[2 * s + 1], mptr[2 * s + 2]};
          if (b[1] - b[0] < kLongSeg && b[2] - b[1] < kLongSeg) {
            gather_groups<2>(T, tix, b, part);
          } else {
            for (int fq = 0; fq < 2; ++fq) {
              const int g = 2 * s + fq;
              double acc = 0.0;
              if (b[fq + 1] - b[fq] >= kLongSeg) {
                for (int j = 0; j < C.n_mlong; ++j) if (lg[C.n_vlong + j] == g) acc = lsum[C.n_vlong + j];
              } else {
                for (int e = b[fq]; e < b[fq + 1]; ++e) acc += T[tix[e]];
              }
              part[fq] = acc;
            }
          }
          double df = part[0], dq = part[1];
          if (lf != 0.0) df += lf;
          if (lq != 0.0) dq += lq;
          lu[dst] = jac_combine(ja, dq, jb, df);
        }
      }
      __syncthreads();
      XG_PHASE(5);
      {
        // level-scheduled refactor; KLU's threshold test (|l_ij| <= 1/0.001,
        // finite nonzero pivots) is evaluated as the entries are produced
        const unsigned short* flp = prog + C.b_flp; const unsigned short* fls = prog + C.b_fls;
        const uint2* fmeta = reinterpret_cast<const uint2*>(prog + C.b_fmeta);
        const unsigned* fops = reinterpret_cast<const unsigned*>(prog + C.b_fops);
        const int wave = tid >> 6, lane = tid & 63, nwaves = nth >> 6;
        int bad = 0;
        for (int l = 0; l < C.nflev; ++l) {
          const int t0 = flp[l], ts = fls[l], t1 = flp[l + 1];
          for (int t = t0 + tid; t < ts; t += nth) {
            const uint2 m = fmeta[t];
            const int p = m.x & 0xFFFF, dv = m.x >> 16, o0 = m.y & 0xFFFF, no = m.y >> 16;
            double v = lu[p];
            const double d = dv != 0xFFFF ? lu[dv] : 1.0;
            for (int o = 0; o < no; o += 4) {
              unsigned pr[4];
              double fa[4], fb[4];
#pragma unroll
              for (int q = 0; q < 4; ++q) pr[q] = o + q < no ? fops[o0 + o + q] : 0u;
#pragma unroll
              for (int q = 0; q < 4; ++q) { fa[q] = lu[pr[q] & 0xFFFF]; fb[q] = lu[pr[q] >> 16]; }
#pragma unroll
              for (int q = 0; q < 4; ++q)
                if (o + q < no) v -= fa[q] * fb[q];
            }
            if (dv != 0xFFFF) {
              v /= d;
              if (!(fabs(v) <= 1.0e3) || !(fabs(d) > 0.0) || !isfinite(d)) bad = 1;
            }
            lu[p] = v;
          }
          for (int j = ts + wave; j < t1; j += nwaves) {
            const uint2 m = fmeta[j];
            const int p = m.x & 0xFFFF, dv = m.x >> 16, o0 = m.y & 0xFFFF, no = m.y >> 16;
            double acc = 0.0;
            // four strides at a time: index pairs, then operands, then the sums in order
            for (int o = lane; o < no; o += 256) {
              unsigned pr[4];
              double fa[4], fb[4];
#pragma unroll
              for (int q = 0; q < 4; ++q) pr[q] = o + 64 * q < no ? fops[o0 + o + 64 * q] : 0u;
#pragma unroll
              for (int q = 0; q < 4; ++q) { fa[q] = lu[pr[q] & 0xFFFF]; fb[q] = lu[pr[q] >> 16]; }
#pragma unroll
              for (int q = 0; q < 4; ++q)
                if (o + 64 * q < no) acc += fa[q] * fb[q];
            }
            for (int sh = 32; sh > 0; sh >>= 1) acc += __shfl_xor(acc, sh, 64);
            if (lane == 0) {
              double v = lu[p] - acc;
              if (dv != 0xFFFF) {
                const double d = lu[dv];
                v /= d;
                if (!(fabs(v) <= 1.0e3) || !(fabs(d) > 0.0) || !isfinite(d)) bad = 1;
              }
              lu[p] = v;
            }
          }
          __syncthreads();
        }
        XG_PHASE(6);
        // the last pivot is never a divisor: check every diagonal
        const unsigned short* dg = prog + C.b_diag;
        for (int i = tid; i < n; i += nth) {
          const double d = lu[dg[i]];
          if (!(fabs(d) > 0.0) || !isfinite(d)) bad = 1;
        }
        if (bad) atomicOr(&sh_int[6], 1);
        __syncthreads();
        use_dense = sh_int[6] != 0;
      }
      XG_PHASE(7);
      if (!use_dense) {
        // L y = P b ran with the refactor (y after the factors); U x = y here
        const unsigned short* pcol = gprog + C.b_pcol;
        const unsigned short* lurp = prog + C.b_lurp; const unsigned short* luci = prog + C.b_luci;
        const unsigned short* dg = prog + C.b_diag;
        const unsigned short* bsp = prog + C.b_bsp; const unsigned short* bsr = prog + C.b_bsr;
        const unsigned short* bss = prog + C.b_bss;
        const int wave = tid >> 6, lane = tid & 63, nwaves = nth >> 6;
        double* y = lu + nlu;
        for (int l = 0; l < C.nbs; ++l) {
          for (int t = bsp[l] + tid; t < bss[l]; t += nth) {
            const int i = bsr[t];
            double v = y[i];
            for (int p = dg[i] + 1; p < lurp[i + 1]; ++p) v -= lu[p] * y[luci[p]];
            y[i] = v / lu[dg[i]];
          }
          for (int j = bss[l] + wave; j < bsp[l + 1]; j += nwaves) {
            const int i = bsr[j];
            double acc = 0.0;
            for (int p = dg[i] + 1 + lane; p < lurp[i + 1]; p += 64) acc += lu[p] * y[luci[p]];
            for (int sh = 32; sh > 0; sh >>= 1) acc += __shfl_xor(acc, sh, 64);
            if (lane == 0) y[i] = (y[i] - acc) / lu[dg[i]];
          }
          __syncthreads();
        }
        for (int j = tid; j < n; j += nth) rhs[pcol[j]] = y[j];
        __syncthreads();
      }
      XG_PHASE(8);
    }
    if (use_dense) {
      if (tid == 0) S.n_dense++;
      const int sing = dense_pivot_solve(A, C, I, ci, prog, T, ja, jb, dc ? S.gcond : 0.0, rhs, red, red_i);
      if (sing && tid == 0) sh_int[2] |= 1;
      __syncthreads();
    }
    // x += dx, weighted update norm; a failed pivot gives a zero update and
    // a linear-solver failure (AmesosSolver: zero solution + status)
    const bool pivot_bad = sh_int[2] & 1;
    double wdx = 0.0;
    int nanf = 0;
    for (int i = tid; i < n; i += nth) {
      const double d = pivot_bad ? 0.0 : rhs[i];
      const double xn = xs[i] + d;
      if (!isfinite(d)) nanf = 1;
      A.x[off + i] = xn;
      const double w = dc ? kDcRel * fmax(0.0, fabs(xn)) + kDcAbs : A.wt[off + i];
      wdx = fmax(wdx, fabs(d) / w);
    }
    if (nanf) atomicOr(&sh_int[2], 2);
    wdx = block_max(wdx, red);
    XG_PHASE(9);
    if (tid == 0) {
      S.wt_dx = wdx;
      S.newton_iter++;
      S.n_iters++;
      S.last_status = sh_int[2] & 1 ? -9 : (sh_int[2] & 2 ? -6 : 0);
    }
    return;
  }

  // ---- 5b. Newton finished for this step
  if (dc) {
    // DC strategy (Interface::spiceStrategy, N_NLS_NOX_Interface.C:321-437):
    // Newton from initJct; on failure gmin stepping (10^p - 10^-4 S on every
    // node-voltage row, p from 4 to -4, step -2); on its failure source
    // stepping (sources scaled from 0 to 1, step 0.2, no initJct); natural
    // continuation with a constant predictor: a failed step is retried at half
    // the step from the last converged point, a converged one grows the step
    // by (1 + a (1 - (iters/20)^2)), a = 0.01 (gmin) or 1 (source).
    if (tid == 0) {
      const int act = dc_strategy_step(S, status);
      sh_int[7] = act;
    }
    __syncthreads();
    const int act = sh_int[7];
    if (act == 2) return;
    if (act == 1) {
      // converged continuation step: it becomes the restart point (xH1 is free during DC)
      for (int i = tid; i < n; i += nth) A.xH1[off + i] = xs[i];
      if (tid == 0) S.sbuf ^= 1;   // the iterate's store/state become the accepted ones
      return;
    }
    if (act == 3) {
      // a strategy starts over from zero with fresh device history
      for (int i = tid; i < n; i += nth) { A.x[off + i] = 0.0; A.xH1[off + i] = 0.0; }
      for (int d = tid; d < C.n_mos; d += nth) {
        const int g = I.dev_off + d;
        for (int j = 0; j < 6; ++j) { A.sto_cur[(size_t)j * NT + g] = 0.0; A.sto_nxt[(size_t)j * NT + g] = 0.0; }
        for (int j = 0; j < 8; ++j) { A.sta_cur[(size_t)j * NT + g] = 0.0; A.sta_nxt[(size_t)j * NT + g] = 0.0; }
      }
      return;
    }
    if (act == 4) {
      for (int i = tid; i < n; i += nth) A.x[off + i] = A.xH1[off + i];
      for (int d = tid; d < C.n_mos; d += nth) {
        const int g = I.dev_off + d;
        for (int j = 0; j < 6; ++j) DS.sto_nxt[(size_t)j * NT + g] = DS.sto_cur[(size_t)j * NT + g];
        for (int j = 0; j < 8; ++j) DS.sta_nxt[(size_t)j * NT + g] = DS.sta_cur[(size_t)j * NT + g];
      }
      return;
    }
    // tranop -> transient: OneStep::initialize (N_TIA_OneStep.C:1684); the
    // operating point's store/state become the accepted ones
    if (tid == 0) S.sbuf ^= 1;
    for (int i = tid; i < n; i += nth) {
      A.xH0[off + i] = xs[i];
      A.qH0[off + i] = A.Q[off + i];
      A.xH1[off + i] = 0.0;
      A.xH2[off + i] = 0.0;
      A.qH2[off + i] = 0.0;
      if (A.pred_mode != 0) A.xP[off + i] = xs[i];   // the operating point starts the history
    }
    if (tid == 0 && A.pred_mode != 0) S.n_acc = 1;
    if (tid == 0) {
      S.mode = 1;
      S.t = 0.0;
      const double stop = next_bp(A, I, S);
      double hh = 0.005 * fabs(stop - S.t);
      if (A.tstep > 0.0) hh = fmin(A.tstep, hh);
      if (A.hmax > 0 && hh > A.hmax) hh = A.hmax;
      S.h = fmax(hh, kHmin);
      S.last_h = S.h; S.h_sum = 2.0 * S.h;
      S.order = 1; S.psi0 = S.h; S.psi1 = 0; S.psi2 = 0;
      S.beginning = 1;
      S.steps_taken = 0;
      sh_d[0] = S.h;
    }
    __syncthreads();
    if (A.method == 1) {
      // Gear12::initialize: both history slots hold the operating point
      for (int i = tid; i < n; i += nth) { A.xH1[off + i] = xs[i]; A.qH1[off + i] = A.Q[off + i]; }
    } else {
      for (int i = tid; i < n; i += nth) A.qH1[off + i] = -(A.F[off + i] - A.B[off + i]) * sh_d[0];
    }
    if (A.rec_max > 0 && tid == 0 && S.n_rec < A.rec_max) A.rec_t[(size_t)I.rec_off + S.n_rec] = 0.0;
    if (A.rec_max > 0 && S.n_rec < A.rec_max)
      for (int i = tid; i < n; i += nth) A.rec_x[((size_t)I.rec_off + S.n_rec) * A.max_n + i] = xs[i];
    __syncthreads();
    if (tid == 0 && A.rec_max > 0) S.n_rec++;
    __syncthreads();
    begin_step(A, S, C, I, red);
    XG_PHASE(10);
    return;
  }

  // transient: local error test (StepErrorControl::evaluateStepError)
  bool ok = status > 0;
  if (ok && !S.beginning) {
    for (int base = tid & ~63; base < n; base += nth) {
      const int i = base + (tid & 63);
      double q2 = 0.0;
      if (i < n) {
        const double q = (xs[i] - A.xn0[off + i]) / S.errwt;
        q2 = q * q;
      }
      const double cs = wave_tree_sum(q2);
      if ((tid & 63) == 0) chs[base >> 6] = cs;
    }
    const double s2 = chunk_total(chs, n);
    if (tid == 0) { S.est = S.ck * sqrt(s2 / n); sh_int[3] = S.est <= 1.0; }
    __syncthreads();
    ok = sh_int[3];
  }
  if (ok) {
    // completeStep (N_TIA_OneStep.C:2124) + updateHistory (:1390)
    if (tid == 0) {
      S.n_steps++; S.nef = 0;
      S.last_h = S.h;
      double newh = S.h;
      double rr = pow(0.5 / (S.est + 0.0001), 1.0 / (S.order + 1.0));
      S.steps_taken += 1;
      if (S.steps_taken >= 2 && S.order == 1) {
        S.order = 2;
        rr = pow(0.5 / (S.est + 0.0001), 1.0 / (S.order + 1.0));
        if (rr <= 1.05) S.order = 1;
      }
      if (rr >= 2.0) newh = 2.0 * S.h;
      else if (rr <= 1) newh = fmax(0.25, fmin(0.9, rr)) * S.h;
      sh_d[1] = newh;
      sh_int[4] = S.order;
    }
    __syncthreads();
    const int new_order = sh_int[4];
    if (A.method == 1) {
      // Gear12::updateHistory (N_TIA_Gear12.C:822-857): shift the values
      for (int i = tid; i < n; i += nth) {
        if (new_order == 2) A.xH2[off + i] = A.xH1[off + i];
        A.qH1[off + i] = A.qH0[off + i];
        A.xH1[off + i] = A.xH0[off + i];
        A.xH0[off + i] = xs[i];
        A.qH0[off + i] = A.Q[off + i];
      }
    } else {
      for (int i = tid; i < n; i += nth) {
        if (new_order == 2) {
          A.xH2[off + i] = A.xH1[off + i];
          A.qH2[off + i] = 1.0 * A.F[off + i] + -1.0 * A.B[off + i];
        }
        const double xv = xs[i], qv = A.Q[off + i];
        A.xH1[off + i] = 1.0 * xv + -1.0 * A.xH0[off + i];
        A.qH1[off + i] = 1.0 * qv + -1.0 * A.qH0[off + i];
        A.xH0[off + i] = xv;
        A.qH0[off + i] = qv;
      }
    }
    if (tid == 0) S.sbuf ^= 1;   // the iterate's store/state become the accepted ones
    if (A.pred_mode != 0) pred_accept(A, I, C, S, xs);
    if (tid == 0) S.t = S.t + S.h;
    __syncthreads();
    if (A.rec_max > 0 && S.n_rec < A.rec_max) {
      if (tid == 0) A.rec_t[(size_t)I.rec_off + S.n_rec] = S.t;
      for (int i = tid; i < n; i += nth) A.rec_x[((size_t)I.rec_off + S.n_rec) * A.max_n + i] = xs[i];
    }
    __syncthreads();
    if (tid == 0) {
      if (A.rec_max > 0 && S.n_rec < A.rec_max) S.n_rec++;
      double newh = fmax(sh_d[1], kHmin);
      if (A.hmax > 0) newh = fmin(newh, A.hmax);
      S.beginning = 0;
      sh_int[5] = 0;
      if (S.t >= A.tstop * (1 - 1e-12)) {
        S.mode = 2;
      } else if (S.bpi < I.n_bp && fabs(S.t - A.bp[I.bp_off + S.bpi]) <= 1e-13 * fmax(1.0, fabs(S.t)) + 1e-24) {
        // breakpoint: restart at order 1 (OneStep::initialize, not at t = 0)
        S.saved_h = newh;
        ++S.bpi;
        const double stop = next_bp(A, I, S);
        double hh = 0.1 * fmin(S.saved_h, fabs(stop - S.t));
        if (A.hmax > 0 && hh > A.hmax) hh = A.hmax;
        S.h = fmax(hh, kHmin);
        S.last_h = S.h; S.h_sum = 2.0 * S.h;
        S.order = 1; S.psi0 = S.h; S.steps_taken = 0; S.beginning = 1;
        sh_int[5] = 1;
        sh_d[2] = S.h;
      } else {
        S.h_sum = newh + S.last_h;
        S.h = newh;
      }
    }
    __syncthreads();
    if (S.mode == 2) { return; }
    if (sh_int[5]) {
      for (int i = tid; i < n; i += nth) {
        if (A.method == 1) {
          A.xH1[off + i] = A.xH0[off + i];
          A.qH1[off + i] = A.qH0[off + i];
        } else {
          A.xH1[off + i] = 0.0;
          A.qH1[off + i] = -(A.F[off + i] - A.B[off + i]) * sh_d[2];
        }
      }
      __syncthreads();
    }
    begin_step(A, S, C, I, red);
    XG_PHASE(10);
    return;
  }
  // rejectStep (N_TIA_OneStep.C:1946)
  if (tid == 0) {
    S.n_rejects++;
    S.nef++;
    if (status <= 0 && S.used_learned) {
      // Newton failed from the network's proposal: retry from the polynomial one
      S.pred_skip = 1;
      S.n_learned_rejects++;
    }
    if (A.pred_mode == 2 && S.feat_valid) pred_score(S, S.used_learned != 0, S.newton_iter);
    S.used_learned = 0;
    const int ord = S.order;
    if (ord >= 1) S.psi0 = S.psi1;
    if (ord >= 2) S.psi1 = S.psi2;
    double newh;
    if (status <= 0) { newh = S.h / 8; S.order = 1; }
    else if (S.nef == 1) {
      // OneStep restarts at order 1; Gear12 keeps its order (N_TIA_Gear12.C:1502-1516)
      if (A.method != 1) S.order = 1;
      const double rr = fmax(0.25, fmin(0.9, pow(0.5 / (S.est + 0.0001), 1.0 / (S.order + 1.0))));
      newh = rr * S.h;
    } else { newh = 0.25 * S.h; S.order = 1; }
    newh = fmax(newh, kHmin);
    if (A.hmax > 0) newh = fmin(newh, A.hmax);
    if (S.nef >= 15 || newh <= kHmin) S.mode = -2;
    S.h_sum = newh + S.last_h;
    S.h = newh;
  }
  __syncthreads();
  if (S.mode < 0) { return; }
  begin_step(A, S, C, I, red);
  XG_PHASE(10);
#undef XG_PHASE
}

// KLDS: the MOSFET class table is staged in LDS (at most 16 entries); TLDS:
// the term array lives in LDS.  Both are template constants so that every
// access compiles to its own memory's instructions (LDS reads wait only on
// LDS, not on the term stores in flight).  Registers are bounded for 3
// wavefronts per SIMD (168 VGPRs, about 34 spilled): each instance's iteration
// is a chain of dependent LDS and HBM round trips, and a third wavefront hides
// more of them than the spills cost (the 4096-ring batch runs about 5 %
// faster than at 2 wavefronts per SIMD with 202 VGPRs and no spills).
template <bool FUSED, bool KLDS, bool TLDS>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3))) k_batch_step(BatchArgs A) {
  if ((int)blockIdx.x >= A.n_inst) return;
  const int inst = A.order[blockIdx.x];
  const InstDesc I = A.inst[inst];
  const ClassDesc C = A.cls[I.cls];
  __shared__ InstState S;
  __shared__ double red[16];
  __shared__ int red_i[16];
  __shared__ int sh_int[8];
  __shared__ double sh_d[8];
  __shared__ double kc_s[KLDS ? kMaxKclsLds * KSTRIDE : 1];
  __shared__ double srcv[kMaxSrc];
  extern __shared__ double dyn[];
  double* xs = dyn;                         // n + 1
  double* rhs = xs + (A.lds_n + 1);         // n
  double* lu = rhs + (A.lds_n + 1);         // nlu factor entries, then the n forward-solve entries
  // class program after the doubles, on a 16-byte boundary (uint4 copies)
  unsigned short* prog = reinterpret_cast<unsigned short*>(dyn + ((2 * (A.lds_n + 1) + A.lds_nlu + 1) & ~1));
  const int tid = threadIdx.x, nth = blockDim.x;
  if (tid == 0) S = A.st[inst];
  __syncthreads();
  if (S.mode >= 2 || S.mode < 0) return;
  // stage the class program (uint16, 16-byte chunks) and the MOSFET class table into LDS
  {
    const uint4* src = reinterpret_cast<const uint4*>(A.blob + C.blob_off);
    uint4* dst = reinterpret_cast<uint4*>(prog);
    const int chunks = C.blob_lds / 8;
    for (int i = tid; i < chunks; i += nth) dst[i] = src[i];
  }
  const double* kcls = A.kcls;
  if (KLDS) {
    for (int i = tid; i < A.n_kcls * KSTRIDE; i += nth) kc_s[i] = A.kcls[i];
    kcls = kc_s;
  }
  __syncthreads();
  // fused form with the term array in LDS: after the program (8-byte
  // aligned); it is written and read within each Newton iteration
  double* T = TLDS ? reinterpret_cast<double*>(prog + ((A.lds_blob + 3) & ~3)) : A.terms + I.term_off;
  const int iters = FUSED ? A.iters : 1;
  for (int it = 0; it < iters; ++it) {
    batch_iteration<FUSED>(A, inst, I, C, S, red, red_i, sh_int, sh_d, xs, rhs, lu, prog, kcls, srcv, T);
    __syncthreads();
    if (S.mode >= 2 || S.mode < 0) break;
  }
  if (tid == 0) A.st[inst] = S;
}

// Device evaluation of every MOSFET of every instance that is iterating
// (DC or transient), one thread per device over the structure-of-arrays
// state: MOSFET1 updateIntermediateVars + the vector and matrix terms,
// written to the device's column of the instance's term array.  Runs before the
// per-instance kernel of the same launch.
__global__ void __launch_bounds__(256, 4) k_batch_eval(BatchArgs A) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.nt_mos) return;
  const int inst = A.dev_inst[g];
  const int mode = A.st[inst].mode;
  if (mode != 0 && mode != 1) return;
  const int newton_iter = A.st[inst].newton_iter;
  const bool first_solve = A.st[inst].cont_first != 0.0;
  const InstDesc& I = A.inst[inst];
  const ClassDesc& C = A.cls[I.cls];
  const int d = g - I.dev_off;
  const int n = C.n, off = I.vec_off;
  const int NT = A.nt_mos;
  const bool dc = mode == 0;
  MosFlags fl;
  fl.dcop = dc; fl.tran = !dc; fl.init_jct = dc && newton_iter == 0 && first_solve; fl.newton_iter = newton_iter;
  fl.tranop = dc;
  const int* lid = A.ci + C.o_lid;
  const int* tmask = A.ci + C.o_tmask;
  double* T = A.terms + I.term_off;
  const double* k = A.kcls + (size_t)A.mcls[g] * KSTRIDE;
  double volt[6], sc[6], sn[6], ac[8];
  for (int j = 0; j < 6; ++j) {
    const int l = lid[j * C.n_mos + d];
    volt[j] = l < n ? A.x[off + l] : 0.0;
  }
  const DevState DS = dev_state(A, A.st[inst].sbuf);
  for (int j = 0; j < 6; ++j) { sc[j] = DS.sto_cur[(size_t)j * NT + g]; sn[j] = DS.sto_nxt[(size_t)j * NT + g]; }
  for (int j = 0; j < 8; ++j) ac[j] = DS.sta_cur[(size_t)j * NT + g];
  const unsigned mlo = (unsigned)tmask[d], mhi = (unsigned)tmask[C.n_mos + d];
  double* Td = T + d;
  const int nm = C.n_mos;
  auto emit = [&](int plane, double v) {
    if (term_on(mlo, mhi, plane)) Td[(size_t)term_rank(mlo, mhi, plane) * nm] = v;
  };
  auto keep = [&](int j, double v) {
    if (j < 6) DS.sto_nxt[(size_t)j * NT + g] = v;
    else DS.sta_nxt[(size_t)(j - 6) * NT + g] = v;
  };
  const int lim = mos1_eval(k, A.tvbi[g], A.beta[g], A.tvto[g], volt, sc, sn, ac, fl, keep, emit);
  if (lim) atomicOr(&A.lim[inst], 1);
}

// Resumes the instances waiting for the network's proposal (mode 3):
// x = x_base + w y, y = sign(g)(2^|g| - 1) from the network's output g (w the
// Newton weight of each unknown; y taken as 0 where |y| <= pred_dead) when
// every |g| <= kPredTrustLog and finite, else x = x_poly (the polynomial
// predictor is the fallback).
__global__ void __launch_bounds__(256) k_pred_apply(BatchArgs A) {
  const int inst = blockIdx.x;
  if (inst >= A.n_inst) return;
  __shared__ int mode, bad;
  if (threadIdx.x == 0) { mode = A.st[inst].mode; bad = 0; }
  __syncthreads();
  if (mode != 3) return;
  const InstDesc I = A.inst[inst];
  const ClassDesc C = A.cls[I.cls];
  const int n = C.n, off = I.vec_off;
  int b = 0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const float y = A.pred_out[(size_t)(I.win_off + r / XG_PRED_WIN) * XG_PRED_WIN + r % XG_PRED_WIN];
    if (!(fabsf(y) <= kPredTrustLog)) b = 1;
  }
  if (b) atomicOr(&bad, 1);
  __syncthreads();
  const bool use = bad == 0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const int w = r / XG_PRED_WIN;
    const double y = pred_sexp(A.pred_out[(size_t)(I.win_off + w) * XG_PRED_WIN + r % XG_PRED_WIN]);
    const double xp = A.xn0[off + r];
    const double yc = fabs(y) > A.pred_dead ? y : 0.0;
    A.x[off + r] = use ? A.xP[4 * A.total_n + off + r] + pred_weight(A.xP[off + r], xp) * yc : xp;
  }
  if (threadIdx.x == 0) {
    InstState& S = A.st[inst];
    S.mode = 1;
    S.used_learned = use ? 1 : 0;
    if (use) S.n_learned++;
  }
}

// Gathers collected training-pair records (src[i] -> row i).
__global__ void k_pair_gather(long long n, const long long* src, const unsigned short* px, const float* pt,
                              const float* pm, unsigned short* ox, float* ot, float* om) {
  const long long i = blockIdx.x;
  if (i >= n) return;
  const long long s = src[i];
  for (int j = threadIdx.x; j < XG_PRED_FEAT; j += blockDim.x) ox[i * XG_PRED_FEAT + j] = px[s * XG_PRED_FEAT + j];
  for (int j = threadIdx.x; j < XG_PRED_WIN; j += blockDim.x) {
    ot[i * XG_PRED_WIN + j] = pt[s * XG_PRED_WIN + j];
    om[i * XG_PRED_WIN + j] = pm[s * XG_PRED_WIN + j];
  }
}

__global__ void k_init_state(int n_inst, InstState* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst) return;
  InstState s;
  memset(&s, 0, sizeof(s));
  s.mode = 0; s.order = 1; s.h = 1e-12;
  s.src_scale = 1.0; s.cont_first = 1.0;
  st[i] = s;
}

// DC strategy starting at a later stage for every instance (the
// counterpart of .OPTIONS NONLIN CONTINUATION=GMIN / SOURCE)
__global__ void k_dc_first_stage(int n_inst, InstState* st, int stage) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst) return;
  InstState& S = st[i];
  S.dc_stage = stage;
  S.cont_k = 0.0;
  if (stage == 1) {
    S.cont_par = 4.0; S.cont_ok = 4.0; S.cont_step = -2.0; S.cont_first = 1.0;
    S.gcond = pow(10.0, 4.0) - pow(10.0, -4.0); S.src_scale = 1.0;
  } else if (stage == 2) {
    S.cont_par = 0.0; S.cont_ok = 0.0; S.cont_step = 0.2; S.cont_first = 0.0;
    S.gcond = 0.0; S.src_scale = 0.0;
  } else {
    S.cont_first = 1.0; S.gcond = 0.0; S.src_scale = 1.0;
  }
}

}  // namespace xg

using namespace xg;

template <class T>
struct DArr {
  T* p = nullptr;
  size_t n = 0;
  ~DArr() { if (p) (void)hipFree(p); }
  hipError_t up(const std::vector<T>& v) {
    n = v.size();
    if (!n) return hipSuccess;
    hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpy(p, v.data(), n * sizeof(T), hipMemcpyHostToDevice);
  }
  hipError_t zeros(size_t cnt) {
    n = cnt;
    if (!n) return hipSuccess;
    hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemset(p, 0, n * sizeof(T));
  }
};

constexpr size_t kMaxLdsBytes = 150 * 1024;   // dynamic LDS per workgroup this engine uses at most

struct xg_batch {
  int n_inst = 0;
  bool started = false;           // a launch has run (the integration method is fixed from then on)
  std::vector<ClassDesc> classes;
  std::vector<InstDesc> insts;
  std::vector<int> cls_n;         // per instance n
  DArr<int> d_order, d_ci, d_mcls, d_src_wave, d_wave_type, d_wave_off;
  DArr<InstDesc> d_inst;
  DArr<ClassDesc> d_cls;
  DArr<InstState> d_st;
  DArr<double> d_kcls, d_tvbi, d_beta, d_tvto, d_sto_cur, d_sto_nxt, d_sta_cur, d_sta_nxt, d_terms;
  DArr<unsigned short> d_blob;
  DArr<unsigned long long> d_mask;
  DArr<double> d_x, d_xn0, d_xH0, d_xH1, d_xH2, d_qH0, d_qH1, d_qH2, d_F, d_Q, d_B, d_F0, d_B0, d_wt;
  DArr<double> d_linF, d_linQ, d_wave_data, d_bp, d_rec_t, d_rec_x, d_dense;
  DArr<unsigned long long> d_prof;
  // learned predictor
  int pred_mode = 0, pred_rows = 0, total_windows = 0, pair_cap = 0;
  long long total_n = 0, total_records = 0;
  xg_mlp* net = nullptr;
  DArr<double> d_xP;
  DArr<unsigned short> d_feat, d_pair_x;
  DArr<float> d_pred_out, d_pair_t, d_pair_m;
  BatchArgs args{};
  hipStream_t stream = nullptr;
  // the LDS groups run side by side on streams of their own (default
  // priority), forked from and joined back into `stream` at every launch
  hipStream_t gstream[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t fork_ev = nullptr, join_ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool concurrent_groups = false;   // chosen by the autotuner
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int block = 256;
  int only_bucket = -1;           // autotuner: launch one bucket only
  int split = 1;                  // 1: separate device-evaluation kernel + per-instance kernel
  size_t shmem = 0;
  // launch buckets: instances grouped by the LDS their class needs, one
  // per-instance launch each with its own dynamic LDS size
  // block 0: the batch's; terms: fused form keeps the term arrays in LDS
  // (shmem_terms bytes, chosen by the autotuner where it fits)
  struct Bucket { int start, count, n, nlu; size_t shmem; int block; int blob = 0, nterm = 0; bool terms = false;
                  size_t shmem_terms = 0; };
  std::vector<Bucket> buckets;
  std::vector<size_t> cls_lds;    // per class: dynamic LDS bytes
  std::vector<int> inst_cls;
  DArr<int> d_dev_inst, d_lim;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_len = 0;
  long long launches = 0;
  double last_seconds = 0;
  double tune_us = 0;
  int order_mode = 1;
  ~xg_batch() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
    for (int j = 0; j < 4; ++j) {
      if (gstream[j]) (void)hipStreamDestroy(gstream[j]);
      if (join_ev[j]) (void)hipEventDestroy(join_ev[j]);
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
  }
  static size_t lds_bytes(int n, int nlu, int blob) {
    return sizeof(double) * (((size_t)2 * (n + 1) + nlu + 1) & ~(size_t)1) + sizeof(unsigned short) * (size_t)blob;
  }
  // mode 0: given order, one launch sized for the largest class; mode 1:
  // instances by LDS need (largest first), up to 4 launches of equal size,
  // each sized for its own largest class
  void set_order(int mode) {
    std::vector<int> ord(n_inst);
    std::iota(ord.begin(), ord.end(), 0);
    buckets.clear();
    if (mode == 1) {
      std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        return cls_lds[inst_cls[a]] > cls_lds[inst_cls[b]];
      });
      const int nb = std::min(4, n_inst);
      for (int k = 0; k < nb; ++k) {
        const int s0 = (int)((long long)n_inst * k / nb), s1 = (int)((long long)n_inst * (k + 1) / nb);
        if (s1 <= s0) continue;
        Bucket bk{s0, s1 - s0, 0, 0, 0, 0};
        int blob = 0, nterm = 0;
        for (int j = s0; j < s1; ++j) {
          const ClassDesc& C = classes[inst_cls[ord[j]]];
          bk.n = std::max(bk.n, C.n); bk.nlu = std::max(bk.nlu, C.nlu + C.n); blob = std::max(blob, C.blob_lds);
          nterm = std::max(nterm, C.tsize);
        }
        bk.shmem = lds_bytes(bk.n, bk.nlu, blob);
        bk.blob = blob; bk.nterm = nterm;
        bk.shmem_terms = lds_bytes(bk.n, bk.nlu, ((blob + 3) & ~3)) + sizeof(double) * (size_t)nterm;
        buckets.push_back(bk);
      }
    } else {
      Bucket bk{0, n_inst, args.max_n, args.max_nlu, shmem, 0};
      for (const ClassDesc& C : classes) { bk.blob = std::max(bk.blob, C.blob_lds); bk.nterm = std::max(bk.nterm, C.tsize); }
      bk.shmem_terms = lds_bytes(bk.n, bk.nlu, ((bk.blob + 3) & ~3)) + sizeof(double) * (size_t)bk.nterm;
      buckets.push_back(bk);
    }
    (void)hipMemcpy(d_order.p, ord.data(), ord.size() * sizeof(int), hipMemcpyHostToDevice);
    order_mode = mode;
  }
  void drop_graph() {
    if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    graph_len = 0;
  }
  // one dispatch of launch group j on stream s
  void launch_bucket(size_t j, hipStream_t s) {
    const Bucket& bk = buckets[j];
    BatchArgs a = args;
    a.n_inst = bk.count; a.order = d_order.p + bk.start; a.lds_n = bk.n; a.lds_nlu = bk.nlu;
    const bool lt = !split && bk.terms;
    a.lds_terms = lt ? 1 : 0; a.lds_blob = bk.blob;
    const size_t shm = lt ? bk.shmem_terms : bk.shmem;
    const int bs = bk.block ? bk.block : block;
    if (split) hipLaunchKernelGGL((k_batch_step<false, false, false>), dim3(bk.count), dim3(bs), bk.shmem, s, a);
    else if (a.n_kcls > kMaxKclsLds) hipLaunchKernelGGL((k_batch_step<true, false, false>), dim3(bk.count), dim3(bs), bk.shmem, s, a);
    else if (lt) hipLaunchKernelGGL((k_batch_step<true, true, true>), dim3(bk.count), dim3(bs), shm, s, a);
    else hipLaunchKernelGGL((k_batch_step<true, true, false>), dim3(bk.count), dim3(bs), shm, s, a);
  }
  bool side_by_side() const {
    return concurrent_groups && only_bucket < 0 && buckets.size() > 1 && buckets.size() <= 4;
  }
  int launch() {
    if (split) {
      if (args.nt_mos > 0)
        hipLaunchKernelGGL(k_batch_eval, dim3((args.nt_mos + 255) / 256), dim3(256), 0, stream, args);
    }
    const bool side = side_by_side();
    if (side && hipEventRecord(fork_ev, stream) != hipSuccess) return XG_E_HIP;
    for (size_t j = 0; j < buckets.size(); ++j) {
      if (only_bucket >= 0 && (int)j != only_bucket) continue;
      hipStream_t s = stream;
      if (side) {
        s = gstream[j];
        if (hipStreamWaitEvent(s, fork_ev, 0) != hipSuccess) return XG_E_HIP;
      }
      launch_bucket(j, s);
      if (side) {
        if (hipEventRecord(join_ev[j], s) != hipSuccess) return XG_E_HIP;
        if (hipStreamWaitEvent(stream, join_ev[j], 0) != hipSuccess) return XG_E_HIP;
      }
    }
    if (pred_mode == 2) {
      const int st = mlp_forward_on_stream(net, pred_rows, d_feat.p, d_pred_out.p, stream);
      if (st != XG_OK) return st;
      hipLaunchKernelGGL(k_pred_apply, dim3(n_inst), dim3(256), 0, stream, args);
    }
    return hipGetLastError() == hipSuccess ? XG_OK : XG_E_HIP;
  }
  // k launches.  With the groups side by side and nothing between launches
  // that spans the groups (no split evaluation, no learned predictor), each
  // group runs its k launches back to back on its own stream between one fork
  // and one join: the groups hold disjoint instances, so a group that finishes
  // a launch early starts its next one instead of waiting for the slowest
  // group at every launch.
  int launch_chunk(int k) {
    if (!side_by_side() || split || pred_mode == 2) {
      for (int i = 0; i < k; ++i) {
        const int st = launch();
        if (st != XG_OK) return st;
      }
      return XG_OK;
    }
    if (hipEventRecord(fork_ev, stream) != hipSuccess) return XG_E_HIP;
    for (size_t j = 0; j < buckets.size(); ++j) {
      if (hipStreamWaitEvent(gstream[j], fork_ev, 0) != hipSuccess) return XG_E_HIP;
      for (int i = 0; i < k; ++i) launch_bucket(j, gstream[j]);
      if (hipEventRecord(join_ev[j], gstream[j]) != hipSuccess) return XG_E_HIP;
      if (hipStreamWaitEvent(stream, join_ev[j], 0) != hipSuccess) return XG_E_HIP;
    }
    return hipGetLastError() == hipSuccess ? XG_OK : XG_E_HIP;
  }
};

extern "C" {

int xg_batch_create(int n_inst, const xg_circuit_desc* descs, double tstop, double tstep, double hmax,
                    int record_max, xg_batch** out) {
  if (n_inst <= 0 || !descs || !out || !(tstop > 0.0) || !std::isfinite(tstop) || !(tstep > 0.0) ||
      !std::isfinite(tstep) || !(hmax >= 0.0) || !std::isfinite(hmax) || record_max < 0)
    return XG_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return XG_E_NOGPU;
  auto b = std::make_unique<xg_batch>();
  b->n_inst = n_inst;
  std::vector<int> ci;                              // concatenated class arrays
  std::map<std::vector<int>, int> cls_of;           // structure key -> class
  std::vector<double> kcls, tvbi, beta, tvto, linF, linQ, wave_data;
  std::vector<int> mcls, dev_inst, src_wave, wave_type, wave_off{0};
  std::vector<unsigned long long> mask;
  std::map<std::vector<double>, int> kindex;
  int dev_off = 0, vec_off = 0, nnz_off = 0, src_off = 0;
  long long dense_total = 0, term_total = 0;
  int max_blob = 0;
  std::vector<unsigned short> blob;
  int max_n = 0, max_nlu = 0;
  std::vector<double> all_bp;
  for (int i = 0; i < n_inst; ++i) {
    HostCircuit hc;
    int st = build_host_circuit(&descs[i], hc);
    if (st != XG_OK) return st;
    // Effective term masks: the record planes that can be nonzero for each
    // device given its class constants (identically-zero terms are neither
    // written nor summed; adding +0.0 would not change any sum).
    std::vector<unsigned long long> emask(hc.n_mos);
    for (int d = 0; d < hc.n_mos; ++d) {
      const double* k = &hc.cls[(size_t)hc.mos_cls[d] * KSTRIDE];
      unsigned long long m = hc.mos_mask[d];
      m &= ~(1ull << R_F_G);                               // old-Meyer gate current term is 0
      if (k[K_GD] == 0.0) m &= ~((1ull << R_F_D) | (1ull << (R_JF0 + 0)) | (1ull << (R_JF0 + 1)) | (1ull << (R_JF0 + 7)));
      if (k[K_GS] == 0.0) m &= ~((1ull << R_F_S) | (1ull << (R_JF0 + 2)) | (1ull << (R_JF0 + 3)) | (1ull << (R_JF0 + 13)));
      const bool no_caps = k[K_OXCAP] == 0.0 && k[K_GSO] == 0.0 && k[K_GDO] == 0.0 && k[K_GBO] == 0.0 &&
                           k[K_CBD] == 0.0 && k[K_CBDSW] == 0.0 && k[K_CBS] == 0.0 && k[K_CBSSW] == 0.0;
      if (no_caps) {
        for (int p : {R_Q_G, R_Q_B, R_Q_DP, R_Q_SP, R_QL_G, R_QL_B, R_QL_DP, R_QL_SP}) m &= ~(1ull << p);
        for (int j = 0; j < 14; ++j) m &= ~(1ull << (R_JQ0 + j));
      }
      emask[d] = m;
    }
    // structure key: pattern + gather maps + node indices + term masks + source rows
    std::vector<int> key;
    for (auto m : emask) { key.push_back((int)(m & 0xffffffffu)); key.push_back((int)(m >> 32)); }
    key.push_back(hc.n); key.push_back(hc.nnz); key.push_back(hc.n_mos); key.push_back((int)hc.srcs.size());
    key.insert(key.end(), hc.rowptr.begin(), hc.rowptr.end());
    key.insert(key.end(), hc.colind.begin(), hc.colind.end());
    key.insert(key.end(), hc.vg_ptr.begin(), hc.vg_ptr.end());
    key.insert(key.end(), hc.vg_ent.begin(), hc.vg_ent.end());
    key.insert(key.end(), hc.mg_ptr.begin(), hc.mg_ptr.end());
    key.insert(key.end(), hc.mg_ent.begin(), hc.mg_ent.end());
    key.insert(key.end(), hc.mos_lid.begin(), hc.mos_lid.end());
    for (auto& s : hc.srcs) { key.push_back(s.kind); key.push_back(s.li_p); key.push_back(s.li_n); }
    for (int k = 0; k < hc.nnz; ++k) key.push_back((hc.linF[k] != 0.0 ? 1 : 0) | (hc.linQ[k] != 0.0 ? 2 : 0));
    int cid;
    auto it = cls_of.find(key);
    if (it == cls_of.end()) {
      cid = (int)b->classes.size();
      cls_of.emplace(key, cid);
      ClassDesc C{};
      C.n = hc.n; C.nnz = hc.nnz; C.n_mos = hc.n_mos; C.n_src = (int)hc.srcs.size();
      LuSymbolic sym;
      // the linear values decide the transversal only through structure here:
      // pass the first instance's Jacobian-like magnitudes (linF + linQ + 1)
      std::vector<double> guide(hc.nnz, 1.0);
      for (int k = 0; k < hc.nnz; ++k) guide[k] = std::fabs(hc.linF[k]) + std::fabs(hc.linQ[k]) * 1e11 + 1e-3;
      try {
        if (lu_analyse(hc.n, hc.rowptr.data(), hc.colind.data(), guide.data(), sym, true) != XG_OK) return XG_E_SETUP;
      } catch (...) {
        return XG_E_SETUP;
      }
      C.nlu = sym.nlu;
      C.nflev = (int)sym.f_lvl_ptr.size() - 1;
      C.nbs = (int)sym.bs_lvl_ptr.size() - 1;
      auto push = [&](const std::vector<int>& v) { int o = (int)ci.size(); ci.insert(ci.end(), v.begin(), v.end()); return o; };
      const int n = hc.n;
      // ---- slot order of the terms (the order the gathers sum them in):
      //      vector terms grouped by (row, vector), matrix terms by (slot,
      //      dFdx/dQdx), each group in device order
      static const int vrow[R_JF0] = {0, 1, 2, 3, 4, 5, 1, 3, 4, 5, 3, 4, 5, 1, 3, 4, 5};
      std::vector<int> vcnt(4 * n + 1, 0), mcnt(2 * hc.nnz + 1, 0);
      auto vgroup = [&](int d, int p) { return 4 * hc.mos_lid[(size_t)6 * d + vrow[p]] + rec_vec_id(p); };
      auto mgroup = [&](int d, int p) {
        return p < R_JQ0 ? 2 * hc.mos_fslot[(size_t)17 * d + (p - R_JF0)] : 2 * hc.mos_qslot[(size_t)14 * d + (p - R_JQ0)] + 1;
      };
      for (int d = 0; d < hc.n_mos; ++d)
        for (int p = 0; p < R_COUNT; ++p) {
          if (!(emask[d] >> p & 1ull)) continue;
          if (p < R_JF0) vcnt[vgroup(d, p)]++; else mcnt[mgroup(d, p)]++;
        }
      std::vector<int> vptr(4 * n + 1, 0), mptr(2 * hc.nnz + 1, 0);
      for (int g = 0; g < 4 * n; ++g) vptr[g + 1] = vptr[g] + vcnt[g];
      mptr[0] = vptr[4 * n];
      for (int g = 0; g < 2 * hc.nnz; ++g) mptr[g + 1] = mptr[g] + mcnt[g];
      C.nterm = mptr[2 * hc.nnz];
      std::vector<int> tpos((size_t)hc.n_mos * R_COUNT, -1);
      bool tix_ok = false;
      {
        std::vector<int> vf(vptr.begin(), vptr.end() - 1), mf(mptr.begin(), mptr.end() - 1);
        for (int d = 0; d < hc.n_mos; ++d)
          for (int p = 0; p < R_COUNT; ++p) {
            if (!(emask[d] >> p & 1ull)) continue;
            tpos[(size_t)d * R_COUNT + p] = p < R_JF0 ? vf[vgroup(d, p)]++ : mf[mgroup(d, p)]++;
          }
      }
      {
        // term array: device d's nonzero planes at rank * n_mos + d; the
        // slot-ordered lists reach them through tix
        std::vector<int> tm((size_t)2 * hc.n_mos);
        int kt = 0;
        for (int d = 0; d < hc.n_mos; ++d) {
          tm[d] = (int)(unsigned)(emask[d] & 0xffffffffull);
          tm[(size_t)hc.n_mos + d] = (int)(unsigned)(emask[d] >> 32);
          kt = std::max(kt, __builtin_popcountll(emask[d]));
        }
        C.kterm = kt;
        C.tsize = std::max(1, kt * hc.n_mos);
        std::vector<int> tix(C.nterm, -1);
        for (int d = 0; d < hc.n_mos; ++d) {
          int rank = 0;
          for (int p = 0; p < R_COUNT; ++p) {
            if (!(emask[d] >> p & 1ull)) continue;
            const int e = tpos[(size_t)d * R_COUNT + p];
            if (e >= 0 && e < C.nterm) tix[e] = rank * hc.n_mos + d;
            ++rank;
          }
        }
        C.o_tmask = push(tm);
        C.o_tix = push(tix);
        tix_ok = (long long)kt * hc.n_mos < (1ll << 30) && kt <= R_COUNT;
        for (int x : tix) tix_ok = tix_ok && x >= 0 && x < C.tsize;
      }
      C.o_vptr = push(vptr); C.o_mptr = push(mptr);
      {
        std::vector<int> lg;
        for (int g = 0; g < 4 * n && (int)lg.size() < kMaxLong; ++g) if (vptr[g + 1] - vptr[g] >= kLongSeg) lg.push_back(g);
        C.n_vlong = (int)lg.size();
        for (int g = 0; g < 2 * hc.nnz && (int)lg.size() < kMaxLong; ++g) if (mptr[g + 1] - mptr[g] >= kLongSeg) lg.push_back(g);
        C.n_mlong = (int)lg.size() - C.n_vlong;
        // a class with more long groups than kMaxLong cannot use this path
        int total_long = 0;
        for (int g = 0; g < 4 * n; ++g) total_long += vptr[g + 1] - vptr[g] >= kLongSeg;
        for (int g = 0; g < 2 * hc.nnz; ++g) total_long += mptr[g + 1] - mptr[g] >= kLongSeg;
        if (total_long > kMaxLong) return XG_E_ARG;
        lg.push_back(0);
        C.o_long = push(lg);
      }
      // per-row source lists (local source index)
      std::vector<int> bgp_h, bge_h;
      bool blob_ok = false;
      {
        std::vector<int> cnt(n + 1, 0);
        for (auto& e : hc.srcs) {
          if (e.li_p < n) cnt[e.li_p]++;
          if (e.kind == 1 && e.li_n < n) cnt[e.li_n]++;
        }
        std::vector<int> ptr(n + 1, 0);
        for (int r = 0; r < n; ++r) ptr[r + 1] = ptr[r] + cnt[r];
        std::vector<int> ent(ptr[n]), fill(ptr.begin(), ptr.end() - 1);
        for (size_t si = 0; si < hc.srcs.size(); ++si) {
          const auto& e = hc.srcs[si];
          if (e.kind == 0) { if (e.li_p < n) ent[fill[e.li_p]++] = (int)si * 2; }
          else {
            if (e.li_p < n) ent[fill[e.li_p]++] = (int)si * 2 + 1;
            if (e.li_n < n) ent[fill[e.li_n]++] = (int)si * 2;
          }
        }
        C.o_bgp = push(ptr); C.o_bge = push(ent);
        bgp_h = ptr; bge_h = ent;
      }
      {
        std::vector<int> lidT((size_t)hc.n_mos * 6);
        for (int d = 0; d < hc.n_mos; ++d)
          for (int j = 0; j < 6; ++j) lidT[(size_t)j * hc.n_mos + d] = hc.mos_lid[(size_t)6 * d + j];
        C.o_lid = push(lidT);
      }
      C.nops = (int)sym.f_ops.size() / 2;
      // ---- uint16 program blob, staged into LDS by every launch
      {
        std::vector<int> fdiv16(sym.f_div.size());
        for (size_t t = 0; t < sym.f_div.size(); ++t) fdiv16[t] = sym.f_div[t] < 0 ? 0xFFFF : sym.f_div[t];
        const size_t b0 = blob.size();
        auto put = [&](const std::vector<int>& v) {
          int o = (int)(blob.size() - b0);
          for (int x : v) blob.push_back((unsigned short)x);
          return o;
        };
        // Refactor program with the forward solve L y = P b folded in: y_i
        // (entry nlu + i) is a target with the operations y_i -= L_ik y_k,
        // scheduled one level after the last L_ik and y_k it reads, so the
        // refactor's level sweep also produces y.  Within each level the
        // targets with >= kLongOps operations go last (a wavefront each);
        // fls[l] = first of them.
        const int nlu0 = C.nlu;
        std::vector<int> plev(nlu0, 0);
        for (size_t l = 0; l + 1 < sym.f_lvl_ptr.size(); ++l)
          for (int t = sym.f_lvl_ptr[l]; t < sym.f_lvl_ptr[l + 1]; ++t) plev[sym.f_tgt[t]] = (int)l;
        int nlev = (int)sym.f_lvl_ptr.size() - 1;
        std::vector<int> ylev(n, 0);
        for (int i = 0; i < n; ++i) {
          int lv = 0;
          for (int q = sym.lu_rowptr[i]; q < sym.lu_diag[i]; ++q)
            lv = std::max(lv, std::max(plev[q], ylev[sym.lu_colind[q]]) + 1);
          ylev[i] = lv;
          nlev = std::max(nlev, lv + 1);
        }
        std::vector<std::vector<int>> ylist(nlev);
        for (int i = 0; i < n; ++i) ylist[ylev[i]].push_back(i);
        std::vector<int> ftg, fop{0}, fops, fdv, fls, flp{0};
        for (int l = 0; l < nlev; ++l) {
          std::vector<int> shrt, lng;   // >= 0: factor target index t; < 0: forward-solve row -1 - i
          if (l + 1 < (int)sym.f_lvl_ptr.size())
            for (int t = sym.f_lvl_ptr[l]; t < sym.f_lvl_ptr[l + 1]; ++t)
              (sym.f_op_ptr[t + 1] - sym.f_op_ptr[t] >= kLongOps ? lng : shrt).push_back(t);
          for (int i : ylist[l]) (sym.lu_diag[i] - sym.lu_rowptr[i] >= kLongOps ? lng : shrt).push_back(-1 - i);
          // targets with no operation and no division keep the value the
          // Jacobian scatter (or the right-hand side) put there: no work
          auto noop = [&](int e) {
            return e >= 0 ? sym.f_op_ptr[e + 1] == sym.f_op_ptr[e] && fdiv16[e] == 0xFFFF
                          : sym.lu_diag[-1 - e] == sym.lu_rowptr[-1 - e];
          };
          shrt.erase(std::remove_if(shrt.begin(), shrt.end(), noop), shrt.end());
          auto add = [&](int e) {
            if (e >= 0) {
              ftg.push_back(sym.f_tgt[e]);
              for (int o = sym.f_op_ptr[e]; o < sym.f_op_ptr[e + 1]; ++o) {
                fops.push_back(sym.f_ops[2 * o]); fops.push_back(sym.f_ops[2 * o + 1]);
              }
              fdv.push_back(fdiv16[e]);
            } else {
              const int i = -1 - e;
              ftg.push_back(nlu0 + i);
              for (int q = sym.lu_rowptr[i]; q < sym.lu_diag[i]; ++q) {
                fops.push_back(q); fops.push_back(nlu0 + sym.lu_colind[q]);
              }
              fdv.push_back(0xFFFF);
            }
            fop.push_back((int)fops.size() / 2);
          };
          for (int e : shrt) add(e);
          fls.push_back((int)ftg.size());
          for (int e : lng) add(e);
          flp.push_back((int)ftg.size());
          if (flp.back() == flp[flp.size() - 2]) { flp.pop_back(); fls.pop_back(); }   // a level of copies only
        }
        C.nflev = (int)flp.size() - 1;
        C.nops = (int)fops.size() / 2;
        // solve rows: long L (forward) / U (backward) parts last in their level
        auto split_rows = [&](const std::vector<int>& lp, const std::vector<int>& rows, bool fwd,
                              std::vector<int>& out, std::vector<int>& split) {
          for (size_t l = 0; l + 1 < lp.size(); ++l) {
            std::vector<int> lng;
            for (int t = lp[l]; t < lp[l + 1]; ++t) {
              const int i = rows[t];
              const int len = fwd ? sym.lu_diag[i] - sym.lu_rowptr[i] : sym.lu_rowptr[i + 1] - sym.lu_diag[i] - 1;
              if (len >= kLongOps) lng.push_back(i); else out.push_back(i);
            }
            split.push_back((int)out.size());
            out.insert(out.end(), lng.begin(), lng.end());
          }
        };
        std::vector<int> bsr, bss;
        split_rows(sym.bs_lvl_ptr, sym.bs_rows, false, bsr, bss);
        // linear entries per row: CSR slots with a nonzero linear value
        std::vector<int> lrp{0}, lk;
        for (int r = 0; r < n; ++r) {
          for (int k = hc.rowptr[r]; k < hc.rowptr[r + 1]; ++k)
            if (hc.linF[k] != 0.0 || hc.linQ[k] != 0.0) lk.push_back(k);
          lrp.push_back((int)lk.size());
        }
        std::vector<int> fmeta;
        for (size_t t = 0; t < ftg.size(); ++t) {
          fmeta.push_back(ftg[t]); fmeta.push_back(fdv[t]);
          fmeta.push_back(fop[t]); fmeta.push_back(fop[t + 1] - fop[t]);
        }
        auto align = [&](size_t a) { while ((blob.size() - b0) % a) blob.push_back(0); };
        align(4);
        C.b_fmeta = put(fmeta);
        align(2);
        C.b_fops = put(fops);
        C.b_flp = put(flp); C.b_fls = put(fls);
        C.b_lurp = put(sym.lu_rowptr); C.b_luci = put(sym.lu_colind);
        C.b_diag = put(sym.lu_diag);
        C.b_bsp = put(sym.bs_lvl_ptr); C.b_bsr = put(bsr); C.b_bss = put(bss);
        // the arrays read once per iteration stay in HBM
        align(8);
        C.blob_lds = (int)(blob.size() - b0);
        C.b_prow = put(sym.perm_row);
        C.b_pcol = put(sym.perm_col); C.b_rp = put(hc.rowptr); C.b_cc = put(hc.colind); C.b_a2lu = put(sym.a2lu);
        C.b_vptr = put(vptr); C.b_mptr = put(mptr); C.b_lrp = put(lrp); C.b_lk = put(lk);
        C.b_bgp = put(bgp_h); C.b_bge = put(bge_h);
        {
          std::vector<int> vflag(hc.n);
          for (int r = 0; r < hc.n; ++r) vflag[r] = hc.var_type[r] == 'V' ? 1 : 0;
          C.b_vrow = put(vflag);
        }
        bool ops_in = true;
        for (int x : fops) ops_in = ops_in && x >= 0 && x < nlu0 + n;
        blob_ok = (int)ftg.size() <= nlu0 + n && (int)fls.size() == C.nflev && (int)flp.size() == C.nflev + 1 &&
                  (int)bsr.size() == n && (int)bss.size() == C.nbs && fop.back() == C.nops &&
                  (int)fops.size() == 2 * C.nops && ops_in && nlu0 + n < 0xFFFF && C.nops < 0xFFFF;
        while ((blob.size() - b0) % 8) blob.push_back(0);
        C.blob_off = (int)b0;
        C.blob_len = (int)(blob.size() - b0);
      }
      // Every index the kernel reads is checked here, on the host, before any launch.
      {
        auto in = [](const std::vector<int>& v, int lo, int hi) {
          for (int x : v) if (x < lo || x >= hi) return false;
          return true;
        };
        bool ok = in(hc.colind, 0, C.n) && in(hc.mos_lid, 0, C.n + 1) && in(sym.a2lu, 0, C.nlu) &&
                  in(sym.lu_colind, 0, C.n) && in(sym.lu_diag, 0, C.nlu) && in(sym.f_tgt, 0, C.nlu) &&
                  in(sym.f_ops, 0, C.nlu) && in(sym.f_div, -1, C.nlu) && in(sym.fs_rows, 0, C.n) &&
                  in(sym.bs_rows, 0, C.n) && in(sym.perm_row, 0, C.n) && in(sym.perm_col, 0, C.n) &&
                  in(tpos, -1, C.nterm) && tix_ok;
        // every value of the uint16 program must fit (0xFFFF is the "no division" sentinel)
        ok = ok && C.nlu < 0xFFFF && (int)sym.f_ops.size() / 2 < 0xFFFF && C.nnz < 0xFFFF && C.n < 0xFFFF;
        // term ranges are uint16 in the program; source values are staged in LDS
        ok = ok && blob_ok && C.nterm < 0xFFFF && C.n_src <= kMaxSrc && 2 * C.nnz < 0xFFFF;
        ok = ok && (int)sym.f_op_ptr.size() == C.nlu + 1 && sym.f_op_ptr[C.nlu] * 2 == (int)sym.f_ops.size();
        ok = ok && sym.f_lvl_ptr.back() == C.nlu && sym.fs_lvl_ptr.back() == C.n && sym.bs_lvl_ptr.back() == C.n;
        ok = ok && (int)hc.rowptr.size() == C.n + 1 && hc.rowptr[C.n] == C.nnz;
        ok = ok && vptr[0] == 0 && mptr[2 * hc.nnz] == C.nterm;
        for (auto& e : hc.srcs) ok = ok && e.li_p <= C.n && e.li_n <= C.n && e.wave >= 0 && e.wave < (int)hc.wave_type.size();
        if (!ok) return XG_E_SETUP;
      }
      max_blob = std::max(max_blob, C.blob_lds);
      b->classes.push_back(C);
      max_n = std::max(max_n, C.n);
      max_nlu = std::max(max_nlu, C.nlu + C.n);
    } else {
      cid = it->second;
    }
    InstDesc D{};
    D.cls = cid; D.dev_off = dev_off; D.vec_off = vec_off; D.nnz_off = nnz_off; D.src_off = src_off;
    D.rec_off = i * std::max(record_max, 0);
    D.dense_off = dense_total;
    dense_total += (long long)hc.n * hc.n;
    D.term_off = term_total;
    term_total += b->classes[cid].tsize;
    b->insts.push_back(D);
    b->cls_n.push_back(hc.n_mos + hc.n);
    // devices
    for (int d = 0; d < hc.n_mos; ++d) {
      std::vector<double> kk(hc.cls.begin() + (size_t)hc.mos_cls[d] * KSTRIDE, hc.cls.begin() + (size_t)(hc.mos_cls[d] + 1) * KSTRIDE);
      auto kt = kindex.find(kk);
      int kid;
      if (kt == kindex.end()) { kid = (int)kindex.size(); kindex.emplace(kk, kid); kcls.insert(kcls.end(), kk.begin(), kk.end()); }
      else kid = kt->second;
      mcls.push_back(kid);
      dev_inst.push_back(i);
    }
    tvbi.insert(tvbi.end(), hc.mos_tvbi.begin(), hc.mos_tvbi.end());
    beta.insert(beta.end(), hc.mos_beta.begin(), hc.mos_beta.end());
    tvto.insert(tvto.end(), hc.mos_tvto.begin(), hc.mos_tvto.end());
    for (auto m : hc.mos_mask) mask.push_back(m);
    linF.insert(linF.end(), hc.linF.begin(), hc.linF.end());
    linQ.insert(linQ.end(), hc.linQ.begin(), hc.linQ.end());
    // waves (renumbered into the batch table) and sources
    const int wbase = (int)wave_type.size();
    for (size_t w = 0; w < hc.wave_type.size(); ++w) {
      wave_type.push_back(hc.wave_type[w]);
      wave_data.insert(wave_data.end(), hc.wave_data.begin() + hc.wave_off[w], hc.wave_data.begin() + hc.wave_off[w + 1]);
      wave_off.push_back((int)wave_data.size());
    }
    for (auto& s : hc.srcs) src_wave.push_back(wbase + s.wave);
    {
      // this instance's own breakpoints (a Xyce run of it alone stops at these)
      std::vector<double> bp;
      for (size_t si = 0; si < hc.srcs.size(); ++si) wave_breakpoints(hc, hc.srcs[si].wave, tstep, tstop, bp);
      bp.push_back(tstop);
      std::sort(bp.begin(), bp.end());
      bp.erase(std::unique(bp.begin(), bp.end()), bp.end());
      b->insts[i].bp_off = (int)all_bp.size();
      b->insts[i].n_bp = (int)bp.size();
      all_bp.insert(all_bp.end(), bp.begin(), bp.end());
    }
    dev_off += hc.n_mos; vec_off += hc.n; nnz_off += hc.nnz; src_off += (int)hc.srcs.size();
  }
  const size_t NT = dev_off;
  if (hipStreamCreate(&b->stream) != hipSuccess) return XG_E_HIP;
  for (int j = 0; j < 4; ++j) {
    if (hipStreamCreate(&b->gstream[j]) != hipSuccess) return XG_E_HIP;
    HIPCHK(hipEventCreateWithFlags(&b->join_ev[j], hipEventDisableTiming));
  }
  HIPCHK(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming));
  HIPCHK(hipEventCreate(&b->ev0)); HIPCHK(hipEventCreate(&b->ev1));
  HIPCHK(b->d_ci.up(ci)); HIPCHK(b->d_cls.up(b->classes)); HIPCHK(b->d_inst.up(b->insts));
  HIPCHK(b->d_dev_inst.up(dev_inst)); HIPCHK(b->d_lim.zeros(n_inst));
  HIPCHK(b->d_kcls.up(kcls)); HIPCHK(b->d_mcls.up(mcls)); HIPCHK(b->d_tvbi.up(tvbi)); HIPCHK(b->d_beta.up(beta));
  HIPCHK(b->d_tvto.up(tvto)); HIPCHK(b->d_mask.up(mask));
  HIPCHK(b->d_sto_cur.zeros(6 * NT)); HIPCHK(b->d_sto_nxt.zeros(6 * NT));
  HIPCHK(b->d_sta_cur.zeros(8 * NT)); HIPCHK(b->d_sta_nxt.zeros(8 * NT));
  HIPCHK(b->d_terms.zeros((size_t)std::max(1LL, term_total)));
  HIPCHK(b->d_blob.up(blob));
  const size_t NV = vec_off;
  for (auto* v : {&b->d_x, &b->d_xn0, &b->d_xH0, &b->d_xH1, &b->d_xH2, &b->d_qH0, &b->d_qH1, &b->d_qH2, &b->d_F,
                  &b->d_Q, &b->d_B, &b->d_F0, &b->d_B0, &b->d_wt})
    HIPCHK(v->zeros(NV));
  HIPCHK(b->d_linF.up(linF)); HIPCHK(b->d_linQ.up(linQ));
  HIPCHK(b->d_src_wave.up(src_wave)); HIPCHK(b->d_wave_type.up(wave_type)); HIPCHK(b->d_wave_off.up(wave_off));
  HIPCHK(b->d_wave_data.up(wave_data)); HIPCHK(b->d_bp.up(all_bp));
  HIPCHK(b->d_st.zeros(n_inst));
  HIPCHK(b->d_dense.zeros((size_t)dense_total));
  k_init_state<<<(n_inst + 255) / 256, 256>>>(n_inst, b->d_st.p);
  HIPCHK(b->d_order.zeros(n_inst));
  if (record_max > 0) {
    HIPCHK(b->d_rec_t.zeros((size_t)n_inst * record_max));
    HIPCHK(b->d_rec_x.zeros((size_t)n_inst * record_max * max_n));
  }
  HIPCHK(hipDeviceSynchronize());
  BatchArgs& A = b->args;
  A.method = 0;
  A.n_inst = n_inst; A.order = b->d_order.p; A.inst = b->d_inst.p; A.st = b->d_st.p; A.cls = b->d_cls.p; A.ci = b->d_ci.p;
  A.nt_mos = (int)NT; A.kcls = b->d_kcls.p; A.mcls = b->d_mcls.p; A.tvbi = b->d_tvbi.p; A.beta = b->d_beta.p;
  A.tvto = b->d_tvto.p; A.mask = b->d_mask.p; A.sto_cur = b->d_sto_cur.p; A.sto_nxt = b->d_sto_nxt.p;
  A.sta_cur = b->d_sta_cur.p; A.sta_nxt = b->d_sta_nxt.p; A.terms = b->d_terms.p; A.blob = b->d_blob.p;
  A.x = b->d_x.p; A.xn0 = b->d_xn0.p; A.xH0 = b->d_xH0.p; A.xH1 = b->d_xH1.p; A.xH2 = b->d_xH2.p; A.qH0 = b->d_qH0.p;
  A.qH1 = b->d_qH1.p; A.qH2 = b->d_qH2.p; A.F = b->d_F.p; A.Q = b->d_Q.p; A.B = b->d_B.p; A.F0 = b->d_F0.p;
  A.B0 = b->d_B0.p; A.wt = b->d_wt.p; A.linF = b->d_linF.p; A.linQ = b->d_linQ.p; A.src_wave = b->d_src_wave.p;
  A.wave_type = b->d_wave_type.p; A.wave_off = b->d_wave_off.p; A.wave_data = b->d_wave_data.p;
  A.bp = b->d_bp.p; A.tstop = tstop; A.tstep = tstep; A.hmax = hmax;
  A.rec_max = std::max(record_max, 0); A.rec_t = b->d_rec_t.p; A.rec_x = b->d_rec_x.p;
  A.max_n = max_n; A.max_nlu = max_nlu; A.max_blob = max_blob;
  A.lds_n = max_n; A.lds_nlu = max_nlu;
  A.n_kcls = (int)(kcls.size() / KSTRIDE); A.iters = 1; A.prof = nullptr;
  A.dense = b->d_dense.p;
  A.pred_mode = 0;
  A.pred_base = 0;
  A.pred_dead = 0.0;
  A.dev_inst = b->d_dev_inst.p; A.lim = b->d_lim.p;
  b->shmem = xg_batch::lds_bytes(max_n, max_nlu, max_blob);
  for (const ClassDesc& C : b->classes) b->cls_lds.push_back(xg_batch::lds_bytes(C.n, C.nlu + C.n, C.blob_lds));
  for (const InstDesc& D : b->insts) b->inst_cls.push_back(D.cls);
  if (b->shmem > kMaxLdsBytes || max_n > 64 * kMaxChunks) return XG_E_ARG;
  b->set_order(1);
  *out = b.release();
  return XG_OK;
}

int xg_batch_destroy(xg_batch* b) {
  if (!b) return XG_E_ARG;
  delete b;
  return XG_OK;
}

int xg_batch_advance(xg_batch* b, int launches, int graph_chunk) {
  if (!b || launches < 0) return XG_E_ARG;
  if (launches > 0) b->started = true;
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  int left = launches;
  if (graph_chunk > 0 && launches >= graph_chunk) {
    if (!b->gexec || b->graph_len != graph_chunk) {
      b->drop_graph();
      HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int st = b->launch_chunk(graph_chunk);
      if (st != XG_OK) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
      HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      b->graph_len = graph_chunk;
    }
    while (left >= graph_chunk) {
      HIPCHK(hipGraphLaunch(b->gexec, b->stream));
      left -= graph_chunk;
    }
  }
  if (left > 0) {
    const int st = b->launch_chunk(left);
    if (st != XG_OK) return st;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  HIPCHK(hipEventSynchronize(b->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_seconds = ms * 1e-3;
  b->launches += launches;
  return XG_OK;
}

int xg_batch_predictor_options(xg_batch* b, int base, double deadzone) {
  if (!b || base < 0 || base > 1 || !(deadzone >= 0.0)) return XG_E_ARG;
  if (b->launches != 0) return XG_E_ARG;
  b->args.pred_base = base;
  b->args.pred_dead = deadzone;
  b->drop_graph();
  return XG_OK;
}

int xg_batch_predictor(xg_batch* b, int mode, xg_mlp* net, int pair_capacity) {
  if (!b || mode < 0 || mode > 2 || pair_capacity < 0) return XG_E_ARG;
  if (b->launches != 0) return XG_E_ARG;            // the history must start at the operating point
  if (mode == 2) {
    int in = 0, out = 0, max_rows = 0;
    if (!net || mlp_dims(net, &in, &out, &max_rows) != XG_OK) return XG_E_ARG;
    if (in != XG_PRED_FEAT || out != XG_PRED_WIN) return XG_E_ARG;
  }
  if (mode == 1 && pair_capacity == 0) return XG_E_ARG;
  // windows and training-pair records per instance
  long long win = 0, rec = 0, tn = 0;
  for (int i = 0; i < b->n_inst; ++i) {
    const int n = b->classes[b->insts[i].cls].n;
    const int nw = (n + XG_PRED_WIN - 1) / XG_PRED_WIN;
    b->insts[i].win_off = (int)win;
    b->insts[i].pair_off = rec;
    win += nw;
    rec += (long long)nw * (mode == 1 ? pair_capacity : 0);
    tn += n;
  }
  const int rows = (int)((win + 63) / 64 * 64);
  if (mode == 2) {
    int in = 0, out = 0, max_rows = 0;
    mlp_dims(net, &in, &out, &max_rows);
    if (rows > max_rows) return XG_E_ARG;
  }
  b->pred_mode = mode; b->net = net; b->pred_rows = rows; b->total_windows = (int)win;
  b->pair_cap = mode == 1 ? pair_capacity : 0; b->total_n = tn; b->total_records = rec;
  if (mode != 0) {
    HIPCHK(b->d_xP.zeros((size_t)5 * tn));   // x_n .. x_{n-3}, x_cubic
    HIPCHK(b->d_feat.zeros((size_t)rows * XG_PRED_FEAT));
    HIPCHK(b->d_pred_out.zeros((size_t)rows * XG_PRED_WIN));
  }
  if (mode == 1) {
    HIPCHK(b->d_pair_x.zeros((size_t)rec * XG_PRED_FEAT));
    HIPCHK(b->d_pair_t.zeros((size_t)rec * XG_PRED_WIN));
    HIPCHK(b->d_pair_m.zeros((size_t)rec * XG_PRED_WIN));
  }
  HIPCHK(hipMemcpy(b->d_inst.p, b->insts.data(), sizeof(InstDesc) * b->n_inst, hipMemcpyHostToDevice));
  BatchArgs& A = b->args;
  A.pred_mode = mode; A.pair_cap = b->pair_cap; A.xP = b->d_xP.p; A.total_n = tn; A.feat = b->d_feat.p;
  A.pred_out = b->d_pred_out.p; A.pair_x = b->d_pair_x.p; A.pair_t = b->d_pair_t.p;
  A.pair_m = b->d_pair_m.p;
  b->drop_graph();
  return XG_OK;
}

int xg_batch_pairs(xg_batch* b, long long max_records, unsigned short* x_dev, float* t_dev, float* m_dev,
                   long long* count) {
  if (!b || !count || b->pred_mode != 1) return XG_E_ARG;
  std::vector<InstState> st(b->n_inst);
  HIPCHK(hipMemcpy(st.data(), b->d_st.p, sizeof(InstState) * b->n_inst, hipMemcpyDeviceToHost));
  std::vector<long long> src;
  for (int i = 0; i < b->n_inst; ++i) {
    const int nw = (b->classes[b->insts[i].cls].n + XG_PRED_WIN - 1) / XG_PRED_WIN;
    for (long long k = 0; k < st[i].n_pairs * nw; ++k) src.push_back(b->insts[i].pair_off + k);
  }
  *count = (long long)src.size();
  if (!x_dev || !t_dev || !m_dev || max_records <= 0) return XG_OK;
  const long long n = std::min<long long>(max_records, (long long)src.size());
  for (long long k = 0; k < n; ++k)
    if (src[k] < 0 || src[k] >= b->total_records) return XG_E_ARG;
  DArr<long long> d_src;
  src.resize(n);
  HIPCHK(d_src.up(src));
  if (n) hipLaunchKernelGGL(k_pair_gather, dim3((unsigned)n), dim3(256), 0, b->stream, n, d_src.p, b->d_pair_x.p,
                            b->d_pair_t.p, b->d_pair_m.p, x_dev, t_dev, m_dev);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  *count = n;
  return XG_OK;
}

int xg_batch_pred_counters(xg_batch* b, double* c) {
  if (!b || !c) return XG_E_ARG;
  std::vector<InstState> st(b->n_inst);
  HIPCHK(hipMemcpy(st.data(), b->d_st.p, sizeof(InstState) * b->n_inst, hipMemcpyDeviceToHost));
  double pairs = 0, learned = 0, rej = 0;
  for (auto& s : st) { pairs += s.n_pairs; learned += s.n_learned; rej += s.n_learned_rejects; }
  c[0] = pairs; c[1] = learned; c[2] = rej; c[3] = b->total_windows; c[4] = b->pred_mode;
  return XG_OK;
}

int xg_batch_counters(xg_batch* b, double* c) {
  if (!b || !c) return XG_E_ARG;
  std::vector<InstState> st(b->n_inst);
  HIPCHK(hipMemcpy(st.data(), b->d_st.p, sizeof(InstState) * b->n_inst, hipMemcpyDeviceToHost));
  double it = 0, ld = 0, acc = 0, rej = 0, done = 0, fail = 0, dense = 0;
  for (auto& s : st) {
    it += s.n_iters; ld += s.n_loads; acc += s.n_steps; rej += s.n_rejects; dense += s.n_dense;
    done += s.mode == 2; fail += s.mode < 0;
  }
  c[0] = it; c[1] = ld; c[2] = acc; c[3] = rej; c[4] = done; c[5] = fail; c[6] = (double)b->launches;
  c[7] = b->last_seconds;
  c[8] = dense;
  return XG_OK;
}

int xg_batch_instance(xg_batch* b, int inst, double* ts, double* x) {
  if (!b || inst < 0 || inst >= b->n_inst) return XG_E_ARG;
  InstState s;
  HIPCHK(hipMemcpy(&s, b->d_st.p + inst, sizeof(s), hipMemcpyDeviceToHost));
  if (ts) {
    ts[0] = s.t; ts[1] = s.h; ts[2] = s.order; ts[3] = s.mode; ts[4] = (double)s.n_iters; ts[5] = (double)s.n_steps;
    ts[6] = (double)s.n_rejects; ts[7] = (double)s.n_loads;
  }
  if (x) {
    const ClassDesc& C = b->classes[b->insts[inst].cls];
    HIPCHK(hipMemcpy(x, b->d_x.p + b->insts[inst].vec_off, sizeof(double) * C.n, hipMemcpyDeviceToHost));
  }
  return XG_OK;
}

int xg_batch_points(xg_batch* b, int inst, int max_points, double* t, double* x) {
  if (!b || inst < 0 || inst >= b->n_inst || b->args.rec_max == 0 || max_points < 0 || !t || !x) return XG_E_ARG;
  InstState s;
  HIPCHK(hipMemcpy(&s, b->d_st.p + inst, sizeof(s), hipMemcpyDeviceToHost));
  const int cnt = (int)std::min<long long>(s.n_rec, std::min(max_points, b->args.rec_max));
  const ClassDesc& C = b->classes[b->insts[inst].cls];
  const size_t base = b->insts[inst].rec_off;
  HIPCHK(hipMemcpy(t, b->d_rec_t.p + base, sizeof(double) * cnt, hipMemcpyDeviceToHost));
  std::vector<double> tmp((size_t)cnt * b->args.max_n);
  if (cnt) HIPCHK(hipMemcpy(tmp.data(), b->d_rec_x.p + base * b->args.max_n, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
  for (int p = 0; p < cnt; ++p) std::copy(tmp.begin() + (size_t)p * b->args.max_n, tmp.begin() + (size_t)p * b->args.max_n + C.n, x + (size_t)p * C.n);
  return cnt;
}

int xg_batch_autotune(xg_batch* b, double* result) {
  if (!b) return XG_E_ARG;
  // Each candidate runs a few real launches (they advance the simulation;
  // every launch does valid Newton iterations); the fastest time per batched
  // iteration wins.  Split form: workgroup size x instance grouping.  Fused
  // form: workgroup size x grouping with 8 iterations per launch, then, with
  // LDS-sized groups, each group's own workgroup size, the iterations per
  // launch (4, 8, 16, 32, 64) and whether the groups run side by side.
  // Workgroup sizes: 128, 256, 512 threads, per group also 64.
  const int blocks[3] = {128, 256, 512};
  // per group also one wavefront per instance (its barriers cost nothing)
  const int gblocks[4] = {64, 128, 256, 512};
  auto time_launches = [&](int warm, int timed, double* us) -> int {
    int st = xg_batch_advance(b, warm, 0);
    if (st != XG_OK) return st;
    st = xg_batch_advance(b, timed, 0);
    if (st != XG_OK) return st;
    *us = b->last_seconds / timed * 1e6;
    return XG_OK;
  };
  struct Cand { double us; int block, order, k; bool conc; std::vector<int> gblocks, gterms; };
  Cand best_split{1e30, b->block, 1, 1, false, {}}, best_fused{1e30, b->block, 1, 8, false, {}};
  b->concurrent_groups = false;
  for (int sp = 1; sp >= 0; --sp) {
    Cand& bc = sp ? best_split : best_fused;
    for (int o = 0; o < 2; ++o) {
      b->set_order(o);
      for (int bs : blocks) {
        b->block = bs;
        b->split = sp;
        b->args.iters = sp ? 1 : 8;
        b->drop_graph();
        double us = 0;
        const int st = time_launches(2, 8, &us);
        if (st != XG_OK) return st;
        us /= b->args.iters;
        if (us < bc.us) { bc.us = us; bc.block = bs; bc.order = o; bc.k = b->args.iters; }
      }
    }
  }
  // fused, grouped by LDS need: per-group workgroup size, iterations per
  // launch, concurrency
  b->split = 0;
  b->block = best_fused.block;
  b->set_order(1);
  {
    double best_total = 1e30;
    int best_kk = 8;
    std::vector<int> best_blocks(b->buckets.size(), best_fused.block), best_terms(b->buckets.size(), 0);
    for (int kk : {4, 8, 16, 32, 64}) {
      b->args.iters = kk;
      double total = 0;
      std::vector<int> bb(b->buckets.size(), best_fused.block), bt(b->buckets.size(), 0);
      for (size_t j = 0; j < b->buckets.size(); ++j) {
        double bj = 1e30;
        for (int bs : gblocks) {
          for (int lt = 0; lt < 2; ++lt) {
            // the term arrays in LDS where they fit a workgroup's LDS
            if (lt && (b->buckets[j].shmem_terms > kMaxLdsBytes || b->args.n_kcls > kMaxKclsLds)) continue;
            b->buckets[j].block = bs;
            b->buckets[j].terms = lt != 0;
            b->only_bucket = (int)j;
            double us = 0;
            const int st = time_launches(1, 4, &us);
            b->only_bucket = -1;
            if (st != XG_OK) return st;
            if (us < bj) { bj = us; bb[j] = bs; bt[j] = lt; }
          }
        }
        b->buckets[j].block = bb[j];
        b->buckets[j].terms = bt[j] != 0;
        total += bj;
      }
      total /= kk;
      if (total < best_total) { best_total = total; best_kk = kk; best_blocks = bb; best_terms = bt; }
    }
    b->args.iters = best_kk;
    for (size_t j = 0; j < b->buckets.size(); ++j) {
      b->buckets[j].block = best_blocks[j];
      b->buckets[j].terms = best_terms[j] != 0;
    }
    double seq = 0, con = 0;
    b->concurrent_groups = false;
    int st = time_launches(1, 4, &seq);
    if (st != XG_OK) return st;
    b->concurrent_groups = true;
    st = time_launches(1, 4, &con);
    if (st != XG_OK) return st;
    const double grouped = std::min(seq, con) / best_kk;
    if (grouped < best_fused.us) {
      best_fused = Cand{grouped, best_fused.block, 1, best_kk, con < seq, best_blocks, best_terms};
    }
  }
  // the winner
  const bool fused = best_fused.us < best_split.us;
  const Cand& w = fused ? best_fused : best_split;
  b->split = fused ? 0 : 1;
  b->block = w.block;
  b->args.iters = w.k;
  b->set_order(w.order);
  b->concurrent_groups = w.conc;
  if (!w.gblocks.empty() && w.gblocks.size() == b->buckets.size())
    for (size_t j = 0; j < b->buckets.size(); ++j) b->buckets[j].block = w.gblocks[j];
  if (!w.gterms.empty() && w.gterms.size() == b->buckets.size())
    for (size_t j = 0; j < b->buckets.size(); ++j) b->buckets[j].terms = w.gterms[j] != 0;
  b->drop_graph();
  b->tune_us = w.us;
  if (result) {
    result[0] = w.block; result[1] = w.order; result[2] = w.us; result[3] = b->split; result[4] = w.k;
    for (int j = 0; j < 4; ++j)
      result[5 + j] = j < (int)b->buckets.size() ? (b->buckets[j].block ? b->buckets[j].block : b->block) : 0;
    result[9] = b->concurrent_groups ? 1 : 0;
  }
  return XG_OK;
}

int xg_batch_set_method(xg_batch* b, int method) {
  if (!b || (method != 0 && method != 1) || b->started) return XG_E_ARG;
  b->args.method = method;
  b->drop_graph();
  return XG_OK;
}

int xg_batch_group_info(xg_batch* b, int group, int* info) {
  if (!b || !info || group < 0) return XG_E_ARG;
  if (group >= (int)b->buckets.size()) return 0;
  const auto& bk = b->buckets[group];
  info[0] = bk.count;
  info[1] = bk.block ? bk.block : b->block;
  info[2] = !b->split && bk.terms ? 1 : 0;
  info[3] = (int)(!b->split && bk.terms ? bk.shmem_terms : bk.shmem);
  info[4] = (int)bk.shmem_terms;
  return (int)b->buckets.size();
}

int xg_batch_dc_strategy(xg_batch* b, int first_stage) {
  if (!b || first_stage < 0 || first_stage > 2 || b->started) return XG_E_ARG;
  k_dc_first_stage<<<(b->n_inst + 255) / 256, 256, 0, b->stream>>>(b->n_inst, b->d_st.p, first_stage);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  return XG_OK;
}

int xg_batch_dc_stage(xg_batch* b, int inst, int* stage) {
  if (!b || inst < 0 || inst >= b->n_inst || !stage) return XG_E_ARG;
  InstState s;
  HIPCHK(hipMemcpy(&s, b->d_st.p + inst, sizeof(s), hipMemcpyDeviceToHost));
  *stage = (int)s.dc_stage;
  return XG_OK;
}

int xg_batch_config(xg_batch* b, int block, int iters_per_launch, int split_eval, int concurrent_groups) {
  if (!b || iters_per_launch < 0 || iters_per_launch > 1024 || split_eval < -1 || split_eval > 2 ||
      concurrent_groups < -1 || concurrent_groups > 1)
    return XG_E_ARG;
  if (concurrent_groups >= 0) b->concurrent_groups = concurrent_groups == 1;
  if (block) {
    if (block != 64 && block != 128 && block != 256 && block != 512) return XG_E_ARG;
    b->block = block;
    for (auto& bk : b->buckets) bk.block = 0;
  }
  if (split_eval >= 0) {
    b->split = split_eval == 1 ? 1 : 0;
    // 2: fused, each group's term arrays in LDS where they fit; 0: in HBM
    for (auto& bk : b->buckets) bk.terms = split_eval == 2 && bk.shmem_terms <= kMaxLdsBytes && b->args.n_kcls <= kMaxKclsLds;
  }
  if (iters_per_launch) b->args.iters = iters_per_launch;
  if (b->split) b->args.iters = 1;
  b->drop_graph();
  return XG_OK;
}

int xg_batch_profile(xg_batch* b, int enable, double* phase_us) {
  if (!b) return XG_E_ARG;
  if (enable && !b->d_prof.p) {
    HIPCHK(b->d_prof.zeros(16));
  }
  if (phase_us && b->d_prof.p) {
    unsigned long long t[16];
    HIPCHK(hipMemcpy(t, b->d_prof.p, sizeof(t), hipMemcpyDeviceToHost));
    int dev = 0, khz = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    for (int k = 0; k < 15; ++k) phase_us[k] = khz > 0 ? (double)t[k] / (double)khz * 1e3 : 0.0;
    phase_us[15] = (double)t[15];
  }
  if (enable >= 0) b->args.prof = enable ? b->d_prof.p : nullptr;
  if (enable > 0) HIPCHK(hipMemset(b->d_prof.p, 0, 16 * sizeof(unsigned long long)));
  b->drop_graph();
  return XG_OK;
}


}  // extern "C"
