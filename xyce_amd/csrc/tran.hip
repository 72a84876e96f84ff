// Device-resident DC operating point and transient of ONE circuit (a power
// grid, a gate array): every vector of the Newton and time-integration loop
// lives in HBM; the host sequences the launches and reads back a few scalars
// per Newton iteration (norms, the limiter flag) to take the decisions.
//
// Reference loop (paths relative to the Xyce source root):
//   Transient::doLoopProcess           src/AnalysisPKG/N_ANP_Transient.C:1184
//     takeAnIntegrationStep_           :3262
//   OneStep predictor / step control   src/TimeIntegrationPKG/N_TIA_OneStep.C:135,1946,2124
//   DampedNewton::solve                src/NonlinearSolverPKG/N_NLS_DampedNewton.C:362-509
//     updateWeights_ :297, converged_ :1191
//   AmesosSolver::doSolve              src/LinearAlgebraServicesPKG/N_LAS_AmesosSolver.C:216
//   DC operating point: Newton from initJct, then gmin stepping, then source
//   stepping (Interface::spiceStrategy, src/NonlinearSolverPKG/N_NLS_NOX_Interface.C:321-437)
//
// Device loads come from the drop-in circuit (kernels.hip), factorisation and
// solves from the drop-in LU (either method, FP64 or FP32 factors with FP64
// refinement), all on device arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "xg_internal.h"

#define HIPCHK(x)                          \
  do {                                     \
    hipError_t e_ = (x);                   \
    if (e_ != hipSuccess) return XG_E_HIP; \
  } while (0)

namespace xg {

// every multiply-add rounded twice, as in the oracle and a baseline x86-64
// build of the reference
#pragma clang fp contract(off)

namespace {

template <class T>
struct DevArr {
  T* p = nullptr;
  size_t n = 0;
  ~DevArr() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t cnt) {
    if (p) { (void)hipFree(p); p = nullptr; }
    n = cnt;
    return cnt ? hipMalloc((void**)&p, cnt * sizeof(T)) : hipSuccess;
  }
};

inline unsigned blocks(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

// ---- fixed-order reductions: kRed blocks each sum a grid-strided slice in
// index order, then a tree in LDS; one block then sums the partials in block
// order.  The result does not depend on timing.
constexpr int kRed = 256;
constexpr int kRedThreads = 256;
enum RedOp { RED_SUMSQ = 0, RED_MAXABS = 1, RED_MAXRATIO = 2, RED_SUMSQ_DIFF_OVER = 3 };

template <int OP>
__device__ __forceinline__ double red_term(int i, const double* a, const double* b, double s) {
  if (OP == RED_SUMSQ) return a[i] * a[i];
  if (OP == RED_MAXABS) return fabs(a[i]);
  if (OP == RED_MAXRATIO) return fabs(a[i]) / b[i];
  const double q = (a[i] - b[i]) / s;      // RED_SUMSQ_DIFF_OVER: sum ((a - b) / s)^2
  return q * q;
}

template <int OP>
__device__ __forceinline__ double red_join(double u, double v) {
  if (OP == RED_MAXABS || OP == RED_MAXRATIO) return (u != u || v != v) ? NAN : fmax(u, v);
  return u + v;
}

template <int OP>
__global__ void __launch_bounds__(kRedThreads) k_red_part(int n, const double* a, const double* b, double s,
                                                          double* part) {
  __shared__ double sm[kRedThreads];
  double acc = 0.0;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += kRed * kRedThreads)
    acc = red_join<OP>(acc, red_term<OP>(i, a, b, s));
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] = red_join<OP>(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

template <int OP>
__global__ void __launch_bounds__(kRed) k_red_fin(const double* part, double* out) {
  __shared__ double sm[kRed];
  sm[threadIdx.x] = part[threadIdx.x];
  __syncthreads();
  for (int w = kRed / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] = red_join<OP>(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

// probe rows: per row r, the compensated sum of w_r[i] * x[i] (each product
// rounded, then summed exactly up to the last (sum, correction) pair)
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__global__ void __launch_bounds__(kRedThreads) k_probe_part(int n, int k, const double* w, const double* x,
                                                            double* part) {
  __shared__ double ss[kRedThreads], sc[kRedThreads];
  for (int r = 0; r < k; ++r) {
    const double* wr = w + (size_t)r * n;
    double s = 0.0, c = 0.0;
    for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += kRed * kRedThreads) {
      double t, e;
      two_sum(s, wr[i] * x[i], t, e);
      s = t; c += e;
    }
    ss[threadIdx.x] = s; sc[threadIdx.x] = c;
    __syncthreads();
    for (int h = kRedThreads / 2; h > 0; h >>= 1) {
      if (threadIdx.x < h) {
        double t, e;
        two_sum(ss[threadIdx.x], ss[threadIdx.x + h], t, e);
        ss[threadIdx.x] = t;
        sc[threadIdx.x] = sc[threadIdx.x] + sc[threadIdx.x + h] + e;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * ((size_t)r * kRed + blockIdx.x)] = ss[0]; part[2 * ((size_t)r * kRed + blockIdx.x) + 1] = sc[0]; }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kRed) k_probe_fin(int k, const double* part, double* out) {
  __shared__ double ss[kRed], sc[kRed];
  for (int r = 0; r < k; ++r) {
    ss[threadIdx.x] = part[2 * ((size_t)r * kRed + threadIdx.x)];
    sc[threadIdx.x] = part[2 * ((size_t)r * kRed + threadIdx.x) + 1];
    __syncthreads();
    for (int h = kRed / 2; h > 0; h >>= 1) {
      if (threadIdx.x < h) {
        double t, e;
        two_sum(ss[threadIdx.x], ss[threadIdx.x + h], t, e);
        ss[threadIdx.x] = t;
        sc[threadIdx.x] = sc[threadIdx.x] + sc[threadIdx.x + h] + e;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) out[r] = ss[0] + sc[0];
    __syncthreads();
  }
}

// ---- vector kernels (one thread per unknown)
__global__ void k_gather_sel(int m, const int* sel, const double* x, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] = x[sel[i]];
}
__global__ void k_copy(int n, const double* a, double* b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
__global__ void k_fillv(size_t n, double* a, double v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_add(int n, const double* d, double* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += d[i];
}
__global__ void k_scale(int n, double s, double* a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] *= s;
}
// DampedNewton::updateWeights_: w = reltol * max(|x_old|, |x|) + abstol
__global__ void k_weights(int n, double rel, double abs_, const double* xo, const double* x, double* w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = rel * fmax(fabs(xo[i]), fabs(x[i])) + abs_;
}
// GStepping::augmentResidual / augmentJacobian on the node-voltage rows
__global__ void k_gmin_rhs(int nv, const int* vrow, double g, const double* x, double* rhs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nv) rhs[vrow[k]] -= g * x[vrow[k]];
}
__global__ void k_gmin_jac(int nv, const int* vdiag, double g, double* J) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nv && vdiag[k] >= 0) J[vdiag[k]] += g;
}
// OneStep::obtainPredictor: xn0 = x_n + beta1 * xHistory[1] (+ beta2 * xHistory[2])
__global__ void k_predict(int n, int order, double b1, double b2, const double* x0, const double* x1,
                          const double* x2, double* xn0, double* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = x0[i];
  v += b1 * x1[i];
  if (order == 2) v += b2 * x2[i];
  xn0[i] = v;
  x[i] = v;
}
// OneStep::updateHistory (difference form), after completeStep set the order
__global__ void k_history(int n, int order, const double* x, const double* Q, const double* F, const double* B,
                          double* x0, double* x1, double* x2, double* q0, double* q1, double* q2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (order == 2) {
    x2[i] = x1[i];
    q2[i] = 1.0 * F[i] + -1.0 * B[i];
  }
  x1[i] = 1.0 * x[i] + -1.0 * x0[i];
  q1[i] = 1.0 * Q[i] + -1.0 * q0[i];
  x0[i] = x[i];
  q0[i] = Q[i];
}
// OneStep::initialize: xHistory[1] = 0, qHistory[1] = -(F - B) * h
__global__ void k_init_hist(int n, double h, const double* F0, const double* B0, double* x1, double* q1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x1[i] = 0.0;
  q1[i] = -(F0[i] - B0[i]) * h;
}

}  // namespace
}  // namespace xg

using namespace xg;

struct xg_tran {
  xg_circuit* c = nullptr;
  xg_lu* lu = nullptr;
  CircuitDev d;
  hipStream_t s = nullptr;
  int n = 0, nnz = 0;
  // device vectors
  DevArr<double> xn0, xH0, xH1, xH2, qH0, qH1, qH2, F0, B0, rhs, J, dx, wt, part, red, zeros;
  DevArr<int> vrow, vdiag;
  int nv = 0;
  // recording: probe rows (k x n) and selected unknowns, or whole solutions
  DevArr<double> probe_w, probe_part, rec_y;
  DevArr<int> sel;
  int probe_k = 0, n_sel = 0, rec_width = 0, rec_max = 0, rec_count = 0;
  std::vector<double> rec_t;
  std::vector<int> rec_order;
  // parameters
  double tstop = 0, tstep = 0, hmax = 0;
  double reltol = 1e-3, abstol = 1e-6, nl_reltol = 1e-2, nl_abstol = 1e-6;
  // integrator state (the oracle's / OneStep's locals)
  bool started = false, finished = false;
  int status = 0;
  std::vector<double> bps;
  size_t bpi = 0;
  double t = 0, h = 0, lastTimeStep = 0, currentTimeStepSum = 0, savedTimeStep = 0, estOverTol = 0;
  double psi[3] = {0, 0, 0};
  int order = 1, usedOrder = 1, nef = 0;
  long long numberOfSteps = 0;
  bool beginning = true;
  // DC continuation
  double gcond = 0.0, src_scale = 1.0;
  bool dc_initjct = true;
  int dc_stage = 0;
  // counters
  double steps = 0, rejects = 0, newton_iters = 0, loads = 0, dc_iters = 0, lu_fail = 0, seconds = 0;
  bool progress = false;   // a line on stderr per DC Newton solve
  bool uic = false;        // .TRAN ... UIC: no operating point, the run starts from x = 0

  ~xg_tran() {
    if (lu) xg_lu_destroy(lu);
    if (c) xg_circuit_destroy(c);
  }
};

namespace {

struct NLP {
  double reltol, abstol, dxtol, rhstol, smallUpdateTol;
  int maxstep;
  bool transient;
};
struct NewtonResult { int iters = 0; int status = 0; };

// reductions into red[slot] (device), no wait
template <int OP>
void reduce(xg_tran* T, int n, const double* a, const double* b, double s, int slot) {
  hipLaunchKernelGGL(k_red_part<OP>, dim3(kRed), dim3(kRedThreads), 0, T->s, n, a, b, s, T->part.p);
  hipLaunchKernelGGL(k_red_fin<OP>, dim3(1), dim3(kRed), 0, T->s, T->part.p, T->red.p + slot);
}

// the circuit's solver state for this load
int set_state(xg_tran* T, int dcop, int newton_iter, int init_jct, double hh, double time) {
  xg_solver_state ss{};
  ss.dcop = dcop; ss.tran = dcop ? 0 : 1; ss.init_jct = init_jct; ss.init_fix = 0;
  ss.tranop = dcop;   // the engine's operating points are a transient's
  ss.newton_iter = newton_iter; ss.order = T->order; ss.h = hh; ss.time = time;
  ss.tstep = T->tstep; ss.tstop = T->tstop;
  return xg_set_solver_state(T->c, &ss);
}

// rhs(it): load at x, residual; red[0] = sum RHS^2, red[1] = max |RHS|
int residual(xg_tran* T, bool dcop, const double* q0, const double* q2) {
  int st = circuit_load(T->c);
  if (st != XG_OK) return st;
  T->loads += 1;
  const int n = T->n;
  if (dcop && T->src_scale != 1.0) hipLaunchKernelGGL(k_scale, dim3(blocks(n)), dim3(256), 0, T->s, n, T->src_scale, T->d.B);
  st = circuit_combine(T->c, q0 ? q0 : T->zeros.p, q2 ? q2 : T->zeros.p, T->rhs.p, T->J.p);
  if (st != XG_OK) return st;
  if (dcop && T->gcond != 0.0) {
    hipLaunchKernelGGL(k_gmin_rhs, dim3(blocks(T->nv)), dim3(256), 0, T->s, T->nv, T->vrow.p, T->gcond, T->d.x, T->rhs.p);
    hipLaunchKernelGGL(k_gmin_jac, dim3(blocks(T->nv)), dim3(256), 0, T->s, T->nv, T->vdiag.p, T->gcond, T->J.p);
  }
  reduce<RED_SUMSQ>(T, n, T->rhs.p, nullptr, 0.0, 0);
  reduce<RED_MAXABS>(T, n, T->rhs.p, nullptr, 0.0, 1);
  HIPCHK(hipGetLastError());
  return XG_OK;
}

// DampedNewton::solve with full steps from the x in the circuit's buffer;
// xcurr = the last accepted solution (transient) or zeros (DC)
int newton(xg_tran* T, bool dcop, double hh, double time, const NLP& p, const double* xcurr, const double* q0,
           const double* q2, NewtonResult* out) {
  const int n = T->n;
  hipStream_t s = T->s;
  double host[4];
  int lim = 0;
  auto fetch = [&](int cnt) -> int {
    HIPCHK(hipMemcpyAsync(host, T->red.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&lim, T->d.limited, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return XG_OK;
  };
  int st = set_state(T, dcop, 0, dcop && T->dc_initjct ? 1 : 0, hh, time);
  if (st != XG_OK) return st;
  st = residual(T, dcop, q0, q2);
  if (st != XG_OK) return st;
  if (p.transient)
    hipLaunchKernelGGL(k_weights, dim3(blocks(n)), dim3(256), 0, s, n, p.reltol, p.abstol, xcurr, T->d.x, T->wt.p);
  if ((st = fetch(2)) != XG_OK) return st;
  double normRHS = std::sqrt(host[0]);
  double normRHS_init = normRHS, normRHS_old = normRHS;
  int nlStep = 0, count = 0;
  double tmpConvRate = 1.0;
  const double eps = std::numeric_limits<double>::epsilon();
  const double convRateMax = 0.5 * std::numeric_limits<double>::max();
  for (;;) {
    nlStep++;
    // the Jacobian was combined with the residual at this x
    HIPCHK(hipStreamSynchronize(s));
    int ls = xg_lu_numeric_dev(T->lu, T->J.p);
    if (ls == XG_OK) {
      ls = xg_lu_solve_dev(T->lu, T->rhs.p, T->dx.p, 0);
    } else if (ls > 0) {
      hipLaunchKernelGGL(k_fillv, dim3(blocks(n)), dim3(256), 0, s, (size_t)n, T->dx.p, 0.0);
      T->lu_fail += 1;
    } else {
      return ls;
    }
    if (ls < 0 && ls != XG_E_NAN) return ls;
    hipLaunchKernelGGL(k_add, dim3(blocks(n)), dim3(256), 0, s, n, T->dx.p, T->d.x);
    if ((st = set_state(T, dcop, nlStep, 0, hh, time)) != XG_OK) return st;
    if ((st = residual(T, dcop, q0, q2)) != XG_OK) return st;
    if (!p.transient)
      hipLaunchKernelGGL(k_weights, dim3(blocks(n)), dim3(256), 0, s, n, p.reltol, p.abstol, T->zeros.p, T->d.x,
                         T->wt.p);
    reduce<RED_MAXRATIO>(T, n, T->dx.p, T->wt.p, 0.0, 2);
    if ((st = fetch(3)) != XG_OK) return st;
    normRHS = std::sqrt(host[0]);
    const bool nanRHS = !std::isfinite(normRHS);
    int status = 0;
    if (ls != 0) status = (ls == XG_E_NAN) ? -6 : -9;
    else if (lim && nlStep < p.maxstep) continue;
    else if (lim) status = -1;
    else if (nanRHS) status = -6;
    else {
      const double maxNormRHS = host[1], wtNormDX = host[2];
      if (normRHS < eps) status = 1;
      else {
        const double normRHS_rel = normRHS / normRHS_init;
        const double resConvRate = normRHS / normRHS_old;
        normRHS_old = normRHS;
        if (maxNormRHS <= p.rhstol && wtNormDX <= p.dxtol) status = 2;
        else if (nlStep >= p.maxstep && p.transient && normRHS_rel <= 0.9 && resConvRate <= 1.0) status = -3;
        else if (wtNormDX <= p.smallUpdateTol) status = 4;
        else if (nlStep >= p.maxstep) status = -1;
        else if (resConvRate > convRateMax) status = -2;
        else {
          if (p.transient && std::fabs(resConvRate - 1.0) <= 1.0e-3) {
            if (count == 0 || resConvRate < tmpConvRate) tmpConvRate = resConvRate;
            count++;
          } else {
            count = 0;
          }
          if (p.transient && count == 5) status = -3;
        }
      }
    }
    if (status != 0) {
      out->iters = nlStep;
      out->status = status;
      return XG_OK;
    }
  }
}

// One LOCA natural-continuation run (gmin or source stepping) from zero.
template <class SetParam>
int dc_continuation(xg_tran* T, double p0, double p1, double step, double min_step, double max_step, double aggr,
                    bool init_jct, SetParam set_param, int* ok) {
  const int n = T->n;
  const NLP p{1e-3, 1e-12, 1.0, 1e-6, 1e-6, 20, false};
  hipLaunchKernelGGL(k_fillv, dim3(blocks(n + 1)), dim3(256), 0, T->s, (size_t)(n + 1), T->d.x, 0.0);
  int st = circuit_history(T->c, 0);
  if (st != XG_OK) return st;
  hipLaunchKernelGGL(k_copy, dim3(blocks(n)), dim3(256), 0, T->s, n, T->d.x, T->xH0.p);   // x_ok
  double par = p0, par_ok = p0;
  bool first = true, done = false;
  for (int k = 0; k < 400 && !done; ++k) {
    set_param(par);
    T->dc_initjct = first && init_jct;
    NewtonResult r;
    if ((st = newton(T, true, 0.0, 0.0, p, T->zeros.p, nullptr, nullptr, &r)) != XG_OK) return st;
    T->dc_iters += r.iters;
    if (T->progress)
      std::fprintf(stderr, "xg_tran DC %s stepping: parameter %.6g, Newton status %d after %d iterations\n",
                   T->dc_stage == 1 ? "gmin" : "source", par, r.status, r.iters);
    if (r.status > 0) {
      first = false;
      par_ok = par;
      hipLaunchKernelGGL(k_copy, dim3(blocks(n)), dim3(256), 0, T->s, n, T->d.x, T->xH0.p);
      if ((st = circuit_history(T->c, 1)) != XG_OK) return st;
      if (par == p1) { done = true; break; }
      const double rr = r.iters / 20.0;
      step *= 1.0 + aggr * (1.0 - rr * rr);
      if (std::fabs(step) > max_step) step = step > 0 ? max_step : -max_step;
    } else {
      hipLaunchKernelGGL(k_copy, dim3(blocks(n)), dim3(256), 0, T->s, n, T->xH0.p, T->d.x);
      if ((st = circuit_history(T->c, 2)) != XG_OK) return st;
      step *= 0.5;
      if (std::fabs(step) < min_step) break;
    }
    par = par_ok + step;
    if ((p1 - par) * (p1 - p0) <= 0.0) par = p1;
  }
  T->dc_initjct = true;
  set_param(p1);
  *ok = done ? 1 : -1;
  return XG_OK;
}

// DC operating point (the oracle's dcop): Newton from initJct, then gmin
// stepping, then source stepping
int dcop(xg_tran* T, int* ok) {
  const int n = T->n;
  const NLP p{1e-3, 1e-12, 1.0, 1e-6, 1e-6, 200, false};
  hipLaunchKernelGGL(k_fillv, dim3(blocks(n + 1)), dim3(256), 0, T->s, (size_t)(n + 1), T->d.x, 0.0);
  int st = circuit_history(T->c, 0);
  if (st != XG_OK) return st;
  T->gcond = 0.0; T->src_scale = 1.0; T->dc_initjct = true; T->dc_stage = 0;
  NewtonResult r;
  if ((st = newton(T, true, 0.0, 0.0, p, T->zeros.p, nullptr, nullptr, &r)) != XG_OK) return st;
  T->dc_iters += r.iters;
  if (T->progress)
    std::fprintf(stderr, "xg_tran DC Newton: status %d after %d iterations\n", r.status, r.iters);
  if (r.status > 0) { *ok = 1; return XG_OK; }
  int good = -1;
  T->dc_stage = 1;
  st = dc_continuation(T, 4.0, -4.0, -2.0, 1e-6, 1e12, 0.01, true,
                       [&](double q) { T->gcond = std::pow(10.0, q) - std::pow(10.0, -4.0); }, &good);
  T->gcond = 0.0;
  if (st != XG_OK) return st;
  if (good > 0) { *ok = 1; return XG_OK; }
  T->dc_stage = 2;
  st = dc_continuation(T, 0.0, 1.0, 0.2, 1e-4, 0.2, 1.0, false, [&](double q) { T->src_scale = q; }, &good);
  T->src_scale = 1.0;
  if (st != XG_OK) return st;
  *ok = good;
  return XG_OK;
}

int record(xg_tran* T, double tt, int ord) {
  if (T->rec_count >= T->rec_max) return XG_OK;
  const int n = T->n;
  double* row = T->rec_y.p + (size_t)T->rec_count * T->rec_width;
  if (T->probe_k == 0 && T->n_sel == 0) {
    hipLaunchKernelGGL(k_copy, dim3(blocks(n)), dim3(256), 0, T->s, n, T->xH0.p, row);
  } else {
    if (T->probe_k > 0) {
      hipLaunchKernelGGL(k_probe_part, dim3(kRed), dim3(kRedThreads), 0, T->s, n, T->probe_k, T->probe_w.p, T->xH0.p,
                         T->probe_part.p);
      hipLaunchKernelGGL(k_probe_fin, dim3(1), dim3(kRed), 0, T->s, T->probe_k, T->probe_part.p, row);
    }
    if (T->n_sel > 0)
      hipLaunchKernelGGL(k_gather_sel, dim3(blocks(T->n_sel)), dim3(256), 0, T->s, T->n_sel, T->sel.p, T->xH0.p,
                         row + T->probe_k);
  }
  HIPCHK(hipGetLastError());
  T->rec_t.push_back(tt);
  T->rec_order.push_back(ord);
  T->rec_count++;
  return XG_OK;
}

double next_bp(xg_tran* T) {
  while (T->bpi < T->bps.size() && T->bps[T->bpi] <= T->t * (1 + 1e-13) + 1e-30) ++T->bpi;
  return T->bpi < T->bps.size() ? T->bps[T->bpi] : T->tstop;
}

// OneStep::initialize (the oracle's `initialize`)
void initialize(xg_tran* T, bool initial) {
  const double hmin = 1e-18, h0_max_factor = 0.005;
  const double stopTime = next_bp(T);
  const double time_to_stop = stopTime - T->t;
  double hh = initial ? h0_max_factor * std::fabs(time_to_stop) : 0.1 * std::min(T->savedTimeStep, std::fabs(time_to_stop));
  if (T->tstep > 0.0 && initial) hh = std::min(T->tstep, hh);
  if (T->hmax > 0 && hh > T->hmax) hh = T->hmax;
  hh = std::max(hh, hmin);
  T->h = hh;
  T->lastTimeStep = T->h;
  T->currentTimeStepSum = 2.0 * T->h;
  hipLaunchKernelGGL(k_init_hist, dim3(blocks(T->n)), dim3(256), 0, T->s, T->n, T->h, T->F0.p, T->B0.p, T->xH1.p,
                     T->qH1.p);
  T->numberOfSteps = 0; T->order = 1; T->usedOrder = 1; T->psi[0] = T->h;
  T->beginning = true;
}

// .TRAN ... UIC / NOOP (N_ANP_Transient.C:879-1000): no operating point; the
// solution starts at zero (no .IC) and one transient load there
// (initializeProblem, loadRHS) gives the devices' state and F, Q, B, from
// which the histories are set as after an operating point
int uic_point(xg_tran* T) {
  const int n = T->n;
  hipLaunchKernelGGL(k_fillv, dim3(blocks(n + 1)), dim3(256), 0, T->s, (size_t)(n + 1), T->d.x, 0.0);
  int st = circuit_history(T->c, 0);
  if (st != XG_OK) return st;
  if ((st = set_state(T, 0, 0, 0, T->tstep, 0.0)) != XG_OK) return st;
  if ((st = circuit_load(T->c)) != XG_OK) return st;
  T->loads += 1;
  T->dc_stage = -1;
  return XG_OK;
}

// DC point (or the UIC start), then the transient's start (the oracle's
// tran() before its loop)
int start(xg_tran* T) {
  int st;
  if (T->uic) {
    if ((st = uic_point(T)) != XG_OK) return st;
  } else {
    int ok = 0;
    st = dcop(T, &ok);
    if (st != XG_OK) return st;
    if (ok <= 0) { T->status = -100; T->finished = true; return XG_OK; }
  }
  const int n = T->n;
  if ((st = circuit_history(T->c, 1)) != XG_OK) return st;
  hipLaunchKernelGGL(k_copy, dim3(blocks(n)), dim3(256), 0, T->s, n, T->d.x, T->xH0.p);
  hipLaunchKernelGGL(k_copy, dim3(blocks(n)), dim3(256), 0, T->s, n, T->d.Q, T->qH0.p);
  hipLaunchKernelGGL(k_copy, dim3(blocks(n)), dim3(256), 0, T->s, n, T->d.F, T->F0.p);
  hipLaunchKernelGGL(k_copy, dim3(blocks(n)), dim3(256), 0, T->s, n, T->d.B, T->B0.p);
  hipLaunchKernelGGL(k_fillv, dim3(blocks(n)), dim3(256), 0, T->s, (size_t)n, T->xH2.p, 0.0);
  hipLaunchKernelGGL(k_fillv, dim3(blocks(n)), dim3(256), 0, T->s, (size_t)n, T->qH2.p, 0.0);
  T->t = 0.0;
  T->bpi = 0;
  initialize(T, true);
  if ((st = record(T, 0.0, 0)) != XG_OK) return st;
  T->started = true;
  return XG_OK;
}

// One attempted step of the transient loop (the oracle's loop body)
int step(xg_tran* T) {
  const double r_min = 0.25, r_max = 0.9, r_hincr = 2.0, r_hincr_test = 2.0, tolAimFac = 0.5, hmin = 1e-18;
  const int n = T->n;
  hipStream_t s = T->s;
  double stopTime = next_bp(T);
  if (T->t + T->h > stopTime) { T->savedTimeStep = T->h; T->h = stopTime - T->t; T->currentTimeStepSum = T->h + T->lastTimeStep; }
  // updateCoeffs
  const double temp1 = T->h;
  if (T->order == 2) T->psi[2] = T->psi[1];
  T->psi[1] = T->psi[0]; T->psi[0] = temp1;
  double beta1, beta2 = 0.0, ck;
  if (T->order == 2) {
    const double temp2 = T->psi[1];
    beta1 = temp1 / temp2 + (temp1 / temp2) * (temp1 / temp2) / 2;
    beta2 = -1.0 * temp1 * temp1 / temp2 / T->psi[2] / 2;
    ck = (T->h / T->currentTimeStepSum) / 3.0;
  } else {
    beta1 = temp1 / T->psi[1];
    ck = T->h / T->currentTimeStepSum;
  }
  hipLaunchKernelGGL(k_predict, dim3(blocks(n)), dim3(256), 0, s, n, T->order, beta1, beta2, T->xH0.p, T->xH1.p,
                     T->xH2.p, T->xn0.p, T->d.x);
  // error weights: reltol * max_i |x_n,i| + abstol (one value for every unknown)
  reduce<RED_MAXABS>(T, n, T->xH0.p, nullptr, 0.0, 3);
  double mx = 0.0;
  HIPCHK(hipMemcpyAsync(&mx, T->red.p + 3, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const double ew = T->reltol * mx + T->abstol;
  const NLP p{T->nl_reltol, T->nl_abstol, 0.33, 1e-2, 1e-6, 20, true};
  NewtonResult ns;
  int st = newton(T, false, T->h, T->t + T->h, p, T->xH0.p, T->qH0.p, T->qH2.p, &ns);
  if (st != XG_OK) return st;
  T->newton_iters += ns.iters;
  bool ok = ns.status > 0;
  if (ok && !T->beginning) {
    reduce<RED_SUMSQ_DIFF_OVER>(T, n, T->d.x, T->xn0.p, ew, 4);
    double ssq = 0.0;
    HIPCHK(hipMemcpyAsync(&ssq, T->red.p + 4, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    T->estOverTol = ck * std::sqrt(ssq / n);
    ok = T->estOverTol <= 1.0;
  }
  if (ok) {
    // completeStep
    T->numberOfSteps++; T->nef = 0; T->steps += 1;
    const double tnew = T->t + T->h;
    double newh = T->h;
    T->lastTimeStep = T->h; T->usedOrder = T->order;
    double rr = tolAimFac / (T->estOverTol + 0.0001);
    rr = std::pow(rr, 1.0 / (T->order + 1.0));
    if (T->numberOfSteps >= 2 && T->order == 1) {
      T->order = 2;
      rr = tolAimFac / (T->estOverTol + 0.0001);
      rr = std::pow(rr, 1.0 / (T->order + 1.0));
      if (rr <= 1.05) T->order = 1;
    }
    if (rr >= r_hincr_test) { rr = r_hincr; newh = rr * T->h; }
    else if (rr <= 1) { rr = std::max(r_min, std::min(r_max, rr)); newh = rr * T->h; }
    hipLaunchKernelGGL(k_history, dim3(blocks(n)), dim3(256), 0, s, n, T->order, T->d.x, T->d.Q, T->d.F, T->d.B,
                       T->xH0.p, T->xH1.p, T->xH2.p, T->qH0.p, T->qH1.p, T->qH2.p);
    if ((st = circuit_history(T->c, 1)) != XG_OK) return st;
    T->t = tnew;
    if (T->t <= T->tstop * (1 + 1e-12) && (st = record(T, T->t, T->usedOrder)) != XG_OK) return st;
    newh = std::max(newh, hmin);
    if (T->hmax > 0) newh = std::min(newh, T->hmax);
    T->beginning = false;
    if (T->bpi < T->bps.size() &&
        std::fabs(T->t - T->bps[T->bpi]) <= 1e-13 * std::max(1.0, std::fabs(T->t)) + 1e-24 &&
        T->t < T->tstop * (1 - 1e-12)) {
      // a breakpoint: restart at order 1 from this point's F and B
      hipLaunchKernelGGL(k_copy, dim3(blocks(n)), dim3(256), 0, s, n, T->d.F, T->F0.p);
      hipLaunchKernelGGL(k_copy, dim3(blocks(n)), dim3(256), 0, s, n, T->d.B, T->B0.p);
      T->savedTimeStep = newh;
      ++T->bpi;
      initialize(T, false);
    } else {
      T->currentTimeStepSum = newh + T->lastTimeStep;
      T->h = newh;
    }
    if (T->t >= T->tstop * (1 - 1e-12)) T->finished = true;
  } else {
    // rejectStep
    T->rejects += 1;
    T->nef++;
    if (T->nef >= 15) { T->status = -101; T->finished = true; return XG_OK; }
    for (int i = 1; i <= T->order; ++i) T->psi[i - 1] = T->psi[i];
    double newh;
    if (ns.status <= 0) { newh = T->h / 8; T->order = 1; }
    else if (T->nef == 1) {
      T->order = 1;
      double rr = tolAimFac / (T->estOverTol + 0.0001);
      rr = std::pow(rr, 1.0 / (T->order + 1.0));
      rr = std::max(r_min, std::min(r_max, rr));
      newh = rr * T->h;
    } else { newh = r_min * T->h; T->order = 1; }
    newh = std::max(newh, hmin);
    if (T->hmax > 0) newh = std::min(newh, T->hmax);
    if (newh <= hmin) { T->status = -102; T->finished = true; return XG_OK; }
    T->currentTimeStepSum = newh + T->lastTimeStep;
    T->h = newh;
    if ((st = circuit_history(T->c, 2)) != XG_OK) return st;
  }
  return XG_OK;
}

}  // namespace

extern "C" {

int xg_tran_create(const xg_circuit_desc* desc, double tstop, double tstep, double hmax, int lu_method,
                   xg_tran** out) {
  if (!desc || !out || !(tstop > 0.0) || !(tstep > 0.0) || !(hmax >= 0.0) || lu_method < 0 || lu_method > 2)
    return XG_E_ARG;
  auto T = std::make_unique<xg_tran>();
  int st = xg_circuit_create(desc, &T->c);
  if (st != XG_OK) return st;
  T->d = circuit_dev(T->c);
  T->s = (hipStream_t)T->d.stream;
  T->n = T->d.n;
  T->nnz = T->d.nnz;
  const HostCircuit& hc = circuit_host(T->c);
  st = xg_lu_symbolic_ex(T->n, hc.rowptr.data(), hc.colind.data(), nullptr, lu_method, &T->lu);
  if (st != XG_OK) return st;
  T->tstop = tstop; T->tstep = tstep; T->hmax = hmax;
  const size_t n = T->n;
  for (auto* a : {&T->xn0, &T->xH0, &T->xH1, &T->xH2, &T->qH0, &T->qH1, &T->qH2, &T->F0, &T->B0, &T->rhs, &T->dx,
                  &T->wt, &T->zeros})
    HIPCHK(a->alloc(n + 1));
  HIPCHK(T->J.alloc(std::max<size_t>(1, T->nnz)));
  HIPCHK(T->part.alloc(kRed));
  HIPCHK(T->red.alloc(8));
  hipLaunchKernelGGL(k_fillv, dim3(blocks(n + 1)), dim3(256), 0, T->s, n + 1, T->zeros.p, 0.0);
  // node-voltage rows and their diagonal slots (gmin stepping)
  std::vector<int> vr, vd;
  for (size_t k = 1; k < hc.node_gid.size(); ++k) {
    const int r = hc.node_gid[k];
    if (r < 0) continue;
    int slot = -1;
    for (int q = hc.rowptr[r]; q < hc.rowptr[r + 1]; ++q)
      if (hc.colind[q] == r) slot = q;
    vr.push_back(r);
    vd.push_back(slot);
  }
  T->nv = (int)vr.size();
  if (T->nv) {
    HIPCHK(T->vrow.alloc(vr.size())); HIPCHK(T->vdiag.alloc(vd.size()));
    HIPCHK(hipMemcpyAsync(T->vrow.p, vr.data(), vr.size() * sizeof(int), hipMemcpyHostToDevice, T->s));
    HIPCHK(hipMemcpyAsync(T->vdiag.p, vd.data(), vd.size() * sizeof(int), hipMemcpyHostToDevice, T->s));
  }
  // breakpoints of every independent source, and tstop
  for (const auto& e : hc.srcs) wave_breakpoints(hc, e.wave, tstep, tstop, T->bps);
  T->bps.push_back(tstop);
  std::sort(T->bps.begin(), T->bps.end());
  T->bps.erase(std::unique(T->bps.begin(), T->bps.end()), T->bps.end());
  HIPCHK(hipStreamSynchronize(T->s));
  *out = T.release();
  return XG_OK;
}

int xg_tran_destroy(xg_tran* T) {
  if (!T) return XG_E_ARG;
  (void)hipStreamSynchronize(T->s);
  delete T;
  return XG_OK;
}

int xg_tran_set_precision(xg_tran* T, int precision, double ir_tol, int ir_max) {
  if (!T) return XG_E_ARG;
  return xg_lu_set_precision(T->lu, precision, ir_tol, ir_max);
}

int xg_tran_set_tolerances(xg_tran* T, double reltol, double abstol, double nl_reltol, double nl_abstol) {
  if (!T || T->started || !(reltol > 0) || !(abstol > 0) || !(nl_reltol > 0) || !(nl_abstol > 0)) return XG_E_ARG;
  T->reltol = reltol; T->abstol = abstol; T->nl_reltol = nl_reltol; T->nl_abstol = nl_abstol;
  return XG_OK;
}

int xg_tran_set_record(xg_tran* T, int k, const double* w, int n_sel, const int* sel, int record_max) {
  if (!T || T->started || k < 0 || n_sel < 0 || record_max < 0 || (k > 0 && !w) || (n_sel > 0 && !sel))
    return XG_E_ARG;
  const size_t n = T->n;
  for (int i = 0; i < n_sel; ++i)
    if (sel[i] < 0 || sel[i] >= T->n) return XG_E_ARG;
  T->probe_k = k;
  T->n_sel = n_sel;
  T->rec_width = (k == 0 && n_sel == 0) ? (int)n : k + n_sel;
  T->rec_max = record_max;
  if (k > 0) {
    HIPCHK(T->probe_w.alloc((size_t)k * n));
    HIPCHK(hipMemcpyAsync(T->probe_w.p, w, sizeof(double) * k * n, hipMemcpyHostToDevice, T->s));
    HIPCHK(T->probe_part.alloc((size_t)2 * k * kRed));
  }
  if (n_sel > 0) {
    HIPCHK(T->sel.alloc(n_sel));
    HIPCHK(hipMemcpyAsync(T->sel.p, sel, sizeof(int) * n_sel, hipMemcpyHostToDevice, T->s));
  }
  HIPCHK(T->rec_y.alloc(std::max<size_t>(1, (size_t)record_max * T->rec_width)));
  HIPCHK(hipStreamSynchronize(T->s));
  return XG_OK;
}

int xg_tran_set_uic(xg_tran* T, int on) {
  if (!T || T->started) return XG_E_ARG;
  T->uic = on != 0;
  return XG_OK;
}

int xg_tran_set_progress(xg_tran* T, int on) {
  if (!T) return XG_E_ARG;
  T->progress = on != 0;
  return XG_OK;
}

int xg_tran_run(xg_tran* T, double t_until, long long max_steps, double* stats) {
  if (!T || max_steps < 0) return XG_E_ARG;
  auto t0 = std::chrono::steady_clock::now();
  int st = XG_OK;
  if (!T->started && !T->finished) st = start(T);
  const double s0 = T->steps + T->rejects;
  while (st == XG_OK && !T->finished && T->t < t_until * (1 - 1e-12) && T->steps + T->rejects - s0 < max_steps)
    st = step(T);
  T->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    double lu[7];
    xg_lu_precision_stats(T->lu, lu);
    const double v[14] = {T->steps, T->rejects, T->newton_iters, T->loads, T->t, (double)T->status,
                          T->finished ? 1.0 : 0.0, T->dc_iters, (double)T->dc_stage, T->seconds, T->lu_fail,
                          lu[3], lu[4], (double)T->rec_count};
    std::memcpy(stats, v, sizeof(v));
  }
  return st;
}

int xg_tran_points(xg_tran* T, int max_points, double* t, int* order, double* y) {
  if (!T || max_points < 0) return XG_E_ARG;
  const int m = std::min(max_points, T->rec_count);
  const size_t width = T->rec_width;
  if (t) std::copy(T->rec_t.begin(), T->rec_t.begin() + m, t);
  if (order) std::copy(T->rec_order.begin(), T->rec_order.begin() + m, order);
  if (y && m > 0) {
    HIPCHK(hipMemcpyAsync(y, T->rec_y.p, sizeof(double) * width * m, hipMemcpyDeviceToHost, T->s));
    HIPCHK(hipStreamSynchronize(T->s));
  }
  return m;
}

int xg_tran_solution(xg_tran* T, double* x) {
  if (!T || !x) return XG_E_ARG;
  HIPCHK(hipMemcpyAsync(x, T->xH0.p, sizeof(double) * T->n, hipMemcpyDeviceToHost, T->s));
  HIPCHK(hipStreamSynchronize(T->s));
  return XG_OK;
}

}  // extern "C"
