// CPU ORACLE — test infrastructure only (see oracle.h).
//
// Loader, residual/Jacobian assembly, KLU-semantics LU, DampedNewton, OneStep
// time integration, and the extern "C" surface the tests load with ctypes.
#include "oracle.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <thread>
#include <atomic>
#include <iterator>

namespace orc {
void qmeyer(double, double, double, double, double, double&, double&, double&, double, double);
double limvds(double, double);
double pnjlim(double, double, double, double, int*);
double fetlim(double, double, double);
}

namespace orc {

// mos1 kernels live in oracle.cpp; declared here to keep the files independent
void mos1_update_ext(Mos1& mi, const SolverState& ss, const Options& opt, const double* x,
                     const double* staCur, const double* stoCur, double* staNext, double* stoNext);
void mos1_vectors_ext(const Mos1& mi, const SolverState& ss, const Options& opt,
                      double* F, double* Q, double* dFdxdVp, double* dQdxdVp);
void mos1_matrices_ext(const Mos1& mi, const SolverState& ss, double* dFdx, double* dQdx);

// KLU-style sparse LU: minimum-degree column order on the pattern of A+A^T
// (standing in for AMD), then left-looking Gilbert-Peierls factorization with
// threshold partial pivoting (KLU's default tol 0.001, the diagonal preferred
// when it passes).  The symbolic order is computed once per pattern and
// reused, as Xyce reuses KLU's analysis.  Parity is on the solution (pinned
// against numpy / scipy in the tests); KLU's own pivot sequence and factor
// bits: parity unpinned (KLU, in Trilinos Amesos / SuiteSparse, is absent).
struct SparseLU {
  int n = 0;
  std::vector<int> q;          // column order
  std::vector<int> Lp, Li, Up, Ui;
  std::vector<double> Lx, Ux;
  std::vector<int> pinv;       // row -> pivot position
  std::vector<int> a_cp, a_ri, a_pos;   // column-wise copy of the pattern (positions into CSR values)
  bool analysed = false;

  void analyse(int nn, const int* rowptr, const int* colind) {
    n = nn;
    std::vector<std::vector<int>> g(n);
    for (int r = 0; r < n; ++r)
      for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
        int c = colind[k];
        if (c != r) { g[r].push_back(c); g[c].push_back(r); }
      }
    for (auto& v : g) { std::sort(v.begin(), v.end()); v.erase(std::unique(v.begin(), v.end()), v.end()); }
    // minimum degree on the explicit elimination graph (ties: lowest index)
    std::vector<char> gone(n, 0);
    q.clear();
    for (int it = 0; it < n; ++it) {
      int best = -1; size_t bd = (size_t)-1;
      for (int v = 0; v < n; ++v) if (!gone[v] && g[v].size() < bd) { bd = g[v].size(); best = v; }
      q.push_back(best); gone[best] = 1;
      std::vector<int> nb = g[best];
      for (int a : nb) {
        auto& ga = g[a];
        ga.erase(std::remove(ga.begin(), ga.end(), best), ga.end());
        std::vector<int> m; std::set_union(ga.begin(), ga.end(), nb.begin(), nb.end(), std::back_inserter(m));
        m.erase(std::remove(m.begin(), m.end(), a), m.end());
        ga.swap(m);
      }
      g[best].clear();
    }
    a_cp.assign(n + 1, 0);
    for (int r = 0; r < n; ++r) for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) a_cp[colind[k] + 1]++;
    for (int c = 0; c < n; ++c) a_cp[c + 1] += a_cp[c];
    a_ri.resize(a_cp[n]); a_pos.resize(a_cp[n]);
    std::vector<int> f(a_cp.begin(), a_cp.end() - 1);
    for (int r = 0; r < n; ++r) for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) { int c = colind[k]; a_ri[f[c]] = r; a_pos[f[c]++] = k; }
    analysed = true;
  }

  // returns 0 ok, k+1 if pivot k is zero
  int factor(const double* vals) {
    const double tol = 0.001;
    pinv.assign(n, -1);
    Lp.assign(1, 0); Up.assign(1, 0); Li.clear(); Lx.clear(); Ui.clear(); Ux.clear();
    std::vector<double> xw(n, 0.0);
    std::vector<int> mark(n, -1), stack, topo, pstack;
    for (int k = 0; k < n; ++k) {
      const int col = q[k];
      // symbolic reach of column col of A in the graph of L (DFS), topological order
      topo.clear();
      for (int p = a_cp[col]; p < a_cp[col + 1]; ++p) {
        int r = a_ri[p];
        if (mark[r] == k) continue;
        // iterative DFS from r
        stack.assign(1, r); pstack.assign(1, 0); mark[r] = k;
        while (!stack.empty()) {
          int j = stack.back();
          int jj = pinv[j];
          bool pushed = false;
          if (jj >= 0) {
            for (int t = Lp[jj] + pstack.back(); t < Lp[jj + 1]; ++t) {
              ++pstack.back();
              int i = Li[t];
              if (mark[i] != k) { mark[i] = k; stack.push_back(i); pstack.push_back(0); pushed = true; break; }
            }
          }
          if (!pushed) { topo.push_back(j); stack.pop_back(); pstack.pop_back(); }
        }
      }
      for (int p = a_cp[col]; p < a_cp[col + 1]; ++p) xw[a_ri[p]] += vals[a_pos[p]];
      // numeric: for rows already pivoted, in reverse topological order
      for (int t = (int)topo.size() - 1; t >= 0; --t) {
        int j = topo[t], jj = pinv[j];
        if (jj < 0) continue;
        double xj = xw[j];
        for (int s = Lp[jj]; s < Lp[jj + 1]; ++s) xw[Li[s]] -= Lx[s] * xj;
      }
      // pivot among non-pivoted rows; prefer the diagonal row of this column
      int piv = -1; double amax = 0.0;
      for (int j : topo) if (pinv[j] < 0 && std::fabs(xw[j]) > amax) { amax = std::fabs(xw[j]); piv = j; }
      if (piv < 0 || !(amax > 0.0) || !std::isfinite(amax)) { for (int j : topo) xw[j] = 0.0; return k + 1; }
      if (pinv[col] < 0 && mark[col] == k && std::fabs(xw[col]) >= tol * amax) piv = col;
      const double pv = xw[piv];
      for (int j : topo) {
        if (pinv[j] >= 0) { Ui.push_back(pinv[j]); Ux.push_back(xw[j]); }
      }
      Ui.push_back(k); Ux.push_back(pv);
      Up.push_back((int)Ui.size());
      pinv[piv] = k;
      for (int j : topo) {
        if (pinv[j] < 0) { Li.push_back(j); Lx.push_back(xw[j] / pv); }
        xw[j] = 0.0;
      }
      Lp.push_back((int)Li.size());
    }
    return 0;
  }

  // solve A x = b; L rows are original row ids, U rows pivot positions
  void solve(const double* b, double* x) const {
    std::vector<double> y(b, b + n), z(n);
    // forward: for k, z[k] = y[row pivoted at k]; then eliminate
    std::vector<int> prow(n);
    for (int r = 0; r < n; ++r) prow[pinv[r]] = r;
    for (int k = 0; k < n; ++k) {
      double zk = y[prow[k]];
      z[k] = zk;
      for (int s = Lp[k]; s < Lp[k + 1]; ++s) y[Li[s]] -= Lx[s] * zk;
    }
    // backward with U stored by columns (U(:,k) entries at pivot rows < k, diag last)
    for (int k = n - 1; k >= 0; --k) {
      double d = Ux[Up[k + 1] - 1];
      double xk = z[k] / d;
      z[k] = xk;
      for (int s = Up[k]; s < Up[k + 1] - 1; ++s) z[Ui[s]] -= Ux[s] * xk;
    }
    for (int k = 0; k < n; ++k) x[q[k]] = z[k];
  }
};

// Hooks through which a caller's device loads and linear solver stand in for
// the oracle's own, so that this Newton / time-integration loop (the
// restatement of the reference's driver) drives them: the arguments are
// those of orc_load and orc_lu_solve.
typedef int (*OrcLoadHook)(void* ctx, const int* ssi, const double* ssd, const double* x, const double* sta_cur,
                           const double* sto_cur, double* sta_next, double* sto_next, double* F, double* Q, double* B,
                           double* dFdxdVp, double* dQdxdVp, double* dFdx, double* dQdx);
typedef int (*OrcSolveHook)(void* ctx, int n, const int* rowptr, const int* colind, const double* vals,
                            const double* b, double* x);

struct Work {
  Circuit* c;
  OrcLoadHook load_hook = nullptr;
  OrcSolveHook solve_hook = nullptr;
  void* hook_ctx = nullptr;
  int n, nnz;
  std::vector<double> x, F, Q, B, dFdxdVp, dQdxdVp, RHS, dFdx, dQdx, J;
  std::vector<double> staCur, staNext, stoCur, stoNext;
  int limited = 0;
  double time = 0, tstep = 0, tstop = 0, bptol = 1e-18;
  // transient tolerances: local error (TIA RELTOL/ABSTOL, N_TIA_TIAParams.C)
  // and Newton weights (NONLIN-TRAN RELTOL/ABSTOL, N_NLS_NLParams.C); the
  // defaults are the reference's
  double tia_reltol = 1e-3, tia_abstol = 1e-6, nl_reltol = 1e-2, nl_abstol = 1e-6;
  SparseLU slu;
  // recorded quantities: empty = the whole solution at each accepted point;
  // else probe_k rows of probe_w (row-major, n each): the products w_r . x
  int probe_k = 0;
  std::vector<double> probe_w;
  int method = 0;   // 0 = OneStep trapezoid, 1 = Gear12 (.OPTIONS TIMEINT METHOD=GEAR)
  // DC continuation (Interface::spiceStrategy, N_NLS_NOX_Interface.C:321-437):
  // conductance added to every node-voltage row (GStepping::augmentResidual /
  // augmentJacobian, N_NLS_NOX_AugmentLinSys_GStepping.C), the scale of every
  // independent source (VSRCSCALE), and whether the first Newton iteration
  // applies initJct (only the first solve of a strategy does)
  double gcond = 0.0, src_scale = 1.0;
  bool dc_initjct = true;
  int dc_strategy = 0;               // 0 plain Newton, 1 gmin stepping, 2 source stepping (the last DC solve)
  int dc_first_stage = 0;            // tests: start the strategy at this stage (0 = the reference's order)
  int uic = 0;                       // .TRAN ... UIC: no operating point, the run starts from x = 0
  int tranop = 0;                    // the operating point being solved is a .TRAN's (tranopFlag)
  std::vector<int> vrow, vdiag;      // node-voltage rows and their diagonal CSR slots
  explicit Work(Circuit* cc) : c(cc) {
    n = c->n; nnz = (int)c->colind.size();
    x.assign(n + 1, 0.0); F.assign(n + 1, 0.0); Q.assign(n + 1, 0.0); B.assign(n + 1, 0.0);
    dFdxdVp.assign(n + 1, 0.0); dQdxdVp.assign(n + 1, 0.0); RHS.assign(n, 0.0);
    dFdx.assign(nnz, 0.0); dQdx.assign(nnz, 0.0); J.assign(nnz, 0.0);
    staCur.assign(c->nsta, 0.0); staNext.assign(c->nsta, 0.0);
    stoCur.assign(c->nsto, 0.0); stoNext.assign(c->nsto, 0.0);
  }
};

// Linear-device matrices, loaded once (CktLoader separateLoad,
// N_LOA_CktLoader.C:504-544): Resistor G stamps (N_DEV_Resistor.C:1306),
// Capacitor C stamps (N_DEV_Capacitor.C:2171), Vsrc +-1 stamps (N_DEV_Vsrc.C:1420).
static void load_linear_matrices(Circuit& c) {
  int nnz = (int)c.colind.size();
  c.linF.assign(nnz, 0.0); c.linQ.assign(nnz, 0.0);
  auto a = [&](std::vector<double>& v, int s, double x) { if (s >= 0) v[s] += x; };
  for (auto& d : c.res) { double G = 1.0 / d.val; a(c.linF, d.slot[0], G); a(c.linF, d.slot[1], -G); a(c.linF, d.slot[2], -G); a(c.linF, d.slot[3], G); }
  for (auto& d : c.cap) { double C = d.val; a(c.linQ, d.slot[0], C); a(c.linQ, d.slot[1], -C); a(c.linQ, d.slot[2], -C); a(c.linQ, d.slot[3], C); }
  for (auto& d : c.vsrc) { a(c.linF, d.slot[0], 1.0); a(c.linF, d.slot[1], -1.0); a(c.linF, d.slot[2], 1.0); a(c.linF, d.slot[3], -1.0); }
  c.lin_loaded = true;
}

// NonlinearEquationLoader::loadRHS (N_LOA_NonlinearEquationLoader.C:374-465):
// zero the DAE vectors, update sources, device updateState, device vector loads.
// Transient uses the separateLoad path (nonlinear devices, then the linear
// matrices' row sums, FilteredMatrix::axpy N_LAS_FilteredMatrix.C:473, then the
// B vector of the sources, DeviceMgr::loadBVectorsforSources); DC loads all
// devices directly.
static void load_rhs(Work& w, const SolverState& ss) {
  Circuit& c = *w.c;
  const int n = w.n;
  if (w.load_hook) {
    // the caller's loads: vectors and matrices at this x, as orc_load returns them
    const int ssi[7] = {ss.dcop, ss.tran, ss.initJct, ss.initFix, ss.newtonIter, ss.order, ss.tranop};
    const double ssd[4] = {ss.h, ss.time, w.tstep, w.tstop};
    w.x[n] = 0.0;
    w.limited = w.load_hook(w.hook_ctx, ssi, ssd, w.x.data(), w.staCur.data(), w.stoCur.data(), w.staNext.data(),
                            w.stoNext.data(), w.F.data(), w.Q.data(), w.B.data(), w.dFdxdVp.data(), w.dQdxdVp.data(),
                            w.dFdx.data(), w.dQdx.data());
    return;
  }
  std::fill(w.F.begin(), w.F.end(), 0.0); std::fill(w.Q.begin(), w.Q.end(), 0.0);
  std::fill(w.B.begin(), w.B.end(), 0.0);
  std::fill(w.dFdxdVp.begin(), w.dFdxdVp.end(), 0.0); std::fill(w.dQdxdVp.begin(), w.dQdxdVp.end(), 0.0);
  w.x[n] = 0.0;
  w.limited = 0;
  for (auto& mi : c.mos) {
    mos1_update_ext(mi, ss, c.opt, w.x.data(), w.staCur.data(), w.stoCur.data(), w.staNext.data(), w.stoNext.data());
    if (mi.limitedFlag) w.limited = 1;
  }
  for (auto& mi : c.mos) mos1_vectors_ext(mi, ss, c.opt, w.F.data(), w.Q.data(), w.dFdxdVp.data(), w.dQdxdVp.data());
  if (!c.lin_loaded) load_linear_matrices(c);
  if (ss.dcop) {
    for (auto& d : c.res) { double i0 = (w.x[d.li_p] - w.x[d.li_n]) / d.val; w.F[d.li_p] += i0; w.F[d.li_n] -= i0; }
    for (auto& d : c.cap) { double q0 = d.val * (w.x[d.li_p] - w.x[d.li_n]); w.Q[d.li_p] += q0; w.Q[d.li_n] -= q0; }
    for (auto& d : c.vsrc) { double ib = w.x[d.li_bra]; w.F[d.li_p] += ib; w.F[d.li_n] += -ib; w.F[d.li_bra] += w.x[d.li_p] - w.x[d.li_n]; }
  } else {
    for (int r = 0; r < n; ++r) {
      double sf = 0.0, sq = 0.0;
      for (int k = c.rowptr[r]; k < c.rowptr[r + 1]; ++k) {
        if (c.linF[k] != 0.0) sf += c.linF[k] * w.x[c.colind[k]];
        if (c.linQ[k] != 0.0) sq += c.linQ[k] * w.x[c.colind[k]];
      }
      w.F[r] += sf; w.Q[r] += sq;
    }
  }
  for (auto& d : c.vsrc) w.B[d.li_bra] += c.waves[d.wave].eval(ss.time, w.tstep, w.tstop, w.bptol);
  for (auto& d : c.isrc) {
    double s = c.waves[d.wave].eval(ss.time, w.tstep, w.tstop, w.bptol);
    w.B[d.li_p] -= s; w.B[d.li_n] += s;
  }
}

// B holds only the independent sources' values: source stepping scales it
static void scale_sources(Work& w) {
  if (w.src_scale == 1.0) return;
  for (int i = 0; i < w.n; ++i) w.B[i] *= w.src_scale;
}

// NonlinearEquationLoader::loadJacobian (:514-558): zero, nonlinear device
// matrices, then FilteredMatrix::addToMatrix of the linear part (:632).
static void load_jacobian(Work& w, const SolverState& ss) {
  if (w.load_hook) return;   // loaded with the vectors at the same x
  Circuit& c = *w.c;
  std::fill(w.dFdx.begin(), w.dFdx.end(), 0.0); std::fill(w.dQdx.begin(), w.dQdx.end(), 0.0);
  for (auto& mi : c.mos) mos1_matrices_ext(mi, ss, w.dFdx.data(), w.dQdx.data());
  if (!c.lin_loaded) load_linear_matrices(c);
  for (int k = 0; k < w.nnz; ++k) {
    if (c.linF[k] != 0.0) w.dFdx[k] += c.linF[k];
    if (c.linQ[k] != 0.0) w.dQdx[k] += c.linQ[k];
  }
}

// OneStep::obtainResidual N_TIA_OneStep.C:219-289 and NoTimeIntegration::obtainResidual
// (N_TIA_NoTimeIntegration.C:161).  Epetra Update(a,A,b,B,c) is c*this + a*A + b*B.
// Gear12::obtainResidual (N_TIA_Gear12.C:208-263): q0 = qHistory[0] (Q_n),
// q2 = qHistory[1] (Q_{n-1}).
static void combine_residual_gear(Work& w, const SolverState& ss, const double* q0, const double* q1) {
  const int n = w.n;
  const double h = ss.h;
  for (int i = 0; i < n; ++i) {
    double r = ss.a0 * w.Q[i] + ss.a1 * q0[i];
    if (ss.order == 2) r = r + ss.a2 * q1[i];
    r = (1.0 / h) * r + 1.0 * w.F[i] + -1.0 * w.B[i];
    r *= -1.0;
    r = r + 1.0 * ((ss.a0 / h) * w.dQdxdVp[i]);
    r = r + 1.0 * w.dFdxdVp[i];
    w.RHS[i] = r;
  }
}

static void combine_residual(Work& w, const SolverState& ss, const double* q0, const double* q2) {
  const int n = w.n;
  if (!ss.dcop && ss.method == 1) { combine_residual_gear(w, ss, q0, q2); return; }
  if (ss.dcop) {
    for (int i = 0; i < n; ++i) {
      double r = 1.0 * w.F[i] + -1.0 * w.B[i];
      r *= -1.0;
      r = r + 1.0 * w.dFdxdVp[i];
      w.RHS[i] = r;
    }
    if (w.gcond != 0.0)
      for (int r : w.vrow) w.RHS[r] -= w.gcond * w.x[r];
    return;
  }
  const double h = ss.h;
  for (int i = 0; i < n; ++i) {
    double r = 1.0 * w.Q[i] + -1.0 * q0[i];
    if (ss.order == 2) {
      r = (1.0 / h) * r + (1.0 / 2.0) * w.F[i] + (-1.0 / 2.0) * w.B[i];
      r = r + (1.0 / 2.0) * q2[i];
    } else {
      r = (1.0 / h) * r + 1.0 * w.F[i] + -1.0 * w.B[i];
    }
    r *= -1.0;
    double dq = (-(-1.0) / h) * w.dQdxdVp[i];
    r = r + 1.0 * dq;
    double fs = (ss.order == 2) ? 1.0 / 2.0 : 1.0;
    r = r + fs * w.dFdxdVp[i];
    w.RHS[i] = r;
  }
}

// OneStep::obtainJacobian :469 -> EpetraMatrix::linearCombo N_LAS_EpetraMatrix.C:629
// (J = a*dQdx + b*dFdx per entry); NoTimeIntegration::obtainJacobian uses a=1e-20, b=1.
static void combine_jacobian(Work& w, const SolverState& ss) {
  double a, b;
  if (ss.dcop) { a = 1.0e-20; b = 1.0; }
  else if (ss.method == 1) { a = ss.a0 / ss.h; b = 1.0; }    // Gear12::obtainJacobian :477-505
  else { a = -(-1.0) / ss.h; b = (ss.order == 2) ? 1.0 / 2.0 : 1.0; }
  for (int k = 0; k < w.nnz; ++k) w.J[k] = a * w.dQdx[k] + b * w.dFdx[k];
  if (ss.dcop && w.gcond != 0.0)
    for (int k : w.vdiag)
      if (k >= 0) w.J[k] += w.gcond;
}

// KLU-semantics direct solve (AmesosSolver::doSolve N_LAS_AmesosSolver.C:216-478:
// numeric factorization every call, singular -> zero solution and failure).
// Restated as dense LU with partial pivoting; KLU itself is third-party
// (Trilinos Amesos) and absent, so LU parity is on the solution: pinned against
// numpy.linalg.solve and scipy.sparse.linalg.splu in the tests.
static int lu_solve_dense(int n, const int* rowptr, const int* colind, const double* vals,
                          const double* b, double* xout) {
  std::vector<double> A((size_t)n * n, 0.0);
  for (int r = 0; r < n; ++r)
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) A[(size_t)r * n + colind[k]] += vals[k];
  std::vector<double> y(b, b + n);
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k; double best = std::fabs(A[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i) { double v = std::fabs(A[(size_t)i * n + k]); if (v > best) { best = v; p = i; } }
    if (!(best > 0.0) || !std::isfinite(best)) { for (int i = 0; i < n; ++i) xout[i] = 0.0; return k + 1; }
    if (p != k) { for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]); std::swap(y[k], y[p]); }
    double piv_v = A[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      double l = A[(size_t)i * n + k];
      if (l == 0.0) continue;
      l /= piv_v; A[(size_t)i * n + k] = l;
      double* ri = &A[(size_t)i * n]; const double* rk = &A[(size_t)k * n];
      for (int j = k + 1; j < n; ++j) ri[j] -= l * rk[j];
      y[i] -= l * y[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = y[i];
    for (int j = i + 1; j < n; ++j) s -= A[(size_t)i * n + j] * xout[j];
    xout[i] = s / A[(size_t)i * n + i];
  }
  for (int i = 0; i < n; ++i) if (!std::isfinite(xout[i])) { for (int j = 0; j < n; ++j) xout[j] = 0.0; return -99; }
  return 0;
}


static int lu_solve_sparse(SparseLU& lu, int n, const int* rowptr, const int* colind, const double* vals,
                           const double* b, double* xout) {
  if (!lu.analysed || lu.n != n) lu.analyse(n, rowptr, colind);
  int st = lu.factor(vals);
  if (st != 0) { for (int i = 0; i < n; ++i) xout[i] = 0.0; return st; }
  lu.solve(b, xout);
  for (int i = 0; i < n; ++i) if (!std::isfinite(xout[i])) { for (int j = 0; j < n; ++j) xout[j] = 0.0; return -99; }
  return 0;
}

struct NLP {
  double reltol, abstol, dxtol, rhstol, smallUpdateTol; int maxstep; bool transient;
};

static double inf_norm(const std::vector<double>& v, int n) { double m = 0; for (int i = 0; i < n; ++i) m = std::max(m, std::fabs(v[i])); return m; }
static double two_norm(const std::vector<double>& v, int n) { double s = 0; for (int i = 0; i < n; ++i) s += v[i] * v[i]; return std::sqrt(s); }

struct NewtonStats { int iters = 0; int status = 0; };

// DampedNewton::solve N_NLS_DampedNewton.C:362-509 with FULL steps
// (fullNewton_ :1147), updateWeights_ :297, converged_ :1191.  The transient
// stopping rule (weights, DeltaXTol 0.33, limiter flags) is pinned by a
// hook-observed run in which every solve must stop exactly where the rule,
// evaluated from the observed iterates, says (tests/test_oracle_pins.py).
static NewtonStats newton(Work& w, SolverState ss, const NLP& p, const std::vector<double>& xcurr,
                          const double* q0, const double* q2, long* n_loads) {
  const int n = w.n;
  NewtonStats st;
  std::vector<double> dx(n), wt(n);
  auto rhs = [&](int it) {
    ss.newtonIter = it;
    ss.initJct = ss.dcop && it == 0 && w.dc_initjct;   // N_DEV_SolverState.C:387-409
    load_rhs(w, ss);
    if (ss.dcop) scale_sources(w);
    combine_residual(w, ss, q0, q2);
    if (n_loads) ++*n_loads;
  };
  auto weights = [&]() {
    for (int i = 0; i < n; ++i) wt[i] = p.reltol * std::max(std::fabs(xcurr[i]), std::fabs(w.x[i])) + p.abstol;
  };
  rhs(0);
  double normRHS = two_norm(w.RHS, n);
  double normRHS_init = normRHS, normRHS_old = normRHS;
  if (p.transient) weights();
  int nlStep = 0, count = 0; double tmpConvRate = 1.0;
  const double eps = std::numeric_limits<double>::epsilon();
  const double convRateMax = 0.5 * std::numeric_limits<double>::max();
  while (true) {
    nlStep++;
    load_jacobian(w, ss);
    combine_jacobian(w, ss);
    int ls = w.solve_hook
                 ? w.solve_hook(w.hook_ctx, n, w.c->rowptr.data(), w.c->colind.data(), w.J.data(), w.RHS.data(), dx.data())
                 : lu_solve_sparse(w.slu, n, w.c->rowptr.data(), w.c->colind.data(), w.J.data(), w.RHS.data(), dx.data());
    for (int i = 0; i < n; ++i) w.x[i] += dx[i];
    rhs(nlStep);
    normRHS = two_norm(w.RHS, n);
    bool nanRHS = !std::isfinite(normRHS);
    if (!p.transient) weights();
    int status = 0;
    if (ls != 0) { status = (ls == -99) ? -6 : -9; }
    else if (w.limited && nlStep < p.maxstep) { status = 0; goto next; }
    else if (w.limited) { status = -1; }
    else if (nanRHS) { status = -6; }
    else {
      double maxNormRHS = inf_norm(w.RHS, n);
      double wtNormDX = 0; for (int i = 0; i < n; ++i) wtNormDX = std::max(wtNormDX, std::fabs(dx[i]) / wt[i]);
      if (normRHS < eps) { status = 1; }
      else {
        double normRHS_rel = normRHS / normRHS_init;
        double resConvRate = normRHS / normRHS_old;
        normRHS_old = normRHS;
        double updateSize = wtNormDX;
        if (maxNormRHS <= p.rhstol && updateSize <= p.dxtol) status = 2;
        else if (nlStep >= p.maxstep && p.transient && normRHS_rel <= 0.9 && resConvRate <= 1.0) status = -3;
        else if (updateSize <= p.smallUpdateTol) status = 4;
        else if (nlStep >= p.maxstep) status = -1;
        else if (resConvRate > convRateMax) status = -2;
        else {
          if (p.transient && std::fabs(resConvRate - 1.0) <= 1.0e-3) {
            if (count == 0 || resConvRate < tmpConvRate) tmpConvRate = resConvRate;
            count++;
          } else count = 0;
          if (p.transient && count == 5) status = -3;
        }
      }
    }
    if (status != 0) { st.iters = nlStep; st.status = status; return st; }
  next:;
  }
}

// Breakpoint list: sources' corners (PulseData/PWLinData getBreakPoints
// N_DEV_SourceData.C:1442,2044) plus the stop time.
static std::vector<double> breakpoints(const Circuit& c, double tstep, double tstop) {
  std::vector<double> bp;
  for (auto& d : c.vsrc) c.waves[d.wave].breakpoints(tstep, tstop, bp);
  for (auto& d : c.isrc) c.waves[d.wave].breakpoints(tstep, tstop, bp);
  bp.push_back(tstop);
  std::sort(bp.begin(), bp.end());
  bp.erase(std::unique(bp.begin(), bp.end()), bp.end());
  return bp;
}

struct TranResult {
  std::vector<double> t, x;   // accepted points
  std::vector<int> ord;       // integration order of the step that ended at each point (0 = DC point)
  long newton_iters = 0, steps = 0, rejects = 0, loads = 0;
  double seconds = 0;
  int status = 0;
};

static double wrms(const double* v, const double* wt, int n) {
  double s = 0; for (int i = 0; i < n; ++i) { double q = v[i] / wt[i]; s += q * q; }
  return std::sqrt(s / n);
}

// DC operating point.  Xyce uses NOX for DCOP (N_NLS_Manager.C:368-376), which
// is not on the path; the oracle runs DampedNewton with the DC defaults and an
// initJct first iteration (N_DEV_SolverState.C:387-409).
// One continuation run (LOCA natural continuation with a constant
// predictor and adaptive steps, as set up by Interface::gminSteppingSolve
// :1116-1185 and sourceSteppingSolve :1284-1345): the parameter moves from
// p0 to p1; each step is a Newton solve (at most 20 iterations) from the last
// converged point; a failed step is retried at half the step, a converged one
// lets the step grow by (1 + a (1 - (iters/20)^2)) up to max_step, with the
// aggressiveness a = 0.01 for gmin and 1 for source stepping; source stepping
// disables initJct (setDisableInitJctFlags).  LOCA is
// third-party (Trilinos); this restates its published natural-continuation
// scheme with the reference's settings.  The continuation and iteration path:
// parity unpinned; the operating point it reaches is pinned (gate logic levels,
// the DC residual, and each stage reaching plain Newton's point).
template <class SetParam>
static int dc_continuation(Work& w, long* iters, double p0, double p1, double step, double min_step, double max_step,
                           double aggressiveness, bool init_jct, SetParam set_param) {
  const int n = w.n;
  SolverState base; base.dcop = 1; base.tran = 0; base.tranop = w.tranop; base.time = 0.0;
  NLP p{1e-3, 1e-12, 1.0, 1e-6, 1e-6, 20, false};
  std::vector<double> xc(n, 0.0);
  std::fill(w.x.begin(), w.x.end(), 0.0);
  std::fill(w.staCur.begin(), w.staCur.end(), 0.0); std::fill(w.stoCur.begin(), w.stoCur.end(), 0.0);
  std::vector<double> x_ok(w.x), sta_ok(w.staCur), sto_ok(w.stoCur);
  double par = p0, par_ok = p0;
  bool first = true, done = false;
  for (int k = 0; k < 400 && !done; ++k) {
    set_param(par);
    w.dc_initjct = first && init_jct;
    NewtonStats ns = newton(w, base, p, xc, nullptr, nullptr, nullptr);
    if (iters) *iters += ns.iters;
    if (ns.status > 0) {
      first = false;
      par_ok = par;
      x_ok = w.x; sta_ok = w.staNext; sto_ok = w.stoNext;
      w.staCur = w.staNext; w.stoCur = w.stoNext;
      if (par == p1) { done = true; break; }
      const double r = ns.iters / 20.0;
      step *= 1.0 + aggressiveness * (1.0 - r * r);
      if (std::fabs(step) > max_step) step = step > 0 ? max_step : -max_step;
    } else {
      std::copy(x_ok.begin(), x_ok.end(), w.x.begin());
      w.staNext = sta_ok; w.stoNext = sto_ok;
      step *= 0.5;
      if (std::fabs(step) < min_step) break;
    }
    par = par_ok + step;
    if ((p1 - par) * (p1 - p0) <= 0.0) par = p1;   // do not step past the end value
  }
  w.dc_initjct = true;
  set_param(p1);
  return done ? 1 : -1;
}

// DC operating point: Interface::spiceStrategy (N_NLS_NOX_Interface.C:321-437):
// Newton from initJct; if that fails, gmin stepping (10^p - 10^-4 siemens on
// every node-voltage row, p from 4 to -4); if that fails, source stepping
// (all independent sources scaled from 0 to 1).
static int dcop(Work& w, long* iters) {
  const int n = w.n;
  SolverState base; base.dcop = 1; base.tran = 0; base.tranop = w.tranop; base.time = 0.0;
  NLP p{1e-3, 1e-12, 1.0, 1e-6, 1e-6, 200, false};
  std::vector<double> xc(n, 0.0);
  std::fill(w.x.begin(), w.x.end(), 0.0);
  w.gcond = 0.0; w.src_scale = 1.0; w.dc_initjct = true;
  w.dc_strategy = 0;
  if (w.dc_first_stage == 0) {
    NewtonStats ns = newton(w, base, p, xc, nullptr, nullptr, nullptr);
    if (iters) *iters += ns.iters;
    if (ns.status > 0) return ns.status;
  }
  if (w.vrow.empty()) {
    const Circuit& c = *w.c;
    for (size_t k = 1; k < c.node_gid.size(); ++k) {
      const int r = c.node_gid[k];
      if (r < 0) continue;
      int slot = -1;
      for (int q = c.rowptr[r]; q < c.rowptr[r + 1]; ++q) if (c.colind[q] == r) slot = q;
      w.vrow.push_back(r);
      w.vdiag.push_back(slot);
    }
  }
  int st = -1;
  if (w.dc_first_stage <= 1) {
    w.dc_strategy = 1;
    st = dc_continuation(w, iters, 4.0, -4.0, -2.0, 1e-6, 1e12, 0.01, true,
                         [&](double q) { w.gcond = std::pow(10.0, q) - std::pow(10.0, -4.0); });
    w.gcond = 0.0;
    if (st > 0) return st;
  }
  w.dc_strategy = 2;
  st = dc_continuation(w, iters, 0.0, 1.0, 0.2, 1e-4, 0.2, 1.0, false, [&](double q) { w.src_scale = q; });
  w.src_scale = 1.0;
  return st;
}

// Transient loop: OneStep::initialize (N_TIA_OneStep.C:1684), updateCoeffs
// (:1584), obtainPredictor (:135), StepErrorControl::evaluateStepError
// (N_TIA_StepErrorControl.C:458, error test skipped on the first step after
// (re)initialization), completeStep (:2124) / rejectStep (:1946), updateHistory
// (:1390).  Error weights: DataStore::setErrorWtVector with NEWLTE=1
// (N_TIA_DataStore.C:1304), estimate ck*WRMS(x - xn0) (N_TIA_OneStep.h:93).
// Breakpoints restart the integrator at order 1.
static void record_point(const Work& w, TranResult& R, double t, const std::vector<double>& x, int ord) {
  R.t.push_back(t);
  R.ord.push_back(ord);
  if (w.probe_k == 0) {
    R.x.insert(R.x.end(), x.begin(), x.begin() + w.n);
    return;
  }
  // compensated (Neumaier) sums: a probe such as the total charge of a large
  // network is a sum of many like terms, whose plain rounding would swamp the
  // small changes a step makes to it
  for (int r = 0; r < w.probe_k; ++r) {
    const double* wr = w.probe_w.data() + (size_t)r * w.n;
    double s = 0.0, comp = 0.0;
    for (int i = 0; i < w.n; ++i) {
      const double v = wr[i] * x[i];
      const double t = s + v;
      comp += std::fabs(s) >= std::fabs(v) ? (s - t) + v : (v - t) + s;
      s = t;
    }
    R.x.push_back(s + comp);
  }
}

static TranResult tran(Work& w, double tstop, double tstep, double hmax, bool record, double t_record_stop) {
  TranResult R;
  Circuit& c = *w.c;
  const int n = w.n;
  w.tstep = tstep; w.tstop = tstop;
  auto t0 = std::chrono::steady_clock::now();
  long dc_iters = 0;
  if (w.uic) {
    // .TRAN ... UIC / NOOP (N_ANP_Transient.C:879-1000): no operating point;
    // the solution starts at zero (no .IC here) and one load there
    // (initializeProblem, loadRHS) gives the devices' state and F, Q, B from
    // which the histories are set as after an operating point
    std::fill(w.x.begin(), w.x.end(), 0.0);
    std::fill(w.staCur.begin(), w.staCur.end(), 0.0); std::fill(w.stoCur.begin(), w.stoCur.end(), 0.0);
    std::fill(w.staNext.begin(), w.staNext.end(), 0.0); std::fill(w.stoNext.begin(), w.stoNext.end(), 0.0);
    SolverState ss; ss.dcop = 0; ss.tran = 1; ss.newtonIter = 0; ss.time = 0.0; ss.h = tstep; ss.method = w.method;
    load_rhs(w, ss);
    w.dc_strategy = -1;
  } else {
    w.tranop = 1;   // tranopFlag: a transient with no time integration yet (N_ANP_AnalysisManager.C:1170-1174)
    int st = dcop(w, &dc_iters);
    w.tranop = 0;
    if (st <= 0) { R.status = -100; return R; }
  }
  // tranop -> transient
  w.staCur = w.staNext; w.stoCur = w.stoNext;
  std::vector<double> xH0(w.x.begin(), w.x.begin() + n), xH1(n, 0.0), xH2(n, 0.0);
  std::vector<double> qH0(w.Q.begin(), w.Q.begin() + n), qH1(n), qH2(n, 0.0);
  std::vector<double> F0(w.F.begin(), w.F.begin() + n), B0(w.B.begin(), w.B.begin() + n);
  const double reltol = w.tia_reltol, abstol = w.tia_abstol;
  const double r_min = 0.25, r_max = 0.9, r_hincr = 2.0, r_hincr_test = 2.0, tolAimFac = 0.5;
  const double h0_safety = 2.0, h0_max_factor = 0.005;
  const double hmin = 1e-18;
  std::vector<double> bps = breakpoints(c, tstep, tstop);
  size_t bpi = 0;
  double t = 0.0;
  int order = 1, usedOrder = 1; long numberOfSteps = 0; int nef = 0;
  double psi[3] = {0, 0, 0};
  double h, lastTimeStep, currentTimeStepSum, savedTimeStep = 0.0, estOverTol = 0.0;
  bool beginning = true;
  auto next_bp = [&]() { while (bpi < bps.size() && bps[bpi] <= t * (1 + 1e-13) + 1e-30) ++bpi; return bpi < bps.size() ? bps[bpi] : tstop; };
  auto initialize = [&](bool initial) {
    // qHistory[1] = -(F - B)*h ; step from delta_x_errorNorm_q1
    std::vector<double> qwt(n);
    for (int i = 0; i < n; ++i) qwt[i] = reltol * std::fabs(qH0[i]) + abstol;
    double stopTime = next_bp();
    double time_to_stop = stopTime - t;
    std::vector<double> dq(n);
    for (int i = 0; i < n; ++i) dq[i] = (F0[i] - B0[i]);
    double hh;
    if (initial) {
      // dnorm_q uses qHistory[1] before it is scaled by -h: it is formed with
      // the trial step, so compute with the (F-B) direction scaled by the
      // candidate below (Xyce computes it on the previous qHistory[1], which is
      // zero at the start of a run).
      hh = h0_max_factor * std::fabs(time_to_stop);
    } else {
      hh = 0.1 * std::min(savedTimeStep, std::fabs(time_to_stop));
    }
    if (tstep > 0.0 && initial) hh = std::min(tstep, hh);
    if (hmax > 0 && hh > hmax) hh = hmax;
    hh = std::max(hh, hmin);
    h = hh;
    lastTimeStep = h; currentTimeStepSum = 2.0 * h;
    if (w.method == 1) {
      // Gear12::initialize (N_TIA_Gear12.C:1272-1285): both history slots hold the current point
      for (int i = 0; i < n; ++i) { xH1[i] = xH0[i]; qH1[i] = qH0[i]; }
    } else {
      for (int i = 0; i < n; ++i) { xH1[i] = 0.0; qH1[i] = -(F0[i] - B0[i]) * h; }
    }
    numberOfSteps = 0; order = 1; usedOrder = 1; psi[0] = h;
    beginning = true;
  };
  initialize(true);
  SolverState ss; ss.dcop = 0; ss.tran = 1;
  NLP p{w.nl_reltol, w.nl_abstol, 0.33, 1e-2, 1e-6, 20, true};
  std::vector<double> xn0(n), xcurr(n), corr(n), errwt(n);
  if (record) record_point(w, R, 0.0, xH0, 0);
  while (t < tstop * (1 - 1e-12)) {
    double stopTime = next_bp();
    if (t + h > stopTime) { savedTimeStep = h; h = stopTime - t; currentTimeStepSum = h + lastTimeStep; }
    // updateCoeffs
    double temp1 = h;
    if (order == 2) psi[2] = psi[1];
    psi[1] = psi[0]; psi[0] = temp1;
    double beta0 = 1.0, beta1, beta2 = 0.0, ck;
    double a0 = 1.0, a1 = -1.0, a2 = 0.0;
    if (w.method == 1) {
      // Gear12::updateCoeffs (N_TIA_Gear12.C:1050-1111): BDF2 predictor and
      // corrector weights for variable steps, or BDF1
      if (order == 2) {
        beta2 = temp1 / psi[2] * (temp1 + psi[1]) / (psi[1] + psi[2]);
        beta1 = -temp1 / psi[1] - beta2 * (psi[1] + psi[2]) / psi[1];
        beta0 = 1.0 - beta2 - beta1;
        a2 = -temp1 / psi[1] * temp1 / (2 * temp1 + psi[1]);
        a1 = 1 - a2;
        a0 = -a1 - a2 * (1 + psi[1] / temp1);
        a2 = a2 / a0;
        a1 = a1 / a0;
        a0 = -1 / a0;
        ck = h / (temp1 + psi[1] + psi[2]);
      } else {
        beta0 = 1.0 + temp1 / psi[1];
        beta1 = -temp1 / psi[1];
        ck = h / (temp1 + psi[1]);
      }
    } else if (order == 2) {
      double temp2 = psi[1];
      beta1 = temp1 / temp2 + (temp1 / temp2) * (temp1 / temp2) / 2;
      beta2 = -1.0 * temp1 * temp1 / temp2 / psi[2] / 2;
      ck = (h / currentTimeStepSum) / 3.0;
    } else {
      beta1 = temp1 / psi[1];
      ck = h / currentTimeStepSum;
    }
    // predictor: OneStep x_n + sum beta_i (difference history); Gear12
    // sum beta_i x_{n-i} (N_TIA_Gear12.C:131-141)
    for (int i = 0; i < n; ++i) {
      double v;
      if (w.method == 1) {
        v = beta0 * xH0[i];
        v = beta1 * xH1[i] + v;
        if (order == 2) v = beta2 * xH2[i] + v;
      } else {
        v = xH0[i];
        v += beta1 * xH1[i];
        if (order == 2) v += beta2 * xH2[i];
      }
      xn0[i] = v;
      xcurr[i] = xH0[i];
    }
    double mx = 0; for (int i = 0; i < n; ++i) mx = std::max(mx, std::fabs(xH0[i]));
    for (int i = 0; i < n; ++i) errwt[i] = reltol * mx + abstol;
    for (int i = 0; i < n; ++i) w.x[i] = xn0[i];
    ss.order = order; ss.h = h; ss.time = t + h;
    ss.method = w.method; ss.a0 = a0; ss.a1 = a1; ss.a2 = a2;
    NewtonStats ns = newton(w, ss, p, xcurr, qH0.data(), w.method == 1 ? qH1.data() : qH2.data(), &R.loads);
    R.newton_iters += ns.iters;
    bool ok = ns.status > 0;
    if (ok && !beginning) {
      for (int i = 0; i < n; ++i) corr[i] = w.x[i] - xn0[i];
      estOverTol = ck * wrms(corr.data(), errwt.data(), n);
      ok = estOverTol <= 1.0;
    }
    if (ok) {
      // completeStep
      numberOfSteps++; nef = 0; R.steps++;
      double tnew = t + h;
      double newh = h;
      double oldLast = lastTimeStep; (void)oldLast;
      lastTimeStep = h; usedOrder = order;
      double rr = tolAimFac / (estOverTol + 0.0001);
      rr = std::pow(rr, 1.0 / (order + 1.0));
      if (numberOfSteps >= 2 && order == 1) {
        order = 2;
        rr = tolAimFac / (estOverTol + 0.0001);
        rr = std::pow(rr, 1.0 / (order + 1.0));
        if (rr <= 1.05) order = 1;
      }
      if (rr >= r_hincr_test) { rr = r_hincr; newh = rr * h; }
      else if (rr <= 1) { rr = std::max(r_min, std::min(r_max, rr)); newh = rr * h; }
      // updateHistory
      if (w.method == 1) {
        // Gear12::updateHistory (N_TIA_Gear12.C:822-857): shift the values
        if (order == 2) xH2 = xH1;
        qH1 = qH0;
        xH1 = xH0;
        for (int i = 0; i < n; ++i) { xH0[i] = w.x[i]; qH0[i] = w.Q[i]; }
      } else {
        if (order == 2) {
          xH2 = xH1;
          for (int i = 0; i < n; ++i) qH2[i] = 1.0 * w.F[i] + -1.0 * w.B[i];
        }
        for (int i = 0; i < n; ++i) {
          xH1[i] = 1.0 * w.x[i] + -1.0 * xH0[i];
          qH1[i] = 1.0 * w.Q[i] + -1.0 * qH0[i];
          xH0[i] = w.x[i]; qH0[i] = w.Q[i];
        }
      }
      w.staCur = w.staNext; w.stoCur = w.stoNext;
      t = tnew;
      if (record && t <= t_record_stop * (1 + 1e-12)) record_point(w, R, t, xH0, usedOrder);
      newh = std::max(newh, hmin);
      if (hmax > 0) newh = std::min(newh, hmax);
      beginning = false;
      // reached a breakpoint: restart at order 1
      if (bpi < bps.size() && std::fabs(t - bps[bpi]) <= 1e-13 * std::max(1.0, std::fabs(t)) + 1e-24 && t < tstop * (1 - 1e-12)) {
        for (int i = 0; i < n; ++i) { F0[i] = w.F[i]; B0[i] = w.B[i]; }
        savedTimeStep = newh;
        ++bpi;
        initialize(false);
      } else {
        currentTimeStepSum = newh + lastTimeStep;
        h = newh;
      }
    } else {
      // rejectStep
      R.rejects++;
      nef++;
      if (nef >= 15) { R.status = -101; break; }
      for (int i = 1; i <= order; ++i) psi[i - 1] = psi[i];
      double newh;
      if (ns.status <= 0) { newh = h / 8; order = 1; }
      else if (nef == 1) {
        // OneStep restarts at order 1; Gear12 keeps its order for the first
        // local-error failure (N_TIA_Gear12.C:1502-1516)
        if (w.method != 1) order = 1;
        double rr = tolAimFac / (estOverTol + 0.0001);
        rr = std::pow(rr, 1.0 / (order + 1.0));
        rr = std::max(r_min, std::min(r_max, rr));
        newh = rr * h;
      } else { newh = r_min * h; order = 1; }
      newh = std::max(newh, hmin);
      if (hmax > 0) newh = std::min(newh, hmax);
      if (newh <= hmin) { R.status = -102; break; }
      currentTimeStepSum = newh + lastTimeStep;
      h = newh;
      w.staNext = w.staCur; w.stoNext = w.stoCur;
    }
  }
  R.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return R;
}

} // namespace orc

// ===========================================================================
// extern "C" surface for the tests (ctypes).
// ===========================================================================
using namespace orc;

struct OrcHandle {
  Circuit c;
  Work* w = nullptr;
  TranResult last;
};

extern "C" {

OrcHandle* orc_create(int n_nodes, int n_dev, const int* dev_type, const int* dev_idx,
                      int n_models, const double* model_par,
                      int n_mos, const int* mos_nodes, const int* mos_model, const double* mos_inst,
                      int n_res, const int* res_nodes, const double* res_val,
                      int n_cap, const int* cap_nodes, const double* cap_val,
                      int n_vsrc, const int* vsrc_nodes, const int* vsrc_wave,
                      int n_isrc, const int* isrc_nodes, const int* isrc_wave,
                      int n_wave, const int* wave_type, const int* wave_off, const double* wave_data,
                      double gmin, double temp) {
  try {
    auto* h = new OrcHandle();
    Circuit& c = h->c;
    c.opt.gmin = gmin; c.opt.temp = temp;
    c.n_nodes = n_nodes;
    for (int i = 0; i < n_dev; ++i) c.order.push_back({dev_type[i], dev_idx[i]});
    c.models.resize(n_models);
    for (int m = 0; m < n_models; ++m) std::memcpy(c.models[m].p, model_par + (size_t)m * MP_COUNT, sizeof(double) * MP_COUNT);
    c.mos.resize(n_mos);
    for (int i = 0; i < n_mos; ++i) {
      for (int k = 0; k < 4; ++k) c.mos[i].node[k] = mos_nodes[4 * i + k];
      c.mos[i].model = mos_model[i];
      std::memcpy(c.mos[i].ip, mos_inst + (size_t)i * IP_COUNT, sizeof(double) * IP_COUNT);
    }
    auto two = [](std::vector<TwoTerm>& v, int cnt, const int* nodes, const double* val, const int* wave) {
      v.resize(cnt);
      for (int i = 0; i < cnt; ++i) {
        v[i].node[0] = nodes[2 * i]; v[i].node[1] = nodes[2 * i + 1];
        v[i].val = val ? val[i] : 0.0; v[i].wave = wave ? wave[i] : -1;
      }
    };
    two(c.res, n_res, res_nodes, res_val, nullptr);
    two(c.cap, n_cap, cap_nodes, cap_val, nullptr);
    two(c.vsrc, n_vsrc, vsrc_nodes, nullptr, vsrc_wave);
    two(c.isrc, n_isrc, isrc_nodes, nullptr, isrc_wave);
    c.waves.resize(n_wave);
    for (int i = 0; i < n_wave; ++i) {
      c.waves[i].type = wave_type[i];
      c.waves[i].p.assign(wave_data + wave_off[i], wave_data + wave_off[i + 1]);
    }
    c.setup();
    h->w = new Work(&c);
    return h;
  } catch (...) {
    return nullptr;
  }
}

void orc_destroy(OrcHandle* h) { if (h) { delete h->w; delete h; } }

// out: n, nnz, nsta, nsto, n_mos
void orc_info(OrcHandle* h, int* out) {
  out[0] = h->c.n; out[1] = (int)h->c.colind.size(); out[2] = h->c.nsta; out[3] = h->c.nsto; out[4] = (int)h->c.mos.size();
}

void orc_pattern(OrcHandle* h, int* rowptr, int* colind, int* node_gid, int* mos_lid, int* mos_fslot,
                 int* mos_qslot, int* mos_stasto) {
  Circuit& c = h->c;
  std::copy(c.rowptr.begin(), c.rowptr.end(), rowptr);
  std::copy(c.colind.begin(), c.colind.end(), colind);
  std::copy(c.node_gid.begin(), c.node_gid.end(), node_gid);
  for (size_t i = 0; i < c.mos.size(); ++i) {
    const Mos1& m = c.mos[i];
    int L[6] = {m.li_d, m.li_g, m.li_s, m.li_b, m.li_dp, m.li_sp};
    for (int k = 0; k < 6; ++k) mos_lid[6 * i + k] = L[k];
    for (int k = 0; k < 17; ++k) mos_fslot[17 * i + k] = m.fslot[k];
    for (int k = 0; k < 14; ++k) mos_qslot[14 * i + k] = m.qslot[k];
    mos_stasto[2 * i] = m.sta0; mos_stasto[2 * i + 1] = m.sto0;
  }
}

// One device load at x: updateState + loadDAEVectors + loadDAEMatrices.
// ssi: dcop tran initJct initFix newtonIter order tranop ; ssd: h time tstep tstop
int orc_load(OrcHandle* h, const int* ssi, const double* ssd, const double* x,
             const double* sta_cur, const double* sto_cur, double* sta_next, double* sto_next,
             double* F, double* Q, double* B, double* dFdxdVp, double* dQdxdVp, double* dFdx, double* dQdx) {
  Work& w = *h->w;
  SolverState ss; ss.dcop = ssi[0]; ss.tran = ssi[1]; ss.initJct = ssi[2]; ss.initFix = ssi[3];
  ss.newtonIter = ssi[4]; ss.order = ssi[5]; ss.tranop = ssi[6]; ss.h = ssd[0]; ss.time = ssd[1];
  w.tstep = ssd[2]; w.tstop = ssd[3];
  std::copy(x, x + w.n, w.x.begin()); w.x[w.n] = 0.0;
  std::copy(sta_cur, sta_cur + w.c->nsta, w.staCur.begin());
  std::copy(sto_cur, sto_cur + w.c->nsto, w.stoCur.begin());
  std::copy(sta_next, sta_next + w.c->nsta, w.staNext.begin());
  std::copy(sto_next, sto_next + w.c->nsto, w.stoNext.begin());
  load_rhs(w, ss);
  load_jacobian(w, ss);
  std::copy(w.staNext.begin(), w.staNext.end(), sta_next);
  std::copy(w.stoNext.begin(), w.stoNext.end(), sto_next);
  std::copy(w.F.begin(), w.F.begin() + w.n, F); std::copy(w.Q.begin(), w.Q.begin() + w.n, Q);
  std::copy(w.B.begin(), w.B.begin() + w.n, B);
  std::copy(w.dFdxdVp.begin(), w.dFdxdVp.begin() + w.n, dFdxdVp);
  std::copy(w.dQdxdVp.begin(), w.dQdxdVp.begin() + w.n, dQdxdVp);
  std::copy(w.dFdx.begin(), w.dFdx.end(), dFdx); std::copy(w.dQdx.begin(), w.dQdx.end(), dQdx);
  return w.limited;
}

// Residual and Jacobian combine (a8) on the vectors of the last orc_load.
void orc_combine(OrcHandle* h, const int* ssi, const double* ssd, const double* q0, const double* q2,
                 double* RHS, double* J) {
  Work& w = *h->w;
  SolverState ss; ss.dcop = ssi[0]; ss.tran = ssi[1]; ss.order = ssi[5]; ss.h = ssd[0]; ss.time = ssd[1];
  combine_residual(w, ss, q0, q2);
  combine_jacobian(w, ss);
  std::copy(w.RHS.begin(), w.RHS.end(), RHS);
  std::copy(w.J.begin(), w.J.end(), J);
}

int orc_lu_solve(int n, const int* rowptr, const int* colind, const double* vals, const double* b, double* x) {
  return lu_solve_dense(n, rowptr, colind, vals, b, x);
}

int orc_lu_solve_sparse(int n, const int* rowptr, const int* colind, const double* vals, const double* b, double* x) {
  SparseLU lu;
  return lu_solve_sparse(lu, n, rowptr, colind, vals, b, x);
}

// CPU baseline: run the full transient of every handle, spread over `threads`
// host threads (each handle runs on one thread).  out[0]=total Newton
// iterations, out[1]=wall seconds, out[2]=accepted steps, out[3]=failures.
int orc_tran_many(int count, OrcHandle** hs, double tstop, double tstep, double hmax, int threads, double* out) {
  std::vector<TranResult> res(count);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  std::atomic<int> next{0};
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (int i = next++; i < count; i = next++) res[i] = tran(*hs[i]->w, tstop, tstep, hmax, false, 0.0);
    });
  for (auto& th : pool) th.join();
  out[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  double it = 0, st = 0, fail = 0;
  for (auto& r : res) { it += r.newton_iters; st += r.steps; fail += r.status != 0; }
  out[0] = it; out[2] = st; out[3] = fail;
  return 0;
}

int orc_limvds(int cnt, const double* vnew, const double* vold, double* out) {
  for (int i = 0; i < cnt; ++i) out[i] = limvds(vnew[i], vold[i]);
  return 0;
}
int orc_pnjlim(int cnt, const double* vnew, const double* vold, const double* vt, const double* vcrit, double* out, int* icheck) {
  for (int i = 0; i < cnt; ++i) out[i] = pnjlim(vnew[i], vold[i], vt[i], vcrit[i], &icheck[i]);
  return 0;
}
int orc_fetlim(int cnt, const double* vnew, const double* vold, const double* vto, double* out) {
  for (int i = 0; i < cnt; ++i) out[i] = fetlim(vnew[i], vold[i], vto[i]);
  return 0;
}
int orc_qmeyer(int cnt, const double* in /*vgs vgd vgb von vdsat phi cox*/, double* out /*capgs capgd capgb*/) {
  for (int i = 0; i < cnt; ++i) {
    const double* a = in + 7 * i;
    qmeyer(a[0], a[1], a[2], a[3], a[4], out[3 * i], out[3 * i + 1], out[3 * i + 2], a[5], a[6]);
  }
  return 0;
}
double orc_wave(OrcHandle* h, int wave, double t, double tstep, double tstop, double bptol) {
  return h->c.waves[wave].eval(t, tstep, tstop, bptol);
}

// DC operating point; returns Newton status, x (n), sta/sto (next).
int orc_dcop(OrcHandle* h, double* x, double* sta, double* sto, double* Q, double* F, double* B) {
  Work& w = *h->w;
  long it = 0;
  int st = dcop(w, &it);
  std::copy(w.x.begin(), w.x.begin() + w.n, x);
  std::copy(w.staNext.begin(), w.staNext.end(), sta);
  std::copy(w.stoNext.begin(), w.stoNext.end(), sto);
  std::copy(w.Q.begin(), w.Q.begin() + w.n, Q);
  std::copy(w.F.begin(), w.F.begin() + w.n, F);
  std::copy(w.B.begin(), w.B.begin() + w.n, B);
  return st;
}

// Full transient.  stats: [steps, rejects, newton_iters, loads, seconds, status, npoints]
int orc_set_hooks(OrcHandle* h, OrcLoadHook load, OrcSolveHook solve, void* ctx) {
  h->w->load_hook = load;
  h->w->solve_hook = solve;
  h->w->hook_ctx = ctx;
  return 0;
}

// Transient tolerances (a tighter run is the accurate reference that the
// tests compare runs with different Newton paths against).
int orc_set_tolerances(OrcHandle* h, double tia_reltol, double tia_abstol, double nl_reltol, double nl_abstol) {
  h->w->tia_reltol = tia_reltol; h->w->tia_abstol = tia_abstol;
  h->w->nl_reltol = nl_reltol; h->w->nl_abstol = nl_abstol;
  return 0;
}

int orc_tran(OrcHandle* h, double tstop, double tstep, double hmax, int record, double t_record_stop, double* stats) {
  Work& w = *h->w;
  h->last = tran(w, tstop, tstep, hmax, record != 0, t_record_stop);
  const TranResult& R = h->last;
  stats[0] = (double)R.steps; stats[1] = (double)R.rejects; stats[2] = (double)R.newton_iters;
  stats[3] = (double)R.loads; stats[4] = R.seconds; stats[5] = R.status; stats[6] = (double)R.t.size();
  return R.status;
}

void orc_tran_points(OrcHandle* h, double* t, double* x) {
  const TranResult& R = h->last;
  std::copy(R.t.begin(), R.t.end(), t);
  std::copy(R.x.begin(), R.x.end(), x);
}

void orc_tran_orders(OrcHandle* h, int* ord) {
  const TranResult& R = h->last;
  std::copy(R.ord.begin(), R.ord.end(), ord);
}

// Record k products w_r . x per accepted point instead of the solution
// (k = 0 restores whole solutions); w is k x n, row-major.
// Which DC strategy produced the last operating point: 0 Newton, 1 gmin
// stepping, 2 source stepping.
int orc_dc_strategy(OrcHandle* h) { return h->w->dc_strategy; }

// Start later DC solves at this stage of the strategy (0 Newton, 1 gmin
// stepping, 2 source stepping): lets the tests reach each stage.
int orc_set_dc_first_stage(OrcHandle* h, int stage) {
  if (stage < 0 || stage > 2) return -1;
  h->w->dc_first_stage = stage;
  return 0;
}

// Later orc_tran calls skip the operating point (.TRAN ... UIC) when on != 0.
int orc_set_uic(OrcHandle* h, int on) {
  h->w->uic = on != 0;
  return 0;
}

// Integration method of later orc_tran calls: 0 = OneStep trapezoid (the
// default), 1 = Gear12.
int orc_set_method(OrcHandle* h, int method) {
  if (method != 0 && method != 1) return -1;
  h->w->method = method;
  return 0;
}

int orc_set_probe(OrcHandle* h, int k, const double* w) {
  if (k < 0 || (k > 0 && !w)) return -1;
  h->w->probe_k = k;
  h->w->probe_w.assign(w, w + (size_t)k * h->w->n);
  return 0;
}

}  // extern "C"
