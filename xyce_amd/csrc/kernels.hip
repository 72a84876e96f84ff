// Drop-in device-side and linear-side kernels for one circuit, and the C ABI
// that a Xyce DeviceMaster / Linear::Solver shim binds (include/xyce_amd.h).
//
// Data layout in HBM (one circuit):
//   MOSFET constants   class table [n_cls][32] + per device tVbi, Beta, tVto
//   node indices       int32 [n_mos][6] (d g s b d' s'; n = ground slot)
//   store / state      SoA planes [6][n_mos] / [8][n_mos]
//   record terms       SoA planes [48][n_mos], only planes whose destination
//                      is not ground are written (per-device 64-bit mask)
//   CSR values         dFdx, dQdx, J [nnz]; LU values [nnz(L+U)]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "mos1_device.h"
#include "xg_internal.h"

#define HIPCHK(x)                                    \
  do {                                               \
    hipError_t e_ = (x);                             \
    if (e_ != hipSuccess) return XG_E_HIP;           \
  } while (0)

namespace xg {

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t cnt) {
    if (p) { (void)hipFree(p); p = nullptr; }
    n = cnt;
    if (cnt == 0) return hipSuccess;
    return hipMalloc((void**)&p, cnt * sizeof(T));
  }
  hipError_t upload(const std::vector<T>& v, hipStream_t s) {
    hipError_t e = alloc(v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
  }
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
struct EvalArgs {
  int n_mos;
  const double* cls; const int* mcls; const double* tvbi; const double* beta; const double* tvto;
  const int* lid; const unsigned long long* mask;
  const double* x;
  const double* sto_cur; const double* sto_nxt_in; const double* sta_cur;
  double* sto_out; double* sta_out;
  double* rec;
  int* limited;
  MosFlags fl;
};

__global__ void __launch_bounds__(256) k_mos1_eval(EvalArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_mos) return;
  const int N = a.n_mos;
  const double* k = a.cls + (size_t)a.mcls[i] * KSTRIDE;
  double volt[6], sc[6], sn[6], ac[8];
  for (int j = 0; j < 6; ++j) volt[j] = a.x[a.lid[6 * i + j]];
  for (int j = 0; j < 6; ++j) { sc[j] = a.sto_cur[(size_t)j * N + i]; sn[j] = a.sto_nxt_in[(size_t)j * N + i]; }
  for (int j = 0; j < 8; ++j) ac[j] = a.sta_cur[(size_t)j * N + i];
  const unsigned long long m = a.mask[i];
  double* rec = a.rec;
  auto emit = [&](int plane, double v) { if (m >> plane & 1ull) rec[(size_t)plane * N + i] = v; };
  // the store and state inputs are all in registers before the first output
  // is written (the outputs may be the input arrays)
  auto keep = [&](int j, double v) {
    if (j < 6) a.sto_out[(size_t)j * N + i] = v;
    else a.sta_out[(size_t)(j - 6) * N + i] = v;
  };
  const int lim = mos1_eval(k, a.tvbi[i], a.beta[i], a.tvto[i], volt, sc, sn, ac, a.fl, keep, emit);
  if (lim) atomicOr(a.limited, 1);
}

struct SrcArgs {
  const int* bg_ptr; const int* bg_ent;  // per row: entries (src*2 + neg)
  const int* src_wave; const int* wave_type; const int* wave_off; const double* wave_data;
  double time, tstep, tstop;
};

__device__ __forceinline__ double src_val(const SrcArgs& s, int src) {
  const int w = s.src_wave[src];
  const int off = s.wave_off[w];
  return wave_value(s.wave_type[w], s.wave_data + off, s.wave_off[w + 1] - off, s.time, s.tstep, s.tstop);
}

// Rows and CSR slots that sum more than this many terms (a supply node
// collects a term from every PMOS) are summed by a whole workgroup.
constexpr int kGatherLong = 64;

// sum over the workgroup in a fixed order (lane-strided partial sums, a
// fixed shuffle tree per wavefront, wavefronts in index order): the same
// bits on every call
__device__ __forceinline__ double wg_sum_fixed(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; ++w) t += sh[w];
  return t;
}

// Per row: F/Q/dFdxdVp/dQdxdVp from the record terms (device order), then the
// linear part's row sums (FilteredMatrix::axpy, N_LAS_FilteredMatrix.C:511-536),
// then B from the sources (DeviceMgr::loadBVectorsforSources).  Long rows are
// left to k_gather_vectors_long.
__global__ void k_gather_vectors(int n, int n_mos, const int* __restrict__ vg_ptr, const int* __restrict__ vg_ent,
                                 const double* __restrict__ rec, const int* __restrict__ rowptr,
                                 const int* __restrict__ colind, const double* __restrict__ linF,
                                 const double* __restrict__ linQ, const double* __restrict__ x, SrcArgs s,
                                 double* F, double* Q, double* B, double* FL, double* QL) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  if (vg_ptr[r + 1] - vg_ptr[r] + rowptr[r + 1] - rowptr[r] > kGatherLong) return;
  double f = 0.0, q = 0.0, fl = 0.0, ql = 0.0;
  for (int e = vg_ptr[r]; e < vg_ptr[r + 1]; ++e) {
    const int ent = vg_ent[e];
    const int dev = ent >> 6, plane = ent & 63;
    const double v = rec[(size_t)plane * n_mos + dev];
    const int vid = d_rec_vec_id(plane);
    if (vid == 0) f += v; else if (vid == 1) q += v; else if (vid == 2) fl += v; else ql += v;
  }
  double sf = 0.0, sq = 0.0;
  for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
    const double xv = x[colind[k]];
    if (linF[k] != 0.0) sf += linF[k] * xv;
    if (linQ[k] != 0.0) sq += linQ[k] * xv;
  }
  f += sf; q += sq;
  double b = 0.0;
  for (int e = s.bg_ptr[r]; e < s.bg_ptr[r + 1]; ++e) {
    const int ent = s.bg_ent[e];
    const double v = src_val(s, ent >> 1);
    if (ent & 1) b -= v; else b += v;
  }
  F[r] = f; Q[r] = q; B[r] = b; FL[r] = fl; QL[r] = ql;
}

// Long rows and slots are summed in two passes: chunks of at most
// kGatherChunk terms, one workgroup each, into partial sums (fixed-order
// workgroup reduction), then one thread per long entry adds its chunks'
// partials in chunk order.  Chunk c covers terms [cb[c], ce[c]) of entry
// ci[c]; kind 0 = record terms, 1 = linear CSR entries (vectors only).
constexpr int kGatherChunk = 2048;

__global__ void __launch_bounds__(256) k_gather_vectors_part(
    int n_mos, const int* __restrict__ ck, const int* __restrict__ cb, const int* __restrict__ ce,
    const int* __restrict__ vg_ent, const double* __restrict__ rec, const int* __restrict__ colind,
    const double* __restrict__ linF, const double* __restrict__ linQ, const double* __restrict__ x, double* part) {
  __shared__ double sh[16];
  const int c = blockIdx.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (ck[c] == 0) {
    for (int e = cb[c] + (int)threadIdx.x; e < ce[c]; e += blockDim.x) {
      const int ent = vg_ent[e];
      const int dev = ent >> 6, plane = ent & 63;
      const double v = rec[(size_t)plane * n_mos + dev];
      const int vid = d_rec_vec_id(plane);
      if (vid == 0) a0 += v; else if (vid == 1) a1 += v; else if (vid == 2) a2 += v; else a3 += v;
    }
  } else {
    for (int k = cb[c] + (int)threadIdx.x; k < ce[c]; k += blockDim.x) {
      const double xv = x[colind[k]];
      if (linF[k] != 0.0) a0 += linF[k] * xv;
      if (linQ[k] != 0.0) a1 += linQ[k] * xv;
    }
  }
  a0 = wg_sum_fixed(a0, sh); a1 = wg_sum_fixed(a1, sh); a2 = wg_sum_fixed(a2, sh); a3 = wg_sum_fixed(a3, sh);
  if (threadIdx.x == 0) {
    part[4 * c] = a0; part[4 * c + 1] = a1; part[4 * c + 2] = a2; part[4 * c + 3] = a3;
  }
}

// one wavefront per long row: record-term chunks [rc[3l], rc[3l+1]), CSR
// chunks [rc[3l+1], rc[3l+2]) of the partial sums (lane-strided, fixed
// shuffle tree), then the sources
__device__ __forceinline__ double wave_sum_fixed(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void __launch_bounds__(64) k_gather_vectors_fin(const int* __restrict__ rows, const int* __restrict__ rc,
                                                          const double* __restrict__ part, SrcArgs s, double* F,
                                                          double* Q, double* B, double* FL, double* QL) {
  const int l = blockIdx.x, lane = threadIdx.x;
  const int r = rows[l];
  double f = 0.0, q = 0.0, fl = 0.0, ql = 0.0, sf = 0.0, sq = 0.0;
  for (int c = rc[3 * l] + lane; c < rc[3 * l + 1]; c += 64) {
    f += part[4 * c]; q += part[4 * c + 1]; fl += part[4 * c + 2]; ql += part[4 * c + 3];
  }
  for (int c = rc[3 * l + 1] + lane; c < rc[3 * l + 2]; c += 64) { sf += part[4 * c]; sq += part[4 * c + 1]; }
  f = wave_sum_fixed(f); q = wave_sum_fixed(q); fl = wave_sum_fixed(fl); ql = wave_sum_fixed(ql);
  sf = wave_sum_fixed(sf); sq = wave_sum_fixed(sq);
  if (lane != 0) return;
  f += sf; q += sq;
  double b = 0.0;
  for (int e = s.bg_ptr[r]; e < s.bg_ptr[r + 1]; ++e) {
    const int ent = s.bg_ent[e];
    const double v = src_val(s, ent >> 1);
    if (ent & 1) b -= v; else b += v;
  }
  F[r] = f; Q[r] = q; B[r] = b; FL[r] = fl; QL[r] = ql;
}

#pragma clang fp contract(off)
__device__ __forceinline__ double jac_combine_k(double a, double dq, double b, double df) { return a * dq + b * df; }
#pragma clang fp contract(fast)

// Per CSR slot: dFdx, dQdx from the record terms, + the linear matrices
// (FilteredMatrix::addToMatrix), optionally J = a*dQdx + b*dFdx
// (EpetraMatrix::linearCombo N_LAS_EpetraMatrix.C:629) scattered to LU storage.
__global__ void k_gather_matrices(int nnz, int n_mos, const int* __restrict__ mg_ptr, const int* __restrict__ mg_ent,
                                  const double* __restrict__ rec, const double* __restrict__ linF,
                                  const double* __restrict__ linQ, double* dFdx, double* dQdx, double a, double b,
                                  double* J, double* lu, const int* __restrict__ a2lu) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  if (mg_ptr[s + 1] - mg_ptr[s] > kGatherLong) return;   // k_gather_matrices_long
  double df = 0.0, dq = 0.0;
  for (int e = mg_ptr[s]; e < mg_ptr[s + 1]; ++e) {
    const int ent = mg_ent[e];
    const int dev = ent >> 6, plane = ent & 63;
    const double v = rec[(size_t)plane * n_mos + dev];
    if (plane < R_JQ0) df += v; else dq += v;
  }
  if (linF[s] != 0.0) df += linF[s];
  if (linQ[s] != 0.0) dq += linQ[s];
  if (dFdx) { dFdx[s] = df; dQdx[s] = dq; }
  if (J || lu) {
    const double j = jac_combine_k(a, dq, b, df);
    if (J) J[s] = j;
    if (lu) lu[a2lu[s]] = j;
  }
}

// Long CSR slots: chunk partial sums of dFdx and dQdx terms ...
__global__ void __launch_bounds__(256) k_gather_matrices_part(int n_mos, const int* __restrict__ cb,
                                                              const int* __restrict__ ce, const int* __restrict__ mg_ent,
                                                              const double* __restrict__ rec, double* part) {
  __shared__ double sh[16];
  const int c = blockIdx.x;
  double df = 0.0, dq = 0.0;
  for (int e = cb[c] + (int)threadIdx.x; e < ce[c]; e += blockDim.x) {
    const int ent = mg_ent[e];
    const int dev = ent >> 6, plane = ent & 63;
    const double v = rec[(size_t)plane * n_mos + dev];
    if (plane < R_JQ0) df += v; else dq += v;
  }
  df = wg_sum_fixed(df, sh);
  dq = wg_sum_fixed(dq, sh);
  if (threadIdx.x == 0) { part[2 * c] = df; part[2 * c + 1] = dq; }
}

// ... then one wavefront per long slot over its chunks [sc[l], sc[l + 1])
__global__ void __launch_bounds__(64) k_gather_matrices_fin(const int* __restrict__ slots, const int* __restrict__ sc,
                                                           const double* __restrict__ part,
                                                           const double* __restrict__ linF,
                                                           const double* __restrict__ linQ, double* dFdx, double* dQdx,
                                                           double a, double b, double* J, double* lu,
                                                           const int* __restrict__ a2lu) {
  const int l = blockIdx.x, lane = threadIdx.x;
  const int s = slots[l];
  double df = 0.0, dq = 0.0;
  for (int c = sc[l] + lane; c < sc[l + 1]; c += 64) { df += part[2 * c]; dq += part[2 * c + 1]; }
  df = wave_sum_fixed(df);
  dq = wave_sum_fixed(dq);
  if (lane != 0) return;
  if (linF[s] != 0.0) df += linF[s];
  if (linQ[s] != 0.0) dq += linQ[s];
  if (dFdx) { dFdx[s] = df; dQdx[s] = dq; }
  if (J || lu) {
    const double j = jac_combine_k(a, dq, b, df);
    if (J) J[s] = j;
    if (lu) lu[a2lu[s]] = j;
  }
}

// OneStep::obtainResidual (N_TIA_OneStep.C:219-289) / NoTimeIntegration (DC).
#pragma clang fp contract(off)
__global__ void k_combine_rhs(int n, int dcop, int order, double h, const double* F, const double* Q, const double* B,
                              const double* FL, const double* QL, const double* q0, const double* q2, double* rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r;
  if (dcop) {
    r = 1.0 * F[i] + -1.0 * B[i];
    r *= -1.0;
    r = r + 1.0 * FL[i];
  } else {
    r = 1.0 * Q[i] + -1.0 * q0[i];
    if (order == 2) {
      r = (1.0 / h) * r + 0.5 * F[i] + -0.5 * B[i];
      r = r + 0.5 * q2[i];
    } else {
      r = (1.0 / h) * r + 1.0 * F[i] + -1.0 * B[i];
    }
    r *= -1.0;
    const double dq = (1.0 / h) * QL[i];
    r = r + 1.0 * dq;
    r = r + (order == 2 ? 0.5 : 1.0) * FL[i];
  }
  rhs[i] = r;
}

__global__ void k_combine_jac(int nnz, double a, double b, const double* dFdx, const double* dQdx, double* J) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  J[s] = a * dQdx[s] + b * dFdx[s];
}
#pragma clang fp contract(fast)

template <class T>
__global__ void k_fill(T* p, size_t n, T v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// LU kernels are templated on the factor type T: double for the plain path,
// float for the FP32 factorisation that FP64 iterative refinement corrects.
template <class T>
__global__ void k_scatter_a(int nnz, const double* vals, const int* a2lu, T* lu) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nnz) lu[a2lu[s]] = (T)vals[s];
}

// Entry-level scheduled LU (Doolittle order per entry: a_ij - sum_k l_ik u_kj,
// k ascending, then / u_jj below the diagonal), one level per barrier.
// One workgroup; LU values in global memory (L2 resident for these sizes).
template <class T>
__global__ void __launch_bounds__(1024) k_lu_factor_1wg(int nlev, const int* __restrict__ lvl_ptr,
                                                        const int* __restrict__ tgt, const int* __restrict__ op_ptr,
                                                        const int* __restrict__ ops, const int* __restrict__ div,
                                                        T* lu) {
  for (int l = 0; l < nlev; ++l) {
    for (int t = lvl_ptr[l] + threadIdx.x; t < lvl_ptr[l + 1]; t += blockDim.x) {
      const int p = tgt[t];
      T v = lu[p];
      for (int o = op_ptr[t]; o < op_ptr[t + 1]; ++o) v -= lu[ops[2 * o]] * lu[ops[2 * o + 1]];
      const int d = div[t];
      if (d >= 0) v /= lu[d];
      lu[p] = v;
    }
    __syncthreads();
  }
}

template <class T>
__global__ void k_lu_factor_level(int t0, int t1, const int* __restrict__ tgt, const int* __restrict__ op_ptr,
                                  const int* __restrict__ ops, const int* __restrict__ div, T* lu) {
  const int t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= t1) return;
  const int p = tgt[t];
  T v = lu[p];
  for (int o = op_ptr[t]; o < op_ptr[t + 1]; ++o) v -= lu[ops[2 * o]] * lu[ops[2 * o + 1]];
  const int d = div[t];
  if (d >= 0) v /= lu[d];
  lu[p] = v;
}

// pivot check: non-finite or zero pivot, or a multiplier beyond KLU's
// threshold (|l_ij| > 1/0.001: KLU would have re-pivoted) -> bad = row + 1
template <class T>
__global__ void k_lu_pivots(int n, const int* lu_rowptr, const int* diag, const T* lu, int* bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = (double)lu[diag[i]];
  const bool hard = !(fabs(d) > 0.0) || !isfinite(d);
  bool b = hard;
  for (int p = lu_rowptr[i]; p < diag[i] && !b; ++p) b = !(fabs((double)lu[p]) <= 1.0e3);
  if (b) atomicMin(bad, i + 1);
  if (hard) atomicOr(bad + 1, 1);   // a zero or non-finite pivot (else only the threshold)
}

// Level-scheduled triangular solves; b is permuted on the way in and x on the
// way out (B = P_r A P_c).  One workgroup; arithmetic in T.
template <class T>
__global__ void __launch_bounds__(1024) k_lu_solve_1wg(int n, const int* __restrict__ lu_rowptr,
                                                       const int* __restrict__ lu_colind, const int* __restrict__ diag,
                                                       const T* __restrict__ lu, int nfl,
                                                       const int* __restrict__ fl_ptr, const int* __restrict__ fl_rows,
                                                       int nbl, const int* __restrict__ bl_ptr,
                                                       const int* __restrict__ bl_rows, const int* __restrict__ prow,
                                                       const int* __restrict__ pcol, const double* __restrict__ b,
                                                       T* y, double* x, int* nan_flag) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) y[i] = (T)b[prow[i]];
  __syncthreads();
  for (int l = 0; l < nfl; ++l) {
    for (int t = fl_ptr[l] + threadIdx.x; t < fl_ptr[l + 1]; t += blockDim.x) {
      const int i = fl_rows[t];
      T v = y[i];
      for (int p = lu_rowptr[i]; p < diag[i]; ++p) v -= lu[p] * y[lu_colind[p]];
      y[i] = v;
    }
    __syncthreads();
  }
  for (int l = 0; l < nbl; ++l) {
    for (int t = bl_ptr[l] + threadIdx.x; t < bl_ptr[l + 1]; t += blockDim.x) {
      const int i = bl_rows[t];
      T v = y[i];
      for (int p = diag[i] + 1; p < lu_rowptr[i + 1]; ++p) v -= lu[p] * y[lu_colind[p]];
      y[i] = v / lu[diag[i]];
    }
    __syncthreads();
  }
  int bad = 0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const double v = (double)y[j];
    if (!isfinite(v)) bad = 1;
    x[pcol[j]] = v;
  }
  if (bad) atomicOr(nan_flag, 1);
}

// transpose solves B^T z = c = U^T L^T z: U^T by the backward levels in reverse
// (w_j = (c_j - sum_i U_ij w_i) / U_jj over column j of U), then L^T by the
// forward levels in reverse (z_i = w_i - sum_j L_ji z_j over column i of L);
// a column's entries are the CSR positions listed in cptr/cpos with their rows
// in crow
__global__ void k_solve_ut_level(int t0, int t1, const int* rows, const int* cptr, const int* cpos, const int* crow,
                                 const int* diag, const double* lu, double* y) {
  const int t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= t1) return;
  const int j = rows[t];
  double v = y[j];
  for (int k = cptr[j]; k < cptr[j + 1]; ++k) v -= lu[cpos[k]] * y[crow[k]];
  y[j] = v / lu[diag[j]];
}
__global__ void k_solve_lt_level(int t0, int t1, const int* rows, const int* cptr, const int* cpos, const int* crow,
                                 const double* lu, double* y) {
  const int t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= t1) return;
  const int i = rows[t];
  double v = y[i];
  for (int k = cptr[i]; k < cptr[i + 1]; ++k) v -= lu[cpos[k]] * y[crow[k]];
  y[i] = v;
}

template <class T>
__global__ void k_solve_fwd_level(int t0, int t1, const int* fl_rows, const int* lu_rowptr, const int* lu_colind,
                                  const int* diag, const T* lu, T* y) {
  const int t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= t1) return;
  const int i = fl_rows[t];
  T v = y[i];
  for (int p = lu_rowptr[i]; p < diag[i]; ++p) v -= lu[p] * y[lu_colind[p]];
  y[i] = v;
}
template <class T>
__global__ void k_solve_bwd_level(int t0, int t1, const int* bl_rows, const int* lu_rowptr, const int* lu_colind,
                                  const int* diag, const T* lu, T* y) {
  const int t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= t1) return;
  const int i = bl_rows[t];
  T v = y[i];
  for (int p = diag[i] + 1; p < lu_rowptr[i + 1]; ++p) v -= lu[p] * y[lu_colind[p]];
  y[i] = v / lu[diag[i]];
}
template <class T>
__global__ void k_permute_in(int n, const int* prow, const double* b, T* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (T)b[prow[i]];
}
template <class T>
__global__ void k_permute_out(int n, const int* pcol, const T* y, double* x, int* nan_flag) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) {
    const double v = (double)y[j];
    x[pcol[j]] = v;
    if (!isfinite(v)) atomicOr(nan_flag, 1);
  }
}

// Iterative refinement (FP64): r = b - A x over the caller's CSR (one row per
// thread), with per-block max|r| and max|b| folded into red[0..1] as ordered
// non-negative doubles (their bit patterns compare like the values).
__global__ void k_residual_f64(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                               const double* __restrict__ a, const double* __restrict__ x,
                               const double* __restrict__ b, double* __restrict__ r,
                               unsigned long long* red) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ar = 0.0, ab = 0.0;
  if (i < n && rp[i + 1] - rp[i] <= kGatherLong) {   // longer rows: k_spmv_part + k_residual_fin
    double s = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) s += a[k] * x[ci[k]];
    const double v = b[i] - s;
    r[i] = v;
    ar = fabs(v);
    ab = fabs(b[i]);
    if (!isfinite(ar)) ar = INFINITY;
  }
  for (int o = 32; o > 0; o >>= 1) {
    ar = fmax(ar, __shfl_xor(ar, o, 64));
    ab = fmax(ab, __shfl_xor(ab, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&red[0], (unsigned long long)__double_as_longlong(ar));
    atomicMax(&red[1], (unsigned long long)__double_as_longlong(ab));
  }
}
// rows of A with more than kGatherLong entries: chunk partial products
// (fixed-order workgroup reduction) ...
__global__ void __launch_bounds__(256) k_spmv_part(const int* __restrict__ cb, const int* __restrict__ ce,
                                                   const int* __restrict__ ci, const double* __restrict__ a,
                                                   const double* __restrict__ x, double* part) {
  __shared__ double sh[16];
  const int c = blockIdx.x;
  double acc = 0.0;
  for (int k = cb[c] + (int)threadIdx.x; k < ce[c]; k += blockDim.x) acc += a[k] * x[ci[k]];
  acc = wg_sum_fixed(acc, sh);
  if (threadIdx.x == 0) part[c] = acc;
}

// ... then one wavefront per long row: its chunks in a fixed order, r = b - Ax
__global__ void __launch_bounds__(64) k_residual_fin(const int* __restrict__ rows, const int* __restrict__ rc,
                                                     const double* __restrict__ part, const double* __restrict__ b,
                                                     double* __restrict__ r, unsigned long long* red) {
  const int l = blockIdx.x, lane = threadIdx.x;
  double acc = 0.0;
  for (int c = rc[l] + lane; c < rc[l + 1]; c += 64) acc += part[c];
  acc = wave_sum_fixed(acc);
  if (lane != 0) return;
  const int i = rows[l];
  const double v = b[i] - acc;
  r[i] = v;
  double ar = fabs(v);
  if (!isfinite(ar)) ar = INFINITY;
  atomicMax(&red[0], (unsigned long long)__double_as_longlong(ar));
  atomicMax(&red[1], (unsigned long long)__double_as_longlong(fabs(b[i])));
}

__global__ void k_axpy_f64(int n, const double* d, double* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += d[i];
}
__global__ void k_axpy_neg_f64(int n, const double* d, double* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] -= d[i];
}

// state permutation between the Xyce layout and the SoA planes
__global__ void k_sta_to_soa(int n_mos, int nslots, const int* off0, const double* xy, double* soa) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_mos) return;
  for (int j = 0; j < nslots; ++j) soa[(size_t)j * n_mos + i] = xy[off0[i] + j];
}
__global__ void k_soa_to_sta(int n_mos, int nslots, const int* off0, const double* soa, double* xy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_mos) return;
  for (int j = 0; j < nslots; ++j) xy[off0[i] + j] = soa[(size_t)j * n_mos + i];
}

inline unsigned nblk(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

// ---- the work area of xg_lu_solve_noalloc: one file-scope array per GPU,
// held from module load until the process ends.  Host code never uses the
// name `scratch` as an address (it is a host-side stand-in there): the
// device address is looked up once per device and kept in scratch_dev.
constexpr size_t kScratchDoubles = (size_t)4 << 20;   // 32 MiB
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];
static std::mutex scratch_mu;
static double* scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (hipGetSymbolAddress,
// once per device, with that device current).
static int scratch_address(double** out) {
  *out = nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return XG_E_HIP;
  if (dev < 0 || dev >= kMaxDevices) return XG_E_ARG;
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return XG_E_HIP;
    scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = scratch_dev[dev];
  return XG_OK;
}

}  // namespace xg

// ---------------------------------------------------------------------------
// host objects
// ---------------------------------------------------------------------------
using namespace xg;

struct xg_lu {
  LuSymbolic s;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  DBuf<int> a2lu, lu_rowptr, lu_colind, lu_diag, f_lvl_ptr, f_tgt, f_op_ptr, f_ops, f_div;
  DBuf<int> fs_ptr, fs_rows, bs_ptr, bs_rows, prow, pcol, flag, arp, aci;
  // column lists of U (above the diagonal) and L (below it) for transpose solves, built on first use
  DBuf<int> ut_ptr, ut_pos, ut_row, lt_ptr, lt_pos, lt_row;
  bool have_t = false;
  DBuf<int> long_rows, lcb, lce, lrc;   // rows of A with many entries: chunks for the residual
  DBuf<double> lpart;
  int n_long = 0, n_lchunks = 0;
  DBuf<double> lu, avals, y, b, x, r, d;
  DBuf<float> lu32, y32;
  DBuf<unsigned long long> red;
  std::vector<int> h_flvl, h_fs, h_bs;
  bool one_wg = true;
  std::vector<int> a_rowptr, a_colind;
  // precision: 0 = FP64 factors; 1 = FP32 factors + FP64 iterative refinement
  int mode = 0;
  double ir_tol = 1e-12;
  int ir_max = 10;
  bool have64 = false;    // FP64 factors of the current values exist (refinement fallback)
  bool soft64 = false;    // those factors have multipliers beyond KLU's threshold: solves refine
  bool repivot_futile = false;  // a re-pivot left threshold failures: keep such factors at once
  double stats[6] = {0, 0, 0, 0, 0, 0};  // last solve: refinements, rel. residual; totals: solves, refinements, fallbacks, FP32 pivot failures
  // method: 1 = level-scheduled entry LU (small systems), 2 = multifrontal
  int method = 1;
  std::unique_ptr<MfLu> mfp;
  // device address of the static work area on this handle's device (set up
  // with the handle; null if the lookup failed)
  double* scratch_work = nullptr;

  // symbolic analysis for the chosen (or automatically picked) method
  int analyse(const double* values, int want_method) {
    LuSymbolic t;
    int st = lu_analyse(s.n ? s.n : (int)a_rowptr.size() - 1, a_rowptr.data(), a_colind.data(), values, t, false, true);
    if (st != XG_OK) return st;
    const int n = t.n;
    int m = want_method;
    if (m == 0) m = (t.flops > 2e7 || n > 65535 || t.nlu > (1 << 24)) ? 2 : 1;
    if (m == 1) {
      st = lu_analyse(n, a_rowptr.data(), a_colind.data(), values, t, true, false);
      if (st != XG_OK) return st;
    }
    s = std::move(t);
    method = m;
    return XG_OK;
  }

  // rows of A longer than kGatherLong (a supply node's) are summed in chunks
  int setup_long_rows() {
    std::vector<int> lr, cb, ce, rc{0};
    const int n = (int)a_rowptr.size() - 1;
    for (int i = 0; i < n; ++i) {
      if (a_rowptr[i + 1] - a_rowptr[i] <= kGatherLong) continue;
      lr.push_back(i);
      for (int k = a_rowptr[i]; k < a_rowptr[i + 1]; k += kGatherChunk) {
        cb.push_back(k);
        ce.push_back(std::min(a_rowptr[i + 1], k + kGatherChunk));
      }
      rc.push_back((int)cb.size());
    }
    n_long = (int)lr.size();
    n_lchunks = (int)cb.size();
    if (n_long) {
      HIPCHK(long_rows.upload(lr, stream)); HIPCHK(lcb.upload(cb, stream)); HIPCHK(lce.upload(ce, stream));
      HIPCHK(lrc.upload(rc, stream)); HIPCHK(lpart.alloc(cb.size()));
    }
    return XG_OK;
  }

  int upload() {
    hipStream_t st = stream;
    {
      const int rc = setup_long_rows();
      if (rc != XG_OK) return rc;
    }
    if (method == 2) {
      HIPCHK(prow.upload(s.perm_row, st)); HIPCHK(pcol.upload(s.perm_col, st));
      HIPCHK(arp.upload(a_rowptr, st)); HIPCHK(aci.upload(a_colind, st));
      HIPCHK(flag.alloc(2)); HIPCHK(red.alloc(2));
      HIPCHK(avals.alloc(s.nnz_a)); HIPCHK(b.alloc(s.n)); HIPCHK(x.alloc(s.n)); HIPCHK(r.alloc(s.n)); HIPCHK(d.alloc(s.n));
      mfp = std::make_unique<MfLu>();
      const int rc = mfp->setup(s, a_rowptr.data(), a_colind.data(), (void*)stream, true);
      if (rc != XG_OK) return rc;
      // the plan holds what the kernels need; drop the host-side structure
      std::vector<std::vector<int>>().swap(s.lcol);
      have64 = false;
      return XG_OK;
    }
    mfp.reset();
    HIPCHK(a2lu.upload(s.a2lu, st)); HIPCHK(lu_rowptr.upload(s.lu_rowptr, st));
    HIPCHK(lu_colind.upload(s.lu_colind, st)); HIPCHK(lu_diag.upload(s.lu_diag, st));
    HIPCHK(f_lvl_ptr.upload(s.f_lvl_ptr, st)); HIPCHK(f_tgt.upload(s.f_tgt, st));
    HIPCHK(f_op_ptr.upload(s.f_op_ptr, st)); HIPCHK(f_ops.upload(s.f_ops, st)); HIPCHK(f_div.upload(s.f_div, st));
    HIPCHK(fs_ptr.upload(s.fs_lvl_ptr, st)); HIPCHK(fs_rows.upload(s.fs_rows, st));
    HIPCHK(bs_ptr.upload(s.bs_lvl_ptr, st)); HIPCHK(bs_rows.upload(s.bs_rows, st));
    HIPCHK(prow.upload(s.perm_row, st)); HIPCHK(pcol.upload(s.perm_col, st));
    HIPCHK(arp.upload(a_rowptr, st)); HIPCHK(aci.upload(a_colind, st));
    HIPCHK(flag.alloc(2)); HIPCHK(red.alloc(2));
    HIPCHK(lu.alloc(s.nlu)); HIPCHK(avals.alloc(s.nnz_a)); HIPCHK(y.alloc(s.n)); HIPCHK(b.alloc(s.n));
    HIPCHK(x.alloc(s.n)); HIPCHK(r.alloc(s.n)); HIPCHK(d.alloc(s.n));
    if (mode == 1) { HIPCHK(lu32.alloc(s.nlu)); HIPCHK(y32.alloc(s.n)); }
    h_flvl = s.f_lvl_ptr; h_fs = s.fs_lvl_ptr; h_bs = s.bs_lvl_ptr;
    have_t = false;   // a new pattern: the transpose column lists are rebuilt when next needed
    const int nlev = (int)s.f_lvl_ptr.size() - 1;
    one_wg = s.nlu <= (1 << 20) && nlev <= 4096 && s.n <= (1 << 16);
    have64 = false;
    return XG_OK;
  }

  // numeric factorisation of the values in `avals`, into FP64 or FP32 storage
  template <class T>
  int factor_device(T* f) {
    const int nlev = (int)h_flvl.size() - 1;
    k_fill<T><<<nblk(s.nlu), 256, 0, stream>>>(f, s.nlu, (T)0);
    k_scatter_a<T><<<nblk(s.nnz_a), 256, 0, stream>>>(s.nnz_a, avals.p, a2lu.p, f);
    if (one_wg) {
      k_lu_factor_1wg<T><<<1, 1024, 0, stream>>>(nlev, f_lvl_ptr.p, f_tgt.p, f_op_ptr.p, f_ops.p, f_div.p, f);
    } else {
      for (int l = 0; l < nlev; ++l) {
        const int t0 = h_flvl[l], t1 = h_flvl[l + 1];
        if (t1 > t0) k_lu_factor_level<T><<<nblk(t1 - t0), 256, 0, stream>>>(t0, t1, f_tgt.p, f_op_ptr.p, f_ops.p, f_div.p, f);
      }
    }
    HIPCHK(hipGetLastError());
    return XG_OK;
  }
  template <class T>
  int check_pivots(const T* f, int* bad_out, int* hard_out) {
    int init[2] = {0x7fffffff, 0};
    HIPCHK(hipMemcpyAsync(flag.p, init, sizeof(init), hipMemcpyHostToDevice, stream));
    k_lu_pivots<T><<<nblk(s.n), 256, 0, stream>>>(s.n, lu_rowptr.p, lu_diag.p, f, flag.p);
    int res[2];
    HIPCHK(hipMemcpyAsync(res, flag.p, sizeof(res), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    *bad_out = res[0] == 0x7fffffff ? 0 : res[0];
    *hard_out = res[1];
    return XG_OK;
  }
  // column lists of the factors' pattern for transpose solves (level path)
  int setup_transpose() {
    const int n = s.n;
    std::vector<int> uptr(n + 1, 0), lptr(n + 1, 0);
    for (int i = 0; i < n; ++i)
      for (int p = s.lu_rowptr[i]; p < s.lu_rowptr[i + 1]; ++p) {
        const int j = s.lu_colind[p];
        if (j > i) uptr[j + 1]++;
        else if (j < i) lptr[j + 1]++;
      }
    for (int j = 0; j < n; ++j) { uptr[j + 1] += uptr[j]; lptr[j + 1] += lptr[j]; }
    std::vector<int> upos(uptr[n]), urow(uptr[n]), lpos(lptr[n]), lrow(lptr[n]);
    std::vector<int> uf(uptr.begin(), uptr.end() - 1), lf(lptr.begin(), lptr.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int p = s.lu_rowptr[i]; p < s.lu_rowptr[i + 1]; ++p) {
        const int j = s.lu_colind[p];
        if (j > i) { upos[uf[j]] = p; urow[uf[j]++] = i; }
        else if (j < i) { lpos[lf[j]] = p; lrow[lf[j]++] = i; }
      }
    hipStream_t st = stream;
    HIPCHK(ut_ptr.upload(uptr, st)); HIPCHK(ut_pos.upload(upos, st)); HIPCHK(ut_row.upload(urow, st));
    HIPCHK(lt_ptr.upload(lptr, st)); HIPCHK(lt_pos.upload(lpos, st)); HIPCHK(lt_row.upload(lrow, st));
    HIPCHK(hipStreamSynchronize(st));
    have_t = true;
    return XG_OK;
  }
  // A^T x = b with the FP64 factors of B = P_r A P_c: B^T (P_r x) = P_c^T b
  int solve_transpose(const double* db, double* dx, int* nan_flag) {
    const int n = s.n;
    if (!have_t) {
      const int rc = setup_transpose();
      if (rc != XG_OK) return rc;
    }
    k_permute_in<double><<<nblk(n), 256, 0, stream>>>(n, pcol.p, db, y.p);
    for (size_t l = h_bs.size() - 1; l-- > 0;)
      if (h_bs[l + 1] > h_bs[l])
        k_solve_ut_level<<<nblk(h_bs[l + 1] - h_bs[l]), 256, 0, stream>>>(h_bs[l], h_bs[l + 1], bs_rows.p, ut_ptr.p,
                                                                          ut_pos.p, ut_row.p, lu_diag.p, lu.p, y.p);
    for (size_t l = h_fs.size() - 1; l-- > 0;)
      if (h_fs[l + 1] > h_fs[l])
        k_solve_lt_level<<<nblk(h_fs[l + 1] - h_fs[l]), 256, 0, stream>>>(h_fs[l], h_fs[l + 1], fs_rows.p, lt_ptr.p,
                                                                          lt_pos.p, lt_row.p, lu.p, y.p);
    k_permute_out<double><<<nblk(n), 256, 0, stream>>>(n, prow.p, y.p, dx, nan_flag);
    HIPCHK(hipGetLastError());
    return XG_OK;
  }
  // solve with the factors f and device vectors db -> dx (arithmetic in T),
  // yy the permuted work vector; *nan_flag collects non-finite results
  template <class T>
  int solve_device(const T* f, T* yy, const double* db, double* dx, int* nan_flag) {
    const int n = s.n;
    if (one_wg) {
      k_lu_solve_1wg<T><<<1, 1024, 0, stream>>>(n, lu_rowptr.p, lu_colind.p, lu_diag.p, f, (int)h_fs.size() - 1,
                                                fs_ptr.p, fs_rows.p, (int)h_bs.size() - 1, bs_ptr.p, bs_rows.p,
                                                prow.p, pcol.p, db, yy, dx, nan_flag);
    } else {
      k_permute_in<T><<<nblk(n), 256, 0, stream>>>(n, prow.p, db, yy);
      for (size_t l = 0; l + 1 < h_fs.size(); ++l)
        if (h_fs[l + 1] > h_fs[l])
          k_solve_fwd_level<T><<<nblk(h_fs[l + 1] - h_fs[l]), 256, 0, stream>>>(h_fs[l], h_fs[l + 1], fs_rows.p, lu_rowptr.p, lu_colind.p, lu_diag.p, f, yy);
      for (size_t l = 0; l + 1 < h_bs.size(); ++l)
        if (h_bs[l + 1] > h_bs[l])
          k_solve_bwd_level<T><<<nblk(h_bs[l + 1] - h_bs[l]), 256, 0, stream>>>(h_bs[l], h_bs[l + 1], bs_rows.p, lu_rowptr.p, lu_colind.p, lu_diag.p, f, yy);
      k_permute_out<T><<<nblk(n), 256, 0, stream>>>(n, pcol.p, yy, dx, nan_flag);
    }
    HIPCHK(hipGetLastError());
    return XG_OK;
  }
  // either method: factor the values in `avals` (*bad = failing pivot + 1,
  // *hard = 1 if a pivot was zero or not finite)
  int factor_any(bool fp32, int* bad, int* hard) {
    if (method == 2) return mfp->factor(avals.p, s.nnz_a, fp32, bad, hard);
    const int st = fp32 ? factor_device<float>(lu32.p) : factor_device<double>(lu.p);
    if (st != XG_OK) return st;
    return fp32 ? check_pivots<float>(lu32.p, bad, hard) : check_pivots<double>(lu.p, bad, hard);
  }
  int solve_any(bool fp32, const double* db, double* dx) {
    if (method == 2) return mfp->solve(db, dx, fp32, flag.p + 1);
    return fp32 ? solve_device<float>(lu32.p, y32.p, db, dx, flag.p + 1)
                : solve_device<double>(lu.p, y.p, db, dx, flag.p + 1);
  }
  // FP64-factor solve whose work vectors and flag are the caller's device
  // memory `work` (work_doubles() doubles, 256-byte aligned) instead of the
  // handle's buffers; *nan_flag is the first word of the work area
  size_t work_doubles() const {
    const size_t flag_d = 32;                       // one int, padded to 256 bytes
    if (method == 2) return flag_d + mfp->solve_work_doubles();
    return flag_d + (((size_t)s.n + 31) & ~(size_t)31);
  }
  int solve_with_work(const double* db, double* dx, double* work) {
    int* nan_flag = reinterpret_cast<int*>(work);
    double* w = work + 32;
    k_fill<<<1, 1, 0, stream>>>(nan_flag, 1, 0);
    if (method == 2) return mfp->solve_ext(db, dx, nan_flag, w);
    return solve_device<double>(lu.p, w, db, dx, nan_flag);
  }
  // FP64 factors whose multipliers exceeded KLU's threshold: solve, then
  // refine x += A^-1 r against A in FP64 while the residual decreases, up to
  // ir_max corrections or ir_tol; the best iterate is kept
  int solve_refined64(const double* db, double* dx) {
    int st = solve_any(false, db, dx);
    if (st != XG_OK) return st;
    double rel = 0.0, prev = INFINITY;
    int it = 0;
    for (;;) {
      st = residual(db, dx, &rel);
      if (st != XG_OK) return st;
      if (rel <= ir_tol || it >= ir_max || !(rel < prev)) break;
      prev = rel;
      st = solve_any(false, r.p, d.p);
      if (st != XG_OK) return st;
      k_axpy_f64<<<nblk(s.n), 256, 0, stream>>>(s.n, d.p, dx);
      ++it;
    }
    if (!(rel < prev) && it > 0) {
      // the last correction did not lower the residual: take it back
      k_axpy_neg_f64<<<nblk(s.n), 256, 0, stream>>>(s.n, d.p, dx);
    }
    stats[0] = it; stats[1] = std::min(rel, prev); stats[2] += 1; stats[3] += it;
    return XG_OK;
  }
  // r = b - A x; returns max|r| / max|b| (0/0 -> 0)
  int residual(const double* db, const double* dx, double* rel) {
    HIPCHK(hipMemsetAsync(red.p, 0, 2 * sizeof(unsigned long long), stream));
    k_residual_f64<<<nblk(s.n), 256, 0, stream>>>(s.n, arp.p, aci.p, avals.p, dx, db, r.p, red.p);
    if (n_long) {
      k_spmv_part<<<n_lchunks, 256, 0, stream>>>(lcb.p, lce.p, aci.p, avals.p, dx, lpart.p);
      k_residual_fin<<<n_long, 64, 0, stream>>>(long_rows.p, lrc.p, lpart.p, db, r.p, red.p);
    }
    unsigned long long h[2];
    HIPCHK(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    double rn, bn;
    std::memcpy(&rn, &h[0], 8);
    std::memcpy(&bn, &h[1], 8);
    *rel = bn > 0.0 ? rn / bn : (rn > 0.0 ? INFINITY : 0.0);
    return XG_OK;
  }
  // FP32 solve + FP64 refinement of device b -> x; falls back to FP64
  // factors when the residual stops decreasing or ir_max corrections did
  // not reach ir_tol (an ill-conditioned system: the FP32 factors' error
  // times the condition number is near 1, or FP64 itself cannot reach ir_tol)
  int solve_refined(const double* db, double* dx) {
    int st = solve_any(true, db, dx);
    if (st != XG_OK) return st;
    double rel = 0.0, prev = INFINITY;
    int it = 0;
    bool fell_back = false;
    for (;;) {
      st = residual(db, dx, &rel);
      if (st != XG_OK) return st;
      if (rel <= ir_tol) break;
      if (it >= ir_max || !(rel < prev)) {
        // refinement does not converge with the FP32 factors: FP64 factors
        HIPCHK(hipMemsetAsync(flag.p + 1, 0, sizeof(int), stream));
        if (!have64) {
          int bad64 = 0, hard64 = 0;
          st = factor_any(false, &bad64, &hard64);
          if (st != XG_OK) return st;
          have64 = true;
          soft64 = bad64 != 0 && !hard64;
        }
        st = solve_any(false, db, dx);
        if (st != XG_OK) return st;
        st = residual(db, dx, &rel);
        if (st != XG_OK) return st;
        fell_back = true;
        break;
      }
      prev = rel;
      st = solve_any(true, r.p, d.p);
      if (st != XG_OK) return st;
      k_axpy_f64<<<nblk(s.n), 256, 0, stream>>>(s.n, d.p, dx);
      ++it;
    }
    stats[0] = it; stats[1] = rel; stats[2] += 1; stats[3] += it; stats[4] += fell_back ? 1 : 0;
    return XG_OK;
  }
};

struct xg_circuit {
  HostCircuit hc;
  hipStream_t stream = nullptr;
  int n_cls = 0;
  DBuf<double> cls, tvbi, beta, tvto;
  DBuf<int> mcls, lid, sta0, sto0;
  DBuf<unsigned long long> mask;
  DBuf<double> sto_cur, sto_nxt, sta_cur, sta_nxt, rec;
  DBuf<double> x, F, Q, B, FL, QL, rhs, q0, q2, dFdx, dQdx, J, linF, linQ;
  DBuf<int> rowptr, colind, vg_ptr, vg_ent, mg_ptr, mg_ent, bg_ptr, bg_ent, src_wave, wave_type, wave_off, limited;
  // long rows / slots: chunk lists (kind, begin, end), per-entry chunk ranges, partial sums
  DBuf<int> long_rows, long_slots, vck, vcb, vce, vrc, mcb, mce, msc;
  DBuf<double> vpart, mpart;
  int n_long_rows = 0, n_long_slots = 0, n_vchunks = 0, n_mchunks = 0;
  DBuf<double> wave_data, xy_sta, xy_sto;
  xg_solver_state ss{};
  // pinned host staging for the loads into host arrays (allocated on first use)
  double* hstage = nullptr;
  size_t hstage_n = 0;
  ~xg_circuit() { if (hstage) (void)hipHostFree(hstage); }
  hipError_t stage(size_t cnt) {
    if (cnt <= hstage_n) return hipSuccess;
    if (hstage) { (void)hipHostFree(hstage); hstage = nullptr; hstage_n = 0; }
    hipError_t e = hipHostMalloc((void**)&hstage, cnt * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) hstage_n = cnt;
    return e;
  }
};

static int circuit_upload(xg_circuit* c) {
  HostCircuit& h = c->hc;
  hipStream_t s = c->stream;
  c->n_cls = (int)(h.cls.size() / KSTRIDE);
  HIPCHK(c->cls.upload(h.cls, s)); HIPCHK(c->tvbi.upload(h.mos_tvbi, s)); HIPCHK(c->beta.upload(h.mos_beta, s));
  HIPCHK(c->tvto.upload(h.mos_tvto, s)); HIPCHK(c->mcls.upload(h.mos_cls, s)); HIPCHK(c->lid.upload(h.mos_lid, s));
  HIPCHK(c->sta0.upload(h.mos_sta0, s)); HIPCHK(c->sto0.upload(h.mos_sto0, s));
  std::vector<unsigned long long> m(h.mos_mask.begin(), h.mos_mask.end());
  HIPCHK(c->mask.upload(m, s));
  const size_t N = h.n_mos;
  HIPCHK(c->sto_cur.alloc(6 * N)); HIPCHK(c->sto_nxt.alloc(6 * N));
  HIPCHK(c->sta_cur.alloc(8 * N)); HIPCHK(c->sta_nxt.alloc(8 * N)); HIPCHK(c->rec.alloc((size_t)R_COUNT * N));
  const int n = h.n, nnz = h.nnz;
  HIPCHK(c->x.alloc(n + 1)); HIPCHK(c->F.alloc(n)); HIPCHK(c->Q.alloc(n)); HIPCHK(c->B.alloc(n));
  HIPCHK(c->FL.alloc(n)); HIPCHK(c->QL.alloc(n)); HIPCHK(c->rhs.alloc(n)); HIPCHK(c->q0.alloc(n)); HIPCHK(c->q2.alloc(n));
  HIPCHK(c->dFdx.alloc(nnz)); HIPCHK(c->dQdx.alloc(nnz)); HIPCHK(c->J.alloc(nnz));
  HIPCHK(c->linF.upload(h.linF, s)); HIPCHK(c->linQ.upload(h.linQ, s));
  HIPCHK(c->rowptr.upload(h.rowptr, s)); HIPCHK(c->colind.upload(h.colind, s));
  HIPCHK(c->vg_ptr.upload(h.vg_ptr, s)); HIPCHK(c->vg_ent.upload(h.vg_ent, s));
  HIPCHK(c->mg_ptr.upload(h.mg_ptr, s)); HIPCHK(c->mg_ent.upload(h.mg_ent, s));
  {
    std::vector<int> lr, ls;
    for (int r = 0; r < n; ++r)
      if (h.vg_ptr[r + 1] - h.vg_ptr[r] + h.rowptr[r + 1] - h.rowptr[r] > kGatherLong) lr.push_back(r);
    for (int q = 0; q < nnz; ++q)
      if (h.mg_ptr[q + 1] - h.mg_ptr[q] > kGatherLong) ls.push_back(q);
    c->n_long_rows = (int)lr.size();
    c->n_long_slots = (int)ls.size();
    std::vector<int> ck, cb, ce, rc, mb, me, sc;
    auto split = [&](int kind, int b0, int b1) {
      for (int a = b0; a < b1; a += kGatherChunk) { ck.push_back(kind); cb.push_back(a); ce.push_back(std::min(b1, a + kGatherChunk)); }
    };
    for (int r : lr) {
      rc.push_back((int)cb.size());
      split(0, h.vg_ptr[r], h.vg_ptr[r + 1]);
      rc.push_back((int)cb.size());
      split(1, h.rowptr[r], h.rowptr[r + 1]);
      rc.push_back((int)cb.size());
    }
    sc.push_back(0);
    for (int q : ls) {
      for (int a = h.mg_ptr[q]; a < h.mg_ptr[q + 1]; a += kGatherChunk) { mb.push_back(a); me.push_back(std::min(h.mg_ptr[q + 1], a + kGatherChunk)); }
      sc.push_back((int)mb.size());
    }
    c->n_vchunks = (int)cb.size();
    c->n_mchunks = (int)mb.size();
    if (!lr.empty()) {
      HIPCHK(c->long_rows.upload(lr, s)); HIPCHK(c->vck.upload(ck, s)); HIPCHK(c->vcb.upload(cb, s));
      HIPCHK(c->vce.upload(ce, s)); HIPCHK(c->vrc.upload(rc, s)); HIPCHK(c->vpart.alloc(4 * (size_t)cb.size()));
    }
    if (!ls.empty()) {
      HIPCHK(c->long_slots.upload(ls, s)); HIPCHK(c->mcb.upload(mb, s)); HIPCHK(c->mce.upload(me, s));
      HIPCHK(c->msc.upload(sc, s)); HIPCHK(c->mpart.alloc(2 * (size_t)mb.size()));
    }
  }
  // per-row source gather lists (source order)
  std::vector<int> bcnt(n + 1, 0);
  for (auto& e : h.srcs) {
    if (e.kind == 0) { if (e.li_p < n) bcnt[e.li_p]++; }
    else { if (e.li_p < n) bcnt[e.li_p]++; if (e.li_n < n) bcnt[e.li_n]++; }
  }
  std::vector<int> bptr(n + 1, 0);
  for (int r = 0; r < n; ++r) bptr[r + 1] = bptr[r] + bcnt[r];
  std::vector<int> bent(bptr[n]), fill(bptr.begin(), bptr.end() - 1), swave;
  for (size_t si = 0; si < h.srcs.size(); ++si) {
    const auto& e = h.srcs[si];
    swave.push_back(e.wave);
    if (e.kind == 0) { if (e.li_p < n) bent[fill[e.li_p]++] = (int)si * 2; }
    else {
      if (e.li_p < n) bent[fill[e.li_p]++] = (int)si * 2 + 1;   // B[pos] -= I
      if (e.li_n < n) bent[fill[e.li_n]++] = (int)si * 2;       // B[neg] += I
    }
  }
  HIPCHK(c->bg_ptr.upload(bptr, s)); HIPCHK(c->bg_ent.upload(bent, s)); HIPCHK(c->src_wave.upload(swave, s));
  HIPCHK(c->wave_type.upload(h.wave_type, s)); HIPCHK(c->wave_off.upload(h.wave_off, s));
  HIPCHK(c->wave_data.upload(h.wave_data, s));
  HIPCHK(c->limited.alloc(1));
  HIPCHK(c->xy_sta.alloc(std::max(1, h.nsta))); HIPCHK(c->xy_sto.alloc(std::max(1, h.nsto)));
  k_fill<<<nblk(n + 1), 256, 0, s>>>(c->x.p, n + 1, 0.0);
  HIPCHK(hipStreamSynchronize(s));
  return XG_OK;
}

static void gather_vectors(xg_circuit* c);
static void gather_matrices(xg_circuit* c);

static SrcArgs src_args(xg_circuit* c) {
  SrcArgs a;
  a.bg_ptr = c->bg_ptr.p; a.bg_ent = c->bg_ent.p; a.src_wave = c->src_wave.p; a.wave_type = c->wave_type.p;
  a.wave_off = c->wave_off.p; a.wave_data = c->wave_data.p;
  a.time = c->ss.time; a.tstep = c->ss.tstep; a.tstop = c->ss.tstop;
  return a;
}

namespace xg {

CircuitDev circuit_dev(xg_circuit* c) {
  CircuitDev d;
  d.n = c->hc.n; d.nnz = c->hc.nnz; d.n_mos = c->hc.n_mos; d.stream = (void*)c->stream;
  d.x = c->x.p; d.F = c->F.p; d.Q = c->Q.p; d.B = c->B.p; d.FL = c->FL.p; d.QL = c->QL.p;
  d.dFdx = c->dFdx.p; d.dQdx = c->dQdx.p; d.limited = c->limited.p;
  return d;
}

const HostCircuit& circuit_host(const xg_circuit* c) { return c->hc; }

int circuit_load(xg_circuit* c) {
  HostCircuit& h = c->hc;
  hipStream_t s = c->stream;
  const int N = h.n_mos;
  k_fill<<<1, 1, 0, s>>>(c->limited.p, 1, 0);
  if (N > 0) {
    EvalArgs a;
    a.n_mos = N; a.cls = c->cls.p; a.mcls = c->mcls.p; a.tvbi = c->tvbi.p; a.beta = c->beta.p; a.tvto = c->tvto.p;
    a.lid = c->lid.p; a.mask = c->mask.p; a.x = c->x.p; a.sto_cur = c->sto_cur.p; a.sto_nxt_in = c->sto_nxt.p;
    a.sta_cur = c->sta_cur.p; a.sto_out = c->sto_nxt.p; a.sta_out = c->sta_nxt.p; a.rec = c->rec.p;
    a.limited = c->limited.p;
    a.fl.dcop = c->ss.dcop; a.fl.tran = c->ss.tran; a.fl.init_jct = c->ss.init_jct; a.fl.newton_iter = c->ss.newton_iter;
    a.fl.tranop = c->ss.tranop;
    k_mos1_eval<<<nblk(N), 256, 0, s>>>(a);
  }
  gather_vectors(c);
  gather_matrices(c);
  HIPCHK(hipGetLastError());
  return XG_OK;
}

int circuit_combine(xg_circuit* c, const double* q0, const double* q2, double* rhs, double* J) {
  HostCircuit& h = c->hc;
  hipStream_t s = c->stream;
  const int n = h.n;
  k_combine_rhs<<<nblk(n), 256, 0, s>>>(n, c->ss.dcop, c->ss.order, c->ss.h, c->F.p, c->Q.p, c->B.p, c->FL.p,
                                        c->QL.p, c->ss.dcop ? c->F.p : q0, c->ss.dcop ? c->F.p : q2, rhs);
  double a, b;
  if (c->ss.dcop) { a = 1.0e-20; b = 1.0; } else { a = 1.0 / c->ss.h; b = c->ss.order == 2 ? 0.5 : 1.0; }
  k_combine_jac<<<nblk(h.nnz), 256, 0, s>>>(h.nnz, a, b, c->dFdx.p, c->dQdx.p, J);
  HIPCHK(hipGetLastError());
  return XG_OK;
}

int circuit_history(xg_circuit* c, int op) {
  const size_t N = c->hc.n_mos;
  if (N == 0) return XG_OK;
  hipStream_t s = c->stream;
  if (op == 0) {
    k_fill<<<nblk(8 * N), 256, 0, s>>>(c->sta_cur.p, 8 * N, 0.0);
    k_fill<<<nblk(8 * N), 256, 0, s>>>(c->sta_nxt.p, 8 * N, 0.0);
    k_fill<<<nblk(6 * N), 256, 0, s>>>(c->sto_cur.p, 6 * N, 0.0);
    k_fill<<<nblk(6 * N), 256, 0, s>>>(c->sto_nxt.p, 6 * N, 0.0);
  } else if (op == 1) {
    HIPCHK(hipMemcpyAsync(c->sta_cur.p, c->sta_nxt.p, sizeof(double) * 8 * N, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(c->sto_cur.p, c->sto_nxt.p, sizeof(double) * 6 * N, hipMemcpyDeviceToDevice, s));
  } else if (op == 2) {
    HIPCHK(hipMemcpyAsync(c->sta_nxt.p, c->sta_cur.p, sizeof(double) * 8 * N, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(c->sto_nxt.p, c->sto_cur.p, sizeof(double) * 6 * N, hipMemcpyDeviceToDevice, s));
  } else {
    return XG_E_ARG;
  }
  HIPCHK(hipGetLastError());
  return XG_OK;
}

}  // namespace xg

extern "C" {

int xg_circuit_create(const xg_circuit_desc* desc, xg_circuit** out) {
  if (!desc || !out) return XG_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return XG_E_NOGPU;
  auto c = std::make_unique<xg_circuit>();
  int st = build_host_circuit(desc, c->hc);
  if (st != XG_OK) return st;
  if (hipStreamCreate(&c->stream) != hipSuccess) return XG_E_HIP;
  st = circuit_upload(c.get());
  if (st != XG_OK) return st;
  c->ss.tran = 1; c->ss.order = 1; c->ss.h = 1e-12;
  *out = c.release();
  return XG_OK;
}

int xg_circuit_destroy(xg_circuit* c) {
  if (!c) return XG_E_ARG;
  hipStream_t s = c->stream;
  delete c;
  if (s) (void)hipStreamDestroy(s);
  return XG_OK;
}

int xg_circuit_sizes(const xg_circuit* c, int* sizes) {
  if (!c || !sizes) return XG_E_ARG;
  sizes[0] = c->hc.n; sizes[1] = c->hc.nnz; sizes[2] = c->hc.nsta; sizes[3] = c->hc.nsto;
  sizes[4] = c->hc.n_mos; sizes[5] = c->hc.n_nodes;
  return XG_OK;
}

int xg_circuit_pattern(const xg_circuit* c, int* rowptr, int* colind, int* node_gid, int* mos_lid, int* mos_fslot,
                       int* mos_qslot, int* mos_stasto) {
  if (!c) return XG_E_ARG;
  const HostCircuit& h = c->hc;
  if (rowptr) std::copy(h.rowptr.begin(), h.rowptr.end(), rowptr);
  if (colind) std::copy(h.colind.begin(), h.colind.end(), colind);
  if (node_gid) std::copy(h.node_gid.begin(), h.node_gid.end(), node_gid);
  if (mos_lid) std::copy(h.mos_lid.begin(), h.mos_lid.end(), mos_lid);
  if (mos_fslot) std::copy(h.mos_fslot.begin(), h.mos_fslot.end(), mos_fslot);
  if (mos_qslot) std::copy(h.mos_qslot.begin(), h.mos_qslot.end(), mos_qslot);
  if (mos_stasto)
    for (int i = 0; i < h.n_mos; ++i) { mos_stasto[2 * i] = h.mos_sta0[i]; mos_stasto[2 * i + 1] = h.mos_sto0[i]; }
  return XG_OK;
}

int xg_set_solver_state(xg_circuit* c, const xg_solver_state* ss) {
  if (!c || !ss) return XG_E_ARG;
  c->ss = *ss;
  return XG_OK;
}

// a2 on host (dev = false) or device (dev = true) arrays in Xyce's layout.
// Only the MOSFETs' state and store entries of sta_next / sto_next are
// written; every other entry keeps the caller's value, as each reference
// master writes only its own devices' entries.
static int update_state_impl(xg_circuit* c, const double* x, const double* sta_cur, const double* sto_cur,
                             double* sta_next, double* sto_next, int* limited, bool dev) {
  if (!c || !x) return XG_E_ARG;
  HostCircuit& h = c->hc;
  hipStream_t s = c->stream;
  const int N = h.n_mos;
  if (N > 0 && (!sta_cur || !sto_cur || !sta_next || !sto_next)) return XG_E_ARG;
  const hipMemcpyKind in = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  HIPCHK(hipMemcpyAsync(c->x.p, x, sizeof(double) * h.n, in, s));
  k_fill<<<1, 1, 0, s>>>(c->x.p + h.n, 1, 0.0);
  if (N > 0) {
    const double* sc = sta_cur;
    const double* oc = sto_cur;
    const double* on = sto_next;
    if (!dev) {
      HIPCHK(hipMemcpyAsync(c->xy_sta.p, sta_cur, sizeof(double) * h.nsta, hipMemcpyHostToDevice, s));
      sc = c->xy_sta.p;
    }
    k_sta_to_soa<<<nblk(N), 256, 0, s>>>(N, 8, c->sta0.p, sc, c->sta_cur.p);
    if (!dev) {
      HIPCHK(hipMemcpyAsync(c->xy_sto.p, sto_cur, sizeof(double) * h.nsto, hipMemcpyHostToDevice, s));
      oc = c->xy_sto.p;
    }
    k_sta_to_soa<<<nblk(N), 256, 0, s>>>(N, 6, c->sto0.p, oc, c->sto_cur.p);
    if (!dev) {
      HIPCHK(hipMemcpyAsync(c->xy_sto.p, sto_next, sizeof(double) * h.nsto, hipMemcpyHostToDevice, s));
      on = c->xy_sto.p;
    }
    k_sta_to_soa<<<nblk(N), 256, 0, s>>>(N, 6, c->sto0.p, on, c->sto_nxt.p);
  }
  k_fill<<<1, 1, 0, s>>>(c->limited.p, 1, 0);
  if (N > 0) {
    EvalArgs a;
    a.n_mos = N; a.cls = c->cls.p; a.mcls = c->mcls.p; a.tvbi = c->tvbi.p; a.beta = c->beta.p; a.tvto = c->tvto.p;
    a.lid = c->lid.p; a.mask = c->mask.p; a.x = c->x.p; a.sto_cur = c->sto_cur.p; a.sto_nxt_in = c->sto_nxt.p;
    a.sta_cur = c->sta_cur.p; a.sto_out = c->sto_nxt.p; a.sta_out = c->sta_nxt.p; a.rec = c->rec.p;
    a.limited = c->limited.p;
    a.fl.dcop = c->ss.dcop; a.fl.tran = c->ss.tran; a.fl.init_jct = c->ss.init_jct; a.fl.newton_iter = c->ss.newton_iter;
    a.fl.tranop = c->ss.tranop;
    k_mos1_eval<<<nblk(N), 256, 0, s>>>(a);
    HIPCHK(hipGetLastError());
    // in-place read of sto_nxt_in vs write of sto_out: each thread reads its
    // own inputs before writing its own outputs, so aliasing is safe.
    if (dev) {
      k_soa_to_sta<<<nblk(N), 256, 0, s>>>(N, 8, c->sta0.p, c->sta_nxt.p, sta_next);
      k_soa_to_sta<<<nblk(N), 256, 0, s>>>(N, 6, c->sto0.p, c->sto_nxt.p, sto_next);
    } else {
      // the caller's other entries come back unchanged
      HIPCHK(hipMemcpyAsync(c->xy_sta.p, sta_next, sizeof(double) * h.nsta, hipMemcpyHostToDevice, s));
      k_soa_to_sta<<<nblk(N), 256, 0, s>>>(N, 8, c->sta0.p, c->sta_nxt.p, c->xy_sta.p);
      HIPCHK(hipMemcpyAsync(sta_next, c->xy_sta.p, sizeof(double) * h.nsta, hipMemcpyDeviceToHost, s));
      k_soa_to_sta<<<nblk(N), 256, 0, s>>>(N, 6, c->sto0.p, c->sto_nxt.p, c->xy_sto.p);
      HIPCHK(hipMemcpyAsync(sto_next, c->xy_sto.p, sizeof(double) * h.nsto, hipMemcpyDeviceToHost, s));
    }
  }
  int lim = 0;
  HIPCHK(hipMemcpyAsync(&lim, c->limited.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (limited) *limited = lim;
  return XG_OK;
}

int xg_update_state(xg_circuit* c, const double* x, const double* sta_cur, const double* sto_cur, double* sta_next,
                    double* sto_next, int* limited) {
  return update_state_impl(c, x, sta_cur, sto_cur, sta_next, sto_next, limited, false);
}

int xg_update_state_dev(xg_circuit* c, const double* x, const double* sta_cur, const double* sto_cur,
                        double* sta_next, double* sto_next, int* limited) {
  return update_state_impl(c, x, sta_cur, sto_cur, sta_next, sto_next, limited, true);
}

// Adds k device arrays of m entries each (this handle's contributions) into
// the caller's arrays: device arrays by a kernel on the handle's stream, host
// arrays through the pinned staging buffer.  Each output entry becomes
// caller's value + contribution, one rounding, as a master's "+=" into a
// vector the loader has zeroed or other devices have filled.
static int accumulate_out(xg_circuit* c, int k, const double* const* src, double* const* dst, size_t m, bool dev) {
  hipStream_t s = c->stream;
  if (dev) {
    for (int i = 0; i < k; ++i)
      if (dst[i] && m) k_axpy_f64<<<nblk(m), 256, 0, s>>>((int)m, src[i], dst[i]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return XG_OK;
  }
  int used = 0;
  for (int i = 0; i < k; ++i) used += dst[i] != nullptr;
  if (used == 0 || m == 0) { HIPCHK(hipStreamSynchronize(s)); return XG_OK; }
  HIPCHK(c->stage((size_t)used * m));
  for (int i = 0, j = 0; i < k; ++i)
    if (dst[i]) HIPCHK(hipMemcpyAsync(c->hstage + (size_t)(j++) * m, src[i], sizeof(double) * m, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int i = 0, j = 0; i < k; ++i) {
    if (!dst[i]) continue;
    const double* st = c->hstage + (size_t)(j++) * m;
    double* d = dst[i];
    for (size_t e = 0; e < m; ++e) d[e] += st[e];
  }
  return XG_OK;
}

static void gather_vectors(xg_circuit* c) {
  HostCircuit& h = c->hc;
  hipStream_t s = c->stream;
  const int n = h.n;
  k_gather_vectors<<<nblk(n), 256, 0, s>>>(n, h.n_mos, c->vg_ptr.p, c->vg_ent.p, c->rec.p, c->rowptr.p, c->colind.p,
                                           c->linF.p, c->linQ.p, c->x.p, src_args(c), c->F.p, c->Q.p, c->B.p, c->FL.p, c->QL.p);
  if (c->n_long_rows) {
    k_gather_vectors_part<<<c->n_vchunks, 256, 0, s>>>(h.n_mos, c->vck.p, c->vcb.p, c->vce.p, c->vg_ent.p, c->rec.p,
                                                       c->colind.p, c->linF.p, c->linQ.p, c->x.p, c->vpart.p);
    k_gather_vectors_fin<<<c->n_long_rows, 64, 0, s>>>(c->long_rows.p, c->vrc.p, c->vpart.p, src_args(c), c->F.p,
                                                       c->Q.p, c->B.p, c->FL.p, c->QL.p);
  }
}

static void gather_matrices(xg_circuit* c) {
  HostCircuit& h = c->hc;
  hipStream_t s = c->stream;
  k_gather_matrices<<<nblk(h.nnz), 256, 0, s>>>(h.nnz, h.n_mos, c->mg_ptr.p, c->mg_ent.p, c->rec.p, c->linF.p,
                                               c->linQ.p, c->dFdx.p, c->dQdx.p, 0.0, 0.0, nullptr, nullptr, nullptr);
  if (c->n_long_slots) {
    k_gather_matrices_part<<<c->n_mchunks, 256, 0, s>>>(h.n_mos, c->mcb.p, c->mce.p, c->mg_ent.p, c->rec.p, c->mpart.p);
    k_gather_matrices_fin<<<c->n_long_slots, 64, 0, s>>>(c->long_slots.p, c->msc.p, c->mpart.p, c->linF.p, c->linQ.p,
                                                         c->dFdx.p, c->dQdx.p, 0.0, 0.0, nullptr, nullptr, nullptr);
  }
}

static int load_vectors_impl(xg_circuit* c, double* F, double* Q, double* B, double* dFdxdVp, double* dQdxdVp,
                             bool dev) {
  if (!c) return XG_E_ARG;
  gather_vectors(c);
  HIPCHK(hipGetLastError());
  const double* src[5] = {c->F.p, c->Q.p, c->B.p, c->FL.p, c->QL.p};
  double* dst[5] = {F, Q, B, dFdxdVp, dQdxdVp};
  return accumulate_out(c, 5, src, dst, (size_t)c->hc.n, dev);
}

static int load_matrices_impl(xg_circuit* c, double* dFdx_vals, double* dQdx_vals, bool dev) {
  if (!c) return XG_E_ARG;
  gather_matrices(c);
  HIPCHK(hipGetLastError());
  const double* src[2] = {c->dFdx.p, c->dQdx.p};
  double* dst[2] = {dFdx_vals, dQdx_vals};
  return accumulate_out(c, 2, src, dst, (size_t)c->hc.nnz, dev);
}

int xg_load_vectors(xg_circuit* c, double* F, double* Q, double* B, double* dFdxdVp, double* dQdxdVp) {
  return load_vectors_impl(c, F, Q, B, dFdxdVp, dQdxdVp, false);
}

int xg_load_vectors_dev(xg_circuit* c, double* F, double* Q, double* B, double* dFdxdVp, double* dQdxdVp) {
  return load_vectors_impl(c, F, Q, B, dFdxdVp, dQdxdVp, true);
}

int xg_load_matrices(xg_circuit* c, double* dFdx_vals, double* dQdx_vals) {
  return load_matrices_impl(c, dFdx_vals, dQdx_vals, false);
}

int xg_load_matrices_dev(xg_circuit* c, double* dFdx_vals, double* dQdx_vals) {
  return load_matrices_impl(c, dFdx_vals, dQdx_vals, true);
}

static int combine_impl(xg_circuit* c, const double* q0, const double* q2, double* rhs, double* jac_vals, bool dev) {
  if (!c) return XG_E_ARG;
  HostCircuit& h = c->hc;
  hipStream_t s = c->stream;
  const int n = h.n;
  const hipMemcpyKind in = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  const hipMemcpyKind out = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (!c->ss.dcop) {
    if (!q0) return XG_E_ARG;
    HIPCHK(hipMemcpyAsync(c->q0.p, q0, sizeof(double) * n, in, s));
    if (q2) HIPCHK(hipMemcpyAsync(c->q2.p, q2, sizeof(double) * n, in, s));
    else k_fill<<<nblk(n), 256, 0, s>>>(c->q2.p, n, 0.0);
  }
  k_combine_rhs<<<nblk(n), 256, 0, s>>>(n, c->ss.dcop, c->ss.order, c->ss.h, c->F.p, c->Q.p, c->B.p, c->FL.p, c->QL.p,
                                        c->q0.p, c->q2.p, c->rhs.p);
  double a, b;
  if (c->ss.dcop) { a = 1.0e-20; b = 1.0; } else { a = 1.0 / c->ss.h; b = c->ss.order == 2 ? 0.5 : 1.0; }
  k_combine_jac<<<nblk(h.nnz), 256, 0, s>>>(h.nnz, a, b, c->dFdx.p, c->dQdx.p, c->J.p);
  HIPCHK(hipGetLastError());
  if (rhs) HIPCHK(hipMemcpyAsync(rhs, c->rhs.p, sizeof(double) * n, out, s));
  if (jac_vals) HIPCHK(hipMemcpyAsync(jac_vals, c->J.p, sizeof(double) * h.nnz, out, s));
  HIPCHK(hipStreamSynchronize(s));
  return XG_OK;
}

int xg_combine(xg_circuit* c, const double* q0, const double* q2, double* rhs, double* jac_vals) {
  return combine_impl(c, q0, q2, rhs, jac_vals, false);
}

int xg_combine_dev(xg_circuit* c, const double* q0, const double* q2, double* rhs, double* jac_vals) {
  return combine_impl(c, q0, q2, rhs, jac_vals, true);
}

// ---------------- LU ----------------
int xg_lu_symbolic_ex(int n, const int* rowptr, const int* colind, const double* values, int method, xg_lu** out) {
  if (n <= 0 || !rowptr || !colind || !out || method < 0 || method > 2) return XG_E_ARG;
  if (!xg::valid_csr(n, rowptr, colind)) return XG_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return XG_E_NOGPU;
  auto lu = std::make_unique<xg_lu>();
  lu->a_rowptr.assign(rowptr, rowptr + n + 1);
  lu->a_colind.assign(colind, colind + rowptr[n]);
  lu->s.n = n;
  try {
    int st = lu->analyse(values, method);
    if (st != XG_OK) return st;
  } catch (...) {
    return XG_E_SETUP;
  }
  if (hipStreamCreate(&lu->stream) != hipSuccess) return XG_E_HIP;
  lu->own_stream = true;
  int st = lu->upload();
  if (st != XG_OK) return st;
  // the no-allocation solve's work area on this device (looked up once per device)
  if (scratch_address(&lu->scratch_work) != XG_OK) lu->scratch_work = nullptr;
  *out = lu.release();
  return XG_OK;
}

int xg_lu_symbolic(int n, const int* rowptr, const int* colind, const double* values, xg_lu** out) {
  return xg_lu_symbolic_ex(n, rowptr, colind, values, 0, out);
}

int xg_lu_destroy(xg_lu* lu) {
  if (!lu) return XG_E_ARG;
  hipStream_t s = lu->own_stream ? lu->stream : nullptr;
  delete lu;
  if (s) (void)hipStreamDestroy(s);
  return XG_OK;
}

// FP64 factors of the values in avals after the FP32 ones (mode 1) or alone:
// threshold-only failures are kept (solves refine), a zero or non-finite
// pivot is reported through *hard
static int lu_factor64(xg_lu* lu, bool fp32_failed, int* bad, int* hard) {
  int st = lu->factor_any(false, bad, hard);
  if (st != XG_OK) return st;
  if (*bad && *hard) return XG_OK;
  if (fp32_failed) lu->stats[5] += 1;
  lu->have64 = true;
  lu->soft64 = *bad != 0;
  return XG_OK;
}

static int lu_numeric_impl(xg_lu* lu, const double* vals, bool dev) {
  if (!lu || !vals) return XG_E_ARG;
  hipStream_t s = lu->stream;
  const hipMemcpyKind in = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  HIPCHK(hipMemcpyAsync(lu->avals.p, vals, sizeof(double) * lu->s.nnz_a, in, s));
  lu->have64 = false;
  lu->soft64 = false;
  int bad = 0, hard = 0, st;
  if (lu->mode == 1) {
    st = lu->factor_any(true, &bad, &hard);
    if (st != XG_OK) return st;
    if (!bad) return XG_OK;
  }
  st = lu_factor64(lu, lu->mode == 1, &bad, &hard);
  if (st != XG_OK) return st;
  if (!bad || (!hard && lu->repivot_futile)) return XG_OK;
  // a pivot that fails KLU's threshold or is zero (KLU would pivot elsewhere):
  // analyse again steering the transversal by the values, then refactor
  std::vector<double> hv;
  if (dev) {
    hv.resize(lu->s.nnz_a);
    HIPCHK(hipMemcpyAsync(hv.data(), lu->avals.p, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  const double* hvals = dev ? hv.data() : vals;
  try {
    if (lu->analyse(hvals, lu->method) != XG_OK) return bad;
  } catch (...) {
    return bad;
  }
  st = lu->upload();
  if (st != XG_OK) return st;
  HIPCHK(hipMemcpyAsync(lu->avals.p, hvals, sizeof(double) * lu->s.nnz_a, hipMemcpyHostToDevice, s));
  if (lu->mode == 1) {
    st = lu->factor_any(true, &bad, &hard);
    if (st != XG_OK) return st;
    if (!bad) return XG_OK;
  }
  st = lu_factor64(lu, lu->mode == 1, &bad, &hard);
  if (st != XG_OK) return st;
  if (bad && !hard) lu->repivot_futile = true;  // kept, refined from now on without a re-pivot
  return (bad && hard) ? bad : XG_OK;
}

int xg_lu_numeric(xg_lu* lu, const double* vals) { return lu_numeric_impl(lu, vals, false); }

int xg_lu_numeric_dev(xg_lu* lu, const double* vals) { return lu_numeric_impl(lu, vals, true); }

static int lu_solve_impl(xg_lu* lu, const double* b, double* x, int transpose, bool dev) {
  if (!lu || !b || !x) return XG_E_ARG;
  // transpose solves: the level-scheduled method with plain FP64 factors
  if (transpose && (lu->method != 1 || lu->mode == 1 || !lu->have64 || lu->soft64)) return XG_E_ARG;
  hipStream_t s = lu->stream;
  const int n = lu->s.n;
  HIPCHK(hipMemcpyAsync(lu->b.p, b, sizeof(double) * n, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  k_fill<<<1, 1, 0, s>>>(lu->flag.p + 1, 1, 0);
  int st;
  if (transpose) st = lu->solve_transpose(lu->b.p, lu->x.p, lu->flag.p + 1);
  else if (lu->mode == 1 && !lu->have64) st = lu->solve_refined(lu->b.p, lu->x.p);
  else if (lu->soft64) st = lu->solve_refined64(lu->b.p, lu->x.p);
  else st = lu->solve_any(false, lu->b.p, lu->x.p);
  if (st != XG_OK) return st;
  int nanf = 0;
  HIPCHK(hipMemcpyAsync(&nanf, lu->flag.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
  if (dev) {
    HIPCHK(hipStreamSynchronize(s));
    // the solve kernels flag every non-finite result
    if (nanf) { k_fill<<<nblk(n), 256, 0, s>>>(x, (size_t)n, 0.0); HIPCHK(hipStreamSynchronize(s)); return XG_E_NAN; }
    HIPCHK(hipMemcpyAsync(x, lu->x.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return XG_OK;
  }
  HIPCHK(hipMemcpyAsync(x, lu->x.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  bool bad = nanf != 0;
  for (int i = 0; i < n && !bad; ++i) bad = !std::isfinite(x[i]);
  if (bad) { std::fill(x, x + n, 0.0); return XG_E_NAN; }
  return XG_OK;
}

int xg_lu_solve(xg_lu* lu, const double* b, double* x, int transpose) {
  return lu_solve_impl(lu, b, x, transpose, false);
}

int xg_lu_solve_dev(xg_lu* lu, const double* b, double* x, int transpose) {
  return lu_solve_impl(lu, b, x, transpose, true);
}

int xg_lu_solve_noalloc(xg_lu* lu, const double* b, double* x) {
  if (!lu || !b || !x) return XG_E_ARG;
  if (!lu->scratch_work) return XG_E_HIP;
  // FP32 factors, and FP64 factors with multipliers beyond the threshold,
  // are refined against A (and FP32 refinement may refactor in FP64): not a
  // plain solve with the current factors
  if ((lu->mode == 1 && !lu->have64) || lu->soft64) return XG_E_ARG;
  if (lu->work_doubles() > kScratchDoubles) return XG_E_ARG;
  hipStream_t s = lu->stream;
  int st = lu->solve_with_work(b, x, lu->scratch_work);
  if (st != XG_OK) return st;
  int nanf = 0;
  HIPCHK(hipMemcpyAsync(&nanf, lu->scratch_work, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (nanf) {
    k_fill<<<nblk(lu->s.n), 256, 0, s>>>(x, (size_t)lu->s.n, 0.0);
    HIPCHK(hipStreamSynchronize(s));
    return XG_E_NAN;
  }
  return XG_OK;
}

int xg_lu_set_precision(xg_lu* lu, int precision, double ir_tol, int ir_max) {
  if (!lu || precision < 0 || precision > 1 || !(ir_tol > 0.0) || ir_max < 0) return XG_E_ARG;
  lu->mode = precision;
  lu->ir_tol = ir_tol;
  lu->ir_max = ir_max;
  if (precision == 1 && lu->method == 1) {
    HIPCHK(lu->lu32.alloc(lu->s.nlu));
    HIPCHK(lu->y32.alloc(lu->s.n));
  }
  lu->have64 = false;
  return XG_OK;
}

int xg_lu_precision_stats(const xg_lu* lu, double* stats) {
  if (!lu || !stats) return XG_E_ARG;
  for (int i = 0; i < 6; ++i) stats[i] = lu->stats[i];
  stats[6] = lu->mode;
  return XG_OK;
}

int xg_lu_info(const xg_lu* lu, double* info) {
  if (!lu || !info) return XG_E_ARG;
  if (lu->method == 2) {
    info[0] = (double)lu->mfp->front_elements();
    info[1] = lu->mfp->levels();
    info[2] = 2.0 * lu->mfp->levels();
    info[3] = lu->mfp->flops();
  } else {
    info[0] = lu->s.nlu;
    info[1] = (double)lu->s.f_lvl_ptr.size() - 1;
    info[2] = (double)(lu->s.fs_lvl_ptr.size() - 1 + lu->s.bs_lvl_ptr.size() - 1);
    info[3] = lu->s.flops;
  }
  info[4] = lu->method;
  info[5] = lu->method == 2 ? lu->mfp->supernodes() : 0;
  return XG_OK;
}

}  // extern "C"
