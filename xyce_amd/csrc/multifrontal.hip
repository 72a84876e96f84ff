// Multifrontal numeric LU for large systems (power-grid meshes, gate
// arrays): the factorisation that AmesosSolver::doSolve
// (src/LinearAlgebraServicesPKG/N_LAS_AmesosSolver.C:216-478) gets from KLU,
// restructured for the GPU.  The symbolic analysis (maximum transversal,
// approximate minimum degree, elimination tree) is the one the level-
// scheduled path uses; here the columns of L are grouped into supernodes,
// each with a dense frontal matrix, and the elimination tree of supernodes
// is processed level by level (all fronts of a level in one launch):
//
//   assemble  F_s = A entries of s's columns + extend-add of the children's
//             Schur complements (the j-th children of a level in launch j,
//             rows split over workgroups: a fixed order, no atomics)
//   factor    dense LU of the pivot block, L21, U12 and the Schur complement
//             F22 -= L21 U12; small fronts one workgroup each, large fronts
//             blocked by 32-column panels (diagonal block in LDS, row-parallel
//             triangular solves for L21 and U12, tiled update over the GPU)
//
// Static pivots, as in the level-scheduled path, with KLU's threshold test
// (|l_ij| <= 1/0.001, finite nonzero pivots) on every column.
//
// Layout in HBM: fronts row-major, one after another (fo[s]); factors stay in
// the fronts (F11 = L11\U11, F21 = L21, F12 = U12); per-front solve vectors
// (vo[s]); index lists in B (permuted) numbering.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "xg_internal.h"

namespace xg {

// ---------------------------------------------------------------------------
// plan (host)
// ---------------------------------------------------------------------------
int mf_plan(const LuSymbolic& s, const int* rowptr, const int* colind, MfPlan& p) {
  const int n = s.n;
  if ((int)s.etree.size() != n || (int)s.lcol.size() != n) return XG_E_SETUP;
  p = MfPlan();
  p.n = n;
  const std::vector<int>& par = s.etree;
  std::vector<int> nchild(n, 0);
  for (int k = 0; k < n; ++k) if (par[k] >= 0) nchild[par[k]]++;
  // fundamental supernodes: k joins k-1's supernode when k is k-1's parent,
  // its only child, and the structures nest exactly
  std::vector<int> snode(n);
  p.sfirst.push_back(0);
  for (int k = 0; k < n; ++k) {
    const bool join = k > 0 && par[k - 1] == k && nchild[k] == 1 &&
                      s.lcol[k - 1].size() == s.lcol[k].size() + 1;
    if (!join && k > 0) p.sfirst.push_back(k);
    snode[k] = (int)p.sfirst.size() - 1;
  }
  p.sfirst.push_back(n);
  const int ns = p.ns = (int)p.sfirst.size() - 1;
  // front index lists: pivots, then the last column's rows below
  p.iptr.assign(ns + 1, 0);
  p.m.resize(ns);
  for (int t = 0; t < ns; ++t) {
    const int last = p.sfirst[t + 1] - 1;
    p.m[t] = (p.sfirst[t + 1] - p.sfirst[t]) + (int)s.lcol[last].size();
    p.iptr[t + 1] = p.iptr[t] + p.m[t];
    p.max_m = std::max<long long>(p.max_m, p.m[t]);
  }
  p.idx.resize(p.iptr[ns]);
  for (int t = 0; t < ns; ++t) {
    int q = p.iptr[t];
    for (int c = p.sfirst[t]; c < p.sfirst[t + 1]; ++c) p.idx[q++] = c;
    for (int r : s.lcol[p.sfirst[t + 1] - 1]) p.idx[q++] = r;
  }
  // parents, children, levels
  p.sparent.assign(ns, -1);
  for (int t = 0; t < ns; ++t) {
    const int pk = par[p.sfirst[t + 1] - 1];
    p.sparent[t] = pk >= 0 ? snode[pk] : -1;
    if (p.sparent[t] >= 0 && p.sparent[t] <= t) return XG_E_SETUP;
  }
  p.cptr.assign(ns + 1, 0);
  for (int t = 0; t < ns; ++t) if (p.sparent[t] >= 0) p.cptr[p.sparent[t] + 1]++;
  for (int t = 0; t < ns; ++t) p.cptr[t + 1] += p.cptr[t];
  p.clist.resize(p.cptr[ns]);
  {
    std::vector<int> f(p.cptr.begin(), p.cptr.end() - 1);
    for (int t = 0; t < ns; ++t) if (p.sparent[t] >= 0) p.clist[f[p.sparent[t]]++] = t;
  }
  std::vector<int> lvl(ns, 0);
  int maxl = 0;
  for (int t = 0; t < ns; ++t) {       // children come before parents (postorder numbering)
    for (int c = p.cptr[t]; c < p.cptr[t + 1]; ++c) lvl[t] = std::max(lvl[t], lvl[p.clist[c]] + 1);
    maxl = std::max(maxl, lvl[t]);
  }
  p.lvl_ptr.assign(maxl + 2, 0);
  for (int t = 0; t < ns; ++t) p.lvl_ptr[lvl[t] + 1]++;
  for (int l = 0; l <= maxl; ++l) p.lvl_ptr[l + 1] += p.lvl_ptr[l];
  p.lvl_s.resize(ns);
  {
    std::vector<int> f(p.lvl_ptr.begin(), p.lvl_ptr.end() - 1);
    for (int t = 0; t < ns; ++t) p.lvl_s[f[lvl[t]]++] = t;
  }
  // extend-add maps: each row below a front's pivots, located in its parent
  p.eptr.assign(ns + 1, 0);
  for (int t = 0; t < ns; ++t) p.eptr[t + 1] = p.eptr[t] + (p.m[t] - (p.sfirst[t + 1] - p.sfirst[t]));
  p.emap.resize(p.eptr[ns]);
  for (int t = 0; t < ns; ++t) {
    const int np = p.sfirst[t + 1] - p.sfirst[t];
    const int nu = p.m[t] - np;
    if (nu == 0) continue;
    const int pa = p.sparent[t];
    if (pa < 0) return XG_E_SETUP;
    const int* pb = p.idx.data() + p.iptr[pa];
    const int* pe = p.idx.data() + p.iptr[pa + 1];
    const int npa = p.sfirst[pa + 1] - p.sfirst[pa];
    for (int a = 0; a < nu; ++a) {
      const int r = p.idx[p.iptr[t] + np + a];
      // the parent's list is pivots (ascending) then rows (ascending)
      int loc = -1;
      if (r >= p.sfirst[pa] && r < p.sfirst[pa + 1]) loc = r - p.sfirst[pa];
      else {
        const int* f = std::lower_bound(pb + npa, pe, r);
        if (f != pe && *f == r) loc = (int)(f - pb);
      }
      if (loc < 0) return XG_E_SETUP;   // structures must nest
      p.emap[p.eptr[t] + a] = loc;
    }
  }
  // front and vector offsets, flops
  p.fo.assign(ns + 1, 0);
  p.vo.assign(ns + 1, 0);
  for (int t = 0; t < ns; ++t) {
    p.fo[t + 1] = p.fo[t] + (long long)p.m[t] * p.m[t];
    p.vo[t + 1] = p.vo[t] + p.m[t];
    const double np = p.sfirst[t + 1] - p.sfirst[t], mm = p.m[t];
    for (double k = 0; k < np; ++k) p.flops += 2.0 * (mm - k - 1) * (mm - k - 1) + (mm - k - 1);
  }
  // A entries -> front elements: entry (i,j) of B goes to the front whose
  // pivots hold min(i,j)
  p.amap.resize(s.nnz_a);
  std::vector<int> pos_in(0);
  for (int r = 0; r < n; ++r) {
    const int i = s.rinv[r];
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      const int j = s.pinv[colind[k]];
      const int t = snode[std::min(i, j)];
      const int* b = p.idx.data() + p.iptr[t];
      const int np = p.sfirst[t + 1] - p.sfirst[t];
      auto local = [&](int g) -> int {
        if (g >= p.sfirst[t] && g < p.sfirst[t + 1]) return g - p.sfirst[t];
        const int* e = p.idx.data() + p.iptr[t + 1];
        const int* f = std::lower_bound(b + np, e, g);
        return (f != e && *f == g) ? (int)(f - b) : -1;
      };
      const int li = local(i), lj = local(j);
      if (li < 0 || lj < 0) return XG_E_SETUP;
      p.amap[k] = p.fo[t] + (long long)li * p.m[t] + lj;
    }
  }
  return XG_OK;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
struct MfDev {
  const int* sfirst; const int* m; const long long* fo; const int* iptr; const int* idx;
  const int* cptr; const int* clist; const int* eptr; const int* emap; const int* vo;
};

template <class T>
__global__ void k_mf_fill(T* F, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) F[i] = (T)0;
}

template <class T>
__global__ void k_mf_scatter(int nnz, const double* vals, const long long* amap, T* F) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) F[amap[k]] += (T)vals[k];
}

// extend-add of one child's Schur complement into its parent: block.x picks
// the child from `list`, block.y a chunk of 16 of its rows.  Children of one
// parent go in separate launches (their j-th child in launch j), so every
// parent element is added to in a fixed order and no two blocks of a launch
// write the same element.
template <class T>
__global__ void __launch_bounds__(256) k_mf_extend_add(MfDev d, const int* list, const int* sparent, T* F) {
  const int ch = list[blockIdx.x];
  const int mc = d.m[ch], npc = d.sfirst[ch + 1] - d.sfirst[ch], nu = mc - npc;
  const int a0 = blockIdx.y * 16;
  if (a0 >= nu) return;
  const int pa = sparent[ch];
  const int ms = d.m[pa];
  T* Fs = F + d.fo[pa];
  const T* Fc = F + d.fo[ch];
  const int* em = d.emap + d.eptr[ch];
  const int a1 = min(nu, a0 + 16);
  for (int a = a0; a < a1; ++a) {
    const long long prow = (long long)em[a] * ms;
    const T* crow = Fc + (long long)(npc + a) * mc + npc;
    for (int b = threadIdx.x; b < nu; b += blockDim.x) Fs[prow + em[b]] += crow[b];
  }
}

template <class T>
__device__ __forceinline__ bool mf_bad_pivot(T d) {
  const double v = (double)d;
  return !(fabs(v) > 0.0) || !isfinite(v);
}

// small fronts: one workgroup per front, unblocked right-looking over the
// pivots, the whole front updated (its trailing block is the Schur complement)
template <class T>
__global__ void __launch_bounds__(256) k_mf_factor_small(MfDev d, const int* list, T* F, int* bad) {
  const int s = list[blockIdx.x];
  const int ms = d.m[s], np = d.sfirst[s + 1] - d.sfirst[s];
  T* Fs = F + d.fo[s];
  __shared__ int flag;
  if (threadIdx.x == 0) flag = 0;
  __syncthreads();
  for (int k = 0; k < np; ++k) {
    const T piv = Fs[(long long)k * ms + k];
    if (mf_bad_pivot(piv)) { if (threadIdx.x == 0) flag = 1; break; }
    for (int i = k + 1 + threadIdx.x; i < ms; i += blockDim.x) {
      const T l = Fs[(long long)i * ms + k] / piv;
      Fs[(long long)i * ms + k] = l;
      if (!(fabs((double)l) <= 1.0e3)) flag = 1;
    }
    __syncthreads();
    const int w = ms - k - 1;
    for (long long e = threadIdx.x; e < (long long)w * w; e += blockDim.x) {
      const int i = k + 1 + (int)(e / w), j = k + 1 + (int)(e % w);
      Fs[(long long)i * ms + j] -= Fs[(long long)i * ms + k] * Fs[(long long)k * ms + j];
    }
    __syncthreads();
  }
  __syncthreads();
  if (threadIdx.x == 0 && flag) atomicMin(bad, d.sfirst[s] + 1);
}

// large fronts, blocked by panels of pw <= 32 columns
// (1) one launch per panel: every workgroup (one wavefront) factors the
// pw x pw diagonal block in LDS (workgroup 0 also stores the factors in
// `dbuf`, which the update launch copies into the front: the other
// workgroups may still be reading the unfactored block), then solves a 64-row block of L21 (x U11 = row) or a 64-column
// block of U12 (L11 x = column), its tile staged through LDS so the global
// loads and stores are coalesced.
template <class T>
__global__ void __launch_bounds__(64) k_mf_panel_solve(T* Fs, int ms, int p0, int pw, int nrb, int col0, T* dbuf,
                                                       int* bad) {
  __shared__ T D[32][33];
  __shared__ T Tl[64][33];
  __shared__ int flag;
  const int t = threadIdx.x;
  if (t == 0) flag = 0;
  for (int e = t; e < pw * pw; e += 64) D[e / pw][e % pw] = Fs[(long long)(p0 + e / pw) * ms + p0 + e % pw];
  __syncthreads();
  for (int k = 0; k < pw; ++k) {
    const T piv = D[k][k];
    if (mf_bad_pivot(piv)) { if (t == 0) flag = 1; break; }
    if (t > k && t < pw) {
      const T l = D[t][k] / piv;
      D[t][k] = l;
      if (!(fabs((double)l) <= 1.0e3)) flag = 1;
      for (int j = k + 1; j < pw; ++j) D[t][j] -= l * D[k][j];
    }
    __syncthreads();
  }
  __syncthreads();
  const int base = p0 + pw;
  if (blockIdx.x == 0) {
    for (int e = t; e < pw * pw; e += 64) dbuf[e] = D[e / pw][e % pw];
  }
  if ((int)blockIdx.x < nrb) {
    // L21 rows [r0, r0 + 64)
    const int r0 = base + blockIdx.x * 64;
    for (int e = t; e < 64 * pw; e += 64) {
      const int rr = e / pw, q = e % pw;
      Tl[rr][q] = r0 + rr < ms ? Fs[(long long)(r0 + rr) * ms + p0 + q] : (T)0;
    }
    __syncthreads();
    if (r0 + t < ms) {
      for (int q = 0; q < pw; ++q) {
        T a = Tl[t][q];
        for (int r = 0; r < q; ++r) a -= Tl[t][r] * D[r][q];
        a /= D[q][q];
        if (!(fabs((double)a) <= 1.0e3)) flag = 1;
        Tl[t][q] = a;
      }
    }
    __syncthreads();
    for (int e = t; e < 64 * pw; e += 64) {
      const int rr = e / pw, q = e % pw;
      if (r0 + rr < ms) Fs[(long long)(r0 + rr) * ms + p0 + q] = Tl[rr][q];
    }
  } else {
    // U12 columns [c0, c0 + 64)
    const int c0 = base + (blockIdx.x - nrb) * 64;
    for (int e = t; e < 64 * pw; e += 64) {
      const int q = e / 64, cc = e % 64;
      Tl[cc][q] = c0 + cc < ms ? Fs[(long long)(p0 + q) * ms + c0 + cc] : (T)0;
    }
    __syncthreads();
    for (int r = 1; r < pw; ++r) {
      T a = Tl[t][r];
      for (int q = 0; q < r; ++q) a -= D[r][q] * Tl[t][q];
      Tl[t][r] = a;
    }
    __syncthreads();
    for (int e = t; e < 64 * pw; e += 64) {
      const int q = e / 64, cc = e % 64;
      if (c0 + cc < ms) Fs[(long long)(p0 + q) * ms + c0 + cc] = Tl[cc][q];
    }
  }
  __syncthreads();
  if (t == 0 && flag && (blockIdx.x == 0 || (int)blockIdx.x < nrb)) atomicMin(bad, col0 + p0 + 1);
}

// (2) trailing update F[i][j] -= sum_q L[i][q] U[q][j], i, j >= p0 + pw,
// 64x64 tiles, the panel's L rows and U columns staged in LDS
template <class T>
__global__ void __launch_bounds__(256) k_mf_update(T* Fs, int ms, int p0, int pw, const T* dbuf) {
  constexpr int TB = 64, KB = 16;
  __shared__ T Ls[TB][KB + 1];
  __shared__ T Us[KB][TB + 1];
  const int base = p0 + pw;
  if (blockIdx.x == 0 && blockIdx.y == 0)    // the panel's diagonal block factors
    for (int e = threadIdx.x; e < pw * pw; e += blockDim.x) Fs[(long long)(p0 + e / pw) * ms + p0 + e % pw] = dbuf[e];
  if (base >= ms) return;
  const int i0 = base + blockIdx.y * TB, j0 = base + blockIdx.x * TB;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;   // 16 x 16 threads, 4 x 4 outputs each
  T acc[4][4];
  for (int a = 0; a < 4; ++a) for (int b = 0; b < 4; ++b) acc[a][b] = (T)0;
  for (int q0 = 0; q0 < pw; q0 += KB) {
    for (int e = threadIdx.x; e < TB * KB; e += 256) {
      const int r = e / KB, q = e % KB;
      const int gi = i0 + r, gq = p0 + q0 + q;
      Ls[r][q] = (gi < ms && q0 + q < pw) ? Fs[(long long)gi * ms + gq] : (T)0;
      const int qq = e / TB, c = e % TB;
      const int gj = j0 + c, gq2 = p0 + q0 + qq;
      Us[qq][c] = (gj < ms && q0 + qq < pw) ? Fs[(long long)gq2 * ms + gj] : (T)0;
    }
    __syncthreads();
    for (int q = 0; q < KB; ++q) {
      T lv[4], uv[4];
      for (int a = 0; a < 4; ++a) lv[a] = Ls[ty * 4 + a][q];
      for (int b = 0; b < 4; ++b) uv[b] = Us[q][tx * 4 + b];
      for (int a = 0; a < 4; ++a) for (int b = 0; b < 4; ++b) acc[a][b] += lv[a] * uv[b];
    }
    __syncthreads();
  }
  for (int a = 0; a < 4; ++a) {
    const int gi = i0 + ty * 4 + a;
    if (gi >= ms) continue;
    for (int b = 0; b < 4; ++b) {
      const int gj = j0 + tx * 4 + b;
      if (gj < ms) Fs[(long long)gi * ms + gj] -= acc[a][b];
    }
  }
}

// forward solve of a level: V_s = b(pivot rows) + children's updates, then
// y = L11^-1 V (unit), V(rows below) -= L21 y; y written to yg
template <class T>
__global__ void __launch_bounds__(256) k_mf_fwd(MfDev d, const int* list, const T* F, T* V, const double* b, T* yg) {
  const int s = list[blockIdx.x];
  const int ms = d.m[s], np = d.sfirst[s + 1] - d.sfirst[s];
  const T* Fs = F + d.fo[s];
  T* Vs = V + d.vo[s];
  const int* ix = d.idx + d.iptr[s];
  for (int t = threadIdx.x; t < ms; t += blockDim.x) Vs[t] = t < np ? (T)b[ix[t]] : (T)0;
  __syncthreads();
  for (int c = d.cptr[s]; c < d.cptr[s + 1]; ++c) {
    const int ch = d.clist[c];
    const int npc = d.sfirst[ch + 1] - d.sfirst[ch], nu = d.m[ch] - npc;
    const T* Vc = V + d.vo[ch];
    const int* em = d.emap + d.eptr[ch];
    for (int a = threadIdx.x; a < nu; a += blockDim.x) Vs[em[a]] += Vc[npc + a];
    __syncthreads();
  }
  for (int k = 0; k < np; ++k) {
    const T yk = Vs[k];
    for (int i = k + 1 + threadIdx.x; i < ms; i += blockDim.x) Vs[i] -= Fs[(long long)i * ms + k] * yk;
    __syncthreads();
  }
  for (int t = threadIdx.x; t < np; t += blockDim.x) yg[ix[t]] = Vs[t];
}

// backward solve of a level (roots first): x = U11^-1 (y - U12 x_below)
template <class T>
__global__ void __launch_bounds__(256) k_mf_bwd(MfDev d, const int* list, const T* F, T* V, T* x) {
  const int s = list[blockIdx.x];
  const int ms = d.m[s], np = d.sfirst[s + 1] - d.sfirst[s];
  const T* Fs = F + d.fo[s];
  T* Vs = V + d.vo[s];
  const int* ix = d.idx + d.iptr[s];
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    T v = x[ix[t]];                 // holds y on entry
    for (int u = np; u < ms; ++u) v -= Fs[(long long)t * ms + u] * x[ix[u]];
    Vs[t] = v;
  }
  __syncthreads();
  for (int k = np - 1; k >= 0; --k) {
    if (threadIdx.x == 0) Vs[k] = Vs[k] / Fs[(long long)k * ms + k];
    __syncthreads();
    const T xk = Vs[k];
    for (int t = threadIdx.x; t < k; t += blockDim.x) Vs[t] -= Fs[(long long)t * ms + k] * xk;
    __syncthreads();
  }
  for (int t = threadIdx.x; t < np; t += blockDim.x) x[ix[t]] = Vs[t];
}

template <class T>
__global__ void k_mf_perm_in(int n, const int* prow, const double* b, double* bp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) bp[i] = b[prow[i]];
}
template <class T>
__global__ void k_mf_perm_out(int n, const int* pcol, const T* xp, double* x, int* nan_flag) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) {
    const double v = (double)xp[j];
    x[pcol[j]] = v;
    if (!isfinite(v)) atomicOr(nan_flag, 1);
  }
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
#define MFCHK(x)                           \
  do {                                     \
    hipError_t e_ = (x);                   \
    if (e_ != hipSuccess) return XG_E_HIP; \
  } while (0)

template <class T>
struct DevVec {
  T* p = nullptr;
  ~DevVec() { if (p) (void)hipFree(p); }
  hipError_t up(const std::vector<T>& v, hipStream_t s) {
    if (p) { (void)hipFree(p); p = nullptr; }
    if (v.empty()) return hipSuccess;
    hipError_t e = hipMalloc((void**)&p, v.size() * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
  }
  hipError_t alloc(size_t n) {
    if (p) { (void)hipFree(p); p = nullptr; }
    return n ? hipMalloc((void**)&p, n * sizeof(T)) : hipSuccess;
  }
};

constexpr int kMfSmall = 160;   // fronts up to this order are factored by one workgroup
constexpr int kMfPanel = 32;

struct MfLu::Impl {
  MfPlan P;
  hipStream_t st = nullptr;
  DevVec<int> sfirst, m, iptr, idx, cptr, clist, eptr, emap, vo, lvl_s, prow, pcol, bad;
  DevVec<long long> fo, amap;
  DevVec<double> F64, V64, y64, b64;
  DevVec<float> F32, V32, y32;
  std::vector<std::vector<int>> small_of_level, large_of_level;
  DevVec<int> small_list, ea_list, sparent;
  DevVec<double> dbuf;            // one panel's diagonal-block factors (FP64 or FP32)
  std::vector<int> small_off;     // per level: offset into small_list
  // extend-add launches: per level, per child index j: [offset, count, max rows]
  struct EaLaunch { int off, count, max_nu; };
  std::vector<std::vector<EaLaunch>> ea_of_level;
  MfDev dev() const {
    MfDev d;
    d.sfirst = sfirst.p; d.m = m.p; d.fo = fo.p; d.iptr = iptr.p; d.idx = idx.p; d.cptr = cptr.p;
    d.clist = clist.p; d.eptr = eptr.p; d.emap = emap.p; d.vo = vo.p;
    return d;
  }
};

MfLu::MfLu() : impl_(new Impl) {}
MfLu::~MfLu() { delete impl_; }

int MfLu::setup(const LuSymbolic& s, const int* rowptr, const int* colind, void* stream, bool want_fp32) {
  Impl& I = *impl_;
  I.st = (hipStream_t)stream;
  int rc = mf_plan(s, rowptr, colind, I.P);
  if (rc != XG_OK) return rc;
  const MfPlan& P = I.P;
  // every index a kernel reads is checked here
  for (int v : P.idx) if (v < 0 || v >= P.n) return XG_E_SETUP;
  for (long long a : P.amap) if (a < 0 || a >= P.fo[P.ns]) return XG_E_SETUP;
  for (int c : P.clist) if (c < 0 || c >= P.ns) return XG_E_SETUP;
  for (int t = 0; t < P.ns; ++t)
    for (int e = P.eptr[t]; e < P.eptr[t + 1]; ++e)
      if (P.sparent[t] < 0 || P.emap[e] < 0 || P.emap[e] >= P.m[P.sparent[t]]) return XG_E_SETUP;
  hipStream_t st = I.st;
  MFCHK(I.sfirst.up(P.sfirst, st)); MFCHK(I.m.up(P.m, st)); MFCHK(I.iptr.up(P.iptr, st)); MFCHK(I.idx.up(P.idx, st));
  MFCHK(I.cptr.up(P.cptr, st)); MFCHK(I.clist.up(P.clist, st)); MFCHK(I.eptr.up(P.eptr, st));
  MFCHK(I.emap.up(P.emap, st)); MFCHK(I.vo.up(P.vo, st)); MFCHK(I.fo.up(P.fo, st)); MFCHK(I.amap.up(P.amap, st));
  MFCHK(I.prow.up(s.perm_row, st)); MFCHK(I.pcol.up(s.perm_col, st));
  MFCHK(I.bad.alloc(2));
  MFCHK(I.dbuf.alloc(32 * 32));
  MFCHK(I.F64.alloc(P.fo[P.ns])); MFCHK(I.V64.alloc(P.vo[P.ns])); MFCHK(I.y64.alloc(P.n)); MFCHK(I.b64.alloc(P.n));
  if (want_fp32) { MFCHK(I.F32.alloc(P.fo[P.ns])); MFCHK(I.V32.alloc(P.vo[P.ns])); MFCHK(I.y32.alloc(P.n)); }
  // per level: small fronts (one launch) and large fronts (blocked, one by one)
  const int nl = (int)P.lvl_ptr.size() - 1;
  std::vector<int> sl;
  I.small_off.assign(nl + 1, 0);
  I.large_of_level.assign(nl, {});
  for (int l = 0; l < nl; ++l) {
    I.small_off[l] = (int)sl.size();
    for (int q = P.lvl_ptr[l]; q < P.lvl_ptr[l + 1]; ++q) {
      const int t = P.lvl_s[q];
      if (P.m[t] <= kMfSmall) sl.push_back(t); else I.large_of_level[l].push_back(t);
    }
  }
  I.small_off[nl] = (int)sl.size();
  MFCHK(I.small_list.up(sl, st));
  // extend-add launches: the j-th child of every front of a level together
  std::vector<int> ea;
  I.ea_of_level.assign(nl, {});
  for (int l = 1; l < nl; ++l) {
    int maxc = 0;
    for (int q = P.lvl_ptr[l]; q < P.lvl_ptr[l + 1]; ++q) {
      const int t = P.lvl_s[q];
      maxc = std::max(maxc, P.cptr[t + 1] - P.cptr[t]);
    }
    for (int j = 0; j < maxc; ++j) {
      Impl::EaLaunch L{(int)ea.size(), 0, 0};
      for (int q = P.lvl_ptr[l]; q < P.lvl_ptr[l + 1]; ++q) {
        const int t = P.lvl_s[q];
        if (P.cptr[t + 1] - P.cptr[t] <= j) continue;
        const int ch = P.clist[P.cptr[t] + j];
        const int nu = P.m[ch] - (P.sfirst[ch + 1] - P.sfirst[ch]);
        if (nu == 0) continue;
        ea.push_back(ch);
        L.count++;
        L.max_nu = std::max(L.max_nu, nu);
      }
      if (L.count) I.ea_of_level[l].push_back(L);
    }
  }
  MFCHK(I.ea_list.up(ea, st));
  MFCHK(I.sparent.up(P.sparent, st));
  MFCHK(I.lvl_s.up(P.lvl_s, st));
  MFCHK(hipStreamSynchronize(st));
  return XG_OK;
}

double MfLu::flops() const { return impl_->P.flops; }
long long MfLu::front_elements() const { return impl_->P.fo.empty() ? 0 : impl_->P.fo[impl_->P.ns]; }
int MfLu::levels() const { return (int)impl_->P.lvl_ptr.size() - 1; }
int MfLu::supernodes() const { return impl_->P.ns; }

template <class T>
static int mf_factor_t(MfLu::Impl& I, const double* vals_dev, int nnz, T* F, int* bad_out) {
  const MfPlan& P = I.P;
  hipStream_t st = I.st;
  const long long tot = P.fo[P.ns];
  hipLaunchKernelGGL(k_mf_fill<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, F, tot);
  hipLaunchKernelGGL(k_mf_scatter<T>, dim3((nnz + 255) / 256), dim3(256), 0, st, nnz, vals_dev, I.amap.p, F);
  const int init[2] = {0x7fffffff, 0};
  MFCHK(hipMemcpyAsync(I.bad.p, init, sizeof(init), hipMemcpyHostToDevice, st));
  const MfDev d = I.dev();
  const int nl = (int)P.lvl_ptr.size() - 1;
  for (int l = 0; l < nl; ++l) {
    const int cnt = P.lvl_ptr[l + 1] - P.lvl_ptr[l];
    (void)cnt;
    for (const auto& L : I.ea_of_level[l])
      hipLaunchKernelGGL(k_mf_extend_add<T>, dim3(L.count, (L.max_nu + 15) / 16), dim3(256), 0, st, d,
                         I.ea_list.p + L.off, I.sparent.p, F);
    const int ns_small = I.small_off[l + 1] - I.small_off[l];
    if (ns_small)
      hipLaunchKernelGGL(k_mf_factor_small<T>, dim3(ns_small), dim3(256), 0, st, d, I.small_list.p + I.small_off[l], F,
                         I.bad.p);
    for (int t : I.large_of_level[l]) {
      T* Fs = F + P.fo[t];
      const int ms = P.m[t], np = P.sfirst[t + 1] - P.sfirst[t];
      for (int p0 = 0; p0 < np; p0 += kMfPanel) {
        const int pw = std::min(kMfPanel, np - p0);
        const int rest = ms - p0 - pw;
        const int nb = (rest + 63) / 64;
        T* dbuf = reinterpret_cast<T*>(I.dbuf.p);
        hipLaunchKernelGGL(k_mf_panel_solve<T>, dim3(std::max(1, 2 * nb)), dim3(64), 0, st, Fs, ms, p0, pw, nb,
                           P.sfirst[t], dbuf, I.bad.p);
        const unsigned tiles = (unsigned)std::max(1, (rest + 63) / 64);
        hipLaunchKernelGGL(k_mf_update<T>, dim3(tiles, tiles), dim3(256), 0, st, Fs, ms, p0, pw, (const T*)dbuf);
      }
    }
  }
  MFCHK(hipGetLastError());
  int res[2];
  MFCHK(hipMemcpyAsync(res, I.bad.p, sizeof(res), hipMemcpyDeviceToHost, st));
  MFCHK(hipStreamSynchronize(st));
  *bad_out = res[0] == 0x7fffffff ? 0 : res[0];
  return XG_OK;
}

template <class T>
static int mf_solve_t(MfLu::Impl& I, const T* F, T* V, T* y, const double* b_dev, double* x_dev, int* nan_flag) {
  const MfPlan& P = I.P;
  hipStream_t st = I.st;
  const MfDev d = I.dev();
  const int n = P.n;
  hipLaunchKernelGGL(k_mf_perm_in<T>, dim3((n + 255) / 256), dim3(256), 0, st, n, I.prow.p, b_dev, I.b64.p);
  const int nl = (int)P.lvl_ptr.size() - 1;
  for (int l = 0; l < nl; ++l) {
    const int cnt = P.lvl_ptr[l + 1] - P.lvl_ptr[l];
    hipLaunchKernelGGL(k_mf_fwd<T>, dim3(cnt), dim3(256), 0, st, d, I.lvl_s.p + P.lvl_ptr[l], F, V, I.b64.p, y);
  }
  for (int l = nl - 1; l >= 0; --l) {
    const int cnt = P.lvl_ptr[l + 1] - P.lvl_ptr[l];
    hipLaunchKernelGGL(k_mf_bwd<T>, dim3(cnt), dim3(256), 0, st, d, I.lvl_s.p + P.lvl_ptr[l], F, V, y);
  }
  hipLaunchKernelGGL(k_mf_perm_out<T>, dim3((n + 255) / 256), dim3(256), 0, st, n, I.pcol.p, y, x_dev, nan_flag);
  MFCHK(hipGetLastError());
  return XG_OK;
}

int MfLu::factor(const double* vals_dev, int nnz, bool fp32, int* bad) {
  Impl& I = *impl_;
  if (fp32) {
    if (!I.F32.p) return XG_E_ARG;
    return mf_factor_t<float>(I, vals_dev, nnz, I.F32.p, bad);
  }
  return mf_factor_t<double>(I, vals_dev, nnz, I.F64.p, bad);
}

int MfLu::solve(const double* b_dev, double* x_dev, bool fp32, int* nan_flag) {
  Impl& I = *impl_;
  if (fp32) {
    if (!I.F32.p) return XG_E_ARG;
    return mf_solve_t<float>(I, I.F32.p, I.V32.p, I.y32.p, b_dev, x_dev, nan_flag);
  }
  return mf_solve_t<double>(I, I.F64.p, I.V64.p, I.y64.p, b_dev, x_dev, nan_flag);
}

}  // namespace xg
