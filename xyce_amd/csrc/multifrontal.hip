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

constexpr int kMfTiny = 64;     // fronts up to this order: one wavefront, the front in LDS
constexpr int kMfPanel = 32;    // panel width of the blocked factorisation (64 spills the register panel)
constexpr int kMfRowsPerLoad = 64 / kMfPanel;   // L21 rows a 64-lane load covers
constexpr int kMfSolveBlock = 1024;

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

// value of v in lane `src` (wave-uniform) of the wavefront
__device__ __forceinline__ double mf_bcast(double v, int src) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), src);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float mf_bcast(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

template <class T>
__device__ __forceinline__ T mf_wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// assembly of the children's Schur complements into a larger front: one
// workgroup per kMfAsmRows rows of the parent (entries {front, first row,
// offset into `alo`}); the children are added one after another in their
// fixed order, so every parent element receives the same sums in the same
// order on every refactor (no atomics).  A child's rows map to increasing
// parent rows; alo holds, per entry and child, the first child row that lands
// at or after the entry's first row.  The four wavefronts take the child rows
// in turn (wave w: rows a0 + w, a0 + w + 4, ...), lanes over 256-column
// chunks; a chunk's parent column positions are loaded once, and each wave
// loads four rows' child and parent values before it stores any, so a row
// costs two memory round trips shared with the wave's other rows.
constexpr int kMfAsmRows = 16;

template <class T>
__global__ void __launch_bounds__(256) k_mf_assemble(MfDev d, const int4* ent, const int* alo, T* F) {
  const int4 e = ent[blockIdx.x];
  const int s = e.x, ms = d.m[s];
  const int r1 = min(ms, e.y + kMfAsmRows);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T* Fs = F + d.fo[s];
  for (int c = d.cptr[s]; c < d.cptr[s + 1]; ++c) {
    const int ch = d.clist[c];
    const int mc = d.m[ch], npc = d.sfirst[ch + 1] - d.sfirst[ch], nu = mc - npc;
    const int* em = d.emap + d.eptr[ch];
    const T* Fc = F + d.fo[ch];
    const int a0 = alo[e.z + (c - d.cptr[s])] + wave;
    for (int c0 = 0; c0 < nu; c0 += 256) {
      const int b0 = c0 + lane;
      int eb[4];
      bool ok[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int b = b0 + 64 * q;
        ok[q] = b < nu;
        eb[q] = ok[q] ? em[b] : 0;
      }
      for (int a = a0; a < nu; a += 16) {
        // rows a, a + 4, a + 8, a + 12 of this wave (those that land in the range)
        int pr[4];
        bool in[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int ak = a + 4 * k;
          pr[k] = ak < nu ? em[ak] : r1;
          in[k] = pr[k] < r1;
        }
        if (!in[0]) break;
        T cv[4][4], pv[4][4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const T* crow = Fc + (long long)(npc + a + 4 * k) * mc + npc;
          const T* prow = Fs + (long long)(in[k] ? pr[k] : 0) * ms;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            cv[k][q] = (in[k] && ok[q]) ? crow[b0 + 64 * q] : (T)0;
            pv[k][q] = (in[k] && ok[q]) ? prow[eb[q]] : (T)0;
          }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          T* prow = Fs + (long long)(in[k] ? pr[k] : 0) * ms;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (in[k] && ok[q]) prow[eb[q]] = pv[k][q] + cv[k][q];
        }
        if (!in[3]) break;
      }
    }
    __syncthreads();
  }
}

template <class T>
__device__ __forceinline__ bool mf_bad_pivot(T d) {
  const double v = (double)d;
  return !(fabs(v) > 0.0) || !isfinite(v);
}

// fronts of order <= 64: one wavefront per front, the front staged in LDS
// (row stride m + 1) and its children's updates added there, then
// right-looking over its pivots with lane i owning row i
template <class T>
__global__ void __launch_bounds__(64) k_mf_factor_tiny(MfDev d, const int* list, T* F, int* bad) {
  extern __shared__ unsigned char mf_sm[];
  T* A = reinterpret_cast<T*>(mf_sm);
  const int s = list[blockIdx.x];
  const int ms = d.m[s], np = d.sfirst[s + 1] - d.sfirst[s];
  const int ld = ms + 1;
  T* Fs = F + d.fo[s];
  const int t = threadIdx.x;
  for (int i0 = 0; i0 < ms; i0 += 16) {
    T v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = (t < ms && i0 + q < ms) ? Fs[(long long)(i0 + q) * ms + t] : (T)0;
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (t < ms && i0 + q < ms) A[(i0 + q) * ld + t] = v[q];
  }
  __syncthreads();
  // the children's Schur complements, one child after another (a child's
  // update has at most ms rows and columns: lane t adds column t)
  for (int c = d.cptr[s]; c < d.cptr[s + 1]; ++c) {
    const int ch = d.clist[c];
    const int mc = d.m[ch], npc = d.sfirst[ch + 1] - d.sfirst[ch], nu = mc - npc;
    const int* em = d.emap + d.eptr[ch];
    const T* Fc = F + d.fo[ch];
    const int eb = t < nu ? em[t] : 0;     // lane t: parent position of the child's row/column t
    const T* cb = Fc + (long long)npc * mc + npc + t;
    int a = 0;
    for (; a + 16 <= nu; a += 16) {
      T v[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = t < nu ? cb[(long long)(a + q) * mc] : (T)0;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int pr = __builtin_amdgcn_readlane(eb, a + q);
        if (t < nu) A[pr * ld + eb] += v[q];
      }
    }
    for (; a + 4 <= nu; a += 4) {
      T v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = t < nu ? cb[(long long)(a + q) * mc] : (T)0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int pr = __builtin_amdgcn_readlane(eb, a + q);
        if (t < nu) A[pr * ld + eb] += v[q];
      }
    }
    for (; a < nu; ++a) {
      const int pr = __builtin_amdgcn_readlane(eb, a);
      if (t < nu) A[pr * ld + eb] += cb[(long long)a * mc];
    }
    __syncthreads();
  }
  bool fail = false, hard = false;
  for (int k = 0; k < np; ++k) {
    const T piv = A[k * ld + k];
    if (mf_bad_pivot(piv)) { fail = true; hard = true; break; }
    if (t > k && t < ms) {
      T* ri = A + t * ld;
      const T* rk = A + k * ld;
      const T l = ri[k] / piv;
      ri[k] = l;
      if (!(fabs((double)l) <= 1.0e3)) fail = true;
      int j = k + 1;
      for (; j + 8 <= ms; j += 8) {
        T u[8], a[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) u[q] = rk[j + q];
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = ri[j + q];
#pragma unroll
        for (int q = 0; q < 8; ++q) ri[j + q] = a[q] - l * u[q];
      }
      for (; j < ms; ++j) ri[j] -= l * rk[j];
    }
    __syncthreads();
  }
  for (int i = 0; i < ms; ++i)
    if (t < ms) Fs[(long long)i * ms + t] = A[i * ld + t];
  if (__any(fail) && t == 0) atomicMin(bad, d.sfirst[s] + 1);
  if (hard && t == 0) atomicOr(bad + 1, 1);
}

// larger fronts, blocked by panels of kMfPanel columns; all fronts of a level
// in one launch per panel index (entries {front, block, dbuf slot, 0}).
// (1) every workgroup (one wavefront) factors the pw x pw diagonal block in
// registers (lane t owns row t, the pivot row broadcast across lanes), then
// solves a 64-row block of L21 (x U11 = row) or a 64-column block of U12
// (L11 x = column).  Block 0 also stores the block's factors in its dbuf
// slot; the update launch copies them into the front (the front's other
// workgroups may still be reading the unfactored block).
template <class T>
__global__ void __launch_bounds__(64) k_mf_panel(MfDev d, const int4* ent, int p0, T* F, T* dbuf, int* bad) {
  __shared__ T D[kMfPanel][kMfPanel + 1];
  __shared__ T Rd[kMfPanel];
  __shared__ T Tl[64][kMfPanel + 1];
  const int4 e = ent[blockIdx.x];
  const int s = e.x, blk = e.y, slot = e.z;
  const int ms = d.m[s], np = d.sfirst[s + 1] - d.sfirst[s];
  const int pw = min(kMfPanel, np - p0);
  const int base = p0 + pw, rest = ms - base, nrb = (rest + 63) / 64;
  T* Fs = F + d.fo[s];
  const int t = threadIdx.x;
  T r[kMfPanel];
#pragma unroll
  for (int q = 0; q < kMfPanel; ++q)
    r[q] = (t < pw && q < pw) ? Fs[(long long)(p0 + t) * ms + p0 + q] : (T)0;
  // (loops keep constant trip counts so that the register arrays stay
  // register-resident after full unrolling)
  bool fail = false, stop = false;
#pragma unroll
  for (int k = 0; k < kMfPanel; ++k) {
    if (k < pw && !stop) {
      const T piv = mf_bcast(r[k], k);
      if (mf_bad_pivot(piv)) {
        fail = true;
        stop = true;
      } else {
        const bool upd = t > k && t < pw;
        const T l = upd ? r[k] / piv : (T)0;
        if (upd) {
          r[k] = l;
          if (!(fabs((double)l) <= 1.0e3)) fail = true;
        }
#pragma unroll
        for (int q = 0; q < kMfPanel; ++q) {
          if (q > k) {
            const T u = mf_bcast(r[q], k);
            if (upd) r[q] -= l * u;
          }
        }
      }
    }
  }
  if (t < kMfPanel) {
#pragma unroll
    for (int q = 0; q < kMfPanel; ++q) D[t][q] = r[q];
    // pivot reciprocals, all at once: the row solves below multiply
    Rd[t] = t < pw ? (T)1 / D[t][t] : (T)0;
  }
  __syncthreads();
  if (blk == 0) {
    T* db = dbuf + (long long)slot * kMfPanel * kMfPanel;
    for (int i = 0; i < pw; ++i)
      if (t < pw) db[i * kMfPanel + t] = D[i][t];
  }
  if (rest > 0 && blk < nrb) {
    // L21 rows [r0, r0 + 64), staged through LDS for coalesced access
    const int r0 = base + blk * 64;
    {
      // all loads issued before the LDS stores (one wait, not one per row)
      T v[kMfPanel];
#pragma unroll
      for (int it = 0; it < kMfPanel; ++it) {
        const int rr = t / kMfPanel + kMfRowsPerLoad * it, q = t % kMfPanel;
        v[it] = (r0 + rr < ms && q < pw) ? Fs[(long long)(r0 + rr) * ms + p0 + q] : (T)0;
      }
#pragma unroll
      for (int it = 0; it < kMfPanel; ++it) Tl[t / kMfPanel + kMfRowsPerLoad * it][t % kMfPanel] = v[it];
    }
    __syncthreads();
    T a[kMfPanel];
#pragma unroll
    for (int q = 0; q < kMfPanel; ++q) a[q] = Tl[t][q];
    const bool valid = r0 + t < ms;
#pragma unroll
    for (int k = 0; k < kMfPanel; ++k) {
      if (k < pw) {
        a[k] = a[k] * Rd[k];
        if (valid && !(fabs((double)a[k]) <= 1.0e3)) fail = true;
#pragma unroll
        for (int q = 0; q < kMfPanel; ++q)
          if (q > k) a[q] -= a[k] * D[k][q];
      }
    }
#pragma unroll
    for (int q = 0; q < kMfPanel; ++q) Tl[t][q] = a[q];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kMfPanel; ++it) {
      const int rr = t / kMfPanel + kMfRowsPerLoad * it, q = t % kMfPanel;
      if (r0 + rr < ms && q < pw) Fs[(long long)(r0 + rr) * ms + p0 + q] = Tl[rr][q];
    }
  } else if (rest > 0) {
    // U12 columns [c0, c0 + 64): lane t owns column c0 + t
    const int c = base + (blk - nrb) * 64 + t;
    const bool valid = c < ms;
    T b[kMfPanel];
#pragma unroll
    for (int q = 0; q < kMfPanel; ++q) b[q] = (valid && q < pw) ? Fs[(long long)(p0 + q) * ms + c] : (T)0;
#pragma unroll
    for (int k = 0; k < kMfPanel; ++k) {
      if (k < pw) {
#pragma unroll
        for (int q = 0; q < kMfPanel; ++q)
          if (q > k) b[q] -= D[q][k] * b[k];
      }
    }
#pragma unroll
    for (int q = 0; q < kMfPanel; ++q)
      if (valid && q < pw) Fs[(long long)(p0 + q) * ms + c] = b[q];
  }
  if (__any(fail) && t == 0) atomicMin(bad, d.sfirst[s] + p0 + 1);
  if (__any(stop) && t == 0) atomicOr(bad + 1, 1);
}

// (2) trailing update F[i][j] -= sum_q L[i][q] U[q][j], i, j >= p0 + pw, one
// 64x64 tile per workgroup (entries {front, tile row, tile column, dbuf
// slot}; tile row -1: the front has nothing below the panel), the panel's L
// rows and U columns staged in LDS.  Tile (0, 0) or the -1 entry copies the
// panel's diagonal-block factors from dbuf into the front.
// Trailing update of the larger fronts on the matrix cores: the 64x64 tile
// of a workgroup as 4 wavefronts x 2 x 2 blocks of 16 x 16, rank-4 products
// per v_mfma_{f64,f32}_16x16x4 (A: lane l holds row l % 16, column l / 16 of
// the 16 x 4 block; B: row l / 16, column l % 16 of the 4 x 16 block;
// accumulator register r of lane l: column l % 16, row l / 16 + 4 r for FP64
// and 4 (l / 16) + r for FP32).
typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
template <class T> struct MfAcc;
template <> struct MfAcc<double> {
  typedef mf_d4 V;
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
  static __device__ __forceinline__ V mma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct MfAcc<float> {
  typedef mf_f4 V;
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
  static __device__ __forceinline__ V mma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <class T>
__global__ void __launch_bounds__(256) k_mf_update(MfDev d, const int4* ent, int p0, T* F, const T* dbuf) {
  constexpr int TB = 64, KB = kMfPanel;
  __shared__ T Ls[TB][KB + 1];
  __shared__ T Us[KB][TB + 1];
  const int4 e = ent[blockIdx.x];
  const int s = e.x;
  const int ms = d.m[s], np = d.sfirst[s + 1] - d.sfirst[s];
  const int pw = min(kMfPanel, np - p0);
  const int base = p0 + pw;
  T* Fs = F + d.fo[s];
  if (e.y <= 0 && e.z <= 0) {
    const T* db = dbuf + (long long)e.w * kMfPanel * kMfPanel;
    for (int q = threadIdx.x; q < pw * kMfPanel; q += blockDim.x) {
      const int i = q / kMfPanel, j = q % kMfPanel;
      if (j < pw) Fs[(long long)(p0 + i) * ms + p0 + j] = db[q];
    }
  }
  if (e.y < 0) return;
  const int i0 = base + e.y * TB, j0 = base + e.z * TB;
  {
    // the L21 rows and U12 columns of the tile (zero beyond the front and the
    // panel), all loads issued before the LDS stores
    T lv[TB * KB / 256], uv[TB * KB / 256];
#pragma unroll
    for (int it = 0; it < TB * KB / 256; ++it) {
      const int x = threadIdx.x + it * 256;
      const int r = x / KB, q = x % KB;
      const int gi = i0 + r;
      lv[it] = (gi < ms && q < pw) ? Fs[(long long)gi * ms + p0 + q] : (T)0;
      const int qq = x / TB, c = x % TB;
      const int gj = j0 + c;
      uv[it] = (gj < ms && qq < pw) ? Fs[(long long)(p0 + qq) * ms + gj] : (T)0;
    }
#pragma unroll
    for (int it = 0; it < TB * KB / 256; ++it) {
      const int x = threadIdx.x + it * 256;
      Ls[x / KB][x % KB] = lv[it];
      Us[x / TB][x % TB] = uv[it];
    }
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32, lr = lane & 15, lk = lane >> 4;
  typedef typename MfAcc<T>::V V;
  V acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = V{};
#pragma unroll
  for (int k = 0; k < KB; k += 4) {
    const T a0 = Ls[wr + lr][k + lk], a1 = Ls[wr + 16 + lr][k + lk];
    const T b0 = Us[k + lk][wc + lr], b1 = Us[k + lk][wc + 16 + lr];
    acc[0][0] = MfAcc<T>::mma(a0, b0, acc[0][0]);
    acc[0][1] = MfAcc<T>::mma(a0, b1, acc[0][1]);
    acc[1][0] = MfAcc<T>::mma(a1, b0, acc[1][0]);
    acc[1][1] = MfAcc<T>::mma(a1, b1, acc[1][1]);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = i0 + wr + 16 * a + MfAcc<T>::row(lk, r), gj = j0 + wc + 16 * b + lr;
        if (gi < ms && gj < ms) Fs[(long long)gi * ms + gj] -= acc[a][b][r];
      }
}

// sums over c in [c0, c1) of F[r + k][c] v[c] for the rows r .. r + nr - 1
// (nr <= 4, wave-uniform) of a row-major front, by one wavefront: every lane
// keeps eight loads in flight (four rows, two column blocks), so the four rows
// share their memory round trips; row k's sum is returned in out[k] in every
// lane
template <class T>
__device__ __forceinline__ void mf_rows_dot4(const T* Fs, long long ms, int r, int nr, const T* v, int c0, int c1,
                                             int lane, T out[4]) {
  const T* row[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) row[k] = Fs + (long long)(r + (k < nr ? k : 0)) * ms;
  T a[4], b[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) { a[k] = (T)0; b[k] = (T)0; }
  int c = c0 + lane;
  for (; c + 64 < c1; c += 128) {
    const T v0 = v[c], v1 = v[c + 64];
    T f0[4], f1[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) { f0[k] = row[k][c]; f1[k] = row[k][c + 64]; }
#pragma unroll
    for (int k = 0; k < 4; ++k) { a[k] += f0[k] * v0; b[k] += f1[k] * v1; }
  }
  if (c < c1) {
    const T v0 = v[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] += row[k][c] * v0;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) out[k] = mf_wave_sum(a[k] + b[k]);
}

// ---- solves.  V holds, per front, the vector of its index list: y on the
// pivot rows, the update for the parent below them.

// sum the children's updates into the front's vector Vl (LDS)
template <class T>
__device__ __forceinline__ void mf_gather_children(const MfDev& d, int s, const T* V, T* Vl) {
  for (int c = d.cptr[s]; c < d.cptr[s + 1]; ++c) {
    const int ch = d.clist[c];
    const int npc = d.sfirst[ch + 1] - d.sfirst[ch], nu = d.m[ch] - npc;
    const T* Vc = V + d.vo[ch];
    const int* em = d.emap + d.eptr[ch];
    for (int a = threadIdx.x; a < nu; a += blockDim.x) Vl[em[a]] += Vc[npc + a];
    __syncthreads();
  }
}

// forward, fronts of order <= 64: one wavefront, lane i owns row i
template <class T>
__global__ void __launch_bounds__(64) k_mf_fwd_tiny(MfDev d, const int* list, const T* F, T* V, const double* b,
                                                    T* yg) {
  __shared__ T Vl[kMfTiny];
  const int s = list[blockIdx.x];
  const int ms = d.m[s], np = d.sfirst[s + 1] - d.sfirst[s];
  const T* Fs = F + d.fo[s];
  const int* ix = d.idx + d.iptr[s];
  const int t = threadIdx.x;
  Vl[t] = t < np ? (T)b[ix[t]] : (T)0;
  __syncthreads();
  mf_gather_children(d, s, V, Vl);
  T v = Vl[t];
  for (int k = 0; k < np; ++k) {
    const T yk = mf_bcast(v, k);
    if (t > k && t < ms) v -= Fs[(long long)t * ms + k] * yk;
  }
  if (t < ms) V[d.vo[s] + t] = v;
  if (t < np) yg[ix[t]] = v;
}

// forward, larger fronts: one workgroup per front, the vector in LDS; the
// unit-lower L11 by 64-row blocks (rows updated by a row-wise product with the
// solved part, then the block's triangle solved by one wavefront from LDS),
// then the rows below the pivots updated by L21 y
template <class T>
__global__ void __launch_bounds__(kMfSolveBlock) k_mf_fwd_big(MfDev d, const int* list, const T* F, T* V,
                                                             const double* b, T* yg, int ldsv) {
  extern __shared__ unsigned char mf_sm[];
  T* Bl = reinterpret_cast<T*>(mf_sm);          // 64 x 65 block
  const int s = list[blockIdx.x];
  T* Vl = ldsv ? Bl + 64 * 65 : V + d.vo[s];    // the front's vector (ms)
  const int ms = d.m[s], np = d.sfirst[s + 1] - d.sfirst[s];
  const T* Fs = F + d.fo[s];
  const int* ix = d.idx + d.iptr[s];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = blockDim.x >> 6;
  for (int i = t; i < ms; i += blockDim.x) Vl[i] = i < np ? (T)b[ix[i]] : (T)0;
  __syncthreads();
  mf_gather_children(d, s, V, Vl);
  for (int b0 = 0; b0 < np; b0 += 64) {
    const int bw = min(64, np - b0);
    if (b0 > 0) {
      for (int rr = 4 * wave; rr < bw; rr += 4 * nw) {
        const int nr = min(4, bw - rr);
        T acc[4];
        mf_rows_dot4(Fs, ms, b0 + rr, nr, Vl, 0, b0, lane, acc);
        if (lane == 0)
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (k < nr) Vl[b0 + rr + k] -= acc[k];
      }
    }
    {
      T v[64 * 64 / kMfSolveBlock];
#pragma unroll
      for (int it = 0; it < 64 * 64 / kMfSolveBlock; ++it) {
        const int x = t + it * kMfSolveBlock, i = x >> 6, j = x & 63;
        v[it] = (i < bw && j < i) ? Fs[(long long)(b0 + i) * ms + b0 + j] : (T)0;
      }
#pragma unroll
      for (int it = 0; it < 64 * 64 / kMfSolveBlock; ++it) {
        const int x = t + it * kMfSolveBlock;
        Bl[(x >> 6) * 65 + (x & 63)] = v[it];
      }
    }
    __syncthreads();
    if (wave == 0) {
      T v = lane < bw ? Vl[b0 + lane] : (T)0;
      for (int k = 0; k < bw; ++k) {
        const T yk = mf_bcast(v, k);
        if (lane > k && lane < bw) v -= Bl[lane * 65 + k] * yk;
      }
      if (lane < bw) Vl[b0 + lane] = v;
    }
    __syncthreads();
  }
  for (int rr = np + 4 * wave; rr < ms; rr += 4 * nw) {
    const int nr = min(4, ms - rr);
    T acc[4];
    mf_rows_dot4(Fs, ms, rr, nr, Vl, 0, np, lane, acc);
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k < nr) Vl[rr + k] -= acc[k];
  }
  __syncthreads();
  if (ldsv)
    for (int i = t; i < ms; i += blockDim.x) V[d.vo[s] + i] = Vl[i];
  for (int i = t; i < np; i += blockDim.x) yg[ix[i]] = Vl[i];
}

// backward, fronts of order <= 64 (roots first): x = U11^-1 (y - U12 x_below)
template <class T>
__global__ void __launch_bounds__(64) k_mf_bwd_tiny(MfDev d, const int* list, const T* F, T* x) {
  __shared__ T Xl[kMfTiny];
  const int s = list[blockIdx.x];
  const int ms = d.m[s], np = d.sfirst[s + 1] - d.sfirst[s];
  const T* Fs = F + d.fo[s];
  const int* ix = d.idx + d.iptr[s];
  const int t = threadIdx.x;
  if (t < ms) Xl[t] = x[ix[t]];           // y on the pivot rows, solved x below
  __syncthreads();
  T v = Xl[t];
  if (t < np) {
    const T* row = Fs + (long long)t * ms;
    int u = np;
    for (; u + 4 <= ms; u += 4) {
      const T f0 = row[u], f1 = row[u + 1], f2 = row[u + 2], f3 = row[u + 3];
      v -= f0 * Xl[u];
      v -= f1 * Xl[u + 1];
      v -= f2 * Xl[u + 2];
      v -= f3 * Xl[u + 3];
    }
    for (; u < ms; ++u) v -= row[u] * Xl[u];
  }
  // each lane's pivot reciprocal first (in parallel), so that the serial
  // sweep multiplies instead of dividing
  const T rd = t < np ? (T)1 / Fs[(long long)t * ms + t] : (T)0;
  for (int k = np - 1; k >= 0; --k) {
    const T xk = mf_bcast(v, k) * mf_bcast(rd, k);
    if (t == k) v = xk;
    else if (t < k) v -= Fs[(long long)t * ms + k] * xk;
  }
  if (t < np) x[ix[t]] = v;
}

// backward, larger fronts: rows minus U12 x_below (row-wise products), then
// U11 by 64-row blocks from the bottom
template <class T>
__global__ void __launch_bounds__(kMfSolveBlock) k_mf_bwd_big(MfDev d, const int* list, const T* F, T* V, T* x,
                                                             int ldsv) {
  extern __shared__ unsigned char mf_sm[];
  T* Bl = reinterpret_cast<T*>(mf_sm);          // 64 x 65 block
  const int s = list[blockIdx.x];
  T* Xl = ldsv ? Bl + 64 * 65 : V + d.vo[s];    // scratch (ms); V is not read by the backward solve
  const int ms = d.m[s], np = d.sfirst[s + 1] - d.sfirst[s];
  const T* Fs = F + d.fo[s];
  const int* ix = d.idx + d.iptr[s];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = blockDim.x >> 6;
  for (int i = t; i < ms; i += blockDim.x) Xl[i] = x[ix[i]];
  __syncthreads();
  for (int rr = 4 * wave; rr < np; rr += 4 * nw) {
    const int nr = min(4, np - rr);
    T acc[4];
    mf_rows_dot4(Fs, ms, rr, nr, Xl, np, ms, lane, acc);
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k < nr) Xl[rr + k] -= acc[k];
  }
  __syncthreads();
  for (int b1 = np; b1 > 0; b1 -= 64) {
    const int b0 = max(0, b1 - 64), bw = b1 - b0;
    if (b1 < np) {
      for (int rr = 4 * wave; rr < bw; rr += 4 * nw) {
        const int nr = min(4, bw - rr);
        T acc[4];
        mf_rows_dot4(Fs, ms, b0 + rr, nr, Xl, b1, np, lane, acc);
        if (lane == 0)
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (k < nr) Xl[b0 + rr + k] -= acc[k];
      }
    }
    {
      T v[64 * 64 / kMfSolveBlock];
#pragma unroll
      for (int it = 0; it < 64 * 64 / kMfSolveBlock; ++it) {
        const int x = t + it * kMfSolveBlock, i = x >> 6, j = x & 63;
        v[it] = (i < bw && j >= i && j < bw) ? Fs[(long long)(b0 + i) * ms + b0 + j] : (T)0;
      }
#pragma unroll
      for (int it = 0; it < 64 * 64 / kMfSolveBlock; ++it) {
        const int x = t + it * kMfSolveBlock;
        Bl[(x >> 6) * 65 + (x & 63)] = v[it];
      }
    }
    __syncthreads();
    if (wave == 0) {
      T v = lane < bw ? Xl[b0 + lane] : (T)0;
      const T rd = lane < bw ? (T)1 / Bl[lane * 65 + lane] : (T)0;
      for (int k = bw - 1; k >= 0; --k) {
        const T xk = mf_bcast(v, k) * mf_bcast(rd, k);
        if (lane == k) v = xk;
        else if (lane < k) v -= Bl[lane * 65 + k] * xk;
      }
      if (lane < bw) Xl[b0 + lane] = v;
    }
    __syncthreads();
  }
  for (int i = t; i < np; i += blockDim.x) x[ix[i]] = Xl[i];
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

// The refactor and the solve are fixed launch sequences for a given plan and
// set of buffers: each is captured once into a hipGraph and replayed.
struct MfGraph {
  const void* key[8] = {};
  hipGraphExec_t exec = nullptr;
};

struct MfLu::Impl {
  MfPlan P;
  hipStream_t st = nullptr;
  std::vector<MfGraph> graphs;
  ~Impl() {
    for (auto& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
  }
  DevVec<int> sfirst, m, iptr, idx, cptr, clist, eptr, emap, vo, prow, pcol, bad;
  DevVec<long long> fo, amap;
  DevVec<double> F64, V64, y64, b64;
  DevVec<float> F32, V32, y32;
  DevVec<int> tiny_list, big_list;
  DevVec<int4> aent;              // assembly entries of the larger fronts {front, first row, alo offset, 0}
  DevVec<int> alo;                // per entry and child: first child row at or after the entry's first row
  DevVec<int4> pent, uent;        // panel and update launch entries
  DevVec<double> dbuf;            // diagonal-block factors, one 32x32 slot per front of a panel launch
  std::vector<int> tiny_off, tiny_maxm, big_off;
  int big_maxm = 0;
  std::vector<int> aoff;          // per level: offset into aent
  // blocked launches: per level, per panel index j
  struct PanelLaunch { int p0, poff, pcnt, uoff, ucnt; };
  std::vector<std::vector<PanelLaunch>> pl_of_level;
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
  for (int t = 0; t < P.ns; ++t) {
    if (P.m[t] < P.sfirst[t + 1] - P.sfirst[t] || P.sfirst[t + 1] <= P.sfirst[t]) return XG_E_SETUP;
    for (int e = P.eptr[t]; e < P.eptr[t + 1]; ++e) {
      if (P.sparent[t] < 0 || P.emap[e] < 0 || P.emap[e] >= P.m[P.sparent[t]]) return XG_E_SETUP;
      if (e > P.eptr[t] && P.emap[e] <= P.emap[e - 1]) return XG_E_SETUP;   // assembly bisects the map
    }
  }
  hipStream_t st = I.st;
  MFCHK(I.sfirst.up(P.sfirst, st)); MFCHK(I.m.up(P.m, st)); MFCHK(I.iptr.up(P.iptr, st)); MFCHK(I.idx.up(P.idx, st));
  MFCHK(I.cptr.up(P.cptr, st)); MFCHK(I.clist.up(P.clist, st)); MFCHK(I.eptr.up(P.eptr, st));
  MFCHK(I.emap.up(P.emap, st)); MFCHK(I.vo.up(P.vo, st)); MFCHK(I.fo.up(P.fo, st)); MFCHK(I.amap.up(P.amap, st));
  MFCHK(I.prow.up(s.perm_row, st)); MFCHK(I.pcol.up(s.perm_col, st));
  MFCHK(I.bad.alloc(2));
  MFCHK(I.F64.alloc(P.fo[P.ns])); MFCHK(I.V64.alloc(P.vo[P.ns])); MFCHK(I.y64.alloc(P.n)); MFCHK(I.b64.alloc(P.n));
  if (want_fp32) { MFCHK(I.F32.alloc(P.fo[P.ns])); MFCHK(I.V32.alloc(P.vo[P.ns])); MFCHK(I.y32.alloc(P.n)); }
  // per level: tiny fronts (one launch), larger fronts (blocked panels, all
  // fronts of the level in each panel launch)
  const int nl = (int)P.lvl_ptr.size() - 1;
  std::vector<int> tl, bl;
  std::vector<int4> pe, ue;
  int max_slots = 1;
  I.tiny_off.assign(nl + 1, 0);
  I.big_off.assign(nl + 1, 0);
  I.tiny_maxm.assign(nl, 0);
  I.pl_of_level.assign(nl, {});
  I.big_maxm = 0;
  for (int l = 0; l < nl; ++l) {
    I.tiny_off[l] = (int)tl.size();
    I.big_off[l] = (int)bl.size();
    int maxp = 0;
    for (int q = P.lvl_ptr[l]; q < P.lvl_ptr[l + 1]; ++q) {
      const int t = P.lvl_s[q];
      if (P.m[t] <= kMfTiny) {
        tl.push_back(t);
        I.tiny_maxm[l] = std::max(I.tiny_maxm[l], P.m[t]);
      } else {
        bl.push_back(t);
        I.big_maxm = std::max(I.big_maxm, P.m[t]);
        maxp = std::max(maxp, (P.sfirst[t + 1] - P.sfirst[t] + kMfPanel - 1) / kMfPanel);
      }
    }
    for (int j = 0; j < maxp; ++j) {
      Impl::PanelLaunch L{j * kMfPanel, (int)pe.size(), 0, (int)ue.size(), 0};
      int slot = 0;
      for (int b = I.big_off[l]; b < (int)bl.size(); ++b) {
        const int t = bl[b];
        const int np = P.sfirst[t + 1] - P.sfirst[t];
        if (np <= L.p0) continue;
        const int pw = std::min(kMfPanel, np - L.p0);
        const int rest = P.m[t] - L.p0 - pw;
        const int nrb = (rest + 63) / 64;
        if (rest > 0) {
          for (int blk = 0; blk < 2 * nrb; ++blk) pe.push_back(make_int4(t, blk, slot, 0));
          for (int ti = 0; ti < nrb; ++ti)
            for (int tj = 0; tj < nrb; ++tj) ue.push_back(make_int4(t, ti, tj, slot));
        } else {
          pe.push_back(make_int4(t, 0, slot, 0));
          ue.push_back(make_int4(t, -1, -1, slot));
        }
        ++slot;
      }
      L.pcnt = (int)pe.size() - L.poff;
      L.ucnt = (int)ue.size() - L.uoff;
      max_slots = std::max(max_slots, slot);
      I.pl_of_level[l].push_back(L);
    }
  }
  I.tiny_off[nl] = (int)tl.size();
  I.big_off[nl] = (int)bl.size();
  MFCHK(I.tiny_list.up(tl, st));
  MFCHK(I.big_list.up(bl, st));
  MFCHK(I.pent.up(pe, st));
  MFCHK(I.uent.up(ue, st));
  MFCHK(I.dbuf.alloc((size_t)max_slots * kMfPanel * kMfPanel));
  // assembly of the larger fronts: 32 parent rows per workgroup (the tiny
  // fronts assemble their children inside their factor kernel)
  std::vector<int4> ae;
  std::vector<int> lo;
  I.aoff.assign(nl + 1, 0);
  for (int l = 0; l < nl; ++l) {
    I.aoff[l] = (int)ae.size();
    for (int b = I.big_off[l]; b < I.big_off[l + 1]; ++b) {
      const int t = bl[b];
      if (P.cptr[t + 1] == P.cptr[t]) continue;
      for (int r0 = 0; r0 < P.m[t]; r0 += kMfAsmRows) {
        ae.push_back(make_int4(t, r0, (int)lo.size(), 0));
        for (int c = P.cptr[t]; c < P.cptr[t + 1]; ++c) {
          const int ch = P.clist[c];
          const int* eb = P.emap.data() + P.eptr[ch];
          const int* ee = P.emap.data() + P.eptr[ch + 1];
          lo.push_back((int)(std::lower_bound(eb, ee, r0) - eb));
        }
      }
    }
  }
  I.aoff[nl] = (int)ae.size();
  MFCHK(I.aent.up(ae, st));
  MFCHK(I.alo.up(lo, st));
  MFCHK(hipStreamSynchronize(st));
  return XG_OK;
}

double MfLu::flops() const { return impl_->P.flops; }
long long MfLu::front_elements() const { return impl_->P.fo.empty() ? 0 : impl_->P.fo[impl_->P.ns]; }
int MfLu::levels() const { return (int)impl_->P.lvl_ptr.size() - 1; }
int MfLu::supernodes() const { return impl_->P.ns; }

template <class T>
__global__ void k_mf_flag_init(int* bad) {
  bad[0] = 0x7fffffff;
  bad[1] = 0;
}

// the launches of one refactor (no host work: the sequence is captured)
template <class T>
static int mf_factor_enqueue(MfLu::Impl& I, const double* vals_dev, int nnz, T* F) {
  const MfPlan& P = I.P;
  hipStream_t st = I.st;
  const long long tot = P.fo[P.ns];
  hipLaunchKernelGGL(k_mf_fill<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, F, tot);
  hipLaunchKernelGGL(k_mf_scatter<T>, dim3((nnz + 255) / 256), dim3(256), 0, st, nnz, vals_dev, I.amap.p, F);
  hipLaunchKernelGGL(k_mf_flag_init<T>, dim3(1), dim3(1), 0, st, I.bad.p);
  const MfDev d = I.dev();
  T* dbuf = reinterpret_cast<T*>(I.dbuf.p);
  const int nl = (int)P.lvl_ptr.size() - 1;
  for (int l = 0; l < nl; ++l) {
    const int na = I.aoff[l + 1] - I.aoff[l];
    if (na) hipLaunchKernelGGL(k_mf_assemble<T>, dim3(na), dim3(256), 0, st, d, I.aent.p + I.aoff[l], I.alo.p, F);
    const int nt = I.tiny_off[l + 1] - I.tiny_off[l];
    if (nt) {
      const int mm = I.tiny_maxm[l];
      hipLaunchKernelGGL(k_mf_factor_tiny<T>, dim3(nt), dim3(64), (size_t)mm * (mm + 1) * sizeof(T), st, d,
                         I.tiny_list.p + I.tiny_off[l], F, I.bad.p);
    }
    for (const auto& L : I.pl_of_level[l]) {
      hipLaunchKernelGGL(k_mf_panel<T>, dim3(L.pcnt), dim3(64), 0, st, d, I.pent.p + L.poff, L.p0, F, dbuf,
                         I.bad.p);
      hipLaunchKernelGGL(k_mf_update<T>, dim3(L.ucnt), dim3(256), 0, st, d, I.uent.p + L.uoff, L.p0, F,
                         (const T*)dbuf);
    }
  }
  MFCHK(hipGetLastError());
  return XG_OK;
}

// Replays the graph captured for `key`, capturing it on first use.
template <class Enqueue>
static int mf_replay(MfLu::Impl& I, const void* const* key, int nkey, Enqueue enqueue) {
  MfGraph* g = nullptr;
  for (auto& c : I.graphs)
    if (std::equal(key, key + nkey, c.key)) { g = &c; break; }
  if (!g) {
    hipGraph_t graph = nullptr;
    MFCHK(hipStreamBeginCapture(I.st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipError_t e = hipStreamEndCapture(I.st, &graph);
    if (rc != XG_OK) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    MFCHK(e);
    MfGraph ng;
    std::copy(key, key + nkey, ng.key);
    e = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MFCHK(e);
    I.graphs.push_back(ng);
    g = &I.graphs.back();
  }
  MFCHK(hipGraphLaunch(g->exec, I.st));
  return XG_OK;
}

template <class T>
static int mf_factor_t(MfLu::Impl& I, const double* vals_dev, int nnz, T* F, int* bad_out, int* hard_out) {
  const void* key[8] = {(const void*)1, vals_dev, F, (const void*)(long long)nnz, nullptr, nullptr, nullptr, nullptr};
  int rc = mf_replay(I, key, 8, [&] { return mf_factor_enqueue<T>(I, vals_dev, nnz, F); });
  if (rc != XG_OK) return rc;
  int res[2];
  MFCHK(hipMemcpyAsync(res, I.bad.p, sizeof(res), hipMemcpyDeviceToHost, I.st));
  MFCHK(hipStreamSynchronize(I.st));
  *bad_out = res[0] == 0x7fffffff ? 0 : res[0];
  *hard_out = res[1];
  return XG_OK;
}

template <class T>
static int mf_solve_enqueue(MfLu::Impl& I, const T* F, T* V, T* y, double* bp, const double* b_dev, double* x_dev,
                            int* nan_flag) {
  const MfPlan& P = I.P;
  hipStream_t st = I.st;
  const MfDev d = I.dev();
  const int n = P.n;
  // the larger fronts keep their vector in LDS when it fits in 64 KB with the
  // 64x65 block, else in their slice of V
  const size_t blk_bytes = 64 * 65 * sizeof(T);
  const bool ldsv = blk_bytes + (size_t)I.big_maxm * sizeof(T) <= 65536;
  const size_t big_lds = blk_bytes + (ldsv ? (size_t)I.big_maxm * sizeof(T) : 0);
  hipLaunchKernelGGL(k_mf_perm_in<T>, dim3((n + 255) / 256), dim3(256), 0, st, n, I.prow.p, b_dev, bp);
  const int nl = (int)P.lvl_ptr.size() - 1;
  for (int l = 0; l < nl; ++l) {
    const int nt = I.tiny_off[l + 1] - I.tiny_off[l], nb = I.big_off[l + 1] - I.big_off[l];
    if (nt)
      hipLaunchKernelGGL(k_mf_fwd_tiny<T>, dim3(nt), dim3(64), 0, st, d, I.tiny_list.p + I.tiny_off[l], F, V,
                         bp, y);
    if (nb)
      hipLaunchKernelGGL(k_mf_fwd_big<T>, dim3(nb), dim3(kMfSolveBlock), big_lds, st, d, I.big_list.p + I.big_off[l],
                         F, V, bp, y, (int)ldsv);
  }
  for (int l = nl - 1; l >= 0; --l) {
    const int nt = I.tiny_off[l + 1] - I.tiny_off[l], nb = I.big_off[l + 1] - I.big_off[l];
    if (nb)
      hipLaunchKernelGGL(k_mf_bwd_big<T>, dim3(nb), dim3(kMfSolveBlock), big_lds, st, d, I.big_list.p + I.big_off[l],
                         F, V, y, (int)ldsv);
    if (nt)
      hipLaunchKernelGGL(k_mf_bwd_tiny<T>, dim3(nt), dim3(64), 0, st, d, I.tiny_list.p + I.tiny_off[l], F, y);
  }
  hipLaunchKernelGGL(k_mf_perm_out<T>, dim3((n + 255) / 256), dim3(256), 0, st, n, I.pcol.p, y, x_dev, nan_flag);
  MFCHK(hipGetLastError());
  return XG_OK;
}

template <class T>
static int mf_solve_t(MfLu::Impl& I, const T* F, T* V, T* y, double* bp, const double* b_dev, double* x_dev,
                      int* nan_flag) {
  const void* key[8] = {(const void*)2, F, V, y, bp, b_dev, x_dev, nan_flag};
  return mf_replay(I, key, 8, [&] { return mf_solve_enqueue<T>(I, F, V, y, bp, b_dev, x_dev, nan_flag); });
}

int MfLu::factor(const double* vals_dev, int nnz, bool fp32, int* bad, int* hard) {
  Impl& I = *impl_;
  if (fp32) {
    if (!I.F32.p) return XG_E_ARG;
    return mf_factor_t<float>(I, vals_dev, nnz, I.F32.p, bad, hard);
  }
  return mf_factor_t<double>(I, vals_dev, nnz, I.F64.p, bad, hard);
}

int MfLu::solve(const double* b_dev, double* x_dev, bool fp32, int* nan_flag) {
  Impl& I = *impl_;
  if (fp32) {
    if (!I.F32.p) return XG_E_ARG;
    return mf_solve_t<float>(I, I.F32.p, I.V32.p, I.y32.p, I.b64.p, b_dev, x_dev, nan_flag);
  }
  return mf_solve_t<double>(I, I.F64.p, I.V64.p, I.y64.p, I.b64.p, b_dev, x_dev, nan_flag);
}

static size_t round32(size_t v) { return (v + 31) & ~(size_t)31; }

size_t MfLu::solve_work_doubles() const {
  const MfPlan& P = impl_->P;
  return round32((size_t)P.vo[P.ns]) + 2 * round32((size_t)P.n);
}

int MfLu::solve_ext(const double* b_dev, double* x_dev, int* nan_flag, double* work) {
  Impl& I = *impl_;
  const MfPlan& P = I.P;
  double* V = work;
  double* y = V + round32((size_t)P.vo[P.ns]);
  double* bp = y + round32((size_t)P.n);
  return mf_solve_t<double>(I, I.F64.p, V, y, bp, b_dev, x_dev, nan_flag);
}

}  // namespace xg
