// Internal layout shared by the host setup code and the HIP kernels.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/xyce_amd.h"

namespace xg {

// ---- per-class MOSFET level-1 constants (everything but VTO/KP-derived terms)
enum MosClassField {
  K_DTYPE = 0, K_GAMMA, K_LAMBDA, K_TPHI, K_VT, K_DSAT, K_SSAT, K_DVCRIT, K_SVCRIT,
  K_CBD, K_CBDSW, K_CBS, K_CBSSW, K_TDEPCAP, K_TBULKPOT, K_MJ, K_MJSW,
  K_F2D, K_F3D, K_F4D, K_F2S, K_F3S, K_F4S, K_OXCAP, K_GSO, K_GDO, K_GBO,
  K_GD, K_GS, K_M, K_GMIN, K_COUNT
};
constexpr int KSTRIDE = 32;   // doubles per class record

// ---- per-device contribution record planes (SoA: plane * n_dev + dev)
// Each plane holds one term exactly as the reference adds it (sign included),
// so a gather that sums the terms in device order reproduces the reference's
// sequential "+=" into Epetra vectors / raw matrix pointers.
enum RecPlane {
  R_F_D = 0, R_F_G, R_F_S, R_F_B, R_F_DP, R_F_SP,     // F vector (6)
  R_Q_G, R_Q_B, R_Q_DP, R_Q_SP,                       // Q vector (4)
  R_FL_B, R_FL_DP, R_FL_SP,                           // dFdxdVp (3)
  R_QL_G, R_QL_B, R_QL_DP, R_QL_SP,                   // dQdxdVp (4)
  R_JF0,                                              // dFdx pointer terms 0..16
  R_JQ0 = R_JF0 + 17,                                 // dQdx pointer terms 0..13
  R_COUNT = R_JQ0 + 14                                // 48
};
// Vector id of a vector plane: 0 F, 1 Q, 2 dFdxdVp, 3 dQdxdVp.
inline int rec_vec_id(int plane) {
  return plane < R_Q_G ? 0 : plane < R_FL_B ? 1 : plane < R_QL_G ? 2 : 3;
}

// State / store slot order (MOSFET1 registerStateLIDs / registerStoreLIDs,
// N_DEV_MOSFET1.C:1556-1609).
enum StaSlot { S_QGS = 0, S_QGD, S_QGB, S_CAPGS, S_CAPGD, S_CAPGB, S_QBD, S_QBS, S_COUNT };
enum StoSlot { O_VBD = 0, O_VBS, O_VGS, O_VDS, O_VON, O_GM, O_COUNT };

// ---- source list entry (B vector)
struct SrcEntry { int wave; int kind; int li_p; int li_n; };  // kind 0 = V (B[bra] += v), 1 = I
// li_p for V sources is the branch row.

// ---- host-side circuit after setup
struct HostCircuit {
  int n = 0, nnz = 0, n_mos = 0, n_nodes = 0, nsta = 0, nsto = 0;
  double gmin = 1e-12;
  // CSR pattern
  std::vector<int> rowptr, colind;
  std::vector<int> node_gid;
  std::vector<char> var_type;          // 'V' / 'I' per unknown
  // MOSFETs
  std::vector<double> cls;             // n_cls * KSTRIDE
  std::vector<int> mos_cls;
  std::vector<double> mos_tvbi, mos_beta, mos_tvto;
  std::vector<int> mos_lid;            // 6 per device: d g s b dp sp  (n = ground)
  std::vector<int> mos_fslot, mos_qslot;
  std::vector<int> mos_sta0, mos_sto0; // Xyce-layout offsets
  std::vector<uint64_t> mos_mask;      // used record planes
  // linear part (CSR-aligned) and sources
  std::vector<double> linF, linQ;
  std::vector<SrcEntry> srcs;
  std::vector<int> wave_type, wave_off;
  std::vector<double> wave_data;
  std::vector<double> breakpoints;     // filled per tran run
  // gather maps: vectors per row, matrices per slot; entry = dev * 64 + plane
  std::vector<int> vg_ptr, vg_ent;     // n + 1
  std::vector<int> mg_ptr, mg_ent;     // nnz + 1
  std::string err;
};

int build_host_circuit(const xg_circuit_desc* d, HostCircuit& hc);
void wave_breakpoints(const HostCircuit& hc, int w, double tstep, double tstop, std::vector<double>& out);

// ---- sparse LU symbolic analysis (host)
struct LuSymbolic {
  int n = 0, nnz_a = 0, nlu = 0;
  std::vector<int> perm_row;     // B row i  <- A row perm_row[i]
  std::vector<int> perm_col;     // B col j  <- A col perm_col[j]
  std::vector<int> a2lu;         // A CSR slot -> LU position
  std::vector<int> lu_rowptr, lu_colind, lu_diag;   // row-major L+U pattern of B
  // factor schedule (entry level)
  std::vector<int> f_lvl_ptr;    // levels + 1 -> index into f_tgt
  std::vector<int> f_tgt;        // LU position of target
  std::vector<int> f_op_ptr;     // per target (same order as f_tgt) + 1
  std::vector<int> f_ops;        // pairs (pos_l, pos_u)
  std::vector<int> f_div;        // per target: diag position to divide by, or -1
  // solve schedules
  std::vector<int> fs_lvl_ptr, fs_rows;   // forward
  std::vector<int> bs_lvl_ptr, bs_rows;   // backward
  double flops = 0;
  int dense_tail = 0;
  // elimination tree and column structure of L (rows below the diagonal,
  // ascending) of B, kept when the multifrontal plan is wanted
  std::vector<int> etree;
  std::vector<std::vector<int>> lcol;
  std::vector<int> rinv, pinv;   // A row -> B row, A col -> B col
};

// want_schedule: build the entry-level factor / solve schedules;
// keep_tree: keep the elimination tree and the L column structure
// A square CSR pattern the symbolic analysis can read: rowptr[0] = 0,
// non-decreasing, every column index in [0, n).
inline bool valid_csr(int n, const int* rowptr, const int* colind) {
  if (n <= 0 || !rowptr || !colind || rowptr[0] != 0) return false;
  for (int r = 0; r < n; ++r)
    if (rowptr[r + 1] < rowptr[r]) return false;
  for (int k = 0; k < rowptr[n]; ++k)
    if (colind[k] < 0 || colind[k] >= n) return false;
  return true;
}

int lu_analyse(int n, const int* rowptr, const int* colind, const double* values, LuSymbolic& s,
               bool want_schedule = true, bool keep_tree = false);

// Multifrontal plan of a symbolic analysis (multifrontal.hip)
struct MfPlan {
  int n = 0, ns = 0;
  std::vector<int> sfirst;         // ns + 1: columns [sfirst[s], sfirst[s+1]) of B
  std::vector<int> sparent;        // parent supernode, -1 at a root
  std::vector<int> m;              // front order (pivots + rows below)
  std::vector<long long> fo;       // front element offsets (row-major fronts), fo[ns] = total
  std::vector<int> iptr, idx;      // front index lists (B indices), pivots first
  std::vector<long long> amap;     // A slot -> front element
  std::vector<int> cptr, clist;    // children of each supernode
  std::vector<int> eptr, emap;     // per supernode: parent-local index of each row below its pivots
  std::vector<int> lvl_ptr, lvl_s; // supernodes by level, leaves first
  std::vector<int> vo;             // per-front vector offsets, vo[ns] = total
  double flops = 0;
  long long max_m = 0;
};
int mf_plan(const LuSymbolic& s, const int* rowptr, const int* colind, MfPlan& p);

// Multifrontal factorisation on the GPU (multifrontal.hip); the stream is a
// hipStream_t, vectors are device pointers in A's numbering.
class MfLu {
 public:
  struct Impl;
  MfLu();
  ~MfLu();
  MfLu(const MfLu&) = delete;
  MfLu& operator=(const MfLu&) = delete;
  int setup(const LuSymbolic& s, const int* rowptr, const int* colind, void* stream, bool want_fp32);
  // *bad = 0, or (first failing pivot column of B) + 1; *hard = 1 if a pivot
  // was zero or not finite (else every failure is a multiplier beyond KLU's
  // threshold)
  int factor(const double* vals_dev, int nnz, bool fp32, int* bad, int* hard);
  int solve(const double* b_dev, double* x_dev, bool fp32, int* nan_flag);
  // FP64 solve whose front vectors, work vector and permuted right-hand side
  // are the caller's device memory `work` (solve_work_doubles() doubles,
  // 256-byte aligned) instead of the plan's buffers
  size_t solve_work_doubles() const;
  int solve_ext(const double* b_dev, double* x_dev, int* nan_flag, double* work);
  double flops() const;
  long long front_elements() const;
  int levels() const;
  int supernodes() const;

 private:
  Impl* impl_;
};

// ---- a drop-in circuit's device state, for the single-circuit transient
// engine (tran.hip); implemented in kernels.hip.  Every call is enqueued on
// the circuit's stream and returns without waiting.
struct CircuitDev {
  int n = 0, nnz = 0, n_mos = 0;
  void* stream = nullptr;                 // hipStream_t
  double* x = nullptr;                    // n + 1 (entry n: the ground dummy, 0)
  const double *F = nullptr, *Q = nullptr, *FL = nullptr, *QL = nullptr;
  double* B = nullptr;                    // source stepping scales it in place
  const double *dFdx = nullptr, *dQdx = nullptr;
  const int* limited = nullptr;           // set by the last load when a junction limiter acted
};
CircuitDev circuit_dev(xg_circuit* c);
const HostCircuit& circuit_host(const xg_circuit* c);
// device evaluation at the circuit's x under its solver state, then the
// vector and matrix gathers (F, Q, B, limiter vectors, dFdx, dQdx)
int circuit_load(xg_circuit* c);
// residual and Jacobian (OneStep / NoTimeIntegration) from the last load into
// device arrays rhs (n) and J (nnz); q0, q2 device arrays (unused at DC)
int circuit_combine(xg_circuit* c, const double* q0, const double* q2, double* rhs, double* J);
// MOSFET state / store history: 0 zero both, 1 current <- next (a step or
// operating point accepted), 2 next <- current (a step rejected)
int circuit_history(xg_circuit* c, int op);

// learned-predictor network (mlp.hip): forward on a caller stream (a
// hipStream_t), and its dimensions
int mlp_forward_on_stream(xg_mlp* net, int rows, const unsigned short* x, float* y, void* stream);
int mlp_dims(const xg_mlp* net, int* in_dim, int* out_dim, int* max_rows);

}  // namespace xg
