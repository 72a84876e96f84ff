// Host-only introspection entry points: the setup and symbolic analysis that
// xg_circuit_create / xg_lu_symbolic run before touching the GPU, exposed so
// the pattern, stamp maps and LU orderings can be checked without a device.
#include <algorithm>

#include "xg_internal.h"

extern "C" {

int xg_host_pattern(const xg_circuit_desc* d, int* sizes, int* rowptr, int* colind, int* node_gid, int* mos_lid,
                    int* mos_fslot, int* mos_qslot, int* mos_stasto) {
  if (!d || !sizes) return XG_E_ARG;
  xg::HostCircuit hc;
  const int st = xg::build_host_circuit(d, hc);
  if (st != XG_OK) return st;
  sizes[0] = hc.n; sizes[1] = hc.nnz; sizes[2] = hc.nsta; sizes[3] = hc.nsto; sizes[4] = hc.n_mos; sizes[5] = hc.n_nodes;
  if (rowptr) std::copy(hc.rowptr.begin(), hc.rowptr.end(), rowptr);
  if (colind) std::copy(hc.colind.begin(), hc.colind.end(), colind);
  if (node_gid) std::copy(hc.node_gid.begin(), hc.node_gid.end(), node_gid);
  if (mos_lid) std::copy(hc.mos_lid.begin(), hc.mos_lid.end(), mos_lid);
  if (mos_fslot) std::copy(hc.mos_fslot.begin(), hc.mos_fslot.end(), mos_fslot);
  if (mos_qslot) std::copy(hc.mos_qslot.begin(), hc.mos_qslot.end(), mos_qslot);
  if (mos_stasto)
    for (int i = 0; i < hc.n_mos; ++i) { mos_stasto[2 * i] = hc.mos_sta0[i]; mos_stasto[2 * i + 1] = hc.mos_sto0[i]; }
  return XG_OK;
}

int xg_host_lu_analyse(int n, const int* rowptr, const int* colind, const double* values, double* info, int* perm_row,
                       int* perm_col, int* lu_rowptr, int* lu_colind, int max_lu) {
  if (n <= 0 || !rowptr || !colind || !info || !xg::valid_csr(n, rowptr, colind)) return XG_E_ARG;
  xg::LuSymbolic s;
  try {
    const int st = xg::lu_analyse(n, rowptr, colind, values, s, true);
    if (st != XG_OK) return st;
  } catch (...) {
    return XG_E_SETUP;
  }
  info[0] = s.nlu;
  info[1] = (double)s.f_lvl_ptr.size() - 1;
  info[2] = (double)(s.fs_lvl_ptr.size() - 1 + s.bs_lvl_ptr.size() - 1);
  info[3] = s.flops;
  if (perm_row) std::copy(s.perm_row.begin(), s.perm_row.end(), perm_row);
  if (perm_col) std::copy(s.perm_col.begin(), s.perm_col.end(), perm_col);
  if (lu_rowptr) std::copy(s.lu_rowptr.begin(), s.lu_rowptr.end(), lu_rowptr);
  if (lu_colind && s.nlu <= max_lu) std::copy(s.lu_colind.begin(), s.lu_colind.end(), lu_colind);
  return XG_OK;
}

int xg_host_mf_plan(int n, const int* rowptr, const int* colind, const double* values, double* info) {
  if (n <= 0 || !rowptr || !colind || !info || !xg::valid_csr(n, rowptr, colind)) return XG_E_ARG;
  xg::LuSymbolic s;
  xg::MfPlan p;
  try {
    int st = xg::lu_analyse(n, rowptr, colind, values, s, false, true);
    if (st != XG_OK) return st;
    st = xg::mf_plan(s, rowptr, colind, p);
    if (st != XG_OK) return st;
  } catch (...) {
    return XG_E_SETUP;
  }
  info[0] = p.ns;
  info[1] = (double)p.lvl_ptr.size() - 1;
  info[2] = (double)p.fo[p.ns];
  info[3] = p.flops;
  info[4] = (double)p.max_m;
  info[5] = s.nlu;
  return XG_OK;
}

}  // extern "C"
