// Host-side circuit setup: MOSFET level-1 model/instance constants, node
// ordering, the fixed CSR pattern and the stamp gather maps.
//
// Reference (paths relative to the Xyce source root):
//   Model::processParams       src/DeviceModelPKG/OpenModels/N_DEV_MOSFET1.C:3939
//   Instance::processParams    N_DEV_MOSFET1.C:383
//   Instance::updateTemperature N_DEV_MOSFET1.C:3487
//   Topology::addDevice        src/TopoManagerPKG/N_TOP_Topology.C:191
//   Graph::insertNode / generateBFT_   src/UtilityPKG/N_UTL_Graph.h:237,832
//   CktGraphBasic::getBFSNodeList      src/TopoManagerPKG/N_TOP_CktGraphBasic.C:181
//   SerialLSUtil::setupNodeGIDs / setupSolnAndStateGIDs / generateRowColData
//                                      src/TopoManagerPKG/N_TOP_SerialLSUtil.C:468,504,620
//   MOSFET1 jacStamp (RD/RS merge)     N_DEV_MOSFET1.C:1293-1337, registerJacLIDs :1663
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <stdexcept>
#include <unordered_map>

#include "xg_internal.h"

namespace xg {

namespace {

constexpr double kQ = 1.6021918e-19;
constexpr double kBoltz = 1.3806226e-23;
constexpr double kRefTemp = 300.15;
constexpr double kKoverQ = kBoltz / kQ;

enum { P_TYPE = 0, P_VTO, P_KP, P_GAMMA, P_PHI, P_LAMBDA, P_RD, P_RS, P_CBD, P_CBS, P_IS, P_PB, P_CGSO,
       P_CGDO, P_CGBO, P_RSH, P_CJ, P_MJ, P_CJSW, P_MJSW, P_JS, P_TOX, P_LD, P_UO, P_FC, P_NSUB, P_NSS,
       P_TPG, P_L, P_W };
enum { I_L = 0, I_W, I_AD, I_AS, I_PD, I_PS, I_NRD, I_NRS, I_M, I_VTO, I_KP };

struct ModelConsts {
  int dtype;
  double tnom, fact1, vtnom, egfet1, pbfact1, oxideCapFactor, vt0, kp, gamma, phi, lambda, uo;
  double rd, rs, cbd, cbs, is, pb, cgso, cgdo, cgbo, rsh, cj, mj, cjsw, mjsw, js, ld, fc, L, W;
  bool cbdGiven, cbsGiven, cjGiven, cjswGiven, rshGiven;
};

inline bool given(const double* p, int k) { return !std::isnan(p[k]); }
inline double val(const double* p, int k, double d) { return given(p, k) ? p[k] : d; }

// MOSFET1 Model::processParams (N_DEV_MOSFET1.C:3939-4019) with the .model
// defaults of the parameter table (:150-360).
ModelConsts model_consts(const double* p) {
  ModelConsts m{};
  m.dtype = p[P_TYPE] < 0 ? -1 : 1;
  m.vt0 = val(p, P_VTO, 0.0); m.kp = val(p, P_KP, 2e-5); m.gamma = val(p, P_GAMMA, 0.0);
  m.phi = val(p, P_PHI, 0.6); m.lambda = val(p, P_LAMBDA, 0.0); m.rd = val(p, P_RD, 0.0);
  m.rs = val(p, P_RS, 0.0); m.cbd = val(p, P_CBD, 0.0); m.cbs = val(p, P_CBS, 0.0);
  m.is = val(p, P_IS, 1e-14); m.pb = val(p, P_PB, 0.8); m.cgso = val(p, P_CGSO, 0.0);
  m.cgdo = val(p, P_CGDO, 0.0); m.cgbo = val(p, P_CGBO, 0.0); m.rsh = val(p, P_RSH, 0.0);
  m.cj = val(p, P_CJ, 0.0); m.mj = val(p, P_MJ, 0.5); m.cjsw = val(p, P_CJSW, 0.0);
  m.mjsw = val(p, P_MJSW, 0.5); m.js = val(p, P_JS, 0.0); m.ld = val(p, P_LD, 0.0);
  m.uo = val(p, P_UO, 600.0); m.fc = val(p, P_FC, 0.5); m.L = val(p, P_L, 1e-4); m.W = val(p, P_W, 1e-4);
  m.cbdGiven = given(p, P_CBD) || m.cbd != 0;
  m.cbsGiven = given(p, P_CBS) || m.cbs != 0;
  m.cjGiven = given(p, P_CJ);
  m.cjswGiven = given(p, P_CJSW);
  m.rshGiven = given(p, P_RSH);
  m.tnom = kRefTemp;
  m.fact1 = m.tnom / kRefTemp;
  m.vtnom = m.tnom * kKoverQ;
  const double kt1 = kBoltz * m.tnom;
  m.egfet1 = 1.16 - (7.02e-4 * m.tnom * m.tnom) / (m.tnom + 1108);
  const double arg1 = -m.egfet1 / (kt1 + kt1) + 1.1150877 / (kBoltz * (kRefTemp + kRefTemp));
  m.pbfact1 = -2 * m.vtnom * (1.5 * std::log(m.fact1) + kQ * arg1);
  const double tox = val(p, P_TOX, 1e-7);
  if (!given(p, P_TOX) || tox == 0) {
    m.oxideCapFactor = 0;
  } else {
    m.oxideCapFactor = 3.9 * 8.854214871e-12 / tox;
    if (!given(p, P_KP)) m.kp = m.uo * m.oxideCapFactor * 1e-4;
    const double nsub = val(p, P_NSUB, 0.0);
    if (given(p, P_NSUB) && nsub * 1e6 > 1.45e16) {
      if (!given(p, P_PHI)) m.phi = std::max(0.1, 2 * m.vtnom * std::log(nsub * 1e6 / 1.45e16));
      const double fermis = m.dtype * .5 * m.phi;
      double wkfng = 3.2;
      const int gate_type = given(p, P_TPG) ? (int)p[P_TPG] : 1;
      if (gate_type != 0) wkfng = 3.25 + .5 * m.egfet1 - m.dtype * gate_type * .5 * m.egfet1;
      const double wkfngs = wkfng - (3.25 + .5 * m.egfet1 + fermis);
      if (!given(p, P_GAMMA)) m.gamma = std::sqrt(2 * 11.70 * 8.854214871e-12 * kQ * nsub * 1e6) / m.oxideCapFactor;
      if (!given(p, P_VTO))
        m.vt0 = wkfngs - val(p, P_NSS, 0.0) * 1e4 * kQ / m.oxideCapFactor + m.dtype * (m.gamma * std::sqrt(m.phi) + m.phi);
    }
  }
  return m;
}

struct InstConsts {
  double cls[KSTRIDE];
  double tvbi, beta, tvto;
};

// Instance::processParams (:383-472) and updateTemperature (:3487-3721).
InstConsts inst_consts(ModelConsts m, const double* ip, double temp, double gmin) {
  if (given(ip, I_VTO)) m.vt0 = ip[I_VTO];
  if (given(ip, I_KP)) m.kp = ip[I_KP];
  const double l = val(ip, I_L, m.L), w = val(ip, I_W, m.W);
  const double ad = val(ip, I_AD, 0.0), as = val(ip, I_AS, 0.0);
  const double pd = val(ip, I_PD, 0.0), ps = val(ip, I_PS, 0.0);
  const double nrd = val(ip, I_NRD, 1.0), nrs = val(ip, I_NRS, 1.0), M = val(ip, I_M, 1.0);
  InstConsts o{};
  double* k = o.cls;
  std::fill(k, k + KSTRIDE, 0.0);
  const double ratio = temp / m.tnom;
  const double vt = temp * kKoverQ;
  const double fact2 = temp / kRefTemp;
  const double kt = temp * kBoltz;
  const double egfet = 1.16 - (7.02e-4 * temp * temp) / (temp + 1108);
  const double arg = -egfet / (kt + kt) + 1.1150877 / (kBoltz * (kRefTemp + kRefTemp));
  const double pbfact = -2 * vt * (1.5 * std::log(fact2) + kQ * arg);
  const double ratio4 = ratio * std::sqrt(ratio);
  const double tkp = m.kp / ratio4;
  const double phio = (m.phi - m.pbfact1) / m.fact1;
  const double tphi = fact2 * phio + pbfact;
  o.tvbi = m.vt0 - m.dtype * (m.gamma * std::sqrt(m.phi)) + .5 * (m.egfet1 - egfet) + m.dtype * .5 * (tphi - m.phi);
  o.tvto = o.tvbi + m.dtype * m.gamma * std::sqrt(tphi);
  const double tsat = m.is * std::exp(-egfet / vt + m.egfet1 / m.vtnom);
  const double tsatd = m.js * std::exp(-egfet / vt + m.egfet1 / m.vtnom);
  const double pbo = (m.pb - m.pbfact1) / m.fact1;
  const double gmaold = (m.pb - pbo) / pbo;
  double cf = 1 / (1 + m.mj * (4e-4 * (m.tnom - kRefTemp) - gmaold));
  double tcbd = m.cbd * cf, tcbs = m.cbs * cf, tcj = m.cj * cf;
  cf = 1 / (1 + m.mjsw * (4e-4 * (m.tnom - kRefTemp) - gmaold));
  double tcjsw = m.cjsw * cf;
  const double tbulk = fact2 * pbo + pbfact;
  const double gmanew = (tbulk - pbo) / pbo;
  cf = (1 + m.mj * (4e-4 * (temp - kRefTemp) - gmanew));
  tcbd *= cf; tcbs *= cf; tcj *= cf;
  cf = (1 + m.mjsw * (4e-4 * (temp - kRefTemp) - gmanew));
  tcjsw *= cf;
  const double tdep = m.fc * tbulk;
  const double r2 = std::sqrt(2.0);
  double dvcrit, svcrit;
  if (m.js == 0 || ad == 0 || as == 0) {
    dvcrit = svcrit = vt * std::log(vt / (r2 * m.is));
  } else {
    dvcrit = vt * std::log(vt / (r2 * m.js * ad));
    svcrit = vt * std::log(vt / (r2 * m.js * as));
  }
  const double a1 = 1 - m.fc;
  const double sarg = std::exp((-m.mj) * std::log(a1));
  const double sargsw = std::exp((-m.mjsw) * std::log(a1));
  auto dep_coeffs = [&](double cz, double czsw, double& f2, double& f3, double& f4) {
    f2 = cz * (1 - m.fc * (1 + m.mj)) * sarg / a1 + czsw * (1 - m.fc * (1 + m.mjsw)) * sargsw / a1;
    f3 = cz * m.mj * sarg / a1 / tbulk + czsw * m.mjsw * sargsw / a1 / tbulk;
    f4 = cz * tbulk * (1 - a1 * sarg) / (1 - m.mj) + czsw * tbulk * (1 - a1 * sargsw) / (1 - m.mjsw) -
         f3 / 2 * (tdep * tdep) - tdep * f2;
  };
  const double czbd = m.cbdGiven ? tcbd : (m.cjGiven ? tcj * ad : 0);
  const double czbdsw = m.cjswGiven ? tcjsw * pd : 0;
  const double czbs = m.cbsGiven ? tcbs : (m.cjGiven ? tcj * as : 0);
  const double czbssw = m.cjswGiven ? tcjsw * ps : 0;
  double f2d, f3d, f4d, f2s, f3s, f4s;
  dep_coeffs(czbd, czbdsw, f2d, f3d, f4d);
  dep_coeffs(czbs, czbssw, f2s, f3s, f4s);
  double gd = 0, gs = 0;
  if (m.rd != 0) gd = 1 / m.rd;
  else if (m.rshGiven) gd = m.rsh != 0 ? 1 / (m.rsh * nrd) : 0;
  if (m.rs != 0) gs = 1 / m.rs;
  else if (m.rshGiven) gs = m.rsh != 0 ? 1 / (m.rsh * nrs) : 0;
  if (l - 2 * m.ld <= 0) throw std::runtime_error("Effective channel length less than zero.");
  const double leff = l - 2 * m.ld;
  double dsat, ssat;
  if (tsatd == 0 || ad == 0 || as == 0) { dsat = tsat; ssat = tsat; }
  else { dsat = tsatd * ad; ssat = tsatd * as; }
  o.beta = tkp * w / leff;
  k[K_DTYPE] = m.dtype; k[K_GAMMA] = m.gamma; k[K_LAMBDA] = m.lambda; k[K_TPHI] = tphi; k[K_VT] = vt;
  k[K_DSAT] = dsat; k[K_SSAT] = ssat; k[K_DVCRIT] = dvcrit; k[K_SVCRIT] = svcrit;
  k[K_CBD] = czbd; k[K_CBDSW] = czbdsw; k[K_CBS] = czbs; k[K_CBSSW] = czbssw;
  k[K_TDEPCAP] = tdep; k[K_TBULKPOT] = tbulk; k[K_MJ] = m.mj; k[K_MJSW] = m.mjsw;
  k[K_F2D] = f2d; k[K_F3D] = f3d; k[K_F4D] = f4d; k[K_F2S] = f2s; k[K_F3S] = f3s; k[K_F4S] = f4s;
  k[K_OXCAP] = m.oxideCapFactor * leff * w;
  k[K_GSO] = m.cgso * w; k[K_GDO] = m.cgdo * w; k[K_GBO] = m.cgbo * leff;
  k[K_GD] = gd; k[K_GS] = gs; k[K_M] = M; k[K_GMIN] = gmin;
  return o;
}

}  // namespace

void wave_breakpoints(const HostCircuit& hc, int w, double tstep, double tstop, std::vector<double>& out) {
  const double* p = hc.wave_data.data() + hc.wave_off[w];
  const int np = hc.wave_off[w + 1] - hc.wave_off[w];
  if (hc.wave_type[w] == XG_WAVE_PULSE) {
    auto d = [&](int i, double x) { return (i < np && !std::isnan(p[i])) ? p[i] : x; };
    const double TD = d(2, 0.0), TR = d(3, tstep), TF = d(4, tstep), PW = d(5, tstop), PER = d(6, tstop);
    for (double base = TD; base < tstop; base += PER) {
      for (double c : {0.0, TR, TR + PW, TR + PW + TF})
        if (base + c > 0 && base + c < tstop) out.push_back(base + c);
      if (PER <= 0) break;
    }
  } else if (hc.wave_type[w] == XG_WAVE_PWL) {
    for (int i = 0; i + 1 < np; i += 2)
      if (p[i] > 0 && p[i] < tstop) out.push_back(p[i]);
  }
}

// Rejects a description whose counts, device indices, model or waveform
// indices or waveform parameter counts are inconsistent, before anything is
// indexed with them (node ids are checked where the node graph is built).
static void validate_desc(const xg_circuit_desc* d) {
  auto bad = [](const char* what) { throw std::runtime_error(std::string("invalid circuit description: ") + what); };
  if (d->n_nodes < 1 || d->n_dev < 0 || d->n_models < 0 || d->n_mos < 0 || d->n_res < 0 || d->n_cap < 0 ||
      d->n_vsrc < 0 || d->n_isrc < 0 || d->n_wave < 0)
    bad("negative count");
  auto need = [&](const void* ptr, int n, const char* what) { if (n > 0 && !ptr) bad(what); };
  need(d->dev_type, d->n_dev, "dev_type"); need(d->dev_idx, d->n_dev, "dev_idx");
  need(d->model_par, d->n_models, "model_par");
  need(d->mos_nodes, d->n_mos, "mos_nodes"); need(d->mos_model, d->n_mos, "mos_model");
  need(d->mos_inst, d->n_mos, "mos_inst");
  need(d->res_nodes, d->n_res, "res_nodes"); need(d->res_val, d->n_res, "res_val");
  need(d->cap_nodes, d->n_cap, "cap_nodes"); need(d->cap_val, d->n_cap, "cap_val");
  need(d->vsrc_nodes, d->n_vsrc, "vsrc_nodes"); need(d->vsrc_wave, d->n_vsrc, "vsrc_wave");
  need(d->isrc_nodes, d->n_isrc, "isrc_nodes"); need(d->isrc_wave, d->n_isrc, "isrc_wave");
  need(d->wave_type, d->n_wave, "wave_type"); need(d->wave_off, d->n_wave, "wave_off");
  need(d->wave_data, d->n_wave, "wave_data");
  const int cnt[5] = {d->n_mos, d->n_res, d->n_cap, d->n_vsrc, d->n_isrc};
  if (d->n_dev != cnt[0] + cnt[1] + cnt[2] + cnt[3] + cnt[4]) bad("n_dev is not the sum of the device counts");
  std::vector<char> seen[5];
  for (int t = 0; t < 5; ++t) seen[t].assign(cnt[t], 0);
  for (int o = 0; o < d->n_dev; ++o) {
    const int t = d->dev_type[o], i = d->dev_idx[o];
    if (t < 0 || t > 4) bad("device type");
    if (i < 0 || i >= cnt[t] || seen[t][i]) bad("device order must list every device once");
    seen[t][i] = 1;
  }
  for (int i = 0; i < d->n_mos; ++i)
    if (d->mos_model[i] < 0 || d->mos_model[i] >= d->n_models) bad("MOSFET model index");
  if (d->n_wave > 0 && d->wave_off[0] != 0) bad("wave_off[0] must be 0");
  for (int w = 0; w < d->n_wave; ++w) {
    const int np = d->wave_off[w + 1] - d->wave_off[w], t = d->wave_type[w];
    const bool ok = (t == XG_WAVE_DC && np >= 1) || (t == XG_WAVE_PULSE && np == 7) ||
                    (t == XG_WAVE_PWL && np >= 2 && np % 2 == 0);
    if (!ok) bad("waveform type or parameter count (DC >= 1, PULSE 7, PWL time/value pairs)");
  }
  for (int i = 0; i < d->n_vsrc; ++i)
    if (d->vsrc_wave[i] < 0 || d->vsrc_wave[i] >= d->n_wave) bad("voltage source waveform index");
  for (int i = 0; i < d->n_isrc; ++i)
    if (d->isrc_wave[i] < 0 || d->isrc_wave[i] >= d->n_wave) bad("current source waveform index");
}

int build_host_circuit(const xg_circuit_desc* d, HostCircuit& hc) {
  try {
    validate_desc(d);
    hc.n_nodes = d->n_nodes;
    hc.n_mos = d->n_mos;
    hc.gmin = d->gmin > 0 ? d->gmin : 1e-12;
    const double temp = d->temp > 0 ? d->temp : kRefTemp;

    // ---- MOSFET constants, deduplicated into classes
    std::vector<ModelConsts> mc(d->n_models);
    for (int m = 0; m < d->n_models; ++m) mc[m] = model_consts(d->model_par + (size_t)m * XG_NMP);
    std::map<std::vector<double>, int> cls_index;
    hc.mos_cls.resize(d->n_mos);
    hc.mos_tvbi.resize(d->n_mos); hc.mos_beta.resize(d->n_mos); hc.mos_tvto.resize(d->n_mos);
    for (int i = 0; i < d->n_mos; ++i) {
      InstConsts ic = inst_consts(mc[d->mos_model[i]], d->mos_inst + (size_t)i * XG_NIP, temp, hc.gmin);
      std::vector<double> key(ic.cls, ic.cls + KSTRIDE);
      auto it = cls_index.find(key);
      int ci;
      if (it == cls_index.end()) {
        ci = (int)cls_index.size();
        cls_index.emplace(key, ci);
        hc.cls.insert(hc.cls.end(), key.begin(), key.end());
      } else {
        ci = it->second;
      }
      hc.mos_cls[i] = ci;
      hc.mos_tvbi[i] = ic.tvbi; hc.mos_beta[i] = ic.beta; hc.mos_tvto[i] = ic.tvto;
    }

    // ---- graph of voltage nodes and devices, inserted in netlist order
    // (Topology::addDevice: each device's nodes first, then the device node
    // adjacent to them; Graph::insertNode back-links every adjacency).
    const int nd = d->n_dev;
    auto dev_nodes = [&](int oi, int* out) -> int {
      const int t = d->dev_type[oi], i = d->dev_idx[oi];
      if (t == XG_DEV_MOS1) { for (int k = 0; k < 4; ++k) out[k] = d->mos_nodes[4 * i + k]; return 4; }
      const int* nn = t == XG_DEV_RES ? d->res_nodes : t == XG_DEV_CAP ? d->cap_nodes
                    : t == XG_DEV_VSRC ? d->vsrc_nodes : d->isrc_nodes;
      out[0] = nn[2 * i]; out[1] = nn[2 * i + 1];
      return 2;
    };
    std::vector<std::vector<int>> adj;
    std::vector<int> gnode;                 // graph index -> (>=0 voltage node id) or -(1+device order index)
    std::vector<int> vindex(d->n_nodes, -1);
    std::unordered_map<int, int> keys1;     // same container the reference iterates for the BFS root
    adj.reserve((size_t)d->n_nodes + nd);
    for (int oi = 0; oi < nd; ++oi) {
      int nl[4];
      const int cnt = dev_nodes(oi, nl);
      for (int k = 0; k < cnt; ++k) {
        if (nl[k] < 0 || nl[k] >= d->n_nodes) throw std::runtime_error("node id out of range");
        if (vindex[nl[k]] >= 0) continue;
        vindex[nl[k]] = (int)adj.size();
        keys1[(int)adj.size()] = 0;
        adj.emplace_back();
        gnode.push_back(nl[k]);
      }
      const int me = (int)adj.size();
      adj.emplace_back();
      for (int k = 0; k < cnt; ++k) { adj[me].push_back(vindex[nl[k]]); adj[vindex[nl[k]]].push_back(me); }
      keys1[me] = 0;
      gnode.push_back(-(1 + oi));
    }
    // BFS from the first key of keys1 (Graph::generateBFT), forests continue
    // from the lowest unvisited non-isolated index.
    std::vector<int> bft;
    bft.reserve(adj.size());
    {
      std::vector<char> seen(adj.size(), 0);
      std::deque<int> q;
      const int start = keys1.begin()->first;
      q.push_back(start); seen[start] = 1; bft.push_back(start);
      size_t root = 0;
      while (!q.empty()) {
        const int cur = q.front(); q.pop_front();
        for (int a : adj[cur]) if (!seen[a]) { seen[a] = 1; q.push_back(a); bft.push_back(a); }
        if (q.empty() && bft.size() != adj.size()) {
          for (; root < adj.size(); ++root)
            if (!adj[root].empty() && !seen[root]) { seen[root] = 1; q.push_back((int)root); bft.push_back((int)root); break; }
        }
      }
    }
    // ---- solution / state / store numbering along the reversed BFS list
    hc.node_gid.assign(d->n_nodes, -1);
    std::vector<int> vs_bra(d->n_vsrc, -1), dp_lid(d->n_mos, -1), sp_lid(d->n_mos, -1);
    hc.mos_sta0.assign(d->n_mos, 0); hc.mos_sto0.assign(d->n_mos, 0);
    int row = 0, sta = 0, sto = 0;
    for (auto it = bft.rbegin(); it != bft.rend(); ++it) {
      const int g = gnode[*it];
      if (g >= 0) {
        if (g == 0) continue;
        hc.node_gid[g] = row++; hc.var_type.push_back('V');
        continue;
      }
      const int oi = -g - 1, t = d->dev_type[oi], i = d->dev_idx[oi];
      if (t == XG_DEV_MOS1) {
        const double* k = &hc.cls[(size_t)hc.mos_cls[i] * KSTRIDE];
        if (k[K_GD] != 0) { dp_lid[i] = row++; hc.var_type.push_back('V'); }
        if (k[K_GS] != 0) { sp_lid[i] = row++; hc.var_type.push_back('V'); }
        hc.mos_sta0[i] = sta; sta += S_COUNT;
        hc.mos_sto0[i] = sto; sto += O_COUNT;
      } else if (t == XG_DEV_VSRC) {
        vs_bra[i] = row++; hc.var_type.push_back('I');
      } else if (t == XG_DEV_CAP) {
        sta += 1;
      }
    }
    const int n = row;
    hc.n = n; hc.nsta = sta; hc.nsto = sto;
    auto lid = [&](int node) { const int g = hc.node_gid[node]; return g < 0 ? n : g; };
    hc.mos_lid.resize((size_t)6 * d->n_mos);
    for (int i = 0; i < d->n_mos; ++i) {
      int* L = &hc.mos_lid[(size_t)6 * i];
      for (int k = 0; k < 4; ++k) L[k] = lid(d->mos_nodes[4 * i + k]);
      L[4] = dp_lid[i] >= 0 ? dp_lid[i] : L[0];
      L[5] = sp_lid[i] >= 0 ? sp_lid[i] : L[2];
    }
    // ---- pattern: union of the device stamps, sorted-unique per row
    std::vector<std::vector<int>> rows(n);
    auto put = [&](int r, int c) { if (r < n && c < n) rows[r].push_back(c); };
    static const int mos_stamp[6][5] = {{0, 4, -1}, {1, 3, 4, 5, -1}, {2, 5, -1}, {1, 3, 4, 5, -1},
                                         {0, 1, 3, 4, 5}, {1, 2, 3, 4, 5}};
    for (auto it = bft.rbegin(); it != bft.rend(); ++it) {
      const int g = gnode[*it];
      if (g >= 0) continue;
      const int oi = -g - 1, t = d->dev_type[oi], i = d->dev_idx[oi];
      if (t == XG_DEV_MOS1) {
        const int* L = &hc.mos_lid[(size_t)6 * i];
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 5 && mos_stamp[r][c] >= 0; ++c) put(L[r], L[mos_stamp[r][c]]);
      } else if (t == XG_DEV_RES || t == XG_DEV_CAP) {
        const int* nn = (t == XG_DEV_RES ? d->res_nodes : d->cap_nodes) + 2 * i;
        const int p = lid(nn[0]), q = lid(nn[1]);
        put(p, p); put(p, q); put(q, p); put(q, q);
      } else if (t == XG_DEV_VSRC) {
        const int p = lid(d->vsrc_nodes[2 * i]), q = lid(d->vsrc_nodes[2 * i + 1]), b = vs_bra[i];
        put(p, b); put(q, b); put(b, p); put(b, q);
      }
    }
    hc.rowptr.assign(n + 1, 0);
    for (int r = 0; r < n; ++r) {
      auto& v = rows[r];
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      hc.rowptr[r + 1] = hc.rowptr[r] + (int)v.size();
    }
    hc.nnz = hc.rowptr[n];
    hc.colind.resize(hc.nnz);
    for (int r = 0; r < n; ++r) std::copy(rows[r].begin(), rows[r].end(), hc.colind.begin() + hc.rowptr[r]);
    rows.clear(); rows.shrink_to_fit();
    auto slot = [&](int r, int c) -> int {
      if (r >= n || c >= n) return -1;
      const int* b = hc.colind.data() + hc.rowptr[r];
      const int* e = hc.colind.data() + hc.rowptr[r + 1];
      const int* f = std::lower_bound(b, e, c);
      if (f == e || *f != c) throw std::runtime_error("stamp entry missing from pattern");
      return (int)(f - hc.colind.data());
    };
    // ---- MOSFET pointer slots (setupPointers order) and used-plane masks
    hc.mos_fslot.resize((size_t)17 * d->n_mos); hc.mos_qslot.resize((size_t)14 * d->n_mos);
    hc.mos_mask.assign(d->n_mos, 0);
    static const int fpair[17][2] = {{0, 0}, {0, 4}, {2, 2}, {2, 5}, {3, 3}, {3, 4}, {3, 5}, {4, 0}, {4, 1},
                                     {4, 3}, {4, 4}, {4, 5}, {5, 1}, {5, 2}, {5, 3}, {5, 4}, {5, 5}};
    static const int qpair[14][2] = {{1, 1}, {1, 3}, {1, 4}, {1, 5}, {3, 1}, {3, 3}, {3, 4}, {3, 5},
                                     {4, 1}, {4, 3}, {4, 4}, {5, 1}, {5, 3}, {5, 5}};
    // vector planes -> local row (d g s b dp sp = 0..5)
    static const int vrow[R_JF0] = {0, 1, 2, 3, 4, 5, 1, 3, 4, 5, 3, 4, 5, 1, 3, 4, 5};
    for (int i = 0; i < d->n_mos; ++i) {
      const int* L = &hc.mos_lid[(size_t)6 * i];
      const double* k = &hc.cls[(size_t)hc.mos_cls[i] * KSTRIDE];
      uint64_t mask = 0;
      for (int p = 0; p < R_JF0; ++p) {
        if (L[vrow[p]] >= n) continue;
        if (p == R_F_D && k[K_GD] == 0.0) continue;
        if (p == R_F_S && k[K_GS] == 0.0) continue;
        mask |= 1ull << p;
      }
      for (int j = 0; j < 17; ++j) {
        const int s = slot(L[fpair[j][0]], L[fpair[j][1]]);
        hc.mos_fslot[(size_t)17 * i + j] = s;
        if (s >= 0) mask |= 1ull << (R_JF0 + j);
      }
      for (int j = 0; j < 14; ++j) {
        const int s = slot(L[qpair[j][0]], L[qpair[j][1]]);
        hc.mos_qslot[(size_t)14 * i + j] = s;
        if (s >= 0) mask |= 1ull << (R_JQ0 + j);
      }
      hc.mos_mask[i] = mask;
    }
    // ---- linear matrices (loaded once, separateLoad) and sources
    hc.linF.assign(hc.nnz, 0.0); hc.linQ.assign(hc.nnz, 0.0);
    auto addv = [&](std::vector<double>& v, int s, double x) { if (s >= 0) v[s] += x; };
    for (int oi = 0; oi < nd; ++oi) {
      const int t = d->dev_type[oi], i = d->dev_idx[oi];
      if (t == XG_DEV_RES || t == XG_DEV_CAP) {
        const int* nn = (t == XG_DEV_RES ? d->res_nodes : d->cap_nodes) + 2 * i;
        const int p = lid(nn[0]), q = lid(nn[1]);
        const double g = t == XG_DEV_RES ? 1.0 / d->res_val[i] : d->cap_val[i];
        auto& v = t == XG_DEV_RES ? hc.linF : hc.linQ;
        addv(v, slot(p, p), g); addv(v, slot(p, q), -g); addv(v, slot(q, p), -g); addv(v, slot(q, q), g);
      }
    }
    for (int oi = 0; oi < nd; ++oi) {
      const int t = d->dev_type[oi], i = d->dev_idx[oi];
      if (t == XG_DEV_VSRC) {
        const int p = lid(d->vsrc_nodes[2 * i]), q = lid(d->vsrc_nodes[2 * i + 1]), b = vs_bra[i];
        addv(hc.linF, slot(p, b), 1.0); addv(hc.linF, slot(q, b), -1.0);
        addv(hc.linF, slot(b, p), 1.0); addv(hc.linF, slot(b, q), -1.0);
      }
    }
    for (int i = 0; i < d->n_vsrc; ++i) hc.srcs.push_back({d->vsrc_wave[i], 0, vs_bra[i], n});
    for (int i = 0; i < d->n_isrc; ++i)
      hc.srcs.push_back({d->isrc_wave[i], 1, lid(d->isrc_nodes[2 * i]), lid(d->isrc_nodes[2 * i + 1])});
    hc.wave_type.assign(d->wave_type, d->wave_type + d->n_wave);
    hc.wave_off.assign(d->wave_off, d->wave_off + d->n_wave + 1);
    hc.wave_data.assign(d->wave_data, d->wave_data + (d->n_wave ? d->wave_off[d->n_wave] : 0));

    // ---- gather maps in device order (the reference's "+=" sequence)
    std::vector<int> vcnt(n + 1, 0), mcnt(hc.nnz + 1, 0);
    for (int i = 0; i < d->n_mos; ++i) {
      const int* L = &hc.mos_lid[(size_t)6 * i];
      const uint64_t mask = hc.mos_mask[i];
      for (int p = 0; p < R_JF0; ++p) if (mask >> p & 1) vcnt[L[vrow[p]]]++;
      for (int j = 0; j < 17; ++j) if (mask >> (R_JF0 + j) & 1) mcnt[hc.mos_fslot[(size_t)17 * i + j]]++;
      for (int j = 0; j < 14; ++j) if (mask >> (R_JQ0 + j) & 1) mcnt[hc.mos_qslot[(size_t)14 * i + j]]++;
    }
    hc.vg_ptr.assign(n + 1, 0); hc.mg_ptr.assign(hc.nnz + 1, 0);
    for (int r = 0; r < n; ++r) hc.vg_ptr[r + 1] = hc.vg_ptr[r] + vcnt[r];
    for (int s = 0; s < hc.nnz; ++s) hc.mg_ptr[s + 1] = hc.mg_ptr[s] + mcnt[s];
    hc.vg_ent.resize(hc.vg_ptr[n]); hc.mg_ent.resize(hc.mg_ptr[hc.nnz]);
    std::vector<int> vfill(hc.vg_ptr.begin(), hc.vg_ptr.end() - 1), mfill(hc.mg_ptr.begin(), hc.mg_ptr.end() - 1);
    for (int i = 0; i < d->n_mos; ++i) {
      const int* L = &hc.mos_lid[(size_t)6 * i];
      const uint64_t mask = hc.mos_mask[i];
      for (int p = 0; p < R_JF0; ++p) if (mask >> p & 1) hc.vg_ent[vfill[L[vrow[p]]]++] = i * 64 + p;
      for (int j = 0; j < 17; ++j) if (mask >> (R_JF0 + j) & 1) hc.mg_ent[mfill[hc.mos_fslot[(size_t)17 * i + j]]++] = i * 64 + R_JF0 + j;
      for (int j = 0; j < 14; ++j) if (mask >> (R_JQ0 + j) & 1) hc.mg_ent[mfill[hc.mos_qslot[(size_t)14 * i + j]]++] = i * 64 + R_JQ0 + j;
    }
    return XG_OK;
  } catch (const std::exception& e) {
    hc.err = e.what();
    return XG_E_SETUP;
  }
}

}  // namespace xg
