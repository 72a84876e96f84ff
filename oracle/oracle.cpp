// CPU ORACLE — test infrastructure only (see oracle.h).
//
// Every function below cites the reference code it restates.  Paths are
// relative to the Xyce source root.  Where a piece of the reference could not
// be pinned by an independent check, the comment says "parity unpinned".
#include "oracle.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <queue>
#include <stdexcept>
#include <unordered_map>

namespace orc {

static const double CONSTQ = 1.6021918e-19;
static const double CONSTboltz = 1.3806226e-23;
static const double CONSTREFTEMP = 300.15;
static const double CONSTKoverQ = CONSTboltz / CONSTQ;
static const double CONSTMAX_EXP_ARG = 100.0;
static inline double nan_() { return std::numeric_limits<double>::quiet_NaN(); }

// ---------------------------------------------------------------------------
// Limiters and Meyer capacitances
// N_DEV_DeviceSupport.C:161 (limvds), :188 (pnjlim), :309 (fetlim), :507 (qmeyer)
// Pinned by tests/test_oracle_devices.py: pnjlim's current matching, fixed
// points, the documented bounds of fetlim / limvds, textbook Meyer values and
// continuity across region boundaries.
// ---------------------------------------------------------------------------
double limvds(double vnew, double vold) {
  if (vold >= 3.5) {
    if (vnew > vold) vnew = std::min(vnew, 3.0 * vold + 2.0);
    else if (vnew < 3.5) vnew = std::max(vnew, 2.0);
  } else {
    if (vnew > vold) vnew = std::min(vnew, 4.0);
    else vnew = std::max(vnew, -0.5);
  }
  return vnew;
}

double pnjlim(double vnew, double vold, double vt, double vcrit, int* icheck) {
  if (vnew > vcrit && std::fabs(vnew - vold) > vt + vt) {
    if (vold > 0) {
      double arg = 1 + (vnew - vold) / vt;
      vnew = (arg > 0) ? vold + vt * std::log(arg) : vcrit;
    } else {
      vnew = vt * std::log(vnew / vt);
    }
    *icheck = 1;
  } else {
    *icheck = 0;
  }
  return vnew;
}

double fetlim(double vnew, double vold, double vto) {
  double vtsthi = std::fabs(2 * (vold - vto)) + 2;
  double vtstlo = vtsthi / 2 + 2;
  double vtox = vto + 3.5;
  double delv = vnew - vold;
  if (vold >= vto) {
    if (vold >= vtox) {
      if (delv <= 0) {
        if (vnew >= vtox) { if (-delv > vtstlo) vnew = vold - vtstlo; }
        else vnew = std::max(vnew, vto + 2.0);
      } else {
        if (delv >= vtsthi) vnew = vold + vtsthi;
      }
    } else {
      if (delv <= 0) vnew = std::max(vnew, vto - 0.5);
      else vnew = std::min(vnew, vto + 4.0);
    }
  } else {
    if (delv <= 0) {
      if (-delv > vtsthi) vnew = vold - vtsthi;
    } else {
      double vtemp = vto + 0.5;
      if (vnew <= vtemp) { if (delv > vtstlo) vnew = vold + vtstlo; }
      else vnew = vtemp;
    }
  }
  return vnew;
}

void qmeyer(double vgs, double vgd, double /*vgb*/, double von, double vdsat,
            double& capgs, double& capgd, double& capgb, double phi, double cox) {
  double vgst = vgs - von;
  if (vgst <= -phi) {
    capgb = cox / 2; capgs = 0; capgd = 0;
  } else if (vgst <= -phi / 2) {
    capgb = -vgst * cox / (2 * phi); capgs = 0; capgd = 0;
  } else if (vgst <= 0) {
    capgb = -vgst * cox / (2 * phi);
    capgs = vgst * cox / (1.5 * phi) + cox / 3;
    capgd = 0;
  } else {
    double vds = vgs - vgd;
    if (vdsat <= vds) {
      capgs = cox / 3; capgd = 0; capgb = 0;
    } else {
      double vddif = 2.0 * vdsat - vds;
      double vddif1 = vdsat - vds;
      double vddif2 = vddif * vddif;
      capgd = cox * (1.0 - vdsat * vdsat / vddif2) / 3;
      capgs = cox * (1.0 - vddif1 * vddif1 / vddif2) / 3;
      capgb = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// Source waveforms: PulseData::updateSource N_DEV_SourceData.C:1168,
// PulseData::initializeSource (defaults TR=TF=tstep, PW=PER=tstop),
// PWLinData::updateSource :1770 (no REPEAT, TD=0).
// Pinned by closed-form checks at and between breakpoints.
// ---------------------------------------------------------------------------
double Wave::eval(double t, double tstep, double tstop, double bptol) const {
  if (type == W_DC) return p[0];
  if (type == W_PULSE) {
    double V1 = p[0], V2 = p[1];
    double TD = std::isnan(p[2]) ? 0.0 : p[2];
    double TR = std::isnan(p[3]) ? tstep : p[3];
    double TF = std::isnan(p[4]) ? tstep : p[4];
    double PW = std::isnan(p[5]) ? tstop : p[5];
    double PER = std::isnan(p[6]) ? tstop : p[6];
    double time = t - TD;
    if (time > PER && PER != 0.0) time -= PER * std::floor(time / PER);
    if (time <= 0 || (time > TR + PW + TF && std::fabs(time - (TR + PW + TF)) > bptol)) return V1;
    if ((time > TR && std::fabs(time - TR) > bptol) &&
        (time < TR + PW || std::fabs(time - (TR + PW)) < bptol)) return V2;
    if (time > 0 && (time < TR || std::fabs(time - TR) < bptol))
      return TR != 0.0 ? V1 + (V2 - V1) * time / TR : V1;
    return TF != 0.0 ? V2 + (V1 - V2) * (time - (TR + PW)) / TF : V2;
  }
  // PWL: p = t0 v0 t1 v1 ...
  int num = (int)p.size() / 2;
  double time = t;
  if (time < 0) return 0.0;
  double t1, t2, v1, v2;
  if (time <= p[2 * (num - 1)]) {
    int loc = 0;
    for (int i = 0; i < num; ++i) if (time < p[2 * i]) { loc = i; break; }
    if (loc == 0 && !(time < p[0])) loc = num - 1;  // time == last point
    if (loc == 0) { t1 = 0.0; v1 = 0.0; }
    else { t1 = p[2 * (loc - 1)]; v1 = p[2 * (loc - 1) + 1]; }
    t2 = p[2 * loc]; v2 = p[2 * loc + 1];
  } else {
    t1 = 0.0; t2 = 1.0; v1 = v2 = p[2 * (num - 1) + 1];
  }
  if (t1 == t2) return v2;
  double len = t2 - t1;
  double sv = (t2 - time) * v1 / len;
  sv += (-t1 + time) * v2 / len;
  return sv;
}

void Wave::breakpoints(double tstep, double tstop, std::vector<double>& out) const {
  if (type == W_PULSE) {
    double TD = std::isnan(p[2]) ? 0.0 : p[2];
    double TR = std::isnan(p[3]) ? tstep : p[3];
    double TF = std::isnan(p[4]) ? tstep : p[4];
    double PW = std::isnan(p[5]) ? tstop : p[5];
    double PER = std::isnan(p[6]) ? tstop : p[6];
    for (double base = TD; base < tstop; base += PER) {
      const double c[4] = {0.0, TR, TR + PW, TR + PW + TF};
      for (double ci : c) if (base + ci > 0 && base + ci < tstop) out.push_back(base + ci);
      if (PER <= 0) break;
    }
  } else if (type == W_PWL) {
    for (size_t i = 0; i + 1 < p.size(); i += 2)
      if (p[i] > 0 && p[i] < tstop) out.push_back(p[i]);
  }
}

// ---------------------------------------------------------------------------
// MOSFET level-1 setup.
// Model::processParams N_DEV_MOSFET1.C:3939, Instance::processParams :383,
// Instance::updateTemperature :3487.  Defaults from the parameter tables :80-382.
// Pinned by the Shichman-Hodges closed-form tests (tVto == VTO and
// Beta == KP*W/(L-2LD) at TNOM, which any slip in these formulas breaks).
// ---------------------------------------------------------------------------
bool Mos1Model::given(int k) const { return !std::isnan(p[k]); }

static double mp_or(const Mos1Model& m, int k, double dflt) { return m.given(k) ? m.p[k] : dflt; }

static void model_process(Mos1Model& m) {
  m.dtype = (m.p[MP_TYPE] < 0) ? -1 : 1;
  m.vt0 = mp_or(m, MP_VTO, 0.0);
  m.transconductance = mp_or(m, MP_KP, 2e-5);
  m.gamma = mp_or(m, MP_GAMMA, 0.0);
  m.phi = mp_or(m, MP_PHI, 0.6);
  m.lambda = mp_or(m, MP_LAMBDA, 0.0);
  m.drainResistance = mp_or(m, MP_RD, 0.0);
  m.sourceResistance = mp_or(m, MP_RS, 0.0);
  m.capBD = mp_or(m, MP_CBD, 0.0);
  m.capBS = mp_or(m, MP_CBS, 0.0);
  m.jctSatCur = mp_or(m, MP_IS, 1e-14);
  m.bulkJctPotential = mp_or(m, MP_PB, 0.8);
  m.gsoCap = mp_or(m, MP_CGSO, 0.0);
  m.gdoCap = mp_or(m, MP_CGDO, 0.0);
  m.gboCap = mp_or(m, MP_CGBO, 0.0);
  m.sheetResistance = mp_or(m, MP_RSH, 0.0);
  m.bulkCapFactor = mp_or(m, MP_CJ, 0.0);
  m.MJ = mp_or(m, MP_MJ, 0.5);
  m.sideWallCapFactor = mp_or(m, MP_CJSW, 0.0);
  m.MJSW = mp_or(m, MP_MJSW, 0.5);
  m.jctSatCurDensity = mp_or(m, MP_JS, 0.0);
  double tox = mp_or(m, MP_TOX, 1e-7);
  m.latDiff = mp_or(m, MP_LD, 0.0);
  m.surfaceMobility = mp_or(m, MP_UO, 600.0);
  m.fwdCapDepCoeff = mp_or(m, MP_FC, 0.5);
  double nsub = mp_or(m, MP_NSUB, 0.0);
  double nss = mp_or(m, MP_NSS, 0.0);
  double tpg = mp_or(m, MP_TPG, 0.0);
  m.capBDGiven = m.given(MP_CBD) || m.capBD != 0;
  m.capBSGiven = m.given(MP_CBS) || m.capBS != 0;
  m.bulkCapFactorGiven = m.given(MP_CJ);
  m.sideWallCapFactorGiven = m.given(MP_CJSW);
  m.tnom = CONSTREFTEMP;

  m.fact1 = m.tnom / CONSTREFTEMP;
  m.vtnom = m.tnom * CONSTKoverQ;
  double kt1 = CONSTboltz * m.tnom;
  m.egfet1 = 1.16 - (7.02e-4 * m.tnom * m.tnom) / (m.tnom + 1108);
  double arg1 = -m.egfet1 / (kt1 + kt1) + 1.1150877 / (CONSTboltz * (CONSTREFTEMP + CONSTREFTEMP));
  m.pbfact1 = -2 * m.vtnom * (1.5 * std::log(m.fact1) + CONSTQ * arg1);

  if (!m.given(MP_TOX) || tox == 0) {
    m.oxideCapFactor = 0;
  } else {
    m.oxideCapFactor = 3.9 * 8.854214871e-12 / tox;
    if (!m.given(MP_KP)) m.transconductance = m.surfaceMobility * m.oxideCapFactor * 1e-4;
    if (m.given(MP_NSUB)) {
      if (nsub * 1e6 > 1.45e16) {
        if (!m.given(MP_PHI)) {
          m.phi = 2 * m.vtnom * std::log(nsub * 1e6 / 1.45e16);
          m.phi = std::max(0.1, m.phi);
        }
        double fermis = m.dtype * .5 * m.phi;
        double wkfng = 3.2;
        int gateType = m.given(MP_TPG) ? (int)tpg : 1;
        if (gateType != 0) {
          double fermig = m.dtype * gateType * .5 * m.egfet1;
          wkfng = 3.25 + .5 * m.egfet1 - fermig;
        }
        double wkfngs = wkfng - (3.25 + .5 * m.egfet1 + fermis);
        if (!m.given(MP_GAMMA))
          m.gamma = std::sqrt(2 * 11.70 * 8.854214871e-12 * CONSTQ * nsub * 1e6) / m.oxideCapFactor;
        if (!m.given(MP_VTO)) {
          double vfb = wkfngs - nss * 1e4 * CONSTQ / m.oxideCapFactor;
          m.vt0 = vfb + m.dtype * (m.gamma * std::sqrt(m.phi) + m.phi);
        }
      }
    }
  }
}

static void inst_process(Mos1& mi, const Options& opt) {
  Mos1Model& m = mi.mdl;
  auto ipor = [&](int k, double d) { return std::isnan(mi.ip[k]) ? d : mi.ip[k]; };
  mi.l = ipor(IP_L, mp_or(m, MP_L, 1e-4));
  mi.w = ipor(IP_W, mp_or(m, MP_W, 1e-4));
  mi.drainArea = ipor(IP_AD, 0.0);
  mi.sourceArea = ipor(IP_AS, 0.0);
  mi.drainPerimeter = ipor(IP_PD, 0.0);
  mi.sourcePerimeter = ipor(IP_PS, 0.0);
  mi.drainSquares = ipor(IP_NRD, 1.0);
  mi.sourceSquares = ipor(IP_NRS, 1.0);
  mi.M = ipor(IP_M, 1.0);
  mi.temp = opt.temp;

  // updateTemperature N_DEV_MOSFET1.C:3487-3721
  double temp = mi.temp, tnom = m.tnom;
  double ratio = temp / tnom;
  mi.vt = temp * CONSTKoverQ;
  double fact2 = temp / CONSTREFTEMP;
  double kt = temp * CONSTboltz;
  double egfet = 1.16 - (7.02e-4 * temp * temp) / (temp + 1108);
  double arg = -egfet / (kt + kt) + 1.1150877 / (CONSTboltz * (CONSTREFTEMP + CONSTREFTEMP));
  double pbfact = -2 * mi.vt * (1.5 * std::log(fact2) + CONSTQ * arg);
  double ratio4 = ratio * std::sqrt(ratio);
  mi.tTransconductance = m.transconductance / ratio4;
  double phio = (m.phi - m.pbfact1) / m.fact1;
  mi.tPhi = fact2 * phio + pbfact;
  mi.tVbi = m.vt0 - m.dtype * (m.gamma * std::sqrt(m.phi)) + .5 * (m.egfet1 - egfet) +
            m.dtype * .5 * (mi.tPhi - m.phi);
  mi.tVto = mi.tVbi + m.dtype * m.gamma * std::sqrt(mi.tPhi);
  mi.tSatCur = m.jctSatCur * std::exp(-egfet / mi.vt + m.egfet1 / m.vtnom);
  mi.tSatCurDens = m.jctSatCurDensity * std::exp(-egfet / mi.vt + m.egfet1 / m.vtnom);
  double pbo = (m.bulkJctPotential - m.pbfact1) / m.fact1;
  double gmaold = (m.bulkJctPotential - pbo) / pbo;
  double capfact = 1 / (1 + m.MJ * (4e-4 * (m.tnom - CONSTREFTEMP) - gmaold));
  double tCbd = m.capBD * capfact, tCbs = m.capBS * capfact, tCj = m.bulkCapFactor * capfact;
  capfact = 1 / (1 + m.MJSW * (4e-4 * (m.tnom - CONSTREFTEMP) - gmaold));
  double tCjsw = m.sideWallCapFactor * capfact;
  mi.tBulkPot = fact2 * pbo + pbfact;
  double gmanew = (mi.tBulkPot - pbo) / pbo;
  capfact = (1 + m.MJ * (4e-4 * (temp - CONSTREFTEMP) - gmanew));
  tCbd *= capfact; tCbs *= capfact; tCj *= capfact;
  capfact = (1 + m.MJSW * (4e-4 * (temp - CONSTREFTEMP) - gmanew));
  tCjsw *= capfact;
  mi.tDepCap = m.fwdCapDepCoeff * mi.tBulkPot;
  const double root2 = std::sqrt(2.0);
  if (m.jctSatCurDensity == 0 || mi.drainArea == 0 || mi.sourceArea == 0) {
    mi.sourceVcrit = mi.drainVcrit = mi.vt * std::log(mi.vt / (root2 * m.jctSatCur));
  } else {
    mi.drainVcrit = mi.vt * std::log(mi.vt / (root2 * m.jctSatCurDensity * mi.drainArea));
    mi.sourceVcrit = mi.vt * std::log(mi.vt / (root2 * m.jctSatCurDensity * mi.sourceArea));
  }
  double czbd = m.capBDGiven ? tCbd : (m.bulkCapFactorGiven ? tCj * mi.drainArea : 0);
  double czbdsw = m.sideWallCapFactorGiven ? tCjsw * mi.drainPerimeter : 0;
  double a1 = 1 - m.fwdCapDepCoeff;
  double sarg = std::exp((-m.MJ) * std::log(a1));
  double sargsw = std::exp((-m.MJSW) * std::log(a1));
  mi.Cbd = czbd; mi.Cbdsw = czbdsw;
  mi.f2d = czbd * (1 - m.fwdCapDepCoeff * (1 + m.MJ)) * sarg / a1 +
           czbdsw * (1 - m.fwdCapDepCoeff * (1 + m.MJSW)) * sargsw / a1;
  mi.f3d = czbd * m.MJ * sarg / a1 / mi.tBulkPot + czbdsw * m.MJSW * sargsw / a1 / mi.tBulkPot;
  mi.f4d = czbd * mi.tBulkPot * (1 - a1 * sarg) / (1 - m.MJ) +
           czbdsw * mi.tBulkPot * (1 - a1 * sargsw) / (1 - m.MJSW) - mi.f3d / 2 * (mi.tDepCap * mi.tDepCap) -
           mi.tDepCap * mi.f2d;
  double czbs = m.capBSGiven ? tCbs : (m.bulkCapFactorGiven ? tCj * mi.sourceArea : 0);
  double czbssw = m.sideWallCapFactorGiven ? tCjsw * mi.sourcePerimeter : 0;
  mi.Cbs = czbs; mi.Cbssw = czbssw;
  mi.f2s = czbs * (1 - m.fwdCapDepCoeff * (1 + m.MJ)) * sarg / a1 +
           czbssw * (1 - m.fwdCapDepCoeff * (1 + m.MJSW)) * sargsw / a1;
  mi.f3s = czbs * m.MJ * sarg / a1 / mi.tBulkPot + czbssw * m.MJSW * sargsw / a1 / mi.tBulkPot;
  mi.f4s = czbs * mi.tBulkPot * (1 - a1 * sarg) / (1 - m.MJ) +
           czbssw * mi.tBulkPot * (1 - a1 * sargsw) / (1 - m.MJSW) - mi.f3s / 2 * (mi.tDepCap * mi.tDepCap) -
           mi.tDepCap * mi.f2s;

  // Instance::processParams N_DEV_MOSFET1.C:409-469
  if (m.drainResistance != 0) mi.drainConductance = 1 / m.drainResistance;
  else if (m.given(MP_RSH)) mi.drainConductance = m.sheetResistance != 0 ? 1 / (m.sheetResistance * mi.drainSquares) : 0;
  else mi.drainConductance = 0;
  if (m.sourceResistance != 0) mi.sourceConductance = 1 / m.sourceResistance;
  else if (m.given(MP_RSH)) mi.sourceConductance = m.sheetResistance != 0 ? 1 / (m.sheetResistance * mi.sourceSquares) : 0;
  else mi.sourceConductance = 0;
  if (mi.l - 2 * m.latDiff <= 0) throw std::runtime_error("Effective channel length less than zero.");
  mi.EffectiveLength = mi.l - 2 * m.latDiff;
  mi.GateSourceOverlapCap = m.gsoCap * mi.w;
  mi.GateDrainOverlapCap = m.gdoCap * mi.w;
  mi.GateBulkOverlapCap = m.gboCap * mi.EffectiveLength;
  mi.OxideCap = m.oxideCapFactor * mi.EffectiveLength * mi.w;
}

// ---------------------------------------------------------------------------
// Instance::updateIntermediateVars N_DEV_MOSFET1.C:2689-3477 and the state /
// store writes of Master::updateState :4241-4333 (old-Meyer back-averaging).
// x has length n+1; x[n] is the ground dummy (always 0).
// Pinned by: closed-form Shichman-Hodges currents per region, diode closed
// form, depletion charge at MJ=0.5, KCL, n<->p mirror, d<->s swap, and the
// finite-difference Jacobian check the reference itself uses (TESTJAC); the
// Meyer back-averaging by the trapezoid-integral property of the gate charges
// over a gate-voltage ramp (tests/test_oracle_pins.py).
// ---------------------------------------------------------------------------
static void mos1_update(Mos1& mi, const SolverState& ss, const Options& opt, const double* x,
                        const double* staCur, const double* stoCur, double* staNext, double* stoNext) {
  const Mos1Model& m = mi.mdl;
  const int dtype = m.dtype;
  double DrainSatCur, SourceSatCur;
  if (mi.tSatCurDens == 0 || mi.drainArea == 0 || mi.sourceArea == 0) {
    DrainSatCur = mi.tSatCur; SourceSatCur = mi.tSatCur;
  } else {
    DrainSatCur = mi.tSatCurDens * mi.drainArea; SourceSatCur = mi.tSatCurDens * mi.sourceArea;
  }
  double Beta = mi.tTransconductance * mi.w / mi.EffectiveLength;
  mi.Vd = x[mi.li_d]; mi.Vg = x[mi.li_g]; mi.Vs = x[mi.li_s]; mi.Vb = x[mi.li_b];
  mi.Vsp = x[mi.li_sp]; mi.Vdp = x[mi.li_dp];
  mi.Vddp = mi.Vd - mi.Vdp;
  mi.Vssp = mi.Vs - mi.Vsp;
  double Vbsp = mi.Vb - mi.Vsp, Vgsp = mi.Vg - mi.Vsp, Vdpsp = mi.Vdp - mi.Vsp;
  mi.Vgb = mi.Vg - mi.Vb;
  double vbs = dtype * Vbsp, vgs = dtype * Vgsp, vds = dtype * Vdpsp;
  double vbd = vbs - vds, vgd = vgs - vds;
  mi.origFlag = 1; mi.limitedFlag = 0;
  mi.vgs_orig = vgs; mi.vds_orig = vds; mi.vbs_orig = vbs; mi.vbd_orig = vbd; mi.vgd_orig = vgd;

  if (ss.initJct && opt.voltlim) {
    vbs = -1; vgs = dtype * mi.tVto; vds = 0; vbd = vbs - vds; vgd = vgs - vds;
  }
  double vbs_old, vbd_old, vgs_old, vds_old, Von;
  if (ss.newtonIter == 0) {
    if (!ss.dcop) {
      vbs_old = stoCur[mi.sto0 + 1]; vbd_old = stoCur[mi.sto0 + 0];
      vgs_old = stoCur[mi.sto0 + 2]; vds_old = stoCur[mi.sto0 + 3];
      Von = dtype * stoCur[mi.sto0 + 4];
    } else {
      vbs_old = vbs; vbd_old = vbd; vgs_old = vgs; vds_old = vds; Von = 0.0;
    }
  } else {
    vbs_old = stoNext[mi.sto0 + 1]; vbd_old = stoNext[mi.sto0 + 0];
    vgs_old = stoNext[mi.sto0 + 2]; vds_old = stoNext[mi.sto0 + 3];
    Von = dtype * stoNext[mi.sto0 + 4];
  }
  double vgd_old = vgs_old - vds_old;
  int Check = 1;
  if (opt.voltlim) {
    {  // no OFF devices: the initFix&&OFF exclusion never applies
      if (vds_old >= 0) {
        vgs = fetlim(vgs, vgs_old, Von);
        vds = vgs - vgd;
        vds = limvds(vds, vds_old);
        vgd = vgs - vds;
      } else {
        vgd = fetlim(vgd, vgd_old, Von);
        vds = vgs - vgd;
        vds = -limvds(-vds, -vds_old);
        vgs = vgd + vds;
      }
      if (vds >= 0.0) {
        vbs = pnjlim(vbs, vbs_old, mi.vt, mi.sourceVcrit, &Check);
        vbd = vbs - vds;
      } else {
        vbd = pnjlim(vbd, vbd_old, mi.vt, mi.drainVcrit, &Check);
        vbs = vbd + vds;
      }
      if (Check == 1) mi.limitedFlag = 1;
    }
  }
  vbd = vbs - vds; vgd = vgs - vds; mi.Vgb = vgs - vbs;
  if (mi.vgs_orig != vgs || mi.vds_orig != vds || mi.vbs_orig != vbs || mi.vbd_orig != vbd || mi.vgd_orig != vgd)
    mi.origFlag = 0;

  // bulk diodes :2919-2942
  if (vbs <= 0) {
    mi.gbs = SourceSatCur / mi.vt; mi.gbs += opt.gmin; mi.cbs = mi.gbs * vbs;
  } else {
    double evbs = std::exp(std::min(CONSTMAX_EXP_ARG, vbs / mi.vt));
    mi.gbs = SourceSatCur * evbs / mi.vt + opt.gmin;
    mi.cbs = SourceSatCur * (evbs - 1) + opt.gmin * vbs;
  }
  if (vbd <= 0) {
    mi.gbd = DrainSatCur / mi.vt; mi.gbd += opt.gmin; mi.cbd = mi.gbd * vbd;
  } else {
    double evbd = std::exp(std::min(CONSTMAX_EXP_ARG, vbd / mi.vt));
    mi.gbd = DrainSatCur * evbd / mi.vt + opt.gmin;
    mi.cbd = DrainSatCur * (evbd - 1) + opt.gmin * vbd;
  }
  mi.mode = (vds >= 0) ? 1 : -1;
  double Vdsat;
  {
    double sarg, arg, vgst, betap;
    double vbx = (mi.mode == 1 ? vbs : vbd);
    if (vbx <= 0) sarg = std::sqrt(mi.tPhi - vbx);
    else { sarg = std::sqrt(mi.tPhi); sarg = sarg - vbx / (sarg + sarg); sarg = std::max(0.0, sarg); }
    Von = (mi.tVbi * dtype) + m.gamma * sarg;
    vgst = (mi.mode == 1 ? vgs : vgd) - Von;
    Vdsat = std::max(vgst, 0.0);
    arg = (sarg <= 0) ? 0 : m.gamma / (sarg + sarg);
    if (vgst <= 0) {
      mi.cdrain = 0; mi.gm = 0; mi.gds = 0; mi.gmbs = 0;
    } else {
      betap = Beta * (1 + m.lambda * (vds * mi.mode));
      if (vgst <= vds * mi.mode) {
        mi.cdrain = betap * vgst * vgst * .5;
        mi.gm = betap * vgst;
        mi.gds = m.lambda * Beta * vgst * vgst * .5;
        mi.gmbs = mi.gm * arg;
      } else {
        mi.cdrain = betap * (vds * mi.mode) * (vgst - .5 * (vds * mi.mode));
        mi.gm = betap * (vds * mi.mode);
        mi.gds = betap * (vgst - (vds * mi.mode)) + m.lambda * Beta * (vds * mi.mode) * (vgst - .5 * (vds * mi.mode));
        mi.gmbs = mi.gm * arg;
      }
    }
  }
  mi.von = dtype * Von;
  mi.vdsat = dtype * Vdsat;
  mi.cd = mi.mode * mi.cdrain - mi.cbd;

  // depletion charges :3103-3220
  auto dep = [&](double v, double C, double Csw, double f2, double f3, double f4, double& q, double& cap, bool bdform) {
    if (C != 0 || Csw != 0) {
      if (v < mi.tDepCap) {
        double a = 1 - v / mi.tBulkPot, sarg, sargsw;
        if (bdform) {
          if (m.MJ == .5 && m.MJSW == .5) sarg = sargsw = 1 / std::sqrt(a);
          else {
            sarg = (m.MJ == .5) ? 1 / std::sqrt(a) : std::exp(-m.MJ * std::log(a));
            sargsw = (m.MJSW == .5) ? 1 / std::sqrt(a) : std::exp(-m.MJSW * std::log(a));
          }
        } else {
          if (m.MJ == m.MJSW) {
            sarg = sargsw = (m.MJ == .5) ? 1 / std::sqrt(a) : std::exp(-m.MJ * std::log(a));
          } else {
            sarg = (m.MJ == .5) ? 1 / std::sqrt(a) : std::exp(-m.MJ * std::log(a));
            sargsw = (m.MJSW == .5) ? 1 / std::sqrt(a) : std::exp(-m.MJSW * std::log(a));
          }
        }
        q = mi.tBulkPot * (C * (1 - a * sarg) / (1 - m.MJ) + Csw * (1 - a * sargsw) / (1 - m.MJSW));
        cap = C * sarg + Csw * sargsw;
      } else {
        if (bdform) { q = f4 + v * (f2 + v * f3 / 2); cap = f2 + v * f3; }
        else { q = f4 + v * (f2 + v * (f3 / 2)); cap = f2 + f3 * v; }
      }
    } else { q = 0; cap = 0; }
  };
  dep(vbs, mi.Cbs, mi.Cbssw, mi.f2s, mi.f3s, mi.f4s, mi.qbs, mi.capbs, false);
  dep(vbd, mi.Cbd, mi.Cbdsw, mi.f2d, mi.f3d, mi.f4d, mi.qbd, mi.capbd, true);

  // Meyer :3241-3352 (old Meyer with back-averaging)
  if (mi.mode > 0) qmeyer(vgs, vgd, mi.Vgb, Von, Vdsat, mi.capgs, mi.capgd, mi.capgb, mi.tPhi, mi.OxideCap);
  else qmeyer(vgd, vgs, mi.Vgb, Von, Vdsat, mi.capgd, mi.capgs, mi.capgb, mi.tPhi, mi.OxideCap);
  if (ss.dcop) {
    mi.Capgs = 2.0 * mi.capgs + mi.GateSourceOverlapCap;
    mi.Capgd = 2.0 * mi.capgd + mi.GateDrainOverlapCap;
    mi.Capgb = 2.0 * mi.capgb + mi.GateBulkOverlapCap;
  } else {
    double capgs_old = staCur[mi.sta0 + 3], capgd_old = staCur[mi.sta0 + 4], capgb_old = staCur[mi.sta0 + 5];
    mi.Capgs = (mi.capgs + capgs_old + mi.GateSourceOverlapCap);
    mi.Capgd = (mi.capgd + capgd_old + mi.GateDrainOverlapCap);
    mi.Capgb = (mi.capgb + capgb_old + mi.GateBulkOverlapCap);
  }
  mi.Capgs *= (mi.Capgs < 0.0) ? -1.0 : 1.0;
  mi.Capgd *= (mi.Capgd < 0.0) ? -1.0 : 1.0;
  mi.Capgb *= (mi.Capgb < 0.0) ? -1.0 : 1.0;

  mi.Idrain = mi.drainConductance * mi.Vddp;
  mi.Isource = mi.sourceConductance * mi.Vssp;
  if (mi.mode >= 0) {
    mi.Gm = mi.gm; mi.Gmbs = mi.gmbs; mi.nrmsum = mi.Gm + mi.Gmbs; mi.revsum = 0;
    mi.cdreq = dtype * mi.cdrain;
  } else {
    mi.Gm = -mi.gm; mi.Gmbs = -mi.gmbs; mi.nrmsum = 0; mi.revsum = -(mi.Gm + mi.Gmbs);
    mi.cdreq = -(dtype) * mi.cdrain;
  }
  mi.vbs = vbs; mi.vgs = vgs; mi.vds = vds; mi.vbd = vbd; mi.vgd = vgd;

  // Master::updateState :4258-4329 — store: vbd vbs vgs vds von gm; state: qgs qgd qgb capgs capgd capgb qbd qbs
  stoNext[mi.sto0 + 0] = mi.vbd; stoNext[mi.sto0 + 1] = mi.vbs; stoNext[mi.sto0 + 2] = mi.vgs;
  stoNext[mi.sto0 + 3] = mi.vds; stoNext[mi.sto0 + 4] = mi.von; stoNext[mi.sto0 + 5] = mi.Gm;
  staNext[mi.sta0 + 3] = mi.capgs; staNext[mi.sta0 + 4] = mi.capgd; staNext[mi.sta0 + 5] = mi.capgb;
  if (ss.dcop) {
    mi.qgs = mi.Capgs * mi.vgs; mi.qgd = mi.Capgd * mi.vgd; mi.qgb = mi.Capgb * mi.Vgb;
  } else {
    mi.qgs = staCur[mi.sta0 + 0]; mi.qgd = staCur[mi.sta0 + 1]; mi.qgb = staCur[mi.sta0 + 2];
    double vgs1 = stoCur[mi.sto0 + 2], vbs1 = stoCur[mi.sto0 + 1], vds1 = stoCur[mi.sto0 + 3];
    double vgb1 = vgs1 - vbs1, vgd1 = vgs1 - vds1;
    mi.qgs += mi.Capgs * (mi.vgs - vgs1);
    mi.qgd += mi.Capgd * (mi.vgd - vgd1);
    mi.qgb += mi.Capgb * ((mi.vgs - mi.vbs) - vgb1);
  }
  staNext[mi.sta0 + 0] = mi.qgs; staNext[mi.sta0 + 1] = mi.qgd; staNext[mi.sta0 + 2] = mi.qgb;
  staNext[mi.sta0 + 6] = mi.qbd; staNext[mi.sta0 + 7] = mi.qbs;
}

// Master::loadDAEVectors N_DEV_MOSFET1.C:4343-4576 (old Meyer branch).
// The limiter-correction vectors are pinned against the device's own stamps
// times the limiters' moves, gmin excluded on the junctions, with the
// reference's bulk-row charge term (gcgb where the stamp has gcbd,
// :4524-4527) stated in tests/test_oracle_pins.py.
static void mos1_load_vectors(const Mos1& mi, const SolverState& ss, const Options& opt,
                              double* F, double* Q, double* dFdxdVp, double* dQdxdVp) {
  const int Dtype = mi.mdl.dtype;
  const double M = mi.M;
  double ceqbs = Dtype * (mi.cbs), ceqbd = Dtype * (mi.cbd), ceqgb = 0.0, ceqgs = 0.0, ceqgd = 0.0;
  double coef;
  if (mi.drainConductance != 0.0) F[mi.li_d] += mi.Idrain * M;
  coef = (ceqgs + ceqgd + ceqgb);
  F[mi.li_g] += coef * M;
  if (mi.sourceConductance != 0.0) F[mi.li_s] += mi.Isource * M;
  coef = ceqbs + ceqbd - ceqgb;
  F[mi.li_b] += coef * M;
  coef = -mi.Idrain - (ceqbd - mi.cdreq + ceqgd);
  F[mi.li_dp] += coef * M;
  coef = -mi.Isource - (ceqbs + mi.cdreq + ceqgs);
  F[mi.li_sp] += coef * M;

  double Qeqbs = Dtype * (mi.qbs), Qeqbd = Dtype * (mi.qbd), Qeqgb = Dtype * (mi.qgb);
  double Qeqgs = Dtype * (mi.qgs), Qeqgd = Dtype * (mi.qgd);
  coef = (Qeqgs + Qeqgd + Qeqgb);
  Q[mi.li_g] += coef * M;
  coef = Qeqbs + Qeqbd - Qeqgb;
  Q[mi.li_b] += coef * M;
  coef = -(Qeqbd + Qeqgd);
  Q[mi.li_dp] += coef * M;
  coef = -(Qeqbs + Qeqgs);
  Q[mi.li_sp] += coef * M;

  if (!mi.origFlag) {
    const double gmin1 = opt.gmin;
    double dvgx = (mi.mode > 0) ? (mi.vgs - mi.vgs_orig) : (mi.vgd - mi.vgd_orig);
    double dvbx = (mi.mode > 0) ? (mi.vbs - mi.vbs_orig) : (mi.vbd - mi.vbd_orig);
    double j4 = Dtype * (+((mi.gbd - gmin1)) * (mi.vbd - mi.vbd_orig) + ((mi.gbs - gmin1)) * (mi.vbs - mi.vbs_orig));
    double j5 = Dtype * (-((mi.gbd - gmin1)) * (mi.vbd - mi.vbd_orig) + mi.gds * (mi.vds - mi.vds_orig) +
                         mi.Gm * dvgx + mi.Gmbs * dvbx);
    double j6 = Dtype * (-((mi.gbs - gmin1)) * (mi.vbs - mi.vbs_orig) - mi.gds * (mi.vds - mi.vds_orig) -
                         mi.Gm * dvgx - mi.Gmbs * dvbx);
    dFdxdVp[mi.li_b] += j4 * M;
    dFdxdVp[mi.li_dp] += j5 * M;
    dFdxdVp[mi.li_sp] += j6 * M;
    double gcgd = 0, gcgs = 0, gcgb = 0, gcbs = 0, gcbd = 0;
    if (ss.tran || ss.tranop) { gcgd = mi.Capgd; gcgs = mi.Capgs; gcgb = mi.Capgb; gcbs = mi.capbs; gcbd = mi.capbd; }
    double q2 = Dtype * (gcgd * (mi.vgd - mi.vgd_orig) + gcgs * (mi.vgs - mi.vgs_orig) +
                         gcgb * (mi.vgs - mi.vgs_orig - mi.vbs + mi.vbs_orig));
    double q4 = Dtype * (-(gcgb) * (mi.vgs - mi.vgs_orig - mi.vbs + mi.vbs_orig) + (gcgb) * (mi.vbd - mi.vbd_orig) +
                         (gcbs) * (mi.vbs - mi.vbs_orig));
    double q5 = Dtype * (-(gcgd) * (mi.vgd - mi.vgd_orig) - (gcbd) * (mi.vbd - mi.vbd_orig));
    double q6 = Dtype * (-gcgs * (mi.vgs - mi.vgs_orig) - (gcbs) * (mi.vbs - mi.vbs_orig));
    dQdxdVp[mi.li_g] += q2 * M;
    dQdxdVp[mi.li_b] += q4 * M;
    dQdxdVp[mi.li_dp] += q5 * M;
    dQdxdVp[mi.li_sp] += q6 * M;
  }
}

// Master::loadDAEMatrices N_DEV_MOSFET1.C:4587-4868 (old Meyer branch).
// Each write goes to the CSR slot its setupPointers pointer (:1774-1917)
// addresses; slot -1 is the ground dummy and is dropped.
static inline void addv(double* v, int s, double x) { if (s >= 0) v[s] += x; }
static void mos1_load_matrices(const Mos1& mi, const SolverState& ss, double* dFdx, double* dQdx) {
  const double M = mi.M;
  const int* f = mi.fslot;
  const double gd = mi.drainConductance, gs = mi.sourceConductance;
  addv(dFdx, f[0], gd * M);
  addv(dFdx, f[1], -(gd * M));
  addv(dFdx, f[2], gs * M);
  addv(dFdx, f[3], -(gs * M));
  addv(dFdx, f[4], (mi.gbs + mi.gbd) * M);
  addv(dFdx, f[5], -(mi.gbd * M));
  addv(dFdx, f[6], -(mi.gbs * M));
  addv(dFdx, f[7], -(gd * M));
  addv(dFdx, f[8], (mi.Gm) * M);
  addv(dFdx, f[9], (-mi.gbd + mi.Gmbs) * M);
  addv(dFdx, f[10], (gd + mi.gds + mi.gbd + mi.revsum) * M);
  addv(dFdx, f[11], (-mi.gds - mi.nrmsum) * M);
  addv(dFdx, f[12], -((mi.Gm) * M));
  addv(dFdx, f[13], -(gs * M));
  addv(dFdx, f[14], -((mi.gbs + mi.Gmbs) * M));
  addv(dFdx, f[15], -((mi.gds + mi.revsum) * M));
  addv(dFdx, f[16], (gs + mi.gds + mi.gbs + mi.nrmsum) * M);
  double gcgd = 0, gcgs = 0, gcgb = 0, gcbs = 0, gcbd = 0;
  if (ss.tran || ss.tranop) { gcgd = mi.Capgd; gcgs = mi.Capgs; gcgb = mi.Capgb; gcbs = mi.capbs; gcbd = mi.capbd; }
  const int* q = mi.qslot;
  addv(dQdx, q[0], (gcgd + gcgs + gcgb) * M);
  addv(dQdx, q[1], -(gcgb * M));
  addv(dQdx, q[2], -(gcgd * M));
  addv(dQdx, q[3], -(gcgs * M));
  addv(dQdx, q[4], -(gcgb * M));
  addv(dQdx, q[5], (+gcbs + gcbd + gcgb) * M);
  addv(dQdx, q[6], -(+gcbd * M));
  addv(dQdx, q[7], -(+gcbs * M));
  addv(dQdx, q[8], -gcgd * M);
  addv(dQdx, q[9], -gcbd * M);
  addv(dQdx, q[10], (+gcbd + gcgd) * M);
  addv(dQdx, q[11], -(gcgs * M));
  addv(dQdx, q[12], -(+gcbs * M));
  addv(dQdx, q[13], (+gcbs + gcgs) * M);
}

// ---------------------------------------------------------------------------
// Topology: graph insertion (Topology::addDevice N_TOP_Topology.C:191),
// Graph::insertNode (N_UTL_Graph.h:237), BFS from keys1_.begin()
// (Graph::generateBFT N_UTL_Graph.h:707, generateBFT_ :832) using the same
// std::unordered_map<int,...> container, reversed (CktGraphBasic::getBFSNodeList
// N_TOP_CktGraphBasic.C:181), GIDs (SerialLSUtil::setupNodeGIDs :468,
// setupSolnAndStateGIDs :504), sorted-unique rows (generateRowColData :620).
// Pinned by a hand-enumerated 3-stage ring pattern.  The BFS start depends on
// libstdc++'s unordered_map iteration order, as in the reference: parity of the
// node ORDER against a real Xyce run is unpinned (no Xyce build exists here).
// ---------------------------------------------------------------------------
int Circuit::slot(int row, int col) const {
  if (row < 0 || col < 0) return -1;
  auto b = colind.begin() + rowptr[row], e = colind.begin() + rowptr[row + 1];
  auto it = std::lower_bound(b, e, col);
  if (it == e || *it != col) throw std::runtime_error("stamp entry not in pattern");
  return (int)(it - colind.begin());
}

void Circuit::setup() {
  for (auto& mdl : models) model_process(mdl);
  for (auto& mi : mos) {
    mi.mdl = models[mi.model];
    if (!std::isnan(mi.ip[IP_VTO])) mi.mdl.vt0 = mi.ip[IP_VTO];
    if (!std::isnan(mi.ip[IP_KP])) mi.mdl.transconductance = mi.ip[IP_KP];
    inst_process(mi, opt);
  }
  // --- graph insertion in netlist order
  std::vector<std::vector<int>> adj;
  std::unordered_map<int, int> keys1;          // graph index -> (unused) key
  std::vector<int> gkind, gref;                // kind: -1 = voltage node (gref=node id), else device order index
  std::map<int, int> vnode_index;
  auto ext_nodes = [&](int oi) {
    std::vector<int> nl;
    auto [t, i] = order[oi];
    if (t == MOS1) for (int k = 0; k < 4; ++k) nl.push_back(mos[i].node[k]);
    else {
      const TwoTerm& d = (t == RES) ? res[i] : (t == CAP) ? cap[i] : (t == VSRC) ? vsrc[i] : isrc[i];
      nl.push_back(d.node[0]); nl.push_back(d.node[1]);
    }
    return nl;
  };
  for (int oi = 0; oi < (int)order.size(); ++oi) {
    std::vector<int> nl = ext_nodes(oi);
    for (int nd : nl) {
      if (vnode_index.count(nd)) continue;
      int idx = (int)adj.size();
      adj.push_back({});
      keys1[idx] = idx;
      vnode_index[nd] = idx;
      gkind.push_back(-1); gref.push_back(nd);
    }
    int idx = (int)adj.size();
    std::vector<int> ids;
    for (int nd : nl) ids.push_back(vnode_index[nd]);
    adj.push_back(ids);
    for (int id : ids) adj[id].push_back(idx);
    keys1[idx] = idx;
    gkind.push_back(oi); gref.push_back(oi);
  }
  // --- BFS (generateBFT_)
  std::vector<int> bft;
  {
    int start = keys1.begin()->first;
    std::vector<char> found(adj.size(), 0);
    std::queue<int> q;
    q.push(start); bft.push_back(start); found[start] = 1;
    unsigned root = 0;
    while (!q.empty()) {
      int cur = q.front(); q.pop();
      for (int a : adj[cur]) if (!found[a]) { q.push(a); bft.push_back(a); found[a] = 1; }
      if (q.empty() && bft.size() != adj.size()) {
        for (; root < adj.size(); ++root)
          if (adj[root].size() > 0 && !found[root]) { q.push(root); bft.push_back(root); found[root] = 1; break; }
      }
    }
  }
  std::reverse(bft.begin(), bft.end());
  // --- solution / state / store GIDs in ordered-list order
  node_gid.assign(n_nodes, -1);
  int row = 0, sta = 0, sto = 0;
  var_type.clear();
  for (int gi : bft) {
    if (gkind[gi] < 0) {
      int nd = gref[gi];
      if (nd == 0) continue;   // ground
      node_gid[nd] = row++; var_type.push_back('V');
    } else {
      auto [t, i] = order[gkind[gi]];
      if (t == MOS1) {
        Mos1& mi = mos[i];
        mi.li_dp = mi.li_sp = -2;
        if (mi.drainConductance != 0) { mi.li_dp = row++; var_type.push_back('V'); }
        if (mi.sourceConductance != 0) { mi.li_sp = row++; var_type.push_back('V'); }
        mi.sta0 = sta; sta += 8; mi.sto0 = sto; sto += 6;
      } else if (t == VSRC) {
        vsrc[i].li_bra = row++; var_type.push_back('I');
      } else if (t == CAP) {
        cap[i].sta0 = sta; sta += 1;
      }
    }
  }
  n = row; nsta = sta; nsto = sto;
  auto lid = [&](int nd) { int g = node_gid[nd]; return g < 0 ? n : g; };
  for (auto& mi : mos) {
    mi.li_d = lid(mi.node[0]); mi.li_g = lid(mi.node[1]); mi.li_s = lid(mi.node[2]); mi.li_b = lid(mi.node[3]);
    if (mi.li_dp == -2) mi.li_dp = mi.li_d;
    if (mi.li_sp == -2) mi.li_sp = mi.li_s;
  }
  for (auto* v : {&res, &cap, &vsrc, &isrc})
    for (auto& d : *v) { d.li_p = lid(d.node[0]); d.li_n = lid(d.node[1]); }
  // --- row/col lists from device stamps, in ordered-list device order
  std::vector<std::vector<int>> rl(n);
  auto put = [&](int r, int c) { if (r < n && c < n) rl[r].push_back(c); };
  for (int gi : bft) {
    if (gkind[gi] < 0) continue;
    auto [t, i] = order[gkind[gi]];
    if (t == MOS1) {
      const Mos1& mi = mos[i];
      // jacStamp_DC_SC rows (d g s b d' s'), merged through li_dp/li_sp
      const int L[6] = {mi.li_d, mi.li_g, mi.li_s, mi.li_b, mi.li_dp, mi.li_sp};
      const std::vector<std::vector<int>> st = {{0, 4}, {1, 3, 4, 5}, {2, 5}, {1, 3, 4, 5}, {0, 1, 3, 4, 5}, {1, 2, 3, 4, 5}};
      for (int r = 0; r < 6; ++r) for (int c : st[r]) put(L[r], L[c]);
    } else if (t == RES || t == CAP) {
      const TwoTerm& d = (t == RES) ? res[i] : cap[i];
      put(d.li_p, d.li_p); put(d.li_p, d.li_n); put(d.li_n, d.li_p); put(d.li_n, d.li_n);
    } else if (t == VSRC) {
      const TwoTerm& d = vsrc[i];
      put(d.li_p, d.li_bra); put(d.li_n, d.li_bra); put(d.li_bra, d.li_p); put(d.li_bra, d.li_n);
    }
  }
  rowptr.assign(n + 1, 0); colind.clear();
  for (int r = 0; r < n; ++r) {
    std::sort(rl[r].begin(), rl[r].end());
    rl[r].erase(std::unique(rl[r].begin(), rl[r].end()), rl[r].end());
    colind.insert(colind.end(), rl[r].begin(), rl[r].end());
    rowptr[r + 1] = (int)colind.size();
  }
  auto sl = [&](int r, int c) { return (r >= n || c >= n) ? -1 : slot(r, c); };
  for (auto& mi : mos) {
    const int d = mi.li_d, g = mi.li_g, s = mi.li_s, b = mi.li_b, dp = mi.li_dp, sp = mi.li_sp;
    const int F[17][2] = {{d, d}, {d, dp}, {s, s}, {s, sp}, {b, b}, {b, dp}, {b, sp}, {dp, d}, {dp, g}, {dp, b},
                          {dp, dp}, {dp, sp}, {sp, g}, {sp, s}, {sp, b}, {sp, dp}, {sp, sp}};
    const int Qs[14][2] = {{g, g}, {g, b}, {g, dp}, {g, sp}, {b, g}, {b, b}, {b, dp}, {b, sp},
                           {dp, g}, {dp, b}, {dp, dp}, {sp, g}, {sp, b}, {sp, sp}};
    for (int k = 0; k < 17; ++k) mi.fslot[k] = sl(F[k][0], F[k][1]);
    for (int k = 0; k < 14; ++k) mi.qslot[k] = sl(Qs[k][0], Qs[k][1]);
  }
  for (auto* v : {&res, &cap}) for (auto& d : *v) {
    d.slot[0] = sl(d.li_p, d.li_p); d.slot[1] = sl(d.li_p, d.li_n);
    d.slot[2] = sl(d.li_n, d.li_p); d.slot[3] = sl(d.li_n, d.li_n);
  }
  for (auto& d : vsrc) {
    d.slot[0] = sl(d.li_p, d.li_bra); d.slot[1] = sl(d.li_n, d.li_bra);
    d.slot[2] = sl(d.li_bra, d.li_p); d.slot[3] = sl(d.li_bra, d.li_n);
  }
  lin_loaded = false;
}

} // namespace orc

namespace orc {
void mos1_update_ext(Mos1& mi, const SolverState& ss, const Options& opt, const double* x,
                     const double* staCur, const double* stoCur, double* staNext, double* stoNext) {
  mos1_update(mi, ss, opt, x, staCur, stoCur, staNext, stoNext);
}
void mos1_vectors_ext(const Mos1& mi, const SolverState& ss, const Options& opt,
                      double* F, double* Q, double* dFdxdVp, double* dQdxdVp) {
  mos1_load_vectors(mi, ss, opt, F, Q, dFdxdVp, dQdxdVp);
}
void mos1_matrices_ext(const Mos1& mi, const SolverState& ss, double* dFdx, double* dQdx) {
  mos1_load_matrices(mi, ss, dFdx, dQdx);
}
}  // namespace orc
