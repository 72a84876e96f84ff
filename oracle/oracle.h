// CPU ORACLE — test infrastructure only.
//
// A plain C++17 FP64 restatement of the reference algorithms on the per-Newton
// hot path of a Xyce transient run.  It is the parity checker for the HIP
// library and the timed CPU baseline.  It is NOT part of the product: only the
// tests, the smoke check and the CPU-baseline leg of the benchmark load it.
//
// Reference files restated (paths relative to the Xyce source root):
//   MOSFET level 1 ............ src/DeviceModelPKG/OpenModels/N_DEV_MOSFET1.C
//   limiters / Meyer caps ..... src/DeviceModelPKG/Core/N_DEV_DeviceSupport.C
//   linear devices ............ N_DEV_Resistor.C, N_DEV_Capacitor.C, N_DEV_Vsrc.C, N_DEV_ISRC.C
//   source waveforms .......... src/DeviceModelPKG/Core/N_DEV_SourceData.C
//   loader (separateLoad) ..... src/LoaderServicesPKG/N_LOA_CktLoader.C
//   residual / Jacobian ....... src/TimeIntegrationPKG/N_TIA_OneStep.C
//   Newton .................... src/NonlinearSolverPKG/N_NLS_DampedNewton.C
//   topology / CSR pattern .... src/TopoManagerPKG/N_TOP_SerialLSUtil.C,
//                               src/UtilityPKG/N_UTL_Graph.h
//
// Parity status of each piece is stated next to its function in oracle.cpp.
#pragma once
#include <cstdint>
#include <vector>
#include <string>

namespace orc {

enum DevType { MOS1 = 0, RES = 1, CAP = 2, VSRC = 3, ISRC = 4 };
enum WaveType { W_DC = 0, W_PULSE = 1, W_PWL = 2 };

// Model parameter slots (NaN in the input means "not given").
enum MP {
  MP_TYPE = 0, MP_VTO, MP_KP, MP_GAMMA, MP_PHI, MP_LAMBDA, MP_RD, MP_RS, MP_CBD, MP_CBS,
  MP_IS, MP_PB, MP_CGSO, MP_CGDO, MP_CGBO, MP_RSH, MP_CJ, MP_MJ, MP_CJSW, MP_MJSW,
  MP_JS, MP_TOX, MP_LD, MP_UO, MP_FC, MP_NSUB, MP_NSS, MP_TPG, MP_L, MP_W, MP_COUNT
};
// Instance parameter slots.
enum IP {
  IP_L = 0, IP_W, IP_AD, IP_AS, IP_PD, IP_PS, IP_NRD, IP_NRS, IP_M, IP_VTO, IP_KP, IP_COUNT
};

struct SolverState {
  int dcop = 0, tran = 0, initJct = 0, initFix = 0;
  int tranop = 0;        // the operating point of a .TRAN (SolverState::tranopFlag)
  int newtonIter = 0;
  int order = 1;
  double h = 0.0;        // current time step
  double time = 0.0;     // t_{n+1}
  // integration method: 0 = OneStep trapezoid, 1 = Gear12 (BDF1/BDF2), whose
  // residual weights the charges by alpha0..alpha2 (N_TIA_Gear12.C:1050-1111)
  int method = 0;
  double a0 = 1.0, a1 = -1.0, a2 = 0.0;
};

struct Options {
  double gmin = 1e-12;
  double temp = 300.15;
  int voltlim = 1;
};

struct Wave {
  int type = W_DC;
  std::vector<double> p;
  double eval(double t, double tstep, double tstop, double bptol) const;
  void breakpoints(double tstep, double tstop, std::vector<double>& out) const;
};

struct Mos1Model {
  int dtype = 1;
  double p[MP_COUNT];
  bool given(int k) const;
  // derived (Model::processParams)
  double tnom, fact1, vtnom, egfet1, pbfact1, oxideCapFactor;
  double vt0, transconductance, gamma, phi, lambda, surfaceMobility;
  double drainResistance, sourceResistance, capBD, capBS, jctSatCur, bulkJctPotential;
  double gsoCap, gdoCap, gboCap, sheetResistance, bulkCapFactor, MJ, sideWallCapFactor, MJSW;
  double jctSatCurDensity, latDiff, fwdCapDepCoeff;
  bool capBDGiven, capBSGiven, bulkCapFactorGiven, sideWallCapFactorGiven;
};

struct Mos1 {
  int node[4];            // netlist node ids d g s b (0 = ground)
  int model;
  double ip[IP_COUNT];
  // setup-derived constants (Instance::processParams / updateTemperature)
  Mos1Model mdl;          // per-device copy (Monte Carlo VTO/KP overrides applied)
  double l, w, drainArea, sourceArea, drainPerimeter, sourcePerimeter, drainSquares, sourceSquares, M;
  double temp, vt, tTransconductance, tPhi, tVbi, tVto, tSatCur, tSatCurDens, tBulkPot, tDepCap;
  double drainVcrit, sourceVcrit, Cbd, Cbdsw, Cbs, Cbssw, f2d, f3d, f4d, f2s, f3s, f4s;
  double drainConductance, sourceConductance, EffectiveLength;
  double GateSourceOverlapCap, GateDrainOverlapCap, GateBulkOverlapCap, OxideCap;
  // solution LIDs (index into x; n == ground dummy)
  int li_d, li_g, li_s, li_b, li_dp, li_sp;
  int sta0, sto0;         // first state / store index
  int fslot[17], qslot[14];
  // intermediate variables (Instance::updateIntermediateVars)
  double Vd, Vg, Vs, Vb, Vdp, Vsp, Vddp, Vssp, Vgb;
  double vbs, vgs, vds, vbd, vgd, vgs_orig, vds_orig, vbs_orig, vbd_orig, vgd_orig;
  double von, vdsat, cdrain, gm, gds, gmbs, gbs, gbd, cbs, cbd, cd, qbs, qbd, capbs, capbd;
  double capgs, capgd, capgb, Capgs, Capgd, Capgb, qgs, qgd, qgb;
  double Gm, Gmbs, nrmsum, revsum, cdreq, Idrain, Isource;
  int mode, origFlag, limitedFlag;
};

struct TwoTerm { int node[2]; double val; int wave; int li_p, li_n, li_bra; int sta0; int slot[4]; };

struct Circuit {
  Options opt;
  int n_nodes = 0;
  std::vector<std::pair<int,int>> order;   // netlist order (type, index)
  std::vector<Mos1Model> models;
  std::vector<Mos1> mos;
  std::vector<TwoTerm> res, cap, vsrc, isrc;
  std::vector<Wave> waves;

  // topology
  int n = 0;                 // number of solution unknowns
  int nsta = 0, nsto = 0;
  std::vector<int> node_gid; // node id -> solution index (-1 ground)
  std::vector<int> rowptr, colind;
  std::vector<char> var_type; // 'V' or 'I'

  // linear (separateLoad) cache in CSR form
  std::vector<double> linF, linQ;
  bool lin_loaded = false;

  void setup();
  int slot(int row, int col) const;   // CSR position or -1
};

// Newton / transient parameters (defaults = Xyce transient defaults)
struct TranParams {
  double tstop = 0, tstep = 0, hmax = 0;
  double reltol = 1e-3, abstol = 1e-6;       // TIA
  double nl_reltol = 1e-2, nl_abstol = 1e-6, dxtol = 0.33, rhstol = 1e-2; int maxNewton = 20;
  double smallUpdateTol = 1e-6;
};

} // namespace orc
