// MOSFET level-1 evaluation on the GPU: one thread per device, structure of
// arrays in HBM.
//
// Restates, for the transient/DC old-Meyer path (newMeyer off, the default
// N_DEV_DeviceOptions.C:88-149):
//   Instance::updateIntermediateVars  src/DeviceModelPKG/OpenModels/N_DEV_MOSFET1.C:2689-3477
//   Master::updateState (state/store)  N_DEV_MOSFET1.C:4241-4333
//   Master::loadDAEVectors             N_DEV_MOSFET1.C:4343-4576
//   Master::loadDAEMatrices            N_DEV_MOSFET1.C:4587-4868
//   DeviceSupport::limvds/pnjlim/fetlim/qmeyer  src/DeviceModelPKG/Core/N_DEV_DeviceSupport.C:161,188,309,507
// Instead of "+=" through raw pointers, each contribution is written as one
// term of a record plane; the stamp gathers then sum the terms per CSR slot in
// the reference's device order, so no atomics are needed and the sums are
// reproducible bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include "xg_internal.h"

namespace xg {

#pragma clang fp contract(off)

__device__ __forceinline__ double d_limvds(double vnew, double vold) {
  if (vold >= 3.5) {
    if (vnew > vold) vnew = fmin(vnew, 3.0 * vold + 2.0);
    else if (vnew < 3.5) vnew = fmax(vnew, 2.0);
  } else {
    if (vnew > vold) vnew = fmin(vnew, 4.0);
    else vnew = fmax(vnew, -0.5);
  }
  return vnew;
}

__device__ __forceinline__ double d_pnjlim(double vnew, double vold, double vt, double vcrit, int* icheck) {
  if (vnew > vcrit && fabs(vnew - vold) > vt + vt) {
    if (vold > 0) {
      const double arg = 1 + (vnew - vold) / vt;
      vnew = arg > 0 ? vold + vt * log(arg) : vcrit;
    } else {
      vnew = vt * log(vnew / vt);
    }
    *icheck = 1;
  } else {
    *icheck = 0;
  }
  return vnew;
}

__device__ __forceinline__ double d_fetlim(double vnew, double vold, double vto) {
  const double vtsthi = fabs(2 * (vold - vto)) + 2;
  const double vtstlo = vtsthi / 2 + 2;
  const double vtox = vto + 3.5;
  const double delv = vnew - vold;
  if (vold >= vto) {
    if (vold >= vtox) {
      if (delv <= 0) {
        if (vnew >= vtox) { if (-delv > vtstlo) vnew = vold - vtstlo; }
        else vnew = fmax(vnew, vto + 2.0);
      } else if (delv >= vtsthi) {
        vnew = vold + vtsthi;
      }
    } else {
      vnew = delv <= 0 ? fmax(vnew, vto - 0.5) : fmin(vnew, vto + 4.0);
    }
  } else {
    if (delv <= 0) {
      if (-delv > vtsthi) vnew = vold - vtsthi;
    } else {
      const double vtemp = vto + 0.5;
      if (vnew <= vtemp) { if (delv > vtstlo) vnew = vold + vtstlo; }
      else vnew = vtemp;
    }
  }
  return vnew;
}

__device__ __forceinline__ void d_qmeyer(double vgs, double vgd, double von, double vdsat, double& capgs,
                                         double& capgd, double& capgb, double phi, double cox) {
  // branch results in locals, written through the references once (stores
  // through the references in every branch get merged into stores through a
  // selected pointer, which keeps the caller's outputs in scratch memory)
  double gs, gd, gb;
  const double vgst = vgs - von;
  if (vgst <= -phi) {
    gb = cox / 2; gs = 0; gd = 0;
  } else if (vgst <= -phi / 2) {
    gb = -vgst * cox / (2 * phi); gs = 0; gd = 0;
  } else if (vgst <= 0) {
    gb = -vgst * cox / (2 * phi);
    gs = vgst * cox / (1.5 * phi) + cox / 3;
    gd = 0;
  } else {
    const double vds = vgs - vgd;
    if (vdsat <= vds) {
      gs = cox / 3; gd = 0; gb = 0;
    } else {
      const double vddif = 2.0 * vdsat - vds;
      const double vddif1 = vdsat - vds;
      const double vddif2 = vddif * vddif;
      gd = cox * (1.0 - vdsat * vdsat / vddif2) / 3;
      gs = cox * (1.0 - vddif1 * vddif1 / vddif2) / 3;
      gb = 0;
    }
  }
  capgs = gs; capgd = gd; capgb = gb;
}

__device__ __forceinline__ int d_rec_vec_id(int plane) {
  return plane < R_Q_G ? 0 : plane < R_FL_B ? 1 : plane < R_QL_G ? 2 : 3;
}

struct MosFlags {
  int dcop, tran, init_jct, newton_iter;
  int tranop;   // the operating point of a transient
};

// Evaluates one device.  volt[6] = V(d g s b d' s'); sto_cur/sto_nxt/sta_cur
// are the device's 6 store / 8 state values; outputs the new store values
// through keep(O_*, value) and state values through keep(6 + S_*, value) as
// soon as they are known, and the record terms through emit(plane, value),
// every plane once and in increasing plane order (the batch engine's term
// store advances one position per nonzero plane).  Returns the limiter flag.
template <class Keep, class Emit>
__device__ __forceinline__ int mos1_eval(const double* __restrict__ k, double tvbi, double beta, double tvto,
                                         const double volt[6], const double sto_cur[6], const double sto_nxt[6],
                                         const double sta_cur[8], const MosFlags fl, Keep keep, Emit emit) {
  const int dtype = (int)k[K_DTYPE];
  const double gamma = k[K_GAMMA], lambda = k[K_LAMBDA], tphi = k[K_TPHI], vt = k[K_VT];
  const double gmin = k[K_GMIN], M = k[K_M], gd = k[K_GD], gs = k[K_GS];
  const double Vd = volt[0], Vg = volt[1], Vs = volt[2], Vb = volt[3], Vdp = volt[4], Vsp = volt[5];
  const double Vddp = Vd - Vdp, Vssp = Vs - Vsp;
  double vbs = dtype * (Vb - Vsp), vgs = dtype * (Vg - Vsp), vds = dtype * (Vdp - Vsp);
  double vbd = vbs - vds, vgd = vgs - vds;
  const double vgs_o = vgs, vds_o = vds, vbs_o = vbs, vbd_o = vbd, vgd_o = vgd;
  if (fl.init_jct) { vbs = -1; vgs = dtype * tvto; vds = 0; vbd = vbs - vds; vgd = vgs - vds; }
  double vbs_old, vbd_old, vgs_old, vds_old, Von;
  if (fl.newton_iter == 0) {
    if (!fl.dcop) {
      vbs_old = sto_cur[O_VBS]; vbd_old = sto_cur[O_VBD]; vgs_old = sto_cur[O_VGS]; vds_old = sto_cur[O_VDS];
      Von = dtype * sto_cur[O_VON];
    } else {
      vbs_old = vbs; vbd_old = vbd; vgs_old = vgs; vds_old = vds; Von = 0.0;
    }
  } else {
    vbs_old = sto_nxt[O_VBS]; vbd_old = sto_nxt[O_VBD]; vgs_old = sto_nxt[O_VGS]; vds_old = sto_nxt[O_VDS];
    Von = dtype * sto_nxt[O_VON];
  }
  const double vgd_old = vgs_old - vds_old;
  int check = 1, limited = 0;
  if (vds_old >= 0) {
    vgs = d_fetlim(vgs, vgs_old, Von);
    vds = vgs - vgd;
    vds = d_limvds(vds, vds_old);
    vgd = vgs - vds;
  } else {
    vgd = d_fetlim(vgd, vgd_old, Von);
    vds = vgs - vgd;
    vds = -d_limvds(-vds, -vds_old);
    vgs = vgd + vds;
  }
  if (vds >= 0.0) {
    vbs = d_pnjlim(vbs, vbs_old, vt, k[K_SVCRIT], &check);
    vbd = vbs - vds;
  } else {
    vbd = d_pnjlim(vbd, vbd_old, vt, k[K_DVCRIT], &check);
    vbs = vbd + vds;
  }
  if (check == 1) limited = 1;
  vbd = vbs - vds; vgd = vgs - vds;
  const double Vgb = vgs - vbs;
  const int orig = !(vgs_o != vgs || vds_o != vds || vbs_o != vbs || vbd_o != vbd || vgd_o != vgd);

  // bulk junction diodes
  double gbs, cbs, gbd, cbd;
  const double ssat = k[K_SSAT], dsat = k[K_DSAT];
  if (vbs <= 0) { gbs = ssat / vt; gbs += gmin; cbs = gbs * vbs; }
  else { const double e = exp(fmin(100.0, vbs / vt)); gbs = ssat * e / vt + gmin; cbs = ssat * (e - 1) + gmin * vbs; }
  if (vbd <= 0) { gbd = dsat / vt; gbd += gmin; cbd = gbd * vbd; }
  else { const double e = exp(fmin(100.0, vbd / vt)); gbd = dsat * e / vt + gmin; cbd = dsat * (e - 1) + gmin * vbd; }
  const int mode = vds >= 0 ? 1 : -1;
  // Shichman-Hodges drain current
  double cdrain, gm, gds, gmbs, Vdsat;
  {
    const double vbx = mode == 1 ? vbs : vbd;
    double sarg;
    if (vbx <= 0) sarg = sqrt(tphi - vbx);
    else { sarg = sqrt(tphi); sarg = sarg - vbx / (sarg + sarg); sarg = fmax(0.0, sarg); }
    Von = (tvbi * dtype) + gamma * sarg;
    const double vgst = (mode == 1 ? vgs : vgd) - Von;
    Vdsat = fmax(vgst, 0.0);
    const double arg = sarg <= 0 ? 0 : gamma / (sarg + sarg);
    if (vgst <= 0) {
      cdrain = 0; gm = 0; gds = 0; gmbs = 0;
    } else {
      const double vm = vds * mode;
      const double betap = beta * (1 + lambda * vm);
      if (vgst <= vm) {
        cdrain = betap * vgst * vgst * .5;
        gm = betap * vgst;
        gds = lambda * beta * vgst * vgst * .5;
        gmbs = gm * arg;
      } else {
        cdrain = betap * vm * (vgst - .5 * vm);
        gm = betap * vm;
        gds = betap * (vgst - vm) + lambda * beta * vm * (vgst - .5 * vm);
        gmbs = gm * arg;
      }
    }
  }
  const double von = dtype * Von;
  // depletion charges
  const double tdep = k[K_TDEPCAP], tbulk = k[K_TBULKPOT], mj = k[K_MJ], mjsw = k[K_MJSW];
  double qbs = 0, capbs = 0, qbd = 0, capbd = 0;
  {
    const double C = k[K_CBS], Csw = k[K_CBSSW];
    if (C != 0 || Csw != 0) {
      if (vbs < tdep) {
        const double a = 1 - vbs / tbulk;
        double sarg, sargsw;
        if (mj == mjsw) {
          sarg = sargsw = (mj == .5) ? 1 / sqrt(a) : exp(-mj * log(a));
        } else {
          sarg = (mj == .5) ? 1 / sqrt(a) : exp(-mj * log(a));
          sargsw = (mjsw == .5) ? 1 / sqrt(a) : exp(-mjsw * log(a));
        }
        qbs = tbulk * (C * (1 - a * sarg) / (1 - mj) + Csw * (1 - a * sargsw) / (1 - mjsw));
        capbs = C * sarg + Csw * sargsw;
      } else {
        qbs = k[K_F4S] + vbs * (k[K_F2S] + vbs * (k[K_F3S] / 2));
        capbs = k[K_F2S] + k[K_F3S] * vbs;
      }
    }
  }
  {
    const double C = k[K_CBD], Csw = k[K_CBDSW];
    if (C != 0 || Csw != 0) {
      if (vbd < tdep) {
        const double a = 1 - vbd / tbulk;
        double sarg, sargsw;
        if (mj == .5 && mjsw == .5) {
          sarg = sargsw = 1 / sqrt(a);
        } else {
          sarg = (mj == .5) ? 1 / sqrt(a) : exp(-mj * log(a));
          sargsw = (mjsw == .5) ? 1 / sqrt(a) : exp(-mjsw * log(a));
        }
        qbd = tbulk * (C * (1 - a * sarg) / (1 - mj) + Csw * (1 - a * sargsw) / (1 - mjsw));
        capbd = C * sarg + Csw * sargsw;
      } else {
        qbd = k[K_F4D] + vbd * (k[K_F2D] + vbd * k[K_F3D] / 2);
        capbd = k[K_F2D] + vbd * k[K_F3D];
      }
    }
  }
  // Meyer capacitances with back-averaging
  double capgs, capgd, capgb;
  const double cox = k[K_OXCAP];
  {
    // one call with the terminals swapped in reverse mode (two calls writing
    // through swapped references would make the compiler keep the three
    // outputs in scratch memory)
    double c_s, c_d;
    d_qmeyer(mode > 0 ? vgs : vgd, mode > 0 ? vgd : vgs, Von, Vdsat, c_s, c_d, capgb, tphi, cox);
    capgs = mode > 0 ? c_s : c_d;
    capgd = mode > 0 ? c_d : c_s;
  }
  double Capgs, Capgd, Capgb;
  if (fl.dcop) {
    Capgs = 2.0 * capgs + k[K_GSO]; Capgd = 2.0 * capgd + k[K_GDO]; Capgb = 2.0 * capgb + k[K_GBO];
  } else {
    Capgs = (capgs + sta_cur[S_CAPGS] + k[K_GSO]);
    Capgd = (capgd + sta_cur[S_CAPGD] + k[K_GDO]);
    Capgb = (capgb + sta_cur[S_CAPGB] + k[K_GBO]);
  }
  Capgs *= (Capgs < 0.0) ? -1.0 : 1.0;
  Capgd *= (Capgd < 0.0) ? -1.0 : 1.0;
  Capgb *= (Capgb < 0.0) ? -1.0 : 1.0;
  const double Idrain = gd * Vddp, Isource = gs * Vssp;
  double Gm, Gmbs, nrmsum, revsum, cdreq;
  if (mode >= 0) { Gm = gm; Gmbs = gmbs; nrmsum = Gm + Gmbs; revsum = 0; cdreq = dtype * cdrain; }
  else { Gm = -gm; Gmbs = -gmbs; nrmsum = 0; revsum = -(Gm + Gmbs); cdreq = -(dtype)*cdrain; }

  // store / state (Master::updateState)
  keep(O_VBD, vbd); keep(O_VBS, vbs); keep(O_VGS, vgs); keep(O_VDS, vds);
  keep(O_VON, von); keep(O_GM, Gm);
  double qgs, qgd, qgb;
  if (fl.dcop) {
    qgs = Capgs * vgs; qgd = Capgd * vgd; qgb = Capgb * Vgb;
  } else {
    qgs = sta_cur[S_QGS]; qgd = sta_cur[S_QGD]; qgb = sta_cur[S_QGB];
    const double vgs1 = sto_cur[O_VGS], vbs1 = sto_cur[O_VBS], vds1 = sto_cur[O_VDS];
    const double vgb1 = vgs1 - vbs1, vgd1 = vgs1 - vds1;
    qgs += Capgs * (vgs - vgs1);
    qgd += Capgd * (vgd - vgd1);
    qgb += Capgb * ((vgs - vbs) - vgb1);
  }
  keep(6 + S_QGS, qgs); keep(6 + S_QGD, qgd); keep(6 + S_QGB, qgb);
  keep(6 + S_CAPGS, capgs); keep(6 + S_CAPGD, capgd); keep(6 + S_CAPGB, capgb);
  keep(6 + S_QBD, qbd); keep(6 + S_QBS, qbs);

  // ---- vector terms (loadDAEVectors)
  const double ceqbs = dtype * (cbs), ceqbd = dtype * (cbd), ceqgb = 0.0, ceqgs = 0.0, ceqgd = 0.0;
  emit(R_F_D, Idrain * M);
  emit(R_F_G, (ceqgs + ceqgd + ceqgb) * M);
  emit(R_F_S, Isource * M);
  emit(R_F_B, (ceqbs + ceqbd - ceqgb) * M);
  emit(R_F_DP, (-Idrain - (ceqbd - cdreq + ceqgd)) * M);
  emit(R_F_SP, (-Isource - (ceqbs + cdreq + ceqgs)) * M);
  const double Qeqbs = dtype * (qbs), Qeqbd = dtype * (qbd), Qeqgb = dtype * (qgb), Qeqgs = dtype * (qgs),
               Qeqgd = dtype * (qgd);
  emit(R_Q_G, (Qeqgs + Qeqgd + Qeqgb) * M);
  emit(R_Q_B, (Qeqbs + Qeqbd - Qeqgb) * M);
  emit(R_Q_DP, (-(Qeqbd + Qeqgd)) * M);
  emit(R_Q_SP, (-(Qeqbs + Qeqgs)) * M);
  double gcgd = 0, gcgs = 0, gcgb = 0, gcbs = 0, gcbd = 0;
  if (fl.tran || fl.tranop) { gcgd = Capgd; gcgs = Capgs; gcgb = Capgb; gcbs = capbs; gcbd = capbd; }
  if (!orig) {
    const double dvgx = mode > 0 ? (vgs - vgs_o) : (vgd - vgd_o);
    const double dvbx = mode > 0 ? (vbs - vbs_o) : (vbd - vbd_o);
    const double j4 = dtype * (+((gbd - gmin)) * (vbd - vbd_o) + ((gbs - gmin)) * (vbs - vbs_o));
    const double j5 = dtype * (-((gbd - gmin)) * (vbd - vbd_o) + gds * (vds - vds_o) + Gm * dvgx + Gmbs * dvbx);
    const double j6 = dtype * (-((gbs - gmin)) * (vbs - vbs_o) - gds * (vds - vds_o) - Gm * dvgx - Gmbs * dvbx);
    emit(R_FL_B, j4 * M); emit(R_FL_DP, j5 * M); emit(R_FL_SP, j6 * M);
    const double q2 = dtype * (gcgd * (vgd - vgd_o) + gcgs * (vgs - vgs_o) + gcgb * (vgs - vgs_o - vbs + vbs_o));
    const double q4 = dtype * (-(gcgb) * (vgs - vgs_o - vbs + vbs_o) + (gcgb) * (vbd - vbd_o) + (gcbs) * (vbs - vbs_o));
    const double q5 = dtype * (-(gcgd) * (vgd - vgd_o) - (gcbd) * (vbd - vbd_o));
    const double q6 = dtype * (-gcgs * (vgs - vgs_o) - (gcbs) * (vbs - vbs_o));
    emit(R_QL_G, q2 * M); emit(R_QL_B, q4 * M); emit(R_QL_DP, q5 * M); emit(R_QL_SP, q6 * M);
  } else {
    emit(R_FL_B, 0.0); emit(R_FL_DP, 0.0); emit(R_FL_SP, 0.0);
    emit(R_QL_G, 0.0); emit(R_QL_B, 0.0); emit(R_QL_DP, 0.0); emit(R_QL_SP, 0.0);
  }
  // ---- matrix terms (loadDAEMatrices, pointer order of setupPointers)
  emit(R_JF0 + 0, gd * M);
  emit(R_JF0 + 1, -(gd * M));
  emit(R_JF0 + 2, gs * M);
  emit(R_JF0 + 3, -(gs * M));
  emit(R_JF0 + 4, (gbs + gbd) * M);
  emit(R_JF0 + 5, -(gbd * M));
  emit(R_JF0 + 6, -(gbs * M));
  emit(R_JF0 + 7, -(gd * M));
  emit(R_JF0 + 8, (Gm)*M);
  emit(R_JF0 + 9, (-gbd + Gmbs) * M);
  emit(R_JF0 + 10, (gd + gds + gbd + revsum) * M);
  emit(R_JF0 + 11, (-gds - nrmsum) * M);
  emit(R_JF0 + 12, -((Gm)*M));
  emit(R_JF0 + 13, -(gs * M));
  emit(R_JF0 + 14, -((gbs + Gmbs) * M));
  emit(R_JF0 + 15, -((gds + revsum) * M));
  emit(R_JF0 + 16, (gs + gds + gbs + nrmsum) * M);
  emit(R_JQ0 + 0, (gcgd + gcgs + gcgb) * M);
  emit(R_JQ0 + 1, -(gcgb * M));
  emit(R_JQ0 + 2, -(gcgd * M));
  emit(R_JQ0 + 3, -(gcgs * M));
  emit(R_JQ0 + 4, -(gcgb * M));
  emit(R_JQ0 + 5, (+gcbs + gcbd + gcgb) * M);
  emit(R_JQ0 + 6, -(+gcbd * M));
  emit(R_JQ0 + 7, -(+gcbs * M));
  emit(R_JQ0 + 8, -gcgd * M);
  emit(R_JQ0 + 9, -gcbd * M);
  emit(R_JQ0 + 10, (+gcbd + gcgd) * M);
  emit(R_JQ0 + 11, -(gcgs * M));
  emit(R_JQ0 + 12, -(+gcbs * M));
  emit(R_JQ0 + 13, (+gcbs + gcgs) * M);
  return limited;
}

// PULSE / PWL / DC source value at t (PulseData::updateSource
// N_DEV_SourceData.C:1168, PWLinData::updateSource :1770).
__device__ __forceinline__ double wave_value(int type, const double* p, int np, double t, double tstep, double tstop) {
  if (type == XG_WAVE_DC) return p[0];
  const double bptol = 1e-18;
  if (type == XG_WAVE_PULSE) {
    const double V1 = p[0], V2 = p[1];
    const double TD = isnan(p[2]) ? 0.0 : p[2];
    const double TR = isnan(p[3]) ? tstep : p[3];
    const double TF = isnan(p[4]) ? tstep : p[4];
    const double PW = isnan(p[5]) ? tstop : p[5];
    const double PER = isnan(p[6]) ? tstop : p[6];
    double time = t - TD;
    if (time > PER && PER != 0.0) time -= PER * floor(time / PER);
    if (time <= 0 || (time > TR + PW + TF && fabs(time - (TR + PW + TF)) > bptol)) return V1;
    if ((time > TR && fabs(time - TR) > bptol) && (time < TR + PW || fabs(time - (TR + PW)) < bptol)) return V2;
    if (time > 0 && (time < TR || fabs(time - TR) < bptol)) return TR != 0.0 ? V1 + (V2 - V1) * time / TR : V1;
    return TF != 0.0 ? V2 + (V1 - V2) * (time - (TR + PW)) / TF : V2;
  }
  const int num = np / 2;
  if (t < 0) return 0.0;
  double t1, t2, v1, v2;
  if (t <= p[2 * (num - 1)]) {
    int loc = num - 1;
    for (int i = 0; i < num; ++i) if (t < p[2 * i]) { loc = i; break; }
    if (loc == 0) { t1 = 0.0; v1 = 0.0; } else { t1 = p[2 * (loc - 1)]; v1 = p[2 * (loc - 1) + 1]; }
    t2 = p[2 * loc]; v2 = p[2 * loc + 1];
  } else {
    t1 = 0.0; t2 = 1.0; v1 = v2 = p[2 * (num - 1) + 1];
  }
  if (t1 == t2) return v2;
  const double len = t2 - t1;
  double sv = (t2 - t) * v1 / len;
  sv += (-t1 + t) * v2 / len;
  return sv;
}

#pragma clang fp contract(fast)

}  // namespace xg
