"""Seeded synthetic circuits and their flat-array description.

A circuit is described the way a netlist would list it: devices in netlist
order, nodes as integers (0 is ground), MOSFET level-1 ``.model`` cards and
instance parameters with NaN meaning "not given" (the reference's
``given("X")`` tests, e.g. src/DeviceModelPKG/OpenModels/N_DEV_MOSFET1.C:3954).
The same arrays feed the HIP library (``xg_circuit_create``) and the CPU
oracle, and ``to_spice`` writes the equivalent Xyce netlist.

Workloads:
  * ``ring``        CMOS ring oscillator, level-1 MOSFETs, seeded load caps,
                    current-pulse kick (single-instance and Monte Carlo forms).
  * ``mc_ring_batch`` uneven Monte Carlo batch of rings.
  * ``power_grid``  seeded resistive mesh with decaps, VDD pads, PWL sinks.
  * ``gate_dag``    level-1 INV/NAND2 array wired as a random DAG.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

NAN = float("nan")

MOS1, RES, CAP, VSRC, ISRC = 0, 1, 2, 3, 4
W_DC, W_PULSE, W_PWL = 0, 1, 2

# model parameter slots (NaN = not given)
MODEL_PARAMS = ["TYPE", "VTO", "KP", "GAMMA", "PHI", "LAMBDA", "RD", "RS", "CBD", "CBS",
                "IS", "PB", "CGSO", "CGDO", "CGBO", "RSH", "CJ", "MJ", "CJSW", "MJSW",
                "JS", "TOX", "LD", "UO", "FC", "NSUB", "NSS", "TPG", "L", "W"]
INST_PARAMS = ["L", "W", "AD", "AS", "PD", "PS", "NRD", "NRS", "M", "VTO", "KP"]
NMP = len(MODEL_PARAMS)
NIP = len(INST_PARAMS)


def model_card(mtype: int, **kw) -> np.ndarray:
    p = np.full(NMP, NAN)
    p[0] = float(mtype)
    for k, v in kw.items():
        p[MODEL_PARAMS.index(k.upper())] = float(v)
    return p


def inst_card(**kw) -> np.ndarray:
    p = np.full(NIP, NAN)
    for k, v in kw.items():
        p[INST_PARAMS.index(k.upper())] = float(v)
    return p


@dataclass
class Circuit:
    """Netlist-ordered device lists; node 0 is ground."""
    n_nodes: int = 1
    order: list = field(default_factory=list)          # (type, index)
    models: list = field(default_factory=list)
    mos_nodes: list = field(default_factory=list)
    mos_model: list = field(default_factory=list)
    mos_inst: list = field(default_factory=list)
    res: list = field(default_factory=list)            # (p, n, R)
    cap: list = field(default_factory=list)            # (p, n, C)
    vsrc: list = field(default_factory=list)           # (p, n, wave)
    isrc: list = field(default_factory=list)           # (p, n, wave)
    waves: list = field(default_factory=list)          # (type, [params])
    names: dict = field(default_factory=dict)          # node id -> name (for netlists)

    def node(self, name: str | None = None) -> int:
        nid = self.n_nodes
        self.n_nodes += 1
        if name is not None:
            self.names[nid] = name
        return nid

    def add_model(self, card: np.ndarray) -> int:
        self.models.append(np.asarray(card, dtype=np.float64))
        return len(self.models) - 1

    def add_mos(self, d, g, s, b, model, inst=None):
        self.order.append((MOS1, len(self.mos_nodes)))
        self.mos_nodes.append((d, g, s, b))
        self.mos_model.append(model)
        self.mos_inst.append(inst_card() if inst is None else np.asarray(inst, dtype=np.float64))

    def add_res(self, p, n, r):
        self.order.append((RES, len(self.res)))
        self.res.append((p, n, float(r)))

    def add_cap(self, p, n, c):
        self.order.append((CAP, len(self.cap)))
        self.cap.append((p, n, float(c)))

    def add_wave(self, wtype, params) -> int:
        self.waves.append((wtype, [float(v) for v in params]))
        return len(self.waves) - 1

    def add_vsrc(self, p, n, wave):
        self.order.append((VSRC, len(self.vsrc)))
        self.vsrc.append((p, n, wave))

    def add_isrc(self, p, n, wave):
        self.order.append((ISRC, len(self.isrc)))
        self.isrc.append((p, n, wave))

    # ------------------------------------------------------------------
    def arrays(self) -> dict:
        """Flat arrays in the layout of ``xg_circuit_desc`` (include/xyce_amd.h)."""
        def i32(a, shape=None):
            a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
            return a.reshape(shape) if shape is not None else a

        def f64(a):
            return np.ascontiguousarray(np.asarray(a, dtype=np.float64))

        wave_off = [0]
        wave_data = []
        for _, p in self.waves:
            wave_data.extend(p)
            wave_off.append(len(wave_data))
        return dict(
            n_nodes=self.n_nodes,
            dev_type=i32([t for t, _ in self.order]),
            dev_idx=i32([i for _, i in self.order]),
            model_par=f64(np.array(self.models).reshape(-1) if self.models else np.zeros(0)),
            n_models=len(self.models),
            mos_nodes=i32(np.array(self.mos_nodes).reshape(-1) if self.mos_nodes else np.zeros(0)),
            mos_model=i32(self.mos_model),
            mos_inst=f64(np.array(self.mos_inst).reshape(-1) if self.mos_inst else np.zeros(0)),
            res_nodes=i32([(p, n) for p, n, _ in self.res]).reshape(-1),
            res_val=f64([r for _, _, r in self.res]),
            cap_nodes=i32([(p, n) for p, n, _ in self.cap]).reshape(-1),
            cap_val=f64([c for _, _, c in self.cap]),
            vsrc_nodes=i32([(p, n) for p, n, _ in self.vsrc]).reshape(-1),
            vsrc_wave=i32([w for _, _, w in self.vsrc]),
            isrc_nodes=i32([(p, n) for p, n, _ in self.isrc]).reshape(-1),
            isrc_wave=i32([w for _, _, w in self.isrc]),
            wave_type=i32([t for t, _ in self.waves]),
            wave_off=i32(wave_off),
            wave_data=f64(wave_data),
            n_mos=len(self.mos_nodes), n_res=len(self.res), n_cap=len(self.cap),
            n_vsrc=len(self.vsrc), n_isrc=len(self.isrc), n_wave=len(self.waves),
        )

    # ------------------------------------------------------------------
    def to_spice(self, title="xyce_amd circuit", tran=None) -> str:
        """Equivalent Xyce netlist (one .model per distinct card; values with 17
        significant digits, so reading it back gives the same doubles)."""
        nm = lambda i: "0" if i == 0 else self.names.get(i, f"n{i}")
        lines = [f"* {title}"]
        for mi, card in enumerate(self.models):
            typ = "nmos" if card[0] > 0 else "pmos"
            kv = " ".join(f"{MODEL_PARAMS[k]}={card[k]:.17g}" for k in range(1, NMP)
                          if not math.isnan(card[k]))
            lines.append(f".model m{mi} {typ} level=1 {kv}")
        for t, i in self.order:
            if t == MOS1:
                d, g, s, b = self.mos_nodes[i]
                inst = self.mos_inst[i]
                kv = " ".join(f"{INST_PARAMS[k]}={inst[k]:.17g}" for k in range(0, 9)
                              if not math.isnan(inst[k]))
                mname = f"m{self.mos_model[i]}"
                if not (math.isnan(inst[9]) and math.isnan(inst[10])):
                    card = self.models[self.mos_model[i]].copy()
                    if not math.isnan(inst[9]):
                        card[1] = inst[9]
                    if not math.isnan(inst[10]):
                        card[2] = inst[10]
                    typ = "nmos" if card[0] > 0 else "pmos"
                    mname = f"mdev{i}"
                    mkv = " ".join(f"{MODEL_PARAMS[k]}={card[k]:.17g}" for k in range(1, NMP)
                                   if not math.isnan(card[k]))
                    lines.append(f".model {mname} {typ} level=1 {mkv}")
                lines.append(f"M{i} {nm(d)} {nm(g)} {nm(s)} {nm(b)} {mname} {kv}")
            elif t == RES:
                p, n, r = self.res[i]
                lines.append(f"R{i} {nm(p)} {nm(n)} {r:.17g}")
            elif t == CAP:
                p, n, c = self.cap[i]
                lines.append(f"C{i} {nm(p)} {nm(n)} {c:.17g}")
            else:
                p, n, w = (self.vsrc if t == VSRC else self.isrc)[i]
                wt, wp = self.waves[w]
                if wt == W_DC:
                    src = f"DC {wp[0]:.17g}"
                elif wt == W_PULSE:
                    src = "PULSE(" + " ".join(f"{v:.17g}" for v in wp if not math.isnan(v)) + ")"
                else:
                    src = "PWL(" + " ".join(f"{v:.17g}" for v in wp) + ")"
                lines.append(f"{'V' if t == VSRC else 'I'}{i} {nm(p)} {nm(n)} {src}")
        if tran:
            lines.append(".TRAN " + " ".join(f"{v:.17g}" for v in tran))
        lines.append(".END")
        return "\n".join(lines) + "\n"


# ----------------------------------------------------------------------
# Workload generators
# ----------------------------------------------------------------------
NMOS_RING = dict(VTO=0.7, KP=110e-6)
PMOS_RING = dict(VTO=-0.7, KP=50e-6)
VDD = 1.8


def ring(stages: int = 101, seed: int = 1, mc_seed: int | None = None, kick_amp: float = 50e-6,
         kick_width: float = 1e-9, kick_delay: float = 0.1e-9) -> Circuit:
    """CMOS ring oscillator of ``stages`` inverters (level-1 MOSFETs).

    NMOS W/L = 2u/0.5u VTO=0.7 KP=110u; PMOS W/L = 4u/0.5u VTO=-0.7 KP=50u;
    VDD = 1.8 V; per-stage load C ~ U[1,5] fF from ``seed``; a current PULSE
    of ``kick_width`` into stage 1 breaks the metastable DC point.
    With ``mc_seed`` every device gets its own VTO ~ N(VTO0, 30 mV) and
    KP ~ N(KP0, 5%) (per-device overrides of the model card; ``to_spice``
    writes them as one ``.model`` per device).
    """
    c = Circuit()
    vdd = c.node("vdd")
    outs = [c.node(f"s{k + 1}") for k in range(stages)]
    rng = np.random.default_rng(seed)
    caps = rng.uniform(1e-15, 5e-15, size=stages)
    mc = np.random.default_rng(mc_seed) if mc_seed is not None else None
    mn = c.add_model(model_card(+1, **NMOS_RING))
    mp = c.add_model(model_card(-1, **PMOS_RING))
    wdc = c.add_wave(W_DC, [VDD])
    c.add_vsrc(vdd, 0, wdc)
    wk = c.add_wave(W_PULSE, [0.0, kick_amp, kick_delay, 20e-12, 20e-12, kick_width, 1.0])
    c.add_isrc(0, outs[0], wk)
    for k in range(stages):
        out = outs[k]
        inp = outs[k - 1]
        if mc is not None:
            vn = NMOS_RING["VTO"] + 0.03 * mc.standard_normal()
            kn = NMOS_RING["KP"] * (1 + 0.05 * mc.standard_normal())
            vp = PMOS_RING["VTO"] + 0.03 * mc.standard_normal()
            kp = PMOS_RING["KP"] * (1 + 0.05 * mc.standard_normal())
            c.add_mos(out, inp, 0, 0, mn, inst_card(W=2e-6, L=0.5e-6, VTO=vn, KP=kn))
            c.add_mos(out, inp, vdd, vdd, mp, inst_card(W=4e-6, L=0.5e-6, VTO=vp, KP=kp))
        else:
            c.add_mos(out, inp, 0, 0, mn, inst_card(W=2e-6, L=0.5e-6))
            c.add_mos(out, inp, vdd, vdd, mp, inst_card(W=4e-6, L=0.5e-6))
        c.add_cap(out, 0, caps[k])
    return c


def mc_stage_counts(n_inst: int = 4096, seed: int = 3, lo: int = 11, hi: int = 201) -> np.ndarray:
    """Odd stage counts uniform on {lo, lo+2, ..., hi}."""
    rng = np.random.default_rng(seed)
    choices = np.arange(lo, hi + 1, 2)
    return rng.choice(choices, size=n_inst)


def mc_ring_batch(n_inst: int = 4096, seed: int = 3, lo: int = 11, hi: int = 201):
    """The Monte Carlo batch: instance i has S_i stages, load caps from seed
    1000+i and per-device VTO/KP from Monte Carlo seed 1000+i."""
    stages = mc_stage_counts(n_inst, seed, lo, hi)
    return [ring(int(s), seed=1000 + i, mc_seed=1000 + i) for i, s in enumerate(stages)]


def power_grid(nx: int = 512, ny: int = 512, seed: int = 2, pad_pitch: int = 64, n_sinks: int = 4096,
               tstop: float = 2e-9) -> Circuit:
    """Resistive mesh: R ~ U[0.05,0.2] ohm per branch, decap ~ U[10,100] fF per
    node, 1.8 V pads every ``pad_pitch`` nodes, ``n_sinks`` PWL current sinks
    with peaks ~ U[0.1,2] mA and edges ~ U[20,200] ps."""
    c = Circuit()
    rng = np.random.default_rng(seed)
    ids = np.array([[c.node(f"g{i}_{j}") for j in range(nx)] for i in range(ny)])
    wdc = c.add_wave(W_DC, [VDD])
    for i in range(0, ny, pad_pitch):
        for j in range(0, nx, pad_pitch):
            c.add_vsrc(int(ids[i, j]), 0, wdc)
    rh = rng.uniform(0.05, 0.2, size=(ny, nx - 1))
    rv = rng.uniform(0.05, 0.2, size=(ny - 1, nx))
    for i in range(ny):
        for j in range(nx):
            if j + 1 < nx:
                c.add_res(int(ids[i, j]), int(ids[i, j + 1]), rh[i, j])
            if i + 1 < ny:
                c.add_res(int(ids[i, j]), int(ids[i + 1, j]), rv[i, j])
    dc = rng.uniform(10e-15, 100e-15, size=(ny, nx))
    for i in range(ny):
        for j in range(nx):
            c.add_cap(int(ids[i, j]), 0, dc[i, j])
    srng = np.random.default_rng(seed + 1)
    nodes = srng.integers(0, nx * ny, size=n_sinks)
    peaks = srng.uniform(0.1e-3, 2e-3, size=n_sinks)
    rise = srng.uniform(20e-12, 200e-12, size=n_sinks)
    fall = srng.uniform(20e-12, 200e-12, size=n_sinks)
    start = srng.uniform(0.0, 0.5 * tstop, size=n_sinks)
    width = srng.uniform(50e-12, 400e-12, size=n_sinks)
    for k in range(n_sinks):
        t0 = start[k]
        pts = [0.0, 0.0, t0, 0.0, t0 + rise[k], peaks[k], t0 + rise[k] + width[k], peaks[k],
               t0 + rise[k] + width[k] + fall[k], 0.0]
        w = c.add_wave(W_PWL, pts)
        nd = int(ids.reshape(-1)[nodes[k]])
        c.add_isrc(nd, 0, w)
    return c


def gate_dag(n_gates: int = 250_000, seed: int = 4, frac_nand: float = 0.5, window: int = 64,
             diffusion_caps: bool = True) -> Circuit:
    """Level-1 INV/NAND2 array wired as a random DAG: gate k draws its inputs
    from the outputs of the last ``window`` gates that still have fan-out
    left (or from primary inputs driven by PULSE sources), fan-out per gate
    ~ U{1..4}; each gate output carries a wire cap ~ U[0.5,5] fF.  The window
    models placement locality: with unbounded random wiring the graph is an
    expander and any direct factorisation fills in almost completely.

    ``diffusion_caps``: every source and drain diffusion is W x 1 um with a
    junction capacitance (CJ 2e-4 F/m^2, CJSW 1e-10 F/m).  Without it the
    node between a NAND2's two series NMOS carries no charge at all (these
    cards have no TOX, so no Meyer or overlap capacitance): when both devices
    are off it is held only by ~3e-12 S of junction leakage and gmin, Newton
    moves it by hundreds of volts (fetlim and limvds clamp it without raising
    the limiter flag, which only pnjlim sets), the uniform local-error weight
    rejects every step and the step size falls below its minimum within a few
    picoseconds, in the oracle as on the GPU."""
    c = Circuit()
    rng = np.random.default_rng(seed)
    vdd = c.node("vdd")
    jcap = dict(CJ=2e-4, CJSW=1e-10) if diffusion_caps else {}
    mn = c.add_model(model_card(+1, **NMOS_RING, **jcap))
    mp = c.add_model(model_card(-1, **PMOS_RING, **jcap))

    def dev(w):
        if not diffusion_caps:
            return inst_card(W=w, L=0.5e-6)
        return inst_card(W=w, L=0.5e-6, AD=w * 1e-6, AS=w * 1e-6, PD=2 * w + 2e-6, PS=2 * w + 2e-6)
    c.add_vsrc(vdd, 0, c.add_wave(W_DC, [VDD]))
    n_pi = max(4, n_gates // 1000)
    pis = []
    for k in range(n_pi):
        nd = c.node(f"pi{k}")
        per = 2e-9 * (1 + k % 7)
        w = c.add_wave(W_PULSE, [0.0, VDD, 0.1e-9 * (k % 11), 50e-12, 50e-12, per / 2 - 50e-12, per])
        c.add_vsrc(nd, 0, w)
        pis.append(nd)
    outs = []
    fanout_left = []
    nand = rng.random(n_gates) < frac_nand
    wire = rng.uniform(0.5e-15, 5e-15, size=n_gates)
    cap_fan = rng.integers(1, 5, size=n_gates)
    pool = list(pis)
    pool_left = [4] * len(pis)
    for k in range(n_gates):
        nin = 2 if nand[k] else 1
        ins = []
        for _ in range(nin):
            j = int(rng.integers(max(0, len(pool) - window), len(pool)))
            ins.append(pool[j])
            pool_left[j] -= 1
            if pool_left[j] <= 0 and len(pool) > 8:
                pool[j] = pool[-1]; pool_left[j] = pool_left[-1]
                pool.pop(); pool_left.pop()
        out = c.node(f"g{k}")
        if nin == 1:
            c.add_mos(out, ins[0], 0, 0, mn, dev(2e-6))
            c.add_mos(out, ins[0], vdd, vdd, mp, dev(4e-6))
        else:
            mid = c.node(f"m{k}")
            c.add_mos(out, ins[0], mid, 0, mn, dev(4e-6))
            c.add_mos(mid, ins[1], 0, 0, mn, dev(4e-6))
            c.add_mos(out, ins[0], vdd, vdd, mp, dev(4e-6))
            c.add_mos(out, ins[1], vdd, vdd, mp, dev(4e-6))
        c.add_cap(out, 0, wire[k])
        outs.append(out)
        pool.append(out)
        pool_left.append(int(cap_fan[k]))
    return c


def inverter_chain(stages: int = 3, vin_wave=None, **model_extra) -> Circuit:
    """Small driven inverter chain used by unit tests; ``model_extra`` adds
    .model parameters to both cards (e.g. TOX, CJ, CGSO to turn on the Meyer
    and junction capacitances)."""
    c = Circuit()
    vdd = c.node("vdd")
    inp = c.node("in")
    mn = c.add_model(model_card(+1, **NMOS_RING, **model_extra))
    mp = c.add_model(model_card(-1, **PMOS_RING, **model_extra))
    c.add_vsrc(vdd, 0, c.add_wave(W_DC, [VDD]))
    w = c.add_wave(W_PULSE, [0.0, VDD, 0.2e-9, 50e-12, 50e-12, 1e-9, 3e-9]) if vin_wave is None else vin_wave
    c.add_vsrc(inp, 0, w)
    prev = inp
    for k in range(stages):
        out = c.node(f"o{k}")
        c.add_mos(out, prev, 0, 0, mn, inst_card(W=2e-6, L=0.5e-6, AD=4e-12, AS=4e-12, PD=8e-6, PS=8e-6))
        c.add_mos(out, prev, vdd, vdd, mp, inst_card(W=4e-6, L=0.5e-6, AD=8e-12, AS=8e-12, PD=12e-6, PS=12e-6))
        c.add_cap(out, 0, 2e-15)
        prev = out
    return c
