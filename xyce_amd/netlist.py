"""Netlist subset reader and .PRN writer/reader.

The reader takes the part of Xyce's netlist language that the hot path
covers (src/IOInterfacePKG/N_IO_NetlistImportTool.h:130 is the reference's
importer) and builds a ``Circuit``:

* elements: ``R``, ``C``, ``V`` and ``I`` (DC value, ``PULSE(...)``,
  ``PWL(...)``), ``M`` (level-1 MOSFET: d g s b model, L W AD AS PD PS NRD NRS
  M);
* ``.MODEL name NMOS|PMOS (LEVEL=1 ...)`` with the level-1 parameters;
* ``.TRAN tstep tstop [tstart [tmax]]``, ``.PRINT TRAN V(node) ...``,
  ``.END``; ``.OPTIONS TIMEINT METHOD=TRAP|GEAR`` (also 7|8, the reference's
  numbering, src/TimeIntegrationPKG/N_TIA_fwd.h:52-58) selects the
  integrator; ``.OPTIONS`` lines may restate the defaults the path runs with
  (TIMEINT RELTOL=1e-3 ABSTOL=1e-6, DEVICE GMIN=1e-12 VOLTLIM=1, LINSOL
  TYPE=KLU) and anything else there is refused, as are the MOSFET ``OFF``
  and ``IC`` instance settings;
* ``*`` comment lines, ``;`` end-of-line comments, ``+`` continuation lines,
  case-insensitive names, SPICE scale suffixes (T G MEG K MIL M U N P F).

Nodes are numbered in order of first appearance, devices keep netlist order
(the order Xyce's topology inserts them in).  Anything else raises
``NetlistError`` naming the line.

The .PRN writer follows Xyce's default transient output
(src/IOInterfacePKG/N_IO_OutputterTimePrn.C): a header row ``Index TIME
V(...)...``, one row per accepted time point in scientific notation, and the
``End of Xyce(TM) Simulation`` line.
"""
from __future__ import annotations

import math
import re

import numpy as np

from .circuits import (INST_PARAMS, MODEL_PARAMS, W_DC, W_PULSE, W_PWL, Circuit, inst_card,
                       model_card)


class NetlistError(ValueError):
    pass


_SCALE = [("MEG", 1e6), ("MIL", 25.4e-6), ("T", 1e12), ("G", 1e9), ("K", 1e3), ("M", 1e-3), ("U", 1e-6),
          ("N", 1e-9), ("P", 1e-12), ("F", 1e-15)]
_NUM = re.compile(r"^([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)([a-zA-Z]*)$")


def value(tok: str) -> float:
    """A SPICE number: 1.5k, 2u, 10fF, 3MEG, 1e-12."""
    m = _NUM.match(tok.strip())
    if not m:
        raise NetlistError(f"not a number: {tok!r}")
    v = float(m.group(1))
    suf = m.group(2).upper()
    for s, f in _SCALE:
        if suf.startswith(s):
            return v * f
    return v


def _logical_lines(text: str):
    """(line number, text) with comments removed and continuations joined."""
    out = []
    for no, raw in enumerate(text.splitlines(), 1):
        line = raw.split(";", 1)[0].rstrip()
        s = line.strip()
        if not s or s.startswith("*"):
            continue
        if s.startswith("+"):
            if not out:
                raise NetlistError(f"line {no}: continuation without a line")
            out[-1] = (out[-1][0], out[-1][1] + " " + s[1:])
        else:
            out.append((no, s))
    return out


def _tokens(s: str):
    # parentheses and '=' become separators; keep function names attached
    s = s.replace("(", " ( ").replace(")", " ) ").replace("=", " = ")
    return s.split()


def _kv(tokens, no):
    """name=value pairs (after the positional part)."""
    out = {}
    i = 0
    while i < len(tokens):
        t = tokens[i]
        if t in ("(", ")"):
            i += 1
            continue
        if i + 1 < len(tokens) and tokens[i + 1] == "=":
            if i + 2 >= len(tokens):
                raise NetlistError(f"line {no}: missing value for {t}")
            out[t.upper()] = value(tokens[i + 2])
            i += 3
        else:
            raise NetlistError(f"line {no}: unexpected {t!r}")
    return out


def _source(tokens, no):
    """Waveform of a V/I element from the tokens after its nodes."""
    toks = [t for t in tokens]
    up = [t.upper() for t in toks]
    for fn, wt in (("PULSE", W_PULSE), ("PWL", W_PWL)):
        if fn in up:
            k = up.index(fn)
            if k + 1 >= len(toks) or toks[k + 1] != "(":
                raise NetlistError(f"line {no}: {fn} needs '(...)'")
            end = toks.index(")", k + 1) if ")" in toks[k + 1:] else -1
            if end < 0:
                raise NetlistError(f"line {no}: unclosed {fn}(")
            vals = [value(t) for t in toks[k + 2:end] if t != ","]
            if wt == W_PULSE:
                if not 2 <= len(vals) <= 7:
                    raise NetlistError(f"line {no}: PULSE takes 2 to 7 values")
                vals = vals + [math.nan] * (7 - len(vals))
            elif len(vals) < 2 or len(vals) % 2:
                raise NetlistError(f"line {no}: PWL takes time/value pairs")
            return wt, vals
    rest = [t for t in toks if t.upper() != "DC"]
    if len(rest) != 1:
        raise NetlistError(f"line {no}: unsupported source specification {' '.join(tokens)!r}")
    return W_DC, [value(rest[0])]


# .OPTIONS settings the path runs with, all of them the reference's defaults
# (TIMEINT RELTOL/ABSTOL: src/TimeIntegrationPKG/N_TIA_TIAParams.C; DEVICE GMIN
# and VOLTLIM, the voltage limiter flag read at N_DEV_MOSFET1.C:2785,2864:
# src/DeviceModelPKG/Core/N_DEV_DeviceOptions.C; LINSOL TYPE=KLU).  A netlist
# may restate them; another value, or any other setting, is refused rather than
# ignored, since the path would not honour it.
_FIXED_OPTIONS = {
    "TIMEINT": {"RELTOL": 1e-3, "ABSTOL": 1e-6},
    "DEVICE": {"GMIN": 1e-12, "VOLTLIM": 1.0},
    "LINSOL": {"TYPE": "KLU"},
}
_METHODS = {"TRAP": "trap", "TRAPEZOIDAL": "trap", "7": "trap", "GEAR": "gear", "8": "gear"}


def _options(tok, no, analysis):
    """One ``.OPTIONS <package> NAME=value ...`` line."""
    if len(tok) < 2:
        raise NetlistError(f"line {no}: .OPTIONS needs a package name")
    pkg = tok[1].upper()
    if pkg not in _FIXED_OPTIONS:
        raise NetlistError(f"line {no}: .OPTIONS {tok[1]} is not supported")
    i = 2
    while i < len(tok):
        if i + 2 >= len(tok) or tok[i + 1] != "=":
            raise NetlistError(f"line {no}: expected NAME=value after .OPTIONS {tok[1]}, got {tok[i]!r}")
        name, raw = tok[i].upper(), tok[i + 2]
        i += 3
        if pkg == "TIMEINT" and name == "METHOD":
            if raw.upper() not in _METHODS:
                raise NetlistError(f"line {no}: integration method {raw!r} is not supported")
            analysis["method"] = _METHODS[raw.upper()]
            continue
        fixed = _FIXED_OPTIONS[pkg]
        if name not in fixed:
            raise NetlistError(f"line {no}: .OPTIONS {pkg} {name} is not supported")
        want = fixed[name]
        ok = raw.upper() == want if isinstance(want, str) else value(raw) == want
        if not ok:
            hint = " (voltage limiting cannot be switched off)" if name == "VOLTLIM" else ""
            raise NetlistError(f"line {no}: .OPTIONS {pkg} {name}={raw} is not supported, "
                               f"only the default {want}{hint}")


def parse_netlist(text: str):
    """Returns (Circuit, analysis) with analysis = {'tran': (tstep, tstop,
    tstart, tmax) or None, 'print': [node names], 'title': str, 'method':
    'trap' | 'gear', 'nodes': {name: index}}."""
    c = Circuit()
    lines = text.splitlines()
    title = lines[0].strip() if lines else ""
    body = "\n".join(["*"] + lines[1:]) if lines else ""    # the first line is the title
    nodes = {"0": 0, "GND": 0}
    models = {}
    pending = []        # MOSFETs waiting for their .model (it may follow them)
    analysis = dict(tran=None, print=[], title=title.lstrip("*").strip(), method="trap")

    def node(name):
        key = name.upper()
        if key not in nodes:
            nodes[key] = c.node(name)
        return nodes[key]

    for no, s in _logical_lines(body):
        tok = _tokens(s)
        head = tok[0].upper()
        if head.startswith("."):
            if head == ".MODEL":
                if len(tok) < 3:
                    raise NetlistError(f"line {no}: .MODEL needs a name and a type")
                typ = tok[2].upper()
                if typ not in ("NMOS", "PMOS"):
                    raise NetlistError(f"line {no}: model type {tok[2]!r} is not supported (NMOS/PMOS)")
                kv = _kv(tok[3:], no)
                level = kv.pop("LEVEL", 1.0)
                if level != 1.0:
                    raise NetlistError(f"line {no}: only LEVEL=1 MOSFETs are supported")
                bad = [k for k in kv if k not in MODEL_PARAMS]
                if bad:
                    raise NetlistError(f"line {no}: unsupported model parameters {bad}")
                models[tok[1].upper()] = (1 if typ == "NMOS" else -1, kv)
            elif head == ".TRAN":
                vals = [value(t) for t in tok[1:]]
                if len(vals) < 2:
                    raise NetlistError(f"line {no}: .TRAN needs tstep and tstop")
                vals = vals + [0.0, 0.0][: 4 - len(vals)]
                analysis["tran"] = tuple(vals[:4])
            elif head == ".PRINT":
                i = 1
                while i < len(tok):
                    if tok[i].upper() == "V" and i + 2 < len(tok) and tok[i + 1] == "(":
                        analysis["print"].append(tok[i + 2])
                        i += 4
                    else:
                        i += 1
            elif head in (".OPTIONS", ".OPTION"):
                _options(tok, no, analysis)
            elif head == ".END":
                break
            else:
                raise NetlistError(f"line {no}: unsupported control line {tok[0]}")
            continue
        kind = head[0]
        if kind in "RC":
            if len(tok) != 4:
                raise NetlistError(f"line {no}: {kind} element takes two nodes and a value")
            p, n = node(tok[1]), node(tok[2])
            (c.add_res if kind == "R" else c.add_cap)(p, n, value(tok[3]))
        elif kind in "VI":
            if len(tok) < 4:
                raise NetlistError(f"line {no}: source needs two nodes and a value")
            p, n = node(tok[1]), node(tok[2])
            wt, vals = _source(tok[3:], no)
            w = c.add_wave(wt, vals)
            (c.add_vsrc if kind == "V" else c.add_isrc)(p, n, w)
        elif kind == "M":
            if len(tok) < 6:
                raise NetlistError(f"line {no}: MOSFET needs d g s b and a model")
            d, g, s_, b = (node(t) for t in tok[1:5])
            for t in tok[6:]:
                if t.upper() in ("OFF", "IC"):
                    # OFF selects the initFix treatment and IC the initial junction
                    # voltages (N_DEV_MOSFET1.C:2785-2796,2864-2867); neither exists here
                    raise NetlistError(f"line {no}: MOSFET {t.upper()} is not supported")
            kv = _kv(tok[6:], no)
            bad = [k for k in kv if k not in INST_PARAMS[:9]]
            if bad:
                raise NetlistError(f"line {no}: unsupported MOSFET parameters {bad}")
            # the device is added now (netlist order); its model index is fixed below
            c.add_mos(d, g, s_, b, -1, inst_card(**kv))
            pending.append((len(c.mos_model) - 1, tok[5].upper(), no))
        else:
            raise NetlistError(f"line {no}: unsupported element {tok[0]!r}")
    index = {}
    for mi, mname, no in pending:
        if mname not in models:
            raise NetlistError(f"line {no}: unknown model {mname!r}")
        if mname not in index:
            typ, kv = models[mname]
            index[mname] = c.add_model(model_card(typ, **kv))
        c.mos_model[mi] = index[mname]
    analysis["nodes"] = {k: v for k, v in nodes.items() if k != "GND"}
    return c, analysis


def write_prn(path, t, columns):
    """Xyce-style .PRN: ``columns`` maps a header (e.g. 'V(OUT)') to values."""
    names = list(columns)
    t = np.asarray(t, dtype=np.float64)
    data = [np.asarray(columns[k], dtype=np.float64) for k in names]
    with open(path, "w") as f:
        f.write("Index".ljust(8) + " " + "TIME".rjust(16) + "".join(" " + k.rjust(16) for k in names) + "\n")
        for i in range(len(t)):
            f.write(f"{i:<8d} {t[i]:16.8e}" + "".join(f" {d[i]:16.8e}" for d in data) + "\n")
        f.write("End of Xyce(TM) Simulation\n")


def read_prn(path):
    """(header list, float array [rows][cols]) of a .PRN file (Xyce's or ours)."""
    rows = []
    with open(path) as f:
        header = f.readline().split()
        for line in f:
            if line.startswith("End of Xyce"):
                break
            if line.strip():
                rows.append([float(v) for v in line.split()])
    return header, np.array(rows)


def run_transient(text: str, prn_path=None, record_max: int = 200_000):
    """Runs a netlist's .TRAN on the GPU (the batched engine with one
    instance: DC operating point, then trapezoid with the reference's step
    control) and returns (t, {'V(node)': values}) for its .PRINT nodes
    (every node when there is none); writes a .PRN file when asked."""
    from . import Batch, host_pattern
    c, an = parse_netlist(text)
    if an["tran"] is None:
        raise NetlistError("no .TRAN line")
    tstep, tstop, tstart, tmax = an["tran"]
    b = Batch([c], tstop, tstep, tmax, record_max=record_max, method=an["method"])
    cnt = b.run_to_end(chunk=512)
    if cnt["failed"]:
        raise RuntimeError("transient failed (DC operating point or time-step control)")
    hp = host_pattern(c)
    t, x = b.points(0, hp["n"])
    gid = hp["node_gid"]
    names = an["print"] or [k for k in an["nodes"] if k != "0"]
    cols = {}
    for nm in names:
        k = an["nodes"].get(nm.upper())
        if k is None:
            raise NetlistError(f".PRINT names an unknown node {nm!r}")
        cols[f"V({nm.upper()})"] = x[:, gid[k]] if k != 0 else np.zeros(len(t))
    keep = t >= tstart
    t = t[keep]
    cols = {k: v[keep] for k, v in cols.items()}
    if prn_path is not None:
        write_prn(prn_path, t, cols)
    return t, cols
