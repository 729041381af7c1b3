"""Readers for the reference's SW2 input decks: `.2dm` mesh, `.bc` cards and
`.hot` initial conditions, turned into an `SW2Problem` plus the run controls
(Newton tolerances, time series, time span, test-case flag).

Only the card subset that the SW2 Newton path consumes is interpreted; any
other physics card that would change the equations (wind, waves, baroclinic
terms, transport, Dirichlet strings, ...) raises `DeckError` rather than
being silently dropped.  Formats follow:

* `.2dm`: ND id x y z / TRI id n1 n2 n3 mat (src/initio/read_geo.c:50),
* `.bc`: cards as parsed by src/initio/read_bc_{OP,IP,MP,FR,NB,SERIES,TC,TEST,STRINGS}.c,
* `.hot`: SMS DATASET blocks IOH / IPH / IV / IPV (src/structs/ssw.c:164-545).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import mesh as _mesh


class DeckError(ValueError):
    pass


class _XY:
    def __init__(self, xyz):
        self.x, self.y = xyz[:, 0], xyz[:, 1]


NOT_QUITE_SMALL = 1.1920928955078125e-07  # FLT_EPSILON, src/include/define.h:28 (IOH zero depths become -NOT_QUITE_SMALL, ssw.c:221-224)


@dataclass
class Series:
    """Piecewise-linear time series (src/structs/sseries.c)."""

    kind: str
    sid: int
    times: np.ndarray
    values: np.ndarray  # [npoints, nvalues]

    def average(self, t0: float, t1: float, col: int = 0) -> float:
        """Mean value over [t0, t1] of the linear interpolant, the quantity
        sseries_setall_ts_valuesAVG hands to the boundary terms
        (src/structs/sseries.c:618)."""
        t, v = self.times, self.values[:, col]
        if t1 <= t0:
            return float(np.interp(t0, t, v))
        pts = np.unique(np.concatenate([[t0, t1], t[(t > t0) & (t < t1)]]))
        vals = np.interp(pts, t, v)
        return float(np.sum(0.5 * (vals[1:] + vals[:-1]) * np.diff(pts)) / (t1 - t0))

    def value_at(self, t: float, col: int = 0) -> float:
        """tc_eval_series at t (src/tools/tc_eval_series.c:5-41) over the
        interval sseries_get_interval picks (src/structs/sseries.c:693-719)."""
        small = 2.220446049250313e-16
        ilast = len(self.times) - 1
        i = 0
        while i < ilast and t >= self.times[i] * (1.0 - small):
            i += 1
        i -= 1
        if i == ilast:
            return float(self.values[ilast, col])
        ltime, rtime = float(self.times[i]), float(self.times[i + 1])
        lv, rv = float(self.values[i, col]), float(self.values[i + 1, col])
        theta = (rtime - t) / (rtime - ltime)
        return theta * lv + (1.0 - theta) * rv

    def is_constant(self, col: int = 0) -> bool:
        return bool(np.all(self.values[:, col] == self.values[0, col]))


@dataclass
class Deck:
    problem: _mesh.SW2Problem
    tol_nonlin: float
    inc_nonlin: float
    max_nonlin_it: int
    max_lin_it: int
    nblock: int
    prec: int
    t0: float
    tf: float
    dt_series: Series
    series: dict
    edge_series: np.ndarray      # series id per 1D element (value source of e1d_val)
    test_case: str | None
    density: float
    cards: list = field(default_factory=list)

    def dt_at(self, t: float) -> float:
        return float(np.interp(t, self.dt_series.times, self.dt_series.values[:, 0]))


# ------------------------------------------------------------------ .2dm
def read_2dm(path):
    nodes, tris = {}, []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "ND":
                nodes[int(tok[1])] = (float(tok[2]), float(tok[3]), float(tok[4]))
            elif tok[0] == "TRI":
                tris.append((int(tok[1]), int(tok[2]), int(tok[3]), int(tok[4]), int(tok[5])))
            elif tok[0] in ("E3T",):
                tris.append((int(tok[1]), int(tok[2]), int(tok[3]), int(tok[4]), int(tok[5])))
            elif tok[0] in ("MESH2D",):
                continue
            elif tok[0] in ("E4Q", "QUAD", "E6T", "E8Q", "E9Q", "TET", "PRISM"):
                raise DeckError(f"{path}: element type {tok[0]} is not a linear triangle")
    n = len(nodes)
    if sorted(nodes) != list(range(1, n + 1)):
        raise DeckError(f"{path}: node ids are not 1..{n}")
    xyz = np.array([nodes[i] for i in range(1, n + 1)], dtype=np.float64)
    tris.sort()
    if [t[0] for t in tris] != list(range(1, len(tris) + 1)):
        raise DeckError(f"{path}: element ids are not 1..{len(tris)}")
    e2d = np.array([[t[1] - 1, t[2] - 1, t[3] - 1] for t in tris], dtype=np.int32)
    mat = np.array([t[4] - 1 for t in tris], dtype=np.int32)
    return xyz, e2d, mat


# ------------------------------------------------------------------ .hot
def read_hot(path, nnodes):
    """Returns {name: array}; scalar sets are [N], vector sets [N, k]."""
    out = {}
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        tok = lines[i].split()
        if tok and tok[0] in ("BEGSCL", "BEGVEC"):
            vec = tok[0] == "BEGVEC"
            name, nd = None, None
            i += 1
            while i < len(lines):
                t = lines[i].split()
                if t and t[0] == "ND":
                    nd = int(t[1])
                elif t and t[0] == "NAME":
                    name = t[1].strip('"').upper()
                elif t and t[0] == "TS":
                    i += 1
                    break
                i += 1
            if nd != nnodes:
                raise DeckError(f"{path}: dataset {name} has {nd} values for {nnodes} nodes")
            rows = []
            while i < len(lines) and len(rows) < nd:
                t = lines[i].split()
                if t:
                    rows.append([float(v) for v in t])
                i += 1
            arr = np.array(rows, dtype=np.float64)
            out[name] = arr if vec else arr[:, 0]
            continue
        i += 1
    return out


# ------------------------------------------------------------------ .bc
# card groups the reference's .bc parser dispatches on (src/initio/read_bc.c)
_CARD_GROUPS = {"OP", "IP", "MP", "FR", "NB", "DB", "EGS", "NDS", "MTS", "SERIES", "TC", "TEST", "SOUT", "FOUT", "END",
                "CN",
                "PC", "NDP", "FLX", "FST", "WER", "WRS", "FLP", "FGT", "SLS", "SDR", "BR", "OFF", "ON", "MID"}
_IGNORED = {("OP", "SW2"), ("OP", "TRN"), ("OP", "INC"), ("MP", "ML"), ("MP", "SRT"), ("SOUT", None),
            ("SERIES", "AWRITE"), ("SERIES", "OUTPUT"), ("PC", None), ("FOUT", None)}


def _series_card(tok, lines, i):
    kind = tok[1]
    sid = int(tok[2])
    npts = int(tok[3])
    rows = []
    j = i + 1
    while len(rows) < npts:
        t = lines[j].split("#")[0].split("!")[0].split()
        j += 1
        if not t:
            continue
        rows.append([float(v) for v in t])
    if kind in ("AWRITE", "OUTPUT"):
        return None, j
    r = np.array(rows, dtype=np.float64)
    return Series(kind, sid, r[:, 0].copy(), r[:, 1:].copy()), j


def read_deck(stem):
    """Builds the SW2 problem of <stem>.2dm / <stem>.bc / <stem>.hot."""
    xyz, e2d, mat = read_2dm(stem + ".2dm")
    N, E = xyz.shape[0], e2d.shape[0]
    with open(stem + ".bc") as f:
        lines = f.read().splitlines()

    op = dict(nblock=1, prec=1, tau_pg=0.5, tau_temporal=0.0)
    ip = dict(tol_nonlin=None, inc_nonlin=None, max_nonlin_it=None, max_lin_it=500)
    mp = dict(g=9.8, density=1000.0, viscosity=1.0e-6, muc=1.0, dtl=0.0, wind=False)
    wind_flag = {}   # material -> wind stress law (MP WND STR)
    windatt = {}     # material -> attenuation (MP WND ATT)
    mat_string = {}
    ev = {}
    eev = {}      # material -> (coef, mode)
    cor = {}      # material -> latitude in degrees
    manning = {}
    series = {}
    nb = {}       # string -> (kind, series id)
    edges = []    # (n1, n2, string)
    t0 = tf = None
    test_case = None
    test_args = []
    cards = []
    i = 0
    while i < len(lines):
        raw = lines[i].split("#")[0]
        tok = raw.split()
        if not tok or tok[0].startswith("!") or tok[0].startswith("#"):
            i += 1
            continue
        c0 = tok[0].upper()
        c1 = tok[1].upper() if len(tok) > 1 else None
        cards.append(" ".join(tok))
        if c0 == "END":
            break
        if c0 == "SERIES":
            s, i = _series_card([c0, c1] + tok[2:], lines, i)
            if s is not None:
                series[s.sid] = s
            continue
        if (c0, c1) in _IGNORED or (c0, None) in _IGNORED:
            pass
        elif c0 == "OP":
            if c1 == "BLK":
                op["nblock"] = int(tok[2])
            elif c1 == "PRE":
                op["prec"] = int(tok[2])
            elif c1 == "TPG":
                op["tau_pg"] = float(tok[2])
            elif c1 == "TEM":
                op["tau_temporal"] = float(tok[2])
            else:
                raise DeckError(f"unsupported card: {raw.strip()}")
        elif c0 == "IP":
            key = {"NTL": "tol_nonlin", "ITL": "inc_nonlin", "NIT": "max_nonlin_it", "MIT": None}.get(c1, "?")
            if key == "?":
                raise DeckError(f"unsupported card: {raw.strip()}")
            if key == "max_nonlin_it":
                ip[key] = int(tok[2])
            elif key:
                ip[key] = float(tok[2])
            # IP MIT is parsed into the model but the single-model solve keeps
            # its own limit of 500 (src/solver/solv_initialize.c:54)
        elif c0 == "MTS":
            mat_string[int(tok[1]) - 1] = int(tok[2])
        elif c0 == "MP":
            if c1 == "EVS":
                ev[int(tok[2]) - 1] = (float(tok[3]), float(tok[4]), float(tok[5]))
                eev.pop(int(tok[2]) - 1, None)
            elif c1 == "EEV":
                mode = int(tok[4])
                if mode not in (1, 2, 3):
                    raise DeckError(f"EEV mode must be 1, 2 or 3: {raw.strip()}")
                eev[int(tok[2]) - 1] = (float(tok[3]), mode)
                ev.pop(int(tok[2]) - 1, None)
            elif c1 == "COR":
                cor[int(tok[2]) - 1] = float(tok[3])
            elif c1 == "MUC":
                mp["muc"] = float(tok[2])
            elif c1 == "MU":
                mp["viscosity"] = float(tok[2])
            elif c1 == "RHO":
                mp["density"] = float(tok[2])
            elif c1 == "G":
                mp["g"] = float(tok[2])
            elif c1 == "DTL":
                mp["dtl"] = max(float(tok[2]), 0.0)
            elif c1 == "WND":
                # read_bc_MP.c:241-276: any MP WND card switches winds on
                mp["wind"] = True
                c2 = tok[2].upper() if len(tok) > 2 else None
                if c2 == "STR":
                    wind_flag[int(tok[3]) - 1] = int(tok[4])
                elif c2 == "ATT":
                    windatt[int(tok[3]) - 1] = float(tok[4])
                else:
                    raise DeckError(f"MP WND must be followed by STR or ATT: {raw.strip()}")
            else:
                raise DeckError(f"unsupported card: {raw.strip()}")
        elif c0 == "FR":
            if c1 != "MNG":
                raise DeckError(f"unsupported card: {raw.strip()}")
            manning[int(tok[2])] = float(tok[3])
        elif c0 == "NB":
            if c1 not in ("OTW", "VEL", "SOURCE", "OUT"):
                raise DeckError(f"unsupported card: {raw.strip()}")
            nb[int(tok[2])] = (c1, int(tok[3]) if len(tok) > 3 else None)
        elif c0 == "EGS":
            edges.append((int(tok[1]) - 1, int(tok[2]) - 1, int(tok[3])))
        elif c0 == "TC":
            if c1 == "T0":
                t0 = float(tok[2])
            elif c1 == "TF":
                tf = float(tok[2])
            else:
                raise DeckError(f"unsupported card: {raw.strip()}")
        elif c0 == "TEST":
            test_case = c1
            test_args = [float(v) for v in tok[2:]]
        elif c0 in _CARD_GROUPS:
            raise DeckError(f"unsupported card: {raw.strip()}")
        # any other first word is not a card and the reference's parser skips it
        i += 1

    # materials: one per string id that MTS maps (material m -> string s)
    nmat = int(mat.max()) + 1
    for m in range(nmat):
        mat_string.setdefault(m, m + 1)
    ks = np.zeros(E, dtype=np.float64)
    src_flag = np.zeros(E, dtype=np.int32)
    src = np.zeros(E, dtype=np.float64)
    for m in range(nmat):
        s = mat_string[m]
        sel = mat == m
        if s in manning:
            ks[sel] = _mesh.manning_to_ks(manning[s], mp["muc"], mp["g"])
        if s in nb and nb[s][0] == "SOURCE":
            ser = series[nb[s][1]]
            if not ser.is_constant():
                raise DeckError("time-varying NB SOURCE series are not supported")
            src_flag[sel] = 1
            src[sel] = ser.values[0, 0]

    mat_turb = np.array([_mesh.TURB_EVS if m in ev else _mesh.TURB_EEV if m in eev else _mesh.TURB_NONE
                         for m in range(nmat)], dtype=np.int32)
    mat_ev = np.array([c for m in range(nmat) for c in ev.get(m, (0.0, 0.0, 0.0))], dtype=np.float64)
    mat_eev_coef = np.array([eev[m][0] if m in eev else 0.0 for m in range(nmat)], dtype=np.float64)
    mat_eev_mode = np.array([eev[m][1] if m in eev else 0 for m in range(nmat)], dtype=np.int32)
    # angular speed per material (src/tools/coriolis.c:3-5)
    mat_cor_speed = np.array([_mesh.coriolis_speed(cor[m]) if m in cor else 0.0 for m in range(nmat)],
                             dtype=np.float64)

    # 1D boundary elements from EGS cards, each with its adjacent triangle
    owner = {}
    for ie, t in enumerate(e2d):
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            owner.setdefault((min(a, b), max(a, b)), ie)
    e1d, e1d_adj, e1d_bc, e1d_val, e1d_ser = [], [], [], [], []
    kinds = {"OTW": _mesh.BCT_PRS_NEU, "VEL": _mesh.BCT_VEL_NEU, "OUT": _mesh.BCT_OUTFLOW}
    for (a, b, s) in edges:
        if s not in nb or nb[s][0] == "SOURCE":
            continue  # an edge string without a natural condition is a wall
        kind, sid = nb[s]
        key = (min(a, b), max(a, b))
        if key not in owner:
            raise DeckError(f"EGS {a + 1} {b + 1}: not an edge of the mesh")
        e1d.append((a, b))
        e1d_adj.append(owner[key])
        e1d_bc.append(kinds[kind])
        e1d_ser.append(sid if sid is not None else -1)
        e1d_val.append(0.0)
    B = len(e1d)

    # hotstart (src/structs/ssw.c:164-545): depths default to 0 when the file
    # holds no IOH; zero depths read from IOH become -NOT_QUITE_SMALL; without
    # IPH/IPV the old and older states copy the current one, with them the
    # older state still copies the current one (ssw.c:510-542)
    hot = read_hot(stem + ".hot", N)
    head = np.zeros(N)
    if "IOH" in hot:
        head = hot["IOH"].copy()
        head[head == 0.0] = -NOT_QUITE_SMALL
    vel = np.zeros(2 * N)
    if "IV" in hot:
        vel[0::2] = hot["IV"][:, 0]
        vel[1::2] = hot["IV"][:, 1]
    old_head = hot["IPH"].copy() if "IPH" in hot else head.copy()
    older_head = head.copy()
    old_vel = vel.copy()
    if "IPV" in hot:
        old_vel[0::2] = hot["IPV"][:, 0]
        old_vel[1::2] = hot["IPV"][:, 1]
    older_vel = vel.copy()
    if test_case == "DAM":
        # testcase_prep_dam2d runs after the hotstart (src/main/adh_func.c:294-298)
        from . import qa
        head = qa.dam2d_initial_head(_XY(xyz),
                                     *test_args[:3])

    dt_series = next((s for s in series.values() if s.kind == "DT"), None)
    if dt_series is None or t0 is None or tf is None:
        raise DeckError("deck needs SERIES DT, TC T0 and TC TF")

    # wind stations (SERIES WIND, read_bc_SERIES.c:101-106): one station
    # weighs 1 on every node (station_node_contrib.c:26-32)
    wind_series = [s for s in series.values() if s.kind == "WIND"]
    if wind_series:
        mp["wind"] = True
    if len(wind_series) > 1:
        raise DeckError("more than one wind station is not supported")
    if any(f != 0 for f in wind_flag.values()):
        raise DeckError("only wind stress law 0 (MP WND STR <mat> 0) is supported")

    p = _mesh.SW2Problem(
        name=stem.rsplit("/", 1)[-1],
        x=np.ascontiguousarray(xyz[:, 0]), y=np.ascontiguousarray(xyz[:, 1]), z=np.ascontiguousarray(xyz[:, 2]),
        e2d=np.ascontiguousarray(e2d.ravel()), e2d_mat=mat, e2d_ks=ks, e2d_src_flag=src_flag, e2d_src=src,
        e1d=np.ascontiguousarray(np.array(e1d, dtype=np.int32).reshape(-1)),
        e1d_adj=np.array(e1d_adj, dtype=np.int32), e1d_bc=np.array(e1d_bc, dtype=np.int32),
        e1d_val=np.array(e1d_val, dtype=np.float64),
        node_bc=np.zeros(N, dtype=np.int32), node_bc_val=np.zeros(3 * N, dtype=np.float64),
        mat_turb=mat_turb, mat_eev_mode=mat_eev_mode, mat_eev_coef=mat_eev_coef,
        mat_ev=mat_ev, mat_cor_speed=mat_cor_speed,
        g=mp["g"], dt=dt_series.values[0, 0], tau_temporal=op["tau_temporal"], tau_pg=op["tau_pg"],
        drying_lower=mp["dtl"], drying_upper=mp["dtl"], viscosity=mp["viscosity"],
        head0=head, vel0=vel, nsteps=0,
        wind=int(mp["wind"]), density=mp["density"],
        mat_windatt=np.array([windatt.get(m, 1.0) for m in range(nmat)], dtype=np.float64),
        meta=dict(old_head=old_head, old_vel=old_vel, older_head=older_head, older_vel=older_vel, deck=stem,
                  latitude=[cor.get(m, 0.0) for m in range(nmat)]),
    )
    deck = Deck(problem=p,
                tol_nonlin=ip["tol_nonlin"] if ip["tol_nonlin"] is not None else -9999999.9,
                inc_nonlin=ip["inc_nonlin"] if ip["inc_nonlin"] is not None else -9999999.9,
                max_nonlin_it=ip["max_nonlin_it"] if ip["max_nonlin_it"] is not None else 10,
                max_lin_it=ip["max_lin_it"], nblock=op["nblock"], prec=op["prec"], t0=t0, tf=tf,
                dt_series=dt_series, series=series, edge_series=np.array(e1d_ser, dtype=np.int32),
                test_case=test_case, density=mp["density"], cards=cards)
    deck.test_args = test_args
    set_edge_values(deck, t0, t0 + p.dt)
    return deck


def wind_values(deck: Deck, t: float) -> np.ndarray:
    """Node wind values (x, y per node) at time t, as meteor_update sets them
    from the wind station series (src/meteor/meteor_update.c:30-50,
    meteor_update_series.c:24-54: the station value times its node weight,
    here 1, added to 0)."""
    N = deck.problem.nnodes
    out = np.zeros(2 * N)
    for s in (s for s in deck.series.values() if s.kind == "WIND"):
        if deck.density > 0.0:
            out[0::2] += s.value_at(t, 0) * 1.0
            out[1::2] += s.value_at(t, 1) * 1.0
    return out


def set_edge_values(deck: Deck, t0: float, t1: float):
    """Boundary values for the step [t0, t1]: series averages per edge."""
    p = deck.problem
    for b in range(p.nelems1d):
        sid = int(deck.edge_series[b])
        if sid >= 0:
            p.e1d_val[b] = deck.series[sid].average(t0, t1)

