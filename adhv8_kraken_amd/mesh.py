"""Synthetic SW2 problems on the reference's structured flume pattern.

The node and element numbering follows the reference's flume mesh generator
(STAND_ALONE_UTILITIES/flumeBuilder/flumeBuilder.c:65-145): nodes column by
column with j fastest (y decreasing from ymax), each cell split along the
diagonal inode -> inode+npy+1 into two counter-clockwise triangles, and the
boundary edge strings written in the same node order as its EGS cards
(flumeBuilder.c:225-261).  Problems are held as plain numpy arrays in the
reference's own layouts (per-node AoS velocity, 3 DOF per node ordered
u, v, h) so that the HIP library and the CPU oracle ingest identical doubles.

Random fields use SplitMix64 (Steele, Lea, Flood 2014) so every
configuration is reproducible from its seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# boundary-condition flags, reference src/include/define.h:375-413
BCT_OUTFLOW = 4
BCT_PRS_DIR = 7
BCT_VEL_DIR = 8
BCT_PRS_NEU = 9
BCT_VEL_NEU = 10
BCT_VEL_PRS_DIR = 16

TURB_NONE, TURB_EVS, TURB_EEV = 0, 1, 2

EARTH_ROTAT = 7.2722e-5  # reference tools/coriolis.c:2


def manning_to_ks(n: float, kn: float = 1.0, g: float = 9.8) -> float:
    """Manning n -> equivalent sand roughness height (reference
    friction/fr_manningsn_to_rheight.c:16-36), evaluated with libm pow."""
    if n < 1.0e-8:
        return 0.0
    return math.pow(8.25 * math.sqrt(g) * n / kn, 6.0)


def coriolis_speed(lat_deg: float) -> float:
    """reference tools/coriolis.c:3-5"""
    return math.sin((math.pi / 180.0) * lat_deg) * 2.0 * EARTH_ROTAT


class SplitMix64:
    """SplitMix64 stream; uniform() returns doubles in [0, 1) from the top 53 bits."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def uniform_array(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        # vectorised SplitMix64 for large per-element fields
        idx = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + idx * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            self.state = (self.state + n * 0x9E3779B97F4A7C15) & self.MASK
        u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u


@dataclass
class SW2Problem:
    """One SW2 instance: mesh, strings, materials, model constants, state."""

    name: str
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    e2d: np.ndarray            # int32 [3E]
    e2d_mat: np.ndarray        # int32 [E]
    e2d_ks: np.ndarray         # float64 [E]
    e2d_src_flag: np.ndarray   # int32 [E]
    e2d_src: np.ndarray        # float64 [E]
    e1d: np.ndarray            # int32 [2B]
    e1d_adj: np.ndarray        # int32 [B]
    e1d_bc: np.ndarray         # int32 [B]
    e1d_val: np.ndarray        # float64 [B]
    node_bc: np.ndarray        # int32 [N]
    node_bc_val: np.ndarray    # float64 [3N]
    mat_turb: np.ndarray
    mat_eev_mode: np.ndarray
    mat_eev_coef: np.ndarray
    mat_ev: np.ndarray
    mat_cor_speed: np.ndarray
    g: float = 9.8
    dt: float = 10.0
    tau_temporal: float = 0.0
    tau_pg: float = 0.5
    drying_lower: float = 0.0
    drying_upper: float = 0.0
    viscosity: float = 1.0e-6
    wind: int = 0              # winds on (MP WND / SERIES WIND)
    mat_windatt: np.ndarray = None   # [M] wind attenuation, None = 1.0
    density: float = 1000.0    # MP RHO
    head0: np.ndarray = None   # [N]
    vel0: np.ndarray = None    # [2N] AoS
    nsteps: int = 1
    meta: dict = field(default_factory=dict)

    @property
    def nnodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def nelems2d(self) -> int:
        return int(self.e2d.shape[0] // 3)

    @property
    def nelems1d(self) -> int:
        return int(self.e1d.shape[0] // 2)

    @property
    def nmat(self) -> int:
        return int(self.mat_turb.shape[0])


def flume(npx: int, npy: int, xmax: float, ymax: float, *, zfun=None, n_manning=0.025,
          n_elem=None, inflow_q=0.05, tail_depth=1.0, h0=1.0, dt=10.0, name="flume",
          viscosity=1.0e-6, tau_pg=0.5, src=None, walls_1d=False, nsteps=1) -> SW2Problem:
    """Structured channel [0,xmax] x [0,ymax] with npx x npy nodes.

    West edge: NB VEL (series value inflow_q m^2/s; the residual uses
    flux = -value, reference fe_sw2_boundary_resid.c:212-215).  East edge:
    NB OTW with tailwater elevation z(xmax, .) + tail_depth.
    """
    dx = xmax / (npx - 1)
    dy = ymax / (npy - 1)
    i = np.repeat(np.arange(npx), npy)
    j = np.tile(np.arange(npy), npx)
    x = 0.0 + i * dx
    y = ymax - j * dy
    if zfun is None:
        z = -(0.0 + 1.0e-4 * x + 0.0 * x * x)
    else:
        z = zfun(x, y)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)

    ci = np.repeat(np.arange(npx - 1), npy - 1)
    cj = np.tile(np.arange(npy - 1), npx - 1)
    inode = ci * npy + cj
    t1 = np.stack([inode, inode + 1, inode + npy + 1], axis=1)
    t2 = np.stack([inode, inode + npy + 1, inode + npy], axis=1)
    e2d = np.empty((2 * inode.shape[0], 3), dtype=np.int32)
    e2d[0::2] = t1
    e2d[1::2] = t2
    E = e2d.shape[0]

    # boundary edges: west (NB VEL) then east (NB OTW), flumeBuilder.c:225-252 order
    jj = np.arange(npy - 1)
    west = np.stack([jj, jj + 1], axis=1)
    west_adj = 2 * (0 * (npy - 1) + jj) + 0
    base = npy * (npx - 1)
    east = np.stack([base + jj, base + jj + 1], axis=1)
    east_adj = 2 * ((npx - 2) * (npy - 1) + jj) + 1
    e1d = [west, east]
    e1d_adj = [west_adj, east_adj]
    zt = z[base + jj]
    tail = float(z[base]) + tail_depth
    e1d_bc = [np.full(npy - 1, BCT_VEL_NEU), np.full(npy - 1, BCT_PRS_NEU)]
    e1d_val = [np.full(npy - 1, inflow_q), np.full(npy - 1, tail)]
    if walls_1d:
        ii = np.arange(npx - 1)
        south = np.stack([npy * ii + (npy - 1), npy * (ii + 1) + (npy - 1)], axis=1)
        south_adj = 2 * (ii * (npy - 1) + (npy - 2)) + 0
        north = np.stack([npy * ii, npy * (ii + 1)], axis=1)
        north_adj = 2 * (ii * (npy - 1) + 0) + 1
        e1d += [south, north]
        e1d_adj += [south_adj, north_adj]
        e1d_bc += [np.full(npx - 1, BCT_VEL_NEU)] * 2
        e1d_val += [np.zeros(npx - 1)] * 2
    e1d = np.ascontiguousarray(np.concatenate(e1d).astype(np.int32).ravel())
    e1d_adj = np.ascontiguousarray(np.concatenate(e1d_adj).astype(np.int32))
    e1d_bc = np.ascontiguousarray(np.concatenate(e1d_bc).astype(np.int32))
    e1d_val = np.ascontiguousarray(np.concatenate(e1d_val).astype(np.float64))
    del zt

    if n_elem is None:
        ks = np.full(E, manning_to_ks(n_manning), dtype=np.float64)
    else:
        ks = np.array([manning_to_ks(float(v)) for v in n_elem], dtype=np.float64) if E < 5000 else \
            np.power(8.25 * math.sqrt(9.8) * np.asarray(n_elem, dtype=np.float64) / 1.0, 6.0)
    N = x.shape[0]
    src_flag = np.zeros(E, dtype=np.int32)
    srcv = np.zeros(E, dtype=np.float64)
    if src is not None:
        src_flag[:] = 1
        srcv[:] = src
    head0 = np.full(N, h0, dtype=np.float64) if np.isscalar(h0) else np.asarray(h0, dtype=np.float64)
    return SW2Problem(
        name=name, x=x, y=y, z=z, e2d=np.ascontiguousarray(e2d.ravel()),
        e2d_mat=np.zeros(E, dtype=np.int32), e2d_ks=ks, e2d_src_flag=src_flag, e2d_src=srcv,
        e1d=e1d, e1d_adj=e1d_adj, e1d_bc=e1d_bc, e1d_val=e1d_val,
        node_bc=np.zeros(N, dtype=np.int32), node_bc_val=np.zeros(3 * N, dtype=np.float64),
        mat_turb=np.array([TURB_EVS], dtype=np.int32), mat_eev_mode=np.array([0], dtype=np.int32),
        mat_eev_coef=np.array([0.0]), mat_ev=np.zeros(3), mat_cor_speed=np.array([0.0]),
        dt=dt, viscosity=viscosity, tau_pg=tau_pg,
        head0=np.ascontiguousarray(head0), vel0=np.zeros(2 * N, dtype=np.float64), nsteps=nsteps,
        meta=dict(npx=npx, npy=npy, xmax=xmax, ymax=ymax, dx=dx, dy=dy),
    )


def channel_config1() -> SW2Problem:
    """400 m x 200 m, 40 x 20 cells (861 nodes, 1600 triangles); bed slope 1e-4,
    Manning n 0.025, h0 = 1 m, 10 m^3/s inflow over 200 m, 1 m tailwater, dt 10 s."""
    return flume(41, 21, 400.0, 200.0, n_manning=0.025, inflow_q=0.05, tail_depth=1.0, h0=1.0,
                 dt=10.0, name="channel-40x20")


def bumpy_channel(ncx: int, ncy: int, seed: int = 2, nbumps: int = 20, dt: float = 5.0,
                  nsteps: int = 10, dxy: float = 10.0, name=None) -> SW2Problem:
    """Channel of ncx x ncy cells (dx = dy = 10 m): bed slope 1e-4 plus seeded
    Gaussian bumps (amplitude U(0,0.3) m, sigma U(5,30) m), Manning n
    U(0.02,0.035) per element, h0 = 1 m, dt = 5 s."""
    rng = SplitMix64(seed)
    L, W = ncx * dxy, ncy * dxy
    bumps = []
    for _ in range(nbumps):
        bumps.append((rng.uniform(0, L), rng.uniform(0, W), rng.uniform(0.0, 0.3), rng.uniform(5.0, 30.0)))

    def zfun(x, y):
        z = -(0.0 + 1.0e-4 * x + 0.0 * x * x)
        for (bx, by, a, s) in bumps:
            z = z + a * np.exp(-((x - bx) ** 2 + (y - by) ** 2) / (2.0 * s * s))
        return z

    E = 2 * ncx * ncy
    n_elem = rng.uniform_array(E, 0.02, 0.035)
    p = flume(ncx + 1, ncy + 1, L, W, zfun=zfun, n_elem=n_elem, inflow_q=0.05, tail_depth=1.0, h0=1.0,
              dt=dt, name=name or f"bumpy-{ncx}x{ncy}", nsteps=nsteps)
    p.meta.update(seed=seed, bumps=bumps)
    return p


def log_uniform_sizes(n: int, lo: float, hi: float, seed: int) -> list:
    """n triangle counts drawn log-uniform in [lo, hi]."""
    rng = SplitMix64(seed)
    return [int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(n)]


def _grid_for(tris: int):
    # 2:1 channel of ncx x ncy cells, two triangles per cell
    ncy = max(2, int(round(math.sqrt(tris / 4.0))))
    return 2 * ncy, ncy


def channel_instance(tris: int, seed: int, dt: float = 5.0, nbumps: int = 20) -> SW2Problem:
    """One channel of the batched ensemble: about `tris` triangles (2:1 cells
    grid, dx = dy = 10 m), bed slope 1e-4 plus seeded Gaussian bumps
    (amplitude U(0,0.3) m, sigma U(5,30) m), one Manning n U(0.015,0.05) and
    an inflow U(5,50) m^3/s for the instance, h0 = 1 m, 1 m tailwater."""
    ncx, ncy = _grid_for(tris)
    rng = SplitMix64(seed)
    L, W = ncx * 10.0, ncy * 10.0
    n = rng.uniform(0.015, 0.05)
    Q = rng.uniform(5.0, 50.0)
    bumps = [(rng.uniform(0, L), rng.uniform(0, W), rng.uniform(0.0, 0.3), rng.uniform(5.0, 30.0))
             for _ in range(nbumps)]

    def zfun(x, y):
        z = -(1.0e-4 * x)
        for (bx, by, a, sg) in bumps:
            z = z + a * np.exp(-((x - bx) ** 2 + (y - by) ** 2) / (2.0 * sg * sg))
        return z

    p = flume(ncx + 1, ncy + 1, L, W, zfun=zfun, n_manning=n, inflow_q=Q / W, tail_depth=1.0, h0=1.0, dt=dt,
              name=f"channel-{ncx}x{ncy}-s{seed}", nsteps=20)
    p.meta.update(seed=seed, manning=n, inflow=Q)
    return p


def channel_ensemble(n: int = 256, lo: int = 5000, hi: int = 200000, seed: int = 4) -> list:
    """Batched channel ensemble: n instances, triangle counts log-uniform in
    [lo, hi], instance k seeded seed * 1000 + k."""
    return [channel_instance(t, seed * 1000 + k) for k, t in enumerate(log_uniform_sizes(n, lo, hi, seed))]


def floodplain_ensemble(n: int = 64, lo: int = 50000, hi: int = 500000, seed: int = 5, dt: float = 2.0) -> list:
    """Batched wet/dry ensemble: n sloped floodplains (cross-slope 1e-3, 40 %
    of nodes dry, rainfall 1e-5 m/s), triangle counts log-uniform in
    [lo, hi], instance k seeded seed * 1000 + k (Manning n and inflow drawn
    per instance)."""
    out = []
    for k, t in enumerate(log_uniform_sizes(n, lo, hi, seed)):
        ncx, ncy = _grid_for(t)
        out.append(floodplain(ncx, ncy, seed=seed * 1000 + k, dt=dt))
    return out


def config_problem(idx: int) -> SW2Problem:
    if idx == 0:
        return channel_config1()
    if idx == 1:
        return bumpy_channel(1000, 500)
    if idx == 2:
        return bumpy_channel(2000, 1000, nsteps=10)
    raise ValueError(idx)


def floodplain(ncx: int, ncy: int, seed: int = 5, dry_frac: float = 0.4, rain: float = 1.0e-5, dt: float = 2.0,
               nsteps: int = 50, n_lo: float = 0.015, n_hi: float = 0.05, inflow_q=None, name=None) -> SW2Problem:
    """Sloped floodplain with wetting and drying: bed slope 1e-4 along x and
    cross-slope 1e-3 along y, initial depth 0 on the highest `dry_frac` of
    nodes (the still-water level sits at the matching bed quantile), rainfall
    `rain` m/s on every element (NB SOURCE), DTL 0.1 m, EVS 0.075 m^2/s."""
    rng = SplitMix64(seed)
    L, W = ncx * 10.0, ncy * 10.0

    def zfun(x, y):
        return -(1.0e-4 * x) + 1.0e-3 * y

    n = rng.uniform(n_lo, n_hi)
    q = rng.uniform(5.0, 50.0) / W if inflow_q is None else inflow_q
    p = flume(ncx + 1, ncy + 1, L, W, zfun=zfun, n_manning=n, inflow_q=q, tail_depth=0.5, h0=0.0, dt=dt,
              name=name or f"floodplain-{ncx}x{ncy}", src=rain, nsteps=nsteps)
    wse0 = float(np.quantile(p.z, 1.0 - dry_frac))
    p.head0 = np.ascontiguousarray(np.maximum(wse0 - p.z, 0.0))
    base = (ncx + 1 - 1) * (ncy + 1)
    p.e1d_val[p.e1d_bc == BCT_PRS_NEU] = wse0
    p.drying_lower = 0.1
    p.drying_upper = 0.1
    p.mat_ev[:] = 0.075
    p.meta.update(seed=seed, wse0=wse0, manning=n, inflow=q)
    return p
