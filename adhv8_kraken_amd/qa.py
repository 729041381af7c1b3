"""The reference's SW2 test-case error reports (src/initio/test_cases.c),
restated so that a run of this library can be compared with the figures in
the reference's committed qa/sw2/*/tri/error.out.last files.

Each writer takes the mesh and the state after a completed time step; the
sums run over nodes in index order, as the reference's loops do.
"""
from __future__ import annotations

import math

import numpy as np

from . import mesh as _mesh


def error_nb(p, head, vel):
    """write_testcase_error_nb (test_cases.c:341-388): mean |u - a|, |v - a|
    with a = 0.1/sqrt(2) m/s, and mean |h - 1|."""
    a = 0.1 / math.sqrt(2.0)
    n = p.nnodes
    ex = ey = eh = 0.0
    for i in range(n):
        ex += abs(vel[2 * i] - a)
        ey += abs(vel[2 * i + 1] - a)
        eh += abs(head[i] - 1.0)
    return np.array([ex / n, ey / n, eh / n])


def slice_nodes_x(p, node1, node2):
    """tl_find_nodes_on_sliceX (src/tools/tl_find_nodes_on_slice.c:8-99):
    nodes on the line through node1 and node2 (0-based), bubble-sorted by y."""
    x0, x1, y0, y1 = p.x[node1], p.x[node2], p.y[node1], p.y[node2]
    ids = []
    for i in range(p.nnodes):
        x, y = p.x[i], p.y[i]
        if abs(x0 - x1) > 1e-6:
            r = y - (y0 * ((x - x1) / (x0 - x1)) + y1 * ((x - x0) / (x1 - x0)))
            if abs(r) < 1e-6:
                ids.append(i)
        elif abs(x - x0) < 1e-6:
            ids.append(i)
    n = len(ids)
    for _ in range(n):
        for j in range(n - 1):
            if p.y[ids[j]] > p.y[ids[j + 1]]:
                ids[j], ids[j + 1] = ids[j + 1], ids[j]
    return ids


def error_xcoriolis(p, head, vel, node1, node2, latitude):
    """write_testcase_error_xcoriolis (test_cases.c:449-500): mean over the
    transect of |dh/dy + f u / g| (geostrophic balance), the loop skipping
    the last node pair as the reference's does, divided by (count - 1)."""
    rads = (3.141592653589793 / 180.0) * latitude
    f = 2.0 * _mesh.EARTH_ROTAT * math.sin(rads)
    ids = slice_nodes_x(p, node1, node2)
    n = len(ids)
    err = 0.0
    for i in range(n - 2):
        a, b = ids[i + 1], ids[i]
        slope = -(f * (vel[2 * a] + vel[2 * b]) / 2.0) / p.g
        dhdy = (head[a] - head[b]) / (p.y[a] - p.y[b])
        err += abs(dhdy - slope)
    return np.array([err / (n - 1)])


def dam2d_initial_head(p, dam_location, hl, hr):
    """testcase_prep_dam2d (test_cases.c:1902-1913)."""
    return np.where(p.x <= dam_location, hl, hr).astype(np.float64)


class Dam2dError:
    """write_testcase_error_dam2d (test_cases.c:1916-2000): mean |h - h_a|
    and |u - u_a| against Ritter's dry-bed dam-break solution, evaluated after
    every step; the maxima over all steps are kept as the reference's
    error_h_tmax/error_u_tmax.  The fan velocity divides by the step's start
    time, as the reference's formula does (0/0 at the first step gives NaN,
    which the maxima skip)."""

    def __init__(self, dam_location, hl, hr):
        self.x0, self.hl, self.hr = dam_location, hl, hr
        self.tmax_h = 0.0
        self.tmax_u = 0.0
        self.last = None

    def __call__(self, p, head, vel, t_prev, dt):
        g = p.g
        c0 = math.sqrt(g * self.hl)
        time = t_prev + dt
        eu = eh = 0.0
        with np.errstate(divide="ignore", invalid="ignore"):
            for i in range(p.nnodes):
                x = p.x[i]
                xa = self.x0 - time * c0
                xb = self.x0 + 2 * time * c0
                if x <= xa:
                    ha, ua = self.hl, 0.0
                elif xa < x <= xb:
                    ha = (4.0 / (9.0 * g)) * (c0 - ((x - self.x0) / (2 * time))) ** 2
                    ua = (2.0 / 3.0) * (float(np.float64(x - self.x0) / np.float64(t_prev)) + c0)
                else:
                    ha, ua = self.hr, 0.0
                eu += abs(vel[2 * i] - ua)
                eh += abs(head[i] - ha)
        eh /= p.nnodes
        eu /= p.nnodes
        if eh > self.tmax_h:
            self.tmax_h = eh
        if eu > self.tmax_u:
            self.tmax_u = eu
        self.last = (eh, eu)
        return np.array([eh, eu, self.tmax_h, self.tmax_u])


def grid_mass_wet(p, head, density, djac):
    """tl_find_grid_mass_elem2d (src/tools/tl_grid_mass.c:20-) for a fully
    wet grid: density * sum over triangles of area * mean nodal depth."""
    e = p.e2d.reshape(-1, 3)
    h = head[e]
    if np.any(h <= 0.0):
        raise ValueError("grid has dry nodes; the wet-only mass formula does not apply")
    m = 0.0
    for ie in range(e.shape[0]):
        m += djac[ie] * (h[ie, 0] + h[ie, 1] + h[ie, 2]) / 3.0
    return density * m


def error_water_source(p, head, density, djac, time):
    """write_testcase_error_water_source (test_cases.c:1784-1818): |grid mass
    - analytic mass| and its ratio to the grid mass, for the deck's rainfall
    of 0.001 and 0.003 m/s on two quarter areas of a 10000 m^2 basin that
    starts 100 m deep."""
    omega1 = omega2 = 2500.0
    ri1, ri2 = 0.001, 0.003
    total_area = 10000.0
    head0 = 100.0
    grid_mass = grid_mass_wet(p, head, density, djac)
    analytic_depth = head0 + ((omega1 * ri1 + omega2 * ri2) / total_area) * time
    analytic_mass = analytic_depth * total_area * density
    err = abs(grid_mass - analytic_mass)
    return np.array([err, err / grid_mass])


SMALL = 2.220446049250313e-16


def _dry_wet_factornew(x, y, h, djac):
    """fe_3m2d_dry_wet_factornew (prototype src/tools/tl_grid_mass.c:5, body
    src/NEW_ADH/tools/tl_grid_mass.c:442-521): the wet fraction of a
    triangle cut at the zero-depth line, and the number of dry vertices."""
    passcode = 0
    total = 0
    for i in range(3):
        d = 1 if h[i] < 0.0 else 0
        total += d
        passcode += d << i
    if total == 3:
        return 0.0, 3
    if total == 0:
        return 1.0, 0
    if total != 1:
        passcode = 7 - passcode
    only = passcode // 2
    nx, ny = [0.0] * 3, [0.0] * 3
    for i in range(3):
        if i != only:
            xsi = h[i] / (h[i] - h[only])
            nx[i] = x[i] + xsi * (x[only] - x[i])      # FIXED_POS, macro.h:756-759
            ny[i] = y[i] + xsi * (y[only] - y[i])
        else:
            nx[i], ny[i] = x[i], y[i]
    elem_djac = djac
    v0x, v0y = nx[1] - nx[0], ny[1] - ny[0]            # TRI_AREA, macro.h:771-776
    v1x, v1y = nx[2] - nx[0], ny[2] - ny[0]
    area = (v0x * v1y - v0y * v1x) / 2.0
    if area / elem_djac < SMALL:
        i = total % 2
        return float(i), (0 if i == 1 else 3)
    factor = float((total >> 1) - (total % 2))
    factor *= area / elem_djac
    return factor + (1.0 if factor < 0 else 0.0), total


def grid_mass(p, head, density, djac):
    """tl_find_grid_mass_elem2d (src/tools/tl_grid_mass.c:20-96): density x
    the water volume, partially dry triangles integrated over their wet
    part with the negative depths counted as positive mass."""
    e = p.e2d.reshape(-1, 3)
    m = 0.0
    for ie in range(e.shape[0]):
        nd = e[ie]
        eh = [float(head[n]) for n in nd]
        neg = [0.0, 0.0, 0.0]
        factor, num_dry = _dry_wet_factornew([float(p.x[n]) for n in nd], [float(p.y[n]) for n in nd], eh,
                                             float(djac[ie]))
        if num_dry == 1:
            for i in range(3):
                if eh[i] < 0.0:
                    neg[i] = eh[i]
            f1, f2 = 1.0, -1.0 * (1.0 - factor)
        elif num_dry == 2:
            for i in range(3):
                if eh[i] < 0.0:
                    neg[i] = eh[i]
                    eh[i] = 0.0
            f1, f2 = factor, 0.0
        else:
            f1, f2 = factor, 0.0
        dj = float(djac[ie])
        # integrate_triangle_f (fe_elem2d_integrations.c:60-62)
        m += (f1 * dj * (eh[0] + eh[1] + eh[2]) / 3.0) + (f2 * dj * (neg[0] + neg[1] + neg[2]) / 3.0)
    return m * density


def error_wet_dry(p, head, vel, initial_depth, density, djac, initial_mass):
    """write_testcase_error_wet_dry (test_cases.c:297-335): mean |u|, mean
    |v| (the water starts and should end at rest), mean |h - h0|, each
    relative to its scale (1 for the velocities, max |h0| for the depth),
    and the grid mass error |mass - mass0| (a closed basin: no boundary or
    source flux, tl_grid_mass.c:99-150) with its ratio to the initial mass."""
    N = p.nnodes
    ex = float(np.sum(np.abs(vel[0::2])))
    ey = float(np.sum(np.abs(vel[1::2])))
    eh = float(np.sum(np.abs(head - initial_depth)))
    max_head = float(np.max(np.abs(initial_depth)))
    if max_head < 1e-6:
        max_head = 1.0
    gm_err = abs(grid_mass(p, head, density, djac) - 0.0 - initial_mass)
    return np.array([ex / N, ex / N / 1.0, ey / N, ey / N / 1.0, eh / N, eh / N / max_head, gm_err,
                     gm_err / initial_mass])


def error_winds(p, head, wind, node1, node2, g, density):
    """write_testcase_error_winds_and_waves (test_cases.c:1844-1892), 2D:
    mean over the transect node1-node2 of |g h_avg dh/ds - |tau|/rho|, the
    wind setup balance, with |tau| from the node wind values."""
    ids = slice_nodes_x(p, node1, node2)
    n = len(ids)
    err = 0.0
    for i in range(n - 1):
        a, b = ids[i + 1], ids[i]
        slope = math.sqrt(wind[2 * a] ** 2 + wind[2 * a + 1] ** 2)
        dist = math.sqrt((p.x[a] - p.x[b]) ** 2 + (p.y[a] - p.y[b]) ** 2)
        dhdx = (head[a] - head[b]) / dist
        err += abs(g * 0.5 * (head[a] + head[b]) * dhdx - slope / density)
    return np.array([err / (n - 1), float(n)])
