"""SW2 Newton steps/sec on one MI355X, FP64 and FP32-refined.

Workload: the 1000 x 500-cell channel (1,000,000 linear triangles, 501,501
nodes, 1.5M DOF): bed slope 1e-4 plus 20 seeded Gaussian bumps, Manning n
U(0.02, 0.035) per element, h0 = 1 m, dt = 5 s; each "step" of the metric is
one Newton iteration (FD Jacobian + assembly + scaling + block-Jacobi
BiCGSTAB + update + residual), iterated to 1e-8 relative residual within each
time step.  Inputs are resident in HBM before the timed region.

Prints one JSON line (the driver's contract) with the FP32-refined form as
`value`, the FP64 form beside it, the SpMV roofline (HIP-event timed inside
the library on its stream) and the CPU baseline (the oracle restatement,
timed on this host on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_form(p, precision, steps, warmup, rtol, graphs=1):
    from adhv8_kraken_amd.sw2 import SW2Model, default_opts, YES
    g = SW2Model(p, default_opts(precision=precision, graphs=graphs))
    # tolerance: 1e-8 of the initial residual of the first step (the reference's NTL is absolute)
    g.shift()
    g.init()
    r0 = float(np.abs(g.resid()).max())
    g.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
    g.set_opts(default_opts(precision=precision, tol_nonlin=rtol * r0, max_nonlin_it=40, graphs=graphs))
    for i in range(warmup):
        g.timestep()
    newton_its = 0
    lin_its = 0
    phases = dict(load=0.0, setup=0.0, solve=0.0, resid=0.0)
    t0 = time.perf_counter()
    for i in range(steps):
        r = g.timestep()
        if r["verdict"] != YES:
            raise RuntimeError(f"time step {i} did not converge: {r}")
        newton_its += r["iterations"]
        lin_its += r["lin_its"]
        for k in phases:
            phases[k] += r["ms_" + k]
        log(f"  precision={precision} step {i}: newton {r['iterations']} lin {r['lin_its']} "
            f"final max|R| {r['hist_max'][-1]:.3e}")
    elapsed = time.perf_counter() - t0
    h, vel, *_ = g.get_state()
    return g, dict(newton_its=newton_its, lin_its=lin_its, seconds=elapsed, r0=r0,
                   newton_per_s=newton_its / elapsed, phases_ms=phases, head=h, vel=vel)


def measured_traffic(precision):
    """HBM bytes per K1 launch from the committed counter summary (rocprofv3
    FETCH_SIZE + WRITE_SIZE passes, scaled by the copy-kernel calibration)."""
    import re
    path = os.path.join(ROOT, "profiles", "traffic_k_bcg_k1_channel1000x500.md")
    key = {0: "k_bcg_k1<double>", 1: "k_bcg_k1<mixed>", 2: "k_bcg_k1<float>"}[precision]
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes " + re.escape(key) + r" (\d+)$", txt, re.M)
    return int(m.group(1)) if m else None


def spmv_roofline(g, p, precision):
    """Algorithmic bytes per launch of the SpMV + block-Jacobi + dot kernel
    (K1) over its measured average duration."""
    N = p.nnodes
    nnz = len(g.pattern()[1])
    mb = 8 if precision == 0 else 4          # matrix / preconditioner entries
    vb = 4 if precision == 2 else 8          # vector entries
    # matrix values (diag + off-diagonal blocks), block columns, row_ptr,
    # preconditioner blocks, x (p), q, Ap written
    algo = 9 * mb * (N + nnz) + 4 * nnz + 4 * (N + 1) + 9 * mb * N + 3 * vb * N * 3
    ms = g.time_kernel(0, reps=50)
    achieved = algo / (ms * 1e-3) / 1e9
    traffic = measured_traffic(precision) if (p.nnodes, nnz) == (501501, 3003000) else None
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel="k_bcg_k1",
                avg_us=round(ms * 1e3, 2), algo_bytes=int(algo))


def cpu_baseline(seconds_budget=20.0):
    """The oracle (C restatement of the reference's serial path, skyline-LU
    preconditioned BiCGSTAB) on a bounded sample: one time step of the same
    channel family at 200 x 100 cells (40,000 triangles), scaled to the
    metric's 1M-triangle mesh by triangle count."""
    from adhv8_kraken_amd import mesh
    from oracle.oracle import Oracle, opts
    p = mesh.bumpy_channel(200, 100, seed=2)
    o = Oracle(p)
    o.shift()
    o.init()
    r0 = float(np.abs(o.resid()).max())
    best = None
    for nblock in (1, 4, 16):
        o.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
        o.shift()
        t0 = time.perf_counter()
        res = o.newton(opts(tol_nonlin=1e-8 * r0, prec=1, nblock=nblock, max_nonlin_it=40))
        dt = time.perf_counter() - t0
        rate = res["iterations"] / dt
        log(f"  cpu oracle nblock={nblock}: {res['iterations']} newton its in {dt:.2f}s ok={res['ok']}")
        if res["ok"] == 1 and (best is None or rate > best[0]):
            best = (rate, nblock, res["iterations"], dt)
        if dt > seconds_budget:
            break
    rate, nblock, its, dt = best
    scale = p.nelems2d / 1.0e6
    return dict(value=round(rate * scale, 4), unit="Newton steps/s", cores=1, kind="port",
                sample=f"oracle on the 200x100-cell channel (40k triangles), 1 time step, OP BLK {nblock}, "
                       f"{its} Newton its in {dt:.2f}s; rate scaled by 40k/1M triangles")


def kernel_mode(p):
    from adhv8_kraken_amd.sw2 import SW2Model, default_opts
    out = {}
    for prec in (2, 1, 0):
        g = SW2Model(p, default_opts(precision=prec, max_nonlin_it=2))
        g.shift()
        g.timestep()
        out[f"k1_prec{prec}_ms"] = g.time_kernel(0, reps=20)
        out[f"jac_prec{prec}_ms"] = g.time_kernel(1, reps=2)
        if prec == 0:
            out["copy_ms"] = g.time_kernel(4, reps=5)
            out["copy_bytes"] = 2 * 8 * 81 * p.nelems2d
        g.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4, help="timed time steps (each several Newton steps)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ncx", type=int, default=1000)
    ap.add_argument("--ncy", type=int, default=500)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graphs", type=int, default=0, help="1: replay captured HIP graphs of Krylov chunks")
    ap.add_argument("--kernels", action="store_true",
                    help="profiling mode: one time step per form, then the K1 kernel, the FD Jacobian "
                         "and the calibration copy launched back to back (for rocprofv3 counter passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding is described in "
              f"DESIGN.md but not run here", file=sys.stderr)
        sys.exit(2)
    from adhv8_kraken_amd import mesh
    log(f"building the {args.ncx}x{args.ncy}-cell channel")
    p = mesh.bumpy_channel(args.ncx, args.ncy, seed=2)
    log(f"  {p.nnodes} nodes, {p.nelems2d} triangles")
    if args.kernels:
        kernel_mode(p)
        return
    res = {}
    roof = {}
    for prec in (0, 1, 2):
        log(f"form precision={prec}")
        g, r = run_form(p, prec, args.steps, args.warmup, 1e-8, args.graphs)
        res[prec] = r
        roof[prec] = spmv_roofline(g, p, prec)
        roof[prec]["jac_us"] = round(g.time_kernel(1, reps=5) * 1e3, 1)
        g.close()
    cpu = None if args.no_cpu else cpu_baseline()
    r32, r64 = res[1], res[0]
    # both forms ran the same steps from the same state: error of the FP32-refined
    # state against the FP64 run, relative to the FP64 field's largest magnitude
    def rel_err(r):
        return dict(head=float(np.abs(r["head"] - r64["head"]).max() / np.abs(r64["head"]).max()),
                    velocity=float(np.abs(r["vel"] - r64["vel"]).max() / np.abs(r64["vel"]).max()))
    err, err32 = rel_err(r32), rel_err(res[2])
    out = {
        "metric": "SW2 Newton steps/sec @1M tris, FP64 vs FP32-refined, 1e-8 rel residual",
        "value": round(r32["newton_per_s"], 3),
        "unit": "Newton steps/s",
        "n_gpus": 1,
        "steps": r32["newton_its"],
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r32["seconds"] / r32["newton_its"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"channel {args.ncx}x{args.ncy} cells, {p.nelems2d} triangles, dt 5 s, "
                               f"{args.steps} time steps, Newton to 1e-8 rel",
                   "nodes": p.nnodes, "triangles": p.nelems2d},
        "fp64": {"value": round(r64["newton_per_s"], 3), "newton_its": r64["newton_its"],
                 "lin_its": r64["lin_its"], "phases_ms": r64["phases_ms"], "roofline": roof[0]},
        "fp32_refined": {"newton_its": r32["newton_its"], "lin_its": r32["lin_its"], "phases_ms": r32["phases_ms"],
                         "max_rel_error_vs_fp64": err,
                         "form": "FP32 Jacobian storage, preconditioner and SpMV; FP64 Krylov vectors, residual, update"},
        "fp32_all": {"value": round(res[2]["newton_per_s"], 3), "newton_its": res[2]["newton_its"],
                     "lin_its": res[2]["lin_its"], "phases_ms": res[2]["phases_ms"],
                     "max_rel_error_vs_fp64": err32, "roofline": roof[2],
                     "form": "FP32 matrix and Krylov vectors; FP64 residual, update"},
        "roofline": roof[1],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
