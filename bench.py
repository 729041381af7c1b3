"""SW2 Newton steps/sec on one MI355X, FP64 and FP32-refined.

Workload: the 1000 x 500-cell channel (1,000,000 linear triangles, 501,501
nodes, 1.5M DOF): bed slope 1e-4 plus 20 seeded Gaussian bumps, Manning n
U(0.02, 0.035) per element, h0 = 1 m, dt = 5 s; each "step" of the metric is
one Newton iteration (FD Jacobian + assembly + scaling + block-Jacobi
BiCGSTAB + update + residual), iterated to 1e-8 relative residual within each
time step.  Inputs are resident in HBM before the timed region.

Prints one JSON line with the FP32-refined form as
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


def run_form(p, precision, steps, warmup, rtol, graphs=1, autotune=True, preconditioner=2):
    from adhv8_kraken_amd.sw2 import SW2Model, default_opts, YES
    g = SW2Model(p, default_opts(precision=precision, graphs=graphs, preconditioner=preconditioner))
    # tolerance: 1e-8 of the initial residual of the first step (the reference's NTL is absolute)
    g.shift()
    g.init()
    r0 = float(np.abs(g.resid()).max())
    g.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
    g.set_opts(default_opts(precision=precision, tol_nonlin=rtol * r0, max_nonlin_it=40, graphs=graphs,
                            preconditioner=preconditioner))
    for i in range(warmup):
        g.timestep()
    tuning = g.autotune(reps=3) if autotune else None
    # the timed steps start again from the initial state: the workload's own
    # time steps, with the device, caches and launch shapes warm
    g.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
    g.set_dt(p.dt)
    newton_its = 0
    lin_its = 0
    phases = dict(load=0.0, setup=0.0, solve=0.0, resid=0.0)
    first = None
    elapsed = 0.0
    for i in range(steps):
        t0 = time.perf_counter()
        r = g.timestep()
        elapsed += time.perf_counter() - t0
        if r["verdict"] != YES:
            raise RuntimeError(f"time step {i} did not converge: {r}")
        newton_its += r["iterations"]
        lin_its += r["lin_its"]
        for k in phases:
            phases[k] += r["ms_" + k]
        if i == 0:   # the state after the first step (outside the timed region), for the oracle comparison
            first = g.get_state()[:2]
        log(f"  precision={precision} pc={preconditioner} step {i}: newton {r['iterations']} lin {r['lin_its']} "
            f"final max|R| {r['hist_max'][-1]:.3e}")
    h, vel, *_ = g.get_state()
    return g, dict(newton_its=newton_its, lin_its=lin_its, seconds=elapsed, r0=r0,
                   newton_per_s=newton_its / elapsed, phases_ms=phases, head=h, vel=vel, tuning=tuning,
                   first=first)


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


def measured_traffic_kernel(name, summary="traffic_k_bcg_k1_channel1000x500.md"):
    """HBM bytes (read + write) per launch of a kernel from a committed
    FETCH_SIZE / WRITE_SIZE table (copy-kernel calibrated), or None."""
    import re
    try:
        txt = open(os.path.join(ROOT, "profiles", summary)).read()
    except OSError:
        return None
    m = re.search(r"^\| `[^`]*" + re.escape(name) + r"[^`]*` \| \d+ \| \d+ \| ([0-9.]+) \| ([0-9.]+) \|", txt, re.M)
    return int(round((float(m.group(1)) + float(m.group(2))) * 1e6)) if m else None


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


# Jacobian layout -> (summary name, kernel name prefix in the counter tables)
JAC_KERNELS = {0: ("k_jac2d_list", "k_jac2d_list<true"), 1: ("k_jac2d_cols", "k_jac2d_cols<true"),
               2: ("k_jac2d_lds_cols", "k_jac2d_lds<true"), 3: ("k_jac2d_lds_group", "k_jac2d_lds<true")}


def summary_gflop(name):
    """The `= **X GFLOP** per launch` figure of a committed counter summary."""
    import re
    try:
        txt = open(os.path.join(ROOT, "profiles", name)).read()
    except OSError:
        return None
    m = re.search(r"= \*\*([0-9.]+) GFLOP\*\* per launch", txt)
    return float(m.group(1)) * 1e9 if m else None


def jacobian_roofline(g, p):
    """FD Jacobian (wet list kernel, the dominant kernel) against the FP64
    vector peak.  `achieved` is the algorithmic FP64 work of a Jacobian --
    every triangle's 18 perturbed residuals evaluated in full, as the
    reference does (the pressure block for the 6 depth perturbations only,
    fe_sw2_body_resid.c:761-862), counted with rocprofv3 FP64 instruction
    counters on the column-uniform kernel that does exactly that (layout 2,
    committed summary) -- over the live HIP-event time of the kernel in use.
    The default kernel executes less (the drag coefficient shared by a
    triangle's u and v columns); its own count is reported beside it."""
    kernel, kname = JAC_KERNELS[g.get_tuning().layout[0]]
    ms = g.time_kernel(1, reps=5)
    flops = executed = traffic = None
    if (p.nnodes, p.nelems2d) == (501501, 1000000):
        flops = summary_gflop("counters_k_jac2d_lds_cols_channel1000x500.md")
        executed = summary_gflop(f"counters_{kernel}_channel1000x500.md")
        traffic = measured_traffic_kernel(kname, f"traffic_{kernel}_channel1000x500.md")
    out = dict(bound="fp64-vector", kernel=kernel, avg_us=round(ms * 1e3, 1), peak=78.6, unit="TFLOP/s",
               traffic=traffic)
    if flops:
        ach = flops / (ms * 1e-3) / 1e12
        out.update(achieved=round(ach, 2), frac=round(ach / 78.6, 4), flops_per_launch=flops,
                   flops_per_triangle=round(flops / p.nelems2d, 1), executed_flops_per_launch=executed)
    return out


def profile_lu_cost(p, nblock):
    """Storage (bytes) and flops of the reference's skyline LU of `nblock`
    node blocks on mesh p: contiguous node ranges of N // nblock nodes, the
    last taking the rest (src/node/node_order.c:64-91); a DOF row's profile
    runs from its leftmost same-block neighbour (solv_blk_profile.c:5-47), and
    the factorisation (solv_blk_load_profile.c:51-79) costs about
    2 * sum(row length^2) flops."""
    N = p.nnodes
    e = np.asarray(p.e2d).reshape(-1, 3)
    inc = N // nblock
    blk = np.minimum(np.arange(N) // inc, nblock - 1)
    mn = np.arange(N)
    for a in range(3):
        for b in range(3):
            same = blk[e[:, a]] == blk[e[:, b]]
            np.minimum.at(mn, e[same, a], e[same, b])
    rowlen = np.repeat(3 * (np.arange(N) - mn), 3) + np.tile(np.arange(3), N)
    return 8 * (2 * int(rowlen.sum()) + 3 * N), 2.0 * float((rowlen.astype(np.float64) ** 2).sum())


def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(p):
    """The oracle (C restatement of the reference's serial path: FD Jacobian,
    assembly, Jacobi scaling, skyline-LU preconditioned BiCGSTAB, Newton) on
    the metric's own mesh, one core.  OP BLK candidates: the reference's
    default 1 (one LU of the whole matrix) and blocks of 256-4096 nodes plus
    one mesh column; the cost model `profile_lu_cost` ranks them and the
    cheapest is run for one full time step (Newton to 1e-8 of R0)."""
    from oracle.oracle import Oracle, opts
    N = p.nnodes
    npy = p.meta["npy"]
    sizes = {"BLK 1": 1}
    for nodes in (256, 512, 1024, 2048, 4096):
        sizes[f"{nodes}-node blocks"] = N // nodes
    sizes[f"one mesh column ({npy} nodes)"] = N // npy
    costs = {k: profile_lu_cost(p, nb) for k, nb in sizes.items()}
    best = min(costs, key=lambda k: costs[k][1])
    nblock = sizes[best]
    o = Oracle(p)
    o.shift()
    o.init()
    r0 = float(np.abs(o.resid()).max())
    o.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
    o.shift()
    t0 = time.perf_counter()
    res = o.newton(opts(tol_nonlin=1e-8 * r0, prec=1, nblock=nblock, max_nonlin_it=40))
    dt = time.perf_counter() - t0
    its = res["iterations"]
    log(f"  cpu oracle OP BLK {nblock} ({best}): {its} newton its in {dt:.2f}s ok={res['ok']} "
        f"(load {res['time_load']:.1f}s, setup {res['time_setup']:.1f}s, solve {res['time_solve']:.1f}s)")
    o_state = o.get_state()[:2]
    model, nproc = host_cpu()
    others = "; ".join(f"{k}: OP BLK {sizes[k]}, {costs[k][0] / 1e9:.2f} GB, {costs[k][1]:.2e} flops"
                       for k in sizes if k != best)
    return dict(value=round(its / dt, 4) if res["ok"] == 1 else None, unit="Newton steps/s", cores=1, kind="port",
                converged=res["ok"] == 1, op_blk=nblock, cpu_model=model, nproc=nproc,
                sample=f"oracle on the {p.meta['npx'] - 1}x{npy - 1}-cell channel itself ({p.nelems2d} triangles), "
                       f"one time step from the initial state, OP BLK {nblock} ({best}: skyline LU of "
                       f"{costs[best][0] / 1e9:.2f} GB, {costs[best][1]:.2e} flops per factorisation), "
                       f"{its} Newton its in {dt:.1f} s (load {res['time_load']:.1f} s, set-up "
                       f"{res['time_setup']:.1f} s, solve {res['time_solve']:.1f} s); not run, by the same "
                       f"cost model: {others}"), o_state


def run_ensemble(ps, precision, steps, warmup, rtol, max_retries=2, autotune=True):
    """Every instance of the ensemble advanced `steps` time steps together
    (one launch per phase for all instances); each instance's nonlinear
    tolerance is rtol times its own opening residual."""
    from adhv8_kraken_amd.sw2 import SW2Ensemble, default_opts, YES
    t0 = time.perf_counter()
    ens = SW2Ensemble(ps, default_opts(precision=precision, max_nonlin_it=40))
    t_build = time.perf_counter() - t0
    r0 = ens.residual_norm()
    for i, p in enumerate(ps):
        ens.set_state(i, p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
        ens.set_tol(i, rtol * r0[i])
    for _ in range(warmup):
        ens.timestep(max_retries=max_retries)
    tuning = ens.autotune(reps=2) if autotune else None
    its = np.zeros(len(ps), dtype=np.int64)
    lin = np.zeros(len(ps), dtype=np.int64)
    fails = 0
    dts = [[] for _ in ps]
    t0 = time.perf_counter()
    for s in range(steps):
        rs = ens.timestep(max_retries=max_retries)
        for i, r in enumerate(rs):
            its[i] += r["iterations"]
            lin[i] += r["lin_its"]
            fails += r["verdict"] != YES
            dts[i].append((r["verdict"], ens.get_dt(i)))
        log(f"  precision={precision} step {s}: newton its {int(sum(r['iterations'] for r in rs))} "
            f"(max {max(r['iterations'] for r in rs)}), failed {sum(r['verdict'] != YES for r in rs)}")
    elapsed = time.perf_counter() - t0
    tris = np.array([p.nelems2d for p in ps], dtype=np.float64)
    states = [ens.get_state(i) for i in range(len(ps))]
    ens.close()
    return dict(seconds=elapsed, build_s=t_build, newton_its=int(its.sum()), lin_its=int(lin.sum()), fails=fails,
                inst_newton_per_s=float(its.sum() / elapsed), tri_newton_per_s=float((its * tris).sum() / elapsed),
                states=states, dts=dts, tuning=tuning)


def run_sequential(ps, precision, steps, rtol, max_retries=2):
    """The same instances one after another through the single-model path."""
    from adhv8_kraken_amd.sw2 import SW2Model, default_opts
    tri_its = 0.0
    total = 0.0
    for p in ps:
        g = SW2Model(p, default_opts(precision=precision))
        g.shift()
        g.init()
        r0 = float(np.abs(g.resid()).max())
        g.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
        g.set_opts(default_opts(precision=precision, tol_nonlin=rtol * r0, max_nonlin_it=40))
        t0 = time.perf_counter()
        for _ in range(steps):
            r = g.timestep(max_retries=max_retries)
            tri_its += r["iterations"] * p.nelems2d
        total += time.perf_counter() - t0
        g.close()
    return dict(seconds=total, tri_newton_per_s=tri_its / total)


def ensemble_mode(args):
    from adhv8_kraken_amd import mesh
    if args.workload == "channel-ensemble":
        ps = mesh.channel_ensemble()
        steps = args.steps if args.steps_set else 20
        forms = (0, 1)
        desc = "256 channels, 5k-200k triangles log-uniform (seed 4), Manning n U(0.015,0.05), inflow U(5,50) m^3/s"
    else:
        ps = mesh.floodplain_ensemble()
        steps = args.steps if args.steps_set else 10   # a bounded sample of the 50-step run
        forms = (1, 0)
        desc = ("64 wet/dry floodplains, 50k-500k triangles log-uniform (seed 5), cross-slope 1e-3, 40% dry, "
                "rain 1e-5 m/s, dt 2 s (cut by 0.25 per failed attempt, up to 2 retries)")
    ntri = sum(p.nelems2d for p in ps)
    log(f"{args.workload}: {len(ps)} instances, {ntri} triangles, {sum(p.nnodes for p in ps)} nodes")
    if args.time_span > 0:
        return time_control_mode(args, ps, forms[0], desc, ntri)
    res = {}
    for prec in forms:
        log(f"form precision={prec}")
        res[prec] = run_ensemble(ps, prec, steps, args.warmup, 1e-8, autotune=not args.no_autotune)
        log(f"  {res[prec]['tri_newton_per_s'] / 1e6:.1f} M triangle-Newton/s")
    err = None
    if 0 in res and 1 in res:
        # instances whose two runs took the same time steps (a failed step
        # cuts dt, after which the difference is time discretisation, not precision)
        eh = ev = 0.0
        same = 0
        for i, ((h64, v64), (h32, v32)) in enumerate(zip(res[0]["states"], res[1]["states"])):
            if res[0]["dts"][i] != res[1]["dts"][i]:
                continue
            same += 1
            eh = max(eh, float(np.abs(h32 - h64).max() / max(np.abs(h64).max(), 1e-300)))
            ev = max(ev, float(np.abs(v32 - v64).max() / max(np.abs(v64).max(), 1e-300)))
        err = dict(head=eh, velocity=ev, instances_compared=same)
    seq = None
    if args.compare > 0:
        sub = ps[: args.compare]
        csteps = 3
        e = run_ensemble(sub, forms[0], csteps, 0, 1e-8)
        q = run_sequential(sub, forms[0], csteps, 1e-8)
        seq = dict(instances=len(sub), steps=csteps, ensemble_tri_newton_per_s=round(e["tri_newton_per_s"]),
                   sequential_tri_newton_per_s=round(q["tri_newton_per_s"]),
                   speedup=round(e["tri_newton_per_s"] / q["tri_newton_per_s"], 2))
    main = res[forms[0]]
    out = {
        "metric": "SW2 ensemble Newton throughput (sum of triangles x Newton iterations per second)",
        "value": round(main["tri_newton_per_s"]),
        "unit": "triangle-Newton steps/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if forms[0] == 0 else "f32+f64", "data": "synthetic",
        "config": {"workload": desc, "instances": len(ps), "triangles": ntri, "time_steps": steps,
                   "newton_tol": "1e-8 of each instance's opening residual"},
        "forms": {str(k): dict(tri_newton_per_s=round(v["tri_newton_per_s"]),
                               instance_newton_per_s=round(v["inst_newton_per_s"], 1),
                               equiv_1M_tri_newton_per_s=round(v["tri_newton_per_s"] / 1e6, 2),
                               newton_its=v["newton_its"], lin_its=v["lin_its"], failed_steps=v["fails"],
                               seconds=round(v["seconds"], 2), build_s=round(v["build_s"], 1),
                               list_kernel_shapes=v["tuning"])
                  for k, v in res.items()},
        "fp32_refined_max_rel_error_vs_fp64": err,
        "vs_sequential_single_model": seq,
    }
    print(json.dumps(out), flush=True)


def direct_mode(args):
    """Instances of the wet/dry floodplain ensemble one at a time with the
    reference's default preconditioner, OP BLK 1 (one LU of each instance's
    whole matrix, the tiled band LU), FP64: a bounded sample (the smallest,
    median and largest instance unless --instances picks others, --steps time
    steps each, default 1) of the systems block Jacobi does not solve."""
    from adhv8_kraken_amd import mesh
    from adhv8_kraken_amd.sw2 import SW2Model, default_opts, YES
    sizes = mesh.log_uniform_sizes(64, 50000, 500000, 5)
    order = np.argsort(sizes)
    pick = ([int(k) for k in args.instances.split(",")] if args.instances else
            [int(order[0]), int(order[len(order) // 2]), int(order[-1])])
    steps = args.steps if args.steps_set else 1
    rows, tri_newton, seconds, fails = [], 0, 0.0, 0
    for k in pick:
        ncx, ncy = mesh._grid_for(sizes[k])
        p = mesh.floodplain(ncx, ncy, seed=5 * 1000 + k, dt=2.0)
        g = SW2Model(p, default_opts(preconditioner=1, nblock=1, max_nonlin_it=40))
        g.shift(); g.init()
        r0 = float(np.abs(g.resid()).max())
        g.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
        g.set_opts(default_opts(preconditioner=1, nblock=1, max_nonlin_it=40, tol_nonlin=1e-8 * r0))
        its = setup_ms = solve_ms = 0
        t0 = time.time()
        verdicts = []
        for _ in range(steps):
            r = g.timestep(max_retries=2)
            verdicts.append(int(r["verdict"] == YES))
            its += r["iterations"]
            setup_ms += r["ms_setup"]
            solve_ms += r["ms_solve"]
        dt = time.time() - t0
        g.close()
        fails += verdicts.count(0)
        tri_newton += p.nelems2d * its
        seconds += dt
        rows.append(dict(instance=k, cells=f"{ncx}x{ncy}", triangles=p.nelems2d, newton_its=its,
                         steps_ok=sum(verdicts), seconds=round(dt, 2),
                         factor_ms_per_newton=round(setup_ms / max(its, 1), 1),
                         solve_ms_per_newton=round(solve_ms / max(its, 1), 1)))
        log(f"  instance {k} ({ncx}x{ncy}): {its} Newton its, {sum(verdicts)}/{steps} steps, {dt:.1f} s")
    out = {
        "metric": "SW2 floodplain instances with OP BLK 1 on the device (triangles x Newton iterations per second)",
        "value": round(tri_newton / seconds), "unit": "triangle-Newton steps/s", "n_gpus": 1, "steps": steps,
        "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "wet/dry floodplain ensemble instances (64 instances, 50k-500k triangles, seed 5), "
                               "one at a time, OP BLK 1 (tiled band LU of the whole matrix), BiCGSTAB, Newton to "
                               "1e-8 of the opening residual, dt 2 s cut by 0.25 per failed attempt (2 retries)",
                   "instances": pick, "time_steps": steps},
        "failed_steps": fails, "instances": rows,
    }
    print(json.dumps(out), flush=True)


def direct_ensemble_mode(args):
    """The whole wet/dry floodplain ensemble (or the --instances subset) with
    the reference's default preconditioner, OP BLK 1, inside the ensemble
    handle: instances grouped by band storage (sw2.lu_groups, at most
    --lu-budget GB of bands per group, like sizes together), each group one
    handle whose tiled-LU launches factor and sweep all its members' bands
    at once; FP64, --steps time steps (default 1).  --compare K also runs the
    K smallest instances one at a time through the single-model path."""
    from adhv8_kraken_amd import mesh
    from adhv8_kraken_amd.sw2 import SW2Ensemble, SW2Model, default_opts, lu_groups, profile_lu_bytes, YES
    ps = mesh.floodplain_ensemble()
    idx = [int(k) for k in args.instances.split(",")] if args.instances else list(range(len(ps)))
    steps = args.steps if args.steps_set else 1
    sub = [ps[i] for i in idx]
    groups = [[idx[j] for j in g] for g in lu_groups(sub, 1, args.lu_budget * 1e9)]
    log(f"floodplain-direct-ensemble: {len(idx)} instances in {len(groups)} groups "
        f"({[len(g) for g in groups]}), {sum(profile_lu_bytes(ps[i]) for i in idx) / 1e9:.1f} GB of bands")

    def opts(tol=1e-8):
        return default_opts(preconditioner=1, nblock=1, max_nonlin_it=40, tol_nonlin=tol)

    rows, tri_newton, seconds, fails = [], 0, 0.0, 0
    for gi, g in enumerate(groups):
        gp = [ps[i] for i in g]
        ens = SW2Ensemble(gp, opts())
        r0 = ens.residual_norm()
        for j, p in enumerate(gp):
            ens.set_state(j, p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
            ens.set_tol(j, 1e-8 * r0[j])
        its = np.zeros(len(g), dtype=np.int64)
        ok = np.zeros(len(g), dtype=np.int64)
        t0 = time.time()
        for st in range(steps):
            rs = ens.timestep(max_retries=2)
            for j, r in enumerate(rs):
                its[j] += r["iterations"]
                ok[j] += r["verdict"] == YES
            log(f"  group {gi} step {st}: {sum(r['iterations'] for r in rs)} Newton its, "
                f"{sum(r['verdict'] == YES for r in rs)}/{len(g)} ok, {time.time() - t0:.1f} s")
        dt = time.time() - t0
        ens.close()
        tn = int(sum(its[j] * gp[j].nelems2d for j in range(len(g))))
        tri_newton += tn
        seconds += dt
        fails += int(steps * len(g) - ok.sum())
        rows.append(dict(instances=g, triangles=int(sum(p.nelems2d for p in gp)),
                         band_gb=round(sum(profile_lu_bytes(p) for p in gp) / 1e9, 1),
                         newton_its=int(its.sum()), max_newton_its=int(its.max()), steps_ok=int(ok.sum()),
                         seconds=round(dt, 2), tri_newton_per_s=round(tn / dt)))
        log(f"  group {gi}: {len(g)} instances, {int(its.sum())} Newton its (max {int(its.max())}), "
            f"{int(ok.sum())}/{steps * len(g)} steps ok, {dt:.1f} s")
    seq = None
    if args.compare > 0:
        small = sorted(idx, key=lambda i: profile_lu_bytes(ps[i]))[: args.compare]
        tn_e, t_e, tn_s, t_s = 0, 0.0, 0, 0.0
        gp = [ps[i] for i in small]
        ens = SW2Ensemble(gp, opts())
        r0 = ens.residual_norm()
        for j, p in enumerate(gp):
            ens.set_state(j, p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
            ens.set_tol(j, 1e-8 * r0[j])
        t0 = time.time()
        rs = ens.timestep(max_retries=2)
        t_e = time.time() - t0
        tn_e = sum(r["iterations"] * p.nelems2d for r, p in zip(rs, gp))
        ens.close()
        for j, p in enumerate(gp):
            m = SW2Model(p, opts())
            m.shift(); m.init()
            m.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
            m.set_opts(opts(1e-8 * r0[j]))
            t0 = time.time()
            r = m.timestep(max_retries=2)
            t_s += time.time() - t0
            tn_s += r["iterations"] * p.nelems2d
            m.close()
        seq = dict(instances=small, ensemble_tri_newton_per_s=round(tn_e / t_e),
                   sequential_tri_newton_per_s=round(tn_s / t_s), speedup=round((tn_e / t_e) / (tn_s / t_s), 2))
        log(f"  {len(small)} smallest instances, one step: ensemble {tn_e / t_e / 1e6:.2f} vs one at a time "
            f"{tn_s / t_s / 1e6:.2f} M triangle-Newton/s")
    out = {
        "metric": "SW2 floodplain ensemble with OP BLK 1 inside the ensemble handle (triangles x Newton iterations "
                  "per second)",
        "value": round(tri_newton / seconds), "unit": "triangle-Newton steps/s", "n_gpus": 1, "steps": steps,
        "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "wet/dry floodplain ensemble (64 instances, 50k-500k triangles, seed 5), OP BLK 1 "
                               "(tiled band LU of each instance's whole matrix), BiCGSTAB, Newton to 1e-8 of each "
                               "instance's opening residual, dt 2 s cut by 0.25 per failed attempt (2 retries), "
                               f"groups of at most {args.lu_budget:g} GB of bands",
                   "instances": len(idx), "time_steps": steps},
        "seconds": round(seconds, 2), "failed_steps": fails, "groups": rows, "vs_one_at_a_time": seq,
    }
    print(json.dumps(out), flush=True)


def time_control_mode(args, ps, precision, desc, ntri):
    """Every instance integrated to t = --time-span with the reference's time
    control, capped at --wall seconds: reports how far the instances got."""
    from adhv8_kraken_amd.sw2 import SW2Ensemble, default_opts
    from adhv8_kraken_amd import timeloop
    ens = SW2Ensemble(ps, default_opts(precision=precision, max_nonlin_it=40))
    r0 = ens.residual_norm()
    for i, p in enumerate(ps):
        ens.set_state(i, p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
        ens.set_tol(i, 1e-8 * r0[i])
    dt_max = ps[0].dt

    def progress(calls, t, failed):
        log(f"  call {calls}: t min {t.min():.4g} s, max {t.max():.4g} s, at the end "
            f"{int((t >= args.time_span).sum())}/{len(t)}, calls with every attempt failed {int(failed.sum())}")
    res = timeloop.run_ensemble(ens, args.time_span, dt_max, wall_s=args.wall, on_call=progress)
    ens.close()
    tris = np.array([p.nelems2d for p in ps], dtype=np.float64)
    out = {
        "metric": "SW2 ensemble under the reference's time control: instances reaching the end time",
        "value": int((res["t"] >= args.time_span).sum()), "unit": "instances", "n_gpus": 1,
        "higher_is_better": True, "data": "synthetic", "dtype": "f64" if precision == 0 else "f32+f64",
        "config": {"workload": desc, "instances": len(ps), "triangles": ntri, "end_time_s": args.time_span,
                   "dt_max_s": dt_max, "wall_cap_s": args.wall},
        "seconds": round(res["seconds"], 1), "calls": res["calls"],
        "t_reached_min_s": float(res["t"].min()), "t_reached_median_s": float(np.median(res["t"])),
        "t_reached_max_s": float(res["t"].max()), "smallest_dt_s": res["min_dt"],
        "attempts": int(res["attempts"].sum()), "calls_with_all_attempts_failed": int(res["failed_calls"].sum()),
        "tri_newton_per_s": float((res["newton"] * tris).sum() / max(res["seconds"], 1e-9)),
    }
    print(json.dumps(out), flush=True)


def kernel_mode(p, forms, jac_layout=None):
    from adhv8_kraken_amd.sw2 import SW2Model, default_opts
    out = {}
    for prec in forms:
        g = SW2Model(p, default_opts(precision=prec, max_nonlin_it=2))
        if jac_layout is not None:
            t = g.get_tuning()
            t.layout[0] = jac_layout
            g.set_tuning(t)
        g.shift()
        g.timestep()
        out[f"k1_prec{prec}_ms"] = g.time_kernel(0, reps=20)
        out[f"jac_prec{prec}_ms"] = g.time_kernel(1, reps=2)
        if prec == forms[-1]:
            out["copy_ms"] = g.time_kernel(4, reps=5)
            out["copy_bytes"] = 2 * 8 * 81 * p.nelems2d
        g.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed time steps (each several Newton steps)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ncx", type=int, default=1000)
    ap.add_argument("--ncy", type=int, default=500)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graphs", type=int, default=0, help="1: replay captured HIP graphs of Krylov chunks")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes of the element list kernels (for clean kernel traces)")
    ap.add_argument("--workload", default="channel",
                    choices=["channel", "channel-ensemble", "floodplain-ensemble", "floodplain-direct",
                             "floodplain-direct-ensemble"],
                    help="channel: the 1M-triangle headline; the others print their own JSON line")
    ap.add_argument("--instances", default="",
                    help="floodplain-direct: comma-separated instance indices (default: smallest, median, "
                         "largest); floodplain-direct-ensemble: the subset to run (default: all 64)")
    ap.add_argument("--lu-budget", type=float, default=200.0,
                    help="floodplain-direct-ensemble: GB of profile-LU bands per ensemble group")
    ap.add_argument("--compare", type=int, default=0,
                    help="ensembles: also run the first K instances one by one through the single-model path")
    ap.add_argument("--time-span", type=float, default=0.0,
                    help="ensembles: run the reference's time control (retry with dt/4 until a step succeeds, "
                         "regrow dt x2 to the deck's dt) for every instance up to this simulated time")
    ap.add_argument("--wall", type=float, default=300.0, help="wall-clock cap of --time-span runs, seconds")
    ap.add_argument("--op-blk1-steps", type=int, default=1,
                    help="time steps of the FP64 OP BLK 1 form (the reference's default preconditioner) run as a "
                         "bounded sample beside the block-Jacobi forms; 0 skips it")
    ap.add_argument("--kernels", action="store_true",
                    help="profiling mode: one time step per form, then the K1 kernel, the FD Jacobian "
                         "and the calibration copy launched back to back (for rocprofv3 counter passes)")
    ap.add_argument("--kernel-forms", default="2,1,0",
                    help="--kernels: the precisions to run (0 FP64, 1 FP32-refined, 2 all-FP32)")
    ap.add_argument("--jac-layout", type=int, default=None,
                    help="--kernels: the wet-list Jacobian layout to run (default: the library's)")
    args = ap.parse_args()
    args.steps_set = args.steps is not None
    if args.steps is None:
        args.steps = 10
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the path is built and measured on one "
              f"GPU; sharding across GPUs (domain decomposition, or instances for ensembles) is not built",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "floodplain-direct":
        direct_mode(args)
        return
    if args.workload == "floodplain-direct-ensemble":
        direct_ensemble_mode(args)
        return
    if args.workload != "channel":
        ensemble_mode(args)
        return
    from adhv8_kraken_amd import mesh
    log(f"building the {args.ncx}x{args.ncy}-cell channel")
    p = mesh.bumpy_channel(args.ncx, args.ncy, seed=2)
    log(f"  {p.nnodes} nodes, {p.nelems2d} triangles")
    if args.kernels:
        kernel_mode(p, [int(f) for f in args.kernel_forms.split(",")], args.jac_layout)
        return
    res = {}
    roof = {}
    # forms: (precision, preconditioner) -- FP64, FP32-refined, all-FP32 with
    # block Jacobi, and the FP32-refined form with the block ILU(0)
    forms = [(0, 2), (1, 2), (2, 2), (1, 3)]
    for prec, pc in forms:
        log(f"form precision={prec} preconditioner={pc}")
        g, r = run_form(p, prec, args.steps, args.warmup, 1e-8, args.graphs, autotune=not args.no_autotune,
                        preconditioner=pc)
        res[(prec, pc)] = r
        if pc == 2:
            roof[prec] = spmv_roofline(g, p, prec)
            roof[prec]["jac_us"] = round(g.time_kernel(1, reps=5) * 1e3, 1)
        if (prec, pc) == (1, 2):
            jroof = jacobian_roofline(g, p)
        g.close()
    # the reference's default preconditioner, OP BLK 1 (one LU of the whole
    # matrix: the tiled band LU), FP64: a bounded sample of the same workload
    op1 = None
    if args.op_blk1_steps > 0:
        log(f"form precision=0 preconditioner=1 nblock=1 (OP BLK 1), {args.op_blk1_steps} time step(s)")
        g, op1 = run_form(p, 0, args.op_blk1_steps, 0, 1e-8, 0, autotune=False, preconditioner=1)
        g.close()
    cpu, o_state = (None, None) if args.no_cpu else cpu_baseline(p)
    r64 = res[(0, 2)]
    # the refined forms' headline is the faster preconditioner
    best_pc = 3 if res[(1, 3)]["newton_per_s"] > res[(1, 2)]["newton_per_s"] else 2
    r32 = res[(1, best_pc)]

    # every form ran the same time steps from the same state: error against the FP64 run, relative to
    # the FP64 field's largest magnitude
    def rel_err(a, b):
        return dict(head=float(np.abs(a[0] - b[0]).max() / np.abs(b[0]).max()),
                    velocity=float(np.abs(a[1] - b[1]).max() / np.abs(b[1]).max()))

    def form_out(r, form):
        out = {"value": round(r["newton_per_s"], 3), "newton_its": r["newton_its"], "lin_its": r["lin_its"],
               "phases_ms": r["phases_ms"], "form": form,
               "max_rel_error_vs_fp64_gpu": rel_err((r["head"], r["vel"]), (r64["head"], r64["vel"]))}
        if o_state is not None:   # the oracle ran the first time step in FP64 from the same state
            out["first_step_max_rel_error_vs_fp64_oracle"] = rel_err(r["first"], o_state)
        return out

    out = {
        "metric": "SW2 Newton steps/sec @1M tris, FP64 vs FP32-refined, 1e-8 rel residual",
        "value": round(r32["newton_per_s"], 3),
        "unit": "Newton steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r32["seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"channel {args.ncx}x{args.ncy} cells, {p.nelems2d} triangles, dt 5 s, "
                               f"{args.steps} time steps, Newton to 1e-8 rel",
                   "nodes": p.nnodes, "triangles": p.nelems2d},
        "time_steps": args.steps,
        "newton_its": r32["newton_its"],
        "ms_per_newton_it": round(1e3 * r32["seconds"] / r32["newton_its"], 3),
        "form": ("FP32 Jacobian storage, preconditioner and SpMV; FP64 Krylov vectors, residual, update; "
                 + ("block ILU(0)" if best_pc == 3 else "3x3 block Jacobi")),
        "fp64": form_out(r64, "FP64, 3x3 block Jacobi"),
        "fp32_refined_block_jacobi": form_out(res[(1, 2)], "FP32 matrix and block-Jacobi inverses, FP64 vectors"),
        "fp32_refined_ilu0": form_out(res[(1, 3)], "FP32 matrix and ILU(0) factors, FP64 vectors"),
        "fp32_all": form_out(res[(2, 2)], "FP32 matrix and Krylov vectors; FP64 residual, update"),
        "fp64_op_blk1": None if op1 is None else {
            "value": round(op1["newton_per_s"], 4), "newton_its": op1["newton_its"], "lin_its": op1["lin_its"],
            "time_steps": args.op_blk1_steps, "phases_ms": op1["phases_ms"],
            "form": "FP64, OP BLK 1: the reference's default preconditioner, one LU of the whole matrix "
                    "(tiled band LU on the device), a bounded sample of the first time step(s)",
            "first_step_max_rel_error_vs_fp64_gpu": rel_err(op1["first"], r64["first"])},
        "list_kernel_shapes": r32["tuning"],
        "roofline": jroof,
        "spmv_roofline": {"fp32_matrix_fp64_vectors": roof[1], "fp64": roof[0], "fp32": roof[2]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)

if __name__ == "__main__":
    main()
