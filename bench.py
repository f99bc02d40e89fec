#!/usr/bin/env python3
"""Throughput benchmark of the MI355X explicit SEM time step.

Workload (the large in-plane throughput configuration): P-SV, 200 km x 100 km,
2000 x 1000 elements, NGLL = 5 (32,012,001 global GLL nodes), 8-layer medium
(vp 4000 -> 7000 m/s) with a seeded 10 % Gaussian random field (2 km correlation),
vs = vp/sqrt(3), rho = 770 + 0.32 vp, explosive Ricker source (f0 = 1 Hz), four
absorbing boundaries.  A "step" is one Newmark step of the whole mesh; inputs are
resident in HBM before the timed region.

Metric: GLL node-updates/s = npoin * steps / seconds (FP64 fields).  The same run is
repeated with the FP32 bulk update (fault/time accumulators in FP64) and reported
under "variants".

Roofline: algorithmic bytes per step B = npoin*(6*ndof+1)*s + nelem*25*(4+2*s)
(d, v, a read and written, one inverse-mass word per node, int32 ibool and two
moduli per element-local point; s = 8 or 4), divided by the step time measured
with HIP events on the solver's stream, against 8 TB/s.

CPU baseline: the reference SEM2DPACK Fortran solver (built from source by
oracle/build_ref.py) on one core, on a 500 x 250-element window of the same medium
(same element size and statistics), timed with its own "CPU time per timestep".
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
CPU_SAMPLE = dict(nx=500, nz=250, lx=50e3, lz=25e3, steps=60)


def alg_bytes(npoin, nelem, ndof, s, m=2):
    return npoin * (6 * ndof + 1) * s + nelem * 25 * (4 + m * s)


def cpu_baseline():
    """Reference Fortran solver on a window of the medium; returns dict or None."""
    exe = os.path.join(ROOT, "oracle", "_ref", "sem2dsolve")
    if not os.path.exists(exe):
        return None
    from sem2dpack_amd import workloads
    S = CPU_SAMPLE
    with tempfile.TemporaryDirectory(prefix="sem2d_cpu_") as tmp:
        grid = workloads.layered_random_grid((0.0, S["lx"]), (0.0, S["lz"]), dx=100.0)
        workloads.write_hete1(os.path.join(tmp, "c3_media.tab"), *grid)
        with open(os.path.join(tmp, "Par.inp"), "w") as fh:
            fh.write(workloads.par_throughput(S["nx"], S["nz"], S["lx"], S["lz"], S["steps"]))
        cmd = f"ulimit -s unlimited; exec taskset -c 0 {exe}"
        t0 = time.time()
        r = subprocess.run(["bash", "-c", cmd], cwd=tmp, capture_output=True, text=True, timeout=600)
        wall = time.time() - t0
        out = r.stdout
        m = re.findall(r"CPU time per timestep[ .]*([0-9.+\-Ee]+)", out)
        if r.returncode != 0 or "FATAL" in out or not m:
            return {"value": None, "error": (out + r.stderr)[-400:]}
        per_step = float(m[-1])
        npoin = (4 * S["nx"] + 1) * (4 * S["nz"] + 1)
    return {"value": npoin / per_step, "unit": "GLL node-updates/s", "cores": 1, "kind": "reference",
            "sample": f"{S['nx']}x{S['nz']}-element window (npoin={npoin}) of the same medium, {S['steps']} "
                      f"steps, reference sem2dsolve flang -O2 on 1 pinned core; {per_step * 1e3:.2f} ms/step "
                      f"(its own CPU-time-per-step report; wall {wall:.1f} s incl. init)"}


def run_gpu(pb, steps, warmup, fp32):
    from sem2dpack_amd import solver
    s = solver.Solver(pb, bulk_fp32=fp32)
    s.step(warmup)
    s.fields()
    t0 = time.perf_counter()
    s.step(steps)
    t1 = time.perf_counter()
    ev_ms = s.last_step_ms()
    d, v, a = s.fields()
    s.close()
    import numpy as np
    finite = bool(np.isfinite(v).all())
    return (t1 - t0), ev_ms, finite


def batch_bytes(npoin, nelem, nflt, ndof=2, s=4):
    """Algorithmic bytes of one batched step: fields and inverse mass per node, int32
    ibool per element-local point (homogeneous scenarios: no per-point moduli) and
    ~0.5 KB per FP64 fault pair."""
    return npoin * (6 * ndof + 1) * s + nelem * 25 * 4 + nflt * 512


def run_batch(args):
    """Rupture ensemble: n scenarios batched (hipGraph-replayed loop, autotuned tile
    kernel) vs the same scenarios stepped one by one, FP32 bulk with FP64 friction.
    The batched time includes the graph capture; set-up (building the scenarios, the
    batch tables, tuning) is outside both timed regions."""
    import numpy as np
    from sem2dpack_amd import solver, workloads
    n, steps = args.scenarios, args.steps
    scen = workloads.rupture_ensemble(n)
    t0 = time.time()
    members, stats = [], []
    for k, sc in enumerate(scen):
        pb = workloads.build_rupture(sc, nt=steps)
        stats.append((pb.npoin, pb.nelem, len(pb.fault["node1"])))
        members.append(solver.Solver(pb, bulk_fp32=True, keep_problem=False))
        if k % 32 == 31:
            print(f"[batch] built {k + 1}/{n} scenarios ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
    build_s = time.time() - t0
    npoin = sum(x[0] for x in stats)
    nelem = sum(x[1] for x in stats)
    nflt = sum(x[2] for x in stats)
    b = solver.Batch(members)
    choice = b.tune(trial_steps=5)
    cfg = b.config()
    print(f"[batch] tuned: {cfg}", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    b.step(steps)
    t_batch = time.perf_counter() - t1
    ev_ms = b.last_step_ms()
    print(f"[batch] {steps} steps in {t_batch:.2f} s", file=sys.stderr, flush=True)
    # the same scenarios, one after the other, each with its own solver
    t_solo = 0.0
    for k, m in enumerate(members):
        m.step(0)  # host-side table set-up, outside the timed region as for the batch
        t2 = time.perf_counter()
        m.step(steps)
        t_solo += time.perf_counter() - t2
        if k % 32 == 31:
            print(f"[batch] one-by-one {k + 1}/{n} ({t_solo:.1f} s)", file=sys.stderr, flush=True)
    # equal accuracy: bitwise identical results
    same = True
    for k in range(n):
        fb, fs = b.fault_state(k), members[k].fault_state()
        same &= all(np.array_equal(fb[q], fs[q]) for q in fs)
        same &= np.array_equal(b.potency(k), members[k].potency())
        if k % 16 == 0:
            same &= all(np.array_equal(x, y) for x, y in zip(b.fields(k), members[k].fields()))
    B = batch_bytes(npoin, nelem, nflt)
    line = {
        "metric": "batched scenarios/s",
        "value": n / t_batch,
        "unit": "scenarios/s",
        "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": 1e3 * t_batch / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 bulk, f64 friction", "data": "synthetic",
        "config": {"workload": f"{n} RSF rupture scenarios, nx~U{{100..400}}, nz=nx/2, NGLL=5, "
                               f"{steps} steps", "npoin_total": npoin, "nelem_total": nelem,
                   "fault_pairs_total": nflt, "build_s": round(build_s, 1), "tile_config": cfg,
                   "tune_choice": choice},
        "one_by_one_s": t_solo,
        "speedup_vs_one_by_one": t_solo / t_batch,
        "bitwise_equal_to_one_by_one": bool(same),
        "node_updates_per_s": npoin * steps / t_batch,
        "roofline": {"bound": "hbm", "achieved": B / (ev_ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": B / (ev_ms * 1e-3) / HBM_PEAK, "traffic": None,
                     "alg_bytes_per_step": B, "event_ms_per_step": ev_ms},
    }
    print(json.dumps(line))


def run_adjoint(args):
    """Waveform gradient with respect to per-element (rho, lambda, mu): one gradient =
    checkpointed forward + segment replays + adjoint sweep.  Reported against a plain
    forward run of the same problem, and checked against central differences along
    random directions (FP64 and FP32 bulk)."""
    import numpy as np
    from sem2dpack_amd import adjoint, solver, workloads
    nt = args.steps
    work = tempfile.mkdtemp(prefix="sem2d_adj_")
    path, back, true = workloads.adjoint_setup(work, nx=400, nz=400, nt=nt, nsrc=16, nrec=64, f0=3.0)
    t0 = time.time()
    model = adjoint.SeismogramModel(path)
    obs = model.forward(*true)
    pb = model.problem(*back)
    model.close()  # its stored forward states are not needed any more
    build_s = time.time() - t0
    print(f"[adjoint] problem: npoin={pb.npoin} nelem={pb.nelem} nt={nt}", file=sys.stderr, flush=True)
    # plain forward (the reference's time loop) on the background model
    s = solver.Solver(pb)
    s.step(5)
    s.close()
    s = solver.Solver(pb)
    t1 = time.perf_counter()
    s.step(nt)
    t_fwd = time.perf_counter() - t1
    fwd_ev_ms = s.last_step_ms() * nt
    s.close()
    # the same with FP32 fields, the reference point of the FP32 gradient
    s = solver.Solver(pb, bulk_fp32=True)
    s.step(5)
    s.close()
    s = solver.Solver(pb, bulk_fp32=True)
    t1 = time.perf_counter()
    s.step(nt)
    t_fwd32 = time.perf_counter() - t1
    fwd32_ev_ms = s.last_step_ms() * nt
    s.close()
    out = {}
    for fp32 in (False, True):
        m = adjoint.SeismogramModel(path, bulk_fp32=fp32)
        g_s = solver.Solver(pb, bulk_fp32=fp32)
        m._state = adjoint.ForwardState(pb, g_s, tuple(np.ascontiguousarray(a) for a in back))
        t2 = time.perf_counter()
        g_s.adjoint_init(nt, 0)
        plan = g_s.adjoint_plan()
        t_init = time.perf_counter() - t2
        g_s.adjoint_forward()
        seis = g_s.seismograms()
        t_fw = time.perf_counter() - t2 - t_init
        grad = m.backward(seis - obs)
        t_grad = time.perf_counter() - t2
        out[fp32] = dict(grad=grad, t=t_grad, ev=m.last_timing, J=0.5 * float(np.sum((seis - obs) ** 2)),
                         plan={"history": plan[0], "segment": plan[1]},
                         host_s={"init": t_init, "forward": t_fw, "backward": t_grad - t_init - t_fw})
        m.close()
        print(f"[adjoint] fp32={fp32}: gradient in {t_grad:.2f} s (forward {t_fwd:.2f} s), plan {plan}, "
              f"device ms {m.last_timing}, wall s {out[fp32]['host_s']}", file=sys.stderr, flush=True)
    # central differences along random directions (FP64 forward)
    rng = np.random.Generator(np.random.PCG64(20260502))
    fdm = adjoint.SeismogramModel(path)
    errs64, errs32 = [], []
    eps = 1e-2
    for k in range(args.fd_dirs):
        dm = [rng.standard_normal(p.size) * p * 1e-2 for p in back]
        Jp = 0.5 * float(np.sum((fdm.forward(*[p + eps * q for p, q in zip(back, dm)]) - obs) ** 2))
        Jm = 0.5 * float(np.sum((fdm.forward(*[p - eps * q for p, q in zip(back, dm)]) - obs) ** 2))
        fd = (Jp - Jm) / (2 * eps)
        for fp32, errs in ((False, errs64), (True, errs32)):
            an = sum(np.dot(a, b) for a, b in zip(out[fp32]["grad"], dm))
            errs.append(abs(an - fd) / abs(fd))
        print(f"[adjoint] direction {k}: FD {fd:.6e} rel err fp64 {errs64[-1]:.2e} fp32 {errs32[-1]:.2e}",
              file=sys.stderr, flush=True)
    fdm.close()
    t_grad = out[False]["t"]
    ev_f, ev_b = out[False]["ev"]
    line = {
        "metric": "adjoint gradients/s", "value": 1.0 / t_grad, "unit": "gradients/s",
        "n_gpus": 1, "steps": nt, "warmup": 0, "ms_per_step": 1e3 * t_grad / nt,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "P-SV 400x400 elements NGLL=5, 16 simultaneous Ricker forces (3 Hz), 64 surface "
                               "receivers, L2 waveform misfit, gradient w.r.t. per-element rho/lambda/mu",
                   "npoin": pb.npoin, "nelem": pb.nelem, "storage_plan": out[False]["plan"],
                   "build_s": round(build_s, 1)},
        "forward_s": t_fwd, "gradient_s": t_grad, "gradient_over_forward": t_grad / t_fwd,
        "device_ms": {"forward_plain": fwd_ev_ms, "forward_storing": ev_f, "backward": ev_b},
        "fp32_bulk": {"gradient_s": out[True]["t"], "forward_s": t_fwd32,
                      "gradient_over_forward": out[True]["t"] / t_fwd32,
                      "gradient_over_fp64_forward": out[True]["t"] / t_fwd,
                      "device_ms": {"forward_plain": fwd32_ev_ms, "forward_storing": out[True]["ev"][0],
                                    "backward": out[True]["ev"][1]},
                      "storage_plan": out[True]["plan"]},
        "fd_check": {"directions": args.fd_dirs, "eps_rel": eps * 1e-2,
                     "max_rel_err_fp64": max(errs64) if errs64 else None,
                     "max_rel_err_fp32": max(errs32) if errs32 else None},
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=["throughput", "batch", "adjoint"], default="throughput",
                    help="throughput: the 2000x1000-element P-SV case (default); batch: the "
                         "rupture ensemble, batched vs one by one; adjoint: waveform gradients")
    ap.add_argument("--fd-dirs", type=int, default=20, help="adjoint: finite-difference directions")
    ap.add_argument("--scenarios", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="time a 1 GiB device-to-device copy (reference for HBM counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 27  # 2^27 doubles = 1 GiB
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / 10
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "ms": 1e3 * t,
                          "GBps_rw": 16 * n / t / 1e9}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not an implemented path.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        args.steps = 3000 if args.steps is None else args.steps
        run_batch(args)
        return
    if args.workload == "adjoint":
        args.steps = 3000 if args.steps is None else args.steps
        run_adjoint(args)
        return
    args.steps = 1000 if args.steps is None else args.steps

    from sem2dpack_amd import workloads
    t0 = time.time()
    pb = workloads.build_throughput(nt=args.steps + args.warmup)
    build_s = time.time() - t0
    cpu = None if args.no_cpu_baseline else cpu_baseline()

    npoin, nelem, ndof = pb.npoin, pb.nelem, pb.ndof
    wall, ev_ms, finite = run_gpu(pb, args.steps, args.warmup, False)
    B64 = alg_bytes(npoin, nelem, ndof, 8)
    value = npoin * args.steps / wall
    ach = B64 / (ev_ms * 1e-3)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "c3_fp64_hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_step")
    line = {
        "metric": "GLL node-updates/s",
        "value": value,
        "unit": "GLL node-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "P-SV 2000x1000 elements NGLL=5, layered random medium, explosion, 4 ABCs",
                   "npoin": npoin, "nelem": nelem, "dt": pb.dt, "bulk": "fp64", "build_s": round(build_s, 1)},
        "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": ach / HBM_PEAK, "traffic": traffic,
                     "alg_bytes_per_step": B64, "event_ms_per_step": ev_ms},
        "cpu_baseline": cpu,
        "finite": finite,
    }
    if not args.no_fp32:
        wall32, ev32, fin32 = run_gpu(pb, args.steps, args.warmup, True)
        B32 = alg_bytes(npoin, nelem, ndof, 4)
        line["variants"] = {"fp32_bulk": {
            "value": npoin * args.steps / wall32, "ms_per_step": 1e3 * wall32 / args.steps,
            "roofline_frac": B32 / (ev32 * 1e-3) / HBM_PEAK, "achieved_GBs": B32 / (ev32 * 1e-3) / 1e9,
            "alg_bytes_per_step": B32, "finite": fin32}}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
