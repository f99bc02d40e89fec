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
(same element size and statistics), 120 steps, timed with its own "CPU time per
timestep"; the host's CPU model and core count are reported with it.

The default run also measures the two other forms of the solver, each at its
configuration's own step count (3000), and reports them under "variants":
  batch    256 rate-and-state rupture scenarios (nx ~ U{100..400}, nz = nx/2) batched
           in one persistent kernel per step under a replayed hipGraph, against the same
           scenarios stepped one by one: scenarios/s, speedup, bitwise equality;
  adjoint  the waveform gradient of a 400 x 400-element P-SV model (16 sources, 64
           receivers): gradients/s and gradient cost over a forward run for the history
           and the sqrt(N)-checkpoint storage plans, FP64 and FP32 bulk, and central
           finite differences along random directions;
  c1       the reference's own CPU-runnable case (10 km box, 100 x 100 elements, point
           force, 2000 steps, 9 receivers) run whole from a fresh solver, against the
           whole reference run on one core, with the seismograms compared.
All host-side set-up (problem building, the CPU baseline) runs before the GPU is first
used.
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
CPU_SAMPLE = dict(nx=500, nz=250, lx=50e3, lz=25e3, steps=120)


def alg_bytes(npoin, nelem, ndof, s, m=2):
    return npoin * (6 * ndof + 1) * s + nelem * 25 * (4 + m * s)


def host_cpu():
    """(model name, logical CPUs of the host, CPUs this process may run on)"""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count(), len(os.sched_getaffinity(0))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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
    model, ncpu, naff = host_cpu()
    return {"value": npoin / per_step, "unit": "GLL node-updates/s", "cores": 1, "kind": "reference",
            "sample": f"{S['nx']}x{S['nz']}-element window (npoin={npoin}) of the same medium, {S['steps']} "
                      f"steps, reference sem2dsolve flang -O2 on 1 pinned core; {per_step * 1e3:.2f} ms/step "
                      f"(its own CPU-time-per-step report; wall {wall:.1f} s incl. init)",
            "steps": S["steps"], "ms_per_step": per_step * 1e3, "cpu_model": model, "nproc": ncpu,
            "cpus_available": naff}


def cpu_c1():
    """Reference sem2dsolve on the whole 2000-step 100x100-element P-SV box (about 20 s on
    one core); returns its timing and its float32 seismograms [nsamp][2][9]."""
    exe = os.path.join(ROOT, "oracle", "_ref", "sem2dsolve")
    if not os.path.exists(exe):
        return None, None
    from sem2dpack_amd import workloads
    import numpy as np
    with tempfile.TemporaryDirectory(prefix="sem2d_c1_") as tmp:
        with open(os.path.join(tmp, "Par.inp"), "w") as fh:
            fh.write(workloads.par_c1())
        t0 = time.time()
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec taskset -c 0 {exe}"], cwd=tmp,
                           capture_output=True, text=True, timeout=600)
        wall = time.time() - t0
        out = r.stdout
        per = re.findall(r"CPU time per timestep[ .]*([0-9.+\-Ee]+)", out)
        tot = re.findall(r"Total solver CPU time[ .]*([0-9.+\-Ee]+)", out)
        if r.returncode != 0 or "FATAL" in out or not per:
            return {"value": None, "error": (out + r.stderr)[-400:]}, None
        seis = np.stack([np.fromfile(os.path.join(tmp, n), np.float32).reshape(9, -1).T
                         for n in ("Ux_sem2d.dat", "Uz_sem2d.dat")], axis=1)
    return {"ms_per_step": float(per[-1]) * 1e3, "solver_s": float(tot[-1]) if tot else None, "wall_s": wall,
            "cores": 1, "kind": "reference",
            "sample": "the whole 2000-step run, reference sem2dsolve flang -O2 on 1 pinned core"}, seis


def run_c1(pb, cpu, ref_seis):
    """The 100x100-element P-SV point-force box: all 2000 steps on the GPU, from a fresh
    solver, with its 9 velocity receivers recorded every step."""
    import numpy as np
    from sem2dpack_amd import solver
    t0 = time.perf_counter()
    s = solver.Solver(pb)
    s.fields()
    t1 = time.perf_counter()
    s.step(pb.nt)
    t2 = time.perf_counter()
    seis = s.seismograms()
    ev_ms = s.last_step_ms()
    npoin = pb.npoin
    s.close()
    r = {"workload": "P-SV 10x10 km box 100x100 elements NGLL=5, point force, 4 ABCs, 2000 steps, 9 receivers",
         "npoin": npoin, "steps": pb.nt, "setup_s": t1 - t0, "run_s": t2 - t1,
         "ms_per_step": 1e3 * (t2 - t1) / pb.nt, "event_ms_per_step": ev_ms,
         "value": npoin * pb.nt / (t2 - t1), "unit": "GLL node-updates/s", "dtype": "f64",
         "finite": bool(np.isfinite(seis).all())}
    if cpu and cpu.get("ms_per_step"):
        cpu = dict(cpu)
        cpu["value"] = npoin / (cpu["ms_per_step"] * 1e-3)
        cpu["unit"] = "GLL node-updates/s"
        r["cpu_baseline"] = cpu
        r["speedup_vs_cpu_run"] = cpu["solver_s"] / (t2 - t1) if cpu.get("solver_s") else None
    if ref_seis is not None and ref_seis.shape == seis.shape:
        # the reference's seismogram files hold float32: agreement to ~1e-7 is all they show
        r["seismo_rel_l2_vs_reference_f32"] = float(np.linalg.norm(seis - ref_seis) / np.linalg.norm(ref_seis))
    return r


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


# arrays of a built problem that the solver does not read (coordinates, material
# fields, raw masses): dropped before a built problem is sent back from a worker
_UNUSED_BY_SOLVER = ("coord", "mass_raw", "mass", "rho", "cp", "cs", "lam", "mu", "dvol", "tags", "perm")


def _build_rupture_job(job):
    from sem2dpack_amd import workloads
    sc, steps = job
    pb = workloads.build_rupture(sc, nt=steps)
    for k in _UNUSED_BY_SOLVER:
        if hasattr(pb, k):
            setattr(pb, k, None)
    return pb


def build_batch_problems(n, steps, workers):
    """The n rupture-ensemble problems, built by a pool of forked workers (host only: call
    before the GPU is first used in this process)."""
    import multiprocessing as mp
    from sem2dpack_amd import workloads
    scen = workloads.rupture_ensemble(n)
    t0 = time.time()
    pbs = []
    jobs = [(sc, steps) for sc in scen]
    if workers > 1:
        with mp.get_context("fork").Pool(workers) as pool:
            for k, pb in enumerate(pool.imap(_build_rupture_job, jobs, chunksize=2)):
                pbs.append(pb)
                if k % 32 == 31:
                    log(f"batch: built {k + 1}/{n} scenarios ({time.time() - t0:.0f} s)")
    else:
        for k, job in enumerate(jobs):
            pbs.append(_build_rupture_job(job))
            if k % 32 == 31:
                log(f"batch: built {k + 1}/{n} scenarios ({time.time() - t0:.0f} s)")
    return pbs, time.time() - t0


def run_batch(pbs, steps, build_s):
    """Rupture ensemble: the scenarios batched (hipGraph-replayed loop, autotuned tile
    kernel) vs the same scenarios stepped one by one, FP32 bulk with FP64 friction.
    The batched time includes the graph capture; set-up (building the scenarios, the
    batch tables, tuning) is outside both timed regions.  Returns the result dict."""
    import numpy as np
    from sem2dpack_amd import solver
    n = len(pbs)
    stats = [(pb.npoin, pb.nelem, len(pb.fault["node1"])) for pb in pbs]
    members = []
    for k in range(n):
        members.append(solver.Solver(pbs[k], bulk_fp32=True, keep_problem=False))
        pbs[k] = None  # the device solver holds what it needs
    npoin = sum(x[0] for x in stats)
    nelem = sum(x[1] for x in stats)
    nflt = sum(x[2] for x in stats)
    b = solver.Batch(members)
    choice = b.tune(trial_steps=5)
    cfg = b.config()
    log(f"batch: {n} scenarios, npoin {npoin}, tuned {cfg}")
    t1 = time.perf_counter()
    b.step(steps)
    t_batch = time.perf_counter() - t1
    ev_ms = b.last_step_ms()
    log(f"batch: {steps} steps in {t_batch:.2f} s")
    # the same scenarios, one after the other, each with its own solver
    t_solo = 0.0
    for k, m in enumerate(members):
        m.step(0)  # host-side table set-up, outside the timed region as for the batch
        t2 = time.perf_counter()
        m.step(steps)
        t_solo += time.perf_counter() - t2
        if k % 32 == 31:
            log(f"batch: one by one {k + 1}/{n} ({t_solo:.1f} s)")
    # equal accuracy: bitwise identical results
    same = True
    for k in range(n):
        fb, fs = b.fault_state(k), members[k].fault_state()
        same &= all(np.array_equal(fb[q], fs[q]) for q in fs)
        same &= np.array_equal(b.potency(k), members[k].potency())
        if k % 16 == 0:
            same &= all(np.array_equal(x, y) for x, y in zip(b.fields(k), members[k].fields()))
    b.close()
    for m in members:
        m.close()
    B = batch_bytes(npoin, nelem, nflt)
    return {
        "metric": "batched scenarios/s", "value": n / t_batch, "unit": "scenarios/s",
        "steps": steps, "ms_per_step": 1e3 * t_batch / steps,
        "dtype": "f32 bulk, f64 friction",
        "workload": f"{n} RSF rupture scenarios, nx~U{{100..400}}, nz=nx/2, NGLL=5, {steps} steps",
        "npoin_total": npoin, "nelem_total": nelem, "fault_pairs_total": nflt, "build_s": round(build_s, 1),
        "tile_config": cfg, "tune_choice": choice,
        "one_by_one_s": t_solo, "batched_s": t_batch,
        "speedup_vs_one_by_one": t_solo / t_batch,
        "bitwise_equal_to_one_by_one": bool(same),
        "node_updates_per_s": npoin * steps / t_batch,
        "roofline": {"bound": "hbm", "achieved": B / (ev_ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": B / (ev_ms * 1e-3) / HBM_PEAK, "traffic": None,
                     "alg_bytes_per_step": B, "event_ms_per_step": ev_ms},
    }


def run_adjoint(nt, fd_dirs):
    """Waveform gradient with respect to per-element (rho, lambda, mu): one gradient =
    forward run storing what the backward sweep needs + the adjoint sweep.  Timed for
    both storage plans (the displacement history when it fits; sqrt(N) checkpoints
    with segment replays) against a plain forward run of the same precision, and
    checked against central differences along random directions.  Returns the dict."""
    import numpy as np
    from sem2dpack_amd import adjoint, solver, workloads
    work = tempfile.mkdtemp(prefix="sem2d_adj_")
    path, back, true = workloads.adjoint_setup(work, nx=400, nz=400, nt=nt, nsrc=16, nrec=64, f0=3.0)
    t0 = time.time()
    model = adjoint.SeismogramModel(path)
    obs = model.forward(*true)
    pb = model.problem(*back)
    model.close()  # its stored forward states are not needed any more
    build_s = time.time() - t0
    log(f"adjoint: problem npoin={pb.npoin} nelem={pb.nelem} nt={nt}")

    def plain_forward(fp32):
        s = solver.Solver(pb, bulk_fp32=fp32)
        s.step(5)
        s.close()
        s = solver.Solver(pb, bulk_fp32=fp32)
        t1 = time.perf_counter()
        s.step(nt)
        t = time.perf_counter() - t1
        ev = s.last_step_ms() * nt
        s.close()
        return t, ev

    t_fwd, fwd_ev = plain_forward(False)
    t_fwd32, fwd32_ev = plain_forward(True)
    seg_ck = int(math.ceil(math.sqrt(nt)))
    out = {}
    for fp32 in (False, True):
        for seg in (0, seg_ck):
            m = adjoint.SeismogramModel(path, bulk_fp32=fp32, seg=seg)
            g_s = solver.Solver(pb, bulk_fp32=fp32)
            st = adjoint.ForwardState(pb, g_s, tuple(np.ascontiguousarray(a) for a in back))
            t2 = time.perf_counter()
            g_s.adjoint_init(nt, seg)
            plan = g_s.adjoint_plan()
            g_s.adjoint_forward()
            seis = g_s.seismograms()
            grad = m.backward(seis - obs, st)
            t_grad = time.perf_counter() - t2
            out[(fp32, seg)] = dict(grad=grad, t=t_grad, ev=m.last_timing,
                                    plan={"history": bool(plan[0]), "segment": plan[1]})
            st.close()
            log(f"adjoint: fp32={fp32} seg={seg}: gradient {t_grad:.3f} s, plan {plan}, device ms {m.last_timing}")
    adjoint.release()
    # central differences along random directions (FP64 forward), against both plans
    rng = np.random.Generator(np.random.PCG64(20260502))
    fdm = adjoint.SeismogramModel(path)
    errs = {k: [] for k in out}
    eps = 1e-2
    for k in range(fd_dirs):
        dm = [rng.standard_normal(p.size) * p * 1e-2 for p in back]
        Jp = 0.5 * float(np.sum((fdm.forward(*[p + eps * q for p, q in zip(back, dm)]) - obs) ** 2))
        Jm = 0.5 * float(np.sum((fdm.forward(*[p - eps * q for p, q in zip(back, dm)]) - obs) ** 2))
        fd = (Jp - Jm) / (2 * eps)
        for key in out:
            an = sum(np.dot(a, b) for a, b in zip(out[key]["grad"], dm))
            errs[key].append(abs(an - fd) / abs(fd))
        log(f"adjoint: direction {k}: FD {fd:.6e}, rel err fp64 history {errs[(False, 0)][-1]:.2e}")
    fdm.close()

    def plan_row(fp32, seg, t_ref):
        r = out[(fp32, seg)]
        return {"gradient_s": r["t"], "gradients_per_s": 1.0 / r["t"], "gradient_over_forward": r["t"] / t_ref,
                "plan": r["plan"], "device_ms": {"forward_storing": r["ev"][0], "backward": r["ev"][1]},
                "fd_max_rel_err": max(errs[(fp32, seg)]) if fd_dirs else None}

    t_best = out[(False, 0)]["t"]
    return {
        "metric": "adjoint gradients/s", "value": 1.0 / t_best, "unit": "gradients/s", "steps": nt,
        "dtype": "f64",
        "workload": "P-SV 400x400 elements NGLL=5, 16 simultaneous Ricker forces (3 Hz), 64 surface receivers, "
                    "L2 waveform misfit, gradient w.r.t. per-element rho/lambda/mu",
        "npoin": pb.npoin, "nelem": pb.nelem, "build_s": round(build_s, 1),
        "forward_s": t_fwd, "forward_device_ms": fwd_ev, "gradient_over_forward": t_best / t_fwd,
        "fp64": {"history": plan_row(False, 0, t_fwd), "checkpoints": plan_row(False, seg_ck, t_fwd)},
        "fp32_bulk": {"forward_s": t_fwd32, "forward_device_ms": fwd32_ev,
                      "history": plan_row(True, 0, t_fwd32), "checkpoints": plan_row(True, seg_ck, t_fwd32)},
        "fd_check": {"directions": fd_dirs, "eps_rel": eps * 1e-2},
    }


def copy_calibration():
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps of the throughput case (default 1000)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=["all", "throughput", "c1", "batch", "adjoint"], default="all",
                    help="all (default): the 2000x1000-element P-SV throughput case with the batch and "
                         "adjoint forms and the 100x100-element box (c1) as variants; or one of them alone")
    ap.add_argument("--form-steps", type=int, default=3000,
                    help="steps of the batch and adjoint forms (their configurations: 3000)")
    ap.add_argument("--fd-dirs", type=int, default=4, help="adjoint: finite-difference directions")
    ap.add_argument("--scenarios", type=int, default=256)
    ap.add_argument("--build-workers", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="time a 1 GiB device-to-device copy (reference for HBM counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not an implemented path.", file=sys.stderr)
        sys.exit(2)
    wl = args.workload
    steps = 1000 if args.steps is None else args.steps

    # ---- host-side set-up, before the GPU is first used ----
    pb = cpu = batch_pbs = None
    build_s = batch_build_s = 0.0
    if wl in ("all", "throughput"):
        from sem2dpack_amd import workloads
        t0 = time.time()
        pb = workloads.build_throughput(nt=steps + args.warmup)
        build_s = time.time() - t0
        log(f"throughput problem built in {build_s:.1f} s")
        if not args.no_cpu_baseline:
            cpu = cpu_baseline()
            log(f"cpu baseline: {cpu and cpu.get('ms_per_step')} ms/step")
    c1_pb = c1_cpu = c1_ref = None
    if wl in ("all", "c1"):
        from sem2dpack_amd import workloads
        c1_pb = workloads.build_c1()
        if not args.no_cpu_baseline:
            c1_cpu, c1_ref = cpu_c1()
            log(f"c1 cpu baseline: {c1_cpu and c1_cpu.get('ms_per_step')} ms/step")
    if wl in ("all", "batch"):
        batch_pbs, batch_build_s = build_batch_problems(args.scenarios, args.form_steps, args.build_workers)
        log(f"batch problems built in {batch_build_s:.1f} s")

    # ---- GPU ----
    if wl == "batch":
        r = run_batch(batch_pbs, args.form_steps, batch_build_s)
        r.update({"n_gpus": 1, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                  "data": "synthetic", "config": {"workload": r.pop("workload")}})
        print(json.dumps(r))
        return
    if wl == "c1":
        r = run_c1(c1_pb, c1_cpu, c1_ref)
        r.update({"metric": "GLL node-updates/s", "n_gpus": 1, "warmup": 0, "higher_is_better": True,
                  "scaling": "weak", "vs_baseline": None, "data": "synthetic",
                  "config": {"workload": r.pop("workload")}})
        print(json.dumps(r))
        return
    if wl == "adjoint":
        r = run_adjoint(args.form_steps, args.fd_dirs)
        r.update({"n_gpus": 1, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                  "data": "synthetic", "config": {"workload": r.pop("workload")}})
        print(json.dumps(r))
        return

    npoin, nelem, ndof = pb.npoin, pb.nelem, pb.ndof
    wall, ev_ms, finite = run_gpu(pb, steps, args.warmup, False)
    log(f"throughput fp64: {1e3 * wall / steps:.3f} ms/step")
    B64 = alg_bytes(npoin, nelem, ndof, 8)
    value = npoin * steps / wall
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
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "P-SV 2000x1000 elements NGLL=5, layered random medium, explosion, 4 ABCs",
                   "npoin": npoin, "nelem": nelem, "dt": pb.dt, "bulk": "fp64", "build_s": round(build_s, 1)},
        "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": ach / HBM_PEAK, "traffic": traffic,
                     "traffic_source": "profiles/c3_fp64_hbm_traffic.json: rocprofv3 FETCH_SIZE + WRITE_SIZE of "
                                       "the step's kernels (scaled by a copy kernel), not measured in this run",
                     "alg_bytes_per_step": B64, "event_ms_per_step": ev_ms},
        "cpu_baseline": cpu,
        "finite": finite,
        "variants": {},
    }
    if not args.no_fp32:
        wall32, ev32, fin32 = run_gpu(pb, steps, args.warmup, True)
        log(f"throughput fp32 bulk: {1e3 * wall32 / steps:.3f} ms/step")
        B32 = alg_bytes(npoin, nelem, ndof, 4)
        line["variants"]["fp32_bulk"] = {
            "value": npoin * steps / wall32, "ms_per_step": 1e3 * wall32 / steps,
            "roofline_frac": B32 / (ev32 * 1e-3) / HBM_PEAK, "achieved_GBs": B32 / (ev32 * 1e-3) / 1e9,
            "alg_bytes_per_step": B32, "finite": fin32}
    pb = None
    if wl == "all":
        line["variants"]["c1"] = run_c1(c1_pb, c1_cpu, c1_ref)
        log(f"c1: {line['variants']['c1']['ms_per_step']:.4f} ms/step")
        c1_pb = None
        line["variants"]["batch"] = run_batch(batch_pbs, args.form_steps, batch_build_s)
        batch_pbs = None
        line["variants"]["adjoint"] = run_adjoint(args.form_steps, args.fd_dirs)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
