"""Hand-written adjoint (sem2d_adjoint_*) vs finite differences.

Problem: 40x40 elements of 100 m, free surface, 3 absorbing sides, 4 simultaneous
vertical Ricker forces (10 Hz), 8 surface velocity receivers, 400 steps.  Misfit
J(m) = 1/2 sum (seis(m) - seis(m_true))^2 over samples, components and receivers,
m = per-element (rho, lambda, mu), evaluated at the smooth background.

Bars: the directional derivative g.dm of the FP64 adjoint gradient agrees with the
central difference (J(m + e dm) - J(m - e dm)) / 2e within 1e-3 (relative), both for
the GPU forward and for the oracle (the C restatement of the reference, on the CPU);
the gradient does not depend on the checkpoint segment length (bitwise); the default
plan (displacement history, no replay) agrees with the checkpoint plan to round-off; the
FP32 bulk gradient is reported against the FP64 one.
"""
import numpy as np
import pytest

from sem2dpack_amd import adjoint, builder, parfile, solver, workloads

pytestmark = pytest.mark.gpu

NT = 400


@pytest.fixture(scope="module")
def prob(tmp_path_factory):
    d = tmp_path_factory.mktemp("adjgpu")
    path, back, true = workloads.adjoint_setup(str(d), nx=40, nz=40, nt=NT, nsrc=4, nrec=8, f0=10.0)
    model = adjoint.SeismogramModel(path)
    obs = model.forward(*true)
    seis0 = model.forward(*back)
    g = model.backward(seis0 - obs)
    return dict(path=path, dir=str(d), back=back, true=true, obs=obs, grad=g, model=model)


def misfit(seis, obs):
    r = seis - obs
    return 0.5 * float(np.sum(r * r))


def directions(back, rng):
    nel = back[0].size
    out = {}
    for name, mask in (("rho", (1, 0, 0)), ("lam", (0, 1, 0)), ("mu", (0, 0, 1)), ("all", (1, 1, 1))):
        out[name] = [m * rng.standard_normal(nel) * p * 1e-2 for m, p in zip(mask, back)]
    return out


def check(g, dm, Jp, Jm, eps, label):
    an = sum(np.dot(a, b) for a, b in zip(g, dm))
    fd = (Jp - Jm) / (2 * eps)
    err = abs(an - fd) / abs(fd)
    print(f"  {label}: adjoint {an:.8e}  FD {fd:.8e}  rel {err:.2e}", flush=True)
    assert err < 1e-3, label


def test_forward_matches_plain_stepping(prob):
    pb = prob["model"].problem(*prob["back"])
    s = solver.Solver(pb)
    s.step(NT)
    seis = s.seismograms()
    model = adjoint.SeismogramModel(prob["path"])
    assert np.array_equal(model.forward(*prob["back"]), seis)


def test_gradient_vs_gpu_central_differences(prob):
    print("\n[adjoint vs GPU central differences]", flush=True)
    model = adjoint.SeismogramModel(prob["path"])
    back, obs, g = prob["back"], prob["obs"], prob["grad"]
    eps = 1e-2
    for label, dm in directions(back, np.random.default_rng(5)).items():
        Jp = misfit(model.forward(*[p + eps * q for p, q in zip(back, dm)]), obs)
        Jm = misfit(model.forward(*[p - eps * q for p, q in zip(back, dm)]), obs)
        check(g, dm, Jp, Jm, eps, label)


def test_gradient_vs_oracle_central_differences(prob, oracle_api):
    """finite differences of the reference's forward (the oracle, on the CPU)"""
    print("\n[adjoint vs oracle central differences]", flush=True)
    par = parfile.read_par(prob["path"])
    back, obs, g = prob["back"], prob["obs"], prob["grad"]

    def J(m):
        pb = builder.build(par, fname_dir=prob["dir"], elem_props=m)
        s = solver.Solver(pb, api=oracle_api, compact=False)
        s.step(NT)
        return misfit(s.seismograms(), obs)
    eps = 1e-2
    dirs = directions(back, np.random.default_rng(9))
    for label in ("all", "rho"):
        dm = dirs[label]
        check(g, dm, J([p + eps * q for p, q in zip(back, dm)]), J([p - eps * q for p, q in zip(back, dm)]),
              eps, "oracle " + label)


def test_checkpoint_segments_do_not_change_the_gradient(prob):
    back, obs = prob["back"], prob["obs"]
    grads = []
    for seg in (7, 64, NT):
        model = adjoint.SeismogramModel(prob["path"], seg=seg)
        seis = model.forward(*back)
        grads.append(model.backward(seis - obs))
        model.close()
    for g in grads[1:]:
        for a, b in zip(g, grads[0]):
            assert np.array_equal(a, b)


def test_history_plan_matches_checkpoints(prob):
    """The default plan keeps every step's forward displacement (no replay) and takes a_n
    and the predicted velocity from differences of it; it agrees with the checkpoint
    plan (stored a_n and v_n) to round-off, in FP64 and with FP32 fields."""
    back, obs = prob["back"], prob["obs"]
    for fp32, tol in ((False, 1e-9), (True, 1e-3)):
        grads = []
        for seg, history in ((0, True), (20, False)):
            model = adjoint.SeismogramModel(prob["path"], seg=seg, bulk_fp32=fp32)
            seis = model.forward(*back)
            assert model._state.solver.adjoint_plan()[0] == history
            grads.append(model.backward(seis - obs))
            model.close()
        errs = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(*grads)]
        print(f"\n  fp32={fp32}: history vs checkpoint plan rel-L2 (rho, lam, mu): {errs}", flush=True)
        assert max(errs) < tol
    adjoint.release()  # the kept history block goes back to the device


def test_fp32_bulk_gradient(prob):
    back, obs, g64 = prob["back"], prob["obs"], prob["grad"]
    model = adjoint.SeismogramModel(prob["path"], bulk_fp32=True)
    seis = model.forward(*back)
    g32 = model.backward(seis - obs)
    errs = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(g32, g64)]
    print(f"\n  FP32 bulk gradient rel-L2 error vs FP64 (rho, lam, mu): {errs}", flush=True)
    assert max(errs) < 5e-2


def test_torch_autograd(prob):
    import torch
    back, obs, g = prob["back"], prob["obs"], prob["grad"]
    model = adjoint.SeismogramModel(prob["path"])
    ts = [torch.tensor(p, dtype=torch.float64, requires_grad=True) for p in back]
    seis = adjoint.seismograms(*ts, model)
    J = 0.5 * ((seis - torch.as_tensor(obs)) ** 2).sum()
    J.backward()
    for t, ref in zip(ts, g):
        assert np.array_equal(t.grad.numpy(), ref)


def test_torch_two_live_graphs(prob):
    """Two forwards through the torch Function before any backward (two misfit terms):
    each output keeps its own forward state, so each backward gives the gradient at its
    own model, equal to a lone forward/backward there."""
    import torch
    back, true, obs = prob["back"], prob["true"], prob["obs"]
    mid = [0.5 * (a + b) for a, b in zip(back, true)]
    model = adjoint.SeismogramModel(prob["path"])
    ta = [torch.tensor(p, dtype=torch.float64, requires_grad=True) for p in back]
    tb = [torch.tensor(p, dtype=torch.float64, requires_grad=True) for p in mid]
    sa = adjoint.seismograms(*ta, model)
    sb = adjoint.seismograms(*tb, model)  # second forward before the first backward
    o = torch.as_tensor(obs)
    (0.5 * ((sb - o) ** 2).sum()).backward()
    (0.5 * ((sa - o) ** 2).sum()).backward()
    lone = adjoint.SeismogramModel(prob["path"])
    ref_b = lone.backward(lone.forward(*mid) - obs)
    lone.close()
    for t, ref in zip(ta, prob["grad"]):
        assert np.array_equal(t.grad.numpy(), ref)
    for t, ref in zip(tb, ref_b):
        assert np.array_equal(t.grad.numpy(), ref)
    assert not any(np.array_equal(x, y) for x, y in zip(prob["grad"], ref_b))
