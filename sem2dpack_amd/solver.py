"""Python driver of the C ABI (``include/sem2d.h``) for a built :class:`~builder.Problem`.

:class:`Solver` plays the role of the reference's time loop (``SRC/main.f90:51-99``):
it creates the device solver from the problem arrays, registers absorbing
boundaries, sources, receivers and the fault in the reference's order, then
advances and fetches.  It always drives the HIP library (``libsem2d_hip.so``); a
missing library raises :class:`~_lib.NativeLibraryMissing`.
"""
import ctypes as C

import numpy as np

from . import capi
from ._lib import load_hip_lib


class SolverError(RuntimeError):
    pass


def _p(a, t=np.float64):
    return a.ctypes.data_as(C.POINTER(C.c_double if t == np.float64 else C.c_int))


class Solver:
    FIELD_CODE = {"D": 0, "V": 1, "A": 2}

    def __init__(self, pb, bulk_fp32=False, api=None, structured=True, compact=True, keep_problem=True,
                 itinfo=0):
        self.api = api if api is not None else load_hip_lib()
        self.pb = pb
        self.ndof, self.npoin, self.nelem = pb.ndof, pb.npoin, pb.nelem
        self._keep = []
        ibool = np.ascontiguousarray(pb.ibool, dtype=np.int32)
        hprime = np.ascontiguousarray(pb.hprime.T, dtype=np.float64)  # [j][i]
        rmass = np.ascontiguousarray(pb.rmass, dtype=np.float64)
        d = capi.Desc()
        d.ndof, d.ngll, d.nelem, d.npoin = pb.ndof, pb.ngll, pb.nelem, pb.npoin
        d.nelast = pb.nelast
        keep = []
        if pb.coef_mode == "shared":
            d.coef_mode = 1
        elif (compact or pb.coefs is None) and getattr(pb, "lam_mu", None) is not None:
            d.coef_mode = 2
        else:
            d.coef_mode = 0
        if d.coef_mode == 2:
            lm = np.ascontiguousarray(pb.lam_mu, dtype=np.float64)
            ge = np.ascontiguousarray(pb.geom, dtype=np.float64)
            wg = np.ascontiguousarray(pb.wgll, dtype=np.float64)
            keep += [lm, ge, wg]
            d.lam_mu, d.geom, d.wgll = _p(lm), _p(ge), _p(wg)
        else:
            coefs = np.ascontiguousarray(pb.coefs, dtype=np.float64)
            keep.append(coefs)
            d.coefs = _p(coefs)
        self.coef_mode = d.coef_mode
        d.ibool, d.hprime, d.rmass = _p(ibool, np.int32), _p(hprime), _p(rmass)
        d.dt, d.beta, d.gamma = pb.dt, pb.beta, pb.gamma
        d.bulk_fp32 = 1 if bulk_fp32 else 0
        cells = None
        if structured and getattr(pb, "mesh", None) is not None and pb.mesh.get("nx", 0) > 0:
            m = pb.mesh
            cells = np.empty(m["nx"] * m["nz"], dtype=np.int32)
            cells[pb.elem_iz * m["nx"] + pb.elem_ix] = np.arange(1, pb.nelem + 1, dtype=np.int32)
            d.nx, d.nz, d.ezflt = m["nx"], m["nz"], max(m["ezflt"], 0)
            d.elem_of_cell = _p(cells, np.int32)
        h = C.c_void_p()
        rc = self.api.create(C.byref(d), C.byref(h))
        if rc != 0:
            raise SolverError(f"{self.api.prefix}create failed ({rc})")
        self.h = h
        if itinfo:
            self.set_itinfo(pb.itinfo if itinfo is True else int(itinfo))
        self._setup()
        if not keep_problem:
            self.pb = None  # the device solver holds what it needs; free the host problem

    # ------------------------------------------------------------------
    def _check(self, rc, what):
        if rc != 0:
            msg = self.api.last_error(self.h)
            raise SolverError(f"{self.api.prefix}{what} failed ({rc}): {msg.decode() if msg else ''}")

    def _setup(self):
        pb = self.pb
        if getattr(pb, "scheme", "newmark") == "leapfrog":
            self._check(self.api.set_time_scheme(self.h, 1), "set_time_scheme")
        if getattr(pb, "scheme", "newmark") == "hht":
            self._check(self.api.set_hht(self.h, float(pb.alpha)), "set_hht")
        per = getattr(pb, "periodic", None)
        if per is not None:
            pm = np.ascontiguousarray(per["master"], dtype=np.int32)
            ps = np.ascontiguousarray(per["slave"], dtype=np.int32)
            self._check(self.api.set_periodic(self.h, len(pm), _p(pm, np.int32), _p(ps, np.int32)), "set_periodic")
        if getattr(pb, "kv_eta", None) is not None:
            eta = np.ascontiguousarray(pb.kv_eta, dtype=np.float64)
            self._check(self.api.set_kelvin_voigt(self.h, _p(eta)), "set_kelvin_voigt")
        if getattr(pb, "beta25", None) is not None:
            beta = np.ascontiguousarray(pb.beta25, dtype=np.float64)
            self._check(self.api.set_25d(self.h, _p(beta)), "set_25d")
        for ab in pb.abso:
            node = np.ascontiguousarray(ab["node"], dtype=np.int32)
            Cm = np.ascontiguousarray(ab["C"].T, dtype=np.float64)  # [ndof][n]
            self._check(self.api.set_abso(self.h, len(node), _p(node, np.int32), _p(Cm)), "set_abso")
            if ab.get("normals") is not None:
                nrm = np.ascontiguousarray(ab["normals"].T, dtype=np.float64)  # [2][n]
                self._check(self.api.set_abso_normals(self.h, len(node), _p(nrm)), "set_abso_normals")
            if ab.get("stacey") is not None:
                enode, K = ab["stacey"]
                self._check(self.api.set_abso_stacey(self.h, len(enode), _p(enode, np.int32), _p(K)),
                            "set_abso_stacey")
                if ab.get("periodic_ends"):
                    self._check(self.api.set_abso_periodic(self.h), "set_abso_periodic")
            if ab.get("wave_V") is not None:
                w = pb.wave
                xz = np.ascontiguousarray(ab["coord"].T, dtype=np.float64)  # [2][n]
                wv = np.ascontiguousarray(ab["wave_V"].T, dtype=np.float64)  # [ndof][n]
                wt = np.ascontiguousarray(ab["wave_T"].T, dtype=np.float64)
                wb = np.ascontiguousarray(ab["wave_B"], dtype=np.float64)
                par = np.array([w["s"][0], w["s"][1], w["x0"], w["z0"], w["tdelay"], w["f0"], w["t0"],
                                w["stf_ampli"]], dtype=np.float64)
                self._check(self.api.set_abso_wave(self.h, len(node), _p(xz), _p(wv), _p(wt), _p(wb), _p(par)),
                            "set_abso_wave")
        lsf = getattr(pb, "lsf", None)
        if lsf is not None:
            n1 = np.ascontiguousarray(lsf["node1"], dtype=np.int32)
            n2 = np.ascontiguousarray(lsf["node2"], dtype=np.int32)
            arrs = [np.ascontiguousarray(lsf[k].T, dtype=np.float64) for k in ("n", "B", "W1", "W2")]
            self._check(self.api.set_lsf(self.h, len(n1), _p(n1, np.int32), _p(n2, np.int32),
                                         *[_p(x) for x in arrs], float(lsf["kt"]), float(lsf["kn"])), "set_lsf")
        for s in pb.sources:
            isrc = C.c_int()
            if s["kind"] == "force":
                dirv = np.ascontiguousarray(s["dir"], dtype=np.float64)
                self._check(self.api.set_force(self.h, int(s["iglob"]), _p(dirv), C.byref(isrc)), "set_force")
            else:
                ixi = np.ascontiguousarray(s["iglob_xi"], dtype=np.int32)
                ieta = np.ascontiguousarray(s["iglob_eta"], dtype=np.int32)
                cxi = np.ascontiguousarray(s["coef_xi"], dtype=np.float64)
                ceta = np.ascontiguousarray(s["coef_eta"], dtype=np.float64)
                self._check(self.api.set_moment(self.h, ixi.shape[0], _p(ixi, np.int32), _p(ieta, np.int32),
                                                _p(cxi), _p(ceta), C.byref(isrc)), "set_moment")
            tab = np.ascontiguousarray(s["stf_table"], dtype=np.float64)
            self._check(self.api.set_stf(self.h, isrc.value, len(tab), _p(tab)), "set_stf")
        # the other boundary conditions in their order: Dirichlet boundaries before or
        # after the fault
        dirs = getattr(pb, "dirichlet", [])
        for dc in dirs:
            if not dc["after_fault"]:
                self._set_dirichlet(dc)
        if pb.fault is not None:
            self._set_fault(pb.fault)
        for dc in dirs:
            if dc["after_fault"]:
                self._set_dirichlet(dc)
        r = pb.receivers
        if r is not None and not r.get("atnode", True):
            el = np.ascontiguousarray(r["elem"], dtype=np.int32)
            w = np.ascontiguousarray(r["interp"], dtype=np.float64)
            self.nrec, self.nsamp = len(el), r["nsamp"]
            self._check(self.api.set_receivers_interp(self.h, len(el), _p(el, np.int32), _p(w), int(r["isamp"]),
                                                      self.FIELD_CODE[r["field"]], int(r["nsamp"])),
                        "set_receivers_interp")
        elif r is not None:
            node = np.ascontiguousarray(r["iglob"], dtype=np.int32)
            self.nrec, self.nsamp = len(node), r["nsamp"]
            self._check(self.api.set_receivers(self.h, len(node), _p(node, np.int32), int(r["isamp"]),
                                               self.FIELD_CODE[r["field"]], int(r["nsamp"])), "set_receivers")

    def _set_dirichlet(self, dc):
        node = np.ascontiguousarray(dc["node"], dtype=np.int32)
        self._check(self.api.set_dirichlet(self.h, len(node), _p(node, np.int32), int(dc["comp"])), "set_dirichlet")

    def _set_fault(self, f):
        nd = self.ndof
        fd = capi.FaultDesc()
        keep = []

        def arr(a, t=np.float64):
            a = np.ascontiguousarray(a, dtype=t)
            keep.append(a)
            return _p(a, t)

        npf = len(f["node1"])
        fd.npoin = npf
        one_side = f.get("node2") is None  # fault on a domain boundary (symmetry)
        fd.node1 = arr(f["node1"], np.int32)
        fd.node2 = None if one_side else arr(f["node2"], np.int32)
        fd.n1 = arr(f["n1"].T)
        fd.B, fd.invM1, fd.Z = arr(f["B"].T), arr(f["invM1"].T), arr(f["Z"].T)
        fd.invM2 = None if one_side else arr(f["invM2"].T)
        fd.T0 = arr(f["T0"].T)
        fd.cohesion = arr(f["cohesion"])
        fd.coefA2V, fd.coefA2D = f["coefA2V"], f["coefA2D"]
        fd.allow_opening = 1 if f["opening"] else 0
        if f["law"] == "TWF":
            fd.law, fd.kind, fd.healing = 3, 1, 0
            names = ()
        elif f["law"] == "SWF":
            fd.law, fd.kind, fd.healing = 1, f["swf_kind"], 1 if f["healing"] else 0
            names = ("dc", "mus", "mud", "alpha", "p")
        else:
            fd.law, fd.kind, fd.healing = 2, f["rsf_kind"], 0
            names = ("dc", "mus", "a", "b", "vstar", "vc")
        for q, nme in enumerate(names):
            fd.p[q] = arr(f[nme])
        fd.theta0 = arr(f["theta"])
        fd.V0 = arr(f["V0"][:, 0])
        fd.oit, fd.oitd, fd.oix1, fd.oixn, fd.oixd = f["oit"], f["oitd"], f["oix1"], f["oixn"], f["oixd"]
        self.nout = (f["oixn"] - f["oix1"]) // f["oixd"] + 1
        self.fault_nsamp = (self.pb.nt - f["oit"]) // f["oitd"] + 1
        self.npf = npf
        self._check(self.api.set_fault(self.h, C.byref(fd), int(self.fault_nsamp)), "set_fault")
        tw = f.get("twf")
        if tw is not None:
            par = arr([tw[k] for k in ("mus", "mud", "mu0", "x", "z", "v", "l", "t", "dc")])
            self._check(self.api.set_fault_twf(self.h, int(tw["kind"]), par, arr(f["coord"].T)), "set_fault_twf")
        if f.get("osides"):
            self._check(self.api.set_fault_sides(self.h, 1), "set_fault_sides")
        if f.get("nor_kind", 1) != 1:
            self._check(self.api.set_fault_normal(self.h, int(f["nor_kind"]), float(f["nor_coef"]),
                                                  float(f["nor_v"])), "set_fault_normal")

    # ------------------------------------------------------------------
    def step(self, n):
        self._check(self.api.step(self.h, int(n)), "step")

    def fields(self):
        sh = (self.ndof, self.npoin)
        d, v, a = np.empty(sh), np.empty(sh), np.empty(sh)
        self._check(self.api.fetch_fields(self.h, _p(d), _p(v), _p(a)), "fetch_fields")
        return d, v, a

    def set_fields(self, d=None, v=None, a=None):
        arrs = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (d, v, a)]
        ptrs = [None if x is None else _p(x) for x in arrs]
        self._check(self.api.set_fields(self.h, *ptrs), "set_fields")

    def seismograms(self):
        out = np.empty((self.nsamp, self.ndof, self.nrec))
        n = C.c_int()
        self._check(self.api.fetch_seismo(self.h, _p(out), C.byref(n)), "fetch_seismo")
        return out[:n.value]

    def fault_records(self):
        out = np.empty((self.fault_nsamp, 6, self.nout))
        n = C.c_int()
        self._check(self.api.fetch_fault(self.h, _p(out), C.byref(n)), "fetch_fault")
        return out[:n.value]

    def fault_sides(self):
        """[nsamp][4*ndof][nout] per-side records (osides = T): d of side 1, side 2, v of
        side 1, side 2, in the fault frame for P-SV."""
        out = np.empty((self.fault_nsamp, 4 * self.ndof, self.nout))
        n = C.c_int()
        self._check(self.api.fetch_fault_sides(self.h, _p(out), C.byref(n)), "fetch_fault_sides")
        return out[:n.value]

    def potency(self):
        """[nrows][2*(ndof+1)] potency rows, one per step from it = 0."""
        n = C.c_int()
        self._check(self.api.fetch_potency(self.h, None, C.byref(n)), "fetch_potency")
        out = np.empty((max(n.value, 1), 2 * (self.ndof + 1)))
        self._check(self.api.fetch_potency(self.h, _p(out), C.byref(n)), "fetch_potency")
        return out[:n.value]

    def fault_state(self):
        nd, npf = self.ndof, self.npf
        D, V = np.empty((nd, npf)), np.empty((nd, npf))
        T, Ts, MU = np.empty((2, npf)), np.empty((2, npf)), np.empty(npf)
        self._check(self.api.fetch_fault_state(self.h, _p(D), _p(V), _p(T), _p(MU), _p(Ts)), "fetch_fault_state")
        return dict(D=D, V=V, T=T, MU=MU, Tstick=Ts)

    def fint(self, d):
        d = np.ascontiguousarray(d, dtype=np.float64)
        f = np.empty_like(d)
        self._check(self.api.compute_fint(self.h, _p(d), _p(f)), "compute_fint")
        return f

    def set_itinfo(self, itinfo):
        """progress rows every itinfo steps (SRC/main.f90:73-76) and the non-finite watchdog;
        0 turns both off"""
        self._check(self.api.set_itinfo(self.h, int(itinfo)), "set_itinfo")

    def info_rows(self):
        """[nrows][4]: it, time, max |v|, max |d| of every progress row so far"""
        n = C.c_int()
        self._check(self.api.fetch_info(self.h, None, C.byref(n)), "fetch_info")
        rows = np.zeros((n.value, 4))
        if n.value:
            self._check(self.api.fetch_info(self.h, _p(rows), C.byref(n)), "fetch_info")
        return rows

    def field_info(self):
        """(max |v|, max |d|, number of non-finite values of d, v, a) of the current fields"""
        vm, dm, nb = C.c_double(), C.c_double(), C.c_int()
        self._check(self.api.field_info(self.h, C.byref(vm), C.byref(dm), C.byref(nb)), "field_info")
        return vm.value, dm.value, nb.value

    def fint_noalloc(self, d):
        """as :meth:`fint`, through sem2d_compute_fint_noalloc (no device allocation)"""
        d = np.ascontiguousarray(d, dtype=np.float64)
        f = np.empty_like(d)
        self._check(self.api.compute_fint_noalloc(self.h, _p(d), _p(f)), "compute_fint_noalloc")
        return f

    def last_step_ms(self):
        ms = C.c_double()
        self._check(self.api.last_step_ms(self.h, C.byref(ms)), "last_step_ms")
        return ms.value

    # -- adjoint (sem2d_adjoint_*) ---------------------------------------
    def adjoint_init(self, nsteps, seg=0):
        self.adj_nsteps = int(nsteps)
        self._check(self.api.adjoint_init(self.h, int(nsteps), int(seg)), "adjoint_init")

    def adjoint_forward(self):
        self._check(self.api.adjoint_forward(self.h), "adjoint_forward")

    def adjoint_backward(self, gseis):
        g = np.ascontiguousarray(gseis, dtype=np.float64)
        if g.shape != (self.nsamp, self.ndof, self.nrec):
            raise ValueError(f"gseis must have shape {(self.nsamp, self.ndof, self.nrec)}")
        self._check(self.api.adjoint_backward(self.h, _p(g)), "adjoint_backward")

    def adjoint_fetch(self):
        """(glam, gmu) per element, gmass per node, gabso [ndof][npoin]."""
        ne = self.nelem
        glam, gmu = np.empty(ne), np.empty(ne)
        gmass, gabso = np.empty(self.npoin), np.empty((self.ndof, self.npoin))
        self._check(self.api.adjoint_fetch(self.h, _p(glam), _p(gmu), _p(gmass), _p(gabso)), "adjoint_fetch")
        return glam, gmu, gmass, gabso

    def adjoint_timing(self):
        f, b = C.c_double(), C.c_double()
        self._check(self.api.adjoint_timing(self.h, C.byref(f), C.byref(b)), "adjoint_timing")
        return f.value, b.value

    def adjoint_plan(self):
        """(history, seg): history True when the forward displacements of every step are
        kept (no replay), seg the checkpoint segment length."""
        hi, sg = C.c_int(), C.c_int()
        self._check(self.api.adjoint_plan(self.h, C.byref(hi), C.byref(sg)), "adjoint_plan")
        return bool(hi.value), sg.value

    def set_tile_config(self, threads, wg_per_cu):
        """Tile-kernel launch configuration (structured grids); results do not depend on it."""
        self._check(self.api.set_tile_config(self.h, int(threads), int(wg_per_cu)), "set_tile_config")

    def close(self):
        if getattr(self, "h", None):
            self.api.destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many scenarios advanced together (``sem2d_batch_*`` of ``include/sem2d.h``).

    ``members`` are :class:`Solver` objects, set up and not yet stepped; they are kept
    alive by the batch.  Every member's results are bitwise those of its own solver.
    """

    def __init__(self, members):
        self.members = list(members)
        self.api = self.members[0].api
        n = len(self.members)
        arr = (C.c_void_p * n)(*[m.h.value for m in self.members])
        h = C.c_void_p()
        rc = self.api.batch_create(n, arr, C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.api.batch_last_error(h) if h.value else b""
            self.close()
            raise SolverError(f"batch_create failed ({rc}): {msg.decode() if msg else ''}")

    def _check(self, rc, what):
        if rc != 0:
            msg = self.api.batch_last_error(self.h)
            raise SolverError(f"batch_{what} failed ({rc}): {msg.decode() if msg else ''}")

    def tune(self, trial_steps=5):
        choice = C.c_int()
        self._check(self.api.batch_tune(self.h, int(trial_steps), C.byref(choice)), "tune")
        return choice.value

    def set_config(self, threads, wg_per_cu):
        self._check(self.api.batch_set_config(self.h, int(threads), int(wg_per_cu)), "set_config")

    def config(self):
        t, w, n = C.c_int(), C.c_int(), C.c_int()
        self._check(self.api.batch_config(self.h, C.byref(t), C.byref(w), C.byref(n)), "config")
        return dict(threads=t.value, workgroups=w.value, ntiles=n.value)

    def step(self, n):
        self._check(self.api.batch_step(self.h, int(n)), "step")

    def fields(self, m):
        s = self.members[m]
        sh = (s.ndof, s.npoin)
        d, v, a = np.empty(sh), np.empty(sh), np.empty(sh)
        self._check(self.api.batch_fetch_fields(self.h, m, _p(d), _p(v), _p(a)), "fetch_fields")
        return d, v, a

    def seismograms(self, m):
        s = self.members[m]
        out = np.empty((s.nsamp, s.ndof, s.nrec))
        n = C.c_int()
        self._check(self.api.batch_fetch_seismo(self.h, m, _p(out), C.byref(n)), "fetch_seismo")
        return out[:n.value]

    def fault_records(self, m):
        s = self.members[m]
        out = np.empty((s.fault_nsamp, 6, s.nout))
        n = C.c_int()
        self._check(self.api.batch_fetch_fault(self.h, m, _p(out), C.byref(n)), "fetch_fault")
        return out[:n.value]

    def fault_state(self, m):
        s = self.members[m]
        nd, npf = s.ndof, s.npf
        D, V = np.empty((nd, npf)), np.empty((nd, npf))
        T, Ts, MU = np.empty((2, npf)), np.empty((2, npf)), np.empty(npf)
        self._check(self.api.batch_fetch_fault_state(self.h, m, _p(D), _p(V), _p(T), _p(MU), _p(Ts)),
                    "fetch_fault_state")
        return dict(D=D, V=V, T=T, MU=MU, Tstick=Ts)

    def potency(self, m):
        """Member m's potency rows, as Solver.potency: its row of step 0, then one per step."""
        n = C.c_int()
        self._check(self.api.batch_fetch_potency(self.h, m, None, C.byref(n)), "fetch_potency")
        out = np.empty((n.value, 2 * (self.members[m].ndof + 1)))
        self._check(self.api.batch_fetch_potency(self.h, m, _p(out), C.byref(n)), "fetch_potency")
        return out

    def last_step_ms(self):
        ms = C.c_double()
        self._check(self.api.batch_last_step_ms(self.h, C.byref(ms)), "last_step_ms")
        return ms.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.api.batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
