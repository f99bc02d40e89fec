"""Gradients of seismogram functionals with respect to per-element rho, lambda, mu.

The reference has no differentiation; this module adds it on top of the drop-in
time step.  :class:`SeismogramModel` maps a model -- constant density and Lame
parameters in every element, arrays over the cells (index iz*nx + ix) -- to the
seismograms of a structured P-SV input (its sources, receivers and boundaries),
and back-propagates dJ/d(seismograms) to dJ/d(model) with the hand-written
adjoint of the HIP library (``sem2d_adjoint_*``: the forward displacement history
kept in device memory when it fits, else sqrt(N) checkpoints and replayed segments;
transposed element operator).

The library returns the sensitivities to the quantities the time step consumes:
lambda and mu of every element (through the element coefficients), the assembled
mass M and the Clayton-Engquist coefficients C (the latter including their share of
the modified mass).  The chain rule to rho, lambda, mu is done here, mirroring how
the builder forms them (``SRC/mat_mass.f90:29-61``, ``SRC/bc_abso.f90:195-207``):

  M(node)   = sum over element points at the node of rho_e * w_i w_j detJ
  C(node,c) = sum over boundary-edge points at the node of z_e,c * w_k |J1D|,
              z = rho cp = sqrt(rho (lambda + 2 mu)) along the normal,
              z = rho cs = sqrt(rho mu) along the boundary.

:func:`seismograms` exposes the map as a ``torch.autograd.Function``.
"""
import numpy as np

from . import builder, parfile, solver


class SeismogramModel:
    """seismograms(rho, lam, mu) of an input file, differentiable by the adjoint.

    ``par``: Par.inp path or text (structured P-SV, fixed ``Dt`` so that every model
    shares the time step); ``fname_dir``: where its data files are read.  ``seg``: 0
    lets the library keep every step's displacement when it fits in device memory
    (else checkpoints every ceil(sqrt(N)) steps); > 0 forces checkpoints every seg steps.
    """

    def __init__(self, par, nsteps=None, fname_dir=None, bulk_fp32=False, seg=0):
        if isinstance(par, str) and "\n" in par:
            par = parfile.parse_par(par)
        elif isinstance(par, str):
            import os
            fname_dir = fname_dir or os.path.dirname(os.path.abspath(par))
            par = parfile.read_par(par)
        self.par, self.fname_dir = par, fname_dir
        self.nsteps, self.bulk_fp32, self.seg = nsteps, bulk_fp32, seg
        self._state = None
        self.last_timing = None

    def problem(self, rho, lam, mu):
        return builder.build(self.par, nsteps=self.nsteps, fname_dir=self.fname_dir, per_element_coefs=False,
                             elem_props=(rho, lam, mu))

    def run_forward(self, rho, lam, mu):
        """(seismograms [nsamp][ndof][nrec] FP64, state): a forward run whose state
        (problem, solver with its stored forward history or checkpoints, model arrays)
        stays alive, independent of every other run, until ``state.close()`` or until the
        state is garbage-collected."""
        rho, lam, mu = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in (rho, lam, mu))
        pb = self.problem(rho, lam, mu)
        if pb.ndof != 2 or pb.receivers is None:
            raise ValueError("SeismogramModel needs a P-SV input with receivers")
        s = solver.Solver(pb, bulk_fp32=self.bulk_fp32)
        s.adjoint_init(pb.nt, self.seg)
        s.adjoint_forward()
        return s.seismograms(), ForwardState(pb, s, (rho, lam, mu))

    def forward(self, rho, lam, mu):
        """Seismograms [nsamp][ndof][nrec] (FP64) of the model; :meth:`backward` without
        a state then differentiates at this model (the previous one's state is freed)."""
        self.close()  # the previous model's stored states go before the new ones are planned
        seis, self._state = self.run_forward(rho, lam, mu)
        return seis

    def backward(self, gseis, state=None):
        """(grho, glam, gmu) over the cells, for dJ/d(seismograms) = gseis, at the model
        of ``state`` (a :meth:`run_forward` state; default: the last :meth:`forward`)."""
        st = state if state is not None else self._state
        if st is None:
            raise RuntimeError("backward before forward")
        if st.solver is None:
            raise RuntimeError("backward on a closed forward state")
        pb, s, (rho, lam, mu) = st.pb, st.solver, st.model
        s.adjoint_backward(gseis)
        self.last_timing = s.adjoint_timing()
        glam_e, gmu_e, gmass, gabso = s.adjoint_fetch()
        return self.chain(pb, rho, lam, mu, glam_e, gmu_e, gmass, gabso)

    @staticmethod
    def chain(pb, rho, lam, mu, glam_e, gmu_e, gmass, gabso):
        """Per-cell gradients from the library's sensitivities (see the module text)."""
        nx = pb.mesh["nx"]
        cell = pb.elem_iz * nx + pb.elem_ix  # builder element -> cell
        r, la, m = rho[cell], lam[cell], mu[cell]
        ib0 = getattr(pb, "_ibool0", None)
        if ib0 is None:
            ib0 = pb._ibool0 = (pb.ibool - 1).astype(np.intp)
        grho_e = np.einsum("eij,eij->e", pb.dvol, np.take(gmass, ib0))
        glam_e = glam_e.copy()
        gmu_e = gmu_e.copy()
        for ab in pb.abso:
            bidx, elem, wk = ab["contrib"]
            node = ab["node"][bidx] - 1
            re, le, me = r[elem], la[elem], m[elem]
            zp = np.sqrt(re * (le + 2.0 * me))
            zs = np.sqrt(re * me)
            for c in range(pb.ndof):
                g = gabso[c, node] * wk
                if c == ab["geo_nor"]:
                    np.add.at(grho_e, elem, g * (le + 2.0 * me) / (2.0 * zp))
                    np.add.at(glam_e, elem, g * re / (2.0 * zp))
                    np.add.at(gmu_e, elem, g * re / zp)
                else:
                    np.add.at(grho_e, elem, g * me / (2.0 * zs))
                    np.add.at(gmu_e, elem, g * re / (2.0 * zs))
        out = [np.zeros(pb.nelem) for _ in range(3)]
        for o, g in zip(out, (grho_e, glam_e, gmu_e)):
            o[cell] = g
        return tuple(out)

    def close(self):
        if self._state is not None:
            self._state.close()
            self._state = None


class ForwardState:
    """One forward run kept for its backward: the problem, the solver (holding the
    forward history or checkpoints in device memory) and the model arrays."""

    def __init__(self, pb, s, model):
        self.pb, self.solver, self.model = pb, s, model

    def close(self):
        if self.solver is not None:
            self.solver.close()
            self.solver = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _make_function():
    import torch

    class _Seismograms(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rho, lam, mu, model):
            arrs = [t.detach().to("cpu", torch.float64).numpy() for t in (rho, lam, mu)]
            seis, state = model.run_forward(*arrs)
            ctx.model = model
            ctx.state = state  # this call's own forward state: several graphs may be alive
            ctx.like = rho
            return torch.as_tensor(seis, dtype=torch.float64, device=rho.device)

        @staticmethod
        def backward(ctx, gseis):
            g = gseis.detach().to("cpu", torch.float64).numpy()
            grads = ctx.model.backward(g, ctx.state)
            like = ctx.like
            return tuple(torch.as_tensor(x, dtype=like.dtype, device=like.device) for x in grads) + (None,)

    return _Seismograms


_FN = None


def release():
    """Return the device block the history plan keeps for the next gradient (the library
    keeps one block per process, so that a sequence of gradients does not pay for
    releasing and mapping it each time)."""
    api = solver.load_hip_lib()
    if api.adjoint_release() != 0:
        raise RuntimeError("sem2d_adjoint_release failed")


def seismograms(rho, lam, mu, model):
    """torch: seismograms of the per-cell model tensors (rho, lam, mu), differentiable."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(rho, lam, mu, model)
