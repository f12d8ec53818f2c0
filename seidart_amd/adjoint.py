"""Gradients of a receiver-trace misfit for the 2-D elastic and EM schemes.

The reference has no adjoint or gradient code; this module adds full
waveform inversion (FWI) gradients on top of the same forward schemes
(legacy/fortran/seismicfdtd.f95:12-385, seismic2; electromagfdtd.f95:10-326,
electromag2).  Pieces:

* :class:`AdjointEngine` runs the forward of every shot of one
  :class:`~seidart_amd.solver.Simulation`, keeping a full state (fields and
  C-PML memory variables) every ``every`` steps.  Its backward pass walks
  the segments last to first: it copies the segment's checkpoint into slot 0
  of a trail, recomputes the segment's states into the trail
  (``sdt_plan_run_trail``: the forward kernel writing each new state into
  the next slot), then runs the hand-written adjoint kernels over the
  segment (``sdt_adjoint_run``).  Memory: one checkpoint per segment plus
  ``every + 1`` trail states per shot.
* The adjoint yields dJ/d(kernel coefficient) per cell in FP64 (the slots of
  ``include/seidart_hip.h``: elastic n11 n13 n15 n33 n35 s15 s35 s55
  dt/rho_x dt/rho_z 1-gx 1-gz; EM caEx cbEx caEz cbEz) and dJ/d(source
  increment) per step.
* :class:`Traces` is the ``torch.autograd.Function`` around it: inputs are
  those coefficient slots and increments, output the traces.
  :func:`elastic2_traces` and :func:`em2_traces` compose it with the
  coefficient setup (:func:`seidart_amd.problem.elastic2_coefs`,
  :func:`seidart_amd.problem.em2_coefs`: averaging stencils, dt / rho, the
  ca / cb of permittivity and conductivity, written with torch operations),
  so torch's autograd carries the chain rule to the per-cell Cij, density,
  attenuation, permittivity and conductivity.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _hip
from .problem import elastic2_coefs, em2_coefs
from .solver import Simulation

__all__ = ["AdjointEngine", "Elastic2Engine", "Traces", "Elastic2Traces", "elastic2_traces",
           "em2_traces", "PARAMS", "PARAM_SLOTS", "NSLOTS"]

NSLOTS = {"elastic2": 12, "em2": 4}
# coefficient slots each per-cell parameter feeds (include/seidart_hip.h order)
PARAM_SLOTS = {
    "elastic2": {"c11": (0,), "c13": (1,), "c15": (2, 5), "c33": (3,), "c35": (4, 6), "c55": (7,),
                 "rho": (8, 9), "gamma_x": (10,), "gamma_z": (11,)},
    "em2": {"e11": (0, 1), "e13": (0, 1, 2, 3), "e33": (2, 3),
            "s11": (0, 1), "s13": (0, 1, 2, 3), "s33": (2, 3)},
}
# per-cell parameters of elastic2_traces / em2_traces
PARAMS = {k: tuple(v) for k, v in PARAM_SLOTS.items()}


class AdjointEngine:
    """Checkpointed forward and adjoint of every shot of 2-D elastic or EM
    problems (one kind per engine).

    ``problems``: one Problem or a list (one plan; shots of all models run in
    the same launches).  ``every``: steps per checkpoint segment.
    """

    def __init__(self, problems, precision: str = "fp32", every: int = 64,
                 device: str = "cuda:0", graph: bool = True, grad_mask: int | None = None,
                 reference: bool = False, overlap: bool = True, all_slots: bool = False):
        """``grad_mask``: bit q = accumulate coefficient slot q.  ``reference``:
        run the unfused three-kernel adjoint (a cross-check of the fused
        kernel; same results to rounding).  ``overlap``: two trail buffers,
        the recomputation of the next segment running on a second stream
        while the adjoint walks the current one.  ``all_slots``: keep the
        optional coefficient slots (c15 / c35 couplings, attenuation) on the
        device even where the problems leave them trivial, so that later
        coefficients may make them non-trivial; otherwise the kernels skip
        them (their gradients are still available)."""
        self.sim = Simulation(problems, precision, device, all_slots=all_slots, all_sources=True)
        # the trail recomputation runs the one-step kernel over the plan's
        # work list beside the adjoint kernel; there the column-strip order
        # (sdt_plan_set_tile_order 1) costs more than it gains in the
        # forward pass, so autotune keeps the default order
        self.sim.TILE_ORDERS = (0,)
        if self.sim.kind not in NSLOTS:
            raise ValueError("the adjoint engine handles elastic2 and em2 problems")
        self.kind = self.sim.kind
        self.nslot = NSLOTS[self.kind]
        self.all_slots_mask = (1 << self.nslot) - 1
        if grad_mask is None:
            grad_mask = self.all_slots_mask
        if every < 1:
            raise ValueError("every must be >= 1")
        sim = self.sim
        self.every = int(every)
        self.graph = graph
        self.nt = sim.nt
        self.nseg = (self.nt + self.every - 1) // self.every
        self.device = sim.device
        self.dtype = sim.dtype
        dev, dt = self.device, self.dtype
        self.grad = []
        for p in sim.problems:
            self.grad.append(torch.zeros((self.nslot, p.nz, p.nx), dtype=torch.float64, device=dev))
        self.ckpt, self.trail, self.lam, self.work, self.gsrc, self.resid, self.init = ([] for _ in range(7))
        adj = (_hip.SdtAdjShot * len(sim.shots))()
        for s, (m, _, b) in enumerate(sim.shots):
            p = sim.problems[m]
            lx, ly, lz = p.slabs()
            nst = b["state"].numel()
            nwk = _hip.adjoint_work_elems(sim.kid, p.nx, p.ny, p.nz, lx, ly, lz)
            self.ckpt.append(torch.zeros((self.nseg, nst), dtype=dt, device=dev))
            self.trail.append(torch.zeros((2 if overlap else 1, self.every + 1, nst), dtype=dt,
                                          device=dev))
            self.lam.append(torch.zeros((2, nst), dtype=dt, device=dev))
            self.work.append(torch.zeros(nwk, dtype=dt, device=dev))
            self.gsrc.append(torch.zeros((self.nt, 3), dtype=torch.float64, device=dev))
            self.resid.append(torch.zeros_like(b["trace"]))
            self.init.append(b["state"].clone())
            A = adj[s]
            A.trail[0] = self.trail[s][0].data_ptr()
            A.trail[1] = self.trail[s][1].data_ptr() if overlap else None
            A.trail_len = self.every + 1
            A.lam[0] = self.lam[s][0].data_ptr()
            A.lam[1] = self.lam[s][1].data_ptr()
            A.work = self.work[s].data_ptr()
            A.resid = self.resid[s].data_ptr()
            A.grad = self.grad[m].data_ptr()
            A.gsrc = self.gsrc[s].data_ptr()
        self._adj = adj
        self.overlap = bool(overlap)
        self._side = torch.cuda.Stream(device=dev) if overlap else None  # default priority
        self.forward_id = 0  # checkpoints belong to the latest forward
        self.reference = bool(reference)
        self.grad_mask = None
        self.set_grad_mask(grad_mask)

    def set_grad_mask(self, mask: int) -> None:
        """Choose the coefficient slots the adjoint accumulates (bit q = slot q)."""
        mask = int(mask) & self.all_slots_mask
        if mask == self.grad_mask:
            return
        torch.cuda.synchronize(self.device)
        flags = _hip.SDT_ADJ_REFERENCE if self.reference else 0
        _hip.check(_hip.lib().sdt_adjoint_setup(self.sim._plan, self._adj, len(self.sim.shots),
                                                mask, flags))
        self.grad_mask = mask

    # ------------------------------------------------------------------
    @property
    def nshots(self) -> int:
        return len(self.sim.shots)

    def set_coefs(self, model: int, slots) -> None:
        """Overwrite model ``model``'s coefficient slots (FP64 tensors or
        arrays of shape (nz, nx)).  Slots the device skips must stay trivial
        (elastic: 0 for the couplings, 1 for 1 - gamma)."""
        for q, v in enumerate(slots):
            dst = self.sim.coef_t[model][q]
            v = torch.as_tensor(v)
            if dst is None:
                if not bool(torch.all(v == (1.0 if q >= 10 else 0.0))):
                    raise ValueError(f"coefficient slot {q} is not on the device (trivial in the "
                                     "problem); build the engine with all_slots=True")
                continue
            dst.copy_(v.to(self.device, self.dtype))

    def set_increments(self, shot: int, inc) -> None:
        """Overwrite shot ``shot``'s per-step source increments: ``inc`` of
        shape (3, nt) (x, y, z; y unused in 2-D)."""
        src = self.sim.shots[shot][2]["src"]
        for c in (0, 2):
            src[c][: self.nt].copy_(torch.as_tensor(inc[c]).to(self.device, self.dtype)[: self.nt])

    def forward(self):
        """Run all steps from the initial state, checkpointing every
        ``every`` steps.  Returns the per-shot device trace buffers
        ([nt, 2, nrec_max], plan precision)."""
        sim = self.sim
        self.forward_id += 1
        for s, (_, _, b) in enumerate(sim.shots):
            b["state"].copy_(self.init[s])
            b["trace"].zero_()
        sim.step = 0
        for g in range(self.nseg):
            for s, (_, _, b) in enumerate(sim.shots):
                self.ckpt[s][g].copy_(b["state"])
            n = min(self.every, self.nt - g * self.every)
            sim.run(n, graph_block=self.every if (self.graph and n == self.every) else 0)
        sim.check()
        return [b["trace"] for _, _, b in sim.shots]

    def backward(self, resid):
        """Adjoint pass for the residuals ``resid`` (per shot, dJ/dtrace of
        shape [nt, 2, nrec]).  Must follow :meth:`forward` with the same
        coefficients.  Returns (grad, gsrc): per model the FP64
        [nslot, nz, nx] coefficient-slot gradient, per shot the FP64 [nt, 3]
        source-increment gradient."""
        sim = self.sim
        L = _hip.lib()
        stream = torch.cuda.current_stream(self.device)
        h = C.c_void_p(stream.cuda_stream)
        for g in self.grad:
            g.zero_()
        for s in range(self.nshots):
            self.lam[s].zero_()
            self.gsrc[s].zero_()
            r = torch.as_tensor(resid[s]).to(self.device, self.dtype)
            self.resid[s].zero_()
            self.resid[s][:, :, : r.shape[2]].copy_(r)
        seg = lambda g: (g * self.every, min(self.nt, (g + 1) * self.every))
        if not self.overlap:
            for g in range(self.nseg - 1, -1, -1):
                t_lo, t_hi = seg(g)
                for s in range(self.nshots):
                    self.trail[s][0][0].copy_(self.ckpt[s][g])
                _hip.check(L.sdt_plan_run_trail(sim._plan, t_lo, t_hi - t_lo, 0, h))
                _hip.check(L.sdt_adjoint_run(sim._plan, t_hi, t_lo, t_lo, 0, h))
            return self.grad, self.gsrc
        # recompute segment g-1 into trail buffer (g-1)%2 on the side stream
        # while the adjoint of segment g reads buffer g%2
        side = self._side
        hs = C.c_void_p(side.cuda_stream)
        side.wait_stream(stream)
        rec_done = {}
        adj_done = {}

        def recompute(g):
            b = g % 2
            if g + 2 in adj_done:  # the adjoint that read this buffer last
                side.wait_event(adj_done[g + 2])
            with torch.cuda.stream(side):
                for s in range(self.nshots):
                    self.trail[s][b][0].copy_(self.ckpt[s][g])
            t_lo, t_hi = seg(g)
            _hip.check(L.sdt_plan_run_trail(sim._plan, t_lo, t_hi - t_lo, b, hs))
            ev = torch.cuda.Event()
            ev.record(side)
            rec_done[g] = ev

        recompute(self.nseg - 1)
        for g in range(self.nseg - 1, -1, -1):
            stream.wait_event(rec_done[g])
            if g > 0:
                recompute(g - 1)
            t_lo, t_hi = seg(g)
            _hip.check(L.sdt_adjoint_run(sim._plan, t_hi, t_lo, t_lo, g % 2, h))
            ev = torch.cuda.Event()
            ev.record(stream)
            adj_done[g] = ev
        return self.grad, self.gsrc

    def adjoint_state(self, shot: int) -> torch.Tensor:
        """lambda_0 = dJ/d(initial state) of a shot after :meth:`backward`
        (fields then psi, the layout of the state)."""
        return self.lam[shot][0]

    def close(self):
        self.sim.close()


Elastic2Engine = AdjointEngine


class Traces(torch.autograd.Function):
    """traces = F(coefficient slots, source increments).

    Inputs after the engine: per model one FP64 tensor [nslot, nz, nx] of
    coefficient slots, then per shot one FP64 tensor [3, nt] of source
    increments.  Output: FP64 traces [nshots, nt, 2, nrec_max].  The
    backward pass is the engine's checkpointed adjoint.
    """

    @staticmethod
    def forward(ctx, engine: Elastic2Engine, *inputs):
        nm = len(engine.sim.problems)
        for m in range(nm):
            engine.set_coefs(m, inputs[m].detach())
        for s in range(engine.nshots):
            engine.set_increments(s, inputs[nm + s].detach())
        tr = engine.forward()
        nrec = max(t.shape[2] for t in tr)
        out = torch.zeros((len(tr), engine.nt, 2, nrec), dtype=torch.float64, device=engine.device)
        for s, t in enumerate(tr):
            out[s, :, :, : t.shape[2]] = t.double()
        ctx.engine = engine
        ctx.forward_id = engine.forward_id
        ctx.devs = [x.device for x in inputs]
        return out.to(inputs[0].device)

    @staticmethod
    def backward(ctx, gout):
        eng = ctx.engine
        if eng.forward_id != ctx.forward_id:
            raise RuntimeError("the engine ran another forward since this graph was built; "
                               "backpropagate before the next forward")
        gout = gout.to(eng.device)
        grad, gsrc = eng.backward([gout[s, :, :, : eng.resid[s].shape[2]] for s in range(eng.nshots)])
        outs = [g.clone().to(d) for g, d in zip(grad, ctx.devs)]
        nm = len(grad)
        outs += [gsrc[s].t().contiguous().to(ctx.devs[nm + s]) for s in range(eng.nshots)]
        return (None, *outs)


def _mask_from(engine, params, kind):
    """Accumulate only the coefficient slots the parameters that need a
    gradient feed."""
    if engine.kind != kind:
        raise ValueError(f"engine runs {engine.kind} problems, not {kind}")
    mask = 0
    for prm in params:
        for k, v in prm.items():
            if torch.is_tensor(v) and v.requires_grad:
                for q in PARAM_SLOTS[kind][k]:
                    mask |= 1 << q
    if mask:
        engine.set_grad_mask(mask)


def _increments(engine, divisor_of):
    """Per shot [3, nt] source increments s(it) dt / divisor(src) (torch)."""
    sim = engine.sim
    incs = []
    for m, s, _ in sim.shots:
        p = sim.problems[m]
        sh = p.shots[s]
        i, _, k = sh.src
        comps = []
        for c in range(3):
            f = sh.force[c]
            div = divisor_of(m, c)
            if f is None or c == 1:
                comps.append(torch.zeros(engine.nt, dtype=torch.float64, device=div.device))
                continue
            f = torch.as_tensor(np.pad(np.asarray(f, dtype=np.float64)[: p.nt], (0, engine.nt - p.nt)),
                                device=div.device)
            comps.append(f * p.dt / div[k, i])
        incs.append(torch.stack(comps))
    return incs


def em2_traces(engine: AdjointEngine, params):
    """EM 2-D receiver traces (Ex, Ez) of every shot as a differentiable
    function of the per-cell relative permittivity and conductivity.

    ``params``: per model a dict of FP64 torch tensors (nz, nx) with keys
    among e11 e13 e33 s11 s13 s33 (absent = the engine's problem values, held
    constant).  Returns FP64 [nshots, nt, 2, nrec_max]; backpropagating runs
    the adjoint.  The source increment divides by e11 / e33 at the source
    cell (electromagfdtd.f95:300-301), so gradients include that term."""
    sim = engine.sim
    if len(params) != len(sim.problems):
        raise ValueError("one parameter dict per model")
    _mask_from(engine, params, "em2")
    slots, Cs = [], []
    for p, prm in zip(sim.problems, params):
        dev = next(iter(prm.values())).device if prm else engine.device
        Cd = {k: prm[k] if k in prm else torch.as_tensor(p.coef[k], dtype=torch.float64, device=dev)
              for k in ("e11", "e13", "e33", "s11", "s13", "s33")}
        slots.append(torch.stack(em2_coefs(Cd, p.dt)))
        Cs.append(Cd)
    incs = _increments(engine, lambda m, c: Cs[m]["e11" if c == 0 else "e33"])
    return Traces.apply(engine, *slots, *incs)


def elastic2_traces(engine: AdjointEngine, params):
    """Receiver traces of every shot as a differentiable function of the
    per-cell model parameters.

    ``params``: per model of the engine a dict of FP64 torch tensors (nz, nx)
    with keys c11 c13 c15 c33 c35 c55 rho and optionally gamma_x gamma_z
    (absent = the engine's problem values, held constant).  Returns FP64
    [nshots, nt, 2, nrec_max]; backpropagating through it runs the adjoint.
    """
    sim = engine.sim
    if len(params) != len(sim.problems):
        raise ValueError("one parameter dict per model")
    _mask_from(engine, params, "elastic2")
    slots = []
    rho_by_model = []
    for m, (p, prm) in enumerate(zip(sim.problems, params)):
        dev = prm["rho"].device
        Cd = {k: prm[k] if k in prm else torch.as_tensor(p.coef[k], dtype=torch.float64, device=dev)
              for k in ("c11", "c13", "c15", "c33", "c35", "c55", "rho")}
        gam = {}
        for a in ("x", "z"):
            if "gamma_" + a in prm:
                gam[a] = prm["gamma_" + a]
            elif a in p.gamma:
                gam[a] = torch.as_tensor(p.gamma[a], dtype=torch.float64, device=dev)
        slots.append(torch.stack(elastic2_coefs(Cd, gam, p.dt, getattr(p, "density_method", "legacy"))))
        rho_by_model.append(Cd["rho"])
    # s(it) * dt / rho(src), seismicfdtd.f95:359-361
    incs = _increments(engine, lambda m, c: rho_by_model[m])
    return Traces.apply(engine, *slots, *incs)


Elastic2Traces = Traces
