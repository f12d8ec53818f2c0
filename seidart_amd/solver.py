"""Device-side run of the FDTD core: allocate buffers, build a plan, step.

PyTorch is used only for device memory and the current stream; every
update is a kernel of libseidart_hip.so reached through its C ABI.  One
:class:`Simulation` holds any number of models of one physics kind (grids
may differ in size) and any number of shots per model; each time step
advances all of them in the same launches (the batched path that replaces
one ``seidartfdtd`` process per shot, src/seidart/simulations/
common_offset.py:143-283).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _hip
from .problem import FIELDS, KINDS, Problem, aux_b_x_half_by_row, em_dbh, field_extent, kernel_coefs, \
    ncomp, source_increments, stress_source_increments

__all__ = ["Simulation", "PRECISIONS"]

PRECISIONS = {"fp64": (_hip.SDT_FP64, torch.float64), "fp32": (_hip.SDT_FP32, torch.float32)}
# kinds whose single-pass step kernel reads state A or B and writes the other
PINGPONG = ("elastic2", "em2")


def _dev_ptr(t: torch.Tensor | None):
    return None if t is None else C.c_void_p(t.data_ptr())


class Simulation:
    """All shots of one or more Problems (same kind) on one GPU."""

    def __init__(self, problems, precision: str = "fp64", device: str = "cuda:0",
                 all_slots: bool = False, all_sources: bool = False):
        """``all_slots``: pass every coefficient slot to the device even where
        the optional ones are trivial (zero c15 / c35 couplings, no
        attenuation), so they can be overwritten later.  ``all_sources``:
        allocate the x and z source series even where a component is
        absent (the adjoint engine writes them)."""
        if isinstance(problems, Problem):
            problems = [problems]
        if not problems:
            raise ValueError("no problems")
        kinds = {p.kind for p in problems}
        if len(kinds) != 1:
            raise ValueError("all problems of a Simulation must share one physics kind")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {list(PRECISIONS)}")
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: the HIP path needs a ROCm device")
        _hip.lib()  # fail early if the library is not built
        self.kind = problems[0].kind
        splits = {bool(getattr(p, "cpml_split", False)) for p in problems}
        if len(splits) != 1:
            raise ValueError("all problems of a Simulation must agree on cpml_split")
        self.split = splits.pop()
        if self.split and self.kind != "elastic2":
            raise ValueError("cpml_split applies to elastic2 only")
        # kind as the C ABI takes it (with the split C-PML flag)
        self.kid = KINDS[self.kind] | (_hip.SDT_SPLIT_CPML if self.split else 0)
        self.prec, self.dtype = PRECISIONS[precision]
        self.precision = precision
        self.device = torch.device(device)
        self.problems = problems
        self.nt = max(p.nt for p in problems)
        self.ncomp = ncomp(self.kind)
        self.names = FIELDS[self.kind]
        self._keep = []  # every tensor a descriptor points to
        self.coef_t = []  # per model: device coefficient tensor per slot (None = skipped)
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=self.dtype, device=self.device)

        models = (_hip.SdtModel * len(problems))()
        shot_list = []
        self.shots = []  # (problem index, shot index, dict of tensors)
        for m, p in enumerate(problems):
            p.validate()
            M = models[m]
            M.nx, M.ny, M.nz = p.nx, p.ny, p.nz
            M.lx, M.ly, M.lz = p.slabs()
            M.dx, M.dy, M.dz, M.dt = p.dx, p.dy, p.dz, p.dt
            if self.kind.startswith("em"):
                M.scal[0] = em_dbh(p.dt)
            host = kernel_coefs(p)
            skip = set()
            if self.kind == "elastic2" and not all_slots:
                # optional slots: c15/c35 couplings all zero, or no attenuation
                if all(not np.any(host[q]) for q in (2, 4, 5, 6)):
                    skip |= {2, 4, 5, 6}
                if all(np.all(host[q] == 1.0) for q in (10, 11)):
                    skip |= {10, 11}
            self.coef_t.append([])
            for q, a in enumerate(host):
                if q in skip:
                    M.coef[q] = None
                    self.coef_t[-1].append(None)
                    continue
                t = T(a)
                self._keep.append(t)
                self.coef_t[-1].append(t)
                M.coef[q] = t.data_ptr()
            for ax, a in enumerate(("x", "y", "z")):
                pr = p.prof.get(a)
                if pr is None:
                    continue
                vals = [1.0 / pr.kappa, pr.a, pr.b, 1.0 / pr.kappa_h, pr.a_h, pr.b_h]
                ts = [T(v) for v in vals]
                self._keep += ts
                for q, t in enumerate(ts):
                    M.prof[ax][q] = t.data_ptr()
            if self.kind == "elastic2":
                aux = T(aux_b_x_half_by_row(p))
                self._keep.append(aux)
                M.aux = aux.data_ptr()
            lx, ly, lz = p.slabs()
            npsi = _hip.psi_elems(self.kid, p.nx, p.ny, p.nz, lx, ly, lz)
            nel = int(np.prod(p.field_shape))
            for s, sh in enumerate(p.shots):
                init = (p.init[s] if p.init else None) or {}
                bufs = {}
                # one contiguous state (fields then psi), the layout of
                # sdt_state_elems; the single-pass ELASTIC2 and EM2 kernels
                # ping-pong between state A and state B
                nst = _hip.state_elems(self.kid, p.nx, p.ny, p.nz, lx, ly, lz)
                for tag in (("", "_b") if self.kind in PINGPONG else ("",)):
                    st = torch.zeros(nst, dtype=self.dtype, device=self.device)
                    bufs["state" + tag] = st
                    for f, n in enumerate(self.names):
                        bufs[n + tag] = st[f * nel:(f + 1) * nel].view(p.field_shape)
                    bufs["psi" + tag] = st[len(self.names) * nel:] if npsi else st[:1]
                for n in self.names:
                    if n in init:
                        # initial values only where the field lives in the
                        # legacy arrays; the staggered EM fields' last column
                        # or row stays zero
                        a = np.asarray(init[n], dtype=np.float64).reshape(p.field_shape)
                        inside = np.zeros_like(a)
                        ext = field_extent(self.kind, n)
                        inside[ext] = a[ext]
                        bufs[n].copy_(T(inside))
                inc = source_increments(p, sh)
                if all_sources and self.kind in ("elastic2", "em2"):
                    inc = [np.zeros(p.nt) if (v is None and c != 1) else v for c, v in enumerate(inc)]
                bufs["src"] = [None if v is None else T(np.pad(v, (0, self.nt - v.size))) for v in inc]
                minc = stress_source_increments(p, sh)
                bufs["msrc"] = [None if v is None else T(np.pad(v, (0, self.nt - v.size))) for v in minc]
                rec = np.asarray(sh.rec, dtype=np.int32).reshape(-1, 3)
                bufs["rec"] = torch.as_tensor(rec, device=self.device) if rec.size else None
                bufs["trace"] = torch.zeros((self.nt, self.ncomp, max(rec.shape[0], 1)),
                                            dtype=self.dtype, device=self.device)
                bufs["nrec"] = rec.shape[0]
                S = _hip.SdtShot()
                S.model = m
                S.isrc, S.jsrc, S.ksrc = (int(v) for v in sh.src)
                for f, n in enumerate(self.names):
                    S.field[f] = bufs[n].data_ptr()
                S.psi = bufs["psi"].data_ptr()
                if self.kind in PINGPONG:
                    for f, n in enumerate(self.names):
                        S.field_b[f] = bufs[n + "_b"].data_ptr()
                    S.psi_b = bufs["psi_b"].data_ptr()
                for c in range(3):
                    S.src[c] = bufs["src"][c].data_ptr() if bufs["src"][c] is not None else None
                for c in range(6):
                    S.msrc[c] = bufs["msrc"][c].data_ptr() if bufs["msrc"][c] is not None else None
                S.nrec = bufs["nrec"]
                S.rec = bufs["rec"].data_ptr() if bufs["rec"] is not None else None
                S.trace = bufs["trace"].data_ptr()
                shot_list.append(S)
                self.shots.append((m, s, bufs))
        shots = (_hip.SdtShot * len(shot_list))(*shot_list)
        torch.cuda.synchronize(self.device)
        plan = C.c_void_p()
        _hip.check(_hip.lib().sdt_plan_create(self.kid, self.prec, models, len(problems), shots,
                                              len(shot_list), self.nt, C.byref(plan)))
        self._plan = plan
        self._models, self._shots = models, shots
        self.step = 0

    # ------------------------------------------------------------------
    def run(self, nsteps: int | None = None, graph_block: int = 0, stream=None):
        """Enqueue the next ``nsteps`` steps (default: to the end)."""
        if nsteps is None:
            nsteps = self.nt - self.step
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        _hip.check(_hip.lib().sdt_plan_run(self._plan, self.step, nsteps, graph_block,
                                           C.c_void_p(stream.cuda_stream)))
        self.step += nsteps
        return self

    def run_noalloc(self, nsteps: int | None = None, stream=None):
        """Advance the next ``nsteps`` steps through the allocation-free entry
        point (``sdt_run_noalloc``: the same kernels and results as ``run``,
        its descriptors and work lists in the library's fixed device
        scratch, no device allocation) on this Simulation's buffers, and
        report instability like ``check``.  Synchronises the stream."""
        if nsteps is None:
            nsteps = self.nt - self.step
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        first = C.c_int(-1)
        per = (C.c_int * len(self.shots))()
        rc = _hip.lib().sdt_run_noalloc(self.kid, self.prec, self._models, len(self.problems),
                                        self._shots, len(self.shots), self.nt, self.step, nsteps,
                                        C.c_void_p(stream.cuda_stream), C.byref(first), per)
        self.step += nsteps
        self.unstable = [int(v) for v in per]
        _hip.check(rc)
        return self

    def profile(self, nsteps: int, stream=None):
        """Run the next ``nsteps`` steps with every kernel bracketed by HIP
        events; return mean ms per launch of (first field kernel, second
        field kernel, receiver gather).  With two steps per launch
        (``steps_per_launch``) the first is the two-step kernel, per launch,
        and the gather is 0 (the kernel writes the receiver samples)."""
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        out = (C.c_double * 3)()
        _hip.check(_hip.lib().sdt_plan_profile(self._plan, self.step, nsteps,
                                               C.c_void_p(stream.cuda_stream), out))
        self.step += nsteps
        return tuple(float(v) for v in out)

    @property
    def steps_per_launch(self) -> int:
        """Time steps one launch of the ELASTIC2 / EM2 step kernel advances
        (2 by default for plans of at least 4M cells per step, else 1;
        ``autotune`` times both); results do not depend on it."""
        v = C.c_int()
        _hip.check(_hip.lib().sdt_plan_get_steps_per_launch(self._plan, C.byref(v)))
        return v.value

    @steps_per_launch.setter
    def steps_per_launch(self, steps: int):
        _hip.check(_hip.lib().sdt_plan_set_steps_per_launch(self._plan, int(steps)))

    # ------------------------------------------------------------------
    TILE_CANDIDATES = tuple((wb, rows) for wb in (4, 2, 1) for rows in (1, 2, 4, 8, 16, 32, 64))
    TILE_ORDERS = (0, 1)  # work-list orders (sdt_plan_set_tile_order)
    AUTOTUNE_RETIME = 3  # shapes autotune times a second time

    def tiles(self):
        """Current (waves per block, rows) of the single-pass step kernel
        (ELASTIC2, EM2)."""
        wb, rows = C.c_int(), C.c_int()
        _hip.check(_hip.lib().sdt_plan_get_tiles(self._plan, C.byref(wb), C.byref(rows)))
        return wb.value, rows.value

    def set_tiles(self, waves_per_block: int, rows: int):
        _hip.check(_hip.lib().sdt_plan_set_tiles(self._plan, int(waves_per_block), int(rows)))
        return self

    @property
    def tile_order(self) -> int:
        """Order of the step kernels' work lists: 0 deals the tiles to the
        XCDs in turn, 1 gives each XCD a run of tiles taken column by column
        (ELASTIC2, EM2); results do not depend on it."""
        v = C.c_int()
        _hip.check(_hip.lib().sdt_plan_get_tile_order(self._plan, C.byref(v)))
        return v.value

    @tile_order.setter
    def tile_order(self, order: int):
        _hip.check(_hip.lib().sdt_plan_set_tile_order(self._plan, int(order)))

    def use(self, key):
        """Apply an ``autotune`` key (steps per launch, waves per block,
        rows, tile order)."""
        spl, wb, rows, order = key
        self.steps_per_launch = spl
        self.tile_order = order
        return self.set_tiles(wb, rows)

    def autotune(self, steps: int = 8, candidates=None):
        """Pick the tile shape of the single-pass step kernel (ELASTIC2, EM2),
        the order of its work list and the steps per launch (2 or 1) that run
        this plan's steps fastest.

        Each candidate runs ``steps`` steps as one hipGraph block (captured
        once, then replayed and timed with HIP events on the stream, receiver
        gather included) from the current state; the state, traces and step
        counter are restored afterwards, so the run continues as if nothing
        had happened.  Returns {(steps_per_launch, wb, rows, order): ms per
        step} and keeps the best (``steps`` is rounded up to a multiple of 4, so
        both modes end a block with the state in the same buffer); the
        ``AUTOTUNE_RETIME`` fastest shapes are timed again over three replays
        and keep their fastest time.  Other kinds have a single shape:
        returns {}."""
        if self.kind not in PINGPONG:
            return {}
        left = self.nt - self.step
        steps = min(-(-max(int(steps), 4) // 4) * 4, left)
        if steps < 1:
            return {}
        spls = (2, 1)
        keep = [(b["state"].clone(), b["trace"].clone()) for _, _, b in self.shots]
        step0 = self.step
        stream = torch.cuda.current_stream(self.device)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

        def restore(traces=False):
            # state back in A, and in B too: the plan may take a one-step run
            # at an odd step from B's copy
            self.step = step0
            for (_, _, b), (a0, t0) in zip(self.shots, keep):
                b["state"].copy_(a0)
                b["state_b"].copy_(a0)
                if traces:
                    b["trace"].copy_(t0)

        def timed(spl, wb, rows, order, reps):
            # capture this shape, then the fastest of `reps` timed replays
            self.use((spl, wb, rows, order))
            self.run(steps, graph_block=steps)
            restore()
            t = float("inf")
            for _ in range(reps):
                ev0.record(stream)
                self.run(steps, graph_block=steps)
                ev1.record(stream)
                ev1.synchronize()
                t = min(t, ev0.elapsed_time(ev1) / steps)
                restore()
            restore(traces=True)
            return t

        res = {}
        for order in self.TILE_ORDERS:
            for spl in spls:
                for wb, rows in (candidates or self.TILE_CANDIDATES):
                    res[(spl, wb, rows, order)] = timed(spl, wb, rows, order, 1)
        # one replay can order shapes a few per cent apart wrongly: the
        # fastest few are timed again, three replays each
        for key in sorted(res, key=res.get)[:self.AUTOTUNE_RETIME]:
            res[key] = min(res[key], timed(*key, 3))
        self.use(min(res, key=res.get))
        torch.cuda.synchronize(self.device)
        return res

    def capture_graph(self, graph_block: int):
        """Capture (and replay once) the hipGraph of ``graph_block`` steps that
        later ``run(..., graph_block=graph_block)`` calls replay, then put the
        state, traces and step counter back as they were: a run after this
        equals a run without it, minus the capture cost.  Returns False
        (doing nothing) when fewer than ``graph_block`` steps remain."""
        if graph_block < 2 or self.nt - self.step < graph_block:
            return False
        keep = [{k: b[k].clone() for k in ("state", "trace")} for _, _, b in self.shots]
        step0 = self.step
        self.run(graph_block, graph_block=graph_block)
        torch.cuda.synchronize(self.device)
        self.step = step0
        for (_, _, b), k in zip(self.shots, keep):
            for n, t in k.items():
                b[n].copy_(t)
            if "state_b" in b:  # B holds the state too (see autotune)
                b["state_b"].copy_(b["state"])
        torch.cuda.synchronize(self.device)
        return True

    def check(self):
        """Synchronise; raise UnstableError if any shot blew up."""
        first = C.c_int(-1)
        per = (C.c_int * len(self.shots))()
        rc = _hip.lib().sdt_plan_check(self._plan, C.byref(first), per)
        self.unstable = [int(v) for v in per]
        _hip.check(rc)
        return self

    def traces(self):
        """Per shot: numpy [nt, ncomp, nrec] (float64 copy)."""
        torch.cuda.synchronize(self.device)
        return [b["trace"][:, :, : b["nrec"]].double().cpu().numpy() for _, _, b in self.shots]

    def field(self, shot: int, name: str) -> np.ndarray:
        torch.cuda.synchronize(self.device)
        return self.shots[shot][2][name].double().cpu().numpy()

    def close(self):
        if getattr(self, "_plan", None):
            _hip.lib().sdt_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
