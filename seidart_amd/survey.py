"""Survey drivers on the batched GPU path.

:class:`CommonOffset` is the common-offset survey of
src/seidart/simulations/common_offset.py:14-283: one source position and one
receiver per trace, the receiver a fixed offset from its source.  The
reference runs one solver process per source (process_source, :143-236) and
reads each receiver back from snapshot files; here every source position is
a shot of a single plan (one launch per step for all of them), the receiver
samples are gathered on the device, and no snapshot files are written.
"""
from __future__ import annotations

import os

import numpy as np
import pandas as pd

from .problem import TRACE_COMPONENTS, Problem, Shot
from .run import problem_from_files, source_index, write_trace_csv

__all__ = ["CommonOffset", "read_xyz"]


def read_xyz(path: str) -> np.ndarray:
    """An (n, 3) x, y, z table with one header row (pandas.read_csv in
    source_receiver_xyz, common_offset.py:103-119), either orientation."""
    a = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)
    if a.shape[1] != 3 and a.shape[0] == 3:
        a = a.T
    return a


class CommonOffset:
    """``CommonOffset(source_file, channel, project_file, receiver_file,
    receiver_indices=False, single_precision=True)`` with ``co_run()``
    filling ``co_image`` (time_steps x n sources) as the reference does.

    The project must have been built (the coefficient and C-PML files
    ``Model.build`` writes sit next to the project file).  ``precision``
    selects the device arithmetic ("fp32" or "fp64")."""

    def __init__(self, source_file: str, channel: str, project_file: str, receiver_file: str,
                 receiver_indices: bool = False, single_precision: bool = True,
                 precision: str = "fp32", density_method: str = "legacy"):
        self.source_file = source_file
        self.channel = channel
        self.project_file = project_file
        self.receiver_file = receiver_file
        self.receiver_indices = receiver_indices
        self.single_precision = single_precision
        self.precision = precision
        self.output_basefile = None
        self.simulation_type = "CO"
        if channel in ("Vx", "Vy", "Vz"):
            self.is_seismic = True
        elif channel in ("Ex", "Ey", "Ez"):
            self.is_seismic = False
        else:
            raise ValueError("channel must be one of Vx Vy Vz Ex Ey Ez")
        self.base = problem_from_files(project_file, seismic=self.is_seismic,
                                       density_method=density_method)
        comps = TRACE_COMPONENTS[self.base.kind]
        if channel not in comps:
            raise ValueError(f"channel {channel} is not a component of a {self.base.kind} run")
        self.component = comps.index(channel)
        self.time_steps = self.base.nt
        self.dt = self.base.dt
        self.source_xyz_all = read_xyz(source_file)
        self.receiver_xyz_all = read_xyz(receiver_file)
        if len(self.source_xyz_all) != len(self.receiver_xyz_all):
            raise ValueError("the source and receiver files must list the same number of points")
        self.co_image = None
        self.timeseries = None

    def receiver_cells(self) -> np.ndarray:
        """Padded receiver indices: xyz / d (or indices) + cpml, truncated
        (source_receiver_xyz, common_offset.py:121-135; process_source :183)."""
        p = self.base
        r = np.asarray(self.receiver_xyz_all, dtype=np.float64)
        if not self.receiver_indices:
            r = r / np.array([p.dx, p.dy, p.dz])
        r = (r + p.cpml).astype(int)
        if not p.is3d:
            r[:, 1] = 0
        return r.astype(np.int32)

    def source_cells(self) -> list:
        """Padded source cells: int(xyz / d) placed the way the solver places
        the project's own source (run.source_index)."""
        p = self.base
        out = []
        for x, y, z in np.asarray(self.source_xyz_all, dtype=np.float64):
            out.append(source_index(int(x / p.dx), int(y / p.dy) if p.is3d else 0, int(z / p.dz),
                                    p.cpml, p.dim))
        return out

    def problem(self) -> Problem:
        """All source positions as shots of one problem on the project's model."""
        p = self.base
        force, moment = p.shots[0].force, p.shots[0].moment
        rec = self.receiver_cells()
        shots = [Shot(src=s, force=force, rec=rec[q:q + 1], moment=moment)
                 for q, s in enumerate(self.source_cells())]
        q = Problem(p.kind, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, p.dt, p.nt, p.coef, p.prof, shots,
                    p.gamma)
        for a in ("density_method", "cpml", "dim", "cpml_split"):
            if hasattr(p, a):
                setattr(q, a, getattr(p, a))
        q.validate()
        return q

    def co_run(self, parallel: bool = False, graph_block: int = 50) -> np.ndarray:
        """Run every source position (all in one batched plan; ``parallel`` is
        accepted for signature compatibility) and return co_image."""
        del parallel
        from .solver import Simulation
        sim = Simulation(self.problem(), precision=self.precision)
        sim.run(graph_block=graph_block).check()
        tr = sim.traces()
        sim.close()
        self.co_image = np.stack([t[:, self.component, 0] for t in tr], axis=1).astype(np.float64)
        self.timeseries = self.co_image
        self.source_xyz = np.asarray(self.source_xyz_all[-1], dtype=np.float64)
        self.save(output_basefile=self.output_basefile)
        return self.co_image

    def save(self, output_basefile: str | None = None) -> str:
        """co_image as a headerless CSV named like Array.save
        (arraybuild.py:1031-1056: project base, channel, last source x-y-z)."""
        if output_basefile:
            fn = output_basefile + ".csv"
            pd.DataFrame(self.timeseries).to_csv(fn, header=False, index=False)  # Array.save's writer
        else:
            base = os.path.splitext(self.project_file)[0]
            fn = write_trace_csv(self.timeseries, base, self.channel, self.source_xyz)
        self.csvfilename = fn
        return fn
