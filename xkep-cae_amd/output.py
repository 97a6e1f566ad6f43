"""Result formats on the output side of the path.

``ExportProcess`` writes a solve's result as the reference's
(xkep_cae/output/export.py:16-89): ``<prefix>_displacement.csv`` (u as one
comma-separated row), ``<prefix>_contact_force.csv`` (one contact-force norm
per line, header ``contact_force_norm``) when there is a history, and
``<prefix>_summary.json`` (converged, n_increments, total_attempts,
max_displacement, elapsed_seconds).  ``SolverResultData`` carries the
reference's fields (core/data.py:197-213); ``solver_result_data`` fills it from
``ContactFrictionSolver.solve``'s result.

``write_run_summary`` writes the per-run YAML record of the reference's
dynamic bending driver (work/three_point_bend/tools/run_dynamic_bend.py:79-98)
for the keys a strand solve has: tag, converged, n_increments,
elapsed_seconds and final_frac, plus any workflow values passed in.
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import yaml


@dataclass
class SolverResultData:
    u: np.ndarray
    converged: bool
    n_increments: int
    total_attempts: int
    displacement_history: list = field(default_factory=list)
    contact_force_history: list = field(default_factory=list)
    load_history: list = field(default_factory=list)
    elapsed_seconds: float = 0.0
    diagnostics: object | None = None
    energy_history: object | None = None
    n_cutbacks: int = 0
    increment_diagnostics: list = field(default_factory=list)


def solver_result_data(res, *, elapsed_seconds: float = 0.0, contact_force_history=None) -> SolverResultData:
    """SolverResultData of a ``ContactFrictionSolver`` result: the load
    fraction of every converged call, cutbacks and attempts as counted there."""
    loads = [c.load_frac for c in res.calls if c.converged]
    return SolverResultData(u=np.asarray(res.u), converged=bool(res.converged), n_increments=int(res.n_increments),
                            total_attempts=int(res.total_attempts), load_history=loads,
                            contact_force_history=list(contact_force_history or []),
                            elapsed_seconds=float(elapsed_seconds), n_cutbacks=int(res.n_cutbacks))


@dataclass(frozen=True)
class ExportConfig:
    solver_result: SolverResultData
    mesh: object
    output_dir: str = "output"
    formats: tuple = ("csv",)
    prefix: str = "result"


@dataclass(frozen=True)
class ExportResult:
    exported_files: list = field(default_factory=list)
    n_steps_exported: int = 0


class ExportProcess:
    """``ExportProcess.process`` (output/export.py:48-89)."""

    name = "Export"
    module = "post"

    def process(self, input_data: ExportConfig) -> ExportResult:
        out = Path(input_data.output_dir)
        out.mkdir(parents=True, exist_ok=True)
        files: list[str] = []
        r = input_data.solver_result
        if "csv" in input_data.formats:
            p = out / f"{input_data.prefix}_displacement.csv"
            np.savetxt(p, np.asarray(r.u).reshape(1, -1), delimiter=",")
            files.append(str(p))
            if r.contact_force_history:
                p = out / f"{input_data.prefix}_contact_force.csv"
                np.savetxt(p, np.array(r.contact_force_history).reshape(-1, 1), delimiter=",",
                           header="contact_force_norm")
                files.append(str(p))
        if "json" in input_data.formats:
            summary = {
                "converged": r.converged,
                "n_increments": r.n_increments,
                "total_attempts": r.total_attempts,
                "max_displacement": float(np.max(np.abs(r.u))),
                "elapsed_seconds": r.elapsed_seconds,
            }
            p = out / f"{input_data.prefix}_summary.json"
            with open(p, "w") as f:
                json.dump(summary, f, indent=2)
            files.append(str(p))
        return ExportResult(exported_files=files, n_steps_exported=len(r.displacement_history))


def write_run_summary(path, tag: str, result: SolverResultData, **values) -> dict:
    """The run record as YAML (block style, unicode kept); float values are
    rounded as the reference's driver rounds them (seconds 1, fractions 4)."""
    data = {
        "tag": tag,
        "converged": bool(result.converged),
        "n_increments": int(result.n_increments),
        "elapsed_seconds": round(float(result.elapsed_seconds), 1),
        "final_frac": round(float(result.load_history[-1]) if result.load_history else 0.0, 4),
    }
    data.update(values)
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w") as f:
        yaml.dump(data, f, default_flow_style=False, allow_unicode=True)
    return data
