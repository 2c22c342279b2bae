"""Run metadata written next to the results (wakis/logger.py:14-97).

Three dictionaries, ``grid``, ``solver`` and ``wakeSolver``, are filled by
``GridFIT3D``, ``SolverFIT3D`` and ``WakeSolver`` as they are built and
written by ``save_logs`` into ``<results_folder>/wakis.log`` in the
reference's layout: a two-line title, then one ``## <name> Logs ##`` section
per dictionary (wake solver, solver, grid) holding the dictionary as
indented JSON, or ``(empty)``.  Values JSON cannot hold are written as
lists (numpy arrays) or as their ``str``.
"""

from __future__ import annotations

import json
import os

import numpy as np


def _jsonable(value):
    if isinstance(value, dict):
        return {k: _jsonable(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [_jsonable(v) for v in value]
    if isinstance(value, np.ndarray):
        return value.tolist()
    try:
        json.dumps(value)
    except (TypeError, ValueError, OverflowError):
        return str(value)
    return value


class Logger:
    def __init__(self):
        self.grid = {}
        self.solver = {}
        self.wakeSolver = {}

    def save_logs(self, results_folder=None):
        """Write ``wakis.log`` into ``results_folder`` (or the wake solver's
        ``results_folder`` entry), creating the folder if needed."""
        if results_folder is not None:
            self.wakeSolver["results_folder"] = results_folder
        folder = self.wakeSolver["results_folder"]
        if not os.path.exists(folder):
            os.mkdir(folder)
        parts = ["Simulation Parameters\n", "=====================\n\n"]
        for title, data in (("WakeSolver Logs", self.wakeSolver), ("Solver Logs", self.solver),
                            ("Grid Logs", self.grid)):
            parts.append(f"\n## {title} ##\n")
            if not data:
                parts.append("(empty)\n")
                continue
            parts.append(json.dumps(_jsonable(data), indent=2, ensure_ascii=False) + "\n")
        with open(os.path.join(folder, "wakis.log"), "w", encoding="utf-8") as fh:
            fh.write("".join(parts))
