"""Convenience: build a masx Solver from generated workloads (numpy dicts)."""
from __future__ import annotations

import numpy as np

from . import native
from . import workloads as W


def describe(systems, *, use_mas=True, global_diag=False, bank=32, levels=0):
    """Solver with the structure of every system added (no device work)."""
    if isinstance(systems, dict):
        systems = [systems]
    s = native.Solver()
    for w in systems:
        mas = bool(use_mas) and w["n_vertices"] > 0
        i = s.add_system(block_rows=w["block_rows"], rows=w["rows"], cols=w["cols"], n_abd=w["n_bodies"],
                         use_mas=mas, global_diag=global_diag, bank=bank, levels=levels)
        if mas:
            start, num, lst = W.neighbors(w["n_vertices"], w["tets"])
            s.set_topology_morton(i, start, num, lst, w["positions"])
            if len(w.get("pairs", [])):
                s.set_contacts(i, w["pairs"])
    return s


def solver_for(systems, *, use_mas=True, global_diag=False, bank=32, levels=0, device_values=True):
    """One Solver holding every system of `systems` (a dict or a list of dicts),
    set up on the device with its values loaded."""
    import torch

    if isinstance(systems, dict):
        systems = [systems]
    s = describe(systems, use_mas=use_mas, global_diag=global_diag, bank=bank, levels=levels)
    s.setup()
    for i, w in enumerate(systems):
        if device_values:
            s.set_values(i, torch.from_numpy(np.ascontiguousarray(w["vals"])).cuda())
        else:
            s.set_values(i, w["vals"])
    torch.cuda.synchronize()
    return s


def rhs_of(systems):
    if isinstance(systems, dict):
        systems = [systems]
    return np.concatenate([w["rhs"] for w in systems])
