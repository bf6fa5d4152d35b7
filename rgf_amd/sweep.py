"""Frequency and Bloch-phase scans over one slab, through the batched solve,
with the results written to disk.

This is the caller the batched entry (``rgf_solve_batched``) exists for: the
reference's examples solve one frequency at a time and pick input channels by
incident angle (examples/example_script.m:43-53); a scan repeats that over a
grid of ``k0dx`` values and Bloch phases.  Results go to ``.npz`` (numpy) or
``.mat`` (MATLAB v5, ``scipy.io.savemat``: S as a cell array).
"""
from __future__ import annotations

import itertools
import os

import numpy as np

from .api import cal_smatrix_RGF_batched, setup_channels

__all__ = ["scan", "channels_at_angles", "save_scan", "load_scan"]


def channels_at_angles(channels: dict, angles_deg) -> np.ndarray:
    """1-based left input channels nearest to the incident angles (degrees), as
    examples/example_script.m:45-48 computes them: the channel index
    interpolated linearly in atan2(kydx, kxdx) over the propagating channels,
    then rounded."""
    side = channels["L"]
    prop = np.asarray(side["ind_prop"]) - 1
    if prop.size == 0:
        return np.zeros(0, dtype=np.int64)
    ang = np.arctan2(np.asarray(channels["kydx"])[prop].real, np.asarray(side["kxdx"])[prop].real)
    idx = np.arange(1, prop.size + 1, dtype=float)
    order = np.argsort(ang, kind="stable")
    want = np.asarray(angles_deg, dtype=float) / 180 * np.pi
    if np.any(want < ang[order[0]]) or np.any(want > ang[order[-1]]):
        raise ValueError("incident angle outside the range of the propagating channels")
    return np.rint(np.interp(want, ang[order], idx[order])).astype(np.int64)


def scan(syst: dict, *, k0dx=None, bc=None, out=("L", "R"), in_=("L", "R"), angles_deg=None,
         precision: str = "fp64", device: int = -1, options: dict | None = None, chunk: int = 64,
         path: str | None = None) -> dict:
    """S over the grid k0dx x bc (every pair) for one slab.

    ``syst`` is a 2D (``yBC``) or 3D (``xyBC``) system as cal_smatrix_RGF /
    cal_smatrix_RGF_3D take it; ``k0dx`` and ``bc`` default to its own value.
    ``angles_deg`` (2D) replaces ``in_`` by the left channels nearest to those
    incident angles at each grid point.  Solves run ``chunk`` at a time
    through the batched entry.  Returns a dict with ``k0dx``, ``bc``, ``S``
    (list, one matrix per grid point, row-major over (k0dx, bc)), ``in_index``
    (the input selection used at each point) and, with ``path``, writes it
    (see save_scan).
    """
    three_d = np.ndim(syst["epsilon"]) == 3
    bc_key = "xyBC" if three_d else "yBC"
    ks = list(np.atleast_1d(syst["k0dx"] if k0dx is None else k0dx))
    bcs = [syst[bc_key]] if bc is None else list(bc)
    if angles_deg is not None and three_d:
        raise ValueError("angles_deg selects 2D channels; pass in_ for 3D")
    points = list(itertools.product(ks, bcs))
    items, sel_in = [], []
    for k, b in points:
        s = dict(syst, k0dx=k)
        s[bc_key] = b
        if angles_deg is not None:
            ch = setup_channels(np.shape(s["epsilon"])[0], b, k, s["epsilon_L"], s.get("epsilon_R", s["epsilon_L"]))
            inn = {"ind_in_L": channels_at_angles(ch, angles_deg)}
        else:
            inn = in_ if isinstance(in_, dict) else list(in_)
        sel_in.append(inn)
        items.append((s, out if isinstance(out, dict) else list(out), inn))
    S = []
    for q in range(0, len(items), max(1, int(chunk))):
        S += cal_smatrix_RGF_batched(items[q:q + chunk], precision=precision, device=device, options=options)
    res = {"k0dx": np.array([p[0] for p in points]), "bc": [p[1] for p in points], "S": S,
           "in_index": [np.asarray(x["ind_in_L"]) if isinstance(x, dict) else np.asarray(x) for x in sel_in]}
    if path:
        save_scan(res, path)
    return res


def _bc_array(bcs):
    if all(isinstance(b, str) for b in bcs):
        return np.array(bcs)
    return np.array([np.atleast_1d(np.asarray(b, dtype=float)) for b in bcs])


def save_scan(res: dict, path: str) -> None:
    """Write a scan: ``.npz`` stores ``k0dx``, ``bc``, ``S_00000``... and
    ``in_00000``... (no pickled objects); ``.mat`` stores ``k0dx``, ``bc`` and
    the cell arrays ``S`` and ``in_index``."""
    ext = os.path.splitext(path)[1].lower()
    n = len(res["S"])
    if ext == ".npz":
        arrs = {"k0dx": np.asarray(res["k0dx"]), "bc": _bc_array(res["bc"])}
        for q in range(n):
            arrs[f"S_{q:05d}"] = np.asarray(res["S"][q])
            arrs[f"in_{q:05d}"] = np.asarray(res["in_index"][q]).astype(str) \
                if np.asarray(res["in_index"][q]).dtype.kind in "US" else np.asarray(res["in_index"][q])
        np.savez(path, **arrs)
    elif ext == ".mat":
        from scipy.io import savemat
        cellS = np.empty((n, 1), dtype=object)
        cellI = np.empty((n, 1), dtype=object)
        for q in range(n):
            cellS[q, 0] = np.asarray(res["S"][q])
            cellI[q, 0] = np.asarray(res["in_index"][q])
        savemat(path, {"k0dx": np.asarray(res["k0dx"]), "bc": _bc_array(res["bc"]), "S": cellS,
                       "in_index": cellI})
    else:
        raise ValueError("path must end in .npz or .mat")


def load_scan(path: str) -> dict:
    """Read a scan written by save_scan (no pickled objects are loaded)."""
    ext = os.path.splitext(path)[1].lower()
    if ext == ".npz":
        with np.load(path, allow_pickle=False) as z:
            n = sum(1 for k in z.files if k.startswith("S_"))
            return {"k0dx": z["k0dx"], "bc": z["bc"], "S": [z[f"S_{q:05d}"] for q in range(n)],
                    "in_index": [z[f"in_{q:05d}"] for q in range(n)]}
    if ext == ".mat":
        from scipy.io import loadmat
        m = loadmat(path)
        return {"k0dx": m["k0dx"].ravel(), "bc": m["bc"], "S": [c for c in m["S"].ravel()],
                "in_index": [c.ravel() for c in m["in_index"].ravel()]}
    raise ValueError("path must end in .npz or .mat")
