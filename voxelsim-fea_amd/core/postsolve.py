"""Post-solve fields on the GPU: the part of the reference's solve worker that
turns u into the result arrays (core/solve_worker.py:358-431).

``final_fields`` returns the same keys as the worker's ``final.pkl`` for the
fields derived from u: von Mises stress in Pa (``stress3d``), its 99th
percentile colour ceiling, safety factor clip(yield / vm, 0, 10), per-element
displacement magnitude (mean over the 8 corner nodes), their maxima, and the
``reliable`` flag (converged and max displacement <= half the bounding-box
diagonal).
"""

import ctypes
import math

import numpy as np

from . import _native
from .multigrid import _dev, _torch

SAFETY_CAP = 10.0


def final_fields(vfea, u_full, E_pa, vsize, yield_pa, converged=True):
    """vfea: a core.fea.VoxelFEA (supplies the grid, nu, active mask and the
    device handle); u_full: full-grid displacement (numpy or torch);
    E_pa: per-element Young's modulus in Pa (reference element order)."""
    torch = _torch()
    nx, ny, nz = vfea.nx, vfea.ny, vfea.nz
    ne = nx * ny * nz
    ud = _dev(u_full)
    Ed = _dev(np.asarray(E_pa, dtype=float) if not isinstance(E_pa, torch.Tensor) else E_pa)
    s3 = torch.empty(ne, dtype=torch.float32, device="cuda")
    f3 = torch.empty_like(s3)
    d3 = torch.empty_like(s3)
    sc = np.zeros(5)
    _native.check(vfea._lib.vx_post_fields(vfea._h, _native.ptr(ud), _native.ptr(Ed),
                                           _native.ptr(vfea._active_dev), float(vsize),
                                           float(yield_pa), _native.ptr(s3), _native.ptr(f3),
                                           _native.ptr(d3), sc.ctypes.data_as(ctypes.c_void_p)),
                  vfea._h)
    bbox = math.sqrt(((np.array([nx, ny, nz]) * vsize) ** 2).sum())
    return {
        "stress3d": s3.cpu().numpy().reshape(nx, ny, nz),
        "stress_vmax": float(sc[0 + 1]),
        "max_stress_pa": float(sc[0]),
        "safety3d": f3.cpu().numpy().reshape(nx, ny, nz),
        "safety_cap": SAFETY_CAP,
        "min_safety_factor": None if math.isnan(sc[4]) else float(sc[4]),
        "displacement3d": d3.cpu().numpy().reshape(nx, ny, nz),
        "displacement_vmax_m": float(sc[3]),
        "max_displacement_m": float(sc[2]),
        "reliable": bool(converged) and float(sc[2]) <= 0.5 * bbox,
    }


def final_payload(vfea, u_full, E_pa, vsize, yield_pa, dims, origin, active3d,
                  n_islands_removed=0, tol=1e-6):
    """The solve worker's complete ``final.pkl`` dict (core/solve_worker.py:411-431):
    the device fields of ``final_fields`` plus the grid, the CG record of the
    solve and the compute label."""
    f = final_fields(vfea, u_full, E_pa, vsize, yield_pa, converged=bool(vfea.last_converged))
    if float(yield_pa) <= 0:
        f["min_safety_factor"] = None
    return {
        "dims": tuple(int(d) for d in dims), "origin": tuple(float(o) for o in origin),
        "vsize": float(vsize),
        "active3d": np.asarray(active3d, dtype=bool).reshape(tuple(int(d) for d in dims)),
        "stress3d": f["stress3d"].astype(np.float32),
        "stress_vmax": f["stress_vmax"],
        "max_stress_pa": f["max_stress_pa"],
        "safety3d": f["safety3d"].astype(np.float32),
        "safety_cap": f["safety_cap"],
        "min_safety_factor": f["min_safety_factor"],
        "yield_strength_pa": float(yield_pa),
        "displacement3d": f["displacement3d"].astype(np.float32),
        "displacement_vmax_m": f["displacement_vmax_m"],
        "max_displacement_m": f["max_displacement_m"],
        "cg_iters": int(vfea.last_iters),
        "cg_converged": bool(vfea.last_converged),
        "cg_resid_history": list(vfea.resid_history),
        "cg_tol": float(tol),
        "reliable": f["reliable"],
        "n_islands_removed": int(n_islands_removed),
        "compute_label": vfea.plan.label,
    }


def write_final(path, payload):
    """Write ``payload`` the way the worker writes final.pkl: to a temporary
    file, then renamed over the target (core/solve_worker.py:189-193)."""
    import os
    import pickle
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        pickle.dump(payload, fh)
    os.replace(tmp, path)
