"""Oracle restatement of the post-solve field arithmetic of the reference's
solve worker (core/solve_worker.py:358-431).  Test infrastructure only."""

import numpy as np

from .fea_ref import build_edof, von_mises

STRESS_PERCENTILE = 99.0      # core/solve_worker.py:32-37
SAFETY_CAP = 10.0             # core/solve_worker.py:388


def post_fields(nx, ny, nz, nu, vsize, u_full, E_stress, active, yield_pa,
                converged):
    """Fields of ``final.pkl`` that derive from u (core/solve_worker.py:358-431).

    ``E_stress`` is the per-element modulus in Pa (not scaled by vsize); the
    von Mises field is divided by ``vsize`` as the worker does (:358)."""
    active = np.asarray(active, dtype=bool).ravel()
    nelem = nx * ny * nz
    vm = von_mises(nx, ny, nz, nu, u_full, E_stress, active) / vsize
    act = vm[active]
    max_stress = float(act.max()) if act.size else 0.0
    vmax = float(np.percentile(act, STRESS_PERCENTILE)) if act.size else 0.0
    vmax = max(vmax, 1e-6)
    mag = np.sqrt((np.asarray(u_full).reshape(-1, 3) ** 2).sum(axis=1))
    max_disp = float(mag.max()) if mag.size else 0.0
    nodes = build_edof(nx, ny, nz)[:, 0::3] // 3
    disp3d = mag[nodes].mean(axis=1).reshape(nx, ny, nz)
    safety = np.zeros(nelem)
    min_sf = None
    if yield_pa > 0 and act.size:
        eps = max(1e-9, 1e-9 * yield_pa)
        ratio = yield_pa / np.maximum(act, eps)
        safety[active] = np.clip(ratio, 0.0, SAFETY_CAP)
        min_sf = float(ratio.min())
    active3d = active.reshape(nx, ny, nz)
    disp_vmax = max(float(disp3d[active3d].max()) if active3d.any() else 0.0, 1e-9)
    bbox = float(np.sqrt(((np.array([nx, ny, nz]) * vsize) ** 2).sum()))
    return {
        "stress3d": vm.reshape(nx, ny, nz).astype(np.float32),
        "stress_vmax": vmax,
        "max_stress_pa": max_stress,
        "safety3d": safety.reshape(nx, ny, nz).astype(np.float32),
        "safety_cap": SAFETY_CAP,
        "min_safety_factor": min_sf,
        "displacement3d": disp3d.astype(np.float32),
        "displacement_vmax_m": disp_vmax,
        "max_displacement_m": max_disp,
        "reliable": bool(converged) and max_disp <= 0.5 * bbox,
        "vm_pa": vm,
    }
