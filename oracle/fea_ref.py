"""Oracle restatement of the reference element kernels, connectivity, island
removal and the ``VoxelFEA`` solve wrapper (core/fea.py).

Test infrastructure only (see ``oracle/__init__.py``).  Conventions, all as in
the reference:

* Elements are x-slowest / z-fastest: ``e = iz + nz*(iy + ny*ix)``
  (core/fea.py:135-139, numpy meshgrid ``indexing='ij'``).
* Nodes are x-fastest: ``n = ix + (nx+1)*(iy + (ny+1)*iz)`` (core/fea.py:129-130).
* DOF ``3*n + c``; element-local node order (-,-,-),(+,-,-),(+,+,-),(-,+,-)
  then the same at z=+ (core/fea.py:33-36, 140-141).
* Voigt order [xx, yy, zz, xy, yz, zx] with engineering shear strains
  (core/fea.py:39-46, 62-74).
"""

import numpy as np

# Local node corner offsets, core/fea.py:140-141 (and the +-1 reference
# coordinates of core/fea.py:33-36 are 2*offset-1).
LOCAL_OFFSETS = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                 (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))


def _iso_C(nu, E=1.0):
    """Isotropic 6x6 material matrix in Voigt order (core/fea.py:38-46)."""
    c = E / ((1.0 + nu) * (1.0 - 2.0 * nu))
    g = (1.0 - 2.0 * nu) / 2.0
    C = np.zeros((6, 6))
    C[:3, :3] = nu
    C[0, 0] = C[1, 1] = C[2, 2] = 1.0 - nu
    C[3, 3] = C[4, 4] = C[5, 5] = g
    return c * C


def _B_at(xi, eta, zeta):
    """Strain-displacement matrix of the unit cube at a reference point.

    Shape-function derivatives in reference coordinates are scaled by the
    inverse Jacobian 1/0.5 (J = 0.5*I, core/fea.py:59-61)."""
    B = np.zeros((6, 24))
    for i, (ox, oy, oz) in enumerate(LOCAL_OFFSETS):
        a, b, c = 2 * ox - 1, 2 * oy - 1, 2 * oz - 1
        dx = a * (1 + b * eta) * (1 + c * zeta) / 8.0 / 0.5
        dy = b * (1 + a * xi) * (1 + c * zeta) / 8.0 / 0.5
        dz = c * (1 + a * xi) * (1 + b * eta) / 8.0 / 0.5
        k = 3 * i
        B[0, k] = dx
        B[1, k + 1] = dy
        B[2, k + 2] = dz
        B[3, k], B[3, k + 1] = dy, dx
        B[4, k + 1], B[4, k + 2] = dz, dy
        B[5, k], B[5, k + 2] = dz, dx
    return B


def hex8_KE(E=1.0, nu=0.3):
    """24x24 unit-cube stiffness by 2x2x2 Gauss quadrature, det J = 1/8.

    Follows core/fea.py:31-76 (same quadrature order: xi outer, zeta inner).
    The reference takes det J from ``np.linalg.det``, which goes through
    log-determinants and returns 0.12500000000000003 for 0.5*I rather than
    1/8; the same call is made here so KE agrees to the last bit."""
    C = _iso_C(nu, E)
    g = 1.0 / np.sqrt(3.0)
    detJ = np.linalg.det(0.5 * np.eye(3))
    KE = np.zeros((24, 24))
    for xi in (-g, g):
        for eta in (-g, g):
            for zeta in (-g, g):
                B = _B_at(xi, eta, zeta)
                KE += (B.T @ C @ B) * detJ
    return KE


def hex8_B_C(nu=0.3):
    """Centroid B (6x24) and unit-E C (6x6), core/fea.py:79-122."""
    return _B_at(0.0, 0.0, 0.0), _iso_C(nu, 1.0)


def node_id(ix, iy, iz, nx, ny):
    """core/fea.py:129-130."""
    return ix + (nx + 1) * (iy + (ny + 1) * iz)


def build_edof(nx, ny, nz):
    """[Ne, 24] int64 global DOF map, core/fea.py:133-148."""
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                             indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    edof = np.empty((ex.size, 24), dtype=np.int64)
    for k, (ox, oy, oz) in enumerate(LOCAL_OFFSETS):
        n = node_id(ex + ox, ey + oy, ez + oz, nx, ny)
        for c in range(3):
            edof[:, 3 * k + c] = 3 * n + c
    return edof


# ---------------------------------------------------------------------------
# Island removal, core/fea.py:165-227
# ---------------------------------------------------------------------------

def seed_from_fixed_dofs(nx, ny, nz, fixed_dofs):
    """Elements touching at least one fixed node (core/fea.py:165-177)."""
    fixed_dofs = np.asarray(fixed_dofs, dtype=np.int64)
    if fixed_dofs.size == 0:
        return np.zeros((nx, ny, nz), dtype=bool)
    is_fixed = np.zeros((nx + 1) * (ny + 1) * (nz + 1), dtype=bool)
    is_fixed[np.unique(fixed_dofs // 3)] = True
    nodes = build_edof(nx, ny, nz)[:, 0::3] // 3
    return is_fixed[nodes].any(axis=1).reshape(nx, ny, nz)


def connected_component_mask(active3d, seed3d):
    """6-connected flood fill by repeated dilation (core/fea.py:180-204)."""
    active3d = np.asarray(active3d, dtype=bool)
    reached = np.asarray(seed3d, dtype=bool) & active3d
    if not reached.any():
        return reached
    while True:
        grown = reached.copy()
        for ax in range(3):
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[ax], hi[ax] = slice(1, None), slice(None, -1)
            grown[tuple(lo)] |= reached[tuple(hi)]
            grown[tuple(hi)] |= reached[tuple(lo)]
        grown &= active3d
        if np.array_equal(grown, reached):
            return reached
        reached = grown


def drop_unsupported_islands(active3d, fixed_dofs):
    """(kept3d, n_removed), core/fea.py:207-227."""
    active3d = np.asarray(active3d, dtype=bool)
    nx, ny, nz = active3d.shape
    kept = connected_component_mask(
        active3d, seed_from_fixed_dofs(nx, ny, nz, fixed_dofs))
    if not kept.any():
        return active3d, 0
    return kept, int(active3d.sum() - kept.sum())


# ---------------------------------------------------------------------------
# VoxelFEA wrapper, core/fea.py:234-626
# ---------------------------------------------------------------------------

class VoxelFEARef:
    """Restatement of ``core.fea.VoxelFEA`` (numpy, single process).

    ``mg_kwargs`` are passed to the multigrid oracle (``max_levels`` etc.).
    ``pool_workers`` > 1 routes the level-0 outer matvec of the MG path and
    every matvec of the Jacobi path through the multi-process CPU pool
    restatement (core/fea.py:307-340, core/multigrid.py:190-196).
    """

    def __init__(self, nx, ny, nz, nu=0.3, e_min=1e-9, active_elems=None,
                 use_multigrid=True, mg_kwargs=None, pool_workers=0,
                 verbose=False):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nelem = nx * ny * nz
        self.ndof_full = 3 * (nx + 1) * (ny + 1) * (nz + 1)
        self.nu, self.e_min = nu, e_min
        self.use_multigrid = use_multigrid
        self.mg_kwargs = dict(mg_kwargs or {})
        self.pool_workers = int(pool_workers)
        self.verbose = verbose
        self._mg = None
        self._pool = None
        self._fixed_dofs_full = np.zeros(0, dtype=np.int64)
        self.KE = hex8_KE(1.0, nu)
        self._B0, self._C0 = hex8_B_C(nu)
        self.edof_full = build_edof(nx, ny, nz)
        if active_elems is None:
            self.active_elems = np.ones(self.nelem, dtype=bool)
        else:
            self.active_elems = np.asarray(active_elems, dtype=bool).ravel()
        # Active-DOF compression, core/fea.py:271-279.
        edof_a = self.edof_full[self.active_elems]
        used = np.unique(edof_a)
        remap = -np.ones(self.ndof_full, dtype=np.int64)
        remap[used] = np.arange(used.size)
        self.used_dofs, self.remap = used, remap
        self.ndof = int(used.size)
        self.edof = remap[edof_a]
        self.free = None
        self.last_iters = 0
        self.last_converged = None
        self.last_resid_ratio = float("nan")
        self.resid_history = []
        self.fell_back = False
        if self.pool_workers > 1:
            from .parallel_cpu_ref import CPUMatVecPoolRef
            self._pool = CPUMatVecPoolRef(self.edof, self.KE, self.ndof,
                                          self.pool_workers)

    def close(self):
        for attr in ("_pool", "_mg"):
            obj = getattr(self, attr)
            if obj is not None:
                try:
                    obj.close()
                except Exception:
                    pass
                setattr(self, attr, None)

    def set_fixed(self, fixed_dofs):
        """core/fea.py:357-369."""
        fixed_dofs = np.asarray(fixed_dofs, dtype=np.int64)
        self._fixed_dofs_full = fixed_dofs
        free = np.ones(self.ndof, dtype=bool)
        fd = self.remap[fixed_dofs]
        free[fd[fd >= 0]] = False
        self.free = free

    def _apply_K(self, u, Evec_a):
        """Active-space K u (core/fea.py:374-386)."""
        if self._pool is not None:
            return self._pool.apply(u)
        ke_ue = (u[self.edof] @ self.KE.T) * Evec_a[:, None]
        return np.bincount(self.edof.ravel(), weights=ke_ue.ravel(),
                           minlength=self.ndof)

    def _diagonal(self, Evec_a):
        """core/fea.py:388-398."""
        if self._pool is not None:
            d = self._pool.diagonal()
        else:
            contrib = Evec_a[:, None] * np.diag(self.KE)[None, :]
            d = np.bincount(self.edof.ravel(), weights=contrib.ravel(),
                            minlength=self.ndof)
        return np.where(d == 0, 1.0, d)

    def solve(self, Evec, f, tol=1e-4, max_cg=None, x0=None,
              progress_cb=None):
        """MG first, Jacobi-PCG on any MG failure (core/fea.py:400-436)."""
        if self.free is None:
            raise RuntimeError("call set_fixed() before solve()")
        if self.use_multigrid:
            u = self._solve_multigrid(Evec, f, tol, max_cg, x0, progress_cb)
            if u is not None:
                return u
            self.fell_back = True
        return self._solve_jacobi(Evec, f, tol=tol, max_cg=max_cg, x0=x0,
                                  progress_cb=progress_cb)

    def _solve_multigrid(self, Evec, f, tol, max_cg, x0, progress_cb):
        """core/fea.py:438-495: the MG solver is built lazily with the fixed
        DOFs of the first solve and sees voids at E*e_min."""
        from .multigrid_ref import MGSolverRef
        if self._mg is None:
            try:
                self._mg = MGSolverRef(self.nx, self.ny, self.nz,
                                       self._fixed_dofs_full, nu=self.nu,
                                       pool_workers=self.pool_workers,
                                       **self.mg_kwargs)
            except Exception as exc:
                print(f"[oracle] MG construction failed: {exc}")
                self.use_multigrid = False
                self._mg = None
                return None
        Evec_full = np.asarray(Evec, dtype=float)
        Evec_mg = np.where(self.active_elems, Evec_full,
                           Evec_full * self.e_min)
        try:
            u = self._mg.solve(Evec_mg, np.asarray(f, dtype=float), tol=tol,
                               max_cg=max_cg, x0=x0, progress_cb=progress_cb)
        except Exception as exc:
            print(f"[oracle] MG solve failed: {exc}; Jacobi fallback")
            self.use_multigrid = False
            try:
                self._mg.close()
            except Exception:
                pass
            self._mg = None
            return None
        self.last_iters = self._mg.last_iters
        self.last_converged = self._mg.last_converged
        self.last_resid_ratio = self._mg.last_resid_ratio
        self.resid_history = list(self._mg.resid_history)
        return u

    def _solve_jacobi(self, Evec, f, tol=1e-4, max_cg=None, x0=None,
                      progress_cb=None):
        """Jacobi-PCG on active DOFs only, no e_min voids
        (core/fea.py:497-593)."""
        if self.free is None:
            raise RuntimeError("call set_fixed() before solve()")
        free = self.free
        Evec_a = np.asarray(Evec, dtype=float)[self.active_elems]
        if self._pool is not None:
            self._pool.set_density(Evec_a)
        f_r = np.asarray(f, dtype=float)[self.used_dofs]
        Minv = 1.0 / self._diagonal(Evec_a)
        self.last_iters = 0
        self.last_converged = False
        self.last_resid_ratio = float("nan")
        self.resid_history = []
        if x0 is None:
            u = np.zeros(self.ndof)
            r = f_r.copy()
        else:
            u = np.where(free, np.asarray(x0, dtype=float)[self.used_dofs], 0.0)
            r = f_r - self._apply_K(u, Evec_a)
        r = np.where(free, r, 0.0)
        z = np.where(free, Minv * r, 0.0)
        p = z.copy()
        rz = float(r @ z)
        fnorm = float(np.linalg.norm(np.where(free, f_r, 0.0)))
        if fnorm == 0:
            self.last_converged = True
            return np.zeros(self.ndof_full)
        max_cg = max_cg or 4 * int(free.sum())
        for it in range(max_cg):
            Kp = np.where(free, self._apply_K(p, Evec_a), 0.0)
            pKp = float(p @ Kp)
            if pKp == 0.0:
                self.last_converged = (self.last_resid_ratio < tol)
                break
            alpha = rz / pKp
            u = u + alpha * p
            r = r - alpha * Kp
            self.last_iters = it + 1
            self.last_resid_ratio = (
                float(np.linalg.norm(np.where(free, r, 0.0))) / fnorm)
            self.resid_history.append(self.last_resid_ratio)
            if progress_cb is not None:
                progress_cb(self.last_iters, self.last_resid_ratio, tol)
            if self.last_resid_ratio < tol:
                self.last_converged = True
                break
            z = np.where(free, Minv * r, 0.0)
            rz_new = float(r @ z)
            p = z.copy() if rz == 0.0 else z + (rz_new / rz) * p
            p = np.where(free, p, 0.0)
            rz = rz_new
        u = np.where(free, u, 0.0)
        u_full = np.zeros(self.ndof_full)
        u_full[self.used_dofs] = u
        return u_full

    def element_strain_energy(self, u_full):
        """core/fea.py:595-602."""
        u = np.asarray(u_full)
        ce = np.zeros(self.nelem)
        ue = u[self.edof_full[self.active_elems]]
        ce[self.active_elems] = np.einsum("ei,ei->e", ue, ue @ self.KE.T)
        return ce

    def element_von_mises_stress(self, u_full, Evec):
        """Centroid von Mises scaled by per-element E (core/fea.py:604-626)."""
        return von_mises(self.nx, self.ny, self.nz, self.nu, u_full, Evec,
                         self.active_elems)


def von_mises(nx, ny, nz, nu, u_full, Evec, active_elems):
    """Per-element centroid von Mises, zero outside ``active_elems``
    (core/fea.py:604-626)."""
    active = np.asarray(active_elems, dtype=bool).ravel()
    vm = np.zeros(nx * ny * nz)
    if not active.any():
        return vm
    B0, C0 = hex8_B_C(nu)
    ue = np.asarray(u_full)[build_edof(nx, ny, nz)[active]]
    stress = ((ue @ B0.T) @ C0.T) * np.asarray(Evec, float)[active][:, None]
    sx, sy, sz, txy, tyz, tzx = stress.T
    vm[active] = np.sqrt(np.maximum(
        0.5 * ((sx - sy) ** 2 + (sy - sz) ** 2 + (sz - sx) ** 2)
        + 3.0 * (txy ** 2 + tyz ** 2 + tzx ** 2), 0.0))
    return vm
