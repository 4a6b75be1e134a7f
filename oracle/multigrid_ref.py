"""Oracle restatement of the geometric-multigrid-preconditioned CG
(core/multigrid.py:38-583).  Test infrastructure only.

Full regular grid, voids carried at ``E*e_min`` by the caller.  Hierarchy halves
while every dimension is even (core/multigrid.py:144-148); level ``l`` scales
its element stiffness by ``2**l`` (:161-162); coarse E is the 2x2x2 mean
(:287-290); coarse free DOFs are injected from the collocated fine node
(:255-278); R = P^T unscaled (:404-416); V(n_smooth, n_smooth) with damped
Jacobi and a 30-sweep coarsest level (:426-436); Fletcher-Reeves PCG with the
stall guard (:439-583).

``coarse="direct"`` replaces the coarsest level's 30 sweeps by the exact
solve of that level's operator on its free DOFs (no counterpart in the
reference, which only smooths there).  The dense matrix is assembled from the
same scaled element stiffness as ``_apply``; ``coarse_direct_factor`` keeps
the sweeps when the free block has more than 1500 DOFs or a Cholesky pivot
is not above 1e-13 times the largest diagonal entry.  Pinned by
tests/test_oracle_mg.py: the dense matrix against ``_apply`` column by
column, and K x = b on the free DOFs for the solve.
"""

import numpy as np

from .fea_ref import hex8_KE, build_edof


def node_dims(nx, ny, nz):
    """core/multigrid.py:38-39."""
    return (nx + 1, ny + 1, nz + 1)


def prolong_op(fine_dims, coarse_dims):
    """Trilinear fine<-coarse interpolation as (idx[Nf,8], w[Nf,8]);
    corners x-fastest, indices clipped at the upper face (core/multigrid.py:42-73)."""
    Lfx, Lfy, Lfz = fine_dims
    Lcx, Lcy, Lcz = coarse_dims
    n = np.arange(Lfx * Lfy * Lfz)
    pos = np.stack([n % Lfx, (n // Lfx) % Lfy, n // (Lfx * Lfy)], axis=1) / 2.0
    base = np.floor(pos).astype(np.int64)
    frac = pos - base
    idx = np.empty((n.size, 8), dtype=np.int64)
    w = np.empty((n.size, 8))
    k = 0
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                j = [np.clip(base[:, a] + o, 0, L - 1)
                     for a, (o, L) in enumerate(((ox, Lcx), (oy, Lcy), (oz, Lcz)))]
                ws = [frac[:, a] if o else 1.0 - frac[:, a]
                      for a, o in enumerate((ox, oy, oz))]
                idx[:, k] = j[0] + Lcx * (j[1] + Lcy * j[2])
                w[:, k] = ws[0] * ws[1] * ws[2]
                k += 1
    return idx, w


def level_dims(nx, ny, nz, max_levels=6, min_elems=4):
    """Grid hierarchy rule, core/multigrid.py:141-148."""
    dims = [(nx, ny, nz)]
    while (len(dims) < max_levels and all(d % 2 == 0 for d in dims[-1])
           and min(dims[-1]) // 2 >= min_elems // 2 and min(dims[-1]) >= 2):
        dims.append(tuple(d // 2 for d in dims[-1]))
    return dims


def coarse_direct_factor(K, free, max_free=1500):
    """Cholesky factor of K restricted to the free DOFs, as ``(free, L)``, or
    None when there are more than ``max_free`` free DOFs or a pivot is not
    above 1e-13 times the largest diagonal entry."""
    idx = np.flatnonzero(free)
    nf = idx.size
    if nf == 0 or nf > max_free:
        return None
    A = K[np.ix_(idx, idx)].copy()
    dmax = np.diag(A).max()
    L = np.zeros_like(A)
    for j in range(nf):
        d = A[j, j] - L[j, :j] @ L[j, :j]
        if not d > 1e-13 * dmax:
            return None
        L[j, j] = np.sqrt(d)
        L[j + 1:, j] = (A[j + 1:, j] - L[j + 1:, :j] @ L[j, :j]) / L[j, j]
    return idx, L


def _forward(L, b):
    """Solve L y = b for lower-triangular L."""
    y = np.zeros_like(b)
    for i in range(b.size):
        y[i] = (b[i] - L[i, :i] @ y[:i]) / L[i, i]
    return y


class StallError(RuntimeError):
    """The MG-PCG stall/divergence guard fired (core/multigrid.py:524-531)."""


class MGSolverRef:
    """Restatement of ``core.multigrid.MGSolver`` on numpy."""

    def __init__(self, nx, ny, nz, fixed_dofs, nu=0.3, n_smooth=2, omega=0.6,
                 min_elems=4, max_levels=6, pool_workers=0, verbose=False,
                 coarse="sweeps", level_smooth=None):
        self.coarse = coarse
        # per-level smoothing steps (level 0 first; 0 or past the list:
        # n_smooth).  The reference smooths every level alike.
        self.level_smooth = list(level_smooth or [])
        self.coarse_chol = None
        self.nx, self.ny, self.nz = nx, ny, nz
        self.ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
        self.n_smooth, self.omega = n_smooth, omega
        self.KE = hex8_KE(1.0, nu)
        self.diagKE = np.diag(self.KE).copy()
        self.levels = []
        for li, (lx, ly, lz) in enumerate(level_dims(nx, ny, nz, max_levels,
                                                     min_elems)):
            self.levels.append({"dims": (lx, ly, lz),
                                "ndof": 3 * (lx + 1) * (ly + 1) * (lz + 1),
                                "edof": build_edof(lx, ly, lz),
                                "g": float(2 ** li)})
        self.prolong = [prolong_op(node_dims(*self.levels[i]["dims"]),
                                   node_dims(*self.levels[i + 1]["dims"]))
                        for i in range(len(self.levels) - 1)]
        self._build_free(fixed_dofs)
        self._pool = None
        if pool_workers and pool_workers > 1:
            from .parallel_cpu_ref import CPUMatVecPoolRef
            lv0 = self.levels[0]
            self._pool = CPUMatVecPoolRef(lv0["edof"], self.KE, lv0["ndof"],
                                          pool_workers)
        self.last_iters = 0
        self.last_converged = False
        self.last_resid_ratio = float("nan")
        self.resid_history = []

    def close(self):
        if self._pool is not None:
            self._pool.close()
            self._pool = None

    def _build_free(self, fixed_dofs):
        """Injection of the fixed mask to coarse levels (core/multigrid.py:255-278)."""
        free = np.ones(self.levels[0]["ndof"], dtype=bool)
        fd = np.asarray(fixed_dofs, dtype=np.int64)
        if fd.size:
            free[fd] = False
        self.levels[0]["free"] = free
        for li in range(1, len(self.levels)):
            Lfx, Lfy, _ = node_dims(*self.levels[li - 1]["dims"])
            Lcx, Lcy, Lcz = node_dims(*self.levels[li]["dims"])
            n = np.arange(Lcx * Lcy * Lcz)
            fine = 2 * (n % Lcx) + Lfx * (2 * ((n // Lcx) % Lcy)
                                          + Lfy * 2 * (n // (Lcx * Lcy)))
            prev = self.levels[li - 1]["free"].reshape(-1, 3)
            self.levels[li]["free"] = prev[fine].reshape(-1).copy()

    def set_density(self, Evec_full):
        """Per-level E (2x2x2 mean), scaled E and Jacobi inverse diagonal
        (core/multigrid.py:281-315)."""
        E = np.asarray(Evec_full, dtype=float)
        for li, lv in enumerate(self.levels):
            if li > 0:
                plx, ply, plz = self.levels[li - 1]["dims"]
                lx, ly, lz = lv["dims"]
                E = (self.levels[li - 1]["Evec"].reshape(plx, ply, plz)
                     .reshape(lx, 2, ly, 2, lz, 2).mean(axis=(1, 3, 5)).ravel())
            lv["Evec"] = E
            scaled = E * lv["g"]
            lv["_scaled"] = scaled
            if li == 0 and self._pool is not None:
                self._pool.set_density(scaled)
            d = np.bincount(lv["edof"].ravel(),
                            weights=(scaled[:, None] * self.diagKE[None, :]).ravel(),
                            minlength=lv["ndof"])
            lv["Minv"] = 1.0 / np.where(d == 0, 1.0, d)
        self.coarse_chol = None
        if (self.coarse == "direct" and len(self.levels) > 1
                and 0 < self.levels[-1]["free"].sum() <= 1500):
            self.coarse_chol = coarse_direct_factor(self.dense_operator(len(self.levels) - 1),
                                                    self.levels[-1]["free"])

    def dense_operator(self, li):
        """The level's K (unmasked), assembled densely from its scaled E."""
        lv = self.levels[li]
        n = lv["ndof"]
        K = np.zeros((n, n))
        for e, dofs in enumerate(lv["edof"]):
            K[np.ix_(dofs, dofs)] += lv["_scaled"][e] * self.KE
        return K

    def _coarse_solve(self, b):
        """Exact solve on the coarsest level's free DOFs, 0 elsewhere."""
        free, L = self.coarse_chol
        x = np.zeros_like(b)
        y = _forward(L, b[free])
        x[free] = _forward(L.T[::-1, ::-1], y[::-1])[::-1]
        return x

    def _Ku(self, li, u):
        lv = self.levels[li]
        ke = (u[lv["edof"]] @ self.KE.T) * lv["_scaled"][:, None]
        return np.bincount(lv["edof"].ravel(), weights=ke.ravel(),
                           minlength=lv["ndof"])

    def _apply(self, li, u):
        """Masked matrix-free K u (core/multigrid.py:318-341)."""
        return np.where(self.levels[li]["free"], self._Ku(li, u), 0.0)

    def _apply_pooled(self, u):
        """Level-0 outer matvec, through the CPU pool when one exists
        (core/multigrid.py:343-365)."""
        if self._pool is None:
            return self._apply(0, u)
        return np.where(self.levels[0]["free"], self._pool.apply(u), 0.0)

    def _smooth(self, li, u, b, iters):
        """Damped Jacobi (core/multigrid.py:367-402, un-pooled branch)."""
        lv = self.levels[li]
        for _ in range(iters):
            r = b - self._apply(li, u)
            u = u + self.omega * np.where(lv["free"], lv["Minv"] * r, 0.0)
        return u

    def _restrict(self, li, r_fine):
        """R = P^T, unscaled, then the coarse free mask (core/multigrid.py:404-416)."""
        idx, wt = self.prolong[li]
        Nc = self.levels[li + 1]["ndof"] // 3
        rf = r_fine.reshape(-1, 3)
        rc = np.zeros((Nc, 3))
        for c in range(3):
            rc[:, c] = np.bincount(idx.ravel(),
                                   weights=(wt * rf[:, c][:, None]).ravel(),
                                   minlength=Nc)
        return np.where(self.levels[li + 1]["free"], rc.reshape(-1), 0.0)

    def _prolongate(self, li, e_coarse):
        """core/multigrid.py:418-424."""
        idx, wt = self.prolong[li]
        ef = (e_coarse.reshape(-1, 3)[idx] * wt[:, :, None]).sum(axis=1)
        return np.where(self.levels[li]["free"], ef.reshape(-1), 0.0)

    def _vcycle(self, li, b):
        """core/multigrid.py:426-436."""
        if li == len(self.levels) - 1:
            if self.coarse_chol is not None and li > 0:
                return self._coarse_solve(b)
            return self._smooth(li, np.zeros_like(b), b, 30)
        ns = self._ns(li)
        u = self._smooth(li, np.zeros_like(b), b, ns)
        rc = self._restrict(li, b - self._apply(li, u))
        u = u + self._prolongate(li, self._vcycle(li + 1, rc))
        return self._smooth(li, u, b, ns)

    def _ns(self, li):
        """Smoothing steps of level li (n_smooth everywhere in the reference,
        core/multigrid.py:426-436)."""
        k = self.level_smooth[li] if li < len(self.level_smooth) else 0
        return k if k else self.n_smooth

    def solve(self, Evec_full, f, tol=1e-4, max_cg=None, x0=None,
              progress_cb=None):
        """PCG with V-cycle preconditioner and stall guard
        (core/multigrid.py:439-583).  Raises ``StallError`` on a stall."""
        self.resid_history = []
        self.set_density(Evec_full)
        free = self.levels[0]["free"]
        f = np.where(free, np.asarray(f, dtype=float), 0.0)
        if x0 is None:
            u = np.zeros(self.ndof)
            r = f.copy()
        else:
            u = np.where(free, np.asarray(x0, dtype=float), 0.0)
            r = f - self._apply_pooled(u)
        r = np.where(free, r, 0.0)
        fnorm = float(np.linalg.norm(f))
        if fnorm == 0.0:
            return np.zeros(self.ndof)
        z = self._vcycle(0, r)
        p = z.copy()
        rz = float(r @ z)
        max_cg = max_cg or 1000
        self.last_iters = 0
        self.last_converged = False
        self.last_resid_ratio = float("nan")
        best = float("inf")
        for it in range(max_cg):
            Ap = self._apply_pooled(p)
            pAp = float(p @ Ap)
            if pAp == 0.0:
                self.last_converged = self.last_resid_ratio < tol
                break
            alpha = rz / pAp
            u = u + alpha * p
            r = r - alpha * Ap
            self.last_iters = it + 1
            self.last_resid_ratio = float(np.linalg.norm(r)) / fnorm
            self.resid_history.append(self.last_resid_ratio)
            if progress_cb is not None:
                progress_cb(self.last_iters, self.last_resid_ratio, tol)
            if self.last_resid_ratio < tol:
                self.last_converged = True
                break
            best = min(best, self.last_resid_ratio)
            if self.last_iters >= 10 and self.last_resid_ratio > 5.0 * best:
                raise StallError(
                    f"MG-PCG stalled/diverged at iter {self.last_iters} "
                    f"(ratio {self.last_resid_ratio:.3e}, best {best:.3e})")
            z = self._vcycle(0, r)
            rz_new = float(r @ z)
            p = z.copy() if rz == 0.0 else z + (rz_new / rz) * p
            rz = rz_new
        return np.where(free, u, 0.0)
