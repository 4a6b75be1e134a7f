"""Linear-elastic voxel FEA solve wrapper on the GPU (API of core/fea.py).

``VoxelFEA`` keeps the reference's constructor, ``set_fixed`` / ``solve`` /
``element_von_mises_stress`` / ``element_strain_energy`` / ``close`` and the
``last_iters`` / ``last_converged`` / ``last_resid_ratio`` / ``resid_history``
fields, and the same solve semantics:

* the multigrid path solves the FULL grid with voids at ``E*e_min``
  (core/fea.py:460-470) and is built lazily with the fixed DOFs of the first
  solve (core/fea.py:441-447);
* any MG construction or solve failure, the stall guard included, is printed
  and falls back to Jacobi-PCG on the active-DOF system (core/fea.py:448-458,
  476-489, 497-593) -- on the device as well.

Inputs and outputs are numpy arrays, as in the reference; every operation
runs in the HIP library.
"""

import ctypes

import numpy as np

from . import _native
from . import backend
from . import compute_plan
from .multigrid import MGSolver, _dev, _require_gpu, _torch


# ---------------------------------------------------------------------------
# element matrices / connectivity (host utilities with the reference's API)
# ---------------------------------------------------------------------------

def _hex8_KE(E=1.0, nu=0.3):
    """24x24 unit-cube stiffness (core/fea.py:31-76), assembled column by
    column from the library's element operator (the kernels' own code)."""
    lib = _native.load()
    KE = np.zeros((24, 24))
    y = np.zeros(24)
    for j in range(24):
        u = np.zeros(24)
        u[j] = 1.0
        lib.vx_element_apply_host(float(nu), u.ctypes.data, y.ctypes.data)
        KE[:, j] = y
    return E * KE


def _hex8_B_C(nu=0.3):
    """Centroid B (entries +-1/4) and unit-E C (core/fea.py:79-122)."""
    signs = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
    B = np.zeros((6, 24))
    for i, (bx, by, bz) in enumerate(signs / 4.0):
        k = 3 * i
        B[0, k], B[1, k + 1], B[2, k + 2] = bx, by, bz
        B[3, k], B[3, k + 1] = by, bx
        B[4, k + 1], B[4, k + 2] = bz, by
        B[5, k], B[5, k + 2] = bz, bx
    c = 1.0 / ((1 + nu) * (1 - 2 * nu))
    C = np.zeros((6, 6))
    C[:3, :3] = c * nu
    C[np.arange(3), np.arange(3)] = c * (1 - nu)
    C[np.arange(3, 6), np.arange(3, 6)] = c * (1 - 2 * nu) / 2
    return B, C


def _node_id(ix, iy, iz, nx, ny):
    return ix + (nx + 1) * (iy + (ny + 1) * iz)


def build_edof(nx, ny, nz):
    """[Ne, 24] int64 DOF map (core/fea.py:133-148).  The GPU kernels compute
    these indices inline; this host table exists for callers of the API."""
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    edof = np.empty((ex.size, 24), dtype=np.int64)
    offs = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
    for k, (ox, oy, oz) in enumerate(offs):
        n = _node_id(ex + ox, ey + oy, ez + oz, nx, ny)
        for c in range(3):
            edof[:, 3 * k + c] = 3 * n + c
    return edof


def _util_handle(nx, ny, nz, nu=0.3):
    lib = _native.load()
    h = ctypes.c_void_p()
    opt = _native.default_options(max_levels=1)
    _native.check(lib.vx_create(nx, ny, nz, float(nu), ctypes.byref(opt), ctypes.byref(h)))
    lib.vx_set_stream(h, ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream))
    return lib, h


def drop_unsupported_islands(active3d, fixed_dofs):
    """(kept3d, n_removed): keep active elements 6-connected to an element
    touching a fixed node (core/fea.py:207-227), flood fill on the GPU."""
    active3d = np.asarray(active3d, dtype=bool)
    nx, ny, nz = active3d.shape
    fd = np.ascontiguousarray(np.asarray(fixed_dofs, dtype=np.int64).ravel())
    lib, h = _util_handle(nx, ny, nz)
    try:
        a = _dev(active3d.ravel(), _torch().uint8)
        k = _torch().empty_like(a)
        nrem = ctypes.c_int64()
        _native.check(lib.vx_drop_islands(h, _native.ptr(a), fd.ctypes.data_as(ctypes.c_void_p),
                                          fd.size, _native.ptr(k), ctypes.byref(nrem)), h)
        kept = k.cpu().numpy().astype(bool).reshape(nx, ny, nz)
    finally:
        lib.vx_destroy(h)
    return kept, int(nrem.value)


class VoxelFEA:
    """Solves K u = f over the active voxels (core/fea.py:234-626) on the GPU."""

    def __init__(self, nx, ny, nz, nu=0.3, e_min=1e-9, active_elems=None,
                 compute_mode="AUTO", cpu_threads=0, verbose=False, use_multigrid=True,
                 mg_options=None):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.nelem = self.nx * self.ny * self.nz
        self.ndof_full = 3 * (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self.e_min, self.nu = e_min, nu
        self._compute_mode, self._cpu_threads = compute_mode, cpu_threads
        self.use_multigrid = use_multigrid
        self.mg_options = dict(mg_options or {})
        self.verbose = verbose
        self._mg = None
        self._fixed_dofs_full = np.zeros(0, dtype=np.int64)
        self.free = None
        if active_elems is None:
            self.active_elems = np.ones(self.nelem, dtype=bool)
        else:
            self.active_elems = np.asarray(active_elems, dtype=bool).ravel()
        self.last_iters = 0
        self.last_converged = None
        self.last_resid_ratio = float("nan")
        self.resid_history = []
        self.plan = compute_plan.choose(self.ndof_full, mode=compute_mode,
                                        cpu_threads=cpu_threads, verbose=verbose)
        _require_gpu(self.plan, "VoxelFEA")
        self.xp = self.plan.xp
        self._lib, self._h = _util_handle(self.nx, self.ny, self.nz, nu)
        self._active_dev = _dev(self.active_elems, _torch().uint8)
        self.KE = _hex8_KE(1.0, nu)
        self._B0, self._C0 = _hex8_B_C(nu)

    # reduced-DOF bookkeeping of the reference (core/fea.py:271-279), on demand
    @property
    def used_dofs(self):
        if not hasattr(self, "_used"):
            nodes = np.zeros((self.nx + 1) * (self.ny + 1) * (self.nz + 1), dtype=bool)
            a = self.active_elems.reshape(self.nx, self.ny, self.nz)
            m = np.zeros((self.nx + 1, self.ny + 1, self.nz + 1), dtype=bool)
            for ox in (0, 1):
                for oy in (0, 1):
                    for oz in (0, 1):
                        m[ox:ox + self.nx, oy:oy + self.ny, oz:oz + self.nz] |= a
            nodes[:] = m.transpose(2, 1, 0).ravel()
            self._used = np.nonzero(np.repeat(nodes, 3))[0].astype(np.int64)
        return self._used

    @property
    def ndof(self):
        return int(self.used_dofs.size)

    def close(self):
        if self._mg is not None:
            try:
                self._mg.close()
            except Exception:
                pass
            self._mg = None
        if getattr(self, "_h", None) is not None:
            self._lib.vx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fixed(self, fixed_dofs):
        """Original (full-grid) DOF ids to constrain (core/fea.py:357-369)."""
        fd = np.ascontiguousarray(np.asarray(fixed_dofs, dtype=np.int64).ravel())
        self._fixed_dofs_full = fd
        _native.check(self._lib.vx_set_fixed(self._h, fd.ctypes.data_as(ctypes.c_void_p),
                                             fd.size), self._h)
        free = np.ones(self.ndof_full, dtype=bool)
        free[fd] = False
        self.free = free[self.used_dofs]

    def solve(self, Evec, f, tol=1e-4, max_cg=None, x0=None, progress_cb=None):
        """MG-PCG, Jacobi-PCG on any MG failure (core/fea.py:400-436).
        Returns the full-length u as numpy."""
        if self.free is None:
            raise RuntimeError("call set_fixed() before solve()")
        if self.use_multigrid:
            u = self._solve_multigrid(Evec, f, tol, max_cg, x0, progress_cb)
            if u is not None:
                return u
        return self._solve_jacobi(Evec, f, tol=tol, max_cg=max_cg, x0=x0,
                                  progress_cb=progress_cb)

    def _solve_multigrid(self, Evec, f, tol, max_cg, x0, progress_cb):
        if self._mg is None:
            try:
                self._mg = MGSolver(self.nx, self.ny, self.nz, self._fixed_dofs_full,
                                    nu=self.nu, compute_mode=self._compute_mode,
                                    cpu_threads=self._cpu_threads, verbose=self.verbose,
                                    **self.mg_options)
            except Exception as exc:
                print(f"[VoxelSim FEA] MGSolver construction failed for "
                      f"{self.nx}x{self.ny}x{self.nz}: {exc}\n[VoxelSim FEA] Falling back "
                      f"to plain Jacobi-PCG (no multigrid preconditioning) for this solve.")
                self.use_multigrid = False
                self._mg = None
                return None
        try:
            self._mg.set_density(np.asarray(Evec, dtype=float), active=self.active_elems,
                                 e_min=self.e_min)
            f_dev = _dev(f)
            u_dev = _torch().empty_like(f_dev)
            x0_dev = None if x0 is None else _dev(x0)
            rc = self._mg.solve_device(f_dev, u_dev, tol=tol, max_cg=max_cg, x0_dev=x0_dev,
                                       progress_cb=progress_cb)
            if rc == _native.VX_ERR_NOT_CONVERGED:
                print(f"[VoxelSim FEA] WARNING: MG-PCG did not converge in "
                      f"{max_cg or 1000} iterations (tol={tol:g}, final residual ratio="
                      f"{self._mg.last_resid_ratio:.3e}, plan={self.plan.label}). Returned "
                      f"displacement may be inaccurate.")
            u = u_dev.cpu().numpy()
        except Exception as exc:
            print(f"[VoxelSim FEA] MGSolver.solve() failed for {self.nx}x{self.ny}x{self.nz}: "
                  f"{exc}\n[VoxelSim FEA] Falling back to plain Jacobi-PCG (no multigrid "
                  f"preconditioning) for this solve.")
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

    def _solve_jacobi(self, Evec, f, tol=1e-4, max_cg=None, x0=None, progress_cb=None):
        """Jacobi-PCG on the active-DOF system, on the device (core/fea.py:497-593)."""
        if self.free is None:
            raise RuntimeError("call set_fixed() before solve()")
        E_dev = _dev(np.asarray(Evec, dtype=float))
        f_dev = _dev(f)
        u_dev = _torch().empty_like(f_dev)
        # warm start: x0 on the reduced free DOFs, r = f - K x0 (core/fea.py:529-535)
        x0_dev = None if x0 is None else _dev(x0)
        hist_cap = 200000
        hist = np.zeros(hist_cap)
        info = _native.vx_info()

        def _cb(it, ratio, t, _u):
            if progress_cb is not None:
                progress_cb(int(it), float(ratio), float(t))

        cb = _native.PROGRESS_CB(_cb) if progress_cb is not None else _native.PROGRESS_CB()
        self._lib.vx_set_stream(self._h, ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream))
        rc = self._lib.vx_solve_jacobi(self._h, _native.ptr(E_dev), _native.ptr(self._active_dev),
                                       _native.ptr(f_dev), _native.ptr(u_dev),
                                       _native.ptr(x0_dev), float(tol), int(max_cg or 0), cb, None,
                                       hist.ctypes.data_as(ctypes.c_void_p), hist_cap,
                                       ctypes.byref(info))
        _native.check(rc, self._h, allow=(_native.VX_ERR_NOT_CONVERGED,))
        self.last_iters = info.iters
        self.last_converged = bool(info.converged)
        self.last_resid_ratio = info.resid_ratio
        self.resid_history = [float(v) for v in hist[:min(info.iters, hist_cap)]]
        if not self.last_converged:
            print(f"[VoxelSim FEA] WARNING: CG did not converge (tol={tol:g}, final residual "
                  f"ratio={self.last_resid_ratio:.3e}). Returned displacement may be inaccurate.")
        return u_dev.cpu().numpy()

    def element_strain_energy(self, u_full):
        """u_e^T KE u_e per element, 0 outside the active set (core/fea.py:595-602)."""
        ud = _dev(u_full)
        ce = _torch().empty(self.nelem, dtype=_torch().float64, device="cuda")
        _native.check(self._lib.vx_strain_energy(self._h, _native.ptr(ud),
                                                 _native.ptr(self._active_dev),
                                                 _native.ptr(ce)), self._h)
        return ce.cpu().numpy()

    def element_von_mises_stress(self, u_full, Evec):
        """Centroid von Mises scaled by per-element E (core/fea.py:604-626)."""
        ud = _dev(u_full)
        Ed = _dev(np.asarray(Evec, dtype=float))
        vm = _torch().empty(self.nelem, dtype=_torch().float64, device="cuda")
        _native.check(self._lib.vx_von_mises(self._h, _native.ptr(ud), _native.ptr(Ed),
                                             _native.ptr(self._active_dev), _native.ptr(vm)),
                      self._h)
        return vm.cpu().numpy()
