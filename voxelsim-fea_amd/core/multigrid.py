"""Geometric-multigrid-preconditioned CG on the GPU (API of
core/multigrid.py:76-583).

``MGSolver`` keeps the reference's constructor and ``solve`` contract --
full regular grid, voids carried by the caller at ``E*e_min``, damped-Jacobi
V(n_smooth, n_smooth) cycle with a 30-sweep coarsest level, Fletcher-Reeves
PCG with the stall guard, numpy arrays in and out -- and runs every operation
in the HIP library (``include/vxfea.h``).  Extra keyword options select the
MI355X forms: ``precision="mixed"`` (FP32 V-cycle and matvec inside FP64
iterative refinement).

``solve_device`` is the device-resident form: torch tensors in HBM in and
out, no host copies (what the benchmark times).
"""

import ctypes

import numpy as np

from . import _native
from . import backend
from . import compute_plan

_PRECISION = {"fp64": _native.VX_FP64, "mixed": _native.VX_MIXED}
_SMOOTHER = {"jacobi": _native.VX_SMOOTH_JACOBI, "chebyshev": _native.VX_SMOOTH_CHEBYSHEV}


# The two solver configurations the benchmark measures (MGSolver keyword
# arguments).  VARIABLE_CYCLE: mixed precision, Chebyshev smoothing on the
# active-DOF system, a 5-level V-cycle with more smoothing steps on the
# latency-bound coarse levels (2, 2, 6, 12 on levels 0-3, 48 sweeps on the
# coarsest), each k-step pass over [0.05 (2/k)^2, 1] lambda_max, inner FP32
# solves to 3e-5.  PLAIN_CYCLE: the same with two steps on every level of a
# 7-level hierarchy, for parts whose plain cycle converges in a few tens of
# iterations (there the extra coarse-level steps cost more than they save).
VARIABLE_CYCLE = dict(precision="mixed", smoother="chebyshev", cheb_ratio=0.05,
                      system="active", autotune=True, inner_tol=3e-5, max_levels=5,
                      level_smooth=[2, 2, 6, 12], coarse_sweeps=48, cheb_ratio_power=2)
PLAIN_CYCLE = dict(precision="mixed", smoother="chebyshev", cheb_ratio=0.05, system="active",
                   autotune=True, inner_tol=1e-5, max_levels=7)


def _torch():
    import torch
    return torch


def _dev(a, dtype=None):
    """numpy / torch -> contiguous CUDA tensor (float64 unless dtype)."""
    torch = _torch()
    dt = torch.float64 if dtype is None else dtype
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=dt).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda").to(dt).contiguous()


def padded_dims(dims, max_levels=6, overhead=1.3):
    """Element dims grown to multiples of 2^k (largest k < max_levels whose
    volume stays within ``overhead`` times the original), so that the
    hierarchy rule (all dims even, core/multigrid.py:144-148) can coarsen k
    times; the dims themselves when no such k exists."""
    vol = float(np.prod(dims))
    for k in range(max_levels - 1, 0, -1):
        m = 2 ** k
        p = tuple(-(-int(d) // m) * m for d in dims)
        if np.prod(p) <= overhead * vol:
            return p
    return tuple(int(d) for d in dims)


# A part grown with void elements to coarsenable dims (padded_dims): its node
# and element arrays embedded at the low corner of the padded grid.
def pad_dof_ids(ids, dims, pdims):
    """DOF ids of a dims grid -> the same DOFs of the padded pdims grid."""
    ids = np.asarray(ids, dtype=np.int64)
    if tuple(dims) == tuple(pdims):
        return ids
    n, c = ids // 3, ids % 3
    lx, ly = dims[0] + 1, dims[1] + 1
    ix, iy, iz = n % lx, (n // lx) % ly, n // (lx * ly)
    return 3 * (ix + (pdims[0] + 1) * (iy + (pdims[1] + 1) * iz)) + c


def pad_elems(t, dims, pdims):
    """Per-element device tensor (reference element order) -> padded grid, 0 outside."""
    if tuple(dims) == tuple(pdims):
        return t
    out = _torch().zeros(tuple(pdims), dtype=t.dtype, device=t.device)
    out[:dims[0], :dims[1], :dims[2]] = t.view(*dims)
    return out.view(-1)


def _node_view(v, dims):
    nx, ny, nz = dims
    return v.view(-1, nz + 1, ny + 1, nx + 1, 3)


def pad_nodes(v, dims, pdims):
    """[..., ndof of dims] device tensor -> [..., ndof of pdims] (zeros outside)."""
    if tuple(dims) == tuple(pdims):
        return v
    nx, ny, nz = dims
    px, py, pz = pdims
    out = _torch().zeros(v.shape[:-1] + (3 * (px + 1) * (py + 1) * (pz + 1),), dtype=v.dtype,
                         device=v.device)
    _node_view(out, pdims)[:, :nz + 1, :ny + 1, :nx + 1] = _node_view(v.contiguous(), dims)
    return out


def unpad_nodes(vp, dims, pdims, out):
    """The part's nodes of a padded-grid vector vp into out."""
    nx, ny, nz = dims
    out.view(-1, nz + 1, ny + 1, nx + 1, 3).copy_(
        _node_view(vp, pdims)[:, :nz + 1, :ny + 1, :nx + 1])


def _require_gpu(plan, what):
    if not plan.on_gpu:
        raise RuntimeError(
            f"{what}: no usable GPU ({backend.gpu_status()}); this build solves on "
            f"MI355X only (compute plan: {plan.label})")


def _level_smooth(steps):
    """Per-level smoothing steps (level 0 first; levels past the list and
    entries of 0 use n_smooth) as the option's 16 bytes."""
    if steps is None:
        return None
    steps = [int(k) for k in steps]
    if len(steps) > 16 or any(k < 0 or k > 255 for k in steps):
        raise ValueError("level_smooth: at most 16 step counts, each 0-255")
    return (ctypes.c_uint8 * 16)(*steps)


class MGSolver:
    """Multigrid-preconditioned CG on the full voxel grid (HIP).

    ``level_smooth`` gives the smoothing steps per level (pre and post, level 0
    first); the reference uses ``n_smooth`` on every level.  With
    ``cheb_ratio_power=p`` a Chebyshev pass of k steps uses the interval
    [cheb_ratio * (n_smooth / k)**p, 1] * lambda_max.  ``graphs=False``
    launches each PCG step's V-cycle kernel by kernel instead of replaying it
    from a captured HIP graph (same kernels, same results).  ``block_steps``
    sets how many FP32 smoothing steps one launch of the LDS-blocked smoother
    does on the small levels below level 0 (0: 3, 1: one launch per step) and
    ``block_edge`` its owned box edge in nodes (0: from the level size).
    ``keep_direction=True`` (mixed precision) lets each FP32 inner solve of the
    iterative refinement continue the previous one's search direction instead
    of restarting CG; ``keep_drift`` > 0 restarts it all the same where the
    refreshed residual moved by more than that relative amount."""

    def __init__(self, nx, ny, nz, fixed_dofs, nu=0.3, xp=None, n_smooth=2, omega=0.6,
                 min_elems=4, max_levels=6, compute_mode="AUTO", cpu_threads=0,
                 verbose=False, precision="fp64", coarse_sweeps=30, inner_tol=1e-5,
                 max_outer=20, z_chunk=0, smoother="jacobi", cheb_ratio=0.3,
                 system="full", e_coarsen="mean", max_batch=1, flexible=False, pad=False,
                 autotune=False, coarse="sweeps", level_smooth=None, graphs=True,
                 cheb_ratio_power=0, block_steps=0, block_edge=0, keep_direction=False, keep_drift=0.0):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.nelem = self.nx * self.ny * self.nz
        self.ndof = 3 * (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self.n_smooth, self.omega, self.verbose = n_smooth, omega, verbose
        self.precision = precision
        if system not in ("full", "active"):
            raise ValueError("system must be 'full' (voids at E*e_min) or 'active'")
        self.system = system
        # pad: solve on a grid grown with void elements to coarsenable dims.
        # Only the active-DOF system is unchanged by that (void elements carry
        # nothing and their DOFs leave the system).
        if pad and system != "active":
            raise ValueError("pad=True needs system='active'")
        self._pdims = padded_dims((self.nx, self.ny, self.nz), max_levels) if pad else \
            (self.nx, self.ny, self.nz)
        self.padded = self._pdims != (self.nx, self.ny, self.nz)
        self.plan = compute_plan.choose(self.ndof, mode=compute_mode, cpu_threads=cpu_threads,
                                        verbose=verbose)
        _require_gpu(self.plan, "MGSolver")
        self.xp = self.plan.xp
        self._lib = _native.load()
        opt = _native.default_options(
            max_levels=max_levels, min_elems=min_elems, n_smooth=n_smooth, omega=omega,
            coarse_sweeps=coarse_sweeps, precision=_PRECISION[precision],
            inner_tol=inner_tol, max_outer=max_outer, z_chunk=z_chunk,
            smoother=_SMOOTHER[smoother], cheb_ratio=cheb_ratio,
            e_coarsen={"mean": 0, "geometric": 1, "harmonic": 2}[e_coarsen],
            max_batch=max_batch, flexible=1 if flexible else 0,
            coarse_mode={"sweeps": _native.VX_COARSE_SWEEPS,
                         "direct": _native.VX_COARSE_DIRECT}[coarse],
            level_smooth=_level_smooth(level_smooth), eager=0 if graphs else 1,
            cheb_ratio_power=int(cheb_ratio_power), block_steps=int(block_steps),
            block_edge=int(block_edge), keep_direction=1 if keep_direction else 0,
            keep_drift=float(keep_drift))
        self.max_batch = max_batch
        self._coarse_sweeps = int(coarse_sweeps)
        self._steps = [int(k) for k in (level_smooth or [])]
        h = ctypes.c_void_p()
        px, py, pz = self._pdims
        _native.check(self._lib.vx_create(px, py, pz, float(nu), ctypes.byref(opt),
                                          ctypes.byref(h)))
        self._h = h
        self._stream_bind()
        self.levels = []
        for li in range(self._lib.vx_num_levels(h)):
            a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
            self._lib.vx_level_dims(h, li, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
            d = (a.value, b.value, c.value)
            self.levels.append({"dims": d, "ndof": 3 * (d[0] + 1) * (d[1] + 1) * (d[2] + 1),
                                "g": float(2 ** li)})
        fd = np.ascontiguousarray(self._pad_dof_ids(np.asarray(fixed_dofs, dtype=np.int64).ravel()))
        _native.check(self._lib.vx_set_fixed(h, fd.ctypes.data_as(ctypes.c_void_p),
                                             fd.size), h)
        self.last_iters = 0
        self.last_converged = False
        self.last_resid_ratio = float("nan")
        self.last_true_resid_ratio = float("nan")
        self.last_outer_iters = 0
        self.last_solve_ms = float("nan")
        self.resid_history = []
        self._autotune = bool(autotune)
        self.tuned = None          # {level: (operator zc, smoother zc)} once tuned
        self.autotune_s = 0.0

    # -- padding ------------------------------------------------------------
    def _pad_dof_ids(self, ids):
        return pad_dof_ids(ids, (self.nx, self.ny, self.nz), self._pdims)

    def _pad_elems(self, a, dtype):
        return pad_elems(_dev(a, dtype), (self.nx, self.ny, self.nz), self._pdims)

    def _pad_nodes(self, v):
        return pad_nodes(v, (self.nx, self.ny, self.nz), self._pdims)

    def _unpad_nodes(self, vp, out):
        if self.padded:
            unpad_nodes(vp, (self.nx, self.ny, self.nz), self._pdims, out)

    # -- plumbing ----------------------------------------------------------
    def _stream_bind(self):
        torch = _torch()
        self._lib.vx_set_stream(self._h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.vx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, allow=()):
        return _native.check(rc, self._h, allow)

    def _unpadded_only(self, what):
        if self.padded:
            raise RuntimeError(f"{what} works on the solver's own grid: not on a padded solver")

    # -- density -----------------------------------------------------------
    def set_density(self, Evec_full, active=None, e_min=1e-9):
        """Per-level E, 2x2x2 coarsening (core/multigrid.py:281-315).  With
        ``active`` (bool per element) the void rule E*e_min of
        core/fea.py:467-469 is applied on the device; with system="active"
        voids carry no stiffness and DOFs of no active element leave the
        system (the system of the reference's Jacobi-PCG path,
        core/fea.py:497-593)."""
        torch = _torch()
        if self.padded:
            self._E_dev = self._pad_elems(Evec_full, torch.float64)
            act = np.ones(self.nelem, np.uint8) if active is None else active
            self._active_dev = self._pad_elems(act, torch.uint8)
        else:
            self._E_dev = _dev(Evec_full)
            self._active_dev = None if active is None else _dev(active, torch.uint8)
        self._stream_bind()
        em = -1.0 if self.system == "active" else float(e_min)
        self._check(self._lib.vx_set_density(self._h, _native.ptr(self._E_dev),
                                              _native.ptr(self._active_dev), em))
        if self._autotune and self.tuned is None:
            # once per solver: the kernels' timing does not depend on the values
            import time
            t0 = time.perf_counter()
            self.tuned = self.autotune()
            self.autotune_s = time.perf_counter() - t0

    # -- solve -------------------------------------------------------------
    def solve_device(self, f_dev, u_dev, tol=1e-4, max_cg=None, x0_dev=None,
                     progress_cb=None, hist_cap=100000):
        """Device-resident solve on the current density; returns the native
        status (0 ok, VX_ERR_NOT_CONVERGED) and fills the last_* fields.
        Raises ``_native.StallError`` when the stall guard fires."""
        self._stream_bind()
        hist = np.zeros(hist_cap)
        info = _native.vx_info()
        user_cb = progress_cb

        def _cb(it, ratio, t, _user):
            if user_cb is not None:
                user_cb(int(it), float(ratio), float(t))

        cb = _native.PROGRESS_CB(_cb) if progress_cb is not None else _native.PROGRESS_CB()
        fp, xp0 = self._pad_nodes(f_dev), None if x0_dev is None else self._pad_nodes(x0_dev)
        up = _torch().empty_like(fp) if self.padded else u_dev
        rc = self._lib.vx_solve(self._h, _native.ptr(fp), _native.ptr(up),
                                _native.ptr(xp0), float(tol), int(max_cg or 0), cb, None,
                                hist.ctypes.data_as(ctypes.c_void_p), hist_cap,
                                ctypes.byref(info))
        self._unpad_nodes(up, u_dev)
        self.last_iters = info.iters
        self.last_converged = bool(info.converged)
        self.last_resid_ratio = info.resid_ratio
        self.last_true_resid_ratio = info.true_resid_ratio
        self.last_outer_iters = info.outer_iters
        self.last_solve_ms = info.solve_ms
        self.last_corrector = {0: "none", 1: "used", 2: "fell back"}[info.corrector]
        self.last_coarse_direct = bool(info.coarse_direct)
        self.resid_history = [float(v) for v in hist[:min(info.iters, hist_cap)]]
        self._check(rc, allow=(_native.VX_ERR_NOT_CONVERGED,))
        return rc

    def solve_batch_device(self, F_dev, U_dev, tol=1e-4, max_cg=None, progress_cb=None,
                           hist_cap=2000):
        """All load cases of F_dev [nrhs, ndof] (nrhs <= max_batch) in one
        batched solve on the current density; U_dev [nrhs, ndof] receives the
        displacements.  Returns the native status; per-case statistics go to
        ``last_batch`` (list of dicts) and ``batch_history``."""
        nrhs = int(F_dev.shape[0])
        self._stream_bind()
        hist = np.zeros(nrhs * hist_cap)
        infos = (_native.vx_info * nrhs)()

        def _cb(it, ratio, t, _user):
            progress_cb(int(it), float(ratio), float(t))

        cb = _native.PROGRESS_CB(_cb) if progress_cb is not None else _native.PROGRESS_CB()
        Fp = self._pad_nodes(F_dev)
        Up = _torch().empty_like(Fp) if self.padded else U_dev
        rc = self._lib.vx_solve_batch(self._h, nrhs, _native.ptr(Fp), _native.ptr(Up),
                                      float(tol), int(max_cg or 0), cb, None,
                                      hist.ctypes.data_as(ctypes.c_void_p), hist_cap, infos)
        self._unpad_nodes(Up, U_dev)
        self.last_batch = [{"iters": i.iters, "converged": bool(i.converged),
                            "outer_iters": i.outer_iters, "resid_ratio": i.resid_ratio,
                            "true_resid_ratio": i.true_resid_ratio,
                            "corrector": i.corrector} for i in infos]
        self.batch_history = [list(hist[q * hist_cap:q * hist_cap + min(infos[q].iters, hist_cap)])
                              for q in range(nrhs)]
        self.last_solve_ms = infos[0].solve_ms if nrhs else float("nan")
        self._check(rc, allow=(_native.VX_ERR_NOT_CONVERGED,))
        return rc

    def solve_jacobi_device(self, E_dev, active_dev, f_dev, u_dev, tol=1e-4, max_cg=None,
                            progress_cb=None, hist_cap=100000, x0_dev=None):
        """Jacobi-PCG on the active-DOF system on the device
        (core/fea.py:497-593), warm-started from ``x0_dev`` when given; same
        fields as solve_device."""
        self._unpadded_only('solve_jacobi_device')
        self._stream_bind()
        hist = np.zeros(hist_cap)
        info = _native.vx_info()

        def _cb(it, ratio, t, _user):
            progress_cb(int(it), float(ratio), float(t))

        cb = _native.PROGRESS_CB(_cb) if progress_cb is not None else _native.PROGRESS_CB()
        rc = self._lib.vx_solve_jacobi(self._h, _native.ptr(E_dev), _native.ptr(active_dev),
                                       _native.ptr(f_dev), _native.ptr(u_dev),
                                       _native.ptr(x0_dev), float(tol), int(max_cg or 0), cb, None,
                                       hist.ctypes.data_as(ctypes.c_void_p), hist_cap,
                                       ctypes.byref(info))
        self.last_iters = info.iters
        self.last_converged = bool(info.converged)
        self.last_resid_ratio = info.resid_ratio
        self.last_true_resid_ratio = info.true_resid_ratio
        self.last_solve_ms = info.solve_ms
        self.resid_history = [float(v) for v in hist[:min(info.iters, hist_cap)]]
        self._check(rc, allow=(_native.VX_ERR_NOT_CONVERGED,))
        return rc

    def solve(self, Evec_full, f, tol=1e-4, max_cg=None, x0=None, progress_cb=None):
        """core/multigrid.py:439-583; returns numpy u (masked)."""
        self.set_density(Evec_full)
        f_dev = _dev(f)
        u_dev = _torch().empty_like(f_dev)
        x0_dev = None if x0 is None else _dev(x0)
        rc = self.solve_device(f_dev, u_dev, tol=tol, max_cg=max_cg, x0_dev=x0_dev,
                               progress_cb=progress_cb)
        if rc == _native.VX_ERR_NOT_CONVERGED:
            print(f"[VoxelSim FEA] WARNING: MG-PCG did not converge in "
                  f"{max_cg or 1000} iterations (tol={tol:g}, final residual ratio="
                  f"{self.last_resid_ratio:.3e}, plan={self.plan.label}). Returned "
                  f"displacement may be inaccurate.")
        if self.verbose:
            print(f"[VoxelSim FEA] MG solve: {self.last_iters} CG iters "
                  f"(converged={self.last_converged}), {self.last_solve_ms:.1f} ms on device")
        return u_dev.cpu().numpy()

    # -- autotuner ------------------------------------------------------------
    def autotune(self, precision=None, reps=5):
        """Pick each level's tile z-chunk by timing (needs the density set);
        returns {level: (operator zc, smoother zc, smoother-with-dot zc,
        two-Chebyshev-steps-from-zero zc, residual-restriction zc)}."""
        prec = _PRECISION[precision or self.precision]
        self._stream_bind()
        self._check(self._lib.vx_autotune(self._h, prec, int(reps)))
        return {li: tuple(self._lib.vx_tuned_zchunk(self._h, li, prec, c) for c in range(5))
                for li in range(len(self.levels))}

    def _level_steps(self, li):
        return self._steps[li] if li < len(self._steps) and self._steps[li] else self.n_smooth

    def block_plans(self):
        """{level: (box edge, steps per launch)} of the FP32 smoothing passes
        that run blocked (levels below 0 whose pass would otherwise be one
        tile launch per step); levels with tile launches are left out."""
        out = {}
        for li in range(1, len(self.levels)):
            k = self._coarse_sweeps if li == len(self.levels) - 1 else self._level_steps(li)
            e, s = ctypes.c_int32(), ctypes.c_int32()
            self._check(self._lib.vx_block_plan(self._h, li, int(k), ctypes.byref(e),
                                                ctypes.byref(s)))
            if e.value > 0:
                out[li] = (e.value, s.value)
        return out

    # -- learned coarse-grid corrector --------------------------------------
    def level_nodes(self, li):
        a, b, c = self.levels[li]["dims"]
        return (a + 1) * (b + 1) * (c + 1)

    def attach_corrector(self, net, level, gain=0.0, enabled=True, iter_budget=0):
        """Use ``net`` (core.learned.CorrectorNet) in the coarse-grid
        correction at ``level`` of the FP32 V-cycle (precision="mixed");
        ``iter_budget`` > 0 falls back to the plain V-cycle after that many
        inner iterations without convergence."""
        self._stream_bind()
        self._check(self._lib.vx_set_corrector(self._h, net.handle, int(level), float(gain),
                                                1 if enabled else 0, int(iter_budget)))
        self._corrector = net

    def detach_corrector(self):
        self._check(self._lib.vx_set_corrector(self._h, None, 0, 0.0, 0, 0))
        self._corrector = None

    def corrector_eval(self, T, V):
        """Per sample: energy-error drop (/ v.t) of the best combination of
        the plain correction and the attached network's, over the plain one."""
        out = np.zeros(int(T.shape[0]))
        self._stream_bind()
        self._check(self._lib.vx_corrector_eval(self._h, int(T.shape[0]), _native.ptr(T),
                                                 _native.ptr(V),
                                                 out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def capture(self, level, cap, every=1):
        """Record the level residuals of the next solves; returns the buffer
        [cap, 3 * nodes] (valid rows: ``capture_count()``)."""
        torch = _torch()
        buf = torch.zeros(int(cap), 3 * self.level_nodes(level), dtype=torch.float32,
                          device="cuda")
        self._check(self._lib.vx_capture(self._h, int(level), _native.ptr(buf), int(cap),
                                         int(every)))
        self._capture_buf = buf
        return buf

    def capture_count(self):
        return int(self._lib.vx_capture_count(self._h))

    def stop_capture(self):
        self._check(self._lib.vx_capture(self._h, 0, None, 0, 1))

    def corrector_pairs(self, T):
        """Plain coarse-grid corrections V, residuals after them R = T - A V
        and v.t for the captured level residuals T (FP32 V-cycle)."""
        torch = _torch()
        V = torch.empty_like(T)
        R = torch.empty_like(T)
        vt = torch.empty(int(T.shape[0]), dtype=torch.float64, device="cuda")
        self._stream_bind()
        self._check(self._lib.vx_corrector_prepare(self._h, int(T.shape[0]), _native.ptr(T),
                                                    _native.ptr(V), _native.ptr(R),
                                                    _native.ptr(vt)))
        return V, R, vt

    def corrector_train(self, T, V, R, vt, batch=4, steps=200, lr=1e-3, seed=0):
        """Adam training of the attached corrector; returns the loss per step."""
        loss = np.zeros(int(steps), dtype=np.float32)
        self._stream_bind()
        self._check(self._lib.vx_corrector_train(self._h, int(T.shape[0]), _native.ptr(T),
                                                  _native.ptr(V), _native.ptr(R),
                                                  _native.ptr(vt), int(batch), int(steps),
                                                  float(lr), int(seed),
                                                  loss.ctypes.data_as(ctypes.c_void_p)))
        return loss

    # -- single operators (parity tests) -----------------------------------
    def _level_vec(self, li, a):
        return _dev(a)

    def _apply(self, li, u, precision="fp64"):
        """Masked K u at level li (core/multigrid.py:318-341)."""
        self._unpadded_only('_apply')
        ud = _dev(u)
        yd = _torch().empty_like(ud)
        self._stream_bind()
        self._check(self._lib.vx_matvec(self._h, li, _PRECISION[precision], _native.ptr(ud),
                                        _native.ptr(yd)))
        return yd.cpu().numpy()

    def _smooth(self, li, u, b, iters, precision="fp64"):
        """``iters`` damped-Jacobi sweeps (core/multigrid.py:367-402), or a
        Chebyshev step of degree ``iters`` with smoother="chebyshev"; u=None
        starts from zero."""
        self._unpadded_only('_smooth')
        bd = _dev(b)
        ud = None if u is None else _dev(u)
        od = _torch().empty_like(bd)
        self._stream_bind()
        self._check(self._lib.vx_smooth(self._h, li, _PRECISION[precision], _native.ptr(ud),
                                        _native.ptr(bd), int(iters), _native.ptr(od)))
        return od.cpu().numpy()

    def _restrict(self, li, r_fine):
        self._unpadded_only('_restrict')
        rf = _dev(r_fine)
        rc = _torch().empty(self.levels[li + 1]["ndof"], dtype=_torch().float64, device="cuda")
        self._stream_bind()
        self._check(self._lib.vx_restrict(self._h, li, _native.ptr(rf), _native.ptr(rc)))
        return rc.cpu().numpy()

    def _prolongate(self, li, e_coarse):
        self._unpadded_only('_prolongate')
        ec = _dev(e_coarse)
        ef = _torch().empty(self.levels[li]["ndof"], dtype=_torch().float64, device="cuda")
        self._stream_bind()
        self._check(self._lib.vx_prolong(self._h, li, _native.ptr(ec), _native.ptr(ef)))
        return ef.cpu().numpy()

    def _vcycle(self, li, b, precision="fp64"):
        self._unpadded_only('_vcycle')
        if li != 0:
            raise ValueError("the V-cycle entry point starts at level 0")
        bd = _dev(b)
        zd = _torch().empty_like(bd)
        self._stream_bind()
        self._check(self._lib.vx_vcycle(self._h, _PRECISION[precision], _native.ptr(bd),
                                        _native.ptr(zd)))
        return zd.cpu().numpy()

    def inv_diag(self, li):
        self._unpadded_only('inv_diag')
        md = _torch().empty(self.levels[li]["ndof"], dtype=_torch().float64, device="cuda")
        self._stream_bind()
        self._check(self._lib.vx_inv_diag(self._h, li, _native.ptr(md)))
        return md.cpu().numpy()

    def time_kernel(self, kind="matvec", level=0, precision="fp64", reps=20):
        """Average device time (ms) of one level operator launch."""
        self._unpadded_only('time_kernel')
        out = ctypes.c_double()
        self._stream_bind()
        self._check(self._lib.vx_time_kernel(self._h, {"matvec": 0, "jacobi": 1, "chebyshev": 2, "chebyshev2_zero": 3,
                                              "residual_restrict": 4, "chebyshev_dot": 5,
                                              "jacobi_dot": 6}[kind], level,
                                             _PRECISION[precision], reps, ctypes.byref(out)))
        return out.value


class PartsSolver:
    """Many parts of uneven size solved together (``vx_solve_parts``).

    The reference solves each part with its own ``VoxelFEA`` / ``MGSolver``
    (core/fea.py:244-305, core/multigrid.py:439-583).  Here every part keeps
    its own grid, hierarchy, fixed DOFs, density and PCG recurrence -- it stops
    at its own tolerance with its own iteration count -- while each operator,
    smoother, transfer and reduction launch covers all the parts of a level.

    ``parts``: a list of (nx, ny, nz); ``fixed``: a list of each part's fixed
    DOF ids.  Options as ``MGSolver`` (``pad=True`` grows each part with void
    elements to coarsenable dims, which needs ``system="active"``; at most
    ``pad_overhead`` times a part's volume).  The tile kernels' z-chunks follow
    one rule over the whole group (``z_chunk`` fixes them); ``autotune=True``
    times candidate z-chunks level by level after the first ``set_density``
    (``zchunks()`` reports them)."""

    def __init__(self, parts, fixed, nu=0.3, n_smooth=2, omega=0.6, min_elems=4, max_levels=6,
                 compute_mode="AUTO", cpu_threads=0, verbose=False, precision="fp64",
                 coarse_sweeps=30, inner_tol=1e-5, max_outer=20, z_chunk=0, smoother="jacobi",
                 cheb_ratio=0.3, system="full", e_coarsen="mean", pad=False, autotune=False,
                 level_smooth=None, graphs=True, cheb_ratio_power=0, pad_overhead=1.3):
        self.dims = [tuple(int(v) for v in d) for d in parts]
        if len(fixed) != len(self.dims):
            raise ValueError("one fixed-DOF list per part")
        if system not in ("full", "active"):
            raise ValueError("system must be 'full' (voids at E*e_min) or 'active'")
        if pad and system != "active":
            raise ValueError("pad=True needs system='active'")
        self.system, self.precision = system, precision
        self.pdims = [padded_dims(d, max_levels, pad_overhead) if pad else d for d in self.dims]
        ndof = sum(3 * (d[0] + 1) * (d[1] + 1) * (d[2] + 1) for d in self.pdims)
        self.plan = compute_plan.choose(ndof, mode=compute_mode, cpu_threads=cpu_threads,
                                        verbose=verbose)
        _require_gpu(self.plan, "PartsSolver")
        self._lib = _native.load()
        opt = _native.default_options(
            max_levels=max_levels, min_elems=min_elems, n_smooth=n_smooth, omega=omega,
            coarse_sweeps=coarse_sweeps, precision=_PRECISION[precision], inner_tol=inner_tol,
            max_outer=max_outer, z_chunk=z_chunk, smoother=_SMOOTHER[smoother],
            cheb_ratio=cheb_ratio, e_coarsen={"mean": 0, "geometric": 1, "harmonic": 2}[e_coarsen],
            level_smooth=_level_smooth(level_smooth), cheb_ratio_power=int(cheb_ratio_power),
            eager=0 if graphs else 1)
        dims = np.ascontiguousarray(np.asarray(self.pdims, dtype=np.int32).reshape(-1))
        h = ctypes.c_void_p()
        rc = self._lib.vx_parts_create(len(self.pdims), dims.ctypes.data_as(ctypes.c_void_p),
                                       float(nu), ctypes.byref(opt), ctypes.byref(h))
        _native.check(rc)
        self._h = h
        self._stream_bind()
        self.levels = [int(self._lib.vx_parts_num_levels(h, i)) for i in range(len(self.dims))]
        ids = [pad_dof_ids(np.asarray(f, dtype=np.int64).ravel(), d, p)
               for f, d, p in zip(fixed, self.dims, self.pdims)]
        counts = np.asarray([i.size for i in ids], dtype=np.int64)
        flat = np.ascontiguousarray(np.concatenate(ids) if ids else np.zeros(0, np.int64))
        self._check(self._lib.vx_parts_set_fixed(h, flat.ctypes.data_as(ctypes.c_void_p),
                                                 counts.ctypes.data_as(ctypes.c_void_p)))
        self.last_parts = []
        self.last_solve_ms = float("nan")
        self._autotune = bool(autotune)
        self.tuned = False
        self.autotune_s = 0.0

    def _stream_bind(self):
        torch = _torch()
        self._lib.vx_parts_set_stream(self._h,
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))

    def _check(self, rc, allow=()):
        if rc == _native.VX_OK or rc in allow:
            return rc
        msg = self._lib.vx_parts_last_error(self._h).decode()
        if rc == _native.VX_ERR_STALL:
            raise _native.StallError(rc, msg)
        raise _native.NativeError(rc, msg)

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.vx_parts_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_density(self, E_list, active_list=None, e_min=1e-9):
        """Per-part E (E*vsize per element) and active masks, packed on the
        device; with system="active" voids carry no stiffness."""
        torch = _torch()
        act = active_list or [None] * len(self.dims)
        E = [pad_elems(_dev(e), d, p) for e, d, p in zip(E_list, self.dims, self.pdims)]
        A = [pad_elems(_dev(np.ones(int(np.prod(d)), np.uint8) if a is None else a, torch.uint8),
                       d, p) for a, d, p in zip(act, self.dims, self.pdims)]
        self._E = torch.cat(E)
        self._A = torch.cat(A)
        self._stream_bind()
        em = -1.0 if self.system == "active" else float(e_min)
        self._check(self._lib.vx_parts_set_density(self._h, _native.ptr(self._E),
                                                   _native.ptr(self._A), em))
        if self._autotune and not self.tuned:
            import time
            t0 = time.perf_counter()
            self.autotune()
            self.autotune_s = time.perf_counter() - t0

    def autotune(self, reps=5):
        """Time the tile kernels' z-chunk candidates level by level (needs
        set_density) and keep the fastest; drops captured graphs."""
        self._stream_bind()
        self._check(self._lib.vx_parts_autotune(self._h, int(reps)))
        self.tuned = True

    def zchunks(self):
        """{part: [z-chunk per level]} of the tile kernels."""
        return {i: [int(self._lib.vx_parts_zchunk(self._h, i, l)) for l in range(n)]
                for i, n in enumerate(self.levels)}

    def solve_device(self, F_list, tol=1e-4, max_cg=None, progress_cb=None, hist_cap=2000):
        """Solve every part; F_list: device float64 vectors, one per part.
        Returns the list of u (device tensors, each part's own dims) and the
        native status; per-part statistics go to ``last_parts``."""
        torch = _torch()
        n = len(self.dims)
        F = torch.cat([pad_nodes(_dev(f), d, p) for f, d, p in zip(F_list, self.dims, self.pdims)])
        U = torch.empty_like(F)
        hist = np.zeros(n * hist_cap)
        infos = (_native.vx_info * n)()

        def _cb(it, ratio, t, _user):
            progress_cb(int(it), float(ratio), float(t))

        cb = _native.PROGRESS_CB(_cb) if progress_cb is not None else _native.PROGRESS_CB()
        self._stream_bind()
        rc = self._lib.vx_solve_parts(self._h, _native.ptr(F), _native.ptr(U), float(tol),
                                      int(max_cg or 0), cb, None,
                                      hist.ctypes.data_as(ctypes.c_void_p), hist_cap, infos)
        out, off = [], 0
        for d, p in zip(self.dims, self.pdims):
            m = 3 * (p[0] + 1) * (p[1] + 1) * (p[2] + 1)
            up = U[off:off + m]
            if tuple(d) == tuple(p):
                out.append(up)
            else:
                u = torch.empty(3 * (d[0] + 1) * (d[1] + 1) * (d[2] + 1), dtype=U.dtype,
                                device=U.device)
                unpad_nodes(up, d, p, u)
                out.append(u)
            off += m
        self.last_parts = [{"iters": i.iters, "converged": bool(i.converged),
                            "outer_iters": i.outer_iters, "resid_ratio": i.resid_ratio,
                            "true_resid_ratio": i.true_resid_ratio, "levels": i.n_levels}
                           for i in infos]
        self.parts_history = [list(hist[q * hist_cap:q * hist_cap + min(infos[q].iters, hist_cap)])
                              for q in range(n)]
        self.last_solve_ms = infos[0].solve_ms
        self._check(rc, allow=(_native.VX_ERR_NOT_CONVERGED,))
        return out, rc
