/*
 * vxfea.h -- C ABI of the MI355X (gfx950) matrix-free hex8 MG-PCG solver.
 *
 * Plain C types only: sizes are int64_t, arrays are raw pointers.  Unless a
 * comment says "host", every array argument is a DEVICE pointer that the
 * caller owns (the Python side passes torch-ROCm allocations).  The library
 * owns only the workspace inside the opaque handle.  All work is enqueued on
 * the handle's stream (vx_set_stream, default: the null stream) and calls
 * that return results to the host synchronise that stream.
 *
 * Conventions are the reference solver's (core/fea.py:129-148):
 *   element e = iz + nz*(iy + ny*ix)   (x slowest, z fastest)
 *   node    n = ix + (nx+1)*(iy + (ny+1)*iz)   (x fastest)
 *   DOF       3*n + c
 *
 * Every call returns VX_OK (0) or one of the negative status codes below;
 * vx_last_error(h) gives a message for the last failure on a handle.
 */
#ifndef VXFEA_H
#define VXFEA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VX_OK = 0,
  VX_ERR_ARG = -1,            /* bad argument / call order */
  VX_ERR_NOMEM = -2,          /* device allocation failed */
  VX_ERR_HIP = -3,            /* HIP runtime error */
  VX_ERR_STALL = -4,          /* MG-PCG stall guard fired (core/multigrid.py:524-531) */
  VX_ERR_NOT_CONVERGED = -5,  /* max_iter reached above tol (a warning: u is valid) */
  VX_ERR_NO_DEVICE = -6
};

/* Precision of the operator / V-cycle work. */
enum { VX_FP64 = 0, VX_MIXED = 1 };
/* Smoother inside the V-cycle. */
enum { VX_SMOOTH_JACOBI = 0, VX_SMOOTH_CHEBYSHEV = 1 };
/* Coarsest-level treatment: coarse_sweeps smoothing steps from zero
 * (core/multigrid.py:428-429), or an FP64 direct solve: the inverse of the
 * coarsest operator on its free DOFs, factored on the host at
 * vx_set_density and applied as a dense FP64 product in every V-cycle (at
 * most 1500 free DOFs, and only when that block is positive definite;
 * otherwise the sweeps are used and vx_info.coarse_direct is 0). */
enum { VX_COARSE_SWEEPS = 0, VX_COARSE_DIRECT = 1 };

typedef struct vx_solver* vx_handle;

/* Solver options.  vx_default_options() fills the reference's defaults:
 * max_levels 6, min_elems 4, n_smooth 2, omega 0.6, 30 coarsest sweeps,
 * e_min 1e-9 (core/multigrid.py:85-87, core/fea.py:244-246). */
typedef struct {
  int32_t struct_size;        /* sizeof(vx_options) */
  int32_t max_levels;
  int32_t min_elems;
  int32_t n_smooth;
  double omega;
  int32_t coarse_sweeps;
  int32_t precision;          /* VX_FP64 | VX_MIXED */
  int32_t smoother;           /* VX_SMOOTH_* */
  int32_t coarse_mode;        /* VX_COARSE_* */
  int32_t cheb_ratio_power;   /* p: a Chebyshev pass of k steps smooths over
                                 [cheb_ratio (n_smooth / k)^p lambda_max, lambda_max]; 0 (the
                                 default): one interval for every pass.  A degree-k polynomial
                                 resolves a spectrum about k^2 times wider, so p = 2 lets the
                                 long passes of a variable V-cycle reach lower modes */
  double cheb_ratio;          /* lower end of the Chebyshev interval / lambda_max */
  double inner_tol;           /* mixed: relative tolerance of each FP32 inner solve */
  int32_t max_outer;          /* mixed: max iterative-refinement steps */
  int32_t z_chunk;            /* node planes per work-group in the tile kernels (0: auto) */
  int32_t e_coarsen;          /* coarse-level E: 0 arithmetic mean of the 8 children
                                 (core/multigrid.py:287-290), 1 geometric, 2 harmonic */
  int32_t max_batch;          /* right-hand sides per vx_solve_batch call (workspace size) */
  int32_t flexible;           /* 1: flexible CG direction update (-z.Ap / p.Ap) also without a
                                 corrector (the default update is Fletcher-Reeves) */
  uint8_t level_smooth[16];   /* smoothing steps of level l (pre and post), l < 16; 0: n_smooth.
                                 The reference smooths every level alike (core/multigrid.py:
                                 426-436).  Counts that grow toward the coarse levels give a
                                 variable V-cycle (Bramble, Pasciak & Xu, Math. Comp. 56, 1991),
                                 whose preconditioner stays uniform for rediscretised coarse
                                 operators; those levels' launches are latency-bound, so their
                                 extra steps cost little.  The coarsest level uses
                                 coarse_sweeps. */
  int32_t eager;              /* 0: each PCG step's V-cycle is replayed from a captured HIP
                                 graph (one submission instead of one per kernel; same
                                 kernels, same results); 1: launched kernel by kernel */
  int32_t block_steps;        /* FP32 smoothing passes of >= 2 steps on the levels below level 0
                                 with at most 40000 nodes: steps per launch of the LDS-blocked
                                 smoother, whose work-groups repeat the halo work of the steps
                                 they fuse (those levels' launches are latency-bound).  0: the
                                 form vx_autotune picked, else blocked on levels of at most
                                 12000 nodes with a size rule; 1: one tile launch per step;
                                 2-8: blocked with that many (at most) */
  int32_t block_edge;         /* owned box edge (nodes) of the blocked smoother's work-groups,
                                 2-16; 0: chosen as block_steps is */
  int32_t keep_direction;     /* mixed, without the corrector: 1 = each FP32 inner solve of the
                                 iterative refinement continues the previous one's search
                                 direction (rescaled to the new residual) instead of restarting
                                 CG, a "reliable update" of the recurrence residual by the FP64
                                 true residual; 0 (default) = restart.  vx_solve and
                                 vx_solve_batch; vx_solve_parts always restarts */
  double keep_drift;          /* keep_direction: a refresh that moves the residual by more than
                                 this (relative to the new residual's norm) restarts the inner
                                 solve all the same; 0: no limit */
} vx_options;

/* Per-solve statistics (mirrors MGSolver.last_* of core/multigrid.py:482-517). */
typedef struct {
  int32_t iters;              /* PCG iterations (mixed: total inner iterations) */
  int32_t converged;          /* 1 if the ratio dropped below tol */
  int32_t outer_iters;        /* mixed: iterative-refinement steps; FP64: 0 */
  int32_t n_levels;
  double resid_ratio;         /* last recurrence ratio ||r||/||f_free|| */
  double true_resid_ratio;    /* ||f_free - A u|| / ||f_free|| recomputed in FP64 */
  double solve_ms;            /* wall time of the solve on the handle's stream */
  int32_t corrector;          /* learned corrector: 0 not attached, 1 used throughout,
                                 2 switched off by the FP64 true-residual check */
  int32_t coarse_direct;      /* 1: the coarsest level was solved directly (VX_COARSE_DIRECT) */
} vx_info;

/* Called synchronously on the caller's thread after every PCG iteration
 * with (iteration, resid_ratio, tol) -- core/multigrid.py:518-519. */
typedef void (*vx_progress_cb)(int32_t iteration, double resid_ratio, double tol,
                               void* user);

void vx_default_options(vx_options* opt);

/* Number of visible devices (0 when no GPU / no driver). */
int vx_device_count(void);

/* Replaces MGSolver.__init__ (core/multigrid.py:85-174) minus the fixed DOFs:
 * builds the level hierarchy of an nx*ny*nz grid with Poisson ratio nu. */
int vx_create(int32_t nx, int32_t ny, int32_t nz, double nu, const vx_options* opt,
              vx_handle* out);
void vx_destroy(vx_handle h);
const char* vx_last_error(vx_handle h);
int vx_set_stream(vx_handle h, void* hip_stream);
int vx_num_levels(vx_handle h);
/* dims of level l (elements) */
int vx_level_dims(vx_handle h, int32_t level, int32_t* nx, int32_t* ny, int32_t* nz);

/* MGSolver._build_free (core/multigrid.py:255-278): fixed_dofs is a HOST array
 * of original DOF ids; coarse masks are injected from the collocated fine node. */
int vx_set_fixed(vx_handle h, const int64_t* fixed_dofs_host, int64_t n_fixed);

/* MGSolver.set_density (core/multigrid.py:281-315) with the caller's void rule
 * of VoxelFEA._solve_multigrid (core/fea.py:467-469): E_dev is the per-element
 * stiffness scale (E*vsize) in reference element order; where active_dev (uint8,
 * may be NULL = all active) is 0 the element gets E*e_min. */
int vx_set_density(vx_handle h, const double* E_dev, const uint8_t* active_dev,
                   double e_min);

/* MGSolver.solve (core/multigrid.py:439-583): f_dev, u_dev, x0_dev (may be NULL)
 * are [3*(nx+1)*(ny+1)*(nz+1)] float64.  Returns VX_ERR_STALL when the stall
 * guard fires (u_dev then holds the last iterate), VX_ERR_NOT_CONVERGED when
 * max_iter ran out.  hist_host (may be NULL) receives up to hist_cap
 * per-iteration ratios. */
int vx_solve(vx_handle h, const double* f_dev, double* u_dev, const double* x0_dev,
             double tol, int32_t max_iter, vx_progress_cb cb, void* user,
             double* hist_host, int32_t hist_cap, vx_info* info);

/* Batched solve: nrhs (<= max_batch) load cases of the same grid, density and
 * fixed DOFs in one pass -- every operator, smoother, transfer and reduction
 * launch covers all load cases (vectors [nrhs][ndof], contiguous).  Each load
 * case follows exactly the recurrence of vx_solve and stops at its own
 * convergence (its updates are frozen from then on).  hist_host (may be NULL)
 * is [nrhs][hist_cap]; infos (may be NULL) receives one vx_info per load case.
 * The progress callback reports the largest ratio among the unconverged cases. */
int vx_solve_batch(vx_handle h, int32_t nrhs, const double* F_dev, double* U_dev, double tol,
                   int32_t max_iter, vx_progress_cb cb, void* user, double* hist_host,
                   int32_t hist_cap, vx_info* infos);

/* VoxelFEA._solve_jacobi (core/fea.py:497-593): Jacobi-preconditioned CG on
 * the active-DOF system (inactive elements dropped, no e_min, loads on nodes
 * of no active element ignored).  E_dev/active_dev in reference element
 * order, f_dev/u_dev full-grid float64; u is 0 on DOFs outside the system.
 * x0_dev (may be NULL) is a full-grid warm start: the iteration starts from
 * x0 on the system's free DOFs with r = f - K x0 there (core/fea.py:529-535);
 * a zero load still returns u = 0.  max_iter <= 0 means 4 * (number of free
 * DOFs). */
int vx_solve_jacobi(vx_handle h, const double* E_dev, const uint8_t* active_dev,
                    const double* f_dev, double* u_dev, const double* x0_dev, double tol,
                    int32_t max_iter, vx_progress_cb cb, void* user, double* hist_host,
                    int32_t hist_cap, vx_info* info);

/* ---------------------------------------------------------------------------
 * Parts of uneven size in one launch per operation.  The reference solves
 * each part with its own VoxelFEA / MGSolver (core/fea.py:244-305,
 * core/multigrid.py:439-583); here every part keeps its own grid, hierarchy,
 * fixed DOFs, density and PCG recurrence (its own scalars, stall guard and
 * convergence), while each operator, smoother, transfer and reduction
 * launch covers the parts of a level, and a converged part's share of a
 * launch returns at once.  Arrays are the parts' arrays packed one after
 * another in part order: elements (E, active) in the reference's element
 * order per part, DOFs (F, U) full-grid per part.
 * ------------------------------------------------------------------------- */
typedef struct vx_parts* vx_parts_handle;

/* nparts grids of element dims dims_host [nparts][3] (host), one Poisson
 * ratio, options as vx_create (coarse_mode VX_COARSE_SWEEPS, no flexible
 * update; max_batch is ignored).  Each part gets the hierarchy rule of
 * core/multigrid.py:141-148. */
int vx_parts_create(int32_t nparts, const int32_t* dims_host, double nu, const vx_options* opt,
                    vx_parts_handle* out);
void vx_parts_destroy(vx_parts_handle h);
const char* vx_parts_last_error(vx_parts_handle h);
int vx_parts_set_stream(vx_parts_handle h, void* hip_stream);
/* number of levels of one part's hierarchy */
int vx_parts_num_levels(vx_parts_handle h, int32_t part);
/* MGSolver._build_free per part: fixed_dofs_host holds every part's DOF ids
 * one part after another, counts_host [nparts] how many each */
int vx_parts_set_fixed(vx_parts_handle h, const int64_t* fixed_dofs_host,
                       const int64_t* counts_host);
/* MGSolver.set_density per part (the void rule of vx_set_density; e_min < 0:
 * the active-DOF system) */
int vx_parts_set_density(vx_parts_handle h, const double* E_dev, const uint8_t* active_dev,
                         double e_min);
/* Tile z-chunks of the parts, level by level (the vx_autotune of a parts
 * handle; no counterpart in the reference, which solves parts one by one):
 * each candidate, applied to all the parts of a level, is timed on reps
 * smoothing steps over them and the fastest kept.  Needs set_fixed and
 * set_density; drops captured graphs; a no-op when options.z_chunk > 0. */
int vx_parts_autotune(vx_parts_handle h, int32_t reps);
/* the z-chunk of one part's tile kernels at one of its levels (> 0), or
 * VX_ERR_ARG */
int vx_parts_zchunk(vx_parts_handle h, int32_t part, int32_t level);
/* Every part solved at once, each to its own tol with the recurrence of its
 * own vx_solve (FP64) or iterative refinement (VX_MIXED).  hist_host (may be
 * NULL) is [nparts][hist_cap]; infos (may be NULL) one vx_info per part.
 * Returns VX_ERR_STALL if a part's stall guard fired (FP64; the message names
 * the part), VX_ERR_NOT_CONVERGED if one did not converge. */
int vx_solve_parts(vx_parts_handle h, const double* F_dev, double* U_dev, double tol,
                   int32_t max_iter, vx_progress_cb cb, void* user, double* hist_host,
                   int32_t hist_cap, vx_info* infos);

/* Single-operator entry points for parity tests.  All float64, level-l sized.
 * vx_matvec: MGSolver._apply  (masked K u, core/multigrid.py:318-341)
 * vx_smooth: MGSolver._smooth (iters damped-Jacobi sweeps, :367-402; with
 *            VX_SMOOTH_CHEBYSHEV: a Chebyshev smoothing step of degree iters)
 * vx_restrict: MGSolver._restrict (R = P^T then coarse mask, :404-416)
 * vx_prolong: MGSolver._prolongate (P e then fine mask, :418-424)
 * vx_vcycle: MGSolver._vcycle(0, b) (:426-436)
 * precision selects the FP64 or FP32 kernels (inputs/outputs stay float64). */
int vx_matvec(vx_handle h, int32_t level, int32_t precision, const double* u, double* y);
int vx_smooth(vx_handle h, int32_t level, int32_t precision, const double* u,
              const double* b, int32_t iters, double* out);
int vx_restrict(vx_handle h, int32_t level, const double* r_fine, double* r_coarse);
int vx_prolong(vx_handle h, int32_t level, const double* e_coarse, double* e_fine);
int vx_vcycle(vx_handle h, int32_t precision, const double* b, double* z);
/* Inverse Jacobi diagonal of level l per DOF (MGSolver.set_density Minv). */
int vx_inv_diag(vx_handle h, int32_t level, double* minv);

/* VoxelFEA.element_von_mises_stress (core/fea.py:604-626): vm[e] for every
 * element in reference order, 0 where active_dev is 0. */
int vx_von_mises(vx_handle h, const double* u_dev, const double* E_dev,
                 const uint8_t* active_dev, double* vm_dev);

/* Post-solve fields of the reference's solve worker (core/solve_worker.py:358-431)
 * from u (full grid) and the per-element modulus E in Pa (reference order):
 * stress3d = von Mises / vsize, safety3d = clip(yield / max(vm, eps), 0, 10) on
 * active elements, disp3d = mean |u| of the 8 corner nodes (all float32, per
 * element).  scalars5_host receives max_stress_pa, stress_vmax (99th percentile
 * of active stress, >= 1e-6), max_displacement_m (all nodes),
 * displacement_vmax_m (active elements, >= 1e-9), min_safety_factor (NaN
 * when yield <= 0). */
int vx_post_fields(vx_handle h, const double* u_dev, const double* E_dev,
                   const uint8_t* active_dev, double vsize, double yield_pa,
                   float* stress3d_dev, float* safety3d_dev, float* disp3d_dev,
                   double* scalars5_host);

/* VoxelFEA.element_strain_energy (core/fea.py:595-602): u_e^T KE u_e with
 * unit E for active elements, 0 elsewhere; reference element order. */
int vx_strain_energy(vx_handle h, const double* u_dev, const uint8_t* active_dev,
                     double* ce_dev);

/* drop_unsupported_islands (core/fea.py:165-227) on the device: kept_dev gets
 * the active elements 6-connected to an element touching a fixed node; if
 * none is reached kept = active and *n_removed = 0. */
int vx_drop_islands(vx_handle h, const uint8_t* active_dev, const int64_t* fixed_dofs_host,
                    int64_t n_fixed, uint8_t* kept_dev, int64_t* n_removed_host);

/* vx_drop_islands without any device allocation.  It has no counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * next to HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same result as vx_drop_islands; fixed_dofs_dev is a DEVICE
 * array of n_fixed DOF ids (checked on the device: an id outside the grid is
 * VX_ERR_ARG).  Its working memory (per-node flags, counters) is one static
 * device array per GPU, held until the process ends, for grids of at most
 * 2^25 = 33,554,432 nodes (32 MiB; 320^3 elements); a larger grid is
 * VX_ERR_ARG.  Calls that overlap on one GPU share that array and race. */
int vx_drop_islands_noalloc(vx_handle h, const uint8_t* active_dev,
                            const int64_t* fixed_dofs_dev, int64_t n_fixed, uint8_t* kept_dev,
                            int64_t* n_removed_host);

/* Timing helper: average duration (ms) of `reps` back-to-back launches of the
 * level-l operator kernel (kind 0 = masked matvec, 1 = damped-Jacobi sweep,
 * 2 = Chebyshev smoothing step, 3 = the first two Chebyshev steps from zero in
 * one launch, 4 = the residual restricted to level l+1 in the same launch,
 * 5 = a Chebyshev step with the dot-product partials of the last step of a
 * pass, 6 = a damped-Jacobi step with those partials, the last step of an
 * FP32 two-step pass from a nonzero start; 3 and 4 FP32 only) in the given
 * precision, measured with HIP events on the handle's stream. */
int vx_time_kernel(vx_handle h, int32_t kind, int32_t level, int32_t precision,
                   int32_t reps, double* avg_ms);

/* ---------------------------------------------------------------------------
 * Learned coarse-grid corrector network (no counterpart in the reference: its
 * V-cycle, core/multigrid.py:426-436, has no learned part).  A 4-level 3D
 * U-Net on a node grid: 14 convolutions 3x3x3 (channels 16 -> c0 -> c1 -> c2
 * -> c2 and back, stride-2 down-sampling, trilinear up-sampling with additive
 * skips), BF16 activations and weights, FP32 accumulation on the matrix
 * cores.  Input and output have 16 channels; tensors are [B][voxel][16]
 * float32 with voxel = x + X*(y + Y*z).  The flat float32 parameter vector
 * holds, per layer, W[27][Cin][Cout] (tap = dx + 3*(dy + 3*dz)) then b[Cout].
 * ------------------------------------------------------------------------- */
typedef struct vx_net* vxn_handle;

/* channels c0, c1, c2 in {16, 32, 64}; parameters start at 0 */
int vxn_create(int32_t c0, int32_t c1, int32_t c2, vxn_handle* out);
void vxn_destroy(vxn_handle n);
const char* vxn_last_error(vxn_handle n);
int vxn_set_stream(vxn_handle n, void* hip_stream);
int64_t vxn_num_params(vxn_handle n);
/* copy the parameters from / to a device float32 vector; set also clears
 * the Adam moments */
int vxn_set_params(vxn_handle n, const float* params_dev);
int vxn_get_params(vxn_handle n, float* params_dev);
/* out = net(in) for B samples on an X*Y*Z node grid; keeps the activations
 * for vxn_backward */
int vxn_forward(vxn_handle n, int32_t B, int32_t X, int32_t Y, int32_t Z, const float* in_dev,
                float* out_dev);
/* gradient of sum(out * dout) w.r.t. the parameters (overwrites grad_dev)
 * for the inputs of the last vxn_forward */
int vxn_backward(vxn_handle n, const float* dout_dev, float* grad_dev);
/* one Adam step (step counts from 1) on the parameters */
int vxn_adam_step(vxn_handle n, const float* grad_dev, float lr, float beta1, float beta2,
                  float eps, int32_t step);

/* Learned coarse-grid corrector inside the V-cycle (no counterpart in the
 * reference; its V-cycle, core/multigrid.py:426-436, has none).  With a net
 * attached and enabled, every VX_MIXED solve's FP32 V-cycle stops descending
 * at `level`: there it pre-smooths and forms the residual t, computes the
 * plain V-cycle's coarse-grid correction v (restrict, coarser levels,
 * prolong) and the network's correction
 *   y[3n+c] = gain * rms(t) * minv[n] * net(t / rms(t), v, features)[n][c]
 * (free components; minv the level's inverse Jacobi diagonal), adds the
 * combination alpha v + beta y of least energy error (a 2x2 Galerkin system
 * per right-hand side, so never worse than v alone at that level), and
 * post-smooths.  The outer CG then uses the flexible direction update.
 * Every iterative-refinement step recomputes the FP64 true residual; if an
 * inner solve with the corrector fails to halve it (or goes non-finite) the
 * corrector is switched off for the rest of that solve and the plain V-cycle
 * continues (vx_info.corrector = 2); so does an inner solve that uses up
 * iter_budget iterations (> 0; e.g. the plain solve's count) without
 * reaching its tolerance.  net = NULL detaches; gain <= 0 picks
 * the largest node count along an axis of the level; enabled = 0 attaches
 * for vx_corrector_prepare / vx_corrector_train only.  The net handle must
 * outlive its use. */
int vx_set_corrector(vx_handle h, vxn_handle net, int32_t level, float gain, int32_t enabled,
                     int32_t iter_budget);

/* What the attached network adds to the plain coarse-grid correction, per
 * sample (t_i, v_i as from vx_corrector_prepare): the drop of the energy
 * error ||A^-1 t - c||_A^2 from c = v to the best c = alpha v + beta y,
 * divided by v.t (>= 0; 0 = nothing gained).  gain_host [nsamp]. */
int vx_corrector_eval(vx_handle h, int32_t nsamp, const float* T_dev, const float* V_dev,
                      double* gain_host);

/* Record, during the following solves, the level-`level` residual after
 * pre-smoothing of every `every`-th V-cycle (first right-hand side) into
 * buf_dev [cap][3 * nodes of the level] float32, up to cap of them (the
 * training inputs of the corrector).  buf_dev = NULL stops. */
int vx_capture(vx_handle h, int32_t level, float* buf_dev, int32_t cap, int32_t every);
int vx_capture_count(vx_handle h);

/* Training pairs for the attached corrector's level: for each of the nsamp
 * residuals t_i (T_dev [nsamp][3 * nodes] float32) the plain V-cycle's
 * coarse-grid correction v_i (V_dev, same shape), the residual it leaves,
 * r_i = t_i - A v_i (R_dev, same shape), and v_i . t_i (vt_dev [nsamp]
 * float64). */
int vx_corrector_prepare(vx_handle h, int32_t nsamp, const float* T_dev, float* V_dev,
                         float* R_dev, double* vt_dev);

/* Train the attached corrector on its level: `steps` Adam steps (learning
 * rate lr with cosine decay to 5 %) on mini-batches of `batch` samples drawn
 * by a seeded generator, minimising the mean over the batch of
 *   (E(alpha v + beta y) - E(v)) / (v.t),   E(c) = ||A^-1 t - c||_A^2,
 * the energy error of the correction the V-cycle applies (the best
 * combination of v and the network's y) relative to v alone (<= 0).  Its
 * gradient, -2 beta (t - alpha A v - beta A y) / (v.t), needs only the
 * level's own operator.  The network must start from non-zero outputs (at
 * y = 0 the gradient vanishes).  loss_host (may be NULL) receives the loss of
 * every step. */
int vx_corrector_train(vx_handle h, int32_t nsamp, const float* T_dev, const float* V_dev,
                       const float* R_dev, const double* vt_dev, int32_t batch, int32_t steps,
                       float lr, uint64_t seed, float* loss_host);

/* ---------------------------------------------------------------------------
 * Voxelisation inside test (core/inside_worker.py:118-162): for each point,
 * the parity of forward crossings of the ray p + t dir (t > 1e-7) with the
 * triangles (Moeller-Trumbore, det tolerance 1e-9, barycentric -1e-7 /
 * 1 + 1e-7), in the reference's floating-point operation order; out[i] = 1
 * inside.  verts_dev [nv][3] float64, faces_dev [nt][3] int64, dir_host [3]
 * (the reference normalises (1, 0.0073301, 0.0031337)).  Work goes on
 * `stream` (NULL: null stream) and the call synchronises it; the temporary
 * triangle bins are allocated and freed inside the call.
 * ------------------------------------------------------------------------- */
int vx_inside_mask(void* stream, const double* verts_dev, int64_t nv, const int64_t* faces_dev,
                   int64_t nt, const double* points_dev, int64_t np, const double* dir_host,
                   uint8_t* out_dev);
/* Same on the points of a voxel grid generated on the device
 * (core/inside_worker.py:87-111): target 0 = voxel centres in element order
 * (x slowest), 1 = nodes (x fastest); origin_host [3] is the grid's min
 * corner. */
int vx_inside_grid(void* stream, const double* verts_dev, int64_t nv, const int64_t* faces_dev,
                   int64_t nt, int32_t nx, int32_t ny, int32_t nz, const double* origin_host,
                   double vsize, int32_t target, const double* dir_host, uint8_t* out_dev);
const char* vx_inside_last_error(void);

/* Autotuner (no counterpart in the reference): for every level, time the
 * tile kernels' z-chunk (node planes per work-group) over a candidate set,
 * for the operator (matvec forms, class 0; timed in the PCG's fused
 * direction-update form), the smoother / residual forms (class 1) and the
 * smoother forms with dot-product partials (class 2), in the given
 * precision, timed with the handle's max_batch right-hand sides per launch,
 * and keep the fastest for the handle's later launches.  In FP32 it also
 * times, on every level below level 0 with at most 40000 nodes, that level's
 * smoothing pass as one tile launch per step against the blocked smoother's
 * box edges and steps per launch, and keeps the fastest form.  Needs
 * vx_set_fixed and vx_set_density; overwrites the handle's work vectors.
 * opt.z_chunk > 0 overrides the z-chunks, opt.block_steps / opt.block_edge
 * > 0 the blocked form. */
int vx_autotune(vx_handle h, int32_t precision, int32_t reps);
/* The z-chunk the handle uses for a level, precision and class (0 the PCG
 * operator, 1 a smoothing step, 2 the last smoothing step with its dot
 * partials, 3 two Chebyshev steps from zero in one launch, 4 the residual
 * restricted to the next level in the same launch: (zc + 1) / 2 coarse planes
 * per work-group). */
int vx_tuned_zchunk(vx_handle h, int32_t level, int32_t precision, int32_t cls);
/* The blocked smoother's plan for an FP32 pass of `steps` steps on a level
 * (one right-hand side): owned box edge and steps per launch, or 0 and 0 when
 * the pass runs one tile launch per step.  No reference counterpart (an
 * MI355X launch-shape choice). */
int vx_block_plan(vx_handle h, int32_t level, int32_t steps, int32_t* edge,
                  int32_t* per_launch);

/* Host-side element operator (the kernels' own code compiled for the CPU):
 * y24 = KE(nu) * u24 for one unit element in the reference's local DOF order
 * (core/fea.py:33-36), E = 1. */
void vx_element_apply_host(double nu, const double* u24, double* y24);

#ifdef __cplusplus
}
#endif
#endif /* VXFEA_H */
