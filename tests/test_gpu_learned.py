"""GPU tests of the learned coarse-grid corrector inside the V-cycle.

The network (BF16 U-Net, core/learned.py) is trained on the GPU on residual
-> plain-correction pairs captured from the solver's own MG-PCG.  The tests
show what the trained network adds over the method without it (the plain
coarse-grid correction) on held-out residuals, that flexible CG with the
corrector meets the same tolerance and solution as the plain solve, and that
the FP64 true-residual check falls back to the plain V-cycle.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPT = dict(precision="mixed", smoother="chebyshev", cheb_ratio=0.1, system="active",
           inner_tol=1e-6)


@pytest.fixture(scope="module")
def setup():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from core import workloads as W
    from core.multigrid import MGSolver
    p = W.two_material(n=32, seed=7)
    E = torch.tensor(p.E_stiffness, device="cuda")
    A = torch.tensor(p.active.astype(np.uint8), device="cuda")
    f = torch.tensor(p.force, device="cuda")
    mg = MGSolver(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, compute_mode="GPU", **OPT)
    mg.set_density(E, active=A)
    u = torch.empty_like(f)
    mg.solve_device(f, u, tol=1e-6)
    assert mg.last_converged
    return dict(torch=torch, p=p, mg=mg, f=f, u_plain=u.clone(), plain_iters=mg.last_iters)


def _pairs(s, level, n=64):
    mg, torch = s["mg"], s["torch"]
    mg.detach_corrector()
    T = mg.capture(level, n, 1)
    u = torch.empty_like(s["f"])
    mg.solve_device(s["f"], u, tol=1e-6)
    mg.stop_capture()
    k = mg.capture_count()
    assert k >= 16
    return T[:k].contiguous()


def test_trained_network_gain_on_held_out_residuals(setup):
    s = setup
    torch, mg = s["torch"], s["mg"]
    from core.learned import CorrectorNet
    level = 1
    print(f"learned corrector: two_material 32^3, level {level}", flush=True)
    T = _pairs(s, level)
    train, held = T[0::2].contiguous(), T[1::2].contiguous()
    net = CorrectorNet(16, 32, 64, seed=3, out_scale=0.1)
    mg.attach_corrector(net, level, enabled=False)
    Vtr, Rtr, vttr = mg.corrector_pairs(train)
    Vh, _, _ = mg.corrector_pairs(held)
    assert bool((vttr > 0).all())
    g_untrained = mg.corrector_eval(held, Vh)
    loss = mg.corrector_train(train, Vtr, Rtr, vttr, batch=4, steps=400, lr=3e-4, seed=5)
    g_trained = mg.corrector_eval(held, Vh)
    zero = CorrectorNet(16, 32, 64, seed=3, out_scale=0.0)
    zero.set_params(torch.zeros(zero.n_params, device="cuda"))
    mg.attach_corrector(zero, level, enabled=False)
    g_scale_only = mg.corrector_eval(held, Vh)
    print(f"  loss {loss[:10].mean():.4f} -> {loss[-20:].mean():.4f}; held-out energy gain over "
          f"the plain correction: rescaled v only {g_scale_only.mean():.4f}, with the untrained "
          f"net {g_untrained.mean():.4f}, trained {g_trained.mean():.4f}", flush=True)
    # the best combination is never worse than the plain correction
    assert g_untrained.min() > -1e-6 and g_trained.min() > -1e-6
    # training lowers the loss and adds what the untrained network does not
    assert loss[-20:].mean() < loss[:10].mean()
    assert g_trained.mean() > g_untrained.mean() + 1e-3
    mg.detach_corrector()


def test_flexible_cg_with_corrector_matches_plain(setup):
    s = setup
    torch, mg = s["torch"], s["mg"]
    from core.learned import train_corrector
    net, rep = train_corrector(mg, s["f"], 1, n_samples=48, steps=200, lr=3e-4, seed=1)
    u = torch.empty_like(s["f"])
    mg.solve_device(s["f"], u, tol=1e-6)
    print(f"  plain {s['plain_iters']} iterations, learned {mg.last_iters} "
          f"({mg.last_corrector}), true residual {mg.last_true_resid_ratio:.2e}", flush=True)
    assert mg.last_corrector in ("used", "fell back")
    assert mg.last_converged and mg.last_true_resid_ratio < 1e-6
    up = s["u_plain"]
    err = float((u - up).abs().max() / up.abs().max())
    assert err < 1e-5, err
    mg.detach_corrector()


def test_true_residual_check_falls_back(setup):
    s = setup
    torch, mg = s["torch"], s["mg"]
    from core.learned import CorrectorNet
    net = CorrectorNet(16, 16, 16, seed=9, out_scale=1.0)
    # a budget of 3 inner iterations cannot reach the inner tolerance
    mg.attach_corrector(net, 1, enabled=True, iter_budget=3)
    u = torch.empty_like(s["f"])
    mg.solve_device(s["f"], u, tol=1e-6)
    assert mg.last_corrector == "fell back"
    assert mg.last_converged and mg.last_true_resid_ratio < 1e-6
    up = s["u_plain"]
    assert float((u - up).abs().max() / up.abs().max()) < 1e-5
    mg.detach_corrector()
    mg.solve_device(s["f"], u, tol=1e-6)
    assert mg.last_corrector == "none"


def test_batched_solve_with_corrector(setup):
    s = setup
    torch = s["torch"]
    from core.learned import CorrectorNet
    from core.multigrid import MGSolver
    p = s["p"]
    mg = MGSolver(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, compute_mode="GPU", max_batch=3, **OPT)
    mg.set_density(torch.tensor(p.E_stiffness, device="cuda"),
                   active=torch.tensor(p.active.astype(np.uint8), device="cuda"))
    F = torch.stack([s["f"], 2.0 * s["f"], -s["f"]])
    U = torch.empty_like(F)
    net = CorrectorNet(16, 32, 64, seed=4, out_scale=0.1)
    mg.attach_corrector(net, 1, enabled=True)
    mg.solve_batch_device(F, U, tol=1e-6)
    assert all(b["converged"] for b in mg.last_batch)
    up = s["u_plain"]
    for q, sc in enumerate((1.0, 2.0, -1.0)):
        err = float((U[q] - sc * up).abs().max() / up.abs().max() / abs(sc))
        assert err < 1e-5, (q, err)
    mg.close()
