"""Padding with void elements (active-DOF system): a part whose edges do not
halve gets a coarsenable grid; the solution on the part is the unpadded one."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_padded_dims_rule():
    from core.multigrid import padded_dims
    assert padded_dims((30, 27, 21)) == (32, 28, 24)
    assert padded_dims((64, 64, 64)) == (64, 64, 64)
    assert padded_dims((34, 102, 120)) == (40, 104, 120)


@pytest.mark.parametrize("precision", ["mixed"])
def test_padded_solve_equals_unpadded(precision):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from core import workloads as W
    from core.multigrid import MGSolver
    (p,) = W.uneven_parts(n_parts=1, lo=21, hi=31, seed=11)
    print(f"padded solve {precision}: part {p.nx}x{p.ny}x{p.nz}", flush=True)
    E = torch.tensor(p.E_stiffness, device="cuda")
    A = torch.tensor(p.active.astype(np.uint8), device="cuda")
    f = torch.tensor(p.force, device="cuda")
    kw = dict(precision=precision, smoother="chebyshev", cheb_ratio=0.1, system="active",
              inner_tol=1e-8)
    out = {}
    for pad in (False, True):
        mg = MGSolver(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, compute_mode="GPU", pad=pad, **kw)
        mg.set_density(E, active=A)
        u = torch.empty_like(f)
        mg.solve_device(f, u, tol=1e-9, max_cg=20000)
        assert mg.last_converged
        out[pad] = (u.cpu().numpy(), mg.last_iters, len(mg.levels))
        mg.close()
    (u0, it0, nl0), (u1, it1, nl1) = out[False], out[True]
    print(f"  unpadded {nl0} levels {it0} its; padded {nl1} levels {it1} its", flush=True)
    assert nl1 > nl0
    assert it1 < it0
    assert np.abs(u1 - u0).max() <= 1e-6 * np.abs(u0).max()
