"""Pins of the configuration fixtures (tests/golden/configs): scaled-down
instances of the benchmark workloads solved by the reference's own
Jacobi-PCG on the active-DOF system (core/fea.py:497-593, tol 1e-10), written
by tests/golden/make_golden.py.

Each fixture is checked here without the reference: its u solves the
active-DOF system assembled by scipy from the oracle's element matrix (true
residual), vanishes off that system, and its von Mises field is the oracle's;
its inputs are exactly what the workload generators produce, so the GPU
parity tests and the benchmark solve the same problems.
"""

import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import GOLDEN
from oracle.fea_ref import build_edof, hex8_KE, von_mises

CFG = os.path.join(GOLDEN, "configs")
NAMES = sorted(f[:-4] for f in os.listdir(CFG) if f.endswith(".npz"))


def load_config(name):
    g = np.load(os.path.join(CFG, name + ".npz"))
    return {k: g[k] for k in g.files}


def active_system(g):
    """(K restricted to free used DOFs as CSR, their ids): the system of the
    reference's Jacobi path -- active elements only, DOFs of active elements,
    fixed DOFs removed (core/fea.py:271-279, 357-369, 497-535)."""
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    act = g["active"].astype(bool)
    edof = build_edof(nx, ny, nz)[act]
    KE = hex8_KE(1.0, float(g["nu"]))
    E = (g["E_pa"] * float(g["vsize"]))[act]
    n = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    rows = np.repeat(edof, 24, axis=1).ravel()
    cols = np.tile(edof, (1, 24)).ravel()
    K = sp.csr_matrix(((E[:, None, None] * KE[None]).ravel(), (rows, cols)), shape=(n, n))
    used = np.zeros(n, bool)
    used[np.unique(edof)] = True
    free = used.copy()
    free[g["fixed_dofs"]] = False
    ids = np.nonzero(free)[0]
    return K[ids][:, ids], ids, used


def test_every_configuration_has_a_fixture():
    assert NAMES == ["c2_voids_block32", "c3_multi_rhs_part32", "c4_uneven_part_a",
                     "c4_uneven_part_b", "c5_two_material24"]


@pytest.mark.parametrize("name", NAMES)
def test_fixture_solves_the_active_system(name):
    g = load_config(name)
    Kff, ids, used = active_system(g)
    F = np.atleast_2d(g["force"])
    U = np.atleast_2d(g["u"])
    for f, u in zip(F, U):
        ff = f[ids]
        res = np.linalg.norm(ff - Kff @ u[ids]) / np.linalg.norm(ff)
        # the reference stopped at a recurrence ratio of 1e-10
        assert res < 1e-9, res
        assert np.all(u[~used] == 0)
        assert np.all(u[g["fixed_dofs"]] == 0)


@pytest.mark.parametrize("name", NAMES)
def test_fixture_von_mises_is_the_oracles(name):
    g = load_config(name)
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    for u, vm in zip(np.atleast_2d(g["u"]), np.atleast_2d(g["vm"])):
        ref = von_mises(nx, ny, nz, float(g["nu"]), u, g["E_pa"], g["active"]) / float(g["vsize"])
        assert np.abs(ref - vm).max() <= 1e-12 * np.abs(vm).max()


def regenerate(name):
    from core import workloads as W
    return {
        "c5_two_material24": lambda: W.two_material(n=24),
        "c2_voids_block32": lambda: W.voids_block(n=32),
        "c3_multi_rhs_part32": lambda: W.multi_rhs_part(n=32, nrhs=2),
        "c4_uneven_part_a": lambda: W.uneven_parts(n_parts=2, lo=33, hi=48, seed=7)[1],
        "c4_uneven_part_b": lambda: W.uneven_parts(n_parts=1, lo=33, hi=48, seed=11)[0],
    }[name]()


@pytest.mark.parametrize("name", NAMES)
def test_generators_reproduce_the_fixture_inputs(name):
    g = load_config(name)
    p = regenerate(name)
    assert (p.nx, p.ny, p.nz) == (int(g["nx"]), int(g["ny"]), int(g["nz"]))
    assert p.nu == float(g["nu"]) and p.vsize == float(g["vsize"])
    assert np.array_equal(p.active, g["active"])
    assert np.array_equal(p.E_pa, g["E_pa"])
    assert np.array_equal(p.fixed_dofs, g["fixed_dofs"])
    assert np.array_equal(p.force, g["force"])
