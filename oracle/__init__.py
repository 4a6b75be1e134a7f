"""CPU oracle for the voxel hex8 MG-PCG solve -- TEST INFRASTRUCTURE ONLY.

This package is a NumPy restatement of the reference solver's algorithm
(core/fea.py, core/multigrid.py, core/parallel_cpu.py and the post-solve
arithmetic of core/solve_worker.py).  It is the checker for the HIP solver in
``voxelsim-fea_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path never
routes through it and never falls back to it.

Pinning (what checks each function against something other than its own
formula; see tests/test_oracle_*.py):

* ``hex8_KE``: re-derived by an independent Hadamard-mode closed form
  (tests/test_oracle_element.py), symmetric, PSD with exactly six zero
  eigenvalues, rigid-body null space, linear in E, uniform-strain patch test.
* ``hex8_B_C``: closed form (entries +-1/4), uniaxial / pure-shear von Mises.
* ``build_edof`` / ``node_id``: brute force on tiny grids, index formulas.
* ``prolong_op``: partition of unity, exact reproduction of linear fields,
  injection at coarse nodes; ``restrict`` is adjoint to ``prolongate``.
* ``MGSolverRef._apply``: equals a scipy-assembled sparse K times u.
* ``MGSolverRef.set_density`` diagonal: equals the assembled diagonal.
* V-cycle: equals the product of dense sweep/transfer matrices on tiny grids.
* ``MGSolverRef.solve`` / ``VoxelFEARef.solve``: converged u equals a sparse
  direct solve; per-iteration ``resid_history`` / ``last_iters`` match golden
  fixtures written by tests/golden/make_golden.py from the reference itself.
* post-solve fields: closed forms (von Mises of uniaxial / shear states,
  safety clip), and golden fixtures.
* ``convnet_ref`` (the learned corrector's network, which has no reference
  counterpart): convolution against a direct loop, prolongation against its
  1-D matrices, adjoint identities, backward against central finite
  differences of the float64 network (tests/test_oracle_net.py).
"""
