"""MI355X-native voxel FEA solve core.

Drop-in for the reference's ``core`` package on the solve path: put
``voxelsim-fea_amd`` on ``sys.path`` (as the reference's worker puts its
add-on directory there, core/solve_worker.py:76-83) and import
``core.fea.VoxelFEA``, ``core.multigrid.MGSolver``, ``core.backend`` and
``core.compute_plan``.  The solve runs in hand-written HIP kernels for gfx950
behind the C ABI of ``include/vxfea.h`` (``lib/libvxfea.so``); torch-ROCm
tensors are only device buffers.
"""
