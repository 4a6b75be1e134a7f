"""Seeded synthetic voxel problems of the five benchmark workloads.

Each generator returns a :class:`Problem` in the reference solver's own
conventions (core/fea.py:129-148, core/solve_worker.py:243-358):

* ``active``: bool per element, element order x-slowest / z-fastest
  (``active3d.ravel()`` of an ``[nx, ny, nz]`` array);
* ``E_pa``: per-element Young's modulus in Pa over the FULL grid (void
  elements keep the modulus of the material field; the solver multiplies them
  by ``e_min``); the stiffness the solver sees is ``E_pa * vsize``;
* ``fixed_dofs``: int64 original DOF ids ``3*n + c``, nodes x-fastest;
* ``force``: float64 ``[ndof]`` in N (or ``[nrhs, ndof]`` for multi-RHS).

Assembly rules follow core/solve_worker.py:243-287: bearings fix whole nodes or
single components, a load is spread evenly over the nodes it reaches (here:
nodes of at least one kept element), and material that cannot reach a support
is dropped before the solve.  The unsupported-island drop used while
generating is scipy's 6-connected labelling; it is only the generator's
acceptance rule (the solve path has its own GPU flood fill).
"""

from dataclasses import dataclass, field

import numpy as np

STEEL = (200e9, 0.30)
ALUMINIUM = (70e9, 0.33)
ABS = (2.2e9, 0.35)
PP = (1.5e9, 0.30)   # PP modulus; one Poisson ratio per solve (core/fea.py:262)


@dataclass
class Problem:
    name: str
    nx: int
    ny: int
    nz: int
    vsize: float
    nu: float
    active: np.ndarray
    E_pa: np.ndarray
    fixed_dofs: np.ndarray
    force: np.ndarray
    tol: float = 1e-6
    max_levels: int = 6
    n_smooth: int = 2
    omega: float = 0.6
    yield_pa: float = 250e6
    meta: dict = field(default_factory=dict)

    @property
    def nelem(self):
        return self.nx * self.ny * self.nz

    @property
    def nnode(self):
        return (self.nx + 1) * (self.ny + 1) * (self.nz + 1)

    @property
    def ndof(self):
        return 3 * self.nnode

    @property
    def E_stiffness(self):
        """Per-element stiffness scale passed to ``solve`` (E*vsize,
        core/solve_worker.py:342-357)."""
        return self.E_pa * self.vsize


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------

def node_index(ix, iy, iz, nx, ny):
    return ix + (nx + 1) * (iy + (ny + 1) * iz)


def element_nodes_mask(active3d):
    """Bool [nz+1, ny+1, nx+1] (node order) of nodes of >= 1 active element."""
    a = np.asarray(active3d, dtype=bool)
    nx, ny, nz = a.shape
    m = np.zeros((nx + 1, ny + 1, nz + 1), dtype=bool)
    for ox in (0, 1):
        for oy in (0, 1):
            for oz in (0, 1):
                m[ox:ox + nx, oy:oy + ny, oz:oz + nz] |= a
    return np.ascontiguousarray(m.transpose(2, 1, 0))


def face_nodes(nx, ny, nz, axis, side):
    """Node ids on the face ``axis`` in {0,1,2} at index 0 or max."""
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1),
                             np.arange(nz + 1), indexing="ij")
    sel = (ix, iy, iz)[axis] == (0 if side == 0 else (nx, ny, nz)[axis])
    return np.sort(node_index(ix[sel], iy[sel], iz[sel], nx, ny)).astype(np.int64)


def fix_nodes(nodes, comps=(0, 1, 2)):
    nodes = np.asarray(nodes, dtype=np.int64)
    return np.unique(np.concatenate([3 * nodes + c for c in comps]))


def spread_load(ndof, nodes, total):
    """Total force vector ``total`` (3,) spread evenly over ``nodes``
    (core/solve_worker.py:262-267)."""
    f = np.zeros(ndof)
    nodes = np.asarray(nodes, dtype=np.int64)
    if nodes.size:
        per = np.asarray(total, dtype=float) / nodes.size
        for c in range(3):
            f[3 * nodes + c] += per[c]
    return f


def supported_part(active3d, fixed_dofs):
    """Elements 6-connected to an element touching a fixed node."""
    from scipy import ndimage
    a = np.asarray(active3d, dtype=bool)
    nx, ny, nz = a.shape
    fixed_nodes = np.unique(np.asarray(fixed_dofs, dtype=np.int64) // 3)
    is_fixed = np.zeros((nx + 1) * (ny + 1) * (nz + 1), dtype=bool)
    is_fixed[fixed_nodes] = True
    fn = is_fixed.reshape(nz + 1, ny + 1, nx + 1).transpose(2, 1, 0)
    seed = np.zeros_like(a)
    for ox in (0, 1):
        for oy in (0, 1):
            for oz in (0, 1):
                seed |= fn[ox:ox + nx, oy:oy + ny, oz:oz + nz]
    seed &= a
    lab, _ = ndimage.label(a)
    keep = np.unique(lab[seed])
    keep = keep[keep > 0]
    return np.isin(lab, keep) & a


def _smooth_field(rng, shape, n_ctrl=8):
    """Smooth random field: cubic B-spline surface over an n_ctrl^3 lattice
    of standard-normal control values (separable tensor product)."""
    coef = rng.standard_normal((n_ctrl + 3,) * 3)

    def basis(n):
        t = (np.arange(n) + 0.5) / n * n_ctrl
        i = np.floor(t).astype(int)
        s = t - i
        B = np.zeros((n, n_ctrl + 3))
        w = [(1 - s) ** 3 / 6, (3 * s ** 3 - 6 * s ** 2 + 4) / 6,
             (-3 * s ** 3 + 3 * s ** 2 + 3 * s + 1) / 6, s ** 3 / 6]
        for k in range(4):
            B[np.arange(n), i + k] = w[k]
        return B

    Bx, By, Bz = (basis(n) for n in shape)
    f = np.tensordot(Bx, coef, axes=(1, 0))
    f = np.tensordot(f, By, axes=(1, 1))          # [nx, c_z, ny]
    f = np.tensordot(f, Bz, axes=(1, 1))          # [nx, ny, nz]
    return f


def _random_solids(rng, shape, n_box=6, n_cyl=4):
    nx, ny, nz = shape
    ix, iy, iz = np.meshgrid(np.arange(nx) + 0.5, np.arange(ny) + 0.5,
                             np.arange(nz) + 0.5, indexing="ij", sparse=True)
    occ = np.zeros(shape, dtype=bool)
    dims = np.array(shape, dtype=float)
    for _ in range(n_box):
        size = rng.uniform(0.25, 0.75, 3) * dims
        lo = rng.uniform(0, 1, 3) * (dims - size)
        hi = lo + size
        occ |= ((ix >= lo[0]) & (ix < hi[0]) & (iy >= lo[1]) & (iy < hi[1])
                & (iz >= lo[2]) & (iz < hi[2]))
    coords = (ix, iy, iz)
    for _ in range(n_cyl):
        ax = int(rng.integers(0, 3))
        a, b = [k for k in range(3) if k != ax]
        r = rng.uniform(0.08, 0.25) * min(dims[a], dims[b])
        ca = rng.uniform(r, dims[a] - r)
        cb = rng.uniform(r, dims[b] - r)
        length = rng.uniform(0.5, 1.0) * dims[ax]
        z0 = rng.uniform(0, dims[ax] - length)
        occ |= (((coords[a] - ca) ** 2 + (coords[b] - cb) ** 2 <= r * r)
                & (coords[ax] >= z0) & (coords[ax] < z0 + length))
    return occ


def _top_patch_nodes(rng, active3d, patch=8):
    """Nodes (of kept elements) on the z=max face inside a random patch x
    patch voxel window (patch+1 nodes a side) that holds at least one such
    node."""
    nx, ny, nz = active3d.shape
    nm = element_nodes_mask(active3d)[nz]              # [ny+1, nx+1]
    for _ in range(1000):
        x0 = int(rng.integers(0, max(1, nx - patch + 1)))
        y0 = int(rng.integers(0, max(1, ny - patch + 1)))
        sub = nm[y0:y0 + patch + 1, x0:x0 + patch + 1]
        if sub.any():
            jy, jx = np.nonzero(sub)
            return node_index(jx + x0, jy + y0, nz, nx, ny).astype(np.int64)
    raise RuntimeError("no loaded top-face patch found")


def _random_part(rng, shape, occ_range, fixed_axis_face=(2, 0)):
    """Random union of 6 boxes + 4 cylinders whose supported part has an
    occupancy in ``occ_range`` and reaches the top face."""
    nx, ny, nz = shape
    fixed = fix_nodes(face_nodes(nx, ny, nz, *fixed_axis_face))
    for _ in range(500):
        occ = _random_solids(rng, shape)
        kept = supported_part(occ, fixed)
        frac = kept.mean()
        if occ_range[0] <= frac <= occ_range[1] and kept[:, :, -1].any():
            return kept, fixed
    raise RuntimeError("random part rejection sampling did not converge")


# ---------------------------------------------------------------------------
# the five workloads
# ---------------------------------------------------------------------------

def cantilever(nx=64, ny=16, nz=16):
    """64x16x16 solid steel cantilever: x=0 nodes fixed, -1000 N z over the
    x=max nodes; 3-level V-cycle, tol 1e-8."""
    vsize = 1e-3
    E, nu = STEEL
    active = np.ones(nx * ny * nz, dtype=bool)
    fixed = fix_nodes(face_nodes(nx, ny, nz, 0, 0))
    load_nodes = face_nodes(nx, ny, nz, 0, 1)
    ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    force = spread_load(ndof, load_nodes, (0.0, 0.0, -1000.0))
    return Problem("cantilever", nx, ny, nz, vsize, nu, active,
                   np.full(nx * ny * nz, E), fixed, force, tol=1e-8,
                   max_levels=3)


def voids_block(n=128, n_spheres=40, seed=2):
    """n^3 aluminium block with spherical voids (radius 4-12 voxels at
    n=128, scaled with n), z=0 fixed, +5000 N x on the z=max face."""
    rng = np.random.default_rng(seed)
    s = n / 128.0
    vsize = 1e-3
    E, nu = ALUMINIUM
    c = (np.arange(n) + 0.5)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij", sparse=True)
    void = np.zeros((n, n, n), dtype=bool)
    for _ in range(n_spheres):
        r = rng.uniform(4, 12) * s
        cx, cy, cz = rng.uniform(0, n, 3)
        void |= (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= r * r
    fixed = fix_nodes(face_nodes(n, n, n, 2, 0))
    kept = supported_part(~void, fixed)
    nm = element_nodes_mask(kept)
    top = np.nonzero(nm[n].ravel())[0]
    load_nodes = (top + n * (n + 1) * (n + 1)).astype(np.int64)
    ndof = 3 * (n + 1) ** 3
    force = spread_load(ndof, load_nodes, (5000.0, 0.0, 0.0))
    return Problem("voids_block", n, n, n, vsize, nu, kept.ravel(),
                   np.full(n ** 3, E), fixed, force,
                   meta={"seed": seed, "occupancy": float(kept.mean())})


def multi_rhs_part(n=96, nrhs=32, seed=3):
    """n^3 random union of 6 boxes + 4 cylinders (occupancy 40-70 %), z=0
    fixed, ``nrhs`` load cases each a 100-5000 N force in a random direction
    on a random 8x8-voxel top-face patch."""
    rng = np.random.default_rng(seed)
    vsize = 1e-3
    E, nu = STEEL
    kept, fixed = _random_part(rng, (n, n, n), (0.40, 0.70))
    ndof = 3 * (n + 1) ** 3
    forces = np.zeros((nrhs, ndof))
    for k in range(nrhs):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        mag = rng.uniform(100.0, 5000.0)
        forces[k] = spread_load(ndof, _top_patch_nodes(rng, kept), mag * d)
    return Problem("multi_rhs_part", n, n, n, vsize, nu, kept.ravel(),
                   np.full(n ** 3, E), fixed, forces,
                   meta={"seed": seed, "occupancy": float(kept.mean()),
                         "nrhs": nrhs})


def uneven_parts(n_parts=16, lo=32, hi=128, seed=4):
    """``n_parts`` random ABS parts, edges uniform in [lo, hi] per axis,
    occupancy 30-70 %, z=0 fixed, one random top-patch load each."""
    rng = np.random.default_rng(seed)
    E, nu = ABS
    parts = []
    for k in range(n_parts):
        shape = tuple(int(v) for v in rng.integers(lo, hi + 1, 3))
        kept, fixed = _random_part(rng, shape, (0.30, 0.70))
        nx, ny, nz = shape
        ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        force = spread_load(ndof, _top_patch_nodes(rng, kept),
                            rng.uniform(100.0, 5000.0) * d)
        parts.append(Problem(f"uneven_part{k}", nx, ny, nz, 1e-3, nu,
                             kept.ravel(), np.full(nx * ny * nz, E), fixed,
                             force, meta={"seed": seed, "index": k,
                                          "occupancy": float(kept.mean())}))
    return parts


def two_material(n=256, seed=5, occupancy=0.55):
    """n^3 two-material part: solid where a smooth random level set is below
    the threshold giving ~55 % occupancy, steel vs PP split by a second
    smooth level set (one Poisson ratio 0.3), x=0 nodes fixed, roller (z
    locked) on x=max, 20 kN -z spread over the z=max nodes of the part."""
    rng = np.random.default_rng(seed)
    vsize = 1e-3
    phi1 = _smooth_field(rng, (n, n, n))
    phi2 = _smooth_field(rng, (n, n, n))
    raw = phi1 < np.quantile(phi1, occupancy)
    E = np.where(phi2 < np.median(phi2), STEEL[0], PP[0])
    fixed_full = fix_nodes(face_nodes(n, n, n, 0, 0))
    roller = fix_nodes(face_nodes(n, n, n, 0, 1), comps=(2,))
    fixed = np.unique(np.concatenate([fixed_full, roller]))
    kept = supported_part(raw, fixed_full)
    nm = element_nodes_mask(kept)
    top = np.nonzero(nm[n].ravel())[0]
    load_nodes = (top + n * (n + 1) * (n + 1)).astype(np.int64)
    ndof = 3 * (n + 1) ** 3
    force = spread_load(ndof, load_nodes, (0.0, 0.0, -20000.0))
    return Problem("two_material", n, n, n, vsize, 0.30, kept.ravel(),
                   E.ravel().astype(float), fixed, force,
                   meta={"seed": seed, "occupancy": float(kept.mean()),
                         "steel_frac": float((E.ravel()[kept.ravel()] == STEEL[0]).mean())})


WORKLOADS = {
    "cantilever": cantilever,
    "voids_block": voids_block,
    "multi_rhs_part": multi_rhs_part,
    "uneven_parts": uneven_parts,
    "two_material": two_material,
}


# ---------------------------------------------------------------------------
# triangle meshes for the voxelisation inside test
# ---------------------------------------------------------------------------
def box_mesh(lo, hi):
    """Closed axis-aligned box: 8 vertices, 12 triangles."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    v = np.array([[x, y, z] for z in (lo[2], hi[2]) for y in (lo[1], hi[1])
                  for x in (lo[0], hi[0])])
    f = np.array([[0, 2, 1], [1, 2, 3], [4, 5, 6], [5, 7, 6], [0, 1, 4], [1, 5, 4],
                  [2, 6, 3], [3, 6, 7], [0, 4, 2], [2, 4, 6], [1, 3, 5], [3, 7, 5]],
                 dtype=np.int64)
    return v, f


def icosphere(level=2, center=(0.0, 0.0, 0.0), radius=1.0):
    """Closed sphere: an icosahedron subdivided ``level`` times
    (20 * 4^level triangles), vertices on the sphere."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
         [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
         [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
         [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    verts = [np.asarray(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(level):
        mid = {}
        nf = []

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                mid[key] = len(verts) - 1
            return mid[key]
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    V = np.asarray(verts) * float(radius) + np.asarray(center, float)
    return V, np.asarray(f, dtype=np.int64)
