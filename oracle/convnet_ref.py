"""CPU restatement of the learned coarse-grid corrector network (numpy).

Test infrastructure only: tests and the bench's CPU leg use it as the checker
of the HIP network kernels; the solver never calls it.

The network has no counterpart in the reference (the reference's V-cycle,
core/multigrid.py:426-436, has no learned part); this module fixes its exact
arithmetic so that the device kernels can be checked against it:

* tensors are ``[B, Z, Y, X, C]`` float arrays (x fastest, channels innermost,
  the device layout ``[B][voxel][C]``), grids are node grids;
* ``conv3``: 3x3x3 taps ``d = dx + 3*(dy + 3*dz)``, zero padding, stride 1 or 2,
  source index ``stride*o + (dx-1)`` per axis, weights ``W[27, Cin, Cout]``;
* ``up``: channel-wise trilinear prolongation from node grid ``(n+1)//2`` to
  ``n`` (even fine index -> weight 1 from coarse ``i/2``; odd -> 0.5 from
  ``(i-1)/2`` and ``(i+1)/2``, coarse indices past the end dropped);
* BF16 points (``bf16=True``): activations and weights enter every product
  rounded to bfloat16 (round to nearest even), products accumulate in float;
  hidden outputs are stored as bfloat16 after bias and ReLU, the last layer's
  output stays float32.  Backward: the pre-activation gradient is rounded to
  bfloat16 before it enters the data and weight gradients.

Pins: ``conv3`` against a direct loop over taps on tiny grids, ``up`` against
the explicit 1-D interpolation matrices, and the whole backward against
central finite differences of the float64 network (no rounding); tests in
tests/test_oracle_net.py.
"""

import numpy as np

# (name, level, cin, cout, stride, relu); level is the grid the OUTPUT lives on
def layer_table(c0=16, c1=32, c2=64, c3=64, cin=16, cout=16):
    assert c3 == c2, "the deepest skip adds level-3 features to level-2 features"
    return [
        ("enc0a", 0, cin, c0, 1, True),
        ("enc0b", 0, c0, c0, 1, True),
        ("down1", 1, c0, c1, 2, True),
        ("enc1", 1, c1, c1, 1, True),
        ("down2", 2, c1, c2, 2, True),
        ("enc2", 2, c2, c2, 1, True),
        ("down3", 3, c2, c3, 2, True),
        ("enc3", 3, c3, c3, 1, True),
        ("dec2", 2, c2, c2, 1, True),
        ("dec2c", 2, c2, c1, 1, True),
        ("dec1", 1, c1, c1, 1, True),
        ("dec1c", 1, c1, c0, 1, True),
        ("dec0", 0, c0, c0, 1, True),
        ("out", 0, c0, cout, 1, False),
    ]


def n_params(table):
    return sum(27 * ci * co + co for _, _, ci, co, _, _ in table)


def unpack(params, table):
    """Flat float32 vector -> [(W[27,Cin,Cout], b[Cout])] in layer order."""
    out, o = [], 0
    for _, _, ci, co, _, _ in table:
        W = params[o:o + 27 * ci * co].reshape(27, ci, co)
        o += 27 * ci * co
        b = params[o:o + co]
        o += co
        out.append((W, b))
    assert o == params.size
    return out


def init_params(table, seed=0, out_scale=0.1):
    """He-normal weights, zero biases; the output layer scaled by out_scale."""
    rng = np.random.default_rng(seed)
    ps = []
    for i, (_, _, ci, co, _, relu) in enumerate(table):
        std = np.sqrt(2.0 / (27 * ci))
        if not relu:
            std *= out_scale
        ps.append(rng.standard_normal(27 * ci * co) * std)
        ps.append(np.zeros(co))
    return np.concatenate(ps).astype(np.float32)


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def grid_dims(dims, levels=4):
    """Node-grid dims (X, Y, Z) of every network level."""
    out = [tuple(dims)]
    for _ in range(levels - 1):
        out.append(tuple((n + 1) // 2 for n in out[-1]))
    return out


def _src_index(n_out, n_src, stride, off):
    """Per-axis source index stride*o + off and its validity."""
    s = stride * np.arange(n_out) + off
    ok = (s >= 0) & (s < n_src)
    return np.where(ok, s, 0), ok


def conv3(x, W, b, stride, out_dims):
    """y[o] = b + sum_d x[stride*o + d - 1] W[d] (zero outside); x [B,Z,Y,X,C]."""
    B = x.shape[0]
    Zo, Yo, Xo = out_dims[2], out_dims[1], out_dims[0]
    Zi, Yi, Xi = x.shape[1:4]
    y = np.zeros((B, Zo, Yo, Xo, W.shape[2]), dtype=np.float64 if x.dtype == np.float64
                 else np.float32)
    for dz in range(3):
        iz, oz = _src_index(Zo, Zi, stride, dz - 1)
        for dy in range(3):
            iy, oy = _src_index(Yo, Yi, stride, dy - 1)
            for dx in range(3):
                ix, ox = _src_index(Xo, Xi, stride, dx - 1)
                m = (oz[:, None, None] & oy[None, :, None] & ox[None, None, :])
                xs = x[:, iz][:, :, iy][:, :, :, ix] * m[None, :, :, :, None]
                y += xs @ W[dx + 3 * (dy + 3 * dz)]
    return y + b


def conv3_adj_data(g, W, stride, in_dims):
    """Adjoint of conv3 w.r.t. x: dx[i] = sum over (o, d) with stride*o+d-1 = i."""
    B = g.shape[0]
    Zo, Yo, Xo = g.shape[1:4]
    Xi, Yi, Zi = in_dims
    dx_ = np.zeros((B, Zi, Yi, Xi, W.shape[1]), dtype=g.dtype)
    for dz in range(3):
        iz, oz = _src_index(Zo, Zi, stride, dz - 1)
        for dy in range(3):
            iy, oy = _src_index(Yo, Yi, stride, dy - 1)
            for dx in range(3):
                ix, ox = _src_index(Xo, Xi, stride, dx - 1)
                contrib = g @ W[dx + 3 * (dy + 3 * dz)].T
                m = (oz[:, None, None] & oy[None, :, None] & ox[None, None, :])
                contrib = contrib * m[None, :, :, :, None]
                # distinct outputs of one tap hit distinct inputs: plain add
                Z_, Y_, X_ = np.meshgrid(iz, iy, ix, indexing="ij")
                dx_[:, Z_[m], Y_[m], X_[m]] += contrib[:, m]
    return dx_


def conv3_adj_weight(x, g, stride):
    """dW[d] = sum_o x[stride*o + d - 1]^T g[o]."""
    Zo, Yo, Xo = g.shape[1:4]
    Zi, Yi, Xi = x.shape[1:4]
    dW = np.zeros((27, x.shape[4], g.shape[4]), dtype=np.float64)
    gf = g.reshape(-1, g.shape[4]).astype(np.float64)
    for dz in range(3):
        iz, oz = _src_index(Zo, Zi, stride, dz - 1)
        for dy in range(3):
            iy, oy = _src_index(Yo, Yi, stride, dy - 1)
            for dx in range(3):
                ix, ox = _src_index(Xo, Xi, stride, dx - 1)
                m = (oz[:, None, None] & oy[None, :, None] & ox[None, None, :])
                xs = x[:, iz][:, :, iy][:, :, :, ix] * m[None, :, :, :, None]
                dW[dx + 3 * (dy + 3 * dz)] = xs.reshape(-1, x.shape[4]).astype(np.float64).T @ gf
    return dW


def interp_matrix(n_fine):
    """1-D prolongation [n_fine, (n_fine+1)//2] of ``up``."""
    nc = (n_fine + 1) // 2
    P = np.zeros((n_fine, nc))
    for i in range(n_fine):
        if i % 2 == 0:
            P[i, i // 2] = 1.0
        else:
            for j in ((i - 1) // 2, (i + 1) // 2):
                if j < nc:
                    P[i, j] = 0.5
    return P


def up(c, fine_dims):
    """Channel-wise trilinear prolongation of c [B,Zc,Yc,Xc,C] to fine_dims (X,Y,Z)."""
    Px, Py, Pz = (interp_matrix(n) for n in fine_dims)
    t = np.einsum("xi,bzyic->bzyxc", Px, c)
    t = np.einsum("yj,bzjxc->bzyxc", Py, t)
    return np.einsum("zk,bkyxc->bzyxc", Pz, t)


def up_adj(f, coarse_dims):
    """Adjoint of ``up``."""
    fine = (f.shape[3], f.shape[2], f.shape[1])
    Px, Py, Pz = (interp_matrix(n) for n in fine)
    assert (Px.shape[1], Py.shape[1], Pz.shape[1]) == tuple(coarse_dims)
    t = np.einsum("xi,bzyxc->bzyic", Px, f)
    t = np.einsum("yj,bzyxc->bzjxc", Py, t)
    return np.einsum("zk,bzyxc->bkyxc", Pz, t)


class UNetRef:
    """The 4-level network; forward keeps the activations for backward."""

    def __init__(self, table=None):
        self.table = table or layer_table()

    def forward(self, params, x, bf=True):
        """x [B,Z,Y,X,Cin] -> float32 y [B,Z,Y,X,Cout] (float64 if bf is False)."""
        T = self.table
        P = unpack(params.astype(np.float32 if bf else np.float64), T)
        rnd = bf16 if bf else (lambda a: a)
        dt = np.float32 if bf else np.float64
        dims = grid_dims((x.shape[3], x.shape[2], x.shape[1]))
        Wr = [(rnd(W).astype(dt), b.astype(dt)) for W, b in P]
        act = {}

        def layer(i, inp):
            _, lev, _, _, stride, relu = T[i]
            y = conv3(inp, Wr[i][0], Wr[i][1], stride, dims[lev])
            if relu:
                y = np.maximum(y, 0)
            return rnd(y.astype(dt)) if relu else y.astype(dt)

        h = [None] * len(T)
        x0 = rnd(x.astype(dt))
        act["in"] = x0
        h[0] = layer(0, x0)
        for i in range(1, 8):
            h[i] = layer(i, h[i - 1])
        a2 = rnd(up(h[7], dims[2]) + h[5])
        h[8] = layer(8, a2)
        h[9] = layer(9, h[8])
        a1 = rnd(up(h[9], dims[1]) + h[3])
        h[10] = layer(10, a1)
        h[11] = layer(11, h[10])
        a0 = rnd(up(h[11], dims[0]) + h[1])
        h[12] = layer(12, a0)
        h[13] = layer(13, h[12])
        act.update(h=h, a2=a2, a1=a1, a0=a0, dims=dims, W=Wr, bf=bf)
        self.act = act
        return h[13]

    def backward(self, dout):
        """Gradient of sum(y * dout) w.r.t. the flat parameter vector."""
        T, A = self.table, self.act
        h, dims, Wr, bf = A["h"], A["dims"], A["W"], A["bf"]
        rnd = bf16 if bf else (lambda a: a)
        dt = np.float32 if bf else np.float64
        inputs = {0: A["in"], 8: A["a2"], 10: A["a1"], 12: A["a0"]}
        grads = [None] * len(T)

        def back(i, dy, need_dx=True):
            _, lev, _, _, stride, relu = T[i]
            g = dy * (h[i] > 0) if relu else dy
            g = rnd(g.astype(dt))
            x = inputs.get(i, h[i - 1])
            dW = conv3_adj_weight(x, g, stride)
            db = g.reshape(-1, g.shape[4]).astype(np.float64).sum(0)
            grads[i] = (dW, db)
            if not need_dx:
                return None
            in_dims = (x.shape[3], x.shape[2], x.shape[1])
            return conv3_adj_data(g, Wr[i][0], stride, in_dims).astype(dt)

        d = back(13, dout.astype(dt))
        d_a0 = back(12, d)
        d_s0 = d_a0
        d = up_adj(d_a0, dims[1]).astype(dt)
        d = back(11, d)
        d_a1 = back(10, d)
        d_s1 = d_a1
        d = up_adj(d_a1, dims[2]).astype(dt)
        d = back(9, d)
        d_a2 = back(8, d)
        d_s2 = d_a2
        d = up_adj(d_a2, dims[3]).astype(dt)
        d = back(7, d)
        d = back(6, d) + d_s2
        d = back(5, d)
        d = back(4, d) + d_s1
        d = back(3, d)
        d = back(2, d) + d_s0
        d = back(1, d)
        back(0, d, need_dx=False)
        return np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
