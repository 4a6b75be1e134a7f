"""Oracle restatement of the multi-process CPU matvec pool
(core/parallel_cpu.py:56-232).  Test and CPU-timing infrastructure only.

Elements are split into contiguous chunks over ``n_workers`` spawned
processes, each pinned to one BLAS thread (core/parallel_cpu.py:56-65).  The
host writes ``p`` into shared memory, every worker computes its chunk's
gather -> GEMM -> scale -> bincount into its own full-length partial, and the
host sums the partials in worker order (core/parallel_cpu.py:170-205).
"""

import multiprocessing as mp
from multiprocessing import shared_memory

import numpy as np

_APPLY, _DIAG, _STOP = 1, 2, -1


def _pin_blas_one_thread():
    try:
        import threadpoolctl
        threadpoolctl.threadpool_limits(limits=1)
    except Exception:
        pass


def _worker(edof, KE, ndof, p_name, out_name, ev_name, n_el, cmd, go, done):
    _pin_blas_one_thread()
    shms = [shared_memory.SharedMemory(name=n) for n in (p_name, out_name, ev_name)]
    try:
        p = np.ndarray((ndof,), np.float64, buffer=shms[0].buf)
        out = np.ndarray((ndof,), np.float64, buffer=shms[1].buf)
        ev = np.ndarray((n_el,), np.float64, buffer=shms[2].buf)
        flat = edof.ravel()
        dKE = np.diag(KE)
        while True:
            go.wait()
            go.clear()
            op = cmd.value
            if op == _STOP:
                break
            if op == _APPLY:
                contrib = (p[edof] @ KE.T) * ev[:, None]
            elif op == _DIAG:
                contrib = ev[:, None] * dKE[None, :]
            else:
                done.set()
                continue
            out[:] = np.bincount(flat, weights=contrib.ravel(), minlength=ndof)
            done.set()
    finally:
        for s in shms:
            s.close()


class CPUMatVecPoolRef:
    """Element-partitioned matvec over persistent spawn workers."""

    def __init__(self, edof, KE, ndof, n_workers):
        edof = np.ascontiguousarray(edof, dtype=np.int64)
        KE = np.ascontiguousarray(KE, dtype=np.float64)
        ne = edof.shape[0]
        bounds = np.linspace(0, ne, max(1, int(n_workers)) + 1).astype(np.int64)
        chunks = [(int(a), int(b)) for a, b in zip(bounds[:-1], bounds[1:]) if b > a]
        if len(chunks) < 2:
            raise ValueError("CPUMatVecPoolRef needs at least 2 non-empty chunks")
        self.ndof = int(ndof)
        self.n_workers = len(chunks)
        self._closed = False
        ctx = mp.get_context("spawn")
        self._p_shm = shared_memory.SharedMemory(create=True, size=self.ndof * 8)
        self._p = np.ndarray((self.ndof,), np.float64, buffer=self._p_shm.buf)
        self._shms, self._outs, self._evs = [], [], []
        self._cmds, self._go, self._done, self._procs = [], [], [], []
        for a, b in chunks:
            o = shared_memory.SharedMemory(create=True, size=self.ndof * 8)
            e = shared_memory.SharedMemory(create=True, size=(b - a) * 8)
            self._shms += [o, e]
            self._outs.append(np.ndarray((self.ndof,), np.float64, buffer=o.buf))
            self._evs.append(np.ndarray((b - a,), np.float64, buffer=e.buf))
            cmd, go, done = ctx.Value("i", 0), ctx.Event(), ctx.Event()
            proc = ctx.Process(target=_worker, daemon=True,
                               args=(edof[a:b], KE, self.ndof, self._p_shm.name,
                                     o.name, e.name, b - a, cmd, go, done))
            proc.start()
            self._cmds.append(cmd)
            self._go.append(go)
            self._done.append(done)
            self._procs.append(proc)

    def _run(self, op):
        for c in self._cmds:
            c.value = op
        for g in self._go:
            g.set()
        for d in self._done:
            d.wait()
            d.clear()

    def _sum(self):
        out = self._outs[0].copy()
        for o in self._outs[1:]:
            out += o
        return out

    def set_density(self, evec):
        evec = np.asarray(evec, dtype=np.float64).ravel()
        pos = 0
        for e in self._evs:
            e[:] = evec[pos:pos + e.shape[0]]
            pos += e.shape[0]

    def apply(self, p):
        self._p[:] = np.asarray(p, dtype=np.float64)
        self._run(_APPLY)
        return self._sum()

    def diagonal(self):
        self._run(_DIAG)
        return self._sum()

    def close(self):
        if self._closed:
            return
        self._closed = True
        try:
            for c in self._cmds:
                c.value = _STOP
            for g in self._go:
                g.set()
            for p in self._procs:
                p.join(timeout=2.0)
                if p.is_alive():
                    p.terminate()
        finally:
            for s in (self._p_shm, *self._shms):
                try:
                    s.close()
                    s.unlink()
                except Exception:
                    pass

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
