"""Ragged batches: many independent Lindblad systems of uneven dimension.

lime propagates one system at a time (oqs.py:1590-1688).  ``RaggedLindblad`` takes a list
of instances of different n and advances all of them together, in ONE kernel launch
(``lime_lindblad_rk4_ragged``): the instances are packed into three pools (complex64
operators, complex128 states, complex64 per-stage workspace), ordered by cost
n^3 (1 + K_d) (largest first), and a persistent grid of workgroups drains that order
through an atomic queue, each instance on one workgroup for every step of the call.  The
launch shape (tile edge, tile groups per workgroup) comes from the autotuner
(``lime_amd.autotune.tune_ragged``); ``capture``/``run`` put the launch in a hipGraph
(torch.cuda.CUDAGraph) and replay it.

``RaggedClasses`` is the earlier organisation, kept as the comparison point: size classes,
one operator batch each on the per-class fused kernels, all captured into one graph.

An instance is a dict with ``H`` (n, n), ``rho0`` (n, n), optional ``c_ops`` (list of
(n, n) jump operators) and optional ``dephasing`` (scalar or length-n rates gamma_i of
L_i = sqrt(gamma_i)|i><i|, given compactly instead of n dense matrices).
"""
import ctypes

import numpy as np
import torch

from . import autotune
from .propagator import LindbladPropagator, device


def fold_host(inst):
    """(H_eff, D, dense L list) of one instance, complex128 NumPy."""
    H = np.asarray(inst["H"], dtype=np.complex128)
    n = H.shape[0]
    gam_d = np.zeros(n)
    D = np.zeros((n, n), dtype=np.complex128)
    G = np.zeros((n, n), dtype=np.complex128)
    if inst.get("dephasing") is not None:
        g = np.broadcast_to(np.asarray(inst["dephasing"], dtype=np.float64), (n,))
        gam_d += g
        D[np.arange(n), np.arange(n)] += g
    dense = []
    for L in inst.get("c_ops") or []:
        L = np.asarray(L, dtype=np.complex128)
        d = np.diagonal(L)
        if np.count_nonzero(L) == np.count_nonzero(d):            # diagonal jump operator
            gam_d += np.abs(d) ** 2
            D += np.outer(d, d.conj())
        else:
            G += L.conj().T @ L
            dense.append(L)
    G[np.arange(n), np.arange(n)] += gam_d
    return H - 0.5j * G, D, dense


class RaggedLindblad:
    """One persistent launch for a list of instances of uneven n (FP32 MFMA, FP64 state)."""

    def __init__(self, instances, tune=True, dt_tune=0.01, verbose=False, max_groups=0):
        from . import _lib
        self.dev = device()
        self._lib = _lib.load()
        self.count = len(instances)
        self.sizes = np.array([np.asarray(i["H"]).shape[0] for i in instances])
        rec = np.zeros(self.count, dtype=[("n", "<i4"), ("n_dense", "<i4"), ("heff", "<i8"),
                                          ("dmat", "<i8"), ("lops", "<i8"), ("rho", "<i8"),
                                          ("work", "<i8")])
        assert rec.dtype.itemsize == ctypes.sizeof(_lib.RaggedInstance)
        ops, rhos = [], []
        o_ops = o_rho = o_work = 0
        cost = np.zeros(self.count)
        self.kd = np.zeros(self.count, dtype=np.int64)
        for i, inst in enumerate(instances):
            heff, dmat, dense = fold_host(inst)
            n, kd = heff.shape[0], len(dense)
            r0 = np.asarray(inst["rho0"], dtype=np.complex128)
            if r0.shape != (n, n):
                raise ValueError(f"instance {i}: rho0 has shape {r0.shape}, H is {n} x {n}")
            rec[i] = (n, kd, o_ops, -1, -1, o_rho, o_work)
            ops.append(heff.ravel())
            o_ops += n * n
            if np.any(dmat != 0):
                rec[i]["dmat"] = o_ops
                ops.append(dmat.ravel())
                o_ops += n * n
            if kd:
                rec[i]["lops"] = o_ops
                for L in dense:
                    ops.append(L.ravel())
                o_ops += kd * n * n
            rhos.append(r0.ravel())
            o_rho += n * n
            o_work += (3 + kd) * n * n
            cost[i] = float(n) ** 3 * (1 + kd)
            self.kd[i] = kd
        self.work_units = float(cost.sum())
        self.rec = rec
        self.ops = torch.as_tensor(np.concatenate(ops).astype(np.complex64), device=self.dev)
        self.rho = torch.as_tensor(np.concatenate(rhos), device=self.dev).contiguous()
        self.work = torch.empty(o_work, dtype=torch.complex64, device=self.dev)
        self.inst = torch.as_tensor(rec.view(np.uint8), device=self.dev)
        order = np.argsort(-cost, kind="stable").astype(np.int32)
        self.order = torch.as_tensor(order, device=self.dev)
        self.queue = torch.zeros(2, dtype=torch.int32, device=self.dev)
        self.max_groups = int(max_groups)
        self.variant = 0
        self.graph = None
        if tune:
            from . import autotune
            autotune.tune_ragged(self, dt_tune, verbose=verbose)

    def variants(self):
        return int(self._lib.lime_ragged_variants())

    @property
    def flops_per_step(self):
        """Canonical algorithmic FLOPs of one RK4 step of the whole batch: 64 n^3 (1 + K_d)."""
        return 64.0 * self.work_units

    def step(self, dt, nsteps, stream=None):
        """Advance every instance by ``nsteps`` RK4 steps: one kernel launch (async)."""
        from . import _lib
        r = _lib.RaggedBatch()
        r.count, r.variant, r.max_groups = self.count, self.variant, self.max_groups
        r.inst, r.order = self.inst.data_ptr(), self.order.data_ptr()
        r.ops, r.rho, r.work, r.queue = (self.ops.data_ptr(), self.rho.data_ptr(),
                                         self.work.data_ptr(), self.queue.data_ptr())
        p = _lib.StepParams()
        p.dt, p.nsteps = float(dt), int(nsteps)
        s = stream if stream is not None else torch.cuda.current_stream(self.dev)
        _lib.check(self._lib.lime_lindblad_rk4_ragged(ctypes.byref(r), ctypes.byref(p),
                                                      ctypes.c_void_p(s.cuda_stream)))

    step_eager = step

    def capture(self, dt, nsteps):
        """Capture one launch of ``nsteps`` steps of every instance into a graph."""
        saved = self.rho.clone()
        self.step(dt, 1)                          # kernel loaded, attributes set
        torch.cuda.synchronize(self.dev)
        self.rho.copy_(saved)
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.step(dt, nsteps)
        self.graph, self.graph_steps, self.graph_dt = g, nsteps, dt
        return g

    def run(self, dt, nsteps, chunk=None):
        """Advance every instance by ``nsteps`` steps by replaying the captured launch."""
        chunk = int(chunk or nsteps)
        if self.graph is None or self.graph_steps != chunk or self.graph_dt != dt:
            self.capture(dt, chunk)
        done = 0
        while done + chunk <= nsteps:
            self.graph.replay()
            done += chunk
        if done < nsteps:
            self.step(dt, nsteps - done)

    def states(self):
        """Current operators as a list in the original instance order (NumPy)."""
        r = self.rho.cpu().numpy()
        return [r[o:o + n * n].reshape(n, n) for n, o in zip(self.rec["n"], self.rec["rho"])]


class RaggedClasses:
    """Size classes on the per-class fused kernels, all captured into one graph."""

    def __init__(self, instances, precision="fp32", tune=True, dt_tune=0.01, verbose=False):
        self.dev = device()
        self.precision = precision
        self.sizes = np.array([np.asarray(i["H"]).shape[0] for i in instances])
        by_n = {}
        for idx, inst in enumerate(instances):
            by_n.setdefault(int(self.sizes[idx]), []).append(idx)
        self.classes = []
        for n, idxs in sorted(by_n.items(), key=lambda kv: -(kv[0] ** 3) * len(kv[1])):
            folded = [fold_host(instances[i]) for i in idxs]
            kd = {len(f[2]) for f in folded}
            if len(kd) != 1:
                raise ValueError(f"instances of n = {n} differ in their number of dense jump operators")
            kd = kd.pop()
            heff = np.stack([f[0] for f in folded])
            dmat = np.stack([f[1] for f in folded])
            lops = np.stack([np.stack(f[2]) for f in folded]) if kd else None
            prop = LindbladPropagator.from_folded(heff, dmat, lops, precision=precision)
            rho = torch.as_tensor(np.stack([np.asarray(instances[i]["rho0"], dtype=np.complex128)
                                            for i in idxs]), device=self.dev).contiguous()
            self.classes.append({"n": n, "idx": idxs, "prop": prop, "rho": rho,
                                 "cost": float(n) ** 3 * (1 + kd) * len(idxs)})
        if tune:
            for c in self.classes:
                autotune.tune(c["prop"], c["rho"], dt_tune, verbose=verbose)
        self.graph = None
        self.graph_steps = None

    @property
    def flops_per_step(self):
        """Canonical algorithmic FLOPs of one RK4 step of the whole batch: 64 n^3 (1 + K_d)."""
        return sum(64.0 * c["n"] ** 3 * (1 + c["prop"].n_dense) * len(c["idx"]) for c in self.classes)

    def _launch_all(self, dt, nsteps):
        main = torch.cuda.current_stream(self.dev)
        for c in self.classes:
            s = c.setdefault("stream", torch.cuda.Stream(self.dev))
            s.wait_stream(main)
            with torch.cuda.stream(s):
                c["prop"].step(c["rho"], dt, nsteps, stream=s)
            main.wait_stream(s)

    def step_eager(self, dt, nsteps):
        self._launch_all(dt, nsteps)

    def capture(self, dt, nsteps):
        """Capture ``nsteps`` RK4 steps of every class into one graph."""
        saved = [c["rho"].clone() for c in self.classes]
        self._launch_all(dt, 1)                  # workspaces and kernel attributes exist
        torch.cuda.synchronize(self.dev)
        for c, r in zip(self.classes, saved):    # the warm-up step is undone
            c["rho"].copy_(r)
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch_all(dt, nsteps)
        self.graph, self.graph_steps, self.graph_dt = g, nsteps, dt
        return g

    def run(self, dt, nsteps, chunk=None):
        """Advance every instance by ``nsteps`` steps by replaying the captured graph."""
        chunk = int(chunk or nsteps)
        if self.graph is None or self.graph_steps != chunk or self.graph_dt != dt:
            self.capture(dt, chunk)
        done = 0
        while done + chunk <= nsteps:
            self.graph.replay()
            done += chunk
        if done < nsteps:
            self.step_eager(dt, nsteps - done)

    def states(self):
        """Current operators as a list in the original instance order (NumPy)."""
        out = [None] * len(self.sizes)
        for c in self.classes:
            r = c["rho"].cpu().numpy()
            for k, i in enumerate(c["idx"]):
                out[i] = r[k]
        return out
