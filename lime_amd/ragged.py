"""Ragged batches: many independent Lindblad systems of uneven dimension in one graph.

lime propagates one system at a time (oqs.py:1590-1688).  ``RaggedLindblad`` takes a list
of instances of different n, groups them into size classes, and advances all of them
together:

* each class is one operator batch on the fused kernels -- resident (one workgroup per
  operator, all steps in one launch) when it fits, otherwise the per-stage tiled kernels;
* every class's launch shape comes from the autotuner (``lime_amd.autotune``);
* the launches of all classes are captured into ONE hipGraph (torch.cuda.CUDAGraph) in
  which each class runs on its own forked stream, largest cost first, so small classes
  fill the CUs the large ones leave idle; ``run`` replays that graph.

An instance is a dict with ``H`` (n, n), ``rho0`` (n, n), optional ``c_ops`` (list of
(n, n) jump operators) and optional ``dephasing`` (scalar or length-n rates gamma_i of
L_i = sqrt(gamma_i)|i><i|, given compactly instead of n dense matrices).
"""
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
