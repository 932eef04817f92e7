"""Batched Lindblad RK4 propagation on the GPU through liblime_hip.so.

``LindbladPropagator`` owns the device-side operator sets of a batch (H_eff, the
Hadamard dissipator D, the dense jump operators) and advances complex128 operators
``rho`` (density matrices or quantum-regression operators) with the fused kernel
``lime_lindblad_rk4``.  It is the engine behind the lime-compatible facade in
``lime_amd.oqs``; it replaces the per-step NumPy loop of lime's ``_lindblad``
(lime/oqs.py:1590-1688) and of the QRT drivers (oqs.py:726-800, 1227-1331).

Right-hand side (all jump operators folded):
    f(X) = -i (H_eff X - X H_eff^dag) + D o X + sum_dense L X L^dag,
    H_eff = H - (i/2) sum_all L^dag L,  D = sum_{diagonal L = diag(l)} l l^dag.
"""
import ctypes

import numpy as np
import torch

from . import _lib

_PREC_FLAGS = {"fp32": _lib.LIME_F_PREC_FP32, "3xbf16": _lib.LIME_F_PREC_3XBF16,
               "fp64": _lib.LIME_F_PREC_FP64}


def device():
    """The GPU the library runs on; raises when none is visible (no CPU fallback)."""
    if not torch.cuda.is_available():
        raise RuntimeError("lime_amd needs an AMD Instinct GPU (ROCm): no GPU is visible")
    return torch.device("cuda", torch.cuda.current_device())


def as_c128(x, dev):
    """Array-like or tensor -> contiguous complex128 tensor on ``dev``."""
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev, dtype=torch.complex128)
    else:
        if hasattr(x, "toarray"):          # scipy.sparse
            x = x.toarray()
        t = torch.as_tensor(np.asarray(x, dtype=np.complex128), device=dev)
    return t.contiguous()


def _check(t, what, shape, dtype, dev):
    """Every buffer handed to the C ABI is indexed by the kernels from the descriptor's
    sizes alone: its shape, dtype, layout and device are checked here, on the host."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch tensor")
    if t.dtype != dtype or not t.is_contiguous() or t.device != dev:
        raise ValueError(f"{what} must be a contiguous {dtype} tensor on {dev}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what} has shape {tuple(t.shape)}, expected {tuple(shape)}")


def _stack_sets(x, nsets, n, dev, what):
    t = as_c128(x, dev)
    if t.dim() == 2:
        t = t.unsqueeze(0)
    if t.shape[-2:] != (n, n):
        raise ValueError(f"{what} has shape {tuple(t.shape)}, expected (..., {n}, {n})")
    if t.shape[0] == 1 and nsets > 1:
        t = t.expand(nsets, n, n)
    if t.shape[0] != nsets:
        raise ValueError(f"{what} has {t.shape[0]} sets, expected {nsets}")
    return t.contiguous()


def fold_jump_operators(H, c_ops, dev=None):
    """Split jump operators into H_eff, the Hadamard dissipator D and dense operators.

    H: (S, N, N) or (N, N); c_ops: list of (S, N, N) or (N, N) operators (shared).
    Returns complex128 device tensors heff (S,N,N), dmat (S,N,N) or None,
    lops (S,Kd,N,N) or None.  A jump operator is *diagonal* when every off-diagonal
    entry of every set is exactly zero.
    """
    dev = dev or device()
    Ht = as_c128(H, dev)
    if Ht.dim() == 2:
        Ht = Ht.unsqueeze(0)
    if Ht.dim() != 3 or Ht.shape[-1] != Ht.shape[-2]:
        raise ValueError(f"H must be (N, N) or (S, N, N), got {tuple(Ht.shape)}")
    S, n = Ht.shape[0], Ht.shape[-1]
    gamma = torch.zeros_like(Ht)
    dmat = None
    dense = []
    for k, c in enumerate(c_ops or []):
        L = _stack_sets(c, S, n, dev, f"c_ops[{k}]")
        gamma = gamma + torch.matmul(L.conj().transpose(-1, -2), L)
        diag = torch.diagonal(L, dim1=-2, dim2=-1)
        off = L - torch.diag_embed(diag)
        if not bool(torch.any(off != 0)):
            term = diag.unsqueeze(-1) * diag.conj().unsqueeze(-2)
            dmat = term if dmat is None else dmat + term
        else:
            dense.append(L)
    heff = (Ht - 0.5j * gamma).contiguous()
    lops = torch.stack(dense, dim=1).contiguous() if dense else None
    return heff, (dmat.contiguous() if dmat is not None else None), lops


class LindbladPropagator:
    """Operator sets for a batch and the kernel call that advances it.

    Parameters
    ----------
    H : (N, N) or (S, N, N) array-like / tensor
    c_ops : list of (N, N) or (S, N, N) jump operators (may be empty/None)
    group : propagated operators per operator set (operator b uses set b // group)
    precision : 'fp32' (FP32 MFMA, default), '3xbf16', 'fp64'
    """

    def __init__(self, H, c_ops=None, group=1, precision="fp32", sandwich=None,
                 force_tiled=False, variant=0):
        """``sandwich``: optional explicit (L, R) lists of (S,N,N) operators contributing
        sum_k L_k X R_k^dag (used instead of folding dense jump operators; H must then
        already be the effective generator).  ``force_tiled`` selects the per-stage tiled
        kernels even for n <= 80."""
        self.dev = device()
        self.force_tiled = bool(force_tiled)
        self.one_sided = False
        self.variant = int(variant)
        self._work = None
        if precision not in _PREC_FLAGS:
            raise ValueError(f"unknown precision {precision!r}; expected one of {sorted(_PREC_FLAGS)}")
        self.precision = precision
        heff, dmat, lops = fold_jump_operators(H, c_ops, self.dev)
        self.nsets, self.n = heff.shape[0], heff.shape[-1]
        self.group = int(group)
        if self.group < 1:
            raise ValueError("group must be >= 1")
        cdt = torch.complex128 if precision == "fp64" else torch.complex64
        self.heff_c128 = heff
        self.heff = heff.to(cdt).contiguous()
        self.dmat = dmat.to(cdt).contiguous() if dmat is not None else None
        self.lops = lops.to(cdt).contiguous() if lops is not None else None
        self.rops = None
        if sandwich is not None:
            Ls, Rs = sandwich
            if len(Ls) != len(Rs):
                raise ValueError("sandwich needs as many left as right factors")
            if self.lops is not None:
                raise ValueError("sandwich terms cannot be combined with dense jump operators")
            if len(Ls):
                self.lops = torch.stack([_stack_sets(x, self.nsets, self.n, self.dev, "L")
                                         for x in Ls], dim=1).to(cdt).contiguous()
                self.rops = torch.stack([_stack_sets(x, self.nsets, self.n, self.dev, "R")
                                         for x in Rs], dim=1).to(cdt).contiguous()
        self.n_dense = 0 if self.lops is None else self.lops.shape[1]
        self._lib = _lib.load()

    @classmethod
    def from_folded(cls, heff, dmat=None, lops=None, rops=None, group=1, precision="fp32",
                    force_tiled=False, one_sided=False):
        """Build directly from H_eff (S,N,N), D (S,N,N), dense L / R (S,K,N,N) tensors.

        ``one_sided``: drop the X H_eff^dag product, f(X) = -i H_eff X + D o X + sum L X R^dag
        -- with H_eff = i G this steps dX/dt = G X (a Liouville-space generator acting on
        the columns of X; per-stage kernels)."""
        self = object.__new__(cls)
        self.dev = device()
        self.precision = precision
        self.one_sided = bool(one_sided)
        self.force_tiled = bool(force_tiled) or self.one_sided
        self.variant = 0
        self._work = None
        self._lib = _lib.load()
        if precision not in _PREC_FLAGS:
            raise ValueError(f"unknown precision {precision!r}; expected one of {sorted(_PREC_FLAGS)}")
        cdt = torch.complex128 if precision == "fp64" else torch.complex64
        heff = as_c128(heff, self.dev)
        if heff.dim() == 2:
            heff = heff.unsqueeze(0)
        if heff.dim() != 3 or heff.shape[-1] != heff.shape[-2]:
            raise ValueError(f"heff must be (N, N) or (S, N, N), got {tuple(heff.shape)}")
        self.nsets, self.n = heff.shape[0], heff.shape[-1]
        self.group = int(group)
        if self.group < 1:
            raise ValueError("group must be >= 1")
        self.heff_c128 = heff
        self.heff = heff.to(cdt).contiguous()
        self.dmat = _stack_sets(dmat, self.nsets, self.n, self.dev, "dmat").to(cdt).contiguous() \
            if dmat is not None else None
        self.lops = self._op_stack(lops, "lops", cdt)
        self.rops = self._op_stack(rops, "rops", cdt)
        if self.rops is not None and (self.lops is None or self.rops.shape != self.lops.shape):
            raise ValueError("rops needs lops of the same shape")
        self.n_dense = 0 if self.lops is None else self.lops.shape[1]
        return self

    def _op_stack(self, x, what, cdt):
        if x is None:
            return None
        t = as_c128(x, self.dev)
        if t.dim() != 4 or t.shape[0] != self.nsets or t.shape[-2:] != (self.n, self.n):
            raise ValueError(f"{what} has shape {tuple(t.shape)}, expected ({self.nsets}, K, {self.n}, {self.n})")
        return t.to(cdt).contiguous()

    @property
    def batch(self):
        return self.nsets * self.group

    def _resident(self):
        """Does ``step`` run the resident kernel (whose ``variant`` is a launch shape, not
        the tile edge of the per-stage kernels)?"""
        return (not self.force_tiled and not self.one_sided and self.precision == "fp32"
                and self.n <= self.max_n())

    def _tile_variant(self):
        """The ``variant`` field for the per-stage kernels (``apply`` always runs them): a
        tuned tile choice carries over, a resident launch shape does not."""
        return 0 if self._resident() else self.variant

    def tiled(self):
        """The same operator batch on the per-stage kernels (shares the device operators)."""
        t = object.__new__(LindbladPropagator)
        t.__dict__.update(self.__dict__)
        t._work = None
        t.variant = self._tile_variant()
        t.force_tiled = True
        return t

    def adjoint(self):
        """The propagator of the adjoint generator f^dag (same batch layout):
        H_eff -> -H_eff^dag, D -> conj(D), (L, R) -> (L^dag, R^dag)."""
        adj = object.__new__(LindbladPropagator)
        adj.__dict__.update(self.__dict__)
        adj._work = None
        adj.variant = self._tile_variant()
        adj.force_tiled = True
        h = lambda t: t.conj().transpose(-1, -2).contiguous()
        adj.heff = (-h(self.heff)).contiguous()
        adj.heff_c128 = (-h(self.heff_c128)).contiguous()
        adj.dmat = self.dmat.conj().contiguous() if self.dmat is not None else None
        adj.lops = h(self.lops) if self.lops is not None else None
        # R = L (rops None) stays implicit: the adjoint's right factors are adj.lops too
        adj.rops = h(self.rops) if self.rops is not None else None
        return adj

    def _flags(self):
        return _PREC_FLAGS[self.precision] | (_lib.LIME_F_ONE_SIDED if self.one_sided else 0)

    def _descriptor(self, B):
        d = _lib.LindbladBatch()
        d.batch, d.n, d.group, d.n_dense = B, self.n, self.group, self.n_dense
        d.eop_group = B
        d.flags = self._flags()
        d.variant = self.variant
        d.heff = self.heff.data_ptr()
        d.dmat = self.dmat.data_ptr() if self.dmat is not None else None
        d.lops = self.lops.data_ptr() if self.lops is not None else None
        d.rops = self.rops.data_ptr() if self.rops is not None else None
        return d

    @property
    def operand_dtype(self):
        return torch.complex128 if self.precision == "fp64" else torch.complex64

    def apply(self, x64, *, z=None, y=None, y64=None, acc=None, c_lx=1.0, c_z=1.0, c_acc=1.0,
              stream=None):
        """One generator application V = f(x) (x in the operand precision) with the epilogue
        y = c_lx V + c_z z, y64 = y in the operand precision, acc += c_acc V (all optional)."""
        B, n = self.batch, self.n
        op_dtype = self.operand_dtype
        _check(x64, "x (operand precision)", (B, n, n), op_dtype, self.dev)
        for name, t in (("z", z), ("y", y), ("acc", acc)):
            if t is not None:
                _check(t, name, (B, n, n), torch.complex128, self.dev)
        if y64 is not None:
            _check(y64, "y64 (operand precision)", (B, n, n), op_dtype, self.dev)
        d = self._descriptor(B)
        d.variant = self._tile_variant()
        if self.n_dense > 0:
            need = B * self.n_dense * self.n * self.n * (16 if self.precision == "fp64" else 8)
            if self._work is None or self._work.numel() < need:
                self._work = torch.empty(need, dtype=torch.uint8, device=self.dev)
            d.work, d.work_bytes = self._work.data_ptr(), self._work.numel()
        pair = lambda c: (ctypes.c_double * 2)(complex(c).real, complex(c).imag)
        s = stream if stream is not None else torch.cuda.current_stream(self.dev)
        ptr = lambda t: t.data_ptr() if t is not None else None
        _lib.check(self._lib.lime_lindblad_apply(
            ctypes.byref(d), x64.data_ptr(), ptr(z), ptr(y), ptr(y64), ptr(acc),
            pair(c_lx), pair(c_z), pair(c_acc), ctypes.c_void_p(s.cuda_stream)))

    def variants(self, n_eops=0):
        """Number of launch-shape variants the library offers for this batch."""
        d = self._descriptor(self.batch)
        d.n_eops = n_eops
        d.flags |= (_lib.LIME_F_FORCE_TILED if self.force_tiled else 0)
        return int(self._lib.lime_lindblad_variants(ctypes.byref(d)))

    def max_n(self):
        return self._lib.lime_lindblad_max_n(self.n_dense, _PREC_FLAGS[self.precision])

    def step(self, rho, dt, nsteps, *, eops_t=None, eop_group=None, obs=None, states=None,
             nout=1, step_offset=0, rec_capacity=0, renorm=False, stream=None, stage_states=None):
        """Advance ``rho`` (B, N, N) complex128 in place by ``nsteps`` RK4 steps (async).

        ``stage_states`` (rec_capacity, 3, B, N, N) complex128 (record-major) receives the
        FP64 stage operands X + h/2 k1, X + h/2 k2, X + h k3 of every recorded step
        (per-stage kernels: the propagator must be ``force_tiled`` for n <= 80)."""
        n, B = self.n, self.batch
        _check(rho, "rho", (B, n, n), torch.complex128, self.dev)
        eop_group = int(eop_group or B)
        if eop_group < 1:
            raise ValueError("eop_group must be >= 1")
        n_e = 0
        if eops_t is not None:
            if eops_t.dim() != 4:
                raise ValueError("eops_t must be (E_sets, n_e, N, N)")
            n_e = int(eops_t.shape[1])
            _check(eops_t, "eops_t", (-(-B // eop_group), n_e, n, n), torch.complex128, self.dev)
        if obs is not None:
            if eops_t is None:
                raise ValueError("obs needs eops_t")
            _check(obs, "obs", (B, int(rec_capacity), n_e), torch.complex128, self.dev)
        if states is not None:
            _check(states, "states", (B, int(rec_capacity), n, n), torch.complex128, self.dev)
        if stage_states is not None:
            _check(stage_states, "stage_states", (int(rec_capacity), 3, B, n, n), torch.complex128, self.dev)
        d = _lib.LindbladBatch()
        d.batch, d.n, d.group, d.n_dense = B, n, self.group, self.n_dense
        d.n_eops = n_e
        d.eop_group = eop_group
        d.flags = self._flags() | (_lib.LIME_F_RENORM_TRACE if renorm else 0) \
            | (_lib.LIME_F_FORCE_TILED if self.force_tiled else 0)
        d.variant = self.variant
        d.heff = self.heff.data_ptr()
        d.dmat = self.dmat.data_ptr() if self.dmat is not None else None
        d.lops = self.lops.data_ptr() if self.lops is not None else None
        d.eops_t = eops_t.data_ptr() if eops_t is not None else None
        d.rho = rho.data_ptr()
        d.obs = obs.data_ptr() if obs is not None else None
        d.states = states.data_ptr() if states is not None else None
        d.rops = self.rops.data_ptr() if self.rops is not None else None
        d.stage_states = stage_states.data_ptr() if stage_states is not None else None
        need = int(self._lib.lime_lindblad_workspace_bytes(ctypes.byref(d)))
        if need > 0:
            if self._work is None or self._work.numel() < need:
                self._work = torch.empty(need, dtype=torch.uint8, device=self.dev)
            d.work = self._work.data_ptr()
            d.work_bytes = self._work.numel()
        p = _lib.StepParams()
        p.dt, p.nsteps, p.nout = float(dt), int(nsteps), int(nout)
        p.step_offset, p.rec_capacity = int(step_offset), int(rec_capacity)
        s = stream if stream is not None else torch.cuda.current_stream(self.dev)
        _lib.check(self._lib.lime_lindblad_rk4(ctypes.byref(d), ctypes.byref(p),
                                               ctypes.c_void_p(s.cuda_stream)))

    def run(self, rho0, dt, nsteps, e_ops=None, nout=1, store_states=False, renorm=False,
            eop_group=None, chunk=None):
        """Propagate ``rho0`` for ``nsteps`` steps, recording after every ``nout`` steps.

        rho0 : (N, N) (broadcast to the batch) or (B, N, N)
        e_ops : list of (N, N) shared observables, or a tensor (E_sets, n_e, N, N) with
                observable set b // eop_group for operator b.
        Returns (rho (B,N,N), obs (B, Nt//nout, n_e) or None, states (B, Nt//nout, N, N) or None),
        all complex128 device tensors.  Records sit at t_r = (r + 1) * nout * dt.
        """
        n, B = self.n, self.batch
        rho = as_c128(rho0, self.dev)
        if rho.dim() == 2:
            rho = rho.unsqueeze(0).expand(B, n, n)
        rho = rho.clone().contiguous()           # the caller's rho0 is never mutated
        eops_t = None
        if e_ops is not None and len(e_ops) > 0:
            if isinstance(e_ops, torch.Tensor) and e_ops.dim() == 4:
                E = as_c128(e_ops, self.dev)
            else:
                E = torch.stack([as_c128(e, self.dev) for e in e_ops]).unsqueeze(0)
            eops_t = E.transpose(-1, -2).contiguous()
            if eop_group is None:
                eop_group = B // E.shape[0]
        nrec = nsteps // nout if nout > 0 else 0
        obs = None
        if eops_t is not None and nrec > 0:
            obs = torch.zeros((B, nrec, eops_t.shape[1]), dtype=torch.complex128, device=self.dev)
        states = None
        if store_states and nrec > 0:
            states = torch.empty((B, nrec, n, n), dtype=torch.complex128, device=self.dev)
        chunk = int(chunk or nsteps)
        done = 0
        while done < nsteps:
            m = min(chunk, nsteps - done)
            self.step(rho, dt, m, eops_t=eops_t, eop_group=eop_group, obs=obs, states=states,
                      nout=nout, step_offset=done, rec_capacity=nrec, renorm=renorm)
            done += m
        return rho, obs, states
