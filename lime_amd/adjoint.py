"""Exact discrete adjoint of the RK4 Lindblad/QRT propagation, behind torch.autograd.

lime has no gradients; this module differentiates the linear-absorption pipeline of the
chain model (site energies eps, nearest-neighbour couplings J, one shared pure-dephasing
rate gamma) through the same RK4 map the forward kernels apply.

Model (Hilbert space = ground state |g> = index 0 + n_sites sites):
    H = sum_j eps_j |j><j| + J_j (|j><j+1| + h.c.),  L_j = sqrt(gamma) |j><j|  (sites only),
    mu = sum_j mu_j (|j><g| + |g><j|),  X_0 = mu |g><g|,  C_k = Tr[mu X_k],  k = 1..Nt.

Forward: the fused kernel (resident for n <= 80, tiled above; launch shape from the
autotuner) records C_k every step, K steps per launch; the state after each launch is the
checkpoint X_{cK} (default K = 64).

Backward, for a real loss with torch's gradient G_k = dL/dC_k^* convention
(dL = Re sum_k conj(G_k) dC_k): the source of step k is g_k = G_k mu^dag, and with the RK4
stages s_1 = X, s_2 = X + h/2 f(s_1), s_3 = X + h/2 f(s_2), s_4 = X + h f(s_3) of step k
(recomputed from the checkpoint of their segment on the per-stage kernels, which record
s_2, s_3, s_4 in FP64 as the step computes them) and lambda' = lambda_{k+1}:
    kappa_4 = h/6 lambda',               sigma_4 = f^dag(kappa_4)
    kappa_3 = h/3 lambda' + h sigma_4,   sigma_3 = f^dag(kappa_3)
    kappa_2 = h/3 lambda' + h/2 sigma_3, sigma_2 = f^dag(kappa_2)
    kappa_1 = h/6 lambda' + h/2 sigma_2, sigma_1 = f^dag(kappa_1)
    lambda_k = lambda' + sigma_1 + sigma_2 + sigma_3 + sigma_4 + g_k
    dL/dtheta += Re sum_i <kappa_i, (df/dtheta) s_i>.
f^dag is the generator with H_eff -> -H_eff^dag, D -> conj(D): the SAME kernels
(``lime_lindblad_apply``); the contractions run in ``lime_adjoint_contract``.
"""
import ctypes

import numpy as np
import torch

from .propagator import LindbladPropagator, device


def chain_generator(eps, J, gamma):
    """H_eff, D (complex128, (B, n, n)) of the chain + ground model on the GPU.

    gamma: a scalar (one rate shared by every site and chain), (B,) (one rate per chain) or
    (1, ns) (one rate per site, shared by the chains)."""
    B, ns = eps.shape
    n = ns + 1
    dev = eps.device
    H = torch.zeros((B, n, n), dtype=torch.complex128, device=dev)
    idx = torch.arange(1, n, device=dev)
    H[:, idx, idx] = eps.to(torch.complex128)
    H[:, idx[:-1], idx[1:]] = J.to(torch.complex128)
    H[:, idx[1:], idx[:-1]] = J.to(torch.complex128)
    g = torch.as_tensor(gamma, dtype=torch.float64, device=dev)
    rates = torch.zeros((B, n), dtype=torch.float64, device=dev)
    if g.dim() == 2:
        if g.shape != (1, ns):
            raise ValueError(f"per-site dephasing rates must have shape (1, {ns}), got {tuple(g.shape)}")
        rates[:, 1:] = g
    else:
        rates[:, 1:] = g.reshape(-1, 1) if g.dim() == 1 else g
    heff = H - 0.5j * torch.diag_embed(rates.to(torch.complex128))
    dmat = torch.diag_embed(rates.to(torch.complex128))
    return heff.contiguous(), dmat.contiguous()


def dipole_operator(mu_sites, n):
    mu = torch.zeros((n, n), dtype=torch.complex128, device=mu_sites.device)
    mu[1:, 0] = mu_sites.to(torch.complex128)
    mu[0, 1:] = mu_sites.to(torch.complex128)
    return mu


class ChainQRT:
    """Forward propagation with checkpoints and the adjoint backward pass."""

    def __init__(self, eps, J, gamma, mu_sites, dt, nt, checkpoint=64, force_tiled=False,
                 precision="fp32", tune=True, backward_precision=None):
        """``precision``: arithmetic of the forward pass and of the recomputation of the
        stage operands; ``backward_precision`` (default: the same): arithmetic of the
        adjoint generator applications f^dag ('fp32' or 'fp64')."""
        self.dev = device()
        self.eps = eps.to(self.dev, torch.float64)
        self.J = J.to(self.dev, torch.float64)
        self.gamma = gamma
        self.B, self.ns = self.eps.shape
        self.n = self.ns + 1
        self.dt, self.nt, self.K = float(dt), int(nt), int(checkpoint)
        heff, dmat = chain_generator(self.eps, self.J, gamma)
        self.prop = LindbladPropagator.from_folded(heff, dmat, force_tiled=force_tiled,
                                                   precision=precision)
        self.x0 = torch.zeros((self.B, self.n, self.n), dtype=torch.complex128, device=self.dev)
        if tune:
            from . import autotune
            autotune.tune(self.prop, self.x0, self.dt)
        bp = backward_precision or precision
        self.adj = (self.prop if bp == precision else
                    LindbladPropagator.from_folded(heff, dmat, force_tiled=True, precision=bp)).adjoint()
        # the backward recompute records the stage operands on the per-stage kernels
        self.rec_prop = self.prop.tiled()
        self.mu = dipole_operator(torch.as_tensor(mu_sites, device=self.dev), self.n)
        self.x0[:, :, 0] = self.mu[:, 0]

    def forward(self):
        """C (B, Nt) complex128 and the checkpoints X_0, X_K, X_2K, ..."""
        B, n = self.B, self.n
        X = self.x0.clone()
        E_t = self.mu.transpose(0, 1).contiguous().reshape(1, 1, n, n)
        obs = torch.zeros((B, self.nt, 1), dtype=torch.complex128, device=self.dev)
        ncp = self.nt // self.K
        cps = torch.empty((B, ncp, n, n), dtype=torch.complex128, device=self.dev) if ncp else None
        # one pass: K steps per launch (C(t) recorded every step), the state after each
        # launch is the next checkpoint
        done = 0
        while done < self.nt:
            m = min(self.K, self.nt - done)
            self.prop.step(X, self.dt, m, eops_t=E_t, eop_group=B, obs=obs, nout=1,
                           step_offset=done, rec_capacity=self.nt)
            done += m
            if m == self.K and done // self.K <= ncp:
                cps[:, done // self.K - 1].copy_(X)
        self.checkpoints = cps
        return obs[:, :, 0]

    def _segment_states(self, c0, length):
        """Recompute steps c0 .. c0+length-1 from the checkpoint at step c0.

        Returns X (length, B, N, N) with X[j] = X_{c0+j} (record-major, so each step's
        states are one contiguous block) and the FP64 stage operands (length, 3, B, N, N) of
        the step leaving X_{c0+j}: X + h/2 k1, X + h/2 k2, X + h k3 -- the values the forward
        step used, recorded by the per-stage kernels."""
        B, n = self.B, self.n
        Xc = self.x0.clone() if c0 == 0 else self.checkpoints[:, c0 // self.K - 1].clone().contiguous()
        out = torch.empty((length, B, n, n), dtype=torch.complex128, device=self.dev)
        out[0] = Xc
        stages = torch.empty((length, 3, B, n, n), dtype=torch.complex128, device=self.dev)
        st = torch.empty((B, length, n, n), dtype=torch.complex128, device=self.dev)
        self.rec_prop.step(Xc, self.dt, length, states=st, stage_states=stages, nout=1,
                           step_offset=0, rec_capacity=length)
        if length > 1:
            out[1:] = st[:, :length - 1].transpose(0, 1)
        return out, stages

    def backward(self, G):
        """G (B, Nt) complex: torch's dL/dC^*.  Returns (g_eps (B,ns), g_J (B,ns-1),
        g_gamma_sites (B,ns): the gradient w.r.t. each site's dephasing rate in each chain)."""
        B, n, h = self.B, self.n, self.dt
        dev = self.dev
        mu_h = self.mu.conj().transpose(0, 1).contiguous()
        Gt = G.to(dev, torch.complex128).t().contiguous()     # (Nt, B): Gt[k] contiguous
        lam = torch.zeros((B, n, n), dtype=torch.complex128, device=dev)
        outa = torch.zeros((B, n), dtype=torch.complex128, device=dev)
        outb = torch.zeros((B, n), dtype=torch.complex128, device=dev)
        outu = torch.zeros((B, max(n - 1, 1)), dtype=torch.complex128, device=dev)
        c128 = lambda: torch.empty((B, n, n), dtype=torch.complex128, device=dev)
        opd = self.adj.operand_dtype
        c64 = lambda: torch.empty((B, n, n), dtype=opd, device=dev)
        k3, k2, k1 = c128(), c128(), c128()
        k3l, k2l, k1l = c64(), c64(), c64()
        mu_h = mu_h.contiguous()
        lib = self.adj._lib
        from . import _lib as L
        stream = torch.cuda.current_stream(dev)
        sp = stream.cuda_stream
        acc = c128()
        k4 = c128()
        k4l = c64()
        flags = L.LIME_F_PREC_FP64 if opd == torch.complex128 else L.LIME_F_PREC_FP32
        pairs = ctypes.c_void_p * 4
        nseg = (self.nt + self.K - 1) // self.K
        for seg in range(nseg - 1, -1, -1):
            c0 = seg * self.K
            length = min(self.K, self.nt - c0)
            Xs, stages = self._segment_states(c0, length)
            for k in range(c0 + length - 1, c0 - 1, -1):
                # lambda' = lambda_{k+1} + source of record k+1 (C_{k+1} = Tr[mu X_{k+1}]);
                # kappa_4 = h/6 lambda' (FP64 and operand precision); acc = lambda'
                L.check(lib.lime_adjoint_seed(B, n, flags, lam.data_ptr(), Gt[k].data_ptr(),
                                              mu_h.data_ptr(), h / 6.0, k4.data_ptr(), k4l.data_ptr(),
                                              acc.data_ptr(), sp))
                X = Xs[k - c0]
                s2, s3, s4 = stages[k - c0]                      # contiguous views
                self.adj.apply(k4l, z=lam, y=k3, y64=k3l, acc=acc, c_lx=h, c_z=h / 3.0, c_acc=1.0)
                self.adj.apply(k3l, z=lam, y=k2, y64=k2l, acc=acc, c_lx=h / 2, c_z=h / 3.0, c_acc=1.0)
                self.adj.apply(k2l, z=lam, y=k1, y64=k1l, acc=acc, c_lx=h / 2, c_z=h / 6.0, c_acc=1.0)
                self.adj.apply(k1l, acc=acc, c_acc=1.0)
                # the four stage contractions in one launch
                kap = pairs(k1.data_ptr(), k2.data_ptr(), k3.data_ptr(), k4.data_ptr())
                sts = pairs(X.data_ptr(), s2.data_ptr(), s3.data_ptr(), s4.data_ptr())
                L.check(lib.lime_adjoint_contract_pairs(B, n, 4, kap, sts, 1.0, outa.data_ptr(),
                                                        outb.data_ptr(), outu.data_ptr(), sp))
                lam, acc = acc, lam                      # lambda_k = acc; the old buffer is reused
        g_eps = outa[:, 1:].imag.clone()
        g_J = outu[:, 1:].imag.clone()
        g_gamma_sites = outb[:, 1:].real.clone()          # d/d(dephasing rate of site j), per chain
        return g_eps, g_J, g_gamma_sites


class ChainAbsorption(torch.autograd.Function):
    """C(t) of the chain + ground model as a differentiable function of (eps, J, gamma).

    forward(eps (B, ns), J (B, ns-1), gamma, mu_sites (ns,), dt, nt, K, precision) -> C (B, nt)
    complex128.  gamma is a scalar (one shared rate: its gradient sums every site of every
    chain), (B,) (a rate per chain) or (1, ns) (a rate per site, shared by the chains).
    precision 'fp32' runs the FP32-MFMA kernels, 'fp64' the FP64-MFMA tiled kernels for both
    passes, 'fp32/fp64' the FP32-MFMA forward and recomputation with FP64-MFMA adjoint
    applications.
    """

    @staticmethod
    def forward(ctx, eps, J, gamma, mu_sites, dt, nt, checkpoint=64, precision="fp32"):
        fwd, _, bwd = precision.partition("/")
        q = ChainQRT(eps.detach(), J.detach(), gamma.detach(), mu_sites, dt, nt, checkpoint,
                     precision=fwd, backward_precision=bwd or fwd)
        C = q.forward()
        ctx.q = q
        ctx.gamma_shape = gamma.shape
        return C

    @staticmethod
    def backward(ctx, grad_C):
        g_eps, g_J, g_sites = ctx.q.backward(grad_C)
        shape = ctx.gamma_shape
        if len(shape) == 0:
            gg = g_sites.sum()
        elif len(shape) == 1:
            gg = g_sites.sum(dim=1)
        else:
            gg = g_sites.sum(dim=0, keepdim=True)
        return g_eps, g_J, gg.reshape(shape), None, None, None, None, None


def spectrum_loss(C, target, dt):
    """L2 mismatch of |FFT| spectra (lime's e^{-i w t} convention, dt scaling)."""
    S = torch.fft.fft(C, dim=-1) * dt
    St = torch.fft.fft(target, dim=-1) * dt
    return ((S.real - St.real) ** 2).sum(dim=-1)
