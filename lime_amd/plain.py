"""Plain FP64 GPU propagator: the same RK4 / H_eff-folded algorithm in complex128 with
batched torch matmuls (rocBLAS zgemm) and separate elementwise kernels -- no fusion, no
persistence.  It is the comparison point ("plain FP64 on GPU") for the fused kernel and
an FP64 GPU reference for sizes the CPU oracle cannot reach.

Per right-hand side: H_eff X and X H_eff^dag (2 zgemm), D o X (elementwise), and for each
dense jump operator (L X) L^dag (2 zgemm) -- the same algorithmic FLOP count as the fused
kernel.  lime's own operation sequence (2 + 5K products per RHS, lime/oqs.py:706-723) is
``rhs_reference_sequence``.
"""
import torch

from .propagator import as_c128, device, fold_jump_operators


class PlainFP64Propagator:
    def __init__(self, H, c_ops=None, group=1):
        self.dev = device()
        heff, dmat, lops = fold_jump_operators(H, c_ops, self.dev)
        self.group = int(group)
        rep = lambda t: t.repeat_interleave(self.group, dim=0).contiguous() if t is not None else None
        self.heff = rep(heff)
        self.heff_h = rep(heff.conj().transpose(-1, -2).contiguous())
        self.dmat = rep(dmat)
        self.lops = rep(lops)
        self.lops_h = rep(lops.conj().transpose(-1, -2).contiguous()) if lops is not None else None
        self.n = heff.shape[-1]

    def rhs(self, X):
        k = torch.bmm(self.heff, X)
        k = k - torch.bmm(X, self.heff_h)
        k = k * (-1j)
        if self.dmat is not None:
            k = k + self.dmat * X
        if self.lops is not None:
            for j in range(self.lops.shape[1]):
                T = torch.bmm(self.lops[:, j].contiguous(), X)
                k = k + torch.bmm(T, self.lops_h[:, j].contiguous())
        return k

    def step(self, X, dt):
        k1 = self.rhs(X)
        k2 = self.rhs(X + (0.5 * dt) * k1)
        k3 = self.rhs(X + (0.5 * dt) * k2)
        k4 = self.rhs(X + dt * k3)
        return X + (dt / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)

    def run(self, X0, dt, nsteps, e_ops_t=None, nout=1):
        """X0 (B,N,N); e_ops_t (B or 1, n_e, N, N) transposed observables. Returns X, obs."""
        X = as_c128(X0, self.dev).clone()
        obs = []
        for s in range(nsteps):
            X = self.step(X, dt)
            if e_ops_t is not None and (s + 1) % nout == 0:
                obs.append((e_ops_t * X.unsqueeze(1)).sum(dim=(-1, -2)))   # Tr[E X] = sum E^T o X
        return X, (torch.stack(obs, dim=1) if obs else None)


def rhs_reference_sequence(X, H, c_ops):
    """lime's liouvillian (oqs.py:706-723) on batched complex128 tensors: 2 + 5K products."""
    k = -1j * (torch.matmul(H, X) - torch.matmul(X, H))
    for L in c_ops:
        Lh = L.conj().transpose(-1, -2)
        LhL = torch.matmul(Lh, L)
        k = k + torch.matmul(L, torch.matmul(X, Lh)) - 0.5 * (torch.matmul(LhL, X) + torch.matmul(X, LhL))
    return k
