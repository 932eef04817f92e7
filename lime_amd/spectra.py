"""Linear absorption spectra (lime/signal/liouville.py:27-127) and the time-domain route.

* ``linear_absorption(omegas, ham, dip, rho0, c_ops, ntrans)`` -- lime's eigen-series
  spectrum: L' = op2sop(H) + i sum D[L] (= i times the Liouvillian),
  S(w) = -2 Im sum_n <<mu|U1_n>> <<U2_n|mu rho0>> / norm_n / (w - lambda_n) over the ntrans
  eigenvalues with the largest real part ('LR', as lime's ARPACK call).  Here L' is
  diagonalised densely on the GPU and the left eigenvectors come from inv(U1), so the
  left/right pairing is exact (lime pairs two separate ARPACK runs by sorting).
* ``absorption_eseries(omegas, L, edip, rho0, ntrans)`` -- the same from a given L'.
* ``absorption_from_correlation(C, t)`` -- the QRT route: the spectrum of <mu(t) mu(0)>
  with lime's FFT convention (lime/fft.py:9-56); ``C`` may be a batch (inhomogeneous
  average over its first axis).
"""
import numpy as np
import torch

from . import fft as _fft
from .propagator import as_c128, device
from .superoperator import liouvillian_dense


def _dense(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return a.toarray() if hasattr(a, "toarray") else np.asarray(a)


def absorption_eseries(omegas, L, edip, rho0, ntrans=1):
    dev = device()
    Lt = as_c128(_dense(L), dev)
    w, U1 = torch.linalg.eig(Lt)
    order = torch.argsort(w.real, descending=True)[:ntrans]
    mu = as_c128(_dense(edip), dev)
    r0 = as_c128(_dense(rho0), dev)
    I = torch.eye(Lt.shape[0], dtype=torch.complex128, device=dev)
    tmp = []
    for idx in order.tolist():
        # left eigenvector U2_n: (L' - lambda_n)^dag U2_n = 0, i.e. the left singular vector
        # of L' - lambda_n for its smallest singular value (robust to degenerate spectra)
        Uu, _, _ = torch.linalg.svd(Lt - w[idx] * I)
        u2 = Uu[:, -1]
        u1 = U1[:, idx]
        norm = torch.vdot(u2, u1)
        tmp.append(torch.vdot(mu.flatten(), u1) * torch.vdot(u2, (mu @ r0).flatten()) / norm)
    tmp = torch.stack(tmp)
    om = torch.as_tensor(np.asarray(omegas, dtype=np.float64), device=dev).to(torch.complex128)
    sig = (tmp[None, :] / (om[:, None] - w[order][None, :])).sum(dim=1)
    return _dense(-2.0 * sig.imag)


def linear_absorption(omegas, ham, dip, rho0, c_ops=None, ntrans=1):
    dev = device()
    H = as_c128(_dense(ham), dev)
    n = H.shape[0]
    Lfull = liouvillian_dense(_dense(ham), [_dense(c) for c in (c_ops or [])], dev)
    I = torch.eye(n, dtype=torch.complex128, device=dev)
    comm = torch.kron(H, I) - torch.kron(I, H.transpose(0, 1).contiguous())
    diss = Lfull + 1j * comm                       # sum of Lindblad dissipators
    Lp = comm + 1j * diss
    return absorption_eseries(omegas, Lp, dip, rho0, ntrans)


def absorption_from_correlation(C, t):
    """Spectrum g(w) of C(t) (lime's convention); a 2D ``C`` is averaged over axis 0."""
    is_t = isinstance(C, torch.Tensor)
    Ct = C if is_t else torch.as_tensor(np.asarray(C), device=device())
    if Ct.dim() == 2:
        Ct = Ct.mean(dim=0)
    g, freq = _fft.fft(Ct, t)
    return (g if is_t else _dense(g)), freq
