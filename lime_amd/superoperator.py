"""lime.superoperator call surface: Liouville-space algebra and the eigen-series solver.

Row-major vectorisation ``vec(rho) = rho.flatten()``, so ``vec(A rho B) = (A (x) B^T) vec(rho)``
(lime/superoperator.py:112-151, 201-247).

* Builders (``operator_to_superoperator``/``op2sop``, ``left``, ``right``, ``kraus``,
  ``lindblad_dissipator``, ``liouvillian``) return ``scipy.sparse`` CSR matrices exactly as
  lime does (superoperator.py:30-59, 201-290): they are set-up objects, not a time loop.
  ``liouvillian_dense`` builds the same generator as a dense complex128 GPU tensor.
* ``Lindblad_solver`` (superoperator.py:456-773, the exact "eseries" solver) diagonalises
  the dense Liouvillian on the GPU (torch.linalg.eig) and evaluates every time series as
  batched GPU products.  Biorthogonal norms are kept complex (lime keeps only their real
  part, :509); ``evolve`` works (lime's raises at ``Result(times=tlist)``, :527).
"""
import math

import numpy as np
import scipy.sparse as sp
import torch

from .mol import Result
from .propagator import as_c128, device


def _dense(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    if sp.issparse(a):
        return a.toarray()
    return np.asarray(a)


def operator_to_vector(rho):
    return _dense(rho).flatten()


def dm2vec(rho):
    if sp.issparse(rho):
        n, m = rho.shape
        return rho.tolil().reshape((n * m, 1))
    return _dense(rho).flatten()


def operator_to_superoperator(a, kind="commutator"):
    a = sp.csr_matrix(_dense(a))
    idm = sp.identity(a.shape[-1], format="csr")
    if kind in ("commutator", "c", "-"):
        return (sp.kron(a, idm) - sp.kron(idm, a.T)).tocsr()
    if kind in ("left", "l"):
        return sp.kron(a, idm).tocsr()
    if kind in ("right", "r"):
        return sp.kron(idm, a.T).tocsr()
    if kind in ("anticommutator", "a", "+"):
        return (sp.kron(a, idm) + sp.kron(idm, a.T)).tocsr()
    raise ValueError("Error: superoperator {} does not exist.".format(kind))


def op2sop(a, kind="commutator"):
    return operator_to_superoperator(a, kind=kind)


to_super = op2sop


def left(a):
    return operator_to_superoperator(a, "left")


def right(a):
    return operator_to_superoperator(a, "right")


def kraus(a):
    a = _dense(a)
    return (right(np.conj(a).T) @ left(a)).tocsr()


def lindblad_dissipator(l):
    l = sp.csr_matrix(_dense(l))
    return (sp.kron(l, l.conj()) - 0.5 * operator_to_superoperator(l.conj().T @ l, "anticommutator")).tocsr()


def liouvillian(H, c_ops):
    L = -1j * operator_to_superoperator(H)
    for c in c_ops or []:
        L = L + lindblad_dissipator(c)
    return L.tocsr()


def liouvillian_dense(H, c_ops, dev=None):
    """The same generator as a dense (N^2, N^2) complex128 tensor built on the GPU."""
    dev = dev or device()
    Ht = as_c128(_dense(H), dev)
    n = Ht.shape[-1]
    I = torch.eye(n, dtype=torch.complex128, device=dev)
    L = -1j * (torch.kron(Ht, I) - torch.kron(I, Ht.transpose(0, 1).contiguous()))
    for c in c_ops or []:
        C = as_c128(_dense(c), dev)
        CdC = (C.conj().transpose(0, 1) @ C).contiguous()
        L = L + torch.kron(C, C.conj().contiguous()) - 0.5 * (torch.kron(CdC, I) + torch.kron(I, CdC.transpose(0, 1).contiguous()))
    return L


def obs(rho, a):
    return np.vdot(operator_to_vector(np.conj(_dense(a)).T), rho)


def trace(rho):
    n = math.isqrt(len(rho))
    return np.vdot(operator_to_vector(np.identity(n)), rho)


def resolvent(omega, L):
    """(omega - L)^-1 (superoperator.py resolvent), inverted on the GPU; NumPy out."""
    dev = device()
    Lt = torch.as_tensor(np.asarray(_dense(L), dtype=np.complex128), device=dev)
    eye = torch.eye(Lt.shape[0], dtype=torch.complex128, device=dev)
    return torch.linalg.inv(omega * eye - Lt).cpu().numpy()


class Lindblad_solver:
    """Exact propagation from the eigenpairs of the Liouvillian, on the GPU."""

    def __init__(self, H, c_ops=None):
        self.H = H
        self.c_ops = c_ops
        self.L = None
        self.dim = _dense(H).shape[-1] ** 2
        self.idv = operator_to_vector(np.identity(_dense(H).shape[-1]))
        self.eigvals = self.right_eigvecs = self.left_eigvecs = self.norm = None

    def liouvillian(self):
        self.L = liouvillian(self.H, self.c_ops)
        return self.L

    def eigenstates(self, k=None):
        """Eigenpairs of the Liouvillian (superoperator.py:490-523) by a dense GPU
        eigendecomposition; left eigenvectors are the rows of inv(V) (biorthonormal, complex
        norms).  With ``k`` the k eigenvalues of largest real part (lime's ``eigs(.., k,
        which='LR')``) are kept, ordered as lime's ``sort`` orders them (ascending by real,
        then imaginary part), and every later call of this solver uses those k modes; lime's
        own k branch returns unbound names.  k >= dim - 1 raises ValueError as in lime."""
        if k is not None and k >= self.dim - 1:
            raise ValueError("k should be < the size of the matrix.")
        dev = device()
        Ld = liouvillian_dense(self.H, self.c_ops, dev)
        w, vr = torch.linalg.eig(Ld)
        vl_h = torch.linalg.inv(vr)                      # rows: left eigenvectors^dag, vl^dag vr = I
        if k is not None:
            wn = _dense(w)
            keep = np.argsort(-wn.real, kind="stable")[:k]
            keep = keep[np.lexsort((wn[keep].imag, wn[keep].real))]
            idx = torch.as_tensor(keep, device=dev)
            w, vr, vl_h = w[idx], vr[:, idx], vl_h[idx, :]
        self.eigvals = w
        self.right_eigvecs = vr
        self.left_eigvecs = vl_h.conj().transpose(0, 1)
        self.norm = torch.ones_like(w)
        if self.L is None:
            self.liouvillian()
        return _dense(w), _dense(vr), _dense(self.left_eigvecs)

    def _ensure(self):
        if self.eigvals is None:
            self.eigenstates()
        return device()

    def _coeff(self, x0):
        dev = self._ensure()
        v = as_c128(operator_to_vector(x0), dev)
        return (self.left_eigvecs.conj().transpose(0, 1) @ v) / self.norm

    def _observe_row(self, a):
        """<<1| left(a) U1  ->  sum_i (a U1_n)_ii  for every eigenvector n."""
        dev = self._ensure()
        n = int(round(math.sqrt(self.dim)))
        A = as_c128(_dense(a), dev)
        U = self.right_eigvecs.transpose(0, 1).reshape(-1, n, n)           # U_n as matrices (kept modes)
        return torch.einsum("ij,mji->m", A, U)

    def evolve(self, rho0, tlist, e_ops):
        dev = self._ensure()
        t = torch.as_tensor(np.asarray(tlist, dtype=np.float64), device=dev)
        c = self._coeff(rho0)
        ph = torch.exp(self.eigvals[:, None] * t[None, :].to(torch.complex128))   # (m, T)
        out = np.zeros((len(tlist), len(e_ops)), dtype=np.complex128)
        for j, e in enumerate(e_ops):
            row = self._observe_row(e)
            out[:, j] = _dense((row * c) @ ph)
        res = Result(times=tlist)
        res.observables = out
        return res

    def correlation_2op_1t(self, rho0, ops, tlist):
        a, b = ops
        dev = self._ensure()
        coeff = self._observe_row(a) * self._coeff(_dense(b) @ _dense(rho0))
        t = torch.as_tensor(np.asarray(tlist, dtype=np.float64), device=dev).to(torch.complex128)
        return _dense(coeff @ torch.exp(self.eigvals[:, None] * t[None, :]))

    def correlation_2op_1w(self, rho0, ops, w):
        a, b = ops
        dev = self._ensure()
        coeff = self._observe_row(a) * self._coeff(_dense(b) @ _dense(rho0))
        om = torch.as_tensor(np.asarray(w, dtype=np.float64), device=dev).to(torch.complex128)
        return _dense((-1.0 / (self.eigvals[None, :] + 1j * om[:, None])) @ coeff)

    def correlation_3op_1t(self, rho0, ops, t):
        a, b, c = ops
        dev = self._ensure()
        coeff = self._observe_row(b) * self._coeff(_dense(c) @ _dense(rho0) @ _dense(a))
        tt = torch.as_tensor(np.asarray(t, dtype=np.float64), device=dev).to(torch.complex128)
        return _dense(coeff @ torch.exp(self.eigvals[:, None] * tt[None, :]))

    def correlation_3op_1w(self, rho0, ops, w):
        a, b, c = ops
        dev = self._ensure()
        coeff = self._observe_row(b) * self._coeff(_dense(c) @ _dense(rho0) @ _dense(a))
        om = torch.as_tensor(np.asarray(w, dtype=np.float64), device=dev).to(torch.complex128)
        return _dense((-1.0 / (self.eigvals[None, :] + 1j * om[:, None])) @ coeff)

    def correlation_3op_2t(self, rho0, ops, tlist, taulist, k=None):
        """<A(t) B(t+tau) C(t)> as (len(tlist), len(taulist)).

        lime returns ``tmp1.T @ coeff @ tmp2`` of shape (len(taulist), len(tlist))
        (superoperator.py:751); the (t, tau) orientation here matches oqs.correlation_3op_2t.
        """
        a, b, c = ops
        dev = self._ensure()
        n = int(round(math.sqrt(self.dim)))
        wb = self._observe_row(b)                                            # (m,)
        U = self.right_eigvecs.transpose(0, 1).reshape(-1, n, n)             # (n_vec, N, N)
        Aop, Cop = as_c128(_dense(a), dev), as_c128(_dense(c), dev)
        CUA = torch.matmul(torch.matmul(Cop, U), Aop).reshape(U.shape[0], self.dim).transpose(0, 1)
        mid = (self.left_eigvecs.conj().transpose(0, 1) @ CUA) / self.norm[:, None]
        c0 = self._coeff(rho0)
        coeff = wb[:, None] * mid * c0[None, :]
        tt = torch.as_tensor(np.asarray(tlist, dtype=np.float64), device=dev).to(torch.complex128)
        tau = torch.as_tensor(np.asarray(taulist, dtype=np.float64), device=dev).to(torch.complex128)
        e_tau = torch.exp(self.eigvals[:, None] * tau[None, :])
        e_t = torch.exp(self.eigvals[:, None] * tt[None, :])
        return _dense((e_tau.transpose(0, 1) @ coeff @ e_t).transpose(0, 1))

    def correlation_4op_2t(self, rho0, ops, tlist, taulist, k=None):
        if len(ops) != 4:
            raise ValueError("Number of operators is not 4.")
        a, b, c, d = ops
        return self.correlation_3op_2t(rho0, [a, _dense(b) @ _dense(c), d], tlist, taulist, k=k)
