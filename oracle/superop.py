"""Liouville-space oracle (TEST INFRASTRUCTURE ONLY).

Row-major vectorisation ``vec(rho) = rho.flatten()`` so that ``vec(A rho B) = (A (x) B^T) vec rho``.

* ``operator_to_superoperator`` -- ``lime/superoperator.py:201-247`` (kinds
  commutator/left/right/anticommutator, ValueError otherwise)
* ``lindblad_dissipator``       -- ``superoperator.py:250-252``
* ``liouvillian``               -- ``superoperator.py:30-59`` (d vec rho / dt = L vec rho)
* ``left`` / ``right`` / ``kraus`` -- ``superoperator.py:257-290``
* ``obs`` / ``trace``           -- ``superoperator.py:314-319``
* ``eseries_*``                 -- ``superoperator.py:456-773`` with complex biorthogonal
  norms (lime keeps only the real part of diag(vl^dag vr), ``:509``).

Pinned by: textbook identities vec(A X B) = (A (x) B^T) vec X checked on random
matrices; scipy.linalg.expm of the Liouvillian; the closed form of dephasing; ``trace`` and
``obs`` against np.trace(rho) and np.trace(a rho) on complex non-Hermitian rho and a.
"""
import numpy as np
import scipy.linalg as sla


def dm2vec(rho):
    return np.asarray(rho).flatten()


def vec2dm(v, n=None):
    v = np.asarray(v)
    n = n or int(round(np.sqrt(v.size)))
    return v.reshape(n, n)


def operator_to_superoperator(a, kind="commutator"):
    a = np.asarray(a, dtype=np.complex128)
    idm = np.identity(a.shape[-1])
    if kind in ("commutator", "c", "-"):
        return np.kron(a, idm) - np.kron(idm, a.T)
    if kind in ("left", "l"):
        return np.kron(a, idm)
    if kind in ("right", "r"):
        return np.kron(idm, a.T)
    if kind in ("anticommutator", "a", "+"):
        return np.kron(a, idm) + np.kron(idm, a.T)
    raise ValueError("Error: superoperator {} does not exist.".format(kind))


op2sop = operator_to_superoperator


def left(a):
    return operator_to_superoperator(a, "left")


def right(a):
    return operator_to_superoperator(a, "right")


def kraus(a):
    a = np.asarray(a)
    return right(np.conj(a).T) @ left(a)


def lindblad_dissipator(l):
    l = np.asarray(l, dtype=np.complex128)
    return np.kron(l, l.conj()) - 0.5 * operator_to_superoperator(np.conj(l).T @ l, "anticommutator")


def liouvillian(H, c_ops):
    L = -1j * operator_to_superoperator(H)
    for c in c_ops or []:
        L = L + lindblad_dissipator(c)
    return L


def obs(rho_vec, a):
    return np.vdot(dm2vec(np.conj(np.asarray(a)).T), rho_vec)


def trace(rho_vec):
    n = int(round(np.sqrt(len(rho_vec))))
    return np.vdot(dm2vec(np.identity(n)), rho_vec)


def rk4_amplification(L, dt):
    """P(dt L) = I + X + X^2/2 + X^3/6 + X^4/24 with X = dt L: one exact RK4 step."""
    X = dt * np.asarray(L)
    I = np.identity(X.shape[0])
    X2 = X @ X
    return I + X + X2 / 2 + X2 @ X / 6 + X2 @ X2 / 24


def propagate_rk4_liouville(L, rho0, dt, Nt):
    """vec rho_n = P(dt L)^n vec rho0 for every n = 1..Nt (returns (Nt, N, N))."""
    P = rk4_amplification(L, dt)
    v = dm2vec(rho0).astype(np.complex128)
    n = int(round(np.sqrt(v.size)))
    out = np.empty((Nt, n, n), dtype=np.complex128)
    for k in range(Nt):
        v = P @ v
        out[k] = v.reshape(n, n)
    return out


def propagate_exact(L, rho0, t):
    v = sla.expm(np.asarray(L) * t) @ dm2vec(rho0).astype(np.complex128)
    return vec2dm(v)


# ------------------------------------------------------------------- eigen-series solver
class Eseries:
    """Exact-in-time Lindblad propagation from the eigenpairs of L (superoperator.py:456-773)."""

    def __init__(self, H, c_ops):
        self.H = np.asarray(H)
        self.n = self.H.shape[0]
        self.L = liouvillian(H, c_ops)
        w, vl, vr = sla.eig(self.L, left=True, right=True)
        self.eigvals = w
        self.U1 = vr
        self.U2 = vl
        self.norm = np.einsum("ij,ij->j", np.conj(vl), vr)   # complex biorthogonal norms
        self.idv = dm2vec(np.identity(self.n))

    def _coeff_in(self, x0):
        return (np.conj(self.U2).T @ dm2vec(x0)) / self.norm

    def evolve(self, rho0, tlist, e_ops):
        c = self._coeff_in(rho0)
        out = np.zeros((len(tlist), len(e_ops)), dtype=np.complex128)
        for i, t in enumerate(tlist):
            rho = self.U1 @ (c * np.exp(self.eigvals * t))
            out[i] = [obs(rho, e) for e in e_ops]
        return out

    def correlation_2op_1t(self, rho0, ops, tlist):
        a, b = ops
        w = np.conj(self.idv) @ (left(a) @ self.U1)
        coeff = w * self._coeff_in(np.asarray(b) @ np.asarray(rho0))
        return np.array([np.sum(np.exp(self.eigvals * t) * coeff) for t in tlist])

    def correlation_2op_1w(self, rho0, ops, omegas):
        a, b = ops
        w = np.conj(self.idv) @ (left(a) @ self.U1)
        coeff = w * self._coeff_in(np.asarray(b) @ np.asarray(rho0))
        return np.array([np.sum(-1.0 / (self.eigvals + 1j * om) * coeff) for om in omegas])

    def correlation_3op_1t(self, rho0, ops, tlist):
        a, b, c = ops
        w = np.conj(self.idv) @ (left(b) @ self.U1)
        coeff = w * self._coeff_in(np.asarray(c) @ np.asarray(rho0) @ np.asarray(a))
        return np.array([np.sum(np.exp(self.eigvals * t) * coeff) for t in tlist])

    def correlation_3op_2t(self, rho0, ops, tlist, taulist):
        """<A(t) B(t+tau) C(t)> as an (len(tlist), len(taulist)) array."""
        a, b, c = ops
        wb = np.conj(self.idv) @ (left(b) @ self.U1)                      # (k,)
        mid = (np.conj(self.U2).T @ (right(a) @ (left(c) @ self.U1))) / self.norm[:, None]
        c0 = self._coeff_in(rho0)
        coeff = wb[:, None] * mid * c0[None, :]                              # (m, n)
        e_tau = np.exp(np.outer(self.eigvals, taulist))                     # (m, tau)
        e_t = np.exp(np.outer(self.eigvals, tlist))                         # (n, t)
        return (e_tau.T @ coeff @ e_t).T
