"""Bloch-Redfield oracle (TEST INFRASTRUCTURE ONLY).

* ``redfield_tensor`` -- ``lime/oqs.py:528-579``: in the eigenbasis of H,
      R = -i op2sop(diag E) - sum_k op2sop(A_k) (left(Lam_k) - right(Lam_k^dag)),
      Lam_k = C_k o A_k, C_k[n, m] = S_k(-(E_n - E_m)) (real), Hermitian a_ops required.
* ``rhs_hilbert``     -- the same generator applied in Hilbert space:
      d rho/dt = -i[E, rho] - sum_k [A_k, Lam_k rho - rho Lam_k^dag]  (4K GEMMs).
* ``propagate``       -- ``lime/oqs.py:373-468`` (``_redfield``): RK4 on row-major vec rho
  in the eigenbasis; states are transformed back with ``v rho v^dag``; observables are
  taken in the eigenbasis (transformed e_ops).
* ``propagator_expm`` -- ``lime/oqs.py:169-209``: U(t) = exp(R t) stacked as (N^2, N^2, Nt).
* ``response_4op_3t`` -- ``lime/oqs.py:277-366``: <<I| a G(t_i) b G(t_j) c G(t_k) d |rho0>>,
  G = -i U.

Pinned by: the Liouville form against the Hilbert form (independent constructions);
the unitary limit (silent bath) of ``propagate`` against site-basis Hilbert-space RK4 and
e^{-iHt} rho0 e^{iHt} at a non-diagonal H, which fixes the direction of the eigenbasis
change on the way in and out;
the population block of the secular Redfield rate matrix against golden-rule rates
2|A_nm|^2 S(w) (closed form); trace conservation; exp(R t) against the eigendecomposition
V e^{Lambda t} V^-1 and the semigroup law; the four-point response against the closed form
of a single two-level Liouville pathway with distinct rates (time ordering, the -i of G).
"""
import numpy as np
from scipy.linalg import eigh

from .lindblad import dag, rk4
from .superop import left, op2sop, right


def isherm(a, tol=1e-8):
    a = np.asarray(a)
    return np.allclose(a, np.conj(a).T, atol=tol)


def bath_matrices(H, a_ops, spectra):
    for a in a_ops:
        if not isherm(a):
            raise TypeError("Operators in a_ops must be Hermitian.")
    evals, evecs = eigh(np.asarray(H))
    W = np.real(evals[:, None] - evals[None, :])
    A = [dag(evecs) @ np.asarray(a) @ evecs for a in a_ops]
    C = [np.vectorize(s)(-W).astype(np.float64) for s in spectra]
    Lam = [C[k] * A[k] for k in range(len(a_ops))]
    return evals, evecs, A, Lam


def redfield_tensor(H, a_ops, spectra):
    evals, evecs, A, Lam = bath_matrices(H, a_ops, spectra)
    R = -1j * op2sop(np.diag(evals))
    for a, l in zip(A, Lam):
        R = R - op2sop(a) @ (left(l) - right(dag(l)))
    return R, evecs


def rhs_hilbert(rho, E, A, Lam):
    out = -1j * (E[:, None] * rho - rho * E[None, :])
    for a, l in zip(A, Lam):
        x = l @ rho - rho @ dag(l)
        out = out - (a @ x - x @ a)
    return out


def propagate(R, evecs, rho0, dt, Nt, e_ops=()):
    """RK4 on vec rho in the eigenbasis (oqs.py:373-468); returns (states, observables)."""
    n = rho0.shape[0]
    rho = (dag(evecs) @ np.asarray(rho0) @ evecs).flatten().astype(np.complex128)
    e_eb = [dag(evecs) @ np.asarray(e) @ evecs for e in e_ops]
    states = np.empty((Nt, n, n), dtype=np.complex128)
    obs = np.zeros((Nt, len(e_ops)), dtype=np.complex128)
    for k in range(Nt):
        rho = rk4(rho, lambda v, M: M @ v, dt, R)
        tmp = rho.reshape(n, n)
        states[k] = evecs @ tmp @ dag(evecs)
        obs[k] = [np.trace(e @ tmp) for e in e_eb]
    return states, obs


def propagator_expm(R, tlist):
    """U(t) = exp(R t) for each t (scipy expm), stacked as (N^2, N^2, Nt)."""
    from scipy.linalg import expm
    R = np.asarray(R)
    return np.stack([expm(R * t) for t in tlist], axis=-1)


def response_4op_3t(R, rho0, ops_super, tlist):
    """<<I| a G(t_i) b G(t_j) c G(t_k) d |rho0>> with G = -i exp(R t), explicit loops
    (oqs.py:277-366 semantics, out[i, j, k])."""
    a, b, c, d = [np.asarray(o) for o in ops_super]
    n = int(round(np.sqrt(R.shape[0])))
    G = -1j * propagator_expm(R, tlist)
    idm = np.identity(n).flatten()
    v0 = d @ np.asarray(rho0).flatten()
    nt = len(tlist)
    out = np.zeros((nt, nt, nt), dtype=np.complex128)
    for k in range(nt):
        x = c @ (G[:, :, k] @ v0)
        for j in range(nt):
            y = b @ (G[:, :, j] @ x)
            for i in range(nt):
                out[i, j, k] = idm @ (a @ (G[:, :, i] @ y))
    return out
