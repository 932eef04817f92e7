"""Bloch-Redfield propagation on the MI355X kernels (drop-in for lime/oqs.py:40-579).

lime builds the Redfield tensor in the eigenbasis of H as an N^2 x N^2 Liouville matrix
(``redfield_tensor``, oqs.py:528-579)

    R = -i op2sop(diag E) - sum_k op2sop(A_k) (left(Lam_k) - right(Lam_k^dag)),
    A_k = v^dag a_k v,  Lam_k = C_k o A_k,  C_k[n, m] = S_k(-(E_n - E_m)),

and steps vec(rho) with RK4 matvecs (``_redfield``, oqs.py:373-472): O(N^4) per RHS.
The same generator in Hilbert space is

    d rho/dt = -i[E, rho] - sum_k [A_k, Lam_k rho - rho Lam_k^dag]
             = -i (G rho - rho G^dag) + sum_k (A_k rho Lam_k^dag + Lam_k rho A_k^dag),
    G = diag(E) - i sum_k A_k Lam_k,

i.e. the Lindblad kernel's right-hand side with H_eff = G and 2K sandwich terms
(L, R) = (A_k, Lam_k), (Lam_k, A_k): 2 + 4K complex GEMMs per RHS, O(N^3).
``Redfield_solver.evolve`` runs exactly that on ``lime_lindblad_rk4``.

``_redfield`` with an explicit Liouville tensor R keeps lime's Liouville-space algorithm
(RK4 on vec rho, R vec rho per stage) as batched GPU products; ``propagator``/``gf``/
``expect``/``correlation_2op_1t``/``correlation_4op_3t`` (oqs.py:145-366) evaluate the
Liouville-space propagator U(t) on the GPU (eigendecomposition for 'SOS', RK4 for 'EOM').
"""
import numpy as np
import scipy.sparse as sp
import torch
from scipy.linalg import eigh

from . import superoperator as so
from .mol import Result
from .propagator import LindbladPropagator, as_c128, device


def _dense(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return a.toarray() if sp.issparse(a) else np.asarray(a)


def isherm(a, tol=1e-8):
    a = _dense(a)
    return np.allclose(a, np.conjugate(a).T, atol=tol)


def _spectral_matrix(spec, W):
    """C[n, m] = S(-W[n, m]); vectorised when the callback accepts arrays."""
    try:
        c = np.asarray(spec(-W), dtype=np.float64)
        if c.shape == W.shape:
            return c
    except Exception:
        pass
    return np.vectorize(lambda x: float(spec(x)))(-W)


def bath_operators(H, a_ops, spectra):
    """Eigen-decomposition and the eigenbasis operators (E, evecs, A_k, Lam_k)."""
    for a in a_ops:
        if not isherm(a):
            raise TypeError("Operators in a_ops must be Hermitian.")
    evals, evecs = eigh(_dense(H))
    W = np.real(evals[:, None] - evals[None, :])
    A = [evecs.conj().T @ _dense(a) @ evecs for a in a_ops]
    C = [_spectral_matrix(s, W) for s in spectra]
    Lam = [C[k] * A[k] for k in range(len(a_ops))]
    return evals, evecs, A, Lam


def redfield_tensor(H, a_ops, spectra, secular=False):
    """(R as scipy CSR, evecs) exactly as lime/oqs.py:528-579 (``secular`` unused there too)."""
    evals, evecs, A, Lam = bath_operators(H, a_ops, spectra)
    R = -1j * so.op2sop(np.diag(evals))
    for a, l in zip(A, Lam):
        R = R - so.op2sop(a) @ (so.left(l) - so.right(l.conj().T))
    return sp.csr_matrix(R), evecs


def liouville_propagator(G, precision="fp64"):
    """Kernel stepping dX/dt = G X for a dense Liouville-space generator G (M x M): the
    one-sided generator of the per-stage kernels with H_eff = i G.  FP64 by default (these
    are lime's complex128 Liouville-space paths)."""
    dev = device()
    Gt = as_c128(_dense(G), dev)
    return LindbladPropagator.from_folded(1j * Gt, one_sided=True, precision=precision)


def _column_chunk(M, budget=1 << 28):
    """RK4 steps per launch when every step's M x M state is recorded (<= 256 MB)."""
    return max(1, budget // (M * M * 16))


def _redfield(R, rho0, evecs=None, Nt=1, dt=0.005, t0=0, e_ops=None, return_result=True,
              precision="fp64"):
    """RK4 on vec rho with a dense Liouville tensor R (oqs.py:373-468 semantics).

    vec rho is column 0 of an M x M operator X (M = N^2, other columns zero), stepped by
    the one-sided kernels (dX/dt = R X) in launches of many steps; Tr[e rho] is the kernel's
    record sum_m W_m0 X_m0 with W_m0 = e[m % N, m // N].  States are transformed back with
    evecs (rho -> v rho v^dag), observables are taken in the eigenbasis, as lime does."""
    dev = device()
    e_ops = [] if e_ops is None else list(e_ops)
    r0 = _dense(rho0)
    n = r0.shape[0]
    M = n * n
    if evecs is not None:
        r0 = evecs.conj().T @ r0 @ evecs
        e_ops = [evecs.conj().T @ _dense(e) @ evecs for e in e_ops]
    prop = liouville_propagator(R, precision)
    X = torch.zeros((1, M, M), dtype=torch.complex128, device=dev)
    X[0, :, 0] = as_c128(r0.flatten(), dev)
    Et = None
    if e_ops:
        Et = torch.zeros((1, len(e_ops), M, M), dtype=torch.complex128, device=dev)
        for i, e in enumerate(e_ops):
            Et[0, i, :, 0] = as_c128(_dense(e).T.flatten(), dev)     # e[m % N, m // N]
    obs = torch.zeros((1, Nt, len(e_ops)), dtype=torch.complex128, device=dev) if e_ops else None
    cols = torch.empty((Nt, M), dtype=torch.complex128, device=dev) if return_result else None
    chunk = _column_chunk(M) if return_result else Nt
    done = 0
    while done < Nt:
        m = min(chunk, Nt - done)
        # records at chunk-local indices 0..m-1 (one launch sequence per chunk)
        st = torch.empty((1, m, M, M), dtype=torch.complex128, device=dev) if return_result else None
        ob = torch.empty((1, m, len(e_ops)), dtype=torch.complex128, device=dev) if e_ops else None
        prop.step(X, dt, m, eops_t=Et, obs=ob, states=st, nout=1, step_offset=0, rec_capacity=m)
        if ob is not None:
            obs[:, done:done + m] = ob
        if st is not None:
            cols[done:done + m] = st[0, :, :, 0]
        done += m
    if not return_result:
        return _dense(X[0, :, 0].reshape(n, n))
    res = Result(dt=dt, Nt=Nt, rho0=rho0, t0=t0)
    res.observables = _dense(obs[0]) if obs is not None else np.zeros((Nt, 0), np.complex128)
    states = cols.reshape(Nt, n, n)
    if evecs is not None:
        V = as_c128(evecs, dev)
        states = torch.matmul(torch.matmul(V, states), V.conj().transpose(0, 1))
    res.rholist = list(_dense(states))
    return res


def getG(L, t, w=None, k=6, domain="time"):
    """Green's function of (i d/dt - L) by eigendecomposition (oqs.py:474-526), on the GPU.

    time:  G(t) = -i U1 diag(exp(-i lambda t)) U1^-1        -> (M, M, Nt)
    freq:  G(w) = U1 diag(1 / (w - lambda)) U1^-1 = (w - L)^-1 -> (M, M, Nw)
    lime's frequency branch contracts with conj(U1^-1)[b, n] instead of U1^-1[n, b], which is
    the resolvent only when U1 is symmetric; here it is the resolvent (w - L)^-1 itself."""
    dev = device()
    Lt = as_c128(_dense(L), dev)
    lam, U1 = torch.linalg.eig(Lt)
    U2 = torch.linalg.inv(U1)
    if domain == "time":
        tt = torch.as_tensor(np.asarray(t, dtype=np.float64), device=dev).to(torch.complex128)
        ph = -1j * torch.exp(-1j * lam[:, None] * tt[None, :])              # (M, Nt)
    elif domain == "freq":
        ww = torch.as_tensor(np.asarray(w, dtype=np.complex128), device=dev)
        ph = 1.0 / (ww[None, :] - lam[:, None])                              # (M, Nw)
    else:
        raise ValueError(f"unknown domain {domain!r}; expected 'time' or 'freq'")
    G = torch.matmul(U1.unsqueeze(0) * ph.t().unsqueeze(1), U2.unsqueeze(0))  # (Nt, M, M)
    return _dense(G.permute(1, 2, 0))


class Redfield_solver:
    """oqs.py:40-370 with the Hilbert-space Redfield generator on the fused kernel."""

    def __init__(self, H, c_ops=None, spectra=None, e_ops=None):
        self.H = H
        self.c_ops = c_ops
        self.R = None
        self.spectra = spectra
        self.evecs = None
        self.dim = _dense(H).shape[0]
        self.U = None
        self.G = None
        self.e_ops = e_ops
        self._ops = None

    def idm(self, sp_=True):
        v = so.dm2vec(sp.identity(self.dim)) if sp_ else so.dm2vec(np.identity(self.dim))
        return v

    def configure(self, H, c_ops, e_ops):
        self.c_ops, self.e_ops, self.H = c_ops, e_ops, H
        self.dim = _dense(H).shape[0]
        self.R = self._ops = None

    def redfield_tensor(self, secular=False):
        if self.spectra is None:
            raise TypeError("Specify the bath spectral function.")
        R, evecs = redfield_tensor(self.H, self.c_ops, self.spectra, secular)
        self.R, self.evecs = R, evecs
        self._ops = bath_operators(self.H, self.c_ops, self.spectra)
        return R, evecs

    def _generator(self):
        if self._ops is None:
            self.redfield_tensor()
        evals, evecs, A, Lam = self._ops
        G = np.diag(evals).astype(np.complex128)
        for a, l in zip(A, Lam):
            G = G - 1j * (a @ l)
        Ls = list(A) + list(Lam)
        Rs = list(Lam) + list(A)
        return G, Ls, Rs, evecs

    def evolve(self, rho0, dt, Nt, evecs=None, e_ops=None, store_states=False, t0=0, nout=1):
        """RK4 in the eigenbasis (oqs.py:66-90 -> 373-468); lime ignores ``evecs`` too."""
        e_ops = [] if e_ops is None else list(e_ops)
        G, Ls, Rs, V = self._generator()
        rho_eb = V.conj().T @ _dense(rho0) @ V
        e_eb = [V.conj().T @ _dense(e) @ V for e in e_ops]
        prop = LindbladPropagator(G, None, sandwich=(Ls, Rs))
        _, obs, states = prop.run(rho_eb, dt, Nt, e_ops=e_eb or None, nout=nout, store_states=True)
        dev = prop.dev
        Vt = as_c128(V, dev)
        back = torch.matmul(torch.matmul(Vt, states[0]), Vt.conj().transpose(0, 1))
        res = Result(dt=dt, Nt=Nt, rho0=rho0, t0=t0, nout=nout)
        res.observables = _dense(obs[0]) if obs is not None else np.zeros((Nt // nout, 0), np.complex128)
        res.rholist = list(_dense(back))
        return res

    # ---------------- Liouville-space propagator and response functions (oqs.py:145-366)
    def propagator(self, t, method="SOS"):
        if self.R is None:
            raise TypeError("Redfield tensor is not computed. Please call redfield_tensor()")
        dev = device()
        Rm = as_c128(_dense(self.R), dev)
        t = np.asarray(t, dtype=np.float64)
        if method == "EOM":
            from .phys import expm
            Ut = torch.stack([as_c128(u, dev) for u in expm(Rm, t)], dim=-1)
        elif method in ("eseries", "SOS"):
            w, U1 = torch.linalg.eig(Rm)
            U2 = torch.linalg.inv(U1)
            ph = torch.exp(w[:, None] * torch.as_tensor(t, device=dev).to(torch.complex128)[None, :])
            Ut = torch.einsum("aj,jk,jb->abk", U1, ph, U2)
        else:
            raise ValueError(f"unknown propagator method {method!r}")
        self.U = _dense(Ut)
        self.G = -1j * self.U
        return self.U

    def gf(self, t, w=None, secular=False, k=1, domain="time", method="EOM"):
        """Liouville-space Green's function G(t) = -i U(t) as (N^2, N^2, Nt) (oqs.py:145-166):
        'EOM' steps dU/dt = R U with the one-sided kernels; 'eseries' / 'diag' /
        'diagonalization' is getG(i R, t) = -i exp(R t) by eigendecomposition.  lime's EOM
        branch fails (``-1j * list`` of ``expm``'s result, and R unbound when the tensor was
        already computed); here both branches return the same array."""
        if self.R is None:
            self.redfield_tensor(secular=secular)
        if method == "EOM":
            return -1j * self.propagator(t, method="EOM")
        if method in ("eseries", "diag", "diagonalization"):
            return getG(1j * _dense(self.R), t)
        raise ValueError(f"unknown method {method!r}")

    def expect(self, rho0, e_ops):
        """<<e|U(t)|rho0>> in the eigenbasis for each e_op (oqs.py:225-252), on the GPU."""
        dev = device()
        V = self.evecs
        rho_eb = as_c128(so.dm2vec(V.conj().T @ _dense(rho0) @ V), dev)
        E = as_c128(np.stack([so.dm2vec(V.conj().T @ _dense(e) @ V) for e in e_ops]), dev)
        U = as_c128(self.U, dev)                                        # (N^2, N^2, Nt)
        rho_t = torch.einsum("abk,b->ak", U, rho_eb)                    # (N^2, Nt)
        return _dense((E @ rho_t).transpose(0, 1))

    def correlation_2op_1t(self, rho0, a, b, tau):
        if self.G is None:
            self.propagator(tau)
        dev = device()
        idm = as_c128(so.dm2vec(np.identity(self.dim)), dev)
        r0 = _dense(rho0)
        r0 = so.dm2vec(r0) if r0.ndim == 2 else r0
        x = as_c128(_dense(b) @ r0, dev)
        Gx = torch.einsum("abk,b->ak", as_c128(self.G, dev), x)          # (N^2, Nt)
        return _dense(idm @ (as_c128(_dense(a), dev) @ Gx))

    def correlation_4op_3t(self, rho0, oplist, signature, tau):
        """<<I| a G(t3) b G(t2) c G(t1) d |rho0>> as an (Nt, Nt, Nt) array (oqs.py:277-366)."""
        if len(oplist) != 4:
            raise ValueError("Number of operators is not 4.")
        a, b, c, d = [so.operator_to_superoperator(_dense(o), s).toarray()
                      for o, s in zip(oplist, signature)]
        if self.G is None:
            self.propagator(tau)
        dev = device()
        G = as_c128(self.G, dev)
        idm = as_c128(so.dm2vec(np.identity(self.dim)), dev)
        rho = as_c128(d @ so.dm2vec(_dense(rho0)), dev)
        tmp = torch.tensordot(G, rho, dims=([1], [0]))                    # (a, k)
        tmp = as_c128(c, dev) @ tmp
        tmp = torch.tensordot(G, tmp, dims=([1], [0]))                    # (a, j, k)
        tmp = torch.tensordot(as_c128(b, dev), tmp, dims=([1], [0]))
        tmp = torch.tensordot(G, tmp, dims=([1], [0]))                    # (a, i, j, k)
        out = torch.einsum("a,ab,bijk->ijk", idm, as_c128(a, dev), tmp)
        return _dense(out)
