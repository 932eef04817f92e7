"""Redfield on the GPU: the Hilbert-space generator on the fused kernel against lime's
Liouville-space formulation restated in oracle/redfield.py."""
import numpy as np
import pytest

from oracle import redfield as orf
from oracle import superop as osu

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rand_herm(rng, n):
    x = rand_c(rng, n, n)
    return (x + x.conj().T) / 2


def rand_dm(rng, n):
    x = rand_c(rng, n, n)
    r = x @ x.conj().T
    return r / np.trace(r)


def ohmic(eta=0.05, wc=2.0, beta=2.0):
    def S(w):
        w = np.asarray(w, dtype=float)
        with np.errstate(divide="ignore", invalid="ignore"):
            val = eta * w * np.exp(-np.abs(w) / wc) / (1 - np.exp(-beta * w))
        return np.where(np.abs(w) < 1e-12, eta / beta, val)
    return S


@pytest.mark.parametrize("n,K", [(2, 1), (5, 2), (24, 2), (100, 1)])
def test_redfield_evolve_matches_liouville_oracle(n, K):
    from lime_amd.oqs import Redfield_solver
    rng = np.random.default_rng(n * 10 + K)
    H = rand_herm(rng, n) / np.sqrt(n)
    a_ops = [rand_herm(rng, n) / np.sqrt(n) for _ in range(K)]
    spectra = [ohmic(), ohmic(0.03, 1.0, 1.0)][:K]
    rho0 = rand_dm(rng, n)
    P = [np.diag(np.eye(n)[i]).astype(complex) for i in range(min(n, 3))]
    dt, nt = 0.05, 40
    s = Redfield_solver(H, c_ops=a_ops, spectra=spectra)
    R, evecs = s.redfield_tensor()
    Rref, _ = orf.redfield_tensor(H, a_ops, spectra)
    assert np.max(np.abs(R.toarray() - Rref)) < 1e-12
    res = s.evolve(rho0, dt, nt, e_ops=P)
    if n <= 24:
        states, obs = orf.propagate(Rref, evecs, rho0, dt, nt, P)
    else:   # Hilbert-form oracle at large n (pinned against the Liouville form above)
        E, V, A, Lam = orf.bath_matrices(H, a_ops, spectra)
        from oracle.lindblad import rk4
        x = V.conj().T @ rho0 @ V
        states, obs = [], []
        for _ in range(nt):
            x = rk4(x, orf.rhs_hilbert, dt, E, A, Lam)
            states.append(V @ x @ V.conj().T)
            obs.append([np.trace((V.conj().T @ p @ V) @ x) for p in P])
        states, obs = np.array(states), np.array(obs)
    err = np.max(np.abs(np.array(res.rholist) - states))
    print(f"redfield n={n} K={K} state err {err:.3e}")
    assert err < 5e-7
    assert np.max(np.abs(res.observables - obs)) < 5e-7


def test_liouville_redfield_function_and_propagators():
    from lime_amd.redfield import Redfield_solver, _redfield
    rng = np.random.default_rng(3)
    n = 3
    H = rand_herm(rng, n)
    a = [rand_herm(rng, n)]
    s = Redfield_solver(H, c_ops=a, spectra=[ohmic()])
    R, evecs = s.redfield_tensor()
    rho0 = rand_dm(rng, n)
    res = _redfield(R, rho0, evecs=evecs, Nt=30, dt=0.05, e_ops=[np.eye(n)])
    states, obs = orf.propagate(R.toarray(), evecs, rho0, 0.05, 30, [np.eye(n)])
    assert np.max(np.abs(np.array(res.rholist) - states)) < 1e-12
    assert np.max(np.abs(res.observables - obs)) < 1e-12
    t = np.linspace(0, 2, 11)
    U_sos = s.propagator(t, "SOS")
    U_ref = orf.propagator_expm(R.toarray(), t)
    assert np.max(np.abs(U_sos - U_ref)) < 1e-10
    tf = np.linspace(0, 2, 201)
    U_eom = s.propagator(tf, "EOM")                       # RK4 with dt = 0.01
    assert np.max(np.abs(U_eom[:, :, ::20] - U_ref)) < 1e-6
    s.propagator(t, "SOS")
    ops = [rand_herm(rng, n) for _ in range(4)]
    sig = "lrcl"
    got = s.correlation_4op_3t(rho0, ops, sig, t[:5])
    sup = [osu.operator_to_superoperator(o, k) for o, k in zip(ops, sig)]
    s.propagator(t[:5], "SOS")
    got = s.correlation_4op_3t(rho0, ops, sig, t[:5])
    ref = orf.response_4op_3t(R.toarray(), rho0, sup, t[:5])
    assert got.shape == (5, 5, 5) and np.max(np.abs(got - ref)) < 1e-9
    # expect / correlation_2op_1t (oqs.py:225-274) contract U(t) on the GPU
    U = s.propagator(t, "SOS")
    V = evecs
    vec = lambda m: m.reshape(-1)
    e = rand_herm(rng, n)
    r_eb = vec(V.conj().T @ rho0 @ V)
    ref_e = np.array([vec(V.conj().T @ e @ V) @ (U_ref[:, :, k] @ r_eb) for k in range(len(t))])
    assert np.max(np.abs(s.expect(rho0, [e])[:, 0] - ref_e)) < 1e-9
    A = osu.operator_to_superoperator(rand_herm(rng, n), "left")
    B = osu.operator_to_superoperator(rand_herm(rng, n), "right")
    idm = vec(np.eye(n))
    ref_c = np.array([idm @ (A @ (-1j * U_ref[:, :, k]) @ (B @ vec(rho0))) for k in range(len(t))])
    assert np.max(np.abs(s.correlation_2op_1t(rho0, A, B, t) - ref_c)) < 1e-9
