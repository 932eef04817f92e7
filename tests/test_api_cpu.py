"""Host-side pieces of the facade that need no GPU: Liouville builders, Result, phys helpers,
and the no-CPU-fallback guarantee."""
import numpy as np
import pytest

from oracle import superop as osu


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_superoperator_builders_match_oracle():
    from lime_amd import superoperator as so
    rng = np.random.default_rng(1)
    n = 4
    H = rand_c(rng, n, n)
    H = H + H.conj().T
    c_ops = [rand_c(rng, n, n), np.diag(rand_c(rng, n))]
    assert np.allclose(so.liouvillian(H, c_ops).toarray(), osu.liouvillian(H, c_ops))
    for kind in ("commutator", "left", "right", "anticommutator", "c", "l", "r", "a", "-", "+"):
        assert np.allclose(so.operator_to_superoperator(H, kind).toarray(),
                           osu.operator_to_superoperator(H, kind))
    assert np.allclose(so.kraus(c_ops[0]).toarray(), osu.kraus(c_ops[0]))
    assert np.allclose(so.lindblad_dissipator(c_ops[0]).toarray(), osu.lindblad_dissipator(c_ops[0]))
    with pytest.raises(ValueError):
        so.op2sop(H, "nope")
    x = rand_c(rng, n, n)
    v = so.dm2vec(x)
    assert np.allclose(so.obs(v, H), np.trace(H @ x))
    assert np.allclose(so.trace(v), np.trace(x))


def test_result_times_fixed():
    from lime_amd.mol import Result
    r = Result(dt=0.1, Nt=10, rho0=None)
    assert np.allclose(r.times, 0.1 * np.arange(1, 11))
    r = Result(dt=0.1, Nt=10, t0=1.0, nout=5)
    assert np.allclose(r.times, [1.5, 2.0])


def test_phys_helpers():
    from lime_amd import phys
    rng = np.random.default_rng(2)
    a, b = rand_c(rng, 5, 5), rand_c(rng, 5, 5)
    assert np.allclose(phys.obs_dm(a, b), np.trace(b @ a))
    assert np.allclose(phys.comm(a, b), a @ b - b @ a)
    assert np.allclose(phys.transform(a, b), b.conj().T @ a @ b)
    x = a.copy()
    phys.rk4(x, lambda r: -1j * r, 0.1)
    assert np.allclose(x, a * (1 + (-0.1j) + (-0.1j) ** 2 / 2 + (-0.1j) ** 3 / 6 + (-0.1j) ** 4 / 24))


def test_no_cpu_fallback_without_gpu():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from lime_amd.oqs import Lindblad_solver
    with pytest.raises(RuntimeError, match="GPU"):
        Lindblad_solver(np.eye(2), []).evolve(np.eye(2) / 2, 0.1, 2)


def test_coherent_states_against_lime_docstring_values():
    """lime's oqs.py:885-892 and :951-957 quote coherent(5, 0.25j) and coherent_dm(3, 0.25j);
    the untruncated Fock coefficients e^{-|a|^2/2} a^n / sqrt(n!) fix a larger space."""
    from math import factorial

    from lime_amd.oqs import coherent, coherent_dm
    psi = coherent(5, 0.25j)
    quoted = np.array([9.69233235e-01, 0.24230831j, -4.28344935e-02, -0.00618204j, 7.80904967e-04])
    assert np.max(np.abs(psi - quoted)) < 1e-8
    dm = coherent_dm(3, 0.25j)
    quoted_dm = np.array([[0.93941695, -0.23480733j, -0.04216943],
                          [0.23480733j, 0.05869011, -0.01054025j],
                          [-0.04216943, 0.01054025j, 0.00189294]])
    assert np.max(np.abs(dm - quoted_dm)) < 1e-8
    assert abs(np.trace(dm) - 1) < 1e-13 and np.allclose(dm, dm.conj().T)
    alpha, N = 0.7 - 0.4j, 40
    exact = np.array([np.exp(-abs(alpha) ** 2 / 2) * alpha ** k / np.sqrt(float(factorial(k))) for k in range(N)])
    assert np.max(np.abs(coherent(N, alpha) - exact)) < 1e-12
    assert abs(np.linalg.norm(coherent(6, 1.5)) - 1) < 1e-13        # truncated: still normalised
