"""Pins the FP64 oracle before anything is compared against it.

Each check is chosen so that a plausible slip (commutator sign, a dropped 1/2 in the
anticommutator, a transposed kron factor, an off-by-one in the time grid, a missing
conjugate) fails at least one of them:

* reference-held golden vectors (lime examples/cor.dat, dm.dat);
* exact discrete RK4 closed form for diagonal H and L;
* an independent Liouville-space (kron) construction, incl. dense L and general H;
* scipy.linalg.expm in the dt -> 0 limit (RK4 is 4th order);
* invariants (trace, Hermiticity, positivity);
* analytic geometric-series transform for the FFT convention.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import lindblad as ol
from oracle import redfield as orf
from oracle import spectra as osp
from oracle import superop as osu


def cavity_model(n=10, kappa=0.25, n_th=2.0):
    a = np.diag(np.sqrt(np.arange(1, n)), 1).astype(np.complex128)
    num = np.diag(np.arange(n)).astype(np.complex128)
    H = np.diag(np.arange(n) * 1.0).astype(np.complex128)
    c_ops = [np.sqrt(kappa * (1 + n_th)) * a, np.sqrt(kappa * n_th) * ol.dag(a)]
    w = np.exp(-np.arange(n) / n_th)
    rho0 = np.diag(w / w.sum()).astype(np.complex128)
    return H, a, num, c_ops, rho0


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rand_herm(rng, n):
    x = rand_c(rng, n, n)
    return (x + x.conj().T) / 2


def rand_dm(rng, n):
    x = rand_c(rng, n, n)
    r = x @ x.conj().T
    return r / np.trace(r)


# ---------------------------------------------------------------- golden vectors
def test_cavity_golden_vectors(golden_dir):
    g = np.load(os.path.join(golden_dir, "cavity_qrt.npz"))
    H, a, num, c_ops, rho0 = cavity_model()
    tlist = np.linspace(0, 10, 200)
    dt = tlist[1] - tlist[0]
    t, cor, dm = ol.correlation_3p_1t(H, rho0, [ol.dag(a), num, a], c_ops, dt, len(tlist))
    assert np.max(np.abs(t - g["t_cor"])) < 1e-12
    assert np.max(np.abs(cor - g["cor"])) < 1e-13
    assert np.max(np.abs(dm - g["dm"])) < 1e-14
    # the H_eff-folded form and the reference operation sequence agree on the same run
    x0 = a @ rho0 @ ol.dag(a)
    _, obs, states = ol.propagate(H, x0, c_ops, dt, 200, [num], store_states=True)
    assert np.max(np.abs(obs[:, 0] - g["cor"])) < 1e-13
    assert np.max(np.abs(np.array(states) - g["dm"])) < 1e-14


def test_cavity_initial_value_closed_form(golden_dir):
    g = np.load(os.path.join(golden_dir, "cavity_qrt.npz"))
    H, a, num, c_ops, rho0 = cavity_model()
    c0 = np.trace(num @ a @ rho0 @ ol.dag(a)).real     # Tr[n a rho0 a^dag] at t = 0
    assert abs(c0 - 3.9327) < 1e-4
    # first recorded sample sits at t = dt (not t = 0): it is already below C(0)
    assert g["cor"][0].real < c0 and abs(g["cor"][0].real - c0) < 0.02


# ---------------------------------------------------------------- closed forms
@pytest.mark.parametrize("form", ["ref", "heff"])
def test_rk4_closed_form_diagonal(form):
    rng = np.random.default_rng(1)
    n, dt, nt = 6, 0.07, 50
    E = rng.standard_normal(n)
    gam = rng.uniform(0.05, 0.3, n)
    H = np.diag(E).astype(np.complex128)
    c_ops = [np.sqrt(g) * np.diag(np.eye(n)[i]) for i, g in enumerate(gam)]
    rho0 = rand_dm(rng, n)
    lam = -1j * (E[:, None] - E[None, :]) - 0.5 * (gam[:, None] + gam[None, :]) + np.diag(gam)
    expect = ol.rk4_stability_factor(lam * dt) ** nt * rho0
    rho, _, _ = ol.propagate(H, rho0, c_ops, dt, nt, form=form)
    assert np.max(np.abs(rho - expect)) < 1e-14


def test_heff_matches_reference_sequence_dense():
    rng = np.random.default_rng(2)
    n = 7
    H = rand_herm(rng, n)
    c_ops = [0.3 * rand_c(rng, n, n) for _ in range(3)] + \
            [np.sqrt(0.1) * np.diag(np.eye(n)[i]) for i in range(n)] + \
            [np.diag(rand_c(rng, n))]                      # collective diagonal operator
    rho = rand_c(rng, n, n)                                 # non-Hermitian (QRT-like)
    Heff, D, dense = ol.split_ops(H, c_ops)
    assert len(dense) == 3
    a = ol.liouvillian_ref(rho, H, c_ops)
    b = ol.rhs_heff(rho, Heff, D, dense)
    assert np.max(np.abs(a - b)) < 1e-12


def test_liouville_kron_convention():
    rng = np.random.default_rng(3)
    n = 5
    A, X, B = rand_c(rng, n, n), rand_c(rng, n, n), rand_c(rng, n, n)
    lhs = osu.dm2vec(A @ X @ B)
    rhs = np.kron(A, B.T) @ osu.dm2vec(X)
    assert np.max(np.abs(lhs - rhs)) < 1e-12
    assert np.allclose(osu.left(A) @ osu.dm2vec(X), osu.dm2vec(A @ X))
    assert np.allclose(osu.right(B) @ osu.dm2vec(X), osu.dm2vec(X @ B))
    assert np.allclose(osu.kraus(A) @ osu.dm2vec(X), osu.dm2vec(A @ X @ A.conj().T))
    with pytest.raises(ValueError):
        osu.operator_to_superoperator(A, "bogus")


@pytest.mark.parametrize("form", ["ref", "heff"])
def test_hilbert_rk4_equals_liouville_rk4(form):
    rng = np.random.default_rng(4)
    n, dt, nt = 6, 0.05, 40
    H = rand_herm(rng, n)
    c_ops = [0.2 * rand_c(rng, n, n), 0.1 * rand_c(rng, n, n),
             np.sqrt(0.05) * np.diag(np.eye(n)[2])]
    rho0 = rand_dm(rng, n)
    L = osu.liouvillian(H, c_ops)
    ref = osu.propagate_rk4_liouville(L, rho0, dt, nt)
    _, _, states = ol.propagate(H, rho0, c_ops, dt, nt, store_states=True, form=form)
    assert np.max(np.abs(np.array(states) - ref)) < 1e-13


def test_rk4_fourth_order_against_expm():
    rng = np.random.default_rng(5)
    n, T = 4, 2.0
    H = rand_herm(rng, n)
    c_ops = [0.3 * rand_c(rng, n, n), np.sqrt(0.2) * np.diag(np.eye(n)[0])]
    rho0 = rand_dm(rng, n)
    exact = osu.propagate_exact(osu.liouvillian(H, c_ops), rho0, T)
    errs = []
    for nt in (40, 80, 160):
        rho, _, _ = ol.propagate(H, rho0, c_ops, T / nt, nt)
        errs.append(np.max(np.abs(rho - exact)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 13 < r1 < 19 and 13 < r2 < 19, errs


def test_invariants_trace_hermiticity_positivity():
    rng = np.random.default_rng(6)
    n = 8
    H = rand_herm(rng, n)
    c_ops = [0.2 * rand_c(rng, n, n)] + [np.sqrt(0.05) * np.diag(np.eye(n)[i]) for i in range(n)]
    rho0 = rand_dm(rng, n)
    _, _, states = ol.propagate(H, rho0, c_ops, 0.02, 300, store_states=True)
    for rho in states[::30]:
        assert abs(np.trace(rho) - 1) < 1e-13
        assert np.max(np.abs(rho - rho.conj().T)) < 1e-14
        assert np.min(np.linalg.eigvalsh((rho + rho.conj().T) / 2)) > -1e-12


# ---------------------------------------------------------------- correlation functions
def test_correlations_against_eseries():
    rng = np.random.default_rng(7)
    n, dt, nt = 4, 0.005, 400
    H = rand_herm(rng, n)
    c_ops = [0.3 * rand_c(rng, n, n), np.sqrt(0.1) * np.diag(np.eye(n)[1])]
    rho0 = rand_dm(rng, n)
    A, B, C = rand_c(rng, n, n), rand_c(rng, n, n), rand_c(rng, n, n)
    es = osu.Eseries(H, c_ops)
    t = dt * np.arange(1, nt + 1)                     # QRT samples sit at t_k = (k+1) dt
    c2 = ol.correlation_2op_1t(H, rho0, A, B, c_ops, dt, nt)
    assert np.max(np.abs(c2 - es.correlation_2op_1t(rho0, [A, B], t))) < 1e-8
    c3 = ol.correlation_3op_1t(H, rho0, A, B, C, c_ops, dt, nt)
    assert np.max(np.abs(c3 - es.correlation_3op_1t(rho0, [A, B, C], t))) < 1e-8
    # eseries observables vs RK4 states
    _, obs, _ = ol.propagate(H, rho0, c_ops, dt, nt, [A])
    assert np.max(np.abs(obs[:, 0] - es.evolve(rho0, t, [A])[:, 0])) < 1e-8


def test_two_time_correlation_against_eseries():
    rng = np.random.default_rng(8)
    n, dt, nt, ntau = 3, 0.01, 30, 40
    H = rand_herm(rng, n)
    c_ops = [0.3 * rand_c(rng, n, n)]
    rho0 = rand_dm(rng, n)
    A, B, C = rand_c(rng, n, n), rand_c(rng, n, n), rand_c(rng, n, n)
    rk = ol.correlation_3op_2t(H, rho0, [A, B, C], c_ops, dt, nt, ntau)
    es = osu.Eseries(H, c_ops).correlation_3op_2t(
        rho0, [A, B, C], dt * np.arange(1, nt + 1), dt * np.arange(1, ntau + 1))
    assert rk.shape == (nt, ntau)
    assert np.max(np.abs(rk - es)) < 1e-6          # RK4 truncation at dt = 0.01


def test_eseries_frequency_domain_matches_time_integral():
    rng = np.random.default_rng(9)
    n = 3
    H = rand_herm(rng, n)
    c_ops = [0.5 * rand_c(rng, n, n)] + [np.eye(n) * 0.0 + np.diag(np.sqrt([0.3, 0.3, 0.3]))]
    rho0 = rand_dm(rng, n)
    A, B = rand_c(rng, n, n), rand_c(rng, n, n)
    es = osu.Eseries(H, c_ops)
    # steady-state part of <A(t)B> does not decay; subtract it before integrating
    w0 = np.argmin(np.abs(es.eigvals))
    t = np.linspace(0, 80, 160001)
    ct = es.correlation_2op_1t(rho0, [A, B], t)
    om = 0.7
    num = np.trapezoid((ct - ct[-1]) * np.exp(1j * om * t), t)
    full = es.correlation_2op_1w(rho0, [A, B], [om])[0]
    # analytic form: sum_n coeff_n * (-1/(lam_n + i om)); drop the zero mode
    lam = es.eigvals
    w = np.conj(es.idv) @ (osu.left(A) @ es.U1)
    coeff = w * es._coeff_in(B @ rho0)
    mask = np.arange(len(lam)) != w0
    assert abs(full - np.sum(-1.0 / (lam + 1j * om) * coeff)) < 1e-10
    assert abs(num - np.sum((-1.0 / (lam + 1j * om) * coeff)[mask])) < 2e-4


# ---------------------------------------------------------------- Redfield
def ohmic(eta=0.05, wc=2.0, beta=2.0):
    def S(w):
        if abs(w) < 1e-12:
            return eta / beta
        return eta * w * np.exp(-abs(w) / wc) / (1 - np.exp(-beta * w))
    return S


def test_redfield_liouville_equals_hilbert_form():
    rng = np.random.default_rng(10)
    n = 5
    H = rand_herm(rng, n)
    a_ops = [rand_herm(rng, n), rand_herm(rng, n)]
    spectra = [ohmic(), ohmic(0.02, 1.0, 1.0)]
    R, evecs = orf.redfield_tensor(H, a_ops, spectra)
    E, _, A, Lam = orf.bath_matrices(H, a_ops, spectra)
    rho = rand_c(rng, n, n)
    lhs = (R @ rho.flatten()).reshape(n, n)
    rhs = orf.rhs_hilbert(rho, E, A, Lam)
    assert np.max(np.abs(lhs - rhs)) < 1e-12
    with pytest.raises(TypeError):
        orf.redfield_tensor(H, [rand_c(rng, n, n)], [ohmic()])


def test_redfield_population_rates_golden_rule():
    rng = np.random.default_rng(11)
    n = 4
    H = rand_herm(rng, n)
    a = rand_herm(rng, n)
    S = ohmic()
    R, evecs = orf.redfield_tensor(H, [a], [S])
    E = np.linalg.eigvalsh(H)
    A = evecs.conj().T @ a @ evecs
    # dP_m/dt contribution from P_n (n != m): R[(m,m),(n,n)] = 2 |A_mn|^2 S(E_n - E_m)
    for m in range(n):
        for k in range(n):
            if m == k:
                continue
            expect = 2 * abs(A[m, k]) ** 2 * S(E[k] - E[m])
            assert abs(R[m * n + m, k * n + k] - expect) < 1e-12


def test_redfield_rk4_trace_and_backtransform():
    rng = np.random.default_rng(12)
    n = 4
    H = rand_herm(rng, n)
    a_ops = [rand_herm(rng, n)]
    R, evecs = orf.redfield_tensor(H, a_ops, [ohmic()])
    rho0 = rand_dm(rng, n)
    states, obs = orf.propagate(R, evecs, rho0, 0.02, 200, [np.eye(n)])
    assert np.max(np.abs(obs[:, 0] - 1)) < 1e-12
    assert abs(np.trace(states[-1]) - 1) < 1e-12
    # the detailed-balance bath drives the populations towards Boltzmann-like order
    pops = np.real(np.diag(evecs.conj().T @ states[-1] @ evecs))
    assert np.all(pops > -1e-9)


def test_redfield_unitary_limit_in_the_site_basis():
    """With a silent bath (S = 0) the Redfield tensor is the bare commutator in the H
    eigenbasis, so ``propagate`` (eigenbasis RK4, states transformed back) must reproduce
    site-basis dynamics built without any basis change: the Hilbert-space RK4 of -i[H, rho]
    (pinned by the cavity golden vectors and the diagonal closed form) step for step, and
    e^{-iHt} rho0 e^{iHt} to the RK4 truncation error.  Observables are taken with
    non-diagonal site operators.  Reversing the basis change (V rho V^dag on the way in,
    V^dag rho V on the way out) fails every check here at a non-diagonal H."""
    rng = np.random.default_rng(13)
    n = 5
    H = rand_herm(rng, n)
    a_ops = [rand_herm(rng, n)]
    R, evecs = orf.redfield_tensor(H, a_ops, [lambda w: 0.0])
    assert np.max(np.abs(evecs - np.eye(n))) > 0.1          # genuinely non-diagonal H
    rho0 = rand_dm(rng, n)
    e_ops = [rand_herm(rng, n), rand_c(rng, n, n)]
    dt, nt = 0.01, 150
    states, obs = orf.propagate(R, evecs, rho0, dt, nt, e_ops)
    _, obs_site, st_site = ol.propagate(H, rho0, [], dt, nt, e_ops=e_ops, store_states=True)
    assert np.max(np.abs(states - np.array(st_site))) < 1e-12
    assert np.max(np.abs(obs - obs_site)) < 1e-12
    for k in (0, 49, nt - 1):
        U = sla.expm(-1j * H * dt * (k + 1))
        exact = U @ rho0 @ U.conj().T
        assert np.max(np.abs(states[k] - exact)) < 1e-7     # RK4 truncation: 1.6e-8 here


def test_superop_trace_and_obs_on_non_hermitian_vectors():
    """``trace`` = Tr rho and ``obs`` = Tr[a rho] (lime superoperator.py:314-319) on a
    complex, non-Hermitian rho and a: a conjugated rho (vdot argument order), a transposed
    or conjugated a, or a column-major vec fails one of these."""
    rng = np.random.default_rng(14)
    n = 4
    rho = rand_c(rng, n, n)
    a = rand_c(rng, n, n)
    v = osu.dm2vec(rho)
    assert abs(osu.trace(v) - np.trace(rho)) < 1e-13
    assert abs(osu.obs(v, a) - np.trace(a @ rho)) < 1e-12
    assert abs(osu.obs(v, a) - np.trace(a.T @ rho)) > 1e-3       # the test can tell a from a^T


# ---------------------------------------------------------------- FFT convention
def test_fft_against_geometric_series():
    n, dt, x0 = 256, 0.1, 0.3
    w0, g = 1.3, 0.2
    x = x0 + dt * np.arange(n)
    f = np.exp(-(1j * w0 + g) * x)
    G, freq = osp.fft(f, x)
    # closed form of sum_k dt exp(-(i w0 + g) x_k) exp(-i w x_k)
    z = np.exp(-(1j * (w0 + freq) + g) * dt)
    expect = dt * np.exp(-(1j * (w0 + freq) + g) * x0) * (1 - z ** n) / (1 - z)
    assert np.max(np.abs(G - expect)) < 1e-12
    # ifft convention: (1/2pi) sum_k f(w_k) e^{+i w t} dw
    h, fr = osp.ifft(f, x)
    zz = np.exp((1j * (fr - w0) - g) * dt)
    expect_i = dt / (2 * np.pi) * np.exp((1j * (fr - w0) - g) * x0) * (1 - zz ** n) / (1 - zz)
    assert np.max(np.abs(h - expect_i)) < 1e-12


def test_dephasing_lorentzian_spectrum():
    """Uncoupled site with dephasing: <mu(t) mu> decays at gamma/2 -> Lorentzian of HWHM gamma/2."""
    eps, gamma, dt, nt = 0.4, 0.1, 0.05, 4000
    # basis: |g>=0, |e>=1
    H = np.diag([0.0, eps]).astype(np.complex128)
    c_ops = [np.sqrt(gamma) * np.diag([0.0, 1.0]).astype(np.complex128)]
    mu = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    rho0 = np.diag([1.0, 0.0]).astype(np.complex128)
    cor = ol.correlation_2op_1t(H, rho0, mu, mu, c_ops, dt, nt)
    t = dt * np.arange(1, nt + 1)
    lam = -1j * eps - gamma / 2
    expect = ol.rk4_stability_factor(lam * dt) ** np.arange(1, nt + 1)
    assert np.max(np.abs(cor - expect)) < 1e-13
    assert np.max(np.abs(cor - np.exp(lam * t))) < 1e-6


def test_driven_samples_field_at_step_end():
    """Diagonal H0 and drive, no dissipation: each element picks up exactly
    R(-i (w_ij - f(t_{k+1}) d_ij) dt) per step -- pins the t_{k+1} sampling and the sign."""
    rng = np.random.default_rng(13)
    n, dt, nt = 4, 0.05, 60
    E = rng.standard_normal(n)
    d = rng.standard_normal(n)
    f = lambda t: 0.7 * np.cos(1.3 * t)
    rho0 = rand_dm(rng, n)
    H = [np.diag(E), [np.diag(d), f]]
    H0_copy = H[0].copy()
    rho, _, _ = ol.propagate_driven(H, rho0, [], dt, nt)
    expect = rho0.copy()
    for k in range(nt):
        t = (k + 1) * dt
        e = E - f(t) * d
        expect = expect * ol.rk4_stability_factor(-1j * (e[:, None] - e[None, :]) * dt)
    assert np.max(np.abs(rho - expect)) < 1e-14
    assert np.array_equal(H[0], H0_copy)                 # H0 is not mutated
    # a constant drive reduces to the time-independent propagation of H0 - c H1
    c_ops = [0.2 * rand_c(rng, n, n)]
    r1, _, _ = ol.propagate_driven([np.diag(E), [np.diag(d), lambda t: 0.3]], rho0, c_ops, dt, 20)
    r2, _, _ = ol.propagate(np.diag(E) - 0.3 * np.diag(d), rho0, c_ops, dt, 20)
    assert np.max(np.abs(r1 - r2)) < 1e-13


def test_propagator_expm_eigen_and_semigroup():
    """exp(R t) against V e^{Lambda t} V^-1 of a random non-normal R, U(0) = 1 and the
    semigroup law U(t1 + t2) = U(t1) U(t2)."""
    rng = np.random.default_rng(14)
    m = 9
    R = 0.3 * rand_c(rng, m, m)
    t = np.array([0.0, 0.4, 1.1, 1.5])
    U = orf.propagator_expm(R, t)
    assert U.shape == (m, m, 4)
    lam, V = np.linalg.eig(R)
    Vi = np.linalg.inv(V)
    for k, tk in enumerate(t):
        expect = (V * np.exp(lam * tk)) @ Vi
        assert np.max(np.abs(U[:, :, k] - expect)) < 1e-11
    assert np.max(np.abs(U[:, :, 0] - np.eye(m))) < 1e-15
    assert np.max(np.abs(U[:, :, 3] - U[:, :, 1] @ U[:, :, 2])) < 1e-11


def test_response_4op_3t_time_ordering_closed_form():
    """Two levels, R diagonal in the row-major Liouville basis with distinct rates, all four
    superoperators sigma_x from the left or right.  The pathway |0><0| -d-> |1><0| (t_k)
    -c-> |1><1| (t_j) -b-> |0><1| (t_i) -a-> |1><1| gives (-i)^3 exp(l01 t_i + l11 t_j + l10 t_k):
    pins which time goes with which propagator, the -i of G and the left/right (kron)
    convention."""
    sx = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    eye = np.eye(2)
    left = np.kron(sx, eye)           # vec(A X) for row-major vec
    right = np.kron(eye, sx.T)        # vec(X A)
    lam = np.array([-0.05, -0.3 - 1.1j, -0.2 + 0.7j, -0.45])   # l00, l01, l10, l11
    R = np.diag(lam)
    rho0 = np.diag([1.0, 0.0]).astype(np.complex128)
    t = np.array([0.0, 0.3, 0.8, 1.7, 2.2])
    out = orf.response_4op_3t(R, rho0, [left, left, right, left], t)
    ti, tj, tk = np.meshgrid(t, t, t, indexing="ij")
    expect = (-1j) ** 3 * np.exp(lam[1] * ti + lam[3] * tj + lam[2] * tk)
    assert np.max(np.abs(out - expect)) < 1e-13
    # R = 0: G = -i for every time, and the response is (-i)^3 Tr[unvec(a b c d vec rho0)]
    rng = np.random.default_rng(15)
    ops = [rand_c(rng, 9, 9) for _ in range(4)]
    rho_r = rand_dm(rng, 3)
    r0 = orf.response_4op_3t(np.zeros((9, 9)), rho_r, ops, np.array([0.0, 1.0]))
    v = ops[0] @ ops[1] @ ops[2] @ ops[3] @ rho_r.reshape(-1)
    assert np.max(np.abs(r0 - (-1j) ** 3 * np.trace(v.reshape(3, 3)))) < 1e-12


def test_fft2_separable_geometric_series():
    """A separable damped exponential on an nx x nx grid transforms to the product of two
    geometric-series closed forms; lime's fft2 takes the y frequencies from nx (lime/fft.py:100),
    which a rectangular grid exposes."""
    n, dx, dy = 64, 0.1, 0.25
    a, b, wa, wb = 0.3, 0.15, 1.2, -0.7
    x = dx * np.arange(n)
    y = dy * np.arange(n)
    f = np.outer(np.exp(-(a + 1j * wa) * x), np.exp(-(b + 1j * wb) * y))
    fx, fy, G = osp.fft2(f, dx, dy)
    kk = np.arange(-n // 2, n // 2)
    assert np.allclose(fx, 2 * np.pi * kk / (n * dx), atol=1e-12)
    assert np.allclose(fy, 2 * np.pi * kk / (n * dy), atol=1e-12)
    zx = np.exp(-(a + 1j * (wa + fx)) * dx)
    zy = np.exp(-(b + 1j * (wb + fy)) * dy)
    expect = np.outer(dx * (1 - zx ** n) / (1 - zx), dy * (1 - zy ** n) / (1 - zy))
    assert np.max(np.abs(G - expect)) < 1e-11
    fx2, fy2, _ = osp.fft2(np.ones((8, 4)), 1.0, 1.0)
    assert len(fx2) == 8 and len(fy2) == 8
