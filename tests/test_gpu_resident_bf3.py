"""GPU parity of the resident 3xBF16 kernel (n = 65, bordered: 64 sites plus the ground
state) against the FP64 oracle.

Every real operand is split x = hi + lo (BF16 pair, round to nearest even) and every real
product is hi*hi + hi*lo + lo*hi in FP32: the representation error is <= 2^-17 |x| per
operand and the dropped lo*lo term <= 2^-16 |a||b| per term, so one right-hand side is
accurate to ~1e-5 relative and, over the short runs here (dt = 0.02, ||H|| = 1), the state
error stays below 1e-5 of max|rho|.  A kernel that kept only hi*hi (a plausible slip) is
~1e-3 off; a wrong sign or a transposed operand in any border product is O(1) off in the
last row or column, which these comparisons cover element by element."""
import ctypes

import numpy as np
import pytest

from oracle import lindblad as ol

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

N = 65


def _prop():
    from lime_amd.propagator import LindbladPropagator
    return LindbladPropagator


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rand_herm(rng, n, scale=1.0):
    x = rand_c(rng, n, n)
    h = (x + x.conj().T) / 2
    return h * (scale / np.max(np.abs(np.linalg.eigvalsh(h))))


def rand_dm(rng, n):
    x = rand_c(rng, n, n)
    r = x @ x.conj().T
    return r / np.trace(r)


def test_runs_on_the_resident_kernel():
    """n = 65 in 3xBF16 without dense operators needs no workspace (the resident kernel)
    and has one launch shape; FORCE_TILED still selects the per-stage kernels."""
    from lime_amd import _lib
    rng = np.random.default_rng(1)
    H = rand_herm(rng, N)
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(N)[i]) for i in range(N)]
    prop = _prop()(H, c_ops, precision="3xbf16")
    d = prop._descriptor(1)
    assert prop._lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) == 0
    assert prop.variants() == 1
    d.flags |= _lib.LIME_F_FORCE_TILED
    assert prop._lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) > 0


@pytest.mark.parametrize("n_e", [1, 2])
def test_matches_oracle_with_records(n_e):
    """Random dense H, on-site dephasing, a density matrix and a non-Hermitian QRT-like
    operator; observables every 2 steps and the stored states against the oracle."""
    rng = np.random.default_rng(40 + n_e)
    H = rand_herm(rng, N)
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(N)[i]) for i in range(N)]
    X0 = np.stack([rand_dm(rng, N), rand_c(rng, N, N) / N])
    A = [rand_herm(rng, N) for _ in range(n_e)]
    dt, nt = 0.02, 10
    prop = _prop()(H, c_ops, group=2, precision="3xbf16")
    rho, obs, states = prop.run(X0, dt, nt, e_ops=A, nout=2, store_states=True)
    for b in range(2):
        ref, ref_obs, ref_states = ol.propagate(H, X0[b], c_ops, dt, nt, A, nout=2, store_states=True)
        scale = np.max(np.abs(ref))
        got = rho[b].cpu().numpy()
        err = np.max(np.abs(got - ref)) / scale
        border = max(np.max(np.abs(got[-1] - ref[-1])), np.max(np.abs(got[:, -1] - ref[:, -1]))) / scale
        print(f"resident 3xbf16 n_e={n_e} b={b}: rel err {err:.3e} (border row/column {border:.3e})")
        assert err < 1e-5
        assert np.max(np.abs(obs[b].cpu().numpy() - ref_obs)) < 1e-5 * np.max(np.abs(ref_obs))
        assert np.max(np.abs(states[b].cpu().numpy() - np.array(ref_states))) < 1e-5 * scale


def test_matches_tiled_3xbf16_and_renormalises():
    """The same split arithmetic on the per-stage kernels agrees to the split rounding;
    trace renormalisation keeps Re Tr rho = 1 and the border corner enters the trace."""
    rng = np.random.default_rng(77)
    H = rand_herm(rng, N)
    c_ops = [np.sqrt(0.05) * np.diag(np.eye(N)[i]) for i in range(N)]
    rho0 = rand_dm(rng, N)
    res = _prop()(H, c_ops, precision="3xbf16")
    til = _prop()(H, c_ops, precision="3xbf16", force_tiled=True)
    r1, _, _ = res.run(rho0, 0.05, 200)
    r2, _, _ = til.run(rho0, 0.05, 200)
    ref, _, _ = ol.propagate(H, rho0, c_ops, 0.05, 200)
    a, t = r1[0].cpu().numpy(), r2[0].cpu().numpy()
    print(f"resident vs tiled 3xbf16 {np.max(np.abs(a - t)):.3e}; vs oracle {np.max(np.abs(a - ref)):.3e}")
    assert np.max(np.abs(a - ref)) < 1e-5
    assert np.max(np.abs(a - t)) < 1e-5
    rr, _, _ = res.run(rho0, 0.05, 200, renorm=True)
    r = rr[0].cpu().numpy()
    assert abs(np.trace(r).real - 1) < 1e-14
    assert np.max(np.abs(r - r.conj().T)) < 1e-5


def test_chain_populations_long_run():
    """64-site chain plus ground state (the bordered layout's purpose), 2000 steps of
    dt = 0.05 from a site excitation: population error against the oracle <= 1e-6."""
    rng = np.random.default_rng(5)
    n = N
    H = np.zeros((n, n), complex)
    H[np.arange(1, n), np.arange(1, n)] = rng.normal(0, 0.05, n - 1)
    for i in range(1, n - 1):
        H[i, i + 1] = H[i + 1, i] = -0.1
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(1, n)]
    rho0 = np.zeros((n, n), complex)
    rho0[1, 1] = 1
    P = [np.diag(np.eye(n)[i]).astype(complex) for i in range(n)]
    dt, nt = 0.05, 2000
    _, ref, _ = ol.propagate(H, rho0, c_ops, dt, nt, P, nout=20)
    _, obs, _ = _prop()(H, c_ops, precision="3xbf16").run(rho0, dt, nt, e_ops=P, nout=20)
    err = np.max(np.abs(obs[0].cpu().numpy() - ref))
    print(f"resident 3xbf16 chain populations over {nt} steps: max error {err:.3e}")
    assert err < 1e-6


def test_chunked_and_strided_records_equal_single_launch():
    """Records every 3rd step and launches of 7 steps: the same bits as one launch (the step
    offset and the record index are carried across launches)."""
    rng = np.random.default_rng(91)
    H = rand_herm(rng, N)
    c_ops = [np.sqrt(0.03) * np.diag(np.eye(N)[i]) for i in range(N)]
    X0 = np.stack([rand_dm(rng, N), rand_c(rng, N, N) / N, rand_dm(rng, N)])
    A = [rand_herm(rng, N), rand_c(rng, N, N)]
    prop = _prop()(H, c_ops, group=3, precision="3xbf16")
    r1, o1, s1 = prop.run(X0, 0.02, 30, e_ops=A, nout=3, store_states=True)
    r2, o2, s2 = prop.run(X0, 0.02, 30, e_ops=A, nout=3, store_states=True, chunk=7)
    assert torch.equal(r1, r2) and torch.equal(o1, o2) and torch.equal(s1, s2)
    ref, ref_obs, _ = ol.propagate(H, X0[1], c_ops, 0.02, 30, A, nout=3)
    assert np.max(np.abs(o1[1].cpu().numpy() - ref_obs)) < 1e-5 * np.max(np.abs(ref_obs))


def test_renormalisation_records_after_scaling():
    """Trace renormalisation with records: Re Tr rho = 1 after every step, the recorded
    states and observables are those of the renormalised state (the border corner included),
    and the physics agrees with the trace-conserving FP64 oracle."""
    rng = np.random.default_rng(123)
    H = rand_herm(rng, N)
    c_ops = [np.sqrt(0.05) * np.diag(np.eye(N)[i]) for i in range(N)]
    rho0 = rand_dm(rng, N)
    E = [np.diag(np.eye(N)[i]).astype(complex) for i in (0, N // 2, N - 1)]
    dt, nt = 0.05, 40
    rho, obs, states = _prop()(H, c_ops, precision="3xbf16").run(rho0, dt, nt, e_ops=E, nout=8,
                                                                  store_states=True, renorm=True)
    r, st, ob = rho[0].cpu().numpy(), states[0].cpu().numpy(), obs[0].cpu().numpy()
    assert abs(np.trace(r).real - 1) < 1e-13
    for k in range(st.shape[0]):
        assert abs(np.trace(st[k]).real - 1) < 1e-13
        for e in range(len(E)):
            assert abs(ob[k, e] - np.trace(E[e] @ st[k])) < 1e-12
    assert np.allclose(st[-1], r, rtol=0, atol=0)
    ref, _, _ = ol.propagate(H, rho0, c_ops, dt, nt)
    assert np.max(np.abs(r - ref)) / np.max(np.abs(ref)) < 2e-6


def test_observable_count_at_the_lds_limit():
    """The record slots of the observables share the kernel's LDS: up to 181 observables run
    on the resident kernel, 182 move to the per-stage kernels (a workspace, the tiled launch
    variants); both record the same values as the oracle."""
    rng = np.random.default_rng(131)
    H = rand_herm(rng, N)
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(N)[i]) for i in range(N)]
    rho0 = rand_dm(rng, N)
    dt, nt = 0.02, 4
    for n_e, resident in ((181, True), (182, False)):
        E = [rand_herm(rng, N) for _ in range(n_e)]
        prop = _prop()(H, c_ops, precision="3xbf16")
        d = prop._descriptor(1)
        d.n_eops = n_e
        d.obs = 256                                  # a non-NULL obs pointer: records requested
        work = prop._lib.lime_lindblad_workspace_bytes(ctypes.byref(d))
        assert (work == 0) == resident, (n_e, work)
        _, obs, _ = prop.run(rho0, dt, nt, e_ops=E, nout=2)
        _, ref, _ = ol.propagate(H, rho0, c_ops, dt, nt, E, nout=2)
        err = np.max(np.abs(obs[0].cpu().numpy() - ref)) / np.max(np.abs(ref))
        print(f"{n_e} observables ({'resident' if resident else 'tiled'}): rel err {err:.3e}")
        assert err < 1e-5
