"""GPU parity of the fused Lindblad RK4 kernel against the FP64 oracle.

Tolerances (max absolute error vs the FP64 oracle) are stated per test.  The kernel
evaluates the right-hand side with complex64 operands on FP32 MFMA (FP32 accumulate)
and keeps the state and its update in FP64, so errors grow like
eps32 * dt * ||L|| * Nt * |rho| at worst.
"""
import os

import numpy as np
import pytest

from oracle import lindblad as ol

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


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


def chain(n, eps, J):
    H = np.diag(eps).astype(np.complex128)
    for i in range(n - 1):
        H[i, i + 1] = H[i + 1, i] = J
    return H


def test_chain_populations_config1():
    """8-site chain, on-site dephasing, rho0 = |0><0|, dt 0.05 x 2000, populations every 10."""
    rng = np.random.default_rng(2024)
    n, dt, nt, nout = 8, 0.05, 2000, 10
    H = chain(n, rng.normal(0, 0.05, n), -0.1)
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    rho0 = np.zeros((n, n), complex)
    rho0[0, 0] = 1
    P = [np.diag(np.eye(n)[i]).astype(complex) for i in range(n)]
    ref_rho, ref_obs, _ = ol.propagate(H, rho0, c_ops, dt, nt, P, nout=nout)
    prop = _prop()(H, c_ops)
    rho, obs, _ = prop.run(rho0, dt, nt, e_ops=P, nout=nout)
    obs = obs[0].cpu().numpy()
    err = np.max(np.abs(obs - ref_obs))
    print(f"config1 max population error {err:.3e}")
    assert obs.shape == (nt // nout, n)
    assert err < 1e-6                       # gate: <= 1e-6 max population error
    assert np.max(np.abs(rho[0].cpu().numpy() - ref_rho)) < 1e-6


@pytest.mark.parametrize("n", [3, 8, 16, 17, 31, 48, 64, 65, 80, 96])
def test_random_dense_hamiltonian_dephasing(n):
    """General Hermitian H, on-site + collective diagonal jump operators, random rho."""
    rng = np.random.default_rng(n)
    H = rand_herm(rng, n)
    c_ops = [np.sqrt(0.05) * np.diag(np.eye(n)[i]) for i in range(n)]
    c_ops.append(0.1 * np.diag(rand_c(rng, n)))        # collective diagonal operator
    rho0 = rand_dm(rng, n)
    dt, nt = 0.02, 100
    ref, _, _ = ol.propagate(H, rho0, c_ops, dt, nt)
    rho, _, _ = _prop()(H, c_ops).run(rho0, dt, nt)
    err = np.max(np.abs(rho[0].cpu().numpy() - ref))
    print(f"n={n} max element error {err:.3e}")
    assert err < 2e-7 * max(1.0, 1.0)


@pytest.mark.parametrize("n", [4, 10, 16, 33, 64, 80])
def test_dense_jump_operators(n):
    rng = np.random.default_rng(100 + n)
    H = rand_herm(rng, n)
    c_ops = [0.1 * rand_c(rng, n, n) / np.sqrt(n) for _ in range(3)]
    c_ops += [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    rho0 = rand_dm(rng, n)
    dt, nt = 0.02, 60
    ref, _, _ = ol.propagate(H, rho0, c_ops, dt, nt)
    prop = _prop()(H, c_ops)
    assert prop.n_dense == 3
    rho, _, _ = prop.run(rho0, dt, nt)
    err = np.max(np.abs(rho[0].cpu().numpy() - ref))
    print(f"n={n} dense-L max element error {err:.3e}")
    assert err < 5e-7


def test_cavity_golden_vectors_on_gpu(golden_dir):
    """lime's examples/cor.dat and dm.dat reproduced by the GPU kernel (dense a, a^dag)."""
    g = np.load(os.path.join(golden_dir, "cavity_qrt.npz"))
    n, kappa, n_th = 10, 0.25, 2.0
    a = np.diag(np.sqrt(np.arange(1, n)), 1).astype(np.complex128)
    num = np.diag(np.arange(n)).astype(np.complex128)
    H = np.diag(np.arange(n) * 1.0).astype(np.complex128)
    c_ops = [np.sqrt(kappa * (1 + n_th)) * a, np.sqrt(kappa * n_th) * a.conj().T]
    w = np.exp(-np.arange(n) / n_th)
    rho0 = np.diag(w / w.sum()).astype(np.complex128)
    dt = 10.0 / 199
    x0 = a @ rho0 @ a.conj().T
    rho, obs, states = _prop()(H, c_ops).run(x0, dt, 200, e_ops=[num], store_states=True)
    cor = obs[0, :, 0].cpu().numpy()
    dm = states[0].cpu().numpy()
    e1 = np.max(np.abs(cor - g["cor"]))
    e2 = np.max(np.abs(dm - g["dm"]))
    print(f"cavity golden: cor err {e1:.3e}, dm err {e2:.3e}")
    assert e1 < 2e-6 and e2 < 5e-7


def test_qrt_non_hermitian_operator_batch_and_groups():
    """Batch of 2 operator sets x 3 QRT operators each (group=3), shared observable set."""
    rng = np.random.default_rng(7)
    n, S, G = 12, 2, 3
    Hs = np.stack([rand_herm(rng, n) for _ in range(S)])
    L = [np.sqrt(0.03) * np.diag(np.eye(n)[i]) for i in range(n)]
    X0 = np.stack([rand_c(rng, n, n) for _ in range(S * G)])
    A = rand_c(rng, n, n)
    dt, nt = 0.05, 80
    prop = _prop()(Hs, L, group=G)
    rho, obs, _ = prop.run(X0, dt, nt, e_ops=[A], nout=4)
    obs = obs.cpu().numpy()
    for b in range(S * G):
        _, ref_obs, _ = ol.propagate(Hs[b // G], X0[b], L, dt, nt, [A], nout=4)
        # |Tr[A X]| ~ 20 here: 1e-6 relative to the largest recorded value
        assert np.max(np.abs(obs[b] - ref_obs)) < 1e-6 * np.max(np.abs(ref_obs))


@pytest.mark.parametrize("n,dense", [(20, True), (33, False), (65, False)])
def test_trace_renormalisation_keeps_unit_trace(n, dense):
    """Resident kernels; n = 33 and 65 without dense operators are the bordered ones (the
    corner element of the border enters the trace)."""
    rng = np.random.default_rng(9 + n)
    H = rand_herm(rng, n)
    c_ops = ([0.2 * rand_c(rng, n, n) / np.sqrt(n)] if dense else []) \
        + [np.sqrt(0.05) * np.diag(np.eye(n)[i]) for i in range(n)]
    rho0 = rand_dm(rng, n)
    rho, _, _ = _prop()(H, c_ops).run(rho0, 0.05, 500, renorm=True)
    r = rho[0].cpu().numpy()
    # Re Tr rho is renormalised to 1 every step; Im Tr rho is roundoff of the FP32 products
    assert abs(np.trace(r).real - 1) < 1e-14
    assert abs(np.trace(r).imag) < 1e-7
    assert np.max(np.abs(r - r.conj().T)) < 1e-6
    ref, _, _ = ol.propagate(H, rho0, c_ops, 0.05, 500)
    assert np.max(np.abs(r - ref)) < 1e-6


def test_chunked_run_equals_single_launch():
    rng = np.random.default_rng(11)
    n = 24
    H = rand_herm(rng, n)
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    rho0 = rand_dm(rng, n)
    P = [np.diag(np.eye(n)[i]).astype(complex) for i in range(3)]
    prop = _prop()(H, c_ops)
    r1, o1, s1 = prop.run(rho0, 0.05, 120, e_ops=P, nout=6, store_states=True)
    r2, o2, s2 = prop.run(rho0, 0.05, 120, e_ops=P, nout=6, store_states=True, chunk=25)
    assert torch.equal(r1, r2) and torch.equal(o1, o2) and torch.equal(s1, s2)


@pytest.mark.parametrize("n,kd,tiled", [(10, 0, True), (33, 2, True), (64, 0, True),
                                         (65, 1, True), (130, 0, False), (150, 2, False)])
def test_tiled_path_matches_oracle(n, kd, tiled):
    """Per-stage tiled kernels (forced at small n, automatic above 80), incl. dense L,
    observables (FP64 atomics) and state records."""
    rng = np.random.default_rng(300 + n)
    H = rand_herm(rng, n)
    c_ops = [0.1 * rand_c(rng, n, n) / np.sqrt(n) for _ in range(kd)]
    c_ops += [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    X0 = np.stack([rand_dm(rng, n), rand_c(rng, n, n)])
    A = [rand_herm(rng, n), np.diag(np.eye(n)[0]).astype(complex)]
    dt, nt = 0.02, 12
    prop = _prop()(H, c_ops, group=2, force_tiled=tiled)
    rho, obs, states = prop.run(X0, dt, nt, e_ops=A, nout=3, store_states=True)
    for b in range(2):
        ref, ref_obs, ref_states = ol.propagate(H, X0[b], c_ops, dt, nt, A, nout=3, store_states=True)
        scale = max(1.0, np.max(np.abs(ref)))
        err = np.max(np.abs(rho[b].cpu().numpy() - ref)) / scale
        print(f"tiled n={n} kd={kd} b={b} rel err {err:.3e}")
        assert err < 2e-7
        assert np.max(np.abs(obs[b].cpu().numpy() - ref_obs)) < 1e-6 * max(1.0, np.max(np.abs(ref_obs)))
        assert np.max(np.abs(states[b].cpu().numpy() - np.array(ref_states))) < 2e-7 * scale


@pytest.mark.parametrize("prec,tol", [("fp32", 5e-7), ("3xbf16", 1e-5)])
def test_sandwich_terms_resident_and_tiled(prec, tol):
    """Explicit (L, R) sandwich terms sum_k L_k X R_k^dag with R != L (3xBF16: the per-stage
    kernels, with L_k and R_k pre-split into their own planes)."""
    from oracle import lindblad as olb
    rng = np.random.default_rng(77)
    for n, tiled in ((12, False), (12, True), (100, False)):
        G = rand_c(rng, n, n) * 0.3
        Ls = [0.2 * rand_c(rng, n, n) / np.sqrt(n) for _ in range(2)]
        Rs = [0.2 * rand_c(rng, n, n) / np.sqrt(n) for _ in range(2)]
        X0 = rand_c(rng, n, n)
        def rhs(x, G=G, Ls=Ls, Rs=Rs):
            out = -1j * (G @ x - x @ G.conj().T)
            for L, R in zip(Ls, Rs):
                out = out + L @ x @ R.conj().T
            return out
        ref = X0.copy()
        for _ in range(10):
            ref = olb.rk4(ref, rhs, 0.05)
        prop = _prop()(G, None, sandwich=(Ls, Rs), force_tiled=tiled, precision=prec)
        rho, _, _ = prop.run(X0, 0.05, 10)
        err = np.max(np.abs(rho[0].cpu().numpy() - ref)) / np.max(np.abs(ref))
        print(f"sandwich n={n} tiled={tiled} {prec} rel err {err:.3e}")
        assert err < tol


@pytest.mark.parametrize("n,kd", [(20, 0), (70, 2), (130, 1)])
def test_fp64_tiled_variant_matches_oracle(n, kd):
    """The FP64-MFMA tiled kernels (FP64 GPU reference path)."""
    rng = np.random.default_rng(500 + n)
    H = rand_herm(rng, n)
    c_ops = [0.1 * rand_c(rng, n, n) / np.sqrt(n) for _ in range(kd)]
    c_ops += [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    rho0 = rand_dm(rng, n)
    A = [rand_herm(rng, n)]
    dt, nt = 0.02, 10
    prop = _prop()(H, c_ops, precision="fp64")
    rho, obs, _ = prop.run(rho0, dt, nt, e_ops=A, nout=2)
    ref, ref_obs, _ = ol.propagate(H, rho0, c_ops, dt, nt, A, nout=2)
    err = np.max(np.abs(rho[0].cpu().numpy() - ref))
    print(f"fp64 n={n} kd={kd} err {err:.3e}")
    assert err < 1e-13
    assert np.max(np.abs(obs[0].cpu().numpy() - ref_obs)) < 1e-12


@pytest.mark.parametrize("n", [17, 20])
def test_batch_not_a_multiple_of_the_xcd_count(n):
    """13 distinct operators (13 = 8 + 5: the workgroup renumbering over the 8 XCDs has a
    full range and a remainder) on the bordered (n = 17) and padded (n = 20) resident
    kernels: every operator lands on its own state and observables, checked against the
    oracle one by one."""
    rng = np.random.default_rng(1300 + n)
    B = 13
    Hs = np.stack([rand_herm(rng, n) for _ in range(B)])
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    X0 = np.stack([rand_c(rng, n, n) for _ in range(B)])
    A = [rand_herm(rng, n)]
    dt, nt = 0.02, 30
    prop = _prop()(Hs, c_ops)
    rho, obs, _ = prop.run(X0, dt, nt, e_ops=A, nout=5)
    rho, obs = rho.cpu().numpy(), obs.cpu().numpy()
    for b in range(B):
        ref, ref_obs, _ = ol.propagate(Hs[b], X0[b], c_ops, dt, nt, A, nout=5)
        scale = max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(rho[b] - ref)) / scale < 2e-7, b
        assert np.max(np.abs(obs[b] - ref_obs)) < 1e-6 * max(1.0, np.max(np.abs(ref_obs))), b


@pytest.mark.parametrize("n,kd", [(8, 1), (30, 2), (48, 0), (64, 2), (80, 0), (96, 0),
                                  (17, 0), (33, 0), (49, 0), (65, 0), (81, 0)])
def test_every_launch_variant_matches_oracle(n, kd):
    """Each resident launch shape (and the tiled tile choices) gives the same physics.
    n = 16 m + 1 without dense operators runs the bordered kernels (core on MFMA, last row
    and column on VALU); the other sizes the padded tiles."""
    rng = np.random.default_rng(900 + n)
    H = rand_herm(rng, n)
    c_ops = [0.1 * rand_c(rng, n, n) / np.sqrt(n) for _ in range(kd)]
    c_ops += [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    X0 = np.stack([rand_dm(rng, n), rand_c(rng, n, n)])
    A = [rand_herm(rng, n)]
    dt, nt = 0.02, 20
    refs = [ol.propagate(H, X0[b], c_ops, dt, nt, A, nout=5) for b in range(2)]
    prop = _prop()(H, c_ops, group=2)
    nv = prop.variants(n_eops=1)
    assert nv >= 1
    for v in range(nv):
        prop.variant = v
        rho, obs, _ = prop.run(X0, dt, nt, e_ops=A, nout=5)
        for b in range(2):
            ref, ref_obs, _ = refs[b]
            scale = max(1.0, np.max(np.abs(ref)))
            err = np.max(np.abs(rho[b].cpu().numpy() - ref)) / scale
            print(f"variant n={n} kd={kd} v={v} b={b} rel err {err:.3e}")
            assert err < 2e-7
            assert np.max(np.abs(obs[b].cpu().numpy() - ref_obs)) < 1e-6 * max(1.0, np.max(np.abs(ref_obs)))


@pytest.mark.parametrize("n,kd", [(20, 0), (65, 0), (70, 2), (130, 1)])
def test_split_bf16_variant_matches_oracle(n, kd):
    """3xBF16 split products: each operand x = hi + lo (BF16 pair, round to nearest even),
    products hi*hi + hi*lo + lo*hi.  Representation error <= 2^-17 |x| per operand and the
    dropped lo*lo term <= 2^-18 |a b|, so one RHS is accurate to ~1e-5 relative; over 10
    steps of dt = 0.02 with ||H|| = 1 the state error stays below 1e-5 * max|rho|.  Keeping
    only hi*hi (a plausible slip) would give ~1e-3."""
    rng = np.random.default_rng(700 + n)
    H = rand_herm(rng, n)
    c_ops = [0.1 * rand_c(rng, n, n) / np.sqrt(n) for _ in range(kd)]
    c_ops += [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    X0 = np.stack([rand_dm(rng, n), rand_c(rng, n, n) / n])
    A = [rand_herm(rng, n)]
    dt, nt = 0.02, 10
    refs = [ol.propagate(H, X0[b], c_ops, dt, nt, A, nout=5) for b in range(2)]
    errs = {}
    # 3xBF16 on both tile edges (32-tiles: T_k passed as split planes; 64-tiles: complex64)
    for key, prec, variant in (("3xbf16", "3xbf16", 1), ("3xbf16/64", "3xbf16", 2), ("fp32", "fp32", 0)):
        prop = _prop()(H, c_ops, group=2, precision=prec, force_tiled=True, variant=variant)
        rho, obs, _ = prop.run(X0, dt, nt, e_ops=A, nout=5)
        errs[key] = []
        for b in range(2):
            ref, ref_obs, _ = refs[b]
            errs[key].append(np.max(np.abs(rho[b].cpu().numpy() - ref)) / np.max(np.abs(ref)))
            if prec == "3xbf16":
                assert np.max(np.abs(obs[b].cpu().numpy() - ref_obs)) < 1e-5 * np.max(np.abs(ref_obs))
    print(f"n={n} kd={kd} rel err 3xbf16 {max(errs['3xbf16']):.3e} / {max(errs['3xbf16/64']):.3e} "
          f"fp32 {max(errs['fp32']):.3e}")
    assert max(errs["3xbf16"]) < 1e-5
    assert max(errs["3xbf16/64"]) < 1e-5
    assert max(errs["fp32"]) < 2e-7


@pytest.mark.parametrize("n,prec", [(20, "fp32"), (130, "fp32"), (130, "fp64"), (100, "3xbf16")])
def test_tiled_trace_renormalisation_records_after_scaling(n, prec):
    """Trace renormalisation on the per-stage kernels: Re Tr rho = 1 after every step, the
    recorded states and observables are those of the renormalised state, and the physics
    agrees with the (trace-conserving) FP64 oracle."""
    rng = np.random.default_rng(40 + n)
    H = rand_herm(rng, n)
    c_ops = [0.2 * rand_c(rng, n, n) / np.sqrt(n)] + [np.sqrt(0.05) * np.diag(np.eye(n)[i]) for i in range(n)]
    rho0 = rand_dm(rng, n)
    E = [np.diag(np.eye(n)[i]).astype(complex) for i in (0, n // 2)]
    dt, nt = 0.05, 40
    prop = _prop()(H, c_ops, precision=prec, force_tiled=True)
    rho, obs, states = prop.run(rho0, dt, nt, e_ops=E, nout=8, store_states=True, renorm=True)
    r = rho[0].cpu().numpy()
    st = states[0].cpu().numpy()
    ob = obs[0].cpu().numpy()
    assert abs(np.trace(r).real - 1) < 1e-13
    for k in range(st.shape[0]):
        assert abs(np.trace(st[k]).real - 1) < 1e-13
        for e in range(len(E)):
            assert abs(ob[k, e] - np.trace(E[e] @ st[k])) < 1e-12
    assert np.allclose(st[-1], r, rtol=0, atol=0)
    ref, _, _ = ol.propagate(H, rho0, c_ops, dt, nt)
    tol = {"fp32": 2e-7, "fp64": 1e-13, "3xbf16": 2e-6}[prec]
    err = np.max(np.abs(r - ref)) / np.max(np.abs(ref))
    print(f"tiled renorm n={n} {prec}: rel err {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("n,n_e", [(33, 1), (65, 1), (65, 2), (48, 1), (65, 65), (65, 130)])
def test_qrt_per_operator_observables(n, n_e):
    """One observable set per operator (eop_group = 1), as in the absorption ensemble: a
    single observable is held in registers for the whole launch, several are read from
    memory at every record; both against the oracle.  At n = 65 the record slots of 65
    observables still fit beside the double-buffered stage operand, those of 130 do not
    (single-buffered kernel)."""
    rng = np.random.default_rng(300 + n + n_e)
    S, G = 2, 3
    Hs = np.stack([rand_herm(rng, n) for _ in range(S)])
    L = [np.sqrt(0.03) * np.diag(np.eye(n)[i]) for i in range(n)]
    X0 = np.stack([rand_c(rng, n, n) / n for _ in range(S * G)])
    E = np.stack([np.stack([rand_c(rng, n, n) for _ in range(n_e)]) for _ in range(S * G)])
    dt, nt = 0.05, 24
    prop = _prop()(Hs, L, group=G)
    rho, obs, _ = prop.run(X0, dt, nt, e_ops=torch.as_tensor(E), nout=3, eop_group=1)
    obs = obs.cpu().numpy()
    for b in range(S * G):
        ref_rho, ref_obs, _ = ol.propagate(Hs[b // G], X0[b], L, dt, nt, list(E[b]), nout=3)
        assert np.max(np.abs(obs[b] - ref_obs)) < 1e-6 * np.max(np.abs(ref_obs))
        assert np.max(np.abs(rho[b].cpu().numpy() - ref_rho)) < 2e-7 * np.max(np.abs(ref_rho))


def test_host_checks_reject_buffers_the_kernel_would_overrun():
    """Every buffer handed to the C ABI is sized from the descriptor alone, so a buffer of
    the wrong shape, dtype or layout is refused on the host before any launch."""
    rng = np.random.default_rng(3)
    n = 10
    H = np.stack([rand_herm(rng, n) for _ in range(2)])
    prop = _prop()(H, [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)], group=3)   # batch 6
    dev = torch.device("cuda")
    rho = torch.zeros((6, n, n), dtype=torch.complex128, device=dev)
    E = torch.zeros((6, 1, n, n), dtype=torch.complex128, device=dev)
    obs = torch.zeros((6, 4, 1), dtype=torch.complex128, device=dev)
    with pytest.raises(ValueError):
        prop.step(rho[:5].contiguous(), 0.1, 1)                                  # batch
    with pytest.raises(ValueError):
        prop.step(rho.to(torch.complex64), 0.1, 1)                               # dtype
    with pytest.raises(ValueError):
        prop.step(rho, 0.1, 1, eops_t=E[:3].contiguous(), eop_group=1, obs=obs, rec_capacity=4)
    with pytest.raises(ValueError):
        prop.step(rho, 0.1, 1, eops_t=E, eop_group=1, obs=obs[:, :2].contiguous(), rec_capacity=4)
    with pytest.raises(ValueError):
        prop.step(rho, 0.1, 1, eops_t=E.transpose(-1, -2), eop_group=1, obs=obs, rec_capacity=4)
    with pytest.raises(ValueError):
        prop.step(rho, 0.1, 1, states=torch.zeros((6, 2, n, n), dtype=torch.complex128, device=dev),
                  rec_capacity=4)
    x = torch.zeros((6, n, n), dtype=torch.complex64, device=dev)
    with pytest.raises(ValueError):
        prop.apply(x[:4].contiguous())
    with pytest.raises(ValueError):
        prop.apply(x, y=torch.zeros((6, n, n), dtype=torch.complex64, device=dev))
    # the well-formed call still runs
    prop.step(rho, 0.1, 4, eops_t=E, eop_group=1, obs=obs, nout=1, rec_capacity=4)
    torch.cuda.synchronize()


@pytest.mark.parametrize("n", [1, 2, 5])
def test_tiny_systems_and_empty_calls(n):
    """n = 1 and 2 (one MFMA tile, mostly padding), zero steps and an empty batch."""
    rng = np.random.default_rng(50 + n)
    H = rand_herm(rng, n) if n > 1 else np.array([[0.3]], dtype=complex)
    c_ops = [np.sqrt(0.1) * np.diag(np.eye(n)[i]) for i in range(n)]
    if n > 1:
        c_ops.append(0.2 * rand_c(rng, n, n))
    X0 = rand_c(rng, n, n)
    prop = _prop()(H, c_ops)
    rho, _, _ = prop.run(X0, 0.05, 40)
    ref, _, _ = ol.propagate(H, X0, c_ops, 0.05, 40)
    assert np.max(np.abs(rho[0].cpu().numpy() - ref)) < 2e-7 * max(1.0, np.max(np.abs(ref)))
    # zero steps: the state is returned unchanged (and never mutated in the caller's copy)
    r0, _, _ = prop.run(X0, 0.05, 0)
    assert np.array_equal(r0[0].cpu().numpy(), X0)
    for tiled in (False, True):
        p2 = _prop()(H, c_ops, force_tiled=tiled)
        r2, _, _ = p2.run(X0, 0.05, 40)
        assert np.max(np.abs(r2[0].cpu().numpy() - ref)) < 2e-7 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [33, 65, 97, 129])
@pytest.mark.parametrize("prec", ["fp32", "fp64", "3xbf16"])
def test_bordered_tiled_layout(n, prec):
    """Per-stage kernels at n = nc + 1 with nc a multiple of the tile: the MFMA tiles cover
    the nc x nc core, border workgroups the last row and column (both tile sizes; with
    observables, state records and stage-operand records)."""
    rng = np.random.default_rng(800 + n)
    H = rand_herm(rng, n)
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    X0 = np.stack([rand_dm(rng, n), rand_c(rng, n, n) / n])
    A = [rand_herm(rng, n)]
    dt, nt = 0.02, 6
    refs = [ol.propagate(H, X0[b], c_ops, dt, nt, A, nout=2, store_states=True) for b in range(2)]
    tol = {"fp32": 2e-7, "fp64": 1e-13, "3xbf16": 2e-6}[prec]
    for v in (1, 2):                              # 32 x 32 and 64 x 64 tiles
        prop = _prop()(H, c_ops, group=2, precision=prec, force_tiled=True, variant=v)
        rho, obs, states = prop.run(X0, dt, nt, e_ops=A, nout=2, store_states=True)
        for b in range(2):
            ref, ref_obs, ref_states = refs[b]
            scale = np.max(np.abs(ref))
            err = np.max(np.abs(rho[b].cpu().numpy() - ref)) / scale
            print(f"bordered tiled n={n} {prec} tile{32 * v} b={b} rel err {err:.3e}")
            assert err < tol
            assert np.max(np.abs(obs[b].cpu().numpy() - ref_obs)) < 10 * tol * np.max(np.abs(ref_obs))
            assert np.max(np.abs(states[b].cpu().numpy() - np.array(ref_states))) < tol * scale
    # stage-operand records: the FP64 stage operands X + h/2 k1, X + h/2 k2, X + h k3
    if prec == "fp64":
        prop = _prop()(H, c_ops, group=2, precision=prec, force_tiled=True)
        Xt = torch.as_tensor(X0, device="cuda").contiguous()
        stg = torch.empty((1, 3, 2, n, n), dtype=torch.complex128, device="cuda")
        prop.step(Xt, dt, 1, stage_states=stg, nout=1, rec_capacity=1)
        for b in range(2):
            f = lambda x: ol.rhs_heff(x, *ol.split_ops(H, c_ops))
            k1 = f(X0[b])
            s2 = X0[b] + 0.5 * dt * k1
            k2 = f(s2)
            s3 = X0[b] + 0.5 * dt * k2
            s4 = X0[b] + dt * f(s3)
            for i, ref in enumerate((s2, s3, s4)):
                assert np.max(np.abs(stg[0, i, b].cpu().numpy() - ref)) < 1e-13


@pytest.mark.parametrize("n,prec,kd", [(129, "fp32", 0), (129, "fp64", 0), (100, "fp32", 2),
                                       (40, "3xbf16", 0)])
def test_tiled_records_are_stored_and_bitwise_reproducible(n, prec, kd):
    """Per-stage kernels store each record (the per-workgroup partials of Tr[E rho] are summed
    in a fixed order): an obs buffer pre-filled with garbage gives the same records as a
    zeroed one, bit for bit, and two runs agree bit for bit; the records match the oracle."""
    rng = np.random.default_rng(900 + n + kd)
    H = rand_herm(rng, n)
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    c_ops += [0.1 * rand_c(rng, n, n) / np.sqrt(n) for _ in range(kd)]
    X0 = np.stack([rand_dm(rng, n), rand_c(rng, n, n) / n, rand_dm(rng, n)])
    E = [rand_herm(rng, n), rand_c(rng, n, n)]
    dt, nt = 0.02, 5
    prop = _prop()(H, c_ops, group=3, precision=prec, force_tiled=True)
    Et = torch.as_tensor(np.stack(E), device="cuda").transpose(-1, -2).contiguous().unsqueeze(0)
    outs = []
    for fill in (0.0, float("nan"), 1e300):
        rho = torch.as_tensor(X0, device="cuda").contiguous()
        obs = torch.full((3, nt, 2), complex(fill, -fill), dtype=torch.complex128, device="cuda")
        prop.step(rho, dt, nt, eops_t=Et, eop_group=3, obs=obs, nout=1, rec_capacity=nt)
        torch.cuda.synchronize()
        outs.append(obs.cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint64), outs[0].view(np.uint64))
    tol = {"fp32": 2e-7, "fp64": 1e-13, "3xbf16": 2e-6}[prec]
    for b in range(3):
        _, ref_obs, _ = ol.propagate(H, X0[b], c_ops, dt, nt, E)
        err = np.max(np.abs(outs[0][b] - ref_obs)) / np.max(np.abs(ref_obs))
        print(f"stored records n={n} {prec} kd={kd} b={b}: rel err {err:.3e}")
        assert err < 10 * tol


@pytest.mark.parametrize("variant", [3, 4])
@pytest.mark.parametrize("prec,tol", [("fp32", 2e-7), ("3xbf16", 2e-6), ("fp64", 1e-13)])
@pytest.mark.parametrize("n,kd", [(40, 1), (100, 2), (65, 0), (129, 0)])
def test_ksplit_stage_matches_oracle(n, kd, prec, tol, variant):
    """Launch variants 3 and 4: the stage products of every 32 x 32 tile split into 2 or 8 K
    ranges of 32-multiples (at n = 40 with 8 ranges the last six are empty), whose partial
    sums lindblad_stage_sum adds in part order before the same epilogue: records, stored
    states and trace renormalisation against the oracle.  n = 65 and 129 run the bordered
    layout (core tiles split, the border workgroups inside lindblad_stage_sum)."""
    rng = np.random.default_rng(950 + n + kd)
    H = rand_herm(rng, n)
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    c_ops += [0.1 * rand_c(rng, n, n) / np.sqrt(n) for _ in range(kd)]
    X0 = np.stack([rand_dm(rng, n), rand_c(rng, n, n) / n])
    A = [rand_herm(rng, n), rand_c(rng, n, n)]
    dt, nt = 0.02, 6
    prop = _prop()(H, c_ops, group=2, precision=prec, force_tiled=True, variant=variant)
    assert prop.variants() == 5
    rho, obs, states = prop.run(X0, dt, nt, e_ops=A, nout=2, store_states=True)
    for b in range(2):
        ref, ref_obs, ref_states = ol.propagate(H, X0[b], c_ops, dt, nt, A, nout=2, store_states=True)
        scale = np.max(np.abs(ref))
        err = np.max(np.abs(rho[b].cpu().numpy() - ref)) / scale
        print(f"k-split v{variant} n={n} kd={kd} {prec} b={b}: rel err {err:.3e}")
        assert err < tol
        assert np.max(np.abs(obs[b].cpu().numpy() - ref_obs)) < 10 * tol * np.max(np.abs(ref_obs))
        assert np.max(np.abs(states[b].cpu().numpy() - np.array(ref_states))) < tol * scale
    r, _, _ = prop.run(X0[0], dt, nt, renorm=True)
    r = r[0].cpu().numpy()
    ref, _, _ = ol.propagate(H, X0[0], c_ops, dt, nt)
    assert abs(np.trace(r).real - 1) < 1e-13
    assert np.max(np.abs(r - ref)) / np.max(np.abs(ref)) < 10 * tol
