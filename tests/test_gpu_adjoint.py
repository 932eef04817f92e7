"""Adjoint gradients of the chain absorption spectrum loss (GPU adjoint kernels) against
(a) an FP64 unrolled reference gradient (torch autograd through a complex128 RK4 on the
CPU), (b) central finite differences (h = 1e-4) of that FP64 forward, and (c) the FP64
forward itself against the oracle.  Gate: <= 1e-5 relative error, for the FP32 and the FP64
kernels alike:

* site energies, couplings and per-site dephasing rates: max error over the parameters
  relative to the largest gradient of that kind;
* the shared dephasing rate: its gradient is the sum of the per-site ones, which cancel
  (the sum can be 100x smaller than its terms), so its error is taken relative to the sum
  of their magnitudes -- the scale the FP32 rounding of the terms sets;
* central differences: per parameter, relative to the parameter's own gradient (or to the
  largest gradient of its kind when its own is that small)."""
import numpy as np
import pytest

from oracle import lindblad as ol

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def ref_forward(eps, J, gamma, mu_sites, dt, nt):
    """complex128 RK4 of the chain + ground model on the CPU (differentiable)."""
    B, ns = eps.shape
    n = ns + 1
    H = torch.zeros((B, n, n), dtype=torch.complex128)
    idx = torch.arange(1, n)
    H = H.clone()
    H[:, idx, idx] = eps.to(torch.complex128)
    H[:, idx[:-1], idx[1:]] = J.to(torch.complex128)
    H[:, idx[1:], idx[:-1]] = J.to(torch.complex128)
    g = torch.zeros(n, dtype=torch.float64)
    g = torch.cat([g[:1], gamma.reshape(-1).expand(ns) if gamma.dim() < 2 else gamma.reshape(ns)])
    Heff = H - 0.5j * torch.diag_embed(g.to(torch.complex128)).unsqueeze(0)
    D = torch.diag_embed(g.to(torch.complex128)).unsqueeze(0)
    mu = torch.zeros((n, n), dtype=torch.complex128)
    mu[1:, 0] = mu_sites
    mu[0, 1:] = mu_sites
    X = torch.zeros((B, n, n), dtype=torch.complex128)
    X[:, :, 0] = mu[:, 0]
    HeffH = Heff.conj().transpose(-1, -2)

    def f(x):
        return -1j * (Heff @ x - x @ HeffH) + D * x

    out = []
    for _ in range(nt):
        k1 = f(X)
        k2 = f(X + 0.5 * dt * k1)
        k3 = f(X + 0.5 * dt * k2)
        k4 = f(X + dt * k3)
        X = X + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append((mu.transpose(0, 1) * X).sum(dim=(-1, -2)))
    return torch.stack(out, dim=1)


def loss_fn(C, target, dt):
    from lime_amd.adjoint import spectrum_loss
    return spectrum_loss(C, target, dt).sum()


def problem(ns, B, seed):
    rng = np.random.default_rng(seed)
    eps = torch.tensor(rng.normal(0, 0.05, (B, ns)))
    J = torch.tensor(rng.uniform(-0.12, -0.08, (B, ns - 1)))
    gamma = torch.tensor(0.02, dtype=torch.float64)
    mu = torch.tensor(rng.standard_normal(ns))
    return eps, J, gamma, mu


@pytest.mark.parametrize("ns,nt,K,prec", [(5, 120, 16, "fp32"), (12, 64, 64, "fp32"), (100, 24, 8, "fp32"),
                                          (12, 64, 16, "fp64"), (100, 24, 8, "fp64"),
                                          (128, 12, 4, "fp32"),       # n = 129: bordered tiles
                                          (12, 64, 16, "fp32/fp64")])
def test_adjoint_gradient_matches_fp64_reference_and_fd(ns, nt, K, prec):
    from lime_amd.adjoint import ChainAbsorption
    B, dt = 2, 0.05
    eps, J, gamma, mu = problem(ns, B, ns)
    eps_t, J_t, gamma_t, _ = problem(ns, B, ns + 1000)
    with torch.no_grad():
        target = ref_forward(eps_t, J_t, gamma_t, mu, dt, nt)
    # (c) the FP64 reference forward agrees with the oracle
    C_ref = ref_forward(eps, J, gamma, mu, dt, nt).detach()
    n = ns + 1
    Hn = np.zeros((n, n), complex)
    Hn[1:, 1:] = np.diag(eps[0].numpy()) + np.diag(J[0].numpy(), 1) + np.diag(J[0].numpy(), -1)
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(1, n)]
    mun = np.zeros((n, n), complex)
    mun[1:, 0] = mun[0, 1:] = mu.numpy()
    x0 = np.zeros((n, n), complex)
    x0[:, 0] = mun[:, 0]
    _, o, _ = ol.propagate(Hn, x0, c_ops, dt, nt, [mun])
    assert np.max(np.abs(o[:, 0] - C_ref[0].numpy())) < 1e-12
    # (a) FP64 unrolled reference gradients: shared and per-site dephasing rates
    e1, j1, g1 = eps.clone().requires_grad_(), J.clone().requires_grad_(), gamma.clone().requires_grad_()
    loss_fn(ref_forward(e1, j1, g1, mu, dt, nt), target, dt).backward()
    gs1 = gamma.reshape(1, 1).expand(1, ns).clone().requires_grad_()
    loss_fn(ref_forward(eps, J, gs1, mu, dt, nt), target, dt).backward()
    site_scale = float(gs1.grad.abs().sum())
    assert abs(float(gs1.grad.sum()) - float(g1.grad)) < 1e-12 * site_scale
    # adjoint on the GPU: shared rate, then per-site rates
    dev = torch.device("cuda")
    e2 = eps.clone().to(dev).requires_grad_()
    j2 = J.clone().to(dev).requires_grad_()
    g2 = gamma.clone().to(dev).requires_grad_()
    C = ChainAbsorption.apply(e2, j2, g2, mu.to(dev), dt, nt, K, prec)
    ctol = 1e-6 if prec.startswith("fp32") else 1e-12
    assert np.max(np.abs(C.detach().cpu().numpy() - C_ref.numpy())) < ctol * float(C_ref.abs().max())
    loss_fn(C, target.to(dev), dt).backward()
    gs2 = gamma.reshape(1, 1).expand(1, ns).clone().to(dev).requires_grad_()
    loss_fn(ChainAbsorption.apply(eps.to(dev), J.to(dev), gs2, mu.to(dev), dt, nt, K, prec),
            target.to(dev), dt).backward()
    for name, ga, gr in (("eps", e2.grad, e1.grad), ("J", j2.grad, j1.grad), ("gamma_site", gs2.grad, gs1.grad)):
        ga, gr = ga.cpu().numpy(), gr.numpy()
        rel = np.max(np.abs(ga - gr)) / np.max(np.abs(gr))
        print(f"ns={ns} {prec} {name}: adjoint vs FP64 autograd rel err {rel:.2e}")
        assert rel < 1e-5
    relg = abs(float(g2.grad) - float(g1.grad)) / site_scale
    print(f"ns={ns} {prec} gamma (shared): err / sum|per-site| {relg:.2e} "
          f"(own relative {abs(float(g2.grad) - float(g1.grad)) / abs(float(g1.grad)):.2e})")
    assert relg < 1e-5
    # (b) central finite differences of the FP64 forward on a few parameters: h = 1e-4, and
    # 1e-5 for the rate (at 1e-4 = 0.5 % of gamma the quotient's O(h^2) error reaches ~1e-5)
    checks = [("eps", (0, 1)), ("eps", (1, ns - 1)), ("J", (0, 0)), ("gamma", ())]
    for name, ix in checks:
        h = 1e-5 if name == "gamma" else 1e-4
        base = {"eps": eps, "J": J, "gamma": gamma}
        vals = []
        for sgn in (1, -1):
            p = {k: v.clone() for k, v in base.items()}
            p[name][ix] += sgn * h
            with torch.no_grad():
                vals.append(float(loss_fn(ref_forward(p["eps"], p["J"], p["gamma"], mu, dt, nt), target, dt)))
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = float({"eps": e2.grad, "J": j2.grad, "gamma": g2.grad}[name][ix])
        rel = abs(ad - fd) / max(abs(fd), 1e-30)
        print(f"ns={ns} {prec} FD {name}{ix}: adjoint {ad:.8e} fd {fd:.8e} rel {rel:.2e}")
        scale = site_scale if name == "gamma" else float(np.max(np.abs({"eps": e1.grad, "J": j1.grad}[name].numpy())))
        assert rel < 1e-5 or abs(ad - fd) < 1e-5 * scale


@pytest.mark.parametrize("n,pairs", [(2, 1), (5, 4), (65, 2), (129, 4), (300, 3), (513, 1)])
def test_contraction_kernel_against_its_definition(n, pairs):
    """lime_adjoint_contract_pairs on random complex128 operators against the contraction
    it is defined as (P = sum_p conj(kappa_p) o s_p; row and column sums, the diagonal, the
    row- and column-pair terms), accumulated onto non-zero outputs, to 1e-12 relative.
    n = 300 and 513 share column-partial slots between waves (several rounds)."""
    import ctypes
    from lime_amd import _lib as L
    lib = L.load()
    rng = np.random.default_rng(n + pairs)
    B, w = 3, 0.37
    K = [rng.standard_normal((B, n, n)) + 1j * rng.standard_normal((B, n, n)) for _ in range(pairs)]
    S = [rng.standard_normal((B, n, n)) + 1j * rng.standard_normal((B, n, n)) for _ in range(pairs)]
    a0 = rng.standard_normal((B, n)) + 1j * rng.standard_normal((B, n))
    b0 = rng.standard_normal((B, n)) + 1j * rng.standard_normal((B, n))
    u0 = rng.standard_normal((B, n - 1)) + 1j * rng.standard_normal((B, n - 1))
    dev = lambda x: torch.as_tensor(x, device="cuda").contiguous()
    Kt, St = [dev(k) for k in K], [dev(x) for x in S]
    oa, ob, ou = dev(a0), dev(b0), dev(u0)
    kap = (ctypes.c_void_p * pairs)(*[t.data_ptr() for t in Kt])
    sts = (ctypes.c_void_p * pairs)(*[t.data_ptr() for t in St])
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(lib.lime_adjoint_contract_pairs(B, n, pairs, kap, sts, w, oa.data_ptr(), ob.data_ptr(),
                                            ou.data_ptr(), sp))
    torch.cuda.synchronize()
    P = sum(np.conj(k) * x for k, x in zip(K, S))
    row, col = P.sum(axis=2), P.sum(axis=1)
    diag = np.einsum("bjj->bj", P)
    rp = sum((np.conj(k[:, :-1, :]) * x[:, 1:, :] + np.conj(k[:, 1:, :]) * x[:, :-1, :]).sum(axis=2)
             for k, x in zip(K, S))
    cp = sum((np.conj(k[:, :, 1:]) * x[:, :, :-1] + np.conj(k[:, :, :-1]) * x[:, :, 1:]).sum(axis=1)
             for k, x in zip(K, S))
    ra = a0 + w * (row - col)
    rb = b0 + w * (diag - 0.5 * (row + col))
    ru = u0 + w * (rp - cp)
    for got, ref in ((oa, ra), (ob, rb), (ou, ru)):
        err = np.max(np.abs(got.cpu().numpy() - ref)) / np.max(np.abs(ref))
        assert err < 1e-12, err
