"""The plain FP64 GPU propagator (batched rocBLAS zgemm + elementwise kernels) -- the
denominator of the speed-up figures and the FP64 GPU reference of the large-n workload --
against the oracle, and lime's 2 + 5K operation sequence against the folded RHS."""
import numpy as np
import pytest

from oracle import lindblad as ol

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rand_herm(rng, n):
    x = rand_c(rng, n, n)
    h = (x + x.conj().T) / 2
    return h / np.max(np.abs(np.linalg.eigvalsh(h)))


@pytest.mark.parametrize("n,kd,group", [(8, 0, 1), (33, 2, 3), (65, 0, 2)])
def test_plain_fp64_matches_oracle(n, kd, group):
    from lime_amd.plain import PlainFP64Propagator
    rng = np.random.default_rng(60 + n)
    S = 2
    Hs = np.stack([rand_herm(rng, n) for _ in range(S)])
    c_ops = [0.1 * rand_c(rng, n, n) / np.sqrt(n) for _ in range(kd)]
    c_ops += [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    X0 = np.stack([rand_c(rng, n, n) / n for _ in range(S * group)])
    E = rand_c(rng, n, n)
    dt, nt = 0.05, 12
    p = PlainFP64Propagator(Hs, c_ops, group=group)
    Et = torch.as_tensor(E.T.copy(), device="cuda").reshape(1, 1, n, n)
    X, obs = p.run(X0, dt, nt, e_ops_t=Et, nout=4)
    for b in range(S * group):
        ref, ref_obs, _ = ol.propagate(Hs[b // group], X0[b], c_ops, dt, nt, [E], nout=4)
        assert np.max(np.abs(X[b].cpu().numpy() - ref)) < 1e-13 * max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(obs[b].cpu().numpy() - ref_obs)) < 1e-12 * max(1.0, np.max(np.abs(ref_obs)))


def test_reference_sequence_equals_folded_rhs():
    from lime_amd.plain import PlainFP64Propagator, rhs_reference_sequence
    rng = np.random.default_rng(5)
    n = 12
    H = rand_herm(rng, n)
    c_ops = [0.2 * rand_c(rng, n, n), np.diag(rng.uniform(0, 0.3, n)).astype(complex)]
    X = rand_c(rng, 3, n, n)
    dev = torch.device("cuda")
    Xt = torch.as_tensor(X, device=dev)
    a = rhs_reference_sequence(Xt, torch.as_tensor(H, device=dev),
                               [torch.as_tensor(c, device=dev) for c in c_ops])
    p = PlainFP64Propagator(np.stack([H] * 3), c_ops)
    b = p.rhs(Xt)
    assert torch.allclose(a, b, rtol=0, atol=1e-13)
    ref = np.stack([ol.liouvillian_ref(X[i], H, c_ops) for i in range(3)])
    assert np.max(np.abs(a.cpu().numpy() - ref)) < 1e-13
