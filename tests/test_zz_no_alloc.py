"""lime_lindblad_rk4_noalloc (static device scratch, no allocation) against
lime_lindblad_rk4 (caller workspace), both called through the C ABI with ctypes.

The same kernels run on the same inputs, so states and records must agree bit for bit.
Two calls in a row, the second with a smaller batch, check that what the first call left
in the shared scratch array does not leak into the second."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rand_herm(rng, n):
    x = rand_c(rng, n, n)
    h = (x + x.conj().T) / 2
    return h / np.max(np.abs(np.linalg.eigvalsh(h)))


def _case(rng, n, batch, kd, precision):
    from lime_amd.propagator import LindbladPropagator
    H = rand_herm(rng, n)
    c_ops = [np.sqrt(0.02) * np.diag(np.eye(n)[i]) for i in range(n)]
    c_ops += [0.1 * rand_c(rng, n, n) / np.sqrt(n) for _ in range(kd)]
    prop = LindbladPropagator(H, c_ops, group=batch, precision=precision, force_tiled=True)
    X0 = torch.as_tensor(np.stack([rand_c(rng, n, n) / n for _ in range(batch)]), device="cuda")
    E = torch.as_tensor(np.stack([rand_herm(rng, n), rand_c(rng, n, n)]), device="cuda")
    return prop, X0.contiguous(), E.transpose(-1, -2).contiguous().unsqueeze(0)


def _run(prop, X0, Et, dt, nt, noalloc):
    from lime_amd import _lib
    lib = _lib.load()
    B, n = prop.batch, prop.n
    rho = X0.clone()
    obs = torch.full((B, nt, Et.shape[1]), complex(7.0, -7.0), dtype=torch.complex128, device="cuda")
    states = torch.full((B, nt, n, n), complex(3.0, 3.0), dtype=torch.complex128, device="cuda")
    d = _lib.LindbladBatch()
    d.batch, d.n, d.group, d.n_dense = B, n, prop.group, prop.n_dense
    d.n_eops, d.eop_group = Et.shape[1], B
    d.flags = {"fp32": 0, "fp64": _lib.LIME_F_PREC_FP64}[prop.precision] | _lib.LIME_F_FORCE_TILED
    d.heff = prop.heff.data_ptr()
    d.dmat = prop.dmat.data_ptr() if prop.dmat is not None else None
    d.lops = prop.lops.data_ptr() if prop.lops is not None else None
    d.eops_t, d.rho, d.obs, d.states = Et.data_ptr(), rho.data_ptr(), obs.data_ptr(), states.data_ptr()
    work = None
    if not noalloc:
        need = int(lib.lime_lindblad_workspace_bytes(ctypes.byref(d)))
        assert 0 < need <= _lib.LIME_NOALLOC_SCRATCH_BYTES
        work = torch.empty(need, dtype=torch.uint8, device="cuda")
        d.work, d.work_bytes = work.data_ptr(), need
    p = _lib.StepParams()
    p.dt, p.nsteps, p.nout, p.step_offset, p.rec_capacity = dt, nt, 1, 0, nt
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    fn = lib.lime_lindblad_rk4_noalloc if noalloc else lib.lime_lindblad_rk4
    _lib.check(fn(ctypes.byref(d), ctypes.byref(p), stream))
    torch.cuda.synchronize()
    return rho.cpu().numpy(), obs.cpu().numpy(), states.cpu().numpy()


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_matches_workspace_entry_point_two_calls(precision):
    rng = np.random.default_rng(77)
    big = _case(rng, 129, 6, 0, precision)           # bordered tiles, observables + states
    small = _case(rng, 100, 2, 2, precision)          # dense jump operators (T_k in scratch)
    for k, (prop, X0, Et) in enumerate((big, small)):
        print(f"[noalloc] call {k}: n={prop.n} batch={prop.batch} kd={prop.n_dense} {precision}", flush=True)
        ref = _run(prop, X0, Et, 0.02, 4, noalloc=False)
        got = _run(prop, X0, Et, 0.02, 4, noalloc=True)
        for r, g in zip(ref, got):
            assert np.array_equal(_bits(r), _bits(g))
        assert np.all(np.isfinite(got[1]))


def test_noalloc_rejects_batches_above_the_scratch_size():
    from lime_amd import _lib
    lib = _lib.load()
    d = _lib.LindbladBatch()
    d.batch, d.n, d.group, d.eop_group = 64, 512, 64, 1
    d.heff, d.rho = 0x1000, 0x2000                     # never dereferenced: rejected first
    p = _lib.StepParams()
    p.dt, p.nsteps, p.nout = 0.01, 1, 1
    assert lib.lime_lindblad_rk4_noalloc(ctypes.byref(d), ctypes.byref(p), None) == _lib.LIME_ERR_ARG
    assert b"LIME_NOALLOC_SCRATCH_BYTES" in lib.lime_last_error()
