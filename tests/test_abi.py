"""CPU-side checks of the C ABI: the library loads, exports every declared symbol, and
rejects bad arguments with the documented status codes (no GPU work is issued)."""
import ctypes
import os
import re

import pytest

from lime_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lime_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lime_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liblime_hip.so is not built (python -m lime_amd.build)")
    return _lib.load()


def test_header_and_python_export_list_agree():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_every_declared_symbol_is_exported(lib):
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_version_and_limits(lib):
    assert b"gfx950" in lib.lime_version()
    assert lib.lime_lindblad_max_n(0, _lib.LIME_F_PREC_FP32) == 80
    assert lib.lime_lindblad_max_n(3, _lib.LIME_F_PREC_FP32) == 64
    d = _lib.LindbladBatch()
    d.batch, d.n, d.group, d.eop_group, d.flags = 2, 20, 1, 1, _lib.LIME_F_PREC_FP64
    assert lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) == 2 * 400 * 16 * 3


def _call(lib, **kw):
    d = _lib.LindbladBatch()
    d.batch, d.n, d.group, d.eop_group = 1, 8, 1, 1
    d.heff = 0x1000
    d.rho = 0x2000
    for k, v in kw.items():
        if hasattr(d, k):
            setattr(d, k, v)
    p = _lib.StepParams()
    p.dt, p.nsteps, p.nout, p.rec_capacity = 0.1, kw.get("nsteps", 1), 1, 0
    return lib.lime_lindblad_rk4(ctypes.byref(d), ctypes.byref(p), None)


def test_argument_validation_without_gpu(lib):
    assert lib.lime_lindblad_rk4(None, None, None) == _lib.LIME_ERR_ARG
    assert _call(lib, batch=0) == _lib.LIME_OK                 # empty batch: no-op
    assert _call(lib, nsteps=0) == _lib.LIME_OK
    assert _call(lib, n=0) == _lib.LIME_ERR_ARG
    assert _call(lib, group=0) == _lib.LIME_ERR_ARG
    assert _call(lib, heff=None) == _lib.LIME_ERR_ARG
    assert _call(lib, n_dense=2) == _lib.LIME_ERR_ARG          # no lops
    assert _call(lib, obs=0x3000) == _lib.LIME_ERR_ARG         # obs without e_ops
    assert _call(lib, flags=_lib.LIME_F_PREC_FP64) == _lib.LIME_ERR_ARG   # tiled: needs workspace
    assert _call(lib, flags=0x30) == _lib.LIME_ERR_UNSUPPORTED
    # stage operands are recorded by the per-stage kernels only
    assert _call(lib, stage_states=0x5000) == _lib.LIME_ERR_UNSUPPORTED
    assert b"FORCE_TILED" in lib.lime_last_error()
    # 3xBF16 runs on the tiled kernels at every n: it needs a workspace too
    assert _call(lib, flags=_lib.LIME_F_PREC_3XBF16) == _lib.LIME_ERR_ARG
    assert b"workspace" in lib.lime_last_error()
    # n > 80 runs the tiled kernels, which need a caller-provided workspace
    assert _call(lib, n=97) == _lib.LIME_ERR_ARG
    assert b"workspace" in lib.lime_last_error()
    # trace renormalisation on the tiled kernels needs 8 bytes per operator more workspace
    plain = 1 * 200 * 200 * 8 * 3
    assert _call(lib, n=200, work=0x4000, work_bytes=plain,
                 flags=_lib.LIME_F_RENORM_TRACE) == _lib.LIME_ERR_ARG
    assert b"workspace" in lib.lime_last_error()
    with pytest.raises(_lib.LimeError):
        _lib.check(_call(lib, n=200))


def _a256(x):
    return (x + 255) // 256 * 256


def test_workspace_query(lib):
    """Regions S0, S1, increment, T_k (operand precision), trace scales, observable
    partials (one slot per stage workgroup of the 32-tile layout), and for 3xBF16 the
    pre-split BF16 planes of H_eff, L_k (and R_k) per operator set (4 planes of n rows of
    n rounded up to 8; T_k in the same planes), each region 256-byte aligned."""
    d = _lib.LindbladBatch()
    d.batch, d.n, d.group, d.eop_group = 4, 64, 1, 1
    assert lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) == 0          # resident
    d.flags = _lib.LIME_F_FORCE_TILED
    assert lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) == 3 * _a256(4 * 64 * 64 * 8)
    d.flags, d.n, d.n_dense = 0, 130, 2
    op = _a256(4 * 130 * 130 * 8)
    assert lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) == 3 * op + _a256(2 * 4 * 130 * 130 * 8)
    d.flags, d.n, d.n_dense = _lib.LIME_F_PREC_3XBF16, 16, 0                 # complex64 workspace
    planes = _a256(4 * 4 * 16 * 16 * 2)                                      # 4 sets of H_eff planes
    assert lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) == 3 * _a256(4 * 16 * 16 * 8) + planes
    d.flags = _lib.LIME_F_PREC_3XBF16 | _lib.LIME_F_RENORM_TRACE             # + trace scales
    assert lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) == 3 * _a256(4 * 16 * 16 * 8) + 256 + planes
    # two operator sets (group 2), n = 20 (row pitch 24), two dense L and separate R_k
    d.flags, d.n, d.n_dense, d.group = _lib.LIME_F_PREC_3XBF16, 20, 2, 2
    d.rops = 256                                                             # never dereferenced
    op = _a256(4 * 20 * 20 * 8)
    assert lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) == \
        3 * op + _a256(4 * 2 * 4 * 20 * 24 * 2) + _a256(2 * (1 + 2 * 2) * 4 * 20 * 24 * 2)
    d.rops, d.group = None, 1
    # observables: n = 129 (bordered 32-tiles: 16 tiles + 9 border workgroups per operator)
    d.flags, d.n, d.n_dense, d.n_eops = _lib.LIME_F_PREC_FP64, 129, 0, 3
    d.obs = 256                                                              # never dereferenced
    assert lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) == \
        3 * _a256(4 * 129 * 129 * 16) + _a256(4 * 25 * 3 * 16)


def test_workspace_of_the_ksplit_variants(lib):
    """Launch variants 3 and 4 (32-tiles, the stage reduction split in 2 / 8 parts) add the
    parts' partial sums (operand precision) to the workspace; every other variant adds
    nothing; five variants on the per-stage path."""
    d = _lib.LindbladBatch()
    d.batch, d.n, d.group, d.eop_group, d.n_dense = 2, 100, 1, 1, 2
    d.flags = _lib.LIME_F_FORCE_TILED
    op = _a256(2 * 100 * 100 * 8)
    base = 3 * op + _a256(2 * 2 * 100 * 100 * 8)
    assert lib.lime_lindblad_variants(ctypes.byref(d)) == 5
    for variant, extra in ((0, 0), (1, 0), (2, 0), (3, _a256(2 * 2 * 100 * 100 * 8)),
                           (4, _a256(8 * 2 * 100 * 100 * 8))):
        d.variant = variant
        assert lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) == base + extra, variant
    d.flags, d.variant = _lib.LIME_F_FORCE_TILED | _lib.LIME_F_PREC_FP64, 4   # complex128 partials
    assert lib.lime_lindblad_workspace_bytes(ctypes.byref(d)) == \
        3 * _a256(2 * 100 * 100 * 16) + _a256(2 * 2 * 100 * 100 * 16) + _a256(8 * 2 * 100 * 100 * 16)
