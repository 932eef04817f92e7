"""Ragged batches in one captured graph: every instance against the oracle, graph replays
against eager launches (several replays, eager work and synchronisations in between)."""
import numpy as np
import pytest

from oracle import lindblad as ol

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def small_ragged(seed=31):
    from lime_amd import configs
    return configs.ragged_gue(seed=seed, count=14, sizes=(8, 24, 40, 64, 100, 130), fro=0.1)


def oracle_state(inst, dt, nt):
    n = inst["H"].shape[0]
    c_ops = list(inst["c_ops"]) + [np.sqrt(inst["dephasing"]) * np.diag(np.eye(n)[i]) for i in range(n)]
    rho, _, _ = ol.propagate(inst["H"], inst["rho0"], c_ops, dt, nt)
    return rho


def test_ragged_graph_matches_oracle_and_eager():
    from lime_amd.ragged import RaggedLindblad
    cfg = small_ragged()
    dt, chunk, reps = 0.02, 5, 3
    eager = RaggedLindblad(cfg["instances"], tune=True, verbose=True)
    graph = RaggedLindblad(cfg["instances"], tune=True)
    graph.capture(dt, chunk)
    for r in range(reps):
        eager.step_eager(dt, chunk)
        torch.cuda.synchronize()
        graph.graph.replay()
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum()          # unrelated eager work
        for ce, cg in zip(eager.classes, graph.classes):
            assert torch.equal(ce["rho"], cg["rho"]), f"replay {r}, n={ce['n']}"
    got = graph.states()
    for i, inst in enumerate(cfg["instances"]):
        ref = oracle_state(inst, dt, chunk * reps)
        err = np.max(np.abs(got[i] - ref))
        print(f"ragged instance {i} n={inst['H'].shape[0]} err {err:.3e}")
        assert err < 2e-7
    assert graph.flops_per_step > 0


def test_ragged_run_chunks_and_remainder():
    from lime_amd.ragged import RaggedLindblad
    cfg = small_ragged(seed=32)
    dt = 0.02
    a = RaggedLindblad(cfg["instances"], tune=False)
    a.run(dt, 11, chunk=4)                 # 2 replays + 3 eager steps
    b = RaggedLindblad(cfg["instances"], tune=False)
    b.step_eager(dt, 11)
    for ca, cb in zip(a.classes, b.classes):
        assert torch.allclose(ca["rho"], cb["rho"], atol=1e-13, rtol=0)
