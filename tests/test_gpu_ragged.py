"""Uneven batches on the GPU.

``RaggedLindblad`` advances every instance in ONE persistent launch: every size class of
the uneven-batch workload (n = 16, 32, ..., 256, four dense jump operators plus on-site
dephasing) against the FP64 oracle under every launch shape; graph replays against eager
launches bit for bit (several replays, eager work and synchronisations in between); the
result independent of how many workgroups drain the queue.  ``RaggedClasses`` (size
classes captured into one graph) is checked the same way."""
import numpy as np
import pytest

from oracle import lindblad as ol

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def ragged(seed, sizes, count=None):
    from lime_amd import configs
    return configs.ragged_gue(seed=seed, count=count or len(sizes), sizes=tuple(sizes), fro=0.1)


def oracle_state(inst, dt, nt):
    n = inst["H"].shape[0]
    c_ops = list(inst["c_ops"]) + [np.sqrt(inst["dephasing"]) * np.diag(np.eye(n)[i]) for i in range(n)]
    rho, _, _ = ol.propagate(inst["H"], inst["rho0"], c_ops, dt, nt)
    return rho


def _one_per_class(seed):
    """One instance of every size class of the workload (n = 16 .. 256 step 16)."""
    return [ragged(seed + n, (n,))["instances"][0] for n in range(16, 257, 16)]


@pytest.fixture(scope="module")
def class_refs():
    inst = _one_per_class(41)
    dt, nt = 0.02, 4
    return inst, dt, nt, [oracle_state(x, dt, nt) for x in inst]


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_persistent_every_size_class_matches_oracle(class_refs, variant):
    from lime_amd.ragged import RaggedLindblad
    inst, dt, nt, refs = class_refs
    rb = RaggedLindblad(inst, tune=False)
    assert rb.variants() == 5
    rb.variant = variant
    rb.step(dt, nt)
    got = rb.states()
    for x, g, ref in zip(inst, got, refs):
        n = x["H"].shape[0]
        err = np.max(np.abs(g - ref))
        print(f"persistent v{variant} n={n}: max error {err:.3e}")
        assert err < 2e-7, n
        assert abs(np.trace(g) - 1) < 1e-6


def test_persistent_graph_replays_equal_eager_and_any_grid():
    from lime_amd.ragged import RaggedLindblad
    cfg = ragged(42, (8, 24, 40, 64, 100, 130), count=14)
    dt, chunk, reps = 0.02, 3, 3
    eager = RaggedLindblad(cfg["instances"], tune=False)
    graph = RaggedLindblad(cfg["instances"], tune=False)
    one = RaggedLindblad(cfg["instances"], tune=False, max_groups=1)   # one workgroup drains all
    graph.capture(dt, chunk)
    for r in range(reps):
        eager.step(dt, chunk)
        torch.cuda.synchronize()
        graph.graph.replay()
        torch.cuda.synchronize()
        one.step(dt, chunk)
        junk = torch.randn(1000, device="cuda").sum()          # unrelated eager work
        torch.cuda.synchronize()
        assert torch.equal(eager.rho, graph.rho), f"replay {r}"
        assert torch.equal(eager.rho, one.rho), f"launch {r}"
        assert int(graph.queue.abs().sum()) == 0                 # the kernel left the queue empty
    got = graph.states()
    for i, x in enumerate(cfg["instances"]):
        err = np.max(np.abs(got[i] - oracle_state(x, dt, chunk * reps)))
        print(f"persistent graph instance {i} n={x['H'].shape[0]}: err {err:.3e}")
        assert err < 2e-7


def test_persistent_run_chunks_and_remainder():
    from lime_amd.ragged import RaggedLindblad
    cfg = ragged(43, (16, 48, 96, 144))
    dt = 0.02
    a = RaggedLindblad(cfg["instances"], tune=False)
    a.run(dt, 11, chunk=4)                 # 2 replays + 3 eager steps
    b = RaggedLindblad(cfg["instances"], tune=False)
    b.step(dt, 11)
    assert torch.equal(a.rho, b.rho)
    with pytest.raises(ValueError):
        bad = dict(cfg["instances"][0])
        bad["rho0"] = np.eye(3)
        RaggedLindblad([bad], tune=False)


def test_class_graph_matches_oracle_and_eager():
    from lime_amd.ragged import RaggedClasses
    cfg = ragged(31, (8, 24, 40, 64, 100, 130), count=14)
    dt, chunk, reps = 0.02, 5, 3
    eager = RaggedClasses(cfg["instances"], tune=True, verbose=True)
    graph = RaggedClasses(cfg["instances"], tune=True)
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
        print(f"ragged classes instance {i} n={inst['H'].shape[0]} err {err:.3e}")
        assert err < 2e-7
    assert graph.flops_per_step > 0
