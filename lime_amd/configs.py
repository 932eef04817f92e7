"""Seeded synthetic inputs for the benchmark and parity workloads.

Every generator uses ``numpy.random.default_rng(seed)`` and returns plain complex128
NumPy arrays, so the FP64 oracle and the GPU path see bit-identical inputs.

Workloads
---------
chain_config     -- 8-site Frenkel chain, on-site dephasing, rho0 = |0><0|
                    (dt 0.05 x 2000, populations every 10 steps)
lattice_spectrum -- 8x8 nearest-neighbour exciton lattice ensemble for linear absorption:
                    Hilbert space = ground state (index 0) + 64 sites, so n = 65;
                    dipole operators mu_a = sum_i mu_{i,a} (|i><g| + |g><i|), a = x, y, z
                    with mu_i unit vectors uniform on the sphere; QRT operators
                    X_a(0) = mu_a |g><g|, C(t) = (1/3) sum_a Tr[mu_a X_a(t)]
ragged_gue       -- instances with N in {16, 32, ..., 256}: GUE H scaled to spectral radius 1,
                    N on-site dephasing operators + 4 dense L_k with ||L_k||_F = 0.1,
                    rho0 = random pure state
chain_gradient   -- 128-site chains (+ ground state) for the spectrum-loss gradient workload
dense_single     -- N = 512 GUE with 16 dense L_k (||L_k||_F = 0.05),
                    rho0 = (1 - eta) I/N + eta psi psi^dag
"""
import numpy as np


def onsite_dephasing(n, gamma, sites):
    """sqrt(gamma) |i><i| for i in sites, as dense (n, n) complex128 arrays."""
    ops = []
    for i in sites:
        L = np.zeros((n, n), dtype=np.complex128)
        L[i, i] = np.sqrt(gamma)
        ops.append(L)
    return ops


def chain_hamiltonian(eps, J):
    n = len(eps)
    H = np.diag(np.asarray(eps, dtype=np.float64)).astype(np.complex128)
    J = np.broadcast_to(np.asarray(J, dtype=np.float64), (n - 1,))
    for i in range(n - 1):
        H[i, i + 1] = H[i + 1, i] = J[i]
    return H


def lattice_hamiltonian(eps, nx, ny, J):
    """Open-boundary nearest-neighbour nx x ny lattice, site index i = x * ny + y."""
    n = nx * ny
    H = np.diag(np.asarray(eps, dtype=np.float64)).astype(np.complex128)
    for x in range(nx):
        for y in range(ny):
            i = x * ny + y
            if x + 1 < nx:
                H[i, i + ny] = H[i + ny, i] = J
            if y + 1 < ny:
                H[i, i + 1] = H[i + 1, i] = J
    return H


def chain_config(seed=1, n=8, J=-0.1, gamma=0.02, sigma=0.05):
    rng = np.random.default_rng(seed)
    H = chain_hamiltonian(rng.normal(0.0, sigma, n), J)
    c_ops = onsite_dephasing(n, gamma, range(n))
    rho0 = np.zeros((n, n), dtype=np.complex128)
    rho0[0, 0] = 1.0
    e_ops = [np.diag(np.eye(n)[i]).astype(np.complex128) for i in range(n)]
    return dict(H=H, c_ops=c_ops, rho0=rho0, e_ops=e_ops, dt=0.05, nt=2000, nout=10, seed=seed)


def unit_vectors(rng, m):
    v = rng.standard_normal((m, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def lattice_spectrum(seed=2, n_real=4096, nx=8, ny=8, J=-0.08, gamma=0.01, sigma=0.05):
    """Ensemble for linear absorption (ground state at index 0, sites 1..nx*ny).

    Returns H (n_real, n, n), c_ops (n_sites shared on-site dephasing operators),
    mu (n_real, 3, n, n) dipole operators, x0 (n_real, 3, n, n) QRT initial operators.
    """
    rng = np.random.default_rng(seed)
    ns = nx * ny
    n = ns + 1
    eps = rng.normal(0.0, sigma, (n_real, ns))
    mu_vec = np.stack([unit_vectors(rng, ns) for _ in range(n_real)])      # (R, ns, 3)
    H = np.zeros((n_real, n, n), dtype=np.complex128)
    base = lattice_hamiltonian(np.zeros(ns), nx, ny, J)
    H[:, 1:, 1:] = base
    idx = np.arange(1, n)
    H[:, idx, idx] = eps
    c_ops = onsite_dephasing(n, gamma, range(1, n))
    mu = np.zeros((n_real, 3, n, n), dtype=np.complex128)
    mu[:, :, 1:, 0] = np.transpose(mu_vec, (0, 2, 1))
    mu[:, :, 0, 1:] = np.transpose(mu_vec, (0, 2, 1))
    x0 = np.zeros_like(mu)
    x0[:, :, :, 0] = mu[:, :, :, 0]                    # mu_a |g><g|
    return dict(H=H, c_ops=c_ops, mu=mu, x0=x0, dt=0.05, nt=4096, n=n, seed=seed,
                gamma=gamma, J=J)


def gue(rng, n, radius=1.0):
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    h = (x + x.conj().T) / 2
    return h * (radius / np.max(np.abs(np.linalg.eigvalsh(h))))


def dense_jump(rng, n, fro):
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return x * (fro / np.linalg.norm(x))


def random_pure_dm(rng, n):
    psi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    psi /= np.linalg.norm(psi)
    return np.outer(psi, psi.conj())


def ragged_gue(seed=3, count=2048, sizes=tuple(range(16, 257, 16)), gamma=0.01, n_dense=4,
               fro=0.1):
    """The n on-site dephasing operators are given compactly (``dephasing`` = gamma); the
    dense jump operators as matrices in ``c_ops``."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.asarray(sizes), size=count)
    inst = []
    for n in ns:
        n = int(n)
        H = gue(rng, n)
        c_ops = [dense_jump(rng, n, fro) for _ in range(n_dense)]
        inst.append(dict(H=H, c_ops=c_ops, dephasing=gamma, rho0=random_pure_dm(rng, n)))
    return dict(instances=inst, sizes=ns, dt=0.02, nt=1000, seed=seed)


def chain_gradient(seed=4, count=256, n_sites=128, gamma=0.02, sigma=0.05):
    rng = np.random.default_rng(seed)
    eps = rng.normal(0.0, sigma, (count, n_sites))
    J = rng.uniform(-0.12, -0.08, (count, n_sites - 1))
    mu = unit_vectors(rng, n_sites)
    return dict(eps=eps, J=J, gamma=gamma, mu=mu, dt=0.05, nt=2000, checkpoint=64, seed=seed)


def dense_single(seed=5, n=512, n_dense=16, fro=0.05, eta=0.5, batch=1):
    rng = np.random.default_rng(seed)
    H = np.stack([gue(rng, n) for _ in range(batch)])
    Ls = [np.stack([dense_jump(rng, n, fro) for _ in range(batch)]) for _ in range(n_dense)]
    psi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    psi /= np.linalg.norm(psi)
    rho0 = (1 - eta) * np.eye(n) / n + eta * np.outer(psi, psi.conj())
    return dict(H=H, c_ops=Ls, rho0=rho0.astype(np.complex128), dt=0.01, nt=10000, eta=eta,
                seed=seed)
