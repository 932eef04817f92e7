"""Array primitives of lime's hot path (lime/phys.py), for NumPy arrays or torch tensors.

They operate on whatever array type they are given (a GPU torch tensor stays on the GPU).
The propagation itself never goes through ``rk4`` here: the solvers call the fused kernel.

* ``rk4``        -- phys.py:636-649 (classic RK4, in-place update of ``rho``)
* ``comm``       -- phys.py:741-743;  ``anticomm`` -- :746-748;  ``dag`` -- :758
* ``obs_dm``     -- phys.py:837-844: Tr[d rho] (no conjugate on d), here in O(N^2)
* ``transform``  -- phys.py:706-718: v^dag A v
* ``isherm``, ``ket2dm``, ``basis``, ``destroy``, ``pauli`` -- phys.py:1429, 579, 879, 615, 773
* ``expm``       -- phys.py:1352-1401 ('EOM': U(t) from dU/dt = A U by RK4, on the GPU
  kernels); ``ldo`` -- phys.py:1409-1425 (A b)
"""
import numpy as np


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def dag(a):
    if _is_torch(a):
        return a.conj().transpose(-1, -2)
    if hasattr(a, "getH"):
        return a.getH()
    return np.conjugate(a).T


def comm(A, B):
    assert A.shape == B.shape
    return A @ B - B @ A


def anticomm(A, B):
    assert A.shape == B.shape
    return A @ B + B @ A


commutator = comm
anticommutator = anticomm
dagger = dag


def obs_dm(rho, d):
    """Tr[d rho] = sum_ij d_ij rho_ji."""
    if _is_torch(rho):
        return (d * rho.transpose(-1, -2)).sum(dim=(-1, -2))
    if hasattr(rho, "toarray"):
        rho = rho.toarray()
    if hasattr(d, "toarray"):
        d = d.toarray()
    return np.sum(np.asarray(d) * np.asarray(rho).T)


def rk4(rho, fun, dt, *args):
    dt2 = dt / 2.0
    k1 = fun(rho, *args)
    k2 = fun(rho + k1 * dt2, *args)
    k3 = fun(rho + k2 * dt2, *args)
    k4 = fun(rho + k3 * dt, *args)
    rho += (k1 + 2 * k2 + 2 * k3 + k4) / 6.0 * dt
    return rho


def transform(A, v):
    return dag(v) @ (A @ v)


def isherm(a, tol=1e-8):
    a = a.toarray() if hasattr(a, "toarray") else np.asarray(a)
    return np.allclose(a, np.conjugate(a).T, atol=tol)


def ket2dm(psi):
    psi = np.asarray(psi).reshape(-1)
    return np.outer(psi, psi.conj())


def basis(N, j):
    b = np.zeros(N, dtype=np.complex128)
    b[j] = 1.0
    return b


def destroy(N):
    return np.diag(np.sqrt(np.arange(1, N)), 1).astype(np.complex128)


def pauli():
    s0 = np.eye(2, dtype=np.complex128)
    sx = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    sy = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
    sz = np.array([[1, 0], [0, -1]], dtype=np.complex128)
    return s0, sx, sy, sz


def ldo(b, A):
    """Linear differential operator A b (phys.py:1409-1425)."""
    return A @ b


def expm(A, t, method="EOM", precision="fp64"):
    """U(t_k) = e^{A t_k} for the times t (phys.py:1352-1401), as lime's list of Nt arrays.

    'EOM': U(t_0) = I and each further entry one RK4 step of dU/dt = A U with
    dt = t[1] - t[0] (lime's ``interval``), i.e. U(t_k) = P(dt A)^k with the RK4 polynomial
    P; the Nt - 1 steps run in one launch sequence of the one-sided per-stage kernels
    (FP64 by default) recording every step.  'SOS' raises NotImplementedError, as in lime.
    """
    if method == "SOS":
        raise NotImplementedError("Method of SOS has not been implemented. Choose from EOM")
    if method != "EOM":
        raise ValueError(f"unknown method {method!r}")
    import torch

    from .redfield import _column_chunk, liouville_propagator
    t = np.asarray(t, dtype=np.float64)
    Nt = len(t)
    prop = liouville_propagator(A, precision)
    M = prop.n
    U = torch.eye(M, dtype=torch.complex128, device=prop.dev).unsqueeze(0).contiguous()
    out = [U[0].cpu().numpy()]
    if Nt > 1:
        dt = float(t[1] - t[0])
        done, chunk = 0, _column_chunk(M)
        while done < Nt - 1:
            m = min(chunk, Nt - 1 - done)
            st = torch.empty((1, m, M, M), dtype=torch.complex128, device=prop.dev)
            prop.step(U, dt, m, states=st, nout=1, step_offset=0, rec_capacity=m)
            out.extend(st[0].cpu().numpy())
            done += m
    return out
