"""FP64 restatement of lime's Lindblad RK4 path (TEST INFRASTRUCTURE ONLY).

Follows, function by function:

* ``rk4``               -- ``lime/phys.py:636-649`` (classic 4-stage RK4, in-place update)
* ``comm/anticomm/dag`` -- ``lime/phys.py:741-759``
* ``obs_dm``            -- ``lime/phys.py:837-844`` (``Tr[d rho]``, no conjugate)
* ``liouvillian_ref``   -- ``lime/oqs.py:706-713`` (the 2 + 5K matmul operation sequence)
* ``lindbladian_ref``   -- ``lime/oqs.py:716-723``
* ``propagate``         -- ``lime/oqs.py:1590-1688`` (``_lindblad``: copy, promote to
                           complex128, observables and states recorded *after* each step)
* ``correlation_2op_1t``-- ``lime/oqs.py:726-800``  (rho <- B rho0, cor[k] = Tr[A rho(t_{k+1})])
* ``correlation_3op_1t``-- ``lime/oqs.py:1227-1246`` (rho <- C rho0 A, observe B)
* ``correlation_3p_1t`` -- ``lime/correlation.py:17-69`` (the generator of cor.dat / dm.dat)
* ``correlation_3op_2t``-- ``lime/oqs.py:1268-1299``

Pinned by: the reference-held golden vectors ``examples/cor.dat``/``dm.dat``
(dissipator with non-diagonal L, RK4, trace observable, QRT initial operator, time grid);
the exact discrete RK4 closed form R(z)^n for diagonal H and L (commutator sign,
dephasing rates); the Liouville-space construction in ``superop`` (commutator with
general H and dense L); scipy ``expm`` in the dt -> 0 limit (O(dt^4) convergence).

``rhs_heff`` is the H_eff-folded form used by the GPU kernels:
    f(rho) = -i (H_eff rho - rho H_eff^dag) + D o rho + sum_dense L rho L^dag
with H_eff = H - (i/2) sum_k L_k^dag L_k, and every *diagonal* jump operator
L = diag(l) contributing D += l l^dag (Hadamard multiplier).  Pinned against
``liouvillian_ref`` on random inputs (an independent operation sequence).
"""
import numpy as np


# --------------------------------------------------------------------------- primitives
def dag(a):
    """Conjugate transpose (lime/phys.py:758)."""
    return np.conjugate(a).T


def comm(a, b):
    """[A, B] = AB - BA (lime/phys.py:741-743)."""
    assert a.shape == b.shape
    return a @ b - b @ a


def anticomm(a, b):
    """{A, B} = AB + BA (lime/phys.py:746-748)."""
    assert a.shape == b.shape
    return a @ b + b @ a


def obs_dm(rho, d):
    """Tr[d rho] without conjugating d (lime/phys.py:837-844)."""
    return np.trace(d @ rho)


def rk4(rho, fun, dt, *args):
    """Classic RK4 with the reference's stage order and in-place update (phys.py:636-649)."""
    dt2 = dt / 2.0
    k1 = fun(rho, *args)
    k2 = fun(rho + k1 * dt2, *args)
    k3 = fun(rho + k2 * dt2, *args)
    k4 = fun(rho + k3 * dt, *args)
    rho += (k1 + 2 * k2 + 2 * k3 + k4) / 6.0 * dt
    return rho


# --------------------------------------------------------------------------- RHS forms
def lindbladian_ref(l, rho):
    """L rho L^dag - 1/2 {L^dag L, rho}, re-forming L^dag L every call (oqs.py:716-723)."""
    return l @ (rho @ dag(l)) - 0.5 * anticomm(dag(l) @ l, rho)


def liouvillian_ref(rho, H, c_ops):
    """-i[H, rho] + sum_k D[L_k] rho with lime's 2 + 5K matmul sequence (oqs.py:706-713)."""
    rhs = -1j * comm(H, rho)
    for c in c_ops:
        rhs = rhs + lindbladian_ref(c, rho)
    return rhs


def is_diagonal(a, tol=0.0):
    off = a - np.diag(np.diag(a))
    return np.max(np.abs(off)) <= tol if off.size else True


def split_ops(H, c_ops):
    """Fold the jump operators into (H_eff, D, dense_L).

    H_eff = H - (i/2) sum_k L_k^dag L_k ; D = sum_{diag L} l l^dag ; dense_L = the rest.
    """
    H = np.asarray(H, dtype=np.complex128)
    n = H.shape[0]
    gamma = np.zeros((n, n), dtype=np.complex128)
    D = np.zeros((n, n), dtype=np.complex128)
    dense = []
    for c in c_ops or []:
        c = np.asarray(c, dtype=np.complex128)
        gamma += dag(c) @ c
        if is_diagonal(c):
            l = np.diag(c)
            D += np.outer(l, np.conj(l))
        else:
            dense.append(c)
    return H - 0.5j * gamma, D, dense


def rhs_heff(rho, Heff, D, dense):
    """-i(H_eff rho - rho H_eff^dag) + D o rho + sum_dense L rho L^dag."""
    out = -1j * (Heff @ rho - rho @ dag(Heff)) + D * rho
    for l in dense:
        out = out + (l @ rho) @ dag(l)
    return out


# --------------------------------------------------------------------------- drivers
def propagate(H, rho0, c_ops, dt, Nt, e_ops=None, nout=1, store_states=False,
              form="heff"):
    """RK4 propagation with _lindblad's recording convention (oqs.py:1590-1688).

    Records after every ``nout``-th step (t = (k+1) dt for recorded step k).  Returns
    (rho_final, observables (Nt//nout, n_e) complex128, states list or None).
    ``form='ref'`` uses lime's 2 + 5K operation sequence, ``'heff'`` the folded form.
    """
    rho = np.array(rho0, dtype=np.complex128, copy=True)
    e_ops = [] if e_ops is None else [np.asarray(e) for e in e_ops]
    if form == "ref":
        fun, args = liouvillian_ref, (np.asarray(H), [np.asarray(c) for c in (c_ops or [])])
    else:
        fun, args = rhs_heff, split_ops(H, c_ops)
    nrec = Nt // nout
    obs = np.zeros((nrec, len(e_ops)), dtype=np.complex128)
    states = [] if store_states else None
    r = 0
    for k in range(Nt):
        rho = rk4(rho, fun, dt, *args)
        if (k + 1) % nout == 0 and r < nrec:
            for i, e in enumerate(e_ops):
                obs[r, i] = obs_dm(rho, e)
            if store_states:
                states.append(rho.copy())
            r += 1
    return rho, obs, states


def correlation_2op_1t(H, rho0, a_op, b_op, c_ops, dt, Nt):
    """<A(t) B> by QRT: rho <- B rho0, cor[k] = Tr[A rho(t_{k+1})] (oqs.py:726-800)."""
    _, obs, _ = propagate(H, np.asarray(b_op) @ np.asarray(rho0), c_ops, dt, Nt, [a_op])
    return obs[:, 0]


def correlation_3op_1t(H, rho0, a_op, b_op, c_op, c_ops, dt, Nt):
    """<A B(t) C>: rho <- C rho0 A, observe B (oqs.py:1227-1246)."""
    x0 = np.asarray(c_op) @ np.asarray(rho0) @ np.asarray(a_op)
    _, obs, _ = propagate(H, x0, c_ops, dt, Nt, [b_op])
    return obs[:, 0]


def correlation_3p_1t(H, rho0, ops, c_ops, dt, Nt):
    """lime/correlation.py:17-69: rho <- C rho0 A; per step cor = Tr[B rho], dm = rho.

    Uses lime's reference operation sequence (``liouvillian_ref``).  Returns (t, cor, dm).
    """
    A, B, C = [np.asarray(o) for o in ops]
    rho = C @ np.asarray(rho0) @ A
    rho = rho.astype(np.complex128)
    t = 0.0
    ts, cor, dm = [], [], []
    for _ in range(Nt):
        t += dt
        rho = rk4(rho, liouvillian_ref, dt, np.asarray(H), [np.asarray(c) for c in c_ops])
        ts.append(t)
        cor.append(np.trace(B @ rho))
        dm.append(rho.copy())
    return np.array(ts), np.array(cor), np.array(dm)


def correlation_3op_2t(H, rho0, ops, c_ops, dt, Nt, Ntau):
    """<A(t) B(t+tau) C(t)> (oqs.py:1268-1299): rho(t_k) list, then C rho A over Ntau."""
    a, b, c = [np.asarray(o) for o in ops]
    _, _, rho_t = propagate(H, rho0, c_ops, dt, Nt, store_states=True)
    out = np.zeros((Nt, Ntau), dtype=np.complex128)
    for i, rho in enumerate(rho_t):
        _, obs, _ = propagate(H, c @ rho @ a, c_ops, dt, Ntau, [b])
        out[i] = obs[:, 0]
    return out


def rk4_stability_factor(z):
    """R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24: one classic RK4 step of y' = lambda y."""
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


def propagate_driven(H, rho0, c_ops, dt, Nt, e_ops=None, t0=0.0):
    """lime's _lindblad_driven (oqs.py:1691-1800) without its H0 aliasing.

    H = [H0, [H1, f1], ...]; H(t) = H0 - sum_i f_i(t) H_i evaluated at t = t0 + (k+1) dt
    and held over the four RK4 stages of step k.  Returns (rho, obs (Nt, n_e), states).
    """
    H0 = np.asarray(H[0], dtype=np.complex128)
    drives = [(np.asarray(h, dtype=np.complex128), f) for h, f in H[1:]]
    e_ops = [] if e_ops is None else [np.asarray(e) for e in e_ops]
    rho = np.array(rho0, dtype=np.complex128, copy=True)
    c_ops = [np.asarray(c) for c in (c_ops or [])]
    obs = np.zeros((Nt, len(e_ops)), dtype=np.complex128)
    states = []
    t = t0
    for k in range(Nt):
        t += dt
        Ht = H0.copy()
        for h, f in drives:
            Ht = Ht - f(t) * h
        rho = rk4(rho, liouvillian_ref, dt, Ht, c_ops)
        obs[k] = [obs_dm(rho, e) for e in e_ops]
        states.append(rho.copy())
    return rho, obs, states
