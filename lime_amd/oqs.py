"""lime.oqs call surface on the MI355X propagator (drop-in for lime/oqs.py's Lindblad path).

Every propagation below runs the fused HIP kernel ``lime_lindblad_rk4`` through
``lime_amd.propagator`` -- there is no NumPy time loop and no CPU fallback.

Reference functions and the decisions taken where lime's behaviour is broken:

* ``_lindblad``            -- oqs.py:1590-1688.  Copies rho0 (never mutated), complex128
  state, records observables/states *after* each step (t_k = t0 + (k+1) dt; the text
  files carry lime's accumulated ``t += dt`` times, ``lime_amd.datfile``).  lime's
  ``Result`` crash (mol.py:92) is fixed (``lime_amd.mol.Result``).  With
  ``return_result=False`` lime writes observables taken *before* the step under the time
  *after* it (oqs.py:1638-1658); here ``obs.dat`` holds the after-step values.
* ``_correlation_2p_1t``   -- oqs.py:726-800: <A(t)B> with rho <- B rho0, cor[k] at
  t = (k+1) dt, the text file ``output`` in lime's ``'{t} {complex} \\n'`` format (written
  after the run rather than per step); an unknown ``method`` raises ValueError instead of
  lime's ``sys.exit``.
* ``Lindblad_solver``      -- oqs.py:1117-1331 (setH / set_c_ops / set_e_ops / configure /
  liouvillian / evolve / evolve_fast / correlation_2op_1t / correlation_3op_1t /
  correlation_4op_1t (the overriding definition, :1301-1315) / correlation_3op_2t /
  correlation_4op_2t).  ``correlation_3op_2t`` runs the Nt x Ntau two-time grid as ONE
  batched launch of Nt operators instead of lime's Nt serial propagations.
* ``liouvillian``/``lindbladian`` -- oqs.py:706-723, single right-hand-side evaluations.

Extensions (lime has no batching): ``H`` may be (B, N, N) and ``rho0`` (B, N, N); keyword
``nout`` records every nout-th step; ``precision`` selects the GEMM arithmetic;
``renorm`` renormalises Tr rho = 1 after each step (density matrices only).
"""
import numpy as np
import torch

from . import datfile
from . import phys
from . import superoperator as superop
from .mol import Result
from .propagator import LindbladPropagator, as_c128, device
from .redfield import Redfield_solver, _redfield, getG, redfield_tensor  # noqa: F401  (lime.oqs names)


def _np(t):
    return t.detach().cpu().numpy()


def _is_batched(H, rho0):
    return np.ndim(H) == 3 or (hasattr(rho0, "ndim") and rho0.ndim == 3)


def _run(H, rho0, c_ops, dt, Nt, e_ops, nout=1, store_states=False, renorm=False,
         precision="fp32"):
    """Propagate one operator or a batch; returns device tensors (rho, obs, states)."""
    dev = device()
    Ht = as_c128(H, dev)
    r0 = as_c128(rho0, dev)
    if Ht.dim() == 2 and r0.dim() == 3:               # shared H, batch of initial operators
        group = r0.shape[0]
    else:
        group = 1
    prop = LindbladPropagator(Ht, c_ops or [], group=group, precision=precision)
    if r0.dim() == 2 and prop.batch > 1:
        r0 = r0.unsqueeze(0).expand(prop.batch, *r0.shape).contiguous()
    return prop.run(r0, dt, Nt, e_ops=e_ops, nout=nout, store_states=store_states, renorm=renorm)


# ------------------------------------------------------------------ right-hand side
def _apply_generator(rho, H, c_ops):
    """One application of the Lindblad generator on the GPU (FP64-MFMA generator kernel,
    ``lime_lindblad_apply``).  rho: (N, N) or (B, N, N); NumPy in -> NumPy out."""
    dev = device()
    as_numpy = not isinstance(rho, torch.Tensor)
    r = as_c128(rho, dev)
    single = r.dim() == 2
    if single:
        r = r.unsqueeze(0)
    Ht = as_c128(H, dev)
    group = r.shape[0] if Ht.dim() == 2 else 1
    prop = LindbladPropagator(Ht, c_ops or [], group=group, precision="fp64")
    if prop.batch != r.shape[0]:
        raise ValueError(f"rho has {r.shape[0]} operators, H has {prop.nsets} sets")
    y = torch.empty_like(r)
    prop.apply(r.contiguous(), y=y, c_lx=1.0)
    out = y[0] if single else y
    return _np(out) if as_numpy else out


def lindbladian(l, rho):
    """L rho L^dag - 1/2 {L^dag L, rho} (oqs.py:716-723), evaluated on the GPU."""
    n = (rho.shape[-1])
    return _apply_generator(rho, np.zeros((n, n), dtype=np.complex128), [l])


def liouvillian(rho, H, c_ops):
    """-i[H, rho] + sum_k D[L_k] rho (oqs.py:706-713), evaluated on the GPU."""
    return _apply_generator(rho, H, c_ops)


# ------------------------------------------------------------------ drivers
def _lindblad(H, rho0, c_ops, e_ops=None, Nt=1, dt=0.005, return_result=True, t0=0.0,
              nout=1, store_states=True, renorm=False, precision="fp32"):
    e_ops = [] if e_ops is None else list(e_ops)
    batched = _is_batched(H, rho0)
    rho, obs, states = _run(H, rho0, c_ops, dt, Nt, e_ops, nout=nout,
                            store_states=(store_states and return_result), renorm=renorm,
                            precision=precision)
    nrec = Nt // nout
    obs_np = _np(obs) if obs is not None else np.zeros((rho.shape[0], nrec, 0), np.complex128)
    if not return_result:
        t = datfile.time_column(dt, Nt, t0=t0, every=nout)[:nrec]
        rows = obs_np[0] if not batched else obs_np.transpose(1, 0, 2).reshape(nrec, -1)
        datfile.write_obs("obs.dat", t, rows)
        out = _np(rho)
        return out if batched else out[0]
    result = Result(dt=dt, Nt=Nt, rho0=rho0, t0=t0, nout=nout)
    if batched:
        result.observables = obs_np
        result.rholist = _np(states) if states is not None else None
    else:
        result.observables = obs_np[0]
        result.rholist = list(_np(states[0])) if states is not None else None
    return result


def _correlation_2p_1t(H, rho0, ops, c_ops, dt, Nt, method="lindblad", output="cor.dat"):
    if method != "lindblad":
        raise ValueError("The method {} has not been implemented yet! Please try lindblad.".format(method))
    A, B = ops
    x0 = as_c128(B, device()) @ as_c128(rho0, device())
    _, obs, _ = _run(H, x0, c_ops, dt, Nt, [A])
    cor = _np(obs)[0, :, 0] if obs.shape[0] == 1 else _np(obs)[:, :, 0]
    if output:
        t = datfile.time_column(dt, Nt)
        if cor.ndim == 1:
            datfile.write_cor(output, t, cor)
        else:                              # a batch: one column per operator
            datfile.write_obs(output, t, cor.T)
    return cor


def _fast_lindblad(H, rho0, c_ops, e_ops=None, Nt=1, dt=0.005):
    """lime's intended fast path (oqs.py:1485-1564): final rho only."""
    rho, _, _ = _run(H, rho0, c_ops, dt, Nt, None)
    out = _np(rho)
    return out if _is_batched(H, rho0) else out[0]


class Lindblad_solver:
    """oqs.py:1117-1331 on the GPU kernel."""

    def __init__(self, H=None, c_ops=None, e_ops=None, precision="fp32"):
        self.c_ops = c_ops
        self.e_ops = e_ops
        self.H = H
        self.precision = precision

    def set_c_ops(self, c_ops):
        self.c_ops = c_ops

    def set_e_ops(self, e_ops):
        self.e_ops = e_ops

    def setH(self, H):
        self.H = H

    def configure(self, c_ops, e_ops):
        self.c_ops = c_ops
        self.e_ops = e_ops

    def liouvillian(self):
        return superop.liouvillian(self.H, self.c_ops)

    def steady_state(self):
        raise NotImplementedError("steady_state is a stub in lime (oqs.py:1149) as well")

    def evolve(self, rho0, dt, Nt, t0=0.0, e_ops=None, return_result=True, nout=1,
               store_states=True, renorm=False):
        if isinstance(self.H, list):
            from .driven import _lindblad_driven
            return _lindblad_driven(self.H, rho0=rho0, c_ops=self.c_ops, e_ops=e_ops, Nt=Nt,
                                    dt=dt, t0=t0)
        return _lindblad(self.H, rho0, c_ops=self.c_ops, e_ops=e_ops, Nt=Nt, dt=dt,
                         return_result=return_result, t0=t0, nout=nout,
                         store_states=store_states, renorm=renorm, precision=self.precision)

    def evolve_fast(self, rho0, dt, Nt, t0=0.0, e_ops=None, return_result=True, fast=True):
        return _fast_lindblad(self.H, rho0, self.c_ops, e_ops=e_ops, Nt=Nt, dt=dt)

    def correlation_2op_1t(self, rho0, a_op, b_op, dt, Nt, output="cor.dat"):
        return _correlation_2p_1t(self.H, rho0, ops=[a_op, b_op], c_ops=self.c_ops, dt=dt,
                                  Nt=Nt, output=output)

    def correlation_3op_1t(self, rho0, oplist, dt=0.005, Nt=1):
        """<A B(t) C>: rho <- C rho0 A, observe B (oqs.py:1227-1246)."""
        a_op, b_op, c_op = oplist
        dev = device()
        x0 = as_c128(c_op, dev) @ as_c128(rho0, dev) @ as_c128(a_op, dev)
        _, obs, _ = _run(self.H, x0, self.c_ops, dt, Nt, [b_op], precision=self.precision)
        o = _np(obs)[:, :, 0]
        return o[0] if o.shape[0] == 1 else o

    def correlation_4op_1t(self, rho0, ops, dt=0.005, nt=1):
        """<A B(t) C(t) D> (the overriding definition, oqs.py:1301-1315)."""
        if len(ops) != 4:
            raise ValueError("Number of operators is not 4.")
        a, b, c, d = ops
        return self.correlation_3op_1t(rho0, [a, b @ c, d], dt, nt)

    def correlation_3op_2t(self, rho0, ops, dt, Nt, Ntau):
        """<A(t) B(t+tau) C(t)> (oqs.py:1268-1299) as (Nt, Ntau) complex128.

        rho(t_k) for k = 1..Nt from one propagation, then all Nt operators C rho(t_k) A
        advance together as one batched launch over Ntau steps observing B.
        """
        dev = device()
        a_op, b_op, c_op = [as_c128(o, dev) for o in ops]
        _, _, states = _run(self.H, rho0, self.c_ops, dt, Nt, None, store_states=True,
                            precision=self.precision)
        X = torch.matmul(torch.matmul(c_op, states[0]), a_op).contiguous()     # (Nt, N, N)
        _, obs, _ = _run(self.H, X, self.c_ops, dt, Ntau, [b_op], precision=self.precision)
        return _np(obs)[:, :, 0]

    def correlation_4op_2t(self, rho0, ops, dt, nt, ntau):
        """<A(t) B(t+tau) C(t+tau) D(t)> (oqs.py:1317-1331)."""
        if len(ops) != 4:
            raise ValueError("Number of operators is not 4.")
        a, b, c, d = ops
        return self.correlation_3op_2t(rho0, [a, b @ c, d], dt, nt, ntau)


class HEOMSolverDL(Lindblad_solver):
    """oqs.py:1335-1431.  Despite its name, lime's class propagates the Lindblad master
    equation (``solve`` -> ``_lindblad``, ``correlation_2op_1t`` -> ``_correlation_2p_1t``,
    ``correlation_3op_2t`` the same two-time driver as ``Lindblad_solver``); it runs on the
    same fused kernel here."""

    def solve(self, rho0, dt, Nt, return_result=True):
        return _lindblad(self.H, rho0, self.c_ops, e_ops=self.e_ops, Nt=Nt, dt=dt,
                         return_result=return_result, precision=self.precision)


class OQS:
    """oqs.py:583-697 (open-quantum-system container).  lime's ``set_hamiltonian``/``setH``
    assign an undefined name and ``correlation_2p_1t`` calls an undefined module function
    (oqs.py:601-607, :694); here they set the Hamiltonian and run ``_correlation_2p_1t``
    (QRT <A(t)B> on the fused kernel, ``output`` written in lime's cor.dat format)."""

    def __init__(self, H, c_ops=None):
        self.hamiltonian = H
        self.h_sys = H
        self.H = H
        self.nstates = np.shape(H)[-1]
        self.e_ops = None
        self.c_ops = c_ops

    def set_hamiltonian(self, h):
        self.hamiltonian = self.h_sys = self.H = h
        self.nstates = np.shape(h)[-1]

    setH = set_hamiltonian

    def set_c_ops(self, c_ops):
        self.c_ops = c_ops

    def set_e_ops(self, e_ops):
        self.e_ops = e_ops

    def configure(self, c_ops, e_ops):
        self.c_ops = c_ops
        self.e_ops = e_ops

    def correlation_2p_1t(self, rho0, ops, dt, Nt, method="lindblad", output="cor.dat"):
        return _correlation_2p_1t(self.hamiltonian, rho0, ops=ops, c_ops=self.c_ops or [],
                                  dt=dt, Nt=Nt, method=method, output=output)


# ------------------------------------------------------------------ state preparation
def coherent(N, alpha):
    """Coherent state D(alpha)|0> in an N-level truncated Fock space (oqs.py:856-922): the
    displacement operator exp(alpha a^dag - conj(alpha) a) of the truncated ladder operator
    applied to the vacuum, so the state is normalised.  Returns a 1-D complex128 array."""
    from scipy.linalg import expm
    a = phys.destroy(N)
    D = expm(alpha * a.conj().T - np.conj(alpha) * a)
    return D @ phys.basis(N, 0)


def coherent_dm(N, alpha):
    """|alpha><alpha| of ``coherent(N, alpha)`` (oqs.py:926-981)."""
    return phys.ket2dm(coherent(N, alpha))
