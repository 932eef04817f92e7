"""lime.correlation call surface on the MI355X propagator (lime/correlation.py:17-69).

``correlation_3p_1t(H, rho0, ops, c_ops, tlist, dyn)`` is lime's older QRT driver: the
propagated operator starts at C rho0 A (ops = [A, B, C]), every RK4 step records
cor = Tr[B rho] and the operator itself, and the two text files lime writes are produced
in lime's formats (``lime_amd.datfile``): ``cor.dat`` ('{t} {cor} ' per line) and
``dm.dat`` (t followed by the n^2 entries of rho, row-major, as Python complex literals),
with t accumulated as lime does (``t += dt`` per step), dt = tlist[1] - tlist[0] and
Nt = len(tlist).  The propagation is the fused kernel; the
files are written after the run rather than per step.  lime passes the right-hand side as
``dyn`` (its Lindblad ``liouvillian``); that is the only dynamics the fused kernel runs, so
any other ``dyn`` raises ValueError.  Returns None, as lime does.
"""
import numpy as np

from . import datfile
from .propagator import LindbladPropagator


def _dense(a):
    return a.toarray() if hasattr(a, "toarray") else np.asarray(a)


def correlation_3p_1t(H, rho0, ops, c_ops, tlist, dyn=None, *args, output="cor.dat",
                      dm_output="dm.dat"):
    if dyn is not None and getattr(dyn, "__name__", "") not in ("liouvillian", "lindblad"):
        raise ValueError("correlation_3p_1t propagates the Lindblad right-hand side only "
                         f"(got dyn={getattr(dyn, '__name__', dyn)!r})")
    A, B, C = (_dense(o).astype(np.complex128) for o in ops)
    H = _dense(H)
    n = H.shape[-1]
    tlist = np.asarray(tlist, dtype=np.float64)
    Nt = len(tlist)
    dt = float(tlist[1] - tlist[0])
    X0 = C @ _dense(rho0).astype(np.complex128) @ A
    prop = LindbladPropagator(H, [_dense(c) for c in (c_ops or [])])
    _, obs, states = prop.run(X0, dt, Nt, e_ops=[B], store_states=True)
    cor = obs[0, :, 0].cpu().numpy()
    dm = states[0].cpu().numpy()
    t = datfile.time_column(dt, Nt)
    datfile.write_cor(output, t, cor)
    datfile.write_dm(dm_output, t, dm)
    return None
