"""Time-dependent (driven) Lindblad propagation: lime's ``_lindblad_driven`` (oqs.py:1691-1800).

H is ``[H0, [H1, f1], [H2, f2], ...]`` and H(t) = H0 - sum_i f_i(t) H_i is evaluated at the
END of each step, t_{k+1} = t0 + (k+1) dt, and held fixed over its four RK4 stages, as in
lime.  lime builds H(t) with ``Ht = H[0]; Ht += ...`` which aliases and mutates the caller's
H0 on every step (oqs.py:1717-1724); here H(t) is built fresh and H0 is never modified.

Each step is one launch of the fused kernel with the updated H_eff(t) (jump operators are
folded once); observables and states are recorded after every step.
"""
import numpy as np
import torch

from .mol import Result
from .propagator import LindbladPropagator, as_c128, device


def _lindblad_driven(H, rho0, c_ops=None, e_ops=None, Nt=1, dt=0.005, t0=0.0,
                     return_result=True):
    dev = device()
    H0 = as_c128(H[0], dev)
    drives = [(as_c128(h, dev), f) for h, f in H[1:]]
    e_ops = [] if e_ops is None else list(e_ops)
    prop = LindbladPropagator(H0, c_ops or [])
    heff0 = prop.heff_c128                              # H0 - (i/2) sum L^dag L, complex128
    n = prop.n
    rho = as_c128(rho0, dev).reshape(1, n, n).clone()
    E_t = torch.stack([as_c128(e, dev) for e in e_ops]).unsqueeze(0).transpose(-1, -2).contiguous() \
        if e_ops else None
    obs = torch.zeros((1, Nt, len(e_ops)), dtype=torch.complex128, device=dev) if e_ops else None
    states = torch.empty((1, Nt, n, n), dtype=torch.complex128, device=dev) if return_result else None
    t = t0
    for k in range(Nt):
        t += dt
        ht = heff0.clone()
        for h, f in drives:
            ht -= complex(f(t)) * h
        prop.heff.copy_(ht.to(prop.heff.dtype))
        prop.step(rho, dt, 1, eops_t=E_t, eop_group=1, obs=obs, states=states, nout=1,
                  step_offset=k, rec_capacity=Nt)
    if not return_result:
        with open("obs.dat", "w") as fh:
            o = obs[0].cpu().numpy() if obs is not None else np.zeros((Nt, 0))
            fmt = "{} " * (len(e_ops) + 1) + "\n"
            for k in range(Nt):
                fh.write(fmt.format(t0 + (k + 1) * dt, *o[k]))
        return rho[0].cpu().numpy()
    res = Result(dt=dt, Nt=Nt, rho0=rho0, t0=t0)
    res.observables = obs[0].cpu().numpy() if obs is not None else np.zeros((Nt, 0), np.complex128)
    res.rholist = list(states[0].cpu().numpy())
    return res
