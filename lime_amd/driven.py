"""Time-dependent (driven) Lindblad propagation: lime's ``_lindblad_driven`` (oqs.py:1691-1800).

H is ``[H0, [H1, f1], [H2, f2], ...]`` (Hermitian H_i, real f_i: the kernels fold the
dissipator into H_eff = H - (i/2) sum L^dag L and form X H_eff^dag, which is lime's X H only
for Hermitian H) and H(t) = H0 - sum_i f_i(t) H_i is evaluated at the END of each step, t_{k+1} = t0 + (k+1) dt, and held fixed over its four RK4 stages, as in
lime.  lime builds H(t) with ``Ht = H[0]; Ht += ...`` which aliases and mutates the caller's
H0 on every step (oqs.py:1717-1724); here H(t) is built fresh and H0 is never modified.

Each step is one launch of the fused kernel with H_eff(t) (jump operators are folded once):
the drive coefficients f_i(t_k) are evaluated on the host up front, and the H_eff(t_k) of a
chunk of steps are formed on the GPU in one product, so a step launches nothing but the
propagation; observables and states are recorded after every step.
"""
import numpy as np
import torch

from . import datfile
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
    # f_i at t_{k+1}, accumulated as lime does (t += dt before each step)
    coef = np.zeros((Nt, len(drives)), np.complex128)
    t = t0
    for k in range(Nt):
        t += dt
        for i, (_, f) in enumerate(drives):
            coef[k, i] = complex(f(t))
    Hd = torch.stack([h for h, _ in drives]).reshape(len(drives), n * n) if drives else None
    chunk = max(1, min(Nt, (64 << 20) // (16 * n * n)))   # H_eff(t) of up to 64 MiB of steps
    for c0 in range(0, Nt, chunk):
        c1 = min(Nt, c0 + chunk)
        if Hd is not None:
            cf = torch.as_tensor(coef[c0:c1], device=dev)
            heff_t = (heff0.reshape(1, n * n) - cf @ Hd).reshape(c1 - c0, n, n).to(prop.heff.dtype)
        for k in range(c0, c1):
            if Hd is not None:
                prop.heff = heff_t[k - c0]
            prop.step(rho, dt, 1, eops_t=E_t, eop_group=1, obs=obs, states=states, nout=1,
                      step_offset=k, rec_capacity=Nt)
    if not return_result:
        # lime's files (oqs.py:1743-1773): obs.dat with the accumulated time column, and an
        # empty den_mat.dat that lime opens and closes without writing
        o = obs[0].cpu().numpy() if obs is not None else np.zeros((Nt, 0), np.complex128)
        datfile.write_obs("obs.dat", datfile.time_column(dt, Nt, t0), o)
        open("den_mat.dat", "w").close()
        return rho[0].cpu().numpy()
    res = Result(dt=dt, Nt=Nt, rho0=rho0, t0=t0)
    res.observables = obs[0].cpu().numpy() if obs is not None else np.zeros((Nt, 0), np.complex128)
    res.rholist = list(states[0].cpu().numpy())
    return res
