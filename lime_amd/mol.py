"""Solver return container (lime/mol.py:78-110) with the intended semantics.

lime's ``Result(dt, Nt, rho0)`` evaluates ``t0 + arange(Nt // nout)`` with ``t0 = nout = None``
(lime/mol.py:92) and raises TypeError on the default solver path.  Here ``t0`` defaults to
0 and ``nout`` to 1, and ``times[r] = t0 + (r + 1) * nout * dt`` -- the instants at which the
solvers record (after each recorded step, lime/oqs.py:1674-1682).
"""
import numpy as np


class Result:
    def __init__(self, description=None, psi0=None, rho0=None, dt=None, Nt=None, times=None,
                 t0=0.0, nout=1):
        self.description = description
        self.dt = dt
        self.timesteps = Nt
        self.observables = None
        self.rholist = None
        self.psilist = [psi0]
        self.psi = None
        self.rho0 = rho0
        self.psi0 = psi0
        self.nout = 1 if nout is None else nout
        t0 = 0.0 if t0 is None else t0
        if times is not None:
            self.times = np.asarray(times)
        elif dt is not None and Nt is not None:
            self.times = t0 + (np.arange(Nt // self.nout) + 1) * dt * self.nout
        else:
            self.times = None

    def expect(self):
        return self.observables
