"""lime's text output files: writers that produce the same bytes, and readers.

lime writes three plain-text files on the propagation path, one line per recorded step:

* ``cor.dat`` -- ``'{} {} \\n'.format(t, c)``: the time and one correlation value
  (lime/oqs.py:778-787 in ``_correlation_2p_1t``; lime/correlation.py:50-64);
* ``dm.dat``  -- ``('{} ' * (n*n + 1) + '\\n').format(t, *ravel(rho))``: the time and the
  operator's n^2 entries, row-major (lime/correlation.py:47-64);
* ``obs.dat`` -- ``('{} ' * (n_e + 1) + '\\n').format(t, *observables)``
  (lime/oqs.py:1630-1658 in ``_lindblad``, :413-440 in ``_redfield``).

Values are formatted by ``str`` of a NumPy complex128 (``(re+imj)``).  The time is not
``(k + 1) * dt``: lime accumulates ``t += dt`` once per step before writing, so line k
carries the floating-point sum of k + 1 copies of dt (added to the start time).
``time_column`` reproduces that sum exactly, which makes the written files
byte-compatible with lime's wherever the values agree.
"""
import numpy as np


def time_column(dt, nsteps, t0=0.0, every=1):
    """Times of the records of an ``nsteps``-step run recorded after every ``every``-th
    step, accumulated the way lime does (``t = t0; t += dt`` per step)."""
    out = []
    t = float(t0)
    dt = float(dt)
    for k in range(int(nsteps)):
        t += dt
        if (k + 1) % every == 0:
            out.append(t)
    return np.array(out, dtype=np.float64)


def _c(x):
    return np.complex128(x)


def write_cor(path, times, values):
    """``cor.dat``: one ``'{t} {value} '`` line per record."""
    with open(path, "w") as f:
        for t, v in zip(times, values):
            f.write("{} {} \n".format(float(t), _c(v)))


def write_dm(path, times, states):
    """``dm.dat``: the time and the row-major entries of each recorded operator."""
    states = np.asarray(states)
    nn = states[0].size if len(states) else 0
    fmt = "{} " * (nn + 1) + "\n"
    with open(path, "w") as f:
        for t, rho in zip(times, states):
            f.write(fmt.format(float(t), *[_c(x) for x in np.ravel(rho)]))


def write_obs(path, times, observables):
    """``obs.dat``: the time and the observables of each record."""
    observables = np.asarray(observables)
    ne = observables.shape[1] if observables.ndim == 2 else 0
    fmt = "{} " * (ne + 1) + "\n"
    with open(path, "w") as f:
        for t, row in zip(times, observables):
            f.write(fmt.format(float(t), *[_c(x) for x in row]))


def _parse_complex(tok):
    return complex(tok.strip().strip("()"))


def _read_rows(path):
    t, rows = [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            t.append(float(parts[0]))
            rows.append([_parse_complex(p) for p in parts[1:]])
    return np.array(t, dtype=np.float64), rows


def read_cor(path):
    """``cor.dat`` -> (times (Nt,), values (Nt,) complex128)."""
    t, rows = _read_rows(path)
    return t, np.array([r[0] for r in rows], dtype=np.complex128)


def read_obs(path):
    """``obs.dat`` -> (times (Nt,), observables (Nt, n_e) complex128)."""
    t, rows = _read_rows(path)
    ne = len(rows[0]) if rows else 0
    return t, np.array(rows, dtype=np.complex128).reshape(len(rows), ne)


def read_dm(path):
    """``dm.dat`` -> (times (Nt,), operators (Nt, n, n) complex128)."""
    t, rows = _read_rows(path)
    a = np.array(rows, dtype=np.complex128)
    n = int(round(np.sqrt(a.shape[1]))) if a.size else 0
    if n * n != (a.shape[1] if a.size else 0):
        raise ValueError(f"{path}: {a.shape[1]} entries per line is not a square operator")
    return t, a.reshape(len(rows), n, n)
