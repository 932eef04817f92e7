"""Convert the reference's Lindblad-RK4 QRT output files into a small .npz fixture.

Source data (read as text, never executed): ``examples/cor.dat`` and
``examples/dm.dat`` of binggu56/lime.  Both were written by
``lime/correlation.py:17-69`` (``correlation_3p_1t``) when driven by
``examples/test_cavity.py:70-94``:

* 10-level damped cavity, ``a`` = superdiagonal sqrt(1..9) (``cavity.py:453-458``),
  ``num = diag(0..9)`` (``cavity.py:492-500``), ``H = diag(0..9)`` (``cavity.py:280-297``);
* ``c_ops = [sqrt(kappa (1 + n_th)) a, sqrt(kappa n_th) a^dag]`` with kappa = 0.25, n_th = 2;
* ``rho0 = diag(exp(-n / n_th)) / sum`` (``test_cavity.py:46-64``);
* ``tlist = linspace(0, 10, 200)`` so dt = 10/199 and Nt = 200;
* ``ops = [a^dag, num, a]``: the propagated operator starts at ``a rho0 a^dag`` and
  after every RK4 step ``cor = Tr[num rho]`` and ``dm = ravel(rho)`` (row-major) are written,
  each line tagged with t, accumulated as ``t += dt`` from 0 (so the k-th line carries the
  floating-point sum of k + 1 copies of dt, not (k + 1) * dt).

The time column of both files is kept verbatim as text (``t_cor_text``, ``t_dm_text``) so
that writers can be checked for byte-compatible output.

Usage (in a container that holds the reference)::

    python tests/golden/make_cavity_fixture.py /root/reference/examples tests/golden/cavity_qrt.npz
"""
import sys

import numpy as np


def _parse_complex(tok: str) -> complex:
    return complex(tok.strip().strip("()"))


def read_cor(path):
    t, cor, text = [], [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            text.append(parts[0])
            t.append(float(parts[0]))
            cor.append(_parse_complex(parts[1]))
    return np.array(t), np.array(cor, dtype=np.complex128), np.array(text)


def read_dm(path):
    t, rows, text = [], [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            text.append(parts[0])
            t.append(float(parts[0]))
            rows.append([_parse_complex(p) for p in parts[1:]])
    return np.array(t), np.array(rows, dtype=np.complex128), np.array(text)


def main(src_dir, out_path):
    t_cor, cor, t_cor_text = read_cor(f"{src_dir}/cor.dat")
    t_dm, dm, t_dm_text = read_dm(f"{src_dir}/dm.dat")
    n = int(round(np.sqrt(dm.shape[1])))
    np.savez_compressed(out_path, t_cor=t_cor, cor=cor, t_dm=t_dm,
                        dm=dm.reshape(len(t_dm), n, n),
                        t_cor_text=t_cor_text, t_dm_text=t_dm_text,
                        kappa=0.25, n_th=2.0, nlevels=n, t_final=10.0, npts=200)
    print(f"wrote {out_path}: cor {cor.shape}, dm {dm.shape} -> ({len(t_dm)}, {n}, {n})")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
