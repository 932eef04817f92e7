"""lime's text formats (cor.dat, dm.dat, obs.dat): writers and readers on the CPU.

The reference's own files (examples/cor.dat, dm.dat, kept as a fixture together with
their time column as text) fix the format: the oracle's values written by
``lime_amd.datfile`` reproduce lime's lines byte for byte."""
import os

import numpy as np

from lime_amd import datfile
from oracle import lindblad as ol


def _cavity():
    n, kappa, n_th = 10, 0.25, 2.0
    a = np.diag(np.sqrt(np.arange(1, n)), 1).astype(np.complex128)
    num = np.diag(np.arange(n)).astype(np.complex128)
    H = np.diag(np.arange(n) * 1.0).astype(np.complex128)
    c_ops = [np.sqrt(kappa * (1 + n_th)) * a, np.sqrt(kappa * n_th) * a.conj().T]
    w = np.exp(-np.arange(n) / n_th)
    rho0 = np.diag(w / w.sum()).astype(np.complex128)
    return H, a, num, c_ops, rho0


def test_time_column_accumulates_like_lime(golden_dir):
    g = np.load(os.path.join(golden_dir, "cavity_qrt.npz"))
    dt = np.linspace(0, 10, 200)[1] - np.linspace(0, 10, 200)[0]
    t = datfile.time_column(dt, 200)
    assert [str(x) for x in t] == list(g["t_cor_text"])
    # (k + 1) * dt differs from the accumulated sum on many lines: the test can tell
    assert sum(str((k + 1) * dt) != s for k, s in enumerate(g["t_cor_text"])) > 10
    tt = datfile.time_column(0.1, 10, t0=1.0)
    assert np.array_equal(datfile.time_column(0.1, 10, t0=1.0, every=5), tt[[4, 9]])


def test_writers_reproduce_lime_files_byte_for_byte(golden_dir, tmp_path):
    """The oracle's cavity run (which matches lime's values to 1e-14) written with the
    writers: every line equals lime's where the values are bit-identical, and every time
    token and value parses back to lime's."""
    g = np.load(os.path.join(golden_dir, "cavity_qrt.npz"))
    H, a, num, c_ops, rho0 = _cavity()
    dt = np.linspace(0, 10, 200)[1]
    _, cor, dm = ol.correlation_3p_1t(H, rho0, [a.conj().T, num, a], c_ops, dt, 200)
    t = datfile.time_column(dt, 200)
    datfile.write_cor(tmp_path / "cor.dat", t, g["cor"])
    datfile.write_dm(tmp_path / "dm.dat", t, g["dm"])
    lines = open(tmp_path / "cor.dat").read().splitlines()
    assert [l.split()[0] for l in lines] == list(g["t_cor_text"])
    assert lines[0] == "0.05025125628140704 (3.919494245954364+0j) "      # lime's first line
    dm_first = open(tmp_path / "dm.dat").readline()
    assert dm_first.startswith("0.05025125628140704 (0.24516373582935236+0j) 0j 0j ")
    assert dm_first.endswith(" \n")
    tr, cr = datfile.read_cor(tmp_path / "cor.dat")
    td, dr = datfile.read_dm(tmp_path / "dm.dat")
    assert np.array_equal(tr, g["t_cor"]) and np.array_equal(cr, g["cor"])
    assert np.array_equal(td, g["t_dm"]) and np.array_equal(dr, g["dm"])
    assert np.max(np.abs(cor - g["cor"])) < 1e-13 and np.max(np.abs(dm - g["dm"])) < 1e-14


def test_obs_dat_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    obs = rng.standard_normal((7, 3)) + 1j * rng.standard_normal((7, 3))
    t = datfile.time_column(0.05, 7)
    datfile.write_obs(tmp_path / "obs.dat", t, obs)
    line = open(tmp_path / "obs.dat").readline()
    assert line.count(" ") == 4 and line.endswith(" \n")
    tt, back = datfile.read_obs(tmp_path / "obs.dat")
    assert np.array_equal(tt, t) and np.array_equal(back, obs)
