"""lime's Fourier-transform convention (TEST INFRASTRUCTURE ONLY).

* ``fft``  -- ``lime/fft.py:9-56``: g(w) = sum_n f(t_n) e^{-i w t_n} dx, fftshifted,
              w = 2 pi fftshift(fftfreq(n, dx)), times the phase e^{-i w x0}.
* ``ifft`` -- ``lime/fft.py:58-78``: g = (1/2pi) sum f e^{+i w t} dx.
* ``fft2`` -- ``lime/fft.py:80-102`` (the y frequencies come from the x length, as there).

Pinned by: the analytic transform of a damped exponential (a Lorentzian) sampled on the
same grid -- a closed form independent of numpy's FFT; for ``fft2`` the product of two such
closed forms for a separable exponential.
"""
import numpy as np


def fft(f, x=None, axis=-1):
    f = np.asarray(f)
    nx = f.shape[axis]
    if x is None:
        x = np.arange(nx)
    x = np.asarray(x)
    dx = x[1] - x[0]
    g = np.fft.fftshift(np.fft.fft(f, axis=axis), axes=(axis,)) * dx
    freq = 2.0 * np.pi * np.fft.fftshift(np.fft.fftfreq(nx, d=dx))
    shape = [1] * g.ndim
    shape[axis] = nx
    g = g * np.exp(-1j * freq * x[0]).reshape(shape)
    return g, freq


def ifft(f, x=None, axis=-1):
    f = np.asarray(f)
    nx = f.shape[axis]
    if x is None:
        x = np.arange(nx)
    x = np.asarray(x)
    dx = x[1] - x[0]
    g = np.fft.ifftshift(np.fft.ifft(f, axis=axis))
    g = g * dx / 2.0 / np.pi * len(x)
    freq = 2.0 * np.pi * np.fft.ifftshift(np.fft.fftfreq(nx, d=dx))
    return g * np.exp(1j * freq * x[0]), freq


def fft2(f, dx=1.0, dy=1.0):
    nx, ny = f.shape
    g = np.fft.fftshift(np.fft.fft2(f)) * dx * dy
    fx = 2.0 * np.pi * np.fft.fftshift(np.fft.fftfreq(nx, d=dx))
    fy = 2.0 * np.pi * np.fft.fftshift(np.fft.fftfreq(nx, d=dy))
    return fx, fy, g
