"""lime's Fourier-transform convention (lime/fft.py:9-102) on the GPU (rocFFT via torch.fft).

* ``fft(f, x)``  : g(w) = sum_n f(x_n) e^{-i w x_n} dx, fftshifted, w = 2 pi fftshift(fftfreq(n, dx)),
                   including the phase e^{-i w x0} of a grid that does not start at 0.
* ``ifft(f, x)`` : g = (1/2pi) sum_n f e^{+i w x_n} dx.
* ``fft2(f, dx, dy)``: 2D transform for two-time correlation maps; lime computes the second
  frequency axis with ``fftfreq(nx, d=dy)`` (lime/fft.py:100) -- kept.

Inputs may be NumPy arrays (returned as NumPy) or torch tensors (returned on their device).
"""
import numpy as np
import torch

from .propagator import device


def _to_tensor(f):
    if isinstance(f, torch.Tensor):
        return f.to(torch.complex128), True
    return torch.as_tensor(np.asarray(f, dtype=np.complex128), device=device()), False


def _out(t, was_tensor):
    return t if was_tensor else t.detach().cpu().numpy()


def fft(f, x=None, axis=-1):
    ft, is_t = _to_tensor(f)
    nx = ft.shape[axis]
    x = np.arange(nx) if x is None else np.asarray(x.cpu() if isinstance(x, torch.Tensor) else x)
    dx = float(x[1] - x[0])
    g = torch.fft.fftshift(torch.fft.fft(ft, dim=axis), dim=axis) * dx
    freq = 2.0 * np.pi * np.fft.fftshift(np.fft.fftfreq(nx, d=dx))
    shape = [1] * g.dim()
    shape[axis] = nx
    phase = torch.as_tensor(np.exp(-1j * freq * x[0]), device=g.device).reshape(shape)
    return _out(g * phase, is_t), freq


def ifft(f, x=None, axis=-1):
    ft, is_t = _to_tensor(f)
    nx = ft.shape[axis]
    x = np.arange(nx) if x is None else np.asarray(x)
    dx = float(x[1] - x[0])
    g = torch.fft.ifftshift(torch.fft.ifft(ft, dim=axis), dim=axis) * (dx / 2.0 / np.pi * len(x))
    freq = 2.0 * np.pi * np.fft.ifftshift(np.fft.fftfreq(nx, d=dx))
    shape = [1] * g.dim()
    shape[axis] = nx
    phase = torch.as_tensor(np.exp(1j * freq * x[0]), device=g.device).reshape(shape)
    return _out(g * phase, is_t), freq


def fft2(f, dx=1.0, dy=1.0):
    ft, is_t = _to_tensor(f)
    nx, ny = ft.shape
    g = torch.fft.fftshift(torch.fft.fft2(ft)) * (dx * dy)
    fx = 2.0 * np.pi * np.fft.fftshift(np.fft.fftfreq(nx, d=dx))
    fy = 2.0 * np.pi * np.fft.fftshift(np.fft.fftfreq(nx, d=dy))
    return fx, fy, _out(g, is_t)
