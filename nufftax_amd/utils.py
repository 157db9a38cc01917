"""Centered FFT and complex <-> 2-channel helpers (nufftax/utils/fft.py:10-27), on torch."""

import torch


def to_complex(x: torch.Tensor) -> torch.Tensor:
    """(..., 2) real/imag -> complex (...)."""
    return torch.complex(x[..., 0], x[..., 1])


def from_complex(x: torch.Tensor) -> torch.Tensor:
    """complex (...) -> (..., 2) real/imag."""
    return torch.stack([x.real, x.imag], dim=-1)


def fftc(x: torch.Tensor, axes=(-2, -1)) -> torch.Tensor:
    """Centered FFT: image space -> k-space."""
    return torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(x, dim=axes), dim=axes), dim=axes)


def ifftc(x: torch.Tensor, axes=(-2, -1)) -> torch.Tensor:
    """Centered inverse FFT: k-space -> image space."""
    return torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(x, dim=axes), dim=axes), dim=axes)
