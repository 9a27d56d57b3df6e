"""Centre-preserving pad / crop / resize and block downsampling (utils/array_ops.py:12-162).

These move memory (views and copies of device tensors); any leading batch
dimensions are carried through.
"""

from __future__ import annotations

import torch

from .. import runtime

__all__ = ["pad_to", "crop_to", "resize", "downsample"]


def _parity(n):
    return "even" if n % 2 == 0 else "odd"


def pad_to(array: torch.Tensor, npixels: int, fill: float = 0.0) -> torch.Tensor:
    """utils/array_ops.py:12-50."""
    n_in = array.shape[-1]
    if n_in % 2 != npixels % 2:
        raise ValueError(
            "Center-preserving padding requires parity consistency, "
            "i.e. even -> even or odd -> odd: "
            f"input array ({n_in} pixels, {_parity(n_in)}) cannot be padded "
            f"to {npixels} pixels ({_parity(npixels)} parity). "
            f"Try npixels={npixels + 1} instead."
        )
    if npixels < n_in:
        raise ValueError(
            f"Cannot pad to smaller size: input has {n_in} pixels, "
            f"target is {npixels} pixels. Padding requires npixels >= {n_in}."
        )
    p = (npixels - n_in) // 2
    return torch.nn.functional.pad(array, (p, p, p, p), value=fill)


def crop_to(array: torch.Tensor, npixels: int) -> torch.Tensor:
    """utils/array_ops.py:53-90."""
    n_in = array.shape[-1]
    if n_in % 2 != npixels % 2:
        raise ValueError(
            "Center-preserving cropping requires parity consistency, "
            "i.e. even -> even or odd -> odd: "
            f"input array ({n_in} pixels, {_parity(n_in)}) cannot be cropped "
            f"to {npixels} pixels ({_parity(npixels)} parity). "
            f"Try npixels={npixels - 1} instead."
        )
    if npixels > n_in:
        raise ValueError(
            f"Cannot crop to larger size: input has {n_in} pixels, "
            f"target is {npixels} pixels. Cropping requires npixels <= {n_in}."
        )
    a, b = (n_in - npixels) // 2, (n_in + npixels) // 2
    return array[..., a:b, a:b]


def resize(array: torch.Tensor, npixels: int, fill: float = 0.0) -> torch.Tensor:
    """utils/array_ops.py:93-121."""
    n_in = array.shape[-1]
    if npixels == n_in:
        return array
    if npixels < n_in:
        return crop_to(array, npixels)
    return pad_to(array, npixels, fill)


def downsample(array: torch.Tensor, n: int, mean: bool = True) -> torch.Tensor:
    """Block mean/sum by n x n.  utils/array_ops.py:124-162."""
    if array.shape[-1] != array.shape[-2]:
        raise ValueError(f"Input array has shape {tuple(array.shape)}, which is not square")
    if array.shape[-1] % n != 0:
        raise ValueError(
            f"Input array has {array.shape[-1]} pixels, which is not divisible by {n}"
        )
    if torch.is_tensor(array) and array.is_cuda and array.dtype == torch.float32:
        # float32 device images: the block-sum pass of the detector kernels
        out = runtime.detector(array, [("downsample", int(n))])
        return out / float(n * n) if mean else out
    m = array.shape[-1] // n
    blocks = array.reshape(*array.shape[:-2], m, n, m, n)
    return blocks.mean((-3, -1)) if mean else blocks.sum((-3, -1))
