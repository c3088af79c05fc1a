"""Default-device helpers, mirroring tedeous/device.py:7-52.

The reference keeps the compute device as torch's global default device.
This package does the same so that user scripts written against the
reference (``solver_device('gpu')`` then building tensors) behave alike.
"""

from typing import Any

import torch


def solver_device(device: str) -> None:
    """Select the global default device (tedeous/device.py:7-23).

    ``'cuda'`` and ``'gpu'`` select the HIP device when one is visible;
    anything else selects the CPU.
    """
    if device in ('cuda', 'gpu') and torch.cuda.is_available():
        torch.set_default_device('cuda')
    else:
        torch.set_default_device('cpu')


def device_type() -> str:
    """Type string of the current default device (tedeous/device.py:49-52)."""
    return torch.tensor([0.]).device.type


def check_device(data: Any) -> torch.Tensor:
    """Move ``data`` to the default device, converting to a tensor if needed
    (tedeous/device.py:26-46)."""
    target = torch.tensor([0.]).device
    if isinstance(data, torch.Tensor):
        return data if data.device == target else data.to(target)
    try:
        return torch.tensor(data).to(target)
    except Exception as exc:  # same error class the reference raises
        raise TypeError(f"Cannot convert data to tensor: {exc}") from exc
