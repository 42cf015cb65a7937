"""Choosing the device (sim/device.py:12-30 of the reference).

The reference runs its Warp kernels on "cuda:0" or "cpu". This package has one compute path, the HIP kernels on
the ROCm GPU (which torch names "cuda"), so `--device cpu` and a machine without a GPU are errors.
"""

import torch

DEVICE_CHOICES = ("auto", "cuda", "cpu")


class DeviceError(Exception):
    """The requested device is not available."""


def init_device(choice: str) -> str:
    """The torch device name for `choice`: "cuda:0" when a ROCm GPU is present."""
    if choice not in DEVICE_CHOICES:
        raise DeviceError(f"--device must be one of {', '.join(DEVICE_CHOICES)}; got {choice!r}")
    if choice == "cpu":
        raise DeviceError("--device cpu is not available: plume_amd runs its kernels on the GPU only")
    if not torch.cuda.is_available():
        raise DeviceError(f"--device {choice} was asked for, but no ROCm GPU is available")
    import plume_amd._lib  # noqa: F401  (fails loudly when libplume_hip.so is missing)

    return "cuda:0"
