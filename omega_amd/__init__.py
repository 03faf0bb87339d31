"""omega_amd: MI355X (gfx950) matrix-free projectors and image updates with OMEGA's projector API.

The compute path is libomega_hip.so (hand-written HIP, C ABI in include/omega_hip.h), bound with
ctypes in :mod:`omega_amd._lib`. PyTorch is used only as the device tensor container.
"""
from .projector import projectorClass  # noqa: F401

__all__ = ["projectorClass"]
