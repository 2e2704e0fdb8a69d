"""Patch packing and device residency (mirrors forcepho/patches/)."""

from .patch import PatchBase, Patch, PSF_COLS, PSF_DTYPE, LinearWCS, scale_at_sky
from .stamp_patch import Stamp, StampPatch, superpixel_order
from .device_patch import DevicePatchMixin, HIPPatchMixin

__all__ = ["PatchBase", "Patch", "PSF_COLS", "PSF_DTYPE", "LinearWCS", "scale_at_sky",
           "Stamp", "StampPatch", "superpixel_order",
           "DevicePatchMixin", "HIPPatchMixin", "GPUPatchMixin", "SimplePatch"]

# forcepho's name for the device mixin (forcepho/patches/__init__.py:12-17)
GPUPatchMixin = HIPPatchMixin


class SimplePatch(StampPatch, HIPPatchMixin):
    """A stamp-built patch resident on the MI355X."""
    pass
