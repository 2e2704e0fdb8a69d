"""Patch packing and device residency (mirrors forcepho/patches/)."""

from .patch import PatchBase, Patch, PSF_COLS, PSF_DTYPE, LinearWCS, scale_at_sky
from .stamp_patch import Stamp, StampPatch, superpixel_order
from .device_patch import DevicePatchMixin, HIPPatchMixin
from .storage import PSFStore, psf_records
from .pixel_patch import PixelPatch, FITSPatch, JWST_BANDS, superpixel_sort

__all__ = ["PatchBase", "Patch", "PSF_COLS", "PSF_DTYPE", "LinearWCS", "scale_at_sky",
           "Stamp", "StampPatch", "superpixel_order",
           "DevicePatchMixin", "HIPPatchMixin", "GPUPatchMixin", "SimplePatch",
           "PSFStore", "psf_records", "PixelPatch", "FITSPatch", "JWST_BANDS",
           "superpixel_sort", "FITSDevicePatch"]

# forcepho's name for the device mixin (forcepho/patches/__init__.py:12-17)
GPUPatchMixin = HIPPatchMixin


class SimplePatch(StampPatch, HIPPatchMixin):
    """A stamp-built patch resident on the MI355X."""
    pass


class FITSDevicePatch(FITSPatch, HIPPatchMixin):
    """A patch read from FITS images, resident on the MI355X."""
    pass
