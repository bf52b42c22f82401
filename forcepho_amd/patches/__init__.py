"""Patch packers and device mix-ins (forcepho/patches/__init__.py composes
``JadesPatch(StorePatch, GPUPatchMixin)`` the same way)."""

from .patch import Patch, PatchBase, pack_exposures, superpixel_order
from .wcs import AffineWCS, TanWCS, WCS, scale_at_sky
from .device_patch import DevicePatchMixin, HIPPatchMixin
from .pixel_patch import DirectPatch, FITSPatch, PixelPatch, StorePatch, JWST_BANDS
from .storage import MetaStore, PixelStore, PSFStore


class HIPPatch(Patch, HIPPatchMixin):
    """A Patch whose arrays live in MI355X HBM."""
    pass


class JadesPatch(StorePatch, HIPPatchMixin):
    """Pixel/meta/PSF stores packed for the MI355X kernels."""
    pass


class SimplePatch(FITSPatch, HIPPatchMixin):
    """FITS exposures packed for the MI355X kernels."""
    pass


__all__ = ["AffineWCS", "TanWCS", "WCS", "Patch", "PatchBase", "pack_exposures",
           "scale_at_sky", "superpixel_order", "DevicePatchMixin", "HIPPatchMixin",
           "HIPPatch", "PixelPatch", "StorePatch", "FITSPatch", "DirectPatch",
           "JadesPatch", "SimplePatch", "PixelStore", "MetaStore", "PSFStore", "JWST_BANDS"]
