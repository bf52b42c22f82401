"""Patch packers and device mix-ins (forcepho/patches/__init__.py composes
``JadesPatch(StorePatch, GPUPatchMixin)`` the same way)."""

from .patch import (AffineWCS, Patch, PatchBase, pack_exposures, scale_at_sky,
                    superpixel_order)
from .device_patch import DevicePatchMixin, HIPPatchMixin


class HIPPatch(Patch, HIPPatchMixin):
    """A Patch whose arrays live in MI355X HBM."""
    pass


__all__ = ["AffineWCS", "Patch", "PatchBase", "pack_exposures", "scale_at_sky",
           "superpixel_order", "DevicePatchMixin", "HIPPatchMixin", "HIPPatch"]
