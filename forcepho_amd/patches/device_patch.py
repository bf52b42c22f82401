"""Device mix-ins: move a packed patch to HBM and drive evaluations.

``HIPPatchMixin`` is the drop-in for ``GPUPatchMixin``
(forcepho/patches/device_patch.py:260-438): same method names and meaning
(``send_to_device``, ``replace_device_meta_ptrs``,
``send_patchstruct_to_device``, ``swap_on_device``, ``retrieve_array``,
``free``, ``get_proposer``) and the shared ``DevicePatchMixin`` workflow
(``prepare_model``, ``subtract_fixed``, ``design_matrix``, ``render``,
lines 23-257).  Device memory belongs to a ``fpho_patch`` object of
libfpho_hip.so; ``device_patch`` is its opaque handle.
"""

import ctypes

import numpy as np

from .. import _lib
from ..limits import psf_struct_dtype

__all__ = ["DevicePatchMixin", "HIPPatchMixin"]


def _blocks(catalog, size):
    """Consecutive rows of ``catalog`` in blocks of at most ``size``."""
    if catalog is None or len(catalog) == 0:
        return []
    return [catalog[i:i + size] for i in range(0, len(catalog), size)]


class DevicePatchMixin:
    """The fitting workflow of forcepho/patches/device_patch.py:23-257 on top
    of the device primitives of a mix-in such as :class:`HIPPatchMixin`.

    Every operation is a sequence of *scene loads* (pack a scene's metadata
    and hand it to the device) and *evaluations*:

    * ``subtract_fixed`` -- upload the pixels once, then for each block of
      big or fixed sources: load it, evaluate (the device writes data -
      model into the residual buffer) and swap data and residual, so the
      device data lose that block; finally load the active scene;
    * ``design_matrix`` -- per active source, load the source alone, read
      its model back and scale it to unit flux (the model is linear in
      flux);
    * ``render`` -- load a catalog and return minus the residual.
    """

    def _load_scene(self, scene):
        """Pack the scene's metadata and make it the device's."""
        self.pack_meta(scene)
        self.replace_device_meta_ptrs()
        self.send_patchstruct_to_device()
        return scene

    def prepare_model(self, active=None, fixed=None, bounds=None, maxactive=15,
                      model_kwargs={}):
        """Subtract the fixed sources, then wrap the patch and the active
        scene in a :class:`~forcepho_amd.model.FastPosterior`
        (device_patch.py:28-72).  ``bounds`` = (lower, upper) parameter
        arrays gives a bounded transform, else the identity.  Returns the
        model and the active scene's parameter vector."""
        from ..model import FastPosterior, Transform
        proposer, scene = self.subtract_fixed(fixed, active=active, maxactive=maxactive)
        self.return_residual = False
        q = np.array(scene.get_all_source_params(), dtype=np.float64)
        if bounds is None:
            limits = {"transform": Transform(len(q))}
        else:
            limits = {"lower": bounds[0], "upper": bounds[1]}
        model = FastPosterior(proposer, scene=scene, patch=self, **limits, **model_kwargs)
        return model, q

    def subtract_fixed(self, fixed, active=None, big=None, maxactive=15, swap_local=False):
        """Remove blocks (at most ``maxactive`` sources each) of ``big`` and
        then ``fixed`` sources from the device-side data
        (device_patch.py:113-183).  Afterwards the device data are the
        residual of every subtracted block and the device metadata describe
        ``active`` -- or, without it (``swap_local`` required), the last
        subtracted block.  Returns (proposer, that scene)."""
        removed = [self.set_scene(b) for b in _blocks(big, maxactive) + _blocks(fixed, maxactive)]
        if active is None and not swap_local:
            raise ValueError("without active sources, pass swap_local=True")
        if not removed and active is None:
            raise ValueError("No scenes to prepare!")
        proposer = self.get_proposer()
        self.return_residual = True
        # the pixels go up with the first scene's metadata
        first = removed[0] if removed else self.set_scene(active)
        self.pack_meta(first)
        self.send_to_device()
        for k, scene in enumerate(removed):
            if k > 0:
                self._load_scene(scene)
            proposer.evaluate_proposal(scene.get_proposal(), patch=self)
            self.swap_on_device()           # data := data - model(scene)
        if active is None:
            return proposer, removed[-1]
        if removed:
            return proposer, self._load_scene(self.set_scene(active))
        return proposer, first

    def design_matrix(self, active=None, proposer=None):
        """Design matrices for linear flux fitting (device_patch.py:
        185-238): per band, X [n_active, n_pix_band] whose row i is source
        i's model at unit flux in that band, and the target data y (the
        device data, fixed sources already subtracted), so that
        y_band ~ fluxes_band . X_band.  Returns (Xes, ys).

        The model comes back as y - residual, whose float32 rounding is
        absolute (a few ulp of the data), so each source is rendered at its
        own fluxes and its rows are divided by them; a zero flux is rendered
        at 1 instead."""
        y = self.retrieve_array("data")
        cuts = np.cumsum(self.band_N_pix)[:-1]
        ys = np.split(y, cuts)
        proposer = proposer or self.get_proposer()
        self.return_residual = True
        active = np.atleast_1d(active)
        Xes = [np.zeros((len(active), n)) for n in self.band_N_pix]
        for i in range(len(active)):
            row = active[i:i + 1].copy()       # a row of its own, not a view
            norm = np.array([float(row[b][0]) or 1.0 for b in self.bandlist])
            for b, f in zip(self.bandlist, norm):
                row[b] = f
            scene = self._load_scene(self.set_scene(row))
            residual = proposer.evaluate_proposal(scene.get_proposal(), patch=self,
                                                  unpack=False)[2]
            for j, model_b in enumerate(np.split(y - residual, cuts)):
                Xes[j][i] = model_b / norm[j]
        return Xes, ys

    def render(self, catalog):
        """Minus the residual of ``catalog``'s scene against the device data
        (device_patch.py:240-254): the model images when the device data
        are zero.  One array per exposure."""
        proposer = self.get_proposer()
        scene = self._load_scene(self.set_scene(catalog))
        self.return_residual = True
        residuals = proposer.evaluate_proposal(scene.get_proposal(), patch=self,
                                               unpack=True)[2]
        return [-r for r in residuals]


class HIPPatchMixin(DevicePatchMixin):
    """Patch data in MI355X HBM through libfpho_hip.so."""

    device = 0

    def get_proposer(self):
        from ..proposal import HIPProposer
        return HIPProposer()

    # -- descriptors -------------------------------------------------------
    def _desc(self, meta_only=False):
        """fpho_patch_desc over this patch's host arrays (kept alive in
        ``self._desc_keep`` until the call returns).  With ``meta_only`` the
        pixel arrays are left out (the library ignores them then)."""
        keep = {}

        def arr(name, dtype, ctype):
            a = np.ascontiguousarray(getattr(self, name), dtype=dtype).reshape(-1)
            keep[name] = a
            return _lib.fptr(a, ctype)

        d = _lib.PatchDesc()
        d.npix = int(self.npix)
        if not meta_only:
            d.data = arr("data", np.float32, ctypes.c_float)
            d.ierr = arr("ierr", np.float32, ctypes.c_float)
            d.xpix = arr("xpix", np.float32, ctypes.c_float)
            d.ypix = arr("ypix", np.float32, ctypes.c_float)
        d.n_bands = int(len(self.band_N))
        d.n_exp = int(len(self.exposure_N))
        d.band_start = arr("band_start", np.int16, ctypes.c_int16)
        d.band_N = arr("band_N", np.int16, ctypes.c_int16)
        d.exposure_start = arr("exposure_start", np.int32, ctypes.c_int32)
        d.exposure_N = arr("exposure_N", np.int32, ctypes.c_int32)
        d.n_sources = int(self.n_sources)
        d.n_radii = int(self.n_radii)
        d.rad2 = arr("rad2", np.float32, ctypes.c_float)
        d.D = arr("D", np.float32, ctypes.c_float)
        d.crpix = arr("crpix", np.float32, ctypes.c_float)
        d.crval = arr("crval", np.float32, ctypes.c_float)
        d.CW = arr("CW", np.float32, ctypes.c_float)
        d.G = arr("G", np.float32, ctypes.c_float)
        d.n_psf_per_source = arr("n_psf_per_source", np.int32, ctypes.c_int32)
        psf = np.ascontiguousarray(self.psfgauss)
        if psf.dtype != psf_struct_dtype:
            psf = np.ascontiguousarray(psf.view(psf_struct_dtype) if
                                       psf.dtype.itemsize == 28 else psf.astype(psf_struct_dtype))
        keep["psfgauss"] = psf
        d.psfgauss = ctypes.c_void_p(psf.ctypes.data)
        d.n_psfgauss = int(len(psf))
        d.psfgauss_start = arr("psfgauss_start", np.int32, ctypes.c_int32)
        self._desc_keep = keep
        return d

    # -- device primitives -------------------------------------------------
    def send_to_device(self):
        """Copy all patch arrays to HBM (forcepho/patches/device_patch.py:
        268-300); allocates the residual if ``return_residual``."""
        _lib.init(self.device)
        self.free()
        handle = ctypes.c_void_p()
        d = self._desc()
        _lib.check(_lib.lib().fpho_patch_create(ctypes.byref(d), int(bool(self.return_residual)),
                                                ctypes.byref(handle)))
        self._desc_keep = None
        self.device_patch = handle
        self._device_npix = int(self.npix)
        self.device_ptrs = {"residual": bool(self.return_residual)}
        self._has_residual = bool(self.return_residual)
        self._dirty_data = False
        return self.device_patch

    def replace_device_meta_ptrs(self):
        """Upload the currently packed metadata (device_patch.py:302-311)."""
        d = self._desc(meta_only=True)
        _lib.check(_lib.lib().fpho_patch_set_meta(self.device_patch, ctypes.byref(d)))
        self._desc_keep = None

    def send_patchstruct_to_device(self):
        """Nothing to do: the library rebuilds its descriptors lazily when the
        metadata or the data/residual assignment changes."""
        return self.device_patch

    def swap_on_device(self):
        """Upload new metadata and make the residual the new data
        (device_patch.py:346-367)."""
        assert getattr(self, "_has_residual", False), \
            "Must instantiate the Patch with `return_residual=True`"
        self.replace_device_meta_ptrs()
        _lib.check(_lib.lib().fpho_patch_swap_data_residual(self.device_patch))
        self._dirty_data = True
        return self.device_patch

    def retrieve_array(self, arrname="residual"):
        """Copy one device pixel array to host (device_patch.py:369-375)."""
        out = np.empty(self.npix, dtype=np.float32)
        _lib.check(_lib.lib().fpho_patch_read(self.device_patch, _lib.ARRAY_IDS[arrname],
                                              _lib.fptr(out, ctypes.c_float)))
        return out

    def set_launch(self, pix_per_lane=4, waves_per_group=4, persistent=0, groups_per_cu=8,
                   cull=1):
        l = _lib.Launch(pix_per_lane, waves_per_group, persistent, groups_per_cu, cull)
        _lib.check(_lib.lib().fpho_patch_set_launch(self.device_patch, ctypes.byref(l)))

    def set_precision(self, mode):
        """"fp32", "fp64_reduce", "mixed" (default) or "mixed_centre" (the
        mixed mode with every pixel offset taken from the FP64 centre)."""
        _lib.check(_lib.lib().fpho_patch_set_precision(self.device_patch,
                                                       _lib.PRECISION[mode]))

    def set_residual_writes(self, on):
        """Start or stop writing the residual at every evaluation (and
        returning it from ``evaluate_proposal``); the residual buffer is
        kept, so ``swap_on_device`` still works after stopping."""
        _lib.check(_lib.lib().fpho_patch_set_residual_writes(self.device_patch,
                                                             int(bool(on))))
        self.return_residual = bool(on)
        if on:
            self._has_residual = True

    def device_status(self, clear=True):
        """FPHO_STATUS_* bits recorded by this patch's stream-ordered
        evaluations since the last cleared read (waits for the device)."""
        w = ctypes.c_uint32()
        _lib.check(_lib.lib().fpho_patch_status(self.device_patch, ctypes.byref(w),
                                                int(bool(clear))))
        return w.value

    def free(self):
        handle = getattr(self, "device_patch", None)
        if handle:
            _lib.lib().fpho_patch_free(handle)
        self.device_patch = None
        self.device_ptrs = {}
