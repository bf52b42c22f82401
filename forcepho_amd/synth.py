"""Seeded synthetic patches for the benchmark and test configurations.

Conventions (demo/demo_utils.py:43-291 of the reference is the recipe):
sky zero point (53, -27) deg; pixel scale 0.03"/pix; per exposure
D = R(theta)/scale pixels per arcsec with theta ~ U[-5, 5] deg, CD = D^-1/3600;
sources uniform in the inner 80% of the patch; flux ~ lognormal(ln 10, 1)
per band; q ~ U[0.3, 0.95]; pa ~ U[-pi/2, pi/2]; nsersic ~ U[1, 5];
rh ~ U[0.05, 0.3]"; PSF mixtures replicated over the 9 Sersic radius bins
(n_psf_per_source = 9 n_psf); noise per band from the SNR rule of
demo_utils.compute_noise_level: sigma_s = (flux_s / 2) / SNR / sqrt(pi
(rh_s/scale)^2), sigma = median over sources.

Random streams: ``numpy.random.default_rng(SeedSequence([config, patch,
purpose]))`` with purpose 0 = patch geometry and truth, 1 = proposals,
2 = noise.

Pixel values are filled by :func:`fill_data` from a model rendered by any
engine (``render(patch, proposal) -> model`` over the packed pixels).
"""

import numpy as np

from .limits import psf_struct_dtype
from .patches import AffineWCS, HIPPatch
from .sources import Galaxy, Scene
from .spline import get_splines

SCALE = 0.03           # arcsec per pixel
SKY_ZERO = (53.0, -27.0)
SNR = 50.0
N_RADII = 9

__all__ = ["PatchSpec", "config_spec", "make_patch", "fill_data", "make_proposals",
           "truth_theta", "theta_bounds", "rng_for", "config4_specs"]


def rng_for(config, patch, purpose):
    return np.random.default_rng(np.random.SeedSequence([int(config), int(patch),
                                                         int(purpose)]))


class PatchSpec:
    """Geometry of one synthetic patch."""

    def __init__(self, n_sources, n_bands, n_exp_per_band, side, n_psf,
                 superpixel=(8, 8), warp_size=64, mask_frac=0.0, psf_per_exposure=False,
                 psf_elliptical=False, psf_sigma=(0.5, 3.0), dither=0.0, pix_offset=0.0,
                 psf_offset=0.0):
        self.n_sources = int(n_sources)
        self.n_bands = int(n_bands)
        self.n_exp_per_band = np.broadcast_to(np.asarray(n_exp_per_band, dtype=int),
                                              (self.n_bands,)).copy()
        self.side = int(side)
        self.n_psf = np.broadcast_to(np.asarray(n_psf, dtype=int), (self.n_bands,)).copy()
        self.superpixel = tuple(superpixel)
        self.warp_size = int(warp_size)
        self.mask_frac = float(mask_frac)
        self.psf_per_exposure = bool(psf_per_exposure)
        self.psf_elliptical = bool(psf_elliptical)
        self.psf_sigma = tuple(psf_sigma)
        self.psf_offset = float(psf_offset)   # PSF component centres U[-o, o] pixels
        self.dither = float(dither)
        self.pix_offset = float(pix_offset)   # added to every pixel coordinate
        self.crowded = False

    @property
    def n_exp(self):
        return int(self.n_exp_per_band.sum())

    @property
    def npix_unpadded(self):
        return self.n_exp * self.side * self.side

    def n_pairs(self):
        """Dense (pixel, image Gaussian) pairs over unpadded pixels."""
        return int(np.sum(self.n_exp_per_band * self.side ** 2 * self.n_sources *
                          self.n_psf * N_RADII))


def config_spec(config):
    """PatchSpec of configurations 1 (plain CPU-runnable) and 2 (drop-in)."""
    if config == 1:
        return PatchSpec(n_sources=2, n_bands=1, n_exp_per_band=1, side=64, n_psf=3,
                         superpixel=(8, 8), warp_size=64)
    if config == 2:
        # 32x32-pixel superpixels, 5% masked pixels, per-exposure mildly
        # elliptical 4-Gaussian PSFs, dithered exposures
        return PatchSpec(n_sources=20, n_bands=8, n_exp_per_band=4, side=128, n_psf=4,
                         superpixel=(32, 32), warp_size=32, mask_frac=0.05,
                         psf_per_exposure=True, psf_elliptical=True, psf_sigma=(0.5, 4.0),
                         dither=2.0)
    raise ValueError(f"no fixed spec for config {config}")


def config4_specs(n_patches=512, config=4, seed_patch0=0):
    """Uneven patches: sources 1 + Poisson(8) capped at 30, bands U{1..10},
    exposures per band U{1..16}, square tiles of side U{32..256} padded to
    64, 3-6 PSF Gaussians per band."""
    specs = []
    for i in range(n_patches):
        rng = rng_for(config, seed_patch0 + i, 3)
        ns = int(min(1 + rng.poisson(8), 30))
        nb = int(rng.integers(1, 11))
        ne = rng.integers(1, 17, size=nb)
        side = int(rng.integers(32, 257))
        npsf = rng.integers(3, 7, size=nb)
        specs.append(PatchSpec(n_sources=ns, n_bands=nb, n_exp_per_band=ne, side=side,
                               n_psf=npsf, superpixel=(8, 8), warp_size=64))
    return specs


def config5_specs(n_patches=4096, crowded_frac=0.05):
    """Mosaic stress: patches from the configuration-4 distribution, and a
    crowded tail (a seeded choice of ``crowded_frac`` of them) of 60
    sources, 12 bands x 2 exposures of 256 x 256 pixels, 3-6 PSF Gaussians
    per band."""
    specs = config4_specs(n_patches, config=5)
    rng = rng_for(5, 0, 4)
    n_crowd = int(round(crowded_frac * n_patches))
    for i in sorted(rng.choice(n_patches, n_crowd, replace=False).tolist()):
        r = rng_for(5, i, 3)
        specs[i] = PatchSpec(n_sources=60, n_bands=12, n_exp_per_band=2, side=256,
                             n_psf=r.integers(3, 7, size=12), superpixel=(8, 8),
                             warp_size=64)
        specs[i].crowded = True
    return specs


def _psf_rows(rng, n_psf, sigma_range, elliptical, offset=0.0):
    amps = rng.dirichlet(np.full(n_psf, 2.0))
    sig = rng.uniform(sigma_range[0], sigma_range[1], size=n_psf)
    rows = np.zeros((N_RADII, n_psf), dtype=psf_struct_dtype)
    for g in range(n_psf):
        if elliptical:
            ar = rng.uniform(0.8, 1.0)
            ang = rng.uniform(0.0, np.pi)
            c, s = np.cos(ang), np.sin(ang)
            R = np.array([[c, -s], [s, c]])
            C = R @ np.diag([sig[g] ** 2, (ar * sig[g]) ** 2]) @ R.T
        else:
            C = np.diag([sig[g] ** 2, sig[g] ** 2])
        rows["amp"][:, g] = amps[g]
        rows["Cxx"][:, g] = C[0, 0]
        rows["Cyy"][:, g] = C[1, 1]
        rows["Cxy"][:, g] = C[0, 1]
    rows["sersic_bin"] = np.arange(N_RADII)[:, None]
    if offset:
        # a component's centre offset is the same in every radius bin
        rows["xcen"] = rng.uniform(-offset, offset, size=n_psf)[None, :]
        rows["ycen"] = rng.uniform(-offset, offset, size=n_psf)[None, :]
    return rows.reshape(-1)


def make_patch(spec, config=1, patch_id=0, return_residual=False, render=None):
    """Build a HIPPatch with packed pixels (ierr set, data zero) and packed
    metadata for the true scene.  Returns (patch, scene, truth_theta, sigma).
    When ``render`` is given, data = model(truth) + noise is filled in."""
    rng = rng_for(config, patch_id, 0)
    ns, nb, side = spec.n_sources, spec.n_bands, spec.side
    bandlist = [f"B{b}" for b in range(nb)]
    ra0, dec0 = SKY_ZERO
    cosd = np.cos(np.deg2rad(dec0))
    half = 0.4 * side * SCALE  # inner 80%, arcsec

    cat = np.zeros(ns, dtype=[("ra", "f8"), ("dec", "f8"), ("q", "f8"), ("pa", "f8"),
                              ("sersic", "f8"), ("rhalf", "f8")] +
                   [(b, "f8") for b in bandlist])
    dx = rng.uniform(-half, half, size=ns)
    dy = rng.uniform(-half, half, size=ns)
    cat["ra"] = ra0 + dx / 3600.0 / cosd
    cat["dec"] = dec0 + dy / 3600.0
    cat["q"] = rng.uniform(0.3, 0.95, size=ns)
    cat["pa"] = rng.uniform(-np.pi / 2, np.pi / 2, size=ns)
    cat["sersic"] = rng.uniform(1.0, 5.0, size=ns)
    cat["rhalf"] = rng.uniform(0.05, 0.3, size=ns)
    for b in bandlist:
        cat[b] = rng.lognormal(np.log(10.0), 1.0, size=ns)

    # exposures: per exposure rotation, dither, PSF
    bands, wcses, psfs, images, ierrs = [], [], [], [], []
    band_psf = {}
    sigmas = np.zeros(nb)
    for b in range(nb):
        fl = np.array([cat[bandlist[b]][s] for s in range(ns)])
        npx = np.pi * (cat["rhalf"] / SCALE) ** 2
        sigmas[b] = np.median((fl / 2.0) / SNR / np.sqrt(npx))
        for e in range(spec.n_exp_per_band[b]):
            theta = np.deg2rad(rng.uniform(-5.0, 5.0))
            c, s = np.cos(theta), np.sin(theta)
            Rm = np.array([[c, -s], [s, c]])
            D = Rm / SCALE
            cd = np.linalg.inv(D) / 3600.0
            dith = rng.uniform(-spec.dither, spec.dither, size=2) if spec.dither else np.zeros(2)
            crpix = np.array([(side - 1) / 2.0, (side - 1) / 2.0]) + dith + spec.pix_offset
            wcses.append(AffineWCS(crpix, (ra0, dec0), cd))
            if spec.psf_per_exposure or b not in band_psf:
                band_psf[b] = _psf_rows(rng, int(spec.n_psf[b]), spec.psf_sigma,
                                        spec.psf_elliptical, spec.psf_offset)
            psfs.append(band_psf[b])
            bands.append(b)
            ie = np.float32(1.0 / sigmas[b])
            if spec.mask_frac > 0:
                ie = np.full((side, side), ie, dtype=np.float32)
                ie[rng.random((side, side)) < spec.mask_frac] = 0.0
            ierrs.append(ie)
            images.append(np.float32(0.0))

    patch = HIPPatch(return_residual=return_residual)
    patch.bandlist = bandlist
    patch.bands = [bandlist[b] for b in bands]
    patch.wcses = wcses
    patch.psfs = psfs
    patch.pack_pix(images, ierrs, bands, superpixel=spec.superpixel,
                   warp_size=spec.warp_size, shapes=[(side, side)] * len(images))
    if spec.pix_offset:
        patch.xpix = (patch.xpix + np.float32(spec.pix_offset)).astype(np.float32)
        patch.ypix = (patch.ypix + np.float32(spec.pix_offset)).astype(np.float32)
    scene = Scene(catalog=cat, filternames=bandlist)
    patch.truth_catalog = cat
    patch.pack_meta(scene)
    theta = scene.get_all_source_params().copy()
    patch.sigma_band = sigmas
    if render is not None:
        fill_data(patch, render(patch, scene.get_proposal()), config, patch_id)
    return patch, scene, theta


def fill_data(patch, model, config, patch_id):
    """data = model + N(0, sigma_band) on real pixels (ierr > 0 or masked
    real pixels); padding pixels keep data = 0."""
    rng = rng_for(config, patch_id, 2)
    data = np.zeros(patch.npix, dtype=np.float64)
    for e in range(len(patch.exposure_N)):
        b = int(np.searchsorted(np.cumsum(patch.band_N), e, side="right"))
        s0, n = patch.exposure_start[e], patch.original_sizes[e]
        noise = rng.normal(0.0, patch.sigma_band[b], size=n)
        data[s0:s0 + n] = model[s0:s0 + n] + noise
    patch.data = data.astype(np.float32)


def theta_bounds(theta, n_bands, side=None):
    """Box bounds for BoundedTransform around a truth vector: fluxes in
    [0, 3 flux + 1], positions +-3 pix, q [0.2, 1], pa [-2, 2],
    nsersic [0.8, 6], rh [0.02, 0.6]."""
    theta = np.asarray(theta, dtype=np.float64).reshape(-1, n_bands + 6)
    lo = np.zeros_like(theta)
    hi = np.zeros_like(theta)
    lo[:, :n_bands] = 0.0
    hi[:, :n_bands] = 3.0 * theta[:, :n_bands] + 1.0
    dpos = 3 * SCALE / 3600.0
    lo[:, n_bands] = theta[:, n_bands] - dpos / np.cos(np.deg2rad(SKY_ZERO[1]))
    hi[:, n_bands] = theta[:, n_bands] + dpos / np.cos(np.deg2rad(SKY_ZERO[1]))
    lo[:, n_bands + 1] = theta[:, n_bands + 1] - dpos
    hi[:, n_bands + 1] = theta[:, n_bands + 1] + dpos
    lo[:, n_bands + 2], hi[:, n_bands + 2] = 0.2, 1.0
    lo[:, n_bands + 3], hi[:, n_bands + 3] = -2.0, 2.0
    lo[:, n_bands + 4], hi[:, n_bands + 4] = 0.8, 6.0
    lo[:, n_bands + 5], hi[:, n_bands + 5] = 0.02, 0.6
    return lo.reshape(-1), hi.reshape(-1)


def make_proposals(theta, n_bands, n_proposals, config=2, patch_id=0):
    """Proposals perturbed from truth: flux * exp(N(0, 0.05)), ra/dec +
    N(0, 0.3 pix), q + N(0, 0.02), pa + N(0, 0.05), nsersic + N(0, 0.1),
    rh + N(0, 0.005), clipped to the bounds of :func:`theta_bounds`."""
    rng = rng_for(config, patch_id, 1)
    theta = np.asarray(theta, dtype=np.float64).reshape(-1, n_bands + 6)
    lo, hi = theta_bounds(theta.reshape(-1), n_bands)
    out = np.empty((n_proposals, theta.size))
    dpos = 0.3 * SCALE / 3600.0
    for k in range(n_proposals):
        t = theta.copy()
        t[:, :n_bands] *= np.exp(rng.normal(0.0, 0.05, size=t[:, :n_bands].shape))
        t[:, n_bands] += rng.normal(0.0, dpos, size=len(t)) / np.cos(np.deg2rad(SKY_ZERO[1]))
        t[:, n_bands + 1] += rng.normal(0.0, dpos, size=len(t))
        t[:, n_bands + 2] += rng.normal(0.0, 0.02, size=len(t))
        t[:, n_bands + 3] += rng.normal(0.0, 0.05, size=len(t))
        t[:, n_bands + 4] += rng.normal(0.0, 0.1, size=len(t))
        t[:, n_bands + 5] += rng.normal(0.0, 0.005, size=len(t))
        out[k] = np.clip(t.reshape(-1), lo + 1e-12, hi - 1e-12)
    return out


def truth_theta(scene):
    return scene.get_all_source_params().copy()
