"""Scene and Galaxy: flat parameter vector <-> proposal structs.

Mirrors the parts of forcepho/sources.py on the likelihood path:
``Scene.set_all_source_params`` / ``get_all_source_params`` /
``get_proposal`` (forcepho/sources.py:46-97) and ``Galaxy.set_params`` /
``get_param_vector`` / ``proposal`` / ``amplitudes`` /
``damplitude_dsersic`` / ``damplitude_drh`` (forcepho/sources.py:559-694).
The amplitude splines are this package's own FP64 implementation
(:mod:`forcepho_amd.spline`).
"""

import numpy as np

from .limits import source_struct_dtype
from .spline import DEFAULT_SPLINEDATA, get_splines

__all__ = ["Scene", "Galaxy"]


class Galaxy:
    """A Sersic source: fluxes[n_bands], ra, dec (degrees, offsets from the
    patch reference), q, pa (radians), sersic, rhalf (arcsec)."""

    SHAPE_COLS = ["ra", "dec", "q", "pa", "sersic", "rhalf"]
    npos = 2
    nshape = 4

    def __init__(self, filters=("band",), filternames=None, splinedata=None,
                 spline_smoothing=None, spline_type="rect", **kwargs):
        self.filternames = list(filternames if filternames is not None else filters)
        self.flux = np.zeros(len(self.filternames))
        self.ra = self.dec = 0.0
        self.q, self.pa = 1.0, 0.0
        self.sersic, self.rhalf = 2.0, 0.1
        self.fixed = False
        self.id = 0
        self.splines = get_splines(splinedata or DEFAULT_SPLINEDATA,
                                   spline_smoothing=spline_smoothing, spline_type=spline_type)
        self.radii = self.splines.radii
        self.proposal_struct = np.zeros(1, dtype=source_struct_dtype)

    @property
    def n_bands(self):
        return len(self.filternames)

    @property
    def n_params(self):
        return self.n_bands + self.npos + self.nshape

    @property
    def n_gauss(self):
        return len(self.radii)

    @property
    def sersic_range(self):
        return self.splines.sersic_range

    def filter_index(self, filtername):
        return self.filternames.index(filtername)

    def set_params(self, parvector, filtername=None):
        """[flux_1..flux_N, ra, dec, q, pa, sersic, rhalf]
        (forcepho/sources.py:559-592)."""
        if filtername is not None:
            nflux, flux_inds = 1, self.filter_index(filtername)
        else:
            nflux, flux_inds = self.n_bands, slice(None)
        assert len(parvector) == nflux + self.npos + self.nshape, \
            "The length of the parameter vector is not appropriate for this source"
        self.flux[flux_inds] = parvector[:nflux]
        self.ra = parvector[nflux]
        self.dec = parvector[nflux + 1]
        self.q = parvector[nflux + 2]
        self.pa = parvector[nflux + 3]
        self.sersic = parvector[nflux + 4]
        self.rhalf = parvector[nflux + 5]

    def get_param_vector(self, filtername=None):
        flux = [self.flux[self.filter_index(filtername)]] if filtername else self.flux
        return np.concatenate([flux, [self.ra, self.dec, self.q, self.pa,
                                      self.sersic, self.rhalf]])

    @property
    def amplitudes(self):
        return self.splines.amplitudes(self.sersic, self.rhalf)

    @property
    def damplitude_dsersic(self):
        return self.splines.damplitude_dsersic(self.sersic, self.rhalf)

    @property
    def damplitude_drh(self):
        return self.splines.damplitude_drh(self.sersic, self.rhalf)

    def proposal(self):
        """One ``source_struct_dtype`` record (forcepho/sources.py:607-621)."""
        p = self.proposal_struct
        p[...] = 0
        p["fluxes"][0, :self.n_bands] = self.flux
        p["ra"] = self.ra
        p["dec"] = self.dec
        p["q"] = self.q
        p["pa"] = self.pa
        p["nsersic"] = self.sersic
        p["rh"] = self.rhalf
        amp, dn, dr = self.splines.evaluate(self.sersic, self.rhalf)
        p["mixture_amplitudes"][0, :self.n_gauss] = amp
        p["damplitude_drh"][0, :self.n_gauss] = dr
        p["damplitude_dnsersic"][0, :self.n_gauss] = dn
        return p


class Scene:
    """An ordered collection of sources and the flat parameter vector."""

    def __init__(self, sources=(), catalog=None, filternames=None, splinedata=None,
                 spline_smoothing=None, spline_type="rect", **kwargs):
        self.sources = list(sources)
        if catalog is not None:
            self.from_catalog(catalog, filternames=filternames, splinedata=splinedata,
                              spline_smoothing=spline_smoothing, spline_type=spline_type)
        for i, s in enumerate(self.sources):
            s.id = i

    def from_catalog(self, catalog, filternames, splinedata=None, spline_smoothing=None,
                     spline_type="rect"):
        """Sources from a structured array with SHAPE_COLS and one flux column
        per filter name."""
        self.sources = []
        for row in catalog:
            g = Galaxy(filternames=filternames, splinedata=splinedata,
                       spline_smoothing=spline_smoothing, spline_type=spline_type)
            g.flux[:] = [row[f] for f in filternames]
            g.ra, g.dec = float(row["ra"]), float(row["dec"])
            g.q, g.pa = float(row["q"]), float(row["pa"])
            g.sersic, g.rhalf = float(row["sersic"]), float(row["rhalf"])
            self.sources.append(g)

    def __len__(self):
        return len(self.sources)

    @property
    def active_sources(self):
        return [s for s in self.sources if not s.fixed]

    @property
    def n_active(self):
        return len(self.active_sources)

    def set_all_source_params(self, Theta):
        start = 0
        for source in self.active_sources:
            end = start + source.n_params
            source.set_params(Theta[start:end])
            start = end

    def get_all_source_params(self, active=True):
        plist = [s.get_param_vector() for s in self.sources if (not s.fixed) or not active]
        return np.concatenate(plist)

    def get_proposal(self, active=True):
        plist = [s.proposal() for s in self.sources if (not s.fixed) == active]
        return np.concatenate(plist)

    @property
    def parameter_names(self):
        names = []
        for s in self.sources:
            names += [f"{f}_{s.id}" for f in s.filternames]
            names += [f"{c}_{s.id}" for c in Galaxy.SHAPE_COLS]
        return names
