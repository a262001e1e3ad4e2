"""A minimal stand-in for the parts of an astropy WCS object that the
Projectionist reads (wcs.wcs.crval / cdelt / crpix / ctype / phi0 / theta0 /
lonpole / lng / lat, and .celestial), so that geometries can be built
without astropy or pixell.  Any object with the same attributes (a real
astropy.wcs.WCS) works with the Projectionist as well.
"""
import math

import numpy as np

__all__ = ['WCS', 'make_wcs', 'car_geometry', 'zenithal_geometry', 'bounding_geometry']

_CYLINDRICAL = ('CAR', 'CEA')
_ZENITHAL = ('TAN', 'ZEA', 'ARC', 'SIN')


class _Params:
    def __init__(self):
        self.crval = np.zeros(2)
        self.cdelt = np.ones(2)
        self.crpix = np.zeros(2)
        self.ctype = ['RA---CAR', 'DEC--CAR']
        self.phi0 = 0.
        self.theta0 = 0.
        self.lonpole = 0.
        self.lng = 0
        self.lat = 1


class WCS:
    """Celestial 2-axis WCS (lon axis first, as in FITS)."""

    def __init__(self):
        self.wcs = _Params()

    @property
    def celestial(self):
        return self

    def __repr__(self):
        w = self.wcs
        return (f'WCS(ctype={w.ctype}, crval={list(w.crval)}, cdelt={list(w.cdelt)}, '
                f'crpix={list(w.crpix)}, lonpole={w.lonpole})')


def make_wcs(proj, crval, cdelt, crpix, lonpole=None):
    """Build a WCS for projection proj ('CAR', 'CEA', 'TAN', 'ZEA', 'ARC',
    'SIN') with (lon, lat) crval and cdelt in degrees and 1-based crpix.
    Cylindrical projections get phi0 = theta0 = 0 and LONPOLE 0 (crval lat
    <= 0) or 180; zenithal ones theta0 = 90 and LONPOLE 180, matching the
    two branches of Projectionist.get_q (python/proj/wcs.py:526-548)."""
    proj = proj.upper()
    w = WCS()
    p = w.wcs
    p.ctype = ['RA---' + proj, 'DEC--' + proj]
    p.crval = np.array(crval, dtype=float)
    p.cdelt = np.array(cdelt, dtype=float)
    p.crpix = np.array(crpix, dtype=float)
    if proj in _CYLINDRICAL:
        p.phi0, p.theta0 = 0., 0.
        p.lonpole = (0. if p.crval[1] <= 0 else 180.) if lonpole is None else lonpole
    elif proj in _ZENITHAL:
        p.phi0, p.theta0 = 0., 90.
        p.lonpole = 180. if lonpole is None else lonpole
    else:
        raise ValueError(f'unsupported projection {proj}')
    return w


def car_geometry(lon_range, lat_range, res, margin=0, proj='CAR'):
    """(shape, wcs) of a cylindrical map with reference point (0, 0),
    cdelt = (-res, +res) degrees, covering lon_range x lat_range (degrees)
    plus `margin` pixels on every side."""
    lon_lo, lon_hi = lon_range
    lat_lo, lat_hi = lat_range
    ix_hi = math.ceil(lon_hi / res)
    ix_lo = math.floor(lon_lo / res)
    iy_lo = math.floor(lat_lo / res)
    iy_hi = math.ceil(lat_hi / res)
    nx = ix_hi - ix_lo + 2 * margin + 1
    ny = iy_hi - iy_lo + 2 * margin + 1
    # Pixel i (0-based) centre: lon = (i + 1 - crpix_x) * cdelt_x.
    crpix_x = margin + 1 + ix_hi
    crpix_y = margin + 1 - iy_lo
    wcs = make_wcs(proj, (0., 0.), (-res, res), (crpix_x, crpix_y))
    return (ny, nx), wcs


def zenithal_geometry(center, half_width, res, proj='TAN'):
    """(shape, wcs) of a square zenithal map centred on center=(lon, lat)
    degrees, +-half_width degrees, cdelt (-res, res)."""
    n = 2 * int(math.ceil(half_width / res)) + 1
    c = (n + 1) / 2
    wcs = make_wcs(proj, center, (-res, res), (c, c))
    return (n, n), wcs


def bounding_geometry(lon, lat, res, margin=4, proj='CAR'):
    """Cylindrical geometry covering sample positions lon, lat (radians)."""
    d = 180 / np.pi
    return car_geometry((float(np.min(lon)) * d, float(np.max(lon)) * d),
                        (float(np.min(lat)) * d, float(np.max(lat)) * d), res,
                        margin=margin, proj=proj)
