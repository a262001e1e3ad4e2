"""Synthetic observations for tests and benchmarks.

A constant-elevation scan (az triangle wave, python/proj/util.py:4-25) seen
in the horizon frame (CelestialSightLine.for_horizon), a focal plane of
detectors uniform in a disk with random polarisation angles, and a CAR map
footprint covering the scan plus a margin:

  small:  el 50 deg, az +-10 deg at 1 deg/s, 200 Hz, 40000 samples,
          100 detectors in a 1 deg disk, CAR 1 arcmin
  large:  the same scan for 1 h (720000 samples), 1000 detectors in a
          1.5 deg disk, CAR 0.5 arcmin (also tiled 100x100 and HEALPix
          NEST nside 4096)
"""
import numpy as np

from .proj import quat
from .proj.coords import CelestialSightLine, FocalPlane, Assembly
from .proj.util import ces
from .proj.wcslite import car_geometry

DEG = np.pi / 180
ARCMIN = DEG / 60

CONFIGS = {
    'small': dict(n_det=100, n_samp=40000, disk_deg=1.0, res_arcmin=1.0),
    'large': dict(n_det=1000, n_samp=720000, disk_deg=1.5, res_arcmin=0.5),
}


def scan(n_samp, fs=200., el_deg=50., az0_deg=0., throw_deg=10., v_deg=1.):
    """(t, az, el) in seconds and radians."""
    t = np.arange(n_samp) / fs
    az, el = ces(el_deg, az0_deg, throw_deg, v_deg, t)
    return t, az * DEG, el * DEG


def focal_plane(n_det, disk_deg, seed=0):
    """Detectors uniform in a disk of diameter disk_deg, gamma ~ U[0, pi)."""
    rng = np.random.default_rng(seed)
    r = (disk_deg / 2) * DEG * np.sqrt(rng.uniform(size=n_det))
    phi = rng.uniform(0, 2 * np.pi, n_det)
    gamma = rng.uniform(0, np.pi, n_det)
    return FocalPlane.from_xieta(r * np.cos(phi), r * np.sin(phi), gamma)


def footprint(el_deg, throw_deg, az0_deg, disk_deg, res_arcmin, margin=4, tile=None):
    """CAR (shape, wcs) covering the scan of a disk focal plane."""
    r = disk_deg / 2
    el_lo, el_hi = el_deg - r, el_deg + r
    lon_pad = r / np.cos(np.radians(el_hi)) * 1.02
    lon_lo = -(az0_deg + throw_deg) - lon_pad
    lon_hi = -(az0_deg - throw_deg) + lon_pad
    return car_geometry((lon_lo, lon_hi), (el_lo - 0.01 * r, el_hi + 0.01 * r),
                        res_arcmin / 60., margin=margin)


class Observation:
    """One synthetic observation: assembly + map geometry + raw arrays."""

    def __init__(self, n_det, n_samp, disk_deg, res_arcmin, el_deg=50., az0_deg=0.,
                 throw_deg=10., seed=0):
        self.t, az, el = scan(n_samp, el_deg=el_deg, az0_deg=az0_deg, throw_deg=throw_deg)
        self.sight = CelestialSightLine.for_horizon(self.t, az, el)
        self.fplane = focal_plane(n_det, disk_deg, seed=seed)
        self.asm = Assembly.attach(self.sight, self.fplane)
        self.shape, self.wcs = footprint(el_deg, throw_deg, az0_deg, disk_deg, res_arcmin)
        self.n_det, self.n_samp = n_det, n_samp
        self.res_arcmin = res_arcmin

    # Raw engine inputs (the Projectionist's q_native for a CAR map at
    # crval (0, 0) is the boresight itself).
    @property
    def bore(self):
        return np.ascontiguousarray(np.asarray(self.sight.Q), dtype=np.float64)

    @property
    def dets(self):
        return np.ascontiguousarray(np.asarray(self.fplane.quats), dtype=np.float64)

    @property
    def resps(self):
        return np.ascontiguousarray(self.fplane.resps, dtype=np.float32)

    def pixelizor_args(self, tile_shape=None, active_tiles=None):
        from .proj.wcs import Projectionist
        if tile_shape is None:
            p = Projectionist.for_geom(self.shape, self.wcs)
        else:
            p = Projectionist.for_tiled(self.shape, self.wcs, tile_shape,
                                        active_tiles=active_tiles)
        return p._get_pixelizor_args()

    def projectionist(self, tile_shape=None, **kw):
        from .proj.wcs import Projectionist
        if tile_shape is None:
            return Projectionist.for_geom(self.shape, self.wcs, **kw)
        return Projectionist.for_tiled(self.shape, self.wcs, tile_shape, **kw)


def make(name='small', n_det=None, n_samp=None, seed=0):
    c = dict(CONFIGS[name])
    if n_det is not None:
        c['n_det'] = n_det
    if n_samp is not None:
        c['n_samp'] = n_samp
    return Observation(c['n_det'], c['n_samp'], c['disk_deg'], c['res_arcmin'], seed=seed)


def signal(n_det, n_samp, seed=1):
    """float32 N(0, 1) signal of shape (n_det, n_samp)."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n_det, n_samp), dtype=np.float32)


def random_map(ncomp, shape, seed=2):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((ncomp,) + tuple(shape))
