"""Boresight and focal-plane inputs of the projection engine:
CelestialSightLine (boresight quaternions per sample), FocalPlane
(detector offset quaternions + (T, P) responses) and Assembly.

Mirrors python/proj/coords.py:12-479.  The high-precision ``az_el``
constructor needs qpoint, which is not part of this build; it raises
ImportError as the reference does when qpoint is absent.
"""
import numpy as np

from . import quat

DEG = np.pi / 180.


class EarthlySite:
    def __init__(self, lon, lat, elev, typical_weather=None):
        """Arguments in degrees E, degrees N, meters above sea level."""
        self.lon, self.lat, self.elev = lon, lat, elev
        self.typical_weather = typical_weather


def _debabyl(deg, arcmin, arcsec):
    return deg + arcmin / 60 + arcsec / 3600


SITES = {
    'act': EarthlySite(-67.7876, -22.9585, 5188.),
    'so_lat': EarthlySite(-_debabyl(67, 47, 15.68), -_debabyl(22, 57, 39.47), 5188.),
    'so_sat1': EarthlySite(-_debabyl(67, 47, 18.11), -_debabyl(22, 57, 36.38), 5188.),
    'so_sat2': EarthlySite(-_debabyl(67, 47, 17.28), -_debabyl(22, 57, 36.35), 5188.),
    'so_sat3': EarthlySite(-_debabyl(67, 47, 16.53), -_debabyl(22, 57, 35.97), 5188.),
}
SITES['so'] = SITES['so_lat']
SITES['_default'] = SITES['so']
DEFAULT_SITE = 'so'

# Naive horizon -> celestial conversion (coords.py:66-69).
ERA_EPOCH = 946684800 + 3600 * 12
ERA_POLY = np.array([6.300387486754831, 4.894961212823756])


class CelestialSightLine:
    """Carries the boresight pointing quaternions in self.Q."""

    @staticmethod
    def decode_site(site=None):
        if site is None:
            site = DEFAULT_SITE
        if isinstance(site, EarthlySite):
            return site
        if site in SITES:
            return SITES[site]
        raise ValueError('Could not decode %s as a Site.' % site)

    @classmethod
    def naive_az_el(cls, t, az, el, roll=0., site=None, weather=None):
        """Horizon (az, el, radians) at unix time t to celestial, ignoring
        precession and refraction (coords.py:99-125)."""
        site = cls.decode_site(site)
        self = cls()
        J = (np.asarray(t) - ERA_EPOCH) / 86400
        era = np.polyval(ERA_POLY, J)
        lst = era + site.lon * DEG
        self.Q = (quat.euler(2, lst) *
                  quat.euler(1, np.pi / 2 - site.lat * DEG) *
                  quat.euler(2, np.pi) *
                  quat.euler(2, -np.asarray(az)) *
                  quat.euler(1, np.pi / 2 - np.asarray(el)) *
                  quat.euler(2, roll))
        return self

    @classmethod
    def az_el(cls, t, az, el, roll=None, site=None, weather=None, **kwargs):
        import qpoint  # noqa: F401  (not available in this build)
        raise NotImplementedError

    @classmethod
    def for_lonlat(cls, lon, lat, psi=0):
        """Boresight directly at celestial (lon, lat) with focal-plane
        rotation psi (coords.py:171-188)."""
        self = cls()
        self.Q = quat.euler(2, lon) * quat.euler(1, np.pi / 2 - np.asarray(lat)) * \
            quat.euler(2, psi)
        return self

    @classmethod
    def for_horizon(cls, t, az, el, roll=None, site=None, weather=None):
        """Horizon coordinates taken as the celestial frame (coords.py:190-209):
        Q = euler(2, -az) euler(1, pi/2 - el) euler(2, roll)."""
        self = cls()
        az, el, t = map(np.asarray, [az, el, t])
        roll = 0. if roll is None else np.asarray(roll)
        self.Q = quat.euler(2, -az) * quat.euler(1, np.pi / 2 - el) * quat.euler(2, roll)
        return self

    def coords(self, fplane=None, output=None):
        """Celestial (lon, lat, cos2psi, sin2psi) per detector and sample,
        through a CAR engine (coords.py:211-237)."""
        from .. import ProjEng_CAR_TQU_NonTiled
        p = ProjEng_CAR_TQU_NonTiled((1, 1, 1., 1., 1., 1.))
        collapse = (fplane is None)
        if collapse:
            fplane = FocalPlane.boresight()
            if output is not None:
                output = output[None]
        output = p.coords(self.Q, fplane.quats, output)
        if collapse:
            output = output[0]
        return output


class FocalPlane:
    """Detector offset quaternions (quats, (ndet, 4)) and float32 (T, P)
    responses (resps, (ndet, 2))."""

    def __init__(self, quats=None, resps=None, dets=None):
        if quats is None:
            quats = []
        self.quats = quat.G3VectorQuat(np.asarray(quats))
        self.resps = np.ones((len(self.quats), 2), np.float32)
        if resps is not None:
            self.resps[:] = resps
        if np.any(~np.isfinite(np.asarray(self.quats))):
            raise ValueError('nan/inf values in detector quaternions')
        if np.any(~np.isfinite(self.resps)):
            raise ValueError('nan/inf values in detector responses')
        self._dets = list(dets) if dets is not None else []

    def coeffs(self):
        return np.array(self.quats, dtype=np.float64)

    @classmethod
    def boresight(cls):
        return cls(quats=[[1, 0, 0, 0]])

    @classmethod
    def from_xieta(cls, xi, eta, gamma=0, T=1, P=1, Q=1, U=0, hwp=False, dets=None):
        """Detectors at focal-plane (xi, eta) with polarisation angle gamma
        and responses; gamma_total = gamma + arctan2(U,Q)/2 and
        P_total = P * hypot(Q, U) (coords.py:305-381)."""
        gamma = gamma + np.arctan2(U, Q) / 2
        P = P * (np.asarray(Q) ** 2 + np.asarray(U) ** 2) ** 0.5
        if hwp:
            gamma = -gamma
        xi, eta, gamma, T, P, _ = np.broadcast_arrays(xi, eta, gamma, T, P, [0])
        quats = quat.rotation_xieta(xi, eta, gamma)
        resps = np.ones((len(quats), 2))
        resps[:, 0] = T
        resps[:, 1] = P
        return cls(quats, resps=resps, dets=dets)

    def __repr__(self):
        return 'FocalPlane(quats=%s,resps=%s)' % (repr(self.coeffs()), repr(self.resps))

    @property
    def ndet(self):
        return len(self.quats)

    def __len__(self):
        return self.ndet

    def __getitem__(self, sel):
        return FocalPlane(quats=self.coeffs()[sel], resps=self.resps[sel])

    def items(self):
        for q, resp in zip(self.quats, self.resps):
            yield q, resp


class Assembly:
    """Boresight + focal plane, as consumed by the Projectionist."""

    def __init__(self, collapse=False):
        self.collapse = collapse

    @classmethod
    def attach(cls, sight_line, fplane):
        self = cls()
        if isinstance(sight_line, np.ndarray):
            self.Q = quat.G3VectorQuat(sight_line) if not isinstance(sight_line, quat.Quat) \
                else sight_line
        else:
            self.Q = sight_line.Q
        self.fplane = fplane
        return self

    @classmethod
    def for_boresight(cls, sight_line):
        self = cls(collapse=True)
        self.Q = sight_line.Q
        self.fplane = FocalPlane.boresight()
        return self

    @property
    def dets(self):
        return self.fplane
