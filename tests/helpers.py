"""Shared test inputs: small synthetic observations for every engine
configuration, with the matching oracle geometry."""
import numpy as np

import oracle
import so3g_amd
from so3g_amd import workloads
from so3g_amd.proj import quat
from so3g_amd.proj.wcs import Projectionist
from so3g_amd.proj.wcslite import make_wcs

DEG = np.pi / 180
FLAT_PROJS = ['CAR', 'CEA', 'TAN', 'ZEA', 'ARC', 'SIN', 'Flat', 'Quat']
TILE_SHAPE = (37, 53)


def native_rotation(proj):
    """q_celestial_to_native for a map centred on the scan (crval lat 50 deg
    for zenithal projections, (0, 0) for cylindrical ones)."""
    if proj in ('TAN', 'ZEA', 'ARC', 'SIN'):
        return Projectionist.get_q(make_wcs(proj, (0., 50.), (-1., 1.), (1., 1.)))
    return quat.quat(1, 0, 0, 0)


class Case:
    """One engine configuration on a small observation.  The flat map
    footprint is fitted to the oracle's projected coordinates, so every
    projection gets a map of about ny x nx pixels around the scan (clip
    moves it so that part of the scan falls off the map)."""

    def __init__(self, proj, comps='TQU', tiled=False, interp=0, n_det=6, n_samp=3000,
                 seed=0, nside=128, nest=True, nside_tile=8, clip=False, ny=90, nx=160,
                 disk_deg=1.0):
        self.proj, self.comps, self.tiled, self.interp = proj, comps, tiled, interp
        self.obs = workloads.Observation(n_det, n_samp, disk_deg, 2.0, seed=seed)
        q_c2n = native_rotation(proj) if proj != 'HP' else quat.quat(1, 0, 0, 0)
        self.bore = np.ascontiguousarray(np.asarray(q_c2n * self.obs.sight.Q), np.float64)
        self.dets = self.obs.dets
        rng = np.random.default_rng(seed + 7)
        self.resps = np.ascontiguousarray(
            np.stack([rng.uniform(0.5, 1.5, n_det), rng.uniform(0.2, 1.0, n_det)], 1),
            dtype=np.float32)
        self.signal = workloads.signal(n_det, n_samp, seed=seed + 1)
        self.det_weights = rng.uniform(0.5, 2, n_det).astype(np.float32)
        self.ncomp = len(comps)
        if proj == 'HP':
            active = None
            if tiled:
                g1 = oracle.geom_from_args('HP', comps, (nside, 1, nside_tile, None))
                active = sorted(set(np.unique(oracle.pixels(g1, self.bore, self.dets)[..., 0])
                                    .tolist()))
            self.pixargs = (nside, int(nest), nside_tile if tiled else None, active)
            name = f'ProjEng_HP_{comps}_{"Tiled" if tiled else "NonTiled"}'
            self.og = oracle.geom_from_args('HP', comps, self.pixargs)
        else:
            g0 = oracle.geom(proj, 3)
            co = oracle.coords(g0, self.bore, self.dets)
            x, y = co[..., 0].ravel(), co[..., 1].ravel()
            m = 3
            cdx = -(x.max() - x.min()) / (nx - 2 * m)
            cdy = (y.max() - y.min()) / (ny - 2 * m)
            # Pixel index floor(c/cdelt + crpix - 1 + 0.5); put x.max at m.
            crx = m + 0.5 - x.max() / cdx
            cry = m + 0.5 - y.min() / cdy
            if clip:
                crx += nx * 0.4
                cry -= ny * 0.3
            args = (ny, nx, cdy, cdx, cry, crx)
            if tiled:
                args += TILE_SHAPE
                n_tile = ((ny + TILE_SHAPE[0] - 1) // TILE_SHAPE[0]) * \
                    ((nx + TILE_SHAPE[1] - 1) // TILE_SHAPE[1])
                args += (list(range(n_tile)),)
            self.pixargs = args
            name = f'ProjEng_{proj}_{comps}_{"Tiled" if tiled else "NonTiled"}' + \
                ('_Bilinear' if interp else '')
            self.og = oracle.geom_from_args(proj, comps, args, interp=interp)
        self.name = name
        self.engine = getattr(so3g_amd, name)(self.pixargs)

    def map_like(self, lead, seed=3):
        """Random map (array or tile list) with leading shape lead."""
        rng = np.random.default_rng(seed)
        z = self.engine.zeros(tuple(lead))
        if isinstance(z, list):
            return [None if t is None else rng.standard_normal(t.shape) for t in z]
        return rng.standard_normal(z.shape)

    def zeros(self, lead):
        return self.engine.zeros(tuple(lead))


def copy_map(m):
    if isinstance(m, list):
        return [None if t is None else np.array(t) for t in m]
    return np.array(m)


def flat_map(m):
    if isinstance(m, list):
        return np.concatenate([t.ravel() for t in m if t is not None])
    return np.asarray(m).ravel()


def all_flat_cases():
    out = []
    for proj in FLAT_PROJS:
        for tiled in (False, True):
            for interp in (0, 1):
                out.append((proj, tiled, interp))
    return out


def _atan_dec(x):
    """atan(x) of a Decimal in the current context: argument halving
    (atan x = 2 atan(x / (1 + sqrt(1 + x^2)))) to |x| < 0.05, then the
    Taylor series."""
    from decimal import Decimal
    m = 0
    while abs(x) > Decimal('0.05'):
        x = x / (1 + (1 + x * x).sqrt())
        m += 1
    s, term, n, x2 = Decimal(0), x, 0, x * x
    eps = Decimal(10) ** -60
    while True:
        t = term / (2 * n + 1)
        if abs(t) < eps:
            break
        s += t if n % 2 == 0 else -t
        term *= x2
        n += 1
    return s * (2 ** m)


def cr_atan2(y, x):
    """atan2 of two doubles, correctly rounded to double: the quadrant
    arithmetic of C99 atan2 on a 70-digit decimal atan, then rounded once
    (float(Decimal) rounds to nearest).  Finite, non-zero arguments."""
    from decimal import Decimal, localcontext
    with localcontext() as ctx:
        ctx.prec = 70
        yd, xd = Decimal(float(y)), Decimal(float(x))
        pi = 4 * _atan_dec(Decimal(1))
        if abs(yd) <= abs(xd):
            a = _atan_dec(abs(yd) / abs(xd))
        else:
            a = pi / 2 - _atan_dec(abs(xd) / abs(yd))
        if xd < 0:
            a = pi - a
        return float(-a if yd < 0 else a)


def cea_atan2_args(bore, dets):
    """The two atan2 arguments of so3g's CEA pointer (Projection.cxx:460-468):
    y = c d - a b, x = c a + d b of the boresight x detector quaternion
    (the oracle's product, bit-exact), in float64 without FMA."""
    gq = oracle.geom('Quat', 4)
    q = oracle.coords(gq, bore, dets)
    a, b, c, d = (q[..., i] for i in range(4))
    return c * d - a * b, c * a + d * b
