"""Boresight quaternions on the GPU against the host implementation of
CelestialSightLine.for_horizon / naive_az_el (products checked against
closed forms in test_oracle_math.py).  The device sin/cos may differ from
the host's by an ulp: agreement to 4e-16."""
import numpy as np
import pytest

from so3g_amd import workloads
from so3g_amd.proj import quat
from so3g_amd.proj.coords import CelestialSightLine

torch = pytest.importorskip('torch')


@pytest.mark.gpu
def test_horizon_and_native(gpu):
    from so3g_amd import boresight
    t, az, el = workloads.scan(50000, el_deg=47., az0_deg=12.)
    rng = np.random.default_rng(0)
    roll = rng.uniform(-0.2, 0.2, len(t))
    for r in (None, 0.3, roll):
        host = np.asarray(CelestialSightLine.for_horizon(t, az, el, roll=r).Q)
        dev = boresight.horizon(az, el, roll=r).cpu().numpy()
        assert np.abs(dev - host).max() < 4e-16
    # premultiplied by a Projectionist's celestial -> native rotation
    obs = workloads.make('small', n_det=2, n_samp=1000)
    p = obs.projectionist()
    Q = CelestialSightLine.for_horizon(t, az, el).Q
    host = np.asarray(p._cache_q_fp_to_native(Q))
    qc2n = np.asarray(p.q_celestial_to_native) if getattr(p, 'q_celestial_to_native', None) \
        is not None else np.array([1., 0, 0, 0])
    dev = boresight.horizon(az, el, q_pre=qc2n).cpu().numpy()
    assert np.abs(dev - host).max() < 4e-16


@pytest.mark.gpu
def test_naive_az_el(gpu):
    from so3g_amd import boresight
    t0 = 1.7e9
    t, az, el = workloads.scan(20000, el_deg=55., az0_deg=-30.)
    t = t + t0
    host = np.asarray(CelestialSightLine.naive_az_el(t, az, el, roll=0.1, site='so').Q)
    dev = boresight.naive_az_el(t, az, el, roll=0.1, site='so').cpu().numpy()
    assert np.abs(dev - host).max() < 4e-15   # |lst| ~ 1e5 rad: its argument rounding dominates


def _rotate(q, v):
    """Rotate vector v by unit quaternions q (..., 4) = (a, b, c, d): the
    standard rotation-matrix form of q v q*."""
    a, b, c, d = (q[..., i] for i in range(4))
    x, y, z = v
    return np.stack([
        (a * a + b * b - c * c - d * d) * x + 2 * (b * c - a * d) * y + 2 * (b * d + a * c) * z,
        2 * (b * c + a * d) * x + (a * a - b * b + c * c - d * d) * y + 2 * (c * d - a * b) * z,
        2 * (b * d - a * c) * x + 2 * (c * d + a * b) * y + (a * a - b * b - c * c + d * d) * z,
    ], -1)


@pytest.mark.gpu
def test_horizon_closed_form(gpu):
    """The device quaternions point the boresight (z axis) at the horizon
    direction (cos el cos az, -cos el sin az, sin el) and turn the x axis by
    the roll about it: (cos r sin el cos az + sin r sin az,
    -cos r sin el sin az + sin r cos az, -cos r cos el); spherical geometry of
    Rz(-az) Ry(pi/2 - el) Rz(roll), independent of the host code."""
    from so3g_amd import boresight
    t, az, el = workloads.scan(30000, el_deg=47., az0_deg=12.)
    roll = np.random.default_rng(3).uniform(-0.3, 0.3, len(t))
    q = boresight.horizon(az, el, roll=roll).cpu().numpy()
    assert np.abs(np.linalg.norm(q, axis=1) - 1).max() < 1e-15
    z = _rotate(q, (0., 0., 1.))
    want_z = np.stack([np.cos(el) * np.cos(az), -np.cos(el) * np.sin(az), np.sin(el)], -1)
    assert np.abs(z - want_z).max() < 2e-15
    xx = _rotate(q, (1., 0., 0.))
    cr, sr = np.cos(roll), np.sin(roll)
    want_x = np.stack([cr * np.sin(el) * np.cos(az) + sr * np.sin(az),
                       -cr * np.sin(el) * np.sin(az) + sr * np.cos(az),
                       -cr * np.cos(el)], -1)
    assert np.abs(xx - want_x).max() < 2e-15


@pytest.mark.gpu
def test_naive_az_el_closed_form(gpu):
    """naive_az_el points the boresight at the textbook horizontal ->
    equatorial conversion: dec = asin(sin el sin lat + cos el cos lat cos az),
    hour angle H = atan2(-sin az cos el, sin el cos lat - cos el sin lat
    cos az), RA = lst - H, with lst = ERA(t) + site longitude."""
    from so3g_amd import boresight
    from so3g_amd.proj.coords import ERA_EPOCH, ERA_POLY, SITES
    t, az, el = workloads.scan(20000, el_deg=55., az0_deg=-30.)
    t = t + 1.7e9
    q = boresight.naive_az_el(t, az, el, roll=0.1, site='so').cpu().numpy()
    site = SITES['so']
    lat = np.radians(site.lat)
    lst = np.polyval(ERA_POLY, (t - ERA_EPOCH) / 86400) + np.radians(site.lon)
    dec = np.arcsin(np.sin(el) * np.sin(lat) + np.cos(el) * np.cos(lat) * np.cos(az))
    H = np.arctan2(-np.sin(az) * np.cos(el),
                   np.sin(el) * np.cos(lat) - np.cos(el) * np.sin(lat) * np.cos(az))
    ra = lst - H
    z = _rotate(q, (0., 0., 1.))
    want = np.stack([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)], -1)
    # lst ~ 1e5 rad: rounding of the rotation angle dominates the agreement
    assert np.abs(z - want).max() < 1e-10
