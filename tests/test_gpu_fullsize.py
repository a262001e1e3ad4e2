"""Full-size pointing parity on the benchmark observation: 1000 detectors x
720000 samples (7.2e8 det*samples; constant-elevation scan at el 50 deg, az
+-10 deg at 1 deg/s, 200 Hz, detectors in a 1.5 deg disk).

Every pixel index and every spin factor that get_pointing_matrix returns is
compared bit for bit with the OpenMP oracle, in detector chunks (the host
holds one chunk at a time), for the three maps of this configuration: CAR
0.5' non-tiled, CAR 0.5' tiled 100 x 100, HEALPix NEST nside 4096.  CEA
0.5' is compared as well: its longitude is the C library's atan2 in the
oracle (glibc) and the correctly rounded atan2 on the device, so the
longitudes are counted ulp by ulp (glibc's last bit differs in about 0.1 %
of such arguments), the differing ones checked against a 70-digit decimal
atan2, and the pixels must still agree everywhere.

The device trig tables (built on the GPU host with its C library at
so3g_init) are checked against the SHA-256 of the tables of the reference's
own so_linterp.h compiled with glibc 2.35.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N_DET, N_SAMP = 1000, 720000
CHUNK = 50


def test_device_trig_tables_match_golden(gpu):
    from so3g_amd import _native
    a, b = _native.get_tables()
    g = np.load(os.path.join(ROOT, 'tests', 'golden', 'trig_tables.npz'))
    assert hashlib.sha256(a.tobytes()).hexdigest() == str(g['sha_asin'])
    assert hashlib.sha256(b.tobytes()).hexdigest() == str(g['sha_atan2'])


@pytest.fixture(scope='module')
def obs():
    from so3g_amd import workloads
    return workloads.Observation(N_DET, N_SAMP, 1.5, 0.5, seed=0)


def _device_inputs(obs, dev):
    import torch
    return [torch.from_numpy(a).to(dev) for a in (obs.bore, obs.dets, obs.resps)]


def _compare_chunks(eng, og, obs, dev, spin=True):
    """Pointing matrix of every detector chunk on the GPU and in the oracle;
    returns (samples compared, on-map samples)."""
    bore_t, dets_t, resp_t = _device_inputs(obs, dev)
    n_on = 0
    for d0 in range(0, N_DET, CHUNK):
        d1 = min(d0 + CHUNK, N_DET)
        if spin:
            gp, gs = eng.pointing_matrix(bore_t, dets_t[d0:d1], resp_t[d0:d1])
            rp, rs = oracle.pointing_matrix(og, obs.bore, obs.dets[d0:d1], obs.resps[d0:d1])
            gs = gs.cpu().numpy()
            nbad = int((gs.view(np.int32) != rs.view(np.int32)).sum())
            assert nbad == 0, f'dets {d0}:{d1}: {nbad} spin factors differ'
        else:
            gp = eng.pixels(bore_t, dets_t[d0:d1])
            rp = oracle.pixels(og, obs.bore, obs.dets[d0:d1])
        gp = gp.cpu().numpy()
        assert gp.shape == rp.shape
        bad = gp != rp
        nbad = int(bad.sum())
        assert nbad == 0, f'dets {d0}:{d1}: {nbad} pixel indices differ, e.g. ' \
            f'{gp[bad][:4]} vs {rp[bad][:4]}'
        n_on += int((rp[..., 0] >= 0).sum())
    return N_DET * N_SAMP, n_on


def test_full_size_car(gpu, obs):
    import so3g_amd
    args = obs.pixelizor_args()
    eng = so3g_amd.ProjEng_CAR_TQU_NonTiled(args)
    n, n_on = _compare_chunks(eng, oracle.geom_from_args('CAR', 'TQU', args), obs, gpu)
    print(f'CAR 0.5\' {obs.shape}: {n} samples, {n_on} on the map, all bit-exact')
    assert n_on == n


def test_full_size_car_tiled(gpu, obs):
    import so3g_amd
    args = obs.pixelizor_args(tile_shape=(100, 100))
    eng = so3g_amd.ProjEng_CAR_TQU_Tiled(args)
    n, n_on = _compare_chunks(eng, oracle.geom_from_args('CAR', 'TQU', args), obs, gpu)
    print(f'CAR 0.5\' tiled 100x100: {n} samples, {n_on} on the map, all bit-exact')
    assert n_on == n


def test_full_size_healpix_nest_4096(gpu, obs):
    import so3g_amd
    args = (4096, 1)
    eng = so3g_amd.ProjEng_HP_TQU_NonTiled(args)
    n, _ = _compare_chunks(eng, oracle.geom_from_args('HP', 'TQU', args), obs, gpu)
    print(f'HEALPix NEST 4096: {n} samples, all bit-exact')


def test_full_size_cea(gpu, obs):
    """CEA 0.5': y = sin(lat), so the map covers sin(el) of the scan."""
    import so3g_amd
    from so3g_amd.proj.wcslite import car_geometry
    from so3g_amd.proj.wcs import Projectionist
    r = 0.75
    lon_pad = r / np.cos(np.radians(50 + r)) * 1.02
    deg = 180 / np.pi
    lat = (deg * np.sin(np.radians(50 - 1.01 * r)), deg * np.sin(np.radians(50 + 1.01 * r)))
    shape, wcs = car_geometry((-10 - lon_pad, 10 + lon_pad), lat, 0.5 / 60, margin=4,
                              proj='CEA')
    args = Projectionist.for_geom(shape, wcs)._get_pixelizor_args()
    eng = so3g_amd.ProjEng_CEA_TQU_NonTiled(args)
    og = oracle.geom_from_args('CEA', 'TQU', args)
    bore_t, dets_t, resp_t = _device_inputs(obs, gpu)
    from helpers import cea_atan2_args, cr_atan2
    n_lon = n_other = max_ulp = n_checked = 0
    for d0 in range(0, N_DET, CHUNK):
        d1 = min(d0 + CHUNK, N_DET)
        gc = eng.coords(bore_t, dets_t[d0:d1]).cpu().numpy()
        rc = oracle.coords(og, obs.bore, obs.dets[d0:d1])
        n_other += int((gc[..., 1:].view(np.int64) != rc[..., 1:].view(np.int64)).sum())
        du = np.abs(gc[..., 0].view(np.int64) - rc[..., 0].view(np.int64))
        n_lon += int((du != 0).sum())
        max_ulp = max(max_ulp, int(du.max()))
        # the first differing samples of each chunk hold the correctly
        # rounded atan2
        for dd, t in np.argwhere(du != 0)[:20]:
            y, x = cea_atan2_args(obs.bore[t:t + 1], obs.dets[d0 + dd:d0 + dd + 1])
            assert gc[dd, t, 0] == cr_atan2(y[0, 0], x[0, 0]), (d0 + dd, t)
            n_checked += 1
        del gc, rc
    print(f'CEA longitude (device correctly rounded atan2 vs glibc atan2): {n_lon} of '
          f'{N_DET * N_SAMP} differ, by at most {max_ulp} ulp; {n_checked} of them checked '
          f'against a 70-digit atan2')
    assert n_other == 0
    assert max_ulp <= 1
    assert n_lon <= 0.01 * N_DET * N_SAMP
    n, n_on = _compare_chunks(eng, og, obs, gpu)
    print(f'CEA 0.5\' {shape}: {n} samples, {n_on} on the map, pixels and spin bit-exact')
    assert n_on == n
