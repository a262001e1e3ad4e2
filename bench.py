"""Benchmark: det*samples/s of one to_map + from_map pass over one large
observation (1000 detectors x 720000 samples = 7.2e8 det*samples; constant-
elevation scan, el 50 deg, az +-10 deg at 1 deg/s, 200 Hz, detectors in a
1.5 deg disk; CAR 0.5 arcmin TQU float64 map, 191 x 2701 pixels).

One "step" = from_map (map -> fresh float32 signal) + to_map (signal ->
float64 TQU map, zeroed first) over the whole observation, with every input
resident in HBM.  The headline path projects through the cached pointing
matrix: get_pointing_matrix once, a PrecompPlan built from it once, then the
plan's from_map / to_map -- the operator pair of every CG iteration, with
the results of ProjEng_Precomp's from_map / to_map.  The plain
ProjEng_Precomp kernels and the on-the-fly path (pointing recomputed in both
kernels, the Projectionist's default) are timed beside it.

Beside the headline, the line carries: the drop-in ProjEng_Precomp class on
the same tensors (its cached plan), the unplanned Precomp kernels, the
on-the-fly path, the small (100 x 40000) observation, the tiled CAR and
HEALPix nside 4096 maps of the same observation, get_pixels /
get_pointing_matrix / to_weights, N^-1 (nmat_detvecs_apply), G3SuperTimestream
decoding, the precision study (100 x 40000 samples into ZEA / TAN 0.25 arcmin
and HEALPix RING nside 8192), the batch of 64 uneven observations in one
tuned launch (against one launch per observation), the map-making solve
(binned, two-level, learned-deflation and U-Net CG to 1e-6), and the CPU
baseline (the oracle on the host cores).

Prints ONE JSON line.  `python bench.py [--steps K] [--warmup W] [--gpus 1]`.

`--pmc-pass` runs, instead of the timing, one plain copy of 1 GiB (the
calibration for rocprofv3's FETCH_SIZE / WRITE_SIZE counters) and one launch
of each headline kernel; the HBM traffic measured that way is read back from
profiles/traffic_large_car_tqu.md into roofline.traffic.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
COPY_BYTES = 1 << 30  # calibration copy of the --pmc-pass run
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'traffic_large_car_tqu.md')
FP64_FILE = os.path.join(ROOT, 'profiles', 'fp64_onthefly_car_tqu.md')
FP64_PEAK_OPS = 39.3e12  # MI355X FP64 vector: 78.6 TFLOP/s = 39.3e12 lane ops/s (FMA = 1 op)


def profile_json(path):
    """The JSON block of a committed rocprofv3 --pmc summary (None when
    absent): HBM bytes per launch of the headline kernels, or FP64 lane
    operations per sample of the on-the-fly kernels."""
    try:
        txt = open(path).read()
        block = txt.split('```json', 1)[1].split('```', 1)[0]
        return json.loads(block)
    except (OSError, IndexError, ValueError):
        return None


def measured_traffic():
    return profile_json(TRAFFIC_FILE)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def plan_bytes(pinfo, n_s):
    """Algorithmic bytes of the planned from_map and to_map kernels (as
    described in main) and the to_map kernel's name."""
    from so3g_amd import _native as N
    n_on, ns, n_vis = pinfo['n_on'], pinfo['stored_spin'], pinfo['n_vis']
    if pinfo['layout'] == 2:
        # windowed layout: per sample a u16 window slot + the stored factors
        # + the float32 signal (written by from_map, read by to_map); per
        # window slot its compact -> flat pixel (4 B) and the float64 TQU
        # values, read once by from_map, read and written once by to_map's
        # flush
        slots = pinfo['window_slots']
        b_from = n_s * (2 + 4 * ns + 4) + slots * (4 + 8 * 3)
        b_to = n_s * (2 + 4 * ns + 4) + slots * (4 + 2 * 8 * 3)
        return b_from, b_to, 'k_tw_to_map<3,true,float,true,1024,true>'
    # u16 pixel offsets plus one int32 base per 512 samples; the wide groups
    # read int32 pixels (and their group number) instead
    b_from = n_s * (2 + 4 * ns + 4) + 4 * ((n_s + 511) // 512) + \
        pinfo['wide_groups'] * (512 * 2 + 4) + 8 * 3 * n_vis
    if pinfo['layout'] == 1:
        slots = N.get_option('plan_tile_rows') * N.get_option('plan_tile_cols')
        b_to = n_on * (2 + 4 * ns + 4) + 12 * ((n_on + 63) // 64) + 4 * pinfo['stretches'] + \
            pinfo['items'] * (16 + 4 * slots) + 2 * 8 * 3 * n_vis
        return b_from, b_to, 'k_tp_to_map<3,true,float,true>'
    b_to = n_on * (4 + 4 * ns + 4) + n_on // 8 + 4 * pinfo['runs'] + 4 * (n_on // 512) + \
        2 * 8 * 3 * n_vis
    return b_from, b_to, 'k_plan_to_map_c<3,true,float,1,true>'


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--n-det', type=int, default=1000)
    ap.add_argument('--n-samp', type=int, default=720000)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-otf', action='store_true', help='skip the on-the-fly leg')
    ap.add_argument('--pmc-pass', action='store_true',
                    help='one calibration copy + one launch of each headline kernel')
    ap.add_argument('--pmc-otf', action='store_true',
                    help='with --pmc-pass: also one launch of each on-the-fly kernel')
    ap.add_argument('--no-small', action='store_true', help='skip the small-observation leg')
    ap.add_argument('--no-g3', action='store_true',
                    help='skip the G3SuperTimestream decoding leg')
    ap.add_argument('--no-studies', action='store_true',
                    help='skip the precision (C3), batched (C4) and map-making (C5) legs')
    ap.add_argument('--only-small', action='store_true',
                    help='time only the small-observation pair and print its JSON')
    ap.add_argument('--only-tiles', action='store_true',
                    help='time only the tiled-plan to_map on the CAR 0.5 arcmin and HEALPix '
                         'NEST 4096 maps of the benchmark observation and print their JSON')
    ap.add_argument('--only-nmat', action='store_true',
                    help='time only nmat_detvecs_apply and print its JSON, then stream the '
                         'same array once (a scale and a sum) as a reference for profiles')
    ap.add_argument('--cpu-full', action='store_true',
                    help='print the full CPU baseline table (oracle, median of 5) and exit')
    return ap.parse_args()


def main():
    args = parse()
    world = int(os.environ.get('WORLD_SIZE', '1'))
    if args.gpus != 1 or world != 1:
        print('bench.py measures one GPU only: this build has no multi-GPU path '
              '(run with --gpus 1).', file=sys.stderr)
        sys.exit(2)

    import torch
    import so3g_amd
    from so3g_amd import _native as N, workloads

    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)
    N.init(0)
    lib = N.load()
    if args.only_small:
        stream = torch.cuda.current_stream().cuda_stream or None
        print(json.dumps(small_config(torch, so3g_amd, N, workloads, dev, stream)), flush=True)
        return
    if args.only_tiles:
        print(json.dumps(tile_scatter_timing(torch, so3g_amd, workloads, dev, args.n_det,
                                             args.n_samp)), flush=True)
        return
    if args.only_nmat:
        print(json.dumps(nmat_timing(torch, so3g_amd, dev, args.n_det, args.n_samp)), flush=True)
        ref = torch.ones((args.n_det, 2 * (args.n_samp // 2 + 1)), dtype=torch.float32,
                         device=dev)
        ref.mul_(1.5)
        float(ref.sum())
        torch.cuda.synchronize()
        return

    t0 = time.time()
    obs = workloads.Observation(args.n_det, args.n_samp, 1.5, 0.5, seed=0)
    p = obs.projectionist()
    pixargs = p._get_pixelizor_args()
    eng = so3g_amd.ProjEng_CAR_TQU_NonTiled(pixargs)
    g = eng._g
    ny, nx = obs.shape
    n_det, n_t = obs.n_det, obs.n_samp
    N_s = n_det * n_t
    bore = torch.from_numpy(obs.bore).to(dev)
    dets = torch.from_numpy(obs.dets).to(dev)
    resp = torch.from_numpy(obs.resps).to(dev)
    log(f'workload built in {time.time() - t0:.1f}s: {n_det} x {n_t}, map {ny}x{nx}')

    stream = torch.cuda.current_stream().cuda_stream or None

    def chk(rc):
        N.check(rc)

    # Signal = from_map(seeded Gaussian random TQU map) + white noise.
    gen = torch.Generator(device=dev).manual_seed(2)
    sky = torch.randn((3, ny, nx), dtype=torch.float64, device=dev, generator=gen)
    sig = torch.empty((n_det, n_t), dtype=torch.float32, device=dev)
    chk(lib.so3g_from_map(g, sky.data_ptr(), None, None, bore.data_ptr(), n_t, dets.data_ptr(),
                          n_det, resp.data_ptr(), sig.data_ptr(), 0, stream))
    sig += torch.randn((n_det, n_t), dtype=torch.float32, device=dev, generator=gen)

    # Cached pointing matrix (outside the timed region, as in a CG solve).
    torch.cuda.synchronize()
    t0 = time.time()
    pix = torch.empty((n_det, n_t, 2), dtype=torch.int32, device=dev)
    spin = torch.empty((n_det, n_t, 3), dtype=torch.float32, device=dev)
    chk(lib.so3g_pointing_matrix(g, bore.data_ptr(), n_t, dets.data_ptr(), n_det,
                                 resp.data_ptr(), pix.data_ptr(), spin.data_ptr(), stream))
    torch.cuda.synchronize()
    t_pmat = time.time() - t0
    log(f'pointing matrix: {t_pmat * 1e3:.1f} ms')

    # Pixel-sorted plan for the scatter (built once per pointing matrix).
    torch.cuda.synchronize()
    t0 = time.time()
    plan = so3g_amd.PrecompPlan(pix, spin, (ny, nx))
    torch.cuda.synchronize()
    t_plan = time.time() - t0
    log(f'plan: {t_plan * 1e3:.1f} ms')

    shape = np.array([ny, nx], np.int64)
    sig_out = torch.empty_like(sig)
    mapo = torch.zeros((3, ny, nx), dtype=torch.float64, device=dev)

    def pc_from():
        chk(lib.so3g_pc_plan_from_map(plan._h, sky.data_ptr(), sig_out.data_ptr(), 0, stream))

    def pc_from_direct():
        chk(lib.so3g_precomp_from_map(sky.data_ptr(), shape.ctypes.data, None, None,
                                      pix.data_ptr(), 2, spin.data_ptr(), 3, n_det, n_t,
                                      sig_out.data_ptr(), 0, stream))

    def pc_to():
        mapo.zero_()
        chk(lib.so3g_pc_plan_to_map(plan._h, mapo.data_ptr(), sig.data_ptr(), None, stream))

    def pc_to_direct():
        mapo.zero_()
        chk(lib.so3g_precomp_to_map(mapo.data_ptr(), shape.ctypes.data, None, None,
                                    pix.data_ptr(), 2, spin.data_ptr(), 3, n_det, n_t,
                                    sig.data_ptr(), None, None, 0, stream))

    def otf_from():
        chk(lib.so3g_from_map(g, sky.data_ptr(), None, None, bore.data_ptr(), n_t,
                              dets.data_ptr(), n_det, resp.data_ptr(), sig_out.data_ptr(), 0,
                              stream))

    def otf_to():
        mapo.zero_()
        chk(lib.so3g_to_map(g, mapo.data_ptr(), None, None, bore.data_ptr(), n_t,
                            dets.data_ptr(), n_det, resp.data_ptr(), sig.data_ptr(), None,
                            None, 0, stream))

    def run(ops, steps, warmup):
        for _ in range(warmup):
            for f in ops:
                f()
        torch.cuda.synchronize()
        ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in ops] for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(steps):
            for i, f in enumerate(ops):
                ev[s][i][0].record()
                f()
                ev[s][i][1].record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        per_op = [float(np.mean([ev[s][i][0].elapsed_time(ev[s][i][1]) for s in range(steps)]))
                  for i in range(len(ops))]
        return wall, per_op

    if args.pmc_pass:
        a = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        pc_from()
        torch.cuda.synchronize()
        pc_to()
        torch.cuda.synchronize()
        if args.pmc_otf:
            otf_from()
            torch.cuda.synchronize()
            otf_to()
            torch.cuda.synchronize()
        print(json.dumps({'pmc_pass': True, 'copy_bytes': COPY_BYTES}), flush=True)
        return

    if args.cpu_full:
        print(json.dumps(cpu_full(obs, pix, spin, sig, sky, pixargs)), flush=True)
        return

    log('timing cached-pointing path ...')
    wall, (ms_from, ms_to) = run([pc_from, pc_to], args.steps, args.warmup)
    ms_step = wall / args.steps * 1e3
    value = N_s / (wall / args.steps)

    # Algorithmic bytes per launch of the planned kernels: what each must
    # move at least from its own storage (see the plan's layout):
    #   from_map: per sample 2 B compact-pixel offset + 8 B (Q, U) factors +
    #             4 B signal written (the T factor is one float per detector),
    #             4 B base per 512 samples (int32 pixels, 4 B per sample, for
    #             the groups whose pixels span more than 65534), plus the
    #             float64 TQU values of the visited pixels read once;
    #   to_map (tiled layout): per on-map sample 2 B tile slot + 8 B (Q, U)
    #             factors + 4 B signal, plus 12 B per 64 positions (start mask,
    #             start count), 4 B per start (stretches and chunk starts) and
    #             16 B per work item, the 8 KB slot-to-pixel table of each item's
    #             tile, and the visited pixels' float64 TQU values read and
    #             written once;
    #   to_map (pixel-run layout): per on-map sample 4 B sample index + 8 B
    #             factors + 4 B signal + 1 bit run start, plus 4 B per pixel run
    #             and per 512-sample group, plus the visited pixels' values
    #             read and written once.
    # The ProjEng_Precomp interface itself streams 24 B per sample (int32
    # pixel pair, 3 float32 factors, float32 signal): `precomp_equivalent`
    # reports the rates in those bytes.
    npix = ny * nx
    pinfo = plan.info()
    n_on, n_runs, ns, n_vis = pinfo['n_on'], pinfo['runs'], pinfo['stored_spin'], pinfo['n_vis']
    b_from, b_to, to_kernel = plan_bytes(pinfo, N_s)
    b_pc = N_s * (4 * 2 + 4 * 3 + 4)
    gbs_from = b_from / (ms_from * 1e-3)
    gbs_to = b_to / (ms_to * 1e-3)
    tr = measured_traffic() or {}
    if ms_to >= ms_from:
        dom = (to_kernel, gbs_to, b_to, tr.get('to_map_bytes'))
    else:
        dom = ('k_plan_from_map_s<3,true,4,1,false>', gbs_from, b_from,
               tr.get('from_map_bytes'))
    # The plain ProjEng_Precomp kernels (no plan), for reference.
    _, (ms_from_direct, ms_to_direct) = run([pc_from_direct, pc_to_direct],
                                            max(1, min(args.steps, 3)), 1)
    # The drop-in ProjEng_Precomp class with these torch tensors: its first
    # call builds and caches the plan, later calls reuse it.
    precomp_api = precomp_api_timing(torch, so3g_amd, pix, spin, sky, sig, mapo)

    otf = None
    if not args.no_otf:
        log('timing on-the-fly path ...')
        steps_otf = max(1, min(args.steps, 3))
        wall_o, (ms_of, ms_ot) = run([otf_from, otf_to], steps_otf, 1)
        otf = {
            'value': N_s / (wall_o / steps_otf),
            'ms_per_step': wall_o / steps_otf * 1e3,
            'ms_from_map': ms_of, 'ms_to_map': ms_ot,
            'hbm_frac_from_map': (N_s * 4 + 32 * n_t + 40 * n_det + 8 * 3 * npix) /
            (ms_of * 1e-3) / HBM_PEAK,
            'hbm_frac_to_map': (N_s * 4 + 32 * n_t + 40 * n_det + 16 * 3 * npix) /
            (ms_ot * 1e-3) / HBM_PEAK,
        }
        f64 = (profile_json(FP64_FILE) or {}).get('fp64_ops_per_sample')
        if f64:
            otf['fp64_ops_per_sample'] = f64
            otf['fp64_frac_from_map'] = f64['from_map'] * N_s / (ms_of * 1e-3) / FP64_PEAK_OPS
            otf['fp64_frac_to_map'] = f64['to_map'] * N_s / (ms_ot * 1e-3) / FP64_PEAK_OPS

    cpu = None if args.no_cpu else cpu_baseline(obs, pix, spin, sig, sky, pixargs)
    small = None if args.no_small else small_config(torch, so3g_amd, N, workloads, dev, stream)
    nmat = None if args.no_small else nmat_timing(torch, so3g_amd, dev, n_det, n_t)
    g3 = None if args.no_g3 else g3super_timing(torch, n_det, n_t, not args.no_cpu)
    variants = None if args.no_small else c2_variants(torch, so3g_amd, N, lib, obs, pixargs, bore,
                                                      dets, resp, sig, dev, stream, run)
    other = None if args.no_small else c2_other_ops(torch, N, lib, g, bore, dets, resp, pix,
                                                    spin, plan, ny, nx, dev, stream, run)
    c3 = c4 = c5 = None
    if not args.no_studies:
        del plan
        torch.cuda.empty_cache()
        c3 = c3_precision()
        c4 = c4_batched()
        c5 = c5_mapmaking()

    out = {
        'metric': 'det*samples/s for to_map+from_map (% HBM roofline)',
        'value': value,
        'unit': 'det*samples/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 products, f64 map accumulation',
        'data': 'synthetic (seeded scan, focal plane, random TQU sky + white noise)',
        'config': {
            'workload': f'single observation {n_det} det x {n_t} samples, CAR 0.5 arcmin '
                        f'TQU {ny}x{nx} non-tiled; step = from_map + to_map through the '
                        f'cached pointing matrix (PrecompPlan from_map and to_map)',
            'n_det': n_det, 'n_samp': n_t, 'map_shape': [3, ny, nx],
            'pointing': 'pointing_matrix and its plan computed once before '
                        f'timing ({t_pmat * 1e3:.1f} ms + {t_plan * 1e3:.1f} ms)',
        },
        'ms_from_map': ms_from,
        'ms_to_map': ms_to,
        'ms_from_map_without_plan': ms_from_direct,
        'ms_to_map_without_plan': ms_to_direct,
        'plan': {'n_on': n_on, 'runs': n_runs, 'stored_spin_per_sample': ns,
                 't_block': pinfo['t_block'], 'bytes': pinfo['bytes'], 'visited_pixels': n_vis,
                 'layout': pinfo['layout'], 'stretches': pinfo['stretches'],
                 'items': pinfo['items'], 'wide_groups': pinfo['wide_groups']},
        'roofline': {
            'bound': 'hbm', 'kernel': dom[0], 'achieved': dom[1] / 1e9,
            'peak': HBM_PEAK / 1e9, 'unit': 'GB/s', 'frac': dom[1] / HBM_PEAK,
            'traffic': dom[3], 'algorithmic_bytes': dom[2],
            'bytes_basis': 'the plan kernel\'s own storage (14 B per sample for the '
                           'gather, 10-16 B for the scatter)',
            'from_map_frac': gbs_from / HBM_PEAK, 'to_map_frac': gbs_to / HBM_PEAK,
            # The same times on the bytes of the cached pointing matrix the
            # ProjEng_Precomp interface streams (int32 pixel pair + 3 float32
            # factors + float32 signal = 24 B per sample): above 1 because
            # the plan re-encodes that matrix in fewer bytes.
            'precomp_equivalent': {
                'bytes_per_sample': 24,
                'bytes_per_launch': b_pc + 8 * 3 * npix,
                'from_map_frac': (b_pc + 8 * 3 * npix) / (ms_from * 1e-3) / HBM_PEAK,
                'to_map_frac': (b_pc + 16 * 3 * npix) / (ms_to * 1e-3) / HBM_PEAK,
                'pair_frac': (2 * b_pc + 24 * 3 * npix) / (ms_step * 1e-3) / HBM_PEAK},
            # The unplanned ProjEng_Precomp kernels on those 24 B per sample.
            'precomp_kernels': {
                'from_map_frac': (b_pc + 8 * 3 * npix) / (ms_from_direct * 1e-3) / HBM_PEAK,
                'to_map_frac': (b_pc + 16 * 3 * npix) / (ms_to_direct * 1e-3) / HBM_PEAK},
        },
        'plan_cost': {
            'pointing_matrix_ms': t_pmat * 1e3, 'plan_build_ms': t_plan * 1e3,
            'saved_ms_per_step': (ms_from_direct + ms_to_direct) - (ms_from + ms_to),
            'break_even_steps': t_plan * 1e3 / max(1e-9, (ms_from_direct + ms_to_direct) -
                                                  (ms_from + ms_to))},
        'precomp_api': precomp_api,
        'on_the_fly': otf,
        'small_config': small,
        'nmat_detvecs_apply': nmat,
        'g3super_decode': g3,
        'c2_variants': variants,
        'c2_other_ops': other,
        'c3_precision': c3,
        'c4_batched': c4,
        'c5_mapmaking': c5,
        'cpu_baseline': cpu,
    }
    # numpy scalars (study outputs) are written as plain numbers
    print(json.dumps(out, default=lambda o: o.item() if hasattr(o, 'item') else str(o)),
          flush=True)


def precomp_api_timing(torch, so3g_amd, pix, spin, sky, sig, mapo, calls=5):
    """ProjEng_Precomp_NonTiled.from_map + to_map with the benchmark's torch
    tensors, wall time per pair (synchronised): the first pair includes
    building the cached plan; later pairs reuse it."""
    pe = so3g_amd.ProjEng_Precomp_NonTiled()
    ts = []
    for _ in range(calls):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pe.from_map(sky, pix, spin)
        mapo.zero_()
        pe.to_map(mapo, pix, spin, sig)
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    pe.clear_plan()
    log(f'ProjEng_Precomp pair: {ts}')
    return {'ms_first_pair': ts[0], 'ms_pair': float(np.median(ts[1:])),
            'det_samples_per_s': pix.shape[0] * pix.shape[1] / (np.median(ts[1:]) * 1e-3)}


def c3_precision():
    """Precision study at its configuration: 100 det x 40000 samples into ZEA
    and TAN 0.25 arcmin and HEALPix RING nside 8192 maps, to_map and
    to_weights with float32 products into float64 (so3g's arithmetic) and
    all-float32, errors relative to float64 products: the largest max and
    rms relative error over the outputs of each case."""
    from so3g_amd.studies import precision
    t0 = time.time()
    r = precision.run(n_det=100, n_samp=40000, res_arcmin=0.25, nside=8192)
    out = {'workload': r['workload'], 'seconds': time.time() - t0, 'cases': {}}
    for cname, c in r['cases'].items():
        row = {'hit_pixels': c['hit_pixels']}
        for ar in ('ref', 'f32'):
            for kind in ('map', 'weights'):
                es = list(c[ar][kind].values())
                row[f'{ar}_{kind}_max_rel'] = max(e['max_rel'] for e in es)
                row[f'{ar}_{kind}_rms_rel'] = max(e['rms_rel'] for e in es)
        out['cases'][cname] = row
    log(f'C3 precision: {out}')
    return out


def c4_batched():
    """Batched observations at their configuration: 64 uneven observations
    (seed 4, 8.9e9 det*samples) into one tiled CAR 0.5 arcmin TQU map; the
    autotuner picks each op's launch shape (detectors per work item, blocks
    per compute unit, and for to_map global atomics or an LDS table of 512 /
    1024 slots per item), then each op is timed as engine calls per
    observation, one batch-kernel launch per observation, one launch over
    the batch, and a HIP-graph replay of that launch (median of 3)."""
    from so3g_amd.studies import batched
    r = batched.run(n_obs=64, seed=4, reps=3)
    out = {k: r[k] for k in ('workload', 'det_samples', 'setup_s', 'tune_s',
                             'to_map_best_by_table', 'from_map_bit_exact',
                             'to_map_max_rel_diff')}
    out['shape'] = r['shape']
    out['timing_ms'] = {op: {m: v['s'] * 1e3 for m, v in modes.items()}
                        for op, modes in r['timing'].items()}
    out['det_samples_per_s'] = {op: {m: v['det_samples_per_s'] for m, v in modes.items()}
                                for op, modes in r['timing'].items()}
    for op, modes in r['timing'].items():
        out[f'{op}_one_launch_vs_per_obs'] = modes['per_obs']['s'] / modes['eager']['s']
        out[f'{op}_graph_vs_per_obs'] = modes['per_obs']['s'] / modes['graph']['s']
        out[f'{op}_one_launch_vs_api'] = modes['api']['s'] / modes['eager']['s']
    log(f'C4 batched: {out}')
    return out


def c5_mapmaking():
    """Map-making at its configuration (100 det x 40000 samples, CAR 1
    arcmin TQU, 1/f noise, Gaussian random sky): wall time of the solve to
    1e-6 relative residual with binned-preconditioned CG (capped at 40000
    iterations), two-level CG, two-level CG with the learned deflation space,
    and flexible CG with the U-Net correction on top of that (falling back
    to it when it stalls); setup, learning and training times listed apart."""
    from so3g_amd.studies import cg
    r = cg.run()
    keep = ('iters', 'seconds', 'final_residual', 'true_residual', 'fallback_at',
            'map_rms_error_TQU', 'a_norm_error_vs_reference', 'map_rms_diff_vs_reference_TQU',
            'ms_per_iter', 'setup_s', 'train_s', 'pairs_s', 'train_steps', 'pairs',
            'loss_first_last', 'learn_s', 'harvest_iters', 'ritz_vectors', 'ritz_values_max')
    out = {'workload': r['workload'], 'reference': r.get('reference')}
    for k in ('binned', 'two_level', 'deflated', 'learned'):
        if k in r:
            out[k] = {kk: v for kk, v in r[k].items() if kk in keep}
            out[k]['converged'] = bool(r[k]['final_residual'] <= 1e-6)
    for k in ('speedup_two_level', 'speedup_deflated_vs_binned', 'speedup_deflated_vs_two_level',
              'speedup_learned_vs_binned', 'speedup_learned_vs_two_level',
              'speedup_learned_vs_deflated'):
        if k in r:
            out[k] = r[k]
    log(f'C5 map-making: {out}')
    return out


def cpu_baseline(obs, pix, spin, sig, sky, pixargs):
    """The oracle (a C++ port of so3g's loops, g++ -O3, no -march) on the
    host cores, on a bounded sample of the same workload: 200 detectors of
    the cached-pointing pair (from_map OpenMP over detectors, to_map single
    threaded as so3g runs it without a thread assignment)."""
    try:
        import oracle
    except Exception as e:  # pragma: no cover
        log(f'cpu baseline unavailable: {e}')
        return None
    log('timing CPU baseline (oracle) ...')
    n = min(200, obs.n_det)
    p = pix[:n].cpu().numpy()
    s = spin[:n].cpu().numpy()
    sg = sig[:n].cpu().numpy()
    m = sky.cpu().numpy()
    og = oracle.geom_from_args('CAR', 'TQU', pixargs)
    oracle.precomp_from_map(og, m, p[:2], s[:2])  # warm the library
    t0 = time.perf_counter()
    oracle.precomp_from_map(og, m, p, s)
    t1 = time.perf_counter()
    out = np.zeros_like(m)
    oracle.precomp_to_map(og, out, p, s, sg)
    t2 = time.perf_counter()
    n_s = n * obs.n_samp
    return {
        'value': n_s / (t2 - t0), 'unit': 'det*samples/s', 'cores': oracle.num_threads(),
        'kind': 'port',
        'sample': f'{n} of {obs.n_det} detectors x {obs.n_samp} samples, cached-pointing '
                  f'from_map (OpenMP, {oracle.num_threads()} threads: {t1 - t0:.2f} s) + '
                  f'to_map (1 thread: {t2 - t1:.2f} s)',
    }


def small_config(torch, so3g_amd, N, workloads, dev, stream, reps=20):
    """The small observation (100 det x 40000 samples, CAR 1 arcmin TQU): the
    planned from_map + to_map pair, warm (everything resident in the 256 MB
    Infinity Cache) and cold (a 1 GiB buffer streamed before each pass to
    flush it).  Times are HIP-event times of the pair on the stream."""
    lib = N.load()
    obs = workloads.make('small')
    eng = so3g_amd.ProjEng_CAR_TQU_NonTiled(obs.pixelizor_args())
    ny, nx = obs.shape
    n_det, n_t = obs.n_det, obs.n_samp
    bore = torch.from_numpy(obs.bore).to(dev)
    dets = torch.from_numpy(obs.dets).to(dev)
    resp = torch.from_numpy(obs.resps).to(dev)
    pix, spin = eng.pointing_matrix(bore, dets, resp)
    plan = so3g_amd.PrecompPlan(pix, spin, (ny, nx))
    gen = torch.Generator(device=dev).manual_seed(3)
    sky = torch.randn((3, ny, nx), dtype=torch.float64, device=dev, generator=gen)
    sig = torch.empty((n_det, n_t), dtype=torch.float32, device=dev)
    m = torch.zeros((3, ny, nx), dtype=torch.float64, device=dev)
    flush_a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    flush_b = torch.empty_like(flush_a)

    def pair():
        # from_map into a fresh signal, to_map into a fresh map (the plan
        # writes the whole map: to_map_set)
        N.check(lib.so3g_pc_plan_from_map(plan._h, sky.data_ptr(), sig.data_ptr(), 0, stream))
        N.check(lib.so3g_pc_plan_to_map_set(plan._h, m.data_ptr(), sig.data_ptr(), None, stream))

    def timed(cold):
        ts = []
        for _ in range(reps):
            if cold:
                flush_b.copy_(flush_a)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            pair()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))

    pair()
    torch.cuda.synchronize()
    warm, cold = timed(False), timed(True)
    n_s = n_det * n_t
    npix = ny * nx
    info = plan.info()
    bf, bt, _ = plan_bytes(info, n_s)
    b = bf + bt
    b_pc = n_s * 24 * 2 + 3 * 8 * npix * 3
    out = {'workload': f'{n_det} det x {n_t} samples, CAR 1 arcmin TQU {ny}x{nx}, planned '
                       'from_map (fresh signal) + to_map (fresh map)',
           'plan_layout': info['layout'], 'plan_items': info['items'],
           'ms_pair_warm': warm, 'ms_pair_cold': cold,
           'det_samples_per_s_warm': n_s / (warm * 1e-3),
           'det_samples_per_s_cold': n_s / (cold * 1e-3),
           'hbm_frac_cold': b / (cold * 1e-3) / HBM_PEAK,
           'hbm_frac_cold_precomp_equivalent': b_pc / (cold * 1e-3) / HBM_PEAK,
           'median_of': reps}
    del flush_a, flush_b
    log(f'small config: warm {warm:.3f} ms, cold {cold:.3f} ms per pair')
    return out


def tile_scatter_timing(torch, so3g_amd, workloads, dev, n_det, n_t, reps=5):
    """The tiled plan's to_map (k_tp_to_map: work items summed in a map tile
    held in LDS) on the benchmark observation's CAR 0.5 arcmin map and on a
    HEALPix NEST nside 4096 map, ms per call (the output map is not zeroed
    in the timed calls)."""
    obs = workloads.Observation(n_det, n_t, 1.5, 0.5, seed=0)
    bore = torch.from_numpy(obs.bore).to(dev)
    dets = torch.from_numpy(obs.dets).to(dev)
    resp = torch.from_numpy(obs.resps).to(dev)
    g = torch.Generator(device=dev).manual_seed(2)
    sig = torch.randn((n_det, n_t), dtype=torch.float32, device=dev, generator=g)
    out = {'workload': f'{n_det} det x {n_t} samples, tiled plans (layout 1)'}
    car = so3g_amd.ProjEng_CAR_TQU_NonTiled(tuple(obs.projectionist()._get_pixelizor_args()))
    hp = so3g_amd.ProjEng_HP_TQU_NonTiled((4096, True))
    for name, eng, shape in (('car', car, tuple(obs.shape)), ('healpix_nest_4096', hp,
                                                             (12 * 4096 ** 2,))):
        pix, spin = eng.pointing_matrix(bore, dets, resp)
        plan = so3g_amd.PrecompPlan(pix, spin, shape, layout=1)
        del pix, spin
        m = torch.zeros((3,) + shape, dtype=torch.float64, device=dev)
        plan.to_map(m, sig)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            plan.to_map(m, sig)
        e1.record()
        torch.cuda.synchronize()
        out[name] = {'ms_to_map': e0.elapsed_time(e1) / reps, 'items': plan.info()['items']}
        del plan, m
        torch.cuda.empty_cache()
    return out


def nmat_timing(torch, so3g_amd, dev, n_det, n_t, nvec=10, nbin=20, reps=5):
    """N^-1 of the detector-vector noise model (nmat_detvecs_apply) on the
    Fourier transform of the benchmark's time streams: ft (n_det, 2 (n_t/2+1))
    float32, nbin log-spaced bins, nvec vectors.  Bytes: ft read twice and
    written once."""
    nfreq = n_t // 2 + 1
    nmode = 2 * nfreq
    edges = np.unique(np.round(np.geomspace(1, nfreq, nbin + 1)).astype(int))
    edges[0] = 0
    bins = np.stack([edges[:-1], edges[1:]], 1).astype(np.int32)
    g = torch.Generator(device=dev).manual_seed(4)
    ft = torch.randn((n_det, nmode), dtype=torch.float32, device=dev, generator=g)
    iD = torch.rand((len(bins), n_det), device=dev, generator=g) + 0.5
    iV = torch.randn((len(bins), n_det, nvec), device=dev, generator=g) / np.sqrt(n_det)
    so3g_amd.nmat_detvecs_apply(ft, bins, iD, iV, 1.0, 1.0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        so3g_amd.nmat_detvecs_apply(ft, bins, iD, iV, 1.0, 1.0)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    b = 3 * 4 * n_det * nmode
    del ft
    log(f'nmat_detvecs_apply: {ms:.3f} ms')
    return {'shape': [n_det, nmode], 'nvec': nvec, 'nbin': len(bins), 'ms': ms,
            'hbm_frac': b / (ms * 1e-3) / HBM_PEAK}


def c2_other_ops(torch, N, lib, g, bore, dets, resp, pix, spin, plan, ny, nx, dev, stream,
                 run, steps=3):
    """The other ops of the C1/C2 list on the benchmark observation (CAR 0.5
    arcmin TQU non-tiled): get_pixels and get_pointing_matrix (pointing
    recomputed; their outputs are 8 and 20 B per det*sample written), and
    to_weights on the fly, through the cached pointing matrix and through
    the plan (ms; the weights map is zeroed first).  Time only."""
    n_det, n_t = pix.shape[0], pix.shape[1]
    n_s = n_det * n_t
    pbuf = torch.empty_like(pix)
    wmap = torch.zeros((3, 3, ny, nx), dtype=torch.float64, device=dev)
    shape = np.array([ny, nx], np.int64)

    def chk(rc):
        N.check(rc)

    def get_pixels():
        chk(lib.so3g_pixels(g, bore.data_ptr(), n_t, dets.data_ptr(), n_det, pbuf.data_ptr(),
                            stream))

    def get_pmat():
        chk(lib.so3g_pointing_matrix(g, bore.data_ptr(), n_t, dets.data_ptr(), n_det,
                                     resp.data_ptr(), pix.data_ptr(), spin.data_ptr(), stream))

    def otf_weights():
        wmap.zero_()
        chk(lib.so3g_to_weights(g, wmap.data_ptr(), None, None, bore.data_ptr(), n_t,
                                dets.data_ptr(), n_det, resp.data_ptr(), None, None, 0, stream))

    def pc_weights():
        wmap.zero_()
        chk(lib.so3g_precomp_to_weights(wmap.data_ptr(), shape.ctypes.data, None, None,
                                        pix.data_ptr(), 2, spin.data_ptr(), 3, n_det, n_t, None,
                                        None, 0, stream))

    def plan_weights():
        wmap.zero_()
        chk(lib.so3g_pc_plan_to_weights(plan._h, wmap.data_ptr(), None, stream))
    _, ms = run([get_pixels, get_pmat, otf_weights, pc_weights, plan_weights], steps, 1)
    del pbuf, wmap
    out = {}
    for name, t, b in (('get_pixels', ms[0], 8), ('get_pointing_matrix', ms[1], 20),
                       ('to_weights_on_the_fly', ms[2], None),
                       ('to_weights_precomp', ms[3], None),
                       ('to_weights_plan', ms[4], None)):
        out[name] = {'ms': t, 'det_samples_per_s': n_s / (t * 1e-3)}
        if b:
            out[name]['hbm_frac_output'] = b * n_s / (t * 1e-3) / HBM_PEAK
    log(f'other ops: {out}')
    return out


def c2_variants(torch, so3g_amd, N, lib, obs, pixargs, bore, dets, resp, sig, dev, stream, run,
                steps=3):
    """The benchmark observation projected into the other maps of the C2
    case: a tiled CAR 0.5 arcmin TQU map (100 x 100 tiles, the tiles the
    observation hits) and a HEALPix NEST nside 4096 TQU map (the boresight
    taken as celestial).  Each is timed as the on-the-fly pair (pointing
    recomputed in both kernels) and the cached-pointing (ProjEng_Precomp)
    pair, and through a PrecompPlan (tiled map: tile pointer lists; HEALPix:
    pixel tables over the visited pixels only); ms per op and det*samples/s
    per pair."""
    n_det, n_t = obs.n_det, obs.n_samp
    N_s = n_det * n_t
    out = {}
    gen = torch.Generator(device=dev).manual_seed(5)
    sig_out = torch.empty_like(sig)

    def chk(rc):
        N.check(rc)

    # -- tiled CAR -------------------------------------------------------
    probe = so3g_amd.ProjEng_CAR_TQU_Tiled(tuple(pixargs) + (100, 100))
    hits = probe.tile_hits(bore, dets)
    active = [i for i, h in enumerate(hits) if h > 0]
    eng = so3g_amd.ProjEng_CAR_TQU_Tiled(tuple(pixargs) + (100, 100, active))
    g = eng._g
    dims = eng._tile_dims()
    sizes = [3 * int(dims[t, 0]) * int(dims[t, 1]) for t in active]
    buf_in = torch.randn(sum(sizes), dtype=torch.float64, device=dev, generator=gen)
    buf_out = torch.zeros(sum(sizes), dtype=torch.float64, device=dev)
    ptr_in = np.zeros(g.n_tile, np.int64)
    ptr_out = np.zeros(g.n_tile, np.int64)
    at = 0
    for t, sz in zip(active, sizes):
        ptr_in[t] = buf_in.data_ptr() + 8 * at
        ptr_out[t] = buf_out.data_ptr() + 8 * at
        at += sz
    d_in = torch.from_numpy(ptr_in).to(dev)
    d_out = torch.from_numpy(ptr_out).to(dev)
    d_dims = torch.from_numpy(np.ascontiguousarray(dims)).to(dev)
    tshape = np.array([g.n_tile, 1], np.int64)
    pix = torch.empty((n_det, n_t, 3), dtype=torch.int32, device=dev)
    spin = torch.empty((n_det, n_t, 3), dtype=torch.float32, device=dev)
    chk(lib.so3g_pointing_matrix(g, bore.data_ptr(), n_t, dets.data_ptr(), n_det,
                                 resp.data_ptr(), pix.data_ptr(), spin.data_ptr(), stream))

    def otf_from():
        chk(lib.so3g_from_map(g, None, d_in.data_ptr(), d_dims.data_ptr(), bore.data_ptr(), n_t,
                              dets.data_ptr(), n_det, resp.data_ptr(), sig_out.data_ptr(), 0,
                              stream))

    def otf_to():
        buf_out.zero_()
        chk(lib.so3g_to_map(g, None, d_out.data_ptr(), d_dims.data_ptr(), bore.data_ptr(), n_t,
                            dets.data_ptr(), n_det, resp.data_ptr(), sig.data_ptr(), None, None,
                            0, stream))

    def pc_from():
        chk(lib.so3g_precomp_from_map(None, tshape.ctypes.data, d_in.data_ptr(), d_dims.data_ptr(),
                                      pix.data_ptr(), 3, spin.data_ptr(), 3, n_det, n_t,
                                      sig_out.data_ptr(), 0, stream))

    def pc_to():
        buf_out.zero_()
        chk(lib.so3g_precomp_to_map(None, tshape.ctypes.data, d_out.data_ptr(), d_dims.data_ptr(),
                                    pix.data_ptr(), 3, spin.data_ptr(), 3, n_det, n_t,
                                    sig.data_ptr(), None, None, 0, stream))
    torch.cuda.synchronize()
    t0 = time.time()
    tplan = so3g_amd.PrecompPlan(pix, spin, tile_dims=dims)
    torch.cuda.synchronize()
    t_tplan = time.time() - t0

    def plan_from():
        chk(lib.so3g_pc_plan_from_map_tiles(tplan._h, ptr_in.ctypes.data, sig_out.data_ptr(), 0,
                                            stream))

    def plan_to():
        buf_out.zero_()
        chk(lib.so3g_pc_plan_to_map_tiles(tplan._h, ptr_out.ctypes.data, sig.data_ptr(), None,
                                          stream))
    res = {'map': f'CAR 0.5 arcmin TQU {g.naxis[0]}x{g.naxis[1]}, 100x100 tiles, '
                  f'{len(active)} of {g.n_tile} active', 'plan_build_s': t_tplan}
    for name, ops in (('on_the_fly', [otf_from, otf_to]), ('precomp', [pc_from, pc_to]),
                      ('plan', [plan_from, plan_to])):
        _, (mf, mt) = run(ops, steps, 1)
        res[name] = {'ms_from_map': mf, 'ms_to_map': mt, 'det_samples_per_s': N_s / ((mf + mt) * 1e-3)}
    out['tiled_car'] = res
    del pix, spin, buf_in, buf_out, tplan
    log(f'tiled CAR: {res}')

    # -- HEALPix NEST nside 4096 ------------------------------------------
    nside = 4096
    eng = so3g_amd.ProjEng_HP_TQU_NonTiled((nside, True))
    g = eng._g
    npix = 12 * nside * nside
    hmap = torch.randn((3, npix), dtype=torch.float64, device=dev, generator=gen)
    hout = torch.zeros((3, npix), dtype=torch.float64, device=dev)
    shape = np.array([npix, 1], np.int64)
    pix = torch.empty((n_det, n_t, 1), dtype=torch.int32, device=dev)
    spin = torch.empty((n_det, n_t, 3), dtype=torch.float32, device=dev)
    chk(lib.so3g_pointing_matrix(g, bore.data_ptr(), n_t, dets.data_ptr(), n_det,
                                 resp.data_ptr(), pix.data_ptr(), spin.data_ptr(), stream))

    def hotf_from():
        chk(lib.so3g_from_map(g, hmap.data_ptr(), None, None, bore.data_ptr(), n_t,
                              dets.data_ptr(), n_det, resp.data_ptr(), sig_out.data_ptr(), 0,
                              stream))

    def hotf_to():
        hout.zero_()
        chk(lib.so3g_to_map(g, hout.data_ptr(), None, None, bore.data_ptr(), n_t,
                            dets.data_ptr(), n_det, resp.data_ptr(), sig.data_ptr(), None, None,
                            0, stream))

    def hpc_from():
        chk(lib.so3g_precomp_from_map(hmap.data_ptr(), shape.ctypes.data, None, None,
                                      pix.data_ptr(), 1, spin.data_ptr(), 3, n_det, n_t,
                                      sig_out.data_ptr(), 0, stream))

    def hpc_to():
        hout.zero_()
        chk(lib.so3g_precomp_to_map(hout.data_ptr(), shape.ctypes.data, None, None,
                                    pix.data_ptr(), 1, spin.data_ptr(), 3, n_det, n_t,
                                    sig.data_ptr(), None, None, 0, stream))
    torch.cuda.synchronize()
    t0 = time.time()
    hplan = so3g_amd.PrecompPlan(pix, spin, (npix,))
    torch.cuda.synchronize()
    t_hplan = time.time() - t0

    def hplan_from():
        chk(lib.so3g_pc_plan_from_map(hplan._h, hmap.data_ptr(), sig_out.data_ptr(), 0, stream))

    def hplan_to():
        hout.zero_()
        chk(lib.so3g_pc_plan_to_map(hplan._h, hout.data_ptr(), sig.data_ptr(), None, stream))
    res = {'map': f'HEALPix NEST nside {nside} TQU ({npix} pixels, 4.8 GB per map)',
           'plan_visited_pixels': hplan.info()['n_vis'], 'plan_build_s': t_hplan}
    # every to_map above zeroes the 4.8 GB output map first; that alone:
    _, (res['ms_map_zero'],) = run([lambda: hout.zero_()], steps, 1)
    for name, ops in (('on_the_fly', [hotf_from, hotf_to]), ('precomp', [hpc_from, hpc_to]),
                      ('plan', [hplan_from, hplan_to])):
        _, (mf, mt) = run(ops, steps, 1)
        res[name] = {'ms_from_map': mf, 'ms_to_map': mt, 'det_samples_per_s': N_s / ((mf + mt) * 1e-3)}
    out['healpix_nest_4096'] = res
    del pix, spin, hmap, hout, hplan
    log(f'HEALPix: {res}')
    return out


def g3super_timing(torch, n_det, n_t, with_cpu, reps=3, quantum=1e-3):
    """G3SuperTimestream.extract of the benchmark's observation size from an
    encoded blob resident on the GPU into a GPU array: float32 channels, one
    quantum each, white noise of 300 quanta on a slow sinusoid (FLAC level
    1, the reference's default; ~1.37 bytes per sample).  The time is the
    whole extract call (host walk of the blob, sync scan, frame decoding).
    CPU baseline: the oracle's decoder (sequential, RFC 9639 restatement)
    over 32 channels on a 16-thread pool."""
    import concurrent.futures
    from so3g_amd import g3super as g3
    rng = np.random.default_rng(11)
    t = np.arange(n_t)
    d = np.empty((n_det, n_t), np.float32)
    for i in range(n_det):
        d[i] = np.round(rng.normal(0, 300, n_t) + 1e4 * np.sin(t * 1e-3 + i)) * quantum
    q = np.full(n_det, quantum)
    ts = g3.G3SuperTimestream([f'c{i}' for i in range(n_det)], t, None, q)
    ts.data = d
    t0 = time.time()
    ts.encode()
    t_enc = time.time() - t0
    nbytes = len(ts.payload)
    dest = torch.empty((n_det, n_t), dtype=torch.float32, device='cuda')
    ts.extract(dest)
    torch.cuda.synchronize()
    stats = g3.last_decode_stats()
    # spot check against the quantised input
    y, r, _ = g3.split(d[:4], q[:4])
    want = ((r + y).astype(np.int32).astype(np.float32).astype(np.float64) * quantum
            ).astype(np.float32)
    exact = bool(np.array_equal(dest[:4].cpu().numpy(), want))
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.time()
        ts.extract(dest)
        torch.cuda.synchronize()
        times.append(time.time() - t0)
    ms = float(np.median(times)) * 1e3
    n = n_det * n_t
    out = {'workload': f'{n_det} float32 channels x {n_t} samples, FLAC level 1 '
                       f'(1152-sample frames), {nbytes / n:.3f} bytes per sample',
           'ms': ms, 'samples_per_s': n / (ms * 1e-3),
           'compressed_GBps': nbytes / (ms * 1e-3) / 1e9,
           'decoded_GBps': 4 * n / (ms * 1e-3) / 1e9,
           'frames': stats['frames'], 'walked_channels': stats['serial'],
           'spot_check_exact': exact, 'encode_s_host': t_enc, 'median_of': reps}
    log(f'g3super decode: {ms:.3f} ms, {n / (ms * 1e-3):.3g} samples/s')
    if with_cpu:
        import oracle
        k = min(32, n_det)
        offs = ts.offsets

        def one(i):
            o = np.zeros((1, n_t), np.float32)
            oracle.g3s_extract(ts.payload, [offs[i], offs[i + 1]], 'float32', n_t, q[i:i + 1], o)
            return o
        with concurrent.futures.ThreadPoolExecutor(16) as ex:
            t0 = time.time()
            list(ex.map(one, range(k)))
            dt = time.time() - t0
        out['cpu_baseline'] = {'value': k * n_t / dt, 'unit': 'samples/s', 'cores': 16,
                               'kind': 'port',
                               'sample': f'{k} of {n_det} channels, oracle decoder, '
                                         f'16 threads: {dt:.2f} s'}
    del dest
    return out


def _stripe_segments(dom, n_thread):
    """Per-thread (det, t0, t1) segment lists from a per-sample domain array."""
    segs = [[] for _ in range(n_thread)]
    for d, row in enumerate(dom):
        edges = np.flatnonzero(np.diff(row)) + 1
        starts = np.concatenate([[0], edges])
        ends = np.concatenate([edges, [len(row)]])
        for a, b in zip(starts, ends):
            k = int(row[a])
            if 0 <= k < n_thread:
                segs[k].append((d, int(a), int(b)))
    return [np.asarray(x, np.int32).reshape(-1, 3) for x in segs]


def cpu_full(obs, pix, spin, sig, sky, pixargs, n_det=50, reps=5):
    """The oracle on the host cores for every operation of the path, on the
    first n_det detectors of the workload: 1 warm-up, median of reps runs.
    'N threads' = OpenMP over detectors (from_map) or over pixel-stripe thread
    domains (to_map, so3g's 'simple' assignment); to_map without a thread
    assignment and to_weights run single-threaded, as so3g does."""
    import platform
    import oracle
    n = min(n_det, obs.n_det)
    bore = obs.bore
    d, r = obs.dets[:n], obs.resps[:n]
    p = pix[:n].cpu().numpy()
    sp = spin[:n].cpu().numpy()
    sg = sig[:n].cpu().numpy()
    m = sky.cpu().numpy()
    og = oracle.geom_from_args('CAR', 'TQU', pixargs)
    nthr = oracle.num_threads()
    dom = oracle.sample_domains(og, bore, d, 0, nthr)
    segs = _stripe_segments(dom, nthr)

    def med(f):
        f()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    z = lambda *lead: np.zeros(lead + m.shape[1:])
    ops = {
        'on-the-fly from_map (N threads)': lambda: oracle.from_map(og, m, bore, d, r),
        'on-the-fly to_map (1 thread)': lambda: oracle.to_map(og, z(3), bore, d, r, sg),
        'on-the-fly to_map (N threads, pixel stripes)':
            lambda: oracle.to_map_threaded(og, z(3), bore, d, r, sg, segs),
        'on-the-fly to_weights (1 thread)': lambda: oracle.to_weights(og, z(3, 3), bore, d, r),
        'precomp from_map (N threads)': lambda: oracle.precomp_from_map(og, m, p, sp),
        'precomp to_map (1 thread)': lambda: oracle.precomp_to_map(og, z(3), p, sp, sg),
        'precomp to_weights (1 thread)': lambda: oracle.precomp_to_weights(og, z(3, 3), p, sp),
    }
    n_s = n * obs.n_samp
    rates = {}
    for k, f in ops.items():
        t = med(f)
        rates[k] = n_s / t
        log(f'{k}: {t:.3f} s, {n_s / t:.3g} det*samples/s')
    model = ''
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        model = platform.processor()
    return {'cpu': model, 'threads': nthr, 'sample': f'{n} detectors x {obs.n_samp} samples',
            'runs': f'1 warm-up, median of {reps}', 'unit': 'det*samples/s', 'rates': rates}


if __name__ == '__main__':
    main()
