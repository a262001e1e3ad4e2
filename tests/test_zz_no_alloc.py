"""so3g_precomp_from_map_noalloc (the allocation-free entry point) against
so3g_precomp_from_map through the C ABI: bit-identical signals on two calls
in a row, the second with a smaller map and fewer samples (so the shared
scratch holds the first call's larger packed map), in both gather paths
(n_t % 4 == 0 packed, odd n_t generic), fresh and accumulating; a map
beyond the scratch raises ValueError and leaves the signal untouched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(torch, dev, n_det, n_t, ny, nx, seed):
    g = torch.Generator(device=dev).manual_seed(seed)
    pix = torch.stack([torch.randint(-1, ny, (n_det, n_t), generator=g, device=dev),
                       torch.randint(0, nx, (n_det, n_t), generator=g, device=dev)],
                      -1).to(torch.int32).contiguous()
    pix[..., 1][pix[..., 0] < 0] = 0
    spin = torch.randn((n_det, n_t, 3), generator=g, device=dev).float().contiguous()
    sky = torch.randn((3, ny, nx), generator=g, device=dev, dtype=torch.float64)
    base = torch.randn((n_det, n_t), generator=g, device=dev).float()
    return pix, spin, sky, base


def test_noalloc_matches_precomp_from_map(gpu):
    import torch
    from so3g_amd import _native as N
    lib = N.load()
    dev = gpu
    for n_det, n_t, ny, nx in ((37, 4000, 300, 2000), (5, 1200, 20, 30), (9, 1001, 40, 50)):
        pix, spin, sky, base = _case(torch, dev, n_det, n_t, ny, nx, n_t)
        shape = np.array([ny, nx], np.int64)
        for acc in (0, 1):
            a = base.clone()
            b = base.clone()
            N.check(lib.so3g_precomp_from_map(sky.data_ptr(), shape.ctypes.data, None, None,
                                              pix.data_ptr(), 2, spin.data_ptr(), 3, n_det,
                                              n_t, a.data_ptr(), acc, None))
            N.check(lib.so3g_precomp_from_map_noalloc(sky.data_ptr(), shape.ctypes.data,
                                                      pix.data_ptr(), spin.data_ptr(), 3,
                                                      n_det, n_t, b.data_ptr(), acc, None))
            torch.cuda.synchronize()
            assert torch.equal(a.view(torch.int32), b.view(torch.int32)), (n_det, n_t, acc)


def test_noalloc_rejects_large_map(gpu):
    import torch
    from so3g_amd import _native as N
    lib = N.load()
    ny, nx = 1025, 1024          # 1049600 pixels > 1048576
    pix = torch.zeros((2, 8, 2), dtype=torch.int32, device=gpu)
    spin = torch.ones((2, 8, 3), dtype=torch.float32, device=gpu)
    sky = torch.zeros((3, ny, nx), dtype=torch.float64, device=gpu)
    sig = torch.full((2, 8), 7.0, device=gpu)
    shape = np.array([ny, nx], np.int64)
    with pytest.raises(ValueError, match='scratch'):
        N.check(lib.so3g_precomp_from_map_noalloc(sky.data_ptr(), shape.ctypes.data,
                                                  pix.data_ptr(), spin.data_ptr(), 3, 2, 8,
                                                  sig.data_ptr(), 0, None))
    assert bool((sig == 7.0).all())
