"""Local statistics (SpatialOps.cpp:72-146): the oracle restatement pinned to
the reference test's known answers and to the compiled reference; the GPU
kernel (rawops_local_statistics) bit-exact against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref as R


def region(output, image, r):
    return O.grow_rect(output, image, r)


def test_oracle_affine_truth():
    """spatial_ops_tests.cpp:49-62: R = x - 12, G = y - 10, B = 2 over image
    (10, 7, 7, 5): window means and population variances have closed forms."""
    image = (10, 7, 7, 5)
    ys, xs = np.mgrid[7:12, 10:17]
    tile = np.stack([xs - 12.0, ys - 10.0, np.full(xs.shape, 2.0)], -1).astype(np.float32)
    for radius in (0, 1, 3, 8):
        mean, var, cnt = O.local_statistics(tile, image, image, image, radius)
        for y in range(5):
            for x in range(7):
                left, right = max(0, x - radius), min(6, x + radius)
                top, bottom = max(0, y - radius), min(4, y + radius)
                nx, ny = right - left + 1, bottom - top + 1
                assert abs(mean[y, x, 0] - ((left + right) / 2 - 2)) < 1e-12
                assert abs(mean[y, x, 1] - ((top + bottom) / 2 - 3)) < 1e-12
                assert abs(var[y, x, 0] - (nx * nx - 1) / 12) < 1e-12
                assert abs(var[y, x, 1] - (ny * ny - 1) / 12) < 1e-12
                assert mean[y, x, 2] == 2 and var[y, x, 2] == 0 and cnt[y, x, 0] == nx * ny


def test_oracle_stability_and_nonfinite():
    """spatial_ops_tests.cpp:88-99: no cancellation at a 1e20 offset; non-finite
    samples are excluded per channel and an empty window has NaN moments."""
    base = np.float32(1e20)
    nxt = np.nextafter(base, np.float32(np.inf))
    delta = float(nxt) - float(base)
    tile = np.array([[[base] * 3, [nxt] * 3, [base] * 3]], np.float32)
    _, var, _ = O.local_statistics(tile, (0, 0, 3, 1), (1, 0, 1, 1), (0, 0, 3, 1), 1)
    exp = delta * delta * 2 / 9
    assert abs(var[0, 0, 0] - exp) / exp < 1e-8
    tile = np.array([[[-np.inf, 0, 1], [np.inf, 2, 1], [np.nan, 4, 1]]], np.float32)
    mean, var, cnt = O.local_statistics(tile, (0, 0, 3, 1), (1, 0, 1, 1), (0, 0, 3, 1), 1)
    assert cnt[0, 0, 0] == 0 and np.isnan(mean[0, 0, 0]) and np.isnan(var[0, 0, 0])
    assert cnt[0, 0, 1] == 3 and mean[0, 0, 1] == 2


def _case(seed, w, h, image, output, radius, special=True):
    rng = np.random.default_rng(seed)
    needed = region(output, image, radius)
    tile = rng.lognormal(-1.0, 2.0, (needed[3], needed[2], 3)).astype(np.float32)
    tile *= rng.choice([-1.0, 1.0], tile.shape).astype(np.float32)
    if special:
        flat = tile.reshape(-1)
        idx = rng.choice(flat.size, flat.size // 50, replace=False)
        flat[idx[0::3]] = np.nan
        flat[idx[1::3]] = np.inf
        flat[idx[2::3]] = -0.0
    return tile, needed


@pytest.mark.skipif(not R.available(), reason="compiled reference not built here")
@pytest.mark.parametrize("radius", [0, 1, 2, 3, 5, 8])
def test_oracle_matches_compiled_reference(radius):
    image = (3, 2, 57, 41)
    for k, output in enumerate([(3, 2, 57, 41), (20, 10, 13, 9), (59, 42, 1, 1), (3, 30, 57, 13)]):
        tile, needed = _case(10 * radius + k, 0, 0, image, output, radius)
        a = O.local_statistics(tile, needed, output, image, radius)
        b = R.local_statistics(tile, needed, output, image, radius)
        for x, y in zip(a, b):
            assert x.tobytes() == y.tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize("radius", [0, 1, 3, 8])
def test_gpu_kernel_bit_exact(radius):
    import torch
    from librawops_amd import _capi as C
    image = (0, 0, 700, 300)
    for k, output in enumerate([(0, 0, 700, 300), (33, 17, 400, 151), (699, 299, 1, 1)]):
        tile, needed = _case(100 * radius + k, 0, 0, image, output, radius)
        exp = O.local_statistics(tile, needed, output, image, radius)
        got = C.local_statistics(torch.from_numpy(tile).cuda(), needed, output, image, radius)
        for g, e in zip(got, exp):
            assert g.cpu().numpy().view(e.dtype).tobytes() == e.tobytes()
