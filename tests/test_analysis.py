"""histogram_rgb (ImageAnalysis.cpp:10-80).  CPU: the oracle restatement
against the known answers of the reference's own test
(tests/image_analysis_tests.cpp:44-114: bin edges over a site-residue probe,
nonfinite values, a single bin holding both endpoints, extreme and tiny
ranges).  GPU: the device histogram (rawops_histogram_rgb) against the oracle
bit for bit on random images with outliers and nonfinite samples over bin
counts 1..65536 (LDS and global counting), banded accumulation, and the
Python API over an ImageGraph (per-channel coverage, option validation)."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def probe_image(bounds, viewport):
    """The reference test's probe node: R = (x % 9) 0.25 - 0.5, G = (y % 9) 0.25 - 0.5, B = NaN."""
    x0, y0, w, h = viewport
    ys, xs = np.mgrid[y0:y0 + h, x0:x0 + w]
    img = np.empty((h, w, 3), np.float32)
    img[..., 0] = (xs % 9).astype(np.float32) * np.float32(0.25) - np.float32(0.5)
    img[..., 1] = (ys % 9).astype(np.float32) * np.float32(0.25) - np.float32(0.5)
    img[..., 2] = np.nan
    return img


def test_oracle_edge_bins():
    # image_analysis_tests.cpp:44-68: viewport (9, 7, 19, 13), 4 bins of [0, 1]
    img = probe_image((7, 5, 31, 23), (9, 7, 19, 13))
    res = O.histogram_rgb(img, 4, 0.0, 1.0)
    for c in range(2):
        expected = [0, 0, 0, 0]
        below = above = 0
        for y in range(7, 20):
            for x in range(9, 28):
                site = (x if c == 0 else y) % 9
                if site < 2:
                    below += 1
                elif site > 6:
                    above += 1
                else:
                    expected[3 if site == 6 else site - 2] += 1
        counts, under, over, nonfinite, mn, mx = res[c]
        assert list(counts) == expected and under == below and over == above and nonfinite == 0
        assert mn == -0.5 and mx == 1.5
    counts, under, over, nonfinite, mn, mx = res[2]
    assert nonfinite == 247 and mn is None and mx is None and counts.sum() == 0


def test_oracle_nonfinite_and_extreme_ranges():
    # image_analysis_tests.cpp:81-96
    inf, nan = np.inf, np.nan
    img = np.array([[-inf, 0, -2], [inf, 1, 2], [nan, 0.5, 0]], np.float32).reshape(1, 3, 3)
    res = O.histogram_rgb(img, 1, 0.0, 1.0)
    assert res[0][3] == 3 and res[0][4] is None
    assert res[1][0][0] == 3 and res[2][1] == 1 and res[2][2] == 1
    res = O.histogram_rgb(img, 65536, -1e300, 1e300)
    assert res[1][0][32768] == 3
    res = O.histogram_rgb(img, 2, 0.0, 5e-324)
    assert res[1][0][0] == 1 and res[1][2] == 2


def test_oracle_planted_mistake_fails():
    # an off-by-one bin rule (floor of (v - lower) / span * (bins - 1)) must miss the reference's edge counts
    img = probe_image((7, 5, 31, 23), (9, 7, 19, 13))
    d = img[..., 0].astype(np.float64).reshape(-1)
    inside = d[(d >= 0) & (d <= 1)]
    wrong = np.bincount(np.minimum(3, (inside * 3).astype(np.uint32)), minlength=4)
    assert list(wrong) != list(O.histogram_rgb(img, 4, 0.0, 1.0)[0][0])


def random_image(h, w, seed):
    rng = np.random.default_rng(seed)
    img = rng.normal(0.4, 0.5, (h, w, 3)).astype(np.float32)
    m = rng.random((h, w, 3))
    img[m < 0.01] = np.nan
    img[(m >= 0.01) & (m < 0.015)] = np.inf
    img[(m >= 0.015) & (m < 0.02)] = -np.inf
    img[(m >= 0.02) & (m < 0.03)] = 1.0      # the closed last bin
    img[(m >= 0.03) & (m < 0.04)] = 0.0      # the first bin's lower edge
    return img


def device_histogram(img, bins, lower, upper, bands=1):
    import torch
    from librawops_amd import _capi as C
    h, w = img.shape[:2]
    t = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    counts = torch.zeros(3 * bins, dtype=torch.int64, device="cuda")
    tallies = torch.zeros(9, dtype=torch.int64, device="cuda")
    extrema = torch.tensor([-1, 0] * 3, dtype=torch.int32, device="cuda")
    step = -(-h // bands)
    for y in range(0, h, step):
        rows = min(step, h - y)
        C.check(C.lib().rawops_histogram_rgb(t[y:y + rows].data_ptr(), w, rows, w, bins, lower, upper,
                                             counts.data_ptr(), tallies.data_ptr(), extrema.data_ptr(), None))
    torch.cuda.synchronize()
    from librawops_amd.analysis import _key_float
    c = counts.cpu().numpy().view(np.uint64).reshape(3, bins)
    tl = tallies.cpu().numpy().view(np.uint64).reshape(3, 3)
    e = extrema.cpu().numpy().view(np.uint32).reshape(3, 2)
    return [(c[k], int(tl[k, 0]), int(tl[k, 1]), int(tl[k, 2]),
             _key_float(int(e[k, 0])) if int(e[k, 0]) != 0xFFFFFFFF else None,
             _key_float(int(e[k, 1])) if int(e[k, 0]) != 0xFFFFFFFF else None) for k in range(3)]


@pytest.mark.gpu
@pytest.mark.parametrize("bins,lower,upper", [(1, 0.0, 1.0), (4, 0.0, 1.0), (256, -0.25, 1.5), (2048, 0.0, 1.0),
                                              (2049, 0.0, 1.0), (65536, -1e300, 1e300), (65536, 0.0, 1.0),
                                              (3, 0.0, 5e-324)])
def test_device_histogram_matches_oracle(bins, lower, upper):
    img = random_image(203, 157, seed=bins)
    want = O.histogram_rgb(img, bins, lower, upper)
    for bands in (1, 3):
        got = device_histogram(img, bins, lower, upper, bands)
        for k in range(3):
            assert np.array_equal(got[k][0], want[k][0]), (k, bands)
            assert got[k][1:] == want[k][1:], (k, bands)


@pytest.mark.gpu
@pytest.mark.parametrize("bins,lower,upper", [(256, 0.0, 1.0), (100, -0.3, 1.7), (1000, 0.1, 0.7), (7, -1.0, 2.5),
                                              (65536, 0.0, 1.0)])
def test_device_histogram_bin_edges(bins, lower, upper):
    """Samples on and one float32 step either side of every bin edge: the
    device bins them by a reciprocal product and divides only near edges, so
    these are the samples where the two could part."""
    k = np.arange(bins + 1, dtype=np.float64)
    edge = (lower + k * ((upper - lower) / bins)).astype(np.float32)
    vals = np.concatenate([edge, np.nextafter(edge, np.float32(-np.inf)), np.nextafter(edge, np.float32(np.inf))])
    w = 256
    n = -(-vals.size // w) * w
    flat = np.resize(vals, n)
    img = np.stack([flat, np.roll(flat, 1), np.roll(flat, 2)], axis=-1).reshape(n // w, w, 3)
    want = O.histogram_rgb(img, bins, lower, upper)
    got = device_histogram(img, bins, lower, upper)
    for c in range(3):
        assert np.array_equal(got[c][0], want[c][0]), c
        assert got[c][1:] == want[c][1:], c


@pytest.mark.gpu
@pytest.mark.parametrize("height,width", [(1500, 2000), (64, 4100), (4000, 50)])
def test_device_histogram_partial_rows(height, width):
    """Images of 64 rows or more at up to 2048 bins: each workgroup stores its
    bins, tallies and extremum keys into a row of a partial array and a second
    kernel sums the rows into the caller's counters.  1500x2000 at 256 bins
    takes the full 1024 workgroups (two rows each, so the last 274 hold no
    pixels) and 64 merge chunks; 64 rows is the fewest that take the path;
    4000x50 at 256 bins gives each of its 75 workgroups 54 narrow rows.  Two
    bands add into the same counters."""
    img = random_image(height, width, seed=height + width)
    for bins, lower, upper in ((256, 0.0, 1.0), (2048, -0.25, 1.5)):
        want = O.histogram_rgb(img, bins, lower, upper)
        for bands in (1, 2):
            got = device_histogram(img, bins, lower, upper, bands)
            for c in range(3):
                assert np.array_equal(got[c][0], want[c][0]), (bins, c, bands)
                assert got[c][1:] == want[c][1:], (bins, c, bands)


@pytest.mark.gpu
@pytest.mark.parametrize("width", [2049, 4100, 6001])
@pytest.mark.parametrize("bins,lower,upper", [(256, 0.0, 1.0), (1000, -0.3, 1.7), (4096, 0.0, 1.0)])
def test_device_histogram_wide_rows(width, bins, lower, upper):
    """Rows wider than one pass of the workgroup (256 threads x 4 pixel
    slots): every pixel slot p >= 1, column batches starting at 1024 and
    beyond, and a ragged last batch.  Bin-edge samples (on, and one float32
    step either side of, every edge) sit at columns 256 p + j, so each slot
    meets them."""
    img = random_image(7, width, seed=width + bins)
    k = np.arange(bins + 1, dtype=np.float64)
    edge = (lower + k * ((upper - lower) / bins)).astype(np.float32)
    vals = np.concatenate([edge, np.nextafter(edge, np.float32(-np.inf)), np.nextafter(edge, np.float32(np.inf))])
    cols = (256 * np.arange(width // 256 + 1)[:, None] + np.arange(8)[None, :]).reshape(-1)
    cols = cols[cols < width]
    rng = np.random.default_rng(bins)
    for r in range(img.shape[0]):
        for c in range(3):
            img[r, cols, c] = rng.choice(vals, cols.size)
    want = O.histogram_rgb(img, bins, lower, upper)
    for bands in (1, 2):
        got = device_histogram(img, bins, lower, upper, bands)
        for c in range(3):
            assert np.array_equal(got[c][0], want[c][0]), (c, bands)
            assert got[c][1:] == want[c][1:], (c, bands)


@pytest.mark.gpu
def test_histogram_rgb_api():
    import librawops_amd as L
    from tests.helpers import random_samples
    md = L.RawMetadata(64, 48, 64, L.BayerPattern(0), 0, 0, L.Rect(0, 0, 64, 48), (510,) * 4, (16383,) * 4)
    img = L.RawImage(md, random_samples(48, 64, seed=3, hi=16383))
    g = L.ImageGraph(img, L.GraphRecipe(exposure_stops=0.5), demosaic=L.MENON, linear_output=True)
    req = L.RenderRequest(L.Rect(5, 3, 50, 40), 16)
    h = L.histogram_rgb(g, req, L.HistogramOptions(64, 0.0, 1.0))
    assert h.pixel_count == 2000 and h.descriptor == g.output().output_descriptor()
    rgb = L.Renderer().render_image(g, req).rgb
    want = O.histogram_rgb(rgb, 64, 0.0, 1.0)
    for k, ch in enumerate(h.channels):
        assert np.array_equal(ch.counts, want[k][0])
        assert (ch.underflow, ch.overflow, ch.nonfinite, ch.minimum, ch.maximum) == want[k][1:]
        assert int(ch.counts.sum()) + ch.underflow + ch.overflow + ch.nonfinite == 2000
    nan, inf = math.nan, math.inf
    for opts in (L.HistogramOptions(0, 0, 1), L.HistogramOptions(65537, 0, 1), L.HistogramOptions(4, 1, 1),
                 L.HistogramOptions(4, 2, 1), L.HistogramOptions(4, nan, 1), L.HistogramOptions(4, 0, inf),
                 L.HistogramOptions(4, -1e308, 1e308)):
        with pytest.raises(L.InvalidArgument):
            L.histogram_rgb(g, req, opts)
    with pytest.raises(L.InvalidArgument):
        L.histogram_rgb(g, L.RenderRequest(L.Rect(0, 0, 0, 1)))
    with pytest.raises(L.OutOfRange):
        L.histogram_rgb(g, L.RenderRequest(L.Rect(60, 40, 8, 8)))
    tok = L.CancellationToken()
    tok.cancel()
    with pytest.raises(L.RenderCancelled):
        L.histogram_rgb(g, req, cancellation=tok)
