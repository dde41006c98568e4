"""The lattice-probing Menon adjoint used by the large-frame GPU gradient test,
checked against the brute-force column-by-column adjoint of the FP64 oracle
(one forward per sample) on small areas, viewport edges and mosaic phases
included."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import menon_adjoint_by_probing


@pytest.mark.parametrize("case", [((31, 27, 0, 0, (0, 0, 31, 27)), (0, 0, 31, 27)),
                                  ((40, 33, 2, 1, (3, 2, 34, 29)), (8, 5, 21, 17))])
def test_probe_matches_brute_force(case):
    (w, h, pat, ph, area), vp = case
    om = O.Bayer(w, h, w, pat, ph, 0, area, (0,) * 4, (1,) * 4)
    rng = np.random.default_rng(w)
    ax, ay, aw, ah = area
    raw = rng.uniform(0.05, 1.0, (ah, aw))
    _, dec = O.menon_f64(om, raw, area)
    wgt = rng.normal(size=(vp[3], vp[2], 3))
    gains, mult = (1.7, 1.0, 1.3), 1.25
    M = np.array([[1.05, -0.03, -0.02], [-0.06, 1.02, 0.04], [0.03, 0.09, 0.88]])
    lin = (M * np.asarray(gains)[None, :]) * mult
    got = menon_adjoint_by_probing(om, dec, wgt, vp, gains, mult, M)
    brute = np.zeros((ah, aw))
    for y in range(ah):
        for x in range(aw):
            e = np.zeros((ah, aw))
            e[y, x] = 1.0
            out = O.menon_f64(om, e, vp, hz=dec)[0]
            brute[y, x] = np.einsum("hwr,rc,hwc->", wgt, lin, out)
    assert np.max(np.abs(got - brute)) <= 1e-12 * np.max(np.abs(brute))
    assert np.count_nonzero(brute) >= vp[2] * vp[3]
