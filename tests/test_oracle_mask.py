"""The billiard-mask oracle (orc_mask) against the reference's own
xy_in_billiard_single_domain (sub_wave.c:5712) over ij_to_xy (469-476).

tests/golden/mask_cases.npz holds the masks three reference builds computed
(97x61 and 1280x720 grids, LAMBDA/MU of the plain workloads) for D_SINAI,
D_EXT_ELLIPSE and D_CIRCLES with a seeded circle list (every fifth circle
inactive).  A planted mistake -- the cell centre instead of its corner, >=
for >, an inactive circle taken as active -- fails the fixtures.
"""
import os

import numpy as np
import pytest

from gen_inputs import MASK_BOX, mask_cases, mask_circles

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIZES = {"mask_97x61_l03": (97, 61), "mask_97x61_l06_m03": (97, 61), "mask_1280x720_l03": (1280, 720)}


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "mask_cases.npz"))


@pytest.mark.parametrize("case", mask_cases(), ids=lambda c: c[0])
@pytest.mark.parametrize("domain", [3, 199, 20])
def test_mask_oracle_matches_reference(oracle, gold, case, domain):
    tag, lam, mu = case
    nx, ny = SIZES[tag]
    got = oracle.mask(nx, ny, MASK_BOX, domain, lam, mu, mask_circles() if domain == 20 else None)
    want = gold["%s/%d" % (tag, domain)]
    assert got.dtype == np.int16 and np.array_equal(got, want)
    assert 0 < got.sum() < got.size


def test_mask_fixtures_catch_planted_mistakes(oracle, gold):
    tag, lam, mu = mask_cases()[2]
    nx, ny = SIZES[tag]
    b = MASK_BOX
    half = (0.5 * (b[1] - b[0]) / nx, 0.5 * (b[3] - b[2]) / ny)
    centre = (b[0] + half[0], b[1] + half[0], b[2] + half[1], b[3] + half[1])
    for dom in (3, 199, 20):
        c = mask_circles() if dom == 20 else None
        assert not np.array_equal(oracle.mask(nx, ny, centre, dom, lam, mu, c), gold["%s/%d" % (tag, dom)])
    # a radius off by 1e-3 (a wrong constant) moves the boundary cells
    assert not np.array_equal(oracle.mask(nx, ny, b, 3, lam * (1 + 1e-3), mu), gold["%s/3" % tag])
    c = mask_circles()
    c[3] = 1.0
    assert not np.array_equal(oracle.mask(nx, ny, b, 20, lam, mu, c), gold["%s/20" % tag])


def test_mask_oracle_refuses_other_domains(oracle):
    with pytest.raises(ValueError):
        oracle.mask(10, 10, MASK_BOX, 1, 0.5)


@pytest.mark.skipif(not __import__("pyoracle").ref_available(), reason="oracle/_ref not built (needs /root/reference)")
def test_mask_oracle_matches_reference_build_other_circles(oracle):
    """direct comparison with the reference build on another circle list"""
    import pyoracle
    rng = np.random.default_rng(3)
    c = pyoracle.circle_list(rng.uniform(-2.2, 2.2, 70), rng.uniform(-1.3, 1.3, 70), rng.uniform(0.01, 0.5, 70),
                             (rng.uniform(0, 1, 70) > 0.3).astype(float))
    for tag, lam, mu in mask_cases():
        R = pyoracle.RefBuild(tag)
        for dom in (3, 199, 20):
            assert np.array_equal(R.mask(dom, c if dom == 20 else None),
                                  oracle.mask(R.nx, R.ny, MASK_BOX, dom, lam, mu, c if dom == 20 else None))
