"""The closed-form billiard domains of the oracle (orc_domain) against the
reference's own xy_in_billiard_single_domain (sub_wave.c:5712-7602).

tests/golden/domain_cases.npz holds the masks three reference builds (dom_a:
97x61, LAMBDA 0.6, MU 0.3, NPOLY 6, APOLY 0.5; dom_b: 160x90 and dom_c:
121x77, each with another value for every macro the domains read) computed
for every code of
gen_inputs.DOMAIN_CODES at the cell corners of ij_to_xy.  Every code but the
two constant domains takes more than one value in one of the builds, and a
planted mistake -- the cell centre instead of the corner, a parameter off by
a part in a thousand -- fails the fixtures.
"""
import os

import numpy as np
import pytest

from gen_inputs import DOMAIN_CODES, MASK_BOX, domain_sets, mask_circles

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CIRCLE_CODES = (20, 201, 202)
CONSTANT = (999, 76, 77)  # D_NOTHING, D_GRADIENT_INDEX_LENS and D_IMAGE are 1 everywhere
# lattices placed from XMIN, YMIN with the box's own spacing: a box shifted
# with the sample points moves them too, so the cell-centre mistake is
# invisible there; NGRIDX off by one is not (test_new_macros_are_pinned)
BOX_LATTICES = (93, 934, 94, 942)


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "domain_cases.npz"))


def oracle_mask(oracle, tag, dom, box=MASK_BOX, **over):
    nx, ny, kw = domain_sets()[tag]
    kw = dict(kw, **over)
    return oracle.domain(nx, ny, box, dom, oracle.domain_params(**kw), mask_circles() if dom in CIRCLE_CODES else None)


@pytest.mark.parametrize("tag", sorted(domain_sets()))
@pytest.mark.parametrize("dom", DOMAIN_CODES)
def test_domain_oracle_matches_reference(oracle, gold, tag, dom):
    got = oracle_mask(oracle, tag, dom)
    assert got.dtype == np.int16 and np.array_equal(got, gold["%s/%d" % (tag, dom)])


def test_fixtures_are_not_trivial(gold):
    for dom in DOMAIN_CODES:
        vals = set()
        for tag in domain_sets():
            v = np.unique(gold["%s/%d" % (tag, dom)])
            if len(v) > 1:
                vals.add(tag)
        assert (len(vals) == 0) == (dom in CONSTANT), dom


def test_fixtures_catch_planted_mistakes(oracle, gold):
    b = MASK_BOX

    def centre(tag):
        nx, ny, _ = domain_sets()[tag]
        hx, hy = 0.5 * (b[1] - b[0]) / nx, 0.5 * (b[3] - b[2]) / ny
        return (b[0] + hx, b[1] + hx, b[2] + hy, b[3] + hy)

    for dom in DOMAIN_CODES:
        if dom in CONSTANT or dom in BOX_LATTICES:
            continue
        hits = [not np.array_equal(oracle_mask(oracle, t, dom, centre(t)), gold["%s/%d" % (t, dom)])
                for t in domain_sets()]
        assert any(hits), dom
    # LAMBDA or MU off by a part in a thousand moves the boundary of most domains
    moved = 0
    for dom in DOMAIN_CODES:
        for tag in domain_sets():
            kw = domain_sets()[tag][2]
            lam = kw.get("lam")
            a = oracle_mask(oracle, tag, dom, lam=lam * 1.001)
            m = oracle_mask(oracle, tag, dom, mu=kw.get("mu") * 1.001)
            if not np.array_equal(a, gold["%s/%d" % (tag, dom)]) or not np.array_equal(m, gold["%s/%d" % (tag, dom)]):
                moved += 1
                break
    assert moved >= 40


def test_new_macros_are_pinned(oracle, gold):
    """the macros only the lattice, tree-beam and waveguide domains read:
    each planted mistake (WALL_WIDTH_ASYM or _B off by 10 %, the two
    swapped, OSCIL_YMID or the Michelson mirror shifted, NGRIDX off by one)
    fails the fixtures of every code that reads it, in at least one build"""
    def fails(dom, **over):
        return any(not np.array_equal(oracle_mask(oracle, t, dom, **{k: v(domain_sets()[t][2]) for k, v in over.items()}),
                                      gold["%s/%d" % (t, dom)]) for t in domain_sets())

    D = oracle.DOMAIN_DEFAULTS
    g = lambda k: (lambda kw: kw.get(k, D[k]))  # noqa: E731
    for dom in (934, 935, 936, 937, 942, 947):
        assert fails(dom, wall_width_asym=lambda kw: 1.1 * g("wall_width_asym")(kw)), dom
    assert fails(942, wall_width_asym_b=lambda kw: 1.1 * g("wall_width_asym_b")(kw))
    for dom in (942, 947):
        assert fails(dom, wall_width_asym=g("wall_width_asym_b"), wall_width_asym_b=g("wall_width_asym")), dom
    assert fails(731, oscil_ymid=lambda kw: g("oscil_ymid")(kw) + 0.05)
    assert fails(741, michelson_position=lambda kw: g("michelson_position")(kw) + 0.05)
    for dom in (93, 934, 935, 936, 937, 94, 942, 947, 99, 991):
        assert fails(dom, ngridx=lambda kw: g("ngridx")(kw) + 1), dom


def test_domain_oracle_refuses_other_codes(oracle):
    with pytest.raises(ValueError):
        oracle.domain(10, 10, MASK_BOX, 53, oracle.domain_params())  # D_MAZE reads a maze list


@pytest.mark.skipif(not __import__("pyoracle").ref_available(), reason="oracle/_ref not built (needs /root/reference)")
def test_domain_oracle_matches_reference_build_other_circles(oracle):
    """direct comparison with the reference build on another circle list"""
    import pyoracle
    rng = np.random.default_rng(8)
    c = pyoracle.circle_list(rng.uniform(-2.2, 2.2, 60), rng.uniform(-1.3, 1.3, 60), rng.uniform(0.01, 0.5, 60),
                             (rng.uniform(0, 1, 60) > 0.3).astype(float))
    for tag, (nx, ny, kw) in domain_sets().items():
        R = pyoracle.RefBuild(tag)
        for dom in CIRCLE_CODES:
            assert np.array_equal(R.mask(dom, c), oracle.domain(nx, ny, MASK_BOX, dom, oracle.domain_params(**kw), c))


def test_comparison_mask_matches_reference(oracle, gold):
    """xy_in_billiard with COMPARISON (sub_wave.c:7593-7602) as the reference
    builds dom_cmp_a (D_CIRCLES above, D_CIRCLES with the second list below)
    and dom_cmp_b (D_TREE above, D_POLYGON below) compute it; the halves
    swapped fail"""
    from gen_inputs import COMPARISON_BUILDS, mask_circles_b
    P = oracle.domain_params(**domain_sets()["dom_a"][2])
    for tag, (d, db) in COMPARISON_BUILDS.items():
        got = oracle.comparison_mask(97, 61, MASK_BOX, d, db, P, mask_circles(), mask_circles_b())
        assert np.array_equal(got, gold[tag + "/mask"]), tag
        swapped = oracle.comparison_mask(97, 61, MASK_BOX, db, d, P, mask_circles_b(), mask_circles())
        assert not np.array_equal(swapped, gold[tag + "/mask"]), tag


# ---- the list-driven domains ---------------------------------------------------

LIST_CIRCLE_CODES = (96, 932, 933, 941)


def list_mask(oracle, tag, dom, lists=None, **over):
    from gen_inputs import list_data
    nx, ny, kw = domain_sets()[tag]
    kw = dict(kw, **over)
    L = list_data(kw.get("npoly", oracle.DOMAIN_DEFAULTS["npoly"])) if lists is None else lists
    return oracle.list_domain(nx, ny, MASK_BOX, dom, oracle.domain_params(**kw), circles=mask_circles(), **L)


@pytest.mark.parametrize("tag", sorted(domain_sets()))
@pytest.mark.parametrize("dom", __import__("gen_inputs").LIST_DOMAIN_CODES)
def test_list_domain_oracle_matches_reference(oracle, gold, tag, dom):
    """xy_in_billiard_single_domain of the three builds with list_data's
    rectangles, rotated rectangles, polygons and polyline in the reference's
    globals (ref_set_lists), and the circle list"""
    got = list_mask(oracle, tag, dom)
    assert got.dtype == np.int16 and np.array_equal(got, gold["%s/list/%d" % (tag, dom)])


def test_list_domains_read_their_lists(oracle, gold):
    """each code depends on the lists it reads: emptying them, or a planted
    mistake in how one is read (rectangle corners not sorted, rotated
    rectangle width doubled, polygon angle in radians, the polyline reversed)
    fails the fixtures in at least one build"""
    from gen_inputs import LIST_DOMAIN_CODES, list_data

    def fails(dom, edit):
        out = []
        for t in domain_sets():
            L = list_data(domain_sets()[t][2].get("npoly", oracle.DOMAIN_DEFAULTS["npoly"]))
            try:
                got = list_mask(oracle, t, dom, edit(L))
            except ValueError:  # a polyline too short for the code is refused
                out.append(True)
                continue
            out.append(not np.array_equal(got, gold["%s/list/%d" % (t, dom)]))
        return any(out)

    def empty(L):
        return {k: v[:, :0] for k, v in L.items()}

    for dom in LIST_DOMAIN_CODES:
        if dom in LIST_CIRCLE_CODES:
            continue  # the circle list is tested below
        assert fails(dom, empty), dom
    sorted_rects = lambda L: dict(L, rects=np.vstack([np.minimum(L["rects"][0], L["rects"][2]),  # noqa: E731
                                                       np.minimum(L["rects"][1], L["rects"][3]),
                                                       np.maximum(L["rects"][0], L["rects"][2]),
                                                       np.maximum(L["rects"][1], L["rects"][3])]))
    for dom in (53, 192, 193):  # the maze and guide tests leave reversed rectangles empty
        assert fails(dom, sorted_rects), dom
    wide = lambda L: dict(L, rectrot=np.vstack([L["rectrot"][:4], 2.0 * L["rectrot"][4:]]))  # noqa: E731
    for dom in (95, 97, 96, 932):
        assert fails(dom, wide), dom
    rad = lambda L: dict(L, polygons=np.vstack([L["polygons"][:3], L["polygons"][3:4] * np.pi / 2,  # noqa: E731
                                                 L["polygons"][4:]]))
    assert fails(40, rad)
    rev = lambda L: dict(L, polyline=np.ascontiguousarray(L["polyline"][:, ::-1]))  # noqa: E731
    for dom in (41, 42, 421, 27):
        assert fails(dom, rev), dom
    # the disc lattices read every circle of the list, radius MU, active or not
    for dom in LIST_CIRCLE_CODES:
        hits = []
        for t, (nx, ny, kw) in domain_sets().items():
            L = list_data(kw.get("npoly", oracle.DOMAIN_DEFAULTS["npoly"]))
            c = mask_circles()
            got = oracle.list_domain(nx, ny, MASK_BOX, dom, oracle.domain_params(**kw), circles=c[:, ::2], **L)
            hits.append(not np.array_equal(got, gold["%s/list/%d" % (t, dom)]))
        assert any(hits), dom


def test_list_domain_oracle_refuses_short_lists(oracle):
    from gen_inputs import list_data
    L = list_data(6)
    P = oracle.domain_params(npoly=6)
    with pytest.raises(ValueError):
        oracle.list_domain(20, 10, MASK_BOX, 421, P, **dict(L, polyline=L["polyline"][:, :9]))
    with pytest.raises(ValueError):
        oracle.list_domain(20, 10, MASK_BOX, 3, P, **L)  # a closed-form code goes through orc_domain
