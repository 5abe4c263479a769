"""Input builders: the reference's coordinate maps, masks and tables."""
import numpy as np

from wavehip import workloads


def test_splitmix64_known_answer():
    # outputs of the published SplitMix64 (Vigna's splitmix64.c) for seed 1234567
    r = workloads.SplitMix64(1234567)
    assert [r.next() for _ in range(3)] == [6457827717110365317, 3203168211198807973,
                                            9817491932198370423]


def test_splitmix64_matches_counter_form():
    from gen_inputs import splitmix64_stream
    r = workloads.SplitMix64(99)
    assert [int(v) for v in splitmix64_stream(99, 5)] == [r.next() for _ in range(5)]


def test_xy_to_ij_and_ij_to_xy():
    # point source of the 2560x1440 lens workload at (-1.5, 0)
    assert workloads.xy_to_ij(-1.5, 0.0, 2560, 1440, -2.0, 2.0, -1.125, 1.125) == (320, 720)
    x, y = workloads.grid_xy(1280, 720, -2.0, 2.0, -1.125, 1.125)
    assert x[0] == -2.0 and y[0] == -1.125 and x[640] == 0.0 and y[360] == 0.0
    assert x[1] == -2.0 + 4.0 / 1280


def test_masks_closed_form():
    x, y = workloads.grid_xy(1280, 720, -2.0, 2.0, -1.125, 1.125)
    m = workloads.mask_sinai(x, y, 0.3)
    assert m[640, 360] == 0 and m[0, 0] == 1
    # cells exactly on the circle x^2 + y^2 = 0.09 are inside the domain? (> is strict)
    i = 640 + 96  # x = 0.3
    assert x[i] * x[i] == 0.09 or abs(x[i] - 0.3) < 1e-15
    assert m[i, 360] == (1 if x[i] * x[i] > 0.3 * 0.3 else 0)
    e = workloads.mask_ext_ellipse(x, y, 0.6, 0.3)
    assert e[640, 360] == 0 and e[640 + 200, 360] == 1
    d = workloads.mask_discs(x, y, [(0.0, 0.0, 0.1), (1.0, 0.5, 0.05)])
    X, Y = np.meshgrid(x, y, indexing="ij")
    ref = np.where(((X ** 2 + Y ** 2) < 0.01) | (((X - 1) ** 2 + (Y - 0.5) ** 2) < 0.0025), 0, 1)
    assert np.array_equal(d, ref)


def test_constant_media_tables_rule():
    xy = np.array([[0, 1, 2, -1]], np.int16)
    tcc, tc, tg = workloads.constant_media_tables(xy, 0.1, 0.05, 1e-4, 2e-4, 0)
    assert list(tc[0]) == [0.0, 0.1, 0.1, 0.1]
    assert list(tcc[0]) == [0.0, 0.1 * 0.1, 0.1 * 0.1, 0.1 * 0.1]
    assert list(tg[0]) == [0.0, 1e-4, 2e-4, 2e-4]
    tcc, tc, tg = workloads.constant_media_tables(xy, 0.1, 0.05, 1e-4, 2e-4, 1)
    assert tc[0, 0] == 0.05 and tcc[0, 0] == 0.05 * 0.05 and tg[0, 0] == 2e-4


def test_workload_shapes_small():
    w = workloads.c3(n=256, nsteps=10)
    assert w["xy_in"].shape == (256, 256) and (w["xy_in"] == 0).any()
    assert np.all(w["phi"][w["xy_in"] == 0] == 0)
    ws = workloads.c4_instances(count=4, nmax=200)
    assert len(ws) == 4 and all(64 <= w["params"]["nx"] <= 200 for w in ws)


def test_laplacian_table_matches_reference_tables():
    """workloads.laplacian_table equals the table init_laplacian_coords builds
    (fixtures exported from the reference builds), and the oracle's builder at
    sizes the fixtures do not cover."""
    import os
    import pyoracle
    from gen_inputs import table_cases
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "table_cases.npz"))
    for (tag, nx, ny, bc) in {c[:4] for c in table_cases()}:
        if bc > 3 and bc != 10:
            continue  # no table edges for B_COND 4
        nn, nb = workloads.laplacian_table(nx, ny, bc, tuple(g[tag + "/lsh"]))
        rn, rb = g[tag + "/nneighb"], g[tag + "/nghb"]
        used = lambda n, b: np.where(n[..., None] > np.arange(4), b, -1)  # noqa: E731
        assert np.array_equal(nn, rn) and np.array_equal(used(nn, nb), used(rn, rb)), tag
    for nx, ny in ((131, 67), (64, 200)):
        lsh = workloads.lshape_lines(nx, ny, 1.5, -2.0, 2.0, -1.125, 1.125)
        for bc in (0, 1, 2, 3, 10):
            a = pyoracle.laplacian_table(pyoracle.make_params(nx, ny, bcond=bc), lsh)
            b = workloads.laplacian_table(nx, ny, bc, lsh)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (nx, ny, bc)


def test_lshape_mask_and_lines():
    x, y = workloads.grid_xy(24, 17, -2.0, 2.0, -1.125, 1.125)
    m = workloads.mask_lshape(x, y, 1.5)
    X, Y = np.meshgrid(x, y, indexing="ij")
    inside = (np.abs(X) <= 1.5) & (np.abs(Y) <= 1.0) & ~((X > 0) & (Y > 0))
    assert np.array_equal(m == 1, inside)
    # the seams of the 24x17 reference build (ref_lshape_lines)
    assert workloads.lshape_lines(24, 17, 1.5, -2.0, 2.0, -1.125, 1.125) == (4, 1, 11, 7, 20, 15)
