"""The detector-acceptance oracle (shower.py:815-865) pinned by geometry:
straight lines with exactly representable slopes, shifted production points,
strict annulus and energy-window boundaries, pz = 0 and pz < 0, and the
weight sums of the Efficiency / TotalWeight methods."""
import numpy as np

from oracle import detector as od


def rec(E, px, py, pz, x=0.0, y=0.0, z=0.0, w=1.0):
    return np.array([E, px, py, pz]), np.array([x, y, z, w])


def arrays(recs):
    p = np.array([a for a, _ in recs])
    r = np.array([b for _, b in recs])
    return p, r


def test_lines_from_origin_at_exact_slopes():
    # slopes px/pz = 1/4, 1/2, 1 reach r_T = z/4, z/2, z at the plane z
    p, r = arrays([rec(1.0, 0.25, 0.0, 1.0), rec(1.0, 0.0, 0.5, 1.0), rec(1.0, 0.6, 0.8, 0.8)])
    win, ok = od.pass_matrix(p, r, [4.0, 8.0], radius=3.0)
    # r_T at z=4: 1, 2, 5 -> first two pass; at z=8: 2, 4, 10 -> first only
    assert win.all()
    assert ok.tolist() == [[True, True, False], [True, False, False]]
    x, y = od.transverse_position(r[2], p[2], 4.0)
    assert (x, y) == (3.0, 4.0)


def test_shifted_production_point():
    # produced at (1, -2, 3) moving along (2, 2, 4): at z = 5, T = 0.5, (x, y) = (2, -1)
    p, r = arrays([rec(5.0, 2.0, 2.0, 4.0, x=1.0, y=-2.0, z=3.0)])
    x, y = od.transverse_position(r[0], p[0], 5.0)
    assert (x, y) == (2.0, -1.0)
    _, ok = od.pass_matrix(p, r, [5.0], radius=np.sqrt(5.0) + 1e-9, inner_radius=np.sqrt(5.0) - 1e-9)
    assert ok[0, 0]


def test_annulus_boundaries_are_strict():
    # r_T = 1 exactly (slope 1/2 at z = 2)
    p, r = arrays([rec(1.0, 0.5, 0.0, 1.0)])
    assert not od.pass_matrix(p, r, [2.0], radius=1.0)[1][0, 0]
    assert od.pass_matrix(p, r, [2.0], radius=np.nextafter(1.0, 2.0))[1][0, 0]
    assert not od.pass_matrix(p, r, [2.0], radius=2.0, inner_radius=1.0)[1][0, 0]
    assert od.pass_matrix(p, r, [2.0], radius=2.0, inner_radius=np.nextafter(1.0, 0.0))[1][0, 0]


def test_on_axis_record_fails_with_zero_inner_radius():
    # the inner cut is strict (shower.py:854): r_T = 0 > 0 is false
    p, r = arrays([rec(1.0, 0.0, 0.0, 1.0)])
    assert not od.pass_matrix(p, r, [1.0], radius=1.0)[1][0, 0]


def test_energy_window_is_strict():
    p, r = arrays([rec(E, 0.125, 0.0, 1.0) for E in (1.0, 2.0, 3.0)])
    win, ok = od.pass_matrix(p, r, [1.0], radius=1.0, energy_cut=(1.0, 3.0))
    assert win.tolist() == [False, True, False]
    assert ok[0].tolist() == [False, True, False]


def test_zero_and_negative_pz():
    # pz = 0: no finite intersection -> never passes; pz < 0 with the plane
    # ahead: the reference extrapolates backwards (T < 0) and counts the hit
    p, r = arrays([rec(1.0, 0.1, 0.0, 0.0), rec(1.0, 0.25, 0.0, -1.0)])
    _, ok = od.pass_matrix(p, r, [4.0], radius=1.5)
    assert ok[0].tolist() == [False, True]
    x, _ = od.transverse_position(r[1], p[1], 4.0)
    assert x == -1.0


def test_methods_and_weight_sums():
    p, r = arrays([rec(1.0, 0.0, 0.03125, 1.0, w=0.5), rec(2.0, 1.0, 0.0, 1.0, w=0.25), rec(3.0, 0.0, 0.1, 1.0, w=2.0),
                   rec(9.0, 0.0, 0.03125, 1.0, w=7.0)])
    z = [1.0, 20.0]
    # r_T at z=1: 1/32, 1, 0.1, 1/32 ; at z=20: 0.625, 20, 2, 0.625 ; the
    # energy window (0.5, 5) drops the last record
    assert [s.tolist() for s in od.detector_cut(p, r, z, 1.5, "Sample", (0.5, 5.0))] == [[0, 1, 2], [0]]
    assert od.detector_cut(p, r, z, 1.5, "SampleW", (0.5, 5.0)).tolist() == [[True, True, True], [True, False, False]]
    assert od.detector_cut(p, r, z, 1.5, "TotalWeight", (0.5, 5.0)) == [2.75, 0.5]
    eff = od.detector_cut(p, r, z, 1.5, "Efficiency", (0.5, 5.0))
    assert eff == [1.0, 0.5 / 2.75]
    assert od.detector_cut(p, r, z, 1.5, "Efficiency", (100.0, 200.0)) == [0.0, 0.0]
    assert od.detector_cut(p, r, z, 1.5, "Sample", (100.0, 200.0)) == [[], []]
    assert od.detector_cut(p, r, z, 1.5, "Other") is None
