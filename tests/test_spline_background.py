"""The spline background of the oracle (js/rietveld_core.js:549-612, a monotone cubic
Hermite spline with Fritsch-Carlson tangents), pinned by what the mathematics fixes:
interpolation of the anchors, exactness on linear data, flat extension, C1 continuity,
no overshoot where the tangent limiter acts, and a closed form."""
import numpy as np

import oracle_lib as O
from powder6_amd import problems as PR


def test_interpolates_the_anchors_and_is_flat_outside():
    xs = np.array([12.0, 30.0, 55.0, 80.0, 118.0])
    ys = np.array([310.0, 250.0, 260.0, 180.0, 175.0])
    assert np.array_equal(O.spline_values(xs, xs, ys), ys)
    out = O.spline_values(np.array([5.0, 11.9, 118.5, 150.0]), xs, ys)
    assert np.array_equal(out, [ys[0], ys[0], ys[-1], ys[-1]])


def test_linear_data_give_a_line():
    xs = np.array([10.0, 17.0, 40.0, 41.0, 90.0, 120.0])
    ys = 3.0 + 0.25 * xs
    t = np.linspace(10.0, 120.0, 1001)
    assert np.allclose(O.spline_values(t, xs, ys), 3.0 + 0.25 * t, rtol=0, atol=1e-11)


def test_closed_form_of_a_peak():
    # anchors (0,0), (1,1), (2,0): tangents 1, 0, -1, no limiting; on [0, 1] the spline is
    # -u^3 + u^2 + u and on [1, 2] its mirror image
    xs, ys = np.array([0.0, 1.0, 2.0]), np.array([0.0, 1.0, 0.0])
    u = np.linspace(0.0, 1.0, 11)
    assert np.allclose(O.spline_values(u, xs, ys), -u ** 3 + u ** 2 + u, atol=1e-14)
    assert np.allclose(O.spline_values(2.0 - u, xs, ys), -u ** 3 + u ** 2 + u, atol=1e-14)


def test_monotone_steep_data_do_not_overshoot():
    # a step between gentle segments: the tangent limiter (tau > 9) must act, and the
    # spline stays monotone and inside the range of each interval's end values
    xs = np.array([10.0, 20.0, 21.0, 22.0, 40.0, 60.0])
    ys = np.array([100.0, 101.0, 180.0, 181.0, 182.0, 400.0])
    t = np.linspace(10.0, 60.0, 20001)
    y = O.spline_values(t, xs, ys)
    assert np.all(np.diff(y) >= -1e-9)
    for i in range(len(xs) - 1):
        sel = (t >= xs[i]) & (t <= xs[i + 1])
        assert y[sel].min() >= ys[i] - 1e-9 and y[sel].max() <= ys[i + 1] + 1e-9


def test_first_derivative_is_continuous_at_the_anchors():
    xs = np.array([10.0, 25.0, 33.0, 70.0, 100.0])
    ys = np.array([200.0, 150.0, 170.0, 90.0, 120.0])
    e = 1e-6
    for x in xs[1:-1]:
        left = (O.spline_values([x], xs, ys)[0] - O.spline_values([x - e], xs, ys)[0]) / e
        right = (O.spline_values([x + e], xs, ys)[0] - O.spline_values([x], xs, ys)[0]) / e
        assert abs(left - right) <= 1e-3 * max(1.0, abs(left))


def test_one_anchor_is_a_constant_and_none_is_zero():
    t = np.linspace(10, 20, 5)
    assert np.array_equal(O.spline_values(t, [15.0], [42.0]), np.full(5, 42.0))
    assert np.array_equal(O.spline_values(t, np.zeros(0), np.zeros(0)), np.zeros(5))


def test_host_spline_matches_oracle():
    """The package's own spline (used for synthetic observations) against the oracle's,
    on anchors where the tangent limiter acts and where it does not."""
    t = np.linspace(5.0, 125.0, 2001)
    xs = np.array([10.0, 20.0, 35.0, 50.0, 80.0, 100.0, 120.0])
    for ys in ([400.0, 50.0, 50.0, 300.0, 20.0, 25.0, 200.0], [5.0, 6.0, 7.5, 30.0, 31.0, 31.5, 80.0]):
        got = PR.spline_background(t, xs, np.array(ys))
        ref = O.spline_values(t, xs, np.array(ys))
        assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_model_pattern_is_spline_at_zero_scale():
    """With the phase scale at zero the oracle's pattern is the spline through the bga
    anchors, flat outside them (js/rietveld_core.js:920-936)."""
    pb = PR.nacl_spline()
    pb.y = np.ones_like(pb.tth)
    m = O.Model(pb)
    x = pb.start_vector()
    x[pb.var_index('p1.scale')] = 0.0
    y = m.pattern(x)
    p = pb.values(x)
    assert np.allclose(y, O.spline_values(pb.tth, pb.bkg_x, p[3:3 + pb.n_bkg]), rtol=0, atol=1e-9)
    assert y[0] == p[3] and y[-1] == p[3 + pb.n_bkg - 1]


def test_plan_rejects_repeated_anchors():
    """Two anchors at one 2-theta leave an interval of zero width (a NaN tangent in the
    reference's arithmetic); the plan refuses them with a message instead."""
    import pytest
    from powder6_amd import pt
    pb = PR.nacl_spline()
    pb.y = np.ones_like(pb.tth)
    pb.bkg_x = pb.bkg_x.copy()
    pb.bkg_x[3] = pb.bkg_x[2]
    with pytest.raises(Exception, match='distinct'):
        pt.Plan(pb, device=False)
