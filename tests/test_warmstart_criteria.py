"""The warm start's judging rules on hand-written curves (CPU only): the time
to both targets needs stress <= the stress target AND the edge-length CV inside
the 1 % band on either side of the converged CV; "stays" needs every later
check inside both; EdgeCv is the coefficient of variation of the undirected
edge lengths (crates/graph-layouts/src/metrics.rs:150-164)."""
import numpy as np

from jumpcannon import warmstart as W


def test_both_targets_need_the_cv_inside_the_band_from_below_too():
    band = (0.99 * 0.40, 1.01 * 0.40)
    # a random start has a low CV (0.37): stress alone is met at 100, the CV band only at 200
    curve = [(0, 0.20, 0.37), (100, 0.12, 0.38), (200, 0.12, 0.399), (300, 0.11, 0.401)]
    assert W.first_within(curve, 0.125, band) == (200, True)


def test_both_targets_from_above_and_leaving_the_band():
    band = (0.99 * 0.40, 1.01 * 0.40)
    curve = [(0, 0.11, 0.55), (50, 0.11, 0.403), (100, 0.13, 0.402), (150, 0.12, 0.401)]
    assert W.first_within(curve, 0.125, band) == (50, False)   # stress leaves the target at 100
    assert W.first_within(curve[:2], 0.125, band) == (50, True)
    assert W.first_within([(0, 0.2, 0.4), (50, 0.13, 0.5)], 0.125, band) == (None, False)


def test_edge_cv_on_a_hand_computed_layout():
    # path 0-1-2 with edge lengths 1 and 3: mean 2, population std 1, CV 0.5
    off = np.array([0, 1, 3, 4], np.uint32)
    nb = np.array([1, 0, 2, 1], np.uint32)
    xyz = np.array([0, 0, 0, 1, 0, 0, 4, 0, 0], np.float32)
    assert W.EdgeCv(off, nb)(xyz) == np.float32(0.5)
    # equal edges: CV 0; fewer than two edges: 0 by definition
    assert W.EdgeCv(off, nb)(np.array([0, 0, 0, 1, 0, 0, 2, 0, 0], np.float32)) == 0.0
    assert W.EdgeCv(np.array([0, 1, 2], np.uint32), np.array([1, 0], np.uint32))(np.zeros(6, np.float32)) == 0.0
