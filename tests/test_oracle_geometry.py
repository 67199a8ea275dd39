"""Pins of the oracle's box and neighbor functions (CPU): closed forms,
invariants, brute force over images and perfect-lattice shell counts."""
import itertools

import numpy as np
import pytest

from oracle import box as obox
from oracle import neighbor as onb
from simulon_amd import lattice as lat


def test_min_image_orthogonal_closed_form():
    H = np.diag([10.0, 20.0, 30.0])
    r = np.array([[6.0, -11.0, 44.0], [5.0, 10.0, -15.0], [15.0, 0.0, 0.0]])
    out = obox.minimum_image(r, H)
    # rint(0.6)=1, rint(-0.55)=-1, rint(1.4667)=1 ; halves round to even: rint(0.5)=0,
    # rint(0.5)=0, rint(-0.5)=-0 ; rint(1.5)=2
    np.testing.assert_allclose(out, [[-4.0, 9.0, 14.0], [5.0, 10.0, -15.0], [-5.0, 0.0, 0.0]])


@pytest.mark.parametrize("shear", [0.0, 0.1, 0.3])
def test_min_image_fractional_invariants(shear):
    L = 20.0
    H = np.array([[L, 0, 0], [shear * L, L, 0], [0.5 * shear * L, 0, L]])
    Hi = np.linalg.inv(H)
    rng = np.random.default_rng(0)
    r = rng.uniform(-3 * L, 3 * L, (2000, 3))
    m = obox.minimum_image(r, H, Hi)
    s = m @ Hi
    assert np.all(np.abs(s) <= 0.5 + 1e-12)                    # folded to the unit cell
    np.testing.assert_allclose(obox.minimum_image(m, H, Hi), m, atol=1e-9)   # idempotent
    n = (r - m) @ Hi                                          # differs by a lattice vector
    np.testing.assert_allclose(n, np.rint(n), atol=1e-9)


def test_min_image_is_shortest_for_short_vectors():
    """For |r| below half the smallest perpendicular width, the folded vector is
    the shortest of the 27 periodic images (brute force)."""
    L = 24.0
    H = np.array([[L, 0, 0], [0.1 * L, L, 0], [0.05 * L, 0, L]])
    w = obox.perpendicular_widths(H).min()
    rng = np.random.default_rng(1)
    base = rng.uniform(-0.45 * w, 0.45 * w, (500, 3))
    base = base[np.linalg.norm(base, axis=1) < 0.45 * w]
    shifts = np.array(list(itertools.product([-1, 0, 1], repeat=3))) @ H
    r = base + shifts[rng.integers(0, 27, len(base))]
    m = obox.minimum_image(r, H)
    cand = r[:, None, :] - shifts[None, :, :]
    best = cand[np.arange(len(r)), np.argmin(np.linalg.norm(cand, axis=2), axis=1)]
    np.testing.assert_allclose(m, best, atol=1e-9)


def test_min_image_fp32_order_is_division_then_rint():
    """FP32 arithmetic is the reference's: r - L * rint(r / L), each op rounded."""
    L = np.float32(36.4)
    r = np.array([[18.2, -18.2000008, 54.6]], np.float32)
    got = obox.minimum_image(r, np.diag([36.4] * 3))
    want = r - L * np.rint(r / L)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("name,pos_H,rl,expected", [
    ("sc", lat.lattice(lat.SC_BASIS, 10, np.diag([36.4] * 3)), 9.5, 80),
    ("bcc", lat.w_bcc(6)[0], 6.5, 64),
    ("fcc", lat.lattice(lat.FCC_BASIS, 6, np.diag([6 * 5.781] * 3)), 9.5, 78),
])
def test_lattice_shell_counts(name, pos_H, rl, expected):
    pos = pos_H
    n = len(pos)
    L = {"sc": 36.4, "bcc": 6 * 3.165, "fcc": 6 * 5.781}[name]
    H = np.diag([L] * 3)
    ei, _ = onb.brute_force_half_list(pos, H, np.float32(rl * rl))
    counts = np.bincount(np.concatenate(ei), minlength=n)
    assert np.all(counts == expected), (counts.min(), counts.max())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cell_list_equals_brute_force(dtype):
    rng = np.random.default_rng(5)
    for H in [np.diag([30.0, 33.0, 36.0]),
              np.array([[30.0, 0, 0], [3.0, 30.0, 0], [1.5, 0, 30.0]]),
              np.diag([20.0, 20.0, 20.0])]:      # only 2 cells per dimension
        pos = (rng.uniform(0, 1, (900, 3)) @ H + rng.normal(0, 5, (900, 3))).astype(dtype)
        rl2 = onb.list_radius_sq(8.5, 1.0)
        a, da = onb.brute_force_half_list(pos, H, rl2)
        b, db = onb.cell_half_list(pos, H, rl2)
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(da, db)


def test_half_list_is_sorted_and_strict():
    pos, H = lat.lj_argon_nvt(seed=2)
    ei, d = onb.brute_force_half_list(pos, H, onb.list_radius_sq(8.5, 1.0))
    assert np.all(ei[0] < ei[1])
    key = ei[0] * len(pos) + ei[1]
    assert np.all(np.diff(key) > 0)
    assert np.all(d < 9.5)


def test_list_radius_is_float32_sum_squared():
    assert onb.list_radius_sq(8.5, 1.0) == np.float32(90.25)
    rl = np.float32(5.4) + np.float32(1.0)
    assert onb.list_radius_sq(5.4, 1.0) == np.float32(rl * rl)


def test_rebuild_trigger_half_skin():
    H = np.diag([20.0] * 3)
    x0 = np.zeros((3, 3))
    x1 = x0.copy()
    x1[1, 0] = 0.49
    assert not onb.needs_rebuild(x1, x0, H, 1.0)
    x1[1, 0] = 0.51
    assert onb.needs_rebuild(x1, x0, H, 1.0)
    x1[1, 0] = 19.8                      # wraps: displacement -0.2
    assert not onb.needs_rebuild(x1, x0, H, 1.0)
