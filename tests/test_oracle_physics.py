"""Pins of the oracle's force fields and integrator (CPU): closed forms,
conservation laws, finite differences and the reference's own physical
regressions on its W-Re potential file."""
import gzip
import os
import shutil

import numpy as np
import pytest

from oracle import box as obox
from oracle import eam as oeam
from oracle import integrate as oint
from oracle import lj as olj
from oracle import neighbor as onb
from simulon_amd import lattice as lat

HERE = os.path.dirname(os.path.abspath(__file__))
EV_A3_TO_BAR = 1_602_176.6208
EV_A3_TO_GPA = 160.21766208


def _lj(pos, H, rc=8.5, mode="raw", eps=0.0104, sig=3.405):
    ei, _ = onb.brute_force_half_list(pos, H, onb.list_radius_sq(rc, 1.0))
    return olj.lj_forward(pos, ei, np.zeros(len(pos), int), [[eps]], [[sig]], rc, H, mode=mode)


# ---- LJ ---------------------------------------------------------------------
@pytest.mark.parametrize("r", [3.2, 3.82, 5.0, 8.4])
def test_lj_two_atoms_closed_form(r):
    H = np.diag([40.0] * 3)
    pos = np.array([[1.0, 2.0, 3.0], [1.0 + r * 0.6, 2.0 + r * 0.8, 3.0]])
    out = _lj(pos, H)
    U, F = olj.lj_two_atom(r, 0.0104, 3.405)
    assert out["energy"] == pytest.approx(U, rel=1e-12)
    # force on atom 0 points from 1 to 0 when repulsive (F > 0)
    np.testing.assert_allclose(out["forces"][0], -F * np.array([0.6, 0.8, 0.0]), rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(out["forces"][1], -out["forces"][0], rtol=1e-12)
    assert out["virial"] == pytest.approx(F * r, rel=1e-12)


def test_lj_beyond_cutoff_is_zero():
    H = np.diag([40.0] * 3)
    pos = np.array([[0.0, 0, 0], [8.6, 0, 0]])
    out = _lj(pos, H)
    assert out["energy"] == 0.0 and np.all(out["forces"] == 0.0)


@pytest.mark.parametrize("mode", ["raw", "shift", "switch"])
def test_lj_newton_virial_and_gradient(mode):
    pos, H = lat.lj_argon_nvt(seed=3)
    pos = pos[:1000]
    out = _lj(pos, H, mode=mode)
    np.testing.assert_allclose(out["forces"].sum(0), 0.0, atol=1e-12)
    assert np.trace(out["virial_tensor"]) == pytest.approx(out["virial"], rel=1e-10)
    np.testing.assert_allclose(out["virial_tensor"], out["virial_tensor"].T, atol=1e-12)
    if mode == "raw":
        return   # raw truncation: E jumps at rc, -dE/dx only away from crossings
    rng = np.random.default_rng(0)
    h = 1e-5
    for i in rng.integers(0, len(pos), 5):
        for a in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, a] += h
            p2[i, a] -= h
            fd = -(_lj(p1, H, mode=mode)["energy"] - _lj(p2, H, mode=mode)["energy"]) / (2 * h)
            assert fd == pytest.approx(out["forces"][i, a], abs=1e-7)


def test_lj_force_shift_continuity():
    """Force-shift form: U(rc) = 0 and F(rc) = 0."""
    H = np.diag([40.0] * 3)
    for r in [8.5 - 1e-7, 8.5 - 1e-6]:
        pos = np.array([[0.0, 0, 0], [r, 0, 0]])
        out = _lj(pos, H, mode="shift")
        assert abs(out["energy"]) < 1e-10
        assert np.abs(out["forces"]).max() < 1e-8


def test_lj_lattice_sum():
    """Perfect simple-cubic lattice: E = N/2 sum over shells of count * U(r)."""
    a, n = 3.64, 10
    pos = lat.lattice(lat.SC_BASIS, n, np.diag([a * n] * 3))
    out = _lj(pos, np.diag([a * n] * 3))
    shells = {}
    for i in range(-3, 4):
        for j in range(-3, 4):
            for k in range(-3, 4):
                d2 = i * i + j * j + k * k
                if 0 < d2 * a * a < 8.5 ** 2:
                    shells[d2] = shells.get(d2, 0) + 1
    e = 0.5 * len(pos) * sum(c * olj.lj_two_atom(a * np.sqrt(d2), 0.0104, 3.405)[0]
                             for d2, c in shells.items())
    assert out["energy"] == pytest.approx(e, rel=1e-12)
    assert np.abs(out["forces"]).max() < 1e-12


# ---- EAM --------------------------------------------------------------------
@pytest.fixture(scope="module")
def syn(tmp_path_factory):
    p = tmp_path_factory.mktemp("eam") / "W_syn.eam.alloy"
    lat.write_w_synthetic_setfl(p)
    f = oeam.EAMFile(p)
    return str(p), f, oeam.EAMTables(f, dtype=np.float64)


@pytest.fixture(scope="module")
def wre(tmp_path_factory):
    d = tmp_path_factory.mktemp("wre")
    dst = d / "WRe_YC2.eam.fs"
    with gzip.open(os.path.join(HERE, "golden", "data", "WRe_YC2.eam.fs.gz"), "rb") as fi, \
            open(dst, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    return str(dst)


def test_eam_parser_layout(syn, wre):
    _, f, _ = syn
    assert f.elements == ["W"] and f.n_r == 5000 and f.cutoff == pytest.approx(5.5)
    g = oeam.EAMFile(wre)
    assert g.is_fs_format and g.elements == ["W", "Re"]
    assert g.dens[0][0] is not g.dens[0][1]          # FS: per-pair densities
    assert g.n_rho == 10000 and g.n_r == 10000 and g.cutoff == pytest.approx(5.4)


def test_eam_tables_match_analytic(syn):
    """Synthetic file: interpolated tables reproduce the defining functions
    within the linear-interpolation bound h^2/8 max|g''|."""
    _, f, tab = syn
    r = np.linspace(2.0, 5.45, 301)
    fa, phia, _ = lat.w_synthetic_functions(r, np.array([1.0]))
    h = tab.dr
    np.testing.assert_allclose(tab.interp_r(r, tab.dens[0, 0]), fa, atol=h * h / 8 * 2.25 * 4.5 + 1e-9)
    np.testing.assert_allclose(tab.interp_r(r, tab.phi[0, 0]), phia, atol=h * h / 8 * 600 + 1e-8)
    rho = np.linspace(2.0, 30.0, 50)
    hr = tab.drho[0]
    # |F''| = 0.625 rho^-1.5 and |F'''| = 0.9375 rho^-2.5 are largest at rho = 2
    np.testing.assert_allclose(tab.interp_rho(rho, 0), -2.5 * np.sqrt(rho),
                               atol=hr * hr / 8 * 0.625 * 2 ** -1.5 * 1.01 + 1e-9)
    np.testing.assert_allclose(tab.interp_rho(rho, 0, deriv=True), -1.25 / np.sqrt(rho),
                               atol=hr * hr / 8 * 0.9375 * 2 ** -2.5 * 1.01 + 1e-8)


def test_eam_two_atoms_closed_form(syn):
    _, _, tab = syn
    H = np.diag([30.0] * 3)
    r = 2.9
    pos = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
    ei, _ = onb.brute_force_half_list(pos, H, onb.list_radius_sq(5.5, 1.0))
    out = oeam.eam_forward(pos, ei, np.zeros(2, int), tab, H)
    f, phi, _ = lat.w_synthetic_functions(np.array([r]), np.array([1.0]))
    rho = f[0]
    E = 2 * (-2.5 * np.sqrt(rho)) + phi[0]
    # linear interpolation bounds on the rho grid: h^2/8 |F''| and h^2/8 |F'''|
    hr = tab.drho[0]
    eF = hr * hr / 8 * 0.625 * rho ** -1.5
    edF = hr * hr / 8 * 0.9375 * rho ** -2.5
    assert abs(out["energy"] - E) <= 2 * eF * 1.01 + 1e-7
    df = -1.5 * rho
    dphi = 0.6 * (-4 * np.exp(-4 * (r - 2.74)) + 4 * np.exp(-2 * (r - 2.74)))
    s = 2 * (-1.25 / np.sqrt(rho)) * df + dphi          # dE/dr
    tol = 2 * edF * abs(df) * 1.01 + 1e-6
    np.testing.assert_allclose(out["forces"][0], [s, 0, 0], atol=tol)   # -dE/dx0 = +s
    np.testing.assert_allclose(out["forces"][1], [-s, 0, 0], atol=tol)


def _analytic_eam(pos, H, rc=5.5):
    """Closed-form EAM of the synthetic functions (no tables): energy and
    forces by the EAM formula with exact f, f', phi, phi', F, F'."""
    ei, _ = onb.brute_force_half_list(pos, H, np.float64(rc * rc))
    i, j = ei
    rij = obox.minimum_image(pos[i] - pos[j], H)
    r = np.linalg.norm(rij, axis=1)
    t = np.clip((r - 5.0) / 0.5, 0, 1)
    S = 1 - 3 * t * t + 2 * t ** 3
    dS = (-6 * t + 6 * t * t) / 0.5
    ef = np.exp(-1.5 * (r - 2.74))
    f, df = ef * S, -1.5 * ef * S + ef * dS
    p0 = 0.6 * (np.exp(-4 * (r - 2.74)) - 2 * np.exp(-2 * (r - 2.74)))
    dp0 = 0.6 * (-4 * np.exp(-4 * (r - 2.74)) + 4 * np.exp(-2 * (r - 2.74)))
    phi, dphi = p0 * S, dp0 * S + p0 * dS
    n = len(pos)
    rho = np.zeros(n)
    np.add.at(rho, i, f)
    np.add.at(rho, j, f)
    E = (-2.5 * np.sqrt(rho)).sum() + phi.sum()
    Fp = -1.25 / np.sqrt(rho)
    s = (Fp[i] + Fp[j]) * df + dphi
    fv = -s[:, None] * rij / r[:, None]
    F = np.zeros((n, 3))
    np.add.at(F, i, fv)
    np.add.at(F, j, -fv)
    return E, F


def test_eam_analytic_reference_is_consistent():
    """The closed-form helper itself: F = -dE/dx by central differences."""
    pos, H = lat.w_bcc(3, seed=3, amp=0.1)
    E, F = _analytic_eam(pos, H)
    rng = np.random.default_rng(2)
    h = 1e-6
    for i in rng.integers(0, len(pos), 4):
        for a in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, a] += h
            p2[i, a] -= h
            fd = -(_analytic_eam(p1, H)[0] - _analytic_eam(p2, H)[0]) / (2 * h)
            assert fd == pytest.approx(F[i, a], abs=1e-7)


def test_eam_tables_reproduce_closed_form(syn):
    """Oracle (uniform tables, linear interpolation) vs the closed form on a
    jittered bcc cell: energy and forces within the O(h^2) table error.
    Forces come from derivative tables (eam_force_cu.py:535-550), so they
    approximate -dE/dx of the smooth functions, not of the interpolated energy."""
    _, _, tab = syn
    pos, H = lat.w_bcc(4, seed=3, amp=0.1)
    ei, _ = onb.brute_force_half_list(pos, H, onb.list_radius_sq(5.5, 1.0))
    out = oeam.eam_forward(pos, ei, np.zeros(len(pos), int), tab, H)
    E, F = _analytic_eam(pos, H)
    np.testing.assert_allclose(out["forces"].sum(0), 0.0, atol=1e-10)
    # W uses rij/(r + 1e-8) (eam_force_cu.py:566), the scalar virial uses r
    assert np.trace(out["virial_tensor"]) == pytest.approx(-out["virial"], rel=1e-8)
    assert out["energy"] == pytest.approx(E, abs=2e-6 * len(pos))
    err = np.abs(out["forces"] - F).max() / np.abs(F).max()
    assert err < 2e-5, err


def _wre_setup(wre, a=3.1652):
    f = oeam.EAMFile(wre)
    tab = oeam.EAMTables(f, 10000, 10000, dtype=np.float64)
    return tab


def test_eam_wre_zero_stress(wre):
    """cuda_test/test_eam_fs_parser_regression.py:28-46: bcc W 4x4x4 at
    a = 3.1652 is stress free to within 20 bar."""
    tab = oeam.EAMTables(oeam.EAMFile(wre), dtype=np.float64)
    pos, H = lat.w_bcc(4, a=3.1652)
    ei, _ = onb.brute_force_half_list(pos, H, onb.list_radius_sq(tab.cutoff, 1.0))
    out = oeam.eam_forward(pos, ei, np.zeros(len(pos), int), tab, H)
    sig = np.diag(out["virial_tensor"]) / obox.volume(H) * EV_A3_TO_BAR
    assert np.max(np.abs(sig)) < 20.0, sig


def test_eam_wre_elastic_constants(wre):
    """cuda_test/test_w_elastic_static.py:54-90: C11 from energy curvature and
    from stress in 350-700 GPa, C12 in 100-350 GPa, C11 estimates within 8%."""
    tab = _wre_setup(wre)
    pos0, H0 = lat.w_bcc(4, a=3.1652)
    frac = pos0 @ np.linalg.inv(H0)
    V0 = obox.volume(H0)
    eps = 1e-3

    def measure(e11):
        S = np.eye(3)
        S[0, 0] = 1 + e11
        H = S @ H0
        pos = frac @ H
        ei, _ = onb.brute_force_half_list(pos, H, onb.list_radius_sq(tab.cutoff, 1.0))
        out = oeam.eam_forward(pos, ei, np.zeros(len(pos), int), tab, H)
        return out["energy"], np.diag(out["virial_tensor"]) / obox.volume(H) * EV_A3_TO_BAR

    e0, _ = measure(0.0)
    ep, sp = measure(eps)
    em, sm = measure(-eps)
    c11_e = (ep + em - 2 * e0) / (V0 * eps * eps) * EV_A3_TO_GPA
    c11_s = -(sp[0] - sm[0]) / (2 * eps) / 1e4
    c12_s = -(sp[1] - sm[1]) / (2 * eps) / 1e4
    assert 350 < c11_e < 700 and 350 < c11_s < 700 and 100 < c12_s < 350
    assert abs(c11_e - c11_s) / abs(c11_e) < 0.08


# ---- integrator -----------------------------------------------------------------
def test_verlet_constant_force_is_exact():
    """With a constant force velocity Verlet is exact: x = x0 + v0 t + a t^2/2."""
    m = np.full(4, 2.0)
    x = np.zeros((4, 3))
    v = np.ones((4, 3))
    f = np.full((4, 3), 0.3)
    dt, n = 0.01, 50
    for _ in range(n):
        vh, x = oint.first_half(x, v, f, m, dt)
        v, _, _ = oint.second_half(vh, f, m, dt)
    t = n * dt
    np.testing.assert_allclose(x, 1.0 * t + 0.5 * 0.15 * t * t, rtol=1e-12)
    np.testing.assert_allclose(v, 1.0 + 0.15 * t, rtol=1e-12)


def test_verlet_harmonic_energy():
    k, m = 2.0, np.array([1.5])
    x = np.array([[1.0, 0.0, 0.0]])
    v = np.zeros((1, 3))
    dt = 0.01
    E0 = 0.5 * k
    Es = []
    for _ in range(5000):
        vh, x = oint.first_half(x, v, -k * x, m, dt)
        v, ke, _ = oint.second_half(vh, -k * x, m, dt)
        Es.append(ke + 0.5 * k * (x * x).sum())
    assert max(abs(e - E0) for e in Es) < 2e-4 * E0     # bounded O(dt^2) error, no drift


def test_langevin_injected_noise_formula():
    rng = np.random.default_rng(3)
    m = np.full(10, 0.004)
    vh = rng.normal(0, 1, (10, 3))
    f = rng.normal(0, 0.1, (10, 3))
    xi = rng.normal(0, 1, (10, 3))
    g, T, dt = 10.0, 300.0, 0.002
    v, _, _ = oint.second_half(vh, f, m, dt, gamma=g, T=T, noise=xi)
    sig = np.sqrt(oint.KB * T / m * (1 - np.exp(-2 * g * dt)))
    want = vh * np.exp(-g * dt) + xi * sig[:, None] + 0.5 * f / m[:, None] * dt
    np.testing.assert_allclose(v, want, rtol=1e-13)


def test_langevin_equipartition():
    """Free particles under the O step equilibrate to <m v^2> = kB T."""
    rng = np.random.default_rng(4)
    n = 20000
    m = np.full(n, 0.004)
    v = np.zeros((n, 3))
    f = np.zeros((n, 3))
    T = 500.0
    for _ in range(300):
        v, _, _ = oint.second_half(v, f, m, 0.002, gamma=10.0, T=T, noise=rng.normal(0, 1, (n, 3)))
    mv2 = (m[:, None] * v * v).mean()
    assert mv2 == pytest.approx(oint.KB * T, rel=0.02)


def test_berendsen_closed_form():
    K = np.diag([1.0, 2.0, 3.0])
    W = np.diag([0.5, 0.5, 0.5])
    V = 1000.0
    mu, P = oint.berendsen_mu(K, W, V, 0.002, 1.0, 1.0)
    P_want = (6.0 + 1.5) / V / 3.0 * EV_A3_TO_BAR
    assert P == pytest.approx(P_want)
    assert mu ** 3 == pytest.approx(1 + 4.57e-5 * 0.002 * (P_want - 1.0))
    mu_hi, _ = oint.berendsen_mu(K * 1e6, W, V, 0.002, 1.0, 1.0)
    assert mu_hi == pytest.approx(1.05)                 # clamped at (1 + mu_max)^3


def test_maxwell_boltzmann_statistics():
    rng = np.random.default_rng(7)
    m = np.full(50000, lat.W_MASS)
    v = oint.maxwell_boltzmann(m, 300.0, rng, np.float64)
    np.testing.assert_allclose((v * m[:, None]).sum(0), 0.0, atol=1e-6)
    var = v.var()
    assert var == pytest.approx(oint.KB * 300.0 / (lat.W_MASS * oint.MB_MASS_CONV), rel=0.02)
