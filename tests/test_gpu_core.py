"""GPU parity of the MD-step kernels against the CPU oracle (float64).

Bars: neighbor pair sets identical; float32 forces within 1e-5 relative
max-atom error of the float64 oracle on the same configuration; energies and
virials within 1e-6 relative; integrator half steps reproduce the reference
formulas with injected noise.
"""
import numpy as np
import pytest
import torch

from helpers import force_error, oracle_eam, oracle_half_list, oracle_lj, pair_set, sym3
from oracle import integrate as oint
from simulon_amd import lattice as lat
from simulon_amd.engine import ReplicaEngine, ReplicaSpec
from simulon_amd.forcefield import EAMForceField, LJForceField

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
VIR_TOL = 1e-5


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _lj_engine(pos, H, ensemble="NVE", dt=0.002, **kw):
    spec = ReplicaSpec(positions=pos, box=H, dt=dt, ensemble=ensemble, temperature=94.4, **kw)
    eng = ReplicaEngine([spec], [lat.AR_MASS], cutoff=8.5, skin=1.0)
    eng.force_field = LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5)
    return eng


def test_library_version():
    from simulon_amd import _lib
    assert _lib.version().startswith("simulon_hip")


def test_pairs_exact_lj_cubic():
    pos, H = lat.lj_argon_nvt(seed=1)
    _print_case("pairs_lj_cubic", n=len(pos))
    eng = _lj_engine(pos, H)
    eng.setup()
    ei, d = eng.half_list()
    pos = eng.positions().cpu().numpy()       # storage (Morton) order
    ref, dref = oracle_half_list(pos, H, 8.5, 1.0)
    assert pair_set(ei.cpu().numpy()) == pair_set(ref)
    assert np.array_equal(ei.cpu().numpy(), ref)
    np.testing.assert_array_equal(d.cpu().numpy(), dref)
    eng.check()


def test_pairs_exact_triclinic():
    pos, H = lat.lj_argon_npt_tri(ncell=6, seed=2)
    _print_case("pairs_triclinic", n=len(pos))
    eng = _lj_engine(pos, H)
    eng.setup()
    ei, _ = eng.half_list()
    pos = eng.positions().cpu().numpy()
    ref, _ = oracle_half_list(pos, H, 8.5, 1.0)
    assert np.array_equal(ei.cpu().numpy(), ref)


def test_lj_forces_vs_fp64():
    pos, H = lat.lj_argon_nvt(seed=0)
    _print_case("lj_forces", n=len(pos))
    eng = _lj_engine(pos, H)
    eng.setup()
    torch.cuda.synchronize()
    pos = eng.positions().cpu().numpy()
    ref = oracle_lj(pos, H, np.zeros(len(pos), int), [[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, 1.0)
    err = force_error(eng.force.cpu().numpy(), ref["forces"])
    th = eng.thermo_host()[0]
    print(f"force err {err:.3e}  E {th[0]:.9f} vs {ref['energy']:.9f}", flush=True)
    assert err <= FORCE_TOL
    assert abs(th[0] - ref["energy"]) <= 1e-6 * abs(ref["energy"])
    # per-atom virial terms are float32 sums of pair terms of both signs:
    # same 1e-5 relative bar as the forces
    assert abs(th[1] - ref["virial"]) <= VIR_TOL * abs(ref["virial"])
    W = sym3(th[2:8])
    assert np.abs(W - ref["virial_tensor"]).max() <= VIR_TOL * np.abs(ref["virial_tensor"]).max()


@pytest.mark.parametrize("mode", ["shift", "switch"])
def test_lj_smoothing_modes(mode):
    pos, H = lat.lj_argon_nvt(seed=3)
    spec = ReplicaSpec(positions=pos, box=H, dt=0.002)
    eng = ReplicaEngine([spec], [lat.AR_MASS], cutoff=8.5, skin=1.0)
    eng.force_field = LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, mode=mode)
    eng.setup()
    pos = eng.positions().cpu().numpy()
    ref = oracle_lj(pos, H, np.zeros(len(pos), int), [[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, 1.0,
                    mode=mode)
    assert force_error(eng.force.cpu().numpy(), ref["forces"]) <= FORCE_TOL
    th = eng.thermo_host()[0]
    assert abs(th[0] - ref["energy"]) <= 1e-6 * abs(ref["energy"])


def test_lj_triclinic_forces():
    pos, H = lat.lj_argon_npt_tri(ncell=8, seed=5)
    _print_case("lj_triclinic", n=len(pos))
    eng = _lj_engine(pos, H)
    eng.setup()
    pos = eng.positions().cpu().numpy()
    ref = oracle_lj(pos, H, np.zeros(len(pos), int), [[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, 1.0)
    err = force_error(eng.force.cpu().numpy(), ref["forces"])
    print(f"tri force err {err:.3e}", flush=True)
    assert err <= FORCE_TOL


@pytest.fixture(scope="module")
def w_setfl(tmp_path_factory):
    p = tmp_path_factory.mktemp("eam") / "W_synthetic.eam.alloy"
    return str(lat.write_w_synthetic_setfl(p))


def test_eam_forces_vs_fp64(w_setfl):
    pos, H = lat.w_bcc(8, seed=7, amp=0.08)
    _print_case("eam_forces", n=len(pos))
    spec = ReplicaSpec(positions=pos, box=H, dt=0.001)
    eng = ReplicaEngine([spec], [lat.W_MASS], cutoff=5.5, skin=1.0)
    eng.force_field = EAMForceField(w_setfl)
    eng.setup()
    pos = eng.positions().cpu().numpy()
    ref = oracle_eam(pos, H, np.zeros(len(pos), int), w_setfl, 1.0)
    err = force_error(eng.force.cpu().numpy(), ref["forces"])
    th = eng.thermo_host()[0]
    print(f"eam force err {err:.3e} E {th[0]:.8f} vs {ref['energy']:.8f}", flush=True)
    assert err <= FORCE_TOL
    assert abs(th[0] - ref["energy"]) <= 1e-6 * abs(ref["energy"])
    assert abs(th[1] - ref["virial"]) <= VIR_TOL * abs(ref["virial"])
    W = sym3(th[2:8])
    assert np.abs(W - ref["virial_tensor"]).max() <= VIR_TOL * np.abs(ref["virial_tensor"]).max()


def test_eam_lds_tables_match_global(w_setfl):
    """A launch of >= 256 tiles takes the persistent kernels with the pair
    table in LDS; a small one the global tables. Same table values and
    arithmetic: the small replica's forces, energy and virial columns agree
    bit for bit between a packed engine (LDS path) and its own engine (global
    path), and the big replica's forces stay within the float64 tolerance on
    a sample of atoms."""
    small, Hs = lat.w_bcc(8, seed=21, amp=0.08)
    big, Hb = lat.w_bcc(40, seed=22, amp=0.08)
    _print_case("eam_lds_tables", n_small=len(small), n_big=len(big))
    mk = lambda p, H: ReplicaSpec(positions=p, box=H, dt=0.001)
    eng = ReplicaEngine([mk(small, Hs), mk(big, Hb)], [lat.W_MASS], cutoff=5.5, skin=1.0,
                        family="global")
    eng.force_field = EAMForceField(w_setfl)
    eng.setup()
    assert eng.ntiles >= 256
    e1 = ReplicaEngine([mk(small, Hs)], [lat.W_MASS], cutoff=5.5, skin=1.0, family="global")
    e1.force_field = EAMForceField(w_setfl)
    e1.setup()
    assert e1.ntiles < 256
    f_pack = eng.to_caller(eng.force).cpu().numpy()
    np.testing.assert_array_equal(f_pack[:len(small)], e1.to_caller(e1.force).cpu().numpy())
    np.testing.assert_array_equal(eng.thermo_host()[0][:8], e1.thermo_host()[0][:8])
    # big replica: FP64 oracle forces on the full system, compared on all atoms
    pos = eng.positions().cpu().numpy()[eng.replica_slice(1)]
    ref = oracle_eam(pos, Hb, np.zeros(len(pos), int), w_setfl, 1.0)
    err = force_error(eng.force.cpu().numpy()[eng.replica_slice(1)], ref["forces"])
    print(f"eam LDS-table force err {err:.3e}", flush=True)
    assert err <= FORCE_TOL


def test_integrator_injected_noise():
    pos, H = lat.lj_argon_nvt(seed=4)
    eng = _lj_engine(pos, H, ensemble="NVT", gamma=10.0)
    eng.setup()
    x0 = eng.positions().cpu().numpy().copy()
    v0 = eng.vel.cpu().numpy().copy()
    f0 = eng.force.cpu().numpy().copy()
    noise = torch.randn((eng.N, 3), generator=torch.Generator().manual_seed(9)).to(eng.device)
    m = np.full(eng.N, np.float32(eng.types_struct.mass[0]), np.float32)
    eng.step(noise=noise)
    torch.cuda.synchronize()
    vh, x1 = oint.first_half(x0, v0, f0, m, 0.002)
    x1_64 = x0 + vh.astype(np.float64) * 0.002
    np.testing.assert_array_equal(eng.vel_half.cpu().numpy(), vh)
    np.testing.assert_allclose(eng.positions().cpu().numpy(), x1_64, rtol=0, atol=1e-12)
    f1 = eng.force.cpu().numpy()
    v1, ke, T = oint.second_half(vh, f1, m, 0.002, gamma=10.0, T=94.4, noise=noise.cpu().numpy())
    np.testing.assert_allclose(eng.vel.cpu().numpy(), v1, rtol=2e-6, atol=1e-6)
    th = eng.thermo_host()[0]
    assert abs(th[8] - ke) <= 1e-5 * ke


def test_batched_equals_single():
    """Three NVT replicas packed in one engine advance exactly as three
    single-replica engines with the same positions, caller-order velocities,
    per-replica noise seeds and pair-kernel family: 30 graph-replayed steps,
    positions and velocities bit for bit."""
    sizes = [3, 4, 5]
    specs = []
    for k, n in enumerate(sizes):
        pos, H = lat.w_bcc(2 * n, seed=10 + k, amp=0.05)
        specs.append(ReplicaSpec(positions=pos, box=H, dt=0.001, ensemble="NVT",
                                 temperature=300 + 100 * k, gamma=5.0))
    eng = ReplicaEngine(specs, [lat.W_MASS], cutoff=5.0, skin=1.0, seed=77)
    eng.force_field = LJForceField([[0.1]], [[2.5]], 5.0)
    eng.setup()
    v_caller = eng.to_caller(eng.vel).cpu().numpy()
    singles = []
    for k, s in enumerate(specs):
        s2 = ReplicaSpec(positions=s.positions, box=s.box,
                         velocities=v_caller[eng.replica_slice(k)], dt=0.001, ensemble="NVT",
                         temperature=s.temperature, gamma=s.gamma, seed=eng.replica_seed(k))
        e1 = ReplicaEngine([s2], [lat.W_MASS], cutoff=5.0, skin=1.0, family=eng.family)
        e1.force_field = LJForceField([[0.1]], [[2.5]], 5.0)
        e1.setup()
        np.testing.assert_array_equal(e1.force.cpu().numpy(),
                                      eng.force.cpu().numpy()[eng.replica_slice(k)])
        singles.append(e1)
    eng.run(30, graph=True)
    for e1 in singles:
        e1.run(30, graph=True)
    x_all = eng.to_caller(eng.positions(unwrap=True)).cpu().numpy()
    v_all = eng.to_caller(eng.vel).cpu().numpy()
    for k, e1 in enumerate(singles):
        sl = eng.replica_slice(k)
        np.testing.assert_array_equal(e1.to_caller(e1.positions(unwrap=True)).cpu().numpy(),
                                      x_all[sl])
        np.testing.assert_array_equal(e1.to_caller(e1.vel).cpu().numpy(), v_all[sl])
    assert [int(x) for x in eng.nbuild.cpu()] == [int(e.nbuild.cpu()[0]) for e in singles]


def test_replica_noise_streams_independent():
    """Two identical NVT replicas in one engine (default seeds) draw
    independent Langevin kicks: after one step v_a - v_b = rff (xi_a - xi_b),
    whose spread is sqrt(2) rff for independent draws and 0 for shared ones."""
    pos, H = lat.w_bcc(10, seed=4, amp=0.05)
    spec = dict(positions=pos, box=H, velocities=np.zeros((len(pos), 3)), dt=0.001,
                ensemble="NVT", temperature=600.0, gamma=10.0)
    eng = ReplicaEngine([ReplicaSpec(**spec), ReplicaSpec(**spec)], [lat.W_MASS], cutoff=5.0,
                        skin=1.0, seed=3)
    eng.force_field = LJForceField([[0.1]], [[2.5]], 5.0)
    eng.setup()
    eng.step()
    v = eng.to_caller(eng.vel).cpu().numpy().astype(np.float64)
    n = len(pos)
    rff = float(eng.replica_state(0).rff[0])
    d = (v[:n] - v[n:]) / rff
    ratio = float(d.std()) / np.sqrt(2.0)
    print(f"kick difference spread / sqrt(2): {ratio:.4f}", flush=True)
    assert 0.96 < ratio < 1.04


def test_nve_trajectory_vs_fp64_oracle():
    """100 graph-replayed NVE steps (raw cutoff) follow the float64 oracle's
    total-energy trajectory; the raw cutoff itself does not conserve energy."""
    from oracle import lj as olj
    from oracle import neighbor as onb
    pos, H = lat.lj_argon_nvt(seed=6)
    eng = _lj_engine(pos, H, ensemble="NVE", dt=0.002, t_init=94.4)
    eng.setup()
    x = eng.positions().cpu().numpy().copy()
    v = eng.vel.cpu().numpy().astype(np.float64)
    m = np.full(len(x), float(eng.types_struct.mass[0]))
    rl2 = onb.list_radius_sq(8.5, 1.0)

    def force(x):
        ei, _ = onb.brute_force_half_list(x, H, rl2)
        return olj.lj_forward(x, ei, np.zeros(len(x), int), [[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, H)

    out = force(x)
    f = out["forces"]
    e_or = []
    for _ in range(100):
        vh = v + 0.5 * f / m[:, None] * 0.002
        x = x + vh * 0.002
        out = force(x)
        f = out["forces"]
        v = vh + 0.5 * f / m[:, None] * 0.002
        e_or.append(out["energy"] + 0.5 * (m[:, None] * v * v).sum())
    e_gpu = []
    for _ in range(100):
        eng.run(1, graph=True, check=False)
        th = eng.thermo_host()[0]
        e_gpu.append(th[0] + th[8])
    eng.check()
    d = np.max(np.abs(np.array(e_gpu) - np.array(e_or)))
    print(f"NVE max |E_gpu - E_fp64| over 100 steps: {d:.3e} eV (drift {e_or[-1]-e_or[0]:.3e})",
          flush=True)
    assert d < 1e-3
    assert int(eng.nbuild.cpu()[0]) >= 2


def test_nve_conservation_force_shift():
    pos, H = lat.lj_argon_nvt(seed=6)
    spec = ReplicaSpec(positions=pos, box=H, dt=0.002, ensemble="NVE", t_init=94.4)
    eng = ReplicaEngine([spec], [lat.AR_MASS], cutoff=8.5, skin=1.0)
    eng.force_field = LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, mode="shift")
    eng.setup()
    th = eng.thermo_host()[0]
    E0 = th[0] + th[8]
    eng.run(500, graph=True)
    th = eng.thermo_host()[0]
    E1 = th[0] + th[8]
    print(f"force-shift NVE 1 ps: E0 {E0:.6f} E1 {E1:.6f} (KE {th[8]:.3f})", flush=True)
    assert abs(E1 - E0) < 2e-3 * th[8]


def test_caller_order_roundtrip():
    pos, H = lat.w_bcc(5, seed=3, amp=0.05)
    eng = ReplicaEngine([ReplicaSpec(positions=pos, box=H)], [lat.W_MASS], cutoff=5.0, skin=1.0)
    back = eng.to_caller(eng.positions()).cpu().numpy()
    np.testing.assert_array_equal(back, pos)
    assert not np.array_equal(eng.positions().cpu().numpy(), pos)   # storage order is Morton


def test_graph_matches_eager():
    pos, H = lat.lj_argon_nvt(seed=8)
    a = _lj_engine(pos, H, ensemble="NVT", gamma=10.0)
    b = _lj_engine(pos, H, ensemble="NVT", gamma=10.0)
    b.vel.copy_(a.vel)
    a.setup()
    b.setup()
    a.run(20, graph=False)
    b.run(20, graph=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(a.pos4.cpu().numpy(), b.pos4.cpu().numpy())
    np.testing.assert_array_equal(a.vel.cpu().numpy(), b.vel.cpu().numpy())


def test_pos_quantize_records():
    """sim_pos_quantize: fractional coordinates of the wrapped positions in
    units of 2^-32 (within one unit of a float64 numpy restatement, modulo
    2^32), the type in w, and the wrapped difference of two records equal to
    the fractional minimum image of core/box.py."""
    rng = np.random.default_rng(4)
    pos, H = lat.lj_argon_npt_tri(ncell=6, seed=4)
    _print_case("pos_quantize", n=len(pos))
    eng = _lj_engine(pos, H)
    eng.setup()
    # move some atoms out of the primary cell (unwrapped coordinates)
    shift = torch.as_tensor(rng.integers(-2, 3, size=(eng.N, 3)) @ np.asarray(H, np.float64),
                            device=eng.device)
    eng.pos4[:, :3] += shift
    eng.quantize()
    torch.cuda.synchronize()
    q = eng.posq.cpu().numpy().astype(np.int64)
    x = eng.pos4[:, :3].cpu().numpy()
    s = x @ np.linalg.inv(np.asarray(H, np.float64))
    s -= np.floor(s)
    ref = np.rint(s * 2.0 ** 32).astype(np.int64) % (1 << 32)
    got = q[:, :3] % (1 << 32)
    d = (got - ref + (1 << 31)) % (1 << 32) - (1 << 31)
    assert np.abs(d).max() <= 1
    assert np.array_equal(q[:, 3], eng.pos4[:, 3].cpu().numpy().astype(np.int64))
    # record differences = fractional minimum image of the float64 displacement
    i = rng.integers(0, eng.N, 500)
    j = rng.integers(0, eng.N, 500)
    dq = ((got[i] - got[j] + (1 << 31)) % (1 << 32) - (1 << 31)) * 2.0 ** -32
    ds = s[i] - s[j]
    ds -= np.round(ds)
    np.testing.assert_allclose(dq, ds, rtol=0, atol=2 * 2.0 ** -32)


@pytest.mark.parametrize("ff", ["lj", "eam"])
def test_tiled_matches_global_list_kernels(ff, w_setfl):
    """Cell-tiled (LDS-staged, float64 positions) kernels vs the global-list
    kernels (fixed-point position records) on the same configuration: forces
    agree to float32 rounding of the displacements and summation order."""
    if ff == "lj":
        pos, H = lat.lj_argon_npt_tri(ncell=8, seed=9)
        mk = lambda: LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5)
        mass, rc = lat.AR_MASS, 8.5
    else:
        pos, H = lat.w_bcc(10, seed=9, amp=0.1)
        mk = lambda: EAMForceField(w_setfl)
        mass, rc = lat.W_MASS, 5.5
    out = {}
    for fam in ("cell", "rows", "global"):
        eng = ReplicaEngine([ReplicaSpec(positions=pos, box=H)], [mass], cutoff=rc, skin=1.0,
                            family=fam)
        eng.force_field = mk()
        eng.setup()
        assert eng.family == fam and eng.use_tiled == (fam != "global")
        out[fam] = (eng.to_caller(eng.force).cpu().numpy(), eng.thermo_host()[0])
    f_g, th_g = out["global"]
    for fam in ("cell", "rows"):
        f_t, th_t = out[fam]
        err = force_error(f_t, f_g)
        print(f"{ff}: {fam} vs global-list force diff {err:.2e}", flush=True)
        assert err < 5e-6
        assert abs(th_t[0] - th_g[0]) <= 1e-7 * abs(th_g[0])
        np.testing.assert_allclose(th_t[2:8], th_g[2:8], rtol=0,
                                   atol=2e-6 * np.abs(th_g[2:8]).max())
    # cell and row kernels visit each list in the same order: equal bit for bit
    assert np.array_equal(out["cell"][0], out["rows"][0])


@pytest.mark.parametrize("fam", ["cell", "rows"])
@pytest.mark.parametrize("ff", ["lj", "eam"])
def test_tiled_stencil_overflow_reads_global_memory(ff, fam, w_setfl):
    """With an LDS column capacity too small for the stencil columns, every
    cell takes the global-memory path through the same column table: same
    forces and sums, bit for bit, and the overflow is reported but not raised."""
    if ff == "lj":
        pos, H = lat.lj_argon_npt_tri(ncell=8, seed=9)
        mk = lambda: LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5)
        mass, rc = lat.AR_MASS, 8.5
    else:
        pos, H = lat.w_bcc(10, seed=9, amp=0.1)
        mk = lambda: EAMForceField(w_setfl)
        mass, rc = lat.W_MASS, 5.5
    out = []
    for small in (False, True):
        eng = ReplicaEngine([ReplicaSpec(positions=pos, box=H)], [mass], cutoff=rc, skin=1.0,
                            family=fam)
        eng.force_field = mk()
        eng.build_list(force=True)
        eng.pick_cap()
        small_cap = 192 if fam == "rows" else 512
        assert eng.use_tiled and eng.cap > small_cap
        if small:
            eng.cap = small_cap
        eng.compute_forces(virial=True)
        eng.reduce(0, 8)
        ov = int(eng.overflow.item())
        assert bool(ov & 2) == small
        eng.check()
        assert eng.profile.get("lds_overflow", 0) == int(small)
        out.append((eng.force.clone(), eng.thermo[0, :8].clone()))
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])


def test_rebuild_mid_step_keeps_atoms_consistent():
    """A rebuild between the two halves permutes every per-atom array (the
    half-step velocities included): an NVE run with frequent rebuilds on the
    tiled path follows the global-list path."""
    pos, H = lat.lj_argon_nvt(seed=12)
    runs = {}
    for fam in ("cell", "rows", "global"):
        eng = ReplicaEngine([ReplicaSpec(positions=pos, box=H, dt=0.004, ensemble="NVE",
                                         t_init=300.0)], [lat.AR_MASS], cutoff=8.5, skin=0.3,
                            family=fam, velocity_seed=5)
        eng.force_field = LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, mode="shift")
        eng.setup()
        eng.run(60, graph=True)
        th = eng.thermo_host()[0]
        runs[fam] = (th[0] + th[8], int(eng.nbuild.cpu()[0]),
                     eng.to_caller(eng.positions(unwrap=True)).cpu().numpy())
    e_g, nb_g, x_g = runs["global"]
    for fam in ("cell", "rows"):
        e_t, nb_t, x_t = runs[fam]
        print(f"E {fam} {e_t:.6f} global {e_g:.6f} rebuilds {nb_t}/{nb_g}", flush=True)
        assert nb_t >= 5
        assert abs(e_t - e_g) < 1e-4 * abs(e_g)
        assert np.abs(x_t - x_g).max() < 1e-3
