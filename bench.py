"""Benchmark: atom-steps/s of the MD step on one MI355X.

Default workload (w_eam_replicas): 64 W EAM replicas of uneven size (2k..128k
atoms, whole bcc cells, sizes drawn log-uniform with seed 3), Langevin at
100..1500 K (gamma 10/ps), dt = 1 fs, synthetic setfl tables (5000 points:
f = exp(-1.5(r-2.74)), phi = 0.6[exp(-4(r-2.74)) - 2 exp(-2(r-2.74))],
F = -2.5 sqrt(rho), cubic taper 5.0-5.5 A), cutoff 5.5 A, skin 1.0 A.

  batched  all replicas in one packed engine, one graph replay per step
  plain    the same replicas advanced one after another, each by its own
           single-system engine (graph-replayed too), as a per-replica scan does

`value` is the batched whole-job atom-steps/s. Also reported: the plain-loop
rate, the speedup, ns/day, the float32 max-atom force error against the
float64 oracle on sampled replicas, the roofline of the dominant kernel
(EAM force pass, timed with HIP events on its stream), and the oracle's CPU
rate on a bounded sample (host cores).

Other workloads (--workload): lj_argon_nvt (1000 Ar), w_eam_nve (54,000 W),
lj_argon_npt_tri (1,048,576 Ar, triclinic, Berendsen + Langevin), mlp_w (the
learned MLP potential trained on EAM frames, run on 128,000 W atoms with EAM
as the judge every 50 steps, against plain EAM).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import tempfile
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from simulon_amd import lattice as lat  # noqa: E402
from simulon_amd.engine import ReplicaEngine, ReplicaSpec  # noqa: E402
from simulon_amd.forcefield import EAMForceField, LJForceField  # noqa: E402

HBM_PEAK_GBS = 8000.0
# run lengths of the configured workloads (timed steps after warm-up)
DEFAULT_STEPS = {"w_eam_replicas": 1000, "lj_argon_nvt": 1000, "w_eam_nve": 2000,
                 "lj_argon_npt_tri": 500, "mlp_w": 500}
METRIC = "atom-steps/s on 1 GPU: plain vs 64-replica batched EAM; force err vs FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, steps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn(steps)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def _traffic_record(workload, kernel):
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)["kernels"].get(f"{workload}/{kernel}")
    except (OSError, ValueError, KeyError):
        return None


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    counter summary (rocprofv3 FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a copy of known size), or None. The
    counter runs are short (5 steps), so their lists are the ones built at the
    start; traffic_ratio gives the measured bytes over that run's algorithmic
    bytes, which carries over to a later list state."""
    rec = _traffic_record(workload, kernel)
    return None if rec is None else int(rec["bytes_per_launch"])


def traffic_ratio(workload, kernel):
    rec = _traffic_record(workload, kernel)
    return None if rec is None else round(rec["bytes_per_launch"] / rec["algorithmic_bytes"], 4)


def step_roofline(natoms, ms_per_step, bytes_per_atom_step):
    """Whole-step algorithmic bytes per atom-step (integrator 88 B + the force
    field's passes) over the step time, against 8 TB/s."""
    ach = natoms * bytes_per_atom_step / (ms_per_step * 1e-3) / 1e9
    return {"bytes_per_atom_step": round(bytes_per_atom_step, 1), "achieved": round(ach, 1),
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4)}


def mean_full_list(eng):
    return float(eng.nnbr.double().mean().item())


def kernel_time_ms(launch, reps=50):
    """Average duration of one launch, HIP events on the launching stream."""
    launch()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        launch()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / reps


def lockstep_rows(eng, counts):
    """Mean over the waves (64 consecutive atoms of a tile) of the longest list
    in the wave: the list rows a lock-step walk visits per atom."""
    c = counts.cpu().numpy()
    ends = np.array([eng.atom_base[r] + eng.natoms[r] for r in range(eng.R)])
    rows = []
    for r, s0 in eng.tiles_np:
        for w0 in range(int(s0), min(int(s0) + 256, int(ends[r])), 64):
            w1 = min(w0 + 64, int(ends[r]))
            rows.append((int(c[w0:w1].max()), w1 - w0))
    rows = np.array(rows)
    return round(float((rows[:, 0] * rows[:, 1]).sum() / rows[:, 1].sum()), 2)


def eam_kernel_roofline(eng, workload="w_eam_replicas", virial=False):
    """Roofline of the EAM passes with the algorithmic bytes per atom:
    pass 1: 12 (x) + 4 nbar (list) + 4 (F' write); pass 2: 12 + 4 nbar + 4 (F'
    read) + 12 (F write), and for the global list, whose force pass also does
    the step's second half (sim_eam_force_second), + 12 (v_half read) + 12 (v
    write), + 124 when graph replays let it do the next step's first half
    too. The dominant (longer) pass is reported; the kernels timed are the
    ones the step launches (row kernels, or the global-list density pass and
    fused force pass), with the virial as asked. Repeating the fused pass is
    idempotent: v is rewritten from the same v_half, forces and noise (the
    step counter does not advance)."""
    from simulon_amd._lib import call, ptr, stream_ptr
    ff = eng.force_field
    nbar = mean_full_list(eng)

    fam = eng.family if eng.use_tiled else "global"
    if fam != "global":
        rho = eng.scratch("eam_rho", torch.float32)
        fp = eng.scratch("eam_fp", torch.float32)
        common = (*(eng.row_args() if fam == "rows" else eng.cell_args()), ff.params)

        def p1():
            call(f"sim_eam_density_{fam}", *common, ptr(rho), ptr(fp), ptr(eng.partial_cell),
                 ptr(eng.overflow), stream_ptr())

        def p2():
            call(f"sim_eam_force_{fam}", *common, ptr(fp), ptr(eng.force), ptr(eng.partial_cell),
                 int(virial), ptr(eng.overflow), stream_ptr())
    else:
        rec = eng.scratch("eam_rec", torch.int32, (eng.N, 4))
        eng.quantize()
        st = eng.stage_arrays()
        stp = ctypes.byref(st) if st is not None else None

        def p1():      # as the step launches it: records with F', no rho / F' arrays
            if stp is None:
                call("sim_eam_density", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles,
                     ptr(eng.posq), ptr(eng.nbr), ptr(eng.nnbr), eng.N, ff.params, None, None,
                     ptr(rec), ptr(eng.partial), stream_ptr())
            else:
                call("sim_eam_density_staged", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles,
                     ptr(eng.posq), ptr(eng.nbr), ptr(eng.nnbr), eng.N, ff.params, None, None,
                     ptr(rec), ptr(eng.partial), stp, stream_ptr())

        # graph replays of two or more steps launch the variant that also does
        # the next step's first half: timed on copies of the positions,
        # records and displacement maxima, with v_half restored afterwards
        fuse = eng.fuses_next_first()
        if fuse:
            nxt_pos, nxt_q = eng.pos4.clone(), eng.posq.clone()
            nxt_d2 = torch.zeros_like(eng.maxd2)
            nxt = (ptr(nxt_pos), ptr(eng.pos_last), ptr(nxt_d2), ptr(nxt_q))
            saved = (eng.vel_half.clone(), eng.vel.clone())
        else:
            nxt = (None, None, None, None)

        def p2():
            call("sim_eam_force_second", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles,
                 ptr(eng.posq), ptr(rec), ptr(eng.nbr), ptr(eng.nnbr), eng.N, ff.params,
                 ptr(eng.force), ptr(eng.partial), int(virial), *eng.second_args(), stp, *nxt,
                 stream_ptr())

    t1, t2 = kernel_time_ms(p1), kernel_time_ms(p2)
    if fam == "global" and fuse:
        eng.vel_half.copy_(saved[0])
        eng.vel.copy_(saved[1])
    extra = {"lockstep_rows": lockstep_rows(eng, eng.nnbr)} if fam == "global" else {}
    b1 = eng.N * (12 + 4 * nbar + 4)
    # + 124 with the next first half: x and the build position read (2 x 32),
    # v_half, x and the record written (12 + 32 + 16)
    b2 = eng.N * (12 + 4 * nbar + 4 + 12 + (24 if fam == "global" else 0)
                  + (124 if fam == "global" and fuse else 0))
    pre = "eam_force" + ("" if fam == "global" else "_" + fam)
    pre1 = "eam_density" + ("" if fam == "global" else "_" + fam)
    name, t, b = (pre, t2, b2) if t2 >= t1 else (pre1, t1, b1)
    ach = b / (t * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(workload, name),
            "traffic_ratio": traffic_ratio(workload, name),
            "bytes_per_launch": int(b), "avg_launch_ms": round(t, 5), "nbar_full": round(nbar, 2),
            "other_pass_ms": round(t1 if t == t2 else t2, 5), **extra}


def lj_kernel_roofline(eng, workload="lj_argon_npt_tri", virial=True):
    """Roofline of the LJ force kernel the step launches (with the virial
    only when the step computes it: NPT); bytes per atom 12 + 4 nbar + 12,
    + 24 (v_half read, v write) for the global-list pass, which does the
    step's second half too (sim_lj_force_second). Repeating it is idempotent:
    v is rewritten from the same v_half, forces and noise (the step counter
    does not advance)."""
    from simulon_amd._lib import call, ptr, stream_ptr
    ff = eng.force_field
    nbar = mean_full_list(eng)

    if eng.use_tiled:
        def k():
            ff.compute(eng, virial)
    else:
        eng.quantize()

        def k():
            call("sim_lj_force_second", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles,
                 ptr(eng.posq), ptr(eng.nbr), ptr(eng.nnbr), eng.N, ff.params, ptr(eng.force),
                 ptr(eng.partial), int(virial), *eng.second_args(), stream_ptr())

    t = kernel_time_ms(k)
    b = eng.N * (12 + 4 * nbar + 12 + (0 if eng.use_tiled else 24))
    ach = b / (t * 1e-3) / 1e9
    name = "lj_force" + (f"_{eng.family}" if eng.use_tiled else "")
    return {"bound": "hbm", "kernel": name, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(workload, name),
            "traffic_ratio": traffic_ratio(workload, name),
            "bytes_per_launch": int(b), "avg_launch_ms": round(t, 5), "nbar_full": round(nbar, 2)}


# ---------------------------------------------------------------------------
def replica_specs(nrep=64, state=None):
    """The 64-replica scan; state: an npz written by --state-out (positions and
    velocities of a finished run, caller order) to start from instead of the
    lattices."""
    specs = [ReplicaSpec(positions=pos, box=H, dt=0.001, ensemble="NVT", temperature=T,
                         gamma=10.0) for pos, H, T in lat.replica_set(nrep)]
    if state is not None:
        st = np.load(state)
        base = np.concatenate([[0], np.cumsum([len(s.positions) for s in specs])])
        if int(base[-1]) != len(st["pos"]):
            raise SystemExit(f"{state}: {len(st['pos'])} atoms, the scan has {base[-1]}")
        for r, s in enumerate(specs):
            s.positions = st["pos"][base[r]:base[r + 1]]
            s.velocities = st["vel"][base[r]:base[r + 1]]
    return specs


def force_error_sample(eng, setfl, reps):
    """float32 GPU forces vs the float64 oracle (oracle/cell_md.c) on the
    replicas `reps`, at the positions the run ended with."""
    from oracle import cell as ocell
    from oracle import eam as oeam
    tab = oeam.EAMTables(oeam.EAMFile(setfl), dtype=np.float64)
    eng.build_list(force=True)
    eng.compute_forces(virial=True)
    torch.cuda.synchronize()
    worst = 0.0
    f_all = eng.force.cpu().numpy()
    x_all = eng.positions().cpu().numpy()
    for r in reps:
        sl = eng.replica_slice(r)
        x = x_all[sl]
        H = eng.box(r)
        ref = ocell.eam(x, H, None, tab, 1.0)["forces"]
        e = float(np.max(np.linalg.norm(f_all[sl] - ref, axis=1)) /
                  np.max(np.linalg.norm(ref, axis=1)))
        log(f"  force err replica {r} (N={len(x)}): {e:.3e}")
        worst = max(worst, e)
    return worst


def host_cpu():
    """(CPU model name, threads this process may run on, os.cpu_count())."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return model, avail, os.cpu_count()


def cpu_baseline_eam(setfl, seconds=15.0, nrep=64):
    """The reference's CPU step in torch float32 (oracle/torch_ref.py: BAOB
    Langevin, skin trigger, edge distances, edge-list EAM with scatter_add_)
    on all the host threads this process may use, run replica by replica as
    the reference's scan does, on a subsample of the scan (its smallest and
    its median replica, about seconds/2 each), each on its half list built
    once (as between list rebuilds)."""
    from oracle import eam as oeam
    from oracle import integrate as oint
    from oracle import neighbor as onb
    from oracle import torch_ref as otr
    model, avail, visible = host_cpu()
    torch.set_num_threads(avail)
    tab = oeam.EAMTables(oeam.EAMFile(setfl), dtype=np.float32)
    rs = lat.replica_set(nrep)
    sizes = np.array([len(p) for p, _, _ in rs])
    pick = [int(np.argmin(sizes)), int(np.argsort(sizes)[len(sizes) // 2])]
    work, el, parts = 0, 0.0, []
    for r in pick:
        pos, H, T = rs[r]
        x = pos.astype(np.float32)
        n = len(x)
        ei, _ = onb.cell_half_list(x, H, onb.list_radius_sq(5.5, 1.0))
        v = oint.maxwell_boltzmann(np.full(n, lat.W_MASS), T, np.random.default_rng(r))
        st = otr.TorchEAMStep(x, v, lat.W_MASS, H, ei, tab, dt=0.001, skin=1.0, gamma=10.0, T=T,
                              seed=r)
        st.step()                                   # warm-up
        steps = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < seconds / len(pick):
            st.step()
            steps += 1
        dt = time.perf_counter() - t0
        work += n * steps
        el += dt
        parts.append(f"replica {r} ({n} atoms, {T:.0f} K): {steps} steps in {dt:.1f} s")
    return {"value": round(work / el, 1), "unit": "atom-steps/s", "cores": torch.get_num_threads(),
            "kind": "port",
            "sample": "the reference's CPU step in torch float32 (BAOB Langevin + skin trigger + "
                      "edge distances + edge-list EAM with scatter_add_), replica by replica: "
                      + "; ".join(parts) + "; each on its half list built once",
            "cpu_model": model, "os_cpu_count": visible, "torch_threads": torch.get_num_threads()}


def cpu_baseline_lj(pos, H, dt, seconds=15.0, mode="raw"):
    """The oracle's float32 LJ + BAOB step (edge list built once, as between
    list rebuilds) on `pos`, timed on the host."""
    from oracle import integrate as oint
    from oracle import lj as olj
    from oracle import neighbor as onb
    x = np.asarray(pos, np.float32)
    n = len(x)
    ei, _ = onb.cell_half_list(x, H, onb.list_radius_sq(8.5, 1.0))
    types = np.zeros(n, int)
    m = oint.internal_mass(np.full(n, lat.AR_MASS))
    rng = np.random.default_rng(0)
    v = oint.maxwell_boltzmann(np.full(n, lat.AR_MASS), 94.4, rng)
    ff = lambda x_: olj.lj_forward(x_, ei, types, [[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, H,
                                   mode=mode)["forces"]
    f = ff(x)
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        vh, x = oint.first_half(x, v, f, m, dt)
        f = ff(x)
        v, _, _ = oint.second_half(vh, f, m, dt, gamma=10.0, T=94.4,
                                   noise=rng.standard_normal((n, 3)).astype(np.float32))
        steps += 1
    el = time.perf_counter() - t0
    # the reference's list build (blocked O(N^2) brute force, gpu_kdtree.py:90-156)
    # on the same atoms, timed once
    tb = time.perf_counter()
    onb.brute_force_half_list(x.astype(np.float32), H, onb.list_radius_sq(8.5, 1.0))
    tb = time.perf_counter() - tb
    return {"value": round(steps * n / el, 1), "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle float32 LJ + BAOB step, {n} Ar atoms, {steps} steps in {el:.1f} s, "
                      f"list built once (host {os.cpu_count()} cores visible, numpy "
                      f"single-threaded)",
            "list_build": {"algorithm": "O(N^2) blocked brute force, the reference's",
                           "atoms": n, "seconds": round(tb, 3)}}


def bench_replicas(args):
    tmp = tempfile.mkdtemp()
    setfl = lat.write_w_synthetic_setfl(os.path.join(tmp, "W_synthetic.eam.alloy"))
    specs = replica_specs(args.replicas, args.state_in)
    natoms = np.array([len(s.positions) for s in specs])
    log(f"replicas {len(specs)} atoms total {natoms.sum()} min {natoms.min()} max {natoms.max()}")
    eng = ReplicaEngine(specs, [lat.W_MASS], cutoff=5.5, skin=args.skin, seed=2024,
                        family=args.family, stage=not args.no_stage)
    eng.force_field = EAMForceField(setfl, device=eng.device)
    eng.setup()
    # every step computes the virial tensor, as the reference's does
    # (core/md_model.py:146-162, core/force/eam_force_cu.py:566-570)
    eng.run(args.warmup, graph=True, virial=True)
    b0 = eng.nbuild.cpu().numpy().copy()
    t = timed(lambda k: eng.run(k, graph=True, virial=True, check=False), args.steps)
    eng.check()
    log(f"pair kernels: {eng.family} (LDS cap {eng.cap}); autotune {eng.profile.get('autotune_ms')}")
    batched = natoms.sum() * args.steps / t
    builds = eng.nbuild.cpu().numpy() - b0
    log(f"batched: {t * 1e3 / args.steps:.3f} ms/step, {batched:.4e} atom-steps/s, "
        f"rebuilds per replica in the timed steps: min {builds.min()} max {builds.max()} "
        f"total {builds.sum()}")
    if args.state_out:
        np.savez(args.state_out, pos=eng.to_caller(eng.positions(unwrap=True)).cpu().numpy(),
                 vel=eng.to_caller(eng.vel).cpu().numpy())
        log(f"state after the timed steps written to {args.state_out}")
    roof = eam_kernel_roofline(eng, virial=True)       # the instance the step launches
    log(f"roofline {roof}")
    # plain per-replica loop on the same GPU
    plain_t = 0.0
    plain_steps = args.plain_steps or args.steps
    for k, s in enumerate(specs):
        e1 = ReplicaEngine([s], [lat.W_MASS], cutoff=5.5, skin=args.skin, seed=2024 + k)
        e1.force_field = eng.force_field
        e1.setup()
        e1.run(2, graph=True, virial=True)
        plain_t += timed(lambda k_: e1.run(k_, graph=True, virial=True, check=False), plain_steps)
        e1.check()
        del e1
    plain = natoms.sum() * plain_steps / plain_t
    log(f"plain loop: {plain_t * 1e3 / plain_steps:.3f} ms per all-replica step, {plain:.4e} "
        f"atom-steps/s, speedup {batched / plain:.2f}x")
    order = np.argsort(natoms)
    sample = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    ferr = force_error_sample(eng, setfl, sample) if not args.no_check else None
    out = {
        "metric": METRIC, "value": round(float(batched), 1), "unit": "atom-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t * 1e3 / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "w_eam_replicas", "replicas": len(specs),
                   "atoms_total": int(natoms.sum()), "atoms_min": int(natoms.min()),
                   "atoms_max": int(natoms.max()), "dt_fs": 1.0, "temperatures_K": [100, 1500],
                   "gamma_per_ps": 10.0, "cutoff": 5.5, "skin": args.skin, "size_seed": 3,
                   "eam_table": "synthetic setfl, 5000 points"},
        "plain_loop_atom_steps_per_s": round(float(plain), 1),
        "speedup_vs_plain": round(float(batched / plain), 3),
        "ns_per_day": round(args.steps * 0.001e-3 * 86400 / t, 3),
        "force_err_vs_fp64": ferr,
        "force_err_replicas": [int(natoms[r]) for r in sample],
        "rebuilds_in_timed_steps": {"per_replica_min": int(builds.min()),
                                    "per_replica_max": int(builds.max()),
                                    "total": int(builds.sum())},
        "virial": "every step (tensor)",
        "roofline": roof,
        "step_roofline": step_roofline(int(natoms.sum()), t * 1e3 / args.steps,
                                       132 + 8 * roof["nbar_full"]),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_eam(setfl, seconds=args.cpu_seconds,
                                               nrep=args.replicas)
    return out


def bench_single(args):
    w = args.workload
    if w == "lj_argon_nvt":
        pos, H = lat.lj_argon_nvt(seed=0)
        spec = ReplicaSpec(positions=pos, box=H, dt=0.002, ensemble="NVT", temperature=94.4,
                           gamma=10.0)
        eng = ReplicaEngine([spec], [lat.AR_MASS], cutoff=8.5, skin=1.0)
        eng.force_field = LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5)
        dt_ps = 0.002
    elif w == "lj_argon_npt_tri":
        pos, H = lat.lj_argon_npt_tri(ncell=args.fcc_cells, seed=0)
        spec = ReplicaSpec(positions=pos, box=H, dt=0.002, ensemble="NPT", temperature=90.0,
                           gamma=10.0, target_pressure=1.0, tau_p=1.0)
        eng = ReplicaEngine([spec], [lat.AR_MASS], cutoff=8.5, skin=1.0)
        eng.force_field = LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5)
        dt_ps = 0.002
    elif w == "w_eam_nve":
        tmp = tempfile.mkdtemp()
        setfl = lat.write_w_synthetic_setfl(os.path.join(tmp, "W_synthetic.eam.alloy"))
        pos, H = lat.w_bcc(30)
        spec = ReplicaSpec(positions=pos, box=H, dt=0.001, ensemble="NVE", temperature=300.0)
        eng = ReplicaEngine([spec], [lat.W_MASS], cutoff=5.5, skin=1.0)
        eng.force_field = EAMForceField(setfl, device=eng.device)
        dt_ps = 0.001
    else:
        raise SystemExit(f"unknown workload {w}")
    eng.setup()
    virial = eng.has_barostat
    eng.run(args.warmup, graph=True, virial=virial)
    t = timed(lambda k: eng.run(k, graph=True, virial=virial, check=False), args.steps)
    eng.check()
    rate = eng.N * args.steps / t
    drift = None
    if spec.ensemble == "NVE":
        # total-energy drift over 2000 further steps (2 ps at 1 fs)
        th = eng.thermo_host()[0]
        e0 = th[0] + th[8]
        eng.run(2000, graph=True, virial=False)
        th = eng.thermo_host()[0]
        drift = ((th[0] + th[8]) - e0) / eng.N / (2000 * dt_ps * 1e-3)
        log(f"NVE: E/atom {e0 / eng.N:.6f} -> {(th[0] + th[8]) / eng.N:.6f} eV, drift "
            f"{drift:.3e} eV/atom/ns")
    roof = lj_kernel_roofline(eng, w, virial=virial) \
        if isinstance(eng.force_field, LJForceField) else eam_kernel_roofline(eng, w, virial=virial)
    cpu = None
    if not args.no_cpu:
        if w == "lj_argon_nvt":
            cpu = cpu_baseline_lj(pos, H, dt_ps, seconds=args.cpu_seconds)
        elif w == "lj_argon_npt_tri":   # a 2,048-atom block of the same lattice and shear
            cpu = cpu_baseline_lj(*lat.lj_argon_npt_tri(ncell=8, seed=0), dt_ps,
                                  seconds=args.cpu_seconds)
        else:
            cpu = cpu_baseline_eam(setfl, seconds=args.cpu_seconds)
    return {"metric": f"atom-steps/s ({w})", "value": round(float(rate), 1), "unit": "atom-steps/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(t * 1e3 / args.steps, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": w, "atoms": eng.N},
            "ns_per_day": round(args.steps * dt_ps * 1e-3 * 86400 / t, 3),
            "rebuilds": int(eng.nbuild.cpu()[0]), "roofline": roof,
            "step_roofline": step_roofline(
                eng.N, t * 1e3 / args.steps,
                (12 + 4 * roof["nbar_full"] + 12 + 88) if isinstance(eng.force_field, LJForceField)
                else 132 + 8 * roof["nbar_full"]),
            "nve_drift_eV_per_atom_per_ns": drift, "cpu_baseline": cpu}


def mlp_kernel_roofline(eng):
    """Roofline of the MLP passes with the algorithmic bytes per atom:
    energy pass 12 (x) + 4 nbar (list) + 128 (dE/dg write); force pass
    12 + 4 nbar + 128 (dE/dg read) + 12 (F write). Also the dense-layer MFMA
    rate of the energy pass: 2 (32*64 + 64*64 + 64) flop forward and
    2 (64*64 + 64*32) + 64 backward to the input per atom, twice for the
    hi/lo activation split."""
    from simulon_amd._lib import call, ptr, stream_ptr
    ff = eng.force_field
    nbar = mean_full_list(eng)
    c = eng.scratch("mlp_c", torch.float32, (eng.N, 32))
    part, _ = ff.partials(eng)
    st = ff.net.struct()
    eng.quantize()

    def p1():
        call("sim_mlp_energy", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq),
             ptr(eng.nbr), ptr(eng.nnbr), eng.N, st, ptr(c), None, None, ptr(part), stream_ptr())

    def p2():
        call("sim_mlp_force", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq),
             ptr(eng.nbr), ptr(eng.nnbr), eng.N, st, ptr(c), ptr(eng.force), ptr(part), 0,
             stream_ptr())

    t1, t2 = kernel_time_ms(p1), kernel_time_ms(p2)
    b1 = eng.N * (12 + 4 * nbar + 128)
    b2 = eng.N * (12 + 4 * nbar + 128 + 12)
    name, t, b = ("mlp_force", t2, b2) if t2 >= t1 else ("mlp_energy", t1, b1)
    ach = b / (t * 1e-3) / 1e9
    flops = eng.N * (2 * (32 * 64 + 64 * 64 + 64) + 2 * (64 * 64 + 64 * 32) + 64)
    return {"bound": "hbm", "kernel": name, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic("mlp_w", name),
            "traffic_ratio": traffic_ratio("mlp_w", name),
            "bytes_per_launch": int(b), "avg_launch_ms": round(t, 5), "nbar_full": round(nbar, 2),
            "other_pass_ms": round(t1 if t == t2 else t2, 5),
            "energy_pass_dense_tflops": round(flops / (t1 * 1e-3) / 1e12, 2)}


def bench_mlp(args):
    """Learned MLP potential vs plain EAM: frames from EAM Langevin runs of
    2000-atom W cells at 300-3000 K, force-matching training on the GPU, then
    a 128,000-atom W cell (40^3 bcc) at 600 K NVT on the MLP with EAM as the
    judge every 50 steps, against the same run on plain EAM."""
    from simulon_amd.mlp import FrameSet, JudgedMLP, MLPForceField, MLPNet, Trainer, \
        generate_frames
    tmp = tempfile.mkdtemp()
    setfl = lat.write_w_synthetic_setfl(os.path.join(tmp, "W_synthetic.eam.alloy"))
    eam = EAMForceField(setfl)
    t0 = time.perf_counter()
    pos, boxes, en, frc, Ts = generate_frames(eam, nframes=args.mlp_frames, ncell=10,
                                              temps=(300.0, 3000.0), nrep=20, equil=0,
                                              stride=20, seed=0)
    t_gen = time.perf_counter() - t0
    # every tenth sample time of every replica is held out
    held = [k for k in range(len(pos)) if (k // 20) % 10 == 9]
    train = [k for k in range(len(pos)) if (k // 20) % 10 != 9]

    def sel(ks):
        return ([pos[k] for k in ks], [boxes[k] for k in ks], [en[k] for k in ks],
                [frc[k] for k in ks])

    batches = [FrameSet(*sel(train[c:c + 100]), lat.W_MASS, 5.5)
               for c in range(0, len(train), 100)]
    test = FrameSet(*sel(held), lat.W_MASS, 5.5)
    log(f"frames: {len(pos)} x {len(pos[0])} atoms in {t_gen:.1f} s; {len(batches)} batches")
    net = MLPNet("cuda", seed=0)
    tr = Trainer(net, wE=0.05, wF=1.0)
    tr.fit_scaling(batches)
    net.init_weights(0, energy_offset=float(np.mean([e / len(p) for e, p in zip(en, pos)])))
    rmse0, _ = tr.force_rmse(test)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = tr.fit(batches, epochs=args.mlp_epochs, lr=3e-3, lr_end=1e-4)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    rmse1, ref = tr.force_rmse(test)
    log(f"training: {args.mlp_epochs} epochs ({args.mlp_epochs * len(batches)} Adam steps) in "
        f"{t_train:.2f} s; held-out force RMSE {rmse0:.4f} -> {rmse1:.4f} eV/A (RMS {ref:.4f})")
    del batches, test
    # 128k-atom inference, plain EAM vs MLP judged by EAM
    p, H = lat.w_bcc(40, seed=0, amp=0.02)
    spec = ReplicaSpec(positions=p, box=H, dt=0.001, ensemble="NVT", temperature=600.0,
                       gamma=10.0)
    e_eam = ReplicaEngine([spec], [lat.W_MASS], cutoff=5.5, skin=1.0, seed=7, velocity_seed=11)
    e_eam.force_field = eam
    e_eam.setup()
    e_eam.run(args.warmup, graph=True, virial=False)
    t_e = timed(lambda k: e_eam.run(k, graph=True, virial=False, check=False), args.steps)
    e_eam.check()
    rate_eam = e_eam.N * args.steps / t_e
    e_mlp = ReplicaEngine([spec], [lat.W_MASS], cutoff=5.5, skin=1.0, seed=7, velocity_seed=11)
    judge = JudgedMLP(e_mlp, MLPForceField(net), eam, every=50, threshold=args.mlp_threshold)
    e_mlp.setup()
    judge.run(50)
    t_m = timed(lambda k: judge.run(k), args.steps)
    rate_mlp = e_mlp.N * args.steps / t_m
    t_pure = None
    if judge.on_mlp:   # the MLP steps alone (graph replays, no judge)
        t_pure = timed(lambda k: e_mlp.run(k, graph=True, virial=False, check=False), args.steps)
    roof = mlp_kernel_roofline(e_mlp) if judge.on_mlp else None
    rm = [h[1] for h in judge.history]
    rr = [h[2] for h in judge.history]
    log(f"EAM {t_e * 1e3 / args.steps:.3f} ms/step; MLP+judge {t_m * 1e3 / args.steps:.3f} ms/step; "
        f"judge RMSE {rm}")
    return {"metric": "atom-steps/s (mlp_w, learned MLP judged by EAM every 50 steps)",
            "value": round(float(rate_mlp), 1), "unit": "atom-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(t_m * 1e3 / args.steps, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "bf16-mfma/f32", "data": "synthetic",
            "config": {"workload": "mlp_w", "atoms": e_mlp.N, "temperature_K": 600.0,
                       "judge_every": 50, "judge_threshold_eV_A": args.mlp_threshold,
                       "train_frames": len(train), "held_out_frames": len(held),
                       "frame_atoms": len(pos[0]), "train_temperatures_K": [300, 3000]},
            "plain_eam_atom_steps_per_s": round(float(rate_eam), 1),
            "mlp_only_atom_steps_per_s": (round(float(e_mlp.N * args.steps / t_pure), 1)
                                          if t_pure else None),
            "mlp_vs_eam": round(float(rate_mlp / rate_eam), 3),
            "ns_per_day": round(args.steps * 0.001e-3 * 86400 / t_m, 3),
            "on_mlp_at_end": judge.on_mlp, "fallback_step": judge.fallback_step,
            "judge_force_rmse_eV_A": round(float(np.mean(rm)), 5) if rm else None,
            "judge_force_rms_eam_eV_A": round(float(np.mean(rr)), 5) if rr else None,
            "held_out_force_rmse_eV_A": [round(rmse0, 5), round(rmse1, 5)],
            "held_out_force_rms_eV_A": round(ref, 5),
            "train_seconds": round(t_train, 3), "frame_gen_seconds": round(t_gen, 2),
            "train_loss_first_last": [hist[0], hist[-1]],
            "roofline": roof}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the workload's configured run length)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="w_eam_replicas")
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--skin", type=float, default=1.0,
                    help="Verlet skin of the replica workload, Angstrom")
    ap.add_argument("--plain-steps", type=int, default=0)
    ap.add_argument("--fcc-cells", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-stage", action="store_true",
                    help="EAM passes read neighbour records from the list, not staged in LDS")
    ap.add_argument("--family", default=None, choices=["cell", "rows", "global"],
                    help="pair-kernel family of the batched engine (default: measured choice)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first copy a device buffer of this many MiB five times (a kernel "
                         "of known traffic, to scale hardware byte counters)")
    ap.add_argument("--state-out", default=None,
                    help="w_eam_replicas: write positions and velocities after the timed steps")
    ap.add_argument("--state-in", default=None,
                    help="w_eam_replicas: start from a --state-out file (profiling the end state)")
    ap.add_argument("--mlp-frames", type=int, default=2000)
    ap.add_argument("--mlp-epochs", type=int, default=200)
    ap.add_argument("--mlp-threshold", type=float, default=0.1)
    args = ap.parse_args()
    if args.steps is None:
        args.steps = DEFAULT_STEPS[args.workload] if args.workload in DEFAULT_STEPS else 200
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (replicas and "
              "domains would shard across GPUs; that part is design only)", file=sys.stderr)
        raise SystemExit(2)
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    if args.copy_probe_mb > 0:
        n = args.copy_probe_mb << 18                    # float32 elements
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 5 copies of {args.copy_probe_mb} MiB (read {args.copy_probe_mb} MiB + "
            f"write {args.copy_probe_mb} MiB each)")
        del src, dst
    if args.workload == "w_eam_replicas":
        out = bench_replicas(args)
    elif args.workload == "mlp_w":
        out = bench_mlp(args)
    else:
        out = bench_single(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
