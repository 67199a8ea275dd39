"""Learned radial-descriptor MLP potential: network state in device memory,
the force-field plug-in, force-matching training on the GPU and the EAM
judge with fallback.

Model (the library's own; the reference's learned potential is an external
model behind the return contract of
machine_learning_potentials/machine_learning_force.py:85-142, which the
backbone wrapper keeps):
    g_ik = sum_j exp(-(r_ij - mu_k)^2 / (2 w^2)) fc(r_ij),  32 centres 1.5..5.5 A,
           w = 0.25 A, fc = (cos(pi r / rc) + 1) / 2, rc = 5.5 A
    E_i  = w3 . silu(W2 silu(W1 (g_i - mean) / std + b1) + b2) + b3
W1 [64x32] and W2 [64x64] are bfloat16 (MFMA operands, float32
accumulation); biases and w3 are float32. Every evaluation, gradient and
optimizer step runs in the HIP kernels of libsimulon_hip.so (sim_mlp_*).
"""
import time

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_ptr
from .engine import ReplicaEngine, ReplicaSpec

NDESC, NH = _lib.MLP_NDESC, _lib.MLP_NH
NPARAM, NGRAD = _lib.MLP_NPARAM, _lib.MLP_NGRAD
# flat parameter layout (sim_mlp_grad / sim_mlp_adam)
SLICES = {"W1": (0, 2048, (NH, NDESC)), "b1": (2048, 2112, (NH,)), "W2": (2112, 6208, (NH, NH)),
          "b2": (6208, 6272, (NH,)), "w3": (6272, 6336, (NH,)), "b3": (6336, 6337, (1,))}


class MLPNet:
    """Network parameters on the device: float32 master copy (flat), the
    bfloat16 weight matrices the kernels read, Adam moments, and the fixed
    descriptor scaling."""

    def __init__(self, device="cuda", seed=0, rc=5.5, mu0=1.5, mu1=5.5, width=0.25):
        self.device = torch.device(device)
        self.rc, self.mu0, self.width = float(rc), float(mu0), float(width)
        self.dmu = (float(mu1) - float(mu0)) / (NDESC - 1)
        f32 = dict(dtype=torch.float32, device=self.device)
        self.master = torch.zeros(NPARAM, **f32)
        self.adam_m = torch.zeros(NPARAM, **f32)
        self.adam_v = torch.zeros(NPARAM, **f32)
        self.grad = torch.zeros(NGRAD, **f32)
        self.w1 = torch.zeros((NH, NDESC), dtype=torch.bfloat16, device=self.device)
        self.w2 = torch.zeros((NH, NH), dtype=torch.bfloat16, device=self.device)
        self.b1 = torch.zeros(NH, **f32)
        self.b2 = torch.zeros(NH, **f32)
        self.w3 = torch.zeros(NH, **f32)
        self.b3 = torch.zeros(1, **f32)
        self.mean = torch.zeros(NDESC, **f32)
        self.inv_std = torch.ones(NDESC, **f32)
        self.adam_t = 0
        self.init_weights(seed)

    # ---- parameters ---------------------------------------------------------------
    def init_weights(self, seed=0, energy_offset=0.0):
        """Glorot-uniform weights, zero biases, b3 = energy_offset (eV/atom)."""
        rng = np.random.default_rng(seed)
        flat = np.zeros(NPARAM, np.float32)

        def glorot(shape):
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            return rng.uniform(-lim, lim, shape).astype(np.float32)

        flat[0:2048] = glorot((NH, NDESC)).ravel()
        flat[2112:6208] = glorot((NH, NH)).ravel()
        flat[6272:6336] = glorot((1, NH)).ravel()
        flat[6336] = energy_offset
        self.load_flat(flat)

    def load_flat(self, flat):
        """Set the master parameters and rewrite the device copies."""
        self.master.copy_(torch.as_tensor(np.asarray(flat, np.float32)))
        self._sync_copies()
        self.adam_m.zero_()
        self.adam_v.zero_()
        self.adam_t = 0

    def _sync_copies(self):
        m = self.master
        self.w1.copy_(m[0:2048].view(NH, NDESC).to(torch.bfloat16))
        self.b1.copy_(m[2048:2112])
        self.w2.copy_(m[2112:6208].view(NH, NH).to(torch.bfloat16))
        self.b2.copy_(m[6208:6272])
        self.w3.copy_(m[6272:6336])
        self.b3.copy_(m[6336:6337])

    def set_scaling(self, mean, std):
        self.mean.copy_(torch.as_tensor(np.asarray(mean, np.float32)))
        self.inv_std.copy_(torch.as_tensor(1.0 / np.asarray(std, np.float32)))

    def params_numpy(self):
        """Parameters as the kernels use them (W1, W2 bfloat16-valued), float64:
        (W1, b1, W2, b2, w3, b3)."""
        return (self.w1.float().double().cpu().numpy(), self.b1.double().cpu().numpy(),
                self.w2.float().double().cpu().numpy(), self.b2.double().cpu().numpy(),
                self.w3.double().cpu().numpy(), self.b3.double().cpu().numpy()[0])

    def struct(self):
        s = _lib.MLPNet()
        s.w1, s.w2 = self.w1.data_ptr(), self.w2.data_ptr()
        s.b1, s.b2, s.w3, s.b3 = (self.b1.data_ptr(), self.b2.data_ptr(), self.w3.data_ptr(),
                                  self.b3.data_ptr())
        s.mean, s.inv_std = self.mean.data_ptr(), self.inv_std.data_ptr()
        s.rc, s.mu0, s.dmu = self.rc, self.mu0, self.dmu
        s.inv_w2 = 1.0 / (self.width * self.width)
        return s

    def save(self, path):
        np.savez(path, master=self.master.cpu().numpy(), mean=self.mean.cpu().numpy(),
                 inv_std=self.inv_std.cpu().numpy(),
                 basis=np.array([self.rc, self.mu0, self.dmu, self.width]))

    @classmethod
    def load(cls, path, device="cuda"):
        d = np.load(path)
        rc, mu0, dmu, width = [float(x) for x in d["basis"]]
        net = cls(device, rc=rc, mu0=mu0, mu1=mu0 + dmu * (NDESC - 1), width=width)
        net.load_flat(d["master"])
        net.mean.copy_(torch.as_tensor(d["mean"]))
        net.inv_std.copy_(torch.as_tensor(d["inv_std"]))
        return net

    # ---- optimizer ------------------------------------------------------------------
    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        """One Adam step from self.grad (sim_mlp_adam), rewriting the bf16 copies."""
        self.adam_t += 1
        call("sim_mlp_adam", ptr(self.master), ptr(self.grad), ptr(self.adam_m), ptr(self.adam_v),
             self.adam_t, float(lr), float(beta1), float(beta2), float(eps), self.struct(),
             stream_ptr())


def _rows4(eng):
    t = getattr(eng, "_mlp_rows4", None)
    if t is None:
        t = eng.rep_tiles * _lib.MLP_ROWS_PER_TILE
        eng._mlp_rows4 = t
    return t


class MLPForceField:
    """Engine plug-in: descriptor + network pass, then the analytic force pass.
    Writes eng.force and the energy/virial partials of every replica."""
    name = "mlp"

    def __init__(self, net: MLPNet):
        self.net = net
        self.cutoff = net.rc

    def partials(self, eng):
        part = eng.scratch("mlp_partial", torch.float64,
                           (eng.ntiles * _lib.MLP_ROWS_PER_TILE, _lib.SIM_NPART))
        return part, _rows4(eng)

    def compute(self, eng, virial=True, gh=None, e_atom=None):
        if eng.cutoff < self.net.rc:
            raise ValueError("the engine's cutoff is below the network's descriptor cutoff")
        c = eng.scratch("mlp_c", torch.float32, (eng.N, NDESC))
        part, _ = self.partials(eng)
        net = self.net.struct()
        s = stream_ptr()
        eng.quantize()
        call("sim_mlp_energy", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq),
             ptr(eng.nbr), ptr(eng.nnbr), eng.N, net, ptr(c), ptr(gh), ptr(e_atom), ptr(part), s)
        call("sim_mlp_force", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq),
             ptr(eng.nbr), ptr(eng.nnbr), eng.N, net, ptr(c), ptr(eng.force), ptr(part),
             int(bool(virial)), s)


# ---- training data ------------------------------------------------------------------

class FrameSet:
    """Frames with reference energies and forces, packed as the replicas of one
    engine (no stepping: the list is built once for the fixed positions)."""

    def __init__(self, positions, boxes, energies, forces, mass_amu, cutoff, device="cuda",
                 skin=0.3):
        specs = [ReplicaSpec(positions=np.asarray(p, np.float64), box=np.asarray(h, np.float64),
                             velocities=np.zeros((len(p), 3)))
                 for p, h in zip(positions, boxes)]
        self.eng = ReplicaEngine(specs, [mass_amu], cutoff, skin, device=device, tiled=False)
        self.eng.build_list(force=True)
        self.eng.check()
        f = torch.as_tensor(np.concatenate([np.asarray(x, np.float32) for x in forces]),
                            device=self.eng.device)
        self.force_ref = self.eng.from_caller(f).contiguous()
        self.e_ref = torch.as_tensor(np.asarray(energies, np.float64), device=self.eng.device)
        self.atom_rep = self.eng.atom_replica().to(torch.int32).contiguous()
        self.natoms = np.array(self.eng.natoms)

    @property
    def nframes(self):
        return self.eng.R


class Trainer:
    """Force-matching training of an MLPNet on FrameSets, on the device:
    sim_mlp_energy -> sim_reduce (frame energies) -> sim_mlp_force ->
    sim_mlp_adjoint -> sim_mlp_grad -> sim_mlp_adam, all stream-ordered."""

    def __init__(self, net: MLPNet, wE=0.1, wF=1.0, max_blocks=1024):
        self.net, self.wE, self.wF = net, float(wE), float(wF)
        self.max_blocks = int(max_blocks)
        self.gpart = torch.zeros((self.max_blocks, NGRAD), dtype=torch.float32, device=net.device)
        self.ff = MLPForceField(net)

    def fit_scaling(self, frames):
        """Descriptor mean and std over every atom of the FrameSet(s) `frames`
        (one kernel pass each with identity scaling). A std below 5% of the
        largest is raised to it, so that descriptors that barely vary in the
        data (centres no neighbor reaches) do not amplify rare values."""
        sets = frames if isinstance(frames, (list, tuple)) else [frames]
        net = self.net
        net.mean.zero_()
        net.inv_std.fill_(1.0)
        s1 = torch.zeros(NDESC, dtype=torch.float64, device=net.device)
        s2 = torch.zeros_like(s1)
        n = 0
        for fs in sets:
            eng = fs.eng
            g = eng.scratch("mlp_gh", torch.float32, (eng.N, NDESC))
            self.ff.compute(eng, virial=False, gh=g)
            gd = g.double()
            s1 += gd.sum(0)
            s2 += (gd * gd).sum(0)
            n += eng.N
        m = s1 / n
        sd = torch.sqrt(torch.clamp(s2 / n - m * m, min=0.0))
        sd = torch.maximum(sd, 0.05 * sd.max()).clamp_min(1e-6)
        net.set_scaling(m.cpu().numpy(), sd.cpu().numpy())

    def loss_and_grad(self, frames: FrameSet):
        """Forward, forces, adjoint and the weight gradient into net.grad
        (last entry: the loss). Returns nothing (no host sync)."""
        eng, net = frames.eng, self.net
        s = stream_ptr()
        st = net.struct()
        c = eng.scratch("mlp_c", torch.float32, (eng.N, NDESC))
        gh = eng.scratch("mlp_gh", torch.float32, (eng.N, NDESC))
        v = eng.scratch("mlp_v", torch.float32, (eng.N, NDESC))
        part, rows = self.ff.partials(eng)
        eng.quantize()
        call("sim_mlp_energy", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq),
             ptr(eng.nbr), ptr(eng.nnbr), eng.N, st, ptr(c), ptr(gh), None, ptr(part), s)
        call("sim_reduce", eng.R, ptr(rows), ptr(part), 0, 1, ptr(eng.red_chunks),
             ptr(eng.thermo), s)
        call("sim_mlp_force", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq),
             ptr(eng.nbr), ptr(eng.nnbr), eng.N, st, ptr(c), ptr(eng.force), ptr(part), 0, s)
        call("sim_mlp_adjoint", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq),
             ptr(eng.nbr), ptr(eng.nnbr), eng.N, st, ptr(eng.force), ptr(frames.force_ref),
             self.wF, ptr(v), s)
        call("sim_mlp_grad", eng.N, ptr(frames.atom_rep), ptr(eng.reps), ptr(eng.thermo),
             ptr(frames.e_ref), self.wE, self.wF, st, ptr(gh), ptr(v), ptr(eng.force),
             ptr(frames.force_ref), ptr(self.gpart), self.max_blocks, ptr(net.grad), s)

    def step(self, frames: FrameSet, lr=1e-3):
        self.loss_and_grad(frames)
        self.net.adam_step(lr)

    def force_rmse(self, frames: FrameSet):
        """RMS of F_mlp - F_ref over all atoms and components (eV/A), and the
        RMS of F_ref itself."""
        eng = frames.eng
        self.ff.compute(eng, virial=False)
        d = (eng.force - frames.force_ref).double()
        return (float(torch.sqrt((d * d).mean())),
                float(torch.sqrt((frames.force_ref.double() ** 2).mean())))

    def fit(self, batches, epochs, lr=2e-3, lr_end=2e-4, log=None):
        """Adam over the batches for `epochs` passes with an exponential
        learning-rate decay from lr to lr_end; returns the per-epoch mean loss."""
        history = []
        nsteps = max(1, epochs * len(batches))
        k = 0
        for ep in range(epochs):
            losses = []
            for b in batches:
                rate = lr * (lr_end / lr) ** (k / max(1, nsteps - 1))
                self.step(b, rate)
                losses.append(self.net.grad[NGRAD - 1].clone())
                k += 1
            history.append(float(torch.stack(losses).mean()) / 1.0)
            if log:
                log(f"epoch {ep}: loss {history[-1]:.5g}")
        return history


def generate_frames(eam_ff, nframes=2000, ncell=10, temps=(300.0, 3000.0), nrep=20,
                    equil=0, stride=20, dt=0.001, gamma=10.0, seed=0, device="cuda"):
    """Frames of bcc W cells (ncell^3 cells, 2 atoms each) sampled from
    Langevin runs of the EAM force field at temperatures evenly spaced over
    `temps`, all replicas advanced together: returns (positions, boxes,
    energies, forces, temperatures) lists in the caller's atom order, frame
    k = sample k // nrep of replica k % nrep. Sampling starts after `equil`
    steps; the default 0 keeps the approach to equilibrium from the lattice
    in the data (MD runs start there too)."""
    from .lattice import w_bcc, W_MASS
    per = int(np.ceil(nframes / nrep))
    Ts = np.linspace(temps[0], temps[1], nrep)
    specs = []
    for r, T in enumerate(Ts):
        pos, H = w_bcc(ncell, seed=seed + r, amp=0.02)
        specs.append(ReplicaSpec(positions=pos, box=H, dt=dt, ensemble="NVT", temperature=T,
                                 gamma=gamma))
    eng = ReplicaEngine(specs, [W_MASS], eam_ff.cutoff, 1.0, device=device, seed=seed + 17,
                        velocity_seed=seed + 29)
    eng.force_field = eam_ff
    eng.setup()
    eng.run(equil, graph=True, virial=False)
    out = ([], [], [], [], [])
    boxes = eng.box_rows().cpu().numpy()
    for k in range(per):
        eng.run(stride, graph=True, virial=False)
        pos = eng.to_caller(eng.positions(torch.float64)).cpu().numpy()
        frc = eng.to_caller(eng.force).cpu().numpy()
        ep = eng.thermo[:, _lib.T_EPOT].cpu().numpy()
        for r in range(eng.R):
            sl = eng.replica_slice(r)
            out[0].append(pos[sl].copy())
            out[1].append(boxes[r].copy())
            out[2].append(float(ep[r]))
            out[3].append(frc[sl].copy())
            out[4].append(float(Ts[r]))
    return tuple(list(x[:nframes]) for x in out)


# ---- EAM judge -----------------------------------------------------------------------

class JudgedMLP:
    """Run an engine on the MLP with plain EAM re-evaluated every `every`
    steps as the judge of force error: when the force RMSE (eV/A, over every
    atom and component) exceeds `threshold`, EAM takes over for the rest of
    the run. The steps between judgements run as graph replays."""

    def __init__(self, eng, mlp_ff, eam_ff, every=50, threshold=0.3):
        self.eng, self.mlp, self.eam = eng, mlp_ff, eam_ff
        self.every, self.threshold = int(every), float(threshold)
        self.history = []          # (step, rmse, rms_ref)
        self.fallback_step = None
        eng.force_field = mlp_ff

    @property
    def on_mlp(self):
        return self.eng.force_field is self.mlp

    def judge(self):
        """Force RMSE of the current forces (MLP) against EAM at the current
        positions; EAM is evaluated into work arrays, the state is untouched."""
        eng = self.eng
        f = eng.scratch("judge_force", torch.float32, (eng.N, 3))
        part = eng.scratch("judge_partial", torch.float64, (eng.ntiles, _lib.SIM_NPART))
        self.eam.compute(eng, virial=False, force=f, partial=part)
        d = (eng.force - f).double()
        rmse = float(torch.sqrt((d * d).mean()))
        ref = float(torch.sqrt((f.double() ** 2).mean()))
        self.history.append((eng.steps_done, rmse, ref))
        return rmse

    def run(self, nsteps, graph=True):
        done = 0
        while done < nsteps:
            k = min(self.every, nsteps - done)
            self.eng.run(k, graph=graph, virial=False)
            done += k
            if self.on_mlp and self.eng.steps_done % self.every == 0:
                if self.judge() > self.threshold:
                    self.fallback_step = self.eng.steps_done
                    self.eng.force_field = self.eam
                    self.eng._graph = None
                    # forces of the new field at the current positions
                    self.eng.compute_forces(virial=False)
        return self
