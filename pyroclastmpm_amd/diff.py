"""Differentiable explicit MPM (USL) on the GPU.

The forward runs pmpm_step (USL) and keeps a checkpoint of the particle state
every `checkpoint_every` steps; the reverse pass recomputes each segment from
its checkpoint and runs pmpm_adjoint_step (hand-written adjoint kernels of
P2G, the grid update, G2P and the return maps, active sets frozen from the
forward) from the last step back. Each forward step writes its output in the
cell order of its input, and the adjoint of step t comes back in the order of
its input, which is the output order of step t-1: no permutation is needed
between steps.

`simulate(model, theta, steps)` is a torch.autograd.Function over theta =
(friction angle [deg], cohesion [Pa], Young's modulus [Pa]) for a
Drucker-Prager material (dp_rheo.cpp:45-70 with M = 6 sin(phi) / (3 - sin(phi)),
Pt = c cot(phi), dilatancy 0) returning the final particle positions in the
particles' original order.
"""
import ctypes as C
import math

import numpy as np
import torch

from . import capi

_F32 = ["x", "v", "stress", "F", "L", "mass", "volume", "volume0", "eps_e", "stress_ref"]


class ParticleSet:
    """Particle state in SoA device tensors, laid out as pmpm_particles."""

    def __init__(self, n, dim, device="cuda"):
        self.n, self.dim = n, dim
        f = lambda r: torch.zeros((r, n), dtype=torch.float32, device=device)
        self.t = {"x": f(dim), "v": f(dim), "stress": f(6), "F": f(dim * dim), "L": f(dim * dim),
                  "mass": f(1), "volume": f(1), "volume0": f(1), "eps_e": f(6), "stress_ref": f(6),
                  "color": torch.zeros(n, dtype=torch.uint8, device=device),
                  "flags": torch.full((n,), capi.PF_ACTIVE, dtype=torch.uint8, device=device),
                  "id": torch.arange(n, dtype=torch.int32, device=device)}

    def view(self):
        t = self.t
        return capi.Particles(self.n, *[t[k].data_ptr() for k in ["x", "v", "stress", "F", "L", "mass", "volume",
                                                                   "volume0", "color", "flags", "id", "eps_e",
                                                                   "stress_ref"]])

    def clone(self):
        o = ParticleSet.__new__(ParticleSet)
        o.n, o.dim = self.n, self.dim
        o.t = {k: v.clone() for k, v in self.t.items()}
        return o


class StepBuffers:
    """The partition (pmpm_spatial) and node arrays (pmpm_nodes) one step
    uses."""

    def __init__(self, n, nn, dim, device="cuda"):
        i32 = lambda m: torch.zeros(m, dtype=torch.int32, device=device)
        self.sp = {"hash": i32(n), "hash_sorted": i32(n), "sorted": i32(n), "cell_start": i32(nn),
                   "cell_end": i32(nn)}
        self.spatial = capi.Spatial(*[self.sp[k].data_ptr() for k in ["hash", "hash_sorted", "sorted",
                                                                      "cell_start", "cell_end"]], None)
        f = lambda m: torch.zeros(m, dtype=torch.float32, device=device)
        self.nd = {"mass": f(nn), "mom": f(dim * nn), "mom_nt": f(dim * nn),
                   "vel": f(capi.node_vel_stride(dim) * nn)}
        # the grid pass's clamp record, kept with the step's nodes for the
        # adjoint of that step
        self.nd["clamp"] = torch.zeros(nn, dtype=torch.uint8, device=device)
        self.nodes = capi.Nodes(*[self.nd[k].data_ptr() for k in ["mass", "mom", "mom_nt", "vel", "clamp"]])


class AdjointSet:
    """Adjoints of the particle state (x, v, stress, F, eps_e), as pmpm_adjoint."""

    def __init__(self, n, dim, device="cuda"):
        self.n, self.dim = n, dim
        f = lambda r: torch.zeros((r, n), dtype=torch.float32, device=device)
        self.t = {"x": f(dim), "v": f(dim), "stress": f(6), "F": f(dim * dim), "eps_e": f(6)}

    def view(self):
        return capi.Adjoint(*[self.t[k].data_ptr() for k in ["x", "v", "stress", "F", "eps_e"]])


def dp_material(phi_deg, cohesion, E, pois, psi_deg=0.0):
    """p[0..4] of a Drucker-Prager material: K, G, M, Pt, eta_bar."""
    phi = math.radians(phi_deg)
    sp = math.sin(phi)
    K = E / (3.0 * (1.0 - 2.0 * pois))
    G = E / (2.0 * (1.0 + pois))
    M = 6.0 * sp / (3.0 - sp)
    Pt = cohesion * math.cos(phi) / sp if sp > 0 else 0.0
    spsi = math.sin(math.radians(psi_deg))
    etab = 6.0 * spsi / (math.sqrt(3.0) * (3.0 + spsi))  # dp_rheo.cpp:65-66
    return K, G, M, Pt, etab


def dp_material_jacobian(phi_deg, cohesion, E, pois):
    """d(K, G, M, Pt) / d(phi [deg], c, E), 4 x 3."""
    phi = math.radians(phi_deg)
    sp, cp = math.sin(phi), math.cos(phi)
    d = math.pi / 180.0
    J = np.zeros((4, 3))
    J[0, 2] = 1.0 / (3.0 * (1.0 - 2.0 * pois))
    J[1, 2] = 1.0 / (2.0 * (1.0 + pois))
    J[2, 0] = 18.0 * cp / (3.0 - sp) ** 2 * d
    J[3, 0] = -cohesion / sp ** 2 * d
    J[3, 1] = cp / sp
    return J


class DiffMPM:
    """A USL simulation with one Drucker-Prager material whose parameters are
    differentiated: gravity and slip walls (NodeDomain mode 1) on every face,
    particles from `positions` at rest."""

    def __init__(self, dim, origin, end, cell_size, positions, ppc, *, density=1500.0, pois=0.3, dt=1e-4,
                 alpha=0.99, shape_function="cubic", gravity=None, checkpoint_every=20, device="cuda",
                 bc_order=("gravity", "walls")):
        self.dim, self.dt, self.alpha, self.pois = dim, dt, alpha, pois
        if sorted(bc_order) != ["gravity", "walls"]:
            raise ValueError("bc_order must order 'gravity' and 'walls'")
        self.bc_order = tuple(bc_order)  # the node conditions' list order (usl.cpp:96-103)
        self.sf = capi.SF[shape_function]
        self.K = int(checkpoint_every)
        self.grid = capi.make_grid(origin, end, cell_size, dim)
        nn = self.grid.num_cells_total
        pos = np.asarray(positions, dtype=np.float32).reshape(-1, dim)
        n = self.n = pos.shape[0]
        p0 = ParticleSet(n, dim, device)
        p0.t["x"].copy_(torch.from_numpy(np.ascontiguousarray(pos.T)))
        for a in range(dim):
            p0.t["F"][a * dim + a] = 1.0
        # particles.cpp:355-358 / 372-393
        h = np.float32(cell_size)
        vol = np.float32(np.float32(float(h) ** dim) / np.float32(ppc))
        p0.t["volume"].fill_(float(vol))
        p0.t["volume0"].fill_(float(vol))
        p0.t["mass"].fill_(float(np.float32(density) * vol))
        self.initial = p0
        self.gravity = list(gravity) if gravity is not None else [0.0] * (dim - 1) + [-9.81]
        self.device = device
        self.work = StepBuffers(n, nn, dim, device)
        self.sp, self.spatial, self.nd, self.nodes = self.work.sp, self.work.spatial, self.work.nd, self.work.nodes
        self.ctx = capi.Context(n, nn)
        self._segment = None  # per-step states and buffers of a recomputed segment, made on first use

    # -------------------------------------------------------------- params
    def params(self, theta):
        phi, c, E = (float(v) for v in theta)
        K, G, M, Pt, etab = dp_material(phi, c, E, self.pois)
        prm = capi.Params()
        prm.dim, prm.shape_fn, prm.dt, prm.alpha = self.dim, self.sf, self.dt, self.alpha
        prm.small_mass_cutoff, prm.node_types_built, prm.wrap_quirk = 1e-9, int(self.sf == capi.SF["cubic"]), 1
        prm.num_materials = 1
        m = prm.materials[0]
        m.kind, m.flags = capi.MAT_DRUCKER_PRAGER, capi.MATF_UPDATE_HISTORY | capi.MATF_APEX_GUARD
        for i, v in enumerate([K, G, M, Pt, etab]):
            m.p[i] = v
        prm.num_bcs = 2
        g, w = prm.bcs[self.bc_order.index("gravity")], prm.bcs[self.bc_order.index("walls")]
        g.kind = capi.BC_GRAVITY
        for a in range(self.dim):
            g.f[a] = self.gravity[a]
        w.kind = capi.BC_NODE_DOMAIN
        for a in range(self.dim):
            w.i[a] = 1
            w.i[3 + a] = 1
        return prm

    # -------------------------------------------------------------- stepping
    def _stream(self):
        return torch.cuda.current_stream().cuda_stream

    def step(self, prm, a, b, work=None):
        w = work or self.work
        capi.check(capi.lib().pmpm_step(self.ctx.ptr, C.byref(prm), C.byref(self.grid), 0, C.byref(a.view()),
                                        C.byref(b.view()), C.byref(w.spatial), C.byref(w.nodes),
                                        self._stream()), "pmpm_step")

    def adjoint_step(self, prm, state, out_bar, in_bar, pbar, work=None, recompute=True):
        """Adjoint of the step from `state`. With recompute=False, `work` must
        hold the partition and nodes that step(prm, state, ., work) left."""
        w = work or self.work
        capi.check(capi.lib().pmpm_adjoint_step(self.ctx.ptr, C.byref(prm), C.byref(self.grid),
                                                C.byref(state.view()), C.byref(out_bar.view()),
                                                C.byref(in_bar.view()), pbar.data_ptr(), int(bool(recompute)),
                                                C.byref(w.spatial), C.byref(w.nodes), self._stream()),
                   "pmpm_adjoint_step")

    def forward(self, theta, steps):
        """Runs `steps` USL steps from the initial state; returns the final
        state and the checkpoints {step: state at the start of that step}."""
        prm = self.params(theta)
        a = self.initial.clone()
        b = ParticleSet(self.n, self.dim, self.device)
        ckpts = {0: a.clone()}
        for t in range(steps):
            self.step(prm, a, b)
            a, b = b, a
            if (t + 1) % self.K == 0 and t + 1 < steps:
                ckpts[t + 1] = a.clone()
        return a, ckpts

    def positions(self, state):
        """Positions [n, dim] in the particles' original order."""
        x = torch.empty((self.n, self.dim), dtype=torch.float32, device=self.device)
        x[state.t["id"].long()] = state.t["x"].t()
        return x

    def _segment_buffers(self):
        """K - 1 particle sets and K partition/node sets, reused by every
        segment: step k of a segment reads state k (state 0 is the checkpoint
        itself, which a USL step only reads) and keeps its partition and nodes
        in work k for the adjoint of the same step."""
        if self._segment is None:
            nn = self.grid.num_cells_total
            states = [ParticleSet(self.n, self.dim, self.device) for _ in range(self.K - 1)]
            work = [self.work] + [StepBuffers(self.n, nn, self.dim, self.device) for _ in range(self.K - 1)]
            self._segment = (states, work)
        return self._segment

    def backward(self, theta, steps, ckpts, final, gx):
        """Adjoint of the final positions (original order) -> d/d(K, G, M, Pt)."""
        prm = self.params(theta)
        # every adjoint step writes all of in_bar: two sets, used in turn
        gbar, ibar = AdjointSet(self.n, self.dim, self.device), AdjointSet(self.n, self.dim, self.device)
        gbar.t["x"].copy_(gx[final.t["id"].long()].t())
        pbar = torch.zeros(4, dtype=torch.float32, device=self.device)
        pool, work = self._segment_buffers()
        starts = sorted(ckpts.keys())
        for s0 in reversed(starts):
            s1 = min(s0 + self.K, steps)
            states = [ckpts[s0]] + pool[:s1 - s0 - 1]
            for k in range(s1 - s0 - 1):
                self.step(prm, states[k], states[k + 1], work[k])
            # the segment's last step was not run forward here: its adjoint
            # recomputes the partition, P2G and grid update itself
            for k in reversed(range(s1 - s0)):
                self.adjoint_step(prm, states[k], gbar, ibar, pbar, work[k], recompute=k == s1 - s0 - 1)
                gbar, ibar = ibar, gbar
        return pbar.double().cpu().numpy()

    def grad_theta(self, theta, pbar):
        J = dp_material_jacobian(float(theta[0]), float(theta[1]), float(theta[2]), self.pois)
        return pbar @ J


class _Simulate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, model, steps):
        th = theta.detach().double().cpu().numpy()
        final, ckpts = model.forward(th, steps)
        ctx.model, ctx.steps, ctx.ckpts, ctx.final, ctx.th = model, steps, ckpts, final, th
        return model.positions(final)

    @staticmethod
    def backward(ctx, gx):
        m = ctx.model
        pbar = m.backward(ctx.th, ctx.steps, ctx.ckpts, ctx.final, gx.contiguous().float())
        g = torch.from_numpy(m.grad_theta(ctx.th, pbar)).to(dtype=torch.float64)
        return g, None, None


def simulate(model, theta, steps):
    """Final positions [n, dim] (original particle order) after `steps` USL
    steps with Drucker-Prager parameters theta = (phi [deg], c [Pa], E [Pa]);
    differentiable with respect to theta."""
    return _Simulate.apply(theta, model, steps)
