"""An ensemble of differentiated scenes: the forward and reverse passes of
many independent columns advanced together on one GPU.

Every scene is a `diff.DiffMPM` (USL, one Drucker-Prager material whose
parameters are differentiated, a checkpoint every K steps). The forward steps
of all scenes go through pmpm_batch_step: the scenes are spread over a few
streams forked from one origin stream, and the multi-scene step can be
captured into a HIP graph per buffer parity and replayed. The reverse pass
walks the segments backwards. It recomputes each segment of every scene with
pmpm_batch_step, then runs the adjoint steps of all scenes through
pmpm_batch_adjoint_step, so the small scenes run side by side there too.
Each scene's gradient is the one its own `diff.simulate` gives.

The reference has no counterpart: it runs one forward-only scene per process
(global_settings.cpp).
"""
import ctypes as C

import numpy as np
import torch

from . import capi
from .batch import column_geometry, sweep_parameters
from .diff import AdjointSet, DiffMPM, ParticleSet


class Ensemble:
    """Scenes (DiffMPM models sharing one checkpoint interval) differentiated
    together through one pmpm_batch with `lanes` streams."""

    def __init__(self, models, lanes=4):
        models = list(models)
        if not models:
            raise ValueError("an ensemble needs at least one scene")
        intervals = {m.K for m in models}
        if len(intervals) != 1:
            raise ValueError("every scene of an ensemble must use the same checkpoint interval")
        self.models = models
        self.K = intervals.pop()
        self.ptr = capi.vp()
        capi.check(capi.lib().pmpm_batch_create(C.byref(self.ptr), int(lanes)), "pmpm_batch_create")
        # origin stream of every batched call (a graph cannot be captured on
        # the NULL stream); it follows and rejoins the caller's stream
        self.stream = torch.cuda.Stream()
        self.A = [m.initial.clone() for m in models]
        self.B = [ParticleSet(m.n, m.dim, m.device) for m in models]

    @property
    def particles(self):
        return sum(m.n for m in self.models)

    # ------------------------------------------------------------ plumbing
    def _run(self, fn):
        """Runs fn() with the ensemble's stream current, ordered after the
        caller's pending work; the caller's stream waits for it after."""
        caller = torch.cuda.current_stream()
        self.stream.wait_stream(caller)
        with torch.cuda.stream(self.stream):
            out = fn()
        caller.wait_stream(self.stream)
        return out

    @staticmethod
    def _scene(m, prm, a, b, work):
        s = capi.Scene()
        s.ctx, s.prm, s.grid = m.ctx.ptr, prm, m.grid
        s.set[0], s.set[1] = a.view(), b.view()
        s.part, s.nodes, s.musl = work.spatial, work.nodes, 0
        return s

    def _forward_scenes(self, prms):
        return (capi.Scene * len(self.models))(*[self._scene(m, p, a, b, m.work)
                                                 for m, p, a, b in zip(self.models, prms, self.A, self.B)])

    def _batch_step(self, arr, cur):
        capi.check(capi.lib().pmpm_batch_step(self.ptr, arr, len(self.models), cur, self.stream.cuda_stream),
                   "pmpm_batch_step")

    def _reset(self):
        for m, a in zip(self.models, self.A):
            for k, v in m.initial.t.items():
                a.t[k].copy_(v)

    # ------------------------------------------------------------ tuning
    def autotune(self, thetas, steps=2):
        """The library's autotuner over the forward step from the initial
        states (the lane count and the per-scene weights it keeps are used by
        both passes); returns its setting."""
        prms = [m.params(th) for m, th in zip(self.models, thetas)]
        t = capi.BatchTuning()

        def go():
            self._reset()
            arr = self._forward_scenes(prms)
            self._batch_step(arr, 0)  # sizes every workspace
            self._batch_step(arr, 1)
            capi.check(capi.lib().pmpm_batch_autotune(self.ptr, arr, len(self.models), 0, int(steps),
                                                      self.stream.cuda_stream, C.byref(t)),
                       "pmpm_batch_autotune")

        self._run(go)
        return {"lanes": t.lanes, "use_graph": t.use_graph, "step_ms": t.step_ms}

    # ------------------------------------------------------------ passes
    def forward(self, thetas, steps, graph=False):
        """`steps` USL steps of every scene from its initial state, theta
        (phi [deg], c [Pa], E [Pa]) per scene. With graph=True, steps after
        the first replay the multi-scene step captured per buffer parity.
        Returns the final states and, per scene, the checkpoints {step: state
        at the start of that step}."""
        prms = [m.params(th) for m, th in zip(self.models, thetas)]
        L, count = capi.lib(), len(self.models)

        def go():
            self._reset()
            arr = self._forward_scenes(prms)
            ckpts = [{0: a.clone()} for a in self.A]
            cur = 0
            for t in range(steps):
                if graph and t > 0:
                    capi.check(L.pmpm_batch_replay(self.ptr, cur, self.stream.cuda_stream), "pmpm_batch_replay")
                else:
                    self._batch_step(arr, cur)
                if graph and t == 0 and steps > 1:
                    # the first step sized every workspace: capture both parities
                    for c in (0, 1):
                        capi.check(L.pmpm_batch_capture(self.ptr, arr, count, c, self.stream.cuda_stream),
                                   "pmpm_batch_capture")
                cur ^= 1
                if (t + 1) % self.K == 0 and t + 1 < steps:
                    for i, s in enumerate(self.A if cur == 0 else self.B):
                        ckpts[i][t + 1] = s.clone()
            finals = [s.clone() for s in (self.A if cur == 0 else self.B)]
            return finals, ckpts

        return self._run(go)

    def backward(self, thetas, steps, ckpts, finals, gxs):
        """Adjoint of the final positions (original particle order; None for
        a scene whose positions do not enter the loss) -> per scene
        d/d(K, G, M, Pt), shape (scenes, 4)."""
        prms = [m.params(th) for m, th in zip(self.models, thetas)]
        L, count = capi.lib(), len(self.models)

        def go():
            gbar = [AdjointSet(m.n, m.dim, m.device) for m in self.models]
            ibar = [AdjointSet(m.n, m.dim, m.device) for m in self.models]
            for g, f, gx in zip(gbar, finals, gxs):
                if gx is not None:
                    g.t["x"].copy_(gx[f.t["id"].long()].t())
            pbar = torch.zeros((count, 4), dtype=torch.float32, device=self.models[0].device)
            pools = [m._segment_buffers() for m in self.models]
            for s0 in reversed(sorted(ckpts[0].keys())):
                seg = min(s0 + self.K, steps) - s0
                states = [[ck[s0]] + pool[0][:seg - 1] for ck, pool in zip(ckpts, pools)]
                for k in range(seg - 1):
                    arr = (capi.Scene * count)(*[self._scene(m, p, st[k], st[k + 1], pool[1][k])
                                                 for m, p, st, pool in zip(self.models, prms, states, pools)])
                    self._batch_step(arr, 0)
                # the segment's last step was not run forward here: its
                # adjoint recomputes the partition, P2G and grid update
                for k in reversed(range(seg)):
                    adj = (capi.AdjointScene * count)()
                    for i, (m, p, st, pool) in enumerate(zip(self.models, prms, states, pools)):
                        a = adj[i]
                        a.ctx, a.prm, a.grid = m.ctx.ptr, p, m.grid
                        a.state, a.out_bar, a.in_bar = st[k].view(), gbar[i].view(), ibar[i].view()
                        a.param_bar = pbar[i].data_ptr()
                        a.recompute = int(k == seg - 1)
                        a.part, a.nodes = pool[1][k].spatial, pool[1][k].nodes
                    capi.check(L.pmpm_batch_adjoint_step(self.ptr, adj, count, self.stream.cuda_stream),
                               "pmpm_batch_adjoint_step")
                    gbar, ibar = ibar, gbar
            return pbar.double().cpu().numpy()

        return self._run(go)

    def grad_theta(self, thetas, pbar):
        return np.stack([m.grad_theta(th, pb) for m, th, pb in zip(self.models, thetas, pbar)])

    def __del__(self):
        try:
            if self.ptr:
                capi.lib().pmpm_batch_destroy(self.ptr)
                self.ptr = capi.vp()
        except Exception:
            pass


class _EnsembleSimulate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, thetas, ens, steps, graph):
        th = thetas.detach().double().cpu().numpy()
        finals, ckpts = ens.forward(th, steps, graph)
        ctx.ens, ctx.steps, ctx.ckpts, ctx.finals, ctx.th = ens, steps, ckpts, finals, th
        out = ens._run(lambda: [m.positions(f) for m, f in zip(ens.models, finals)])
        for x in out:  # allocated on the ensemble's stream, read on the caller's
            x.record_stream(torch.cuda.current_stream())
        return tuple(out)

    @staticmethod
    def backward(ctx, *gxs):
        ens = ctx.ens
        gxs = [None if g is None else g.contiguous().float() for g in gxs]
        pbar = ens.backward(ctx.th, ctx.steps, ctx.ckpts, ctx.finals, gxs)
        g = torch.from_numpy(ens.grad_theta(ctx.th, pbar)).to(dtype=torch.float64)
        return g, None, None, None


def simulate_ensemble(ens, thetas, steps, graph=False):
    """Final positions of every scene (a tuple of [n_i, 3] tensors in each
    scene's original particle order) after `steps` USL steps with
    Drucker-Prager parameters thetas[i] = (phi [deg], c [Pa], E [Pa]);
    differentiable with respect to thetas (scenes x 3)."""
    return _EnsembleSimulate.apply(thetas, ens, steps, graph)


def column_ensemble(count=64, seed=1, n_min=16000, n_max=1000000, per_axis=2, dt=1e-5, checkpoint_every=20,
                    device="cuda"):
    """The multi-scene column sweep as differentiated scenes: the sweep's
    geometry (particle counts, cell sizes) with DiffMPM's USL step, cubic
    B-splines, gravity and slip walls. Returns the models and their thetas
    (phi, 0, E) as a (count, 3) float64 tensor."""
    models, thetas = [], []
    for i, (n, h, phi, E) in enumerate(sweep_parameters(count, seed, n_min, n_max)):
        pos, end = column_geometry(n, h, per_axis, 0.1, seed * 1000 + i)
        models.append(DiffMPM(3, [0.0, 0.0, 0.0], end, h, pos, per_axis ** 3, dt=dt,
                              checkpoint_every=checkpoint_every, device=device))
        thetas.append([phi, 0.0, E])
    return models, torch.tensor(thetas, dtype=torch.float64)
