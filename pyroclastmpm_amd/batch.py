"""Many independent scenes stepped together on one GPU (pmpm_batch_*).

The reference runs one scene per process through its process globals
(global_settings.cpp); a parameter sweep or an optimisation over many
columns runs them one after another. Here every scene keeps its own buffers,
grid and context, and one batched step advances all of them: the scenes are
spread over a few HIP streams (longest-processing-time split of their step
times), so the small ones run side by side instead of leaving most of the GPU
idle, and the whole multi-scene step can be captured into a HIP graph and
replayed.

`column_batch` builds the sweep this is measured on: Drucker-Prager column
collapses (MUSL, cubic B-splines, slip walls, gravity) with particle counts
log-uniform in [16k, 1M], cell size uniform in [0.01, 0.02] m, friction angle
uniform in [20, 40] deg and Young's modulus uniform in [5e5, 2e6] Pa.
"""
import ctypes as C
import math

import numpy as np
import torch

from . import capi
from .diff import dp_material

_SET_KEYS = ["x", "v", "stress", "F", "L", "mass", "volume", "volume0", "color", "flags", "id", "eps_e",
             "stress_ref"]


def _lattice(lo, hi, spacing, jitter, rng):
    axes = [np.arange(a + 0.5 * spacing, b - 1e-12, spacing) for a, b in zip(lo, hi)]
    grids = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([g.ravel() for g in grids], axis=1)
    if jitter > 0:
        pts = pts + rng.uniform(-jitter * spacing, jitter * spacing, size=pts.shape)
    return pts.astype(np.float32)


def column_geometry(n_target, h, per_axis=2, jitter=0.1, seed=0):
    """Particle positions (n, 3) and domain end of a column with the 1.0 x 0.5
    x 0.5 m benchmark column's proportions (column 0.2 x 0.5 x 0.4 of the
    domain), scaled to about n_target particles at per_axis^3 per cell of
    size h."""
    spacing = h / per_axis
    s = (n_target * spacing ** 3 / (0.2 * 0.5 * 0.4)) ** (1.0 / 3.0)
    col = (0.2 * s, 0.5 * s, 0.4 * s)
    end = (1.0 * s, 0.5 * s, 0.5 * s)
    return _lattice([0.0, 0.0, 0.0], col, spacing, jitter, np.random.default_rng(seed)), end


def sweep_parameters(count=64, seed=1, n_min=16000, n_max=1000000):
    """(n_target, h, phi, E) of each scene of the sweep, drawn with
    numpy.random.default_rng(seed): particle counts log-uniform in
    [n_min, n_max], h uniform in [0.01, 0.02] m, friction angle uniform in
    [20, 40] deg, Young's modulus uniform in [5e5, 2e6] Pa."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = float(math.exp(rng.uniform(math.log(n_min), math.log(n_max))))
        h = float(rng.uniform(0.01, 0.02))
        phi = float(rng.uniform(20.0, 40.0))
        E = float(rng.uniform(5e5, 2e6))
        out.append((n, h, phi, E))
    return out


class ColumnScene:
    """One 3D Drucker-Prager column collapse (the benchmark column's physics
    at another size and material), with both particle buffer sets, nodes,
    partition and context on the device."""

    def __init__(self, n_target, h, phi, E, cohesion=0.0, per_axis=2, dt=1e-5, jitter=0.1, seed=0,
                 device="cuda"):
        pos, end = column_geometry(n_target, h, per_axis, jitter, seed)
        self.n = n = pos.shape[0]
        self.h, self.phi, self.E = h, phi, E
        self.grid = capi.make_grid([0.0, 0.0, 0.0], end, h, 3)
        nn = self.grid.num_cells_total
        vol = np.float32(np.float32(h) ** 3 / np.float32(per_axis ** 3))
        f = lambda r, fill=0.0: torch.full((r, n), fill, dtype=torch.float32, device=device)

        def make_set(with_x):
            t = {"x": torch.from_numpy(np.ascontiguousarray(pos.T)).to(device) if with_x else f(3), "v": f(3),
                 "stress": f(6), "F": f(9), "mass": f(1, float(np.float32(1500.0) * vol)), "volume": f(1, float(vol)),
                 "volume0": f(1, float(vol)), "eps_e": f(6),
                 "color": torch.zeros(n, dtype=torch.uint8, device=device),
                 "flags": torch.full((n,), capi.PF_ACTIVE, dtype=torch.uint8, device=device),
                 "id": torch.arange(n, dtype=torch.int32, device=device)}
            for a in range(3):
                t["F"][4 * a] = 1.0
            return t

        self.sets = [make_set(True), make_set(False)]
        i32 = lambda m: torch.zeros(m, dtype=torch.int32, device=device)
        self.sp = {"hash": i32(n), "hash_sorted": i32(n), "sorted": i32(n), "cell_start": i32(nn),
                   "cell_end": i32(nn)}
        fz = lambda m: torch.zeros(m, dtype=torch.float32, device=device)
        self.nd = {"mass": fz(nn), "mom": fz(3 * nn), "mom_nt": fz(3 * nn), "vel": fz(capi.node_vel_stride(3) * nn)}
        self.ctx = capi.Context(n, nn)
        K, G, M, Pt, etab = dp_material(phi, cohesion, E, 0.3)
        prm = capi.Params()
        prm.dim, prm.shape_fn, prm.dt, prm.alpha = 3, capi.SF["cubic"], dt, 0.99
        prm.small_mass_cutoff, prm.node_types_built, prm.wrap_quirk = 1e-9, 1, 1
        prm.num_materials = 1
        m = prm.materials[0]
        m.kind, m.flags = capi.MAT_DRUCKER_PRAGER, capi.MATF_UPDATE_HISTORY | capi.MATF_APEX_GUARD
        for i, v in enumerate([K, G, M, Pt, etab]):
            m.p[i] = v
        prm.num_bcs = 2
        prm.bcs[0].kind = capi.BC_GRAVITY
        prm.bcs[0].f[2] = -9.81
        prm.bcs[1].kind = capi.BC_NODE_DOMAIN
        for a in range(3):
            prm.bcs[1].i[a] = 1
            prm.bcs[1].i[3 + a] = 1
        self.prm = prm

    def _view(self, t):
        ptr = lambda k: t[k].data_ptr() if k in t else None
        return capi.Particles(self.n, *[ptr(k) for k in _SET_KEYS])

    def scene(self):
        s = capi.Scene()
        s.ctx = self.ctx.ptr
        s.prm = self.prm
        s.grid = self.grid
        s.set[0] = self._view(self.sets[0])
        s.set[1] = self._view(self.sets[1])
        s.part = capi.Spatial(*[self.sp[k].data_ptr() for k in ["hash", "hash_sorted", "sorted", "cell_start",
                                                                  "cell_end"]], None)
        s.nodes = capi.Nodes(*[self.nd[k].data_ptr() for k in ["mass", "mom", "mom_nt", "vel"]])
        s.musl = 1
        return s

    def positions(self, cur):
        """Positions (n, 3) of buffer set `cur`, in the particles' original order."""
        t = self.sets[cur]
        out = torch.empty_like(t["x"])
        out[:, t["id"].long()] = t["x"]
        return out.T


def column_batch(count=64, seed=1, n_min=16000, n_max=1000000, per_axis=2, dt=1e-5, device="cuda"):
    """The multi-scene sweep: `count` columns, parameters drawn with
    numpy.random.default_rng(seed)."""
    return [ColumnScene(n, h, phi, E, per_axis=per_axis, dt=dt, seed=seed * 1000 + i, device=device)
            for i, (n, h, phi, E) in enumerate(sweep_parameters(count, seed, n_min, n_max))]


class Batch:
    """A pmpm_batch over a list of scenes; `cur` is the buffer set the next
    step reads."""

    def __init__(self, scenes, lanes=4):
        self.scenes = scenes
        self.arr = (capi.Scene * len(scenes))(*[s.scene() for s in scenes])
        self.ptr = capi.vp()
        capi.check(capi.lib().pmpm_batch_create(C.byref(self.ptr), int(lanes)), "pmpm_batch_create")
        # the batch's origin stream (a graph cannot be captured on the NULL
        # stream); it follows and rejoins the caller's current stream
        self.stream = torch.cuda.Stream()
        self.cur = 0
        self.graphs = False

    @property
    def particles(self):
        return sum(s.n for s in self.scenes)

    def _on_stream(self, fn):
        caller = torch.cuda.current_stream()
        self.stream.wait_stream(caller)
        fn(self.stream.cuda_stream)
        caller.wait_stream(self.stream)

    def step(self, steps=1):
        L, n = capi.lib(), len(self.scenes)
        for _ in range(steps):
            if self.graphs:
                self._on_stream(lambda s: capi.check(L.pmpm_batch_replay(self.ptr, self.cur, s), "pmpm_batch_replay"))
            else:
                self._on_stream(lambda s: capi.check(L.pmpm_batch_step(self.ptr, self.arr, n, self.cur, s),
                                                     "pmpm_batch_step"))
            self.cur ^= 1

    def capture(self):
        """Captures both parities; later steps replay them."""
        L, n = capi.lib(), len(self.scenes)
        for c in (0, 1):
            self._on_stream(lambda s: capi.check(L.pmpm_batch_capture(self.ptr, self.arr, n, c, s),
                                                 "pmpm_batch_capture"))
        self.graphs = True

    def tuning(self):
        t = capi.BatchTuning()
        capi.check(capi.lib().pmpm_batch_get_tuning(self.ptr, C.byref(t)), "pmpm_batch_get_tuning")
        return {"lanes": t.lanes, "use_graph": t.use_graph, "step_ms": t.step_ms}

    def autotune(self, steps=2):
        """Runs the library's autotuner (advances every scene), keeps its
        setting, and captures graphs when the setting replays them."""
        t = capi.BatchTuning()
        L, n = capi.lib(), len(self.scenes)
        self._on_stream(lambda s: capi.check(L.pmpm_batch_autotune(self.ptr, self.arr, n, self.cur, int(steps), s,
                                                                   C.byref(t)), "pmpm_batch_autotune"))
        self.graphs = False
        if t.use_graph:
            self.capture()
        return {"lanes": t.lanes, "use_graph": t.use_graph, "step_ms": t.step_ms}

    def __del__(self):
        try:
            if self.ptr:
                capi.lib().pmpm_batch_destroy(self.ptr)
                self.ptr = capi.vp()
        except Exception:
            pass
