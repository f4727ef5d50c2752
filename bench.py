#!/usr/bin/env python
"""Benchmark of the explicit MPM step on one MI355X.

Workload (the metric's configuration): 3D granular column collapse,
Drucker-Prager (phi 30 deg, psi 0, c 0), MUSL, cubic B-splines, domain
1.0 x 0.5 x 0.5 m at h = 0.01 m (101 x 51 x 51 nodes), column
0.2 x 0.5 x 0.4 m filled at 27 particles per cell with +-10% seeded jitter
(1,080,000 particles), gravity, slip walls, dt = 1e-5 s.

A step is one full MUSL update of all particles: spatial sort, P2G, node
BCs, first map + re-map, node BCs, second map + return map. Inputs are
resident in HBM before the timed region. `value` times pmpm_step (the C ABI's
whole-step entry point) on one stream; a second timed region issues the same
step entry point by entry point (pmpm_partition, pmpm_p2g, pmpm_grid_bc,
pmpm_musl_map, pmpm_grid_bc, pmpm_musl_g2p) with HIP events on the launch
stream around each (the per-entry split); a third runs pmpm_step with the
library's kernel timing on (a HIP event pair around every kernel, on its
launch stream), which gives each kernel's mean launch time. The roofline
kernel is the one with the largest share of the step.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/s at 1M DP particles; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per particle of each entry point (FP32, 3D, cubic,
# Drucker-Prager, MUSL): the particle fields each one must read and write.
BYTES_PER_PARTICLE = {
    "partition": 56,     # hash: x 12 in, key 4 out; 10-bit pass 1: key 4 (histogram) + key 4 in, key 4 +
                         # slot 4 out; pass 2: key 4 (histogram) + key/slot 8 in, 8 out; ranges: key 4
    "p2g": 64,           # sorted 4, x 12, v 12, m 4, V 4, stress 24, + node flush
    "musl_map": 56,      # fused first map + re-map: sorted 4, x 12, v 12, m 4 in; x 12, v 12 out
    "musl_g2p": 252,     # in: sorted 4 x 12 F 36 V0 4 stress 24 eps 24 m 4 colour 1 flags 1 id 4
                         # out: F 36 V 4 stress 24 eps 24 L 36 m 4 V0 4 colour 1 flags 1 id 4
}
# Algorithmic HBM bytes of each kernel of the step, per particle and per grid
# node of one launch (the compulsory reads and writes of a fused design).
KERNEL_BYTES = {
    "hash_kernel<3>": (16, 36),             # x 12 in, key 4 out; clears cell ranges 8 + node m, P, Q 28
    "radix_hist_kernel<10>": (4, 0),        # key 4
    "radix_scatter_kernel<10>": (16, 0),    # key + slot 8 in, 8 out
    "cell_range_kernel": (4, 0),            # sorted key 4
    "p2g_mfma_kernel<2, 0>": (56, 0),       # x 12, v 12, m 4, V 4, stress 24 (P2G of the particle state)
    "p2g_mfma_kernel<2, 1>": (36, 0),       # re-map: sorted 4, hash 4, x 12, m 4, v' 12
    "g2p_kernel<3, 2, 1>": (52, 0),         # first map: sorted 4, x 12, v 12 in; x 12, v 12 out
    "g2p_kernel<3, 2, 2>": (252, 0),        # second map + return map (entry musl_g2p above)
    "grid_kernel<3> gather 1": (0, 88),     # m, P, Q in 28; m, P, Q out 28, velocity record 32
    "grid_kernel<3> gather 2": (0, 72),     # m, P, Q in 28; Q out 12, velocity record 32
}
# The step's algorithmic bytes per particle (compulsory traffic of a fused
# step: sort 16, P2G 56, MUSL map 52, second map + return map 188 + L 36,
# grid ~23).
STEP_BYTES_PER_PARTICLE = 335
# HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE)
# and scaled by the same counters on a plain device copy of known size.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_kernels_dp_column_1080k_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--per-axis", type=int, default=3, help="particles per cell per axis (3 -> 27 ppc)")
    ap.add_argument("--cell-size", type=float, default=0.01)
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps timed for the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward time ratio")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-cell-size", type=float, default=0.02,
                    help="cell size of the differentiated column (0.02 -> 40,000 particles at 2 per axis)")
    ap.add_argument("--adjoint-per-axis", type=int, default=2)
    ap.add_argument("--no-batched", action="store_true", help="skip the multi-scene sweep")
    ap.add_argument("--batched-scenes", type=int, default=64)
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the differentiated multi-scene sweep")
    ap.add_argument("--ensemble-steps", type=int, default=40)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="before the workload, run 3 plain device copies of MIB MiB (calibrates HBM byte counters)")
    return ap.parse_args()


class Workload:
    """The 3D Drucker-Prager column collapse held in torch tensors on the GPU,
    driven through the C ABI (ctypes)."""

    def __init__(self, per_axis=3, h=0.01, seed=0):
        import torch
        from pyroclastmpm_amd import capi
        from tests import scenarios as S

        self.torch, self.capi = torch, capi
        spec = S.column(h=h, per_axis=per_axis, dt=1e-5, jitter=0.1, seed=seed)
        self.spec = spec
        pos = spec.positions.astype(np.float32)
        self.n = n = pos.shape[0]
        dev = torch.device("cuda")
        self.grid = capi.make_grid(spec.origin, spec.end, h, 3)
        nn = self.nn = self.grid.num_cells_total
        # particles.cpp:355-358: (Real)pow(h, D) / (Real)ppc
        vol = np.float32(np.float32(float(np.float32(h)) ** 3) / np.float32(per_axis ** 3))

        def f32(shape, fill=0.0):
            return torch.full(shape, fill, dtype=torch.float32, device=dev)

        def set_(fill_x):
            d = {}
            d["x"] = torch.from_numpy(np.ascontiguousarray(pos.T)).to(dev).clone() if fill_x else f32((3, n))
            d["v"] = f32((3, n))
            d["stress"] = f32((6, n))
            F = f32((9, n))
            F[0] = 1.0
            F[4] = 1.0
            F[8] = 1.0
            d["F"] = F
            d["L"] = f32((9, n))
            d["mass"] = f32((n,), float(np.float32(1500.0) * vol))
            d["volume"] = f32((n,), float(vol))
            d["volume0"] = f32((n,), float(vol))
            d["color"] = torch.zeros(n, dtype=torch.uint8, device=dev)
            d["flags"] = torch.full((n,), capi.PF_ACTIVE, dtype=torch.uint8, device=dev)
            d["id"] = torch.arange(n, dtype=torch.int32, device=dev)
            d["eps_e"] = f32((6, n))
            d["stress_ref"] = f32((6, n))
            return d

        self.sets = [set_(True), set_(False)]
        self.cur = 0
        self.nodes_t = {k: f32((nn,) if k == "mass" else (3 * nn,)) for k in ["mass", "mom", "mom_nt"]}
        self.nodes_t["vel"] = f32((capi.node_vel_stride(3) * nn,))
        self.sp_t = {k: torch.zeros(n, dtype=torch.int32, device=dev) for k in ["hash", "hash_sorted", "sorted"]}
        self.sp_t["cell_start"] = torch.zeros(nn, dtype=torch.int32, device=dev)
        self.sp_t["cell_end"] = torch.zeros(nn, dtype=torch.int32, device=dev)
        self.ctx = capi.Context(n, nn)
        self.nodes = capi.Nodes(*[self.nodes_t[k].data_ptr() for k in ["mass", "mom", "mom_nt", "vel"]])
        self.spatial = capi.Spatial(self.sp_t["hash"].data_ptr(), self.sp_t["hash_sorted"].data_ptr(),
                                    self.sp_t["sorted"].data_ptr(), self.sp_t["cell_start"].data_ptr(),
                                    self.sp_t["cell_end"].data_ptr(), None)
        self.views = [self._view(s) for s in self.sets]
        prm = capi.Params()
        prm.dim, prm.shape_fn, prm.dt, prm.alpha = 3, capi.SF["cubic"], 1e-5, 0.99
        prm.small_mass_cutoff, prm.node_types_built, prm.wrap_quirk = 1e-9, 1, 1
        p = spec.materials[0][1]
        E, nu = np.float32(p["E"]), np.float32(p["pois"])
        K = E / (np.float32(3.0) * (np.float32(1.0) - np.float32(2.0) * nu))
        G = E / (np.float32(2.0) * (np.float32(1.0) + nu))
        prm.num_materials = 1
        m = prm.materials[0]
        m.kind, m.flags = capi.MAT_DRUCKER_PRAGER, capi.MATF_UPDATE_HISTORY | capi.MATF_APEX_GUARD
        for i, val in enumerate([K, G, p["M"], p["Pt"], 0.0, 0.0]):
            m.p[i] = float(val)
        prm.num_bcs = 2
        prm.bcs[0].kind = capi.BC_GRAVITY
        prm.bcs[0].f[2] = -9.81
        prm.bcs[1].kind = capi.BC_NODE_DOMAIN
        for a in range(3):
            prm.bcs[1].i[a] = 1
            prm.bcs[1].i[3 + a] = 1
        self.prm = prm
        self.stream = torch.cuda.current_stream().cuda_stream

    def _view(self, d):
        return self.capi.Particles(self.n, *[d[k].data_ptr() for k in ["x", "v", "stress", "F", "L", "mass",
                                                                          "volume", "volume0", "color", "flags",
                                                                          "id", "eps_e", "stress_ref"]])

    def calls(self):
        """The step as (name, thunk) entry-point calls, in order."""
        L, capi = self.capi.lib(), self.capi
        a, b = self.views[self.cur], self.views[self.cur ^ 1]
        g, prm, sp, nd, s = C.byref(self.grid), C.byref(self.prm), C.byref(self.spatial), C.byref(self.nodes), \
            self.stream
        nb = self.prm.num_bcs
        return [
            ("partition", lambda: L.pmpm_partition(self.ctx.ptr, 3, g, a.x, self.n, sp, s)),
            ("p2g", lambda: L.pmpm_p2g(self.ctx.ptr, prm, g, C.byref(a), sp, nd, s)),
            ("grid_bc", lambda: L.pmpm_grid_bc(prm, g, nd, 0, nb, 0, s)),
            ("musl_map", lambda: L.pmpm_musl_map(self.ctx.ptr, prm, g, C.byref(a), C.byref(b), sp, nd, s)),
            ("grid_bc_remap", lambda: L.pmpm_grid_bc(prm, g, nd, 0, nb, 1, s)),
            ("musl_g2p", lambda: L.pmpm_musl_g2p(prm, g, C.byref(a), C.byref(b), sp, nd, s)),
        ]

    def step(self, events=None):
        for i, (name, fn) in enumerate(self.calls()):
            if events is not None:
                events[i][0].record()
            rc = fn()
            if rc != 0:
                raise RuntimeError(f"{name} failed with code {rc}")
            if events is not None:
                events[i][1].record()
        self.cur ^= 1

    def fused_step(self):
        L = self.capi.lib()
        a, b = self.views[self.cur], self.views[self.cur ^ 1]
        rc = L.pmpm_step(self.ctx.ptr, C.byref(self.prm), C.byref(self.grid), 1, C.byref(a), C.byref(b),
                         C.byref(self.spatial), C.byref(self.nodes), self.stream)
        if rc != 0:
            raise RuntimeError(f"pmpm_step failed with code {rc}")
        self.cur ^= 1


def cpu_baseline(spec, steps):
    """The oracle (FP32, one thread, the reference's node-gather loop
    structure) on the same workload for `steps` steps after initialisation."""
    from oracle import _oracle
    from tests import scenarios as S

    sim = S.build_oracle(_oracle, spec, "f32")
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    dt = time.perf_counter() - t0
    n = len(spec.positions)
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": n * steps / dt, "unit": "particle-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP32 single thread (g++ -O3), {steps} steps of the full {n}-particle workload, "
                      f"{dt:.1f} s, on {model} ({os.cpu_count()} logical CPUs visible, 1 used)"}


def adjoint_ratio(steps, every=20, h=0.02, per_axis=2):
    """The differentiable 3D Drucker-Prager column: by default h = 0.02
    (51 x 26 x 26 nodes), 8 particles per cell (40,000 particles), USL, cubic,
    dt = 1e-4 (1e-5 below h = 0.02), checkpoints every 20 steps. Loss = mean squared distance of the final
    positions to a target run (phi 32 deg, c 50 Pa, E 1.2e6); gradient with
    respect to (phi, c, E) from (28 deg, 0, 1e6). Returns the backward time per
    step (segment recomputation included) over the forward time per step."""
    import torch
    from pyroclastmpm_amd.diff import DiffMPM, simulate
    from tests import scenarios as S

    dt = 1e-4 if h >= 0.02 else 1e-5
    spec = S.column(h=h, per_axis=per_axis, dt=dt, jitter=0.1, seed=0, musl=False)
    model = DiffMPM(3, spec.origin, spec.end, h, spec.positions, per_axis ** 3, dt=dt, checkpoint_every=every)
    with torch.no_grad():
        target = simulate(model, torch.tensor([32.0, 50.0, 1.2e6], dtype=torch.float64), steps)
    theta = torch.tensor([28.0, 0.0, 1e6], dtype=torch.float64, requires_grad=True)
    x = simulate(model, theta, min(steps, 2 * every))  # warm-up (kernels, workspaces)
    ((x - target) ** 2).mean().backward()
    theta.grad = None
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    torch.cuda.synchronize()
    ev[0].record()
    x = simulate(model, theta, steps)
    ev[1].record()
    loss = ((x - target) ** 2).mean()
    loss.backward()
    ev[2].record()
    torch.cuda.synchronize()
    fwd, bwd = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    return {"metric": "adjoint/forward time ratio", "value": round(bwd / fwd, 3), "unit": "ratio",
            "higher_is_better": False, "forward_ms_per_step": round(fwd / steps, 4),
            "backward_ms_per_step": round(bwd / steps, 4), "steps": steps, "checkpoint_every": every,
            "particles": model.n, "loss": float(loss.detach()), "grad": [float(v) for v in theta.grad],
            "config": f"3D Drucker-Prager column, h {h}, {per_axis ** 3} per cell, USL, cubic, dt {dt}"}


def batched_throughput(count, steps):
    """The multi-scene sweep: `count` Drucker-Prager columns (MUSL, cubic,
    dt 1e-5) with particle counts log-uniform in [16k, 1M], h uniform in
    [0.01, 0.02] m, friction angle in [20, 40] deg, E in [5e5, 2e6] Pa (seed 1),
    stepped together by pmpm_batch_step after the library's autotuner picked
    the stream count and eager or graph replay. Also times the same scenes one
    after another on one stream (lanes 1, eager) for comparison."""
    import torch
    from pyroclastmpm_amd.batch import Batch, column_batch

    scenes = column_batch(count, seed=1)
    b = Batch(scenes, lanes=4)
    b.step(2)  # sizes every workspace
    tuning = b.autotune(steps=2)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(batch):
        torch.cuda.synchronize()
        ev[0].record()
        batch.step(steps)
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / steps

    ms = timed(b)
    serial = Batch(scenes, lanes=1)
    serial.cur = b.cur
    ms_serial = timed(serial)
    n = b.particles
    sizes = sorted(s.n for s in scenes)
    return {"metric": "particle-updates/s, multi-scene column sweep", "value": round(n / (ms * 1e-3), 1),
            "unit": "particle-updates/s", "higher_is_better": True, "scenes": count, "particles": n,
            "smallest": sizes[0], "largest": sizes[-1], "steps": steps, "ms_per_step": round(ms, 4),
            "tuning": tuning, "one_stream_eager_ms_per_step": round(ms_serial, 4),
            "speedup_vs_one_stream": round(ms_serial / ms, 3),
            "config": "3D Drucker-Prager columns, MUSL, cubic, 8 per cell, dt 1e-5, seed 1"}


def ensemble_throughput(count, steps, every=20):
    """The multi-scene sweep differentiated: `count` columns with the sweep's
    geometry and materials (seed 1), USL, cubic, dt 1e-5, checkpoints every
    `every` steps; per-scene loss = mean squared displacement of the final
    positions; gradient with respect to each scene's (phi, c, E). The
    ensemble runs its forward through pmpm_batch_step (lanes and graph replay
    from the library's autotuner) and its reverse pass through
    pmpm_batch_adjoint_step. Compared with the same scenes run one after
    another, each through diff.simulate and its backward."""
    import torch
    from pyroclastmpm_amd.diff import simulate
    from pyroclastmpm_amd.ensemble import Ensemble, column_ensemble, simulate_ensemble

    models, thetas = column_ensemble(count, seed=1, dt=1e-5, checkpoint_every=every)
    x0s = [m.positions(m.initial) for m in models]
    ens = Ensemble(models, lanes=4)
    tuning = ens.autotune(thetas)
    graph = bool(tuning["use_graph"])
    th = thetas.clone().requires_grad_(True)
    xs = simulate_ensemble(ens, th, 2, graph=graph)  # sizes the segment pools and adjoint workspaces
    sum(((x - x0) ** 2).mean() for x, x0 in zip(xs, x0s)).backward()
    th.grad = None
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    torch.cuda.synchronize()
    ev[0].record()
    xs = simulate_ensemble(ens, th, steps, graph=graph)
    ev[1].record()
    sum(((x - x0) ** 2).mean() for x, x0 in zip(xs, x0s)).backward()
    ev[2].record()
    torch.cuda.synchronize()
    fwd, bwd = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    g_ens = th.grad.clone()
    # the same scenes one after another, each alone
    rel = 0.0
    torch.cuda.synchronize()
    ev[0].record()
    grads = []
    for i, m in enumerate(models):
        t = thetas[i].clone().requires_grad_(True)
        ((simulate(m, t, steps) - x0s[i]) ** 2).mean().backward()
        grads.append(t.grad)
    ev[1].record()
    torch.cuda.synchronize()
    seq = ev[0].elapsed_time(ev[1])
    for i, g in enumerate(grads):
        rel = max(rel, float(((g_ens[i] - g).abs() / g.abs().clamp_min(1e-30)).max()))
    n = ens.particles
    return {"metric": "particle-updates/s, differentiated multi-scene column sweep (forward + adjoint)",
            "value": round(n * steps / ((fwd + bwd) * 1e-3), 1), "unit": "particle-updates/s",
            "higher_is_better": True, "scenes": count, "particles": n, "steps": steps, "checkpoint_every": every,
            "tuning": tuning, "forward_ms_per_step": round(fwd / steps, 4),
            "backward_ms_per_step": round(bwd / steps, 4), "adjoint_forward_ratio": round(bwd / fwd, 3),
            "one_after_another_ms_per_step": round(seq / steps, 4),
            "speedup_vs_one_after_another": round(seq / (fwd + bwd), 3),
            "max_rel_grad_difference_vs_alone": rel,
            "config": "3D Drucker-Prager columns, USL, cubic, 8 per cell, dt 1e-5, seed 1"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is not built)", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    if args.copy_probe:
        src = torch.ones(args.copy_probe * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"bench: copy probe {args.copy_probe} MiB read + {args.copy_probe} MiB written per launch",
              file=sys.stderr, flush=True)
        del src, dst
    wl = Workload(args.per_axis, args.cell_size)
    n = wl.n
    print(f"bench: column collapse, {n} particles, grid {list(wl.grid.num_cells)}, warmup {args.warmup}, "
          f"steps {args.steps}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        wl.fused_step()
    torch.cuda.synchronize()
    # timed region 1: the step through pmpm_step (the value)
    t_start, t_end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_start.record()
    for k in range(args.steps):
        wl.fused_step()
    t_end.record()
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    ms_per_step = total_ms / args.steps
    value = n * args.steps / (total_ms / 1e3)
    # timed region 2: the same step issued entry point by entry point, HIP
    # events on the launch stream around each (the per-entry split and the
    # roofline kernel's launch time)
    ncalls = len(wl.calls())
    ev = [[[torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)] for _ in range(ncalls)]
          for _ in range(args.steps)]
    e_start, e_end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e_start.record()
    for k in range(args.steps):
        wl.step(ev[k])
    e_end.record()
    torch.cuda.synchronize()
    entries_ms_per_step = e_start.elapsed_time(e_end) / args.steps
    names = [c[0] for c in wl.calls()]
    per_call = {nm: sum(ev[k][i][0].elapsed_time(ev[k][i][1]) for k in range(args.steps)) / args.steps
                for i, nm in enumerate(names)}
    print("bench: per-entry-point mean ms: " + json.dumps({k: round(v, 4) for k, v in per_call.items()}),
          file=sys.stderr, flush=True)
    # timed region 3: the fused step again, every kernel bracketed by HIP
    # events on its launch stream (the library's kernel timing)
    capi = wl.capi
    capi.timing_read()
    capi.timing_enable(True)
    for k in range(args.steps):
        wl.fused_step()
    capi.timing_enable(False)
    timed = capi.timing_read()
    kernels = {}
    for name, (cnt, ms) in timed.items():
        avg_ms = ms / cnt
        row = {"launches_per_step": round(cnt / args.steps, 3), "avg_us": round(avg_ms * 1e3, 2),
               "share": round(ms / sum(v[1] for v in timed.values()), 4)}
        bp = KERNEL_BYTES.get(name)
        if bp is not None:
            b = bp[0] * n + bp[1] * wl.nn
            row["bytes_per_launch"] = b
            row["gbs"] = round(b / (avg_ms * 1e-3) / 1e9, 1)
            row["frac"] = round(row["gbs"] / HBM_PEAK_GBS, 4)
        kernels[name] = row
    print("bench: kernels " + json.dumps(kernels), file=sys.stderr, flush=True)
    # the roofline kernel: the one with the largest share of the step time
    dom = max(kernels, key=lambda k: kernels[k]["share"])
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f).get(dom)
        if t and t.get("particles") == n:
            traffic = t["bytes_per_launch"]
    d = kernels[dom]
    roof = {"bound": "hbm", "kernel": dom, "achieved": d.get("gbs"), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": d.get("frac"), "traffic": traffic, "bytes_per_launch": d.get("bytes_per_launch"),
            "kernel_us": d["avg_us"], "share_of_step": d["share"],
            "bytes_per_particle": KERNEL_BYTES.get(dom, (None, None))[0],
            "step_frac": round(STEP_BYTES_PER_PARTICLE * n / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(wl.spec, args.cpu_steps)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"value": None, "error": str(e)}
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "3D Drucker-Prager MUSL column collapse, cubic B-spline, "
                               f"{n} particles (27 per cell), grid 101x51x51, dt 1e-5",
                   "particles": n, "cell_size": args.cell_size, "solver": "MUSL", "shape_function": "cubic",
                   "material": "Drucker-Prager phi=30 psi=0 c=0", "entry_point": "pmpm_step",
                   "per_entry_point_ms": per_call, "entry_by_entry_ms_per_step": round(entries_ms_per_step, 4)},
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    if not args.no_adjoint:
        try:
            line["adjoint"] = adjoint_ratio(args.adjoint_steps, h=args.adjoint_cell_size,
                                            per_axis=args.adjoint_per_axis)
            print("bench: adjoint " + json.dumps(line["adjoint"]), file=sys.stderr, flush=True)
        except Exception as e:  # reported, the throughput line stands on its own
            line["adjoint"] = {"value": None, "error": str(e)}
    if not args.no_batched:
        try:
            line["batched"] = batched_throughput(args.batched_scenes, args.batched_steps)
            print("bench: batched " + json.dumps(line["batched"]), file=sys.stderr, flush=True)
        except Exception as e:  # reported, the throughput line stands on its own
            line["batched"] = {"value": None, "error": str(e)}
    if not args.no_ensemble:
        try:
            line["ensemble"] = ensemble_throughput(args.batched_scenes, args.ensemble_steps)
            print("bench: ensemble " + json.dumps(line["ensemble"]), file=sys.stderr, flush=True)
        except Exception as e:  # reported, the throughput line stands on its own
            line["ensemble"] = {"value": None, "error": str(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
