"""GPU parity of the particle- and node-side boundary conditions beyond
gravity and NodeDomain: PlanarDomain (frictional walls acting on the particle
forces before P2G, planardomain_inline.h:56-138) and BodyForce (node moments,
bodyforce_inline.h:41-59), against the CPU oracle on the same inputs, plus the
reference's vibrating-bar acceptance run (BodyForce "fixed" end,
tests/examples/vibrating_bar/test_vibrating_bar.py:126-166).

Bars: multi-step runs within 1e-4 field-scaled relative error of the FP32
oracle; node arrays after one step within 2e-5; the vibrating bar's
centre-of-mass velocity within the reference's MSE < 1e-3 of the analytic
mode, and within 1e-3 of the oracle's trace (relative to its amplitude).
"""
import math
import sys

import numpy as np
import pytest

from tests import scenarios as S

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"\n[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _compare_run(oracle_mod, spec, steps, tol, label):
    _banner(f"{label}: n={len(spec.positions)} steps={steps}")
    pm, solver = S.build_product(spec)
    sim = S.build_oracle(oracle_mod, spec, "f32")
    solver.solve_nsteps(steps)
    for _ in range(steps):
        sim.step()
    got, ref = S.product_state(solver), S.oracle_state(sim)
    errs = {k: S.max_rel_err(got[k], ref[k]) for k in S.FIELDS}
    print(f"[gpu-test] {label} field-scaled errors: {errs}", flush=True)
    for k, e in errs.items():
        assert e <= tol, (label, k, e)
    return solver, sim


def _walls(dim, f0, f1):
    return ("PlanarDomain", {"face0_friction": f0[:dim], "face1_friction": f1[:dim]})


def _column_planar(shape_function, musl, per_axis=2):
    """A DP column on frictional planar walls (no node walls): its particles
    against x = 0, y = 0 and the floor sit within one particle radius of the
    wall (half the particles overlap it at 10% jitter)."""
    spec = S.column(h=0.05, per_axis=per_axis, dt=1e-4, domain=(1.0, 0.5, 0.5), shape_function=shape_function,
                    musl=musl)
    spec.bcs = [("Gravity", {"gravity": [0.0, 0.0, -9.81]}), _walls(3, [20.0, 15.0, 30.0], [5.0, 10.0, 0.0])]
    return spec


@pytest.mark.parametrize("musl", [False, True])
def test_planar_domain_3d_cubic_parity(oracle_mod, musl):
    """typed cubic 3D: the wall force enters the matrix-core P2G's force product
    and the wrap-around sums"""
    spec = _column_planar("cubic", musl)
    first = S.build_oracle(oracle_mod, spec, "f32")
    first.reset()
    first.apply_particle_bcs()
    touching = np.count_nonzero(np.abs(np.asarray(first.get("forces_external"))).sum(axis=1))
    assert touching > 100, "the walls must push on particles in the first step"
    _compare_run(oracle_mod, spec, 10, 1e-4, f"planar walls DP column cubic {'MUSL' if musl else 'USL'}")


@pytest.mark.parametrize("sf", ["quadratic", "linear"])
def test_planar_domain_3d_valu_parity(oracle_mod, sf):
    spec = _column_planar(sf, False)
    _compare_run(oracle_mod, spec, 10, 1e-4, f"planar walls DP column {sf} USL")


@pytest.mark.parametrize("sf", ["cubic", "linear"])
def test_planar_domain_2d_parity(oracle_mod, sf):
    h = 0.05
    pos = S.box_lattice([0.0, 0.0], [0.3, 0.2], h / 2, jitter=0.1, seed=5)
    vel = np.tile([0.4, -0.2], (len(pos), 1))
    spec = S.Spec(dim=2, origin=[0.0, 0.0], end=[1.0, 1.0], cell_size=h, shape_function=sf, dt=1e-4, ppc=4,
                  positions=pos, velocities=vel,
                  materials=[("LinearElastic", {"density": 1000.0, "E": 1e5, "pois": 0.3})],
                  bcs=[("Gravity", {"gravity": [0.0, -9.81]}), _walls(2, [25.0, 35.0], [0.0, 0.0])])
    _compare_run(oracle_mod, spec, 20, 1e-4, f"planar walls 2D block {sf}")


def _body_force_block(oracle_mod, dim, sf, musl):
    h = 0.05
    hi = [0.3, 0.25, 0.2][:dim]
    pos = S.box_lattice([0.0] * dim, hi, h / 2, jitter=0.1, seed=6)
    g = np.zeros(dim)
    g[-1] = -9.81
    spec = S.Spec(dim=dim, origin=[0.0] * dim, end=[1.0] * dim, cell_size=h, shape_function=sf, dt=1e-4,
                  ppc=2 ** dim, positions=pos, velocities=np.tile(np.linspace(0.1, 0.3, dim), (len(pos), 1)),
                  materials=[("LinearElastic", {"density": 1000.0, "E": 1e5, "pois": 0.3})], musl=musl)
    nn = len(S.build_oracle(oracle_mod, spec, "f32").get("node_masses"))  # the grid's node count
    n_ax = round(nn ** (1.0 / dim))  # a cube of nodes
    assert n_ax ** dim == nn
    ids = np.stack(np.unravel_index(np.arange(nn), [n_ax] * dim, order="F"), axis=1)  # x fastest
    rng = np.random.default_rng(8)
    fixed_mask = ids[:, 0] == 1                      # the node column next to the x = 0 wall
    mom_mask = (ids[:, -1] <= 2) & ~fixed_mask      # the lowest node layers
    frc_mask = (ids[:, 0] >= 3) & (ids[:, 0] <= 5)  # a band in x
    fixed_vals = np.zeros((nn, dim))
    fixed_vals[:, -1] = 1e-4
    spec.bcs = [("Gravity", {"gravity": list(g)}),
                ("BodyForce", {"mode": "moments", "values": rng.normal(0, 1e-4, (nn, dim)), "mask": mom_mask}),
                ("BodyForce", {"mode": "forces", "values": rng.normal(0, 1e-4, (nn, dim)), "mask": frc_mask}),
                ("BodyForce", {"mode": "fixed", "values": fixed_vals, "mask": fixed_mask})]
    return spec


@pytest.mark.parametrize("dim,sf,musl", [(2, "cubic", False), (2, "linear", True), (3, "cubic", False),
                                         (3, "cubic", True), (3, "quadratic", False), (1, "linear", False)])
def test_body_force_parity(oracle_mod, dim, sf, musl):
    spec = _body_force_block(oracle_mod, dim, sf, musl)
    _compare_run(oracle_mod, spec, 10, 1e-4, f"body forces {dim}D {sf} {'MUSL' if musl else 'USL'}")


@pytest.mark.parametrize("dim", [2, 3])
def test_body_force_node_arrays(oracle_mod, dim):
    """One USL step: masses, moments and forward moments of every node (the
    masked nodes without mass included: a body force acts whatever the mass)."""
    spec = _body_force_block(oracle_mod, dim, "cubic", False)
    _banner(f"body force node arrays {dim}D")
    pm, solver = S.build_product(spec)
    sim = S.build_oracle(oracle_mod, spec, "f32")
    solver.solve_nsteps(1)
    sim.step()
    nodes = solver.nodes
    for name, got, key in [("masses", nodes.masses, "node_masses"), ("moments", nodes.moments, "node_moments"),
                           ("moments_nt", nodes.moments_nt, "node_moments_nt")]:
        a = np.asarray(got, dtype=np.float64).reshape(np.asarray(sim.get(key)).shape)
        e = S.max_rel_err(a, np.asarray(sim.get(key), dtype=np.float64))
        print(f"[gpu-test] node {name}: {e}", flush=True)
        assert e <= 2e-5, (name, e)


@pytest.mark.parametrize("musl", [False, True])
def test_vibrating_bar_acceptance(oracle_mod, musl):
    """The reference's vibrating bar (USL and MUSL): centre-of-mass velocity
    every 20 steps over 50 s against v0/(beta0 L) cos(beta0 c t), MSE < 1e-3
    (test_vibrating_bar.py:150-166), and against the oracle's trace."""
    spec, m = S.vibrating_bar(musl=musl)
    total = int(50.0 / m["dt"])
    _banner(f"vibrating bar {'MUSL' if musl else 'USL'}: {total} steps")
    pm, solver = S.build_product(spec)
    sim = S.build_oracle(oracle_mod, spec, "f32")
    mass = np.asarray(solver.particles.masses, dtype=np.float64)
    vc, vo, ts = [], [], []
    for i in range(0, total, 20):
        v = np.asarray(solver.particles.velocities, dtype=np.float64)[:, 0]
        vc.append(float((v * mass).sum() / mass.sum()))
        w = np.asarray(sim.get("velocities"), dtype=np.float64)[:, 0]
        vo.append(float((w * mass).sum() / mass.sum()))
        ts.append(i * m["dt"])
        n = min(20, total - i)
        solver.solve_nsteps(n)
        for _ in range(n):
            sim.step()
    ts, vc, vo = map(np.asarray, (ts, vc, vo))
    amp = m["v0"] / (m["beta0"] * m["L"])
    exact = np.cos(m["beta0"] * math.sqrt(m["E"] / m["rho"]) * ts) * amp
    mse = float(((exact - vc) ** 2).mean())
    dev = float(np.abs(vc - vo).max()) / amp
    print(f"[gpu-test] vibrating bar mse {mse:.3e}, max deviation from the oracle {dev:.3e} of the amplitude",
          flush=True)
    assert mse < 1e-3
    assert dev < 1e-3
