"""GPU checks of the adjoint of the USL step (pmpm_adjoint_step) and of the
differentiable solver built on it (pyroclastmpm_amd.diff).

  * one step from a yielding state, all state and parameter directions:
    <adjoint(c), delta> against the central finite difference of
    <c, step(S + eps delta)> computed by the FP64 oracle (relative error
    <= 2e-3; the adjoint runs in FP32);
    The same in 3D on the 40,000-particle differentiable column, in the
    adjoint P2G layout that size picks and in the one-lane layout of large
    launches, and in 2D with the node conditions listed walls first;
  * several steps with checkpoint segments: d<c, x_final>/d(phi, c, E) against
    central finite differences of the FP64 oracle (relative step 1e-6,
    relative error <= 1e-3) and, as a coarse check, of the GPU forward itself
    (relative step 1e-3; the FP32 forward's roundoff makes that difference
    quotient good to a few percent only, so the bar is 10%).
"""
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


def _column_2d(h=0.05, per_axis=2, seed=3):
    pos = S.box_lattice([0.0, 0.0], [0.2, 0.3], h / per_axis, jitter=0.1, seed=seed)
    return pos, dict(dim=2, origin=[0.0, 0.0], end=[0.6, 0.4], cell_size=h, ppc=per_axis ** 2)


def _sym6_to_mat(s6, D):
    idx = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 0): 3, (0, 2): 4, (2, 0): 4, (1, 2): 5, (2, 1): 5}
    n = s6.shape[1]
    m = np.zeros((n, 3, 3))
    for (i, j), k in idx.items():
        m[:, i, j] = s6[k]
    return m


def _mat_to_sym6(m):
    return np.stack([m[:, 0, 0], m[:, 1, 1], m[:, 2, 2], m[:, 0, 1], m[:, 0, 2], m[:, 1, 2]])


def _add_bcs(sim, model, D):
    for b in model.bc_order:  # the model's list order (usl.cpp:96-103)
        if b == "gravity":
            sim.add_bc("Gravity", {"gravity": model.gravity})
        else:
            sim.add_bc("NodeDomain", {"face0_mode": [1] * D, "face1_mode": [1] * D})


def _column_3d(h=0.02, per_axis=2, seed=3, col=(0.2, 0.5, 0.4), end=(1.0, 0.5, 0.5)):
    """The differentiable column's geometry: at h = 0.02 and 8 per cell,
    40,000 particles on 51 x 26 x 26 nodes."""
    pos = S.box_lattice([0.0, 0.0, 0.0], list(col), h / per_axis, jitter=0.1, seed=seed)
    return pos, dict(dim=3, origin=[0.0, 0.0, 0.0], end=list(end), cell_size=h, ppc=per_axis ** 3)


def _oracle_for(oracle_mod, model, geo, pos, theta, state):
    """FP64 oracle holding `state` (numpy dict, the state's particle order)."""
    from pyroclastmpm_amd.diff import dp_material

    D = geo["dim"]
    K, G, M, Pt, _ = dp_material(*theta, model.pois)
    sim = getattr(oracle_mod, f"Sim{D}D_f64")(geo["origin"], geo["end"], geo["cell_size"], "cubic", model.dt,
                                             geo["ppc"])
    sim.set_particles(state["x"].T, np.zeros_like(state["x"].T), None, None)
    sim.add_material("DP_rheo", {"density": 1500.0, "E": float(theta[2]), "pois": model.pois, "M": M, "Pt": Pt,
                                 "dilatancy_angle": 0.0})
    _add_bcs(sim, model, D)
    sim.musl = False
    sim.alpha = model.alpha
    sim.initialize()
    sim.set_material(0, "bulk_modulus", K)
    sim.set_material(0, "shear_modulus", G)
    return sim


def _load(sim, st, D):
    n = st["x"].shape[1]
    sim.set("positions", st["x"].T)
    sim.set("velocities", st["v"].T)
    sim.set("stresses", _sym6_to_mat(st["stress"], D))
    sim.set("F", st["F"].T.reshape(n, D, D))
    sim.set("volumes", st["volume"][0])
    sim.set("volumes_original", st["volume0"][0])
    sim.set("masses", st["mass"][0])
    sim.set_material(0, "eps_e", _sym6_to_mat(st["eps_e"], D))
    sim.set_material(0, "stress_ref", np.zeros((n, 3, 3)))


def _outputs(sim, D):
    n = len(sim.get("positions"))
    eps = np.asarray(sim.get_material(0, "eps_e"), dtype=np.float64).reshape(n, 3, 3)
    return {"x": np.asarray(sim.get("positions")).T, "v": np.asarray(sim.get("velocities")).T,
            "F": np.asarray(sim.get("F")).reshape(n, D * D).T,
            "stress": _mat_to_sym6(np.asarray(sim.get("stresses")).reshape(n, 3, 3)),
            "eps_e": _mat_to_sym6(eps)}


def _branch_margins(sim, st, theta, model, D):
    """Distance of each particle's Drucker-Prager trial state from the branch
    boundaries of its return map (elastic/cone: Phi = 0; cone/apex with
    eta_bar = 0: p = Pt), relative to the trial's scale, for the step the FP64
    oracle takes from `st` (the trial strain is rebuilt from the oracle's
    velocity gradient as dp_rheo_inline.h:123-200 forms it, stress_ref = 0).
    The FP32 forward of the adjoint and the FP64 oracle can put a particle
    that lies within rounding of a boundary on different sides; such a
    particle's loss weight is set to zero so the comparison is of the same
    piecewise map."""
    from pyroclastmpm_amd.diff import dp_material

    K, G, M, Pt, _ = dp_material(*theta, model.pois)
    _load(sim, st, D)
    sim.step()
    n = st["x"].shape[1]
    L = np.zeros((n, 3, 3))
    L[:, :D, :D] = np.asarray(sim.get("velocity_gradient"), dtype=np.float64).reshape(n, D, D)
    eps = _sym6_to_mat(st["eps_e"], D)
    dt = model.dt
    de = 0.5 * (L + L.transpose(0, 2, 1)) * dt
    W = 0.5 * (L - L.transpose(0, 2, 1)) * dt
    et = eps + de - W @ eps + eps @ W
    if D < 3:
        et[:, 2, :] = 0.0
        et[:, :, 2] = 0.0
    ev = np.trace(et, axis1=1, axis2=2)
    p = K * ev
    dev = et - ev[:, None, None] / 3.0 * np.eye(3)
    if D < 3:
        dev[:, 2, 2] = 0.0
    s = 2.0 * G * dev
    sq = np.sqrt(0.5 * np.sum(s * s, axis=(1, 2)))
    eta = M / np.sqrt(3.0)
    phi = sq + eta * (p - Pt)
    scale = sq + eta * (np.abs(p) + abs(Pt)) + 1e-30
    return np.minimum(np.abs(phi) / scale, np.abs(p - Pt) / (np.abs(p) + abs(Pt) + 1e-30))


def _one_step_check(oracle_mod, pos, geo, bc_order=("gravity", "walls"), label="", gravity=None, keep=None):
    from pyroclastmpm_amd import capi
    from pyroclastmpm_amd.diff import AdjointSet, DiffMPM, ParticleSet
    import ctypes as C

    D = geo["dim"]
    # the soft column of the multi-step test after 30 steps: yielding
    model = DiffMPM(D, geo["origin"], geo["end"], geo["cell_size"], pos, geo["ppc"], dt=1e-3, bc_order=bc_order,
                    gravity=gravity)
    theta = (28.0, 200.0, 1e5)
    _banner(f"one-step adjoint: {D}D DP column, n={model.n}, 30 warm-up steps, node conditions {bc_order} {label}")
    state, _ = model.forward(theta, 30)
    torch.cuda.synchronize()
    st = {k: v.double().cpu().numpy() for k, v in state.t.items() if k not in ("color", "flags", "id")}
    n = model.n
    rng = np.random.default_rng(0)
    c = {"x": rng.standard_normal((D, n)), "v": rng.standard_normal((D, n)) * 1e-2,
         "F": rng.standard_normal((D * D, n)), "stress": rng.standard_normal((6, n)) * 1e-4,
         "eps_e": rng.standard_normal((6, n)) * 10.0}
    if D == 2:  # zz, xz, yz are not state in 2D
        for k in ("stress", "eps_e"):
            c[k][[2, 4, 5]] = 0.0
    sim = _oracle_for(oracle_mod, model, geo, pos, theta, st)
    near = _branch_margins(sim, st, theta, model, D) < 1e-4
    if keep is not None:  # loss on a subset of the particles only
        near = near | ~keep(st["x"])
    for k in c:
        c[k][:, near] = 0.0
    print(f"[gpu-test] {int(near.sum())} of {n} particles within 1e-4 of a return-map branch boundary: "
          "loss weight 0", flush=True)
    # output order of the step = cell order of `state`
    prm = model.params(theta)
    L = capi.lib()
    capi.check(L.pmpm_partition(model.ctx.ptr, D, C.byref(model.grid), state.t["x"].data_ptr(), n,
                                C.byref(model.spatial), torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    order = model.sp["sorted"].long().cpu().numpy()
    ob = AdjointSet(n, D)
    for k in c:
        ob.t[k].copy_(torch.from_numpy(c[k][:, order]).float())
    ib = AdjointSet(n, D)
    pbar = torch.zeros(4, dtype=torch.float32, device="cuda")
    model.adjoint_step(prm, state, ob, ib, pbar)
    torch.cuda.synchronize()
    adj = {k: v.double().cpu().numpy() for k, v in ib.t.items()}
    pb = pbar.double().cpu().numpy()

    V0 = st["volume0"][0]

    def ell(stp, params=None):
        _load(sim, stp, D)
        if params is not None:
            for name, val in params.items():
                sim.set_material(0, name, val)
        sim.step()
        o = _outputs(sim, D)
        return sum(float(np.sum(c[k] * o[k])) for k in c)

    delta = {"x": rng.standard_normal((D, n)) * 1e-3, "v": rng.standard_normal((D, n)) * 1e-3,
             "F": rng.standard_normal((D * D, n)) * 1e-3, "stress": rng.standard_normal((6, n)) * 10.0,
             "eps_e": rng.standard_normal((6, n)) * 1e-6}
    if D == 2:
        for k in ("stress", "eps_e"):
            delta[k][[2, 4, 5]] = 0.0
    results = {}
    for field in ["x", "v", "F", "stress", "eps_e"]:
        eps = 1e-4
        sp, sm = dict(st), dict(st)
        sp[field] = st[field] + eps * delta[field]
        sm[field] = st[field] - eps * delta[field]
        if field == "F":  # V = det(F) V0
            for s_ in (sp, sm):
                Fm = s_["F"].T.reshape(n, D, D)
                s_["volume"] = (np.linalg.det(Fm) * V0)[None, :]
        fd = (ell(sp) - ell(sm)) / (2 * eps)
        ad = float(np.sum(adj[field] * delta[field]))
        results[field] = (fd, ad)
    names = ["bulk_modulus", "shear_modulus", "M", "Pt"]
    base = [sim_p for sim_p in [prm.materials[0].p[i] for i in range(4)]]
    for i, name in enumerate(names):
        h = 1e-6 * max(abs(base[i]), 1.0)
        fd = (ell(st, {name: base[i] + h}) - ell(st, {name: base[i] - h})) / (2 * h)
        results[name] = (fd, float(pb[i]))
    print(f"[gpu-test] one-step adjoint {D}D {bc_order} {label} (fd, adjoint): {results}", flush=True)
    # the forward the adjoint linearises against the FP64 oracle's, same state
    out = ParticleSet(n, D)
    model.step(prm, state, out)
    torch.cuda.synchronize()
    go = {k: v.double().cpu().numpy() for k, v in out.t.items() if k in ("x", "v", "F", "stress", "eps_e")}
    sid = state.t["id"].long().cpu().numpy()
    inv = np.empty_like(sid)
    inv[sid] = np.arange(n)
    idx = inv[out.t["id"].long().cpu().numpy()]  # out slot j = slot idx[j] of `state`
    _load(sim, st, D)
    sim.step()
    oo = _outputs(sim, D)
    ferr = {k: S.max_rel_err(go[k], oo[k][:, idx]) for k in go}
    print(f"[gpu-test] one-step forward vs FP64 oracle (field-scaled): {ferr}", flush=True)
    for k, (fd, ad) in results.items():
        assert abs(fd - ad) <= 2e-3 * max(abs(fd), 1e-12) + 1e-12, (k, fd, ad)
    assert results["M"][0] != 0.0 and results["Pt"][0] != 0.0, "the state must yield"


def test_one_step_adjoint_matches_fp64_oracle(oracle_mod):
    pos, geo = _column_2d()
    _one_step_check(oracle_mod, pos, geo)


def test_one_step_adjoint_walls_before_gravity(oracle_mod):
    """Node conditions listed [NodeDomain, Gravity]: the forward clamps, then
    adds gravity; the adjoint must test each clamp on the value it saw."""
    pos, geo = _column_2d()
    _one_step_check(oracle_mod, pos, geo, bc_order=("walls", "gravity"))


@pytest.mark.parametrize("lanes", [None, "1"])
def test_one_step_adjoint_3d_matches_fp64_oracle(oracle_mod, monkeypatch, lanes):
    """The 3D kernels (LDS-staged adjoint G2P, adjoint P2G in the layout this
    launch size picks: sixteen lanes per particle at 40,000 particles; and the
    one-lane layout used from 196,608 particles, forced) on the 40,000-particle
    differentiable column, every state and parameter direction."""
    if lanes:
        monkeypatch.setenv("PMPM_ADJOINT_P2G_LANES", lanes)
    pos, geo = _column_3d()
    _one_step_check(oracle_mod, pos, geo, label=f"lanes {lanes or 'by size'}")


def test_multi_step_parameter_gradient_matches_finite_differences(oracle_mod):
    from pyroclastmpm_amd.diff import DiffMPM, simulate

    pos, geo = _column_2d(seed=5)
    D = geo["dim"]
    # a soft column (E = 1e5, elastic strain rho g H / E ~ 4%) over 60 ms, so
    # it yields and the friction and cohesion gradients are not zero
    # (elastic wave speed 8 m/s: dt 1e-3 < h / c = 6e-3)
    steps = 60
    model = DiffMPM(D, geo["origin"], geo["end"], geo["cell_size"], pos, geo["ppc"], dt=1e-3, checkpoint_every=25)
    theta0 = np.array([28.0, 200.0, 1e5])
    _banner(f"multi-step gradient: 2D DP column, n={model.n}, {steps} steps, checkpoints every 25")
    rng = np.random.default_rng(1)
    cw = torch.from_numpy(rng.standard_normal((model.n, D))).float().cuda()
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    x = simulate(model, th, steps)
    loss = (x * cw).sum()
    loss.backward()
    g_ad = th.grad.numpy()

    def f_gpu(t):
        with torch.no_grad():
            return float((simulate(model, torch.tensor(t, dtype=torch.float64), steps) * cw).sum())

    g_fd = np.zeros(3)
    for i in range(3):
        h = 1e-3 * abs(theta0[i]) if theta0[i] != 0 else 1e-3
        tp, tm = theta0.copy(), theta0.copy()
        tp[i] += h
        tm[i] -= h
        g_fd[i] = (f_gpu(tp) - f_gpu(tm)) / (2 * h)
    # FP64 oracle: the same steps from the same initial state
    from pyroclastmpm_amd.diff import dp_material

    cw_np = cw.double().cpu().numpy()

    def f_oracle(t):
        _, _, M, Pt, _ = dp_material(*t, model.pois)
        sim = getattr(oracle_mod, f"Sim{D}D_f64")(geo["origin"], geo["end"], geo["cell_size"], "cubic", model.dt,
                                                 geo["ppc"])
        sim.set_particles(pos)
        sim.add_material("DP_rheo", {"density": 1500.0, "E": float(t[2]), "pois": model.pois, "M": M, "Pt": Pt,
                                     "dilatancy_angle": 0.0})
        sim.add_bc("Gravity", {"gravity": model.gravity})
        sim.add_bc("NodeDomain", {"face0_mode": [1] * D, "face1_mode": [1] * D})
        sim.musl = False
        sim.alpha = model.alpha
        sim.initialize()
        for _ in range(steps):
            sim.step()
        return float(np.sum(np.asarray(sim.get("positions")) * cw_np))

    g_64 = np.zeros(3)
    for i in range(3):
        h = 1e-6 * abs(theta0[i]) if theta0[i] != 0 else 1e-6
        tp, tm = theta0.copy(), theta0.copy()
        tp[i] += h
        tm[i] -= h
        g_64[i] = (f_oracle(tp) - f_oracle(tm)) / (2 * h)
    print(f"[gpu-test] d<c,x>/d(phi,c,E) adjoint {g_ad} FD(GPU, FP32) {g_fd} FD(oracle, FP64) {g_64}", flush=True)
    for i in range(3):
        assert abs(g_ad[i] - g_64[i]) <= 1e-3 * abs(g_64[i]) + 1e-15, (i, g_ad, g_64)
        assert abs(g_ad[i] - g_fd[i]) <= 1e-1 * abs(g_fd[i]) + 1e-15, (i, g_ad, g_fd)
    assert all(abs(v) > 0 for v in g_64), g_64  # the column yields: every parameter matters


@pytest.mark.parametrize("lanes", [None, "1"])
def test_multi_step_parameter_gradient_3d_matches_fp64_oracle(oracle_mod, monkeypatch, lanes):
    """d<c, x_final>/d(phi, c, E) over 20 steps with a checkpoint every 10
    (recomputed segments) on a 3D column, against central finite differences
    of the FP64 oracle (relative step 1e-6) to 1e-3."""
    from pyroclastmpm_amd.diff import DiffMPM, dp_material, simulate

    if lanes:
        monkeypatch.setenv("PMPM_ADJOINT_P2G_LANES", lanes)
    pos, geo = _column_3d(h=0.025, per_axis=2, seed=7, col=(0.2, 0.2, 0.3), end=(0.6, 0.3, 0.4))
    D, steps = 3, 20
    model = DiffMPM(D, geo["origin"], geo["end"], geo["cell_size"], pos, geo["ppc"], dt=1e-3, checkpoint_every=10)
    theta0 = np.array([28.0, 200.0, 1e5])
    _banner(f"multi-step gradient: 3D DP column, n={model.n}, {steps} steps, checkpoints every 10, "
            f"lanes {lanes or 'by size'}")
    rng = np.random.default_rng(2)
    cw = torch.from_numpy(rng.standard_normal((model.n, D))).float().cuda()
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    ((simulate(model, th, steps) * cw).sum()).backward()
    g_ad = th.grad.numpy()
    cw_np = cw.double().cpu().numpy()

    def f_oracle(t):
        _, _, M, Pt, _ = dp_material(*t, model.pois)
        sim = oracle_mod.Sim3D_f64(geo["origin"], geo["end"], geo["cell_size"], "cubic", model.dt, geo["ppc"])
        sim.set_particles(pos)
        sim.add_material("DP_rheo", {"density": 1500.0, "E": float(t[2]), "pois": model.pois, "M": M, "Pt": Pt,
                                     "dilatancy_angle": 0.0})
        _add_bcs(sim, model, D)
        sim.musl = False
        sim.alpha = model.alpha
        sim.initialize()
        for _ in range(steps):
            sim.step()
        return float(np.sum(np.asarray(sim.get("positions")) * cw_np))

    g_64 = np.zeros(3)
    for i in range(3):
        h = 1e-6 * abs(theta0[i])
        tp, tm = theta0.copy(), theta0.copy()
        tp[i] += h
        tm[i] -= h
        g_64[i] = (f_oracle(tp) - f_oracle(tm)) / (2 * h)
    print(f"[gpu-test] 3D d<c,x>/d(phi,c,E) adjoint {g_ad} FD(oracle, FP64) {g_64}", flush=True)
    assert all(abs(v) > 0 for v in g_64), g_64
    for i in range(3):
        assert abs(g_ad[i] - g_64[i]) <= 1e-3 * abs(g_64[i]) + 1e-15, (i, g_ad, g_64)


def test_adjoint_with_kept_step_buffers_matches_recompute():
    """pmpm_adjoint_step(recompute=0) on the partition and nodes pmpm_step left
    equals recompute=1 bit for bit (the step's node sums do not depend on
    scheduling)."""
    from pyroclastmpm_amd.diff import AdjointSet, DiffMPM, StepBuffers, ParticleSet

    spec = S.column(h=0.04, per_axis=2, dt=1e-4, jitter=0.1, seed=2, musl=False)
    model = DiffMPM(3, spec.origin, spec.end, 0.04, spec.positions, 8, dt=1e-4)
    theta = (30.0, 20.0, 1e6)
    _banner(f"adjoint kept buffers vs recompute: 3D DP column, n={model.n}, 12 warm-up steps")
    state, _ = model.forward(theta, 12)
    prm = model.params(theta)
    n = model.n
    kept = StepBuffers(n, model.grid.num_cells_total, 3)
    nxt = ParticleSet(n, 3)
    model.step(prm, state, nxt, kept)
    g = torch.Generator(device="cuda").manual_seed(4)
    ob = AdjointSet(n, 3)
    for k, v in ob.t.items():
        v.copy_(torch.randn(v.shape, device="cuda", generator=g))
    res = []
    for recompute in (True, False):
        ib = AdjointSet(n, 3)
        pbar = torch.zeros(4, dtype=torch.float32, device="cuda")
        model.adjoint_step(prm, state, ob, ib, pbar, work=None if recompute else kept, recompute=recompute)
        torch.cuda.synchronize()
        res.append((ib, pbar))
    for k in ob.t:
        assert torch.equal(res[1][0].t[k], res[0][0].t[k]), k
    pa, pb = res[1][1], res[0][1]
    assert torch.equal(pa, pb), (pa, pb)
    assert float(pb.abs().max()) > 0


def test_adjoint_p2g_lane_layouts_agree(monkeypatch):
    """The adjoint P2G's layouts (one, four or sixteen lanes per particle,
    chosen by launch size) give the same adjoint step (field-scaled relative
    error <= 1e-5: the layouts add the window's entries in different orders).
    PMPM_ADJOINT_P2G_LANES forces the layout."""
    spec = S.column(h=0.04, per_axis=2, dt=1e-4, jitter=0.1, seed=3, musl=False)
    _lane_layouts_agree(monkeypatch, spec, "3D DP column")


def test_adjoint_p2g_lane_layouts_agree_long_runs(monkeypatch):
    """The same on a column 38 cells long in x: the one-lane layout stages
    the node records of runs of at most 32 x cells (a run ends at every
    32-aligned x block), so here runs split and several rounds of staging
    meet in one workgroup; the cells at x = 0 also take wrap-around entries."""
    spec = S.column(h=0.04, per_axis=2, dt=1e-4, jitter=0.1, seed=4, musl=False, domain=(1.6, 0.24, 0.32),
                    col=(1.5, 0.16, 0.2))
    _lane_layouts_agree(monkeypatch, spec, "3D DP column 38 cells long")


def _lane_layouts_agree(monkeypatch, spec, label):
    from pyroclastmpm_amd.diff import AdjointSet, DiffMPM, StepBuffers, ParticleSet

    model = DiffMPM(3, spec.origin, spec.end, 0.04, spec.positions, 8, dt=1e-4)
    theta = (30.0, 20.0, 1e6)
    _banner(f"adjoint P2G lanes 1 / 4 / 16: {label}, n={model.n}, 12 warm-up steps")
    state, _ = model.forward(theta, 12)
    prm = model.params(theta)
    n = model.n
    kept = StepBuffers(n, model.grid.num_cells_total, 3)
    model.step(prm, state, ParticleSet(n, 3), kept)
    g = torch.Generator(device="cuda").manual_seed(5)
    ob = AdjointSet(n, 3)
    for k, v in ob.t.items():
        v.copy_(torch.randn(v.shape, device="cuda", generator=g))
    res = []
    for lanes in ("1", "4", "16"):
        monkeypatch.setenv("PMPM_ADJOINT_P2G_LANES", lanes)
        ib = AdjointSet(n, 3)
        pbar = torch.zeros(4, dtype=torch.float32, device="cuda")
        model.adjoint_step(prm, state, ob, ib, pbar, work=kept, recompute=False)
        torch.cuda.synchronize()
        res.append((ib, pbar))
    monkeypatch.delenv("PMPM_ADJOINT_P2G_LANES")
    for other in res[1:]:
        for k in ob.t:
            a, b = res[0][0].t[k], other[0].t[k]
            e = float((a - b).abs().max()) / max(float(b.abs().max()), 1e-30)
            assert e <= 1e-5, (k, e)
        assert torch.equal(res[0][1], other[1]), (res[0][1], other[1])
    assert float(res[0][0].t["x"].abs().max()) > 0
