"""Pins of the oracle functions that the reference's unit tests leave open:
the spin terms of the Drucker-Prager and mu(I) rate forms, and the rigid
level-set contact (nearest rigid particle, mass-gradient normal, stick rule).

The spin checks do not restate the update formula. They build the velocity
gradient from an explicit rigid-rotation velocity field v = w x x by finite
differences, in the reference's convention L_ij = d v_j / d x_i
(usl_inline.h:313 accumulates dpsi * v^T), and compare the update with the
rotation Q = exp(dt [w]x) of the body (Rodrigues' formula):
  * Drucker-Prager (dp_rheo_inline.h:148-156) co-rotates the elastic strain:
    eps' = Q eps Q^T + O(dt^2);
  * local mu(I) (localrheo_inline.h:110-116) counter-rotates the stress under
    that convention: sigma' = Q^T sigma Q + O(dt^2) (a reference quirk: its
    Jaumann term -sigma W + W sigma is written for the transposed gradient).
A flipped sign, a dropped term or a transposed operand moves the result by
O(dt) and fails.

The rigid level-set check is a brute force over all particles (no cell lists,
no stored shape-function table) of rigidbodylevelset_inline.h:65-199 and
rigidbodylevelset.cpp:152-217, with shape-function gradients evaluated from
their closed forms, on a tiny 2D scene, plus a physical check that a body
driven into a static rigid layer sticks and one moving away does not."""
import math

import numpy as np
import pytest

from tests.conftest import oracle_sim


def _skew(w):
    return np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])


def _rodrigues(w):
    th = float(np.linalg.norm(w))
    K = _skew(w / th)
    return np.eye(3) + math.sin(th) * K + (1.0 - math.cos(th)) * K @ K


def _commutator(w, A):
    """|K A - A K| for K = [w / |w|]x: co- and counter-rotation by the angle
    t differ by 2 t |K A - A K| to first order."""
    K = _skew(w / np.linalg.norm(w))
    return float(np.linalg.norm(K @ A - A @ K))


def _velgrad_reference_convention(vel, x, dim, eps=1e-4):
    """L_ij = d v_j / d x_i by central differences of the velocity field."""
    L = np.zeros((dim, dim))
    for i in range(dim):
        d = np.zeros(dim)
        d[i] = eps
        L[i] = (vel(x + d) - vel(x - d)) / (2 * eps)
    return L


def _one(oracle_mod, dim, dt):
    s = oracle_sim(oracle_mod, dim, "f64", [0.0] * dim, [1.0] * dim, 0.1, "linear", dt, 1)
    s.set_particles(np.full((1, dim), 0.45))
    return s


DP = {"density": 1500.0, "E": 1e6, "pois": 0.3, "M": 1.2, "Pt": 0.0}


def test_drucker_prager_rate_form_corotates_elastic_strain_3d(oracle_mod):
    dt = 1e-3
    om = np.array([0.3, -0.5, 0.8])
    L = _velgrad_reference_convention(lambda x: np.cross(om, x), np.array([0.4, 0.5, 0.6]), 3)
    Q = _rodrigues(om * dt)
    eps0 = np.array([[-2e-3, 4e-4, 1e-4], [4e-4, -1e-3, -3e-4], [1e-4, -3e-4, -1.5e-3]])
    s = _one(oracle_mod, 3, dt)
    s.add_material("DP_rheo", dict(DP))  # rate form: is_velgrad_strain_increment = 0
    s.initialize()
    s.set_material(0, "eps_e", eps0[None])
    s.set("velocity_gradient", L[None])
    s.stress_update()
    ep = s.get_material(0, "eps_e")[0]
    turn = float(np.linalg.norm(om)) * dt
    co, counter = np.linalg.norm(ep - Q @ eps0 @ Q.T), np.linalg.norm(ep - Q.T @ eps0 @ Q)
    assert co < 2 * turn ** 2 * np.linalg.norm(eps0), (co, counter)
    assert counter > turn * _commutator(om, eps0), (co, counter)
    # invariants of the rotated strain: trace exactly, deviatoric norm to O(dt^2)
    assert abs(np.trace(ep) - np.trace(eps0)) < 1e-15
    dev = lambda e: e - np.trace(e) / 3 * np.eye(3)
    assert abs(np.linalg.norm(dev(ep)) - np.linalg.norm(dev(eps0))) < 2 * turn ** 2 * np.linalg.norm(eps0)
    # the state stayed elastic, so the stress is Hooke's law of the new strain
    # (stress_ref = 0 after initialize): K tr(eps) I + 2G dev(eps)
    E, nu = 1e6, 0.3
    K, G = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    np.testing.assert_allclose(s.get("stresses")[0], K * np.trace(ep) * np.eye(3) + 2 * G * dev(ep),
                               rtol=1e-9, atol=1e-9)


def test_drucker_prager_rate_form_corotates_elastic_strain_2d(oracle_mod):
    dt, om = 1e-3, 0.7
    L = _velgrad_reference_convention(lambda x: om * np.array([-x[1], x[0]]), np.array([0.4, 0.5]), 2)
    th = om * dt
    Q = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    J = np.array([[0.0, -1.0], [1.0, 0.0]])
    eps0 = np.array([[-2e-3, 5e-4], [5e-4, -0.5e-3]])
    s = _one(oracle_mod, 2, dt)
    s.add_material("DP_rheo", dict(DP))
    s.initialize()
    e3 = np.zeros((3, 3))
    e3[:2, :2] = eps0
    s.set_material(0, "eps_e", e3[None])
    s.set("velocity_gradient", L[None])
    s.stress_update()
    ep = s.get_material(0, "eps_e")[0][:2, :2]
    co, counter = np.linalg.norm(ep - Q @ eps0 @ Q.T), np.linalg.norm(ep - Q.T @ eps0 @ Q)
    assert co < 2 * th ** 2 * np.linalg.norm(eps0), (co, counter)
    assert counter > th * np.linalg.norm(eps0 @ J - J @ eps0), (co, counter)


def test_drucker_prager_increment_form_has_no_spin(oracle_mod):
    """With is_velgrad_strain_increment the gradient is the strain increment
    itself (dp_rheo_inline.h:146-147): a skew 'increment' adds nothing."""
    W = _skew(np.array([1e-3, -2e-3, 0.5e-3]))
    eps0 = np.diag([-1e-3, -2e-3, -1.5e-3])
    s = _one(oracle_mod, 3, 1.0)
    prm = dict(DP)
    prm["is_velgrad_strain_increment"] = 1
    s.add_material("DP_rheo", prm)
    s.initialize()
    s.set_material(0, "eps_e", eps0[None])
    s.set("velocity_gradient", W[None])
    s.stress_update()
    np.testing.assert_allclose(s.get_material(0, "eps_e")[0], eps0 + W, atol=1e-18)


def test_local_rheology_spin_counter_rotates_stress(oracle_mod):
    dt = 1e-3
    om = np.array([0.3, -0.5, 0.8])
    L = _velgrad_reference_convention(lambda x: np.cross(om, x), np.array([0.4, 0.5, 0.6]), 3)
    Q = _rodrigues(om * dt)
    sig0 = -1000.0 * np.eye(3) + np.array([[-100.0, 50.0, 0.0], [50.0, 80.0, 20.0], [0.0, 20.0, 20.0]])
    s = _one(oracle_mod, 3, dt)
    s.add_material("LocalGranularRheology", {"density": 1500.0, "E": 1e6, "pois": 0.3, "I0": 0.279, "mu_s": 0.38,
                                             "mu_2": 0.64, "rho_c": 900.0})
    s.set("stresses", sig0[None])
    s.initialize()
    s.set("velocity_gradient", L[None])
    s.stress_update()
    sp = s.get("stresses")[0]
    turn = float(np.linalg.norm(om)) * dt
    counter, co = np.linalg.norm(sp - Q.T @ sig0 @ Q), np.linalg.norm(sp - Q @ sig0 @ Q.T)
    assert counter < 2 * turn ** 2 * np.linalg.norm(sig0), (counter, co)
    assert co > turn * _commutator(om, sig0), (counter, co)
    # pressure exactly, shear invariant to O(dt^2) (elastic branch: tau <= mu_s p)
    assert abs(np.trace(sp) - np.trace(sig0)) < 1e-9
    dev = lambda a: a - np.trace(a) / 3 * np.eye(3)
    tau = lambda a: math.sqrt(0.5 * np.sum(dev(a) ** 2))
    assert tau(sig0) < 0.38 * 1000.0
    assert abs(tau(sp) - tau(sig0)) < 2 * turn ** 2 * np.linalg.norm(sig0)


# ------------------------------------------------------------ rigid level set
def _cubic_type(i, n):
    # nodes.cpp:135-173: 1 on the boundary, 2 next to it, 3 inside (4 never)
    if i == 0 or i == n - 1:
        return 1
    if i == 1:
        return 2
    return 3


def _window(sf):
    # 2D windows, y outer and x inner (global_settings.cpp:137-152)
    lo, w = (0, 2) if sf == "linear" else (-1, 4)
    return [(lo + k % w, lo + k // w) for k in range(w * w)]


def _entry_grad(oracle_mod, sf, t, b, off, nx, ny, inv_h):
    """d N / d x of the window entry `off` of a particle at cell coordinate t
    in bin b, from the closed forms (tent function; typed cubic B-spline)."""
    node = (b[0] + off[0], b[1] + off[1])
    r = (t[0] - node[0], t[1] - node[1])
    if sf == "linear":
        N = [1 - abs(r[0]), 1 - abs(r[1])]
        dN = [-math.copysign(inv_h, r[0]) if r[0] != 0 else 0.0, -math.copysign(inv_h, r[1]) if r[1] != 0 else 0.0]
        if abs(r[0]) > 1 or abs(r[1]) > 1:
            return np.zeros(2)
    else:
        ty = (_cubic_type(node[0], nx), _cubic_type(node[1], ny))
        N = [oracle_mod.cubic_N(r[0], ty[0]), oracle_mod.cubic_N(r[1], ty[1])]
        dN = [oracle_mod.cubic_dN(r[0], inv_h, ty[0]), oracle_mod.cubic_dN(r[1], inv_h, ty[1])]
    return np.array([dN[0] * N[1], N[0] * dN[1]])


def _rigid_scene(oracle_mod, sf, rigid_mode, omega, body_v, seed=3):
    rng = np.random.default_rng(seed)
    h, dt = 0.1, 1e-3
    xs = np.arange(0.3, 0.7, 0.05) + 0.025
    ys = np.arange(0.1, 0.4, 0.05) + 0.025
    body = np.array([(x, y) for y in ys for x in xs]) + rng.uniform(-0.01, 0.01, (len(xs) * len(ys), 2))
    rig = np.array([(x, 0.445) for x in np.arange(0.25, 0.76, 0.05)]) + rng.uniform(-0.01, 0.01, (11, 2))
    vel = np.tile(np.asarray(body_v, float), (len(body), 1)) + rng.uniform(-0.05, 0.05, (len(body), 2))
    s = oracle_sim(oracle_mod, 2, "f64", [0.0, 0.0], [1.0, 1.0], h, sf, dt, 4)
    s.set_particles(body, vel, None, rig, None, None)
    s.add_material("LinearElastic", {"density": 1000.0, "E": 1e4, "pois": 0.3})
    s.add_bc("RigidBodyLevelSet", {"mode": rigid_mode, "euler_angles_per_second": [omega, 0.0], "COM": [0.5, 0.5]})
    s.initialize()
    s.reset()
    s.partition()
    s.calculate_shape_function()
    s.p2g()
    return s


@pytest.mark.parametrize("sf", ["linear", "cubic"])
def test_rigid_levelset_matches_brute_force(oracle_mod, sf):
    omega, dt, h = 2.0, 1e-3, 0.1
    s = _rigid_scene(oracle_mod, sf, 1, omega, (0.0, 0.3))
    nx, ny = s.grid_info()["num_cells"]
    nn = nx * ny
    inv_h = 1.0 / h
    nsn = s.grid_info()["nsn"]
    x0 = s.get("positions")
    v0 = s.get("velocities")
    m = s.get("masses")
    nm = s.get("node_masses")
    P0, Q0 = s.get("node_moments"), s.get("node_moments_nt")
    nbody = len(x0) - 11
    rigid = np.arange(len(x0)) >= nbody
    s.apply_node_bcs(False)
    P1, Q1 = s.get("node_moments"), s.get("node_moments_nt")

    # rigid velocities of the loop rotation: angle step de = -omega dt
    de = -omega * dt
    R = np.array([[math.cos(de), -math.sin(de)], [math.sin(de), math.cos(de)]])
    com = np.array([0.5, 0.5])
    vb_all = v0.copy()
    vb_all[rigid] = ((x0[rigid] - com) @ R.T - (x0[rigid] - com)) / dt
    np.testing.assert_allclose(s.get("velocities")[rigid], vb_all[rigid], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(s.get("positions")[rigid], x0[rigid] + vb_all[rigid] * dt, rtol=0, atol=1e-14)

    t = x0 * inv_h  # origin 0
    bins = np.trunc(t).astype(int)
    hashes = bins[:, 0] + bins[:, 1] * nx
    lin_win = _window("linear")
    full_win = _window(sf)
    P, Q = P0.copy(), Q0.copy()
    contacts = 0
    for node in range(nn):
        if not nm[node] > 1e-9:
            continue
        nb = (node % nx, node // nx)
        normal = np.zeros(2)
        best, best_id = 999999999999999.0, -1
        for sid, off in enumerate(lin_win):
            sb = (nb[0] - off[0], nb[1] - off[1])  # P2G window = negated G2P window
            hsh = sb[0] + sb[1] * nx
            if hsh < 0 or hsh >= nn:
                continue
            for p in np.nonzero(hashes == hsh)[0]:  # stable order: ascending index
                if rigid[p]:
                    d2 = float(np.sum((t[p] - np.array(sb)) ** 2))
                    if d2 < best:
                        best, best_id = d2, p
                else:
                    # the reference indexes dpsi with the linear window size
                    # whatever the spline (rigidbodylevelset_inline.h:147-148)
                    flat = p * len(lin_win) + sid
                    q, k = flat // nsn, flat % nsn
                    normal += m[p] * _entry_grad(oracle_mod, sf, t[q], bins[q], full_win[k], nx, ny, inv_h)
        if best_id < 0:
            continue
        nrm = np.linalg.norm(normal)
        if nrm > 0:
            normal = normal / nrm
        vb = vb_all[best_id]
        if np.dot(P0[node] / nm[node] - vb, normal) >= 0:
            P[node] = vb * nm[node]
            contacts += 1
        if np.dot(Q0[node] / nm[node] - vb, normal) >= 0:
            Q[node] = vb * nm[node]
    assert contacts > 0 and np.any(np.all(P == P0, axis=1) & (nm > 1e-9))
    np.testing.assert_allclose(P1, P, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(Q1, Q, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("direction,sticks", [((0.0, 0.5), True), ((0.0, -0.5), False)])
def test_rigid_levelset_body_into_static_layer(oracle_mod, direction, sticks):
    """A body driven up into a static rigid layer sticks to it (node
    velocity 0 wherever a rigid particle is in reach); driven away, no node
    changes. STATIC mode: rigid velocity 0 (identity rotation)."""
    s = _rigid_scene(oracle_mod, "linear", 0, 0.0, direction, seed=4)
    nm = s.get("node_masses")
    P0 = s.get("node_moments")
    s.apply_node_bcs(False)
    P1 = s.get("node_moments")
    live = nm > 1e-9
    changed = np.any(P1 != P0, axis=1) & live
    if sticks:
        assert changed.sum() >= 5
        np.testing.assert_allclose(P1[changed], 0.0, atol=0)
        # the contact nodes are those next to the rigid layer (y = 0.4 and 0.5)
        ys = (np.nonzero(changed)[0] // 11) * 0.1
        assert set(np.round(ys, 6)) <= {0.4, 0.5}
    else:
        assert not changed.any()
