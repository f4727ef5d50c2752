"""Pin the CPU oracle where the reference's goldens leave it open: closed forms,
invariants and special cases chosen so that a dropped term, a wrong sign or a
transposed operand fails. (FP64 oracle unless stated.)"""
import math

import numpy as np
import pytest

from tests.conftest import oracle_sim


def one_particle(mod, dim=3, prec="f64", dt=0.1):
    s = oracle_sim(mod, dim, prec, [0.0] * dim, [1.0] * dim, 0.1, "linear", dt, 1)
    s.set_particles(np.full((1, dim), 0.45))
    return s


# ------------------------------------------------------------ shape functions
def test_quadratic_closed_form(oracle_mod):
    m = oracle_mod
    # textbook quadratic B-spline values
    for r, v in [(0.0, 0.75), (0.5, 0.5), (-0.5, 0.5), (1.0, 0.125), (1.5, 0.0), (0.25, 0.6875), (1.25, 0.03125)]:
        assert abs(m.quad_N(r) - v) < 1e-15, (r, m.quad_N(r), v)
    # derivative = central difference of N (inv_h = 1)
    for r in np.linspace(-1.45, 1.45, 59):
        if min(abs(abs(r) - 0.5), abs(abs(r) - 1.5)) < 1e-3:
            continue
        fd = (m.quad_N(r + 1e-6) - m.quad_N(r - 1e-6)) / 2e-6
        assert abs(m.quad_dN(r, 1.0) - fd) < 1e-6


def test_cubic_interior_is_textbook_bspline(oracle_mod):
    m = oracle_mod
    for r in np.linspace(-2.2, 2.2, 89):
        a = abs(r)
        ref = (2 / 3 - a * a + 0.5 * a ** 3) if a < 1 else ((2 - a) ** 3 / 6 if a < 2 else 0.0)
        assert abs(m.cubic_N(r, 3) - ref) < 1e-14
        if min(abs(a - 1), abs(a - 2), a) > 1e-3:
            fd = (m.cubic_N(r + 1e-6, 3) - m.cubic_N(r - 1e-6, 3)) / 2e-6
            assert abs(m.cubic_dN(r, 1.0, 3) - fd) < 1e-6
    # untyped nodes (node types not built) carry no weight
    assert m.cubic_N(0.3, 0) == 0.0 and m.cubic_dN(0.3, 1.0, 0) == 0.0


@pytest.mark.parametrize("sf", ["linear", "quadratic", "cubic"])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_partition_of_unity_and_linear_reproduction(oracle_mod, sf, dim):
    """Interior particles: sum psi = 1, sum dpsi = 0, sum psi x_i = x_p,
    sum dpsi (x) x_i = I (cubic: away from the upper boundary, where the
    reference never assigns node type 4)."""
    rng = np.random.default_rng(1)
    h = 0.1
    x = rng.uniform(0.25, 0.7, size=(50, dim))
    s = oracle_sim(oracle_mod, dim, "f64", [0.0] * dim, [1.0] * dim, h, sf, 0.1, 1)
    s.set_particles(x)
    s.partition()
    s.calculate_shape_function()
    nsn = s.grid_info()["nsn"]
    psi = s.get("psi").reshape(len(x), nsn)
    dpsi = s.get("dpsi").reshape(len(x), nsn, dim)
    bins = s.get("bins")
    lo = 0 if sf == "linear" else -1
    w = 2 if sf == "linear" else 4
    offs = []
    for k in range(nsn):
        if dim == 1:
            offs.append([lo + k])
        elif dim == 2:
            offs.append([lo + k % w, lo + k // w])
        else:
            offs.append([lo + (k // w) % w, lo + k // (w * w), lo + k % w])
    offs = np.array(offs)
    for p in range(len(x)):
        xi = (bins[p] + offs) * h
        assert abs(psi[p].sum() - 1) < 1e-12
        np.testing.assert_allclose(dpsi[p].sum(axis=0), 0, atol=1e-10)
        np.testing.assert_allclose((psi[p][:, None] * xi).sum(axis=0), x[p], atol=1e-12)
        np.testing.assert_allclose(np.einsum("ka,kb->ab", dpsi[p], xi), np.eye(dim), atol=1e-10)


# ------------------------------------------------------------ transfers
def _scene(oracle_mod, dim, sf, n=400, prec="f64", seed=5, density=1000.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.3, 0.7, size=(n, dim))
    v = rng.normal(size=(n, dim))
    s = oracle_sim(oracle_mod, dim, prec, [0.0] * dim, [1.0] * dim, 0.05, sf, 1e-3, 4)
    s.set_particles(x, v)
    s.add_material("LinearElastic", {"density": density, "E": 1e4, "pois": 0.3})
    s.initialize()
    st = rng.normal(size=(n, 3, 3))
    st = 0.5 * (st + st.transpose(0, 2, 1))
    s.set("stresses", st)
    return s, x, v


@pytest.mark.parametrize("sf", ["linear", "quadratic", "cubic"])
@pytest.mark.parametrize("dim", [2, 3])
def test_p2g_conserves_mass_momentum_and_internal_force(oracle_mod, sf, dim):
    s, x, v = _scene(oracle_mod, dim, sf)
    s.reset()
    s.partition()
    s.calculate_shape_function()
    s.p2g()
    m = s.get("masses")
    nm = s.get("node_masses")
    # nodes at or below the reference's mass cutoff (1e-9) are left empty
    # (usl_inline.h:167-170), so the sum may lose up to cutoff per node
    cut = 1e-9 * len(nm)
    assert abs(nm.sum() - m.sum()) <= cut
    vmax = np.abs(v).max()
    np.testing.assert_allclose(s.get("node_moments").sum(axis=0), (m[:, None] * v).sum(axis=0), rtol=0,
                               atol=1e-11 + cut * vmax)
    # mom_nt = mom + f dt and sum_i f_i = -sum_p V sigma sum_i grad N = 0;
    # heavy particles keep every touched node above the mass cutoff
    s, x, v = _scene(oracle_mod, dim, sf, density=1e12)
    s.set("velocities", np.zeros_like(v))
    s.reset()
    s.partition()
    s.calculate_shape_function()
    s.p2g()
    f = (s.get("node_moments_nt") - s.get("node_moments")) / s.dt
    np.testing.assert_allclose(f.sum(axis=0), 0, atol=1e-9 * np.abs(f).max())


@pytest.mark.parametrize("sf", ["linear", "cubic"])
def test_g2p_uniform_velocity_is_preserved(oracle_mod, sf):
    """PIC (alpha 0) and FLIP (alpha 1) both return a uniform, stress-free
    velocity field unchanged; L = 0 and F stays I."""
    for alpha in (0.0, 1.0):
        s, x, _ = _scene(oracle_mod, 3, sf)
        n = len(x)
        s.set("velocities", np.tile([0.3, -0.2, 0.1], (n, 1)))
        s.set("stresses", np.zeros((n, 3, 3)))
        s.alpha = alpha
        s.reset()
        s.partition()
        s.calculate_shape_function()
        s.p2g()
        s.g2p(0)
        np.testing.assert_allclose(s.get("velocities"), np.tile([0.3, -0.2, 0.1], (n, 1)), atol=1e-12)
        np.testing.assert_allclose(s.get("velocity_gradient"), 0, atol=1e-10)
        np.testing.assert_allclose(s.get("positions"), x + 1e-3 * np.array([0.3, -0.2, 0.1]), atol=1e-12)


def test_g2p_pic_flip_blend(oracle_mod):
    """v_new = (1-a) v_pic + a (v + dv): the alpha=0.25 result is the same
    blend of the alpha=0 and alpha=1 results."""
    out = {}
    for alpha in (0.0, 1.0, 0.25):
        s, x, v = _scene(oracle_mod, 2, "linear")
        s.alpha = alpha
        s.reset(); s.partition(); s.calculate_shape_function(); s.p2g(); s.g2p(0)
        out[alpha] = s.get("velocities")
    np.testing.assert_allclose(out[0.25], 0.75 * out[0.0] + 0.25 * out[1.0], atol=1e-12)


# ------------------------------------------------------------ materials
def _single(oracle_mod, kind, params, L, stress=None, dim=3, dt=0.1, volume=None, mass=None):
    s = one_particle(oracle_mod, dim, "f64", dt)
    s.add_material(kind, params)
    if stress is not None:
        s.set("stresses", np.asarray(stress)[None])
    s.initialize()
    if volume is not None:
        s.set("volumes", np.array([volume]))
    if mass is not None:
        s.set("masses", np.array([mass]))
    s.set("velocity_gradient", np.asarray(L)[None])
    s.stress_update()
    return s


def test_linear_elastic_reference_form(oracle_mod):
    """sigma += (lambda + 2G) deps (the reference's lambda*deps*I is a matrix
    product, linearelastic_inline.h:74-75), not lambda tr(deps) I + 2G deps."""
    E, nu, dt = 0.1, 0.1, 0.1
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    L = np.array([[0.1, 0.2, 0.0], [0.0, 0.05, 0.0], [0.3, 0.0, -0.1]])
    s = _single(oracle_mod, "LinearElastic", {"density": 1000.0, "E": E, "pois": nu}, L)
    de = 0.5 * (L + L.T) * dt
    np.testing.assert_allclose(s.get("stresses")[0], (lam + 2 * G) * de, atol=1e-15)
    # the stale reference test's isotropic value (0.0113636) must not appear
    s = _single(oracle_mod, "LinearElastic", {"density": 1000.0, "E": E, "pois": nu}, 0.1 * np.eye(3))
    np.testing.assert_allclose(s.get("stresses")[0][0, 0], 0.00102273, atol=1e-8)


def test_newton_fluid_closed_form(oracle_mod):
    K, gam, mu = 2e4, 7.0, 0.5
    L = np.array([[0.1, 0.3, 0.0], [0.1, -0.2, 0.0], [0.0, 0.0, 0.05]])
    s = _single(oracle_mod, "NewtonFluid", {"density": 1000.0, "viscosity": mu, "bulk_modulus": K, "gamma": gam}, L)
    V0 = s.get("volumes_original")[0]
    s.set("volumes", np.array([0.9 * V0]))
    s.stress_update()
    D = 0.5 * (L + L.T)
    p = -K * ((1 / 0.9) ** gam - 1)
    ref = p * np.eye(3) + (2 * mu / 3) * (D - np.trace(D) / 3 * np.eye(3))
    np.testing.assert_allclose(s.get("stresses")[0], ref, rtol=1e-12)


def test_local_rheology_branches(oracle_mod):
    p = {"density": 1500.0, "E": 1e6, "pois": 0.3, "I0": 0.279, "mu_s": 0.38, "mu_2": 0.64, "rho_c": 900.0,
         "particle_diameter": 0.001, "particle_density": 2450.0}
    E, nu, dt = 1e6, 0.3, 1e-3
    G = 0.5 * E / (1 + nu)
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    sig0 = -1000.0 * np.eye(3)
    # elastic (tau <= mu_s p): Jaumann trial kept
    L = np.array([[0.0, 0.01, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    s = _single(oracle_mod, "LocalGranularRheology", p, L, stress=sig0, dt=dt)
    D, W = 0.5 * (L + L.T), 0.5 * (L - L.T)
    trial = sig0 + dt * (2 * G * D + lam * np.trace(D) * np.eye(3) - sig0 @ W + W @ sig0)
    np.testing.assert_allclose(s.get("stresses")[0], trial, rtol=1e-12)
    # yielding: tau' is the smaller root of t^2 + B t + H = 0, direction kept
    L = np.array([[0.0, 50.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    s = _single(oracle_mod, "LocalGranularRheology", p, L, stress=sig0, dt=dt)
    D, W = 0.5 * (L + L.T), 0.5 * (L - L.T)
    trial = sig0 + dt * (2 * G * D + lam * np.trace(D) * np.eye(3) - sig0 @ W + W @ sig0)
    pt = -np.trace(trial) / 3
    s0 = trial + pt * np.eye(3)
    tau = math.sqrt(0.5 * np.sum(s0 * s0))
    assert tau > 0.38 * pt
    out = s.get("stresses")[0]
    so = out + pt * np.eye(3)
    tau_n = math.sqrt(0.5 * np.sum(so * so))
    al = G * 0.279 * dt * math.sqrt(pt / 2450.0) / 0.005
    B = -(0.64 * pt + tau + al)
    H = 0.64 * pt * tau + 0.38 * pt * al
    assert abs(tau_n * tau_n + B * tau_n + H) < 1e-6 * tau * tau
    assert tau_n <= (-B - math.sqrt(B * B - 4 * H)) / 2 + 1e-9
    np.testing.assert_allclose(so / tau_n, s0 / tau, atol=1e-9)
    # tension -> stress-free
    s = _single(oracle_mod, "LocalGranularRheology", p, np.zeros((3, 3)), stress=1000.0 * np.eye(3), dt=dt)
    np.testing.assert_allclose(s.get("stresses")[0], 0, atol=0)


def _dp(M=1.2, Pt=0.0, psi=0.0, E=1e6, nu=0.3, guard=1):
    return {"density": 1500.0, "E": E, "pois": nu, "M": M, "Pt": Pt, "dilatancy_angle": psi, "apex_guard": guard}


def test_drucker_prager_elastic_cone_apex(oracle_mod):
    E, nu, dt = 1e6, 0.3, 1.0
    K, G = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    eta = 1.2 / math.sqrt(3)
    # elastic: small compression, increment used directly (velgrad = strain increment)
    de = np.diag([-1e-4, -1e-4, -1e-4]) + np.array([[0, 1e-5, 0], [1e-5, 0, 0], [0, 0, 0]])
    s = one_particle(oracle_mod, 3, "f64", dt)
    prm = _dp()
    prm["is_velgrad_strain_increment"] = 1
    s.add_material("DP_rheo", prm)
    s.initialize()
    s.set("velocity_gradient", de[None])
    s.stress_update()
    ref = K * np.trace(de) * np.eye(3) + 2 * G * (de - np.trace(de) / 3 * np.eye(3))
    np.testing.assert_allclose(s.get("stresses")[0], ref, rtol=1e-10)
    np.testing.assert_allclose(s.get_material(0, "eps_e")[0], de, rtol=1e-12)
    # cone: large shear under compression -> back on the yield surface,
    # deviatoric direction kept, pressure unchanged when psi = 0
    de = np.diag([-1e-3, -1e-3, -1e-3]) + np.array([[0, 5e-3, 0], [5e-3, 0, 0], [0, 0, 0]])
    s = one_particle(oracle_mod, 3, "f64", dt)
    s.add_material("DP_rheo", prm)
    s.initialize()
    s.set("velocity_gradient", de[None])
    s.stress_update()
    out = s.get("stresses")[0]
    p = np.trace(out) / 3
    sdev = out - p * np.eye(3)
    J2 = 0.5 * np.sum(sdev * sdev)
    assert abs(math.sqrt(J2) + eta * (p - 0.0)) < 1e-6 * abs(p)
    np.testing.assert_allclose(p, K * np.trace(de), rtol=1e-12)
    tr_dev = 2 * G * (de - np.trace(de) / 3 * np.eye(3))
    np.testing.assert_allclose(sdev / np.linalg.norm(sdev), tr_dev / np.linalg.norm(tr_dev), atol=1e-12)
    # apex: tension with psi = 0 -> p = Pt, s = 0 (guarded); unguarded
    # reproduces the reference's 0*inf = NaN
    de = np.diag([1e-3, 1e-3, 1e-3])
    for guard in (1, 0):
        s = one_particle(oracle_mod, 3, "f64", dt)
        prm2 = dict(prm)
        prm2["apex_guard"] = guard
        prm2["Pt"] = 10.0
        s.add_material("DP_rheo", prm2)
        s.initialize()
        s.set("velocity_gradient", de[None])
        s.stress_update()
        out = s.get("stresses")[0]
        if guard:
            np.testing.assert_allclose(out, 10.0 * np.eye(3), atol=1e-12)
        else:
            assert np.isnan(out).any()
    # apex with dilatancy: p = Pt exactly by the closed form
    s = one_particle(oracle_mod, 3, "f64", dt)
    prm3 = dict(prm)
    prm3["dilatancy_angle"] = 10.0
    prm3["Pt"] = 10.0
    s.add_material("DP_rheo", prm3)
    s.initialize()
    s.set("velocity_gradient", de[None])
    s.stress_update()
    np.testing.assert_allclose(s.get("stresses")[0], 10.0 * np.eye(3), rtol=1e-10)


def test_drucker_prager_friction_form():
    """M = 6 sin(phi) / (3 - sin(phi)) matches Mohr-Coulomb in triaxial
    compression: sqrt(J2) = -eta p at yield with c = 0."""
    from tests.scenarios import dp_params
    M, Pt = dp_params(30.0, 0.0)
    assert abs(M - 1.2) < 1e-12 and Pt == 0.0
    M, Pt = dp_params(32.0, 50.0)
    assert abs(Pt - 50.0 / math.tan(math.radians(32.0))) < 1e-9
    # triaxial compression at yield: s1 = s2 = -p0, s3 = -p0 - q
    phi = math.radians(32.0)
    sig3 = -100.0
    sig1 = sig3 * (1 + math.sin(phi)) / (1 - math.sin(phi))  # MC, c = 0 (compression negative)
    st = np.diag([sig1, sig3, sig3])
    p = np.trace(st) / 3
    sd = st - p * np.eye(3)
    assert abs(math.sqrt(0.5 * np.sum(sd * sd)) + M / math.sqrt(3) * p) < 1e-9 * abs(p)


# ------------------------------------------------------------ boundary conditions
def test_node_domain_fixed_and_slip(oracle_mod):
    s = oracle_sim(oracle_mod, 2, "f64", [0.0, 0.0], [1.0, 1.0], 0.5, "linear", 1.0, 1)
    s.set_particles(np.array([[0.1, 0.1]]))
    nm = np.ones(9)
    mom = np.tile([-1.0, 2.0], (9, 1))
    s.set("node_masses", nm)
    s.set("node_moments", mom)
    s.set("node_moments_nt", mom)
    s.add_bc("NodeDomain", {"face0_mode": [1, 0], "face1_mode": [1, 1]})
    s.apply_node_bcs()
    out = s.get("node_moments_nt")
    ids = s.get("node_ids")
    for i in range(9):
        x, y = ids[i]
        exp = np.array([-1.0, 2.0])
        if x == 0:
            exp[0] = max(0.0, exp[0])
        elif x == 2:
            exp[0] = min(0.0, exp[0])
        if y == 0:
            exp[:] = 0.0
        elif y == 2:
            exp[1] = min(0.0, exp[1])
        np.testing.assert_allclose(out[i], exp)


def test_box_sdf_contact(oracle_mod):
    """A node inside a static box, moving toward it, takes the box velocity; a
    node moving away (release) keeps its own."""
    s = oracle_sim(oracle_mod, 3, "f64", [0.0] * 3, [1.0] * 3, 0.5, "linear", 1.0, 1)
    s.set_particles(np.array([[0.1, 0.1, 0.1]]))
    nm = np.ones(27)
    ids = s.get("node_ids")
    mom = np.zeros((27, 3))
    centre = np.array([0.5, 0.5, 0.95])
    for i in range(27):
        mom[i] = [0.0, 0.0, 1.0] if ids[i][2] == 1 else [0.0, 0.0, -1.0]
    s.set("node_masses", nm)
    s.set("node_moments", mom)
    s.set("node_moments_nt", mom)
    s.add_bc("RigidBodySDF", {"half_extents": [1.0, 1.0, 0.1], "center": list(centre), "velocity": [0.0, 0.0, 0.0],
                              "band": 0.6})
    s.apply_node_bcs()
    out = s.get("node_moments_nt")
    for i in range(27):
        z = ids[i][2] * 0.5
        phi = abs(z - 0.95) - 0.1
        if phi < 0.6 and z < 0.95:
            # below the box: moving up (toward it) sticks, moving down releases
            exp = [0, 0, 0] if ids[i][2] == 1 else [0, 0, -1.0]
            np.testing.assert_allclose(out[i], exp, atol=1e-12)
