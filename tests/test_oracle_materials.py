"""Pin the oracle's von Mises, Mohr-Coulomb and modified Cam-Clay return maps
(vonmises_inline.h, mohrcoulomb_inline.h, modifiedcamclay_inline.h) on single
material points, by what the plasticity theory fixes independently of the
code (de Souza Neto, Peric & Owen 2008, ch. 7-8; Borja & Lee 1990):

  * elastic branch: sigma = K tr(e) I + 2G dev(e) (Cam-Clay: the secant
    pressure p_prev / (1 - v/kappa de_v));
  * J2 radial return: q = sigma_y + H acc on return, pressure and deviatoric
    direction kept, acc increment (q_tr - sigma_y) / (3G + H);
  * Mohr-Coulomb main plane: sigma_1 - sigma_3 + (sigma_1 + sigma_3) sin(phi)
    = 2 c cos(phi) on return, principal directions kept; the apex of a
    dilatant cone: sigma = c cot(phi) I; the eigen-decomposition against
    numpy's;
  * Cam-Clay: the returned state lies on the updated ellipse and satisfies
    the associated volumetric flow rule, to the Newton tolerance.
Each check fails if the branch or a sign of the restatement is wrong.
"""
import math

import numpy as np
import pytest

from tests.test_oracle_closed_forms import one_particle


def _point(oracle_mod, kind, prm, de, stress=None, dim=3, volume=None, pc=None):
    s = one_particle(oracle_mod, dim, "f64", 1.0)
    p = dict(prm)
    p["is_velgrad_strain_increment"] = 1
    s.add_material(kind, p)
    if stress is not None:
        s.set("stresses", np.asarray(stress, dtype=np.float64)[None])
    s.initialize()
    if volume is not None:
        s.set("volumes", np.array([volume]))
    if pc is not None:
        s.set_material(0, "pc", np.array([pc]))
    s.set("velocity_gradient", np.asarray(de)[None])
    s.stress_update()
    return s


E, NU = 1e7, 0.25
K, G = E / (3 * (1 - 2 * NU)), E / (2 * (1 + NU))


def _dev(a):
    return a - np.trace(a) / 3 * np.eye(3)


def _q(sig):
    s = _dev(sig)
    return math.sqrt(1.5 * np.sum(s * s))


@pytest.mark.parametrize("H", [0.0, 2e5])
def test_von_mises_elastic_and_radial_return(oracle_mod, H):
    prm = {"density": 1000.0, "E": E, "pois": NU, "yield_stress": 1e4, "H": H}
    de = np.array([[1e-5, 2e-6, 0], [2e-6, -3e-6, 0], [0, 0, 1e-6]])
    s = _point(oracle_mod, "VonMises", prm, de)
    np.testing.assert_allclose(s.get("stresses")[0], K * np.trace(de) * np.eye(3) + 2 * G * _dev(de), rtol=1e-12)
    assert s.get_material(0, "alpha")[0] == 0.0
    de = np.array([[2e-3, 1e-3, 0], [1e-3, -1e-3, 5e-4], [0, 5e-4, 4e-4]])
    s = _point(oracle_mod, "VonMises", prm, de)
    out = s.get("stresses")[0]
    tr = K * np.trace(de) * np.eye(3) + 2 * G * _dev(de)
    q_tr = _q(tr)
    dg = (q_tr - 1e4) / (3 * G + H)
    acc = s.get_material(0, "alpha")[0]
    np.testing.assert_allclose(acc, dg, rtol=1e-9)
    np.testing.assert_allclose(_q(out), 1e4 + H * acc, rtol=1e-9)  # on the hardened surface
    np.testing.assert_allclose(np.trace(out), np.trace(tr), rtol=1e-12)  # pressure kept
    np.testing.assert_allclose(_dev(out) / _q(out), _dev(tr) / q_tr, atol=1e-12)  # radial
    # elastic strain of the result: s / 2G + tr(e) / 3 I
    np.testing.assert_allclose(s.get_material(0, "eps_e")[0], _dev(out) / (2 * G) + np.trace(de) / 3 * np.eye(3),
                               rtol=1e-9, atol=1e-15)


def test_jacobi_eigen_matches_numpy(oracle_mod):
    """Random states below a high cohesion return their trial stress through
    the principal-stress yield test: the stress and its eigenvalues against
    the elastic law and numpy.linalg.eigvalsh (the principal frame of the
    plastic returns is checked in the cases below)."""
    rng = np.random.default_rng(5)
    prm = {"density": 1000.0, "E": E, "pois": NU, "cohesion": 1e9, "friction_angle": 30.0, "dilatancy_angle": 5.0,
           "H": 0.0}
    for _ in range(20):
        a = rng.normal(0, 1e-4, (3, 3))
        de = 0.5 * (a + a.T)
        s = _point(oracle_mod, "MohrCoulomb", prm, de)
        out = s.get("stresses")[0]
        ref = K * np.trace(de) * np.eye(3) + 2 * G * _dev(de)
        np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-6)
        np.testing.assert_allclose(np.sort(np.linalg.eigvalsh(out)), np.sort(np.linalg.eigvalsh(ref)), rtol=1e-9)


def test_mohr_coulomb_main_plane(oracle_mod):
    phi, c = 30.0, 2e3
    prm = {"density": 1000.0, "E": E, "pois": NU, "cohesion": c, "friction_angle": phi, "dilatancy_angle": 10.0,
           "H": 0.0}
    # p = -1e4, deviator (1.6e4, 0, -1.6e4): outside the cone, sigma_2 in between
    de = np.diag([1.5e-3, -5e-4, -2.5e-3]) + np.array([[0, 3e-5, 0], [3e-5, 0, 0], [0, 0, 0]])
    s = _point(oracle_mod, "MohrCoulomb", prm, de)
    out = s.get("stresses")[0]
    w, V = np.linalg.eigh(out)
    s1, s2, s3 = w[2], w[1], w[0]
    sp, cp = math.sin(math.radians(phi)), math.cos(math.radians(phi))
    assert abs((s1 - s3) + (s1 + s3) * sp - 2 * c * cp) < 1e-6 * abs(s3)  # on the plane
    assert s1 >= s2 >= s3
    tr = K * np.trace(de) * np.eye(3) + 2 * G * _dev(de)
    wt, Vt = np.linalg.eigh(tr)
    for k in range(3):  # principal directions kept
        assert abs(abs(np.dot(V[:, k], Vt[:, k])) - 1.0) < 1e-9
    # trial was outside the cone
    assert (wt[2] - wt[0]) + (wt[2] + wt[0]) * sp - 2 * c * cp > 0
    assert s.get_material(0, "alpha")[0] > 0


def test_mohr_coulomb_apex(oracle_mod):
    phi, c = 30.0, 1e3
    prm = {"density": 1000.0, "E": E, "pois": NU, "cohesion": c, "friction_angle": phi, "dilatancy_angle": 20.0,
           "H": 0.0}
    de = np.diag([1e-3, 1e-3, 1e-3])  # pure tension far past the apex
    s = _point(oracle_mod, "MohrCoulomb", prm, de)
    out = s.get("stresses")[0]
    cot = 1.0 / math.tan(math.radians(phi))
    np.testing.assert_allclose(out, c * cot * np.eye(3), rtol=1e-9)


def test_modified_cam_clay_elastic_and_on_surface(oracle_mod):
    M, lam, kap, Vs, Rr, Pt = 1.2, 0.2, 0.04, 0.5, 2.0, 0.0
    prm = {"density": 1800.0, "E": 1e7, "pois": 0.3, "M": M, "lam": lam, "kap": kap, "Vs": Vs, "R": Rr, "Pt": Pt,
           "beta": 1.0}
    p0 = 5e4
    st0 = -p0 * np.eye(3)
    vol = 1.0
    G_ = 1e7 / (2 * 1.3)
    # elastic: small isotropic compression inside the ellipse
    de = -1e-5 * np.eye(3)
    s = _point(oracle_mod, "ModifiedCamClay", prm, de, stress=st0, volume=vol)
    vk = (vol / Vs) / kap
    p_tr = p0 / (1 - vk * (-np.trace(de)))
    np.testing.assert_allclose(s.get("stresses")[0], -p_tr * np.eye(3), rtol=1e-9)
    pc0 = s.get_material(0, "pc")[0]
    np.testing.assert_allclose(pc0, Rr * p0, rtol=1e-12)  # modifiedcamclay.cpp:88-95
    # plastic: large shear at the same pressure -> on the updated ellipse
    de = np.array([[0, 6e-3, 0], [6e-3, 0, 0], [0, 0, 0]])
    s = _point(oracle_mod, "ModifiedCamClay", prm, de, stress=st0, volume=vol)
    out = s.get("stresses")[0]
    p = -np.trace(out) / 3
    q = _q(out)
    pc = s.get_material(0, "pc")[0]
    ps = 0.5 * (pc + Pt)
    f = (ps - p - Pt) ** 2 + (q / M) ** 2 - ps ** 2
    assert abs(f) < 1e-5 * ps ** 2, (f, ps)
    # the trial lay outside the initial ellipse
    ps0 = 0.5 * (pc0 + Pt)
    q_tr = _q(2 * G_ * de)
    assert (ps0 - p0 - Pt) ** 2 + (q_tr / M) ** 2 - ps0 ** 2 > 0
    # associated flow: the plastic volumetric strain follows dF/dp, and pc
    # hardens with it (dense of critical: p < p_s -> compaction)
    epv = s.get_material(0, "alpha")[0]
    assert (p > ps) == (epv > 0)
    np.testing.assert_allclose(pc, pc0 * (1 + (vol / Vs) / (lam - kap) * epv), rtol=1e-6)
