"""Pin the oracle's BodyForce and PlanarDomain restatements.

BodyForce: the reference's unit-test values (tests/unit/test_bodyforce.h,
transcribed in tests/golden/reference_unit_goldens.json) in 1D/2D/3D.
PlanarDomain: the reference's own test (tests/unit/test_planardomain.h) has
no assertion, so the contact force is pinned by hand-worked cases of the
penalty law of de Vaucorbeil & Nguyen (2021) as planardomain_inline.h
states it: a particle of radius r = V^(1/D)/2 overlapping a wall by
o = r - gap gets (m / dt^2) o (n - mu (v - (n.v) n)), mu the friction angle
in radians used as the coefficient; no force without overlap; corner
particles add the forces of both walls.
"""
import json
import math
import os

import numpy as np
import pytest

from tests.conftest import oracle_sim

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_unit_goldens.json")))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("mode", ["moments", "fixed"])
def test_body_force_goldens(oracle_mod, dim, prec, mode):
    g = G["body_force"]
    s = oracle_sim(oracle_mod, dim, prec, [0.0] * dim, [1.0] * dim, g["cell_size"], "linear", 1.0, 1)
    s.set_particles(np.full((1, dim), 0.1))
    nn = len(s.get("node_masses"))
    vals = np.zeros((nn, dim))
    vals[:2] = np.array(g["values"][str(dim)])
    mask = np.zeros(nn)
    mask[:2] = g["mask"]
    s.add_bc("BodyForce", {"mode": mode, "values": vals, "mask": mask})
    start = vals.copy() if mode == "moments" else np.zeros((nn, dim))
    s.set("node_moments", start)
    s.set("node_moments_nt", start)
    s.apply_node_bcs(False)
    mom, mom_nt = s.get("node_moments"), s.get("node_moments_nt")
    tol = g["tol"]
    if mode == "moments":
        np.testing.assert_allclose(mom[0], vals[0], atol=tol)  # masked out: unchanged
        np.testing.assert_allclose(mom[1], g["moments_mode"]["factor_node1"] * vals[1], atol=tol)
        np.testing.assert_allclose(mom_nt[:2], vals[:2], atol=tol)  # the forward moments are not touched
    else:
        np.testing.assert_allclose(mom_nt[0], 0.0, atol=tol)
        np.testing.assert_allclose(mom_nt[1], vals[1], atol=tol)
        np.testing.assert_allclose(mom[:2], 0.0, atol=tol)  # code, not the outdated test expectation
    # every other node: no mask, nothing changes
    np.testing.assert_allclose(mom[2:], start[2:], atol=0)
    np.testing.assert_allclose(mom_nt[2:], start[2:], atol=0)


def test_body_force_remap_pass_applies_only_fixed(oracle_mod):
    s = oracle_sim(oracle_mod, 2, "f64", [0.0, 0.0], [1.0, 1.0], 0.5, "linear", 1.0, 1)
    s.set_particles(np.full((1, 2), 0.1))
    nn = len(s.get("node_masses"))
    vals = np.tile([1.0, -2.0], (nn, 1))
    s.add_bc("BodyForce", {"mode": "moments", "values": vals, "mask": np.ones(nn)})
    s.add_bc("BodyForce", {"mode": "fixed", "values": 3 * vals, "mask": np.ones(nn)})
    s.set("node_moments", np.zeros((nn, 2)))
    s.set("node_moments_nt", np.zeros((nn, 2)))
    s.apply_node_bcs(True)  # the MUSL re-map pass
    np.testing.assert_array_equal(s.get("node_moments"), 0.0)
    np.testing.assert_array_equal(s.get("node_moments_nt"), 3 * vals)


def _planar_sim(mod, prec, dim, pos, vel, vol, mass, dt, f0, f1, end=1.0):
    s = oracle_sim(mod, dim, prec, [0.0] * dim, [end] * dim, 0.1, "linear", dt, 1)
    s.set_particles(np.atleast_2d(pos), np.atleast_2d(vel))
    s.add_bc("PlanarDomain", {"face0_friction": f0, "face1_friction": f1})
    s.set("volumes", np.atleast_1d(vol))
    s.set("masses", np.atleast_1d(mass))
    s.reset()
    s.apply_particle_bcs()
    return np.asarray(s.get("forces_external"), dtype=np.float64)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_planar_domain_face0_with_friction(oracle_mod, prec):
    # radius 0.05 (V = 1e-3 in 3D), 0.02 from the x = 0 wall: overlap 0.03;
    # m / dt^2 = 2 / 0.01 = 200, scale 6; friction 30 deg = pi/6 rad on x0
    f = _planar_sim(oracle_mod, prec, 3, [0.02, 0.5, 0.5], [-1.0, 2.0, 3.0], 1e-3, 2.0, 0.1,
                    [30.0, 0.0, 0.0], [0.0, 0.0, 0.0])[0]
    mu = math.pi / 6
    expect = [6.0, -6.0 * mu * 2.0, -6.0 * mu * 3.0]  # inward push, friction against the tangential velocity
    np.testing.assert_allclose(f, expect, rtol=2e-6 if prec == "f32" else 1e-12)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_planar_domain_face1_and_corner(oracle_mod, prec):
    # 2D, V = 4e-4 -> r = 0.01; at (0.995, 0.004): 0.005 from x1, 0.004 from y0
    f = _planar_sim(oracle_mod, prec, 2, [0.995, 0.004], [0.5, -0.25], 4e-4, 1.0, 0.01,
                    [0.0, 10.0], [20.0, 0.0])[0]
    k = 1.0 / 0.01 ** 2
    o_x1, o_y0 = 0.01 - 0.005, 0.01 - 0.004
    mu_y0, mu_x1 = math.radians(10.0), math.radians(20.0)
    fy0 = k * o_y0 * np.array([-mu_y0 * 0.5, 1.0])   # floor: up, friction against +x motion
    fx1 = k * o_x1 * np.array([-1.0, -mu_x1 * -0.25])  # right wall: push -x, friction against -y motion
    np.testing.assert_allclose(f, fy0 + fx1, rtol=5e-5 if prec == "f32" else 1e-9)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_planar_domain_no_overlap_no_force(oracle_mod, dim):
    vol = 1e-3 ** dim  # r = 5e-4
    f = _planar_sim(oracle_mod, "f64", dim, [0.5] * dim, [1.0] * dim, vol, 3.0, 0.1, [45.0] * dim, [45.0] * dim)
    np.testing.assert_array_equal(f, 0.0)
    # just inside the radius of the x = 0 wall (1D: a bar particle)
    f = _planar_sim(oracle_mod, "f64", dim, [4e-4] + [0.5] * (dim - 1), [0.0] * dim, vol, 3.0, 0.1, [0.0] * dim,
                    [0.0] * dim)
    assert f[0, 0] == pytest.approx(3.0 / 0.01 * 1e-4, rel=1e-9)
    assert np.all(f[0, 1:] == 0.0)
