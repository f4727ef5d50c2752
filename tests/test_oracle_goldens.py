"""Pin the CPU oracle against the reference's own unit-test goldens.

Every expected number comes from tests/golden/reference_unit_goldens.json,
which transcribes the reference's C++ unit tests (file:line in the fixture).
Both oracle precisions are checked: f32 (the reference's FP32 build) and f64.
"""
import json
import os

import numpy as np
import pytest

from tests.conftest import oracle_sim

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_unit_goldens.json")))
DIMS = [1, 2, 3]
PRECS = ["f32", "f64"]


def _usl_setup(mod, dim, prec, musl=False):
    g = G["usl_stepwise"]
    d = g[str(dim)]
    s = oracle_sim(mod, dim, prec, [0.0] * dim, [1.0] * dim, g["cell_size"], "linear", g["dt"], g["ppc"])
    pos = np.array([d["pos"], d["pos"]])
    s.set_particles(pos, np.ones((2, dim)))
    s.add_material("LinearElastic", {"density": 1000, "E": 0.1, "pois": 0.1})
    s.musl = musl
    s.initialize()
    s.set("masses", np.array(g["masses"]))
    s.set("volumes", np.array(g["volumes"]))
    s.set("volumes_original", np.array(g["volumes"]))
    st = np.zeros((2, 3, 3))
    st[0] = 1.0
    s.set("stresses", st)
    s.reset()
    s.partition()
    s.calculate_shape_function()
    s.p2g()
    return s, d, g["tol"]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dim", DIMS)
def test_usl_stepwise(oracle_mod, dim, prec):
    s, d, tol = _usl_setup(oracle_mod, dim, prec)
    nm = s.get("node_masses")
    mom = s.get("node_moments")
    np.testing.assert_allclose(nm[:2], d["node_mass"], atol=tol)
    np.testing.assert_allclose(mom[:2], d["node_moment"], atol=tol)
    s.g2p(0)
    np.testing.assert_allclose(s.get("volumes")[0], d["volume"], atol=tol)
    np.testing.assert_allclose(s.get("positions")[0], d["position"], atol=tol)
    np.testing.assert_allclose(s.get("velocities")[0], d["velocity"], atol=tol)
    np.testing.assert_allclose(np.diagonal(s.get("F")[0]), d["F_diag"], atol=tol)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dim", DIMS)
def test_musl_double_mapping(oracle_mod, dim, prec):
    s, _, _ = _usl_setup(oracle_mod, dim, prec, musl=True)
    d = G["musl_solve"][str(dim)]
    tol = G["musl_solve"]["tol"]
    s.g2p(1)
    np.testing.assert_allclose(s.get("velocities")[0], d["velocity"], atol=tol)
    np.testing.assert_allclose(s.get("positions")[0], d["position"], atol=tol)
    s.p2g_double_mapping()
    np.testing.assert_allclose(s.get("node_moments_nt")[:2], d["node_moment_nt"], atol=tol)
    s.g2p(2)
    np.testing.assert_allclose(s.get("volumes")[0], d["volume"], atol=tol)
    np.testing.assert_allclose(np.diagonal(s.get("F")[0]), d["F_diag"], atol=tol)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dim", DIMS)
@pytest.mark.parametrize("kind", ["shape_linear", "shape_cubic"])
def test_shape_functions(oracle_mod, dim, prec, kind):
    g = G[kind]
    d = g[str(dim)]
    sf = "linear" if kind == "shape_linear" else "cubic"
    s = oracle_sim(oracle_mod, dim, prec, [0.0] * dim, [1.0] * dim, g["cell_size"], sf, 0.1, 1)
    s.set_particles(np.array([d["pos"]]))
    s.partition()
    s.calculate_shape_function()
    psi = s.get("psi")
    dpsi = s.get("dpsi")
    assert abs(psi.sum() - 1.0) < 1e-4
    for k, v in d["psi"].items():
        assert abs(psi[int(k)] - v) < g["tol"], (k, psi[int(k)], v)
    for k, v in d["dpsi"].items():
        np.testing.assert_allclose(dpsi[int(k)], v, atol=g["tol"])


@pytest.mark.parametrize("dim", DIMS)
def test_spatial_partition(oracle_mod, dim):
    g = G["spatial_partition"]
    d = g[str(dim)]
    s = oracle_sim(oracle_mod, dim, "f32", [0.0] * dim, [1.0] * dim, g["cell_size"], "linear", 0.1, 1)
    s.set_particles(np.array(d["pos"]))
    info = s.grid_info()
    assert info["num_cells"] == d["num_cells"]
    assert info["num_cells_total"] == d["num_cells_total"]
    s.partition()
    np.testing.assert_array_equal(s.get("bins"), d["bins"])
    np.testing.assert_array_equal(s.get("hash_unsorted"), d["hash_unsorted"])
    np.testing.assert_array_equal(s.get("hash_sorted"), d["hash_sorted"])
    np.testing.assert_array_equal(s.get("sorted_index"), d["sorted_index"])
    cs, ce = s.get("cell_start"), s.get("cell_end")
    for c in range(d["num_cells_total"]):
        exp = d["cell_ranges"].get(str(c), [0, 0])
        assert [cs[c], ce[c]] == exp, (c, cs[c], ce[c], exp)


@pytest.mark.parametrize("dim", DIMS)
def test_nodes_constructor(oracle_mod, dim):
    g = G["nodes_constructor"]
    d = g[str(dim)]
    s = oracle_sim(oracle_mod, dim, "f32", [0.0] * dim, [1.0] * dim, g["cell_size"], "linear", 0.1, 1)
    info = s.grid_info()
    assert info["num_cells"] == d["num_cells"] and info["num_cells_total"] == d["num_cells_total"]
    ids = s.get("node_ids")
    n = d["num_cells"]
    for idx in range(d["num_cells_total"]):
        exp = [idx % n[0]]
        if dim > 1:
            exp.append((idx // n[0]) % n[1])
        if dim > 2:
            exp.append(idx // (n[0] * n[1]))
        assert list(ids[idx]) == exp


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dim", DIMS)
def test_particle_volumes_masses(oracle_mod, dim, prec):
    g = G["particles_volumes_masses"]
    d = g[str(dim)]
    s = oracle_sim(oracle_mod, dim, prec, [0.0] * dim, [1.0] * dim, g["cell_size"], "linear", 0.1, g["ppc"])
    s.set_particles(np.array(d["pos"]))
    s.calculate_initial_volumes()
    np.testing.assert_allclose(s.get("volumes"), d["volume"], atol=g["tol"])
    s.calculate_initial_masses(0, g["density"])
    np.testing.assert_allclose(s.get("masses"), d["mass"], atol=g["tol"])
    s.set("volumes", np.full(len(d["pos"]), 100.0))
    s.calculate_initial_volumes()
    np.testing.assert_allclose(s.get("volumes"), 100.0, atol=g["tol"])


@pytest.mark.parametrize("dim", DIMS)
@pytest.mark.parametrize("ramp", [False, True])
def test_gravity(oracle_mod, dim, ramp):
    g = G["gravity"]
    s = oracle_sim(oracle_mod, dim, "f32", [0.0] * dim, [1.0] * dim, g["cell_size"], "linear", g["dt"], 1)
    s.set_particles(np.full((1, dim), 0.1))
    nm = s.get("node_masses")
    nm[0], nm[1] = 1.0, 2.0
    s.set("node_masses", nm)
    if not ramp:
        s.add_bc("Gravity", {"gravity": [g["no_ramp"]["gravity"]] * dim})
        s.apply_node_bcs()
        np.testing.assert_allclose(s.get("node_moments_nt")[:2, 0], g["no_ramp"]["moment_nt"], atol=g["tol"])
    else:
        r = g["ramp"]
        s.add_bc("Gravity", {"gravity": [0.0] * dim, "is_ramp": 1, "ramp_step": r["ramp_step"],
                             "gravity_end": [r["gravity_end"]] * dim})
        s.global_step = 1
        s.apply_node_bcs()
        np.testing.assert_allclose(s.get("node_moments_nt")[:2, 0], r["after_step1"], atol=g["tol"])
        s.global_step = 2
        s.apply_node_bcs()
        np.testing.assert_allclose(s.get("node_moments_nt")[:2, 0], r["after_step2"], atol=g["tol"])
