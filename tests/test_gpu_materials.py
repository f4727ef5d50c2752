"""GPU parity of the materials whose return maps run after G2P (von Mises,
Mohr-Coulomb, modified Cam-Clay; vonmises_inline.h, mohrcoulomb_inline.h,
modifiedcamclay_inline.h) against the CPU oracle on the same inputs, in
scenes where they yield (checked on the oracle's plastic history). Bar:
field-scaled relative error <= 1e-4 against the FP32 oracle after the stated
steps, and the plastic history (accumulated / volumetric plastic strain,
preconsolidation pressure) to the same bar.
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


def _block(dim, sf, material, musl, squeeze=0.4, stress=None, seed=9):
    h = 0.05
    hi = [0.3, 0.3, 0.25][:dim]
    pos = S.box_lattice([0.0] * dim, hi, h / 2, jitter=0.1, seed=seed)
    c = np.asarray(hi) / 2
    vel = -squeeze * (pos - c)  # converging flow: compression and shear
    vel[:, 0] += 0.3 * (pos[:, -1] - c[-1])
    g = [0.0] * dim
    g[-1] = -9.81
    st = None
    if stress is not None:
        st = np.tile(stress, (len(pos), 1, 1))
    return S.Spec(dim=dim, origin=[0.0] * dim, end=[1.0] * dim, cell_size=h, shape_function=sf, dt=1e-4,
                  ppc=2 ** dim, positions=pos, velocities=vel, stresses=st, materials=[material],
                  bcs=[("Gravity", {"gravity": g}),
                       ("NodeDomain", {"face0_mode": [1] * dim, "face1_mode": [1] * dim})], musl=musl)


def _run(oracle_mod, spec, steps, label, hist_keys):
    _banner(f"{label}: n={len(spec.positions)} steps={steps}")
    pm, solver = S.build_product(spec)
    sim = S.build_oracle(oracle_mod, spec, "f32")
    solver.solve_nsteps(steps)
    for _ in range(steps):
        sim.step()
    got, ref = S.product_state(solver), S.oracle_state(sim)
    errs = {k: S.max_rel_err(got[k], ref[k]) for k in S.FIELDS}
    # plastic history, in the reference's particle order
    ph = solver.particles.plastic_history
    for k in hist_keys:
        row = {"alpha": 0, "pc": 1}[k]
        errs[k] = S.max_rel_err(np.asarray(ph[row], dtype=np.float64), np.asarray(sim.get_material(0, k)))
    yielded = int(np.count_nonzero(np.asarray(sim.get_material(0, "alpha"))))
    print(f"[gpu-test] {label} field-scaled errors: {errs}; {yielded} particles with plastic strain", flush=True)
    assert yielded > 0, "the scene must yield"
    for k, e in errs.items():
        assert e <= 1e-4, (label, k, e)


VM = ("VonMises", {"density": 1000.0, "E": 1e6, "pois": 0.3, "yield_stress": 300.0, "H": 2e4})


@pytest.mark.parametrize("sf,musl", [("cubic", False), ("cubic", True), ("quadratic", False)])
def test_von_mises_3d_parity(oracle_mod, sf, musl):
    _run(oracle_mod, _block(3, sf, VM, musl), 15, f"von Mises 3D {sf} {'MUSL' if musl else 'USL'}", ["alpha"])


MC = ("MohrCoulomb", {"density": 1500.0, "E": 1e6, "pois": 0.3, "cohesion": 10.0, "friction_angle": 20.0,
                      "dilatancy_angle": 5.0, "H": 0.0})


@pytest.mark.parametrize("dim,sf,musl", [(2, "cubic", False), (2, "linear", True), (3, "cubic", False),
                                         (3, "cubic", True)])
def test_mohr_coulomb_parity(oracle_mod, dim, sf, musl):
    """a block pulled apart (diverging flow): the tension and shear returns"""
    _run(oracle_mod, _block(dim, sf, MC, musl, squeeze=-1.0), 15,
         f"Mohr-Coulomb {dim}D {sf} {'MUSL' if musl else 'USL'}", ["alpha"])


MCC = ("ModifiedCamClay", {"density": 1800.0, "E": 5e6, "pois": 0.3, "M": 1.0, "lam": 0.2, "kap": 0.05,
                           "Vs": 0.0003, "R": 1.5, "Pt": 0.0, "beta": 1.0})


@pytest.mark.parametrize("dim,sf,musl", [(2, "cubic", False), (3, "cubic", False), (3, "cubic", True),
                                         (2, "quadratic", True)])
def test_modified_cam_clay_parity(oracle_mod, dim, sf, musl):
    st = -2e3 * np.eye(3)  # isotropic preload: pc = R p0
    _run(oracle_mod, _block(dim, sf, MCC, musl, squeeze=2.0, stress=st), 15,
         f"modified Cam-Clay {dim}D {sf} {'MUSL' if musl else 'USL'}", ["alpha", "pc"])
