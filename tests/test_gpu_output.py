"""GPU tests of the output writers (particles.cpp:209-320, nodes.cpp:214-244,
output.cpp:132-190): Solver.run(N, freq) writes particles<step>.csv/.vtp at
the reference's steps, the CSV has the reference's columns ("Velocity:0",
"Mass", ...) and values equal to the solver's getters, the .vtp arrays decode
to the same values, and the reference's own acceptance post-processing of
those CSVs (test_two_disks.py:137-157: KE0 = 2.51 +- 10%, |dKE|/KE0 < 0.2)
holds on them."""
import base64
import struct
import sys
import xml.etree.ElementTree as ET

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


def _read_csv(path):
    import pandas as pd
    return pd.read_csv(path, delimiter=",")


def _vtp_arrays(path):
    """{name: (ncomp, values)} of the point data, plus 'Points' and 'Time'."""
    root = ET.parse(path).getroot()
    assert root.get("type") == "PolyData" and root.get("header_type") == "UInt64"
    out = {}
    dtypes = {"Float64": "<f8", "Float32": "<f4", "Int64": "<i8"}
    for da in root.iter("DataArray"):
        raw = base64.b64decode(da.text or "")
        nbytes = struct.unpack("<Q", raw[:8])[0]
        vals = np.frombuffer(raw[8:8 + nbytes], dtype=dtypes[da.get("type")])
        out[da.get("Name")] = (int(da.get("NumberOfComponents", "1")), vals)
    return out


def test_run_writes_reference_csv_and_vtp(tmp_path):
    spec = S.two_disc()
    pm, solver = S.build_product(spec)
    out = str(tmp_path / "out")
    pm.set_global_output_directory(out)
    solver.particles.set_output_formats(["csv", "vtk"])
    solver.nodes.set_output_formats(["csv"])
    _banner("output: two-disc run(300, 100) with csv + vtk")
    solver.run(300, 100)
    files = sorted(p.name for p in (tmp_path / "out").iterdir())
    for step in (0, 100, 200, 300):
        assert f"particles{step}.csv" in files and f"particles{step}.vtp" in files, files
    assert "nodes0.csv" in files and "nodes100.csv" not in files  # nodes.cpp:221-224
    p = solver.particles
    pos, vel = np.asarray(p.positions), np.asarray(p.velocities)
    st, F, m = np.asarray(p.stresses), np.asarray(p.F), np.asarray(p.masses)
    df = _read_csv(tmp_path / "out" / "particles300.csv")
    cols = list(df.columns)
    assert cols[:5] == ["isRigid", "Positions:0", "Positions:1", "Velocity:0", "Velocity:1"], cols
    assert cols[-5:] == ["Pressure", "q", "|q|/p", "dev strain", "vol strain"], cols
    np.testing.assert_allclose(df[["Velocity:0", "Velocity:1"]].to_numpy(), vel, rtol=1e-7, atol=1e-12)
    np.testing.assert_allclose(df[["Positions:0", "Positions:1"]].to_numpy(), pos, rtol=1e-7, atol=1e-12)
    np.testing.assert_allclose(df["Mass"].to_numpy(), m, rtol=1e-7)
    # Stress in Eigen's column-major order (symmetric), DeformationMatrix too
    np.testing.assert_allclose(df[[f"Stress:{k}" for k in range(9)]].to_numpy(),
                               st.transpose(0, 2, 1).reshape(-1, 9), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(df[[f"DeformationMatrix:{k}" for k in range(4)]].to_numpy(),
                               F.transpose(0, 2, 1).reshape(-1, 4), rtol=1e-7, atol=1e-9)
    # post-processed pressure and von Mises q (particles.cpp:243-256)
    pres = (st[:, 0, 0] + st[:, 1, 1]) / 3.0
    s = st - pres[:, None, None] * np.eye(3)
    q = np.sqrt(1.5 * np.sum(s * s, axis=(1, 2)))
    np.testing.assert_allclose(df["Pressure"].to_numpy(), pres, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(df["q"].to_numpy(), q, rtol=1e-5, atol=1e-6)
    # the .vtp of the same step decodes to the same arrays
    arr = _vtp_arrays(tmp_path / "out" / "particles300.vtp")
    nc, v = arr["Velocity"]
    assert nc == 2
    np.testing.assert_allclose(v.reshape(-1, 2), vel, rtol=1e-7, atol=1e-12)
    pts = arr["Points"][1].reshape(-1, 3)
    np.testing.assert_allclose(pts[:, :2], pos, rtol=0, atol=0)
    assert np.all(pts[:, 2] == 0)
    t = arr["Time"][1][0]
    assert abs(t - 300 * spec.dt) < 1e-6, t  # step * dt, dt held in FP32
    # the reference's acceptance post-processing on the written CSVs
    KE = []
    for step in (0, 100, 200, 300):
        d = _read_csv(tmp_path / "out" / f"particles{step}.csv")
        KE.append(float((0.5 * d["Mass"] * (d["Velocity:0"] ** 2 + d["Velocity:1"] ** 2)).sum()))
    assert abs(KE[0] - 2.51) <= 0.1 * 2.51, KE
    assert abs((KE[-1] - KE[0]) / KE[0]) < 0.2, KE


def test_rigid_particles_excluded_and_unknown_format_rejected(tmp_path):
    disc = S.create_circle([0.5, 0.45], 0.2, 0.05, 4)
    rod = np.stack([np.linspace(0.3037, 0.6961, 17), np.full(17, 0.6913)], axis=1)
    spec = S.Spec(dim=2, origin=[0.0, 0.0], end=[1.0, 1.0], cell_size=0.05, shape_function="linear", dt=1e-3, ppc=4,
                  positions=disc, rigid_positions=rod,
                  materials=[("LinearElastic", {"density": 1000.0, "E": 1000.0, "pois": 0.3})],
                  bcs=[("RigidBodyLevelSet", {"mode": 0})])
    pm, solver = S.build_product(spec)
    pm.set_global_output_directory(str(tmp_path))
    solver.particles.set_output_formats(["csv"])
    solver.particles.output_vtk()
    df = _read_csv(tmp_path / "particles0.csv")
    assert len(df) == len(disc) and (df["isRigid"] == 0).all()  # particles.h:248
    solver.particles.set_output_formats(["obj"])
    with pytest.raises(ValueError, match="obj"):
        solver.particles.output_vtk()
