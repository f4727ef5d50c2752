"""GPU checks of the batched form (pmpm_batch_*): a batched step of several
scenes on several streams, eager or replayed from a captured HIP graph, gives
each scene the state that stepping it alone gives.

Bar: positions, velocities and stresses within a field-scaled relative error
of 1e-5 of the scene stepped alone (the only differences come from the
order of the global float atomics of the wrap-around deposits)."""
import ctypes as C
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _banner(msg):
    print(f"\n[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _scenes(seed=7):
    from pyroclastmpm_amd.batch import ColumnScene
    # four small columns of different sizes, cell sizes and materials
    cfg = [(1500, 0.02, 25.0, 8e5), (6000, 0.015, 32.0, 1.5e6), (2500, 0.02, 38.0, 6e5), (9000, 0.018, 28.0, 2e6)]
    return [ColumnScene(n, h, phi, E, dt=1e-4, seed=seed + i) for i, (n, h, phi, E) in enumerate(cfg)]


def _step_alone(sc, cur):
    from pyroclastmpm_amd import capi
    s = sc.scene()
    out = s.set[cur ^ 1]
    part, nodes = s.part, s.nodes
    capi.check(capi.lib().pmpm_step(s.ctx, C.byref(s.prm), C.byref(s.grid), 1, C.byref(s.set[cur]), C.byref(out),
                                    C.byref(part), C.byref(nodes), torch.cuda.current_stream().cuda_stream),
               "pmpm_step")


def _err(a, b):
    scale = max(float(b.abs().max()), 1e-30)
    return float((a - b).abs().max()) / scale


def _compare(batch_scene, ref_scene, cur, label):
    torch.cuda.synchronize()
    for k in ["x", "v", "stress"]:
        a, b = batch_scene.sets[cur][k], ref_scene.sets[cur][k]
        e = _err(a, b)
        assert e <= TOL, (label, k, e)
        # same cell order on both sides
    assert torch.equal(batch_scene.sets[cur]["id"], ref_scene.sets[cur]["id"]), label


def test_batched_eager_steps_match_scenes_alone():
    from pyroclastmpm_amd.batch import Batch
    _banner("batch eager: 4 columns, 3 lanes, 4 steps vs pmpm_step per scene")
    got, ref = _scenes(), _scenes()
    b = Batch(got, lanes=3)
    for k in range(4):
        b.step()
        for sc in ref:
            _step_alone(sc, k & 1)
        for i in range(len(got)):
            _compare(got[i], ref[i], b.cur, f"eager step {k} scene {i}")


def test_graph_replays_match_eager_every_time():
    from pyroclastmpm_amd.batch import Batch
    _banner("batch graph: capture both parities, 4 replays with eager work between")
    got, ref = _scenes(11), _scenes(11)
    b = Batch(got, lanes=2)
    b.step(2)  # sizes every workspace before the capture
    for k in range(2):
        for sc in ref:
            _step_alone(sc, k & 1)
    b.capture()
    for k in range(4):
        b.step()  # replay
        torch.cuda.synchronize()
        # unrelated eager work on the same stream between replays
        junk = torch.ones(1 << 16, device="cuda").cumsum(0)
        torch.cuda.synchronize()
        del junk
        for sc in ref:
            _step_alone(sc, k & 1)
        for i in range(len(got)):
            _compare(got[i], ref[i], b.cur, f"replay {k} scene {i}")


def test_autotune_keeps_state_consistent():
    from pyroclastmpm_amd.batch import Batch
    _banner("batch autotune: 4 columns, up to 2 lanes, 2 steps per candidate")
    got, ref = _scenes(23), _scenes(23)
    b = Batch(got, lanes=2)
    b.step(2)
    t = b.autotune(steps=2)
    assert 1 <= t["lanes"] <= 2 and t["use_graph"] in (0, 1) and t["step_ms"] > 0
    # the autotuner ran 2 steps alone + 2 lane counts x 2 modes x 2 steps
    total = 2 + 2 + 2 * 2 * 2
    for k in range(total):
        for sc in ref:
            _step_alone(sc, k & 1)
    assert b.cur == 0
    for i in range(len(got)):
        _compare(got[i], ref[i], b.cur, f"after autotune scene {i}")
    # and it keeps stepping with the chosen setting
    b.step(2)
    for k in range(2):
        for sc in ref:
            _step_alone(sc, k & 1)
    for i in range(len(got)):
        _compare(got[i], ref[i], b.cur, f"tuned steps scene {i}")
