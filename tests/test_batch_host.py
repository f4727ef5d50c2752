"""CPU checks of the batched form's host logic: the longest-processing-time
placement of scenes on streams (pmpm_batch_assign) and argument validation.
No GPU is touched."""
import ctypes as C
import itertools

import numpy as np
import pytest

from pyroclastmpm_amd import capi


def _assign(weights, lanes):
    w = (C.c_float * len(weights))(*weights)
    out = (C.c_int32 * len(weights))()
    rc = capi.lib().pmpm_batch_assign(w, len(weights), lanes, out)
    assert rc == capi.PMPM_OK
    return list(out)


def _makespan(weights, lane_of, lanes):
    load = [0.0] * lanes
    for w, l in zip(weights, lane_of):
        load[l] += w
    return max(load)


def _optimum(weights, lanes):
    # brute force over all placements (tiny inputs only)
    return min(_makespan(weights, p, lanes) for p in itertools.product(range(lanes), repeat=len(weights)))


def test_lpt_hand_case():
    # decreasing order 5, 4, 3, 2, 2, 1 on two lanes: 5->0, 4->1, 3->1, 2->0,
    # 2->0 (tie at 7: lowest lane), 1->1
    assert _assign([5, 1, 4, 3, 2, 2], 2) == [0, 1, 1, 1, 0, 0]


def test_single_lane_and_more_lanes_than_scenes():
    assert _assign([3.0, 1.0, 2.0], 1) == [0, 0, 0]
    # each scene alone, heaviest first onto the lowest free lane
    assert _assign([3.0, 1.0, 2.0], 4) == [0, 2, 1]


@pytest.mark.parametrize("lanes", [2, 3])
def test_lpt_within_graham_bound(lanes):
    # Graham (1969): LPT makespan <= (4/3 - 1/(3m)) OPT on m identical machines
    rng = np.random.default_rng(5)
    for _ in range(40):
        w = [float(x) for x in rng.integers(1, 50, size=7)]
        lane_of = _assign(w, lanes)
        assert all(0 <= l < lanes for l in lane_of)
        opt = _optimum(w, lanes)
        assert _makespan(w, lane_of, lanes) <= (4.0 / 3.0 - 1.0 / (3 * lanes)) * opt + 1e-9


def test_lpt_tight_instance():
    # the classic worst case for m = 2: 3, 3, 2, 2, 2 -> LPT 7, OPT 6
    w = [3.0, 3.0, 2.0, 2.0, 2.0]
    lane_of = _assign(w, 2)
    assert _makespan(w, lane_of, 2) == 7.0 and _optimum(w, 2) == 6.0


def test_invalid_arguments_rejected():
    L = capi.lib()
    w = (C.c_float * 2)(1.0, 2.0)
    out = (C.c_int32 * 2)()
    assert L.pmpm_batch_assign(w, 2, 0, out) == capi.PMPM_ERR_INVALID
    assert L.pmpm_batch_assign(w, 2, capi.MAX_LANES + 1, out) == capi.PMPM_ERR_INVALID
    assert L.pmpm_batch_assign(None, 2, 2, out) == capi.PMPM_ERR_INVALID
    assert L.pmpm_batch_assign(w, 0, 2, None) == capi.PMPM_OK
    assert L.pmpm_batch_step(None, None, 0, 0, None) == capi.PMPM_ERR_INVALID
    assert L.pmpm_batch_replay(None, 0, None) == capi.PMPM_ERR_INVALID
    assert L.pmpm_batch_destroy(None) == capi.PMPM_OK


def test_scene_struct_matches_header_layout():
    # pmpm_scene: ctx, params, grid, particles[2], spatial, nodes, musl
    s = capi.Scene
    assert s.ctx.offset == 0
    assert s.prm.offset == C.sizeof(C.c_void_p)
    assert s.set.size == 2 * C.sizeof(capi.Particles)
    assert s.musl.offset == s.nodes.offset + C.sizeof(capi.Nodes)


_SCENE_FIELDS = {
    "pmpm_scene": (capi.Scene, ["ctx", "prm", "grid", "set", "part", "nodes", "musl"]),
    "pmpm_adjoint_scene": (capi.AdjointScene, ["ctx", "prm", "grid", "state", "out_bar", "in_bar", "param_bar",
                                               "recompute", "part", "nodes"]),
}


def test_scene_structs_match_header_offsets(tmp_path):
    # offsetof/sizeof of the header's structs, from a C program built against
    # include/pmpm.h, against the ctypes mirrors
    import os
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pmpm.h"', "int main(void) {"]
    for name, (_, fields) in _SCENE_FIELDS.items():
        lines.append(f'  printf("{name} sizeof %zu\\n", sizeof({name}));')
        for f in fields:
            lines.append(f'  printf("{name} {f} %zu\\n", offsetof({name}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for name, (cls, fields) in _SCENE_FIELDS.items():
        assert got[(name, "sizeof")] == C.sizeof(cls), name
        for f in fields:
            assert got[(name, f)] == getattr(cls, f).offset, (name, f)


def test_batch_adjoint_rejects_bad_arguments():
    L = capi.lib()
    assert L.pmpm_batch_adjoint_step(None, None, 0, None) == capi.PMPM_ERR_INVALID
