"""pmpm_step_noalloc (the step without a context, on the library's file-scope
scratch array) against pmpm_step with a context: every particle, node and
partition array bitwise equal after several steps, on two scenes in a row
(the second smaller, so the scratch holds the first one's leftovers), and a
step too large for the scratch rejected before any launch."""
import ctypes as C
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


class _Scene:
    """A 3D Drucker-Prager column held in device tensors, stepped through the
    C ABI (two particle sets, one partition, one node set)."""

    def __init__(self, h, per_axis, musl):
        from pyroclastmpm_amd import capi

        self.capi = capi
        spec = S.column(h=h, per_axis=per_axis, dt=1e-4, jitter=0.1, seed=4, musl=musl)
        pos = spec.positions.astype(np.float32)
        n = self.n = pos.shape[0]
        dev = torch.device("cuda")
        self.grid = capi.make_grid(spec.origin, spec.end, h, 3)
        nn = self.grid.num_cells_total
        vol = np.float32(np.float32(float(np.float32(h)) ** 3) / np.float32(per_axis ** 3))

        def f32(shape, fill=0.0):
            return torch.full(shape, fill, dtype=torch.float32, device=dev)

        def pset(x):
            d = {"x": x, "v": f32((3, n)), "stress": f32((6, n)), "F": f32((9, n)), "L": f32((9, n)),
                 "mass": f32((n,), float(np.float32(1500.0) * vol)), "volume": f32((n,), float(vol)),
                 "volume0": f32((n,), float(vol)), "color": torch.zeros(n, dtype=torch.uint8, device=dev),
                 "flags": torch.full((n,), capi.PF_ACTIVE, dtype=torch.uint8, device=dev),
                 "id": torch.arange(n, dtype=torch.int32, device=dev), "eps_e": f32((6, n)),
                 "stress_ref": f32((6, n))}
            for a in range(3):
                d["F"][4 * a] = 1.0
            return d

        x0 = torch.from_numpy(np.ascontiguousarray(pos.T)).to(dev)
        self.sets = [pset(x0.clone()), pset(f32((3, n)))]
        self.nd = {"mass": f32((nn,)), "mom": f32((3 * nn,)), "mom_nt": f32((3 * nn,)),
                   "vel": f32((capi.node_vel_stride(3) * nn,))}
        self.sp = {k: torch.zeros(n, dtype=torch.int32, device=dev) for k in ["hash", "hash_sorted", "sorted"]}
        self.sp["cell_start"] = torch.zeros(nn, dtype=torch.int32, device=dev)
        self.sp["cell_end"] = torch.zeros(nn, dtype=torch.int32, device=dev)
        self.nodes = capi.Nodes(*[self.nd[k].data_ptr() for k in ["mass", "mom", "mom_nt", "vel"]])
        self.spatial = capi.Spatial(*[self.sp[k].data_ptr() for k in ["hash", "hash_sorted", "sorted", "cell_start",
                                                                      "cell_end"]], None)
        self.views = [capi.Particles(n, *[d[k].data_ptr() for k in ["x", "v", "stress", "F", "L", "mass", "volume",
                                                                    "volume0", "color", "flags", "id", "eps_e",
                                                                    "stress_ref"]]) for d in self.sets]
        prm = capi.Params()
        prm.dim, prm.shape_fn, prm.dt, prm.alpha = 3, capi.SF["cubic"], 1e-4, 0.99
        prm.small_mass_cutoff, prm.node_types_built, prm.wrap_quirk = 1e-9, 1, 1
        p = spec.materials[0][1]
        E, nu = np.float32(p["E"]), np.float32(p["pois"])
        prm.num_materials = 1
        m = prm.materials[0]
        m.kind, m.flags = capi.MAT_DRUCKER_PRAGER, capi.MATF_UPDATE_HISTORY | capi.MATF_APEX_GUARD
        K = E / (np.float32(3.0) * (np.float32(1.0) - np.float32(2.0) * nu))
        G = E / (np.float32(2.0) * (np.float32(1.0) + nu))
        for i, val in enumerate([K, G, p["M"], p["Pt"], 0.0, 0.0]):
            m.p[i] = float(val)
        prm.num_bcs = 2
        prm.bcs[0].kind = capi.BC_GRAVITY
        prm.bcs[0].f[2] = -9.81
        prm.bcs[1].kind = capi.BC_NODE_DOMAIN
        for a in range(6):
            prm.bcs[1].i[a] = 1
        self.prm, self.musl, self.cur = prm, int(musl), 0

    def step(self, ctx=None):
        L = self.capi.lib()
        a, b = self.views[self.cur], self.views[self.cur ^ 1]
        s = torch.cuda.current_stream().cuda_stream
        if ctx is None:
            rc = L.pmpm_step_noalloc(C.byref(self.prm), C.byref(self.grid), self.musl, C.byref(a), C.byref(b),
                                     C.byref(self.spatial), C.byref(self.nodes), s)
        else:
            rc = L.pmpm_step(ctx.ptr, C.byref(self.prm), C.byref(self.grid), self.musl, C.byref(a), C.byref(b),
                             C.byref(self.spatial), C.byref(self.nodes), s)
        self.capi.check(rc, "step")
        self.cur ^= 1

    def arrays(self):
        torch.cuda.synchronize()
        out = {f"p.{k}": v.cpu() for k, v in self.sets[self.cur].items()}
        out.update({f"n.{k}": v.cpu() for k, v in self.nd.items()})
        out.update({f"s.{k}": v.cpu() for k, v in self.sp.items()})
        return out


@pytest.mark.parametrize("musl", [True, False])
def test_noalloc_step_matches_context_step_twice(musl):
    from pyroclastmpm_amd import capi

    for h, per_axis in [(0.025, 2), (0.05, 2)]:  # the second scene is smaller
        ref, got = _Scene(h, per_axis, musl), _Scene(h, per_axis, musl)
        _banner(f"noalloc vs context step: column n={ref.n}, grid {list(ref.grid.num_cells)}, "
                f"{'MUSL' if musl else 'USL'}, 6 steps")
        ctx = capi.Context(ref.n, ref.grid.num_cells_total)
        for _ in range(6):
            ref.step(ctx)
            got.step(None)
        A, B = ref.arrays(), got.arrays()
        for k in A:
            assert torch.equal(A[k], B[k]), (h, k)
        assert float(A["p.x"].abs().max()) > 0


def test_noalloc_rejects_a_step_larger_than_its_scratch():
    """A grid whose P2G tile partials alone exceed the scratch: rejected on the
    host with PMPM_ERR_WORKSPACE, nothing launched (the node arrays are sized
    for the grid all the same)."""
    from pyroclastmpm_amd import capi

    sc = _Scene(0.05, 2, True)
    g = capi.make_grid([0, 0, 0], [1.0, 1.0, 0.6], 0.01, 3)  # 101 x 101 x 61 nodes
    nn = g.num_cells_total
    dev = torch.device("cuda")
    nd = {"mass": torch.zeros(nn, device=dev), "mom": torch.zeros(3 * nn, device=dev),
          "mom_nt": torch.zeros(3 * nn, device=dev), "vel": torch.zeros(8 * nn, device=dev)}
    nodes = capi.Nodes(*[nd[k].data_ptr() for k in ["mass", "mom", "mom_nt", "vel"]])
    cs, ce = torch.zeros(nn, dtype=torch.int32, device=dev), torch.zeros(nn, dtype=torch.int32, device=dev)
    sp = capi.Spatial(sc.sp["hash"].data_ptr(), sc.sp["hash_sorted"].data_ptr(), sc.sp["sorted"].data_ptr(),
                      cs.data_ptr(), ce.data_ptr(), None)
    rc = capi.lib().pmpm_step_noalloc(C.byref(sc.prm), C.byref(g), 1, C.byref(sc.views[0]), C.byref(sc.views[1]),
                                      C.byref(sp), C.byref(nodes), torch.cuda.current_stream().cuda_stream)
    assert rc == capi.PMPM_ERR_WORKSPACE
    assert b"PMPM_NOALLOC_SCRATCH_BYTES" in capi.lib().pmpm_last_error()
    torch.cuda.synchronize()
