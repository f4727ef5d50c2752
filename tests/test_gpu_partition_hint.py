"""GPU parity of the partition from a sorted hint (the previous call's sorted
keys in hash_sorted) against the oracle's stable sort by key
(spatialpartition.cpp:113-117, thrust::stable_sort_by_key) and against the
full radix path (a context created with PMPM_PARTITION_RADIX=1).

Bar: bit-exact sorted slots, sorted keys, cell ranges and out-of-grid count
for every hint: the hint may change only the speed, never the result.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from tests.conftest import oracle_sim

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"\n[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _context(n, ncells, radix_only):
    """A context on the radix path, or on the hint path at every particle
    count (PMPM_PARTITION_HINT_MIN=0; by default small counts take the radix
    path)."""
    from pyroclastmpm_amd import capi
    env = {"PMPM_PARTITION_RADIX": "1" if radix_only else "0", "PMPM_PARTITION_HINT_MIN": "0"}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return capi.Context(n, ncells)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def _partition(ctx, g, dim, x, hint):
    """pmpm_partition with hash_sorted holding `hint` on entry."""
    from pyroclastmpm_amd import capi
    n = x.shape[0]
    dev = torch.device("cuda")
    xs = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
    hs = torch.from_numpy(hint.astype(np.uint32).view(np.int32)).to(dev)
    t = {"hash": torch.zeros(n, dtype=torch.int32, device=dev), "hash_sorted": hs,
         "sorted": torch.full((n,), -1, dtype=torch.int32, device=dev)}
    cs = torch.full((g.num_cells_total,), 7, dtype=torch.int32, device=dev)  # stale values to overwrite
    ce = torch.full_like(cs, 7)
    sp = capi.Spatial(t["hash"].data_ptr(), t["hash_sorted"].data_ptr(), t["sorted"].data_ptr(), cs.data_ptr(),
                      ce.data_ptr(), None)
    capi.check(capi.lib().pmpm_partition(ctx.ptr, dim, C.byref(g), xs.data_ptr(), n, C.byref(sp), None))
    torch.cuda.synchronize()
    out = C.c_int32(-1)
    capi.check(capi.lib().pmpm_context_out_of_grid(ctx.ptr, None, C.byref(out)))
    return {"hash": t["hash"].cpu().numpy().astype(np.int64),
            "hash_sorted": t["hash_sorted"].cpu().numpy().view(np.uint32).astype(np.int64),
            "sorted": t["sorted"].cpu().numpy(), "cell_start": cs.cpu().numpy(), "cell_end": ce.cpu().numpy(),
            "out_of_grid": out.value}


def _oracle(oracle_mod, dim, end, h, x, ncells):
    """The oracle's cell keys (its binning, spatialpartition_inline.h:49-83)
    with keys outside [0, ncells) set to ncells (such particles sort last and
    join no cell), then the stable sort by key and the cell ranges."""
    s = oracle_sim(oracle_mod, dim, "f32", [0.0] * dim, end, h, "linear", 0.1, 1)
    s.set_particles(x.astype(np.float64))
    s.partition()
    key = np.asarray(s.get("hash_unsorted"), dtype=np.int64).copy()
    key[(key < 0) | (key >= ncells)] = ncells
    order = np.argsort(key, kind="stable")
    ks = key[order]
    cs = np.zeros(ncells, np.int64)
    ce = np.zeros(ncells, np.int64)
    inside = ks < ncells
    first = np.ones(len(ks), bool)
    first[1:] = ks[1:] != ks[:-1]
    last = np.ones(len(ks), bool)
    last[:-1] = ks[1:] != ks[:-1]
    j = np.arange(len(ks))
    cs[ks[first & inside]] = j[first & inside]
    ce[ks[last & inside]] = j[last & inside] + 1
    if not np.any(key >= ncells):  # in-grid scenes: the oracle's own sort and ranges
        np.testing.assert_array_equal(order, s.get("sorted_index"))
        np.testing.assert_array_equal(cs, s.get("cell_start"))
        np.testing.assert_array_equal(ce, s.get("cell_end"))
    return {"hash": key, "hash_sorted": ks, "sorted": order, "cell_start": cs, "cell_end": ce,
            "out_of_grid": int(np.sum(~inside))}


def _check(got, want, what):
    for k in ["hash", "hash_sorted", "sorted", "cell_start", "cell_end"]:
        np.testing.assert_array_equal(got[k], want[k], err_msg=f"{what}: {k}")
    assert got["out_of_grid"] == want["out_of_grid"], what


def _scene(rng, dim, n, h, outside):
    end = [1.0] * dim
    x = rng.uniform(0.0, 0.999, size=(n, dim)).astype(np.float32)
    x[: n // 10] = (np.floor(x[: n // 10] / h) * h).astype(np.float32)  # on cell faces
    if outside:
        k = rng.choice(n, size=outside, replace=False)
        # past the end along the slowest axis: the linear cell index leaves
        # the grid (the only bound the binning checks)
        x[k, dim - 1] = rng.uniform(1.1, 1.5, size=outside).astype(np.float32)
    return x, end


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("hint", ["zeros", "garbage", "sorted_unrelated", "previous_step"])
def test_partition_any_hint_bit_exact(oracle_mod, dim, hint):
    """Every hint gives the oracle's stable sort, with out-of-grid particles
    and a particle count that is no multiple of the workgroup size."""
    from pyroclastmpm_amd import capi
    rng = np.random.default_rng(11 + dim)
    n = {2: 50001, 3: 150003}[dim]
    h = {2: 0.02, 3: 0.025}[dim]
    x, end = _scene(rng, dim, n, h, outside=n // 50)
    g = capi.make_grid([0, 0, 0], end + [0] * (3 - dim), h, dim)
    ncells = g.num_cells_total
    want = _oracle(oracle_mod, dim, end, h, x, ncells)
    if hint == "zeros":
        hv = np.zeros(n, np.int64)
    elif hint == "garbage":
        hv = rng.integers(0, 2 ** 32 - 1, size=n)
    elif hint == "sorted_unrelated":
        hv = np.sort(rng.integers(0, ncells + 1, size=n))
    else:
        # sorted keys of a slightly different scene (3% of the keys changed);
        # the particles themselves are in random order here
        hv = want["hash"].copy()
        moved = rng.choice(n, size=n // 33, replace=False)
        hv[moved] = np.clip(hv[moved] + rng.choice([-1, 1, -ncells // 7], size=moved.size), 0, ncells)
        hv = np.sort(hv)
    _banner(f"hint partition dim={dim} n={n} hint={hint}")
    ctx = _context(n, ncells, radix_only=False)
    _check(_partition(ctx, g, dim, x, hv), want, f"hint {hint}")
    ctx_r = _context(n, ncells, radix_only=True)
    _check(_partition(ctx_r, g, dim, x, hv), want, f"radix path, hint {hint}")


@pytest.mark.parametrize("moving", [0.02, 0.06])
def test_partition_hint_stepping(oracle_mod, moving):
    """Steps as pmpm_step drives it: the particles are reordered into the
    last partition's cell order, a few move, and the next partition reads the
    last sorted keys as its hint. Most of the moving particles change cell:
    some 3,500 movers (sorted in one workgroup) or some 10,000 (the
    multi-block radix passes)."""
    from pyroclastmpm_amd import capi
    rng = np.random.default_rng(5)
    dim, n, h = 3, 200000, 0.02
    x, end = _scene(rng, dim, n, h, outside=0)
    g = capi.make_grid([0, 0, 0], end + [0] * (3 - dim), h, dim)
    ctx = _context(n, g.num_cells_total, radix_only=False)
    hint = rng.integers(0, 2 ** 32 - 1, size=n)  # first call: whatever the buffer holds
    for step in range(4):
        _banner(f"hint stepping ({moving:.0%} moving): step {step}")
        got = _partition(ctx, g, dim, x, hint)
        _check(got, _oracle(oracle_mod, dim, end, h, x, g.num_cells_total), f"step {step}")
        # reorder into cell order, then move some particles by up to a cell
        x = x[got["sorted"]]
        hint = got["hash_sorted"]
        k = rng.choice(n, size=int(n * moving), replace=False)
        x[k] = np.clip(x[k] + rng.uniform(-h, h, size=(k.size, dim)).astype(np.float32), 0.0, 0.999)


def test_partition_hint_big_outside_bucket(oracle_mod):
    """Most particles outside the grid (one very large bucket) with an
    unsorted hint: every particle is a mover."""
    from pyroclastmpm_amd import capi
    rng = np.random.default_rng(9)
    dim, n, h = 2, 40000, 0.05
    x, end = _scene(rng, dim, n, h, outside=30000)
    g = capi.make_grid([0, 0, 0], end + [0] * (3 - dim), h, dim)
    ctx = _context(n, g.num_cells_total, radix_only=False)
    _banner("hint partition: 30000 of 40000 outside, unsorted hint")
    _check(_partition(ctx, g, dim, x, rng.integers(0, 2 ** 32 - 1, size=n)), _oracle(oracle_mod, dim, end, h, x, g.num_cells_total), "big outside bucket")
