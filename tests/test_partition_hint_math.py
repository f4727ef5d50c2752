"""CPU check of the placement rule the partition-from-a-hint kernels
implement (partition.hip: hint_classify_kernel, hint_scan_kernel,
small_sort_kernel, hint_place_kernel), against numpy's stable argsort, which
is thrust::stable_sort_by_key's order (spatialpartition.cpp:113-117).

With a sorted hint h, the particles whose key k[i] equals h[i] (stayers) are
in (key, slot) order among themselves; the rest (movers) are sorted by
(key, slot) and merged per cell:
    stayer i   -> out_start[c] + (S[i] - stay_before[c]) + movers of c before i
    mover r    -> out_start[c] + (r - mover_before[c])
                  + clamp(S[i] - stay_before[c], 0, stay_count[c])
with S[i] the stayers before slot i. An unsorted hint makes every particle a
mover. The GPU tests (test_gpu_partition_hint.py) hold the kernels to the
same result; this test pins the rule itself, including the hints that must
not change it (zero, garbage, sorted but unrelated)."""
import numpy as np
import pytest


def place_by_hint(k, h, ncells):
    n = len(k)
    hint_sorted = n < 2 or bool(np.all(h[:-1] <= h[1:]))
    stay = (k == h) if hint_sorted else np.zeros(n, bool)
    S = np.concatenate([[0], np.cumsum(stay)])[:n]
    cnt_stay = np.bincount(k[stay], minlength=ncells + 1)
    cnt_move = np.bincount(k[~stay], minlength=ncells + 1)
    out_start = np.concatenate([[0], np.cumsum(cnt_stay + cnt_move)])[:-1]
    stay_before = np.concatenate([[0], np.cumsum(cnt_stay)])[:-1]
    mover_before = out_start - stay_before
    movers = np.nonzero(~stay)[0]                       # compacted in slot order
    q = movers - S[movers]                              # their places in that list
    assert np.array_equal(q, np.arange(len(movers)))
    ms = movers[np.argsort(k[movers], kind="stable")]   # sorted by (key, slot)
    dest = np.empty(n, np.int64)
    c = k[ms]
    r = np.arange(len(ms))
    dest[ms] = out_start[c] + (r - mover_before[c]) + np.clip(S[ms] - stay_before[c], 0, cnt_stay[c])
    for i in np.nonzero(stay)[0]:
        ci = k[i]
        seg = ms[mover_before[ci]:mover_before[ci] + cnt_move[ci]]
        dest[i] = out_start[ci] + (S[i] - stay_before[ci]) + int(np.searchsorted(seg, i))
    order = np.empty(n, np.int64)
    order[dest] = np.arange(n)
    return order


@pytest.mark.parametrize("hint", ["previous", "zeros", "garbage", "sorted_unrelated"])
def test_placement_rule_is_the_stable_sort(hint):
    rng = np.random.default_rng({"previous": 1, "zeros": 2, "garbage": 3, "sorted_unrelated": 4}[hint])
    for _ in range(40):
        ncells = int(rng.choice([1, 7, 300, 5000]))
        n = int(rng.integers(1, 3000))
        prev = np.sort(rng.integers(0, ncells + 1, n))      # last step's sorted keys
        k = prev.copy()                                      # the particles in that order
        moved = rng.integers(0, n, int(rng.integers(0, n // 4 + 1)))
        k[moved] = np.clip(k[moved] + rng.integers(-3, 4, moved.size), 0, ncells)
        k[rng.integers(0, n, n // 20)] = ncells              # some leave the grid
        h = {"previous": prev, "zeros": np.zeros(n, np.int64),
             "garbage": rng.integers(0, 2 ** 32 - 1, n),
             "sorted_unrelated": np.sort(rng.integers(0, ncells + 1, n))}[hint]
        np.testing.assert_array_equal(place_by_hint(k, h, ncells), np.argsort(k, kind="stable"))


def test_a_planted_mistake_fails():
    """Dropping the movers' term from a stayer's place breaks the order."""
    rng = np.random.default_rng(9)
    n, ncells = 500, 40
    prev = np.sort(rng.integers(0, ncells, n))
    k = prev.copy()
    k[rng.integers(0, n, 60)] = rng.integers(0, ncells, 60)
    stay = k == prev
    S = np.concatenate([[0], np.cumsum(stay)])[:n]
    cnt_stay = np.bincount(k[stay], minlength=ncells + 1)
    cnt_move = np.bincount(k[~stay], minlength=ncells + 1)
    out_start = np.concatenate([[0], np.cumsum(cnt_stay + cnt_move)])[:-1]
    stay_before = np.concatenate([[0], np.cumsum(cnt_stay)])[:-1]
    wrong = out_start[k[stay]] + (S[stay] - stay_before[k[stay]])  # without "movers of c before i"
    right = np.argsort(np.argsort(k, kind="stable"))[stay]          # true places of the stayers
    assert not np.array_equal(wrong, right)
