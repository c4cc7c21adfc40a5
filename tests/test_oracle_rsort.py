"""The oracle's restatement of Rust's unstable sort (oracle/ora_rsort.c), which
decides the order of nodes with equal coordinates in the nested dissection
(src/sparse.rs:108: `set.sort_unstable_by(...)`).

No Rust toolchain exists here, so the restatement is pinned by what the
published algorithm fixes and by mathematics:
  * the result is sorted and a permutation, on inputs full of ties;
  * slices of at most 20 go through insertion sort, so they equal a stable sort;
  * a slice that is already non-descending is left as it is, a strictly
    descending one is reversed (the run detection of ipnsort);
  * the two networks of the small sort sort every 0/1 input (0-1 principle),
    which a wrong or missing comparator pair fails;
  * the product's own restatement (rustyfem_amd/csrc/rust_sort.hpp, written
    separately) gives the same nested-dissection labels on meshes where most
    coordinates tie (tests/test_abi_host.py, tests/test_oracle_symbolic.py).
"""
import itertools

import numpy as np
import pytest


@pytest.fixture(scope="module")
def ors(ora):
    return ora


def _check_sorted_perm(v, key, n):
    assert sorted(v.tolist()) == list(range(n))
    assert np.all(np.diff(key[v]) >= 0)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 17, 18, 20, 21, 31, 32, 33, 63, 64, 65, 100, 511, 513, 4096])
@pytest.mark.parametrize("distinct", [1, 2, 3, 7, 1000])
def test_sorted_permutation_with_ties(ors, n, distinct):
    rng = np.random.default_rng(n * 31 + distinct)
    key = rng.integers(0, distinct, size=n)
    v = ors.rust_sort_unstable(np.arange(n), key)
    _check_sorted_perm(v, key, n)


@pytest.mark.parametrize("pattern", ["sawtooth", "organ", "few_runs", "descending_ties", "plateaus"])
def test_patterns(ors, pattern):
    n = 3000
    i = np.arange(n)
    key = {"sawtooth": i % 17, "organ": np.minimum(i, n - i), "few_runs": (i // 400) % 3,
           "descending_ties": (n - i) // 3, "plateaus": (i // 50) * 7 % 11}[pattern]
    v = ors.rust_sort_unstable(i, key)
    _check_sorted_perm(v, key, n)


def test_short_slices_are_stable(ors):
    rng = np.random.default_rng(5)
    for n in range(2, 21):
        for _ in range(50):
            key = rng.integers(0, 3, size=n)
            v = ors.rust_sort_unstable(np.arange(n), key)
            assert v.tolist() == sorted(range(n), key=lambda x: key[x])


def test_runs_are_kept_or_reversed(ors):
    n = 200
    key = np.repeat(np.arange(20), 10)  # non-descending with ties: left as it is
    items = np.arange(n)
    assert np.array_equal(ors.rust_sort_unstable(items, key), items)
    key2 = np.arange(n)[::-1].copy()  # strictly descending: reversed
    assert np.array_equal(ors.rust_sort_unstable(items, key2), items[::-1])
    # descending with ties is no strictly descending run: it is sorted by the
    # quicksort, and tied items need not come out reversed
    key3 = np.repeat(np.arange(20)[::-1], 10)
    v = ors.rust_sort_unstable(items, key3)
    _check_sorted_perm(v, key3, n)


@pytest.mark.parametrize("width", [9, 13])
def test_networks_sort_every_binary_input(ors, width):
    for bits in itertools.product((0, 1), repeat=width):
        key = np.array(bits, dtype=np.int64)
        v = ors.rust_sort_network(np.arange(width), key)
        assert np.all(np.diff(key[v]) >= 0), bits


def test_networks_are_not_stable(ors):
    # the tie order of a network is its own: the small sort is no stable sort,
    # which is what makes the restated tie order differ from a stable one
    key = np.zeros(13, dtype=np.int64)
    key[0] = 1
    v = ors.rust_sort_network(np.arange(13), key)
    assert v[-1] == 0
    stable = sorted(range(13), key=lambda x: key[x])
    assert v.tolist() != stable
