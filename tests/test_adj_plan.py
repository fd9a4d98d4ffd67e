"""The binomial checkpoint schedule's cost (bgk_adj_plan, host arithmetic of
the library) against a brute-force dynamic programme.

Model: reversing l steps from a held state with s free snapshot slots costs
T(l, s) = min over the first split m of m + T(l - m, s - 1) + T(m, s)
forward steps (T(1, s) = 0, T(l, 0) = l (l - 1)/2).  The library stores the
states of the schedule's first branch during the recorded forward run, so
its backward pass recomputes T(N, S - 1) minus the last stored step; the
test rebuilds that branch from the programme (first minimal split) and
checks the library's count, which pins both its closed form of T and its
schedule.  Runs without a GPU."""
from functools import lru_cache

import pytest

from bgk_pcm_amd.lattice import adj_plan


@lru_cache(None)
def T(l, s):
    if l <= 1:
        return 0
    if s == 0:
        return l * (l - 1) // 2
    return min(m + T(l - m, s - 1) + T(m, s) for m in range(1, l))


def split(l, s):
    return min(range(1, l), key=lambda m: (m + T(l - m, s - 1) + T(m, s), m))


def chain_end(n, S):
    a, free = 0, S - 1
    while n - a >= 2 and free >= 1:
        a += split(n - a, free)
        free -= 1
    return a


@pytest.mark.parametrize("S", [1, 2, 3, 4, 6])
def test_plan_is_the_optimal_schedule_minus_the_stored_branch(S):
    for n in list(range(1, 40)) + [57, 90]:
        want = T(n, S - 1) - chain_end(n, S)
        assert adj_plan(n, S) == want, (n, S, adj_plan(n, S), want)


def test_plan_of_the_large_adjoint_configuration():
    """5000 steps: two checkpoint levels (289 / 17) recompute every step
    twice; a binomial schedule with 88 snapshots about 1.2 times."""
    r = adj_plan(5000, 88)
    assert 1.0 * 5000 < r < 1.3 * 5000, r
    assert adj_plan(5000, 5000) == 0  # every state stored: nothing to recompute
