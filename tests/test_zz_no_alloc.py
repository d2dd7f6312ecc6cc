"""peps_count_prefix_no_alloc against peps_count_prefix: the no-allocation
entry point (every device buffer a fixed slice of the library's static
scratch array) must give the existing entry point's counts, call after call
on the same scratch, and refuse boards whose rows do not fit it."""
import pytest

import _oracle as o
import peps_mi355x as pm
from peps_mi355x.abi import PEPS_E_MEMORY

pytestmark = pytest.mark.gpu


def show(*a):
    print(*a, flush=True)


def test_no_alloc_matches_count_prefix_on_consecutive_calls():
    # Two calls in a row, the second smaller: the scratch keeps the first
    # call's contents, which the second must not see.
    cases = [(14, []), (11, [3]), (14, []), (9, [1, 3]), (17, [0, 5, 9]), (8, [])]
    for n, cols in cases:
        show("no_alloc", n, cols)
        got = pm.count_prefix_no_alloc(n, cols)
        assert got == pm.count_prefix(n, cols), (n, cols)
        assert got == o.count_prefix_dfs(n, cols), (n, cols)


def test_no_alloc_edge_cases_and_formats():
    assert pm.count_prefix_no_alloc(0, []) == 1
    assert pm.count_prefix_no_alloc(1, []) == 1
    assert pm.count_prefix_no_alloc(3, []) == 0
    assert pm.count_prefix_no_alloc(8, [3, 3]) == 0          # attacking prefix
    assert pm.count_prefix_no_alloc(5, [0, 2, 4, 1, 3]) == 1  # complete board
    with pytest.raises(pm.PepsError, match="outside the board"):
        pm.count_prefix_no_alloc(8, [8])
    with pytest.raises(pm.PepsError, match="N <= 21"):
        pm.count_prefix_no_alloc(22, [])
    # lane records, three lanes, forced packed-weight fallback
    for opts in (pm.PepsOptions(record_format=1), pm.PepsOptions(crt_lanes=3),
                 pm.PepsOptions(record_format=0, packed_weight_bits=2)):
        assert pm.count_prefix_no_alloc(13, [], opts) == 73712, opts
        assert pm.count_prefix_no_alloc(12, [2], opts) == pm.count_prefix(12, [2], opts), opts


def test_no_alloc_rejects_rows_beyond_the_scratch():
    with pytest.raises(pm.PepsError, match="no-allocation scratch") as ei:
        pm.count_prefix_no_alloc(17, [])
    assert ei.value.code == PEPS_E_MEMORY
    # the engine is still usable after a refusal
    assert pm.count_prefix_no_alloc(10, []) == 724
    with pytest.raises(pm.PepsError, match="graph_mode"):
        pm.count_prefix_no_alloc(10, [], pm.PepsOptions(graph_mode=1))
