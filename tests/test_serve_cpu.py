"""Host logic of the serving loop: the pool's range allocator and the
seeded request stream (no GPU)."""
import numpy as np

from tcadhip.serve import RangeAllocator, nodes_for_budget, request_stream


def test_range_allocator_first_fit_and_merge():
    A = RangeAllocator(100)
    a = A.alloc(30)
    b = A.alloc(30)
    c = A.alloc(30)
    assert (a, b, c) == (0, 30, 60)
    assert A.alloc(20) is None
    A.release(b, 30)
    assert A.used == 60
    assert A.alloc(40) is None            # the hole is 30 + 10 at the end, not contiguous
    d = A.alloc(25)
    assert d == 30
    A.release(a, 30)
    A.release(c, 30)
    A.release(d, 25)
    assert A.free == [(0, 100)] and A.used == 0


def test_budget_counts_pool_and_workspace():
    per_node = nodes_for_budget(10 ** 9)
    assert 500_000 < per_node < 1_000_000


def test_request_stream_is_seeded_and_bounded():
    r1 = request_stream(12, rate=2.0, seed=4)
    r2 = request_stream(12, rate=2.0, seed=4)
    assert [r.spec.name for r in r1] == [r.spec.name for r in r2]
    assert [r.arrival for r in r1] == [r.arrival for r in r2]
    assert np.all(np.diff([r.arrival for r in r1]) > 0)
    for r in r1:
        assert 5 <= len(r.spec.sweep) <= 40
        assert r.spec.N <= 200_000
        assert r.kind in ("diode", "mosfet", "trigate")
