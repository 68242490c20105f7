"""Host logic of roughpy_amd.streams.Partition (no GPU): the reference's
partition tests (tests/intervals/test_partitions.py:9-24) and known answers
of intervals/src/partition.cpp (constructor filtering, sub-intervals,
refine_midpoints, mesh, insert_intermediate, merge)."""
import pickle

import pytest

from roughpy_amd.streams import Partition, RealInterval


def test_partition_trim_intermediates():
    p = Partition(RealInterval(0.0, 1.0), intermediates=range(10))
    assert p.intermediates() == []


def test_partition_pickle_roundtrip():
    p = Partition(RealInterval(0.5, 1.6), [0.9, 1.3])
    assert pickle.loads(pickle.dumps(p)) == p


def test_subintervals_and_sorting():
    p = Partition(RealInterval(0.0, 1.0), [0.5, 0.25, 1.0, 0.0, -1.0])
    assert p.intermediates() == [0.25, 0.5]
    assert len(p) == 3
    assert [(p[i].inf, p[i].sup) for i in range(3)] == [(0, 0.25), (0.25, 0.5), (0.5, 1.0)]
    with pytest.raises(IndexError):
        p[3]
    q = Partition(RealInterval(2.0, 3.0))
    assert len(q) == 1 and (q[0].inf, q[0].sup) == (2.0, 3.0)


def test_refine_midpoints_and_mesh():
    p = Partition(RealInterval(0.0, 1.0), [0.5])
    r = p.refine_midpoints()
    assert r.intermediates() == [0.25, 0.5, 0.75]
    assert Partition(RealInterval(0.0, 1.0), [0.1, 0.5]).mesh() == pytest.approx(0.1)
    assert Partition(RealInterval(0.0, 2.0)).refine_midpoints().intermediates() == [1.0]


def test_insert_intermediate():
    p = Partition(RealInterval(0.0, 1.0), [0.5])
    p.insert_intermediate(0.25)
    p.insert_intermediate(0.75)
    assert p.intermediates() == [0.25, 0.5, 0.75]
    with pytest.raises(ValueError):
        p.insert_intermediate(1.5)


def test_merge():
    a = Partition(RealInterval(0.0, 1.0), [0.5])
    b = Partition(RealInterval(0.25, 2.0), [0.5, 1.5])
    m = a.merge(b)
    assert (m.inf, m.sup) == (0.0, 2.0)
    assert m.intermediates() == [0.25, 0.5, 1.0, 1.5]
    # disjoint: both end points between them become cuts
    c = Partition(RealInterval(3.0, 4.0))
    m2 = a.merge(c)
    assert (m2.inf, m2.sup) == (0.0, 4.0)
    assert m2.intermediates() == [1.0, 3.0]
    # merging with itself keeps it
    assert a.merge(a) == a


@pytest.mark.parametrize("a,b,want", [
    # touching intervals: the shared end point is one cut
    (((0.0, 1.0), []), ((1.0, 2.0), []), (0.0, 2.0, [1.0])),
    # nested: intermediates of both and the inner interval's ends
    (((0.0, 3.0), [0.5, 2.5]), ((1.0, 2.0), [1.5]), (0.0, 3.0, [0.5, 1.0, 1.5, 2.0, 2.5])),
    # equal infs: no inner inf; the smaller sup coincides with an intermediate
    (((0.0, 2.0), [1.0]), ((0.0, 3.0), [2.0]), (0.0, 3.0, [1.0, 2.0])),
    # an empty interval at the other's sup still cuts there (partition.cpp:186-196)
    (((3.0, 3.0), []), ((1.0, 3.0), []), (1.0, 3.0, [3.0])),
    # a repeated intermediate keeps its multiplicity
    (((0.0, 1.0), [0.5, 0.5]), ((0.0, 1.0), [0.5]), (0.0, 1.0, [0.5, 0.5])),
    # disjoint: only the end points facing the gap (intermediates dropped)
    (((0.0, 1.0), [0.5]), ((3.0, 4.0), [3.5]), (0.0, 4.0, [1.0, 3.0])),
    (((3.0, 4.0), [3.5]), ((0.0, 1.0), [0.5]), (0.0, 4.0, [1.0, 3.0])),
])
def test_merge_cases(a, b, want):
    """Known answers of intervals::Partition::merge
    (intervals/src/partition.cpp:135-250), worked through its loop by hand."""
    p = Partition(RealInterval(*a[0]), a[1])
    q = Partition(RealInterval(*b[0]), b[1])
    m = p.merge(q)
    assert (m.inf, m.sup, m.intermediates()) == want
