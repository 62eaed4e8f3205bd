"""Pins of the oracle's space model (oracle/spaces.cpp): the order in which
ODE's sweep-and-prune, simple and hash spaces call the near callback.

Each quirk of the reference has a hand-derived case whose expected sequence
is written out literally, chosen so that the obvious simplification (a plain
stable sort, a fresh sort every call, an insertion-ordered list, list order
instead of hash-chain order, list order instead of the level rule) fails it:

* the OPCODE radix sort (collision_sapspace.cpp:654-894): negative keys that
  tie come out in reverse order; the previous call's ranks are kept when the
  keys read in that order are still non-decreasing; a change of count
  invalidates them (RaixSortContext::ReallocateRanksIfNecessary :111-125);
* GeomList / DirtyList (collision_sapspace.cpp:300-421): a moved geom is
  swap-removed and appended at the next collide, which reorders the
  infinite x finite pairs (:517-532);
* the simple and hash spaces put new and moved geoms at the front of their
  list (collision_space.cpp:205-209, collision_kernel.h:189-198);
* the hash space (collision_space.cpp:421-614) emits a pair from the
  lower-level geom, walks each hash chain newest node first, and checks the
  big boxes after all hashed pairs.

Random inputs are then checked against the textbook definitions: a sort by
x-min and a forward scan while x-min <= x-max (sweep and prune), the n^2 loop
over the list (simple), and the closed-AABB-overlap set (hash).
"""
import numpy as np
import pytest

import oracle as O

XYZ = (0) | (1 << 2) | (2 << 4)
ZYX = (2) | (1 << 2) | (0 << 4)
INF = np.float32(np.inf)


def box(x0, x1, y0=0.0, y1=1.0, z0=0.0, z1=1.0):
    return [x0, x1, y0, y1, z0, z1]


def zplane():
    # the AABB of the plane z = 0 (plane.cpp:81-105): x, y unbounded, z <= 0
    return [-INF, INF, -INF, INF, -INF, 0.0]


def run(space, aabbs, body=None, **kw):
    a = np.array(aabbs, np.float32)
    body = np.arange(len(a), dtype=np.int32) if body is None else np.array(body, np.int32)
    return space.collide(a, body, **kw)


# ------------------------------------------------------------- radix sort
def test_radix_sort_distinct_keys_is_argsort():
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 100, 1000):
        keys = rng.permutation(np.linspace(-50, 50, n).astype(np.float32))
        r = O.RadixModel()
        assert r.sort(keys) == [int(i) for i in np.argsort(keys, kind="stable")]


def test_radix_sort_reverses_tied_negative_keys():
    # the sign pass fills negative buckets from their end: tied negatives come
    # out in reverse index order, tied positives in index order
    r = O.RadixModel()
    assert r.sort([1.0, -2.0, 1.0, -2.0, 0.5]) == [3, 1, 4, 0, 2]
    assert r.valid()


def test_radix_sort_keeps_previous_ranks_while_they_stay_ordered():
    r = O.RadixModel()
    assert r.sort([2.0, 1.0]) == [1, 0]
    # read in the previous order the keys (1, 1) do not descend: kept as [1, 0]
    # although a fresh sort of tied positives would give [0, 1]
    assert r.sort([1.0, 1.0]) == [1, 0]
    # -0.0 compares equal to +0.0: an index-ordered input is returned as is,
    # and the ranks stay invalid
    r2 = O.RadixModel()
    assert r2.sort([0.0, -0.0]) == [0, 1]
    assert not r2.valid()
    assert r2.sort([1.0, -1.0]) == [1, 0]
    assert r2.valid()


def test_radix_sort_count_change_invalidates_ranks():
    r = O.RadixModel()
    assert r.sort([2.0, 1.0, 3.0]) == [1, 0, 2]
    # two keys now: the ranks are invalid, index order is checked and kept
    assert r.sort([5.0, 5.0]) == [0, 1]
    assert not r.valid()
    assert r.sort([5.0, 4.0]) == [1, 0]


# -------------------------------------------------------- sweep and prune
def test_sap_first_call_and_moved_geom_reorders_infinite_pairs():
    s = O.SpaceModel("sap", XYZ)
    for g in range(4):
        s.add(g)
    ab = [box(0.0, 1.0), box(0.5, 1.5), box(1.2, 2.0), zplane()]
    ab[3][4] = -INF
    # plane 3 is infinite on x: it pairs, untested, with every finite geom in
    # GeomList order after the pruned pairs
    assert run(s, ab) == [(0, 1), (1, 2), (3, 0), (3, 1), (3, 2)]
    # geom 0 moves: swap-removed (the last geom, 3, takes its slot) and
    # appended at the next collide: GeomList 3, 1, 2, 0
    s.moved(0)
    assert s.order() == [0, 3, 1, 2]
    assert run(s, ab) == [(0, 1), (1, 2), (3, 1), (3, 2), (3, 0)]
    # a moved geom is dirty until the next collide: moving it again changes nothing
    s.moved(1)
    s.moved(1)
    s.moved(2)
    assert s.order() == [1, 2, 3, 0]
    assert run(s, ab) == [(0, 1), (1, 2), (3, 0), (3, 1), (3, 2)]


def test_sap_tied_negative_xmin_pairs_in_radix_order():
    s = O.SpaceModel("sap", XYZ)
    for g in range(3):
        s.add(g)
    ab = [box(-1.0, 0.0), box(-1.0, 0.5), box(-3.0, -0.5)]
    # sorted: 2 (-3), then the tied -1 keys reversed: 1, 0
    assert run(s, ab) == [(2, 1), (2, 0), (1, 0)]


def test_sap_axis_order_prunes_on_the_first_axis():
    s = O.SpaceModel("sap", ZYX)
    for g in range(3):
        s.add(g)
    # z is the sorting axis: z-min 2, 0, 1 -> order 1, 2, 0
    ab = [box(0, 1, 0, 1, 2.0, 3.0), box(0, 1, 0, 1, 0.0, 2.5), box(0, 1, 0, 1, 1.0, 2.2)]
    assert run(s, ab) == [(1, 2), (1, 0), (2, 0)]


def test_sap_filters_same_body_and_category_bits():
    s = O.SpaceModel("sap", XYZ)
    for g in range(4):
        s.add(g)
    ab = [box(0, 1), box(0.1, 1), box(0.2, 1), box(0.3, 1)]
    body = [0, 0, 1, -1]
    cat = np.array([1, 1, 2, 4], np.uint64)
    col = np.array([~np.uint64(0), ~np.uint64(0), 1, 2], np.uint64)
    got = s.collide(np.array(ab, np.float32), np.array(body, np.int32), cat=cat, col=col)
    # (0,1) same body; (2,3): cat2&col3 = 2&2 -> kept; (0,3): 1&2=0, 4&all -> kept
    assert got == [(0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def test_sap_random_matches_sort_and_scan_definition():
    rng = np.random.default_rng(7)
    for trial in range(20):
        n = int(rng.integers(2, 60))
        lo = rng.uniform(-5, 5, size=(n, 3)).astype(np.float32)
        ext = rng.uniform(0.1, 2.0, size=(n, 3)).astype(np.float32)
        ab = np.zeros((n + 2, 6), np.float32)
        ab[:n, 0::2] = lo
        ab[:n, 1::2] = lo + ext
        ab[n] = zplane()
        ab[n + 1] = zplane()
        body = np.arange(n + 2, dtype=np.int32)
        body[n:] = -1
        s = O.SpaceModel("sap", XYZ)
        for g in range(n + 2):
            s.add(g)
        got = s.collide(ab, body)
        # the definition: finite geoms sorted by x-min (distinct keys here),
        # each paired with the later ones while their x-min <= its x-max and
        # y, z overlap; then infinite x infinite and infinite x finite
        fin = list(range(n))
        srt = sorted(fin, key=lambda i: ab[i, 0])
        want = []
        for k, a in enumerate(srt):
            for b in srt[k + 1:]:
                if not ab[b, 0] <= ab[a, 1]:
                    break
                if ab[a, 3] >= ab[b, 2] and ab[b, 3] >= ab[a, 2] and ab[a, 5] >= ab[b, 4] and ab[b, 5] >= ab[a, 4]:
                    want.append((a, b))
        want.append((n, n + 1))
        want += [(n, i) for i in fin] + [(n + 1, i) for i in fin]
        assert got == want, trial


# ------------------------------------------------------------------ simple
def test_simple_space_front_insertion_and_move_to_front():
    s = O.SpaceModel("simple")
    for g in range(3):
        s.add(g)
    ab = [box(0, 1)] * 3
    assert s.order() == [2, 1, 0]
    assert run(s, ab) == [(2, 1), (2, 0), (1, 0)]
    s.moved(0)
    assert run(s, ab) == [(0, 2), (0, 1), (2, 1)]
    s.moved(1)
    s.moved(2)
    s.moved(1)  # already dirty: stays where it is
    assert s.order() == [2, 1, 0]
    assert run(s, ab) == [(2, 1), (2, 0), (1, 0)]
    s.remove(1)
    assert run(s, ab) == [(2, 0)]


def test_simple_space_random_matches_n2_loop_over_the_list():
    rng = np.random.default_rng(3)
    n = 40
    lo = rng.uniform(-3, 3, size=(n, 3)).astype(np.float32)
    ab = np.zeros((n, 6), np.float32)
    ab[:, 0::2] = lo
    ab[:, 1::2] = lo + 1.5
    s = O.SpaceModel("simple")
    for g in range(n):
        s.add(g)
    for g in rng.permutation(n)[:15]:
        s.collide(ab, np.arange(n, dtype=np.int32))  # clean, then move one
        s.moved(int(g))
    order = s.order()
    got = s.collide(ab, np.arange(n, dtype=np.int32))
    want = []
    for i, a in enumerate(order):
        for b in order[i + 1:]:
            x, y = ab[a], ab[b]
            if not (x[0] > y[1] or x[1] < y[0] or x[2] > y[3] or x[3] < y[2] or x[4] > y[5] or x[5] < y[4]):
                want.append((a, b))
    assert got == want


# -------------------------------------------------------------------- hash
def test_hash_chain_order_is_newest_node_first():
    s = O.SpaceModel("hash")
    for g in range(3):
        s.add(g)  # list: 2, 1, 0
    e = 0.05
    ab = [box(e, 0.95, e, 0.95, e, 0.95), box(0.5, 1.4, e, 0.95, e, 0.95), box(0.6, 1.5, e, 0.95, e, 0.95)]
    # all level 0 (largest side 0.9); geom 2 searches first and meets, in
    # cell (0,0,0), the chain 0, 1, 2 (0 inserted last)
    assert run(s, ab) == [(2, 0), (2, 1), (1, 0)]


def test_hash_pair_comes_from_the_lower_level_geom():
    s = O.SpaceModel("hash")
    s.add(0)
    s.add(1)  # list: 1, 0
    ab = [box(0.1, 0.3, 0.1, 0.3, 0.1, 0.3), box(0.0, 1.8, 0.0, 1.8, 0.0, 1.8)]
    # 0 is level -2, 1 is level 1: only 0 searches up to level 1
    assert run(s, ab) == [(0, 1)]


def test_hash_big_boxes_after_hashed_pairs():
    s = O.SpaceModel("hash")
    for g in range(3):
        s.add(g)  # list: 2, 1, 0
    ab = [zplane(), box(-0.5, 0.5, -0.5, 0.5, -0.2, 0.8), box(0.0, 1.0, 0.0, 1.0, -0.1, 0.9)]
    assert run(s, ab) == [(2, 1), (2, 0), (1, 0)]


def test_hash_levels_clamp_and_maxlevel_sends_geoms_to_the_big_list():
    # with levels (0, 0) the 1.8 m box (level 1) is a big box, checked after
    # the hashed pairs
    s = O.SpaceModel("hash", levels=(0, 0))
    for g in range(3):
        s.add(g)  # list: 2, 1, 0
    ab = [box(0.0, 1.8, 0.0, 1.8, 0.0, 1.8), box(0.1, 0.3, 0.1, 0.3, 0.1, 0.3), box(0.2, 0.4, 0.2, 0.4, 0.2, 0.4)]
    assert run(s, ab) == [(2, 1), (2, 0), (1, 0)]
    s2 = O.SpaceModel("hash")
    for g in range(3):
        s2.add(g)
    # default levels: 0 is hashed at level 1 and found by 2 and 1 (level -2)
    assert run(s2, ab) == [(2, 1), (2, 0), (1, 0)]


def test_hash_random_set_and_searcher_rule():
    rng = np.random.default_rng(11)
    n = 60
    lo = rng.uniform(-3, 3, size=(n, 3)).astype(np.float32)
    size = np.exp(rng.uniform(np.log(0.05), np.log(3.0), size=(n, 1))).astype(np.float32)
    ab = np.zeros((n + 1, 6), np.float32)
    ab[:n, 0::2] = lo
    ab[:n, 1::2] = lo + size * rng.uniform(0.5, 1.0, size=(n, 3)).astype(np.float32)
    ab[n] = zplane()
    s = O.SpaceModel("hash")
    for g in range(n + 1):
        s.add(g)
    order = s.order()
    pos = {g: i for i, g in enumerate(order)}
    got = s.collide(ab, np.arange(n + 1, dtype=np.int32))
    want = set()
    for a in range(n + 1):
        for b in range(a + 1, n + 1):
            x, y = ab[a], ab[b]
            if not (x[0] > y[1] or x[1] < y[0] or x[2] > y[3] or x[3] < y[2] or x[4] > y[5] or x[5] < y[4]):
                want.add((a, b))
    assert {tuple(sorted(p)) for p in got} == want
    assert len(got) == len(want)

    def level(g):
        q = max(ab[g, 1] - ab[g, 0], ab[g, 3] - ab[g, 2], ab[g, 5] - ab[g, 4])
        return max(int(np.frexp(np.float32(q))[1]), -3) if np.isfinite(q) else 99

    hashed = [p for p in got if p[1] != n]
    for a, b in hashed:
        # the searcher is the lower-level geom, or the earlier one in the list
        assert level(a) < level(b) or (level(a) == level(b) and pos[a] < pos[b])
    # searchers come in list order; the plane's pairs come last, in list order
    assert [pos[a] for a, _ in hashed] == sorted(pos[a] for a, _ in hashed)
    tail = got[len(hashed):]
    assert [b for _, b in tail] == [n] * len(tail)
    assert [pos[a] for a, _ in tail] == sorted(pos[a] for a, _ in tail)
