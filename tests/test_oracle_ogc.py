"""Pins the oracle's OGC detection, BVH, candidate cache, colouring and
prediction to the reference's tests (tests/ogc.rs:100-292, tests/ogc_bvh.rs,
tests/ogc_cache.rs, src/ogc/world.rs:562-633, src/vbd/coloring.rs:45-137,
src/vbd/prediction.rs:56-87)."""
import math

import numpy as np
import pytest

import oracle as O
from clothcraft_amd import scenes


def lcg(seed, scale=1.0, shift=0.0):
    """The tests' LCG: seed = seed*6364136223846793005 + 1; (seed >> 32) / u32::MAX."""
    state = [seed]

    def rnd():
        state[0] = (state[0] * 6364136223846793005 + 1) % (1 << 64)
        return (state[0] >> 32) / 4294967295.0 * scale + shift
    return rnd


def test_folded_surface_preserves_multiple_layers_and_excludes_false_edge_contact():
    x = np.array([[0, 0.1, 0.1], [-1, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    w = O.World(x, [[1, 2, 3], [2, 1, 4]], 1)
    edges = [tuple(e) for e in w.edges()]
    e = edges.index((1, 2))
    assert not w.edge_feasible(x[0], e, x)
    s = w.detect(x, 0.25, radius=0.2)
    assert int((s["kind_ids"][:, 0] == 2).sum()) == 2


def test_shared_feature_contact_is_not_duplicated():
    x = np.array([[0, -0.03, 0.04], [-1, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, -1]], float)
    w = O.World(x, [[1, 2, 3], [2, 1, 4]], 1)
    s = w.detect(x, 0.2, radius=0.1)
    assert len(s["kind_ids"]) == 1 and s["kind_ids"][0, 0] == 1


def test_reverse_vertex_face_and_edge_edge_bounds_protect_face_interiors():
    x = np.array([[-1, -1, 0.1], [1, -1, 0.1], [0, 1, 0.1], [0, 0, 0], [0.01, 0, -1], [0, 0.01, -1]], float)
    w = O.World(x, [[0, 1, 2], [3, 4, 5]], 3)
    s = w.detect(x, 0.2, radius=0.15)
    assert np.all(s["radii"] <= 0.045 + 1e-12)
    for i in range(3):
        p = x[i] - np.array([0, 0, 10.0])
        d = p - s["centers"][i]
        b = s["radii"][i]
        q = s["centers"][i] + d * (b / np.linalg.norm(d))
        assert q[2] >= 0.055 - 1e-12
    x = np.array([[-1, 0, 0.1], [1, 0, 0.1], [-1, -1, 0.1], [0, -1, 0], [0, 1, 0], [1, 1, 0]], float)
    w = O.World(x, [[0, 1, 2], [3, 4, 5]], 6)
    s = w.detect(x, 0.2, radius=0.15)
    assert np.any(s["kind_ids"][:, 0] == 3)
    assert s["radii"][0] <= 0.045 + 1e-12 and s["radii"][1] <= 0.045 + 1e-12
    moved = x.copy()
    for i in range(6):
        p = x[i] + (np.array([0, 0, -10.0]) if i < 3 else np.array([0, 0, 10.0]))
        d = p - s["centers"][i]
        moved[i] = s["centers"][i] + d * (s["radii"][i] / np.linalg.norm(d))
    w.validate_start(moved)


def test_bounds_are_cumulative_until_detection_and_settings_fail_closed():
    x = np.array([[0, 0, 0.1], [-1, -1, 0], [1, -1, 0], [0, 1, 0]], float)
    w = O.World(x, [[1, 2, 3]], 1)
    s = w.detect(x, 0.2, radius=0.15)
    c, b = s["centers"][0], s["radii"][0]

    def trunc(p):
        d = p - c
        L = np.linalg.norm(d)
        return c + d * (b / L) if L > b else p
    p = trunc(x[0] - np.array([0, 0, 0.04]))
    q = trunc(p - np.array([0, 0, 0.04]))
    assert abs(q[2] - 0.055) < 1e-12
    with pytest.raises(O.OracleError):
        w.detect(x, 0.01, radius=0.15)
    with pytest.raises(O.OracleError):
        w.detect(x, 0.2, radius=0.15, relaxation=0.5)


def test_initial_intersection_is_rejected():
    x = np.array([[0, 0, -1], [0, 0, 1], [2, 0, 0], [-1, -1, 0], [1, -1, 0], [0, 1, 0]], float)
    w = O.World(x, [[0, 1, 2], [3, 4, 5]], 3)
    with pytest.raises(O.OracleError) as e:
        w.validate_start(x)
    assert "initial intersection" in e.value.message


def test_query_radius_changes_bounds_without_activating_distant_forces():
    x = np.array([[0, 0, 0.02], [-1, -1, 0], [1, -1, 0], [0, 1, 0]], float)
    w = O.World(x, [[1, 2, 3]], 1)
    s = w.detect(x, 0.03, radius=0.01)
    assert len(s["kind_ids"]) == 1 and s["active_contacts"] == 0
    assert w.contacts_value(x, s, 0.01, 1e5) == 0.0
    assert abs(s["radii"][0] - 0.009) < 1e-12


def test_dormant_contact_activates_inside_the_same_original_trust_region():
    x = np.array([[0, 0, 0.00305], [-1, -1, 0], [1, -1, 0], [0, 1, 0]], float)
    w = O.World(x, [[1, 2, 3]], 1)
    s = w.detect(x, 0.006, radius=0.003)
    assert s["active_contacts"] == 0 and w.contacts_value(x, s, 0.003, 1e5) == 0.0
    p = np.array([0, 0, 0.0025])
    assert np.linalg.norm(p - s["centers"][0]) <= s["radii"][0]
    x[0] = p
    expected = 0.5 * 1e5 * 0.0005 ** 2
    assert abs(w.contacts_value(x, s, 0.003, 1e5) - expected) < 1e-12
    s2 = w.detect(x, 0.006, radius=0.003)
    assert abs(w.contacts_value(x, s, 0.003, 1e5) - w.contacts_value(x, s2, 0.003, 1e5)) < 1e-12


def test_static_edges_before_moving_edges_are_not_missed():
    x = np.array([[-1, 0, 0.1], [1, 0, 0.1], [-1, -1, 0.1], [0, -1, 0], [0, 1, 0], [1, 1, 0]], float)
    a = O.World(x, [[0, 1, 2], [3, 4, 5]], 3).detect(x, 0.2, radius=0.15)
    b = O.World(x, [[3, 4, 5], [0, 1, 2]], 3).detect(x, 0.2, radius=0.15)
    assert np.array_equal(a["radii"], b["radii"])
    assert a["edge_edge_candidates"] == b["edge_edge_candidates"]
    assert len(a["kind_ids"]) == len(b["kind_ids"])


def test_shared_vertex_near_parallel_edges_do_not_create_false_initial_intersections():
    pts = [[0.194131180644, 0.943517088890, 0.091018132865],
           [0.183311969042, 0.970439314842, 0.101929388940],
           [0.197149559855, 0.954595983028, 0.076061360538],
           [0.195560514927, 0.972076416016, 0.068995960057],
           [0.188853010535, 0.980777919292, 0.085684590042]]
    x = np.array(pts, np.float32).astype(np.float64)
    w = O.World(x, [[0, 1, 2], [3, 2, 4]], 5)
    w.validate_start(x)


# ------------------------------------------------------------------- BVH
def _bounds(lo, hi):
    return np.concatenate([lo, hi])


def test_dual_tree_pairs_match_exhaustive_search_without_duplicates():
    """tests/ogc_bvh.rs:4-67."""
    rnd = lcg(53)
    a = []
    for _ in range(129):
        p = np.array([rnd(), rnd(), rnd()])
        a.append(_bounds(np.minimum(p, p + 0.03), np.maximum(p, p + 0.03)))
    b = []
    for _ in range(97):
        p = np.array([rnd(), rnd(), rnd()])
        b.append(_bounds(p, p))
    a, b = np.array(a), np.array(b)
    da = np.array([i % 3 == 1 for i in range(len(a))])
    db = np.array([i % 2 == 1 for i in range(len(b))])
    ta, tb = O.Bvh(a, da), O.Bvh(b, db)

    def d2(u, v):
        return O.bounds_distance_squared(u, v)
    for frame in range(4):
        if frame > 0:
            for i in range(len(a)):
                if da[i]:
                    dlt = np.array([rnd() - 0.5, rnd() - 0.5, rnd() - 0.5])
                    a[i, :3] += dlt
                    a[i, 3:] += dlt
            ta.refit_dynamic(a)
        for radius in (0.0, 0.05, 0.2, 2.0):
            pairs = ta.pairs_with(tb, radius)
            assert len(pairs) == len(set(pairs))
            exp = {(i, j) for i in range(len(a)) for j in range(len(b))
                   if (da[i] or db[j]) and d2(a[i], b[j]) <= radius * radius}
            assert set(pairs) == exp
            sp = ta.self_pairs(radius)
            assert all(i < j for i, j in sp) and len(sp) == len(set(sp))
            exp = {(i, j) for i in range(len(a)) for j in range(i + 1, len(a))
                   if (da[i] or da[j]) and d2(a[i], a[j]) <= radius * radius}
            assert set(sp) == exp


def test_bounds_distance_never_discards_a_close_primitive_pair():
    """tests/ogc_bvh.rs:69-96."""
    rnd = lcg(19, 4.0, -2.0)
    for _ in range(2000):
        p = np.array([[rnd(), rnd(), rnd()] for _ in range(4)])
        q, _, _ = O.closest_triangle(p[3], p[:3])
        vf = O.bounds_distance_squared(_bounds(p[3], p[3]), _bounds(p[:3].min(0), p[:3].max(0)))
        assert vf <= np.sum((p[3] - q) ** 2) + 1e-12
        s, t = O.closest_segments(p[0], p[1], p[2], p[3])
        ee = O.bounds_distance_squared(_bounds(p[:2].min(0), p[:2].max(0)), _bounds(p[2:].min(0), p[2:].max(0)))
        assert ee <= np.sum((p[0] + (p[1] - p[0]) * s - p[2] - (p[3] - p[2]) * t) ** 2) + 1e-12
    a = _bounds(np.zeros(3), np.ones(3))
    assert O.bounds_distance_squared(a, a) == 0.0
    assert O.bounds_distance_squared(a, _bounds(np.ones(3), np.ones(3))) == 0.0
    assert O.bounds_distance_squared(a, _bounds(np.full(3, 2.0), np.full(3, 2.0))) == 3.0


def test_partitioned_bvh_matches_exhaustive_queries_after_large_deformations():
    """tests/ogc_bvh.rs:98-143."""
    rnd = lcg(7)
    boxes = []
    for _ in range(513):
        p = np.array([rnd(), rnd(), rnd()])
        boxes.append(_bounds(p, p + 0.04))
    boxes = np.array(boxes)
    dyn = np.array([i % 3 == 0 for i in range(len(boxes))])
    t = O.Bvh(boxes, dyn)

    def overlaps(q, b):
        return np.all(q[:3] <= b[3:]) and np.all(q[3:] >= b[:3])
    for frame in range(4):
        if frame > 0:
            for i in range(len(boxes)):
                if dyn[i]:
                    dlt = np.array([rnd() - 0.5, rnd() - 0.5, rnd() - 0.5])
                    boxes[i, :3] += dlt
                    boxes[i, 3:] += dlt
            t.refit_dynamic(boxes)
        for _ in range(50):
            p = np.array([rnd(), rnd(), rnd()])
            q = _bounds(p - 0.15, p + 0.15)
            for only in (False, True):
                exp = [i for i in range(len(boxes)) if overlaps(q, boxes[i]) and (not only or dyn[i])]
                assert sorted(t.query(q, only)) == exp


# ----------------------------------------------------------- cache/fresh
def test_temporal_candidates_match_fresh_detection_through_motion_and_radius_changes():
    """tests/ogc_cache.rs:5-88."""
    x = np.array([[-1, -1, 0.06], [1, -1, 0.06], [1, 1, 0.06], [-1, 1, 0.06],
                  [-1.1, -1.1, 0], [1.1, -1.1, 0], [1.1, 1.1, 0], [-1.1, 1.1, 0]], float)
    original = O.World(x, [[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]], 4, [[0, 1]])
    cached = original.clone()
    reused = rebuilt = 0
    for frame in range(50):
        for i in range(4):
            x[i, 0] += 0.3 if frame == 12 else 0.0007 * (i + 1)
            x[i, 2] += 0.0002 * (i + 1)
        rq = 0.21 if frame >= 25 else 0.16
        a = cached.detect(x, rq, radius=0.1)
        b = original.clone().detect(x, rq, radius=0.1)
        reused += not a["broad_phase_rebuilt"]
        rebuilt += a["broad_phase_rebuilt"]
        assert np.array_equal(a["radii"], b["radii"])
        assert a["minimum_distance"] == b["minimum_distance"]
        assert a["vertex_face_candidates"] == b["vertex_face_candidates"]
        assert a["edge_edge_candidates"] == b["edge_edge_candidates"]
        assert np.array_equal(a["kind_ids"], b["kind_ids"]) and np.array_equal(a["eps"], b["eps"])
        assert np.array_equal(a["centers"], x[:4])
    assert reused > 30 and rebuilt >= 3
    x[4, 0] += 0.01
    with pytest.raises(O.OracleError):
        cached.detect(x, 0.21, radius=0.1)


def test_cached_broad_phase_catches_a_pair_entering_the_query_radius():
    x = np.array([[0, 0, 0.065], [-1, -1, 0], [1, -1, 0], [0, 1, 0]], float)
    w = O.World(x, [[1, 2, 3]], 1)
    a = w.detect(x, 0.05, radius=0.02)
    assert a["vertex_face_candidates"] == 0
    x[0, 2] = 0.047
    b = w.detect(x, 0.05, radius=0.02)
    assert not b["broad_phase_rebuilt"]
    assert b["vertex_face_candidates"] == 1
    assert abs(b["radii"][0] - 0.45 * 0.047) < 1e-14
    assert len(b["kind_ids"]) == 1 and b["active_contacts"] == 0


@pytest.mark.parametrize("offset", [(0.0, 0.0, 0.0), (1e6, -2e6, 3e6)])
def test_current_boxes_preserve_exact_contacts_and_balls_during_cache_reuse(offset):
    """src/ogc/world.rs:566-633: pruned == unpruned detection."""
    off = np.array(offset)
    rest, tris = [], []
    for layer in range(3):
        for y in range(5):
            for xx in range(5):
                rest.append(off + np.array([xx * 0.008, y * 0.008, layer * 0.003]))
        for y in range(4):
            for xx in range(4):
                a = layer * 25 + y * 5 + xx
                tris += [[a, a + 1, a + 5], [a + 1, a + 6, a + 5]]
    rest = np.array(rest)
    pruned = O.World(rest, tris, 50, [[0, 25], [1, 26]])
    ref = pruned.clone()
    rej = reuse = rebuilds = 0
    for frame in range(24):
        x = rest.copy()
        for i in range(50):
            a = frame * 0.15 + i * 0.31
            x[i] += np.array([math.sin(a) * 0.0004, math.cos(a) * 0.0003, math.sin(a) * 0.0002])
            if frame >= 18:
                x[i, 0] += 0.004
        rq = 0.006 if frame <= 11 else (0.009 if frame <= 15 else 0.004)
        a = pruned.detect(x, rq, radius=0.003)
        b = ref.detect(x, rq, radius=0.003, prune=False)
        assert np.array_equal(a["radii"], b["radii"])
        assert a["minimum_distance"] == b["minimum_distance"]
        assert a["active_contacts"] == b["active_contacts"]
        assert np.array_equal(a["kind_ids"], b["kind_ids"])
        assert np.array_equal(a["eps"].view(np.uint64), b["eps"].view(np.uint64))
        assert np.array_equal(a["incident"], b["incident"])
        rej += a["aabb_rejections"]
        rebuilds += a["broad_phase_rebuilt"]
        reuse += not a["broad_phase_rebuilt"]
    assert rej > 0 and reuse > 0 and rebuilds > 1


# -------------------------------------------------------------- colouring
def test_incidence_coloring_matches_expanded_graph_and_has_no_conflicts():
    """src/vbd/coloring.rs:52-137, through a solver whose static graph is the
    test's graph: 9 moving vertices in 3 stacked triangles plus a fixed one."""
    x = []
    for i in range(4):
        z = 0.002 * i
        x += [[0, 0, z], [1, 0, z], [0, 1, z]]
    x = np.array(x, float)
    w = O.World(x, [[0, 1, 2], [3, 4, 5], [6, 7, 8], [9, 10, 11]], 9)
    c = w.detect(x, 0.012)
    assert len(c["kind_ids"]) > 0
    graph = [[1], [0, 2], [1], [], [6], [], [4], [], []]
    grab = dict(vertices=(0, 4, 8), weights=(0.2, 0.3, 0.5), target=(0, 0, 0), stiffness=100.0)

    def greedy(full):
        assigned = [None] * 9
        out = []
        for v in range(9):
            col = 0
            while any(assigned[i] == col for i in full[v]):
                col += 1
            if col == len(out):
                out.append([])
            out[col].append(v)
            assigned[v] = col
        return out, assigned

    for g in (None, grab):
        full = [list(r) for r in graph]
        for row in c["kind_ids"]:
            k = int(row[0])
            cnt = [2, 3, 4, 4][k]
            vs = [int(v) for v in row[1:1 + cnt]]
            for a in vs:
                for b in vs:
                    if a != b and a < 9 and b < 9:
                        full[a].append(b)
        if g:
            for a in g["vertices"]:
                for b in g["vertices"]:
                    if a != b:
                        full[a].append(b)
        exp, assigned = greedy(full)
        got = _oracle_colors(graph, c, g)
        assert got == exp
        for v in range(9):
            for o in full[v]:
                assert assigned[v] != assigned[o]
    assert _oracle_colors([[] for _ in range(9)], dict(kind_ids=np.zeros((0, 5), np.uint32)), None) == [list(range(9))]


def _oracle_colors(graph, contacts, grab):
    ki = contacts["kind_ids"]
    counts = np.array([[2, 3, 4, 4][int(r[0])] for r in ki], np.uint32)
    return O.colors(graph, ki, counts, grab)


def _lowbias32(v):
    v &= 0xFFFFFFFF
    v ^= v >> 16
    v = (v * 0x7FEB352D) & 0xFFFFFFFF
    v ^= v >> 15
    v = (v * 0x846CA68B) & 0xFFFFFFFF
    v ^= v >> 16
    return v


def _jones_plassmann(graph, rows, counts, grab, col0):
    """The fast colouring run as the device runs it: synchronous rounds, each
    colouring the vertices whose uncoloured neighbours all come later in the
    priority order (static colour ascending, then hash descending), from the
    colours held at the round's start."""
    n = len(graph)
    nbrs = [set(r) for r in graph]
    for row, k in zip(rows, counts):
        vs = [int(v) for v in row[1:1 + k] if v < n]
        for a in vs:
            nbrs[a].update(b for b in vs if b != a)
    if grab:
        gv = [v for v in grab["vertices"] if v < n]
        for a in gv:
            nbrs[a].update(b for b in gv if b != a)
    key = [(-col0[v], _lowbias32(v)) for v in range(n)]
    col = [None] * n
    rounds = 0
    while any(c is None for c in col):
        rounds += 1
        start = list(col)
        for v in range(n):
            if start[v] is not None:
                continue
            if any(start[u] is None and key[u] > key[v] for u in nbrs[v]):
                continue
            used = {start[u] for u in nbrs[v] if start[u] is not None}
            c = 0
            while c in used:
                c += 1
            col[v] = c
    return col, rounds


def test_fast_colouring_equals_jones_plassmann_rounds_and_is_proper():
    """Extension (CV_FLAG_FAST_COLORING): the oracle's sequential statement
    (greedy in the order of the static smallest-last colour classes, ties by
    decreasing hash) equals synchronous Jones-Plassmann rounds simulated
    independently here, on random graphs with random contact cliques and a
    grab; the colouring is proper, and without contacts it is the
    smallest-last greedy colouring itself."""
    rng = np.random.default_rng(11)
    for trial in range(12):
        n = int(rng.integers(20, 90))
        graph = [set() for _ in range(n)]
        for _ in range(int(rng.integers(n, 3 * n))):
            a, b = (int(v) for v in rng.integers(0, n, 2))
            if a != b:
                graph[a].add(b)
                graph[b].add(a)
        graph = [sorted(r) for r in graph]
        m = int(rng.integers(1, 4 * n))
        rows = np.zeros((m, 5), np.uint32)
        counts = np.zeros(m, np.uint32)
        for i in range(m):
            k = int(rng.integers(2, 5))
            rows[i, 0] = k - 2 if k < 4 else int(rng.integers(2, 4))
            rows[i, 1:1 + k] = rng.choice(n + 6, k, replace=False)  # ids >= n: avatar
            counts[i] = k
        grab = dict(vertices=tuple(int(v) for v in rng.choice(n, 3, replace=False)), weights=(0.2, 0.3, 0.5),
                    target=(0, 0, 0), stiffness=1.0) if trial % 2 else None
        none = np.zeros((0, 5), np.uint32)
        sl = O.colors(graph, none, np.zeros(0, np.uint32), None, mode=1)
        col0 = [0] * n
        for c, vs in enumerate(sl):
            for v in vs:
                col0[v] = c
        assert O.colors(graph, none, np.zeros(0, np.uint32), None, mode=2) == sl
        jp0, _ = _jones_plassmann(graph, none, [], None, col0)
        assert jp0 == col0
        got = O.colors(graph, rows, counts, grab, mode=2)
        jp, rounds = _jones_plassmann(graph, rows, counts, grab, col0)
        exp = [[v for v in range(n) if jp[v] == c] for c in range(max(jp) + 1)]
        assert got == exp, trial
        assert rounds >= len(sl)
        col = jp
        for v in range(n):
            assert all(col[u] != col[v] for u in graph[v])
        for row, k in zip(rows, counts):
            vs = [int(v) for v in row[1:1 + k] if v < n]
            assert len({col[v] for v in vs}) == len(vs)
        if grab:
            assert len({col[v] for v in grab["vertices"]}) == 3


# ------------------------------------------------------------- prediction
def test_adaptive_guess_preserves_damped_fall_and_supported_rest_and_invalidates_edits():
    """src/vbd/prediction.rs:60-87."""
    s = scenes.default_settings(dt=1 / 60)
    st = scenes.State(np.array([[0.0, 1.0, 0.0]]), np.zeros((1, 3)), 0.0)
    p = O.Prediction()
    g = np.array(s["gravity"])
    for _ in range(5):
        nxt = p.guess(st, s)
        exp = (st.v[0] * s["dt"] + g * s["dt"] ** 2) / (1 + s["damping"] * s["dt"])
        assert np.linalg.norm(nxt[0] - st.x[0] - exp) < 1e-12
        old = st.v.copy()
        st.v = np.ascontiguousarray((nxt - st.x) / s["dt"])
        st.x = np.ascontiguousarray(nxt)
        st.time += s["dt"]
        p.record(old, st, s)
    st.v = np.zeros((1, 3))
    p.record(np.zeros((1, 3)), st, s)
    assert np.array_equal(p.guess(st, s), st.x)
    st.x[0, 0] += 0.01
    assert p.guess(st, s)[0, 1] < st.x[0, 1]
