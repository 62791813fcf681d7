"""The maxLoops branches of the traversal, pinned by a third, pure-Python
restatement written line by line from the reference (numpy float32 scalars:
IEEE single arithmetic, no contraction), independent of the C++ oracle:

  * sample_a: opencl/treeletRenderer_rgba.cl:100-178 (intersectAABB) and
    :201-382 (sample): maxLoops reports the current parent as a hit;
  * travers_c: src/incore/CpuRaycasterSingleRay3.cpp:158-387 (traversSvo):
    maxLoops is a miss; its root-box test is taken as intersectAABB's (gloost
    is not vendored).

A camera 2000 units away makes the epsilon vanish against t (its ulp is
~1e-4 > 1e-5), so `tmin + epsilon` rounds to `tmin`: rays entering a child
through a midplane pick the child behind it, whose exit t equals tmin, and
repeat that step until maxLoops.  For those rays and a sample of ordinary
ones the oracle's node, t, depth, loop and push counts must equal the
restatement's."""
import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import workloads

F = np.float32
MISS = 0xFFFFFFFF


def intersect_aabb(o, d):
    t_near, t_far, r = F(-10000000.0), F(10000000.0), F(0.5)
    for i in range(3):
        if d[i] == 0 and (-r > o[i] or r < o[i]):
            return None
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = (-r - o[i]) / d[i]
            t2 = (r - o[i]) / d[i]
        if t1 > t2:
            t1, t2 = t2, t1
        if t1 > t_near:
            t_near = t1
        if t2 < t_far:
            t_far = t2
        if t_near > t_far or t_far < F(0):
            return None
    return t_near, t_far


def nth_child(masks, first, c):
    index = first
    for i in range(c):
        index += (masks >> (i + 8)) & 1
    return index & 0xFFFFFFFF


def sign(x):
    return F(1) if x > 0 else (F(-1) if x < 0 else x)


def _child_box(o, d, rc, mo, eps, tmin, center, scale_exp2):
    half = scale_exp2 * F(0.5)
    p = [(o[a] + (tmin + eps) * d[a]) - center[a] for a in range(3)]
    ci = 4 * int(p[0] > 0) + 2 * int(p[1] > 0) + int(p[2] > 0)
    cc = [F(-0.5 + bool(ci & (4 >> a))) * scale_exp2 + center[a] for a in range(3)]
    t0, t1 = [], []
    for a in range(3):
        lo = rc[a] * (cc[a] - half) + mo[a]
        hi = rc[a] * (cc[a] + half) + mo[a]
        if lo > hi:
            lo, hi = hi, lo
        t0.append(lo)
        t1.append(hi)
    # OpenCL max(x, y) = y if x < y else x; min(x, y) = y if y < x else x
    mx = lambda x, y: y if x < y else x
    mn = lambda x, y: y if y < x else x
    return ci, cc, mx(t0[0], mx(t0[1], t0[2])), mn(t1[0], mn(t1[1], t1[2]))


def sample_a(nodes, o, d, ratio):
    """treeletRenderer_rgba.cl sample(), relative indexing on the incore buffer."""
    box = intersect_aabb(o, d)
    if box is None:
        return {"node": MISS, "loops": 0, "pushes": 0}
    t_min, t_max = box
    scale_max, eps, min_normal = 13, F(0.00001), F(0.0001)
    d = [min_normal * sign(x) if abs(x) < min_normal else x for x in d]
    rc = [F(1.0) / x for x in d]
    mo = [-o[a] * rc[a] for a in range(3)]
    scale = scale_max - 1
    stack = {scale: [0, t_min, t_max, [F(0), F(0), F(0)]]}
    loops = pushes = 0
    max_loops = (scale_max + 1) * (scale_max + 1) * 5
    while scale < scale_max:
        parent = stack[scale]
        scale_exp2 = F(2.0) ** F(scale - scale_max)
        loops += 1
        if loops == max_loops:
            return {"node": parent[0], "t": parent[1], "depth": scale_max - scale, "loops": loops, "pushes": pushes}
        if parent[2] < F(0):
            scale += 1
            continue
        if abs(parent[1] - parent[2]) > eps:
            ci, cc, tc_min, tc_max = _child_box(o, d, rc, mo, eps, parent[1], parent[3], scale_exp2)
            first, masks = int(nodes[parent[0], 0]), int(nodes[parent[0], 1])
            if (masks >> (ci + 8)) & 1:
                child = (parent[0] + nth_child(masks, first, ci)) & 0xFFFFFFFF
                if (masks >> ci) & 1:
                    return {"node": child, "t": parent[1], "depth": scale_max - scale, "loops": loops,
                            "pushes": pushes}
                if ratio * tc_min > scale_exp2 or scale_max - scale == 12:
                    return {"node": child, "t": parent[1], "depth": scale_max - scale + 1, "loops": loops,
                            "pushes": pushes}
                parent[1] = tc_max
                scale -= 1
                pushes += 1
                stack[scale] = [child, tc_min, tc_max, cc]
                continue
            parent[1] = tc_max
        else:
            scale += 1
    return {"node": MISS, "t": t_max, "loops": loops, "pushes": pushes}


def travers_c(nodes, o, d, ratio, max_depth):
    """CpuRaycasterSingleRay3::traversSvo in FP32, relative indexing."""
    box = intersect_aabb(o, d)
    if box is None:
        return {"node": MISS, "loops": 0, "pushes": 0}
    t_min, t_max = box
    scale_max, eps = max_depth, F(0.00001)
    d = [eps * sign(x) * F(10.0) if abs(x) < eps else x for x in d]
    rc = [F(1.0) / x for x in d]
    mo = [-o[a] * rc[a] for a in range(3)]
    scale = scale_max - 1
    stack = {scale: [0, t_min, t_max, [F(0), F(0), F(0)]]}
    loops = pushes = 0
    max_loops = (max_depth + 1) * (max_depth + 1) * 5
    refetch = True
    while scale < scale_max:
        loops += 1
        if loops == max_loops:
            return {"node": MISS, "loops": loops, "pushes": pushes}
        if refetch:
            p_idx, p_tmin, p_tmax, p_center = stack[scale]
            scale_exp2 = F(2.0) ** F(scale - scale_max)
            refetch = False
        if p_tmax < F(0) or abs(p_tmin - p_tmax) < eps:
            scale += 1
            refetch = True
            continue
        if p_tmin < p_tmax:
            ci, cc, tc_min, tc_max = _child_box(o, d, rc, mo, eps, p_tmin, p_center, scale_exp2)
            first, masks = int(nodes[p_idx, 0]), int(nodes[p_idx, 1])
            if (masks >> (ci + 8)) & 1:
                child = (p_idx + nth_child(masks, first, ci)) & 0xFFFFFFFF
                if (masks >> ci) & 1 or ratio * tc_max > scale_exp2:
                    return {"node": child, "t": p_tmin, "depth": scale_max - scale, "loops": loops, "pushes": pushes}
                stack[scale][1] = tc_max
                scale -= 1
                pushes += 1
                stack[scale] = [child, tc_min, tc_max, cc]
                refetch = True
                continue
            p_tmin = tc_max
        else:
            scale += 1
            refetch = True
    return {"node": MISS, "loops": loops, "pushes": pushes}


@pytest.fixture(scope="module")
def far_setup(tmp_path_factory):
    import oracle
    sc = workloads.menger_scene(depth=6, level=3, treelet_bytes=4096)
    assert sc.info["num_treelets"] > 1
    base = str(tmp_path_factory.mktemp("maxloops") / "menger6")
    sc.write(base)
    ib = workloads.incore_bytes_for(sc)
    ora = oracle.OracleScene(base, ib)
    mm = svo.TreeletMemoryManagerHip(sc, ib, device=-1, resident_all=True)
    nodes, _ = mm.getIncoreBuffer()
    eye = np.array([0.7, 0.45, 0.55])
    eye = eye / np.linalg.norm(eye) * 2000.0
    cam = svo.look_at(tuple(eye), (0.01, 0.02, 0.03), (0.0, 1.0, 0.0), 0.09, 64, 64, far=4000.0)
    return sc, ora, nodes, cam


@pytest.mark.parametrize("policy", ["treelet", "cpu3"])
def test_maxloops_rays_match_python_restatement(far_setup, policy):
    import oracle
    sc, ora, nodes, cam = far_setup
    ref = ora.render(cam, policy)
    dirs = oracle.ray_dirs(cam)
    o = [F(v) for v in cam.position[:3]]
    depth = sc.info["max_depth"]
    max_loops = (14 * 14 * 5) if policy == "treelet" else (depth + 1) ** 2 * 5
    stuck = np.argwhere(ref["loops"] == max_loops)
    assert len(stuck) > 40, "the far camera must drive rays into maxLoops"
    rng = np.random.default_rng(3)
    others = np.argwhere(ref["loops"] < max_loops)
    picks = [tuple(p) for p in stuck] + [tuple(p) for p in others[rng.choice(len(others), 40, replace=False)]]
    for (y, x) in picks:
        d = [F(v) for v in dirs[y, x]]
        got = sample_a(nodes, o, d, F(cam.t_scale_ratio)) if policy == "treelet" else \
            travers_c(nodes, o, d, F(cam.t_scale_ratio), depth)
        where = f"pixel ({x}, {y}), policy {policy}"
        assert got["node"] == int(ref["node"][y, x]), where
        assert got["loops"] == int(ref["loops"][y, x]), where
        assert got["pushes"] == int(ref["pushes"][y, x]), where
        if got["node"] != MISS:
            assert np.float32(got["t"]).view(np.uint32) == ref["t"][y, x].view(np.uint32), where
            assert got["depth"] == int(ref["depth"][y, x]), where
    if policy == "treelet":  # maxLoops returns the parent as a hit
        assert all(ref["node"][y, x] != MISS for (y, x) in stuck)
    else:  # and the CPU caster a miss
        assert all(ref["node"][y, x] == MISS for (y, x) in stuck)


def _inner_nodes(nodes):
    """Every inner node reachable from the root of a stitched incore buffer
    (relative child offsets): index -> (tree depth, centre), in float64.  The
    root is centred at the origin with edge 1; child c of a node of depth k
    lies at centre + (bit - 0.5) * 2^-(k+1) per axis (x = bit 2)."""
    out = {0: (0, np.zeros(3))}
    stack = [0]
    while stack:
        i = stack.pop()
        k, c = out[i]
        first, masks = int(nodes[i, 0]), int(nodes[i, 1])
        for ci in range(8):
            if not (masks >> (8 + ci)) & 1 or (masks >> ci) & 1:
                continue  # empty or leaf
            j = (i + nth_child(masks, first, ci)) & 0xFFFFFFFF
            bits = np.array([(ci >> 2) & 1, (ci >> 1) & 1, ci & 1], np.float64)
            out[j] = (k + 1, c + (bits - 0.5) * 2.0 ** -(k + 1))
            stack.append(j)
    return out


def _check_maxloops_hits(ref, dirs, origin, inner, stuck, max_loops):
    """Geometric checks of the policy-A maxLoops result, independent of the
    traversal's arithmetic: the reported node is an inner node of the tree,
    reported with its own depth (the depth of its children, as sample()
    counts scales) and centre; the point o + t*d lies in its box and on one
    of its midplanes or faces (a stalled step sits on a boundary of the
    child it keeps selecting), and the loop count is the closed form
    (scaleMax+1)^2 * 5.
    Returns the list of failures."""
    bad = []
    for (y, x) in stuck:
        n = int(ref["node"][y, x])
        where = f"pixel ({x}, {y})"
        if int(ref["loops"][y, x]) != max_loops:
            bad.append(f"{where}: loops {int(ref['loops'][y, x])} != {max_loops}")
        if n not in inner:
            bad.append(f"{where}: node {n} is not an inner node")
            continue
        k, c = inner[n]
        if int(ref["depth"][y, x]) != k + 1:
            bad.append(f"{where}: depth {int(ref['depth'][y, x])}, node depth {k}")
        if not np.array_equal(ref["center"][y, x].astype(np.float64), c):
            bad.append(f"{where}: centre {ref['center'][y, x]} != {c}")
        p = origin + float(ref["t"][y, x]) * dirs[y, x].astype(np.float64)
        half = 2.0 ** -(k + 1)
        tol = 1e-3 * half + 1e-3  # t ~ 2000 in float32: p is good to ~2.5e-4
        if np.any(np.abs(p - c) > half + tol):
            bad.append(f"{where}: o + t d = {p} outside the box of node {n}")
        rel = np.abs(p - c)  # a stalled step sits on a midplane or a face of the node
        if np.min(np.minimum(rel, np.abs(rel - half))) > tol:
            bad.append(f"{where}: o + t d = {p} on no plane of node {n}")
    return bad


def test_maxloops_hit_is_the_stalled_parent(far_setup):
    """Policy A reports the parent node at maxLoops
    (opencl/treeletRenderer_rgba.cl:257,268-277).  On every far-camera ray
    that reaches maxLoops the oracle's result passes the geometric checks;
    planting a wrong node (the reported node's parent, or a sibling) or an
    off-by-one loop count in the oracle's result fails them."""
    import oracle
    sc, ora, nodes, cam = far_setup
    ref = ora.render(cam, "treelet")
    dirs = oracle.ray_dirs(cam)
    origin = np.array(cam.position[:3], np.float64)
    max_loops = 14 * 14 * 5  # scaleMax 13 for depth <= 12
    stuck = [tuple(p) for p in np.argwhere(ref["loops"] == max_loops)]
    assert len(stuck) > 40
    inner = _inner_nodes(nodes)
    assert not _check_maxloops_hits(ref, dirs, origin, inner, stuck, max_loops)
    parent_of = {}
    for i, (k, c) in inner.items():
        first, masks = int(nodes[i, 0]), int(nodes[i, 1])
        for ci in range(8):
            if (masks >> (8 + ci)) & 1 and not (masks >> ci) & 1:
                parent_of[(i + nth_child(masks, first, ci)) & 0xFFFFFFFF] = i
    y, x = stuck[0]
    for plant in ("parent", "sibling", "loops"):
        bad = {k: v.copy() for k, v in ref.items()}
        for (yy, xx) in stuck:
            n = int(ref["node"][yy, xx])
            if plant == "parent":
                bad["node"][yy, xx] = parent_of[n]
            elif plant == "sibling":
                p = parent_of[n]
                sibs = [j for j, q in parent_of.items() if q == p and j != n]
                bad["node"][yy, xx] = sibs[0] if sibs else parent_of[n]
            else:
                bad["loops"][yy, xx] = max_loops - 1
        fails = _check_maxloops_hits(bad, dirs, origin, inner, stuck, max_loops)
        assert len(fails) >= len(stuck), f"planted {plant}: only {len(fails)} of {len(stuck)} rays caught"
