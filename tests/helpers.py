"""Pure-Python views of treelet bundles used by the CPU tests."""
import numpy as np

VALID_SHIFT = 8


def valid(masks, c):
    return (int(masks) >> (VALID_SHIFT + c)) & 1


def leaf(masks, c):
    return (int(masks) >> c) & 1


def nth_child_loop(masks, first, c):
    """getNthChildIndex as the reference writes it (src/CpuSvoNode.cpp:145-152):
    add one per valid child below c, modulo 2^32."""
    index = first
    for i in range(c):
        index += valid(masks, i)
    return index & 0xFFFFFFFF


def treelets(scene):
    info = scene.info
    return [scene.treelet(g) for g in range(info["num_treelets"])]


def morton(i, j, k, depth):
    c = 0
    for b in range(depth):
        c |= ((i >> b) & 1) << (3 * b + 2)
        c |= ((j >> b) & 1) << (3 * b + 1)
        c |= ((k >> b) & 1) << (3 * b)
    return c


def morton_np(i, j, k, depth):
    i, j, k = (np.asarray(a, dtype=np.uint64) for a in (i, j, k))
    c = np.zeros(i.shape, np.uint64)
    for b in range(depth):
        one = np.uint64(1)
        c |= ((i >> np.uint64(b)) & one) << np.uint64(3 * b + 2)
        c |= ((j >> np.uint64(b)) & one) << np.uint64(3 * b + 1)
        c |= ((k >> np.uint64(b)) & one) << np.uint64(3 * b)
    return c


def walk_leaves(tls, depth):
    """Morton codes (and colour) of every octree leaf reached by walking the
    treelets from the root through sub-treelet anchors."""
    out = {}
    stack = [(0, 0, 0, 0)]  # (gid, local, code, level)
    while stack:
        gid, loc, code, lvl = stack.pop()
        h, nodes, attr = tls[gid]
        first, masks = int(nodes[loc, 0]), int(nodes[loc, 1])
        for c in range(8):
            if not valid(masks, c):
                continue
            ch = (loc + nth_child_loop(masks, first, c)) & 0xFFFFFFFF
            ccode = (code << 3) | c
            if leaf(masks, c):
                sub = int(nodes[ch, 0])
                if lvl + 1 == depth:
                    assert sub == 0, "scene-depth leaf with a sub-treelet"
                    out[ccode] = int(attr[ch, 1])
                else:
                    assert sub > 0, "anchor above scene depth without sub-treelet"
                    stack.append((sub, 0, ccode, lvl + 1))
            else:
                stack.append((gid, ch, ccode, lvl + 1))
    return out
