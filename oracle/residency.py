"""Residency and analysis bookkeeping of the reference, restated in Python
(test infrastructure; the product never imports it).

* ``Residency``: TreeletMemoryManager's slot table, free-slot stack,
  visibility, freeIncorePosition and the insert / remove bookkeeping
  (src/TreeletMemoryManager.cpp:303-380, 392-443, 455-548, 560-700,
  src/TreeletMemoryManager.h:72-89).  Only the bookkeeping: which treelet
  sits in which slot.  ``expected_incore`` rebuilds the incore node buffer
  that the stitch rule gives for a slot table.
* ``aggregate_records``: the host half of RenderPassAnalyse::
  performAnalysePass (src/RenderPassAnalyse.cpp:248-356) over per-pixel
  analysis records, in the reference's own data-structure order (a map per
  set, a multiset ordered by descending error, the 12 MiB cap).

One deliberate difference: the reference's freeIncorePosition keeps its
probe cursor in a function-level static shared by every manager and never
reset; here each Residency keeps its own cursor starting at 1 (the same as
the reference for one manager in one process).
"""
from __future__ import annotations

from collections import defaultdict
from typing import Dict, List, Optional, Tuple

import numpy as np

MAX_VISIBILITY = 8  # src/TreeletMemoryManager.cpp:47


class Residency:
    def __init__(self, headers: np.ndarray, num_slots: int):
        """headers: (num_treelets, 8) uint32 treelet headers (num_nodes,
        num_leaves, mem_size, gid, parent_gid, parent_leaf_pos,
        parent_leaf_idx, parent_leaf_parent_pos)."""
        self.parent = [int(p) for p in headers[:, 4]]
        self.num_slots = num_slots
        self.search_pos = 1
        self.evictions = 0
        self.reset()

    # resetIncoreBuffer (:303-380)
    def reset(self) -> None:
        n = self.num_slots
        self.slot_gid = [-1] * n
        self.vis = [0] * n
        self.err = [100000.0] * n
        self.treelet_slot: Dict[int, int] = {}
        self.children = defaultdict(set)
        self.free = [n - i for i in range(1, n + 1)]  # std::stack, top = last
        assert self.insert(0, MAX_VISIBILITY, 999.0)

    # insertTreeletIntoIncoreBuffer (:455-548)
    def insert(self, gid: int, visibility: int, error: float) -> bool:
        if gid >= len(self.parent):
            return False
        if not self.free and not self.free_position():
            return False
        slot = self.free.pop()
        self.slot_gid[slot], self.vis[slot], self.err[slot] = gid, visibility, error
        self.treelet_slot[gid] = slot
        if gid != 0:
            self.children[self.parent[gid]].add(gid)
        return True

    # freeIncorePosition (:560-653)
    def free_position(self) -> bool:
        freeable = 0
        tests = 0
        for _ in range(self.num_slots):
            s = self.search_pos
            if self.vis[s] < 0 and self.slot_gid[s] != -1:
                p = self.parent[self.slot_gid[s]]
                if self.vis[self.treelet_slot[p]] < 0:
                    freeable = self.slot_gid[s]
                    break
            self.search_pos += 1
            tests += 1
            if self.search_pos == self.num_slots:
                self.search_pos = 1
            if tests == 100:
                break
        if freeable == 0:
            return False
        queue = [freeable]
        to_clear = []
        while queue:
            p = queue.pop(0)
            for c in sorted(self.children[p]):  # std::set iterates in ascending order
                queue.append(c)
                to_clear.append(c)
            self.children[p] = set()
        self.children[freeable] = set()
        for c in to_clear:
            slot = self.treelet_slot.pop(c)
            assert slot != 0
            self.free.append(slot)
            self.slot_gid[slot], self.vis[slot], self.err[slot] = -1, 0, 100000.0
            self.evictions += 1
        self.remove(freeable)
        self.evictions += 1
        return True

    # removeTreeletFromIncoreBuffer (:665-700)
    def remove(self, gid: int) -> None:
        slot = self.treelet_slot.pop(gid)
        self.slot_gid[slot], self.vis[slot], self.err[slot] = -1, 0, 100000.0
        self.free.append(slot)
        self.children[self.parent[gid]].discard(gid)

    # updateClientSideIncoreBuffer (:392-443)
    def update_client_side(self, visible_old: List[Tuple[int, float]], visible_new: List[Tuple[int, float]]) -> int:
        for i in range(1, self.num_slots):
            self.vis[i] -= 1
        for gid, _ in visible_old:
            self.vis[self.treelet_slot[gid]] = MAX_VISIBILITY
        inserted = 0
        for gid, err in visible_new:
            if self.insert(gid, MAX_VISIBILITY, err):
                inserted += 1
        return inserted


def aggregate_records(records: Dict[str, np.ndarray], slot_gid: List[int], parent: List[int], nodes_per_treelet: int,
                      treelet_bytes: int):
    """RenderPassAnalyse::performAnalysePass, host half (:248-356).  records:
    rec_kind / rec_err / rec_sub / node arrays of the analysis frame (any
    shape, row-major order is the feedback-buffer order).  Returns
    (visible_old, visible_new) as lists of (gid, error): old in ascending
    gid, new by descending error (ties in ascending gid), capped."""
    kind = records["rec_kind"].ravel()
    err = records["rec_err"].ravel()
    sub = records["rec_sub"].ravel()
    node = records["node"].ravel()
    new: Dict[int, float] = {}
    old: Dict[int, float] = {}
    for i in np.nonzero(kind)[0]:
        e = float(err[i])
        s = int(sub[i])
        if s:
            if s not in new or e > new[s]:
                new[s] = e
        g = slot_gid[int(node[i]) // nodes_per_treelet]
        while g != 0:  # addNodeAndItsParentTreeletsToVisibleTreelets
            if g not in old or e > old[g]:
                old[g] = e
            g = parent[g]
        old[0] = 99.0
    visible_new = sorted(sorted(new.items()), key=lambda kv: -kv[1])  # stable: ties keep ascending gid
    cap = int((12.0 * 1048576.0) / treelet_bytes)
    if len(visible_new) > cap:
        visible_new = visible_new[:cap]
    return sorted(old.items()), visible_new


def expected_incore(scene, slot_gid: List[int], nodes_per_treelet: int) -> np.ndarray:
    """Incore node buffer (num_slots*N, 2) uint32 that the reference's stitch
    rule gives for a slot table: every resident treelet copied into its slot,
    each resident non-root treelet's root copied over its anchor leaf with
    first_child = (slot*N + 1) - anchor (mod 2^32) and the leaf bit cleared in
    the anchor's parent (src/TreeletMemoryManager.cpp:488-539).  Slots that
    are free hold whatever was there before; only resident slots are
    compared."""
    N = nodes_per_treelet
    buf = np.zeros((len(slot_gid) * N, 2), np.uint32)
    where = {g: s for s, g in enumerate(slot_gid) if g >= 0}
    for s, g in enumerate(slot_gid):
        if g < 0:
            continue
        h, nodes, _ = scene.treelet(g)
        buf[s * N:s * N + h["num_nodes"]] = nodes
    for s, g in enumerate(slot_gid):
        if g <= 0:
            continue
        h, nodes, _ = scene.treelet(g)
        ps = where[h["parent_gid"]]
        leaf = ps * N + h["parent_leaf_pos"]
        lp = ps * N + h["parent_leaf_parent_pos"]
        buf[leaf] = nodes[0]
        buf[leaf, 0] = (s * N + 1 - leaf) & 0xFFFFFFFF
        buf[lp, 1] &= ~np.uint32(1 << h["parent_leaf_idx"])
    return buf
