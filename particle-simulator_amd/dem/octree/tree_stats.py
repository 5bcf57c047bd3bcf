"""Statistics of a built octree (host side).

Same outputs as the reference's ``compute_tree_stats`` / ``print_tree_stats``
(dem/octree/tree_stats.py:17-201), read from the tree's field views: node and
leaf counts, leaf fill (min/max/mean/median over non-empty leaves and a fill
histogram with edges {0, mpl/4, mpl/2, mpl, max+1}), depth and nodes per level
(depth of every node recovered from the parent levels).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class TreeStatistics:
    tree_depth: int
    total_nodes: int
    max_nodes: int
    internal_nodes: int
    total_leaves: int
    non_empty_leaves: int
    empty_leaves: int
    total_particles: int
    mpl: int
    min_particles_per_leaf: int
    max_particles_per_leaf: int
    avg_particles_per_leaf: float
    median_particles_per_leaf: float
    leaf_fill_histogram: dict[str, int]
    nodes_per_level: list[tuple[int, int]]
    build_time_ms: float | None
    compute_time_ms: float | None


def _node_depths(octree) -> np.ndarray:
    """Depth of every node: children of parent level k sit at depth k+1."""
    n_nodes = int(octree.node_count[None])
    depths = np.zeros(n_nodes, dtype=np.int32)
    offsets = list(octree._parent_level_offsets)
    pc = int(octree.parent_count[None])
    if not offsets or pc == 0:
        return depths
    pfc = octree.parent_first_child.to_numpy()[:pc].astype(np.int64)
    bounds = offsets + [pc]
    for lev in range(len(offsets)):
        firsts = pfc[bounds[lev]: bounds[lev + 1]]
        firsts = firsts[firsts >= 0]
        if firsts.size == 0:
            continue
        kids = (firsts[:, None] + np.arange(8)).ravel()
        depths[kids[kids < n_nodes]] = lev + 1
    return depths


def _histogram(leaf_counts: np.ndarray, mpl: int) -> dict[str, int]:
    hist: dict[str, int] = {}
    if leaf_counts.size == 0:
        return hist
    top = int(leaf_counts.max())
    edges = sorted({0, max(mpl // 4, 1), max(mpl // 2, 1), mpl, top + 1})
    counts, _ = np.histogram(leaf_counts, bins=edges)
    hist["[0]"] = int((leaf_counts == 0).sum())
    for k, c in enumerate(counts):
        lo, hi = edges[k], edges[k + 1]
        if lo == 0 and hi == 1:
            continue
        if lo == 0:
            hist[f"[1, {hi})"] = int(c) - hist["[0]"]
        elif hi == top + 1 and lo > mpl:
            hist[f"({mpl}, ...]"] = int(c)
        else:
            close = "]" if hi in (mpl, top + 1) else ")"
            hist[f"[{lo}, {hi}{close}"] = int(c)
    return hist


def compute_tree_stats(octree, build_time_ms: float | None = None,
                       compute_time_ms: float | None = None) -> TreeStatistics:
    n_nodes = int(octree.node_count[None])
    mpl = int(octree.mpl_field[None])
    first_child = octree.nodes.first_child.to_numpy()[:n_nodes]
    count = octree.nodes.count.to_numpy()[:n_nodes]
    is_leaf = first_child < 0
    leaf_counts = count[is_leaf]
    filled = leaf_counts[leaf_counts > 0]
    depths = _node_depths(octree)
    per_level = np.bincount(depths) if n_nodes else np.zeros(0, dtype=np.int64)
    return TreeStatistics(
        tree_depth=int(depths.max()) + 1 if n_nodes > 0 else 0,
        total_nodes=n_nodes,
        max_nodes=octree.max_nodes,
        internal_nodes=int((~is_leaf).sum()),
        total_leaves=int(is_leaf.sum()),
        non_empty_leaves=int(filled.size),
        empty_leaves=int(is_leaf.sum()) - int(filled.size),
        total_particles=int(octree.num_particles[None]),
        mpl=mpl,
        min_particles_per_leaf=int(filled.min()) if filled.size else 0,
        max_particles_per_leaf=int(filled.max()) if filled.size else 0,
        avg_particles_per_leaf=float(filled.mean()) if filled.size else 0.0,
        median_particles_per_leaf=float(np.median(filled)) if filled.size else 0.0,
        leaf_fill_histogram=_histogram(leaf_counts, mpl),
        nodes_per_level=[(int(k), int(c)) for k, c in enumerate(per_level) if c > 0],
        build_time_ms=build_time_ms,
        compute_time_ms=compute_time_ms,
    )


def print_tree_stats(stats: TreeStatistics) -> None:
    pct = stats.total_nodes / stats.max_nodes * 100 if stats.max_nodes > 0 else 0
    print("=== Octree Statistics ===")
    print(f"  Tree depth:          {stats.tree_depth}")
    print(f"  Total nodes:         {stats.total_nodes} / {stats.max_nodes} ({pct:.1f}%)")
    print(f"    Internal:          {stats.internal_nodes}")
    print(f"    Leaves:            {stats.total_leaves} "
          f"(non-empty: {stats.non_empty_leaves}, empty: {stats.empty_leaves})")
    print(f"  Total particles:     {stats.total_particles}")
    print(f"  mpl:                 {stats.mpl}")
    print(f"  Particles/leaf:      min={stats.min_particles_per_leaf}, "
          f"max={stats.max_particles_per_leaf}, avg={stats.avg_particles_per_leaf:.1f}, "
          f"median={stats.median_particles_per_leaf:.1f}")
    if stats.build_time_ms is not None:
        print(f"  Build time:          {stats.build_time_ms:.1f} ms")
    if stats.compute_time_ms is not None:
        print(f"  Compute time:        {stats.compute_time_ms:.1f} ms")
    if stats.leaf_fill_histogram:
        print("\n  Leaf fill histogram:")
        for label, c in stats.leaf_fill_histogram.items():
            print(f"    {label:16s} {c}")
    if stats.nodes_per_level:
        print("\n  Nodes per level:")
        for level, c in stats.nodes_per_level:
            print(f"    Level {level}:  {c:>8}")
    print()
