"""Host views of the device octree's node arrays.

The reference's ``OctreeNode`` (dem/octree/octree_node.py:11-154) is a set of
Taichi fields (SoA) that callers read with ``.to_numpy()`` and ``[None]``:
tests (tests/octree/test_flat_tree.py), ``tree_stats`` (:84-90), the
visualiser and the PBM collision counter.  Here the arrays live in the tree's
device workspace; these views copy them to numpy on demand, with the same
names and shapes (max_nodes,).
"""

from __future__ import annotations

import numpy as np


class ScalarField:
    """A 0-d field: ``f[None]`` reads / writes the value (Taichi ``shape=()``)."""

    def __init__(self, value, dtype=None):
        self._dtype = dtype
        self.value = value

    def __getitem__(self, key):
        return self.value

    def __setitem__(self, key, value):
        self.value = self._dtype(value) if self._dtype else value

    def to_numpy(self):
        return np.array(self.value)


class DeviceField:
    """Read-only view of one named array of a built tree.

    ``component`` selects one column of an interleaved (k, count) array, e.g.
    ``min_x`` is column 0 of the (3, max_nodes) bbox-minimum array."""

    def __init__(self, tree, name: str, dtype, width: int = 1, component: int | None = None,
                 length: str = "max_nodes"):
        self._tree = tree
        self._name = name
        self._dtype = np.dtype(dtype)
        self._width = width
        self._component = component
        self._length = length

    @property
    def shape(self):
        n = self._tree._field_length(self._length)
        if self._width > 1 and self._component is None:
            return (n, self._width)
        return (n,)

    def to_numpy(self) -> np.ndarray:
        n = self._tree._field_length(self._length)
        raw = self._tree._read_field(self._name, self._dtype)
        if self._width > 1:
            raw = raw[: n * self._width].reshape(n, self._width)
            return raw[:, self._component].copy() if self._component is not None else raw.copy()
        return raw[:n].copy()

    def __getitem__(self, idx):
        return self.to_numpy()[idx]


class OctreeNode:
    """Structure-of-arrays node storage of a ``FlatOctree`` (device-resident)."""

    def __init__(self, tree, max_nodes: int):
        self.max_nodes = max_nodes
        f8, i4 = np.float64, np.int32
        self.min_x = DeviceField(tree, "min", f8, 3, 0)
        self.min_y = DeviceField(tree, "min", f8, 3, 1)
        self.min_z = DeviceField(tree, "min", f8, 3, 2)
        self.max_x = DeviceField(tree, "max", f8, 3, 0)
        self.max_y = DeviceField(tree, "max", f8, 3, 1)
        self.max_z = DeviceField(tree, "max", f8, 3, 2)
        self.R3_sum = DeviceField(tree, "R3_sum", f8)
        self.R3_cx = DeviceField(tree, "R3_c", f8, 3, 0)
        self.R3_cy = DeviceField(tree, "R3_c", f8, 3, 1)
        self.R3_cz = DeviceField(tree, "R3_c", f8, 3, 2)
        self.force_sum_x = DeviceField(tree, "force_sum", f8, 3, 0)
        self.force_sum_y = DeviceField(tree, "force_sum", f8, 3, 1)
        self.force_sum_z = DeviceField(tree, "force_sum", f8, 3, 2)
        for k, ab in enumerate(("xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz")):
            setattr(self, f"fmom_{ab}", DeviceField(tree, "fmom", f8, 9, k))
        self.max_radius = DeviceField(tree, "max_radius", f8)
        self.first_child = DeviceField(tree, "first_child", i4)
        self.next = DeviceField(tree, "next", i4)
        self.leaf_start = DeviceField(tree, "leaf_start", i4)
        self.count = DeviceField(tree, "count", i4)
