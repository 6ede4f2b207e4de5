"""Dimension trees of the HT format.

A tree is described by per-node arrays (lo, hi, left, right): node t covers the modes
[lo[t], hi[t]); node 0 is the root; leaves hold one mode; the left child of an interior
node holds its lower modes. Nodes are numbered breadth-first, which reproduces the
reference's heap numbering for its hard-coded trees
(third_order_tensor_decompsition/Python/HT_TensorLy.py:23,
fourth_order_tensor_decomposition/single_GPU/Python/HT_4d_TensorLy.py:23), so the
reference's rank arrays k[] (k[0] = 1) can be passed unchanged.
"""
from dataclasses import dataclass, field
from typing import List


@dataclass
class Tree:
    lo: List[int]
    hi: List[int]
    left: List[int]
    right: List[int]
    spec: object = field(default=None, compare=False)

    @property
    def n_nodes(self):
        return len(self.lo)

    @property
    def order(self):
        return self.hi[0]

    def is_leaf(self, t):
        return self.left[t] < 0

    def leaves(self):
        return [t for t in range(self.n_nodes) if self.is_leaf(t)]

    def interior(self):
        return [t for t in range(self.n_nodes) if not self.is_leaf(t)]

    def leaf_of_mode(self, m):
        for t in self.leaves():
            if self.lo[t] == m:
                return t
        raise KeyError(m)

    def validate(self):
        """Raise ValueError unless this is a binary contiguous tree covering every mode
        (the same rules the library's ht_validate_tree enforces)."""
        n = self.n_nodes
        d = self.order
        if not (2 <= d <= 8) or n != 2 * d - 1:
            raise ValueError("a tree over d modes has 2d-1 nodes (2 <= d <= 8)")
        if self.lo[0] != 0:
            raise ValueError("the root covers every mode")
        seen = [0] * d
        parent = [-1] * n
        for t in range(n):
            lo, hi, l, r = self.lo[t], self.hi[t], self.left[t], self.right[t]
            if not (0 <= lo < hi <= d):
                raise ValueError(f"node {t}: bad mode range [{lo},{hi})")
            if (l < 0) != (r < 0):
                raise ValueError(f"node {t}: needs 0 or 2 children")
            if l < 0:
                if hi != lo + 1:
                    raise ValueError(f"leaf {t} must hold one mode")
                seen[lo] += 1
            else:
                if not (0 < l < n and 0 < r < n) or parent[l] >= 0 or parent[r] >= 0:
                    raise ValueError(f"node {t}: bad children")
                parent[l] = parent[r] = t
                if not (self.lo[l] == lo and self.hi[l] == self.lo[r] and self.hi[r] == hi):
                    raise ValueError(f"node {t}: children must split [{lo},{hi}) contiguously")
        if seen != [1] * d:
            raise ValueError("every mode must be exactly one leaf")
        if any(parent[t] < 0 for t in range(1, n)):
            raise ValueError("every non-root node needs a parent")
        return self


def from_spec(spec):
    """Tree from a nested tuple of 0-based modes, e.g. ((0, 1), (2, 3))."""
    nodes = []
    queue = [spec]
    while queue:
        s = queue.pop(0)
        nodes.append(s)
        if isinstance(s, tuple):
            if len(s) != 2:
                raise ValueError("tree must be binary")
            queue.extend(s)
    index = {id(s): i for i, s in enumerate(nodes)}

    def span(s):
        if isinstance(s, tuple):
            return span(s[0])[0], span(s[1])[1]
        return s, s + 1

    lo, hi, left, right = [], [], [], []
    for s in nodes:
        a, b = span(s)
        lo.append(a)
        hi.append(b)
        if isinstance(s, tuple):
            left.append(index[id(s[0])])
            right.append(index[id(s[1])])
        else:
            left.append(-1)
            right.append(-1)
    return Tree(lo, hi, left, right, spec=spec).validate()


def reference_3d():
    """{1,2,3} -> {1,2},{3} -> {1},{2}: the tree of htd (3P/HT_TensorLy.py:23)."""
    return from_spec(((0, 1), 2))


def first_vs_rest_3d():
    """{1,2,3} -> {1},{2,3} -> {2},{3}."""
    return from_spec((0, (1, 2)))


def balanced_4d():
    """{1,2,3,4} -> {1,2},{3,4} -> leaves: the tree of htd4 (4P/HT_4d_TensorLy.py:23)."""
    return from_spec(((0, 1), (2, 3)))


def balanced(d):
    """Balanced tree over d modes (left half gets the extra mode)."""
    def build(lo, hi):
        if hi - lo == 1:
            return lo
        mid = lo + (hi - lo + 1) // 2
        return (build(lo, mid), build(mid, hi))
    return from_spec(build(0, d))
