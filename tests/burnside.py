"""Independent counts of symmetry-unique index tuples by Burnside's lemma.

The unique four-index (two-index) tuples of a symmetric n^3 grid are the
orbits of translation classes of 4-point (2-point) configurations whose
per-axis span is at most n-1, under O_h (signed axis permutations, 48) times
the label swaps i<->j, k<->l, (ij)<->(kl) (8; for two-index tuples: i<->j, 2).
|orbits| = (1/|G|) sum_g |Fix(g)|, and Fix factorises over the cycles of the
axis permutation: along a cycle of length L with sign product S the per-axis
class x (a label vector modulo a common shift) must satisfy
S * x o rho - x = const with rho = pi^L.
No code path of the library is used here.
"""
import itertools

import numpy as np


def _label_group(arity):
    if arity == 2:
        return [(0, 1), (1, 0)]
    gens = [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]
    group = {(0, 1, 2, 3)}
    frontier = list(group)
    while frontier:
        nxt = []
        for p in frontier:
            for g in gens:
                q = tuple(p[g[a]] for a in range(4))
                if q not in group:
                    group.add(q)
                    nxt.append(q)
        frontier = nxt
    return sorted(group)


def _axis_classes(n, arity):
    """All label vectors with min 0 and max <= n-1 (one per translation class)."""
    pts = np.array(list(itertools.product(range(n), repeat=arity)), dtype=np.int64)
    return pts[pts.min(axis=1) == 0]


def _compose_power(p, L):
    q = tuple(range(len(p)))
    for _ in range(L):
        q = tuple(p[q[a]] for a in range(len(p)))
    return q


def orbit_count(n, arity=4):
    X = _axis_classes(n, arity)
    labels = _label_group(arity)
    cache = {}

    def fixed(S, rho):
        key = (S, rho)
        if key not in cache:
            y = S * X[:, list(rho)] - X
            cache[key] = int(np.sum(np.all(y == y[:, :1], axis=1)))
        return cache[key]

    total = 0
    for sigma in itertools.permutations(range(3)):
        # cycles of the axis permutation
        seen, cycles = set(), []
        for r in range(3):
            if r in seen:
                continue
            cyc, q = [], r
            while q not in seen:
                seen.add(q)
                cyc.append(q)
                q = sigma[q]
            cycles.append(cyc)
        for signs in itertools.product((1, -1), repeat=3):
            for pi in labels:
                f = 1
                for cyc in cycles:
                    S = int(np.prod([signs[r] for r in cyc]))
                    f *= fixed(S, _compose_power(pi, len(cyc)))
                total += f
    order = 48 * len(labels)
    assert total % order == 0
    return total // order


def axis_class_count(n):
    """Number of per-axis four-point translation classes (the per-axis key count)."""
    return n ** 4 - (n - 1) ** 4
