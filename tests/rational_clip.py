"""Independent exact volume of (tetrahedron ∩ axis-aligned box) in rational arithmetic.

Used only to pin the oracle's clip: it shares no step with the reference's lookup-table clipping.
The intersection is the polytope {x : n_f . x <= c_f} of the 4 tet face half-spaces and the 6 box
half-spaces.  Its vertices are enumerated exactly (every triple of planes solved by Cramer's rule
in Fractions, kept when it satisfies all half-spaces); the volume is the sum over facets of the
pyramids from the vertex centroid to a fan triangulation of the facet polygon, whose vertices are
ordered around the facet centroid with exact cross-product comparisons.
"""

from __future__ import annotations

import functools
import itertools
from fractions import Fraction as F


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _det3(a, b, c):
    return _dot(a, _cross(b, c))


def tet_halfspaces(v):
    hs = []
    for l in range(4):
        i, j, k = [q for q in range(4) if q != l]
        n = _cross(_sub(v[j], v[i]), _sub(v[k], v[i]))
        c = _dot(n, v[i])
        if _dot(n, v[l]) > c:
            n = (-n[0], -n[1], -n[2])
            c = -c
        hs.append((n, c))
    return hs


def box_halfspaces(lo=(0, 0, 0), hi=(1, 1, 1)):
    hs = []
    for a in range(3):
        e = [F(0)] * 3
        e[a] = F(1)
        hs.append((tuple(e), F(hi[a])))
        hs.append((tuple(-x for x in e), -F(lo[a])))
    return hs


def _canonical(hs):
    """Scales each half-space so max |n_i| = 1 and drops duplicates (coinciding facets)."""
    out = {}
    for n, c in hs:
        s = max(abs(x) for x in n)
        if s == 0:
            continue
        key = (tuple(x / s for x in n), c / s)
        out[key] = key
    return list(out)


def polytope_volume(hs):
    hs = _canonical(hs)
    verts = set()
    for (n1, c1), (n2, c2), (n3, c3) in itertools.combinations(hs, 3):
        d = _det3(n1, n2, n3)
        if d == 0:
            continue
        # Cramer's rule for rows n1,n2,n3 and right-hand side (c1,c2,c3)
        x = []
        for a in range(3):
            m = [list(n1), list(n2), list(n3)]
            m[0][a], m[1][a], m[2][a] = c1, c2, c3
            x.append(_det3(tuple(m[0]), tuple(m[1]), tuple(m[2])) / d)
        p = tuple(x)
        if all(_dot(n, p) <= c for n, c in hs):
            verts.add(p)
    if len(verts) < 4:
        return F(0)
    verts = list(verts)
    g = tuple(sum(p[a] for p in verts) / len(verts) for a in range(3))
    vol = F(0)
    for n, c in hs:
        face = [p for p in verts if _dot(n, p) == c]
        if len(face) < 3:
            continue
        fc = tuple(sum(p[a] for p in face) / len(face) for a in range(3))
        ref = _sub(face[0], fc)

        def half(p):
            r = _sub(p, fc)
            cr = _dot(_cross(ref, r), n)
            if cr > 0:
                return 0
            if cr < 0:
                return 1
            return 0 if _dot(ref, r) > 0 else 1

        def cmp(p, q):
            hp, hq = half(p), half(q)
            if hp != hq:
                return hp - hq
            cr = _dot(_cross(_sub(p, fc), _sub(q, fc)), n)
            return -1 if cr > 0 else (1 if cr < 0 else 0)

        face.sort(key=functools.cmp_to_key(cmp))
        for k in range(1, len(face) - 1):
            vol += abs(_det3(_sub(face[0], g), _sub(face[k], g), _sub(face[k + 1], g))) / 6
    return vol


def tet_box_volume(v, lo=(0, 0, 0), hi=(1, 1, 1)):
    """Exact |tet(v) ∩ box| for rational (or integer) vertex coordinates."""
    v = [tuple(F(x) for x in p) for p in v]
    if _det3(_sub(v[1], v[0]), _sub(v[2], v[0]), _sub(v[3], v[0])) == 0:
        return F(0)
    return polytope_volume(tet_halfspaces(v) + box_halfspaces(lo, hi))
