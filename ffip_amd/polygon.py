"""Point-in-polygon test of the reference's pybind module (app/ffip.cpp:354-388).

``check_inside(req_pts, poly_pts)`` is True where a 2-D point lies strictly
inside the polygon given by its vertices in order (CGAL's ON_BOUNDED_SIDE); points
on an edge or a vertex are not inside.  The decision is exact for the double
inputs, as CGAL's predicates are: edge crossings and boundary contact come from
the sign of an orientation determinant, evaluated in floating point and
re-evaluated with exact rationals wherever its error bound leaves the sign in
doubt.  The polygon is taken as simple (no self-intersections), as CGAL's
bounded_side requires.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

# |a*b - c*d| over doubles: the rounded result is within this relative bound of
# the sum of the magnitudes of the terms (two products, one difference and the
# two coordinate differences feeding them)
_ORIENT_REL = 8 * np.finfo(float).eps


def _orient(ax, ay, bx, by, qx, qy):
    """Sign of (b - a) x (q - a): +1 left of a->b, -1 right, 0 on the line."""
    dx1, dy1 = bx - ax, by - ay
    dx2, dy2 = qx - ax, qy - ay
    l, r = dx1 * dy2, dy1 * dx2
    det = l - r
    s = np.sign(det).astype(np.int8)
    doubt = np.abs(det) <= _ORIENT_REL * (np.abs(l) + np.abs(r))
    if doubt.any():
        for q in zip(*np.nonzero(doubt)):
            exact = ((Fraction(float(bx[q])) - Fraction(float(ax[q]))) * (Fraction(float(qy[q])) - Fraction(float(ay[q])))
                     - (Fraction(float(by[q])) - Fraction(float(ay[q]))) * (Fraction(float(qx[q])) - Fraction(float(ax[q]))))
            s[q] = (exact > 0) - (exact < 0)
    return s


def check_inside(req_pts, poly_pts, chunk=65536):
    """Boolean array of shape req_pts.shape[:-1]: point strictly inside the polygon."""
    req = np.asarray(req_pts, dtype=float)
    poly = np.asarray(poly_pts, dtype=float)
    if req.ndim < 2 or poly.ndim < 2:
        raise ValueError("ndim is not bigger than 1")
    if req.shape[-1] != 2:
        raise ValueError("req_pts.shape[-1] is not 2")
    if poly.shape[-1] != 2:
        raise ValueError("poly_pts.shape[-1] is not 2")
    pts = req.reshape(-1, 2)
    ver = poly.reshape(-1, 2)
    out = np.zeros(pts.shape[0], dtype=bool)
    if ver.shape[0] < 3:
        return out.reshape(req.shape[:-1])
    ax, ay = ver[:, 0][None, :], ver[:, 1][None, :]
    nxt = np.roll(ver, -1, axis=0)
    bx, by = nxt[:, 0][None, :], nxt[:, 1][None, :]
    for s in range(0, pts.shape[0], chunk):
        qx = pts[s:s + chunk, 0][:, None]
        qy = pts[s:s + chunk, 1][:, None]
        shape = (qx.shape[0], ver.shape[0])
        o = _orient(*(np.broadcast_to(v, shape) for v in (ax, ay, bx, by, qx, qy)))
        # on an edge: collinear and within the edge's bounding box
        on = ((o == 0) & (qx >= np.minimum(ax, bx)) & (qx <= np.maximum(ax, bx))
              & (qy >= np.minimum(ay, by)) & (qy <= np.maximum(ay, by)))
        # crossings of the ray from q towards +x: upward edges with q on their
        # left, downward edges with q on their right (half-open in y)
        up = (ay <= qy) & (by > qy) & (o > 0)
        down = (by <= qy) & (ay > qy) & (o < 0)
        odd = (np.count_nonzero(up | down, axis=1) & 1).astype(bool)
        out[s:s + chunk] = odd & ~on.any(axis=1)
    return out.reshape(req.shape[:-1])
