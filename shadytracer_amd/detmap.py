"""The recorded hits on a detector as an illuminance map (optics/detmap.js detectorMap :104-213).

The plane is fitted to the hits: the normal is the direction they spread least along (the smallest principal axis of
their positions' covariance); when it lies within a degree of a world axis, the map is laid out along the other two
world axes. Cells are square, over the hits' bounding box in the plane (or a given range about their centre), `cells`
along the longer side; a cell's illuminance is the flux that landed in it over its area in square metres. The
statistics are over the lit cells (those hits landed in): max, mean, min, uniformity min / mean.
"""
import numpy as np


def fit_plane(P):
    """(centre, u, v, normal) of the plane the points P [n, 3] spread in; u, v along world axes when the normal is
    within a degree of one."""
    c = P.mean(axis=0)
    w, V = np.linalg.eigh(np.cov((P - c).T) if len(P) > 1 else np.eye(3))
    n = V[:, 0]
    axis = int(np.argmax(np.abs(n)))
    if abs(n[axis]) > np.cos(np.radians(1.0)):
        n = np.eye(3)[axis] * np.sign(n[axis])
        others = [k for k in range(3) if k != axis]
        u, v = np.eye(3)[others[0]], np.eye(3)[others[1]]
    else:
        u = V[:, 2]
        v = np.cross(n, u)
    return c, u, v, n


def detector_map(hits, cells=64, rng=None, metres_per_unit=1.0):
    """hits: [k, 8] (position xyz, flux, direction xyz, material) as Engine.optics_run returns them. rng: None, or
    {'a': (min, max), 'b': (min, max)} in the map's coordinates about the hits' centre. Returns a dict with E [ny, nx]
    (lx; row 0 at the -v edge), the cell size (m), the frame, and the lit cells' statistics."""
    H = np.asarray(hits, float)
    P, F = H[:, :3], H[:, 3]
    c, u, v, n = fit_plane(P)
    a, b = (P - c) @ u, (P - c) @ v
    if rng:
        (a0, a1), (b0, b1) = rng['a'], rng['b']
    else:
        a0, a1, b0, b1 = a.min(), a.max(), b.min(), b.max()
    side = max(a1 - a0, b1 - b0) / cells
    nx, ny = max(1, int(np.ceil((a1 - a0) / side - 1e-4))), max(1, int(np.ceil((b1 - b0) / side - 1e-4)))
    ix = np.clip(((a - a0) / side).astype(int), 0, nx - 1)
    iy = np.clip(((b - b0) / side).astype(int), 0, ny - 1)
    inside = (a >= a0) & (a <= a1) & (b >= b0) & (b <= b1)
    flux = np.zeros((ny, nx))
    count = np.zeros((ny, nx), int)
    np.add.at(flux, (iy[inside], ix[inside]), F[inside])
    np.add.at(count, (iy[inside], ix[inside]), 1)
    area = (side * metres_per_unit) ** 2
    E = flux / area
    lit = count > 0
    out = {'E': E, 'count': count, 'cell': side * metres_per_unit, 'nx': nx, 'ny': ny, 'centre': c, 'u': u, 'v': v,
           'normal': n, 'range': {'a': (a0, a1), 'b': (b0, b1)}, 'total': float(F[inside].sum()),
           'flatness': float(np.sqrt(np.mean(((P - c) @ n) ** 2)) / max(a1 - a0, b1 - b0, 1e-30))}
    if lit.any():
        Em = E[lit]
        out.update(max=float(Em.max()), meanLit=float(Em.mean()), minLit=float(Em.min()),
                   uniformity=float(Em.min() / Em.mean()) if Em.mean() > 0 else None,
                   hitsPerLitCell=float(count[lit].mean()))
    return out
