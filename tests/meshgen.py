"""Reference-exact Cartesian mesh for tests.

parcop.setup shifts node bounds to faces (pyranda/parcop/parcop.f90:46-56),
patch.dx is (xnf-x1f)/nx (pyranda/parcop/patch.f90:83-85) and the grid is
x1f + (2i-1)/2*dx (pyranda/parcop/mesh.f90:173-177). Matching this rounding
is what lets the oracle reproduce the reference-held scalars to ~1e-11.
"""
import numpy as np


def axis(x1, xn, n):
    du = (xn - x1) / max(n - 1, 1)
    x1f = x1 - du / 2.0
    xnf = xn + du / 2.0
    d = (xnf - x1f) / n
    x = x1f + ((2.0 * np.arange(1, n + 1) - 1.0) * 0.5) * d
    return x, d


def mesh(dom):
    """dom: [(x1,xn,n), (y1,yn,n), (z1,zn,n)] -> (X,Y,Z), (dx,dy,dz)."""
    xs, ds = [], []
    for x1, xn, n in dom:
        x, d = axis(x1, xn, n)
        xs.append(x)
        ds.append(d)
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    return [np.asfortranarray(a) for a in (X, Y, Z)], tuple(ds)
