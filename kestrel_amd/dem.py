"""Bicubic resampling of elevation data (Bicubic, src/Interp2d.f90:214-282).

TileHeightData (src/dem.f90:320-340), which GetHeights (src/dem.f90:360-393)
calls for a DEM topography, gives each vertex the mean of four bicubic
evaluations, at the vertex offset by +-dx/2 and +-dy/2, each point mapped
into the elevation grid's 1-based data coordinates (tile_heights).  bicubic()
is the kernel of that resampling: on the data cell holding a point,
f(t, u) = sum a_pq t^p u^q is fixed by f, f_x, f_y and f_xy at the cell's four
corners, the derivatives taken by central differences over the neighbouring
samples; the 16 coefficients are M times the 16 corner values.  Reading the
DEM file itself (GDAL, src/dem.f90 and src/GeoTiffRead.f90) is not part of
this package: the caller passes the elevation samples, the raster's origin
and pixel size as the file gives them.
"""
from __future__ import annotations

import numpy as np

# corners (t, u) in the order of the corner values: (0,0), (1,0), (0,1), (1,1)
_CORNERS = ((0, 0), (1, 0), (0, 1), (1, 1))


def _coefficient_matrix() -> np.ndarray:
    """M: the inverse of the map from the 16 coefficients a_pq (index 4p + q)
    to the corner values [f, f_t, f_u, f_tu at the four corners] of
    f(t, u) = sum a_pq t^p u^q.  Its entries are small integers."""
    def dpow(s, n, d):     # d-th derivative of s^n at s in {0, 1}
        if d > n:
            return 0.0
        c = 1.0
        for k in range(d):
            c *= n - k
        return c * (s ** (n - d) if n - d > 0 else 1.0)
    B = np.zeros((16, 16))
    for blk, (dt, du) in enumerate(((0, 0), (1, 0), (0, 1), (1, 1))):
        for c, (t, u) in enumerate(_CORNERS):
            for p in range(4):
                for q in range(4):
                    B[4 * blk + c, 4 * p + q] = dpow(t, p, dt) * dpow(u, q, du)
    M = np.linalg.inv(B)
    Mi = np.rint(M)
    assert np.abs(M - Mi).max() < 1e-9
    return Mi


M = _coefficient_matrix()


def bicubic(z: np.ndarray, x, y) -> np.ndarray:
    """Bicubic of z (z[i-1, j-1] = the reference's z(i, j), i along x) at the
    1-based data coordinates (x, y); the points need one sample of margin on
    every side of their data cell."""
    z = np.asarray(z, dtype=np.float64)
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    y = np.atleast_1d(np.asarray(y, dtype=np.float64))
    x0 = np.floor(x).astype(np.int64)
    x1 = np.ceil(x).astype(np.int64)
    y0 = np.floor(y).astype(np.int64)
    y1 = np.ceil(y).astype(np.int64)
    nx, ny = z.shape
    if (x0 < 2).any() or (x1 + 1 > nx).any() or (y0 < 2).any() or (y1 + 1 > ny).any():
        raise ValueError("bicubic: points need one sample of margin around their data cell")

    def Z(i, j):
        return z[i - 1, j - 1]

    ex0 = (x1 - x0 + 1).astype(np.float64)   # x-derivative spans
    ex1 = (x1 + 1 - x0).astype(np.float64)
    ey0 = (y1 - y0 + 1).astype(np.float64)
    ey1 = (y1 + 1 - y0).astype(np.float64)
    f = [Z(x0, y0), Z(x1, y0), Z(x0, y1), Z(x1, y1),
         (Z(x1, y0) - Z(x0 - 1, y0)) / ex0, (Z(x1 + 1, y0) - Z(x0, y0)) / ex1,
         (Z(x1, y1) - Z(x0 - 1, y1)) / ex0, (Z(x1 + 1, y1) - Z(x0, y1)) / ex1,
         (Z(x0, y1) - Z(x0, y0 - 1)) / ey0, (Z(x1, y1) - Z(x1, y0 - 1)) / ey0,
         (Z(x0, y1 + 1) - Z(x0, y0)) / ey1, (Z(x1, y1 + 1) - Z(x1, y0)) / ey1,
         ((Z(x1, y1) - Z(x0 - 1, y1)) - (Z(x1, y0 - 1) - Z(x0 - 1, y0 - 1))) / ex0 / ey0,
         ((Z(x1 + 1, y1) - Z(x0, y1)) - (Z(x1 + 1, y0 - 1) - Z(x0, y0 - 1))) / ex1 / ey0,
         ((Z(x1, y1 + 1) - Z(x0 - 1, y1 + 1)) - (Z(x1, y0) - Z(x0 - 1, y0))) / ex0 / ey1,
         ((Z(x1 + 1, y1 + 1) - Z(x0, y1 + 1)) - (Z(x1 + 1, y0) - Z(x0, y0))) / ex1 / ey1]
    # a = M f, each row summed in column order
    a = []
    for r in range(16):
        acc = M[r, 0] * f[0]
        for c in range(1, 16):
            acc = acc + M[r, c] * f[c]
        a.append(acc)
    t = np.where(x == x0, 0.0, (x - x0) / np.where(x1 == x0, 1, x1 - x0).astype(np.float64))
    u = np.where(y == y0, 0.0, (y - y0) / np.where(y1 == y0, 1, y1 - y0).astype(np.float64))
    ans = np.zeros_like(x)
    for p in range(3, -1, -1):
        ans = t * ans + ((a[4 * p + 3] * u + a[4 * p + 2]) * u + a[4 * p + 1]) * u + a[4 * p]
    return ans


def tile_heights(z: np.ndarray, xv, yv, dx: float, dy: float, centre=(0.0, 0.0),
                 origin=(0.0, 0.0), pixel=(1.0, 1.0)) -> np.ndarray:
    """TileHeightData (src/dem.f90:320-340) at vertices (xv[i], yv[j]) of a
    domain centred at `centre` (its UTM easting and northing): the four points
    centre + (xv +- dx/2, yv +- dy/2), taken in the order (-,-), (-,+), (+,-),
    (+,+), each mapped to data coordinates (x - origin) / pixel + 1 (pixel
    widths as the raster gives them, negative for a north-up raster's y) and
    evaluated by bicubic(); the vertex takes 0.25 times their sum, added in
    that order.  Returns b0[len(xv), len(yv)] (index i along x)."""
    xv = np.asarray(xv, dtype=np.float64)
    yv = np.asarray(yv, dtype=np.float64)
    X, Y = np.meshgrid(xv, yv, indexing="ij")
    h = []
    for a in (0, 1):
        for b in (0, 1):
            px = centre[0] + X + ((a + 0.0) - 0.5) * dx
            py = centre[1] + Y + ((b + 0.0) - 0.5) * dy
            ix = (px - origin[0]) / pixel[0] + 1
            iy = (py - origin[1]) / pixel[1] + 1
            h.append(bicubic(z, ix.ravel(), iy.ravel()).reshape(X.shape))
    return 0.25 * (((h[0] + h[1]) + h[2]) + h[3])
