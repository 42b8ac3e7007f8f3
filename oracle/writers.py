"""The text writers of the reference's registration layer restated over plain arrays: Sphere.save_csv,
save_spectrum and save_zonal_mean (cpp/main.cpp:357-371, :328-335, :373-383) and Plane.save_csv, save_row and
save_column (cpp/main.cpp:464-493), each number printed by ``g`` = "%.10g" (cpp/main.cpp:84).

TEST INFRASTRUCTURE ONLY: the tests build the expected text of a file from the arrays a model returns and compare
it with what the library wrote.  The arithmetic (degree conversions, centre averages, corner vorticity, row sums) is
the reference's, in its order, in IEEE double precision, so equal arrays give equal bytes.
"""
from __future__ import annotations

import math

K_PI = 3.14159265358979323846  # cpp/main.cpp:31-32
K_DEG = K_PI / 180.0


def g(x: float) -> str:
    """cpp/main.cpp:84: snprintf "%.10g"."""
    return "%.10g" % float(x)


def sphere_csv(mu, h, u, v, zeta, q, hs) -> str:
    """cpp/main.cpp:357-371.  mu: Gauss nodes (south to north); the grids are nlat x nlon arrays.  lat(j) =
    asin(mu_j) and lon(k) = 2 pi k / nlon (sphere.cpp:120-121), printed in degrees."""
    J, K = len(mu), len(h[0])
    out = []
    for j in range(J):
        lat = g(math.asin(float(mu[j])) / K_DEG)
        for k in range(K):
            lon = g(2 * K_PI * k / K / K_DEG)
            out.append(" ".join([lon, lat, g(h[j][k]), g(u[j][k]), g(v[j][k]), g(zeta[j][k]), g(q[j][k]), g(hs[j][k])]) + "\n")
        out.append("\n")  # blank line between latitude rows (gnuplot pm3d)
    return "".join(out)


def sphere_spectrum(ke, pe) -> str:
    """cpp/main.cpp:328-335: "n ke pe" for n = 1 .. len(ke) - 1."""
    return "".join(f"{n} {g(ke[n])} {g(pe[n])}\n" for n in range(1, len(ke)))


def sphere_zonal_mean(mu, h, u) -> str:
    """cpp/main.cpp:373-383: per latitude the sequential sums of h and u over longitude, over nlon."""
    out = []
    K = len(h[0])
    for j in range(len(mu)):
        sh = su = 0.0
        for k in range(K):
            sh += float(h[j][k])
            su += float(u[j][k])
        out.append(f"{g(math.asin(float(mu[j])) / K_DEG)} {g(sh / K)} {g(su / K)}\n")
    return "".join(out)


class _Plane:
    def __init__(self, nx, ny, lx, ly, h, u, v):
        self.nx, self.ny, self.lx, self.ly = nx, ny, lx, ly
        self.h, self.u, self.v = h, u, v

    def dx(self):  # plane.hpp:26-31
        return self.lx / self.nx

    def dy(self):
        return self.ly / self.ny

    def x_center(self, i):
        return (i + 0.5) * self.dx()

    def y_center(self, j):
        return (j + 0.5) * self.dy()

    def zeta_at_corner(self, i, j):  # plane.cpp:220-226
        nx, ny = self.nx, self.ny
        im = (i + nx - 1) % nx
        dvdx = (float(self.v[j][i % nx]) - float(self.v[j][im])) / self.dx()
        uj = float(self.u[j][i % nx]) if j < ny else float(self.u[ny - 1][i % nx])
        ujm = float(self.u[j - 1][i % nx]) if j > 0 else float(self.u[0][i % nx])
        return dvdx - (uj - ujm) / self.dy()


def plane_csv(nx, ny, lx, ly, h, u, v) -> str:
    """cpp/main.cpp:464-477: "x y h u v zeta" per cell centre, u and v averaged to the centre, zeta the mean of the
    four corners; a blank line after each row.  h, u: ny x nx; v: (ny + 1) x nx."""
    p = _Plane(nx, ny, lx, ly, h, u, v)
    out = []
    for j in range(ny):
        for i in range(nx):
            uc = 0.5 * (float(u[j][i]) + float(u[j][(i + 1) % nx]))
            vc = 0.5 * (float(v[j][i]) + float(v[j + 1][i]))
            z = 0.25 * (p.zeta_at_corner(i, j) + p.zeta_at_corner(i + 1, j) + p.zeta_at_corner(i, j + 1) + p.zeta_at_corner(i + 1, j + 1))
            out.append(" ".join([g(p.x_center(i)), g(p.y_center(j)), g(h[j][i]), g(uc), g(vc), g(z)]) + "\n")
        out.append("\n")
    return "".join(out)


def plane_row(nx, ny, lx, ly, h, row) -> str:
    """cpp/main.cpp:479-485: "x h" along one row."""
    p = _Plane(nx, ny, lx, ly, h, None, None)
    return "".join(f"{g(p.x_center(i))} {g(h[row][i])}\n" for i in range(nx))


def plane_column(nx, ny, lx, ly, h, column) -> str:
    """cpp/main.cpp:487-493: "y h" along one column."""
    p = _Plane(nx, ny, lx, ly, h, None, None)
    return "".join(f"{g(p.y_center(j))} {g(h[j][column])}\n" for j in range(ny))
