"""CPU restatement of the reference's visualisation writers -- test
infrastructure only (the product path is zj_vtk_write / zj_pvtu_write).

Follows src/io/visualization.cpp:
  * subdivideElement :285-386 -- quadrilateral lattice (n+1)^2 points at
    xi = -1 + 2i/n, eta = -1 + 2j/n, two triangles per cell {v00, v10, v11},
    {v00, v11, v01}; triangle lattice (n+1)(n+2)/2 points at (i/n, j/n) by
    rows, triangles {v0, v1, v2} and {v1, v3, v2}; coordinates by
    ElementGeometry::refToPhys (basis.cpp:882-893, shape functions
    :717-773); every lattice point carries sol_pts[0] (:327, :360), or a
    zero State when the element has no solution points;
  * writeVTU :109-222, writeLegacy :8-107, writePVTU :247-283;
  * IdealGas::pressure / machNumber (include/physics/euler.hpp:16-52).
Number formatting: writeVTU streams doubles with the ostream defaults, which
is printf "%.6g"; writeLegacy sets std::scientific and setprecision(12)
before the first point, which is "%.12e" from then on.  Python's float
%-formatting is correctly rounded, as glibc's printf is.  NaN is printed with
its sign ("-nan" when the sign bit is set), as glibc does.

The pins of this restatement are in tests/test_io.py: closed-form lattice
coordinates on affine elements, point/cell counts, XML well-formedness, and
agreement of the values with an independent evaluation of the shape
functions and of the gas relations.
"""
import math

import numpy as np


def _shape(kind, order, xi, eta):
    """ElementGeometry::shapeFunc (basis.cpp:717-773) for the orders used:
    triangles 1 (3 nodes) and 2 (6 nodes), quadrilaterals 1 (4) and 2 (9)."""
    if kind == 0:
        L0, L1, L2 = 1.0 - xi - eta, xi, eta
        if order == 1:
            return [L0, L1, L2]
        return [L0 * (2 * L0 - 1), L1 * (2 * L1 - 1), L2 * (2 * L2 - 1), 4 * L0 * L1, 4 * L1 * L2, 4 * L2 * L0]
    if order == 1:
        return [0.25 * (1 - xi) * (1 - eta), 0.25 * (1 + xi) * (1 - eta), 0.25 * (1 + xi) * (1 + eta),
                0.25 * (1 - xi) * (1 + eta)]
    if order == 2:
        return [0.25 * xi * (xi - 1) * eta * (eta - 1), 0.25 * xi * (xi + 1) * eta * (eta - 1),
                0.25 * xi * (xi + 1) * eta * (eta + 1), 0.25 * xi * (xi - 1) * eta * (eta + 1),
                0.5 * (1 - xi * xi) * eta * (eta - 1), 0.5 * xi * (xi + 1) * (1 - eta * eta),
                0.5 * (1 - xi * xi) * eta * (eta + 1), 0.5 * xi * (xi - 1) * (1 - eta * eta),
                (1 - xi * xi) * (1 - eta * eta)]
    raise ValueError("geometric order not restated")


def subdivide(kind, order, nodes_xy, state, n):
    """Lattice points, triangles (local ids) and per-point states of one element."""
    pts, tris = [], []
    if kind == 1:
        for j in range(n + 1):
            for i in range(n + 1):
                pts.append((-1.0 + 2.0 * i / n, -1.0 + 2.0 * j / n))
        for j in range(n):
            for i in range(n):
                v00 = j * (n + 1) + i
                v10, v01 = v00 + 1, (j + 1) * (n + 1) + i
                v11 = v01 + 1
                tris += [(v00, v10, v11), (v00, v11, v01)]
    else:
        for j in range(n + 1):
            for i in range(n + 1 - j):
                pts.append((float(i) / n, float(j) / n))
        for j in range(n):
            row = j * (n + 1) - j * (j - 1) // 2
            nxt = (j + 1) * (n + 1) - (j + 1) * j // 2
            for i in range(n - j):
                tris.append((row + i, row + i + 1, nxt + i))
                if i < n - j - 1:
                    tris.append((row + i + 1, nxt + i + 1, nxt + i))
    xy = []
    for xi, eta in pts:
        N = _shape(kind, order, xi, eta)
        x = y = 0.0
        for k, w in enumerate(N):
            x += w * nodes_xy[k][0]
            y += w * nodes_xy[k][1]
        xy.append((x, y))
    return xy, tris, [state] * len(pts)


def pressure(s, gamma):
    rho = s[0]
    u, v, E = _div(s[1], rho), _div(s[2], rho), _div(s[3], rho)
    ke = 0.5 * (u * u + v * v)
    return (gamma - 1.0) * rho * (E - ke)


def _fmax(a, b):
    """C fmax: a NaN operand yields the other one."""
    if math.isnan(a):
        return b
    if math.isnan(b):
        return a
    return max(a, b)


def mach(s, gamma):
    rho = s[0]
    u, v = _div(s[1], rho), _div(s[2], rho)
    vel = math.sqrt(u * u + v * v)
    c = math.sqrt(gamma * _fmax(pressure(s, gamma), 1e-10) / _fmax(rho, 1e-10))
    return vel / c


def _div(a, b):
    """IEEE a / b (Python raises on b == 0)."""
    return float(np.float64(a) / np.float64(b)) if b == 0 else a / b


def fmt_g(x):
    if math.isnan(x):
        return "-nan" if math.copysign(1.0, x) < 0 else "nan"
    return "%.6g" % x


def fmt_e12(x):
    if math.isnan(x):
        return "-nan" if math.copysign(1.0, x) < 0 else "nan"
    return "%.12e" % x


def lattice(xy, elements, kinds, orders, states, n, gamma):
    """All lattice points, triangles and derived fields of a mesh.
    elements: node-id lists; states: per element sol_pts[0] (or None)."""
    P, T, S = [], [], []
    for e, nodes in enumerate(elements):
        st = states[e] if states[e] is not None else (0.0, 0.0, 0.0, 0.0)
        p, t, s = subdivide(kinds[e], orders[e], [xy[k] for k in nodes], st, n)
        off = len(P)
        P += p
        S += s
        T += [(a + off, b + off, c + off) for a, b, c in t]
    with np.errstate(all="ignore"):
        rho = [s[0] for s in S]
        pres = [pressure(s, gamma) for s in S]
        mac = [mach(s, gamma) for s in S]
        vel = [(_div(s[1], s[0]), _div(s[2], s[0])) for s in S]
    return P, T, rho, pres, mac, vel


def vtu_text(P, T, rho, pres, mac, vel):
    o = ['<?xml version="1.0"?>\n',
         '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
         '  <UnstructuredGrid>\n',
         f'    <Piece NumberOfPoints="{len(P)}" NumberOfCells="{len(T)}">\n',
         '      <Points>\n',
         '        <DataArray type="Float64" NumberOfComponents="3" format="ascii">\n']
    o += [f"          {fmt_g(x)} {fmt_g(y)} 0.0\n" for x, y in P]
    o += ['        </DataArray>\n', '      </Points>\n', '      <Cells>\n',
          '        <DataArray type="Int32" Name="connectivity" format="ascii">\n']
    o += [f"          {a} {b} {c}\n" for a, b, c in T]
    o += ['        </DataArray>\n', '        <DataArray type="Int32" Name="offsets" format="ascii">\n']
    o += [f"          {(i + 1) * 3}\n" for i in range(len(T))]
    o += ['        </DataArray>\n', '        <DataArray type="UInt8" Name="types" format="ascii">\n']
    o += ["          5\n"] * len(T)
    o += ['        </DataArray>\n', '      </Cells>\n',
          '      <PointData Scalars="Density" Vectors="Velocity">\n']
    for name, arr in (("Density", rho), ("Pressure", pres), ("Mach", mac)):
        o.append(f'        <DataArray type="Float64" Name="{name}" format="ascii">\n')
        o += [f"          {fmt_g(v)}\n" for v in arr]
        o.append('        </DataArray>\n')
    o.append('        <DataArray type="Float64" Name="Velocity" NumberOfComponents="3" format="ascii">\n')
    o += [f"          {fmt_g(u)} {fmt_g(v)} 0.0\n" for u, v in vel]
    o += ['        </DataArray>\n', '      </PointData>\n', '    </Piece>\n', '  </UnstructuredGrid>\n',
          '</VTKFile>\n']
    return "".join(o)


def legacy_text(P, T, rho, pres, mac, vel):
    o = ["# vtk DataFile Version 3.0\n", "Zhijian FR Solver Output\n", "ASCII\n", "DATASET UNSTRUCTURED_GRID\n",
         f"POINTS {len(P)} double\n"]
    o += [f"{fmt_e12(x)} {fmt_e12(y)} 0.0\n" for x, y in P]
    o.append(f"CELLS {len(T)} {len(T) * 4}\n")
    o += [f"3 {a} {b} {c}\n" for a, b, c in T]
    o.append(f"CELL_TYPES {len(T)}\n")
    o += ["5\n"] * len(T)
    o.append(f"POINT_DATA {len(P)}\n")
    for name, arr in (("Density", rho), ("Pressure", pres), ("Mach", mac)):
        o += [f"SCALARS {name} double 1\n", "LOOKUP_TABLE default\n"]
        o += [f"{fmt_e12(v)}\n" for v in arr]
    o.append("VECTORS Velocity double\n")
    o += [f"{fmt_e12(u)} {fmt_e12(v)} 0.0\n" for u, v in vel]
    return "".join(o)


def pvtu_text(pieces):
    o = ['<?xml version="1.0"?>\n',
         '<VTKFile type="PUnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
         '  <PUnstructuredGrid GhostLevel="0">\n',
         '    <PPointData Scalars="Density" Vectors="Velocity">\n',
         '      <PDataArray type="Float64" Name="Density"/>\n',
         '      <PDataArray type="Float64" Name="Pressure"/>\n',
         '      <PDataArray type="Float64" Name="Mach"/>\n',
         '      <PDataArray type="Float64" Name="Velocity" NumberOfComponents="3"/>\n',
         '    </PPointData>\n',
         '    <PPoints>\n',
         '      <PDataArray type="Float64" NumberOfComponents="3"/>\n',
         '    </PPoints>\n']
    o += [f'    <Piece Source="{p}"/>\n' for p in pieces]
    o += ['  </PUnstructuredGrid>\n', '</VTKFile>\n']
    return "".join(o)
