"""Solution output in the reference's formats.

- Line plots (ProbeOutput::WriteSolutionAtLineProbes, ProbeOutput.cpp:488-594
  with the probe setup of :361-460): evenly spaced points from p0 to p1,
  trilinear interpolation of cell-centre values (MathTools
  trilinear_interpolation.h:50-62), one row per point with the columns
  coordinate, density, velocity x/y/z, pressure, temperature, material ID and
  the level set(s), printed as %16.8e.
- VTK rectilinear-grid snapshots (.vtr, XML, cell data), readable by ParaView
  and VisIt like the reference's solution snapshots.

Arrays are the host layout of include/m2c_amd.h: ghosted bricks indexed
[k][j][i] (V with a trailing component axis).
"""
import base64
import math

import numpy as np

from .configs import _axis


def _locate(coords, p, lo_ok):
    """Cell index i (ghosted numbering, -1 = lower ghost) and the fraction
    between centre i and i+1 containing p, as the probe setup searches it
    (first i with p < x[i+1])."""
    # coords[0] is the lower ghost centre (index -1)
    n = len(coords)
    for q in range(0, n - 1):
        if p < coords[q + 1]:
            i = q - 1
            return i, (p - coords[q]) / (coords[q + 1] - coords[q])
    return None, None


def _interp(arr, comp, ijk, xi, shape_in):
    """InterpolateSolutionAtProbe (ProbeOutput.cpp:849-883) on a ghosted
    [k][j][i](,comp) array; ghost cells outside the domain in two or more
    directions (edges and corners, never filled) are replaced by the average
    of the valid corners."""
    i, j, k = ijk
    nx, ny, nz = shape_in
    vals, valid = [], []
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                ii, jj, kk = i + di, j + dj, k + dk
                outside = (ii < 0 or ii >= nx) + (jj < 0 or jj >= ny) + (kk < 0 or kk >= nz)
                ok = outside <= 1
                v = arr[kk + 1, jj + 1, ii + 1] if comp is None else arr[kk + 1, jj + 1, ii + 1, comp]
                vals.append(float(v) if ok else 0.0)
                valid.append(ok)
    # order: c000 c100 c010 c110 c001 c101 c011 c111 (i fastest)
    c000, c100, c010, c110, c001, c101, c011, c111 = vals
    nvalid = sum(valid)
    if nvalid < 8:
        avg = (c000 + c100 + c010 + c110 + c001 + c101 + c011 + c111) / nvalid
        c = [v if ok else avg for v, ok in zip(vals, valid)]
        c000, c100, c010, c110, c001, c101, c011, c111 = c
    c00 = c000 * (1.0 - xi[0]) + c100 * xi[0]
    c01 = c001 * (1.0 - xi[0]) + c101 * xi[0]
    c10 = c010 * (1.0 - xi[0]) + c110 * xi[0]
    c11 = c011 * (1.0 - xi[0]) + c111 * xi[0]
    c0 = c00 * (1.0 - xi[1]) + c10 * xi[1]
    c1 = c01 * (1.0 - xi[1]) + c11 * xi[1]
    return c0 * (1.0 - xi[2]) + c1 * xi[2]


def line_plot_rows(grid, V, ID, phis, p0, p1, npoints, temperature=None):
    """Rows (coordinate, rho, vx, vy, vz, p, T, ID, phi...) of a line probe.
    `temperature(rho, p, id)` gives T (default 0, the reference's value when
    no heat capacity is specified)."""
    n = (grid.nx, grid.ny, grid.nz)
    ax = [_axis(grid.nx, grid.x0, grid.xmax), _axis(grid.ny, grid.y0, grid.ymax), _axis(grid.nz, grid.z0, grid.zmax)]
    d = [(p1[a] - p0[a]) / (npoints - 1) for a in range(3)]
    h = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    rows = []
    for q in range(npoints):
        p = [p0[a] + q * d[a] for a in range(3)]
        if any(p[a] < ax[a][0] or p[a] >= ax[a][-1] for a in range(3)):
            rows.append([q * h] + [0.0] * (7 + len(phis)))
            continue
        ijk, xi = [], []
        for a in range(3):
            i, f = _locate(ax[a], p[a], True)
            ijk.append(i)
            xi.append(f)
        vals = [_interp(V, c, ijk, xi, n) for c in range(5)]
        myid = _interp(ID, None, ijk, xi, n)
        T = 0.0 if temperature is None else temperature(vals[0], vals[4], myid)
        row = [q * h] + vals[:4] + [vals[4], T, myid] + [_interp(ph, None, ijk, xi, n) for ph in phis]
        rows.append(row)
    return h, rows


def write_line_plot(path, grid, V, ID, phis, p0, p1, npoints, time, time_step, temperature=None):
    """A line-plot file in the reference's format (ProbeOutput.cpp:518-579)."""
    h, rows = line_plot_rows(grid, V, ID, phis, p0, p1, npoints, temperature)
    with open(path, "w") as f:
        f.write("## Line: (%e, %e, %e) -> (%e, %e, %e)\n" % (p0[0], p0[1], p0[2], p1[0], p1[1], p1[2]))
        f.write("## Number of points: %d (h = %e)\n" % (npoints, h))
        f.write("## Time: %e, Time step: %d.\n" % (time, time_step))
        f.write("## Coordinate  |  Density  |  Velocity (Vx,Vy,Vz)  |  Pressure  |  Temperature  |  Material ID  "
                "|  LevelSet(s)\n")
        for r in rows:
            line = "%16.8e  %16.8e  %16.8e  %16.8e  %16.8e  %16.8e  %16.8e  %16.8e" % tuple(r[:8])
            for v in r[8:]:
                line += "%16.8e  " % v
            f.write(line + "\n")
    return rows


def read_line_plot(path):
    """Header lines and the numeric rows of a line-plot file."""
    head, rows = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("##"):
                head.append(line.rstrip("\n"))
            elif line.strip():
                rows.append([float(x) for x in line.split()])
    return head, np.array(rows)


def _b64(arr):
    raw = np.ascontiguousarray(arr).tobytes()
    return base64.b64encode(np.uint32(len(raw)).tobytes() + raw).decode()


def write_vtr(path, grid, V, ID, phis=(), time=0.0):
    """A VTK rectilinear-grid snapshot: node coordinates at the cell faces,
    cell data density, velocity, pressure, material ID and level set(s)."""
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    xs = np.linspace(grid.x0, grid.xmax, nx + 1)
    ys = np.linspace(grid.y0, grid.ymax, ny + 1)
    zs = np.linspace(grid.z0, grid.zmax, nz + 1)
    inner = (slice(1, -1),) * 3
    fields = [("density", V[inner + (0,)], 1), ("velocity", V[inner + (slice(1, 4),)], 3),
              ("pressure", V[inner + (4,)], 1), ("materialid", ID[inner].astype(np.float64), 1)]
    for q, ph in enumerate(phis):
        fields.append(("levelset%d" % q, ph[inner], 1))
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="RectilinearGrid" version="0.1" byte_order="LittleEndian" '
                'header_type="UInt32">\n')
        f.write('<RectilinearGrid WholeExtent="0 %d 0 %d 0 %d">\n' % (nx, ny, nz))
        f.write('<FieldData><DataArray type="Float64" Name="TIME" NumberOfTuples="1" format="ascii">%.17g'
                '</DataArray></FieldData>\n' % time)
        f.write('<Piece Extent="0 %d 0 %d 0 %d">\n<CellData>\n' % (nx, ny, nz))
        for name, a, nc in fields:
            f.write('<DataArray type="Float64" Name="%s" NumberOfComponents="%d" format="binary">%s</DataArray>\n'
                    % (name, nc, _b64(np.asarray(a, dtype="<f8"))))
        f.write('</CellData>\n<Coordinates>\n')
        for c in (xs, ys, zs):
            f.write('<DataArray type="Float64" format="binary">%s</DataArray>\n' % _b64(c.astype("<f8")))
        f.write('</Coordinates>\n</Piece>\n</RectilinearGrid>\n</VTKFile>\n')
