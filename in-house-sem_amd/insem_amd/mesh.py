"""Structured macro-mesh templates: a rectangle, horizontal layers with
curved interfaces, and the axisymmetric borehole (the templates of the
reference's python/insem/mesh/templates.py).

The templates produce the reference's node lattices, cell order (row by
row, x fastest), biquadratic node order, boundary tags and region ids, so
that a model built here numbers its elements -- and therefore sums its
assembly -- as the reference's does.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Sequence

import numpy as np

from .core import MacroMesh

BOTTOM = 1  # z = z_min
RIGHT = 2   # x = x_max
TOP = 3     # z = z_max
LEFT = 4    # x = x_min


def _axis_breaks(spec, n, axis):
    """Cell boundaries along one axis: explicit increasing breakpoints, or
    n equal cells of (min, max)."""
    v = np.asarray(spec, dtype=float)
    if n is not None:
        if v.shape != (2,) or not v[1] > v[0]:
            raise ValueError(f"with n{axis} given, {axis} is (min, max) with min < max, not {spec!r}")
        if n < 1:
            raise ValueError(f"n{axis} is at least 1, not {n}")
        return np.linspace(v[0], v[1], n + 1)
    if v.ndim != 1 or v.size < 2 or not np.all(np.diff(v) > 0):
        raise ValueError(f"{axis} breakpoints increase strictly and number at least 2")
    return v


def _refined(b, quadratic):
    """Node positions along an axis: the breakpoints, and for biquadratic
    cells the midpoints between them too."""
    if not quadratic:
        return b
    mid = 0.5 * (b[:-1] + b[1:])
    return np.insert(b, np.arange(1, b.size), mid)


# Biquadratic node k of a cell as an offset on the cell's 3 x 3 node
# lattice: corners counter-clockwise, face midpoints 0..3, centre.
_Q9_OFFSETS = ((0, 0), (2, 0), (2, 2), (0, 2), (1, 0), (2, 1), (1, 2), (0, 1), (1, 1))
_Q4_OFFSETS = ((0, 0), (1, 0), (1, 1), (0, 1))


def _mesh_on_lattice(x_nodes, z_nodes, quadratic, material, coords):
    """Cells of a node lattice with columns at x_nodes (nx,) and node
    heights z_nodes (nz, nx), row by row with x fastest; the outer faces
    tagged BOTTOM, RIGHT, TOP, LEFT."""
    if np.any(np.diff(z_nodes, axis=0) <= 0):
        raise ValueError("node heights must increase upwards in every column (no layer of zero thickness)")
    step = 2 if quadratic else 1
    nx_nodes = x_nodes.size
    ncx, ncz = (nx_nodes - 1) // step, (z_nodes.shape[0] - 1) // step
    nodes = np.column_stack([np.tile(x_nodes, z_nodes.shape[0]), z_nodes.ravel()])
    base_j, base_i = np.divmod(np.arange(ncx * ncz), ncx)
    offsets = _Q9_OFFSETS if quadratic else _Q4_OFFSETS
    cells = np.column_stack([(step * base_i + oi) + nx_nodes * (step * base_j + oj) for oi, oj in offsets])
    # boundary faces, listed cell by cell and face by face within a cell
    sides = [(np.arange(ncx), 0, BOTTOM), (np.arange(ncz) * ncx + ncx - 1, 1, RIGHT),
             (np.arange(ncx) + ncx * (ncz - 1), 2, TOP), (np.arange(ncz) * ncx, 3, LEFT)]
    faces = np.concatenate([np.column_stack([c, np.full(c.size, f), np.full(c.size, tag)]) for c, f, tag in sides])
    faces = faces[np.lexsort((faces[:, 1], faces[:, 0]))]
    return MacroMesh("quad9" if quadratic else "quad4", coords, nodes, cells.astype(np.int32),
                     np.asarray(material, dtype=np.int32), faces.astype(np.int32))


def rectangle(x, z, nx=None, nz=None, material=0, quadratic=False, coords="cartesian") -> MacroMesh:
    """A rectangle cut at the x and z breakpoints (or into nx by nz equal
    cells); material is one region id or a function of a cell's centre."""
    xb, zb = _axis_breaks(x, nx, "x"), _axis_breaks(z, nz, "z")
    x_nodes = _refined(xb, quadratic)
    z_nodes = np.repeat(_refined(zb, quadratic)[:, None], x_nodes.size, axis=1)
    n_cells = (xb.size - 1) * (zb.size - 1)
    if callable(material):
        xc, zc = 0.5 * (xb[:-1] + xb[1:]), 0.5 * (zb[:-1] + zb[1:])
        regions = np.array([material(xc[i], zc[j]) for j in range(zc.size) for i in range(xc.size)], dtype=np.int32)
    else:
        regions = np.full(n_cells, int(material), dtype=np.int32)
    return _mesh_on_lattice(x_nodes, z_nodes, quadratic, regions, coords)


def layered(x, interfaces, nz, regions, nx=None, quadratic=False, coords="cartesian") -> MacroMesh:
    """Layers between interfaces (bottom to top; each a constant or a
    function z(x)), each cut into nz rows whose node heights blend the two
    bounding interfaces linearly in every node column."""
    xb = _axis_breaks(x, nx, "x")
    n_layers = len(interfaces) - 1
    if n_layers < 1:
        raise ValueError("layers need at least two interfaces")
    if len(regions) != n_layers:
        raise ValueError(f"{n_layers} layers take {n_layers} region ids, not {len(regions)}")
    rows_per_layer = [int(nz)] * n_layers if np.isscalar(nz) else [int(n) for n in nz]
    if len(rows_per_layer) != n_layers or min(rows_per_layer) < 1:
        raise ValueError(f"nz is a positive integer or {n_layers} positive integers")
    x_nodes = _refined(xb, quadratic)

    def heights(iface):
        z = iface(x_nodes) if callable(iface) else iface
        return np.broadcast_to(np.asarray(z, dtype=float), x_nodes.shape).astype(float)

    levels = [heights(f) for f in interfaces]
    step = 2 if quadratic else 1
    z_rows: List[np.ndarray] = [levels[0]]
    for k in range(n_layers):
        n = step * rows_per_layer[k]
        for r in range(1, n + 1):
            t = r / n
            z_rows.append((1.0 - t) * levels[k] + t * levels[k + 1])
    ncx = xb.size - 1
    cell_regions = np.repeat(np.asarray(regions, dtype=np.int32), [ncx * n for n in rows_per_layer])
    return _mesh_on_lattice(x_nodes, np.vstack(z_rows), quadratic, cell_regions, coords)


@dataclass
class BoreholeMesh:
    mesh: MacroMesh
    regions: Dict[str, int]
    boundaries: Dict[str, int]
    r: np.ndarray
    z: np.ndarray


class _Grading:
    """Radial cell boundaries from r0 out to r1: cells growing by `growth`
    from `first`, capped at `cap`; a remainder of at most 1.5 cells becomes
    the last cell."""

    def __init__(self, first, growth, cap):
        self.first, self.growth, self.cap = first, growth, cap

    def breaks(self, r0, r1):
        out, r, size = [r0], r0, self.first
        while r1 - r > 1e-12 * max(1.0, abs(r1)):
            size = min(size, self.cap)
            if r1 - r <= 1.5 * size:
                out.append(r1)
                break
            r += size
            out.append(r)
            size *= self.growth
        return out


def _even_cells(lo, hi, size):
    """Breakpoints after lo that cut [lo, hi] into the fewest equal cells
    no longer than size (up to rounding)."""
    n = max(1, int(np.ceil((hi - lo) / size - 1e-9)))
    return list(np.linspace(lo, hi, n + 1)[1:]), (hi - lo) / n


def borehole(z, annuli, formation_radius, beds=(), dz=0.05, dr=None, growth=1.3, fluid_cells=2,
             quadratic=False) -> BoreholeMesh:
    """Axisymmetric borehole model: annuli (name, outer radius) around the
    axis -- the first is the fluid column of fluid_cells cells, the others
    cut into cells of at most dz -- then the formation, graded from the last
    annulus (cells growing by `growth` up to dr, default 4 dz) out to
    formation_radius, in horizontal beds (z_bottom, name) cut into cells of
    at most dz.  Region ids number the annulus and bed names from 1 in order
    of first appearance."""
    z_lo, z_hi = float(z[0]), float(z[1])
    if not annuli:
        raise ValueError("a borehole needs at least its fluid column among the annuli")
    outer = [float(r) for _, r in annuli]
    if any(b <= a for a, b in zip([0.0] + outer, outer)) or not formation_radius > outer[-1]:
        raise ValueError("annulus radii must increase from the axis and stay inside formation_radius")
    beds = list(beds) if beds else [(z_lo, "formation")]
    bottoms = [float(b[0]) for b in beds]
    if abs(bottoms[0] - z_lo) > 1e-12 * max(1.0, abs(z_lo)) or any(a >= b for a, b in zip(bottoms, bottoms[1:])):
        raise ValueError("beds start at the model's bottom and go upwards (increasing z_bottom)")
    dr = 4 * dz if dr is None else dr

    r_breaks = list(np.linspace(0.0, outer[0], fluid_cells + 1))
    last_cell = outer[0] / fluid_cells
    for inner, out in zip(outer, outer[1:]):
        more, last_cell = _even_cells(inner, out, dz)
        r_breaks += more
    r_breaks += _Grading(min(last_cell * growth, dr), growth, dr).breaks(outer[-1], formation_radius)[1:]

    z_breaks = [z_lo]
    for bottom, top in zip(bottoms, bottoms[1:] + [z_hi]):
        z_breaks += _even_cells(bottom, top, dz)[0]

    ids: Dict[str, int] = {}
    for name in [a for a, _ in annuli] + [b for _, b in beds]:
        ids.setdefault(name, len(ids) + 1)
    bed_bottoms = np.asarray(bottoms)

    def region_of(rc, zc):
        for name, r_out in annuli:
            if rc < r_out:
                return ids[name]
        return ids[beds[int(np.searchsorted(bed_bottoms, zc, side="right")) - 1][1]]

    m = rectangle(r_breaks, z_breaks, material=region_of, quadratic=quadratic, coords="axisymmetric")
    return BoreholeMesh(m, ids, {"bottom": BOTTOM, "outer": RIGHT, "top": TOP, "axis": LEFT}, np.asarray(r_breaks),
                        np.asarray(z_breaks))


__all__: Sequence[str] = ("BOTTOM", "RIGHT", "TOP", "LEFT", "rectangle", "layered", "borehole", "BoreholeMesh")
