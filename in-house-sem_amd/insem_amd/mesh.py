"""Structured macro-mesh templates (reference: python/insem/mesh/templates.py:1-306).

Same node lattices, cell order (row by row, x fastest), boundary tags and
region ids as the reference templates, so a model built here numbers its
elements -- and therefore sums its assembly -- exactly as the reference does.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from .core import MacroMesh

BOTTOM = 1  # z = z_min
RIGHT = 2   # x = x_max
TOP = 3     # z = z_max
LEFT = 4    # x = x_min


def _breaks(spec, n, name):
    v = np.asarray(spec, dtype=float)
    if n is None:
        if v.ndim != 1 or v.size < 2 or np.any(np.diff(v) <= 0):
            raise ValueError(f"{name} breakpoints must be strictly increasing with at least 2 entries")
        return v
    if v.shape != (2,) or v[1] <= v[0]:
        raise ValueError(f"{name} must be (min, max) with min < max when n{name} is given")
    if n < 1:
        raise ValueError(f"n{name} must be >= 1")
    return np.linspace(v[0], v[1], n + 1)


def _with_midpoints(b, quadratic):
    if not quadratic:
        return b
    out = np.empty(2 * b.size - 1)
    out[0::2] = b
    out[1::2] = 0.5 * (b[:-1] + b[1:])
    return out


def _lattice_mesh(xl, zl, quadratic, material, coords):
    """Cells of a node lattice: xl (mx,) columns, zl (mz, mx) heights."""
    st = 2 if quadratic else 1
    mx = xl.size
    ncx, ncz = (mx - 1) // st, (zl.shape[0] - 1) // st
    if np.any(np.diff(zl, axis=0) <= 0):
        raise ValueError("layer interfaces must be strictly increasing in z at every x (no zero-thickness layers)")
    nodes = np.column_stack([np.broadcast_to(xl, zl.shape).ravel(), zl.ravel()])
    jj, ii = np.meshgrid(np.arange(ncz), np.arange(ncx), indexing="ij")
    a, b = (st * ii).ravel(), (st * jj).ravel()
    node = lambda i, j: i + mx * j  # noqa: E731
    if quadratic:
        cells = np.column_stack([node(a, b), node(a + 2, b), node(a + 2, b + 2), node(a, b + 2), node(a + 1, b),
                                 node(a + 2, b + 1), node(a + 1, b + 2), node(a, b + 1), node(a + 1, b + 1)])
    else:
        cells = np.column_stack([node(a, b), node(a + 1, b), node(a + 1, b + 1), node(a, b + 1)])
    bnd = []
    for j in range(ncz):
        for i in range(ncx):
            c = i + ncx * j
            if j == 0:
                bnd.append((c, 0, BOTTOM))
            if i == ncx - 1:
                bnd.append((c, 1, RIGHT))
            if j == ncz - 1:
                bnd.append((c, 2, TOP))
            if i == 0:
                bnd.append((c, 3, LEFT))
    return MacroMesh("quad9" if quadratic else "quad4", coords, nodes, cells.astype(np.int32),
                     np.asarray(material, dtype=np.int32), np.asarray(bnd, dtype=np.int32).reshape(-1, 3))


def rectangle(x, z, nx=None, nz=None, material=0, quadratic=False, coords="cartesian") -> MacroMesh:
    """Structured mesh of a rectangle (templates.py:99-135)."""
    xs, zs = _breaks(x, nx, "x"), _breaks(z, nz, "z")
    ncx, ncz = xs.size - 1, zs.size - 1
    xl = _with_midpoints(xs, quadratic)
    zl = np.repeat(_with_midpoints(zs, quadratic)[:, None], xl.size, axis=1)
    if callable(material):
        xc, zc = 0.5 * (xs[:-1] + xs[1:]), 0.5 * (zs[:-1] + zs[1:])
        mat = np.array([material(xc[i], zc[j]) for j in range(ncz) for i in range(ncx)], dtype=np.int32)
    else:
        mat = np.broadcast_to(np.asarray(material, dtype=np.int32), (ncx * ncz,)).copy()
    return _lattice_mesh(xl, zl, quadratic, mat, coords)


def layered(x, interfaces, nz, regions, nx=None, quadratic=False, coords="cartesian") -> MacroMesh:
    """Layers between interfaces (constants or z(x) functions), node heights
    blended linearly per column (templates.py:138-196)."""
    xs = _breaks(x, nx, "x")
    k = len(interfaces) - 1
    if k < 1:
        raise ValueError("need at least two interfaces (one layer)")
    if len(regions) != k:
        raise ValueError(f"{k} layers need {k} region ids, got {len(regions)}")
    counts = [int(nz)] * k if np.isscalar(nz) else [int(n) for n in nz]
    if len(counts) != k or min(counts) < 1:
        raise ValueError(f"nz must be a positive int or {k} positive ints")
    xl = _with_midpoints(xs, quadratic)

    def height(f):
        if callable(f):
            return np.broadcast_to(np.asarray(f(xl), dtype=float), xl.shape).astype(float)
        return np.full(xl.shape, float(f))

    lev = [height(f) for f in interfaces]
    st = 2 if quadratic else 1
    rows = [lev[0]]
    for layer in range(k):
        nr = st * counts[layer]
        for r in range(1, nr + 1):
            t = r / nr
            rows.append((1.0 - t) * lev[layer] + t * lev[layer + 1])
    ncx = xs.size - 1
    mat = np.concatenate([np.full(ncx * counts[layer], regions[layer], dtype=np.int32) for layer in range(k)])
    return _lattice_mesh(xl, np.vstack(rows), quadratic, mat, coords)


@dataclass
class BoreholeMesh:
    mesh: MacroMesh
    regions: Dict[str, int]
    boundaries: Dict[str, int]
    r: np.ndarray
    z: np.ndarray


def _graded(r0, r1, first, growth, max_size):
    out, r, size = [r0], r0, first
    while r1 - r > 1e-12 * max(1.0, abs(r1)):
        size = min(size, max_size)
        if r1 - r <= 1.5 * size:
            out.append(r1)
            break
        r += size
        out.append(r)
        size *= growth
    return out


def borehole(z, annuli, formation_radius, beds=(), dz=0.05, dr=None, growth=1.3, fluid_cells=2,
             quadratic=False) -> BoreholeMesh:
    """Axisymmetric borehole: annuli around the axis, then a radially graded
    formation in horizontal beds (templates.py:227-306)."""
    z0, z1 = float(z[0]), float(z[1])
    if not annuli:
        raise ValueError("borehole needs at least the fluid column in `annuli`")
    radii = [float(r) for _, r in annuli]
    if any(b <= a for a, b in zip([0.0] + radii, radii)) or formation_radius <= radii[-1]:
        raise ValueError("annulus radii must increase and stay below formation_radius")
    beds = list(beds) or [(z0, "formation")]
    if abs(beds[0][0] - z0) > 1e-12 * max(1.0, abs(z0)) or any(b[0] >= c[0] for b, c in zip(beds, beds[1:])):
        raise ValueError("beds must start at z_min and be sorted by increasing z_bottom")
    dr = 4 * dz if dr is None else dr
    rb = list(np.linspace(0.0, radii[0], fluid_cells + 1))
    last = radii[0] / fluid_cells
    for inner, outer in zip(radii, radii[1:]):
        n = max(1, int(np.ceil((outer - inner) / dz - 1e-9)))
        rb += list(np.linspace(inner, outer, n + 1)[1:])
        last = (outer - inner) / n
    rb += _graded(radii[-1], formation_radius, min(last * growth, dr), growth, dr)[1:]
    tops = [b[0] for b in beds[1:]] + [z1]
    zb = [z0]
    for bottom, top in zip([b[0] for b in beds], tops):
        n = max(1, int(np.ceil((top - bottom) / dz - 1e-9)))
        zb += list(np.linspace(bottom, top, n + 1)[1:])
    names: Dict[str, int] = {}
    for name in [a for a, _ in annuli] + [b for _, b in beds]:
        names.setdefault(name, len(names) + 1)
    bottoms = np.array([b[0] for b in beds])

    def region(rc, zc):
        for name, outer in annuli:
            if rc < outer:
                return names[name]
        return names[beds[int(np.searchsorted(bottoms, zc, side="right")) - 1][1]]

    m = rectangle(rb, zb, material=region, quadratic=quadratic, coords="axisymmetric")
    return BoreholeMesh(m, names, {"bottom": BOTTOM, "outer": RIGHT, "top": TOP, "axis": LEFT}, np.asarray(rb),
                        np.asarray(zb))
