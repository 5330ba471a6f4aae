"""Synthetic workloads of the FEM solve path (structured unit-cube meshes).

The reference reads Gmsh meshes and property files; it ships no generator.
These builders define the inputs of the benchmark configurations with the
library's own host utilities (mesh_box, field_*), so that every consumer
(tests, bench, the oracle) sees identical seeded data.

Seeds: 1 fields, 2 inclusions, 3 MsFEM cell sizes, 4 training fields.
"""
from __future__ import annotations

import numpy as np

from . import (box_face_nodes, box_face_tris, field_grf, field_inclusions, field_lognormal, mesh_box)

FACE_XM, FACE_XP, FACE_YM, FACE_YP, FACE_ZM, FACE_ZP = range(6)


def elastic_cube(n, E=1e9, nu=0.3, pressure=1e6, field=None, seed=1):
    """Unit cube, n^3 hexes x 6 tets; bottom face z=0 clamped, uniform
    pressure on the top face as traction (0,0,-pressure).

    field None: homogeneous E; "grf": two-phase thresholded Gaussian field
    (256 modes, correlation length 0.05), E in {1e9, 2e11}, material id = phase.
    """
    xyz, tets = mesh_box(n)
    if field is None:
        Ev = np.full(len(tets), float(E))
        material = np.zeros(len(tets), dtype=np.int32)
    elif field == "grf":
        Ev, material = field_grf(xyz, tets, 256, 0.05, seed, 1e9, 2e11)
    else:
        raise ValueError(field)
    bottom = box_face_nodes(n, n, n, FACE_ZM)
    top_tris = box_face_tris(n, n, n, FACE_ZP)
    traction = np.tile(np.array([0.0, 0.0, -pressure]), (len(top_tris), 1))
    return dict(problem="elastic", n=n, xyz=xyz, tets=tets, E=Ev, nu=nu, material=material, fixed_nodes=bottom,
                fixed_values=np.zeros((len(bottom), 3)), tris=top_tris, traction=traction)


def laplace_inclusions(n, count=200, rmin=0.02, rmax=0.06, seed=2, sig_bg=1.0, sig_in=100.0):
    """Unit cube scalar Laplace: permittivity sig_bg with `count` seeded
    spherical inclusions of sig_in; potential 0 at z=0 and 1 at z=1 (two
    Dirichlet tags, in that order, as the ElectroStatic driver lists them)."""
    xyz, tets = mesh_box(n)
    sigma = field_inclusions(tets, xyz, count, rmin, rmax, seed, sig_bg, sig_in)
    bottom = box_face_nodes(n, n, n, FACE_ZM)
    top = box_face_nodes(n, n, n, FACE_ZP)
    fixed = np.concatenate([bottom, top]).astype(np.int32)
    vals = np.concatenate([np.zeros(len(bottom)), np.ones(len(top))])
    return dict(problem="scalar", n=n, xyz=xyz, tets=tets, sigma=sigma, fixed_nodes=fixed, fixed_values=vals)


def cell_sizes(nc, n_lo=6, n_hi=10, seed=3):
    """Fine subdivision per coarse cell, drawn U{n_lo..n_hi} (stream 5)."""
    u = _uniforms(nc ** 3, seed, 5)
    return (n_lo + np.floor(u * (n_hi - n_lo + 1))).astype(np.int32).clip(n_lo, n_hi)


def _uniforms(count, seed, stream):
    """The library's splitmix64 uniform generator, vectorised in numpy."""
    idx = np.arange(count, dtype=np.uint64)
    return _splitmix_uniform(np.uint64(seed), np.uint64(stream), idx)


def _mix64(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _splitmix_uniform(seed, stream, idx):
    with np.errstate(over="ignore"):
        key = _mix64(seed * np.uint64(0x100000001B3) + stream)
        h = _mix64(key ^ (idx * np.uint64(0xD6E8FEB86659FD93)))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def msfem_cells(nc, n_lo=6, n_hi=10, n_fixed=None, seed_n=3, seed_E=1, mu=float(np.log(1e9)), sigma=1.0):
    """Coarse nc^3 grid on the unit cube.  Cell c (x fastest) has n_c fine
    hexes per edge, U{n_lo..n_hi} from seed_n unless n_fixed is given, and a
    lognormal Young's modulus exp(mu + sigma N(0,1)) per fine tet (stream
    1000 + c of seed_E, tets in mesh_box order)."""
    ncell = nc ** 3
    cell_n = np.full(ncell, n_fixed, np.int32) if n_fixed else cell_sizes(nc, n_lo, n_hi, seed_n)
    idx = np.arange(ncell)
    origin = np.stack([idx % nc, (idx // nc) % nc, idx // (nc * nc)], 1) / nc
    E = np.concatenate([field_lognormal(6 * int(n) ** 3, seed_E, 1000 + c, 0, mu, sigma)
                        for c, n in enumerate(cell_n)])
    return dict(nc=nc, cell_n=cell_n, origin=origin, h=1.0 / nc, E=E, nu=0.3)


def cell_E(p, c):
    off = int((6 * p["cell_n"][:c].astype(np.int64) ** 3).sum())
    return p["E"][off:off + 6 * int(p["cell_n"][c]) ** 3]


def msfem_global_E(p):
    """Per-tet modulus of the whole box's fine mesh mesh_box(nc*n) when every
    cell has the same n: global hex (I,J,K) lies in cell (I//n, J//n, K//n)
    at local hex (I%n, J%n, K%n), with the same 6-tet split."""
    nc = p["nc"]
    n = int(p["cell_n"][0])
    assert (p["cell_n"] == n).all(), "cells must share one subdivision"
    N = nc * n
    I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    I, J, K = I.T.ravel(), J.T.ravel(), K.T.ravel()  # hex order x fastest
    cell = (I // n) + nc * ((J // n) + nc * (K // n))
    local = (I % n) + n * ((J % n) + n * (K % n))
    idx = (cell[:, None] * 6 * n ** 3 + 6 * local[:, None] + np.arange(6)[None, :]).ravel()
    return p["E"][idx]


def convection_diffusion(m, peclet, seed=0):
    """Non-symmetric test system: upwind 5-point -Laplace + c.grad on an
    m x m interior grid of the unit square (CSR, Dirichlet 0 built in),
    c = (peclet, peclet/2), and a N(0,1) right-hand side from seed."""
    h = 1.0 / (m + 1)
    cx, cy = peclet, 0.5 * peclet
    rows, cols, vals = [], [], []
    for j in range(m):
        for i in range(m):
            r = i + m * j
            entries = {r: 4.0 / h ** 2 + (abs(cx) + abs(cy)) / h}
            for di, dj, coef in ((-1, 0, -1 / h ** 2 - max(cx, 0) / h), (1, 0, -1 / h ** 2 + min(cx, 0) / h),
                                 (0, -1, -1 / h ** 2 - max(cy, 0) / h), (0, 1, -1 / h ** 2 + min(cy, 0) / h)):
                ii, jj = i + di, j + dj
                if 0 <= ii < m and 0 <= jj < m:
                    entries[ii + m * jj] = coef
            for c in sorted(entries):
                rows.append(r)
                cols.append(c)
                vals.append(entries[c])
    n = m * m
    rowptr = np.searchsorted(np.array(rows), np.arange(n + 1)).astype(np.int32)
    b = np.random.default_rng(seed).normal(size=n)
    return rowptr, np.array(cols, np.int32), np.array(vals), b
