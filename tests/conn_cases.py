"""Two-block cases cut from one block, for the interblock connection tests.

A distorted box grid is cut normal to k; each half is then re-indexed by one
of the 24 proper rotations of the index cube (an axis permutation with
reversals, determinant +1, so volumes stay positive).  The cut face of a
half lands on any of its six surfaces, and the two halves' patch axes meet
in any of the eight relative orientations, lower/upper, lower/lower and
upper/upper pairings included.  Because the ghost cells of a connection
continue the partner block, the residual of every cell must equal the uncut
block's to rounding, whatever the rotations.
"""
import itertools
import math
import os

import numpy as np

from aither_amd import cases

SURF_NAMES = ("imin", "imax", "jmin", "jmax", "kmin", "kmax")
FREESTREAM = dict(p=101325.0, rho=1.225, vel=(100.0, 20.0, 10.0))


def _parity(perm):
    p, s = list(perm), 1
    for a in range(3):
        for b in range(a + 1, 3):
            if p[a] > p[b]:
                s = -s
    return s


def rotations():
    """The 24 proper rotations as (perm, flips): new axis q is old axis
    perm[q], reversed when flips[q]."""
    out = []
    for perm in itertools.permutations(range(3)):
        for flips in itertools.product((False, True), repeat=3):
            if _parity(perm) * (-1) ** sum(flips) == 1:
                out.append((perm, flips))
    return out


def rotate_nodes(a, rot):
    perm, flips = rot
    b = a.transpose(perm)
    for q in range(3):
        if flips[q]:
            b = np.flip(b, axis=q)
    return np.ascontiguousarray(b)


def rotate_cell(cell, dims_old, rot):
    """old cell index (i, j, k) in a block of dims_old cells -> new index"""
    perm, flips = rot
    return tuple((dims_old[perm[q]] - 1 - cell[perm[q]]) if flips[q] else cell[perm[q]] for q in range(3))


def cut_surface(rot, lower):
    """surface type (1..6) that the old k-lower (lower=True) or k-upper face
    becomes under rot"""
    perm, flips = rot
    q = perm.index(2)
    lo = lower != flips[q]
    return 2 * q + (1 if lo else 2)


def base_grid(n=(6, 5, 8)):
    ni, nj, nk = n
    i, j, k = np.meshgrid(np.arange(ni + 1), np.arange(nj + 1), np.arange(nk + 1), indexing="ij")
    h = 0.1
    x = i * h + 0.02 * np.sin(0.9 * j) * np.cos(0.7 * k)
    y = j * h + 0.015 * np.sin(0.8 * k + 0.3 * i)
    z = k * h * (1.0 + 0.05 * np.sin(0.6 * i + 0.2 * j))
    return x, y, z


def state(xc, yc, zc):
    rho = 1.225 * (1.0 + 0.08 * np.sin(7.0 * xc + 3.0 * yc) * np.cos(5.0 * zc))
    u = 100.0 + 25.0 * np.sin(6.0 * yc + 2.0 * zc)
    v = 20.0 + 15.0 * np.cos(5.0 * xc - 4.0 * zc)
    w = 10.0 + 12.0 * np.sin(4.0 * xc + 6.0 * yc)
    p = 101325.0 * (1.0 + 0.05 * np.cos(6.0 * xc) * np.sin(5.0 * yc + 3.0 * zc))
    return rho, u, v, w, p


def _inp(name, nblocks_bcs, recon, timeint, solver="lusgs", sweeps=1, extra="", eq="euler"):
    fs = FREESTREAM
    return f"""gridName: {name}
equationSet: {eq}
timeIntegration: {timeint}
cflStart: 2.0
cflMax: 2.0
faceReconstruction: {recon}
limiter: vanAlbada
iterations: 10
referenceTemperature: 288.15
referenceDensity: 1.225
initialConditions: <icState(tag=-1; file={name}.cloud)>
matrixSolver: {solver}
matrixSweeps: {sweeps}
matrixRelaxation: 1.0
{extra}boundaryStates: <characteristic(tag=1; pressure={fs['p']!r}; density={fs['rho']!r}; velocity=[{fs['vel'][0]!r}, {fs['vel'][1]!r}, {fs['vel'][2]!r}])>
boundaryConditions: {len(nblocks_bcs)}
""" + "".join(nblocks_bcs)


def write_pair(outdir, name, rot0=None, rot1=None, cut=4, recon="thirdOrder", timeint="explicitEuler",
               solver="lusgs", sweeps=1, n=(6, 5, 8), eq="euler", scale=1.0):
    """Write the uncut case (rot0 is None) or the two-block case; returns
    (inp path, cell map).  The cell map takes an uncut cell (i, j, k) to
    (block, (i', j', k')) of the written case."""
    os.makedirs(outdir, exist_ok=True)
    x, y, z = (scale * a for a in base_grid(n))
    xc, yc, zc = cases.cell_centres(x, y, z)
    rho, u, v, w, p = state(xc, yc, zc)
    cases.write_cloud(os.path.join(outdir, name + ".cloud"), xc, yc, zc, rho, u, v, w, p)
    ni, nj, nk = n
    if rot0 is None:
        cases.write_plot3d(os.path.join(outdir, name + ".xyz"), [(x, y, z)])
        bcs = [cases._surfaces_box(ni, nj, nk, {s: ("characteristic", 1) for s in SURF_NAMES})]

        def cmap(c):
            return 0, tuple(c)
    else:
        halves = [tuple(a[:, :, :cut + 1] for a in (x, y, z)), tuple(a[:, :, cut:] for a in (x, y, z))]
        dims = [(ni, nj, cut), (ni, nj, nk - cut)]
        rots = [rot0, rot1]
        blocks, newdims = [], []
        for h, rot, d in zip(halves, rots, dims):
            blocks.append(tuple(rotate_nodes(a, rot) for a in h))
            newdims.append(tuple(d[rot[0][q]] for q in range(3)))
        cases.write_plot3d(os.path.join(outdir, name + ".xyz"), blocks)
        s0 = cut_surface(rot0, lower=False)  # block 0's cut is its old k-upper face
        s1 = cut_surface(rot1, lower=True)
        bcs = []
        for b, (sb, sp) in enumerate(((s0, s1), (s1, s0))):
            types = {s: ("characteristic", 1) for s in SURF_NAMES}
            types[SURF_NAMES[sb - 1]] = ("interblock", 1000 * sp + (1 - b))
            bcs.append(cases._surfaces_box(*newdims[b], types))

        def cmap(c):
            i, j, k = c
            if k < cut:
                return 0, rotate_cell((i, j, k), dims[0], rot0)
            return 1, rotate_cell((i, j, k - cut), dims[1], rot1)
    path = os.path.join(outdir, name + ".inp")
    with open(path, "w") as f:
        f.write(_inp(name, bcs, recon, timeint, solver, sweeps, eq=eq))
    return path, cmap


def orientation_set():
    """(rot0, rot1) pairs covering every surface type on both sides and every
    relative orientation of the patch axes"""
    rots = rotations()
    seen, out = set(), []
    for r0 in rots:
        for r1 in rots:
            s0, s1 = cut_surface(r0, False), cut_surface(r1, True)
            # relative orientation: how block 1's patch axes map onto block 0's
            key = (s0, s1, _relative(r0, r1))
            if key not in seen:
                seen.add(key)
                out.append((r0, r1))
    return out


def _relative(r0, r1):
    # the in-plane old axes are i (0) and j (1); record where each lands and
    # whether it is reversed, on both sides
    def inplane(rot):
        perm, flips = rot
        return tuple((perm.index(a), flips[perm.index(a)]) for a in (0, 1))
    return inplane(r0), inplane(r1)
