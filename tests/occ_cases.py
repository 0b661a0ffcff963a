"""Surfaces and grids shared by the occupancy tests."""
import numpy as np

from pyhexopt import meshgen as mg


def sphere(level=3, radius=1.0, centre=(0.0, 0.0, 0.0), flip=False):
    V, F = mg.icosphere(level)
    V = V * radius + np.asarray(centre)
    if flip:
        F = F[:, [0, 2, 1]]
    return V, F.astype(np.int32)


def merge(*meshes):
    Vs, Fs, off = [], [], 0
    for V, F in meshes:
        Vs.append(V)
        Fs.append(F + off)
        off += V.shape[0]
    return np.concatenate(Vs), np.concatenate(Fs).astype(np.int32)


def cube_aligned_grid():
    """Unit cube surface (4 x 4 quads per side) and a voxel grid whose centre lines pass
    through the cube's vertices, edges and diagonals in xy (x, y in {-0.25, ..., 1.25} step
    0.125), with no centre on z = 0 or z = 1."""
    V, F = mg.cube_surface(n=4, lo=0.0, hi=1.0)
    origin = np.array([-0.3125, -0.3125, -0.1])
    return V, F.astype(np.int32), dict(nx=13, ny=13, nz=11, origin=origin, spacing=0.125)


def voxel_centres(nx, ny, nz, origin, spacing):
    """Voxel centres (i + 0.5) * spacing + origin, shaped (nz, ny, nx, 3)."""
    x = (np.arange(nx) + 0.5) * spacing + origin[0]
    y = (np.arange(ny) + 0.5) * spacing + origin[1]
    z = (np.arange(nz) + 0.5) * spacing + origin[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([X, Y, Z], axis=-1)


def random_cells(rng, n, grid, extents=(1, 2, 4, 8)):
    """n cells (x, y, z, extent) aligned to their extent inside [0, grid)^3."""
    e = rng.choice(np.asarray(extents), size=n)
    k = np.stack([rng.integers(0, grid // ee, size=3) for ee in e])
    return np.concatenate([k * e[:, None], e[:, None]], axis=1).astype(np.int32)


def cell_centres(cells, origin, spacing):
    """Cell centres the way compute_sign(OctreeGrid) forms them."""
    c = cells.astype(np.float64)
    half = 0.5 * spacing * c[:, 3]
    cx = origin[0] + spacing * c[:, 0] + half
    cy = origin[1] + spacing * c[:, 1] + half
    cz = origin[2] + spacing * c[:, 2] + half
    return np.stack([cx, cy, cz], axis=1)
