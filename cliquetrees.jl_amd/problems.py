"""Deterministic generators for the workloads the engine is measured on.

Every matrix is regenerated from its seed (numpy PCG64), never stored:
  * 2-D 5-point Dirichlet Laplacian (diag 4, -1 per grid neighbour),
  * 3-D 7-point weighted graph Laplacian, w_e ~ U[0.5, 1.5], plus shift*I,
  * 3-D 27-point weighted stencil, diagonal = sum of weights + shift,
  * KKT matrices [H A'; A -reg*I] with H a 3-D 7-point Laplacian,
  * batches of random-weight 2-D / 3-D grid Laplacians of uneven size.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _grid_edges(nx, ny, nz, offsets):
    """Edges (u, v) with u < v of a box grid for the given neighbour offsets."""
    idx = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    us, vs = [], []
    for dx, dy, dz in offsets:
        xs = slice(max(0, -dx), nx - max(0, dx))
        ys = slice(max(0, -dy), ny - max(0, dy))
        zs = slice(max(0, -dz), nz - max(0, dz))
        xd = slice(max(0, dx), nx - max(0, -dx))
        yd = slice(max(0, dy), ny - max(0, -dy))
        zd = slice(max(0, dz), nz - max(0, -dz))
        a = idx[zs, ys, xs].ravel()
        b = idx[zd, yd, xd].ravel()
        us.append(a)
        vs.append(b)
    u = np.concatenate(us)
    v = np.concatenate(vs)
    lo = np.minimum(u, v)
    hi = np.maximum(u, v)
    return lo, hi


OFF7 = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
OFF27 = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
         if (dz, dy, dx) > (0, 0, 0)]


def _laplacian(n, lo, hi, w, shift, dirichlet_diag=None):
    deg = np.zeros(n)
    np.add.at(deg, lo, w)
    np.add.at(deg, hi, w)
    diag = deg + shift if dirichlet_diag is None else np.full(n, float(dirichlet_diag))
    rows = np.concatenate([lo, hi, np.arange(n)])
    cols = np.concatenate([hi, lo, np.arange(n)])
    vals = np.concatenate([-w, -w, diag])
    A = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    A.sort_indices()
    return A


def laplacian2d_dirichlet(nx, ny):
    """2-D 5-point Dirichlet Laplacian: diag 4, -1 per neighbour (SPD)."""
    lo, hi = _grid_edges(nx, ny, 1, [(1, 0, 0), (0, 1, 0)])
    return _laplacian(nx * ny, lo, hi, np.ones(len(lo)), 0.0, dirichlet_diag=4.0)


def laplacian3d_7pt(nx, ny, nz, seed, shift=1e-3, wlo=0.5, whi=1.5):
    """3-D 7-point graph Laplacian with i.i.d. U[wlo, whi] edge weights + shift*I."""
    rng = np.random.default_rng(seed)
    lo, hi = _grid_edges(nx, ny, nz, OFF7)
    w = rng.uniform(wlo, whi, size=len(lo))
    return _laplacian(nx * ny * nz, lo, hi, w, shift)


def stencil3d_27pt(nx, ny, nz, seed, shift=1e-3, wlo=0.5, whi=1.5):
    """3-D 27-point stencil, weights U[wlo, whi] on the 26-neighbourhood,
    diagonal = sum of incident weights + shift (strictly diagonally dominant)."""
    rng = np.random.default_rng(seed)
    lo, hi = _grid_edges(nx, ny, nz, OFF27)
    w = rng.uniform(wlo, whi, size=len(lo))
    return _laplacian(nx * ny * nz, lo, hi, w, shift)


def grid_laplacian_random(dims, seed, shift=1e-3):
    """Random-weight grid Laplacian for a 2-D (nx, ny) or 3-D (nx, ny, nz) grid."""
    if len(dims) == 2:
        nx, ny, nz = dims[0], dims[1], 1
        offs = [(1, 0, 0), (0, 1, 0)]
    else:
        nx, ny, nz = dims
        offs = OFF7
    rng = np.random.default_rng(seed)
    lo, hi = _grid_edges(nx, ny, nz, offs)
    w = rng.uniform(0.5, 1.5, size=len(lo))
    return _laplacian(nx * ny * nz, lo, hi, w, shift)


def near_cubic_dims(n, dim):
    """Grid dims with product close to n, as equal as possible."""
    if dim == 2:
        a = max(1, int(round(np.sqrt(n))))
        b = max(1, int(round(n / a)))
        return (a, b)
    a = max(1, int(round(n ** (1.0 / 3.0))))
    b = max(1, int(round(np.sqrt(n / a))))
    c = max(1, int(round(n / (a * b))))
    return (a, b, c)


def batch_specs(count, seed, nmin=50, nmax=2000):
    """Sizes and grid dims of the uneven batch: n ~ U{nmin..nmax}, half 2-D, half 3-D."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=count)
    specs = []
    for k, n in enumerate(ns):
        dim = 2 if k % 2 == 0 else 3
        specs.append(near_cubic_dims(int(n), dim))
    return specs


def kkt_matrix(nh_dims, seed, reg=1e-6, per_row=3):
    """Quasi-definite KKT matrix [H A'; A -reg*I]: H a 3-D 7-point Laplacian
    (weights U[0.5,1.5], shift 1e-3), A m x n_H with m = n_H // 4 and
    ``per_row`` N(0,1) entries per row at uniform random columns."""
    nx, ny, nz = nh_dims
    H = laplacian3d_7pt(nx, ny, nz, seed)
    nh = H.shape[0]
    m = max(1, nh // 4)
    rng = np.random.default_rng(seed + 7919)
    rows = np.repeat(np.arange(m), per_row)
    cols = rng.integers(0, nh, size=m * per_row)
    vals = rng.standard_normal(m * per_row)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(m, nh))
    A.sum_duplicates()
    K = sp.bmat([[H, A.T], [A, -reg * sp.identity(m)]], format="csc")
    K.sort_indices()
    return K, nh, m


def kkt_value_update(K, nh, seed):
    """One value update of a KKT pattern: diag(H) rescaled by U[0.1, 10]."""
    rng = np.random.default_rng(seed)
    K = K.copy().tocsc()
    s = rng.uniform(0.1, 10.0, size=nh)
    d = K.diagonal()
    d[:nh] = d[:nh] * s
    K.setdiag(d)
    K.sort_indices()
    return K
