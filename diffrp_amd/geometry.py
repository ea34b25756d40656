"""
Mesh inputs of the rasterisation path: icosphere, normals, face soup, seeded noise.

These are the input generators diffrp's scenes are built from
(``diffrp/scene/objects.py:71-98``, ``diffrp/utils/geometry.py:5-65``).
The icosphere replaces ``trimesh.creation.icosphere`` (used by the README scene,
``README.md:53-63``): a subdivided regular icosahedron whose vertices are pushed
onto the sphere after every subdivision step.  Vertex *order* differs from
trimesh's; vertex *positions* (as a set) are the same construction, which is all
coverage and shading depend on.

Face ordering is hierarchical: the four children of face ``f`` are faces
``4f .. 4f+3``, so triangles adjacent in index are adjacent on the sphere.
"""
import math
from typing import Tuple

import numpy as np
import torch


def icosahedron() -> Tuple[np.ndarray, np.ndarray]:
    """Unit-sphere regular icosahedron (float64 verts [12,3], int64 faces [20,3], CCW outward).

    Vertices are the cyclic permutations of (0, ±1, ±phi), normalised.
    Faces are every triple of mutually adjacent vertices (edge length 2 before
    normalisation), oriented so the geometric normal points away from the origin.
    """
    phi = (1.0 + math.sqrt(5.0)) / 2.0
    v = []
    for a in (-1.0, 1.0):
        for b in (-phi, phi):
            v.append((0.0, a, b))
            v.append((a, b, 0.0))
            v.append((b, 0.0, a))
    v = np.array(v, dtype=np.float64)
    d = np.linalg.norm(v[:, None] - v[None], axis=-1)
    adj = np.abs(d - 2.0) < 1e-9
    faces = []
    for i in range(12):
        for j in range(i + 1, 12):
            if not adj[i, j]:
                continue
            for k in range(j + 1, 12):
                if adj[i, k] and adj[j, k]:
                    n = np.cross(v[j] - v[i], v[k] - v[i])
                    faces.append((i, j, k) if np.dot(n, v[i] + v[j] + v[k]) > 0 else (i, k, j))
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    return v, np.array(faces, dtype=np.int64)


def _subdivide(verts: np.ndarray, faces: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Split every triangle into four at its edge midpoints (shared midpoints deduplicated)."""
    nv = len(verts)
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])  # [3F, 2]
    key = np.minimum(e[:, 0], e[:, 1]) * nv + np.maximum(e[:, 0], e[:, 1])
    uniq, inv = np.unique(key, return_inverse=True)
    a = uniq // nv
    b = uniq % nv
    mid = (verts[a] + verts[b]) * 0.5
    f = len(faces)
    m01, m12, m20 = (inv[:f] + nv), (inv[f:2 * f] + nv), (inv[2 * f:] + nv)
    v0, v1, v2 = faces[:, 0], faces[:, 1], faces[:, 2]
    children = np.stack([
        np.stack([v0, m01, m20], -1),
        np.stack([m01, v1, m12], -1),
        np.stack([m20, m12, v2], -1),
        np.stack([m01, m12, m20], -1),
    ], 1).reshape(-1, 3)
    return np.concatenate([verts, mid]), children


def icosphere(subdivisions: int = 3, radius: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    """Icosphere with ``10*4^s+2`` vertices and ``20*4^s`` faces (float64 verts, int64 faces).

    Same construction as ``trimesh.creation.icosphere`` (the README scene's mesh):
    after each subdivision every vertex is moved radially onto the sphere.
    """
    v, f = icosahedron()
    v = v * radius
    for _ in range(subdivisions):
        v, f = _subdivide(v, f)
        n = np.linalg.norm(v, axis=-1, keepdims=True)
        v = v + (v / n) * (radius - n)
    return v, f


def compute_face_normals(verts: torch.Tensor, faces: torch.Tensor, normalize: bool = False) -> torch.Tensor:
    """cross(v1 - v0, v2 - v0) per face, CCW (``diffrp/utils/geometry.py:26-41``)."""
    fl = faces.long()
    n = torch.linalg.cross(verts[fl[:, 1]] - verts[fl[:, 0]], verts[fl[:, 2]] - verts[fl[:, 0]])
    return torch.nn.functional.normalize(n, dim=-1) if normalize else n


def compute_vertex_normals(verts: torch.Tensor, faces: torch.Tensor) -> torch.Tensor:
    """Normalised sum of unnormalised incident face normals (``diffrp/utils/geometry.py:44-65``)."""
    fn = compute_face_normals(verts, faces)
    fl = faces.long()
    vn = torch.zeros_like(verts)
    vn = vn.index_add(0, fl[:, 0], fn).index_add(0, fl[:, 1], fn).index_add(0, fl[:, 2], fn)
    return torch.nn.functional.normalize(vn, dim=-1)


def make_face_soup(verts: torch.Tensor, tris: torch.Tensor, face_normals: torch.Tensor):
    """Flat-shading soup: vertex ``3f+i`` is corner ``i`` of face ``f`` (``diffrp/utils/geometry.py:5-23``)."""
    fl = tris.long()
    return (
        verts[fl].reshape(-1, 3),
        torch.arange(fl.numel(), device=tris.device, dtype=torch.int32).reshape(tris.shape),
        face_normals[:, None, :].expand(-1, 3, -1).reshape(-1, 3),
    )


def noisy_icosphere(subdivisions: int, radius: float, sigma: float, seed: int = 0):
    """Icosphere as float32 torch tensors with ``N(0, sigma)`` vertex noise from a CPU generator."""
    v, f = icosphere(subdivisions, radius)
    verts = torch.from_numpy(v).float()
    if sigma > 0:
        g = torch.Generator().manual_seed(seed)
        verts = verts + sigma * torch.randn(verts.shape, generator=g, dtype=torch.float32)
    return verts.contiguous(), torch.from_numpy(f).int().contiguous()


def fibonacci_directions(n: int) -> np.ndarray:
    """``n`` unit directions on a Fibonacci sphere: y_i = 1-(2i+1)/n, phi_i = i*pi*(3-sqrt 5)."""
    i = np.arange(n, dtype=np.float64)
    y = 1.0 - (2.0 * i + 1.0) / n
    r = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(phi), y, r * np.sin(phi)], -1)
