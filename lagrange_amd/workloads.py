"""Deterministic synthetic meshes and query sets for the winding-number workloads.

Everything here is seeded through one counter-based generator (SplitMix64 finaliser over
``seed``, ``stream`` and element index) so that the same bytes come out on any machine
with the same numpy.  The reference benchmark draws uniform points in a mesh's bounding
box (modules/winding/tests/test_fast_winding_number.cpp:92-106); the workloads below
follow that pattern at larger scale:

* ``icosphere(nu)``            closed geodesic sphere, 20*nu^2 triangles.
* ``displaced_sphere(...)``    radius 1 displaced by a seeded sum of spherical harmonics.
* ``torus(...)``               parametric torus, optionally displaced.
* ``grid_queries`` / ``uniform_queries`` / ``near_surface_queries``.

Positions are returned as float64 (the generator's precision); callers convert to
float32 exactly as the engine does (FastWindingNumber.cpp:40-45).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        z ^= z >> np.uint64(30)
        z *= _C1
        z ^= z >> np.uint64(27)
        z *= _C2
        z ^= z >> np.uint64(31)
    return z


def rng_u64(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n 64-bit words of the counter-based stream (seed, stream)."""
    base = _mix64(np.array([(seed * 0x100000001B3 + stream * 0x5851F42D4C957F2D) & 0xFFFFFFFFFFFFFFFF],
                           dtype=np.uint64))[0]
    idx = np.arange(offset, offset + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = base + (idx + np.uint64(1)) * _GOLD
    return _mix64(z)


def rng_uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n float64 values uniform in [0, 1) with a 24-bit mantissa (exactly representable in fp32)."""
    return (rng_u64(seed, stream, n, offset) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


# ----------------------------------------------------------------------------------------
# meshes
# ----------------------------------------------------------------------------------------

@dataclass
class Mesh:
    vertices: np.ndarray  # (V, 3) float64
    facets: np.ndarray    # (F, 3) int64
    name: str = ""

    @property
    def num_vertices(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def num_facets(self) -> int:
        return int(self.facets.shape[0])

    def bbox(self):
        used = np.unique(self.facets.reshape(-1)) if self.num_facets else np.zeros(0, np.int64)
        v = self.vertices[used] if used.size else self.vertices
        return v.min(axis=0), v.max(axis=0)

    def f32(self):
        """float32 positions and int32 indices, the engine's internal copy."""
        return (np.ascontiguousarray(self.vertices, dtype=np.float32),
                np.ascontiguousarray(self.facets, dtype=np.int32))


def _icosahedron():
    phi = (1.0 + math.sqrt(5.0)) / 2.0
    v = []
    for a in (-1.0, 1.0):
        for b in (-phi, phi):
            v.append((0.0, a, b))
            v.append((a, b, 0.0))
            v.append((b, 0.0, a))
    v = np.array(v, dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    # faces = all triples of mutually nearest neighbours, oriented outward
    d = np.linalg.norm(v[:, None, :] - v[None, :, :], axis=2)
    edge = d[d > 1e-9].min()
    adj = np.abs(d - edge) < 1e-9
    faces = []
    for i in range(12):
        for j in range(i + 1, 12):
            if not adj[i, j]:
                continue
            for k in range(j + 1, 12):
                if adj[i, k] and adj[j, k]:
                    n = np.cross(v[j] - v[i], v[k] - v[i])
                    faces.append((i, j, k) if np.dot(n, v[i] + v[j] + v[k]) > 0 else (i, k, j))
    return v, np.array(faces, dtype=np.int64)


def icosphere(nu: int, radius: float = 1.0) -> Mesh:
    """Geodesic sphere: every icosahedron face split into nu^2 triangles, projected to the sphere.

    Vertex count 10*nu^2+2, facet count 20*nu^2.  nu = 2^k equals k rounds of midpoint
    subdivision followed by projection.  Outward orientation.
    """
    cv, cf = _icosahedron()
    ii, jj = np.meshgrid(np.arange(nu + 1), np.arange(nu + 1), indexing="ij")
    mask = (ii + jj) <= nu
    ii = ii[mask]
    jj = jj[mask]
    nloc = ii.size
    # local lattice point (i, j) of face (A, B, C) has integer weights (nu-i-j, i, j)
    w = np.stack([nu - ii - jj, ii, jj], axis=1)  # (nloc, 3)
    corners = cf[:, None, :].repeat(nloc, axis=1)  # (20, nloc, 3)
    weights = np.broadcast_to(w[None], corners.shape).copy()
    # canonical key: sort corner ids, zero weights map to corner id 12 (sentinel)
    cid = np.where(weights > 0, corners, 12)
    order = np.argsort(cid, axis=2, kind="stable")
    cid_s = np.take_along_axis(cid, order, axis=2)
    w_s = np.take_along_axis(np.where(weights > 0, weights, 0), order, axis=2)
    base = nu + 1
    key = (((cid_s[..., 0] * 13 + cid_s[..., 1]) * 13 + cid_s[..., 2]) * base + w_s[..., 0]) * base + w_s[..., 1]
    key = key.reshape(-1)
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    # positions from the canonical (sorted) representation so duplicates coincide exactly
    cs = cid_s.reshape(-1, 3)[first]
    ws = w_s.reshape(-1, 3)[first].astype(np.float64)
    cvx = np.vstack([cv, np.zeros((1, 3))])
    p = (cvx[cs[:, 0]] * ws[:, 0:1] + cvx[cs[:, 1]] * ws[:, 1:2] + cvx[cs[:, 2]] * ws[:, 2:3]) / nu
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    p *= radius
    gid = inv.reshape(20, nloc)
    # local index lookup for (i, j)
    lut = -np.ones((nu + 1, nu + 1), dtype=np.int64)
    lut[ii, jj] = np.arange(nloc)
    tris = []
    a_i, a_j = np.meshgrid(np.arange(nu), np.arange(nu), indexing="ij")
    up = (a_i + a_j) <= nu - 1
    ui, uj = a_i[up], a_j[up]
    t_up = np.stack([lut[ui, uj], lut[ui + 1, uj], lut[ui, uj + 1]], axis=1)
    dn = (a_i + a_j) <= nu - 2
    di, dj = a_i[dn], a_j[dn]
    t_dn = np.stack([lut[di + 1, dj], lut[di + 1, dj + 1], lut[di, dj + 1]], axis=1)
    loc = np.vstack([t_up, t_dn])
    tris = gid[:, loc].reshape(-1, 3)
    return Mesh(p, tris.astype(np.int64), name=f"icosphere_nu{nu}")


def _sph_legendre_norm(lmax: int, x: np.ndarray) -> dict:
    """Fully normalised associated Legendre functions Pbar_l^m(x) for 0<=m<=l<=lmax."""
    out = {}
    s = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    for m in range(lmax + 1):
        # P_m^m = (2m-1)!! s^m, unnormalised, then normalise by sqrt((2l+1)/(4pi) (l-m)!/(l+m)!)
        pmm = np.ones_like(x)
        for k in range(1, m + 1):
            pmm = pmm * (2 * k - 1) * s
        prev2 = None
        prev = pmm
        for l in range(m, lmax + 1):
            if l == m:
                cur = pmm
            elif l == m + 1:
                cur = x * (2 * m + 1) * pmm
            else:
                cur = ((2 * l - 1) * x * prev - (l + m - 1) * prev2) / (l - m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
            out[(l, m)] = cur * norm
            if l > m:
                prev2, prev = prev, cur
            else:
                prev2, prev = None, cur
    return out


def harmonic_field(dirs: np.ndarray, seed: int, n_terms: int = 32, lmax: int = 16) -> np.ndarray:
    """Seeded sum of n_terms real spherical harmonics (degree 1..lmax, random order and phase)."""
    u = rng_uniform(seed, 101, 3 * n_terms).reshape(n_terms, 3)
    ls = 1 + np.minimum((u[:, 0] * lmax).astype(np.int64), lmax - 1)
    ms = np.minimum((u[:, 1] * (ls + 1)).astype(np.int64), ls)
    phases = u[:, 2] * 2.0 * math.pi
    z = np.clip(dirs[:, 2], -1.0, 1.0)
    az = np.arctan2(dirs[:, 1], dirs[:, 0])
    P = _sph_legendre_norm(int(ls.max()), z)
    s = np.zeros(dirs.shape[0])
    for l, m, ph in zip(ls, ms, phases):
        s += P[(int(l), int(m))] * np.cos(m * az + ph)
    return s


def displaced_sphere(nu: int, seed: int, amplitude: float = 0.15, n_terms: int = 32,
                     lmax: int = 16, delete_fraction: float = 0.0) -> Mesh:
    """Sphere of radius 1 displaced radially by r = 1 + amplitude * S / max|S|.

    With delete_fraction > 0, that fraction of facets is removed (seeded), which opens
    holes and gives fractional winding numbers.
    """
    m = icosphere(nu)
    dirs = m.vertices / np.linalg.norm(m.vertices, axis=1, keepdims=True)
    s = harmonic_field(dirs, seed, n_terms, lmax)
    r = 1.0 + amplitude * s / np.abs(s).max()
    v = dirs * r[:, None]
    f = m.facets
    if delete_fraction > 0.0:
        keys = rng_u64(seed, 202, f.shape[0])
        order = np.argsort(keys, kind="stable")
        ndel = int(round(delete_fraction * f.shape[0]))
        keep = np.sort(order[ndel:])
        f = f[keep]
    return Mesh(v, f, name=f"displaced_sphere_nu{nu}_s{seed}")


def torus(n_major: int, n_minor: int, R: float = 1.0, r: float = 0.3, seed: int = 0,
          amplitude: float = 0.0) -> Mesh:
    """Parametric torus (outward orientation), 2*n_major*n_minor triangles.

    With amplitude > 0 the tube radius is modulated by a seeded product of sinusoids.
    """
    u = np.arange(n_major) * (2.0 * math.pi / n_major)
    v = np.arange(n_minor) * (2.0 * math.pi / n_minor)
    U, Vv = np.meshgrid(u, v, indexing="ij")
    rr = np.full(U.shape, r)
    if amplitude > 0.0:
        c = rng_uniform(seed, 303, 4)
        k1, k2 = 1 + int(c[0] * 6), 1 + int(c[1] * 4)
        rr = r * (1.0 + amplitude * np.sin(k1 * U + 2 * math.pi * c[2]) * np.sin(k2 * Vv + 2 * math.pi * c[3]))
    x = (R + rr * np.cos(Vv)) * np.cos(U)
    y = (R + rr * np.cos(Vv)) * np.sin(U)
    z = rr * np.sin(Vv)
    verts = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(n_major), np.arange(n_minor), indexing="ij")
    i2 = (i + 1) % n_major
    j2 = (j + 1) % n_minor
    a = (i * n_minor + j).reshape(-1)
    b = (i2 * n_minor + j).reshape(-1)
    c_ = (i2 * n_minor + j2).reshape(-1)
    d = (i * n_minor + j2).reshape(-1)
    f = np.vstack([np.stack([a, b, c_], 1), np.stack([a, c_, d], 1)])
    return Mesh(verts, f.astype(np.int64), name=f"torus_{n_major}x{n_minor}")


# ----------------------------------------------------------------------------------------
# query sets
# ----------------------------------------------------------------------------------------

def _scaled_box(lo, hi, scale):
    c = 0.5 * (lo + hi)
    h = 0.5 * (hi - lo) * scale
    return c - h, c + h


def grid_queries(lo, hi, n: int, scale: float = 1.2) -> np.ndarray:
    """n^3 points on a regular grid over the box scaled about its centre; x fastest, ends inclusive."""
    a, b = _scaled_box(np.asarray(lo, np.float64), np.asarray(hi, np.float64), scale)
    t = np.linspace(0.0, 1.0, n)
    xs = a[0] + (b[0] - a[0]) * t
    ys = a[1] + (b[1] - a[1]) * t
    zs = a[2] + (b[2] - a[2]) * t
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    return np.stack([X, Y, Z], axis=-1).reshape(-1, 3).astype(np.float32)


def uniform_queries(lo, hi, n: int, seed: int, scale: float = 1.0, stream: int = 404) -> np.ndarray:
    a, b = _scaled_box(np.asarray(lo, np.float64), np.asarray(hi, np.float64), scale)
    u = rng_uniform(seed, stream, 3 * n).reshape(n, 3)
    return (a + (b - a) * u).astype(np.float32)


def near_surface_queries(mesh: Mesh, n: int, seed: int, lo_exp: float = -7.0, hi_exp: float = -2.0,
                         stream: int = 505) -> np.ndarray:
    """Points offset from area-weighted surface samples along the face normal.

    Offset = +-diag * 10^U(lo_exp, hi_exp), sign 50/50.  Computed in float64 and rounded
    to float32 at the end; the oracle must be evaluated at the rounded point.
    """
    V = mesh.vertices
    F = mesh.facets
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    cr = np.cross(b - a, c - a)
    area2 = np.linalg.norm(cr, axis=1)
    nrm = cr / np.maximum(area2, 1e-300)[:, None]
    cdf = np.cumsum(area2)
    cdf /= cdf[-1]
    u = rng_uniform(seed, stream, 5 * n).reshape(n, 5)
    t = np.minimum(np.searchsorted(cdf, u[:, 0], side="right"), F.shape[0] - 1)
    s1 = np.sqrt(u[:, 1])
    bu = 1.0 - s1
    bv = s1 * (1.0 - u[:, 2])
    bw = s1 * u[:, 2]
    p = a[t] * bu[:, None] + b[t] * bv[:, None] + c[t] * bw[:, None]
    lo, hi = mesh.bbox()
    diag = float(np.linalg.norm(hi - lo))
    d = diag * np.power(10.0, lo_exp + (hi_exp - lo_exp) * u[:, 3])
    sgn = np.where(u[:, 4] < 0.5, -1.0, 1.0)
    q = p + (sgn * d)[:, None] * nrm[t]
    return q.astype(np.float32)


# ----------------------------------------------------------------------------------------
# the five named workloads
# ----------------------------------------------------------------------------------------

def small_icosphere_workload(n_queries: int = 1_000_000, seed: int = 1):
    """Level-5 icosphere (10,242 V / 20,480 F), points uniform in [-1.5, 1.5]^3."""
    m = icosphere(32)
    q = uniform_queries(np.full(3, -1.5), np.full(3, 1.5), n_queries, seed)
    return m, q


def large_displaced_mesh(seed: int = 2, nu: int = 224) -> Mesh:
    """Displaced sphere, nu=224 -> 1,003,520 facets, 5% deleted (953,344 facets)."""
    return displaced_sphere(nu, seed, amplitude=0.15, n_terms=32, lmax=16, delete_fraction=0.05)


def large_exact_workload(grid_n: int = 256, seed: int = 2, nu: int = 224):
    """~1M-triangle open displaced sphere, grid_n^3 queries on a grid over bbox x1.2."""
    m = large_displaced_mesh(seed, nu)
    lo, hi = m.bbox()
    return m, grid_queries(lo, hi, grid_n, 1.2)


def batched_workload(seed: int = 3, n_meshes: int = 64, f_lo: float = 1e3, f_hi: float = 5e5,
                     q_lo: float = 1e4, q_hi: float = 2e6):
    """n_meshes spheres/tori with log-uniform facet and query counts; points in bbox x1.3."""
    u = rng_uniform(seed, 606, 4 * n_meshes).reshape(n_meshes, 4)
    meshes, queries = [], []
    for i in range(n_meshes):
        nf = math.exp(math.log(f_lo) + (math.log(f_hi) - math.log(f_lo)) * u[i, 1])
        nq = int(math.exp(math.log(q_lo) + (math.log(q_hi) - math.log(q_lo)) * u[i, 2]))
        if u[i, 0] < 0.5:
            nu = max(7, int(round(math.sqrt(nf / 20.0))))
            m = displaced_sphere(nu, seed * 1000 + i, amplitude=0.15, n_terms=8, lmax=6)
        else:
            nv = max(6, int(round(math.sqrt(nf / 2.0 * 0.3))))
            nm = max(6, int(round(nf / 2.0 / nv)))
            m = torus(nm, nv, 1.0, 0.3, seed=seed * 1000 + i, amplitude=0.2)
        lo, hi = m.bbox()
        meshes.append(m)
        queries.append(uniform_queries(lo, hi, nq, seed * 1000 + i, scale=1.3))
    return meshes, queries


def near_surface_workload(n_queries: int = 4_000_000, seed: int = 4, nu: int = 100):
    """200,000-facet displaced sphere, points offset 1e-7..1e-2 x diag along normals."""
    m = displaced_sphere(nu, seed, amplitude=0.15, n_terms=32, lmax=16)
    return m, near_surface_queries(m, n_queries, seed)
