"""Seeded synthetic meshes in lattice space, for the benchmark workloads.

Only IEEE-exact operations run on arrays (numpy +, -, *, /, sqrt); every
transcendental value comes from Python's math module (the C library), one
scalar at a time.  The float32 vertex arrays are therefore the same on any
x86-64 host with this image, which the golden hashes in tests/golden rely on.
"""
from __future__ import annotations

import hashlib
import math

import numpy as np

_ICO_FACES = [
    (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
    (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
    (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
    (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1),
]


def _random_rotation(seed: int) -> np.ndarray:
    """Uniform random rotation (Shoemake's quaternion from three uniforms)."""
    u1, u2, u3 = np.random.default_rng(seed).random(3).tolist()
    s1, s2 = math.sqrt(1 - u1), math.sqrt(u1)
    qx, qy = s1 * math.sin(2 * math.pi * u2), s1 * math.cos(2 * math.pi * u2)
    qz, qw = s2 * math.sin(2 * math.pi * u3), s2 * math.cos(2 * math.pi * u3)
    return np.array([
        [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
        [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
        [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)],
    ])


def _apply(Rm: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Row-wise R @ v written elementwise (no BLAS, no fused multiply-add)."""
    out = np.empty_like(v)
    for r in range(3):
        out[:, r] = (Rm[r, 0] * v[:, 0] + Rm[r, 1] * v[:, 1]) + Rm[r, 2] * v[:, 2]
    return out


def icosphere(subdiv: int = 5, diameter: float = 16.0, center=(48.0, 40.0, 32.0),
              seed: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """Icosahedron refined `subdiv` times (20 * 4^subdiv triangles), projected
    to the sphere, rotated by a seeded random rotation, scaled and moved."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [tuple(float(c) for c in v) for v in verts]

    def unit(v):
        n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
        return (v[0] / n, v[1] / n, v[2] / n)

    verts = [unit(v) for v in verts]
    faces = list(_ICO_FACES)
    for _ in range(subdiv):
        cache: dict[tuple[int, int], int] = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                va, vb = verts[a], verts[b]
                verts.append(unit(((va[0] + vb[0]) / 2, (va[1] + vb[1]) / 2,
                                   (va[2] + vb[2]) / 2)))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = _apply(_random_rotation(seed), np.array(verts, dtype=np.float64))
    r = diameter / 2.0
    pos = np.empty_like(v)
    for k in range(3):
        pos[:, k] = center[k] + r * v[:, k]
    return pos.astype(np.float32).reshape(-1), np.array(faces, dtype=np.uint32).reshape(-1)


def _spow(x: float, p: float) -> float:
    return math.copysign(math.pow(abs(x), p), x)


def vehicle(seed: int = 7, nu: int = 300, nv: int = 280, wheel_segments: int = 2000,
            length: float = 96.0, nose_x: float = 64.0, y_center: float = 80.0,
            ground_z: float = 1.0, wheel_radius: float = 8.0
            ) -> tuple[np.ndarray, np.ndarray, dict]:
    """Superellipsoid body plus four y-axis cylinder wheels resting on the
    plane z = ground_z.  With the defaults: 2*nu*nv + 16*wheel_segments =
    200 000 triangles.  Returns (positions, indices, shape parameters)."""
    rng = np.random.default_rng(seed)
    e_xy, e_z, bw, bh, clear = rng.random(5).tolist()
    e_xy = 2.5 + 1.5 * e_xy          # horizontal exponent in [2.5, 4]
    e_z = 2.5 + 1.5 * e_z            # vertical exponent in [2.5, 4]
    a = length / 2.0
    b = 18.0 + 4.0 * bw              # half width  [18, 22]
    c = 12.0 + 4.0 * bh              # half height [12, 16]
    clearance = 4.0 + 2.0 * clear    # body bottom above the ground [4, 6]
    x0, y0 = nose_x + a, y_center
    z0 = ground_z + clearance + c

    pos: list[tuple[float, float, float]] = []
    tri: list[tuple[int, int, int]] = []
    # body: nv latitude rings between the poles, nu longitudes
    pos.append((x0, y0, z0 - c))                     # south pole
    for j in range(nv):
        eta = -math.pi / 2 + math.pi * (j + 1) / (nv + 1)
        ce, se = math.cos(eta), math.sin(eta)
        for i in range(nu):
            om = -math.pi + 2 * math.pi * i / nu
            co, so = math.cos(om), math.sin(om)
            pos.append((x0 + a * _spow(ce, 2 / e_z) * _spow(co, 2 / e_xy),
                        y0 + b * _spow(ce, 2 / e_z) * _spow(so, 2 / e_xy),
                        z0 + c * _spow(se, 2 / e_z)))
    pos.append((x0, y0, z0 + c))                     # north pole
    north = len(pos) - 1
    ring = lambda j, i: 1 + j * nu + (i % nu)  # noqa: E731
    for i in range(nu):
        tri.append((0, ring(0, i + 1), ring(0, i)))
    for j in range(nv - 1):
        for i in range(nu):
            tri.append((ring(j, i), ring(j, i + 1), ring(j + 1, i + 1)))
            tri.append((ring(j, i), ring(j + 1, i + 1), ring(j + 1, i)))
    for i in range(nu):
        tri.append((north, ring(nv - 1, i), ring(nv - 1, i + 1)))

    # wheels: cylinders along y, width 8, touching the ground plane
    wz = ground_z + wheel_radius
    half_w = 4.0
    for sx in (-1, 1):
        for sy in (-1, 1):
            cx = x0 + sx * 0.62 * a
            cy = y0 + sy * (b - 3.0)
            base = len(pos)
            for side in (-1, 1):
                y = cy + side * half_w
                for k in range(wheel_segments):
                    th = 2 * math.pi * k / wheel_segments
                    pos.append((cx + wheel_radius * math.cos(th), y, wz + wheel_radius * math.sin(th)))
            pos.append((cx, cy - half_w, wz))
            pos.append((cx, cy + half_w, wz))
            c0, c1 = len(pos) - 2, len(pos) - 1
            n = wheel_segments
            for k in range(n):
                k1 = (k + 1) % n
                a0, a1, b0, b1 = base + k, base + k1, base + n + k, base + n + k1
                tri.append((a0, a1, b1))
                tri.append((a0, b1, b0))
                tri.append((c0, a1, a0))
                tri.append((c1, b0, b1))
    p = np.array(pos, dtype=np.float64).astype(np.float32).reshape(-1)
    idx = np.array(tri, dtype=np.uint32).reshape(-1)
    info = {"seed": seed, "e_xy": e_xy, "e_z": e_z, "a": a, "b": b, "c": c,
            "center": (x0, y0, z0), "clearance": clearance, "n_tri": len(tri)}
    return p, idx, info


def mesh_digest(pos: np.ndarray, idx: np.ndarray) -> str:
    """sha256 of the float32 positions followed by the uint32 indices."""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(pos, dtype=np.float32).tobytes())
    h.update(np.ascontiguousarray(idx, dtype=np.uint32).tobytes())
    return h.hexdigest()
