"""Membrane meshes for the configurations: icospheres (12 -> 42 -> 162 ->
642 -> 2562 vertices), ellipsoids, the Evans-Fung biconcave red cell, and
gmsh v2 ASCII read / write in the layout the reference's reader expects
(src/fsi/ParticleShapeFactory.hh:32-83: $Nodes with 1-based ids, then
$Elements whose last three integers are the triangle's 1-based nodes).

Setup-time host code (numpy); triangles are oriented outward (positive
signed volume).
"""
import numpy as np


def icosphere(level, radius=1.0):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    verts = [tuple(p) for p in v]
    for _ in range(level):
        cache = {}
        new_f = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = (np.array(verts[a]) + np.array(verts[b])) / 2.0
                m /= np.linalg.norm(m)
                verts.append(tuple(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_f += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = np.array(new_f, np.int64)
    v = np.array(verts) * radius
    return orient_outward(v, f.astype(np.int32))


def signed_volume(v, f):
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0


def orient_outward(v, f):
    if signed_volume(v, f) < 0:
        f = f[:, [0, 2, 1]].copy()
    return v, np.ascontiguousarray(f, np.int32)


def ellipsoid(level, axes):
    v, f = icosphere(level, 1.0)
    return v * np.asarray(axes, float), f


def biconcave(level, radius):
    """Evans-Fung red-cell shape z = +-(R/2) sqrt(1 - r^2) (C0 + C1 r^2 +
    C2 r^4), r = rho / R, with C0 = 0.207, C1 = 2.003, C2 = -1.123
    (Evans & Fung, Microvasc. Res. 4 (1972) 335), mapped from the unit
    icosphere (x, y kept, z replaced)."""
    v, f = icosphere(level, 1.0)
    r2 = np.clip(v[:, 0] ** 2 + v[:, 1] ** 2, 0, 1)
    c0, c1, c2 = 0.207, 2.003, -1.123
    z = 0.5 * np.sqrt(1 - r2) * (c0 + c1 * r2 + c2 * r2 * r2)
    out = np.stack([v[:, 0], v[:, 1], np.sign(v[:, 2]) * z], axis=1) * radius
    return orient_outward(out, f)


def rotation_matrix(axis, angle):
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def write_gmsh(path, v, f):
    with open(path, "w") as fh:
        fh.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(v))
        for i, p in enumerate(v):
            fh.write("%d %.17g %.17g %.17g\n" % (i + 1, p[0], p[1], p[2]))
        fh.write("$EndNodes\n$Elements\n%d\n" % len(f))
        for i, t in enumerate(f):
            fh.write("%d 2 2 0 1 %d %d %d\n" % (i + 1, t[0] + 1, t[1] + 1, t[2] + 1))
        fh.write("$EndElements\n")


def read_gmsh(path):
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    i = lines.index("$Nodes")
    nn = int(lines[i + 1])
    v = np.zeros((nn, 3))
    for k in range(nn):
        parts = lines[i + 2 + k].split()
        v[int(parts[0]) - 1] = [float(x) for x in parts[1:4]]
    j = lines.index("$Elements")
    ne = int(lines[j + 1])
    tris = []
    for k in range(ne):
        parts = [int(x) for x in lines[j + 2 + k].split()]
        tris.append([parts[-3] - 1, parts[-2] - 1, parts[-1] - 1])
    return v, np.array(tris, np.int32)
