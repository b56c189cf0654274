"""CPU restatement (numpy float32) of the mesh-body samplers, test
infrastructure only:

* Sphere::surface_sampling (scene/sphere.cpp:135-174) over
  IcoSphereMesh(radius, 3) (mesh/icospheremesh.cpp:10-140), with the
  reference's quirks (corner k read at float offset k, the zero third
  cross-product term, a fresh default-seeded engine per triangle);
* the generator: libstdc++'s default_random_engine is minstd_rand0
  (x <- 16807 x mod 2^31 - 1, seed 1), and uniform_real_distribution<float>
  over generate_canonical<float> is float(x - 1) / 2^31;
* Model::surface_sampling's voxelization (scene/model.cpp:170-214 with the
  vendored voxelizer, external/voxelizer/voxelizer.h:629-760): vx_voxelize
  restated in float32 (loops, candidate snapping, 1e-7 duplicate tolerance,
  4096-bucket table order), with its separating-axis triangle-box test
  pinned by float64 clipping of each triangle to each candidate box.
"""
import numpy as np

f32 = np.float32


class MinStdRand0:
    """std::minstd_rand0: the C++ standard fixes its 10000th output from the
    default seed at 1043618065 ([rand.predef])."""

    def __init__(self, seed=1):
        self.x = seed

    def __call__(self):
        self.x = (16807 * self.x) % 2147483647
        return self.x

    def canonical(self):
        return f32(f32(self() - 1) / f32(2147483648.0))


def icosphere(radius, level):
    t = f32((1.0 + np.sqrt(5.0)) * 0.5)
    one = f32(1)
    v = [(-one, t, f32(0)), (one, t, f32(0)), (-one, -t, f32(0)), (one, -t, f32(0)),
         (f32(0), -one, t), (f32(0), one, t), (f32(0), -one, -t), (f32(0), one, -t),
         (t, f32(0), -one), (t, f32(0), one), (-t, f32(0), -one), (-t, f32(0), one)]
    v = [np.array(p, np.float32) for p in v]
    f = [0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
         3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1]
    half = f32(0.5)
    for _ in range(level):
        nv = list(v)
        nf = []
        for k in range(0, len(f), 3):
            a, b, c = f[k], f[k + 1], f[k + 2]
            nv.append((v[a] + v[b]) * half)
            ab = len(nv) - 1
            nv.append((v[b] + v[c]) * half)
            bc = len(nv) - 1
            nv.append((v[c] + v[a]) * half)
            ca = len(nv) - 1
            nf += [a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca]
        v, f = nv, nf
    out = []
    r = f32(radius)
    for p in v:
        ln = np.sqrt(f32(f32(p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]))
        inv = f32(f32(1) / ln)
        out.append(np.array([p[0] * inv * r, p[1] * inv * r, p[2] * inv * r], np.float32))
    return np.array(out, np.float32).reshape(-1), np.array(f, np.int64)


def sphere_sampling(radius, spacing):
    V, I = icosphere(radius, 3)
    s = f32(spacing)
    area_particle = f32(np.pi * float(s) ** 2)
    out = []
    one = f32(1)
    for i in range(0, len(I), 3):
        a, b, c = (V[I[i + k]:I[i + k] + 3] for k in range(3))
        ab, ac = a - b, a - c
        c0 = f32(ab[1] * ac[2] - ab[2] * ac[1])
        c1 = f32(ab[2] * ac[0] - ab[0] * ac[2])
        c2 = f32(ab[0] * ac[1] - ab[0] * ac[1])
        area = f32(0.5 * np.sqrt(float(c0) ** 2 + float(c1) ** 2 + float(c2) ** 2))
        n = int(f32(2) * np.ceil(f32(area / area_particle)))
        g = MinStdRand0()
        for _ in range(n):
            r1 = np.sqrt(g.canonical())
            r2 = g.canonical()
            w0, w1, w2 = f32(one - r1), f32(r1 * f32(one - r2)), f32(r2 * r1)
            out.append(np.array([f32(f32(w0 * a[k] + w1 * b[k]) + w2 * c[k]) for k in range(3)], np.float32))
    for k in range(0, len(V), 3):
        out.append(V[k:k + 3])
    out = np.array(out, np.float32)
    return np.concatenate([out, np.ones((out.shape[0], 1), np.float32)], axis=1)


def read_obj(path):
    """Positions and triangles of an OBJ file (faces fan-triangulated),
    deduplicated by (position, normal) index pair in order of first use."""
    pos, tris, seen, verts = [], [], {}, []
    nn = 0
    for line in open(path):
        p = line.split()
        if not p:
            continue
        if p[0] == "v":
            pos.append([float(x) for x in p[1:4]])
        elif p[0] == "vn":
            nn += 1
        elif p[0] == "f":
            fv = []
            for tok in p[1:]:
                parts = tok.split("/")
                vi = int(parts[0])
                ni = int(parts[2]) if len(parts) > 2 and parts[2] else 0
                fv.append((vi - 1 if vi > 0 else len(pos) + vi, ni - 1 if ni > 0 else (nn + ni if ni < 0 else -1)))
            for k in range(1, len(fv) - 1):
                for key in (fv[0], fv[k], fv[k + 1]):
                    if key not in seen:
                        seen[key] = len(verts)
                        verts.append(pos[key[0]])
                    tris.append(seen[key])
    return np.array(verts, np.float64), np.array(tris, np.int64).reshape(-1, 3)


def _clip_touches(tri, lo, hi):
    """Does the triangle meet the closed box [lo, hi]? Clips the polygon
    against the six planes (Sutherland-Hodgman) in float64."""
    poly = [np.array(p, np.float64) for p in tri]
    for axis in range(3):
        for side, bound in ((1, lo[axis]), (-1, hi[axis])):
            out = []
            n = len(poly)
            for k in range(n):
                a, b = poly[k], poly[(k + 1) % n]
                da, db = side * (a[axis] - bound), side * (b[axis] - bound)
                if da >= 0:
                    out.append(a)
                if (da >= 0) != (db >= 0):
                    t = da / (da - db)
                    out.append(a + t * (b - a))
            poly = out
            if not poly:
                return False
    return True


# ----------------------------------------------------------------------------
# vx_voxelize restated in float32 (external/voxelizer/voxelizer.h:629-760), for
# the exact sample count and order of Model::surface_sampling
# (scene/model.cpp:170-214). The separating-axis test is written here from the
# published algorithm (Akenine-Moller 2001) as the voxelizer states it
# (voxelizer.h:330-508: nine edge axes with its vertex pairs, the box's three
# face normals, the triangle plane), vectorised over a triangle's candidate
# boxes; the clipping check above pins it wherever a box is not within
# rounding of touching the triangle.
def _sat(v1, v2, v3, h):
    """Triangle (v1, v2, v3, relative to each box centre; arrays (m, 3)
    float32) against boxes of half sizes h (m, 3): bool (m,)."""
    e1, e2, e3 = v2 - v1, v3 - v2, v1 - v3
    ok = np.ones(v1.shape[0], bool)
    X, Y, Z = 0, 1, 2

    def axis(pa, pb, rad):
        lo, hi = np.minimum(pa, pb), np.maximum(pa, pb)
        return ~((lo > rad) | (hi < -rad))

    for e, (xa, xb), (ya, yb), (za, zb) in (
            (e1, (v1, v3), (v1, v3), (v2, v3)),
            (e2, (v1, v3), (v1, v3), (v1, v2)),
            (e3, (v1, v2), (v1, v2), (v2, v3))):
        fx, fy, fz = np.abs(e[:, X]), np.abs(e[:, Y]), np.abs(e[:, Z])
        a, b = e[:, Z], e[:, Y]
        ok &= axis(a * xa[:, Y] - b * xa[:, Z], a * xb[:, Y] - b * xb[:, Z], fz * h[:, Y] + fy * h[:, Z])
        a, b = e[:, Z], e[:, X]
        ok &= axis(-a * ya[:, X] + b * ya[:, Z], -a * yb[:, X] + b * yb[:, Z], fz * h[:, X] + fx * h[:, Z])
        a, b = e[:, Y], e[:, X]
        ok &= axis(a * za[:, X] - b * za[:, Y], a * zb[:, X] - b * zb[:, Y], fy * h[:, X] + fx * h[:, Y])
    for k in range(3):
        lo = np.minimum(np.minimum(v1[:, k], v2[:, k]), v3[:, k])
        hi = np.maximum(np.maximum(v1[:, k], v2[:, k]), v3[:, k])
        ok &= ~((lo > h[:, k]) | (hi < -h[:, k]))
    n = np.stack([e1[:, Y] * e2[:, Z] - e1[:, Z] * e2[:, Y], e1[:, Z] * e2[:, X] - e1[:, X] * e2[:, Z],
                  e1[:, X] * e2[:, Y] - e1[:, Y] * e2[:, X]], 1)
    d = -((n[:, X] * v1[:, X] + n[:, Y] * v1[:, Y]) + n[:, Z] * v1[:, Z])
    vmin = np.where(n > 0, -h, h)
    vmax = np.where(n > 0, h, -h)
    dmin = ((n[:, X] * vmin[:, X] + n[:, Y] * vmin[:, Y]) + n[:, Z] * vmin[:, Z]) + d
    dmax = ((n[:, X] * vmax[:, X] + n[:, Y] * vmax[:, Y]) + n[:, Z] * vmax[:, Z]) + d
    return ok & ~(dmin > 0) & (dmax >= 0)


def _bucket(c):
    """vx__vertex_hash: each coordinate times its prime (the int prime
    rounded to float), truncated to a 64-bit integer, xor, mod 4096."""
    h = 0
    for v, prime in zip(c, (73856093, 19349663, 83492791)):
        p = f32(v) * f32(prime)
        h ^= int(np.trunc(np.float64(p))) & 0xFFFFFFFFFFFFFFFF
    return h % 4096


def voxelize(verts, tris, spacing, precision):
    """Model::surface_sampling's vertices: 8 corners per kept voxel, voxels in
    the 4096-bucket table's order (bucket, then insertion), float32 (n, 4)."""
    s, prec = f32(spacing), f32(precision)
    hs = f32(s * f32(0.5))
    table = [[] for _ in range(4096)]
    V = verts.astype(np.float32)

    def snap(p, lower):
        v = f32(f32(p + f32(f32(-1.0 if p < 0 else 1.0) * s) * f32(0.5)) / s)
        return f32((np.floor(v) if lower else np.ceil(v)) * s)

    def steps(lo, hi):
        out, x = [], lo
        while x < hi:
            out.append(x)
            x = f32(x + s)
        return np.array(out, np.float32)

    for t in tris:
        p1, p2, p3 = V[t[0]], V[t[1]], V[t[2]]
        ab, ac = p2 - p1, p3 - p1
        a0 = f32(ab[1] * ac[2] - ab[2] * ac[1])
        a1 = f32(ab[2] * ac[0] - ab[0] * ac[2])
        a2 = f32(ab[0] * ac[1] - ab[1] * ac[0])
        area = f32(np.sqrt(f32(f32(a0 * a0 + a1 * a1) + a2 * a2)) * f32(0.5))
        if float(area) < 1e-7:
            continue
        lo = [snap(min(min(p1[k], p2[k]), p3[k]), True) for k in range(3)]
        hi = [snap(max(max(p1[k], p2[k]), p3[k]), False) for k in range(3)]
        xs, ys, zs = steps(lo[0], hi[0]), steps(lo[1], hi[1]), steps(lo[2], hi[2])
        if not (len(xs) and len(ys) and len(zs)):
            continue
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        g = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)  # loop order x, y, z
        bmin, bmax = g - hs, g + hs
        c = (bmin + bmax) * f32(0.5)
        h = np.abs(bmax - bmin) * f32(0.5) + prec
        keep = _sat(p1[None, :] - c, p2[None, :] - c, p3[None, :] - c, h)
        for cc in c[keep]:
            b = table[_bucket(cc)]
            if not any(abs(float(q[0]) - float(cc[0])) < 1e-7 and abs(float(q[1]) - float(cc[1])) < 1e-7 and
                       abs(float(q[2]) - float(cc[2])) < 1e-7 for q in b):
                b.append(cc)
    corners = np.array([(-hs, hs, hs), (-hs, -hs, hs), (hs, -hs, hs), (hs, hs, hs),
                        (-hs, hs, -hs), (-hs, -hs, -hs), (hs, -hs, -hs), (hs, hs, -hs)], np.float32)
    out = [corners + cc for b in table for cc in b]
    if not out:
        return np.zeros((0, 4), np.float32)
    out = np.concatenate(out).astype(np.float32)
    return np.concatenate([out, np.ones((out.shape[0], 1), np.float32)], 1)


def sat_against_clipping(verts, tris, spacing, precision, margin=1e-6):
    """Every candidate box of every triangle (the voxelizer's loops) is
    classified in float64 by clipping: 'in' when the triangle meets the box
    shrunk by `margin`, 'out' when it misses the box grown by `margin`,
    otherwise ambiguous (within rounding of touching). Returns the counts
    (agree_in, agree_out, disagree, ambiguous) of the float32 separating-axis
    test against that classification."""
    s, prec = f32(spacing), f32(precision)
    hs = f32(s * f32(0.5))
    V = verts.astype(np.float32)
    agree_in = agree_out = disagree = ambiguous = 0

    def snap(p, lower):
        v = f32(f32(p + f32(f32(-1.0 if p < 0 else 1.0) * s) * f32(0.5)) / s)
        return f32((np.floor(v) if lower else np.ceil(v)) * s)

    def steps(lo, hi):
        out, x = [], lo
        while x < hi:
            out.append(x)
            x = f32(x + s)
        return np.array(out, np.float32)

    for t in tris:
        p1, p2, p3 = V[t[0]], V[t[1]], V[t[2]]
        tri = np.stack([p1, p2, p3]).astype(np.float64)
        if np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0])) * 0.5 < 1e-7:
            continue
        lo = [snap(min(min(p1[k], p2[k]), p3[k]), True) for k in range(3)]
        hi = [snap(max(max(p1[k], p2[k]), p3[k]), False) for k in range(3)]
        xs, ys, zs = steps(lo[0], hi[0]), steps(lo[1], hi[1]), steps(lo[2], hi[2])
        if not (len(xs) and len(ys) and len(zs)):
            continue
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        g = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
        bmin, bmax = g - hs, g + hs
        c = (bmin + bmax) * f32(0.5)
        h = np.abs(bmax - bmin) * f32(0.5) + prec
        sat = _sat(p1[None, :] - c, p2[None, :] - c, p3[None, :] - c, h)
        for cc, hh, ov in zip(c.astype(np.float64), h.astype(np.float64), sat):
            inner = _clip_touches(tri, cc - hh + margin, cc + hh - margin)
            outer = _clip_touches(tri, cc - hh - margin, cc + hh + margin)
            if inner:
                agree_in += bool(ov)
                disagree += not ov
            elif not outer:
                agree_out += not ov
                disagree += bool(ov)
            else:
                ambiguous += 1
    return agree_in, agree_out, disagree, ambiguous
