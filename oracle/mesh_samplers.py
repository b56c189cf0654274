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
  vendored voxelizer, external/voxelizer/voxelizer.h:629-760): the set of
  voxel centres whose box, grown by the precision, meets a triangle, checked
  here by clipping each triangle to the box (a different algorithm from the
  separating-axis test the library uses).
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


def voxel_centres(verts, tris, spacing, precision):
    """Lattice indices of the voxels whose box grown by precision meets a
    triangle, among each triangle's candidates. The candidates follow the
    voxelizer (voxelizer.h, vx__map_to_voxel and the loops of vx_voxelize):
    per axis the extent is snapped outwards as floor / ceil((p + sign(p) s/2)
    / s) * s in float32, and centres are stepped from the lower bound by
    x += s in float32 while x < upper bound, which can add one centre past
    the exact range. The overlap test itself is the float64 clipping above.
    Zero-area triangles are skipped."""
    s = f32(spacing)
    half = float(s) / 2 + float(f32(precision))
    found = set()

    def snap(p, lower):
        p = f32(p)
        v = f32(f32(p + f32(f32(-1.0 if p < 0 else 1.0) * s) * f32(0.5)) / s)
        return f32((np.floor(v) if lower else np.ceil(v)) * s)

    def steps(lo, hi):
        out, x = [], lo
        while x < hi:
            out.append(float(x))
            x = f32(x + s)
        return out

    for t in tris:
        tri32 = verts[t].astype(np.float32)
        tri = tri32.astype(np.float64)
        if np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0])) * 0.5 < 1e-7:
            continue
        lo = [snap(v, True) for v in tri32.min(0)]
        hi = [snap(v, False) for v in tri32.max(0)]
        for x in steps(lo[0], hi[0]):
            for y in steps(lo[1], hi[1]):
                for z in steps(lo[2], hi[2]):
                    c = np.array([x, y, z])
                    if _clip_touches(tri, c - half, c + half):
                        found.add(tuple(np.rint(c / float(s)).astype(int)))
    return found
