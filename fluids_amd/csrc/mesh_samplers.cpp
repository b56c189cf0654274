// Surface samplers of the mesh bodies: Sphere::surface_sampling
// (scene/sphere.cpp:135-174) over IcoSphereMesh(radius, 3)
// (mesh/icospheremesh.cpp), and Model::surface_sampling
// (scene/model.cpp:170-214) over the OBJ mesh (mesh/modelmesh.cpp) and the
// voxelizer the reference vendors (external/voxelizer/voxelizer.h:629-760,
// MIT). The arithmetic follows the reference's float / double evaluation
// order, since sample counts and positions depend on it.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "solver.h"

namespace sph {

namespace {

struct P3 {
    float x, y, z;
};
inline P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 operator*(float s, P3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline P3 operator*(P3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }

void put(std::vector<float4>& v, P3 p) { v.push_back(make_float4(p.x, p.y, p.z, 1.0f)); }

}  // namespace

// IcoSphereMesh(radius, level): the 12-vertex icosahedron, each level
// splitting every face into four with three new (unshared) midpoints, then
// every vertex normalised (v * (1 / |v|)) and scaled by the radius.
void icosphere(float radius, unsigned level, std::vector<float>& verts, std::vector<uint16_t>& idx) {
    const float t = (float)((1.0 + std::sqrt(5.0)) * 0.5);
    std::vector<P3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    // faces of the icosahedron: five around vertex 0, five adjacent, five
    // around vertex 3, five adjacent
    std::vector<uint32_t> f = {0, 11, 5,  0, 5,  1,  0, 1,  7,  0, 7,  10, 0, 10, 11,
                               1, 5,  9,  5, 11, 4,  11, 10, 2, 10, 7, 6,  7, 1,  8,
                               3, 9,  4,  3, 4,  2,  3, 2,  6,  3, 6,  8,  3, 8,  9,
                               4, 9,  5,  2, 4,  11, 6, 2,  10, 8, 6,  7,  9, 8,  1};
    for (unsigned l = 0; l < level; ++l) {
        std::vector<uint32_t> nf;
        std::vector<P3> nv = v;
        for (size_t k = 0; k < f.size(); k += 3) {
            const uint32_t a = f[k], b = f[k + 1], c = f[k + 2];
            nv.push_back((v[a] + v[b]) * 0.5f);
            const uint32_t ab = (uint32_t)nv.size() - 1;
            nv.push_back((v[b] + v[c]) * 0.5f);
            const uint32_t bc = (uint32_t)nv.size() - 1;
            nv.push_back((v[c] + v[a]) * 0.5f);
            const uint32_t ca = (uint32_t)nv.size() - 1;
            const uint32_t faces[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), faces, faces + 12);
        }
        f.swap(nf);
        v.swap(nv);
    }
    verts.clear();
    for (const P3& p : v) {
        float len = std::sqrt((p.x * p.x + p.y * p.y) + p.z * p.z);
        P3 n = p * (1.0f / len);
        P3 s = n * radius;
        verts.push_back(s.x);
        verts.push_back(s.y);
        verts.push_back(s.z);
    }
    idx.assign(f.begin(), f.end());  // mesh_index is 16-bit in the reference
}

// Sphere::surface_sampling, with the reference's quirks: a triangle corner
// is read at float offset `index` of the vertex array (not 3*index), the
// third cross-product term of the area is ab.x*ac.y - ab.x*ac.y (= 0), and a
// fresh default-seeded std::default_random_engine (minstd_rand0 in
// libstdc++) is built for every triangle. The mesh vertices follow.
std::vector<float4> sample_sphere(float radius, float spacing) {
    if (!(radius > 0.0f) || !(spacing > 0.0f)) throw Error(SPH_ERR_INVALID, "sphere sampling: radius and spacing > 0");
    std::vector<float> V;
    std::vector<uint16_t> I;
    icosphere(radius, 3, V, I);
    std::vector<float4> out;
    const float area_particle = (float)(M_PI * std::pow(spacing, 2));
    auto corner = [&](uint16_t k) { return P3{V[k], V[k + 1], V[k + 2]}; };
    for (size_t i = 0; i < I.size(); i += 3) {
        P3 a = corner(I[i]), b = corner(I[i + 1]), c = corner(I[i + 2]);
        P3 ab = a - b, ac = a - c;
        float c0 = ab.y * ac.z - ab.z * ac.y;
        float c1 = ab.z * ac.x - ab.x * ac.z;
        float c2 = ab.x * ac.y - ab.x * ac.y;
        float area_triangle = (float)(0.5f * std::sqrt(std::pow(c0, 2) + std::pow(c1, 2) + std::pow(c2, 2)));
        int n = (int)(2 * std::ceil(area_triangle / area_particle));
        std::minstd_rand0 gen;
        std::uniform_real_distribution<float> rnd(0, 1);
        for (int k = 0; k < n; ++k) {
            float r1 = std::sqrt(rnd(gen));
            float r2 = rnd(gen);
            put(out, ((1.0f - r1) * a + (r1 * (1.0f - r2)) * b) + (r2 * r1) * c);
        }
    }
    for (size_t i = 0; i + 2 < V.size(); i += 3) put(out, P3{V[i], V[i + 1], V[i + 2]});
    return out;
}

// --------------------------------------------------------------------- OBJ
namespace {

// The OBJ loader's number parser (tinyobj tryParseDouble, vendored in the
// reference as external/tinyobj/tiny_obj_loader.h): integer digits
// accumulated as m = 10 m + d, fraction digits added as d * 10^-k, then
// sign * ldexp(m * 5^e, e); the double is cast to float.
float obj_number(const char*& s) {
    while (*s == ' ' || *s == '\t') ++s;
    const char* start = s;
    double m = 0.0;
    int e = 0, read = 0;
    char sign = '+', esign = '+';
    if (*s == '+' || *s == '-') sign = *s++;
    else if (!(*s >= '0' && *s <= '9')) goto fail;
    while (*s >= '0' && *s <= '9') {
        m = m * 10 + (*s - '0');
        ++s;
        ++read;
    }
    if (read == 0) goto fail;
    if (*s == '.') {
        ++s;
        int k = 1;
        while (*s >= '0' && *s <= '9') {
            m += (*s - '0') * std::pow(10.0, -k);
            ++k;
            ++s;
        }
    }
    if (*s == 'e' || *s == 'E') {
        ++s;
        if (*s == '+' || *s == '-') esign = *s++;
        int r = 0;
        while (*s >= '0' && *s <= '9') {
            e = e * 10 + (*s - '0');
            ++s;
            ++r;
        }
        if (r == 0) goto fail;
        if (esign == '-') e = -e;
    }
    return (float)((sign == '+' ? 1 : -1) * std::ldexp(m * std::pow(5.0, e), e));
fail:
    s = start;
    while (*s && *s != ' ' && *s != '\t' && *s != '\r' && *s != '\n') ++s;
    return 0.0f;
}

}  // namespace

// ModelMesh (mesh/modelmesh.cpp): positions and normals of an OBJ file,
// one output vertex per distinct (position index, normal index) pair in
// order of first use, faces fan-triangulated.
void load_obj(const std::string& path, std::vector<float>& verts, std::vector<uint32_t>& idx) {
    std::ifstream f(path);
    if (!f.good()) throw Error(SPH_ERR_PARSE, "Could not load model: " + path);
    std::vector<float> pos, nrm;
    std::map<std::pair<int, int>, int> seen;
    verts.clear();
    idx.clear();
    std::string line;
    auto fix_index = [](int i, int n) { return i > 0 ? i - 1 : (i < 0 ? n + i : -1); };
    while (std::getline(f, line)) {
        const char* s = line.c_str();
        while (*s == ' ' || *s == '\t') ++s;
        if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
            s += 2;
            for (int k = 0; k < 3; ++k) pos.push_back(obj_number(s));
        } else if (s[0] == 'v' && s[1] == 'n' && (s[2] == ' ' || s[2] == '\t')) {
            s += 3;
            for (int k = 0; k < 3; ++k) nrm.push_back(obj_number(s));
        } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
            s += 2;
            std::vector<std::pair<int, int>> fv;
            while (*s) {
                while (*s == ' ' || *s == '\t') ++s;
                if (!*s || *s == '\r') break;
                int vi = (int)std::strtol(s, (char**)&s, 10), ni = 0;
                if (*s == '/') {
                    ++s;
                    if (*s != '/') (void)std::strtol(s, (char**)&s, 10);
                    if (*s == '/') {
                        ++s;
                        ni = (int)std::strtol(s, (char**)&s, 10);
                    }
                }
                while (*s && *s != ' ' && *s != '\t') ++s;
                fv.push_back({fix_index(vi, (int)pos.size() / 3), fix_index(ni, (int)nrm.size() / 3)});
            }
            for (size_t k = 1; k + 1 < fv.size(); ++k) {
                const std::pair<int, int> tri[3] = {fv[0], fv[k], fv[k + 1]};
                for (const auto& p : tri) {
                    if (p.first < 0 || 3 * p.first + 2 >= (int)pos.size())
                        throw Error(SPH_ERR_PARSE, "OBJ face names a missing vertex: " + path);
                    auto it = seen.find(p);
                    if (it != seen.end()) {
                        idx.push_back(it->second);
                    } else {
                        int id = (int)verts.size() / 3;
                        for (int c = 0; c < 3; ++c) verts.push_back(pos[3 * p.first + c]);
                        seen[p] = id;
                        idx.push_back(id);
                    }
                }
            }
        }
    }
    if (idx.empty()) throw Error(SPH_ERR_PARSE, "OBJ has no faces: " + path);
}

// ----------------------------------------------------------------- voxels
namespace {

// Separating-axis test of a triangle against an axis-aligned box (Akenine-
// Moller, "Fast 3D triangle-box overlap testing", 2001): the nine axes
// edge x box-axis, the box's three face normals, the triangle's normal. The
// pair of vertices projected for each edge axis, and the comparison forms,
// are those of the published algorithm, which the voxelizer uses.
bool tri_box_overlap(P3 c, P3 h, P3 a, P3 b, P3 d) {
    P3 v0 = a - c, v1 = b - c, v2 = d - c;
    P3 e0 = v1 - v0, e1 = v2 - v1, e2 = v0 - v2;
    auto test = [](float p, float q, float r) {
        float lo = p < q ? p : q, hi = p < q ? q : p;
        return !(lo > r || hi < -r);
    };
    float fx, fy, fz;
    // edge e0: axes x (v0, v2), y (v0, v2), z (v1, v2)
    fx = std::fabs(e0.x); fy = std::fabs(e0.y); fz = std::fabs(e0.z);
    if (!test(e0.z * v0.y - e0.y * v0.z, e0.z * v2.y - e0.y * v2.z, fz * h.y + fy * h.z)) return false;
    if (!test(-e0.z * v0.x + e0.x * v0.z, -e0.z * v2.x + e0.x * v2.z, fz * h.x + fx * h.z)) return false;
    {
        float p1 = e0.y * v1.x - e0.x * v1.y, p2 = e0.y * v2.x - e0.x * v2.y;
        float lo = p2 < p1 ? p2 : p1, hi = p2 < p1 ? p1 : p2, r = fy * h.x + fx * h.y;
        if (lo > r || hi < -r) return false;
    }
    // edge e1: axes x (v0, v2), y (v0, v2), z (v0, v1)
    fx = std::fabs(e1.x); fy = std::fabs(e1.y); fz = std::fabs(e1.z);
    if (!test(e1.z * v0.y - e1.y * v0.z, e1.z * v2.y - e1.y * v2.z, fz * h.y + fy * h.z)) return false;
    if (!test(-e1.z * v0.x + e1.x * v0.z, -e1.z * v2.x + e1.x * v2.z, fz * h.x + fx * h.z)) return false;
    if (!test(e1.y * v0.x - e1.x * v0.y, e1.y * v1.x - e1.x * v1.y, fy * h.x + fx * h.y)) return false;
    // edge e2: axes x (v0, v1), y (v0, v1), z (v1, v2)
    fx = std::fabs(e2.x); fy = std::fabs(e2.y); fz = std::fabs(e2.z);
    if (!test(e2.z * v0.y - e2.y * v0.z, e2.z * v1.y - e2.y * v1.z, fz * h.y + fy * h.z)) return false;
    if (!test(-e2.z * v0.x + e2.x * v0.z, -e2.z * v1.x + e2.x * v1.z, fz * h.x + fx * h.z)) return false;
    {
        float p1 = e2.y * v1.x - e2.x * v1.y, p2 = e2.y * v2.x - e2.x * v2.y;
        float lo = p2 < p1 ? p2 : p1, hi = p2 < p1 ? p1 : p2, r = fy * h.x + fx * h.y;
        if (lo > r || hi < -r) return false;
    }
    // box face normals: the triangle's extent on each axis
    auto extent = [](float p, float q, float r, float& lo, float& hi) {
        lo = hi = p;
        if (q < lo) lo = q;
        if (q > hi) hi = q;
        if (r < lo) lo = r;
        if (r > hi) hi = r;
    };
    float lo, hi;
    extent(v0.x, v1.x, v2.x, lo, hi);
    if (lo > h.x || hi < -h.x) return false;
    extent(v0.y, v1.y, v2.y, lo, hi);
    if (lo > h.y || hi < -h.y) return false;
    extent(v0.z, v1.z, v2.z, lo, hi);
    if (lo > h.z || hi < -h.z) return false;
    // triangle plane against the box
    P3 n = {e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x};
    float dd = -(n.x * v0.x + n.y * v0.y + n.z * v0.z);
    P3 vmin, vmax;
    const float nn[3] = {n.x, n.y, n.z}, hh[3] = {h.x, h.y, h.z};
    float mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = nn[k] > 0.0f ? -hh[k] : hh[k];
        mx[k] = nn[k] > 0.0f ? hh[k] : -hh[k];
    }
    vmin = {mn[0], mn[1], mn[2]};
    vmax = {mx[0], mx[1], mx[2]};
    if (n.x * vmin.x + n.y * vmin.y + n.z * vmin.z + dd > 0.0f) return false;
    return n.x * vmax.x + n.y * vmax.y + n.z * vmax.z + dd >= 0.0f;
}

float voxel_edge(float p, float size, bool lower) {
    float v = (p + (p < 0.0f ? -1.0f : 1.0f) * size * 0.5f) / size;
    return (lower ? std::floor(v) : std::ceil(v)) * size;
}

// bucket of a voxel centre in the voxelizer's 4096-entry table: each
// coordinate times a prime, truncated to a 64-bit integer (two's complement
// for negative values, as x86-64 converts), xor-ed together
size_t voxel_bucket(P3 p) {
    auto part = [](float v, float prime) { return (uint64_t)(int64_t)(v * prime); };
    uint64_t a = part(p.x, 73856093.0f), b = part(p.y, 19349663.0f), c = part(p.z, 83492791.0f);
    return (size_t)((a ^ b ^ c) % 4096u);
}

}  // namespace

// vx_voxelize(mesh, s, s, s, precision) followed by the vertices of the
// output mesh: every voxel whose box (grown by `precision`) a triangle
// touches, once per distinct centre, emitted table bucket by bucket in
// insertion order, 8 corners per voxel.
std::vector<float4> voxel_corners(const std::vector<float>& V, const std::vector<uint32_t>& I, float s,
                                  float precision) {
    const float hs = s * 0.5f;
    std::vector<std::vector<P3>> table(4096);
    auto vert = [&](uint32_t k) { return P3{V[3 * k], V[3 * k + 1], V[3 * k + 2]}; };
    for (size_t t = 0; t + 2 < I.size(); t += 3) {
        P3 p1 = vert(I[t]), p2 = vert(I[t + 1]), p3 = vert(I[t + 2]);
        P3 ab = p2 - p1, ac = p3 - p1;
        float a0 = ab.y * ac.z - ab.z * ac.y, a1 = ab.z * ac.x - ab.x * ac.z, a2 = ab.x * ac.y - ab.y * ac.x;
        float area = std::sqrt(std::pow(a0, 2.f) + std::pow(a1, 2.f) + std::pow(a2, 2.f)) * 0.5f;
        if (area < 0.0000001) continue;
        auto mx = [](float a, float b) { return a > b ? a : b; };
        auto mn = [](float a, float b) { return a > b ? b : a; };
        P3 lo = {mn(mn(mn(INFINITY, p1.x), p2.x), p3.x), mn(mn(mn(INFINITY, p1.y), p2.y), p3.y),
                 mn(mn(mn(INFINITY, p1.z), p2.z), p3.z)};
        P3 hi = {mx(mx(mx(-INFINITY, p1.x), p2.x), p3.x), mx(mx(mx(-INFINITY, p1.y), p2.y), p3.y),
                 mx(mx(mx(-INFINITY, p1.z), p2.z), p3.z)};
        lo = {voxel_edge(lo.x, s, true), voxel_edge(lo.y, s, true), voxel_edge(lo.z, s, true)};
        hi = {voxel_edge(hi.x, s, false), voxel_edge(hi.y, s, false), voxel_edge(hi.z, s, false)};
        for (float x = lo.x; x < hi.x; x += s)
            for (float y = lo.y; y < hi.y; y += s)
                for (float z = lo.z; z < hi.z; z += s) {
                    P3 bmin = {x - hs, y - hs, z - hs}, bmax = {x + hs, y + hs, z + hs};
                    P3 c = (bmin + bmax) * 0.5f;
                    P3 h = {std::fabs(bmax.x - bmin.x) * 0.5f + precision, std::fabs(bmax.y - bmin.y) * 0.5f + precision,
                            std::fabs(bmax.z - bmin.z) * 0.5f + precision};
                    if (!tri_box_overlap(c, h, p1, p2, p3)) continue;
                    auto& bucket = table[voxel_bucket(c)];
                    bool dup = false;
                    for (const P3& q : bucket)
                        if (std::fabs(q.x - c.x) < 0.0000001 && std::fabs(q.y - c.y) < 0.0000001 &&
                            std::fabs(q.z - c.z) < 0.0000001) {
                            dup = true;
                            break;
                        }
                    if (!dup) bucket.push_back(c);
                }
    }
    const P3 corners[8] = {{-hs, hs, hs}, {-hs, -hs, hs}, {hs, -hs, hs}, {hs, hs, hs},
                           {-hs, hs, -hs}, {-hs, -hs, -hs}, {hs, -hs, -hs}, {hs, hs, -hs}};
    std::vector<float4> out;
    for (const auto& bucket : table)
        for (const P3& c : bucket)
            for (const P3& o : corners) put(out, o + c);
    return out;
}

// Model::surface_sampling: the voxel corners of the model's mesh at the
// particle spacing, with a precision of spacing / 10.
std::vector<float4> sample_model(const std::string& obj_path, float spacing) {
    if (!(spacing > 0.0f)) throw Error(SPH_ERR_INVALID, "model sampling: spacing must be > 0");
    std::vector<float> V;
    std::vector<uint32_t> I;
    load_obj(obj_path, V, I);
    return voxel_corners(V, I, spacing, spacing / 10);
}

}  // namespace sph
