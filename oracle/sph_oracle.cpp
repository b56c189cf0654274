// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference SPH step (paivett/fluids), used as the
// checker for the HIP library. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product never does.
//
// Arithmetic is pinned so that integer outputs and, with the same operation
// order, fp32 fields are reproducible bit for bit:
//   * IEEE fp32, round to nearest, no FMA contraction (-ffp-contract=off),
//     correctly rounded '/' and sqrtf (the reference builds with
//     -cl-fast-relaxed-math, whose results are implementation-defined; this
//     file fixes one IEEE interpretation of them);
//   * fast_length(r) == length(r) == sqrtf(dot(r,r));
//   * normalize(r) == fast_normalize(r) == r * (1 / length(r)), and 0 for
//     r == 0 (OpenCL 1.2 defines fast_normalize(p) as p * half_rsqrt(dot(p,p))
//     and leaves the zero vector unchanged; the reciprocal square root is
//     taken here as two correctly rounded operations);
//   * pown(x,2) = x*x, pown(x,3) = (x*x)*x, pown(x,7) = ((((((x*x)*x)*x)*x)*x)*x);
//   * dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z. The w component drops out:
//     every position on the path carries w == 1 and every velocity, force and
//     normal w == 0 (boxvolume.cpp:71, boundaryhandler.cpp:132,
//     boundary.cl:123, pcisph.cl:279-280), so all w terms of the reference's
//     float4 dot products are exactly +0;
//   * OpenCL max(x,y) = (x < y) ? y : x, min(x,y) = (y < x) ? y : x,
//     clamp(x,lo,hi) = min(max(x,lo),hi), sign() as in OpenCL 1.2.
//
// Every function cites the reference file:line it follows. Quirks Q1-Q21 of
// the reference are reproduced, not repaired, except the WCSPH build error
// (normals.cl uses PARTICLE_MASS, which only the PCISPH build defines:
// wcsphsimluation.cpp:368-370 vs normals.cl:69). The repair defines
// PARTICLE_MASS for the WCSPH program as the WCSPH particle mass printed to
// six decimals, exactly as the PCISPH build prints its own mass.
//
// A double-precision twin re-evaluates every field of one step on the same
// sorted state and the same neighbour lists, so the fp32 rounding error of
// each field can be stated.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

struct f4 { float x, y, z, w; };
template <class T> struct V3 { T x, y, z; };

// --------------------------------------------------------------------------
// Pinned scalar helpers (OpenCL 1.2 built-in semantics)
// --------------------------------------------------------------------------
template <class T> static inline T cl_max(T a, T b) { return (a < b) ? b : a; }
template <class T> static inline T cl_min(T a, T b) { return (b < a) ? b : a; }
template <class T> static inline T cl_clamp(T x, T lo, T hi) { return cl_min(cl_max(x, lo), hi); }
template <class T> static inline T cl_sign(T x) {
    if (x > T(0)) return T(1);
    if (x < T(0)) return T(-1);
    if (x != x) return T(0);
    return x;  // keeps the sign of zero
}
template <class T> static inline T sqr(T x) { return x * x; }
template <class T> static inline T cube(T x) { return (x * x) * x; }
template <class T> static inline T pown7(T x) {
    T r = x * x;
    r = r * x; r = r * x; r = r * x; r = r * x; r = r * x;
    return r;
}
static inline float  csqrt(float x) { return sqrtf(x); }
static inline double csqrt(double x) { return std::sqrt(x); }

template <class T> static inline V3<T> v3(const f4& a) { return {T(a.x), T(a.y), T(a.z)}; }
template <class T> static inline V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> static inline V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> static inline V3<T> mul(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> static inline V3<T> smul(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> static inline V3<T> vdiv(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> static inline T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> static inline T length(V3<T> a) { return csqrt(dot(a, a)); }
template <class T> static inline V3<T> normalize(V3<T> a) {
    T d = dot(a, a);
    if (d == T(0)) return a;
    T inv = T(1) / csqrt(d);
    return {a.x * inv, a.y * inv, a.z * inv};
}
template <class T> static inline V3<T> cl_clamp3(V3<T> a, T lo, T hi) {
    return {cl_clamp(a.x, lo, hi), cl_clamp(a.y, lo, hi), cl_clamp(a.z, lo, hi)};
}

// -D constant printing: opencl/clcompiler.cpp:50-54 (fixed, 6 decimals, "f").
static float define_literal(float v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.6f", (double)v);
    return std::strtof(buf, nullptr);
}

// --------------------------------------------------------------------------
// Settings (settings/physicssettings.h:5-25, settings/simulationsettings.h:9-39)
// --------------------------------------------------------------------------
struct Settings {
    float time_step, max_vel, fluid_particle_radius, fluid_support_radius;
    int pcisph_max_iterations;
    float pcisph_error_ratio;
    int method;  // 0 = WCSPH, 1 = PCISPH
    float rest_density, k_viscosity, gravity, gas_stiffness, surface_tension;
    int flags;  // bit 0: WCSPH density evaluates W_POLY6 on r^2 (repairs Q7)
};
static const int FIX_WCSPH_DENSITY = 1;

// --------------------------------------------------------------------------
// Grid: fluid/simulation/grid.cpp:8-12,61-68 and kernels/grid.h
// --------------------------------------------------------------------------
struct Grid {
    float size = 5.0f;       // Grid(5.0, h): pcisphsimluation.cpp:54, wcsphsimluation.cpp:26
    float cell_size = 0.f;   // full fp32 support radius (kernel argument)
    int cps = 0;             // ceil(size / cell) in fp32 (grid.cpp:64)
    long long cells = 0;     // pow(cps, 3.0) (grid.cpp:65)
    float sqr_support = 0.f; // CELL_SIZE*CELL_SIZE with the 6-decimal literal (grid.cl:3, grid.cpp:86)

    void reset(float h) {
        cell_size = h;
        cps = (int)std::ceil(size / cell_size);
        cells = (long long)std::pow((double)cps, 3.0);
        float lit = define_literal(cell_size);
        sqr_support = lit * lit;
    }
};

// split_by_3 with OpenCL 32-bit shift semantics (kernels/morton.h:37-46, Q1):
// x << 32 on a 32-bit uint is x << 0, and the 64-bit mask literals are
// truncated to their low 32 bits on assignment back to uint.
static inline uint32_t split_by_3(uint32_t a) {
    uint32_t x = a & 0x1fffffu;
    x = (x | x) & 0x0000ffffu;
    x = (x | (x << 16)) & 0xff0000ffu;
    x = (x | (x << 8)) & 0x0f00f00fu;
    x = (x | (x << 4)) & 0xc30c30c3u;
    x = (x | (x << 2)) & 0x49249249u;
    return x;
}
// kernels/morton.h:50-56
static inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return split_by_3(x) | (split_by_3(y) << 1) | (split_by_3(z) << 2);
}

// get_grid_coordinates: kernels/grid.h:74-91
static inline void grid_coords(const f4& p, const Grid& g, int c[3]) {
    float half = g.size / 2.0f;
    c[0] = (int)floorf((p.x + half) / g.cell_size);
    c[1] = (int)floorf((p.y + half) / g.cell_size);
    c[2] = (int)floorf((p.z + half) / g.cell_size);
    for (int k = 0; k < 3; ++k) c[k] = (c[k] + g.cps) % g.cps;
}
// CELL_ID: kernels/grid.h:7
static inline long long cell_id(long long x, long long y, long long z, int cps) {
    return (long long)cps * ((long long)cps * z + y) + x;
}

// CELL_NEIGH_OFFSET: kernels/grid.h:15-42. Own cell first, then the 26 others
// in lexicographic (dx, dy, dz) order.
struct Offsets {
    int d[27][3];
    Offsets() {
        d[0][0] = d[0][1] = d[0][2] = 0;
        int k = 1;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    if (!dx && !dy && !dz) continue;
                    d[k][0] = dx; d[k][1] = dy; d[k][2] = dz; ++k;
                }
    }
};
static const Offsets OFF;

static const int LIST_CAP = 50;  // NEIGH_LIST_MAX_LENGTH: kernels/common.h:4

// Sparse stand-in for the dense int2[cps^3] interval array: an absent cell
// reads as [0,0), exactly as the zero-filled array does (grid.cpp:41).
struct Ranges {
    std::unordered_map<long long, std::pair<int, int>> m;
    long long edge_reads = 0;  // reads outside [0, cps^3): undefined in the reference
    std::pair<int, int> get(long long id, long long cells) const {
        if (id < 0 || id >= cells) return {0, 0};
        auto it = m.find(id);
        return it == m.end() ? std::pair<int, int>{0, 0} : it->second;
    }
};

// compute_hashes: kernels/grid.cl:20-39 ; clsort (stable LSD radix sort by
// key, opencl/algorithms/clsort.h:12-35 -> clogs) ; clshuffle gather
// (opencl/algorithms/clshuffle.h:23-35) ; compute_cell_intervals
// (kernels/grid.cl:60-123).
struct SortResult {
    std::vector<uint32_t> keys;   // sorted keys
    std::vector<int> perm;        // mask: sorted i -> original index
};

static SortResult hash_and_sort(const std::vector<f4>& pos, const Grid& g) {
    const int n = (int)pos.size();
    std::vector<uint32_t> keys(n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        int c[3];
        grid_coords(pos[i], g, c);
        keys[i] = morton3((uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2]);
    }
    SortResult r;
    r.perm.resize(n);
    std::iota(r.perm.begin(), r.perm.end(), 0);
    std::stable_sort(r.perm.begin(), r.perm.end(),
                     [&](int a, int b) { return keys[a] < keys[b]; });
    r.keys.resize(n);
    for (int i = 0; i < n; ++i) r.keys[i] = keys[r.perm[i]];
    return r;
}

static void build_ranges(const std::vector<f4>& sorted_pos, const std::vector<uint32_t>& keys,
                         const Grid& g, Ranges& R) {
    R.m.clear();
    const int n = (int)sorted_pos.size();
    auto lin = [&](const f4& p) {
        int c[3];
        grid_coords(p, g, c);
        return cell_id(c[0], c[1], c[2], g.cps);
    };
    for (int i = 0; i < n; ++i) {
        long long c0 = lin(sorted_pos[i]);
        if (i == n - 1) {
            R.m[c0].second = i + 1;
        } else if (keys[i] < keys[i + 1]) {
            long long c1 = lin(sorted_pos[i + 1]);
            R.m[c0].second = i + 1;
            R.m[c1].first = i + 1;
        }
    }
}

// compute_neigh_list: kernels/grid.cl:145-181 (cap 50, own cell first,
// column-major list[k*N + i], strict r2 < CELL_SIZE^2, self included).
static void build_lists(const std::vector<f4>& ref, const std::vector<f4>& nb, const Ranges& R,
                        const Grid& g, std::vector<int>& list, std::vector<int>& cnt,
                        long long* edge_reads) {
    const int n = (int)ref.size();
    list.assign((size_t)LIST_CAP * n, -1);
    cnt.assign(n, 0);
    long long edges = 0;
#pragma omp parallel for schedule(static) reduction(+ : edges)
    for (int i = 0; i < n; ++i) {
        int c[3];
        grid_coords(ref[i], g, c);
        V3<float> pi = v3<float>(ref[i]);
        int len = 0;
        for (int o = 0; len < LIST_CAP && o < 27; ++o) {
            long long id = cell_id(c[0] + OFF.d[o][0], c[1] + OFF.d[o][1], c[2] + OFF.d[o][2], g.cps);
            if (id < 0 || id >= g.cells) ++edges;
            auto iv = R.get(id, g.cells);
            for (int j = iv.first; len < LIST_CAP && j < iv.second; ++j) {
                V3<float> r = sub(pi, v3<float>(nb[j]));
                float r2 = dot(r, r);
                if (r2 < g.sqr_support) {
                    list[(size_t)len * n + i] = j;
                    ++len;
                }
            }
        }
        cnt[i] = len;
    }
    if (edge_reads) *edge_reads += edges;
}

// --------------------------------------------------------------------------
// Host-side constants
// --------------------------------------------------------------------------
// MullerConstants: fluid/simulation/mullerconstants.cpp:4-38. The file calls
// the global ::pow, i.e. pow(double, double), and evaluates in double.
static float muller_default_eval(float h) { return (float)(365.0f / (64.0f * M_PI * ::pow((double)h, 9.0))); }
static float muller_default_grad(float h) { return (float)(-945.0f / (32.0f * M_PI * ::pow((double)h, 9.0))); }
static float muller_pressure_grad(float h) { return (float)(-45.0f / (M_PI * ::pow((double)h, 6.0))); }
static float muller_viscosity_lapl(float h) { return (float)(45.0f / (M_PI * ::pow((double)h, 6.0))); }

struct PciConsts {  // pcisphsimluation.cpp:230-255, 275-278, 501-505
    float mass;           // full fp32 host mass (rho0 (2r)^3 / 1.15)
    float mass_d;         // PARTICLE_MASS literal (Q5)
    float h_d;            // SUPPORT_RADIUS literal
    float rho0_d;         // REST_DENSITY literal
    float st_main, st_term;
    float default_eval, default_grad, visc_lapl, pressure_grad;
    float delta;          // density scale factor (Q16)
};
struct WcConsts {   // wcsphsimluation.cpp:165-202, 368-370
    float mass, h, sqr_h;
    float poly6_eval, poly6_grad, spiky_grad, visc_lapl;
    float st_main, st_term;
    float mass_d, h_d;  // literals of normals.cl (Q8 repair: PARTICLE_MASS defined)
};
struct BdryConsts {  // boundaryhandler.cpp:144, 307-309, 411
    float h;           // handler's support radius (full fp32)
    float h_d;         // SUPPORT_RADIUS literal of boundary.cl
    float poly6_eval;  // MullerConstants::default_eval(h)
    float spiky_grad;  // -45/(pi h^6) evaluated with std::pow(float,int) -> double
    float phi_coeff;   // 1.0 (pcisphsimluation.cpp:57-61, boundaryhandler.h:36)
};

// _deduce_density_scale_factor: pcisphsimluation.cpp:429-469 (Q16).
static float deduce_delta(float h, float r, float dt, float mass, float rho0) {
    float spiky = muller_pressure_grad(h);
    float poly6g = muller_default_grad(h);
    float beta = 2.0f * powf((dt * mass) / rho0, 2.0f);
    float h2 = h * h;
    float vs[3] = {0.f, 0.f, 0.f}, vss[3] = {0.f, 0.f, 0.f};
    float vdv = 0.f;
    for (float x = -h - r; x <= h + r; x += 2.0f * r)
        for (float y = -h - r; y <= h + r; y += 2.0f * r)
            for (float z = -h - r; z <= h + r; z += 2.0f * r) {
                float r2 = (x * x + y * y + z * z);
                float rn = sqrtf(r2);
                if (r2 < h2) {
                    float sf = (float)((double)(spiky * (h - rn) * (h - rn)) * (1.0 / (double)rn));
                    float f = poly6g * (h2 - r2) * (h2 - r2);
                    float sv[3] = {sf * x, sf * y, sf * z};
                    float v[3] = {f * x, f * y, f * z};
                    for (int k = 0; k < 3; ++k) {
                        vs[k] += v[k];
                        vss[k] += sv[k];
                        vdv += v[k] * sv[k];
                    }
                }
            }
    float sdot = 0.f;
    for (int k = 0; k < 3; ++k) sdot += vss[k] * vs[k];
    return -1.0f / (beta * (-sdot - vdv));
}

static PciConsts pci_consts(const Settings& s) {
    PciConsts c;
    float r = s.fluid_particle_radius, h = s.fluid_support_radius;
    c.mass = s.rest_density * powf(2.0f * r, 3.0f) / 1.15f;
    c.st_main = (float)(32.0f / (M_PI * (double)powf(h, 9.0f)));
    c.st_term = powf(h, 6.0f) / 64.0f;
    c.mass_d = define_literal(c.mass);
    c.h_d = define_literal(h);
    c.rho0_d = define_literal(s.rest_density);
    c.default_eval = muller_default_eval(h);
    c.default_grad = muller_default_grad(h);
    c.visc_lapl = muller_viscosity_lapl(h);
    c.pressure_grad = muller_pressure_grad(h);
    c.delta = deduce_delta(h, r, s.time_step, c.mass, s.rest_density);
    return c;
}

static WcConsts wc_consts(const Settings& s) {
    WcConsts c;
    float r = s.fluid_particle_radius, h = s.fluid_support_radius;
    c.mass = s.rest_density * powf(2.0f * r, 3.0f) / 1.0f;
    c.h = h;
    c.sqr_h = h * h;
    c.poly6_eval = (float)(315.0 / (64.0 * M_PI * std::pow((double)h, 9)));
    c.poly6_grad = (float)(-945.0 / (32.0 * M_PI * std::pow((double)h, 9)));
    c.spiky_grad = (float)(-45.0 / (M_PI * std::pow((double)h, 6)));
    c.visc_lapl = (float)(45.0 / (M_PI * std::pow((double)h, 6)));
    c.st_main = (float)(32.0f / (M_PI * (double)powf(h, 9.0f)));
    c.st_term = powf(h, 6.0f) / 64.0f;
    c.mass_d = define_literal(c.mass);
    c.h_d = define_literal(h);
    return c;
}

static BdryConsts bdry_consts(const Settings& s) {
    BdryConsts c;
    c.h = s.fluid_support_radius;
    c.h_d = define_literal(c.h);
    c.poly6_eval = muller_default_eval(c.h);
    c.spiky_grad = (float)(-45.0f / (M_PI * std::pow((double)c.h, 6)));
    c.phi_coeff = 1.0f;
    return c;
}

// --------------------------------------------------------------------------
// Boundary bodies
// --------------------------------------------------------------------------
struct Body {
    int origin = 0, count = 0;
    bool can_move = false;
    float mass = 0.f;
    float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float com[3] = {0, 0, 0};
    float lin[3] = {0, 0, 0}, ang[3] = {0, 0, 0};
    float force[3] = {0, 0, 0}, torque[3] = {0, 0, 0};  // clsum results
    double force64[3] = {0, 0, 0}, torque64[3] = {0, 0, 0};
};

// compute_boundary_phi: kernels/boundary.cl:4-31, brute force per body (Q14).
static void boundary_phi_brute(const f4* p, int n, const BdryConsts& c, float rho0, float* phi) {
    float h2 = sqr(c.h_d);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        V3<float> pi = v3<float>(p[i]);
        float delta = 0;
        for (int j = 0; j < n; ++j) {
            V3<float> r = sub(pi, v3<float>(p[j]));
            float r2 = dot(r, r);
            if (r2 < h2) delta += cube(sqr(c.h_d) - r2);
        }
        phi[i] = (rho0 / (delta * c.poly6_eval)) / c.phi_coeff;
    }
}
// Same sum, evaluated through a cell grid: candidates j are collected from
// the 27 cells around i and summed in increasing j, which is the order of the
// brute-force loop. Used for bodies too large for O(B^2) on the CPU.
static void boundary_phi_grid(const f4* p, int n, const BdryConsts& c, float rho0, float* phi) {
    float h2 = sqr(c.h_d);
    float cell = c.h_d;
    std::unordered_map<long long, std::vector<int>> cells;
    auto key = [&](int x, int y, int z) { return ((long long)x * 2000003LL + y) * 2000003LL + z; };
    std::vector<int> cx(n), cy(n), cz(n);
    for (int i = 0; i < n; ++i) {
        cx[i] = (int)floorf(p[i].x / cell);
        cy[i] = (int)floorf(p[i].y / cell);
        cz[i] = (int)floorf(p[i].z / cell);
        cells[key(cx[i], cy[i], cz[i])].push_back(i);
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        std::vector<int> cand;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = cells.find(key(cx[i] + dx, cy[i] + dy, cz[i] + dz));
                    if (it != cells.end()) cand.insert(cand.end(), it->second.begin(), it->second.end());
                }
        std::sort(cand.begin(), cand.end());
        V3<float> pi = v3<float>(p[i]);
        float delta = 0;
        for (int j : cand) {
            V3<float> r = sub(pi, v3<float>(p[j]));
            float r2 = dot(r, r);
            if (r2 < h2) delta += cube(sqr(c.h_d) - r2);
        }
        phi[i] = (rho0 / (delta * c.poly6_eval)) / c.phi_coeff;
    }
}

// transform_particles: kernels/boundary.cl:33-59 (position part; the
// velocities it writes are never read on the path).
static f4 transform_point(const f4& raw, const Body& b) {
    V3<float> p = v3<float>(raw);
    V3<float> m0 = {b.R[0], b.R[1], b.R[2]}, m1 = {b.R[3], b.R[4], b.R[5]}, m2 = {b.R[6], b.R[7], b.R[8]};
    f4 o;
    o.x = dot(m0, p) + b.com[0];
    o.y = dot(m1, p) + b.com[1];
    o.z = dot(m2, p) + b.com[2];
    o.w = 1.0f;
    return o;
}

// --------------------------------------------------------------------------
// Field storage
// --------------------------------------------------------------------------
template <class T> struct Fields {
    std::vector<T> density, pressure, err;
    std::vector<V3<T>> normal, f_other, f_press, pos_pred, accel;
    std::vector<V3<T>> out_pos, out_vel, out_vel_half;  // integrator outputs, sorted index
    std::vector<V3<T>> b_force, b_torque;               // per boundary particle (unsorted)
    int iterations = 0;
};

enum FieldId {
    F_KEYS = 0, F_PERM, F_NCOUNT, F_NLIST, F_BNCOUNT, F_BNLIST,
    F_POS_SORTED, F_VEL_SORTED, F_DENSITY, F_PRESSURE, F_ERR, F_NORMAL,
    F_FORCE_OTHER, F_FORCE_PRESSURE, F_POS_PRED, F_ACCEL,
    F_POS, F_VEL, F_VEL_HALF, F_BPOS_SORTED, F_BPHI_SORTED, F_BKEYS, F_BPERM,
    F_BPHI, F_BFORCE, F_BTORQUE, F_RANGES, F_BRANGES, F_VEL_HALF_SORTED,
    F_COUNT_
};

// --------------------------------------------------------------------------
// The simulation (PCISPHSimulation / WCSPHSimulation restated)
// --------------------------------------------------------------------------
struct Sim {
    Settings S{};
    Grid grid;
    PciConsts pc{};
    WcConsts wc{};
    BdryConsts bc{};
    float container[3] = {0, 0, 0};
    float g[3] = {0, 0, 0};
    bool twin = false;
    int force_iterations = -1;  // for the twin: replay the fp32 iteration count

    std::vector<std::vector<f4>> volumes;
    std::vector<f4> pos, vel, vel_half;  // "unsorted" buffers (stored order)

    std::vector<Body> bodies;
    std::vector<f4> b_raw, b_unsorted;
    std::vector<float> b_phi;

    // per-step products
    SortResult fs, bs;
    Ranges fr, br;
    std::vector<f4> pos_s, vel_s, vel_half_s, b_pos_s;
    std::vector<float> b_phi_s;
    std::vector<int> nl, nc, bnl, bnc;
    Fields<float> F;
    Fields<double> D;
    long long edge_reads = 0;
    int steps = 0;

    int N() const { return (int)pos.size(); }
    int B() const { return (int)b_raw.size(); }

    void init_params() {
        grid.reset(S.fluid_support_radius);
        pc = pci_consts(S);
        wc = wc_consts(S);
        bc = bdry_consts(S);
        g[0] = 0; g[1] = -S.gravity; g[2] = 0;
    }

    // BoundaryHandler::_sample_surfaces -> _compute_boundary_phi -> sync(true)
    // (boundaryhandler.cpp:119-153, 238-294)
    void compute_phi() {
        b_phi.assign(B(), 0.f);
        for (auto& b : bodies) {
            if (b.count <= 4096)
                boundary_phi_brute(&b_raw[b.origin], b.count, bc, S.rest_density, &b_phi[b.origin]);
            else
                boundary_phi_grid(&b_raw[b.origin], b.count, bc, S.rest_density, &b_phi[b.origin]);
        }
        b_unsorted = b_raw;
        boundary_sync(true);
    }

    // BoundaryHandler::sync + _sort_positions (boundaryhandler.cpp:216-236, 251-294)
    void boundary_sync(bool all) {
        if (B() == 0) return;
        for (auto& b : bodies) {
            if (b.can_move || all) {
                for (int i = 0; i < b.count; ++i) b_unsorted[b.origin + i] = transform_point(b_raw[b.origin + i], b);
            }
        }
        bs = hash_and_sort(b_unsorted, grid);
        b_pos_s.resize(B());
        b_phi_s.resize(B());
        for (int i = 0; i < B(); ++i) {
            b_pos_s[i] = b_unsorted[bs.perm[i]];
            b_phi_s[i] = b_phi[bs.perm[i]];
        }
        build_ranges(b_pos_s, bs.keys, grid, br);
    }

    void neighbour_search() {
        fs = hash_and_sort(pos, grid);
        const int n = N();
        pos_s.resize(n); vel_s.resize(n);
        for (int i = 0; i < n; ++i) {
            pos_s[i] = pos[fs.perm[i]];
            vel_s[i] = vel[fs.perm[i]];
        }
        if (S.method == 0) {
            vel_half_s.resize(n);
            for (int i = 0; i < n; ++i) vel_half_s[i] = vel_half[fs.perm[i]];
        }
        build_ranges(pos_s, fs.keys, grid, fr);
        build_lists(pos_s, pos_s, fr, grid, nl, nc, &edge_reads);
        if (B() > 0) {
            build_lists(pos_s, b_pos_s, br, grid, bnl, bnc, &edge_reads);
        } else {
            bnl.assign((size_t)LIST_CAP * n, -1);
            bnc.assign(n, 0);  // build_neighbourhood fills zeros: boundaryhandler.cpp:368-370
        }
    }

    // ---------------------------------------------------------------- PCISPH
    template <class T> void pcisph_fields(Fields<T>& O, int fixed_iters) {
        const int n = N();
        const bool has_b = B() > 0;
        const T H = T(pc.h_d), H2 = sqr(H);
        const T M = T(pc.mass_d), RHO0 = T(pc.rho0_d);
        const T dt = T(S.time_step), maxv = T(S.max_vel);
        const T cx = T(container[0]), cy = T(container[1]), cz = T(container[2]);
        const V3<T> G = {T(g[0]), T(g[1]), T(g[2])};
        O.density.assign(n, 0); O.pressure.assign(n, 0); O.err.assign(n, 0);
        O.normal.assign(n, {0, 0, 0}); O.f_other.assign(n, {0, 0, 0});
        O.f_press.assign(n, {0, 0, 0}); O.pos_pred.assign(n, {0, 0, 0});
        O.out_pos.assign(n, {0, 0, 0}); O.out_vel.assign(n, {0, 0, 0});

        // compute_initial_density: pcisph.cl:33-102
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            V3<T> pi = v3<T>(pos_s[i]);
            T d = 0, bd = 0;
            for (int k = 0; k < nc[i]; ++k) {
                V3<T> r = sub(pi, v3<T>(pos_s[nl[(size_t)k * n + i]]));
                d += cube(H2 - dot(r, r));
            }
            if (has_b)
                for (int k = 0; k < bnc[i]; ++k) {
                    int j = bnl[(size_t)k * n + i];
                    V3<T> r = sub(pi, v3<T>(b_pos_s[j]));
                    bd += cube(H2 - dot(r, r)) * T(b_phi_s[j]);
                }
            O.density[i] = T(pc.default_eval) * ((d * M) + bd);
        }
        // compute_normals: normals.cl:23-70
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            V3<T> pi = v3<T>(pos_s[i]);
            V3<T> nn = {0, 0, 0};
            for (int k = 0; k < nc[i]; ++k) {
                int j = nl[(size_t)k * n + i];
                V3<T> r = sub(pi, v3<T>(pos_s[j]));
                T q = sqr(H2 - dot(r, r));
                nn = add(nn, vdiv(mul(r, q), O.density[j]));
            }
            V3<T> o = smul(T(pc.default_grad), nn);
            o = mul(o, M);
            O.normal[i] = mul(o, H);
        }
        // compute_initial_forces: pcisph.cl:134-222
        const T k_visc = T(S.k_viscosity), sigma = T(S.surface_tension);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            V3<T> pi = v3<T>(pos_s[i]), vi = v3<T>(vel_s[i]);
            T di = O.density[i];
            V3<T> ni = O.normal[i];
            V3<T> fv = {0, 0, 0}, fcoh = {0, 0, 0}, fcurv = {0, 0, 0};
            for (int k = 0; k < nc[i]; ++k) {
                int j = nl[(size_t)k * n + i];
                V3<T> r = sub(pi, v3<T>(pos_s[j]));
                T l = length(r);
                T dj = O.density[j];
                fv = add(fv, mul(sub(v3<T>(vel_s[j]), vi), (H - l) / dj));
                T corr = (T(2) * RHO0) / (di + dj);
                fcurv = add(fcurv, mul(sub(ni, O.normal[j]), corr));
                // st_kernel: surface_tension.h:23-36
                T stk;
                if (l <= T(0.5f) * H) {
                    T t1 = H - l;
                    T t2 = cube(t1) * cube(l);
                    stk = T(2) * t2 - T(pc.st_term);
                } else if (l <= H) {
                    T t1 = H - l;
                    stk = cube(t1) * cube(l);
                } else {
                    stk = 0;
                }
                fcoh = add(fcoh, mul(mul(normalize(r), stk), corr));
            }
            T sv = (k_visc * sqr(M)) * (T(pc.visc_lapl) / di);
            fv = mul(fv, sv);
            fcoh = mul(fcoh, M * T(pc.st_main));
            T st = (-sigma) * M;
            V3<T> ft = smul(st, add(fcurv, fcoh));
            O.f_other[i] = add(fv, ft);
        }
        // correction loop: pcisphsimluation.cpp:129-148
        const T delta = T(pc.delta);
        const int min_iter = 3, max_iter = S.pcisph_max_iterations;
        const T wpg = T(pc.pressure_grad), wde = T(pc.default_eval);
        const T eta = T(S.pcisph_error_ratio), rho0h = T(S.rest_density);
        int it = 0;
        for (int i = 1; i <= max_iter; ++i) {
            it = i;
            // predict_pos: pcisph.cl:302-344
#pragma omp parallel for schedule(static)
            for (int p = 0; p < n; ++p) {
                V3<T> x = v3<T>(pos_s[p]), v = v3<T>(vel_s[p]);
                V3<T> acc = add(G, vdiv(add(O.f_press[p], O.f_other[p]), M));
                v = add(v, cl_clamp3(mul(acc, dt), -maxv, maxv));
                x = add(x, mul(v, dt));
                T vn = length(v);
                V3<T> cp = x;
                cp.z = cl_clamp(x.z, -cz / T(2), cz / T(2));
                cp.x = cl_clamp(x.x, -cx / T(2), cx / T(2));
                cp.y = cl_clamp(x.y, -cy / T(2), T(10));
                T dd = length(sub(cp, x));
                if (dd > T(0) && vn > T(0)) x = cp;
                O.pos_pred[p] = x;
            }
            // update_pressure: pcisph.cl:349-418
            T mx = 0;
#pragma omp parallel for schedule(static) reduction(max : mx)
            for (int p = 0; p < n; ++p) {
                V3<T> xi = O.pos_pred[p];
                T pd = 0, pdb = 0;
                for (int k = 0; k < nc[p]; ++k) {
                    V3<T> r = sub(xi, O.pos_pred[nl[(size_t)k * n + p]]);
                    pd += cube(H2 - dot(r, r));
                }
                pd *= wde * M;
                if (has_b) {
                    for (int k = 0; k < bnc[p]; ++k) {
                        int j = bnl[(size_t)k * n + p];
                        V3<T> r = sub(xi, v3<T>(b_pos_s[j]));
                        pdb += cube(H2 - dot(r, r)) * T(b_phi_s[j]);
                    }
                    pdb *= wde;
                }
                T dv = cl_max(T(0), (pd + pdb) - RHO0);
                O.err[p] = dv;
                O.pressure[p] += dv * delta;
                if (dv > mx) mx = dv;
            }
            // compute_pressure_force: pcisph.cl:420-504
#pragma omp parallel for schedule(static)
            for (int p = 0; p < n; ++p) {
                V3<T> xi = v3<T>(pos_s[p]);
                T di = O.density[p];
                T C = O.pressure[p] / sqr(di);
                V3<T> f = {0, 0, 0}, fb = {0, 0, 0};
                for (int k = 0; k < nc[p]; ++k) {
                    int j = nl[(size_t)k * n + p];
                    V3<T> r = sub(xi, v3<T>(pos_s[j]));
                    T l = length(r);
                    if (l >= T(1e-5f)) {
                        V3<T> gw = mul(mul(r, T(1) / l), sqr(H - l));
                        f = add(f, smul(C + O.pressure[j] / sqr(O.density[j]), gw));
                    }
                }
                f = mul(f, (-sqr(M)) * wpg);
                if (has_b) {
                    for (int k = 0; k < bnc[p]; ++k) {
                        int j = bnl[(size_t)k * n + p];
                        V3<T> r = sub(xi, v3<T>(b_pos_s[j]));
                        T l = length(r);
                        V3<T> gw = mul(mul(r, T(1) / l), sqr(H - l));
                        fb = add(fb, smul(T(b_phi_s[j]) * (T(2) * C), gw));
                    }
                    fb = mul(fb, (-M) * wpg);
                }
                O.f_press[p] = add(f, fb);
            }
            if (fixed_iters > 0) {
                if (i >= fixed_iters) break;
            } else if (i >= min_iter) {
                // _get_max_density_variation + exit test (Q9)
                if ((mx - rho0h) / rho0h < eta) break;
            }
        }
        O.iterations = it;
    }

    // predict_vel_n_pos (PCISPH): pcisph.cl:240-284
    template <class T> void pcisph_integrate(Fields<T>& O) {
        const int n = N();
        const T M = T(pc.mass_d), dt = T(S.time_step), maxv = T(S.max_vel);
        const T cx = T(container[0]), cy = T(container[1]), cz = T(container[2]);
        const V3<T> G = {T(g[0]), T(g[1]), T(g[2])};
#pragma omp parallel for schedule(static)
        for (int p = 0; p < n; ++p) {
            V3<T> x = v3<T>(pos_s[p]), v = v3<T>(vel_s[p]);
            V3<T> acc = add(G, vdiv(add(O.f_press[p], O.f_other[p]), M));
            v = add(v, cl_clamp3(mul(acc, dt), -maxv, maxv));
            x = add(x, mul(v, dt));
            T vn = length(v);
            V3<T> cp = x;
            cp.z = cl_clamp(x.z, -cz / T(2), cz / T(2));
            cp.x = cl_clamp(x.x, -cx / T(2), cx / T(2));
            cp.y = cl_clamp(x.y, -cy / T(2), T(10));
            V3<T> dvec = sub(cp, x);
            T dd = length(dvec);
            if (dd > T(0) && vn > T(0)) {
                V3<T> nrm = normalize(V3<T>{cl_sign(dvec.x), cl_sign(dvec.y), cl_sign(dvec.z)});
                T rest = dd / (dt * vn);
                T s = (T(1) + rest) * dot(v, nrm);
                v = sub(v, smul(s, nrm));
                x = cp;
            }
            O.out_pos[p] = x;
            O.out_vel[p] = v;
        }
    }

    // ----------------------------------------------------------------- WCSPH
    template <class T> void wcsph_fields(Fields<T>& O) {
        const int n = N();
        const bool has_b = B() > 0;
        const T h = T(wc.h);
        O.density.assign(n, 0); O.pressure.assign(n, 0);
        O.normal.assign(n, {0, 0, 0}); O.accel.assign(n, {0, 0, 0});
        O.out_pos.assign(n, {0, 0, 0}); O.out_vel.assign(n, {0, 0, 0}); O.out_vel_half.assign(n, {0, 0, 0});
        // compute_density_pressure: wcsph.cl:32-79 (Q7: W_POLY6 of the distance)
        const T k_gas = T(S.gas_stiffness), rho0 = T(S.rest_density);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            V3<T> pi = v3<T>(pos_s[i]);
            T d = 0, bd = 0;
            // Q7: the reference passes the distance where W_POLY6 expects r^2
            // (wcsph.cl:61-62,70-71); FIX_WCSPH_DENSITY passes r^2 instead.
            const bool fix = (S.flags & FIX_WCSPH_DENSITY) != 0;
            for (int k = 0; k < nc[i]; ++k) {
                V3<T> r = sub(pi, v3<T>(pos_s[nl[(size_t)k * n + i]]));
                T r2 = dot(r, r);
                T rn = fix ? r2 : csqrt(r2);
                d += cube(h * h - rn);
            }
            d *= T(wc.poly6_eval) * T(wc.mass);
            for (int k = 0; k < bnc[i]; ++k) {
                int j = bnl[(size_t)k * n + i];
                V3<T> r = sub(pi, v3<T>(b_pos_s[j]));
                T r2 = dot(r, r);
                T rn = fix ? r2 : csqrt(r2);
                bd += cube(h * h - rn) * T(b_phi_s[j]);
            }
            d += T(wc.poly6_eval) * bd;
            O.density[i] = d;
            O.pressure[i] = cl_max(T(0), k_gas * (pown7(d / rho0) - T(1)));
        }
        (void)has_b;
        // compute_normals: normals.cl:23-70 with the WCSPH literals (Q8 repair)
        const T Hd = T(wc.h_d), H2d = sqr(Hd), Md = T(wc.mass_d);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            V3<T> pi = v3<T>(pos_s[i]);
            V3<T> nn = {0, 0, 0};
            for (int k = 0; k < nc[i]; ++k) {
                int j = nl[(size_t)k * n + i];
                V3<T> r = sub(pi, v3<T>(pos_s[j]));
                T q = sqr(H2d - dot(r, r));
                nn = add(nn, vdiv(mul(r, q), O.density[j]));
            }
            V3<T> o = smul(T(wc.poly6_grad), nn);
            o = mul(o, Md);
            O.normal[i] = mul(o, Hd);
        }
        // compute_acceleration: wcsph.cl:81-173
        const T mass = T(wc.mass), sg = T(wc.spiky_grad), vl = T(wc.visc_lapl);
        const T k_visc = T(S.k_viscosity), sigma = T(S.surface_tension);
        const V3<T> G = {T(g[0]), T(g[1]), T(g[2])};
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            V3<T> pi = v3<T>(pos_s[i]), vi = v3<T>(vel_s[i]);
            T di = O.density[i], pri = O.pressure[i];
            V3<T> ni = O.normal[i];
            T C = pri / sqr(di);
            V3<T> bacc = {0, 0, 0};
            for (int k = 0; k < bnc[i]; ++k) {
                int j = bnl[(size_t)k * n + i];
                V3<T> r = sub(pi, v3<T>(b_pos_s[j]));
                T rn = length(r);
                V3<T> t = smul(-T(b_phi_s[j]), r);
                t = mul(t, sg / rn);
                t = mul(t, sqr(h - rn));
                t = mul(t, T(2));
                t = mul(t, C);
                bacc = add(bacc, t);
            }
            V3<T> pacc = {0, 0, 0}, vacc = {0, 0, 0}, fcurv = {0, 0, 0}, fcoh = {0, 0, 0};
            for (int k = 0; k < nc[i]; ++k) {
                int j = nl[(size_t)k * n + i];
                T dj = O.density[j], prj = O.pressure[j];
                V3<T> r = sub(pi, v3<T>(pos_s[j]));
                T rn = length(r);
                if (di > T(0) && dj > T(0) && (double)rn > 0.00001) {
                    V3<T> t = mul(mul(mul(r, sg / rn), sqr(h - rn)), C + prj / sqr(dj));
                    pacc = add(pacc, t);
                    vacc = add(vacc, vdiv(mul(sub(v3<T>(vel_s[j]), vi), h - rn), dj));
                    T corr = (T(2) * rho0) / (di + dj);
                    fcurv = add(fcurv, mul(sub(ni, O.normal[j]), corr));
                    T stk;
                    if (rn <= T(0.5f) * h) {
                        T t1 = h - rn;
                        T t2 = cube(t1) * cube(rn);
                        stk = T(2) * t2 - T(wc.st_term);
                    } else if (rn <= h) {
                        T t1 = h - rn;
                        stk = cube(t1) * cube(rn);
                    } else {
                        stk = 0;
                    }
                    fcoh = add(fcoh, mul(mul(normalize(r), stk), corr));
                }
            }
            pacc = mul(pacc, -mass);
            vacc = mul(vacc, vl * k_visc);
            fcoh = mul(fcoh, T(wc.st_main));
            V3<T> ast = smul(-sigma, add(fcurv, fcoh));
            O.accel[i] = add(add(add(add(G, ast), pacc), vacc), bacc);
        }
    }

    // predict_vel_n_pos (WCSPH): wcsph.cl:176-224 (Q12)
    template <class T> void wcsph_integrate(Fields<T>& O) {
        const int n = N();
        const T dt = T(S.time_step), maxv = T(S.max_vel);
        const T cx = T(container[0]), cy = T(container[1]), cz = T(container[2]);
#pragma omp parallel for schedule(static)
        for (int p = 0; p < n; ++p) {
            V3<T> x = v3<T>(pos_s[p]);
            V3<T> vh = v3<T>(vel_half_s[p]);
            V3<T> a = O.accel[p];
            vh = add(vh, smul(dt, a));
            x = add(x, smul(dt, cl_clamp3(add(vh, smul(dt, a)), -maxv, maxv)));
            T vn = length(vh);
            V3<T> cp = x;
            cp.z = cl_clamp(x.z, -cz / T(2), cz / T(2));
            cp.x = cl_clamp(x.x, -cx / T(2), cx / T(2));
            cp.y = cl_clamp(x.y, -cy / T(2), T(10));
            V3<T> dvec = sub(cp, x);
            T dd = length(dvec);
            if (dd > T(0) && vn > T(0)) {
                V3<T> nrm = normalize(V3<T>{cl_sign(dvec.x), cl_sign(dvec.y), cl_sign(dvec.z)});
                T rest = (T(1) * dd) / (dt * vn);
                T s = (T(1) + rest) * dot(vh, nrm);
                vh = sub(vh, smul(s, nrm));
                x = cp;
            }
            V3<T> vt = add(vh, mul(smul(dt, a), T(0.5f)));
            O.out_pos[p] = x;
            O.out_vel_half[p] = cl_clamp3(vh, -maxv, maxv);
            O.out_vel[p] = cl_clamp3(vt, -maxv, maxv);
        }
    }

    // compute_fluid_force + clsum: boundary.cl:61-118, boundaryhandler.cpp:296-349
    template <class T> void fluid_forces(Fields<T>& O) {
        O.b_force.assign(B(), {0, 0, 0});
        O.b_torque.assign(B(), {0, 0, 0});
        if (B() == 0) return;
        const T H = T(bc.h_d), sg = T(bc.spiky_grad);
        const T mass = T(S.method == 1 ? pc.mass : wc.mass);
        for (auto& b : bodies) {
            if (!(b.mass > 0.0f)) continue;
            const V3<T> c = {T(b.com[0]), T(b.com[1]), T(b.com[2])};
#pragma omp parallel for schedule(static)
            for (int q = 0; q < b.count; ++q) {
                int bi = b.origin + q;
                const f4& pb = b_unsorted[bi];
                V3<T> pi = v3<T>(pb);
                int cc[3];
                grid_coords(pb, grid, cc);
                V3<T> f = {0, 0, 0};
                int cnt = 0;
                for (int o = 0; cnt < LIST_CAP && o < 27; ++o) {
                    long long id = cell_id(cc[0] + OFF.d[o][0], cc[1] + OFF.d[o][1], cc[2] + OFF.d[o][2], grid.cps);
                    auto iv = fr.get(id, grid.cells);
                    for (int j = iv.first; cnt < 50 && j < iv.second; ++j) {
                        // membership decided in fp32 for both precisions
                        V3<float> r32 = sub(v3<float>(pos_s[j]), v3<float>(pb));
                        float rn32 = length(r32);
                        if (rn32 < bc.h_d) {
                            V3<T> r = sub(v3<T>(pos_s[j]), pi);
                            T rn = length(r);
                            T C = O.pressure[j];
                            T dj = O.density[j];
                            C = C / sqr(dj);
                            ++cnt;
                            V3<T> t = mul(mul(mul(mul(r, sg / rn), sqr(H - rn)), T(2)), C);
                            f = add(f, t);
                        }
                    }
                }
                f = mul(f, mass * T(b_phi[bi]));
                V3<T> a = sub(pi, c);
                O.b_force[bi] = f;
                O.b_torque[bi] = {a.y * f.z - a.z * f.y, a.z * f.x - a.x * f.z, a.x * f.y - a.y * f.x};
            }
            // clsum (Boost.Compute reduce; its summation order is not pinned).
            // fp64: summed in index order. fp32: one pinned order, that of a
            // 64-lane butterfly: within each run of 64 samples (past the
            // body's end, +0), level by level, x[l] += x[l ^ off] for
            // off = 32, 16, ..., 1, keeping x[0]; the four run sums of each
            // 256 samples are added left to right, and those block sums left
            // to right starting from +0.
            if constexpr (sizeof(T) == 8) {
                double sf[3] = {0, 0, 0}, st[3] = {0, 0, 0};
                for (int q = 0; q < b.count; ++q) {
                    int bi = b.origin + q;
                    sf[0] += (double)O.b_force[bi].x; sf[1] += (double)O.b_force[bi].y; sf[2] += (double)O.b_force[bi].z;
                    st[0] += (double)O.b_torque[bi].x; st[1] += (double)O.b_torque[bi].y;
                    st[2] += (double)O.b_torque[bi].z;
                }
                for (int k = 0; k < 3; ++k) { b.force64[k] = sf[k]; b.torque64[k] = st[k]; }
            } else {
                float acc[6] = {0, 0, 0, 0, 0, 0};
                for (int blk = 0; blk * 256 < b.count; ++blk) {
                    float bsum[6] = {0, 0, 0, 0, 0, 0};
                    for (int w = 0; w < 4; ++w) {
                        float x[6][64];
                        for (int l = 0; l < 64; ++l) {
                            const int q = blk * 256 + w * 64 + l;
                            const bool in = q < b.count;
                            const int bi = b.origin + q;
                            x[0][l] = in ? (float)O.b_force[bi].x : 0.0f;
                            x[1][l] = in ? (float)O.b_force[bi].y : 0.0f;
                            x[2][l] = in ? (float)O.b_force[bi].z : 0.0f;
                            x[3][l] = in ? (float)O.b_torque[bi].x : 0.0f;
                            x[4][l] = in ? (float)O.b_torque[bi].y : 0.0f;
                            x[5][l] = in ? (float)O.b_torque[bi].z : 0.0f;
                        }
                        for (int c = 0; c < 6; ++c) {
                            for (int off = 32; off > 0; off >>= 1) {
                                float y[64];
                                for (int l = 0; l < 64; ++l) y[l] = x[c][l] + x[c][l ^ off];
                                for (int l = 0; l < 64; ++l) x[c][l] = y[l];
                            }
                            bsum[c] = w == 0 ? x[c][0] : bsum[c] + x[c][0];
                        }
                    }
                    for (int c = 0; c < 6; ++c) acc[c] += bsum[c];
                }
                for (int k = 0; k < 3; ++k) { b.force[k] = acc[k]; b.torque[k] = acc[3 + k]; }
            }
        }
    }

    // One full step: PCISPHSimulation::_simulate_pcisph_step
    // (pcisphsimluation.cpp:72-165) or WCSPHSimulation::simulate
    // (wcsphsimluation.cpp:40-106).
    void step(int max_iter_override = -1) {
        boundary_sync(false);
        neighbour_search();
        const int n = N();
        if (S.method == 1) {
            int saved = S.pcisph_max_iterations;
            if (max_iter_override > 0) S.pcisph_max_iterations = max_iter_override;
            pcisph_fields<float>(F, force_iterations);
            fluid_forces<float>(F);
            pcisph_integrate<float>(F);
            if (twin) {
                pcisph_fields<double>(D, F.iterations);
                fluid_forces<double>(D);
                pcisph_integrate<double>(D);
            }
            S.pcisph_max_iterations = saved;
            for (int p = 0; p < n; ++p) {
                pos[p] = {F.out_pos[p].x, F.out_pos[p].y, F.out_pos[p].z, 1.0f};
                vel[p] = {F.out_vel[p].x, F.out_vel[p].y, F.out_vel[p].z, 0.0f};
            }
        } else {
            wcsph_fields<float>(F);
            fluid_forces<float>(F);
            wcsph_integrate<float>(F);
            if (twin) {
                wcsph_fields<double>(D);
                fluid_forces<double>(D);
                wcsph_integrate<double>(D);
            }
            for (int p = 0; p < n; ++p) {
                pos[p] = {F.out_pos[p].x, F.out_pos[p].y, F.out_pos[p].z, pos_s[p].w};
                vel[p] = {F.out_vel[p].x, F.out_vel[p].y, F.out_vel[p].z, vel_s[p].w};
                vel_half[p] = {F.out_vel_half[p].x, F.out_vel_half[p].y, F.out_vel_half[p].z, vel_half_s[p].w};
            }
        }
        ++steps;
    }

    // PCISPHSimulation::_initialize_solver / WCSPHSimulation::_initialize_solver
    // (pcisphsimluation.cpp:532-551, wcsphsimluation.cpp:393-405)
    void initialize_solver() {
        init_params();
        pos.clear();
        for (auto& v : volumes) pos.insert(pos.end(), v.begin(), v.end());
        vel.assign(pos.size(), {0, 0, 0, 0});
        vel_half.assign(pos.size(), {0, 0, 0, 0});
        if (S.method == 1 && N() > 0) {
            // relaxation step, then zero velocities (Q21)
            bool tw = twin;
            twin = false;
            step(10000);
            twin = tw;
            steps = 0;
            vel.assign(pos.size(), {0, 0, 0, 0});
        }
    }
};

}  // namespace oracle

// ----------------------------------------------------------------------------
// C API for the tests (ctypes)
// ----------------------------------------------------------------------------
using namespace oracle;

extern "C" {

typedef struct {
    float time_step, max_vel, fluid_particle_radius, fluid_support_radius;
    int pcisph_max_iterations;
    float pcisph_error_ratio;
    int method;
    float rest_density, k_viscosity, gravity, gas_stiffness, surface_tension;
    int flags;
} oracle_settings;

void* oracle_create(const oracle_settings* s) {
    Sim* sim = new Sim();
    std::memcpy(&sim->S, s, sizeof(Settings));
    sim->init_params();
    return sim;
}
void oracle_destroy(void* h) { delete (Sim*)h; }
void oracle_set_twin(void* h, int on) { ((Sim*)h)->twin = on != 0; }
void oracle_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int oracle_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void oracle_set_rect_limits(void* h, float w, float hh, float d) {
    Sim* s = (Sim*)h;
    s->container[0] = w; s->container[1] = hh; s->container[2] = d;
}
// FluidSimulation::add_volume with a fixed particle set (FluidVolume::particles)
void oracle_add_volume(void* h, const float* xyzw, int n) {
    Sim* s = (Sim*)h;
    std::vector<f4> v(n);
    for (int i = 0; i < n; ++i) v[i] = {xyzw[4 * i], xyzw[4 * i + 1], xyzw[4 * i + 2], 1.0f};
    s->volumes.push_back(v);
    s->initialize_solver();
}
// BoundaryHandler::add_boundary -> _sample_surfaces (raw samples given)
int oracle_add_boundary(void* h, const float* xyzw, int n, int can_move, float mass,
                        const float* R, const float* com) {
    Sim* s = (Sim*)h;
    Body b;
    b.origin = s->B();
    b.count = n;
    b.can_move = can_move != 0;
    b.mass = mass;
    if (R) std::memcpy(b.R, R, sizeof b.R);
    if (com) std::memcpy(b.com, com, sizeof b.com);
    for (int i = 0; i < n; ++i) s->b_raw.push_back({xyzw[4 * i], xyzw[4 * i + 1], xyzw[4 * i + 2], 1.0f});
    s->bodies.push_back(b);
    s->compute_phi();
    return (int)s->bodies.size() - 1;
}
void oracle_set_body_state(void* h, int id, const float* R, const float* com, const float* lin, const float* ang) {
    Body& b = ((Sim*)h)->bodies[id];
    if (R) std::memcpy(b.R, R, sizeof b.R);
    if (com) std::memcpy(b.com, com, sizeof b.com);
    if (lin) std::memcpy(b.lin, lin, sizeof b.lin);
    if (ang) std::memcpy(b.ang, ang, sizeof b.ang);
}
void oracle_get_body_wrench(void* h, int id, float* f, float* t, double* f64, double* t64) {
    Body& b = ((Sim*)h)->bodies[id];
    for (int k = 0; k < 3; ++k) {
        if (f) f[k] = b.force[k];
        if (t) t[k] = b.torque[k];
        if (f64) f64[k] = b.force64[k];
        if (t64) t64[k] = b.torque64[k];
    }
}
void oracle_set_state(void* h, const float* pos, const float* vel, const float* vel_half, int n) {
    Sim* s = (Sim*)h;
    s->pos.resize(n); s->vel.resize(n); s->vel_half.resize(n);
    for (int i = 0; i < n; ++i) {
        s->pos[i] = {pos[4 * i], pos[4 * i + 1], pos[4 * i + 2], 1.0f};
        s->vel[i] = vel ? f4{vel[4 * i], vel[4 * i + 1], vel[4 * i + 2], 0.0f} : f4{0, 0, 0, 0};
        s->vel_half[i] = vel_half ? f4{vel_half[4 * i], vel_half[4 * i + 1], vel_half[4 * i + 2], 0.0f} : f4{0, 0, 0, 0};
    }
}
void oracle_step(void* h, int nsteps) {
    for (int k = 0; k < nsteps; ++k) ((Sim*)h)->step();
}
int oracle_particle_count(void* h) { return ((Sim*)h)->N(); }
int oracle_boundary_count(void* h) { return ((Sim*)h)->B(); }
int oracle_iterations(void* h) { return ((Sim*)h)->F.iterations; }
long long oracle_edge_reads(void* h) { return ((Sim*)h)->edge_reads; }
int oracle_grid_cps(void* h) { return ((Sim*)h)->grid.cps; }
float oracle_delta(void* h) { return ((Sim*)h)->pc.delta; }

void oracle_constants(void* h, float* out, int n) {
    Sim* s = (Sim*)h;
    float v[] = {s->pc.mass, s->pc.mass_d, s->pc.h_d, s->pc.rho0_d, s->pc.st_main, s->pc.st_term,
                 s->pc.default_eval, s->pc.default_grad, s->pc.visc_lapl, s->pc.pressure_grad, s->pc.delta,
                 s->wc.mass, s->wc.poly6_eval, s->wc.poly6_grad, s->wc.spiky_grad, s->wc.visc_lapl,
                 s->wc.st_main, s->wc.st_term, s->wc.mass_d, s->wc.h_d,
                 s->bc.h_d, s->bc.poly6_eval, s->bc.spiky_grad, s->grid.sqr_support, (float)s->grid.cps};
    for (int i = 0; i < n && i < (int)(sizeof v / sizeof v[0]); ++i) out[i] = v[i];
}

// Copies a field. fp32 / int32 fields go to out (4-byte elements); for
// vector fields 4 components per element (w as stored). twin=1 reads the
// double-precision twin into out64 (vector fields: 3 components).
// Returns the number of elements, or -1.
int oracle_get_field(void* h, int id, void* out, int twin) {
    Sim* s = (Sim*)h;
    const int n = s->N(), nb = s->B();
    auto put4 = [&](const std::vector<f4>& v) {
        if (out) std::memcpy(out, v.data(), v.size() * sizeof(f4));
        return (int)v.size();
    };
    auto put3f = [&](const std::vector<V3<float>>& v, float w) {
        float* o = (float*)out;
        if (o) for (size_t i = 0; i < v.size(); ++i) { o[4 * i] = v[i].x; o[4 * i + 1] = v[i].y; o[4 * i + 2] = v[i].z; o[4 * i + 3] = w; }
        return (int)v.size();
    };
    auto put3d = [&](const std::vector<V3<double>>& v) {
        double* o = (double*)out;
        if (o) for (size_t i = 0; i < v.size(); ++i) { o[3 * i] = v[i].x; o[3 * i + 1] = v[i].y; o[3 * i + 2] = v[i].z; }
        return (int)v.size();
    };
    auto puts = [&](const std::vector<float>& v) {
        if (out) std::memcpy(out, v.data(), v.size() * 4);
        return (int)v.size();
    };
    auto putd = [&](const std::vector<double>& v) {
        if (out) std::memcpy(out, v.data(), v.size() * 8);
        return (int)v.size();
    };
    auto puti = [&](const std::vector<int>& v) {
        if (out) std::memcpy(out, v.data(), v.size() * 4);
        return (int)v.size();
    };
    auto putu = [&](const std::vector<uint32_t>& v) {
        if (out) std::memcpy(out, v.data(), v.size() * 4);
        return (int)v.size();
    };
    auto ranges = [&](const Ranges& R) {
        // (cell id, start, end) triples of non-empty cells, sorted by start
        std::vector<std::array<long long, 3>> t;
        for (auto& kv : R.m) if (kv.second.second > kv.second.first) t.push_back({kv.first, kv.second.first, kv.second.second});
        std::sort(t.begin(), t.end(), [](auto& a, auto& b) { return a[1] < b[1]; });
        int* o = (int*)out;
        if (o) for (size_t i = 0; i < t.size(); ++i) { o[3 * i] = (int)t[i][0]; o[3 * i + 1] = (int)t[i][1]; o[3 * i + 2] = (int)t[i][2]; }
        return (int)t.size();
    };
    if (twin) {
        auto& D = s->D;
        switch (id) {
            case F_DENSITY: return putd(D.density);
            case F_PRESSURE: return putd(D.pressure);
            case F_ERR: return putd(D.err);
            case F_NORMAL: return put3d(D.normal);
            case F_FORCE_OTHER: return put3d(D.f_other);
            case F_FORCE_PRESSURE: return put3d(D.f_press);
            case F_POS_PRED: return put3d(D.pos_pred);
            case F_ACCEL: return put3d(D.accel);
            case F_POS: return put3d(D.out_pos);
            case F_VEL: return put3d(D.out_vel);
            case F_VEL_HALF: return put3d(D.out_vel_half);
            case F_BFORCE: return put3d(D.b_force);
            case F_BTORQUE: return put3d(D.b_torque);
            default: return -1;
        }
    }
    auto& F = s->F;
    (void)n; (void)nb;
    switch (id) {
        case F_KEYS: return putu(s->fs.keys);
        case F_PERM: return puti(s->fs.perm);
        case F_NCOUNT: return puti(s->nc);
        case F_NLIST: return puti(s->nl);
        case F_BNCOUNT: return puti(s->bnc);
        case F_BNLIST: return puti(s->bnl);
        case F_POS_SORTED: return put4(s->pos_s);
        case F_VEL_SORTED: return put4(s->vel_s);
        case F_VEL_HALF_SORTED: return put4(s->vel_half_s);
        case F_DENSITY: return puts(F.density);
        case F_PRESSURE: return puts(F.pressure);
        case F_ERR: return puts(F.err);
        case F_NORMAL: return put3f(F.normal, 0.f);
        case F_FORCE_OTHER: return put3f(F.f_other, 0.f);
        case F_FORCE_PRESSURE: return put3f(F.f_press, 0.f);
        case F_POS_PRED: return put3f(F.pos_pred, 1.f);
        case F_ACCEL: return put3f(F.accel, 0.f);
        case F_POS: return put4(s->pos);
        case F_VEL: return put4(s->vel);
        case F_VEL_HALF: return put4(s->vel_half);
        case F_BPOS_SORTED: return put4(s->b_pos_s);
        case F_BPHI_SORTED: return puts(s->b_phi_s);
        case F_BKEYS: return putu(s->bs.keys);
        case F_BPERM: return puti(s->bs.perm);
        case F_BPHI: return puts(s->b_phi);
        case F_BFORCE: return put3f(F.b_force, 0.f);
        case F_BTORQUE: return put3f(F.b_torque, 0.f);
        case F_RANGES: return ranges(s->fr);
        case F_BRANGES: return ranges(s->br);
        default: return -1;
    }
}

// Stand-alone pieces for the unit tests.
uint32_t oracle_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }
float oracle_define_literal(float v) { return define_literal(v); }
float oracle_deduce_delta(float h, float r, float dt, float mass, float rho0) { return deduce_delta(h, r, dt, mass, rho0); }
void oracle_boundary_phi(const float* xyzw, int n, float h, float rho0, int use_grid, float* phi) {
    Settings s{};
    s.fluid_support_radius = h;
    BdryConsts c = bdry_consts(s);
    std::vector<f4> p(n);
    for (int i = 0; i < n; ++i) p[i] = {xyzw[4 * i], xyzw[4 * i + 1], xyzw[4 * i + 2], 1.0f};
    if (use_grid) boundary_phi_grid(p.data(), n, c, rho0, phi);
    else boundary_phi_brute(p.data(), n, c, rho0, phi);
}

// ----------------------------------------------------------------------------
// Samplers (fp32 accumulators exactly as the reference loops them, Q20/Q21)
// ----------------------------------------------------------------------------
// BoxVolume::particles: fluid/simulation/boxvolume.cpp:11-40
int oracle_box_volume(const float* size, const float* center, float radius, float* out, int cap) {
    float spacing = 2 * radius;
    int px = (int)std::ceil(size[0] / spacing), py = (int)std::ceil(size[1] / spacing), pz = (int)std::ceil(size[2] / spacing);
    float dx = size[0] / px, dy = size[1] / py, dz = size[2] / pz;
    int n = 0;
    for (int j = 0; j < py; ++j)
        for (int i = 0; i < px; ++i)
            for (int k = 0; k < pz; ++k) {
                if (out && n < cap) {
                    out[4 * n + 0] = (float)((double)(i * dx + center[0]) - size[0] / 2.0);
                    out[4 * n + 1] = (float)((double)(j * dy + center[1]) - size[1] / 2.0);
                    out[4 * n + 2] = (float)((double)(k * dz + center[2]) - size[2] / 2.0);
                    out[4 * n + 3] = 1.0f;
                }
                ++n;
            }
    return n;
}
static int push(float* out, int cap, int n, float x, float y, float z) {
    if (out && n < cap) { out[4 * n] = x; out[4 * n + 1] = y; out[4 * n + 2] = z; out[4 * n + 3] = 1.0f; }
    return n + 1;
}
// FishTank::surface_sampling: scene/fishtank.cpp:108-141 (tank at the origin,
// identity transform)
int oracle_fishtank(float width, float height, float depth, float s, float* out, int cap) {
    int n = 0;
    for (float x = -width / 2; x <= width / 2; x += s)
        for (float z = -depth / 2; z <= depth / 2; z += s) n = push(out, cap, n, x, -height / 2, z);
    float y0 = -height / 2 + s;
    for (float y = y0; y <= height / 2; y += s) {
        for (float x = -width / 2; x <= width / 2; x += s) {
            n = push(out, cap, n, x, y, -depth / 2);
            n = push(out, cap, n, x, y, depth / 2);
        }
        for (float z = -depth / 2; z <= depth / 2; z += s) {
            n = push(out, cap, n, width / 2, y, z);
            n = push(out, cap, n, -width / 2, y, z);
        }
    }
    return n;
}
// Cube::surface_sampling: scene/cube.cpp:136-170 (body-local samples)
int oracle_cube(float side, float s, float* out, int cap) {
    float width = side - s, height = side - s, depth = side - s;
    int n = 0;
    for (float x = -width / 2; x <= width / 2; x += s)
        for (float z = -depth / 2; z <= depth / 2; z += s) {
            n = push(out, cap, n, x, -height / 2, z);
            n = push(out, cap, n, x, height / 2, z);
        }
    float y0 = -height / 2 + s;
    for (float y = y0; y <= height / 2; y += s) {
        for (float x = -width / 2; x <= width / 2; x += s) {
            n = push(out, cap, n, x, y, -depth / 2);
            n = push(out, cap, n, x, y, depth / 2);
        }
        for (float z = -depth / 2; z <= depth / 2; z += s) {
            n = push(out, cap, n, width / 2, y, z);
            n = push(out, cap, n, -width / 2, y, z);
        }
    }
    return n;
}
// Wall::surface_sampling: scene/wall.cpp:142-152
int oracle_wall(float width, float height, float s, float* out, int cap) {
    int n = 0;
    for (float x = -width / 2; x <= width / 2; x += s)
        for (float y = -height / 2; y <= height / 2; y += s) n = push(out, cap, n, x, y, 0.0f);
    return n;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// WCSPH in double precision over double state, with its own neighbour
// search, for finite differences of a loss on the final state (the gradient
// checks of the adjoint). It restates wcsph_fields / wcsph_integrate above
// (wcsph.cl:32-224) with the repaired density when S.flags says so. Lists
// are rebuilt each step from the double positions with the cutoff
// r^2 < CELL_SIZE^2 (grid.cl:3) and are uncapped: the checks use scenes whose
// lists stay below 50. Particles keep their id order; list order is
// ascending id, which changes nothing but double rounding.
// ---------------------------------------------------------------------------
namespace {
struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double dclamp(double x, double lo, double hi) { return std::min(std::max(x, lo), hi); }
inline double dsign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }
inline D3 dclamp3(D3 a, double lo, double hi) { return {dclamp(a.x, lo, hi), dclamp(a.y, lo, hi), dclamp(a.z, lo, hi)}; }
}  // namespace

extern "C" int oracle_wcsph_rollout64(const oracle_settings* so, double W, double H, double Dp, const double* pos0,
                                      const double* vel0, const double* vh0, int n, int nsteps, double k_visc,
                                      double k_gas, double sigma, double* pos_out, double* vel_out, double* vh_out,
                                      int* max_count) {
    Settings S;
    std::memcpy(&S, so, sizeof(Settings));
    const WcConsts c = wc_consts(S);
    const double h = c.h, h2 = h * h, M = c.mass, rho0 = S.rest_density, dt = S.time_step, vmax = S.max_vel;
    const double cut2 = (double)(define_literal(h) * define_literal(h));
    const double c_rho = (double)c.poly6_eval * M;
    const double c_n = (double)c.poly6_grad * (double)c.mass_d * (double)c.h_d;
    const double H2n = (double)c.h_d * (double)c.h_d;
    const double spiky = c.spiky_grad, visc = (double)c.visc_lapl * k_visc, st_main = c.st_main,
                 st_term = c.st_term;
    const bool fix = (S.flags & FIX_WCSPH_DENSITY) != 0;
    const D3 g = {0.0, -(double)S.gravity, 0.0};
    std::vector<D3> x(n), v(n), vh(n), nrm(n), acc(n);
    std::vector<double> rho(n), p(n);
    for (int i = 0; i < n; ++i) {
        x[i] = {pos0[3 * i], pos0[3 * i + 1], pos0[3 * i + 2]};
        v[i] = vel0 ? D3{vel0[3 * i], vel0[3 * i + 1], vel0[3 * i + 2]} : D3{0, 0, 0};
        vh[i] = vh0 ? D3{vh0[3 * i], vh0[3 * i + 1], vh0[3 * i + 2]} : D3{0, 0, 0};
    }
    // neighbour lists in CSR form over a dense cell box rebuilt every step
    std::vector<int> start(n + 1), cnt(n), nbr;
    std::vector<int> cell_of(n), cell_start, order(n);
    int maxc = 0;
    for (int step = 0; step < nsteps; ++step) {
        double lo[3] = {1e300, 1e300, 1e300};
        for (int i = 0; i < n; ++i) {
            lo[0] = std::min(lo[0], x[i].x);
            lo[1] = std::min(lo[1], x[i].y);
            lo[2] = std::min(lo[2], x[i].z);
        }
        long long c0[3], nc[3] = {1, 1, 1};
        for (int k = 0; k < 3; ++k) c0[k] = (long long)std::floor(lo[k] / h) - 1;
        for (int i = 0; i < n; ++i) {
            const double q[3] = {x[i].x, x[i].y, x[i].z};
            for (int k = 0; k < 3; ++k) nc[k] = std::max(nc[k], (long long)std::floor(q[k] / h) - c0[k] + 2);
        }
        const long long ncell = nc[0] * nc[1] * nc[2];
        if (ncell > (1LL << 28)) return -1;
        auto cidx = [&](const D3& q, int dx, int dy, int dz) {
            long long cx = (long long)std::floor(q.x / h) - c0[0] + dx;
            long long cy = (long long)std::floor(q.y / h) - c0[1] + dy;
            long long cz = (long long)std::floor(q.z / h) - c0[2] + dz;
            return (cz * nc[1] + cy) * nc[0] + cx;
        };
        cell_start.assign(ncell + 1, 0);
        for (int i = 0; i < n; ++i) {
            cell_of[i] = (int)cidx(x[i], 0, 0, 0);
            ++cell_start[cell_of[i] + 1];
        }
        for (long long k = 0; k < ncell; ++k) cell_start[k + 1] += cell_start[k];
        {
            std::vector<int> fill(cell_start.begin(), cell_start.end() - 1);
            for (int i = 0; i < n; ++i) order[fill[cell_of[i]]++] = i;
        }
        auto walk = [&](int i, auto&& f) {
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        long long k = cidx(x[i], dx, dy, dz);
                        for (int e = cell_start[k]; e < cell_start[k + 1]; ++e) {
                            int j = order[e];
                            D3 r = x[i] - x[j];
                            if (ddot(r, r) < cut2) f(j);
                        }
                    }
        };
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            int c = 0;
            walk(i, [&](int) { ++c; });
            cnt[i] = c;
        }
        start[0] = 0;
        for (int i = 0; i < n; ++i) start[i + 1] = start[i] + cnt[i];
        nbr.resize(start[n]);
#pragma omp parallel for schedule(static) reduction(max : maxc)
        for (int i = 0; i < n; ++i) {
            int o = start[i];
            walk(i, [&](int j) { nbr[o++] = j; });
            std::sort(nbr.begin() + start[i], nbr.begin() + start[i + 1]);
            maxc = std::max(maxc, cnt[i]);
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            double d = 0.0;
            for (int e = start[i]; e < start[i + 1]; ++e) {
                const int j = nbr[e];
                D3 r = x[i] - x[j];
                double s2 = ddot(r, r);
                double s = fix ? s2 : std::sqrt(s2);
                double t = h2 - s;
                d += t * t * t;
            }
            rho[i] = d * c_rho;
            double y = rho[i] / rho0;
            p[i] = std::max(0.0, k_gas * (std::pow(y, 7) - 1.0));
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            D3 a = {0, 0, 0};
            for (int e = start[i]; e < start[i + 1]; ++e) {
                const int j = nbr[e];
                D3 r = x[i] - x[j];
                double w = H2n - ddot(r, r);
                a = a + r * (w * w / rho[j]);
            }
            nrm[i] = a * c_n;
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            D3 pacc = {0, 0, 0}, vacc = {0, 0, 0}, fcurv = {0, 0, 0}, fcoh = {0, 0, 0};
            const double di = rho[i], Ci = p[i] / (di * di);
            for (int e = start[i]; e < start[i + 1]; ++e) {
                const int j = nbr[e];
                D3 r = x[i] - x[j];
                double q = std::sqrt(ddot(r, r));
                double dj = rho[j];
                if (!(di > 0 && dj > 0 && q > 0.00001)) continue;
                double hq = h - q;
                pacc = pacc + r * (spiky / q * hq * hq * (Ci + p[j] / (dj * dj)));
                vacc = vacc + (v[j] - v[i]) * (hq / dj);
                double K = 2.0 * rho0 / (di + dj);
                fcurv = fcurv + (nrm[i] - nrm[j]) * K;
                double st = 0.0;
                if (q <= 0.5 * h) st = 2.0 * (hq * hq * hq) * (q * q * q) - st_term;
                else if (q <= h) st = (hq * hq * hq) * (q * q * q);
                fcoh = fcoh + r * (st / q * K);
            }
            acc[i] = g + (fcurv + fcoh * st_main) * (-sigma) + pacc * (-M) + vacc * visc;
        }
        const double lo_x = -W / 2, hi_x = W / 2, lo_y = -H / 2, lo_z = -Dp / 2, hi_z = Dp / 2;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            D3 a = acc[i];
            D3 w = vh[i] + a * dt;
            D3 xp = x[i] + dclamp3(w + a * dt, -vmax, vmax) * dt;
            double vn = std::sqrt(ddot(w, w));
            D3 cp = {dclamp(xp.x, lo_x, hi_x), dclamp(xp.y, lo_y, 10.0), dclamp(xp.z, lo_z, hi_z)};
            D3 dv = cp - xp;
            double dd = std::sqrt(ddot(dv, dv));
            if (dd > 0 && vn > 0) {
                D3 sg = {dsign(dv.x), dsign(dv.y), dsign(dv.z)};
                D3 nn = sg * (1.0 / std::sqrt(ddot(sg, sg)));
                double rest = dd / (dt * vn);
                w = w - nn * ((1.0 + rest) * ddot(w, nn));
                xp = cp;
            }
            x[i] = xp;
            v[i] = dclamp3(w + a * (0.5 * dt), -vmax, vmax);
            vh[i] = dclamp3(w, -vmax, vmax);
        }
    }
    for (int i = 0; i < n; ++i) {
        if (pos_out) { pos_out[3 * i] = x[i].x; pos_out[3 * i + 1] = x[i].y; pos_out[3 * i + 2] = x[i].z; }
        if (vel_out) { vel_out[3 * i] = v[i].x; vel_out[3 * i + 1] = v[i].y; vel_out[3 * i + 2] = v[i].z; }
        if (vh_out) { vh_out[3 * i] = vh[i].x; vh_out[3 * i + 1] = vh[i].y; vh_out[3 * i + 2] = vh[i].z; }
    }
    if (max_count) *max_count = maxc;
    return 0;
}
