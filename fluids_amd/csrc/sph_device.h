// Device helpers of the gfx950 SPH kernels.
//
// Arithmetic is IEEE fp32 with no contraction (the library is compiled with
// -ffp-contract=off and correctly rounded division and square root), and every
// expression keeps the association of the reference OpenCL expression it
// restates, so a kernel produces the same bits as any other IEEE evaluation of
// that expression. Built-in semantics are those of OpenCL 1.2:
//   max(x,y) = x < y ? y : x      min(x,y) = y < x ? y : x
//   clamp(x,lo,hi) = min(max(x,lo),hi)
//   pown(x,2) = x*x, pown(x,3) = (x*x)*x, pown(x,7) left-to-right products
//   fast_length(r) = length(r) = sqrt(dot(r,r))
//   normalize(r) = fast_normalize(r) = r * (1 / length(r)), and r itself for
//     r == 0 (OpenCL 1.2 defines fast_normalize(p) as p * half_rsqrt(dot(p, p));
//     the reciprocal square root is taken as two correctly rounded operations)
// The w component is left out of all arithmetic: positions carry w == 1 and
// velocities, forces and normals w == 0 on the whole path
// (boxvolume.cpp:71, boundaryhandler.cpp:132, pcisph.cl:279-280), so every w
// term of the reference's float4 dot products is exactly +0.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

constexpr int LIST_CAP = 50;      // NEIGH_LIST_MAX_LENGTH (kernels/common.h:4)
constexpr int MAXERR_SLOTS = 16;  // ring of per-correction max density errors
// Each correction's maximum is merged into ERR_LANES words (workgroup b into
// word b mod ERR_LANES) and read back by one load per lane and a wave
// maximum: thousands of workgroups then never queue on one address.
constexpr int ERR_LANES = 64;
constexpr int GATHER_BATCH = 8;  // neighbour loads issued together per thread
constexpr int BOUND_BATCH = 4;   // boundary-neighbour loads issued together per thread
// Largest fluid or boundary particle count of one scene (or one batch): the
// neighbour walk reads candidates through buffer views with 32-bit byte
// offsets.
constexpr int MAX_SCENE_PARTICLES = 1 << 27;

// A buffer view of an array of `bytes` bytes (raw, no stride): loads at an
// offset past the end read zero. Offsets are 32-bit; the launchers keep the
// particle counts below MAX_SCENE_PARTICLES so that 16-byte records fit.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_view(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float4 load4(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, 0));
}
// The first three words of a 16-byte record (its w is not fetched).
__device__ __forceinline__ float4 load3(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff) {
    typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));
    const u32x3 v = __builtin_amdgcn_raw_buffer_load_b96(r, off, soff, 0);
    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), 0.0f);
}
__device__ __forceinline__ float load1(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
}

// Uniform grid restated from kernels/grid.h and fluid/simulation/grid.cpp:61-68.
// Cell ranges live in a dense box: all cps cells in y (the 5 m grid is
// periodic and the integrators allow y up to 10 m), [bx0, bx0+bnx) in x and
// [bz0, bz0+bnz) in z. Entries are int2 {start, end}; {0, 0} is an empty cell,
// as in the reference's zero-filled int2[cps^3]. Instead of zero-filling the
// whole box every step, the hash kernel clears the cells named by the previous
// step's sorted keys before new ranges are written.
struct GridDev {
    float half;         // size / 2 = 2.5f
    float cell;         // cell size: full fp32 support radius
    float sqr_support;  // CELL_SIZE*CELL_SIZE with the 6-decimal literal (grid.cl:3)
    int cps;            // cells per side
    long long cells;    // cps^3
    int bx0, bz0, bnx, bnz;
};

// Per-step device state written by the kernels.
struct DevState {
    unsigned int maxerr[MAXERR_SLOTS][ERR_LANES];  // float bits of max density error per correction (ring)
    int done_iter;                      // correction at which the loop ended (0: running)
    int iterations;
    int out_of_grid;
    int edge_reads;
    int nonfinite;                      // non-finite positions met by the hash kernel
};

__device__ __forceinline__ float cl_max(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float cl_min(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float cl_clamp(float x, float lo, float hi) { return cl_min(cl_max(x, lo), hi); }
__device__ __forceinline__ float cl_sign(float x) {
    if (x > 0.0f) return 1.0f;
    if (x < 0.0f) return -1.0f;
    if (x != x) return 0.0f;
    return x;
}
__device__ __forceinline__ float sqr(float x) { return x * x; }
__device__ __forceinline__ float cube(float x) { return (x * x) * x; }
__device__ __forceinline__ float pown7(float x) {
    float r = x * x;
    r = r * x; r = r * x; r = r * x; r = r * x; r = r * x;
    return r;
}

struct V3 { float x, y, z; };
__device__ __forceinline__ V3 v3(float4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) {
    float d = dot(a, a);
    if (d == 0.0f) return a;
    float inv = 1.0f / sqrtf(d);
    return {a.x * inv, a.y * inv, a.z * inv};
}
__device__ __forceinline__ V3 clamp3(V3 a, float lo, float hi) {
    return {cl_clamp(a.x, lo, hi), cl_clamp(a.y, lo, hi), cl_clamp(a.z, lo, hi)};
}
__device__ __forceinline__ float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// Division and square root of one kernel build. Math<false> is the IEEE
// arithmetic above (correctly rounded, bit-reproducible). Math<true> follows
// the reference's -cl-fast-relaxed-math builds (pcisphsimluation.cpp:496,
// normals through grid.cpp:82): a / b = a * rcp(b), sqrt and 1/sqrt by the
// hardware's v_rcp_f32, v_sqrt_f32 and v_rsq_f32 (each within about 1 ulp);
// additions and products stay uncontracted.
template <bool FAST> struct Math {
    static __device__ __forceinline__ float div(float a, float b) {
        if constexpr (FAST) return a * __builtin_amdgcn_rcpf(b);
        else return a / b;
    }
    static __device__ __forceinline__ float sqrt(float x) {
        if constexpr (FAST) return __builtin_amdgcn_sqrtf(x);
        else return sqrtf(x);
    }
    static __device__ __forceinline__ V3 div(V3 a, float s) {
        if constexpr (FAST) {
            const float inv = __builtin_amdgcn_rcpf(s);
            return {a.x * inv, a.y * inv, a.z * inv};
        } else {
            return a / s;
        }
    }
    static __device__ __forceinline__ float length(V3 a) { return sqrt(dot(a, a)); }
    static __device__ __forceinline__ V3 normalize(V3 a) {
        if constexpr (FAST) {
            float d = dot(a, a);
            if (d == 0.0f) return a;
            float inv = __builtin_amdgcn_rsqf(d);
            return {a.x * inv, a.y * inv, a.z * inv};
        } else {
            return sph::normalize(a);
        }
    }
};

// get_grid_coordinates (kernels/grid.h:74-91)
__device__ __forceinline__ void grid_coords(float4 p, const GridDev& g, int& cx, int& cy, int& cz) {
    cx = (int)floorf((p.x + g.half) / g.cell);
    cy = (int)floorf((p.y + g.half) / g.cell);
    cz = (int)floorf((p.z + g.half) / g.cell);
    cx = (cx + g.cps) % g.cps;
    cy = (cy + g.cps) % g.cps;
    cz = (cz + g.cps) % g.cps;
}

// split_by_3 / morton_3d_encode with OpenCL 32-bit semantics (kernels/morton.h:37-56):
// x << 32 on a 32-bit uint shifts by 0, and the 64-bit masks keep their low 32 bits.
__device__ __forceinline__ uint32_t split_by_3(uint32_t a) {
    uint32_t x = a & 0xffffu;
    x = (x | (x << 16)) & 0xff0000ffu;
    x = (x | (x << 8)) & 0x0f00f00fu;
    x = (x | (x << 4)) & 0xc30c30c3u;
    x = (x | (x << 2)) & 0x49249249u;
    return x;
}
__device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return split_by_3(x) | (split_by_3(y) << 1) | (split_by_3(z) << 2);
}

// Dense-box index of an in-grid cell, or -1 when it lies outside the box.
__device__ __forceinline__ long long box_index(int x, int y, int z, const GridDev& g) {
    int lx = x - g.bx0, lz = z - g.bz0;
    if ((unsigned)lx >= (unsigned)g.bnx || (unsigned)lz >= (unsigned)g.bnz) return -1;
    return ((long long)lz * g.cps + y) * g.bnx + lx;
}

// cell_interval[CELL_ID(x, y, z)] for a neighbour coordinate that may lie one
// cell outside [0, cps) (the reference does not wrap neighbour offsets,
// grid.cl:165-166). CELL_ID is linear, so such a coordinate names another
// in-range cell unless the linear id leaves [0, cps^3), where the reference
// reads out of bounds; that case reads as empty and is counted.
__device__ __forceinline__ int2 cell_range(const int2* __restrict__ R, int x, int y, int z, const GridDev& g,
                                           int& edge) {
    int c = g.cps;
    if ((unsigned)x >= (unsigned)c || (unsigned)y >= (unsigned)c || (unsigned)z >= (unsigned)c) {
        long long L = (long long)c * ((long long)c * z + y) + x;
        if (L < 0 || L >= g.cells) {
            ++edge;
            return make_int2(0, 0);
        }
        long long cc = (long long)c * c;
        z = (int)(L / cc);
        y = (int)((L / c) % c);
        x = (int)(L % c);
    }
    long long b = box_index(x, y, z, g);
    if (b < 0) return make_int2(0, 0);
    return R[b];
}

// Inverse of split_by_3 for 10-bit coordinates (the Morton key is the
// textbook interleave there).
__device__ __forceinline__ uint32_t compact_by_3(uint32_t x) {
    x &= 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030c30c3u;
    x = (x ^ (x >> 4)) & 0x0300f00fu;
    x = (x ^ (x >> 8)) & 0xff0000ffu;
    x = (x ^ (x >> 16)) & 0x000003ffu;
    return x;
}

// Interior cells have all 27 neighbours in [0, cps)^3 without wrapping.
__device__ __forceinline__ bool interior_cell(int x, int y, int z, int cps) {
    return x >= 1 && y >= 1 && z >= 1 && x <= cps - 2 && y <= cps - 2 && z <= cps - 2;
}

// Every bit of a 27-cell mask in CELL_NEIGH_OFFSET order.
constexpr uint32_t ALL_CELLS = (1u << 27) - 1u;

// CELL_NEIGH_OFFSET (kernels/grid.h:15-42): own cell, then the 26 others in
// lexicographic (dx, dy, dz) order.
__device__ __forceinline__ void neigh_offset(int o, int& dx, int& dy, int& dz) {
    if (o == 0) { dx = dy = dz = 0; return; }
    int k = o - 1;
    if (k >= 13) ++k;  // skip (0,0,0), which is lexicographic position 13
    dx = k / 9 - 1;
    dy = (k / 3) % 3 - 1;
    dz = k % 3 - 1;
}
// Inverse of neigh_offset: the CELL_NEIGH_OFFSET index of (dx, dy, dz).
__device__ __forceinline__ int neigh_index(int dx, int dy, int dz) {
    const int k = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
    return k == 13 ? 0 : (k < 13 ? k + 1 : k);
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD owns a contiguous run of particles (Morton-adjacent neighbours then
// share that XCD's L2). Bijective on [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Largest of v (0 <= v < 64) over the active lanes of the wave; every lane
// gets the same value.
__device__ __forceinline__ int wave_max_small(int v) {
    int m = 0;
#pragma unroll
    for (int b = 5; b >= 0; --b) {
        const int c = m | (1 << b);
        if (__ballot(v >= c)) m = c;
    }
    return m;
}

// Entries k0 .. k0 + B - 1 of particle i's column-major list
// (list[k * stride + i]); an entry at or past len reads as dflt. `rows` is
// the wave's largest len and k0 < rows, so every row read is one that some
// lane of the wave uses (rows past it re-read row rows - 1). The loads carry
// no branch, so all B of them are in flight together; the row offset is the
// same for the whole wave. A dflt shared by the wave keeps the gathers of
// unused entries on one cache line.
template <int B>
__device__ __forceinline__ void list_batch(const int* __restrict__ list, size_t stride, int i, int len, int rows,
                                           int k0, int dflt, int (&jj)[B]) {
#pragma unroll
    for (int u = 0; u < B; ++u) {
        const int k = k0 + u < rows ? k0 + u : rows - 1;
        const int j = list[(size_t)k * stride + i];
        jj[u] = k0 + u < len ? j : dflt;
    }
}

// Sum over the 64 lanes of a wave by butterfly shuffles (lane 0's value is
// ((v0 + v32) + (v16 + v48)) + ..., a fixed order). All lanes must be active.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// An index every active lane of the wave holds the same value of.
__device__ __forceinline__ int wave_first(int v) { return __builtin_amdgcn_readfirstlane(v); }

}  // namespace sph
