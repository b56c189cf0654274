// Neighbour search on gfx950: Morton cell keys, stable LSD radix sort,
// gather into sorted order with cell ranges, and the capped neighbour lists
// fused with the density sum that follows them in the reference.
//
// Restates kernels/grid.cl (compute_hashes :20-39, compute_cell_intervals
// :60-123, compute_neigh_list :145-181), opencl/algorithms/clsort.h:12-35
// (stable key/value sort), opencl/algorithms/clshuffle.h:23-35 (gather) and
// the density kernels that consume the lists (pcisph.cl:33-102,
// wcsph.cl:32-79).
#include "kernels.h"
#include "sph_device.h"

namespace sph {

// ---------------------------------------------------------------- hashing
// compute_hashes (grid.cl:20-39). The values array of the sort is the
// identity; the first sort pass generates it instead of reading it.
__global__ __launch_bounds__(256) void k_hash(int n, const float4* __restrict__ pos, GridDev g,
                                              uint32_t* __restrict__ keys, DevState* st) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos[i];
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) atomicAdd(&st->nonfinite, 1);
    int cx, cy, cz;
    grid_coords(p, g, cx, cy, cz);
    keys[i] = morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
    if (box_index(cx, cy, cz, g) < 0) atomicAdd(&st->out_of_grid, 1);
}

// ------------------------------------------------------------- radix sort
// Stable LSD radix sort of (u32 key, i32 value), 8-bit digits. One pass is
// histogram -> exclusive scan (digit-major) -> stable scatter. The scatter
// ranks each 256-key round of a tile with wave64 ballots (the lanes sharing a
// digit are the AND of 8 bit-ballots), then orders waves through LDS, so the
// output keeps input order within equal digits.
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;
constexpr int RADIX = 256;

__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                             int ntiles, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[RADIX];
    int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    int base = blockIdx.x * SORT_TILE;
#pragma unroll 4
    for (int k = 0; k < SORT_ITEMS; ++k) {
        int idx = base + k * SORT_THREADS + t;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & (RADIX - 1)], 1u);
    }
    __syncthreads();
    hist[(size_t)t * ntiles + blockIdx.x] = h[t];
}

// Exclusive scan of len words in place, one workgroup of 1024 threads: each
// thread scans a contiguous chunk, chunk totals are scanned in LDS.
__global__ __launch_bounds__(1024) void k_scan_exclusive(uint32_t* __restrict__ a, int len) {
    __shared__ uint32_t s[1024];
    int t = threadIdx.x;
    int chunk = (len + 1023) / 1024;
    int b = t * chunk;
    int e = b + chunk < len ? b + chunk : len;
    uint32_t sum = 0;
    for (int i = b; i < e; ++i) sum += a[i];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = (t >= off) ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t run = (t > 0) ? s[t - 1] : 0u;
    for (int i = b; i < e; ++i) {
        uint32_t v = a[i];
        a[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                const int* __restrict__ vin,
                                                                uint32_t* __restrict__ kout, int* __restrict__ vout,
                                                                int n, int shift, int ntiles,
                                                                const uint32_t* __restrict__ offs) {
    __shared__ uint32_t base_off[RADIX];
    __shared__ uint32_t run[RADIX];
    __shared__ uint32_t wcnt[SORT_THREADS / 64][RADIX];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    base_off[t] = offs[(size_t)t * ntiles + blockIdx.x];
    run[t] = 0;
#pragma unroll
    for (int w = 0; w < SORT_THREADS / 64; ++w) wcnt[w][t] = 0;
    __syncthreads();
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const int base = blockIdx.x * SORT_TILE;
    for (int k = 0; k < SORT_ITEMS; ++k) {
        int idx = base + k * SORT_THREADS + t;
        bool valid = idx < n;
        uint32_t key = valid ? kin[idx] : 0u;
        int val = valid ? (vin ? vin[idx] : idx) : 0;
        uint32_t d = (key >> shift) & (RADIX - 1);
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t bb = __ballot(valid && bit);
            m &= bit ? bb : ~bb;
        }
        int rank = __popcll(m & lt_mask);
        int cnt = __popcll(m);
        bool leader = valid && rank == 0;
        if (leader) wcnt[wave][d] = (uint32_t)cnt;
        __syncthreads();
        uint32_t pos = 0;
        if (valid) {
            uint32_t pre = run[d];
            for (int w = 0; w < wave; ++w) pre += wcnt[w][d];
            pos = base_off[d] + pre + (uint32_t)rank;
        }
        __syncthreads();
        if (leader) {
            bool highest = true;
            for (int w = wave + 1; w < SORT_THREADS / 64; ++w)
                if (wcnt[w][d]) highest = false;
            if (highest) {
                uint32_t tot = 0;
                for (int w = 0; w < SORT_THREADS / 64; ++w) tot += wcnt[w][d];
                run[d] += tot;
            }
        }
        __syncthreads();
        if (leader) wcnt[wave][d] = 0;
        if (valid && pos < (uint32_t)n) {  // always true for a consistent histogram
            kout[pos] = key;
            vout[pos] = val;
        }
    }
}

int sort_passes(int key_bits) { return key_bits <= 0 ? 1 : (key_bits + 7) / 8; }

size_t sort_hist_words(int n) {
    int ntiles = (n + SORT_TILE - 1) / SORT_TILE;
    return (size_t)RADIX * (ntiles > 0 ? ntiles : 1);
}

void launch_radix_sort(uint32_t* keys_a, uint32_t* keys_b, int* vals_a, int* vals_b, uint32_t* hist, int n,
                       int key_bits, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    int passes = sort_passes(key_bits);
    int ntiles = (n + SORT_TILE - 1) / SORT_TILE;
    // The caller hashed into keys_b when passes is odd, keys_a when even;
    // either way the result lands in (keys_a, vals_a).
    uint32_t* kin = (passes % 2) ? keys_b : keys_a;
    uint32_t* kout = (passes % 2) ? keys_a : keys_b;
    int* vin = nullptr;
    int* vout = (passes % 2) ? vals_a : vals_b;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        PROF_BEGIN(prof, "radix_hist", s);
        hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(SORT_THREADS), 0, s, kin, n, shift, ntiles, hist);
        PROF_END(prof, s);
        PROF_BEGIN(prof, "radix_scan", s);
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, hist, RADIX * ntiles);
        PROF_END(prof, s);
        PROF_BEGIN(prof, "radix_scatter", s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(SORT_THREADS), 0, s, kin, vin, kout, vout, n, shift,
                           ntiles, hist);
        PROF_END(prof, s);
        // next pass reads what this one wrote
        uint32_t* kt = kin; kin = kout; kout = kt;
        int* nv = (vout == vals_a) ? vals_b : vals_a;
        vin = vout;
        vout = nv;
    }
}

void launch_hash(int n, const float4* pos, const GridDev& g, uint32_t* keys, DevState* st, hipStream_t s,
                 Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "hash", s);
    hipLaunchKernelGGL(k_hash, dim3((n + 255) / 256), dim3(256), 0, s, n, pos, g, keys, st);
    PROF_END(prof, s);
}

// ------------------------------------------------------ gather + ranges
// clshuffle of positions/velocities (pcisphsimluation.cpp:89-90,
// wcsphsimluation.cpp:57-59) fused with compute_cell_intervals
// (grid.cl:60-123): a cell's range starts where the sorted key changes and
// ends before the next change.
__device__ __forceinline__ void write_range(int i, int n, const uint32_t* __restrict__ keys_s, float4 p,
                                            int4* __restrict__ R, const GridDev& g, int epoch) {
    uint32_t key = keys_s[i];
    bool first = (i == 0) || keys_s[i - 1] != key;
    bool last = (i == n - 1) || keys_s[i + 1] != key;
    if (first || last) {
        int cx, cy, cz;
        grid_coords(p, g, cx, cy, cz);
        long long b = box_index(cx, cy, cz, g);
        if (b >= 0) {
            if (first) *reinterpret_cast<int2*>(&R[b]) = make_int2(i, epoch);
            if (last) R[b].z = i + 1;
        }
    }
}

__global__ __launch_bounds__(256) void k_gather_fluid(int n, const int* __restrict__ perm,
                                                      const uint32_t* __restrict__ keys_s,
                                                      const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                      const float4* __restrict__ vh, float4* __restrict__ pos_s,
                                                      float4* __restrict__ vel_s, float4* __restrict__ vh_s,
                                                      int4* __restrict__ R, GridDev g, int epoch) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int j = perm[i];
    if ((unsigned)j >= (unsigned)n) j = i;  // a permutation never leaves [0, n)
    float4 p = pos[j];
    pos_s[i] = p;
    vel_s[i] = vel[j];
    if (vh) vh_s[i] = vh[j];
    write_range(i, n, keys_s, p, R, g, epoch);
}

__global__ __launch_bounds__(256) void k_gather_boundary(int n, const int* __restrict__ perm,
                                                         const uint32_t* __restrict__ keys_s,
                                                         const float4* __restrict__ pos, const float* __restrict__ phi,
                                                         float4* __restrict__ pos_s, float* __restrict__ phi_s,
                                                         int4* __restrict__ R, GridDev g, int epoch) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int j = perm[i];
    if ((unsigned)j >= (unsigned)n) j = i;
    float4 p = pos[j];
    pos_s[i] = p;
    phi_s[i] = phi[j];
    write_range(i, n, keys_s, p, R, g, epoch);
}

void launch_gather_fluid(int n, const int* perm, const uint32_t* keys_s, const float4* pos, const float4* vel,
                         const float4* vh, float4* pos_s, float4* vel_s, float4* vh_s, int4* R, const GridDev& g,
                         int epoch, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "gather_ranges", s);
    hipLaunchKernelGGL(k_gather_fluid, dim3((n + 255) / 256), dim3(256), 0, s, n, perm, keys_s, pos, vel, vh, pos_s,
                       vel_s, vh_s, R, g, epoch);
    PROF_END(prof, s);
}

void launch_gather_boundary(int n, const int* perm, const uint32_t* keys_s, const float4* pos, const float* phi,
                            float4* pos_s, float* phi_s, int4* R, const GridDev& g, int epoch, hipStream_t s,
                            Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "gather_boundary", s);
    hipLaunchKernelGGL(k_gather_boundary, dim3((n + 255) / 256), dim3(256), 0, s, n, perm, keys_s, pos, phi, pos_s,
                       phi_s, R, g, epoch);
    PROF_END(prof, s);
}

// ----------------------------------------------- lists fused with density
// compute_neigh_list (grid.cl:145-181): own cell first, then CELL_NEIGH_OFFSET
// order, sorted order inside a cell, strict r2 < CELL_SIZE^2, cap 50,
// column-major list[k*N + i]. The density kernels read exactly these r2 in
// exactly this order, so the sum is accumulated while the list is written.
template <bool WCSPH>
__global__ __launch_bounds__(256) void k_lists_density(int n, const float4* __restrict__ pos_s,
                                                       const int4* __restrict__ fR, int fepoch, int nb,
                                                       const float4* __restrict__ bpos_s,
                                                       const float* __restrict__ bphi_s, const int4* __restrict__ bR,
                                                       int bepoch, GridDev g, DensityParams P,
                                                       int* __restrict__ nl, int* __restrict__ nc,
                                                       int* __restrict__ bnl, int* __restrict__ bnc,
                                                       float* __restrict__ density, float* __restrict__ pressure,
                                                       DevState* st) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p4 = pos_s[i];
    V3 pi = v3(p4);
    int cx, cy, cz;
    grid_coords(p4, g, cx, cy, cz);
    int edge = 0;
    float d = 0.0f;
    int len = 0;
    for (int o = 0; len < LIST_CAP && o < 27; ++o) {
        int dx, dy, dz;
        neigh_offset(o, dx, dy, dz);
        int2 iv = cell_range(fR, fepoch, cx + dx, cy + dy, cz + dz, g, edge);
        for (int j = iv.x; len < LIST_CAP && j < iv.y; ++j) {
            V3 r = pi - v3(pos_s[j]);
            float r2 = dot(r, r);
            if (r2 < g.sqr_support) {
                nl[(size_t)len * n + i] = j;
                ++len;
                if (WCSPH) {
                    float rn = P.fix_density ? r2 : sqrtf(r2);
                    d += cube(P.h * P.h - rn);
                } else {
                    d += cube(P.H2 - r2);
                }
            }
        }
    }
    nc[i] = len;
    float bd = 0.0f;
    int blen = 0;
    if (nb > 0) {
        for (int o = 0; blen < LIST_CAP && o < 27; ++o) {
            int dx, dy, dz;
            neigh_offset(o, dx, dy, dz);
            int2 iv = cell_range(bR, bepoch, cx + dx, cy + dy, cz + dz, g, edge);
            for (int j = iv.x; blen < LIST_CAP && j < iv.y; ++j) {
                V3 r = pi - v3(bpos_s[j]);
                float r2 = dot(r, r);
                if (r2 < g.sqr_support) {
                    bnl[(size_t)blen * n + i] = j;
                    ++blen;
                    if (WCSPH) {
                        float rn = P.fix_density ? r2 : sqrtf(r2);
                        bd += cube(P.h * P.h - rn) * bphi_s[j];
                    } else {
                        bd += cube(P.H2 - r2) * bphi_s[j];
                    }
                }
            }
        }
    }
    bnc[i] = blen;
    if (WCSPH) {
        // compute_density_pressure (wcsph.cl:64-78)
        d *= P.w_eval * P.M;
        d += P.w_eval * bd;
        density[i] = d;
        pressure[i] = cl_max(0.0f, P.k_gas * (pown7(d / P.rho0) - 1.0f));
    } else {
        // compute_initial_density (pcisph.cl:101)
        density[i] = P.w_eval * ((d * P.M) + bd);
    }
    if (edge) atomicAdd(&st->edge_reads, edge);
}

void launch_lists_density(bool wcsph, int n, const float4* pos_s, const int4* fR, int fepoch, int nb,
                          const float4* bpos_s, const float* bphi_s, const int4* bR, int bepoch, const GridDev& g,
                          const DensityParams& P, int* nl, int* nc, int* bnl, int* bnc, float* density,
                          float* pressure, DevState* st, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "lists_density", s);
    if (wcsph)
        hipLaunchKernelGGL(k_lists_density<true>, dim3((n + 255) / 256), dim3(256), 0, s, n, pos_s, fR, fepoch, nb,
                           bpos_s, bphi_s, bR, bepoch, g, P, nl, nc, bnl, bnc, density, pressure, st);
    else
        hipLaunchKernelGGL(k_lists_density<false>, dim3((n + 255) / 256), dim3(256), 0, s, n, pos_s, fR, fepoch, nb,
                           bpos_s, bphi_s, bR, bepoch, g, P, nl, nc, bnl, bnc, density, pressure, st);
    PROF_END(prof, s);
}

// ------------------------------------------------------------------ normals
// compute_normals (normals.cl:23-70)
__global__ __launch_bounds__(256) void k_normals(int n, const float4* __restrict__ pos_s,
                                                 const float* __restrict__ density, const int* __restrict__ nl,
                                                 const int* __restrict__ nc, float H2, float w_grad, float M, float H,
                                                 float4* __restrict__ normal) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 pi = v3(pos_s[i]);
    V3 acc = {0.0f, 0.0f, 0.0f};
    int len = nc[i];
    for (int k = 0; k < len; ++k) {
        int j = nl[(size_t)k * n + i];
        V3 r = pi - v3(pos_s[j]);
        float q = sqr(H2 - dot(r, r));
        acc = acc + (r * q) / density[j];
    }
    V3 o = w_grad * acc;
    o = o * M;
    o = o * H;
    normal[i] = f4(o, 0.0f);
}

void launch_normals(int n, const float4* pos_s, const float* density, const int* nl, const int* nc, float H2,
                    float w_grad, float M, float H, float4* normal, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "normals", s);
    hipLaunchKernelGGL(k_normals, dim3((n + 255) / 256), dim3(256), 0, s, n, pos_s, density, nl, nc, H2, w_grad, M, H,
                       normal);
    PROF_END(prof, s);
}

// ------------------------------------------------------------- step state
__global__ void k_step_begin(DevState* st) {
    int t = threadIdx.x;
    if (t < MAXERR_SLOTS) st->maxerr[t] = 0u;
    if (t == 0) {
        st->done_iter = 0;
        st->iterations = 0;
        st->out_of_grid = 0;
        st->edge_reads = 0;
        st->nonfinite = 0;
    }
}

void launch_step_begin(DevState* st, hipStream_t s) {
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, s, st);
}

// Fills n words with v (used instead of hipMemsetAsync so that a captured
// step holds kernel nodes only).
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

void launch_fill_u32(void* p, size_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t*)p, n, v);
}

// --------------------------------------------------------------- statistics
__global__ __launch_bounds__(256) void k_stats(int n, const int* __restrict__ nc, const int* __restrict__ bnc,
                                               const float4* __restrict__ pos, unsigned long long* __restrict__ acc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long a = 0, b = 0, of = 0, ob = 0, nf = 0;
    if (i < n) {
        a = (unsigned long long)nc[i];
        b = (unsigned long long)bnc[i];
        of = nc[i] >= LIST_CAP;
        ob = bnc[i] >= LIST_CAP;
        float4 p = pos[i];
        nf = !(isfinite(p.x) && isfinite(p.y) && isfinite(p.z));
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        of += __shfl_xor(of, off);
        ob += __shfl_xor(ob, off);
        nf += __shfl_xor(nf, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&acc[0], a);
        atomicAdd(&acc[1], b);
        atomicAdd(&acc[2], of);
        atomicAdd(&acc[3], ob);
        atomicAdd(&acc[4], nf);
    }
}

void launch_stats(int n, const int* nc, const int* bnc, const float4* pos, unsigned long long* acc, hipStream_t s) {
    launch_fill_u32(acc, 20, 0u, s);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_stats, dim3((n + 255) / 256), dim3(256), 0, s, n, nc, bnc, pos, acc);
}

}  // namespace sph
