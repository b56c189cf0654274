// Neighbour search on gfx950: Morton cell keys, stable LSD radix sort,
// gather into sorted order with cell ranges, and the capped neighbour lists
// fused with the density sum that follows them in the reference.
//
// Restates kernels/grid.cl (compute_hashes :20-39, compute_cell_intervals
// :60-123, compute_neigh_list :145-181), opencl/algorithms/clsort.h:12-35
// (stable key/value sort), opencl/algorithms/clshuffle.h:23-35 (gather) and
// the density kernels that consume the lists (pcisph.cl:33-102,
// wcsph.cl:32-79).
#include <algorithm>
#include <stdexcept>

#include "batch.h"
#include "kernels.h"
#include "sph_device.h"

namespace sph {

// ---------------------------------------------------------------- hashing
// compute_hashes (grid.cl:20-39). The values array of the sort is the
// identity; the first sort pass generates it instead of reading it. Before
// writing the new key, each thread clears the cell range named by the
// previous step's sorted key at its index (prev may alias keys).
__device__ __forceinline__ void clear_range(uint32_t pk, const GridDev& g, int2* __restrict__ R) {
    int px = (int)compact_by_3(pk), py = (int)compact_by_3(pk >> 1), pz = (int)compact_by_3(pk >> 2);
    long long b = py < g.cps ? box_index(px, py, pz, g) : -1;
    if (b >= 0) R[b] = make_int2(0, 0);
}

__device__ __forceinline__ uint32_t cell_key(float4 p, const GridDev& g, DevState* st) {
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) atomicAdd(&st->nonfinite, 1);
    int cx, cy, cz;
    grid_coords(p, g, cx, cy, cz);
    if (box_index(cx, cy, cz, g) < 0) atomicAdd(&st->out_of_grid, 1);
    return morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
}

__device__ __forceinline__ void hash_one(int i, const float4* __restrict__ pos, const GridDev& g, uint32_t* keys,
                                         const uint32_t* prev, int2* __restrict__ R, DevState* st) {
    if (R) clear_range(prev[i], g, R);
    keys[i] = cell_key(pos[i], g, st);
}

__global__ __launch_bounds__(256) void k_hash(int n, const float4* __restrict__ pos, GridDev g,
                                              uint32_t* keys, const uint32_t* prev, int2* __restrict__ R,
                                              DevState* st) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    hash_one(i, pos, g, keys, prev, R, st);
}

// ------------------------------------------------------------- radix sort
// Stable LSD radix sort of (u32 key, i32 value), 8- to 10-bit digits (the
// fewest passes over the key's bits). One pass is
//   hist    per-tile digit counts, stored digit-major hist[d][tile]
//   rowscan one workgroup per digit: exclusive scan of its row, row total
//   scatter each tile adds the exclusive prefix of the row totals to its
//           row entries and writes its keys stably
// The scatter ranks each 256-key round of a tile with wave64 ballots (the
// lanes sharing a digit are the AND of one ballot per digit bit), then orders waves
// through LDS, so the output keeps input order within equal digits.
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 8;  // keys per thread of a tile (2048-key tiles)
// Up to SMALL_SORT_N keys a tile holds one key per thread: a small scene then
// spreads each pass over more workgroups with one round each, instead of a
// few workgroups running their rounds one after another.
constexpr int SMALL_SORT_N = 1 << 16;
constexpr int RADIX_MAX = 1024;

static int sort_items(int n) { return n <= SMALL_SORT_N ? 1 : SORT_ITEMS; }

// Inclusive scan of one value per thread over a 256-thread workgroup.
__device__ __forceinline__ uint32_t block_incl_scan_256(uint32_t v, uint32_t* wtot) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    if (lane == 63) wtot[wave] = v;
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < wave; ++w) add += wtot[w];
    __syncthreads();
    return v + add;
}

// With SID the digit is the scene id of the value (sid[value]) instead of a
// key digit: one such pass after the key passes sorts a packed batch of
// scenes by (scene, key), i.e. each scene's segment by its own keys.
template <int BITS, bool SID>
__device__ __forceinline__ uint32_t radix_digit(uint32_t key, int val, int shift, const int* __restrict__ sid) {
    if (SID) return (uint32_t)sid[val];
    return (key >> shift) & ((1u << BITS) - 1u);
}

template <int BITS, bool SID, int ITEMS>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                             const int* __restrict__ vin, int n, int shift,
                                                             int ntiles, uint32_t* __restrict__ hist,
                                                             const int* __restrict__ sid) {
    constexpr int RADIX = 1 << BITS;
    __shared__ uint32_t h[RADIX];
    const int t = threadIdx.x;
    for (int d = t; d < RADIX; d += SORT_THREADS) h[d] = 0;
    __syncthreads();
    int base = blockIdx.x * (SORT_THREADS * ITEMS);
#pragma unroll 4
    for (int k = 0; k < ITEMS; ++k) {
        int idx = base + k * SORT_THREADS + t;
        if (idx < n) {
            int val = SID ? (vin ? vin[idx] : idx) : 0;
            atomicAdd(&h[radix_digit<BITS, SID>(SID ? 0u : keys[idx], val, shift, sid)], 1u);
        }
    }
    __syncthreads();
    for (int d = t; d < RADIX; d += SORT_THREADS) hist[(size_t)d * ntiles + blockIdx.x] = h[d];
}

__global__ __launch_bounds__(SORT_THREADS) void k_radix_rowscan(uint32_t* __restrict__ hist, int ntiles,
                                                                uint32_t* __restrict__ totals) {
    __shared__ uint32_t wtot[4];
    __shared__ uint32_t chunk_total;
    uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (int base = 0; base < ntiles; base += SORT_THREADS) {
        int idx = base + threadIdx.x;
        uint32_t v = idx < ntiles ? row[idx] : 0u;
        uint32_t incl = block_incl_scan_256(v, wtot);
        if (idx < ntiles) row[idx] = carry + incl - v;
        if (threadIdx.x == SORT_THREADS - 1) chunk_total = incl;
        __syncthreads();
        carry += chunk_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

template <int BITS, bool SID, int ITEMS>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                const int* __restrict__ vin,
                                                                uint32_t* __restrict__ kout, int* __restrict__ vout,
                                                                int n, int shift, int ntiles,
                                                                const uint32_t* __restrict__ offs,
                                                                const uint32_t* __restrict__ totals,
                                                                const int* __restrict__ sid) {
    constexpr int RADIX = 1 << BITS;
    constexpr int PER = RADIX / SORT_THREADS;  // digits per thread for the table set-up
    constexpr int WAVES = SORT_THREADS / 64;
    __shared__ uint32_t base_off[RADIX];
    __shared__ uint32_t run[RADIX];
    __shared__ uint32_t wcnt[WAVES][RADIX];
    __shared__ uint32_t wtot[4];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // digit bases: exclusive prefix of the row totals (PER consecutive digits per thread)
    uint32_t tsum = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) tsum += totals[t * PER + k];
    uint32_t pre = block_incl_scan_256(tsum, wtot) - tsum;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        int d = t * PER + k;
        base_off[d] = offs[(size_t)d * ntiles + blockIdx.x] + pre;
        pre += totals[d];
        run[d] = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) wcnt[w][d] = 0;
    }
    __syncthreads();
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const int base = blockIdx.x * (SORT_THREADS * ITEMS);
    for (int k = 0; k < ITEMS; ++k) {
        int idx = base + k * SORT_THREADS + t;
        bool valid = idx < n;
        uint32_t key = valid ? kin[idx] : 0u;
        int val = valid ? (vin ? vin[idx] : idx) : 0;
        uint32_t d = valid ? radix_digit<BITS, SID>(key, val, shift, sid) : 0u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
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
            uint32_t p = run[d];
            for (int w = 0; w < wave; ++w) p += wcnt[w][d];
            pos = base_off[d] + p + (uint32_t)rank;
        }
        __syncthreads();
        if (leader) {
            bool highest = true;
            for (int w = wave + 1; w < WAVES; ++w)
                if (wcnt[w][d]) highest = false;
            if (highest) {
                uint32_t sum = 0;
                for (int w = 0; w < WAVES; ++w) sum += wcnt[w][d];
                run[d] += sum;
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

// Digit width per pass: the fewest passes over key_bits, 8..10 bits each.
static int digit_bits(int key_bits, int& passes) {
    if (key_bits <= 0) key_bits = 1;
    passes = (key_bits + 9) / 10;
    int bits = (key_bits + passes - 1) / passes;
    return bits < 8 ? 8 : bits;
}

int sort_passes(int key_bits) {
    int p;
    digit_bits(key_bits, p);
    return p;
}

static int sort_tiles(int n) {
    const int tile = SORT_THREADS * sort_items(n);
    return (n + tile - 1) / tile;
}

size_t sort_hist_words(int n) {
    int ntiles = sort_tiles(n);
    return (size_t)RADIX_MAX * (ntiles > 0 ? ntiles : 1) + RADIX_MAX;
}

template <int BITS, bool SID, int ITEMS>
static void radix_pass_items(const uint32_t* kin, const int* vin, uint32_t* kout, int* vout, uint32_t* hist,
                             uint32_t* totals, int n, int ntiles, int shift, const int* sid, hipStream_t s,
                             Profiler* prof) {
    PROF_BEGIN(prof, "radix_hist", s);
    hipLaunchKernelGGL((k_radix_hist<BITS, SID, ITEMS>), dim3(ntiles), dim3(SORT_THREADS), 0, s, kin, vin, n, shift,
                       ntiles, hist, sid);
    PROF_END(prof, s);
    PROF_BEGIN(prof, "radix_rowscan", s);
    hipLaunchKernelGGL(k_radix_rowscan, dim3(1 << BITS), dim3(SORT_THREADS), 0, s, hist, ntiles, totals);
    PROF_END(prof, s);
    PROF_BEGIN(prof, "radix_scatter", s);
    hipLaunchKernelGGL((k_radix_scatter<BITS, SID, ITEMS>), dim3(ntiles), dim3(SORT_THREADS), 0, s, kin, vin, kout,
                       vout, n, shift, ntiles, hist, totals, sid);
    PROF_END(prof, s);
}

template <int BITS, bool SID>
static void radix_pass(const uint32_t* kin, const int* vin, uint32_t* kout, int* vout, uint32_t* hist,
                       uint32_t* totals, int n, int ntiles, int shift, const int* sid, hipStream_t s,
                       Profiler* prof) {
    if (sort_items(n) == 1)
        radix_pass_items<BITS, SID, 1>(kin, vin, kout, vout, hist, totals, n, ntiles, shift, sid, s, prof);
    else
        radix_pass_items<BITS, SID, SORT_ITEMS>(kin, vin, kout, vout, hist, totals, n, ntiles, shift, sid, s, prof);
}

int sort_passes(int key_bits, int sid_bits) { return sort_passes(key_bits) + (sid_bits > 0 ? 1 : 0); }

void launch_radix_sort(uint32_t* keys_a, uint32_t* keys_b, int* vals_a, int* vals_b, uint32_t* hist, int n,
                       int key_bits, hipStream_t s, Profiler* prof, const int* sid, int sid_bits,
                       const int* vals_in) {
    if (n <= 0) return;
    int passes;
    const int bits = digit_bits(key_bits, passes);
    if (sid && sid_bits > 10) throw std::invalid_argument("radix sort: more than 1024 segments");
    const int total = passes + (sid ? 1 : 0);
    const int ntiles = sort_tiles(n);
    uint32_t* totals = hist + (size_t)RADIX_MAX * ntiles;
    // The caller hashed into keys_b when the pass count is odd, keys_a when
    // even; either way the result lands in (keys_a, vals_a).
    uint32_t* kin = (total % 2) ? keys_b : keys_a;
    uint32_t* kout = (total % 2) ? keys_a : keys_b;
    const int* vin = vals_in;
    int* vout = (total % 2) ? vals_a : vals_b;
    for (int p = 0; p < total; ++p) {
        const bool scene_pass = sid && p == passes;
        int shift = bits * p;
        if (scene_pass) {
            if (sid_bits <= 8) radix_pass<8, true>(kin, vin, kout, vout, hist, totals, n, ntiles, 0, sid, s, prof);
            else if (sid_bits == 9) radix_pass<9, true>(kin, vin, kout, vout, hist, totals, n, ntiles, 0, sid, s, prof);
            else radix_pass<10, true>(kin, vin, kout, vout, hist, totals, n, ntiles, 0, sid, s, prof);
        } else if (bits == 8) {
            radix_pass<8, false>(kin, vin, kout, vout, hist, totals, n, ntiles, shift, nullptr, s, prof);
        } else if (bits == 9) {
            radix_pass<9, false>(kin, vin, kout, vout, hist, totals, n, ntiles, shift, nullptr, s, prof);
        } else {
            radix_pass<10, false>(kin, vin, kout, vout, hist, totals, n, ntiles, shift, nullptr, s, prof);
        }
        uint32_t* kt = kin; kin = kout; kout = kt;
        int* nv = (vout == vals_a) ? vals_b : vals_a;
        vin = vout;
        vout = nv;
    }
}

void launch_hash(int n, const float4* pos, const GridDev& g, uint32_t* keys, const uint32_t* prev, int2* R,
                 DevState* st, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "hash", s);
    hipLaunchKernelGGL(k_hash, dim3((n + 255) / 256), dim3(256), 0, s, n, pos, g, keys, prev, R, st);
    PROF_END(prof, s);
}

// ------------------------------------------- sort of a nearly sorted step
// The same stable sort of (key, index) pairs, for keys k that differ little
// from a non-decreasing array P of the same length (the solver's particles
// are stored in the previous step's sorted order, so P is that step's sorted
// keys and k[i] != P[i] only where particle i changed cell). With
//   moved(i)  = k[i] != P[i],
//   p_m       = #{x : (P[x], x) < (k[m], m)}   (insertion point of moved m
//               among all pairs of P; P's pairs increase strictly with x),
//   H[p]      = moved elements with insertion point p (0 <= p <= n),
//   X         = exclusive scan of D[x] = H[x] - moved(x) over 0 .. n,
// the sorted rank of an unmoved i is i + X[i] + H[i] (the unmoved before it
// plus the moved below it), and that of a moved m is p_m + X[p_m] plus m's
// rank by (key, index) among the moved elements with the same insertion
// point (its group). Any non-decreasing P gives the exact stable sort; a P
// close to the result makes the work proportional to the moved elements.
//
// The insertion point of a particle that moved into a cell occupied in P
// comes from P's cell range (the range box the previous step's gather
// wrote), checked against P itself; into an empty cell, from the range end
// of the nearest occupied Morton predecessor among the next few keys, also
// checked; otherwise (or without a range box) from a search of P through
// its samples S (every 64th key) and S2 (every 4096th), which every sort
// writes with its output.
//
// Kernels:
//   k_resort_flag    keys, moved mask, insertion points, per-tile sums of D
//                    and of the moved count; the moved lanes of a wave with
//                    one insertion point (a subgroup) reserve their slots
//                    with one atomic and take them in (key, index) order
//   k_resort_tiles   (more than RESORT_SUM_TILES tiles) one workgroup: the
//                    exclusive scans of the tile sums
//   k_resort_place   each tile's prefix (from that scan, or the sum of the
//                    tile sums before it); the scans of D and of the moved
//                    flags inside the tile; the
//                    unmoved elements to their ranks; each group's base rank
//                    and size; the moved list
//   k_resort_moved   groups that came from one subgroup straight to their
//                    ranks; the others gathered by slot; the cells the
//                    moved elements left cleared in the range box
//   k_resort_groups  gathered groups: up to RESORT_SMALL_GROUP ranked by
//                    counting, larger ones sorted by a workgroup each
// Counters return to zero by the end.
constexpr int RESORT_MIN_TILE_SHIFT = 10;  // scan tiles of 1024 positions (256 threads x 4) ...
constexpr int RESORT_MAX_TILES = 4096;     // ... or larger, so that there are at most this many
// Up to this many tiles each k_resort_place workgroup sums the tile sums
// before its own (no scan launch); beyond, k_resort_tiles scans them once.
constexpr int RESORT_SUM_TILES = 1024;
constexpr int RESORT_SMALL_GROUP = 64;     // groups ranked by their list up to this size
constexpr int RESORT_LDS_GROUP = 4096;     // larger groups sort in LDS up to this size, else in place
constexpr int RESORT_LIST_BLOCKS = 512;    // workgroups that walk the list of moved elements
constexpr int RESORT_BIG_BLOCKS = 64;      // workgroups that sort the groups of more than RESORT_SMALL_GROUP
constexpr int RESORT_PRED = 4;             // Morton predecessors tried for a move into an empty cell
// The tile sums are updated with atomics from every wave; atomics on one
// cache line queue up behind each other, so each tile's counter has a
// 128-byte line to itself.
constexpr int RESORT_PAD = 32;

static int resort_tile_shift(int n) {
    int sh = RESORT_MIN_TILE_SHIFT;
    while ((n >> sh) + 1 > RESORT_MAX_TILES) ++sh;
    return sh;
}
static int resort_tiles(int n) { return (n >> resort_tile_shift(n)) + 1; }  // positions 0 .. n

ResortSizes resort_sizes(int n) {
    ResortSizes z;
    const size_t N = (size_t)(n > 0 ? n : 1);
    z.tiles = (size_t)resort_tiles(n);
    z.padded = z.tiles << resort_tile_shift(n);
    z.mask_words = (N + 63) / 64;
    z.elems = N;
    z.big = N / (RESORT_SMALL_GROUP + 1) + 1;
    z.s1 = N / 64 + 1;
    z.s2 = N / 4096 + 1;
    return z;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte sizes of the work arrays in the order resort_carve lays them out.
constexpr int RESORT_PARTS = 16;
static void resort_parts(int n, size_t out[RESORT_PARTS]) {
    const ResortSizes z = resort_sizes(n);
    const size_t sizes[RESORT_PARTS] = {z.mask_words * 8, z.padded * 4,  z.padded * 8, z.tiles * 4 * RESORT_PAD,
                                        z.tiles * 4 * RESORT_PAD, z.elems * 16, z.elems * 4, z.elems * 4,
                                        z.elems * 4,      z.big * 4,     4,            4,
                                        z.s1 * 4,         z.s2 * 4,      z.tiles * 4,  z.tiles * 4};
    for (int k = 0; k < RESORT_PARTS; ++k) out[k] = sizes[k];
}

size_t resort_bytes(int n) {
    size_t parts[RESORT_PARTS], total = 0;
    resort_parts(n, parts);
    for (size_t b : parts) total += align256(b);
    return total;
}

ResortWork resort_carve(void* block, int n) {
    size_t parts[RESORT_PARTS];
    resort_parts(n, parts);
    unsigned char* q = static_cast<unsigned char*>(block);
    void* a[RESORT_PARTS];
    for (int k = 0; k < RESORT_PARTS; ++k) {
        a[k] = q;
        q += align256(parts[k]);
    }
    ResortWork w;
    w.moved = static_cast<unsigned long long*>(a[0]);
    w.H = static_cast<int*>(a[1]);
    w.grp = static_cast<int2*>(a[2]);
    w.T = static_cast<int*>(a[3]);
    w.TF = static_cast<int*>(a[4]);
    w.rec = static_cast<int4*>(a[5]);
    w.mlist = static_cast<int*>(a[6]);
    w.tk = static_cast<uint32_t*>(a[7]);
    w.ti = static_cast<int*>(a[8]);
    w.big = static_cast<int*>(a[9]);
    w.nbig = static_cast<int*>(a[10]);
    w.nmoved = static_cast<int*>(a[11]);
    w.S = static_cast<uint32_t*>(a[12]);
    w.S2 = static_cast<uint32_t*>(a[13]);
    w.Tx = static_cast<int*>(a[14]);
    w.TFx = static_cast<int*>(a[15]);
    return w;
}

void resort_clear(const ResortWork& w, int n, hipStream_t s) {
    const ResortSizes z = resort_sizes(n);
    launch_fill_u32(w.H, z.padded, 0u, s);
    launch_fill_u32(w.T, z.tiles * RESORT_PAD, 0u, s);
    launch_fill_u32(w.TF, z.tiles * RESORT_PAD, 0u, s);
}

// Every sort writes the samples of its output (rank r, key k).
__device__ __forceinline__ void put_sample(const ResortWork& w, int r, uint32_t k) {
    if ((r & 63) == 0) w.S[r >> 6] = k;
    if ((r & 4095) == 0) w.S2[r >> 12] = k;
}

__global__ __launch_bounds__(256) void k_resort_sample(const uint32_t* __restrict__ P, int n, ResortWork w) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j * 64 < n) put_sample(w, j * 64, P[j * 64]);
}

void launch_resort_samples(const uint32_t* sorted, int n, const ResortWork& w, hipStream_t s) {
    if (n <= 0) return;
    const int m = (n + 63) / 64;
    hipLaunchKernelGGL(k_resort_sample, dim3((m + 255) / 256), dim3(256), 0, s, sorted, n, w);
}

// (v, x) < (k, m) for the pair at position x of P holding v
__device__ __forceinline__ bool below(uint32_t v, int x, uint32_t k, int m) { return v < k || (v == k && x < m); }

// Insertion point of (k, m) by search: the first sample of S2 that is not
// below, then of S inside that stretch, then of P inside the 64 positions
// left. P's pairs below (k, m) form a prefix, so each level is a binary
// search; the samples are small enough to stay in the caches.
__device__ int insertion_search(const uint32_t* __restrict__ P, const ResortWork& w, int n, uint32_t k, int m) {
    int lo = 0, hi = (n + 4095) >> 12;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (below(w.S2[mid], mid << 12, k, m)) lo = mid + 1;
        else hi = mid;
    }
    if (lo == 0) return 0;
    int jlo = ((lo - 1) << 6) + 1, jhi = min(lo << 6, (n + 63) >> 6);
    while (jlo < jhi) {
        const int mid = (jlo + jhi) >> 1;
        if (below(w.S[mid], mid << 6, k, m)) jlo = mid + 1;
        else jhi = mid;
    }
    int xlo = ((jlo - 1) << 6) + 1, xhi = min(jlo << 6, n);
    while (xlo < xhi) {
        const int mid = (xlo + xhi) >> 1;
        if (below(P[mid], mid, k, m)) xlo = mid + 1;
        else xhi = mid;
    }
    return xlo;
}

// P's range of cell key k from the range box ((0, 0) when the cell is
// outside the box or empty).
__device__ __forceinline__ int2 key_range(uint32_t k, const int2* __restrict__ R, const GridDev& g) {
    const int x = (int)compact_by_3(k), y = (int)compact_by_3(k >> 1), z = (int)compact_by_3(k >> 2);
    const long long b = y < g.cps ? box_index(x, y, z, g) : -1;
    return b >= 0 ? R[b] : make_int2(0, 0);
}

// Insertion point of (k, m) inside a segment [lo, hi) of P by search:
// galloping outward from m (a particle that changed cell usually keeps a
// nearby rank), then binary inside the bracket. Invariant: the pairs below
// lo are below (k, m), the pair at hi (or the segment's end) is not.
__device__ int segment_search(const uint32_t* __restrict__ P, int lo, int hi, uint32_t k, int m) {
    const int end = hi;
    if (k > P[m]) {
        lo = m + 1;
        for (long long b = 1;; b <<= 1) {
            const long long x = (long long)m + b;
            if (x >= end) break;
            if (below(P[x], (int)x, k, m)) {
                lo = (int)x + 1;
            } else {
                hi = (int)x;
                break;
            }
        }
    } else {
        const int start = lo;
        hi = m;
        for (long long b = 1;; b <<= 1) {
            const long long x = (long long)m - b;
            if (x < start) break;
            if (below(P[x], (int)x, k, m)) {
                lo = (int)x + 1;
                break;
            }
            hi = (int)x;
        }
    }
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (below(P[mid], mid, k, m)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Insertion point of (k, m) inside the segment [lo, hi) of P that holds m
// (the whole array for one scene, the scene's particles in a batch): from
// P's range of k when P holds k (the range is accepted only if P confirms
// both ends), else from the end of the nearest occupied predecessor's range
// when P confirms that k falls right there, else by search (through the
// samples for a whole array, galloping inside a segment). The confirming
// reads are issued together (indices clamped into the segment; the tests
// only use the ones that apply).
__device__ int insertion_point(const uint32_t* __restrict__ P, const ResortWork& w, int n, int lo, int hi, uint32_t k,
                               int m, const int2* __restrict__ R, const GridDev& g) {
    if (R) {
        const int2 r = key_range(k, R, g);
        if (r.y > r.x && r.x >= lo && r.y <= hi) {
            const uint32_t a0 = P[r.x > lo ? r.x - 1 : lo], a1 = P[r.x], a2 = P[r.y - 1];
            const uint32_t a3 = P[r.y < hi ? r.y : hi - 1];
            if ((a1 == k) & (a2 == k) & ((r.x == lo) | (a0 < k)) & ((r.y == hi) | (a3 > k)))
                return m < r.x ? r.x : (m > r.y ? r.y : m);
        } else if (r.y <= r.x) {
            int2 q[RESORT_PRED];
#pragma unroll
            for (int t = 0; t < RESORT_PRED; ++t)
                q[t] = (uint32_t)(t + 1) <= k ? key_range(k - (uint32_t)(t + 1), R, g) : make_int2(0, 0);
            int p = -1;
#pragma unroll
            for (int t = RESORT_PRED - 1; t >= 0; --t)
                if (q[t].y > q[t].x) p = q[t].y;  // the nearest occupied predecessor wins
            if (p > lo && p <= hi) {
                const uint32_t a0 = P[p - 1], a1 = P[p < hi ? p : hi - 1];
                if ((a0 < k) & ((p == hi) | (a1 > k))) return p;
            }
        }
    }
    return (lo == 0 && hi == n) ? insertion_search(P, w, n, k, m) : segment_search(P, lo, hi, k, m);
}

// HASH: the keys are the cells of pos (compute_hashes, with k_hash's
// counters) and are written to keys; otherwise keys holds them. R (may be
// null): the range box of P's cells. With sid and sc (a batch), particle i
// belongs to scene sc[sid[i]]: its grid, its range box (R + roff) and its
// segment [lo, hi), inside which it stays.
template <bool HASH>
__global__ __launch_bounds__(256) void k_resort_flag(int n, int tile_shift, const float4* __restrict__ pos, GridDev g0,
                                                     uint32_t* __restrict__ keys, const uint32_t* __restrict__ P,
                                                     const int2* __restrict__ R0, DevState* st, ResortWork w,
                                                     const int* __restrict__ sid, const BatchScene* __restrict__ sc) {
    const int i = (HASH ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x) * blockDim.x + threadIdx.x;
    bool moved = false;
    uint32_t k = 0, pk = 0;
    GridDev g = g0;
    const int2* R = R0;
    int lo = 0, hi = n;
    if (i < n) {
        if (sid) {
            const BatchScene& S = sc[sid[i]];
            g = S.g;
            R = R0 ? R0 + S.roff : nullptr;
            lo = S.lo;
            hi = S.hi;
        }
        pk = P[i];
        if (HASH) {
            k = cell_key(pos[i], g, st);
            keys[i] = k;
        } else {
            k = keys[i];
        }
        moved = k != pk;
    }
    const uint64_t b = __ballot(moved);
    const int lane = threadIdx.x & 63;
    const int i0 = i - lane;  // blocks are whole waves, so i0 is 64-aligned
    if (lane == 0 && i0 < n) w.moved[i0 >> 6] = b;
    if (!b) return;
    const int p = moved ? insertion_point(P, w, n, lo, hi, k, i, R, g) : -1;
    // The wave's moved lanes with the same insertion point form a subgroup:
    // its first lane (the leader) reserves the subgroup's slots with one
    // atomic, and each member's slot is the leader's base plus the member's
    // rank by (key, index) inside the subgroup (lanes are in index order).
    int leader = lane, cnt = 0, rank = 0;
    for (uint64_t todo = b; todo;) {  // wave-uniform
        const int l = __ffsll((long long)todo) - 1;
        const int pl = __shfl(p, l, 64);
        const uint32_t kl = __shfl(k, l, 64);
        const bool mine = moved && p == pl;
        const uint64_t same = __ballot(mine);
        const bool one_key = __ballot(mine && k != kl) == 0ull;
        if (mine) {
            leader = l;
            cnt = __popcll(same);
            if (one_key) {
                rank = __popcll(same & ((1ull << lane) - 1ull));
            } else {
                rank = 0;
                for (uint64_t m = same; m;) {
                    const int j = __ffsll((long long)m) - 1;
                    m &= m - 1;
                    const uint32_t kj = __shfl(k, j, 64);  // every lane of `same` takes part
                    rank += (kj < k || (kj == k && j < lane)) ? 1 : 0;
                }
            }
        }
        todo &= ~same;
    }
    int base = 0;
    if (moved && leader == lane) base = atomicAdd(&w.H[p], cnt);
    base = __shfl(base, leader, 64);
    // tile sums: the wave's own tile gets -(moved lanes) plus its subgroups
    // that insert there, in one atomic; subgroups inserting elsewhere add
    // their own counts
    const int own = i0 >> tile_shift;
    const bool lead = moved && leader == lane;
    const bool home = lead && (p >> tile_shift) == own;
    int d = home ? cnt : 0;
    for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off, 64);
    if (lane == 0) {
        atomicAdd(&w.T[own * RESORT_PAD], d - __popcll(b));
        atomicAdd(&w.TF[own * RESORT_PAD], __popcll(b));
    }
    if (lead && !home) atomicAdd(&w.T[(p >> tile_shift) * RESORT_PAD], cnt);
    if (moved) {
        // z: the subgroup's size when it took the group's first slots, else 0
        // (a group whose size equals that number came from one wave, ranked)
        w.rec[i] = make_int4(p, base + rank, base == 0 ? cnt : 0, (int)pk);
    }
}

// Inclusive scan of one value per thread over a 1024-thread workgroup
// (wsum: 16 words of LDS).
__device__ __forceinline__ int block_incl_scan_1024(int v, int* wsum) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    int add = 0;
    for (int q = 0; q < wave; ++q) add += wsum[q];
    __syncthreads();
    return v + add;
}

// Sums of the padded counters a[t * RESORT_PAD] and b[t * RESORT_PAD] over
// t < m (every thread of the workgroup gets them).
__device__ __forceinline__ int2 block_sum2(const int* __restrict__ a, const int* __restrict__ b, int m,
                                           int2* part) {
    int x = 0, y = 0;
    for (int t = threadIdx.x; t < m; t += blockDim.x) {
        x += a[t * RESORT_PAD];
        y += b[t * RESORT_PAD];
    }
    for (int off = 32; off > 0; off >>= 1) {
        x += __shfl_xor(x, off, 64);
        y += __shfl_xor(y, off, 64);
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = make_int2(x, y);
    __syncthreads();
    int2 r = make_int2(0, 0);
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
        r.x += part[q].x;
        r.y += part[q].y;
    }
    __syncthreads();
    return r;
}

// Exclusive scans of the padded tile sums (D into Tx, the moved count into
// TFx) and the total moved count, in one workgroup; leaves the sums zero for
// the next sort and clears the count of large groups.
__global__ __launch_bounds__(1024) void k_resort_tiles(int ntiles, ResortWork w) {
    __shared__ int wsum[16];
    __shared__ int tot[2];
    const int t = threadIdx.x;
    int carry = 0, carry_f = 0;
    for (int base = 0; base < ntiles; base += 1024) {
        const int idx = base + t;
        const int v = idx < ntiles ? w.T[idx * RESORT_PAD] : 0;
        const int f = idx < ntiles ? w.TF[idx * RESORT_PAD] : 0;
        const int sv = block_incl_scan_1024(v, wsum);
        const int sf = block_incl_scan_1024(f, wsum);
        if (idx < ntiles) {
            w.Tx[idx] = carry + sv - v;
            w.TFx[idx] = carry_f + sf - f;
            w.T[idx * RESORT_PAD] = 0;
            w.TF[idx * RESORT_PAD] = 0;
        }
        if (t == 1023) {
            tot[0] = sv;
            tot[1] = sf;
        }
        __syncthreads();
        carry += tot[0];
        carry_f += tot[1];
        __syncthreads();
    }
    if (t == 0) {
        *w.nbig = 0;
        *w.nmoved = carry_f;
    }
}

// One tile of 2^tile_shift positions, `per` consecutive per thread (a
// multiple of 4): the unmoved elements written to their ranks, the moved
// ones to the moved list (index order), and for each group its base rank
// p + X[p], size and list head (grp[p]), with H and the list head reset.
// H and the list heads are padded to whole tiles.
__global__ __launch_bounds__(256) void k_resort_place(int n, int tile_shift, int ntiles, int scanned,
                                                      const uint32_t* __restrict__ keys, ResortWork w,
                                                      uint32_t* __restrict__ sk, int* __restrict__ perm) {
    __shared__ uint32_t wtot[4];
    __shared__ int2 part[4];
    const int per = (1 << tile_shift) >> 8;
    const int base = (blockIdx.x << tile_shift) + threadIdx.x * per;
    // the tile sums before this tile: scanned by k_resort_tiles, or summed here
    const int2 before = scanned ? make_int2(w.Tx[blockIdx.x], w.TFx[blockIdx.x])
                                : block_sum2(w.T, w.TF, blockIdx.x, part);
    int s = 0, nf = 0;
    for (int c = 0; c < per; c += 4) {
        const int4 hv = *reinterpret_cast<const int4*>(w.H + base + c);
        const uint32_t fl = base + c < n ? (uint32_t)(w.moved[(base + c) >> 6] >> ((base + c) & 63)) & 0xfu : 0u;
        s += hv.x + hv.y + hv.z + hv.w - __popc(fl);
        nf += __popc(fl);
    }
    const int incl = (int)block_incl_scan_256((uint32_t)s, wtot);
    const int incl_f = (int)block_incl_scan_256((uint32_t)nf, wtot);
    int x = before.x + incl - s;
    int f = before.y + incl_f - nf;
    if (!scanned) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *w.nbig = 0;
        if (blockIdx.x == ntiles - 1 && threadIdx.x == 255) *w.nmoved = before.y + incl_f;
    }
    for (int c = 0; c < per; c += 4) {
        const int4 hv = *reinterpret_cast<const int4*>(w.H + base + c);
        const int h[4] = {hv.x, hv.y, hv.z, hv.w};
        const uint32_t fl = base + c < n ? (uint32_t)(w.moved[(base + c) >> 6] >> ((base + c) & 63)) & 0xfu : 0u;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = base + c + u;
            const bool mv = (fl >> u) & 1u;
            if (h[u] > 0) {
                w.grp[idx] = make_int2(idx + x, h[u]);
                w.H[idx] = 0;
            }
            if (mv) {
                w.mlist[f] = idx;
                ++f;
            } else if (idx < n) {
                const int r = idx + x + h[u];
                if ((unsigned)r < (unsigned)n) {  // always, for consistent work arrays
                    const uint32_t k = keys[idx];
                    sk[r] = k;
                    perm[r] = idx;
                    put_sample(w, r, k);
                }
            }
            x += h[u] - (int)mv;
        }
    }
}

__device__ __forceinline__ bool pair_less(uint32_t ka, int ia, uint32_t kb, int ib) {
    return ka < kb || (ka == kb && ia < ib);
}

// Ascending bitonic network on G = 2^m positions (the form in which every
// comparator puts the smaller pair at the lower index: the first merge step
// of each size pairs a with a ^ (size - 1), the others a with a + stride).
// Positions at or past g hold +infinity: their comparators are skipped.
template <class Load, class Store>
__device__ void bitonic_pairs(int g, int G, Load load, Store store) {
    for (int size = 2; size <= G; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int e = threadIdx.x; e < (G >> 1); e += blockDim.x) {
                const int a = 2 * e - (e & (stride - 1));
                const int c = stride == (size >> 1) ? (a ^ (size - 1)) : a + stride;
                if (c >= g) continue;
                uint32_t ka, kc;
                int ia, ic;
                load(a, ka, ia);
                load(c, kc, ic);
                if (pair_less(kc, ic, ka, ia)) {
                    store(a, kc, ic);
                    store(c, ka, ia);
                }
            }
            __syncthreads();
        }
    }
}

// One large group (gathered at tk/ti[b .. b + g)) sorted by the workgroup,
// in LDS when it fits, and written to its ranks.
__device__ void sort_group(int b, int g, const ResortWork& w, uint32_t* __restrict__ sk, int* __restrict__ perm,
                           uint32_t* lk, int* li) {
    uint32_t* gk = w.tk + b;
    int* gi = w.ti + b;
    int G = 1;
    while (G < g) G <<= 1;
    if (G <= RESORT_LDS_GROUP) {
        for (int u = threadIdx.x; u < g; u += blockDim.x) {
            lk[u] = gk[u];
            li[u] = gi[u];
        }
        __syncthreads();
        bitonic_pairs(g, G, [&](int a, uint32_t& k, int& v) { k = lk[a]; v = li[a]; },
                      [&](int a, uint32_t k, int v) { lk[a] = k; li[a] = v; });
        for (int u = threadIdx.x; u < g; u += blockDim.x) {
            sk[b + u] = lk[u];
            perm[b + u] = li[u];
            put_sample(w, b + u, lk[u]);
        }
    } else {
        bitonic_pairs(g, G, [&](int a, uint32_t& k, int& v) { k = gk[a]; v = gi[a]; },
                      [&](int a, uint32_t k, int v) { gk[a] = k; gi[a] = v; });
        for (int u = threadIdx.x; u < g; u += blockDim.x) {
            sk[b + u] = gk[u];
            perm[b + u] = gi[u];
            put_sample(w, b + u, gk[u]);
        }
    }
    __syncthreads();  // the LDS is read by every thread before the next group overwrites it
}

// A group came from one wave's subgroup when that subgroup took its first
// slots and is as large as the group: its slots are then the ranks.
__device__ __forceinline__ bool ranked_group(int4 rc, int2 e) { return rc.z == e.y; }

// Moved elements (the list, walked by a fixed grid): elements of groups
// that came from one wave go straight to their ranks; the others are
// gathered at base + slot for k_resort_groups (and a group of more than
// RESORT_SMALL_GROUP is listed for a workgroup sort). With a range box, each
// moved element clears the range of the cell it left (cells that keep
// particles are rewritten by the gather that follows).
__global__ __launch_bounds__(256) void k_resort_moved(int n, int ntiles, int scanned, int big_cap,
                                                      const uint32_t* __restrict__ keys, ResortWork w,
                                                      uint32_t* __restrict__ sk, int* __restrict__ perm,
                                                      int2* __restrict__ R, GridDev g,
                                                      const int* __restrict__ sid, const BatchScene* __restrict__ sc) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    if (!scanned) {  // the tile sums back to zero (k_resort_tiles clears them when it runs)
        for (int t = gid; t < ntiles; t += stride) {
            w.T[t * RESORT_PAD] = 0;
            w.TF[t * RESORT_PAD] = 0;
        }
    }
    const int nm = *w.nmoved;
    for (int q = gid; q < nm; q += stride) {
        const int i = w.mlist[q];
        if ((unsigned)i >= (unsigned)n) continue;  // never, for consistent work arrays
        const int4 rc = w.rec[i];  // (insertion point, slot, first subgroup's size or 0, key in the hint)
        if (R) {
            if (sid) {
                const BatchScene& S = sc[sid[i]];
                clear_range((uint32_t)rc.w, S.g, R + S.roff);
            } else {
                clear_range((uint32_t)rc.w, g, R);
            }
        }
        const int2 e = w.grp[rc.x];
        const int b = e.x, gs = e.y;
        if (gs < 1 || b < 0 || b > n - gs || (unsigned)rc.y >= (unsigned)gs) continue;
        const uint32_t k = keys[i];
        if (ranked_group(rc, e)) {
            sk[b + rc.y] = k;
            perm[b + rc.y] = i;
            put_sample(w, b + rc.y, k);
            continue;
        }
        w.tk[b + rc.y] = k;
        w.ti[b + rc.y] = i;
        if (gs > RESORT_SMALL_GROUP && rc.y == 0) {
            const int at = atomicAdd(w.nbig, 1);
            if (at < big_cap) w.big[at] = rc.x;
        }
    }
}

// The gathered groups: up to RESORT_SMALL_GROUP, each member counts the
// members below it (the first list_blocks workgroups walk the moved list);
// larger ones are sorted by one workgroup each (the other workgroups).
__global__ __launch_bounds__(256) void k_resort_groups(int n, int list_blocks, int big_cap,
                                                       const uint32_t* __restrict__ keys, ResortWork w,
                                                       uint32_t* __restrict__ sk, int* __restrict__ perm) {
    __shared__ uint32_t lk[RESORT_LDS_GROUP];
    __shared__ int li[RESORT_LDS_GROUP];
    if ((int)blockIdx.x < list_blocks) {
        const int nm = *w.nmoved;
        for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nm; q += list_blocks * blockDim.x) {
            const int i = w.mlist[q];
            if ((unsigned)i >= (unsigned)n) continue;
            const int4 rc = w.rec[i];
            const int2 e = w.grp[rc.x];
            const int b = e.x, gs = e.y;
            if (gs < 1 || b < 0 || b > n - gs || ranked_group(rc, e) || gs > RESORT_SMALL_GROUP) continue;
            const uint32_t k = keys[i];
            int r = 0;
            for (int u = b; u < b + gs; ++u) r += pair_less(w.tk[u], w.ti[u], k, i) ? 1 : 0;
            sk[b + r] = k;
            perm[b + r] = i;
            put_sample(w, b + r, k);
        }
        return;
    }
    const int nb = min(*w.nbig, big_cap);
    for (int q = blockIdx.x - list_blocks; q < nb; q += gridDim.x - list_blocks) {
        const int2 e = w.grp[w.big[q]];
        if (e.y < 1 || e.x < 0 || e.x > n - e.y) continue;  // uniform over the workgroup
        sort_group(e.x, e.y, w, sk, perm, lk, li);
    }
}

// keys_out/perm: the sorted keys and the permutation. P (read) and keys_out
// may be the same array: P is last read by k_resort_flag.
static void resort_rest(int n, const uint32_t* keys, uint32_t* keys_out, int* perm, const ResortWork& w, int2* R,
                        const GridDev& g, const int* sid, const BatchScene* sc, hipStream_t s, Profiler* prof) {
    const int sh = resort_tile_shift(n), ntiles = resort_tiles(n);
    const int list_blocks = std::min(RESORT_LIST_BLOCKS, (n + 255) / 256);
    const int scanned = ntiles > RESORT_SUM_TILES ? 1 : 0;
    if (scanned) {
        PROF_BEGIN(prof, "resort_scan", s);
        hipLaunchKernelGGL(k_resort_tiles, dim3(1), dim3(1024), 0, s, ntiles, w);
        PROF_END(prof, s);
    }
    PROF_BEGIN(prof, "resort_place", s);
    hipLaunchKernelGGL(k_resort_place, dim3(ntiles), dim3(256), 0, s, n, sh, ntiles, scanned, keys, w, keys_out,
                       perm);
    PROF_END(prof, s);
    const int big_cap = (int)resort_sizes(n).big;
    PROF_BEGIN(prof, "resort_moved", s);
    hipLaunchKernelGGL(k_resort_moved, dim3(list_blocks), dim3(256), 0, s, n, ntiles, scanned, big_cap, keys, w,
                       keys_out, perm, R, g, sid, sc);
    PROF_END(prof, s);
    PROF_BEGIN(prof, "resort_groups", s);
    hipLaunchKernelGGL(k_resort_groups, dim3(list_blocks + RESORT_BIG_BLOCKS), dim3(256), 0, s, n, list_blocks,
                       big_cap, keys, w, keys_out, perm);
    PROF_END(prof, s);
}

void launch_hash_resort(int n, const float4* pos, const GridDev& g, uint32_t* keys, uint32_t* sorted, int* perm,
                        int2* R, DevState* st, const ResortWork& w, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "hash_moved", s);
    hipLaunchKernelGGL(k_resort_flag<true>, dim3((n + 255) / 256), dim3(256), 0, s, n, resort_tile_shift(n), pos, g,
                       keys, sorted, R, st, w, nullptr, nullptr);
    PROF_END(prof, s);
    resort_rest(n, keys, sorted, perm, w, R, g, nullptr, nullptr, s, prof);
}

void launch_batch_hash_resort(int n, const int* sid, const BatchScene* sc, const float4* pos, uint32_t* keys,
                              uint32_t* sorted, int* perm, int2* R, DevState* st, const ResortWork& w, hipStream_t s,
                              Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "batch_hash_moved", s);
    hipLaunchKernelGGL(k_resort_flag<true>, dim3((n + 255) / 256), dim3(256), 0, s, n, resort_tile_shift(n), pos,
                       GridDev{}, keys, sorted, R, st, w, sid, sc);
    PROF_END(prof, s);
    resort_rest(n, keys, sorted, perm, w, R, GridDev{}, sid, sc, s, prof);
}

void launch_resort(int n, const uint32_t* keys, uint32_t* sorted, int* perm, const ResortWork& w, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_resort_flag<false>, dim3((n + 255) / 256), dim3(256), 0, s, n, resort_tile_shift(n), nullptr,
                       GridDev{}, const_cast<uint32_t*>(keys), sorted, nullptr, nullptr, w, nullptr, nullptr);
    resort_rest(n, keys, sorted, perm, w, nullptr, GridDev{}, nullptr, nullptr, s, nullptr);
}

// ------------------------------------------------------ gather + ranges
// clshuffle of positions/velocities (pcisphsimluation.cpp:89-90,
// wcsphsimluation.cpp:57-59) fused with compute_cell_intervals
// (grid.cl:60-123): a cell's range starts where the sorted key changes and
// ends before the next change.
__device__ __forceinline__ void write_range(int i, int lo, int hi, const uint32_t* __restrict__ keys_s, float4 p,
                                            int2* __restrict__ R, const GridDev& g) {
    uint32_t key = keys_s[i];
    bool first = (i == lo) || keys_s[i - 1] != key;
    bool last = (i == hi - 1) || keys_s[i + 1] != key;
    if (first || last) {
        int cx, cy, cz;
        grid_coords(p, g, cx, cy, cz);
        long long b = box_index(cx, cy, cz, g);
        if (b >= 0) {
            if (first) R[b].x = i;
            if (last) R[b].y = i + 1;
        }
    }
}

// sorted index i of the segment [lo, hi) that holds one simulation
__device__ __forceinline__ void gather_one(int i, int lo, int hi, const int* __restrict__ perm,
                                           const uint32_t* __restrict__ keys_s, const float4* __restrict__ pos,
                                           const float4* __restrict__ vel, const float4* __restrict__ vh,
                                           float4* __restrict__ pos_s, float4* __restrict__ vel_s,
                                           float4* __restrict__ vh_s, int2* __restrict__ R, const GridDev& g) {
    int j = perm[i];
    if (j < lo || j >= hi) j = i;  // a permutation never leaves its segment
    float4 p = pos[j];
    pos_s[i] = p;
    vel_s[i] = vel[j];
    if (vh) vh_s[i] = vh[j];
    write_range(i, lo, hi, keys_s, p, R, g);
}

__global__ __launch_bounds__(256) void k_gather_fluid(int n, const int* __restrict__ perm,
                                                      const uint32_t* __restrict__ keys_s,
                                                      const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                      const float4* __restrict__ vh, float4* __restrict__ pos_s,
                                                      float4* __restrict__ vel_s, float4* __restrict__ vh_s,
                                                      int2* __restrict__ R, GridDev g) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gather_one(i, 0, n, perm, keys_s, pos, vel, vh, pos_s, vel_s, vh_s, R, g);
}

// The gather of a step the adjoint records: also keeps the permutation and
// carries the particle ids to the new order (ids_out[i] = ids_in[perm[i]],
// -1 for an index outside [0, n)).
__global__ __launch_bounds__(256) void k_gather_fluid_tape(int n, const int* __restrict__ perm,
                                                           const uint32_t* __restrict__ keys_s,
                                                           const float4* __restrict__ pos,
                                                           const float4* __restrict__ vel,
                                                           const float4* __restrict__ vh, float4* __restrict__ pos_s,
                                                           float4* __restrict__ vel_s, float4* __restrict__ vh_s,
                                                           int2* __restrict__ R, GridDev g, int* __restrict__ perm_out,
                                                           const int* __restrict__ ids_in,
                                                           int* __restrict__ ids_out) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gather_one(i, 0, n, perm, keys_s, pos, vel, vh, pos_s, vel_s, vh_s, R, g);
    const int j = perm[i];
    perm_out[i] = j;
    ids_out[i] = (unsigned)j < (unsigned)n ? ids_in[j] : -1;
}

__global__ __launch_bounds__(256) void k_gather_boundary(int n, const int* __restrict__ perm,
                                                         const uint32_t* __restrict__ keys_s,
                                                         const float4* __restrict__ pos, const float* __restrict__ phi,
                                                         float4* __restrict__ pos_s, float* __restrict__ phi_s,
                                                         int2* __restrict__ R, GridDev g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int j = perm[i];
    if ((unsigned)j >= (unsigned)n) j = i;
    float4 p = pos[j];
    pos_s[i] = p;
    phi_s[i] = phi[j];
    write_range(i, 0, n, keys_s, p, R, g);
}

void launch_gather_fluid(int n, const int* perm, const uint32_t* keys_s, const float4* pos, const float4* vel,
                         const float4* vh, float4* pos_s, float4* vel_s, float4* vh_s, int2* R, const GridDev& g,
                         hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "gather_ranges", s);
    hipLaunchKernelGGL(k_gather_fluid, dim3((n + 255) / 256), dim3(256), 0, s, n, perm, keys_s, pos, vel, vh, pos_s,
                       vel_s, vh_s, R, g);
    PROF_END(prof, s);
}

void launch_gather_fluid_tape(int n, const int* perm, const uint32_t* keys_s, const float4* pos, const float4* vel,
                              const float4* vh, float4* pos_s, float4* vel_s, float4* vh_s, int2* R, const GridDev& g,
                              int* perm_out, const int* ids_in, int* ids_out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_fluid_tape, dim3((n + 255) / 256), dim3(256), 0, s, n, perm, keys_s, pos, vel, vh,
                       pos_s, vel_s, vh_s, R, g, perm_out, ids_in, ids_out);
}

void launch_gather_boundary(int n, const int* perm, const uint32_t* keys_s, const float4* pos, const float* phi,
                            float4* pos_s, float* phi_s, int2* R, const GridDev& g, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "gather_boundary", s);
    hipLaunchKernelGGL(k_gather_boundary, dim3((n + 255) / 256), dim3(256), 0, s, n, perm, keys_s, pos, phi, pos_s,
                       phi_s, R, g);
    PROF_END(prof, s);
}

// Marks, for every interior cell, which of its 27 neighbour cells hold
// boundary particles: bit o (CELL_NEIGH_OFFSET order) of the cell's mask. A
// fluid particle in an interior cell reads the boundary ranges of the marked
// cells only; with no bit set its boundary list is empty and it reads none.
__global__ __launch_bounds__(256) void k_mark_boundary(int n, const float4* __restrict__ bpos_s,
                                                       const uint32_t* __restrict__ keys_s, GridDev g,
                                                       uint32_t* __restrict__ near) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i > 0 && keys_s[i - 1] == keys_s[i]) return;  // one thread per occupied cell
    int cx, cy, cz;
    grid_coords(bpos_s[i], g, cx, cy, cz);
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                int x = cx - dx, y = cy - dy, z = cz - dz;
                if (!interior_cell(x, y, z, g.cps)) continue;
                long long b = box_index(x, y, z, g);
                if (b >= 0) atomicOr(&near[b], 1u << neigh_index(dx, dy, dz));
            }
}

void launch_mark_boundary(int n, const float4* bpos_s, const uint32_t* keys_s, const GridDev& g, uint32_t* near,
                          size_t near_words, hipStream_t s, Profiler* prof) {
    launch_fill_u32(near, near_words, 0u, s);
    if (n <= 0) return;
    PROF_BEGIN(prof, "mark_boundary", s);
    hipLaunchKernelGGL(k_mark_boundary, dim3((n + 255) / 256), dim3(256), 0, s, n, bpos_s, keys_s, g, near);
    PROF_END(prof, s);
}

// ----------------------------------------------- lists fused with density
// compute_neigh_list (grid.cl:145-181): own cell first, then CELL_NEIGH_OFFSET
// order, sorted order inside a cell, strict r2 < CELL_SIZE^2, cap 50,
// column-major list[k*N + i]. The density kernels read exactly these r2 in
// exactly this order, so the sum is accumulated while the list is written.
// Each cell's range is fetched while the previous cell is processed, and a
// cell's candidates are loaded WALK_BATCH at a time, so several loads are in
// flight per thread; they are consumed in order, which keeps both the list
// and the sum identical.
template <bool WCSPH>
__device__ __forceinline__ void accumulate_density(float r2, float& d, const DensityParams& P, float w,
                                                   bool weighted) {
    float t;
    if (WCSPH) {
        float rn = P.fix_density ? r2 : sqrtf(r2);
        t = cube(P.h * P.h - rn);
    } else {
        t = cube(P.H2 - r2);
    }
    d += weighted ? t * w : t;
}

// Candidates per batch. A cell holds about 8 particles at h = 2 x spacing; a
// batch runs for the whole wave while any lane has candidates left, so when
// one of the wave's cells holds 9-12, batches of 4 waste fewer slots than
// batches of 8 (1M throughput scene: walk 215 -> 202 us; 3 and 6 are slower).
constexpr int WALK_BATCH = 4;
// The split walk of small scenes (k_lists_density_split) keeps batches of 8.
constexpr int SPLIT_BATCH = 8;
constexpr int LISTS_THREADS = 64;
// List slots kept in LDS until the walk ends; slots past it are stored
// straight to the list. Lanes of a wave find neighbours at different times,
// so storing each entry when it is found writes a different list row per lane
// and every row reaches HBM as many partial lines. From LDS the rows go out
// one at a time, 64 consecutive entries per wave.
constexpr int STAGE_SLOTS = 28;

template <int TPB, int SLOTS>
struct ListSink {
    int* stage;  // this thread's column of the LDS stage, stride TPB
    int* out;    // list + i, stride n
    int n;
    __device__ __forceinline__ void put(int k, int j) const {
        if (k < SLOTS)
            stage[k * TPB] = j;
        else
            out[(size_t)k * n] = j;
    }
    __device__ __forceinline__ void flush(int len) const {
        const int m = len < SLOTS ? len : SLOTS;
        for (int k = 0; k < m; ++k) out[(size_t)k * n] = stage[k * TPB];
    }
};


// The hits of one batch of candidates j0 .. j0 + WALK_BATCH - 1, in order.
// With CHECKS the batch may run past the cell's end (rem candidates left)
// or past the list cap; without, the caller has established that neither
// can happen for any lane of the wave.
template <bool CHECKS, bool WCSPH, bool PHI, class Sink>
__device__ __forceinline__ void batch_hits(V3 pi, int j0, int rem, const float4 (&q)[WALK_BATCH],
                                           const float (&w)[WALK_BATCH], const GridDev& g, const DensityParams& P,
                                           const Sink& list, int& len, float& d) {
#pragma unroll
    for (int u = 0; u < WALK_BATCH; ++u) {
        V3 r = pi - v3(q[u]);
        float r2 = dot(r, r);
        bool hit = r2 < g.sqr_support;
        if (CHECKS) hit = hit && u < rem && len < LIST_CAP;
        if (hit) {
            list.put(len, j0 + u);
            ++len;
            accumulate_density<WCSPH>(r2, d, P, w[u], PHI);
        }
    }
}

// The walk over the 27 cells. Candidates are loaded WALK_BATCH at a time
// through a buffer view, from one base offset per batch (the per-slot offset
// is a scalar operand), 12 bytes each (w is not needed); loads past the
// cell's end read neighbours or zero and are never used. The distance of every slot is computed unconditionally;
// the range, cap and radius tests only guard the list entry and the density
// term, and the range and cap tests are skipped for batches where the whole
// wave has a full batch left and room for it under the cap. Bit o of `cells`
// (CELL_NEIGH_OFFSET order) selects the cells that are read; the others count
// as empty without a range read, which callers use only around interior
// cells, whose 27 cells all lie inside the grid.
template <bool WCSPH, bool PHI, class Sink>
__device__ __forceinline__ int walk_cells(V3 pi, int cx, int cy, int cz, const int2* __restrict__ R,
                                          const float4* __restrict__ P4, const float* __restrict__ phi,
                                          int count, const GridDev& g, const DensityParams& P,
                                          const Sink& list, float& d, int& edge, uint32_t cells = ALL_CELLS) {
    const __amdgpu_buffer_rsrc_t rp = buffer_view(P4, (uint32_t)count * 16u);
    const __amdgpu_buffer_rsrc_t rw = buffer_view(phi, PHI ? (uint32_t)count * 4u : 0u);
    int len = 0;
    int2 cur = (cells & 1u) ? cell_range(R, cx, cy, cz, g, edge) : make_int2(0, 0);  // own cell first
    for (int o = 0; o < 27 && len < LIST_CAP; ++o) {
        // fetch the next cell's range while this cell is processed
        int2 nxt = make_int2(0, 0);
        if (o + 1 < 27) {
            int dx, dy, dz;
            neigh_offset(o + 1, dx, dy, dz);
            if ((cells >> (o + 1)) & 1u) nxt = cell_range(R, cx + dx, cy + dy, cz + dz, g, edge);
        }
        const int a = cur.x, b = cur.y < count ? cur.y : count;  // ranges never exceed count
        for (int j0 = a; j0 < b && len < LIST_CAP; j0 += WALK_BATCH) {
            const int rem = b - j0;
            float4 q[WALK_BATCH];
            float w[WALK_BATCH];
#pragma unroll
            for (int u = 0; u < WALK_BATCH; ++u) {
                q[u] = load3(rp, (uint32_t)j0 * 16u, 16u * u);
                w[u] = PHI ? load1(rw, (uint32_t)j0 * 4u, 4u * u) : 1.0f;
            }
            if (__all(rem >= WALK_BATCH && len <= LIST_CAP - WALK_BATCH))
                batch_hits<false, WCSPH, PHI>(pi, j0, rem, q, w, g, P, list, len, d);
            else
                batch_hits<true, WCSPH, PHI>(pi, j0, rem, q, w, g, P, list, len, d);
        }
        cur = nxt;
    }
    return len;
}

// One particle of the lists + density pass. `count` bounds the indices a
// range may name (the end of the particle's segment); `stride` is the list
// stride (the total particle count).
template <bool WCSPH, int TPB = LISTS_THREADS, int SLOTS = STAGE_SLOTS>
__device__ __forceinline__ void lists_density_one(
    int i, int stride, int count, const float4* __restrict__ pos_s, const int2* __restrict__ fR, int nb,
    const float4* __restrict__ bpos_s, const float* __restrict__ bphi_s, const int2* __restrict__ bR,
    const uint32_t* __restrict__ bnear, const GridDev& g, const DensityParams& P, int* __restrict__ nl,
    int* __restrict__ nc, int* __restrict__ bnl, int* __restrict__ bnc, float* __restrict__ density,
    float* __restrict__ pressure, float4* __restrict__ prho, DevState* st, int* stage_col) {
    float4 p4 = pos_s[i];
    V3 pi = v3(p4);
    int cx, cy, cz;
    grid_coords(p4, g, cx, cy, cz);
    int edge = 0;
    float d = 0.0f;
    // each thread reads back only the LDS column it wrote, so no barrier is needed
    const ListSink<TPB, SLOTS> fl{stage_col, nl + i, stride};
    int len = walk_cells<WCSPH, false>(pi, cx, cy, cz, fR, pos_s, nullptr, count, g, P, fl, d, edge);
    fl.flush(len);
    nc[i] = len;
    float bd = 0.0f;
    int blen = 0;
    if (nb > 0) {
        // an interior cell walks only the neighbour cells that hold boundary
        // particles (none of its 27 cells is outside the grid, so skipping
        // the others leaves the edge-read count unchanged)
        uint32_t cells = ALL_CELLS;
        if (interior_cell(cx, cy, cz, g.cps)) {
            long long b = box_index(cx, cy, cz, g);
            if (b >= 0) cells = bnear[b];
        }
        if (cells != 0) {
            const ListSink<TPB, SLOTS> bl{stage_col, bnl + i, stride};
            blen = walk_cells<WCSPH, true>(pi, cx, cy, cz, bR, bpos_s, bphi_s, nb, g, P, bl, bd, edge, cells);
            bl.flush(blen);
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
        d = P.w_eval * ((d * P.M) + bd);
        density[i] = d;
    }
    prho[i] = make_float4(p4.x, p4.y, p4.z, d);
    if (edge) atomicAdd(&st->edge_reads, edge);
}

template <bool WCSPH>
__global__ __launch_bounds__(LISTS_THREADS, 8) void k_lists_density(
    int n, const float4* __restrict__ pos_s, const int2* __restrict__ fR, int nb, const float4* __restrict__ bpos_s,
    const float* __restrict__ bphi_s, const int2* __restrict__ bR, const uint32_t* __restrict__ bnear, GridDev g,
    DensityParams P, int* __restrict__ nl, int* __restrict__ nc, int* __restrict__ bnl, int* __restrict__ bnc,
    float* __restrict__ density, float* __restrict__ pressure, float4* __restrict__ prho, DevState* st) {
    __shared__ int stage[STAGE_SLOTS * LISTS_THREADS];
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lists_density_one<WCSPH>(i, n, n, pos_s, fR, nb, bpos_s, bphi_s, bR, bnear, g, P, nl, nc, bnl, bnc, density,
                             pressure, prho, st, stage + threadIdx.x);
}

// ------------------------------------------- lists + density, split walk
// A scene too small to fill the GPU with one thread per particle leaves
// most SIMDs idle while each thread walks its 27 cells one after the other.
// There each particle's walk is split over SPLIT consecutive lanes: lane q
// walks cells [q * SPLIT_CELLS, ...) of CELL_NEIGH_OFFSET order with the
// same tests as walk_cells and keeps, in LDS, its first LIST_CAP hits (index
// and r^2), its hit count per cell and its edge reads per cell. The
// particle's first lane then concatenates the parts in cell order up to the
// cap, adds the density terms in that order, and counts the edge reads the
// single walk makes (cell 0's, and cell o + 1's while fewer than LIST_CAP
// hits precede cell o). Lists, densities and counters equal
// k_lists_density's; the boundary walk is the first lane's.
constexpr int SPLIT = 4;
constexpr int SPLIT_CELLS = 7;  // cells per part; the last part has 6
constexpr int SPLIT_THREADS = 64;
constexpr int SPLIT_MAX_N = 1 << 15;  // the split walk serves scenes up to this many particles

template <bool WCSPH>
__global__ __launch_bounds__(SPLIT_THREADS) void k_lists_density_split(
    int n, const float4* __restrict__ pos_s, const int2* __restrict__ fR, int nb, const float4* __restrict__ bpos_s,
    const float* __restrict__ bphi_s, const int2* __restrict__ bR, const uint32_t* __restrict__ bnear, GridDev g,
    DensityParams P, int* __restrict__ nl, int* __restrict__ nc, int* __restrict__ bnl, int* __restrict__ bnc,
    float* __restrict__ density, float* __restrict__ pressure, float4* __restrict__ prho, DevState* st) {
    __shared__ int hj[SPLIT_THREADS][LIST_CAP];
    __shared__ float hr[SPLIT_THREADS][LIST_CAP];
    __shared__ int hcnt[SPLIT_THREADS][SPLIT_CELLS];
    __shared__ int hedge[SPLIT_THREADS], hstored[SPLIT_THREADS];
    const int tid = threadIdx.x;
    const int t = xcd_block(blockIdx.x, gridDim.x) * SPLIT_THREADS + tid;
    const int i = t / SPLIT, q = t % SPLIT;
    const bool valid = i < n;
    float4 p4 = make_float4(0.f, 0.f, 0.f, 0.f);
    int cx = 0, cy = 0, cz = 0;
    if (valid) {
        p4 = pos_s[i];
        grid_coords(p4, g, cx, cy, cz);
        const V3 pi = v3(p4);
        const __amdgpu_buffer_rsrc_t rp = buffer_view(pos_s, (uint32_t)n * 16u);
        const int o0 = q * SPLIT_CELLS, o1 = o0 + SPLIT_CELLS < 27 ? o0 + SPLIT_CELLS : 27;
        int stored = 0, emask = 0;
        int dx, dy, dz, e = 0;
        neigh_offset(o0, dx, dy, dz);
        int2 cur = cell_range(fR, cx + dx, cy + dy, cz + dz, g, e);
        emask |= e;
        for (int o = o0; o < o1; ++o) {
            int2 nxt = make_int2(0, 0);
            if (o + 1 < o1) {  // fetch the next cell's range while this cell is processed
                e = 0;
                neigh_offset(o + 1, dx, dy, dz);
                nxt = cell_range(fR, cx + dx, cy + dy, cz + dz, g, e);
                emask |= e << (o + 1 - o0);
            }
            const int a = cur.x, b = cur.y < n ? cur.y : n;
            int cnt = 0;
            for (int j0 = a; j0 < b; j0 += SPLIT_BATCH) {
                float4 c4[SPLIT_BATCH];
#pragma unroll
                for (int u = 0; u < SPLIT_BATCH; ++u) c4[u] = load4(rp, (uint32_t)j0 * 16u, 16u * u);
#pragma unroll
                for (int u = 0; u < SPLIT_BATCH; ++u) {
                    V3 r = pi - v3(c4[u]);
                    float r2 = dot(r, r);
                    if (j0 + u < b && r2 < g.sqr_support) {
                        if (stored < LIST_CAP) {
                            hj[tid][stored] = j0 + u;
                            hr[tid][stored] = r2;
                            ++stored;
                        }
                        ++cnt;
                    }
                }
            }
            hcnt[tid][o - o0] = cnt;
            cur = nxt;
        }
        hedge[tid] = emask;
        hstored[tid] = stored;
    }
    __syncthreads();
    if (!valid || q != 0) return;
    // the single walk's list, density and edge reads, from the parts in order
    int len = 0, edge = 0, before = 0;  // before: hits in the cells before cell o
    float d = 0.0f;
    for (int p = 0; p < SPLIT && len < LIST_CAP; ++p) {
        const int lp = tid + p, m = hstored[lp];
        for (int k = 0; k < m && len < LIST_CAP; ++k) {
            nl[(size_t)len * n + i] = hj[lp][k];
            ++len;
            accumulate_density<WCSPH>(hr[lp][k], d, P, 1.0f, false);
        }
    }
    edge = hedge[tid] & 1;  // cell 0's range is always read
    for (int o = 0; o < 27 && before < LIST_CAP; ++o) {
        const int lp = tid + o / SPLIT_CELLS, c = o % SPLIT_CELLS;
        if (o + 1 < 27) {
            const int lq = tid + (o + 1) / SPLIT_CELLS;
            edge += (hedge[lq] >> ((o + 1) % SPLIT_CELLS)) & 1;
        }
        before += hcnt[lp][c];
    }
    nc[i] = len;
    const V3 pi = v3(p4);
    float bd = 0.0f;
    int blen = 0;
    if (nb > 0) {
        uint32_t cells = ALL_CELLS;  // as in lists_density_one
        if (interior_cell(cx, cy, cz, g.cps)) {
            long long bb = box_index(cx, cy, cz, g);
            if (bb >= 0) cells = bnear[bb];
        }
        if (cells != 0) {
            const ListSink<SPLIT_THREADS, 0> bl{nullptr, bnl + i, n};
            blen = walk_cells<WCSPH, true>(pi, cx, cy, cz, bR, bpos_s, bphi_s, nb, g, P, bl, bd, edge, cells);
        }
    }
    bnc[i] = blen;
    if (WCSPH) {
        d *= P.w_eval * P.M;
        d += P.w_eval * bd;
        density[i] = d;
        pressure[i] = cl_max(0.0f, P.k_gas * (pown7(d / P.rho0) - 1.0f));
    } else {
        d = P.w_eval * ((d * P.M) + bd);
        density[i] = d;
    }
    prho[i] = make_float4(p4.x, p4.y, p4.z, d);
    if (edge) atomicAdd(&st->edge_reads, edge);
}

void launch_lists_density(bool wcsph, int n, const float4* pos_s, const int2* fR, int nb, const float4* bpos_s,
                          const float* bphi_s, const int2* bR, const uint32_t* bnear, const GridDev& g,
                          const DensityParams& P, int* nl, int* nc, int* bnl, int* bnc, float* density,
                          float* pressure, float4* prho, DevState* st, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "lists_density", s);
    if (n <= SPLIT_MAX_N) {
        const int sblocks = (int)(((long long)n * SPLIT + SPLIT_THREADS - 1) / SPLIT_THREADS);
        if (wcsph)
            hipLaunchKernelGGL(k_lists_density_split<true>, dim3(sblocks), dim3(SPLIT_THREADS), 0, s, n, pos_s, fR,
                               nb, bpos_s, bphi_s, bR, bnear, g, P, nl, nc, bnl, bnc, density, pressure, prho, st);
        else
            hipLaunchKernelGGL(k_lists_density_split<false>, dim3(sblocks), dim3(SPLIT_THREADS), 0, s, n, pos_s, fR,
                               nb, bpos_s, bphi_s, bR, bnear, g, P, nl, nc, bnl, bnc, density, pressure, prho, st);
        PROF_END(prof, s);
        return;
    }
    const int blocks = (n + LISTS_THREADS - 1) / LISTS_THREADS;
    if (wcsph)
        hipLaunchKernelGGL(k_lists_density<true>, dim3(blocks), dim3(LISTS_THREADS), 0, s, n, pos_s, fR, nb, bpos_s, bphi_s, bR,
                           bnear, g, P, nl, nc, bnl, bnc, density, pressure, prho, st);
    else
        hipLaunchKernelGGL(k_lists_density<false>, dim3(blocks), dim3(LISTS_THREADS), 0, s, n, pos_s, fR, nb, bpos_s, bphi_s, bR,
                           bnear, g, P, nl, nc, bnl, bnc, density, pressure, prho, st);
    PROF_END(prof, s);
}

// ------------------------------------------------------------------ normals
// Neighbour records gathered together per thread in compute_normals (one
// 16-byte record per neighbour; 16 measured best at 4M, equal at 1M).
constexpr int NORMALS_BATCH = 16;

// compute_normals (normals.cl:23-70); prho holds (x, y, z, density).
template <bool F = false>
__device__ __forceinline__ void normals_one(int i, int n, const float4* __restrict__ prho,
                                            const int* __restrict__ nl, const int* __restrict__ nc, float H2,
                                            float w_grad, float M, float H, float4* __restrict__ normal) {
    V3 pi = v3(prho[i]);
    V3 acc = {0.0f, 0.0f, 0.0f};
    const int len = nc[i];
    const int* L = nl + i;
    const __amdgpu_buffer_rsrc_t rp = buffer_view(prho, (uint32_t)n * 16u);
    for (int k0 = 0; k0 < len; k0 += NORMALS_BATCH) {
        int jj[NORMALS_BATCH];
        float4 q[NORMALS_BATCH];
#pragma unroll
        for (int u = 0; u < NORMALS_BATCH; ++u) jj[u] = k0 + u < len ? L[(size_t)(k0 + u) * n] : i;
#pragma unroll
        for (int u = 0; u < NORMALS_BATCH; ++u) q[u] = load4(rp, (uint32_t)jj[u] * 16u, 0u);
#pragma unroll
        for (int u = 0; u < NORMALS_BATCH; ++u) {
            if (k0 + u < len) {
                V3 r = pi - v3(q[u]);
                float qq = sqr(H2 - dot(r, r));
                acc = acc + Math<F>::div(r * qq, q[u].w);
            }
        }
    }
    V3 o = w_grad * acc;
    o = o * M;
    o = o * H;
    normal[i] = f4(o, 0.0f);
}

template <bool F>
__global__ __launch_bounds__(256) void k_normals(int n, const float4* __restrict__ prho,
                                                 const int* __restrict__ nl, const int* __restrict__ nc, float H2,
                                                 float w_grad, float M, float H, float4* __restrict__ normal) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    normals_one<F>(i, n, prho, nl, nc, H2, w_grad, M, H, normal);
}

void launch_normals(int n, const float4* prho, const int* nl, const int* nc, float H2, float w_grad, float M,
                    float H, float4* normal, hipStream_t s, Profiler* prof, bool fast, Shape sh) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "normals", s);
    hipLaunchKernelGGL(fast ? k_normals<true> : k_normals<false>, dim3((n + sh.block - 1) / sh.block),
                       dim3(sh.block), 0, s, n, prho, nl, nc, H2, w_grad, M, H, normal);
    PROF_END(prof, s);
}

// ------------------------------------------------------------- step state
__global__ void k_step_begin(DevState* st) {
    int t = threadIdx.x;
    for (int k = t; k < MAXERR_SLOTS * ERR_LANES; k += blockDim.x) (&st->maxerr[0][0])[k] = 0u;
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
// At most STATS_BLOCKS workgroups stride over the particles; each sums its
// five counters in registers, across its waves in LDS, and adds them to acc
// with one atomic per counter. (One atomic per wave put 5 x n/64 atomics on
// acc's single cache line, and they execute one at a time: 0.94 ms at 1M.)
constexpr int STATS_BLOCKS = 256;

__global__ __launch_bounds__(256) void k_stats(int n, const int* __restrict__ nc, const int* __restrict__ bnc,
                                               const float4* __restrict__ pos, unsigned long long* __restrict__ acc) {
    __shared__ unsigned long long part[4][5];
    unsigned long long v[5] = {0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int a = nc[i], b = bnc[i];
        const float4 p = pos[i];
        v[0] += (unsigned long long)a;
        v[1] += (unsigned long long)b;
        v[2] += a >= LIST_CAP;
        v[3] += b >= LIST_CAP;
        v[4] += !(isfinite(p.x) && isfinite(p.y) && isfinite(p.z));
    }
#pragma unroll
    for (int c = 0; c < 5; ++c)
        for (int off = 32; off > 0; off >>= 1) v[c] += __shfl_xor(v[c], off);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < 5; ++c) part[threadIdx.x >> 6][c] = v[c];
    __syncthreads();
    if (threadIdx.x < 5) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w][threadIdx.x];
        atomicAdd(&acc[threadIdx.x], t);
    }
}

void launch_stats(int n, const int* nc, const int* bnc, const float4* pos, unsigned long long* acc, hipStream_t s) {
    launch_fill_u32(acc, 20, 0u, s);
    if (n <= 0) return;
    const int blocks = std::min((n + 255) / 256, STATS_BLOCKS);
    hipLaunchKernelGGL(k_stats, dim3(blocks), dim3(256), 0, s, n, nc, bnc, pos, acc);
}


// ------------------------------------------------------------ batched scenes
// Kernels over a packed batch of scenes (batch.h): each particle looks up its
// scene's grid and parameters and runs the same per-particle function as the
// single-scene kernel. A grid smaller than one thread per particle strides
// over the particles (persistent blocks).
#define BATCH_LOOP(n)                                                                         \
    for (int base_ = xcd_block(blockIdx.x, gridDim.x) * blockDim.x; base_ < (n);             \
         base_ += gridDim.x * blockDim.x)                                                     \
        for (int i = base_ + threadIdx.x; i < (n); i = (n))

__global__ __launch_bounds__(256) void k_batch_hash(int n, const int* __restrict__ sid,
                                                    const BatchScene* __restrict__ sc,
                                                    const float4* __restrict__ pos, uint32_t* keys,
                                                    const uint32_t* prev, int2* __restrict__ R, int clear,
                                                    DevState* st) {
    BATCH_LOOP(n) {
        const BatchScene& S = sc[sid[i]];
        hash_one(i, pos, S.g, keys, prev, clear ? R + S.roff : nullptr, st);
    }
}

__global__ __launch_bounds__(256) void k_batch_gather(int n, const int* __restrict__ sid,
                                                      const BatchScene* __restrict__ sc, const int* __restrict__ perm,
                                                      const uint32_t* __restrict__ keys_s,
                                                      const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                      const float4* __restrict__ vh, float4* __restrict__ pos_s,
                                                      float4* __restrict__ vel_s, float4* __restrict__ vh_s,
                                                      int2* __restrict__ R) {
    BATCH_LOOP(n) {
        const BatchScene& S = sc[sid[i]];
        gather_one(i, S.lo, S.hi, perm, keys_s, pos, vel, vh, pos_s, vel_s, vh_s, R + S.roff, S.g);
    }
}

template <int TPB, int SLOTS>
__global__ __launch_bounds__(TPB) void k_batch_lists_density(int n, const int* __restrict__ sid,
                                                             const BatchScene* __restrict__ sc,
                                                             const float4* __restrict__ pos_s,
                                                             const int2* __restrict__ R, int* __restrict__ nl,
                                                             int* __restrict__ nc, int* __restrict__ bnc,
                                                             float* __restrict__ density,
                                                             float* __restrict__ pressure, float4* __restrict__ prho,
                                                             DevState* st) {
    __shared__ int stage[SLOTS > 0 ? SLOTS * TPB : 1];
    BATCH_LOOP(n) {
        const BatchScene& S = sc[sid[i]];
        lists_density_one<true, TPB, SLOTS>(i, n, S.hi, pos_s, R + S.roff, 0, nullptr, nullptr, nullptr, nullptr,
                                            S.g, S.D, nl, nc, nullptr, bnc, density, pressure, prho, st,
                                            stage + threadIdx.x);
    }
}

__global__ __launch_bounds__(256) void k_batch_normals(int n, const int* __restrict__ sid,
                                                       const BatchScene* __restrict__ sc,
                                                       const float4* __restrict__ prho, const int* __restrict__ nl,
                                                       const int* __restrict__ nc, float4* __restrict__ normal) {
    BATCH_LOOP(n) {
        const BatchScene& S = sc[sid[i]];
        normals_one(i, n, prho, nl, nc, S.nH2, S.nw_grad, S.nM, S.nH, normal);
    }
}

void launch_batch_hash(int n, const int* sid, const BatchScene* sc, const float4* pos, uint32_t* keys,
                       const uint32_t* prev, int2* R, bool clear, DevState* st, const LaunchCfg& c, hipStream_t s,
                       Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "batch_hash", s);
    hipLaunchKernelGGL(k_batch_hash, dim3(batch_blocks(n, c)), dim3(c.block), 0, s, n, sid, sc, pos, keys, prev, R,
                       clear ? 1 : 0, st);
    PROF_END(prof, s);
}

void launch_batch_gather(int n, const int* sid, const BatchScene* sc, const int* perm, const uint32_t* keys_s,
                         const float4* pos, const float4* vel, const float4* vh, float4* pos_s, float4* vel_s,
                         float4* vh_s, int2* R, const LaunchCfg& c, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "batch_gather_ranges", s);
    hipLaunchKernelGGL(k_batch_gather, dim3(batch_blocks(n, c)), dim3(c.block), 0, s, n, sid, sc, perm, keys_s, pos,
                       vel, vh, pos_s, vel_s, vh_s, R);
    PROF_END(prof, s);
}

template <int TPB, int SLOTS>
static void lists_variant(int n, const int* sid, const BatchScene* sc, const float4* pos_s, const int2* R, int* nl,
                          int* nc, int* bnc, float* density, float* pressure, float4* prho, DevState* st,
                          const LaunchCfg& c, hipStream_t s) {
    hipLaunchKernelGGL((k_batch_lists_density<TPB, SLOTS>), dim3(batch_blocks(n, c)), dim3(TPB), 0, s, n, sid, sc,
                       pos_s, R, nl, nc, bnc, density, pressure, prho, st);
}

void launch_batch_lists_density(int n, const int* sid, const BatchScene* sc, const float4* pos_s, const int2* R,
                                int* nl, int* nc, int* bnc, float* density, float* pressure, float4* prho,
                                DevState* st, const LaunchCfg& c, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "batch_lists_density", s);
    const bool staged = c.stage > 0;
    if (c.block == 64)
        staged ? lists_variant<64, 24>(n, sid, sc, pos_s, R, nl, nc, bnc, density, pressure, prho, st, c, s)
               : lists_variant<64, 0>(n, sid, sc, pos_s, R, nl, nc, bnc, density, pressure, prho, st, c, s);
    else if (c.block == 128)
        staged ? lists_variant<128, 24>(n, sid, sc, pos_s, R, nl, nc, bnc, density, pressure, prho, st, c, s)
               : lists_variant<128, 0>(n, sid, sc, pos_s, R, nl, nc, bnc, density, pressure, prho, st, c, s);
    else if (c.block == 256)
        staged ? lists_variant<256, 24>(n, sid, sc, pos_s, R, nl, nc, bnc, density, pressure, prho, st, c, s)
               : lists_variant<256, 0>(n, sid, sc, pos_s, R, nl, nc, bnc, density, pressure, prho, st, c, s);
    else
        throw std::invalid_argument("batched lists: block must be 64, 128 or 256");
    PROF_END(prof, s);
}

void launch_batch_normals(int n, const int* sid, const BatchScene* sc, const float4* prho, const int* nl,
                          const int* nc, float4* normal, const LaunchCfg& c, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "batch_normals", s);
    hipLaunchKernelGGL(k_batch_normals, dim3(batch_blocks(n, c)), dim3(c.block), 0, s, n, sid, sc, prho, nl, nc,
                       normal);
    PROF_END(prof, s);
}

}  // namespace sph
