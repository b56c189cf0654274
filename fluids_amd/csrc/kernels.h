// Host-side launchers of the gfx950 SPH kernels and their parameter blocks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstddef>

#include <string>
#include <vector>

namespace sph {

struct GridDev;
struct DevState;

// Per-kernel HIP-event timing on the launching stream. Events are recorded
// around each launch while enabled and resolved lazily.
class Profiler {
public:
    bool enabled = false;
    std::string filter;  // when set, only launches of this kernel are timed
    void begin(const char* name, hipStream_t s);
    void end(hipStream_t s);
    void resolve();  // waits for pending events and accumulates their times
    void reset();
    struct Entry {
        std::string name;
        long long launches = 0;
        double total_ms = 0.0;
    };
    std::vector<Entry> entries;
    ~Profiler();

private:
    struct Pending {
        int entry;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    int current_ = -1;
    bool active_ = false;
    hipEvent_t cur_a_ = nullptr;
    hipEvent_t get_event();
};

#define PROF_BEGIN(p, name, s) \
    do {                        \
        if ((p) && (p)->enabled) (p)->begin(name, s); \
    } while (0)
#define PROF_END(p, s) \
    do {               \
        if ((p) && (p)->enabled) (p)->end(s); \
    } while (0)

// Density pass parameters: PCISPH uses H2 (SUPPORT_RADIUS literal squared),
// M = PARTICLE_MASS literal, w_eval = MullerConstants::default_eval;
// WCSPH uses h (full support radius argument), M = particle mass argument,
// w_eval = poly6_eval, rho0 and k_gas for the Tait pressure.
struct DensityParams {
    float h, H2, M, w_eval, rho0, k_gas;
    int fix_density;
};

// PCISPH constants as the reference passes them (literals for -D values,
// full fp32 for kernel arguments).
struct PciParams {
    float H, H2, M, rho0_d;          // SUPPORT_RADIUS, its square, PARTICLE_MASS, REST_DENSITY literals
    float w_eval, w_grad, visc_lapl, press_grad, st_main, st_term;
    float k_visc, sigma, delta;
    float dt, max_vel, gx, gy, gz;
    float lo_x, hi_x, lo_y, lo_z, hi_z;  // container clamp: (-c)/2 and c/2; y is clamped to [lo_y, 10]
    float rho0_host, eta;                // exit test operands (host floats)
    int min_iter, max_iter;
    int fast;                            // 1: Math<true> builds of the kernels (SPH_FAST_MATH)
};

struct WcParams {
    float h, M, spiky_grad, visc_lapl, k_visc, sigma, rho0, st_main, st_term;
    float dt, max_vel, gx, gy, gz;
    float lo_x, hi_x, lo_y, lo_z, hi_z;
    int fast;  // 1: Math<true> build of the single-scene kernels (SPH_FAST_MATH); batches use the IEEE build
};

// Launch shape of a single-scene kernel: threads per block (64, 128 or 256)
// and, for the grid-stride kernels, blocks (0: one thread per particle).
// Shapes change timing only, never results.
struct Shape {
    int block = 256;
    int grid = 0;
};

// grid.hip
int sort_passes(int key_bits);
// passes of a sort whose last pass orders by scene id (sid_bits > 0)
int sort_passes(int key_bits, int sid_bits);
size_t sort_hist_words(int n);
void launch_hash(int n, const float4* pos, const GridDev& g, uint32_t* keys, const uint32_t* prev, int2* R,
                 DevState* st, hipStream_t s, Profiler* prof);

// Work arrays of the sort of a nearly sorted key array (grid.hip). One
// device block of resort_bytes(n) bytes holds them all: resort_carve lays
// them out, resort_clear zeroes the counters (H, T, TF); every sort leaves
// them so.
struct ResortWork {
    unsigned long long* moved;  // mask_words: bit i set when key i differs from the hint
    int* H;                     // padded: moved elements per insertion point
    int2* grp;                  // padded: (base rank, size) of each insertion point's group
    int* T;                     // tiles x 32 (one counter per 128-byte line): per-tile sums of H - moved
    int* TF;                    // tiles x 32: per-tile moved counts
    int* Tx;                    // tiles: exclusive scan of the sums of H - moved
    int* TFx;                   // tiles: exclusive scan of the moved counts
    int4* rec;                  // elems: per moved element (insertion point, slot, size of its wave's
                                //        subgroup if that took the group's first slots else 0, key in the hint)
    int* mlist;                 // elems: the moved elements in index order
    uint32_t* tk;               // elems: groups gathered at base + slot (keys)
    int* ti;                    // elems: (indices)
    int* big;                   // big: insertion points of the groups sorted by a workgroup
    int* nbig;                  // 1
    int* nmoved;                // 1
    uint32_t* S;                // s1: every 64th key of the last sort's output
    uint32_t* S2;               // s2: every 4096th
};
struct ResortSizes {
    size_t tiles = 0, padded = 0, mask_words = 0, elems = 0, big = 0, s1 = 0, s2 = 0;
};
ResortSizes resort_sizes(int n);
size_t resort_bytes(int n);
ResortWork resort_carve(void* block, int n);
void resort_clear(const ResortWork& w, int n, hipStream_t s);
// Writes the samples (w.S, w.S2) of a sorted key array produced by another
// sort, so that it can serve as the next hint.
void launch_resort_samples(const uint32_t* sorted, int n, const ResortWork& w, hipStream_t s);
// The step's keys (hashed from pos into keys, with k_hash's counters) sorted
// with `sorted` as the hint: on entry the previous sorted keys (any
// non-decreasing array of n keys whose samples w holds), on exit the sorted
// keys (and their samples); perm receives the stable order. R (may be null):
// the range box of the hint's cells, used to place moved keys and left with
// the ranges of the cells that lost all their particles cleared, as
// launch_hash clears them. Equal to launch_hash + launch_radix_sort.
void launch_hash_resort(int n, const float4* pos, const GridDev& g, uint32_t* keys, uint32_t* sorted, int* perm,
                        int2* R, DevState* st, const ResortWork& w, hipStream_t s, Profiler* prof);
// The same sort of given keys.
void launch_resort(int n, const uint32_t* keys, uint32_t* sorted, int* perm, const ResortWork& w, hipStream_t s);
// Stable sort of n (key, index) pairs by the low key_bits of the keys; with
// sid (scene id of each index, < 2^sid_bits) a final pass orders the result
// by scene, so each scene's contiguous segment is sorted by its own keys.
// vals_in (optional) holds the values to carry; without it they are the
// indices 0 .. n-1.
void launch_radix_sort(uint32_t* keys_a, uint32_t* keys_b, int* vals_a, int* vals_b, uint32_t* hist, int n,
                       int key_bits, hipStream_t s, Profiler* prof, const int* sid = nullptr, int sid_bits = 0,
                       const int* vals_in = nullptr);
void launch_gather_fluid(int n, const int* perm, const uint32_t* keys_s, const float4* pos, const float4* vel,
                         const float4* vh, float4* pos_s, float4* vel_s, float4* vh_s, int2* R, const GridDev& g,
                         hipStream_t s, Profiler* prof);
// launch_gather_fluid that also writes perm_out = perm and
// ids_out[i] = ids_in[perm[i]] (the adjoint's tape and particle ids).
void launch_gather_fluid_tape(int n, const int* perm, const uint32_t* keys_s, const float4* pos, const float4* vel,
                              const float4* vh, float4* pos_s, float4* vel_s, float4* vh_s, int2* R, const GridDev& g,
                              int* perm_out, const int* ids_in, int* ids_out, hipStream_t s);
void launch_gather_boundary(int n, const int* perm, const uint32_t* keys_s, const float4* pos, const float* phi,
                            float4* pos_s, float* phi_s, int2* R, const GridDev& g, hipStream_t s, Profiler* prof);
void launch_mark_boundary(int n, const float4* bpos_s, const uint32_t* keys_s, const GridDev& g, uint32_t* near,
                          size_t near_words, hipStream_t s, Profiler* prof);
void launch_lists_density(bool wcsph, int n, const float4* pos_s, const int2* fR, int nb, const float4* bpos_s,
                          const float* bphi_s, const int2* bR, const uint32_t* bnear, const GridDev& g,
                          const DensityParams& P, int* nl, int* nc, int* bnl, int* bnc, float* density,
                          float* pressure, float4* prho, DevState* st, hipStream_t s, Profiler* prof);
// fast: the Math<true> build (SPH_FAST_MATH)
void launch_normals(int n, const float4* prho, const int* nl, const int* nc, float H2, float w_grad, float M,
                    float H, float4* normal, hipStream_t s, Profiler* prof, bool fast = false, Shape sh = Shape{});
void launch_step_begin(DevState* st, hipStream_t s);
void launch_fill_u32(void* p, size_t n, uint32_t v, hipStream_t s);
void launch_stats(int n, const int* nc, const int* bnc, const float4* pos, unsigned long long* acc, hipStream_t s);

// sort.hip: the body of sph_sort_pairs (throws sph::Error)
void sort_pairs(uint32_t* keys, int32_t* values, int n, int key_bits, void* stream);
// the body of sph_sort_pairs_noalloc (work arrays in the library's scratch)
void sort_pairs_noalloc(uint32_t* keys, int32_t* values, int n, int key_bits, void* stream);
// the body of sph_sort_near
void sort_near(const uint32_t* hint, const uint32_t* keys, uint32_t* sorted_keys, int32_t* perm, int n,
               void* stream);

// pcisph.hip
void launch_forces_predict(int n, const float4* prho, const float4* vel_s, const float4* normal, const int* nl,
                           const int* nc, const PciParams& P, float4* f_other, float4* pos_pred, hipStream_t s,
                           Profiler* prof, Shape sh = Shape{});
void launch_update_pressure(int n, int iter, const float4* pos_pred, const float4* prho, const int* nl,
                            const int* nc, const int* bnl, const int* bnc, const float4* bpos_s, const float* bphi_s,
                            int has_b, const PciParams& P, float* err, float* pressure, float4* pc, DevState* st,
                            hipStream_t s, Profiler* prof, Shape sh = Shape{});
void launch_pressure_force(int n, int iter, const float4* pc, const float4* vel_s, const float4* f_other,
                           const int* nl, const int* nc, const int* bnl, const int* bnc, const float4* bpos_s,
                           const float* bphi_s, int has_b, const PciParams& P, float4* f_press, float4* pos_pred,
                           float4* pos_out, float4* vel_out, DevState* st, hipStream_t s, Profiler* prof,
                           Shape sh = Shape{});
// Corrections first .. last, one launch per phase on a grid-stride grid of at
// most 256 workgroups; each returns at once when the loop already ended.
void launch_corrections(int n, int first, int last, const float4* prho, const float4* vel_s, const float4* f_other,
                        const int* nl, const int* nc, const int* bnl, const int* bnc, const float4* bpos_s,
                        const float* bphi_s, int has_b, const PciParams& P, float* err, float* pressure,
                        float4* pc, float4* f_press, float4* pos_pred, float4* pos_out, float4* vel_out,
                        DevState* st, hipStream_t s, Profiler* prof);
void launch_pci_integrate_only(int n, const float4* pos_s, const float4* vel_s, const float4* f_other,
                               const PciParams& P, float4* pos_out, float4* vel_out, hipStream_t s,
                               Profiler* prof);

// wcsph.hip
void launch_accel_integrate(int n, const float4* prho, const float4* vel_s, const float4* vh_s,
                            const float* pressure, const float4* normal, const int* nl,
                            const int* nc, const int* bnl, const int* bnc, const float4* bpos_s,
                            const float* bphi_s, const WcParams& P, float4* accel, float4* pos_out,
                            float4* vel_out, float4* vh_out, hipStream_t s, Profiler* prof);

// boundary.hip
void launch_transform(int n, const float4* raw, float4* out, const float R[9], const float com[3], hipStream_t s,
                      Profiler* prof);
void launch_phi(int n, const float4* p, float H_d, float w_eval, float rho0, float phi_coeff, float* phi,
                hipStream_t s, Profiler* prof);
int fluid_force_blocks(int count);
void launch_fluid_force(int count, int origin, const float4* b_unsorted, const float* b_phi, int n,
                        const float4* prho, const float* pressure, const int2* fR,
                        const GridDev& g, float H, float spiky_grad, float mass, const float com[3],
                        float4* b_force, float4* b_torque, float* partial, DevState* st, hipStream_t s,
                        Profiler* prof);
void launch_wrench_reduce(int nbodies, const int* offsets, const int* nblocks, const float* partial, float* wrench,
                          hipStream_t s, Profiler* prof);

}  // namespace sph
