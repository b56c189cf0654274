// Batched WCSPH: many independent scenes packed into one particle array and
// advanced by one set of launches per step.
//
// Scene s owns the contiguous segment [lo, hi) of every per-particle array.
// The neighbour search sorts the whole array by (scene, Morton key): the key
// passes of the radix sort run over all particles and one last pass orders by
// scene id, so each segment ends up sorted by its own keys exactly as a
// single-scene run sorts it. Each scene has its own grid (cell size, dense
// range box at offset roff) and its own step parameters; the kernels look
// them up by the particle's scene id, and evaluate the same per-particle
// functions as the single-scene kernels, so a scene advanced in a batch
// produces the bits it produces alone.
#pragma once

#include "kernels.h"
#include "sph_device.h"

namespace sph {

struct BatchScene {
    GridDev g;
    DensityParams D;
    WcParams P;
    float nH2, nw_grad, nM, nH;  // compute_normals constants
    long long roff;              // offset of the scene's range box
    int lo, hi;                  // particle segment
};

// Launch shape of one batched kernel: threads per block (64, 128 or 256),
// blocks (0: one thread per particle; otherwise a persistent grid-stride
// grid of that many blocks) and, for the lists kernel, the LDS-staged list
// slots (0 or 24).
struct LaunchCfg {
    int block = 256;
    int grid = 0;
    int stage = 24;
};

void launch_batch_hash(int n, const int* sid, const BatchScene* sc, const float4* pos, uint32_t* keys,
                       const uint32_t* prev, int2* R, bool clear, DevState* st, const LaunchCfg& c, hipStream_t s,
                       Profiler* prof);
// The batch's keys sorted by (scene, key) with `sorted` (the last step's
// sorted keys; every scene keeps its segment) as the hint, placing only the
// keys that changed cell; the batched form of launch_hash_resort.
void launch_batch_hash_resort(int n, const int* sid, const BatchScene* sc, const float4* pos, uint32_t* keys,
                              uint32_t* sorted, int* perm, int2* R, DevState* st, const ResortWork& w, hipStream_t s,
                              Profiler* prof);
void launch_batch_gather(int n, const int* sid, const BatchScene* sc, const int* perm, const uint32_t* keys_s,
                         const float4* pos, const float4* vel, const float4* vh, float4* pos_s, float4* vel_s,
                         float4* vh_s, int2* R, const LaunchCfg& c, hipStream_t s, Profiler* prof);
void launch_batch_lists_density(int n, const int* sid, const BatchScene* sc, const float4* pos_s, const int2* R,
                                int* nl, int* nc, int* bnc, float* density, float* pressure, float4* prho,
                                DevState* st, const LaunchCfg& c, hipStream_t s, Profiler* prof);
void launch_batch_normals(int n, const int* sid, const BatchScene* sc, const float4* prho, const int* nl,
                          const int* nc, float4* normal, const LaunchCfg& c, hipStream_t s, Profiler* prof);
void launch_batch_accel_integrate(int n, const int* sid, const BatchScene* sc, const float4* prho,
                                  const float4* vel_s, const float4* vh_s, const float* pressure,
                                  const float4* normal, const int* nl, const int* nc, const int* bnc, float4* accel,
                                  float4* pos_out, float4* vel_out, float4* vh_out, const LaunchCfg& c,
                                  hipStream_t s, Profiler* prof);

// blocks of a launch of n particles under c
inline int batch_blocks(int n, const LaunchCfg& c) {
    int full = (n + c.block - 1) / c.block;
    return c.grid > 0 && c.grid < full ? c.grid : full;
}

}  // namespace sph
