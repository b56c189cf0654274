// Boundary handling and two-way rigid coupling on gfx950, restating
// kernels/boundary.cl and BoundaryHandler (fluid/simulation/boundaryhandler.cpp).
#include "kernels.h"
#include "sph_device.h"

namespace sph {

// transform_particles (boundary.cl:33-59), position part. The particle
// velocities it also writes are never read on the path.
__global__ __launch_bounds__(256) void k_transform(int n, const float4* __restrict__ raw, float4* __restrict__ out,
                                                   float r0, float r1, float r2, float r3, float r4, float r5,
                                                   float r6, float r7, float r8, float cx, float cy, float cz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 p = v3(raw[i]);
    V3 m0 = {r0, r1, r2}, m1 = {r3, r4, r5}, m2 = {r6, r7, r8};
    out[i] = make_float4(dot(m0, p) + cx, dot(m1, p) + cy, dot(m2, p) + cz, 1.0f);
}

void launch_transform(int n, const float4* raw, float4* out, const float R[9], const float com[3], hipStream_t s,
                      Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "transform", s);
    hipLaunchKernelGGL(k_transform, dim3((n + 255) / 256), dim3(256), 0, s, n, raw, out, R[0], R[1], R[2], R[3], R[4],
                       R[5], R[6], R[7], R[8], com[0], com[1], com[2]);
    PROF_END(prof, s);
}

// compute_boundary_phi (boundary.cl:4-31): psi = (rho0 / (sum W * w_eval)) / coeff
// over all samples of one body, brute force in sample order like the
// reference. The samples are streamed through LDS in tiles.
constexpr int PHI_TILE = 256;
__global__ __launch_bounds__(PHI_TILE) void k_phi(int n, const float4* __restrict__ p, float H2, float w_eval,
                                                  float rho0, float phi_coeff, float* __restrict__ phi) {
    __shared__ float4 tile[PHI_TILE];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    V3 pi = i < n ? v3(p[i]) : V3{0.0f, 0.0f, 0.0f};
    float delta = 0.0f;
    for (int base = 0; base < n; base += PHI_TILE) {
        int j = base + threadIdx.x;
        tile[threadIdx.x] = j < n ? p[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        __syncthreads();
        int m = n - base < PHI_TILE ? n - base : PHI_TILE;
        for (int k = 0; k < m; ++k) {
            V3 r = pi - v3(tile[k]);
            float r2 = dot(r, r);
            if (r2 < H2) delta += cube(H2 - r2);
        }
        __syncthreads();
    }
    if (i < n) phi[i] = (rho0 / (delta * w_eval)) / phi_coeff;
}

void launch_phi(int n, const float4* p, float H_d, float w_eval, float rho0, float phi_coeff, float* phi,
                hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "boundary_phi", s);
    float H2 = H_d * H_d;
    hipLaunchKernelGGL(k_phi, dim3((n + PHI_TILE - 1) / PHI_TILE), dim3(PHI_TILE), 0, s, n, p, H2, w_eval, rho0,
                       phi_coeff, phi);
    PROF_END(prof, s);
}

// compute_fluid_force (boundary.cl:61-118): for each sample of a body with
// mass > 0, the pressure force of the fluid found by walking the fluid cells
// (cap 50), and its torque about the body's centre. Each workgroup then sums
// its force and torque in a fixed order (wave butterfly shuffles, then the
// wave sums in wave order) into one partial (6 floats);
// k_wrench_reduce adds the partials of a body in index order. Both orders are
// fixed, so the wrench is reproducible run to run (the reference's clsum
// order is that of Boost.Compute's reduce and is not reproduced).
constexpr int FF_THREADS = 256;
int fluid_force_blocks(int count) { return (count + FF_THREADS - 1) / FF_THREADS; }

__global__ __launch_bounds__(FF_THREADS) void k_fluid_force(int count, int origin,
                                                            const float4* __restrict__ b_unsorted,
                                                            const float* __restrict__ b_phi, int n,
                                                            const float4* __restrict__ prho,
                                                            const float* __restrict__ pressure,
                                                            const int2* __restrict__ fR, GridDev g,
                                                            float H, float spiky_grad, float mass, float cx, float cy,
                                                            float cz, float4* __restrict__ b_force,
                                                            float4* __restrict__ b_torque, float* __restrict__ partial,
                                                            DevState* st) {
    __shared__ float red[6][FF_THREADS / 64];
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    V3 f = {0.0f, 0.0f, 0.0f}, tq = {0.0f, 0.0f, 0.0f};
    if (q < count) {
        int bi = origin + q;
        float4 pb = b_unsorted[bi];
        V3 pi = v3(pb);
        int cxi, cyi, czi;
        grid_coords(pb, g, cxi, cyi, czi);
        int edge = 0, cnt = 0;
        for (int o = 0; cnt < LIST_CAP && o < 27; ++o) {
            int dx, dy, dz;
            neigh_offset(o, dx, dy, dz);
            int2 iv = cell_range(fR, cxi + dx, cyi + dy, czi + dz, g, edge);
            int end = iv.y < n ? iv.y : n;  // ranges never exceed n
            for (int j = iv.x; cnt < LIST_CAP && j < end; ++j) {
                float4 q = prho[j];
                V3 r = v3(q) - pi;
                float rn = length(r);
                if (rn < H) {
                    float C = pressure[j] / sqr(q.w);
                    ++cnt;
                    f = f + (((r * (spiky_grad / rn)) * sqr(H - rn)) * 2.0f) * C;
                }
            }
        }
        f = f * (mass * b_phi[bi]);
        V3 a = pi - V3{cx, cy, cz};
        tq = {a.y * f.z - a.z * f.y, a.z * f.x - a.x * f.z, a.x * f.y - a.y * f.x};
        b_force[bi] = f4(f, 0.0f);
        b_torque[bi] = f4(tq, 0.0f);
        if (edge) atomicAdd(&st->edge_reads, edge);
    }
    // workgroup sums: butterfly shuffles within each wave, then the four wave
    // sums added in wave order by one thread
    float v[6] = {f.x, f.y, f.z, tq.x, tq.y, tq.z};
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] = wave_sum(v[c]);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (lane == 0)
        for (int c = 0; c < 6; ++c) red[c][wave] = v[c];
    __syncthreads();
    if (t < 6) {
        float s = red[t][0];
        for (int w = 1; w < FF_THREADS / 64; ++w) s += red[t][w];
        partial[(size_t)blockIdx.x * 8 + t] = s;
    }
}

void launch_fluid_force(int count, int origin, const float4* b_unsorted, const float* b_phi, int n,
                        const float4* prho, const float* pressure, const int2* fR, const GridDev& g, float H, float spiky_grad, float mass, const float com[3],
                        float4* b_force, float4* b_torque, float* partial, DevState* st, hipStream_t s,
                        Profiler* prof) {
    if (count <= 0) return;
    PROF_BEGIN(prof, "fluid_force", s);
    hipLaunchKernelGGL(k_fluid_force, dim3(fluid_force_blocks(count)), dim3(FF_THREADS), 0, s, count, origin,
                       b_unsorted, b_phi, n, prho, pressure, fR, g, H, spiky_grad, mass, com[0],
                       com[1], com[2], b_force, b_torque, partial, st);
    PROF_END(prof, s);
}

__global__ void k_wrench_reduce(int nbodies, const int* __restrict__ offsets, const int* __restrict__ nblocks,
                                const float* __restrict__ partial, float* __restrict__ wrench) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbodies) return;
    float acc[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < nblocks[b]; ++k)
        for (int c = 0; c < 6; ++c) acc[c] += partial[(size_t)(offsets[b] + k) * 8 + c];
    for (int c = 0; c < 6; ++c) wrench[b * 8 + c] = acc[c];
}

void launch_wrench_reduce(int nbodies, const int* offsets, const int* nblocks, const float* partial, float* wrench,
                          hipStream_t s, Profiler* prof) {
    if (nbodies <= 0) return;
    PROF_BEGIN(prof, "wrench_reduce", s);
    hipLaunchKernelGGL(k_wrench_reduce, dim3((nbodies + 63) / 64), dim3(64), 0, s, nbodies, offsets, nblocks, partial,
                       wrench);
    PROF_END(prof, s);
}

}  // namespace sph
