// PCISPH kernels on gfx950, restating kernels/pcisph.cl and the correction
// loop of PCISPHSimulation::_simulate_pcisph_step (pcisphsimluation.cpp:72-165).
//
// Launch structure of one step (after neighbour search and density):
//   forces_predict       compute_initial_forces (pcisph.cl:134-222) + the
//                        first predict_pos (pcisph.cl:302-344, F_p = 0)
//   per correction i:
//     update_pressure    pcisph.cl:349-418, plus the max of the density error
//                        (the reference's clmax, pcisphsimluation.cpp:471-476)
//                        reduced over each wave and merged with one atomic
//     pressure_force     pcisph.cl:420-504, then either the next predict_pos
//                        or, when this correction is the last one, the final
//                        predict_vel_n_pos (pcisph.cl:240-284)
// The exit test (pcisphsimluation.cpp:142-147) is evaluated on the device by
// every thread of pressure_force from the reduced maximum, so the loop runs
// without a host round trip. Every correction gets one launch per phase; those
// past the ones that always run use small grid-stride grids and return at once
// when the loop has already ended (the usual case with the reference's
// parameters). A loop cap above a few corrections (the relaxation step's
// 10000) is enqueued in chunks, with the host reading the exit flag between
// chunks.
#include "kernels.h"
#include "solver.h"
#include "sph_device.h"

namespace sph {

// acc = g + (pressure_force + other_force) / PARTICLE_MASS (pcisph.cl:260,321)
template <bool F>
__device__ __forceinline__ V3 pci_accel(V3 fp, V3 fo, const PciParams& P) {
    using M = Math<F>;
    return {P.gx + M::div(fp.x + fo.x, P.M), P.gy + M::div(fp.y + fo.y, P.M), P.gz + M::div(fp.z + fo.z, P.M)};
}

// predict_pos body (pcisph.cl:319-343)
template <bool F>
__device__ __forceinline__ V3 pci_predict(V3 x, V3 v, V3 acc, const PciParams& P) {
    v = v + clamp3(acc * P.dt, -P.max_vel, P.max_vel);
    x = x + v * P.dt;
    float vn = Math<F>::length(v);
    V3 cp = {cl_clamp(x.x, P.lo_x, P.hi_x), cl_clamp(x.y, P.lo_y, 10.0f), cl_clamp(x.z, P.lo_z, P.hi_z)};
    float d = Math<F>::length(cp - x);
    if (d > 0.0f && vn > 0.0f) x = cp;
    return x;
}

// predict_vel_n_pos body (pcisph.cl:258-283)
template <bool F>
__device__ __forceinline__ void pci_integrate(V3& x, V3& v, V3 acc, const PciParams& P) {
    using M = Math<F>;
    v = v + clamp3(acc * P.dt, -P.max_vel, P.max_vel);
    x = x + v * P.dt;
    float vn = M::length(v);
    V3 cp = {cl_clamp(x.x, P.lo_x, P.hi_x), cl_clamp(x.y, P.lo_y, 10.0f), cl_clamp(x.z, P.lo_z, P.hi_z)};
    V3 dv = cp - x;
    float d = M::length(dv);
    if (d > 0.0f && vn > 0.0f) {
        V3 nrm = M::normalize(V3{cl_sign(dv.x), cl_sign(dv.y), cl_sign(dv.z)});
        float rest = M::div(d, P.dt * vn);
        float k = (1.0f + rest) * dot(v, nrm);
        v = v - k * nrm;
        x = cp;
    }
}

// surface tension spline st_kernel (kernels/surface_tension.h:23-36)
__device__ __forceinline__ float st_kernel(float r, float h, float c) {
    if (r <= 0.5f * h) {
        float t1 = h - r;
        float t2 = cube(t1) * cube(r);
        return 2.0f * t2 - c;
    } else if (r <= h) {
        float t1 = h - r;
        return cube(t1) * cube(r);
    }
    return 0.0f;
}

// Neighbours whose records (position and density, velocity, normal) are
// gathered together per thread in forces_predict (6 measured best: 4 and 8
// were slower at 1M).
constexpr int FORCE_BATCH = 6;

template <bool F>
__global__ __launch_bounds__(256) void k_forces_predict(int n, const float4* __restrict__ prho,
                                                        const float4* __restrict__ vel_s,
                                                        const float4* __restrict__ normal,
                                                        const int* __restrict__ nl, const int* __restrict__ nc,
                                                        PciParams P, float4* __restrict__ f_other,
                                                        float4* __restrict__ pos_pred) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 pr = prho[i];
    V3 pi = v3(pr), vi = v3(vel_s[i]);
    float di = pr.w;
    V3 ni = v3(normal[i]);
    V3 fv = {0.0f, 0.0f, 0.0f}, fcurv = {0.0f, 0.0f, 0.0f}, fcoh = {0.0f, 0.0f, 0.0f};
    const int len = nc[i];
    const int* L = nl + i;
    const __amdgpu_buffer_rsrc_t rp = buffer_view(prho, (uint32_t)n * 16u);
    const __amdgpu_buffer_rsrc_t rv = buffer_view(vel_s, (uint32_t)n * 16u);
    const __amdgpu_buffer_rsrc_t rn = buffer_view(normal, (uint32_t)n * 16u);
    for (int k0 = 0; k0 < len; k0 += FORCE_BATCH) {
        int jj[FORCE_BATCH];
        float4 qp[FORCE_BATCH], qv[FORCE_BATCH], qn[FORCE_BATCH];
#pragma unroll
        for (int u = 0; u < FORCE_BATCH; ++u) jj[u] = k0 + u < len ? L[(size_t)(k0 + u) * n] : i;
#pragma unroll
        for (int u = 0; u < FORCE_BATCH; ++u) {
            const uint32_t off = (uint32_t)jj[u] * 16u;
            qp[u] = load4(rp, off, 0u);
            qv[u] = load3(rv, off, 0u);
            qn[u] = load3(rn, off, 0u);
        }
#pragma unroll
        for (int u = 0; u < FORCE_BATCH; ++u) {
            if (k0 + u < len) {
                V3 r = pi - v3(qp[u]);
                float l = Math<F>::length(r);
                float dj = qp[u].w;
                fv = fv + (v3(qv[u]) - vi) * Math<F>::div(P.H - l, dj);
                float corr = Math<F>::div(2.0f * P.rho0_d, di + dj);
                fcurv = fcurv + (ni - v3(qn[u])) * corr;
                fcoh = fcoh + (Math<F>::normalize(r) * st_kernel(l, P.H, P.st_term)) * corr;
            }
        }
    }
    fv = fv * ((P.k_visc * sqr(P.M)) * Math<F>::div(P.visc_lapl, di));
    fcoh = fcoh * (P.M * P.st_main);
    V3 ft = ((-P.sigma) * P.M) * (fcurv + fcoh);
    V3 fo = fv + ft;
    f_other[i] = f4(fo, 0.0f);
    // first predict_pos: the pressure force buffer is zero-filled at this
    // point (pcisphsimluation.cpp:126)
    V3 zero = {0.0f, 0.0f, 0.0f};
    pos_pred[i] = f4(pci_predict<F>(pi, vi, pci_accel<F>(zero, fo, P), P), 1.0f);
}

// update_pressure also packs (x, y, z, p/rho^2) of every particle for the
// pressure force that follows: p_j/SQR(rho_j) of pcisph.cl:482 and
// C = p_i/SQR(rho_i) of pcisph.cl:452 are this same value.
// One correction's update_pressure over the whole grid (grid-stride loop).
template <bool F>
__device__ __forceinline__ void update_pressure_phase(int n, int iter, const float4* __restrict__ pos_pred,
                                                      const float4* __restrict__ prho,
                                                      const int* __restrict__ nl, const int* __restrict__ nc,
                                                      const int* __restrict__ bnl, const int* __restrict__ bnc,
                                                      const float4* __restrict__ bpos_s,
                                                      const float* __restrict__ bphi_s, int has_b, const PciParams& P,
                                                      float* __restrict__ err, float* __restrict__ pressure,
                                                      float4* __restrict__ pc, DevState* st) {
    __shared__ uint32_t wmax[4];
    const int stride = gridDim.x * blockDim.x;
    float dvmax = 0.0f;
    for (int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        V3 xi = v3(pos_pred[i]);
        float pd = 0.0f, pdb = 0.0f;
        const int len = nc[i];
        const int rows = wave_max_small(len);
        const int dflt = wave_first(i);
        const __amdgpu_buffer_rsrc_t rpp = buffer_view(pos_pred, (uint32_t)n * 16u);
        int jj[GATHER_BATCH];
        if (rows > 0) list_batch(nl, n, i, len, rows, 0, dflt, jj);
        for (int k0 = 0; k0 < rows; k0 += GATHER_BATCH) {
            float4 q[GATHER_BATCH];
#pragma unroll
            for (int u = 0; u < GATHER_BATCH; ++u) q[u] = load3(rpp, (uint32_t)jj[u] * 16u, 0u);
            if (k0 + GATHER_BATCH < rows) list_batch(nl, n, i, len, rows, k0 + GATHER_BATCH, dflt, jj);
#pragma unroll
            for (int u = 0; u < GATHER_BATCH; ++u) {
                V3 r = xi - v3(q[u]);
                float t = pd + cube(P.H2 - dot(r, r));
                if (k0 + u < len) pd = t;
            }
        }
        pd *= P.w_eval * P.M;
        if (has_b) {
            const int blen = bnc[i];
            const int brows = wave_max_small(blen);
            for (int k0 = 0; k0 < brows; k0 += GATHER_BATCH) {
                int jj[GATHER_BATCH];
                float4 q[GATHER_BATCH];
                float w[GATHER_BATCH];
                list_batch(bnl, n, i, blen, brows, k0, 0, jj);  // past blen: boundary particle 0, unused
#pragma unroll
                for (int u = 0; u < GATHER_BATCH; ++u) {
                    q[u] = bpos_s[jj[u]];
                    w[u] = bphi_s[jj[u]];
                }
#pragma unroll
                for (int u = 0; u < GATHER_BATCH; ++u) {
                    V3 r = xi - v3(q[u]);
                    float t = pdb + cube(P.H2 - dot(r, r)) * w[u];
                    if (k0 + u < blen) pdb = t;
                }
            }
            pdb *= P.w_eval;
        }
        float dv = cl_max(0.0f, (pd + pdb) - P.rho0_d);
        dvmax = dv > dvmax ? dv : dvmax;
        err[i] = dv;
        float p0 = (iter == 1) ? 0.0f : pressure[i];  // zero-filled per step (pcisphsimluation.cpp:118)
        float p = p0 + dv * P.delta;
        pressure[i] = p;
        float4 pr = prho[i];
        pc[i] = make_float4(pr.x, pr.y, pr.z, Math<F>::div(p, sqr(pr.w)));
    }
    // dv >= +0, so its bits order like the values
    uint32_t bits = __float_as_uint(dvmax);
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t o = __shfl_xor(bits, off);
        bits = o > bits ? o : bits;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = bits;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = wmax[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = wmax[w] > m ? wmax[w] : m;
        atomicMax(&st->maxerr[iter & (MAXERR_SLOTS - 1)][blockIdx.x & (ERR_LANES - 1)], m);
    }
}

template <bool F>
__global__ __launch_bounds__(256) void k_update_pressure(int n, int iter, const float4* __restrict__ pos_pred,
                                                         const float4* __restrict__ prho,
                                                         const int* __restrict__ nl, const int* __restrict__ nc,
                                                         const int* __restrict__ bnl, const int* __restrict__ bnc,
                                                         const float4* __restrict__ bpos_s,
                                                         const float* __restrict__ bphi_s, int has_b, PciParams P,
                                                         float* __restrict__ err, float* __restrict__ pressure,
                                                         float4* __restrict__ pc, DevState* st) {
    int done = st->done_iter;
    if (done != 0 && done < iter) return;
    update_pressure_phase<F>(n, iter, pos_pred, prho, nl, nc, bnl, bnc, bpos_s, bphi_s, has_b, P, err, pressure, pc,
                             st);
}

// One correction's pressure force and prediction (or final integration)
// over the whole grid. Reads the maximum error that every block of the
// update_pressure phase has merged; block 0 records the end of the loop.
template <bool F>
__device__ __forceinline__ void pressure_force_phase(
    int n, int iter, const float4* __restrict__ pc, const float4* __restrict__ vel_s,
    const float4* __restrict__ f_other, const int* __restrict__ nl, const int* __restrict__ nc,
    const int* __restrict__ bnl, const int* __restrict__ bnc, const float4* __restrict__ bpos_s,
    const float* __restrict__ bphi_s, int has_b, const PciParams& P, float4* __restrict__ f_press,
    float4* __restrict__ pos_pred, float4* __restrict__ pos_out, float4* __restrict__ vel_out, DevState* st) {
    // the maximum over the merged words: one per lane, then the wave's
    // maximum (non-negative float bits order like the values)
    uint32_t mbits = __hip_atomic_load(&st->maxerr[iter & (MAXERR_SLOTS - 1)][threadIdx.x & (ERR_LANES - 1)],
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t o = __shfl_xor(mbits, off);
        mbits = o > mbits ? o : mbits;
    }
    float mx = __uint_as_float(mbits);
    bool last = iter >= P.max_iter || (iter >= P.min_iter && ((mx - P.rho0_host) / P.rho0_host < P.eta));
    if (blockIdx.x == 0 && threadIdx.x == 0 && last) {
        st->done_iter = iter;
        st->iterations = iter;
    }
    if (blockIdx.x == 0 && threadIdx.x < ERR_LANES) st->maxerr[(iter + 1) & (MAXERR_SLOTS - 1)][threadIdx.x] = 0u;
    const int stride = gridDim.x * blockDim.x;
    for (int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    float4 ci = pc[i];
    V3 xi = v3(ci);
    float C = ci.w;
    V3 f = {0.0f, 0.0f, 0.0f}, fb = {0.0f, 0.0f, 0.0f};
    const int len = nc[i];
    const int* L = nl + i;
    const __amdgpu_buffer_rsrc_t rpc = buffer_view(pc, (uint32_t)n * 16u);
    for (int k0 = 0; k0 < len; k0 += GATHER_BATCH) {
        int jj[GATHER_BATCH];
        float4 q[GATHER_BATCH];
#pragma unroll
        for (int u = 0; u < GATHER_BATCH; ++u) jj[u] = k0 + u < len ? L[(size_t)(k0 + u) * n] : i;
#pragma unroll
        for (int u = 0; u < GATHER_BATCH; ++u) q[u] = load4(rpc, (uint32_t)jj[u] * 16u, 0u);
#pragma unroll
        for (int u = 0; u < GATHER_BATCH; ++u) {
            if (k0 + u < len) {
                V3 r = xi - v3(q[u]);
                float l = Math<F>::length(r);
                if (l >= 1e-5f) {
                    V3 gw = (r * Math<F>::div(1.0f, l)) * sqr(P.H - l);
                    f = f + (C + q[u].w) * gw;
                }
            }
        }
    }
    f = f * ((-sqr(P.M)) * P.press_grad);
    if (has_b) {
        // boundary terms in list order; indices past blen read boundary
        // particle 0 and are not used
        const int blen = bnc[i];
        for (int k0 = 0; k0 < blen; k0 += BOUND_BATCH) {
            int jj[BOUND_BATCH];
            float4 q[BOUND_BATCH];
            float w[BOUND_BATCH];
#pragma unroll
            for (int u = 0; u < BOUND_BATCH; ++u) jj[u] = k0 + u < blen ? bnl[(size_t)(k0 + u) * n + i] : 0;
#pragma unroll
            for (int u = 0; u < BOUND_BATCH; ++u) {
                q[u] = bpos_s[jj[u]];
                w[u] = bphi_s[jj[u]];
            }
#pragma unroll
            for (int u = 0; u < BOUND_BATCH; ++u) {
                if (k0 + u < blen) {
                    V3 r = xi - v3(q[u]);
                    float l = Math<F>::length(r);
                    V3 gw = (r * Math<F>::div(1.0f, l)) * sqr(P.H - l);
                    fb = fb + (w[u] * (2.0f * C)) * gw;
                }
            }
        }
        fb = fb * ((-P.M) * P.press_grad);
    }
    V3 fp = f + fb;
    f_press[i] = f4(fp, 0.0f);
    V3 vi = v3(vel_s[i]);
    V3 acc = pci_accel<F>(fp, v3(f_other[i]), P);
    if (!last) {
        pos_pred[i] = f4(pci_predict<F>(xi, vi, acc, P), 1.0f);
    } else {
        V3 x = xi, v = vi;
        pci_integrate<F>(x, v, acc, P);
        pos_out[i] = f4(x, 1.0f);
        vel_out[i] = f4(v, 0.0f);
    }
}
}

template <bool F>
__global__ __launch_bounds__(256) void k_pressure_force(
    int n, int iter, const float4* __restrict__ pc, const float4* __restrict__ vel_s,
    const float4* __restrict__ f_other, const int* __restrict__ nl, const int* __restrict__ nc,
    const int* __restrict__ bnl, const int* __restrict__ bnc, const float4* __restrict__ bpos_s,
    const float* __restrict__ bphi_s, int has_b, PciParams P, float4* __restrict__ f_press,
    float4* __restrict__ pos_pred, float4* __restrict__ pos_out, float4* __restrict__ vel_out, DevState* st) {
    int done = st->done_iter;
    if (done != 0 && done < iter) return;
    pressure_force_phase<F>(n, iter, pc, vel_s, f_other, nl, nc, bnl, bnc, bpos_s, bphi_s, has_b, P, f_press, pos_pred,
                            pos_out, vel_out, st);
}

// Final integration with a zero pressure force, for a loop of zero corrections.
template <bool F>
__global__ __launch_bounds__(256) void k_integrate_only(int n, const float4* __restrict__ pos_s,
                                                        const float4* __restrict__ vel_s,
                                                        const float4* __restrict__ f_other, PciParams P,
                                                        float4* __restrict__ pos_out, float4* __restrict__ vel_out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 x = v3(pos_s[i]), v = v3(vel_s[i]);
    V3 zero = {0.0f, 0.0f, 0.0f};
    pci_integrate<F>(x, v, pci_accel<F>(zero, v3(f_other[i]), P), P);
    pos_out[i] = f4(x, 1.0f);
    vel_out[i] = f4(v, 0.0f);
}

// Corrections up to min_iter always run and get one thread per particle;
// later ones run in k_correction_tail.
static int correction_blocks(int n) { return (n + 255) / 256; }
// Blocks of a grid-stride correction launch: a fixed grid, or one thread per
// particle.
static int shape_blocks(int n, Shape sh) {
    const int full = (n + sh.block - 1) / sh.block;
    return sh.grid > 0 && sh.grid < full ? sh.grid : full;
}

// Grid of the launches for corrections past the ones that always run: they
// usually return at once (the exit test ended the loop), so a small
// grid-stride grid keeps an idle launch cheap.
constexpr int TAIL_BLOCKS = 256;

void launch_forces_predict(int n, const float4* prho, const float4* vel_s, const float4* normal, const int* nl,
                           const int* nc, const PciParams& P, float4* f_other, float4* pos_pred, hipStream_t s,
                           Profiler* prof, Shape sh) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "forces_predict", s);
    hipLaunchKernelGGL(P.fast ? k_forces_predict<true> : k_forces_predict<false>, dim3((n + sh.block - 1) / sh.block),
                       dim3(sh.block), 0, s, n, prho, vel_s, normal, nl, nc, P, f_other, pos_pred);
    PROF_END(prof, s);
}

void launch_update_pressure(int n, int iter, const float4* pos_pred, const float4* prho, const int* nl,
                            const int* nc, const int* bnl, const int* bnc, const float4* bpos_s, const float* bphi_s,
                            int has_b, const PciParams& P, float* err, float* pressure, float4* pc, DevState* st,
                            hipStream_t s, Profiler* prof, Shape sh) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "update_pressure", s);
    hipLaunchKernelGGL(P.fast ? k_update_pressure<true> : k_update_pressure<false>, dim3(shape_blocks(n, sh)),
                       dim3(sh.block), 0, s, n, iter, pos_pred, prho, nl, nc, bnl, bnc, bpos_s, bphi_s, has_b, P, err,
                       pressure, pc, st);
    PROF_END(prof, s);
}

void launch_pressure_force(int n, int iter, const float4* pc, const float4* vel_s, const float4* f_other,
                           const int* nl, const int* nc, const int* bnl, const int* bnc, const float4* bpos_s,
                           const float* bphi_s, int has_b, const PciParams& P, float4* f_press, float4* pos_pred,
                           float4* pos_out, float4* vel_out, DevState* st, hipStream_t s, Profiler* prof,
                           Shape sh) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "pressure_force", s);
    hipLaunchKernelGGL(P.fast ? k_pressure_force<true> : k_pressure_force<false>, dim3(shape_blocks(n, sh)),
                       dim3(sh.block), 0, s, n, iter, pc, vel_s, f_other, nl, nc, bnl, bnc, bpos_s, bphi_s, has_b, P,
                       f_press, pos_pred, pos_out, vel_out, st);
    PROF_END(prof, s);
}

void launch_corrections(int n, int first, int last, const float4* prho, const float4* vel_s, const float4* f_other,
                        const int* nl, const int* nc, const int* bnl, const int* bnc, const float4* bpos_s,
                        const float* bphi_s, int has_b, const PciParams& P, float* err, float* pressure,
                        float4* pc, float4* f_press, float4* pos_pred, float4* pos_out, float4* vel_out,
                        DevState* st, hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    const int blocks = correction_blocks(n) < TAIL_BLOCKS ? correction_blocks(n) : TAIL_BLOCKS;
    for (int iter = first; iter <= last; ++iter) {
        PROF_BEGIN(prof, "update_pressure_tail", s);
        hipLaunchKernelGGL(P.fast ? k_update_pressure<true> : k_update_pressure<false>, dim3(blocks), dim3(256), 0,
                           s, n, iter, pos_pred, prho, nl, nc, bnl, bnc, bpos_s, bphi_s, has_b, P, err, pressure, pc,
                           st);
        PROF_END(prof, s);
        PROF_BEGIN(prof, "pressure_force_tail", s);
        hipLaunchKernelGGL(P.fast ? k_pressure_force<true> : k_pressure_force<false>, dim3(blocks), dim3(256), 0, s,
                           n, iter, pc, vel_s, f_other, nl, nc, bnl, bnc, bpos_s, bphi_s, has_b, P, f_press, pos_pred,
                           pos_out, vel_out, st);
        PROF_END(prof, s);
    }
}

void launch_pci_integrate_only(int n, const float4* pos_s, const float4* vel_s, const float4* f_other,
                               const PciParams& P, float4* pos_out, float4* vel_out, hipStream_t s,
                               Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "integrate", s);
    hipLaunchKernelGGL(P.fast ? k_integrate_only<true> : k_integrate_only<false>, dim3((n + 255) / 256), dim3(256), 0,
                       s, n, pos_s, vel_s, f_other, P, pos_out, vel_out);
    PROF_END(prof, s);
}

}  // namespace sph
