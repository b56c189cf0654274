// WCSPH kernels on gfx950, restating kernels/wcsph.cl. compute_acceleration
// (wcsph.cl:81-173) and predict_vel_n_pos (wcsph.cl:176-224) are fused: the
// integrator reads only the particle's own acceleration, so one pass computes
// the acceleration, stores it and advances the particle. Density, Tait
// pressure and the lists come from k_lists_density<true> (grid.hip), the
// normals from k_normals.
#include "batch.h"
#include "kernels.h"
#include "sph_device.h"

namespace sph {

__device__ __forceinline__ float st_kernel_wc(float r, float h, float c) {
    // st_kernel (kernels/surface_tension.h:23-36)
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

// Neighbours whose four records are gathered together per thread in the
// acceleration kernel (6 measured best for the batch of scenes: 4 and 8
// were slower).
constexpr int WC_BATCH = 6;

// One particle of compute_acceleration + predict_vel_n_pos; n is the list
// stride (the total particle count). F: the Math<true> build (SPH_FAST_MATH).
template <bool F = false>
__device__ __forceinline__ void accel_integrate_one(
    int i, int n, const float4* __restrict__ prho, const float4* __restrict__ vel_s, const float4* __restrict__ vh_s,
    const float* __restrict__ pressure, const float4* __restrict__ normal, const int* __restrict__ nl,
    const int* __restrict__ nc, const int* __restrict__ bnl, const int* __restrict__ bnc,
    const float4* __restrict__ bpos_s, const float* __restrict__ bphi_s, const WcParams& P,
    float4* __restrict__ accel, float4* __restrict__ pos_out, float4* __restrict__ vel_out,
    float4* __restrict__ vh_out) {
    float4 p4 = prho[i];
    V3 pi = v3(p4), vi = v3(vel_s[i]);
    float di = p4.w, pri = pressure[i];
    V3 ni = v3(normal[i]);
    using M = Math<F>;
    float C = M::div(pri, sqr(di));
    const float h = P.h;
    V3 bacc = {0.0f, 0.0f, 0.0f};
    // boundary terms in list order; indices past blen read boundary particle
    // 0 and are not used
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
                V3 r = pi - v3(q[u]);
                float rn = M::length(r);
                V3 t = (-w[u]) * r;
                t = t * M::div(P.spiky_grad, rn);
                t = t * sqr(h - rn);
                t = t * 2.0f;
                t = t * C;
                bacc = bacc + t;
            }
        }
    }
    V3 pacc = {0.0f, 0.0f, 0.0f}, vacc = {0.0f, 0.0f, 0.0f}, fcurv = {0.0f, 0.0f, 0.0f}, fcoh = {0.0f, 0.0f, 0.0f};
    const int len = nc[i];
    const int* L = nl + i;
    const __amdgpu_buffer_rsrc_t rp = buffer_view(prho, (uint32_t)n * 16u);
    const __amdgpu_buffer_rsrc_t rv = buffer_view(vel_s, (uint32_t)n * 16u);
    const __amdgpu_buffer_rsrc_t rn = buffer_view(normal, (uint32_t)n * 16u);
    const __amdgpu_buffer_rsrc_t rq = buffer_view(pressure, (uint32_t)n * 4u);
    for (int k0 = 0; k0 < len; k0 += WC_BATCH) {
        int jj[WC_BATCH];
        float4 qp[WC_BATCH], qv[WC_BATCH], qn[WC_BATCH];
        float qpr[WC_BATCH];
#pragma unroll
        for (int u = 0; u < WC_BATCH; ++u) jj[u] = k0 + u < len ? L[(size_t)(k0 + u) * n] : i;
#pragma unroll
        for (int u = 0; u < WC_BATCH; ++u) {
            const uint32_t j = (uint32_t)jj[u];
            qp[u] = load4(rp, j * 16u, 0u);
            qv[u] = load3(rv, j * 16u, 0u);
            qn[u] = load3(rn, j * 16u, 0u);
            qpr[u] = load1(rq, j * 4u, 0u);
        }
#pragma unroll
        for (int u = 0; u < WC_BATCH; ++u) {
            if (k0 + u < len) {
                float dj = qp[u].w, prj = qpr[u];
                V3 r = pi - v3(qp[u]);
                float rn = M::length(r);
                // wcsph.cl:152 compares against the double literal 0.00001
                if (di > 0.0f && dj > 0.0f && (double)rn > 0.00001) {
                    pacc = pacc + ((r * M::div(P.spiky_grad, rn)) * sqr(h - rn)) * (C + M::div(prj, sqr(dj)));
                    vacc = vacc + M::div((v3(qv[u]) - vi) * (h - rn), dj);
                    float corr = M::div(2.0f * P.rho0, di + dj);
                    fcurv = fcurv + (ni - v3(qn[u])) * corr;
                    fcoh = fcoh + (M::normalize(r) * st_kernel_wc(rn, h, P.st_term)) * corr;
                }
            }
        }
    }
    pacc = pacc * (-P.M);
    vacc = vacc * (P.visc_lapl * P.k_visc);
    fcoh = fcoh * P.st_main;
    V3 ast = (-P.sigma) * (fcurv + fcoh);
    V3 g = {P.gx, P.gy, P.gz};
    V3 a = (((g + ast) + pacc) + vacc) + bacc;
    accel[i] = f4(a, 0.0f);

    // predict_vel_n_pos (wcsph.cl:194-223), the leap-frog with the reference's
    // double use of dt*acc (Q12)
    V3 x = pi;
    V3 vh = v3(vh_s[i]);
    vh = vh + P.dt * a;
    x = x + P.dt * clamp3(vh + P.dt * a, -P.max_vel, P.max_vel);
    float vn = M::length(vh);
    V3 cp = {cl_clamp(x.x, P.lo_x, P.hi_x), cl_clamp(x.y, P.lo_y, 10.0f), cl_clamp(x.z, P.lo_z, P.hi_z)};
    V3 dvec = cp - x;
    float d = M::length(dvec);
    if (d > 0.0f && vn > 0.0f) {
        V3 nrm = M::normalize(V3{cl_sign(dvec.x), cl_sign(dvec.y), cl_sign(dvec.z)});
        float rest = M::div(1.0f * d, P.dt * vn);
        float kk = (1.0f + rest) * dot(vh, nrm);
        vh = vh - kk * nrm;
        x = cp;
    }
    V3 vt = vh + (P.dt * a) * 0.5f;
    pos_out[i] = f4(x, 1.0f);  // pos.w stays 1 (wcsph.cl:202 adds dt * 0 to it)
    vh_out[i] = f4(clamp3(vh, -P.max_vel, P.max_vel), 0.0f);
    vel_out[i] = f4(clamp3(vt, -P.max_vel, P.max_vel), 0.0f);
}

template <bool F>
__global__ __launch_bounds__(256) void k_accel_integrate(
    int n, const float4* __restrict__ prho, const float4* __restrict__ vel_s, const float4* __restrict__ vh_s,
    const float* __restrict__ pressure, const float4* __restrict__ normal, const int* __restrict__ nl,
    const int* __restrict__ nc, const int* __restrict__ bnl, const int* __restrict__ bnc,
    const float4* __restrict__ bpos_s, const float* __restrict__ bphi_s, WcParams P, float4* __restrict__ accel,
    float4* __restrict__ pos_out, float4* __restrict__ vel_out, float4* __restrict__ vh_out) {
    int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    accel_integrate_one<F>(i, n, prho, vel_s, vh_s, pressure, normal, nl, nc, bnl, bnc, bpos_s, bphi_s, P, accel,
                           pos_out, vel_out, vh_out);
}

void launch_accel_integrate(int n, const float4* prho, const float4* vel_s, const float4* vh_s,
                            const float* pressure, const float4* normal, const int* nl, const int* nc,
                            const int* bnl, const int* bnc, const float4* bpos_s, const float* bphi_s,
                            const WcParams& P, float4* accel, float4* pos_out, float4* vel_out, float4* vh_out,
                            hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "accel_integrate", s);
    hipLaunchKernelGGL(P.fast ? k_accel_integrate<true> : k_accel_integrate<false>, dim3((n + 255) / 256), dim3(256),
                       0, s, n, prho, vel_s, vh_s, pressure, normal, nl, nc, bnl, bnc, bpos_s, bphi_s, P, accel,
                       pos_out, vel_out, vh_out);
    PROF_END(prof, s);
}


// One step's acceleration and integration over a packed batch of scenes
// (batch.h); no scene has boundary particles.
__global__ __launch_bounds__(256) void k_batch_accel_integrate(
    int n, const int* __restrict__ sid, const BatchScene* __restrict__ sc, const float4* __restrict__ prho,
    const float4* __restrict__ vel_s, const float4* __restrict__ vh_s, const float* __restrict__ pressure,
    const float4* __restrict__ normal, const int* __restrict__ nl, const int* __restrict__ nc,
    const int* __restrict__ bnc, float4* __restrict__ accel, float4* __restrict__ pos_out,
    float4* __restrict__ vel_out, float4* __restrict__ vh_out) {
    for (int base = xcd_block(blockIdx.x, gridDim.x) * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        int i = base + threadIdx.x;
        if (i >= n) break;
        const BatchScene& S = sc[sid[i]];
        accel_integrate_one(i, n, prho, vel_s, vh_s, pressure, normal, nl, nc, nullptr, bnc, nullptr, nullptr, S.P,
                            accel, pos_out, vel_out, vh_out);
    }
}

void launch_batch_accel_integrate(int n, const int* sid, const BatchScene* sc, const float4* prho,
                                  const float4* vel_s, const float4* vh_s, const float* pressure,
                                  const float4* normal, const int* nl, const int* nc, const int* bnc, float4* accel,
                                  float4* pos_out, float4* vel_out, float4* vh_out, const LaunchCfg& c,
                                  hipStream_t s, Profiler* prof) {
    if (n <= 0) return;
    PROF_BEGIN(prof, "batch_accel_integrate", s);
    hipLaunchKernelGGL(k_batch_accel_integrate, dim3(batch_blocks(n, c)), dim3(c.block), 0, s, n, sid, sc, prho,
                       vel_s, vh_s, pressure, normal, nl, nc, bnc, accel, pos_out, vel_out, vh_out);
    PROF_END(prof, s);
}

}  // namespace sph
