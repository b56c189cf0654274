// Reverse-mode (adjoint) kernels of one WCSPH step (wcsph.cl:32-224 as
// restated in grid.hip and wcsph.hip), for gradients of a loss on the final
// state with respect to the initial state and to k_viscosity, the gas
// stiffness and the surface tension.
//
// A step maps the stored state (x, v, v_half) to the next one through the
// sort permutation, the neighbour lists, and (in sorted order i)
//   rho_i = c_rho sum_j (h^2 - s_ij)^3,   s = |r|^2 (or |r| without the
//           density repair), r = x_i - x_j, j over the list of i
//   p_i   = max(0, k_gas ((rho_i / rho0)^7 - 1))
//   n_i   = c_n sum_j r (H^2 - |r|^2)^2 / rho_j
//   a_i   = g - sigma (F_curv + st_main F_coh) - M P_i + c_v V_i
//   P_i   = sum_j phi(q) r (C_i + p_j / rho_j^2),  phi = spiky (h - q)^2 / q,
//           C_i = p_i / rho_i^2, q = |r|
//   V_i   = sum_j (v_j - v_i)(h - q) / rho_j
//   F_curv = sum_j K_ij (n_i - n_j),  F_coh = sum_j K_ij r/q st(q),
//           K_ij = 2 rho0 / (rho_i + rho_j)
// (pair terms of a_i need rho_i, rho_j > 0 and q > 1e-5), then the
// leap-frog with the clamp box. The lists are the fixed, piecewise-constant
// topology of the step; every kernel vanishes at q = h, so the loss is
// continuous across list changes.
//
// Lists are symmetric (j in list(i) iff i in list(j): the cutoff test and the
// 27-cell walk are symmetric while no list reaches the cap), so each thread
// gathers every derivative that lands on its own particle: from the terms of
// its own a_i, n_i, rho_i ("i side") and from the terms of its neighbours'
// a_j, n_j, rho_j that involve it ("j side"). No atomics, and the result does
// not depend on scheduling. A capped list breaks the symmetry; such a step is
// counted and the backward pass reports an error.
#include "adjoint_dev.h"
#include "kernels.h"
#include "sph_device.h"

namespace sph {

namespace {

__device__ __forceinline__ V3 zero3() { return {0.0f, 0.0f, 0.0f}; }
__device__ __forceinline__ V3 mask3(V3 v, V3 x, float lo, float hi) {
    return {(x.x >= lo && x.x <= hi) ? v.x : 0.0f, (x.y >= lo && x.y <= hi) ? v.y : 0.0f,
            (x.z >= lo && x.z <= hi) ? v.z : 0.0f};
}

// surface-tension spline st(q) (kernels/surface_tension.h:23-36) and dst/dq
__device__ __forceinline__ void st_and_slope(float q, float h, float c, float& st, float& dst) {
    if (q <= 0.5f * h) {
        float t = h - q;
        st = 2.0f * (cube(t) * cube(q)) - c;
        dst = 2.0f * (-3.0f * sqr(t) * cube(q) + 3.0f * cube(t) * sqr(q));
    } else if (q <= h) {
        float t = h - q;
        st = cube(t) * cube(q);
        dst = -3.0f * sqr(t) * cube(q) + 3.0f * cube(t) * sqr(q);
    } else {
        st = 0.0f;
        dst = 0.0f;
    }
}

template <int NT>
__device__ __forceinline__ void block_sum3(float a, float b, float c, float* out) {
    __shared__ float red[3][NT];
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    red[2][threadIdx.x] = c;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) out[k] = red[k][0];
}

}  // namespace

constexpr int ADJ_THREADS = 256;
// The neighbour loops of the adjoint kernels run on ADJ_SPLIT lanes per
// particle (entries k = part, part + ADJ_SPLIT, ...), whose sums are then
// combined: at the few tens of thousands of particles an adjoint runs on,
// one lane per particle leaves most of the GPU idle while each lane walks its
// list. (The adjoints are accurate to fp32 rounding; unlike the forward pass
// they are not bit-for-bit tied to an order of summation.)
constexpr int ADJ_SPLIT = 4;

__device__ __forceinline__ float split_sum(float v) {
    for (int off = 1; off < ADJ_SPLIT; off <<= 1) v += __shfl_xor(v, off, ADJ_SPLIT);
    return v;
}
__device__ __forceinline__ void split_sum(V3& v) {
    v.x = split_sum(v.x);
    v.y = split_sum(v.y);
    v.z = split_sum(v.z);
}

// ---------------------------------------------------------- K1: integrate
// Adjoint of predict_vel_n_pos (wcsph.cl:194-223) for sorted particle i.
// gx, gv, gvh: adjoints of the step's outputs at index i. Writes the adjoint
// of a_i, of the sorted input position (its direct part) and of the sorted
// input half velocity.
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_integrate(int n, AdjTape T, WcParams P,
                                                               const float4* __restrict__ g_pos,
                                                               const float4* __restrict__ g_vel,
                                                               const float4* __restrict__ g_vh,
                                                               float4* __restrict__ abar, float4* __restrict__ xbar,
                                                               float4* __restrict__ vhbar) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float dt = P.dt, vmax = P.max_vel;
    V3 xs = v3(T.prho[i]), vhs = v3(T.vh_s[i]), a = v3(T.accel[i]);
    // replay the forward values the branches depend on
    V3 vh = vhs + dt * a;
    V3 u = vh + dt * a;
    V3 x = xs + dt * clamp3(u, -vmax, vmax);
    float vn = length(vh);
    V3 cp = {cl_clamp(x.x, P.lo_x, P.hi_x), cl_clamp(x.y, P.lo_y, 10.0f), cl_clamp(x.z, P.lo_z, P.hi_z)};
    V3 dvec = cp - x;
    float d = length(dvec);
    bool hit = d > 0.0f && vn > 0.0f;
    V3 nrm = zero3(), vh2 = vh;
    float rest = 0.0f;
    if (hit) {
        nrm = normalize(V3{cl_sign(dvec.x), cl_sign(dvec.y), cl_sign(dvec.z)});
        rest = (1.0f * d) / (dt * vn);
        float kk = (1.0f + rest) * dot(vh, nrm);
        vh2 = vh - kk * nrm;
    }
    V3 vt = vh2 + (dt * a) * 0.5f;
    // outputs: pos = hit ? cp : x; vh_out = clamp(vh2); v_out = clamp(vt)
    V3 gX = v3(g_pos[i]);
    V3 gVt = mask3(v3(g_vel[i]), vt, -vmax, vmax);
    V3 gVh2 = mask3(v3(g_vh[i]), vh2, -vmax, vmax) + gVt;
    V3 gA = (0.5f * dt) * gVt;
    V3 gx_pre, gvh;
    if (hit) {
        // x2 = cp(x): unclamped components pass through
        V3 keep = {(x.x >= P.lo_x && x.x <= P.hi_x) ? 1.0f : 0.0f, (x.y >= P.lo_y && x.y <= 10.0f) ? 1.0f : 0.0f,
                   (x.z >= P.lo_z && x.z <= P.hi_z) ? 1.0f : 0.0f};
        // vh2 = vh - (1 + d / (dt vn)) (vh . nrm) nrm
        float gk = -dot(gVh2, nrm);
        gvh = gVh2 + (gk * (1.0f + rest)) * nrm;
        float grest = gk * dot(vh, nrm);
        float gd = grest / (dt * vn);
        float gvn = -grest * d / (dt * vn * vn);
        gvh = gvh + (gvn / vn) * vh;
        V3 gdvec = (gd / d) * dvec;  // d = |cp - x|
        V3 gcp = gX + gdvec;
        gx_pre = V3{keep.x * gcp.x, keep.y * gcp.y, keep.z * gcp.z} - gdvec;
    } else {
        gvh = gVh2;
        gx_pre = gX;
    }
    // x = xs + dt clamp(u); u = vh + dt a; vh = vhs + dt a
    V3 gu = mask3(dt * gx_pre, u, -vmax, vmax);
    gvh = gvh + gu;
    gA = gA + dt * gu + dt * gvh;
    abar[i] = f4(gA, 0.0f);
    xbar[i] = f4(gx_pre, 0.0f);
    vhbar[i] = f4(gvh, 0.0f);
}

// ------------------------------------------------------- K2: acceleration
// Adjoint of compute_acceleration (wcsph.cl:81-173) without boundary terms.
// Accumulates into xbar (positions), writes vbar, rbar (density, partial),
// pbar, nbar, and per-block partial sums of d/dk_visc and d/dsigma.
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_accel(int n, AdjTape T, WcParams P,
                                                           const float4* __restrict__ abar,
                                                           float4* __restrict__ xbar, float4* __restrict__ vbar,
                                                           float* __restrict__ rbar, float* __restrict__ pbar,
                                                           float4* __restrict__ nbar, float* __restrict__ partial) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / ADJ_SPLIT, part = t % ADJ_SPLIT;
    float g_kv = 0.0f, g_sig = 0.0f;
    V3 gx = zero3(), gv = zero3(), gn = zero3();
    float gr = 0.0f, gp = 0.0f, gC = 0.0f, di = 1.0f, Ci = 0.0f;
    if (i < n) {
        const float h = P.h, M = P.M, sm = P.st_main, sig = P.sigma, cv = P.visc_lapl * P.k_visc;
        float4 pi4 = T.prho[i];
        V3 xi = v3(pi4), vi = v3(T.vel_s[i]), ni = v3(T.normal[i]), ai = v3(abar[i]);
        di = pi4.w;
        const float pi = T.pressure[i];
        Ci = pi / sqr(di);
        const int len = T.nc[i];
        for (int k = part; k < len; k += ADJ_SPLIT) {
            int j = T.nl[(size_t)k * n + i];
            float4 pj4 = T.prho[j];
            float dj = pj4.w;
            V3 r = xi - v3(pj4);
            float q = length(r);
            if (!(di > 0.0f && dj > 0.0f && (double)q > 0.00001)) continue;
            float pj = T.pressure[j];
            V3 vj = v3(T.vel_s[j]), nj = v3(T.normal[j]), aj = v3(abar[j]);
            float Cj = pj / sqr(dj);
            float hq = h - q;
            float phi = P.spiky_grad * sqr(hq) / q;
            float dphi = P.spiky_grad * (-2.0f * hq / q - sqr(hq) / sqr(q));
            V3 G = phi * r;
            float K = (2.0f * P.rho0) / (di + dj);
            float dK = -2.0f * P.rho0 / sqr(di + dj);
            float st, dst;
            st_and_slope(q, h, P.st_term, st, dst);
            V3 rh = r / q;
            // ---- i side: terms of a_i, adjoint ai
            {
                float cP = Ci + pj / sqr(dj);
                V3 Jphi_a = phi * ai + (dphi / q) * dot(r, ai) * r;
                gx = gx + (-M * cP) * Jphi_a;
                gC += -M * dot(G, ai);
                V3 dv = vj - vi;
                float dva = dot(dv, ai);
                gv = gv + (-cv * hq / dj) * ai;
                gx = gx + (-cv * dva / dj) * rh;  // through q
                g_kv += P.visc_lapl * hq * dva / dj;
                V3 dn = ni - nj;
                float dna = dot(dn, ai);
                gn = gn + (-sig * K) * ai;
                float gK = -sig * dna;
                V3 psi = rh * st;
                float psia = dot(psi, ai);
                V3 Jpsi_a = (st / q) * (ai - dot(rh, ai) * rh) + dst * dot(rh, ai) * rh;
                gx = gx + (-sig * sm * K) * Jpsi_a;
                gK += -sig * sm * psia;
                gr += gK * dK;
                g_sig += -K * dna - sm * K * psia;
            }
            // ---- j side: terms of a_j that involve particle i, adjoint aj
            {
                float cPj = Cj + pi / sqr(di);
                V3 Jphi_a = phi * aj + (dphi / q) * dot(r, aj) * r;
                gx = gx + (M * cPj) * Jphi_a;
                float gD = M * dot(G, aj);  // d/d(p_i / rho_i^2)
                gp += gD / sqr(di);
                gr += -2.0f * pi * gD / cube(di);
                V3 dv = vi - vj;
                float dva = dot(dv, aj);
                gv = gv + (cv * hq / di) * aj;
                gr += -cv * hq * dva / sqr(di);
                gx = gx + (-cv * dva / di) * rh;
                V3 dn = nj - ni;
                gn = gn + (sig * K) * aj;
                float gK = -sig * dot(dn, aj);
                V3 Jpsi_a = (st / q) * (aj - dot(rh, aj) * rh) + dst * dot(rh, aj) * rh;
                gx = gx + (sig * sm * K) * Jpsi_a;
                gK += sig * sm * st * dot(rh, aj);
                gr += gK * dK;
            }
        }
    }
    split_sum(gx);
    split_sum(gv);
    split_sum(gn);
    gr = split_sum(gr);
    gp = split_sum(gp);
    gC = split_sum(gC);
    if (i < n && part == 0) {
        gp += gC / sqr(di);
        gr += -2.0f * Ci * gC / di;
        xbar[i] = f4(v3(xbar[i]) + gx, 0.0f);
        vbar[i] = f4(gv, 0.0f);
        nbar[i] = f4(gn, 0.0f);
        rbar[i] = gr;
        pbar[i] = gp;
    }
    block_sum3<ADJ_THREADS>(g_kv, g_sig, 0.0f, partial + 3 * blockIdx.x);
}

// ------------------------------------------------- K3: normals + pressure
// Adjoint of compute_normals (normals.cl:23-70) and of the Tait pressure
// (wcsph.cl:75-77); completes rbar and writes the d/dk_gas partials.
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_normals(int n, AdjTape T, float cn, float H2, float rho0,
                                                             float k_gas, const float4* __restrict__ nbar,
                                                             const float* __restrict__ pbar,
                                                             float4* __restrict__ xbar, float* __restrict__ rbar,
                                                             float* __restrict__ partial) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / ADJ_SPLIT, part = t % ADJ_SPLIT;
    float g_kg = 0.0f;
    V3 gx = zero3();
    float gr = 0.0f, di = 1.0f;
    if (i < n) {
        float4 pi4 = T.prho[i];
        V3 xi = v3(pi4), gni = v3(nbar[i]);
        di = pi4.w;
        const int len = T.nc[i];
        for (int k = part; k < len; k += ADJ_SPLIT) {
            int j = T.nl[(size_t)k * n + i];
            if (j == i) continue;  // r = x_i - x_i does not depend on x_i
            float4 pj4 = T.prho[j];
            float dj = pj4.w;
            V3 r = xi - v3(pj4);
            float w = H2 - dot(r, r);
            V3 gnj = v3(nbar[j]);
            // i side: n_i += cn r w^2 / rho_j
            gx = gx + (cn / dj) * (sqr(w) * gni - 4.0f * w * dot(r, gni) * r);
            // j side: n_j += cn (-r) w^2 / rho_i
            gx = gx - (cn / di) * (sqr(w) * gnj - 4.0f * w * dot(r, gnj) * r);
            gr += cn * sqr(w) * dot(r, gnj) / sqr(di);
        }
    }
    split_sum(gx);
    gr = split_sum(gr);
    if (i < n && part == 0) {
        float p = T.pressure[i];
        if (p > 0.0f) {
            float x = di / rho0;
            float x6 = cube(x) * cube(x);
            gr += pbar[i] * k_gas * 7.0f * x6 / rho0;
            g_kg = pbar[i] * (pown7(x) - 1.0f);
        }
        xbar[i] = f4(v3(xbar[i]) + gx, 0.0f);
        rbar[i] = rbar[i] + gr;
    }
    block_sum3<ADJ_THREADS>(0.0f, 0.0f, g_kg, partial + 3 * blockIdx.x);
}

// ------------------------------------------------------- K4: density
// Adjoint of the density sum (wcsph.cl:58-73), then the sorted-order
// adjoints go back to the step's stored order through the permutation.
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_density(int n, AdjTape T, float c_rho, float h2, int fix,
                                                             const float* __restrict__ rbar,
                                                             const float4* __restrict__ xbar,
                                                             const float4* __restrict__ vbar,
                                                             const float4* __restrict__ vhbar,
                                                             float4* __restrict__ gx_out, float4* __restrict__ gv_out,
                                                             float4* __restrict__ gvh_out, int* __restrict__ capped) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / ADJ_SPLIT, part = t % ADJ_SPLIT;
    V3 gx = zero3();
    if (i < n) {
    float4 pi4 = T.prho[i];
    V3 xi = v3(pi4);
    float gri = rbar[i];
    const int len = T.nc[i];
    for (int k = part; k < len; k += ADJ_SPLIT) {
        int j = T.nl[(size_t)k * n + i];
        if (j == i) continue;
        V3 r = xi - v3(T.prho[j]);
        float r2 = dot(r, r);
        float g = gri + rbar[j];
        if (fix) {
            // (h^2 - r^2)^3: d/dx_i = -6 (h^2 - r^2)^2 r
            gx = gx + (-6.0f * c_rho * sqr(h2 - r2) * g) * r;
        } else {
            float q = sqrtf(r2);
            if (q > 0.0f) gx = gx + (-3.0f * c_rho * sqr(h2 - q) * g / q) * r;
        }
    }
    }
    split_sum(gx);
    if (i >= n || part != 0) return;
    if (T.nc[i] >= LIST_CAP) atomicAdd(capped, 1);
    gx = v3(xbar[i]) + gx;
    int s = T.perm[i];
    gx_out[s] = f4(gx, 0.0f);
    gv_out[s] = vbar[i];
    gvh_out[s] = vhbar[i];
}

// Sums the per-block parameter partials of one step in a fixed order into
// double accumulators: one wave per parameter, lane l adding partials
// l, l + 64, ..., then a butterfly over the lanes.
constexpr int ADJ_PARAMS = 3;
__global__ __launch_bounds__(64 * ADJ_PARAMS) void k_adj_params(int nblocks, const float* __restrict__ pa,
                                                                const float* __restrict__ pb,
                                                                double* __restrict__ acc) {
    const int c = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double s = 0.0;
    for (int b = lane; b < nblocks; b += 64) s += (double)pa[3 * b + c] + (double)pb[3 * b + c];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) acc[c] += s;
}

int adj_blocks(int n) { return (n + ADJ_THREADS - 1) / ADJ_THREADS; }
int adj_split_blocks(int n) { return (int)(((long long)n * ADJ_SPLIT + ADJ_THREADS - 1) / ADJ_THREADS); }

void launch_adjoint_step(int n, const AdjTape& T, const WcParams& P, float cn, float nH2, float c_rho, float h2,
                         int fix, float rho0, float k_gas, const float4* gx, const float4* gv, const float4* gvh,
                         AdjWork& W, float4* gx_out, float4* gv_out, float4* gvh_out, double* params, int* capped,
                         hipStream_t s) {
    if (n <= 0) return;
    const int nb = adj_blocks(n), ns = adj_split_blocks(n);
    hipLaunchKernelGGL(k_adj_integrate, dim3(nb), dim3(ADJ_THREADS), 0, s, n, T, P, gx, gv, gvh, W.abar, W.xbar,
                       W.vhbar);
    hipLaunchKernelGGL(k_adj_accel, dim3(ns), dim3(ADJ_THREADS), 0, s, n, T, P, W.abar, W.xbar, W.vbar, W.rbar,
                       W.pbar, W.nbar, W.part_a);
    hipLaunchKernelGGL(k_adj_normals, dim3(ns), dim3(ADJ_THREADS), 0, s, n, T, cn, nH2, rho0, k_gas, W.nbar, W.pbar,
                       W.xbar, W.rbar, W.part_b);
    hipLaunchKernelGGL(k_adj_density, dim3(ns), dim3(ADJ_THREADS), 0, s, n, T, c_rho, h2, fix, W.rbar, W.xbar,
                       W.vbar, W.vhbar, gx_out, gv_out, gvh_out, capped);
    hipLaunchKernelGGL(k_adj_params, dim3(1), dim3(64 * ADJ_PARAMS), 0, s, ns, W.part_a, W.part_b, params);
}

}  // namespace sph
