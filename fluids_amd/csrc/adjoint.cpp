// Host side of the WCSPH adjoint: a forward driver over the same kernels and
// parameters as WCSPHSimulation (so its trajectory is the solver's, bit for
// bit), checkpoints every C steps, and the backward sweep that recomputes
// each segment's tape from its checkpoint and runs the adjoint kernels of
// adjoint.hip in reverse step order.
#include "adjoint_host.h"

#include <chrono>
#include <cstring>

namespace sph {

WcAdjoint::WcAdjoint(const Settings& s, const float container[3], const float* pos, const float* vel,
                     const float* vh, int n, int checkpoint_every)
    : s_(s), n_(n), every_(checkpoint_every) {
    if (s.method != SPH_WCSPH) throw Error(SPH_ERR_INVALID, "adjoint: WCSPH only");
    if (n <= 0 || !pos) throw Error(SPH_ERR_INVALID, "adjoint: no particles");
    if (every_ < 1) throw Error(SPH_ERR_INVALID, "adjoint: checkpoint interval must be >= 1");
    for (int k = 0; k < 3; ++k) {
        if (!(container[k] > 0.0f)) throw Error(SPH_ERR_INVALID, "adjoint: container sides must be > 0");
        container_[k] = container[k];
    }
    k_visc_ = s.k_viscosity;
    k_gas_ = s.gas_stiffness;
    sigma_ = s.surface_tension;
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    grid_.reset(s.fluid_support_radius);
    std::vector<float4> corners;
    for (int sx = -1; sx <= 1; sx += 2)
        for (int sz = -1; sz <= 1; sz += 2)
            corners.push_back(make_float4(sx * container_[0] / 2, 0.0f, sz * container_[2] / 2, 1.0f));
    grid_.include(corners, 2);
    const size_t N = n_;
    d_pos.resize(N); d_vel.resize(N); d_vh.resize(N); d_pos_s.resize(N); d_density.resize(N);
    d_keys_a.resize(N); d_keys_b.resize(N); d_vals_a.resize(N); d_vals_b.resize(N);
    d_hist.resize(sort_hist_words(n_));
    d_bnc.resize(N); d_ids_a.resize(N); d_ids_b.resize(N);
    d_state.resize(1);
    d_g[0].resize(3 * N); d_g[1].resize(3 * N);
    d_w_abar.resize(N); d_w_xbar.resize(N); d_w_vbar.resize(N); d_w_vhbar.resize(N); d_w_nbar.resize(N);
    d_w_rbar.resize(N); d_w_pbar.resize(N);
    d_part_a.resize(3 * (size_t)adj_split_blocks(n_)); d_part_b.resize(3 * (size_t)adj_split_blocks(n_));
    d_params.resize(3); d_capped.resize(1);
    hip_check(hipMemsetAsync(d_bnc.get(), 0, N * sizeof(int), stream_), "memset");
    hip_check(hipMemsetAsync(d_keys_a.get(), 0, N * sizeof(uint32_t), stream_), "memset");
    d_resort.resize(resort_bytes(n_));
    resort_ = resort_carve(d_resort.get(), n_);
    resort_clear(resort_, n_, stream_);
    for (int k = 0; k < every_; ++k) tape_.emplace_back(new Slot());
    set_initial(pos, vel, vh);
}

WcAdjoint::~WcAdjoint() {
    if (stream_) hipStreamDestroy(stream_);
}

void WcAdjoint::set_initial(const float* pos, const float* vel, const float* vh) {
    if (!pos) throw Error(SPH_ERR_INVALID, "adjoint: positions are required");
    init_pos_.resize(n_);
    init_vel_.assign(n_, make_float4(0, 0, 0, 0));
    init_vh_.assign(n_, make_float4(0, 0, 0, 0));
    for (int i = 0; i < n_; ++i) {
        init_pos_[i] = make_float4(pos[4 * i], pos[4 * i + 1], pos[4 * i + 2], 1.0f);
        if (vel) init_vel_[i] = make_float4(vel[4 * i], vel[4 * i + 1], vel[4 * i + 2], 0.0f);
        if (vh) init_vh_[i] = make_float4(vh[4 * i], vh[4 * i + 1], vh[4 * i + 2], 0.0f);
    }
    if (grid_.include(init_pos_, 2) || !d_ranges.get()) {
        d_ranges.release();
        d_ranges.resize(std::max<size_t>(grid_.box_cells(), 1));
    }
    steps_ = 0;
    taped_segment_ = -1;
}

void WcAdjoint::set_params(float k_visc, float k_gas, float sigma) {
    k_visc_ = k_visc;
    k_gas_ = k_gas;
    sigma_ = sigma;
    steps_ = 0;
    taped_segment_ = -1;
}

WcStepParams WcAdjoint::params() const {
    Settings s = s_;
    s.k_viscosity = k_visc_;
    s.gas_stiffness = k_gas_;
    s.surface_tension = sigma_;
    return wcsph_step_params(s, container_);
}

void WcAdjoint::Slot::alloc(int n) {
    if ((int)nc.size() == n) return;
    const size_t N = n;
    prho.resize(N); vel_s.resize(N); vh_s.resize(N); normal.resize(N); accel.resize(N);
    pressure.resize(N); nl.resize(N * LIST_CAP); nc.resize(N); perm.resize(N);
}

void WcAdjoint::upload_initial() {
    const size_t B = n_ * sizeof(float4);
    hip_check(hipMemcpyAsync(d_pos.get(), init_pos_.data(), B, hipMemcpyHostToDevice, stream_), "upload");
    hip_check(hipMemcpyAsync(d_vel.get(), init_vel_.data(), B, hipMemcpyHostToDevice, stream_), "upload");
    hip_check(hipMemcpyAsync(d_vh.get(), init_vh_.data(), B, hipMemcpyHostToDevice, stream_), "upload");
    std::vector<int> ids(n_);
    for (int i = 0; i < n_; ++i) ids[i] = i;
    hip_check(hipMemcpyAsync(d_ids_a.get(), ids.data(), n_ * sizeof(int), hipMemcpyHostToDevice, stream_), "upload");
    // a fresh range box: the first step has no previous keys to clear by
    hip_check(hipMemsetAsync(d_ranges.get(), 0, d_ranges.size() * sizeof(int2), stream_), "memset");
    fresh_ = true;
    sorted_hint_ = false;
}

// One WCSPHSimulation::simulate (solver.cpp) writing its intermediate values
// into tape slot `slot`.
void WcAdjoint::run_step(int slot, const WcStepParams& W) {
    Slot& T = *tape_[slot];
    T.alloc(n_);
    hipStream_t s = stream_;
    launch_step_begin(d_state.get(), s);
    int2* clear = fresh_ ? nullptr : d_ranges.get();
    if (sorted_hint_) {
        // the solver's sort: particles are stored in the last step's order,
        // so only the keys that changed cell are placed
        launch_hash_resort(n_, d_pos.get(), grid_.dev, d_keys_b.get(), d_keys_a.get(), d_vals_a.get(), clear,
                           d_state.get(), resort_, s, nullptr);
    } else {
        const int passes = sort_passes(grid_.key_bits);
        uint32_t* hash_out = (passes % 2) ? d_keys_b.get() : d_keys_a.get();
        launch_hash(n_, d_pos.get(), grid_.dev, hash_out, d_keys_a.get(), clear, d_state.get(), s, nullptr);
        launch_radix_sort(d_keys_a.get(), d_keys_b.get(), d_vals_a.get(), d_vals_b.get(), d_hist.get(), n_,
                          grid_.key_bits, s, nullptr);
        launch_resort_samples(d_keys_a.get(), n_, resort_, s);
    }
    fresh_ = false;
    sorted_hint_ = true;
    // the gather also records the permutation and moves the particle ids
    // into the new order (into d_ids_b, swapped in below)
    launch_gather_fluid_tape(n_, d_vals_a.get(), d_keys_a.get(), d_pos.get(), d_vel.get(), d_vh.get(), d_pos_s.get(),
                             T.vel_s.get(), T.vh_s.get(), d_ranges.get(), grid_.dev, T.perm.get(), d_ids_a.get(),
                             d_ids_b.get(), s);
    launch_lists_density(true, n_, d_pos_s.get(), d_ranges.get(), 0, nullptr, nullptr, nullptr, nullptr, grid_.dev,
                         W.D, T.nl.get(), T.nc.get(), nullptr, d_bnc.get(), d_density.get(), T.pressure.get(),
                         T.prho.get(), d_state.get(), s, nullptr);
    launch_normals(n_, T.prho.get(), T.nl.get(), T.nc.get(), W.nH2, W.nw_grad, W.nM, W.nH, T.normal.get(), s,
                   nullptr);
    launch_accel_integrate(n_, T.prho.get(), T.vel_s.get(), T.vh_s.get(), T.pressure.get(), T.normal.get(),
                           T.nl.get(), T.nc.get(), nullptr, d_bnc.get(), nullptr, nullptr, W.P, T.accel.get(),
                           d_pos.get(), d_vel.get(), d_vh.get(), s, nullptr);
    d_ids_a.swap(d_ids_b);
}

void WcAdjoint::save_checkpoint(int k) {
    while ((int)ck_.size() <= k) ck_.emplace_back(new Checkpoint());
    Checkpoint& c = *ck_[k];
    if ((int)c.ids.size() != n_) {
        c.pos.resize(n_); c.vel.resize(n_); c.vh.resize(n_); c.ids.resize(n_);
    }
    const size_t B = n_ * sizeof(float4);
    hip_check(hipMemcpyAsync(c.pos.get(), d_pos.get(), B, hipMemcpyDeviceToDevice, stream_), "copy");
    hip_check(hipMemcpyAsync(c.vel.get(), d_vel.get(), B, hipMemcpyDeviceToDevice, stream_), "copy");
    hip_check(hipMemcpyAsync(c.vh.get(), d_vh.get(), B, hipMemcpyDeviceToDevice, stream_), "copy");
    hip_check(hipMemcpyAsync(c.ids.get(), d_ids_a.get(), n_ * sizeof(int), hipMemcpyDeviceToDevice, stream_), "copy");
}

void WcAdjoint::load_checkpoint(int k) {
    Checkpoint& c = *ck_[k];
    const size_t B = n_ * sizeof(float4);
    hip_check(hipMemcpyAsync(d_pos.get(), c.pos.get(), B, hipMemcpyDeviceToDevice, stream_), "copy");
    hip_check(hipMemcpyAsync(d_vel.get(), c.vel.get(), B, hipMemcpyDeviceToDevice, stream_), "copy");
    hip_check(hipMemcpyAsync(d_vh.get(), c.vh.get(), B, hipMemcpyDeviceToDevice, stream_), "copy");
    hip_check(hipMemcpyAsync(d_ids_a.get(), c.ids.get(), n_ * sizeof(int), hipMemcpyDeviceToDevice, stream_), "copy");
    sorted_hint_ = false;  // the restored positions belong to an earlier order
}

void WcAdjoint::forward(int nsteps) {
    if (nsteps < 1) throw Error(SPH_ERR_RANGE, "adjoint: at least one step");
    const WcStepParams W = params();
    auto t0 = std::chrono::steady_clock::now();
    upload_initial();
    for (int t = 0; t < nsteps; ++t) {
        if (t % every_ == 0) save_checkpoint(t / every_);
        run_step(t % every_, W);
    }
    // the particle ids of the final state, by stored index: d_ids_a is
    // overwritten when backward recomputes an earlier segment
    final_ids_.resize(n_);
    hip_check(hipMemcpyAsync(final_ids_.data(), d_ids_a.get(), n_ * sizeof(int), hipMemcpyDeviceToHost, stream_),
              "download");
    hip_check(hipStreamSynchronize(stream_), "sync");
    forward_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    steps_ = nsteps;
    taped_segment_ = (nsteps - 1) / every_;
    final_valid_ = true;
}

void WcAdjoint::final_state(float* pos, float* vel, float* vh) {
    if (steps_ == 0) throw Error(SPH_ERR_STATE, "adjoint: run forward first");
    std::vector<float4> p(n_), v(n_), w(n_);
    std::vector<int> ids(n_);
    // the stored arrays hold the end of the forward pass until backward runs
    if (!final_valid_) throw Error(SPH_ERR_STATE, "adjoint: the final state was overwritten by a backward pass");
    hip_check(hipMemcpyAsync(p.data(), d_pos.get(), n_ * sizeof(float4), hipMemcpyDeviceToHost, stream_), "download");
    hip_check(hipMemcpyAsync(v.data(), d_vel.get(), n_ * sizeof(float4), hipMemcpyDeviceToHost, stream_), "download");
    hip_check(hipMemcpyAsync(w.data(), d_vh.get(), n_ * sizeof(float4), hipMemcpyDeviceToHost, stream_), "download");
    hip_check(hipMemcpyAsync(ids.data(), d_ids_a.get(), n_ * sizeof(int), hipMemcpyDeviceToHost, stream_), "download");
    hip_check(hipStreamSynchronize(stream_), "sync");
    for (int i = 0; i < n_; ++i) {
        int id = ids[i];
        if (id < 0 || id >= n_) throw Error(SPH_ERR_STATE, "adjoint: particle ids corrupted");
        if (pos) std::memcpy(pos + 4 * id, &p[i], sizeof(float4));
        if (vel) std::memcpy(vel + 4 * id, &v[i], sizeof(float4));
        if (vh) std::memcpy(vh + 4 * id, &w[i], sizeof(float4));
    }
}

void WcAdjoint::backward(const float* gx, const float* gv, const float* gvh, float* g_pos0, float* g_vel0,
                         float* g_vh0, double* g_params) {
    if (steps_ == 0) throw Error(SPH_ERR_STATE, "adjoint: run forward first");
    auto t0 = std::chrono::steady_clock::now();
    const WcStepParams W = params();
    // adjoint of the final state, by stored index (the ids kept by forward,
    // so that a repeated backward maps the cotangent the same way)
    const std::vector<int>& ids = final_ids_;
    if ((int)ids.size() != n_) throw Error(SPH_ERR_STATE, "adjoint: run forward first");
    std::vector<float4> g(3 * (size_t)n_, make_float4(0, 0, 0, 0));
    for (int i = 0; i < n_; ++i) {
        int id = ids[i];
        if (id < 0 || id >= n_) throw Error(SPH_ERR_STATE, "adjoint: particle ids corrupted");
        if (gx) g[i] = make_float4(gx[4 * id], gx[4 * id + 1], gx[4 * id + 2], 0.0f);
        if (gv) g[n_ + i] = make_float4(gv[4 * id], gv[4 * id + 1], gv[4 * id + 2], 0.0f);
        if (gvh) g[2 * (size_t)n_ + i] = make_float4(gvh[4 * id], gvh[4 * id + 1], gvh[4 * id + 2], 0.0f);
    }
    int cur = 0;
    hip_check(hipMemcpyAsync(d_g[cur].get(), g.data(), g.size() * sizeof(float4), hipMemcpyHostToDevice, stream_),
              "upload");
    hip_check(hipMemsetAsync(d_params.get(), 0, 3 * sizeof(double), stream_), "memset");
    hip_check(hipMemsetAsync(d_capped.get(), 0, sizeof(int), stream_), "memset");
    AdjWork Wk{d_w_abar.get(), d_w_xbar.get(), d_w_vbar.get(), d_w_vhbar.get(), d_w_nbar.get(),
               d_w_rbar.get(), d_w_pbar.get(), d_part_a.get(), d_part_b.get()};
    const float cn = (W.nw_grad * W.nM) * W.nH;
    const float c_rho = W.D.w_eval * W.D.M;
    const float h2 = W.D.h * W.D.h;
    const int nseg = (steps_ + every_ - 1) / every_;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int a = seg * every_, b = std::min(steps_, a + every_);
        if (seg != taped_segment_) {
            load_checkpoint(seg);
            for (int t = a; t < b; ++t) run_step(t - a, W);
            taped_segment_ = seg;
        }
        final_valid_ = false;
        for (int t = b - 1; t >= a; --t) {
            Slot& S = *tape_[t - a];
            AdjTape T{S.prho.get(), S.vel_s.get(), S.vh_s.get(), S.pressure.get(), S.normal.get(), S.accel.get(),
                      S.nl.get(), S.nc.get(), S.perm.get()};
            float4* gin = d_g[cur].get();
            float4* gout = d_g[1 - cur].get();
            launch_adjoint_step(n_, T, W.P, cn, W.nH2, c_rho, h2, W.D.fix_density, W.D.rho0, W.D.k_gas, gin,
                                gin + n_, gin + 2 * (size_t)n_, Wk, gout, gout + n_, gout + 2 * (size_t)n_,
                                d_params.get(), d_capped.get(), stream_);
            cur = 1 - cur;
        }
    }
    hip_check(hipMemcpyAsync(g.data(), d_g[cur].get(), g.size() * sizeof(float4), hipMemcpyDeviceToHost, stream_),
              "download");
    double prm[3];
    int capped = 0;
    hip_check(hipMemcpyAsync(prm, d_params.get(), sizeof prm, hipMemcpyDeviceToHost, stream_), "download");
    hip_check(hipMemcpyAsync(&capped, d_capped.get(), sizeof capped, hipMemcpyDeviceToHost, stream_), "download");
    hip_check(hipStreamSynchronize(stream_), "sync");
    backward_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (capped > 0)
        throw Error(SPH_ERR_RANGE, "adjoint: " + std::to_string(capped) +
                                       " neighbour lists reached the cap of 50; the adjoint needs uncapped lists");
    // the state before step 0 is in particle-id order
    for (int i = 0; i < n_; ++i)
        for (int c = 0; c < 3; ++c) {
            if (g_pos0) g_pos0[4 * i + c] = (&g[i].x)[c];
            if (g_vel0) g_vel0[4 * i + c] = (&g[n_ + i].x)[c];
            if (g_vh0) g_vh0[4 * i + c] = (&g[2 * (size_t)n_ + i].x)[c];
        }
    for (int i = 0; i < n_; ++i) {
        if (g_pos0) g_pos0[4 * i + 3] = 0.0f;
        if (g_vel0) g_vel0[4 * i + 3] = 0.0f;
        if (g_vh0) g_vh0[4 * i + 3] = 0.0f;
    }
    if (g_params) {
        g_params[0] = prm[0];  // k_viscosity
        g_params[1] = prm[2];  // gas_stiffness
        g_params[2] = prm[1];  // surface_tension
    }
}

}  // namespace sph
