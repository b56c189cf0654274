// Hand-written adjoint of the WCSPH step (host): forward with checkpoints,
// backward with per-segment recomputation (see adjoint.hip for the kernels).
#pragma once

#include <memory>
#include <vector>

#include "adjoint_dev.h"
#include "solver.h"

namespace sph {

class WcAdjoint {
public:
    // WCSPH settings, a clamp-box container, and the initial state by particle
    // id (xyzw; vel / vh may be NULL for zero). No boundary particles.
    WcAdjoint(const Settings& s, const float container[3], const float* pos, const float* vel, const float* vh,
              int n, int checkpoint_every);
    ~WcAdjoint();
    WcAdjoint(const WcAdjoint&) = delete;
    WcAdjoint& operator=(const WcAdjoint&) = delete;

    void set_initial(const float* pos, const float* vel, const float* vh);
    void set_params(float k_visc, float k_gas, float sigma);
    // runs nsteps from the initial state, keeping a checkpoint every C steps
    void forward(int nsteps);
    // final state by particle id (valid after forward, before backward)
    void final_state(float* pos, float* vel, float* vh);
    // gradients of a loss with adjoint (gx, gv, gvh) on the final state (by
    // id; NULL = 0): with respect to the initial state (by id) and to
    // (k_viscosity, gas_stiffness, surface_tension)
    void backward(const float* gx, const float* gv, const float* gvh, float* g_pos0, float* g_vel0, float* g_vh0,
                  double* g_params);
    int particle_count() const { return n_; }
    int steps() const { return steps_; }
    double forward_ms() const { return forward_ms_; }
    double backward_ms() const { return backward_ms_; }

private:
    struct Slot {
        DeviceArray<float4> prho, vel_s, vh_s, normal, accel;
        DeviceArray<float> pressure;
        DeviceArray<int> nl, nc, perm;
        void alloc(int n);
    };
    struct Checkpoint {
        DeviceArray<float4> pos, vel, vh;
        DeviceArray<int> ids;
    };
    WcStepParams params() const;
    void upload_initial();
    void run_step(int slot, const WcStepParams& W);
    void save_checkpoint(int k);
    void load_checkpoint(int k);

    Settings s_;
    float container_[3] = {0, 0, 0};
    int n_ = 0, every_ = 20, steps_ = 0, taped_segment_ = -1;
    bool fresh_ = true, final_valid_ = false;
    bool sorted_hint_ = false;  // d_keys_a holds the last step's sorted keys for the positions in d_pos
    float k_visc_ = 0, k_gas_ = 0, sigma_ = 0;
    double forward_ms_ = 0, backward_ms_ = 0;
    Grid grid_;
    hipStream_t stream_ = nullptr;
    std::vector<float4> init_pos_, init_vel_, init_vh_;
    std::vector<int> final_ids_;  // particle id at each stored index after forward
    std::vector<std::unique_ptr<Slot>> tape_;
    std::vector<std::unique_ptr<Checkpoint>> ck_;

    DeviceArray<float4> d_pos, d_vel, d_vh, d_pos_s;
    DeviceArray<float> d_density;
    DeviceArray<uint32_t> d_keys_a, d_keys_b, d_hist;
    DeviceArray<int> d_vals_a, d_vals_b, d_bnc, d_ids_a, d_ids_b, d_capped;
    DeviceArray<int2> d_ranges;
    DeviceArray<DevState> d_state;
    DeviceArray<unsigned char> d_resort;  // work arrays of the moved-keys sort
    ResortWork resort_;
    DeviceArray<float4> d_g[2];
    DeviceArray<float4> d_w_abar, d_w_xbar, d_w_vbar, d_w_vhbar, d_w_nbar;
    DeviceArray<float> d_w_rbar, d_w_pbar, d_part_a, d_part_b;
    DeviceArray<double> d_params;
};

}  // namespace sph
