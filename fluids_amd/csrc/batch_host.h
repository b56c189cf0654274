// Batched WCSPH solver (host): many independent scenes, one packed particle
// array, one set of launches per step (see batch.h for the device layout).
#pragma once

#include <string>
#include <vector>

#include "batch.h"
#include "solver.h"

namespace sph {

class BatchWCSPH {
public:
    enum Kernel { K_HASH = 0, K_GATHER, K_LISTS, K_NORMALS, K_ACCEL, K_COUNT };

    BatchWCSPH();
    ~BatchWCSPH();
    BatchWCSPH(const BatchWCSPH&) = delete;
    BatchWCSPH& operator=(const BatchWCSPH&) = delete;

    // a WCSPH scene: its settings, its container (set_rect_limits) and its
    // particles (the concatenated FluidVolume samples); returns the scene id
    int add_scene(const Settings& s, const float container[3], const float* xyzw, int n);
    void finalize();
    void step(int nsteps);
    void synchronize();
    void set_graph(bool on);
    void set_config(int kernel, const LaunchCfg& c);
    const LaunchCfg& config(int kernel) const { return cfg_[kernel]; }
    double tuned_ms(int kernel) const { return tuned_ms_[kernel]; }
    static const char* kernel_name(int k);
    // times each kernel's launch shapes on the batch and keeps the fastest;
    // with a cache file, reuses a result for the same device and batch size.
    // Returns 1 when the cache supplied the shapes.
    int autotune(const char* cache_path, int trials);

    int scene_count() const { return (int)scenes_.size(); }
    int scene_particles(int scene) const;
    int total_particles() const { return n_; }
    long long range_cells() const;
    // which: 0 positions, 1 velocities, 2 half-step velocities (float4, stored
    // order), 3 density (float, sorted order), 4 neighbour counts (int)
    void copy_scene(int which, int scene, float* out, int n);
    void set_scene_state(int which, int scene, const float* in, int n);
    Profiler& profiler() { return prof_; }
    hipStream_t stream() const { return stream_; }
    DevState* device_state() { return d_state.get(); }

private:
    struct HostScene {
        Settings s;
        float container[3];
        int n = 0;
        std::vector<float4> pos;
    };
    void require_final() const;
    void check_scene(int scene) const;
    void enqueue_step(bool clear);
    void drop_graph();
    std::string tune_key() const;
    bool load_tuning(const std::string& path);
    void save_tuning(const std::string& path) const;

    std::vector<HostScene> host_;
    std::vector<BatchScene> scenes_;
    int n_ = 0, key_bits_ = 1, sid_bits_ = 8;
    bool finalized_ = false, fresh_ = true, use_graph_ = true;
    // d_keys_a holds the last step's sorted keys for the positions in d_pos:
    // the next step places only the keys that changed cell
    bool sorted_hint_ = false;
    LaunchCfg cfg_[K_COUNT];
    double tuned_ms_[K_COUNT] = {0, 0, 0, 0, 0};
    hipStream_t stream_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    Profiler prof_;

    DeviceArray<float4> d_pos, d_vel, d_vh, d_pos_s, d_vel_s, d_vh_s, d_prho, d_normal, d_accel;
    DeviceArray<float> d_density, d_pressure;
    DeviceArray<uint32_t> d_keys_a, d_keys_b, d_hist;
    DeviceArray<int> d_vals_a, d_vals_b, d_nl, d_nc, d_bnc, d_sid;
    DeviceArray<int2> d_ranges;
    DeviceArray<BatchScene> d_scenes;
    DeviceArray<DevState> d_state;
    DeviceArray<unsigned char> d_resort;  // work arrays of the moved-keys sort
    ResortWork resort_;
};

}  // namespace sph
