// Host side of the gfx950 SPH solver. The class layout mirrors the
// reference's simulation layer: FluidSimulation (fluidsimulation.h:23-92) with
// PCISPHSimulation and WCSPHSimulation, a uniform Grid (grid.h) and a
// BoundaryHandler (boundaryhandler.h); the OpenCL runtime wrappers are
// replaced by one HIP stream per simulation and the kernels of kernels.h.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "kernels.h"
#include "sph.h"
#include "sph_device.h"

namespace sph {

struct Settings {  // PhysicsSettings + SimulationSettings (settings/*.h)
    float time_step = 0.01f, max_vel = 50.0f, fluid_particle_radius = 0.008f, fluid_support_radius = 0.032f;
    int pcisph_max_iterations = 7;
    float pcisph_error_ratio = 0.01f;
    int method = 0;
    float rest_density = 998.29f, k_viscosity = 0.0035105f, gravity = 9.8f, gas_stiffness = 3.0f,
          surface_tension = 0.0728f;
    int flags = 0;
};

class Error : public std::runtime_error {
public:
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);

// Owning device allocation.
template <class T> class DeviceArray {
public:
    DeviceArray() = default;
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    ~DeviceArray() { release(); }
    void resize(size_t n) {
        if (n == n_ && p_) return;
        release();
        if (n) hip_check(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
        n_ = n;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    void swap(DeviceArray& o) {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// -D literal of a float (opencl/clcompiler.cpp:50-54): fixed, 6 decimals.
float define_literal(float v);

// FluidVolume (fluidvolume.h:7-12)
class FluidVolume {
public:
    virtual ~FluidVolume() = default;
    virtual std::vector<float4> particles(float particle_radius) const = 0;
};
class FixedVolume : public FluidVolume {
public:
    explicit FixedVolume(std::vector<float4> p) : p_(std::move(p)) {}
    std::vector<float4> particles(float) const override { return p_; }

private:
    std::vector<float4> p_;
};
class BoxVolume : public FluidVolume {  // boxvolume.cpp:11-40
public:
    BoxVolume(const float size[3], const float center[3]);
    std::vector<float4> particles(float particle_radius) const override;

private:
    float size_[3], center_[3];
};

// Samplers (restated from scene/*.cpp)
std::vector<float4> sample_box_volume(const float size[3], const float center[3], float radius);
std::vector<float4> sample_fishtank(float w, float h, float d, float spacing);
std::vector<float4> sample_cube(float side, float spacing);
std::vector<float4> sample_wall(float w, float h, float spacing);

// Uniform grid (grid.cpp:8-12, 61-68) and the dense range box.
class Grid {
public:
    GridDev dev{};
    float size = 5.0f;
    float cell_size = 0.0f;
    int cps = 0;
    int key_bits = 0;
    bool have_box = false;
    int x0 = 0, x1 = -1, z0 = 0, z1 = -1;  // inclusive cell bounds of the box
    long long version = 0;                 // bumped whenever the box layout changes

    void reset(float h);
    void cell_of(float4 p, int& cx, int& cy, int& cz) const;
    // grows the box to contain the cells of pts (+ margin); returns true if it changed
    bool include(const std::vector<float4>& pts, int margin);
    size_t box_cells() const { return have_box ? (size_t)(x1 - x0 + 1) * cps * (z1 - z0 + 1) : 0; }
    void update_dev();
};

struct Body {
    int origin = 0, count = 0;
    bool can_move = false;
    float mass = 0.0f;
    float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float com[3] = {0, 0, 0}, lin[3] = {0, 0, 0}, ang[3] = {0, 0, 0};
};

class SimulationBase;

// BoundaryHandler (boundaryhandler.cpp): sampled rigid surfaces, psi, the
// boundary grid and the fluid wrench on each body.
class BoundaryHandler {
public:
    explicit BoundaryHandler(SimulationBase& sim) : sim_(sim) {}
    int add_boundary(const std::vector<float4>& local, bool can_move, float mass, const float R[9],
                     const float com[3]);
    void set_body_state(int id, const float* R, const float* com, const float* lin, const float* ang);
    void resample();                // psi of all bodies + sync(true)
    void sync(bool all);            // transform moving bodies, re-sort when needed
    void apply_fluid_forces();      // per-body wrench of the current step
    void wrench(int id, float f[3], float t[3]);
    int count() const { return (int)raw_.size(); }
    int body_count() const { return (int)bodies_.size(); }
    std::vector<float4> transformed_samples() const;
    void realloc_ranges();

    // device
    DeviceArray<float4> d_raw, d_unsorted, d_pos_s, d_force, d_torque;
    DeviceArray<float> d_phi, d_phi_s, d_partial, d_wrench;
    DeviceArray<uint32_t> d_keys_a, d_keys_b, d_hist;
    DeviceArray<int> d_vals_a, d_vals_b, d_offsets, d_nblocks;
    DeviceArray<int2> d_ranges;
    DeviceArray<uint32_t> d_near;  // per box cell: bit o set when neighbour cell o holds boundary particles
    int sorts = 0;
    long long box_version = -1;   // grid box layout d_ranges was built for
    std::vector<Body> bodies_;

private:
    void sort_all();
    SimulationBase& sim_;
    std::vector<float4> raw_;
    std::vector<float> wrench_host_;
    bool wrench_valid_ = false;
    int moving_ = 0;
};

// FluidSimulation (fluidsimulation.h:23-92)
class FluidSimulation {
public:
    virtual ~FluidSimulation() = default;
    virtual void simulate() = 0;
    virtual void reset(const Settings& s) = 0;
    virtual int particle_count() const = 0;
    virtual int add_boundary(const std::vector<float4>& local, bool can_move, float mass, const float R[9],
                             const float com[3]) = 0;
    virtual void set_rect_limits(float w, float h, float d) = 0;
    virtual int boundary_particle_count() const = 0;
    virtual void add_volume(std::shared_ptr<FluidVolume> v) = 0;
};

// State and neighbour search shared by both solvers.
class SimulationBase : public FluidSimulation {
public:
    explicit SimulationBase(const Settings& s);
    ~SimulationBase() override;

    int particle_count() const override { return n_; }
    int boundary_particle_count() const override { return boundary_.count(); }
    void set_rect_limits(float w, float h, float d) override;
    int add_boundary(const std::vector<float4>& local, bool can_move, float mass, const float R[9],
                     const float com[3]) override;
    void add_volume(std::shared_ptr<FluidVolume> v) override;
    void reset(const Settings& s) override;

    void step_dt(float dt);
    void set_body_state(int id, const float* R, const float* com, const float* lin, const float* ang) {
        boundary_.set_body_state(id, R, com, lin, ang);
    }
    void body_wrench(int id, float f[3], float t[3]) { boundary_.wrench(id, f, t); }
    int body_count() const { return boundary_.body_count(); }

    void get_state(int which, float* xyzw, int n);
    void set_state(int which, const float* xyzw, int n);
    long long get_field(int field, void* out, long long cap);
    void stats(sph_step_stats& out);
    void synchronize();

    hipStream_t stream() const { return stream_; }
    Profiler& profiler() { return prof_; }
    const Settings& settings() const { return s_; }
    Grid& grid() { return grid_; }

    // device fluid buffers (public for the boundary handler)
    DeviceArray<float4> d_pos, d_vel, d_vh, d_pos_s, d_vel_s, d_vh_s;
    DeviceArray<float4> d_normal, d_f_other, d_f_press, d_pos_pred, d_accel;
    DeviceArray<float4> d_prho;  // (x, y, z, density) of the sorted particles
    DeviceArray<float4> d_pc;    // (x, y, z, pressure / density^2), PCISPH
    DeviceArray<float> d_density, d_pressure, d_err;
    DeviceArray<uint32_t> d_keys_a, d_keys_b, d_hist;
    DeviceArray<int> d_vals_a, d_vals_b, d_nl, d_nc, d_bnl, d_bnc;
    DeviceArray<int2> d_ranges;
    // work arrays of the sort that places only the keys that changed cell
    DeviceArray<unsigned char> d_resort;
    DeviceArray<DevState> d_state;
    DeviceArray<unsigned long long> d_acc;
    float container[3] = {0, 0, 0};
    // frames advanced by simulate() and the simulated time they cover
    long long frames_done = 0;
    double sim_time = 0.0;
    bool have_container = false;

protected:
    virtual void initialize_params() = 0;
    virtual void after_initialize() {}
    void initialize_solver();
    void alloc_fluid(int n);
    void neighbour_search();
    void ensure_box(const std::vector<float4>& fluid_pts);
    void ensure_ranges();
    void require_ready() const;

    Settings s_;
    Grid grid_;
    BoundaryHandler boundary_;
    std::vector<std::shared_ptr<FluidVolume>> volumes_;
    int n_ = 0;
    bool fresh_ranges_ = true;  // the fluid range box holds no ranges yet
    bool sorted_hint_ = false;  // d_keys_a holds a non-decreasing array of n_ keys (a sort's output)
    ResortWork resort_;         // views into d_resort
    long long box_version_ = -1;  // grid box layout d_ranges was built for
    hipStream_t stream_ = nullptr;
    Profiler prof_;
    friend class BoundaryHandler;
};

class PCISPHSimulation : public SimulationBase {
public:
    explicit PCISPHSimulation(const Settings& s);
    void simulate() override {
        simulate_step(3, s_.pcisph_max_iterations);
        ++frames_done;
        sim_time += s_.time_step;
    }
    float density_scale_factor() const { return delta_; }
    PciParams params() const { return P_; }

    // Launch shapes of the step's list kernels (the single-scene form of the
    // reference's launch autotuning, opencl/clkernel.cpp:85-151): shapes
    // change timing only, never results.
    enum TunedKernel { T_NORMALS, T_FORCES, T_UPDATE, T_PFORCE, T_COUNT };
    static const char* tuned_name(int k);
    Shape shape(int k) const { return shapes_[k]; }
    void set_shape(int k, Shape sh);
    // Times each kernel's candidate shapes over `trials` steps from the
    // current state, keeps the fastest, restores the state; with cache_path,
    // earlier results for the same device and scene size are reused (returns
    // 1) or the new ones appended (returns 0).
    int autotune(const char* cache_path, int trials);

protected:
    void initialize_params() override;
    void after_initialize() override;
    void simulate_step(int min_iter, int max_iter);
    std::string tune_key() const;
    Shape shapes_[T_COUNT];
    void deduce_density_scale_factor();
    void fill_params();
    float mass_ = 0.0f, st_main_ = 0.0f, st_term_ = 0.0f, delta_ = 0.0f;
    PciParams P_{};
};

class WCSPHSimulation : public SimulationBase {
public:
    explicit WCSPHSimulation(const Settings& s);
    void simulate() override;

protected:
    void initialize_params() override;
    float mass_ = 0.0f;
};

// Kernel arguments of one WCSPH step (density, normals, acceleration).
struct WcStepParams {
    DensityParams D;
    WcParams P;
    float nH2, nw_grad, nM, nH;
};
WcStepParams wcsph_step_params(const Settings& s, const float container[3]);

// FluidSimulationFactory::build_simulation (fluidsimulationfactory.h:13-40)
std::unique_ptr<SimulationBase> build_simulation(const Settings& s);

// Settings::load (settings/settings.cpp:14-59)
Settings load_settings(const std::string& path);

// Fluid particle mass of a method, as the reference's host computes it.
float fluid_mass(const Settings& s);

}  // namespace sph
