// extern "C" entry points of libsph (declared in include/sph.h). Every entry
// point converts exceptions into status codes; the message of the last
// failure of the calling thread is kept for sph_last_error().
#include <cmath>
#include <cstring>
#include <string>

#include "adjoint_host.h"
#include "batch_host.h"
#include "rigid.h"
#include "output.h"
#include "scene.h"
#include "solver.h"
#include "sph.h"

using namespace sph;

struct sph_handle {
    std::unique_ptr<SimulationBase> sim;
};

struct sph_adjoint {
    std::unique_ptr<WcAdjoint> adj;
};

struct sph_batch {
    BatchWCSPH batch;
};

struct sph_rigid_world {
    std::unique_ptr<RigidWorld> own;
    RigidWorld* world = nullptr;  // own.get(), or a world owned by a scene
};

struct sph_frames {
    std::unique_ptr<FrameWriter> w;
};

struct sph_perf_log {
    std::unique_ptr<PerfLog> p;
};

struct sph_scene {
    std::unique_ptr<SceneRun> run;
    sph_handle handle;
    sph_rigid_world world;
};

namespace {
thread_local std::string g_last_error;

template <class F> int guard(F&& f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return SPH_ERR_INVALID;
    } catch (...) {
        g_last_error = "unknown error";
        return SPH_ERR_INVALID;
    }
}

Settings to_settings(const sph_settings* s) {
    Settings o;
    o.time_step = s->time_step;
    o.max_vel = s->max_vel;
    o.fluid_particle_radius = s->fluid_particle_radius;
    o.fluid_support_radius = s->fluid_support_radius;
    o.pcisph_max_iterations = s->pcisph_max_iterations;
    o.pcisph_error_ratio = s->pcisph_error_ratio;
    o.method = s->method;
    o.rest_density = s->rest_density;
    o.k_viscosity = s->k_viscosity;
    o.gravity = s->gravity;
    o.gas_stiffness = s->gas_stiffness;
    o.surface_tension = s->surface_tension;
    o.flags = s->flags;
    return o;
}

void from_settings(const Settings& o, sph_settings* s) {
    s->time_step = o.time_step;
    s->max_vel = o.max_vel;
    s->fluid_particle_radius = o.fluid_particle_radius;
    s->fluid_support_radius = o.fluid_support_radius;
    s->pcisph_max_iterations = o.pcisph_max_iterations;
    s->pcisph_error_ratio = o.pcisph_error_ratio;
    s->method = o.method;
    s->rest_density = o.rest_density;
    s->k_viscosity = o.k_viscosity;
    s->gravity = o.gravity;
    s->gas_stiffness = o.gas_stiffness;
    s->surface_tension = o.surface_tension;
    s->flags = o.flags;
}

void validate(const sph_settings* s) {
    if (!s) throw Error(SPH_ERR_INVALID, "settings is NULL");
    if (s->method != SPH_WCSPH && s->method != SPH_PCISPH) throw Error(SPH_ERR_INVALID, "Unknown simulation method!");
    if (!(s->fluid_support_radius > 0.0f) || !(s->fluid_particle_radius > 0.0f))
        throw Error(SPH_ERR_INVALID, "radii must be positive");
    if (!(s->time_step > 0.0f)) throw Error(SPH_ERR_INVALID, "time_step must be positive");
    if (s->flags & ~(SPH_FIX_WCSPH_DENSITY | SPH_FAST_MATH | SPH_FULL_SORT)) throw Error(SPH_ERR_INVALID, "unknown flag bits");
}

// Batches and the adjoint run the IEEE kernels only (the adjoint's
// derivatives are those of the IEEE step).
void validate_ieee_only(const sph_settings* s, const char* who) {
    validate(s);
    if (s->flags & SPH_FAST_MATH) throw Error(SPH_ERR_INVALID, std::string(who) + ": SPH_FAST_MATH is not supported");
}

SimulationBase& sim_of(sph_handle* h) {
    if (!h || !h->sim) throw Error(SPH_ERR_INVALID, "handle is NULL");
    return *h->sim;
}

std::vector<float4> to_f4(const float* xyzw, int n, float w) {
    if (n < 0) throw Error(SPH_ERR_RANGE, "negative count");
    if (n > 0 && !xyzw) throw Error(SPH_ERR_INVALID, "data is NULL");
    std::vector<float4> v(n);
    for (int i = 0; i < n; ++i) v[i] = make_float4(xyzw[4 * i], xyzw[4 * i + 1], xyzw[4 * i + 2], w);
    return v;
}

int copy_samples(const std::vector<float4>& v, float* out, int cap) {
    if (out)
        for (int i = 0; i < (int)v.size() && i < cap; ++i) {
            out[4 * i] = v[i].x;
            out[4 * i + 1] = v[i].y;
            out[4 * i + 2] = v[i].z;
            out[4 * i + 3] = v[i].w;
        }
    return (int)v.size();
}
}  // namespace

extern "C" {

const char* sph_last_error(void) { return g_last_error.c_str(); }

const char* sph_version(void) { return "sph-gfx950 0.1"; }

int sph_default_settings(sph_settings* out) {
    return guard([&] {
        if (!out) throw Error(SPH_ERR_INVALID, "out is NULL");
        from_settings(Settings{}, out);
        return SPH_OK;
    });
}

int sph_load_settings(const char* path, sph_settings* out) {
    return guard([&] {
        if (!path || !out) throw Error(SPH_ERR_INVALID, "NULL argument");
        from_settings(load_settings(path), out);
        return SPH_OK;
    });
}

sph_handle* sph_create(const sph_settings* settings) {
    sph_handle* h = nullptr;
    int rc = guard([&] {
        validate(settings);
        auto sim = build_simulation(to_settings(settings));
        h = new sph_handle{std::move(sim)};
        return SPH_OK;
    });
    return rc == SPH_OK ? h : nullptr;
}

int sph_destroy(sph_handle* h) {
    return guard([&] {
        delete h;
        return SPH_OK;
    });
}

int sph_reset(sph_handle* h, const sph_settings* settings) {
    return guard([&] {
        validate(settings);
        sim_of(h).reset(to_settings(settings));
        return SPH_OK;
    });
}

int sph_set_rect_limits(sph_handle* h, float width, float height, float depth) {
    return guard([&] {
        sim_of(h).set_rect_limits(width, height, depth);
        return SPH_OK;
    });
}

int sph_add_volume(sph_handle* h, const float* xyzw, int n) {
    return guard([&] {
        auto v = to_f4(xyzw, n, 1.0f);
        sim_of(h).add_volume(std::make_shared<FixedVolume>(std::move(v)));
        return SPH_OK;
    });
}

int sph_add_box_volume(sph_handle* h, const float size[3], const float center[3]) {
    return guard([&] {
        if (!size || !center) throw Error(SPH_ERR_INVALID, "NULL argument");
        sim_of(h).add_volume(std::make_shared<BoxVolume>(size, center));
        return SPH_OK;
    });
}

int sph_add_boundary(sph_handle* h, const float* local_xyzw, int n, int can_move, float mass, const float R[9],
                     const float com[3]) {
    return guard([&] {
        auto v = to_f4(local_xyzw, n, 1.0f);
        return sim_of(h).add_boundary(v, can_move != 0, mass, R, com);
    });
}

int sph_set_body_state(sph_handle* h, int body, const float R[9], const float com[3], const float lin_vel[3],
                       const float ang_vel[3]) {
    return guard([&] {
        sim_of(h).set_body_state(body, R, com, lin_vel, ang_vel);
        return SPH_OK;
    });
}

int sph_get_body_wrench(sph_handle* h, int body, float force[3], float torque[3]) {
    return guard([&] {
        float f[3], t[3];
        sim_of(h).body_wrench(body, f, t);
        for (int k = 0; k < 3; ++k) {
            if (force) force[k] = f[k];
            if (torque) torque[k] = t[k];
        }
        return SPH_OK;
    });
}

int sph_step(sph_handle* h) {
    return guard([&] {
        sim_of(h).simulate();
        return SPH_OK;
    });
}

int sph_step_n(sph_handle* h, int n) {
    return guard([&] {
        if (n < 0) throw Error(SPH_ERR_RANGE, "negative step count");
        auto& s = sim_of(h);
        for (int k = 0; k < n; ++k) s.simulate();
        return SPH_OK;
    });
}

int sph_step_dt(sph_handle* h, float dt) {
    return guard([&] {
        if (!(dt > 0.0f)) throw Error(SPH_ERR_INVALID, "dt must be positive");
        sim_of(h).step_dt(dt);
        return SPH_OK;
    });
}

int sph_synchronize(sph_handle* h) {
    return guard([&] {
        sim_of(h).synchronize();
        return SPH_OK;
    });
}

int sph_particle_count(sph_handle* h) {
    return guard([&] { return sim_of(h).particle_count(); });
}

int sph_boundary_particle_count(sph_handle* h) {
    return guard([&] { return sim_of(h).boundary_particle_count(); });
}

int sph_get_positions(sph_handle* h, float* xyzw, int n) {
    return guard([&] {
        sim_of(h).get_state(0, xyzw, n);
        return SPH_OK;
    });
}
int sph_set_positions(sph_handle* h, const float* xyzw, int n) {
    return guard([&] {
        sim_of(h).set_state(0, xyzw, n);
        return SPH_OK;
    });
}
int sph_get_velocities(sph_handle* h, float* xyzw, int n) {
    return guard([&] {
        sim_of(h).get_state(1, xyzw, n);
        return SPH_OK;
    });
}
int sph_set_velocities(sph_handle* h, const float* xyzw, int n) {
    return guard([&] {
        sim_of(h).set_state(1, xyzw, n);
        return SPH_OK;
    });
}
int sph_get_half_velocities(sph_handle* h, float* xyzw, int n) {
    return guard([&] {
        sim_of(h).get_state(2, xyzw, n);
        return SPH_OK;
    });
}
int sph_set_half_velocities(sph_handle* h, const float* xyzw, int n) {
    return guard([&] {
        sim_of(h).set_state(2, xyzw, n);
        return SPH_OK;
    });
}

int sph_get_field(sph_handle* h, int field, void* out, long long cap) {
    return guard([&] { return (int)sim_of(h).get_field(field, out, cap); });
}

int sph_get_step_stats(sph_handle* h, sph_step_stats* out) {
    return guard([&] {
        if (!out) throw Error(SPH_ERR_INVALID, "out is NULL");
        sim_of(h).stats(*out);
        return SPH_OK;
    });
}

int sph_device_state(sph_handle* h, void** pos, void** vel, void** vel_half) {
    return guard([&] {
        auto& s = sim_of(h);
        if (pos) *pos = s.d_pos.get();
        if (vel) *vel = s.d_vel.get();
        if (vel_half) *vel_half = s.d_vh.get();
        return SPH_OK;
    });
}

void* sph_stream(sph_handle* h) { return (h && h->sim) ? (void*)h->sim->stream() : nullptr; }

int sph_set_profiling(sph_handle* h, int on) {
    return guard([&] {
        sim_of(h).profiler().enabled = on != 0;
        return SPH_OK;
    });
}

int sph_set_profiling_filter(sph_handle* h, const char* kernel) {
    return guard([&] {
        sim_of(h).profiler().filter = kernel ? kernel : "";
        return SPH_OK;
    });
}


int sph_get_kernel_times(sph_handle* h, char* names, int* launches, double* total_ms, int cap) {
    return guard([&] {
        auto& p = sim_of(h).profiler();
        p.resolve();
        int n = (int)p.entries.size();
        for (int k = 0; k < n && k < cap; ++k) {
            if (names) {
                std::strncpy(names + 48 * k, p.entries[k].name.c_str(), 47);
                names[48 * k + 47] = 0;
            }
            if (launches) launches[k] = (int)p.entries[k].launches;
            if (total_ms) total_ms[k] = p.entries[k].total_ms;
        }
        return n;
    });
}

int sph_reset_kernel_times(sph_handle* h) {
    return guard([&] {
        sim_of(h).profiler().reset();
        return SPH_OK;
    });
}

int sph_sample_box_volume(const float size[3], const float center[3], float radius, float* out, int cap) {
    return guard([&] {
        if (!size || !center) throw Error(SPH_ERR_INVALID, "NULL argument");
        return copy_samples(sample_box_volume(size, center, radius), out, cap);
    });
}
int sph_sample_fishtank(float width, float height, float depth, float spacing, float* out, int cap) {
    return guard([&] {
        if (!(spacing > 0.0f)) throw Error(SPH_ERR_INVALID, "spacing must be positive");
        return copy_samples(sample_fishtank(width, height, depth, spacing), out, cap);
    });
}
int sph_sample_cube(float side, float spacing, float* out, int cap) {
    return guard([&] {
        if (!(spacing > 0.0f)) throw Error(SPH_ERR_INVALID, "spacing must be positive");
        return copy_samples(sample_cube(side, spacing), out, cap);
    });
}
int sph_sample_wall(float width, float height, float spacing, float* out, int cap) {
    return guard([&] {
        if (!(spacing > 0.0f)) throw Error(SPH_ERR_INVALID, "spacing must be positive");
        return copy_samples(sample_wall(width, height, spacing), out, cap);
    });
}

}  // extern "C"

// ------------------------------------------------------------ rigid bodies
namespace {
RigidWorld& world_of(sph_rigid_world* w) {
    if (!w || !w->world) throw Error(SPH_ERR_INVALID, "rigid world is NULL");
    return *w->world;
}
}  // namespace

extern "C" {

sph_rigid_world* sph_rigid_create(float gravity_y) {
    sph_rigid_world* out = nullptr;
    int rc = guard([&] {
        if (!std::isfinite(gravity_y)) throw Error(SPH_ERR_INVALID, "gravity must be finite");
        out = new sph_rigid_world();
        out->own.reset(new RigidWorld(gravity_y));
        out->world = out->own.get();
        return SPH_OK;
    });
    return rc == SPH_OK ? out : nullptr;
}

int sph_rigid_destroy(sph_rigid_world* w) {
    return guard([&] {
        if (w && !w->own) throw Error(SPH_ERR_INVALID, "this rigid world belongs to a scene");
        delete w;
        return SPH_OK;
    });
}

int sph_rigid_set_container(sph_rigid_world* w, float width, float height, float depth, const float center[3]) {
    return guard([&] {
        world_of(w).set_container(width, height, depth, center);
        return SPH_OK;
    });
}

int sph_rigid_add_box(sph_rigid_world* w, float mass, const float half[3], const float pos[3], const float quat[4]) {
    return guard([&] { return world_of(w).add_box(mass, half, pos, quat); });
}

int sph_rigid_add_sphere(sph_rigid_world* w, float mass, float radius, const float pos[3]) {
    return guard([&] { return world_of(w).add_sphere(mass, radius, pos); });
}

int sph_rigid_add_static(sph_rigid_world* w, const float pos[3], const float quat[4]) {
    return guard([&] { return world_of(w).add_static(pos, quat); });
}

int sph_rigid_apply_force_n_torque(sph_rigid_world* w, int body, const float f[3], const float t[3]) {
    return guard([&] {
        world_of(w).apply_force_n_torque(body, f, t);
        return SPH_OK;
    });
}

int sph_rigid_step(sph_rigid_world* w, float dt) {
    return guard([&] {
        world_of(w).step(dt);
        return SPH_OK;
    });
}

int sph_rigid_get_state(sph_rigid_world* w, int body, float R[9], float com[3], float quat[4], float lin_vel[3],
                        float ang_vel[3]) {
    return guard([&] {
        RigidWorld& W = world_of(w);
        const RigidBody& b = W.body(body);
        W.transform(body, R, com);
        for (int k = 0; k < 4; ++k)
            if (quat) quat[k] = b.q[k];
        for (int k = 0; k < 3; ++k) {
            if (lin_vel) lin_vel[k] = b.lin[k];
            if (ang_vel) ang_vel[k] = b.ang[k];
        }
        return SPH_OK;
    });
}

int sph_rigid_set_state(sph_rigid_world* w, int body, const float pos[3], const float quat[4], const float lin_vel[3],
                        const float ang_vel[3]) {
    return guard([&] {
        RigidWorld& W = world_of(w);
        W.set_pose(body, pos, quat);
        W.set_velocity(body, lin_vel, ang_vel);
        return SPH_OK;
    });
}

int sph_rigid_body_count(sph_rigid_world* w) {
    return guard([&] { return world_of(w).count(); });
}

int sph_coupled_step(sph_handle* h, sph_rigid_world* w, int npairs, const int* sph_ids, const int* rigid_ids) {
    return guard([&] {
        SimulationBase& S = sim_of(h);
        RigidWorld& W = world_of(w);
        if (npairs < 0 || (npairs > 0 && (!sph_ids || !rigid_ids))) throw Error(SPH_ERR_INVALID, "bad body pairs");
        for (int k = 0; k < npairs; ++k) {
            float R[9], com[3];
            W.transform(rigid_ids[k], R, com);
            const RigidBody& b = W.body(rigid_ids[k]);
            S.set_body_state(sph_ids[k], R, com, b.lin, b.ang);
        }
        S.simulate();
        for (int k = 0; k < npairs; ++k) {
            if (!(W.body(rigid_ids[k]).mass > 0.0f)) continue;
            float f[3], t[3];
            S.body_wrench(sph_ids[k], f, t);
            W.apply_force_n_torque(rigid_ids[k], f, t);
        }
        W.step(S.settings().time_step);
        return SPH_OK;
    });
}

}  // extern "C"

// ------------------------------------------------------------ batched WCSPH
namespace {
BatchWCSPH& batch_of(sph_batch* b) {
    if (!b) throw Error(SPH_ERR_INVALID, "batch is NULL");
    return b->batch;
}

int copy_kernel_times(Profiler& p, char* names, int* launches, double* total_ms, int cap) {
    p.resolve();
    int n = (int)p.entries.size();
    for (int k = 0; k < n && k < cap; ++k) {
        if (names) {
            std::strncpy(names + 48 * k, p.entries[k].name.c_str(), 47);
            names[48 * k + 47] = 0;
        }
        if (launches) launches[k] = (int)p.entries[k].launches;
        if (total_ms) total_ms[k] = p.entries[k].total_ms;
    }
    return n;
}
}  // namespace

extern "C" {

sph_batch* sph_batch_create(void) {
    sph_batch* out = nullptr;
    int rc = guard([&] {
        out = new sph_batch();
        return SPH_OK;
    });
    return rc == SPH_OK ? out : nullptr;
}

int sph_batch_destroy(sph_batch* b) {
    return guard([&] {
        delete b;
        return SPH_OK;
    });
}

int sph_batch_add_scene(sph_batch* b, const sph_settings* s, float width, float height, float depth,
                        const float* xyzw, int n) {
    return guard([&] {
        validate_ieee_only(s, "batch");
        const float c[3] = {width, height, depth};
        return batch_of(b).add_scene(to_settings(s), c, xyzw, n);
    });
}

int sph_batch_finalize(sph_batch* b) {
    return guard([&] {
        batch_of(b).finalize();
        return SPH_OK;
    });
}

int sph_batch_autotune(sph_batch* b, const char* cache_path, int trials) {
    return guard([&] { return batch_of(b).autotune(cache_path, trials); });
}

int sph_batch_get_launch(sph_batch* b, int kernel, int* block, int* grid, int* stage) {
    return guard([&] {
        if (kernel < 0 || kernel >= BatchWCSPH::K_COUNT) throw Error(SPH_ERR_RANGE, "batch: no such kernel");
        const LaunchCfg& c = batch_of(b).config(kernel);
        if (block) *block = c.block;
        if (grid) *grid = c.grid;
        if (stage) *stage = c.stage;
        return SPH_OK;
    });
}

int sph_batch_set_launch(sph_batch* b, int kernel, int block, int grid, int stage) {
    return guard([&] {
        LaunchCfg c;
        c.block = block;
        c.grid = grid;
        c.stage = stage;
        batch_of(b).set_config(kernel, c);
        return SPH_OK;
    });
}

int sph_batch_set_graph(sph_batch* b, int on) {
    return guard([&] {
        batch_of(b).set_graph(on != 0);
        return SPH_OK;
    });
}

int sph_batch_step(sph_batch* b, int nsteps) {
    return guard([&] {
        batch_of(b).step(nsteps);
        return SPH_OK;
    });
}

int sph_batch_synchronize(sph_batch* b) {
    return guard([&] {
        batch_of(b).synchronize();
        return SPH_OK;
    });
}

int sph_batch_scene_count(sph_batch* b) {
    return guard([&] { return batch_of(b).scene_count(); });
}

int sph_batch_particle_count(sph_batch* b, int scene) {
    return guard([&] { return batch_of(b).scene_particles(scene); });
}

int sph_batch_total_particles(sph_batch* b) {
    return guard([&] { return batch_of(b).total_particles(); });
}

int sph_batch_get_positions(sph_batch* b, int scene, float* xyzw, int n) {
    return guard([&] {
        batch_of(b).copy_scene(0, scene, xyzw, n);
        return SPH_OK;
    });
}

int sph_batch_get_velocities(sph_batch* b, int scene, float* xyzw, int n) {
    return guard([&] {
        batch_of(b).copy_scene(1, scene, xyzw, n);
        return SPH_OK;
    });
}

int sph_batch_get_half_velocities(sph_batch* b, int scene, float* xyzw, int n) {
    return guard([&] {
        batch_of(b).copy_scene(2, scene, xyzw, n);
        return SPH_OK;
    });
}

int sph_batch_set_positions(sph_batch* b, int scene, const float* xyzw, int n) {
    return guard([&] {
        batch_of(b).set_scene_state(0, scene, xyzw, n);
        return SPH_OK;
    });
}

int sph_batch_set_velocities(sph_batch* b, int scene, const float* xyzw, int n) {
    return guard([&] {
        batch_of(b).set_scene_state(1, scene, xyzw, n);
        return SPH_OK;
    });
}

int sph_batch_get_density(sph_batch* b, int scene, float* out, int n) {
    return guard([&] {
        batch_of(b).copy_scene(3, scene, out, n);
        return SPH_OK;
    });
}

int sph_batch_get_neighbour_counts(sph_batch* b, int scene, int* out, int n) {
    return guard([&] {
        batch_of(b).copy_scene(4, scene, reinterpret_cast<float*>(out), n);
        return SPH_OK;
    });
}

int sph_batch_set_profiling(sph_batch* b, int on, const char* kernel) {
    return guard([&] {
        Profiler& p = batch_of(b).profiler();
        p.resolve();
        p.reset();
        p.filter = kernel ? kernel : "";
        p.enabled = on != 0;
        return SPH_OK;
    });
}

int sph_batch_get_kernel_times(sph_batch* b, char* names, int* launches, double* total_ms, int cap) {
    return guard([&] { return copy_kernel_times(batch_of(b).profiler(), names, launches, total_ms, cap); });
}

}  // extern "C"

// ------------------------------------------------------------- adjoint
namespace {
WcAdjoint& adjoint_of(sph_adjoint* a) {
    if (!a || !a->adj) throw Error(SPH_ERR_INVALID, "adjoint is NULL");
    return *a->adj;
}
}  // namespace

extern "C" {

sph_adjoint* sph_adjoint_create(const sph_settings* s, float width, float height, float depth, const float* pos,
                                const float* vel, const float* vel_half, int n, int checkpoint_every) {
    sph_adjoint* out = nullptr;
    int rc = guard([&] {
        validate_ieee_only(s, "adjoint");
        const float c[3] = {width, height, depth};
        std::unique_ptr<WcAdjoint> a(new WcAdjoint(to_settings(s), c, pos, vel, vel_half, n, checkpoint_every));
        out = new sph_adjoint{std::move(a)};
        return SPH_OK;
    });
    return rc == SPH_OK ? out : nullptr;
}

int sph_adjoint_destroy(sph_adjoint* a) {
    return guard([&] {
        delete a;
        return SPH_OK;
    });
}

int sph_adjoint_set_initial(sph_adjoint* a, const float* pos, const float* vel, const float* vel_half) {
    return guard([&] {
        adjoint_of(a).set_initial(pos, vel, vel_half);
        return SPH_OK;
    });
}

int sph_adjoint_set_params(sph_adjoint* a, float k_viscosity, float gas_stiffness, float surface_tension) {
    return guard([&] {
        adjoint_of(a).set_params(k_viscosity, gas_stiffness, surface_tension);
        return SPH_OK;
    });
}

int sph_adjoint_forward(sph_adjoint* a, int nsteps) {
    return guard([&] {
        adjoint_of(a).forward(nsteps);
        return SPH_OK;
    });
}

int sph_adjoint_final_state(sph_adjoint* a, float* pos, float* vel, float* vel_half) {
    return guard([&] {
        adjoint_of(a).final_state(pos, vel, vel_half);
        return SPH_OK;
    });
}

int sph_adjoint_backward(sph_adjoint* a, const float* g_pos, const float* g_vel, const float* g_vel_half,
                         float* g_pos0, float* g_vel0, float* g_vel_half0, double* g_params) {
    return guard([&] {
        adjoint_of(a).backward(g_pos, g_vel, g_vel_half, g_pos0, g_vel0, g_vel_half0, g_params);
        return SPH_OK;
    });
}

int sph_adjoint_timings(sph_adjoint* a, double* forward_ms, double* backward_ms) {
    return guard([&] {
        if (forward_ms) *forward_ms = adjoint_of(a).forward_ms();
        if (backward_ms) *backward_ms = adjoint_of(a).backward_ms();
        return SPH_OK;
    });
}

}  // extern "C"

// ------------------------------------------------------------ scene files
namespace sph {
std::vector<float4> sample_sphere(float radius, float spacing);
std::vector<float4> sample_model(const std::string& obj_path, float spacing);
}

namespace {
sph_scene* scene_of(sph_scene* sc) {
    if (!sc || !sc->run) throw Error(SPH_ERR_INVALID, "scene is NULL");
    return sc;
}
}  // namespace

extern "C" {

int sph_sample_sphere(float radius, float spacing, float* out, int cap) {
    return guard([&] { return copy_samples(sample_sphere(radius, spacing), out, cap); });
}

int sph_sample_model(const char* obj_path, float spacing, float* out, int cap) {
    return guard([&] {
        if (!obj_path) throw Error(SPH_ERR_INVALID, "model path is NULL");
        return copy_samples(sample_model(obj_path, spacing), out, cap);
    });
}

sph_scene* sph_scene_load(const char* path, const sph_settings* s, const char* base_dir) {
    sph_scene* out = nullptr;
    int rc = guard([&] {
        validate(s);
        if (!path) throw Error(SPH_ERR_INVALID, "scene path is NULL");
        std::unique_ptr<sph_scene> sc(new sph_scene());
        sc->run.reset(new SceneRun(path, to_settings(s), base_dir ? base_dir : ""));
        sc->handle.sim = std::move(sc->run->sim_owner());
        sc->world.world = &sc->run->world();
        out = sc.release();
        return SPH_OK;
    });
    return rc == SPH_OK ? out : nullptr;
}

int sph_scene_destroy(sph_scene* sc) {
    return guard([&] {
        delete sc;
        return SPH_OK;
    });
}

sph_handle* sph_scene_simulation(sph_scene* sc) {
    sph_handle* out = nullptr;
    guard([&] {
        out = &scene_of(sc)->handle;
        return SPH_OK;
    });
    return out;
}

sph_rigid_world* sph_scene_world(sph_scene* sc) {
    sph_rigid_world* out = nullptr;
    guard([&] {
        out = &scene_of(sc)->world;
        return SPH_OK;
    });
    return out;
}

int sph_scene_step(sph_scene* sc, int nsteps) {
    return guard([&] {
        if (nsteps < 0) throw Error(SPH_ERR_RANGE, "negative step count");
        for (int k = 0; k < nsteps; ++k) scene_of(sc)->run->step();
        return SPH_OK;
    });
}

int sph_scene_body_count(sph_scene* sc) {
    return guard([&] { return (int)scene_of(sc)->run->bodies().size(); });
}

int sph_scene_body(sph_scene* sc, int k, char* type, char* name, int* sph_id, int* rigid_id, float* mass,
                   int* samples) {
    return guard([&] {
        const auto& b = scene_of(sc)->run->bodies();
        if (k < 0 || k >= (int)b.size()) throw Error(SPH_ERR_RANGE, "no scene body " + std::to_string(k));
        const auto& e = b[k];
        if (type) {
            std::strncpy(type, e.body.type.c_str(), 15);
            type[15] = 0;
        }
        if (name) {
            std::strncpy(name, e.body.name.c_str(), 63);
            name[63] = 0;
        }
        if (sph_id) *sph_id = e.sph_id;
        if (rigid_id) *rigid_id = e.rigid_id;
        if (mass) *mass = e.body.mass;
        if (samples) *samples = e.samples;
        return SPH_OK;
    });
}

}  // extern "C"

// ------------------------------------------------------------------ output
extern "C" {

sph_frames* sph_frames_open(sph_handle* h, const char* path, int fields) {
    sph_frames* out = nullptr;
    int rc = guard([&] {
        if (!path) throw Error(SPH_ERR_INVALID, "frames path is NULL");
        std::unique_ptr<FrameWriter> w(new FrameWriter(sim_of(h), path, fields));
        out = new sph_frames{std::move(w)};
        return SPH_OK;
    });
    return rc == SPH_OK ? out : nullptr;
}

int sph_frames_write(sph_frames* f) {
    return guard([&] {
        if (!f || !f->w) throw Error(SPH_ERR_INVALID, "frames is NULL");
        f->w->write();
        return SPH_OK;
    });
}

int sph_frames_close(sph_frames* f) {
    return guard([&] {
        if (!f) return SPH_OK;
        std::unique_ptr<sph_frames> own(f);
        long long n = f->w ? f->w->close() : 0;
        return (int)n;
    });
}

sph_perf_log* sph_perf_open(sph_handle* h, const char* path) {
    sph_perf_log* out = nullptr;
    int rc = guard([&] {
        if (!path) throw Error(SPH_ERR_INVALID, "performance log path is NULL");
        std::unique_ptr<PerfLog> p(new PerfLog(sim_of(h), path));
        out = new sph_perf_log{std::move(p)};
        return SPH_OK;
    });
    return rc == SPH_OK ? out : nullptr;
}

int sph_perf_frame(sph_perf_log* p) {
    return guard([&] {
        if (!p || !p->p) throw Error(SPH_ERR_INVALID, "performance log is NULL");
        p->p->frame();
        return SPH_OK;
    });
}

int sph_perf_close(sph_perf_log* p) {
    return guard([&] {
        long long n = (p && p->p) ? p->p->lines() : 0;
        delete p;
        return (int)n;
    });
}

}  // extern "C"

// ------------------------------------------------------------ autotuning
namespace {
PCISPHSimulation* pcisph_of(sph_handle* h) {
    if (!h) throw Error(SPH_ERR_INVALID, "null handle");
    auto* p = dynamic_cast<PCISPHSimulation*>(h->sim.get());
    if (!p) throw Error(SPH_ERR_STATE, "launch shapes are tuned for PCISPH handles");
    return p;
}
}  // namespace

extern "C" int sph_autotune(sph_handle* h, const char* cache_path, int trials) {
    return guard([&] { return pcisph_of(h)->autotune(cache_path, trials); });
}

extern "C" int sph_get_launch(sph_handle* h, int kernel, int* block, int* grid) {
    return guard([&] {
        auto* p = pcisph_of(h);
        if (kernel < 0 || kernel >= PCISPHSimulation::T_COUNT) throw Error(SPH_ERR_RANGE, "no such tuned kernel");
        Shape sh = p->shape(kernel);
        if (block) *block = sh.block;
        if (grid) *grid = sh.grid;
        return SPH_OK;
    });
}

extern "C" int sph_set_launch(sph_handle* h, int kernel, int block, int grid) {
    return guard([&] {
        pcisph_of(h)->set_shape(kernel, Shape{block, grid});
        return SPH_OK;
    });
}

// ------------------------------------------------------------------ sort
extern "C" int sph_sort_pairs(uint32_t* keys, int32_t* values, int n, int key_bits, void* stream) {
    return guard([&] {
        sort_pairs(keys, values, n, key_bits, stream);
        return SPH_OK;
    });
}

extern "C" int sph_sort_pairs_noalloc(uint32_t* keys, int32_t* values, int n, int key_bits, void* stream) {
    return guard([&] {
        sort_pairs_noalloc(keys, values, n, key_bits, stream);
        return SPH_OK;
    });
}

extern "C" int sph_sort_near(const uint32_t* hint, const uint32_t* keys, uint32_t* sorted_keys, int32_t* perm,
                             int n, void* stream) {
    return guard([&] {
        sort_near(hint, keys, sorted_keys, perm, n, stream);
        return SPH_OK;
    });
}
