// Host side of the gfx950 SPH solver (see solver.h).
#include "solver.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

namespace sph {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(SPH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

float define_literal(float v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.6f", (double)v);
    return std::strtof(buf, nullptr);
}

// ------------------------------------------------------------------ Profiler
hipEvent_t Profiler::get_event() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
}
void Profiler::begin(const char* name, hipStream_t s) {
    active_ = filter.empty() || filter == name;
    if (!active_) return;
    int idx = -1;
    for (size_t k = 0; k < entries.size(); ++k)
        if (entries[k].name == name) idx = (int)k;
    if (idx < 0) {
        entries.push_back({name, 0, 0.0});
        idx = (int)entries.size() - 1;
    }
    current_ = idx;
    cur_a_ = get_event();
    hip_check(hipEventRecord(cur_a_, s), "hipEventRecord");
}
void Profiler::end(hipStream_t s) {
    if (!active_) return;
    active_ = false;
    hipEvent_t b = get_event();
    hip_check(hipEventRecord(b, s), "hipEventRecord");
    pending_.push_back({current_, cur_a_, b});
    if (pending_.size() > 4096) resolve();
}
void Profiler::resolve() {
    for (auto& p : pending_) {
        hip_check(hipEventSynchronize(p.b), "hipEventSynchronize");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
        entries[p.entry].launches += 1;
        entries[p.entry].total_ms += ms;
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pending_.clear();
}
void Profiler::reset() {
    resolve();
    entries.clear();
}
Profiler::~Profiler() {
    for (auto& p : pending_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : pool_) (void)hipEventDestroy(e);
}

// ------------------------------------------------------------------ samplers
// BoxVolume::particles (boxvolume.cpp:11-40): fp32 lattice, order y, x, z.
std::vector<float4> sample_box_volume(const float size[3], const float center[3], float radius) {
    std::vector<float4> out;
    float spacing = 2 * radius;
    int px = (int)std::ceil(size[0] / spacing);
    int py = (int)std::ceil(size[1] / spacing);
    int pz = (int)std::ceil(size[2] / spacing);
    float dx = size[0] / px, dy = size[1] / py, dz = size[2] / pz;
    out.reserve((size_t)px * py * pz);
    for (int j = 0; j < py; ++j)
        for (int i = 0; i < px; ++i)
            for (int k = 0; k < pz; ++k) {
                float4 p;
                p.x = (float)((double)(i * dx + center[0]) - size[0] / 2.0);
                p.y = (float)((double)(j * dy + center[1]) - size[1] / 2.0);
                p.z = (float)((double)(k * dz + center[2]) - size[2] / 2.0);
                p.w = 1.0f;
                out.push_back(p);
            }
    return out;
}

static inline void put(std::vector<float4>& v, float x, float y, float z) { v.push_back(make_float4(x, y, z, 1.0f)); }

// FishTank::surface_sampling (scene/fishtank.cpp:108-141): open-top tank,
// float accumulators.
std::vector<float4> sample_fishtank(float width, float height, float depth, float s) {
    std::vector<float4> v;
    for (float x = -width / 2; x <= width / 2; x += s)
        for (float z = -depth / 2; z <= depth / 2; z += s) put(v, x, -height / 2, z);
    float y0 = -height / 2 + s;
    for (float y = y0; y <= height / 2; y += s) {
        for (float x = -width / 2; x <= width / 2; x += s) {
            put(v, x, y, -depth / 2);
            put(v, x, y, depth / 2);
        }
        for (float z = -depth / 2; z <= depth / 2; z += s) {
            put(v, width / 2, y, z);
            put(v, -width / 2, y, z);
        }
    }
    return v;
}

// Cube::surface_sampling (scene/cube.cpp:136-170): faces shrunk by one spacing.
std::vector<float4> sample_cube(float side, float s) {
    std::vector<float4> v;
    float width = side - s, height = side - s, depth = side - s;
    for (float x = -width / 2; x <= width / 2; x += s)
        for (float z = -depth / 2; z <= depth / 2; z += s) {
            put(v, x, -height / 2, z);
            put(v, x, height / 2, z);
        }
    float y0 = -height / 2 + s;
    for (float y = y0; y <= height / 2; y += s) {
        for (float x = -width / 2; x <= width / 2; x += s) {
            put(v, x, y, -depth / 2);
            put(v, x, y, depth / 2);
        }
        for (float z = -depth / 2; z <= depth / 2; z += s) {
            put(v, width / 2, y, z);
            put(v, -width / 2, y, z);
        }
    }
    return v;
}

// Wall::surface_sampling (scene/wall.cpp:142-152)
std::vector<float4> sample_wall(float width, float height, float s) {
    std::vector<float4> v;
    for (float x = -width / 2; x <= width / 2; x += s)
        for (float y = -height / 2; y <= height / 2; y += s) put(v, x, y, 0.0f);
    return v;
}

BoxVolume::BoxVolume(const float size[3], const float center[3]) {
    for (int k = 0; k < 3; ++k) {
        size_[k] = size[k];
        center_[k] = center[k];
    }
}
std::vector<float4> BoxVolume::particles(float r) const { return sample_box_volume(size_, center_, r); }

// ---------------------------------------------------------------------- Grid
void Grid::reset(float h) {
    cell_size = h;
    cps = (int)std::ceil(size / cell_size);  // grid.cpp:64 (fp32 ceil)
    if (cps < 1 || cps > 1024)
        throw Error(SPH_ERR_INVALID, "support radius gives " + std::to_string(cps) +
                                         " cells per side; the 10-bit Morton key needs 1..1024");
    int b = 0;
    while ((1 << b) < cps) ++b;
    key_bits = 3 * b;
    have_box = false;
    x0 = z0 = 0;
    x1 = z1 = -1;
    ++version;
    update_dev();
}

void Grid::cell_of(float4 p, int& cx, int& cy, int& cz) const {
    float half = size / 2.0f;
    cx = (int)std::floor((p.x + half) / cell_size);
    cy = (int)std::floor((p.y + half) / cell_size);
    cz = (int)std::floor((p.z + half) / cell_size);
    cx = (cx + cps) % cps;
    cy = (cy + cps) % cps;
    cz = (cz + cps) % cps;
}

bool Grid::include(const std::vector<float4>& pts, int margin) {
    int nx0 = have_box ? x0 : cps, nx1 = have_box ? x1 : -1;
    int nz0 = have_box ? z0 : cps, nz1 = have_box ? z1 : -1;
    bool any = have_box;
    for (const auto& p : pts) {
        if (!(std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z))) continue;
        int cx, cy, cz;
        cell_of(p, cx, cy, cz);
        if (cx < 0 || cz < 0) continue;
        nx0 = std::min(nx0, std::max(0, cx - margin));
        nx1 = std::max(nx1, std::min(cps - 1, cx + margin));
        nz0 = std::min(nz0, std::max(0, cz - margin));
        nz1 = std::max(nz1, std::min(cps - 1, cz + margin));
        any = true;
    }
    if (!any) return false;
    bool changed = !have_box || nx0 != x0 || nx1 != x1 || nz0 != z0 || nz1 != z1;
    x0 = nx0; x1 = nx1; z0 = nz0; z1 = nz1;
    have_box = true;
    if (changed) ++version;
    update_dev();
    return changed;
}

void Grid::update_dev() {
    dev.half = size / 2.0f;
    dev.cell = cell_size;
    float lit = define_literal(cell_size);
    dev.sqr_support = lit * lit;  // grid.cl:3 (CELL_SIZE*CELL_SIZE)
    dev.cps = cps;
    dev.cells = (long long)std::pow((double)cps, 3.0);
    dev.bx0 = x0;
    dev.bz0 = z0;
    dev.bnx = have_box ? x1 - x0 + 1 : 0;
    dev.bnz = have_box ? z1 - z0 + 1 : 0;
}

// ------------------------------------------------------------ BoundaryHandler
static void mat_apply(const Body& b, const float4& p, float4& o) {
    // transform_particles: dot(row, p) + com (boundary.cl:48-51)
    o.x = ((b.R[0] * p.x + b.R[1] * p.y) + b.R[2] * p.z) + b.com[0];
    o.y = ((b.R[3] * p.x + b.R[4] * p.y) + b.R[5] * p.z) + b.com[1];
    o.z = ((b.R[6] * p.x + b.R[7] * p.y) + b.R[8] * p.z) + b.com[2];
    o.w = 1.0f;
}

std::vector<float4> BoundaryHandler::transformed_samples() const {
    std::vector<float4> out(raw_.size());
    for (const auto& b : bodies_)
        for (int i = 0; i < b.count; ++i) mat_apply(b, raw_[b.origin + i], out[b.origin + i]);
    return out;
}

int BoundaryHandler::add_boundary(const std::vector<float4>& local, bool can_move, float mass, const float R[9],
                                  const float com[3]) {
    Body b;
    b.origin = (int)raw_.size();
    b.count = (int)local.size();
    b.can_move = can_move;
    b.mass = mass;
    if (R) std::memcpy(b.R, R, sizeof b.R);
    if (com) std::memcpy(b.com, com, sizeof b.com);
    for (const auto& p : local) raw_.push_back(make_float4(p.x, p.y, p.z, 1.0f));
    bodies_.push_back(b);
    resample();
    return (int)bodies_.size() - 1;
}

void BoundaryHandler::set_body_state(int id, const float* R, const float* com, const float* lin, const float* ang) {
    if (id < 0 || id >= (int)bodies_.size()) throw Error(SPH_ERR_RANGE, "no such body");
    Body& b = bodies_[id];
    if (R) std::memcpy(b.R, R, sizeof b.R);
    if (com) std::memcpy(b.com, com, sizeof b.com);
    if (lin) std::memcpy(b.lin, lin, sizeof b.lin);
    if (ang) std::memcpy(b.ang, ang, sizeof b.ang);
    if (b.can_move) {
        // keep the range box large enough for the body's new placement
        std::vector<float4> pts(b.count);
        for (int i = 0; i < b.count; ++i) mat_apply(b, raw_[b.origin + i], pts[i]);
        if (sim_.grid_.include(pts, 2)) {
            sim_.d_ranges.release();
            d_ranges.release();
        }
    }
}

void BoundaryHandler::realloc_ranges() {
    size_t cells = sim_.grid_.box_cells();
    if (d_ranges.size() != cells || !d_ranges.get() || box_version != sim_.grid_.version) {
        d_ranges.release();
        d_ranges.resize(cells);
        d_near.resize(cells);
        box_version = sim_.grid_.version;
        sorts = 0;
        if (cells) hip_check(hipMemsetAsync(d_ranges.get(), 0, cells * sizeof(int2), sim_.stream_), "hipMemsetAsync");
    }
}

// _sample_surfaces -> _alloc_buffers -> _compute_boundary_phi -> sync(true)
// (boundaryhandler.cpp:119-153, 155-214, 238-249)
void BoundaryHandler::resample() {
    hipStream_t s = sim_.stream_;
    const Settings& S = sim_.s_;
    int B = count();
    if (B > MAX_SCENE_PARTICLES)
        throw Error(SPH_ERR_RANGE, "at most " + std::to_string(MAX_SCENE_PARTICLES) + " boundary particles per scene");
    moving_ = 0;
    for (auto& b : bodies_) moving_ += b.can_move ? 1 : 0;
    d_raw.resize(B);
    d_unsorted.resize(B);
    d_pos_s.resize(B);
    d_force.resize(B);
    d_torque.resize(B);
    d_phi.resize(B);
    d_phi_s.resize(B);
    d_keys_a.resize(B);
    d_keys_b.resize(B);
    d_vals_a.resize(B);
    d_vals_b.resize(B);
    d_hist.resize(sort_hist_words(B));
    d_ranges.release();
    sorts = 0;
    if (B == 0) return;
    hip_check(hipMemcpyAsync(d_raw.get(), raw_.data(), B * sizeof(float4), hipMemcpyHostToDevice, s), "upload");
    hip_check(hipMemsetAsync(d_force.get(), 0, B * sizeof(float4), s), "memset");
    hip_check(hipMemsetAsync(d_torque.get(), 0, B * sizeof(float4), s), "memset");
    // psi from the raw samples of each body (boundary.cl:4-31)
    float h = S.fluid_support_radius;
    float h_d = define_literal(h);
    float w_eval = (float)(365.0f / (64.0f * M_PI * ::pow((double)h, 9.0)));  // MullerConstants::default_eval
    for (auto& b : bodies_)
        launch_phi(b.count, d_raw.get() + b.origin, h_d, w_eval, S.rest_density, 1.0f, d_phi.get() + b.origin, s,
                   &sim_.prof_);
    // per-body partial-sum layout of the fluid wrench
    std::vector<int> offs, nblk;
    int total = 0;
    for (auto& b : bodies_) {
        offs.push_back(total);
        int nb = fluid_force_blocks(b.count);
        nblk.push_back(nb);
        total += nb;
    }
    d_partial.resize((size_t)std::max(total, 1) * 8);
    d_wrench.resize((size_t)bodies_.size() * 8);
    d_offsets.resize(bodies_.size());
    d_nblocks.resize(bodies_.size());
    hip_check(hipMemcpyAsync(d_offsets.get(), offs.data(), offs.size() * sizeof(int), hipMemcpyHostToDevice, s),
              "upload");
    hip_check(hipMemcpyAsync(d_nblocks.get(), nblk.data(), nblk.size() * sizeof(int), hipMemcpyHostToDevice, s),
              "upload");
    hip_check(hipMemsetAsync(d_wrench.get(), 0, bodies_.size() * 8 * sizeof(float), s), "memset");
    hip_check(hipStreamSynchronize(s), "sync");  // raw_ / offs host storage
    wrench_valid_ = false;
    if (sim_.grid_.include(transformed_samples(), 2)) {
        sim_.d_ranges.release();
        d_ranges.release();
    }
    sync(true);
}

void BoundaryHandler::sort_all() {
    hipStream_t s = sim_.stream_;
    int B = count();
    Grid& g = sim_.grid_;
    realloc_ranges();
    int passes = sort_passes(g.key_bits);
    uint32_t* hash_out = (passes % 2) ? d_keys_b.get() : d_keys_a.get();
    // the previous sort's keys (in d_keys_a) name the cells to clear; none on
    // the first sort after an allocation, whose box is freshly zeroed
    launch_hash(B, d_unsorted.get(), g.dev, hash_out, d_keys_a.get(), sorts > 0 ? d_ranges.get() : nullptr,
                sim_.d_state.get(), s, &sim_.prof_);
    launch_radix_sort(d_keys_a.get(), d_keys_b.get(), d_vals_a.get(), d_vals_b.get(), d_hist.get(), B, g.key_bits,
                      s, &sim_.prof_);
    launch_gather_boundary(B, d_vals_a.get(), d_keys_a.get(), d_unsorted.get(), d_phi.get(), d_pos_s.get(),
                           d_phi_s.get(), d_ranges.get(), g.dev, s, &sim_.prof_);
    launch_mark_boundary(B, d_pos_s.get(), d_keys_a.get(), g.dev, d_near.get(), d_near.size(), s, &sim_.prof_);
    ++sorts;
}

// BoundaryHandler::sync (boundaryhandler.cpp:251-294). The reference re-sorts
// every boundary particle every step; the sort of an unchanged sample set
// gives the same order, so the sort is redone only when a body moved.
void BoundaryHandler::sync(bool all) {
    int B = count();
    if (B == 0) return;
    hipStream_t s = sim_.stream_;
    bool any = false;
    for (auto& b : bodies_) {
        if (b.can_move || all) {
            launch_transform(b.count, d_raw.get() + b.origin, d_unsorted.get() + b.origin, b.R, b.com, s, &sim_.prof_);
            any = true;
        }
    }
    if (any || !d_ranges.get() || sorts == 0) sort_all();
}

// apply_fluid_forces (boundaryhandler.cpp:296-349) for bodies with mass > 0.
void BoundaryHandler::apply_fluid_forces() {
    int B = count();
    if (B == 0) return;
    hipStream_t s = sim_.stream_;
    const Settings& S = sim_.s_;
    float h = S.fluid_support_radius;
    float H = define_literal(h);
    float spiky = (float)(-45.0f / (M_PI * std::pow((double)h, 6)));
    float mass = fluid_mass(S);
    std::vector<int> offs;
    int total = 0;
    bool any = false;
    for (auto& b : bodies_) {
        offs.push_back(total);
        total += fluid_force_blocks(b.count);
    }
    for (size_t k = 0; k < bodies_.size(); ++k) {
        const Body& b = bodies_[k];
        if (!(b.mass > 0.0f)) continue;
        any = true;
        launch_fluid_force(b.count, b.origin, d_unsorted.get(), d_phi.get(), sim_.n_, sim_.d_prho.get(),
                           sim_.d_pressure.get(), sim_.d_ranges.get(), sim_.grid_.dev, H, spiky, mass, b.com,
                           d_force.get(), d_torque.get(),
                           d_partial.get() + (size_t)offs[k] * 8, sim_.d_state.get(), s, &sim_.prof_);
    }
    if (any) launch_wrench_reduce((int)bodies_.size(), d_offsets.get(), d_nblocks.get(), d_partial.get(),
                                  d_wrench.get(), s, &sim_.prof_);
    wrench_valid_ = false;
}

void BoundaryHandler::wrench(int id, float f[3], float t[3]) {
    if (id < 0 || id >= (int)bodies_.size()) throw Error(SPH_ERR_RANGE, "no such body");
    if (!wrench_valid_) {
        wrench_host_.assign(bodies_.size() * 8, 0.0f);
        if (!bodies_.empty()) {
            hip_check(hipMemcpyAsync(wrench_host_.data(), d_wrench.get(), bodies_.size() * 8 * sizeof(float),
                                     hipMemcpyDeviceToHost, sim_.stream_),
                      "download");
            hip_check(hipStreamSynchronize(sim_.stream_), "sync");
        }
        wrench_valid_ = true;
    }
    bool active = bodies_[id].mass > 0.0f;
    for (int k = 0; k < 3; ++k) {
        f[k] = active ? wrench_host_[id * 8 + k] : 0.0f;
        t[k] = active ? wrench_host_[id * 8 + 3 + k] : 0.0f;
    }
}

// ------------------------------------------------------------ SimulationBase
SimulationBase::SimulationBase(const Settings& s) : s_(s), boundary_(*this) {
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    d_state.resize(1);
    d_acc.resize(10);
    hip_check(hipMemsetAsync(d_state.get(), 0, sizeof(DevState), stream_), "memset");
    grid_.reset(s_.fluid_support_radius);  // Grid(5.0, h)
}

SimulationBase::~SimulationBase() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    prof_.reset();
    if (stream_) (void)hipStreamDestroy(stream_);
}

float fluid_mass(const Settings& s) {
    if (s.method == SPH_PCISPH) return s.rest_density * powf(2.0f * s.fluid_particle_radius, 3.0f) / 1.15f;
    return s.rest_density * powf(2.0f * s.fluid_particle_radius, 3.0f) / 1.0f;
}

void SimulationBase::set_rect_limits(float w, float h, float d) {
    container[0] = w;
    container[1] = h;
    container[2] = d;
    have_container = true;
    std::vector<float4> corners;
    for (int sx = -1; sx <= 1; sx += 2)
        for (int sz = -1; sz <= 1; sz += 2) corners.push_back(make_float4(sx * w / 2, 0.0f, sz * d / 2, 1.0f));
    ensure_box(corners);
}

int SimulationBase::add_boundary(const std::vector<float4>& local, bool can_move, float mass, const float R[9],
                                 const float com[3]) {
    return boundary_.add_boundary(local, can_move, mass, R, com);
}

void SimulationBase::add_volume(std::shared_ptr<FluidVolume> v) {
    if (s_.method == SPH_PCISPH && !have_container)
        throw Error(SPH_ERR_STATE, "set_rect_limits must precede add_volume: the PCISPH relaxation step clamps to it");
    volumes_.push_back(std::move(v));
    initialize_solver();
}

void SimulationBase::reset(const Settings& s) {
    if (s.method != s_.method) throw Error(SPH_ERR_INVALID, "reset cannot change the method");
    s_ = s;
    grid_.reset(s_.fluid_support_radius);
    if (have_container) set_rect_limits(container[0], container[1], container[2]);
    d_ranges.release();
    boundary_.d_ranges.release();
    if (boundary_.count() > 0) boundary_.resample();
    initialize_solver();
}

void SimulationBase::alloc_fluid(int n) {
    if (n > MAX_SCENE_PARTICLES)
        throw Error(SPH_ERR_RANGE, "at most " + std::to_string(MAX_SCENE_PARTICLES) + " fluid particles per scene");
    n_ = n;
    size_t N = (size_t)std::max(n, 1);
    d_pos.resize(N); d_vel.resize(N); d_pos_s.resize(N); d_vel_s.resize(N);
    d_normal.resize(N); d_f_other.resize(N); d_f_press.resize(N); d_pos_pred.resize(N);
    d_prho.resize(N); d_pc.resize(N);
    fresh_ranges_ = true;
    d_density.resize(N); d_pressure.resize(N); d_err.resize(N);
    d_keys_a.resize(N); d_keys_b.resize(N); d_vals_a.resize(N); d_vals_b.resize(N);
    d_hist.resize(sort_hist_words(n));
    sorted_hint_ = false;
    d_resort.resize(resort_bytes(n));
    resort_ = resort_carve(d_resort.get(), n);
    resort_clear(resort_, n, stream_);
    d_nl.resize(N * LIST_CAP); d_nc.resize(N); d_bnl.resize(N * LIST_CAP); d_bnc.resize(N);
    if (s_.method == SPH_WCSPH) {
        d_vh.resize(N); d_vh_s.resize(N); d_accel.resize(N);
    }
    hip_check(hipMemsetAsync(d_nc.get(), 0, N * sizeof(int), stream_), "memset");
    hip_check(hipMemsetAsync(d_bnc.get(), 0, N * sizeof(int), stream_), "memset");
    hip_check(hipMemsetAsync(d_f_press.get(), 0, N * sizeof(float4), stream_), "memset");
    hip_check(hipMemsetAsync(d_pressure.get(), 0, N * sizeof(float), stream_), "memset");
    hip_check(hipMemsetAsync(d_err.get(), 0, N * sizeof(float), stream_), "memset");
}

void SimulationBase::ensure_box(const std::vector<float4>& pts) {
    if (grid_.include(pts, 2)) {
        d_ranges.release();
        boundary_.d_ranges.release();
    }
}

// (Re)allocates the range boxes after the box grew; called before each step.
void SimulationBase::ensure_ranges() {
    size_t cells = grid_.box_cells();
    if (!d_ranges.get() || d_ranges.size() != cells || box_version_ != grid_.version) {
        d_ranges.release();
        d_ranges.resize(cells);
        if (cells) hip_check(hipMemsetAsync(d_ranges.get(), 0, cells * sizeof(int2), stream_), "memset");
        fresh_ranges_ = true;
        box_version_ = grid_.version;
    }
    if (boundary_.count() > 0 &&
        (!boundary_.d_ranges.get() || boundary_.box_version != grid_.version))
        boundary_.sync(true);
}

// _initialize_solver (pcisphsimluation.cpp:532-551, wcsphsimluation.cpp:393-405)
void SimulationBase::initialize_solver() {
    grid_.reset(s_.fluid_support_radius);
    if (have_container) {
        std::vector<float4> corners;
        for (int sx = -1; sx <= 1; sx += 2)
            for (int sz = -1; sz <= 1; sz += 2)
                corners.push_back(make_float4(sx * container[0] / 2, 0.0f, sz * container[2] / 2, 1.0f));
        grid_.include(corners, 2);
    }
    if (boundary_.count() > 0) grid_.include(boundary_.transformed_samples(), 2);
    d_ranges.release();
    boundary_.d_ranges.release();
    initialize_params();
    std::vector<float4> particles;
    for (auto& v : volumes_) {
        auto p = v->particles(s_.fluid_particle_radius);
        particles.insert(particles.end(), p.begin(), p.end());
    }
    for (auto& p : particles) p.w = 1.0f;
    alloc_fluid((int)particles.size());
    ensure_box(particles);
    ensure_ranges();
    if (n_ > 0) {
        hip_check(hipMemcpyAsync(d_pos.get(), particles.data(), n_ * sizeof(float4), hipMemcpyHostToDevice, stream_),
                  "upload");
        hip_check(hipMemsetAsync(d_vel.get(), 0, n_ * sizeof(float4), stream_), "memset");
        if (d_vh.get()) hip_check(hipMemsetAsync(d_vh.get(), 0, n_ * sizeof(float4), stream_), "memset");
        hip_check(hipStreamSynchronize(stream_), "sync");
    }
    after_initialize();
}

void SimulationBase::require_ready() const {
    if (!have_container) throw Error(SPH_ERR_STATE, "set_rect_limits must be called before stepping");
}

// hash -> stable sort -> gather + ranges (pcisphsimluation.cpp:79-97)
void SimulationBase::neighbour_search() {
    // the previous step's sorted keys (d_keys_a) name the cells to clear
    int2* clear = fresh_ranges_ ? nullptr : d_ranges.get();
    if (sorted_hint_ && !(s_.flags & SPH_FULL_SORT)) {
        // particles are stored in the previous step's sorted order: only the
        // keys that changed cell are placed; the result is the same sort
        launch_hash_resort(n_, d_pos.get(), grid_.dev, d_keys_b.get(), d_keys_a.get(), d_vals_a.get(), clear,
                           d_state.get(), resort_, stream_, &prof_);
    } else {
        const int passes = sort_passes(grid_.key_bits);
        uint32_t* hash_out = (passes % 2) ? d_keys_b.get() : d_keys_a.get();
        launch_hash(n_, d_pos.get(), grid_.dev, hash_out, d_keys_a.get(), clear, d_state.get(), stream_, &prof_);
        launch_radix_sort(d_keys_a.get(), d_keys_b.get(), d_vals_a.get(), d_vals_b.get(), d_hist.get(), n_,
                          grid_.key_bits, stream_, &prof_);
        if (!(s_.flags & SPH_FULL_SORT)) launch_resort_samples(d_keys_a.get(), n_, resort_, stream_);
    }
    fresh_ranges_ = false;
    sorted_hint_ = n_ > 0;
    launch_gather_fluid(n_, d_vals_a.get(), d_keys_a.get(), d_pos.get(), d_vel.get(),
                        s_.method == SPH_WCSPH ? d_vh.get() : nullptr, d_pos_s.get(), d_vel_s.get(),
                        s_.method == SPH_WCSPH ? d_vh_s.get() : nullptr, d_ranges.get(), grid_.dev, stream_, &prof_);
}

void SimulationBase::step_dt(float dt) {
    if (dt > 0.0f && dt != s_.time_step) {
        s_.time_step = dt;
        initialize_params();
    }
    simulate();
}

void SimulationBase::synchronize() { hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

void SimulationBase::get_state(int which, float* xyzw, int n) {
    if (n != n_) throw Error(SPH_ERR_RANGE, "size mismatch: " + std::to_string(n) + " vs " + std::to_string(n_));
    if (which == 2 && s_.method != SPH_WCSPH) throw Error(SPH_ERR_STATE, "buffer not available for this method");
    if (n == 0) return;  // a scene without fluid particles has no buffers yet
    DeviceArray<float4>* src = which == 0 ? &d_pos : which == 1 ? &d_vel : &d_vh;
    if (!src->get()) throw Error(SPH_ERR_STATE, "buffer not available for this method");
    hip_check(hipMemcpyAsync(xyzw, src->get(), n * sizeof(float4), hipMemcpyDeviceToHost, stream_), "download");
    synchronize();
}

void SimulationBase::set_state(int which, const float* xyzw, int n) {
    if (n != n_) throw Error(SPH_ERR_RANGE, "size mismatch: " + std::to_string(n) + " vs " + std::to_string(n_));
    if (which == 2 && s_.method != SPH_WCSPH) throw Error(SPH_ERR_STATE, "buffer not available for this method");
    if (n == 0) return;  // a scene without fluid particles has no buffers yet
    DeviceArray<float4>* dst = which == 0 ? &d_pos : which == 1 ? &d_vel : &d_vh;
    if (!dst->get()) throw Error(SPH_ERR_STATE, "buffer not available for this method");
    std::vector<float4> v(n);
    for (int i = 0; i < n; ++i)
        v[i] = make_float4(xyzw[4 * i], xyzw[4 * i + 1], xyzw[4 * i + 2], which == 0 ? 1.0f : 0.0f);
    if (which == 0) {
        ensure_box(v);
        // positions set from outside may lie anywhere: the next step sorts
        // every key instead of placing the moved ones into the last order
        sorted_hint_ = false;
    }
    hip_check(hipMemcpyAsync(dst->get(), v.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream_), "upload");
    synchronize();
}

template <class T> static long long copy_out(const DeviceArray<T>& a, size_t n, void* out, long long cap,
                                             hipStream_t s) {
    if (!out || n == 0) return (long long)n;
    if ((long long)n > cap) throw Error(SPH_ERR_RANGE, "output buffer too small");
    if (!a.get()) throw Error(SPH_ERR_STATE, "field not available");
    if (n) hip_check(hipMemcpyAsync(out, a.get(), n * sizeof(T), hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "sync");
    return (long long)n;
}

long long SimulationBase::get_field(int field, void* out, long long cap) {
    size_t n = (size_t)n_;
    size_t B = (size_t)boundary_.count();
    hipStream_t s = stream_;
    switch (field) {
        case SPH_FIELD_KEYS: return copy_out(d_keys_a, n, out, cap, s);
        case SPH_FIELD_PERM: return copy_out(d_vals_a, n, out, cap, s);
        case SPH_FIELD_NCOUNT: return copy_out(d_nc, n, out, cap, s);
        case SPH_FIELD_BNCOUNT: return copy_out(d_bnc, n, out, cap, s);
        case SPH_FIELD_NLIST:
        case SPH_FIELD_BNLIST: {
            size_t m = n * LIST_CAP;
            if (!out) return (long long)m;
            if ((long long)m > cap) throw Error(SPH_ERR_RANGE, "output buffer too small");
            const auto& L = field == SPH_FIELD_NLIST ? d_nl : d_bnl;
            const auto& Cn = field == SPH_FIELD_NLIST ? d_nc : d_bnc;
            std::vector<int> cnt(n);
            if (n) {
                hip_check(hipMemcpyAsync(out, L.get(), m * sizeof(int), hipMemcpyDeviceToHost, s), "download");
                hip_check(hipMemcpyAsync(cnt.data(), Cn.get(), n * sizeof(int), hipMemcpyDeviceToHost, s), "download");
            }
            hip_check(hipStreamSynchronize(s), "sync");
            int* o = (int*)out;
            for (size_t k = 0; k < (size_t)LIST_CAP; ++k)
                for (size_t i = 0; i < n; ++i)
                    if ((int)k >= cnt[i]) o[k * n + i] = -1;
            return (long long)m;
        }
        case SPH_FIELD_POS_SORTED: return copy_out(d_pos_s, n, out, cap, s);
        case SPH_FIELD_VEL_SORTED: return copy_out(d_vel_s, n, out, cap, s);
        case SPH_FIELD_VEL_HALF_SORTED: return copy_out(d_vh_s, n, out, cap, s);
        case SPH_FIELD_DENSITY: return copy_out(d_density, n, out, cap, s);
        case SPH_FIELD_PRESSURE: return copy_out(d_pressure, n, out, cap, s);
        case SPH_FIELD_ERR: return copy_out(d_err, n, out, cap, s);
        case SPH_FIELD_NORMAL: return copy_out(d_normal, n, out, cap, s);
        case SPH_FIELD_FORCE_OTHER: return copy_out(d_f_other, n, out, cap, s);
        case SPH_FIELD_FORCE_PRESSURE: return copy_out(d_f_press, n, out, cap, s);
        case SPH_FIELD_POS_PRED: return copy_out(d_pos_pred, n, out, cap, s);
        case SPH_FIELD_ACCEL: return copy_out(d_accel, n, out, cap, s);
        case SPH_FIELD_POS: return copy_out(d_pos, n, out, cap, s);
        case SPH_FIELD_VEL: return copy_out(d_vel, n, out, cap, s);
        case SPH_FIELD_VEL_HALF: return copy_out(d_vh, n, out, cap, s);
        case SPH_FIELD_BPOS_SORTED: return copy_out(boundary_.d_pos_s, B, out, cap, s);
        case SPH_FIELD_BPHI_SORTED: return copy_out(boundary_.d_phi_s, B, out, cap, s);
        case SPH_FIELD_BKEYS: return copy_out(boundary_.d_keys_a, B, out, cap, s);
        case SPH_FIELD_BPERM: return copy_out(boundary_.d_vals_a, B, out, cap, s);
        case SPH_FIELD_BPHI: return copy_out(boundary_.d_phi, B, out, cap, s);
        case SPH_FIELD_BFORCE: return copy_out(boundary_.d_force, B, out, cap, s);
        case SPH_FIELD_BTORQUE: return copy_out(boundary_.d_torque, B, out, cap, s);
        case SPH_FIELD_RANGES:
        case SPH_FIELD_BRANGES: {
            // (reference cell id, start, end) of the occupied cells, read
            // back from the device range box at the cells of the sorted keys
            bool fl = field == SPH_FIELD_RANGES;
            size_t m = fl ? n : B;
            const auto& K = fl ? d_keys_a : boundary_.d_keys_a;
            const auto& P = fl ? d_pos_s : boundary_.d_pos_s;
            const auto& R = fl ? d_ranges : boundary_.d_ranges;
            std::vector<uint32_t> keys(m);
            std::vector<float4> pos(m);
            std::vector<int2> box(R.size());
            if (m) {
                hip_check(hipMemcpyAsync(keys.data(), K.get(), m * 4, hipMemcpyDeviceToHost, s), "download");
                hip_check(hipMemcpyAsync(pos.data(), P.get(), m * 16, hipMemcpyDeviceToHost, s), "download");
                hip_check(hipMemcpyAsync(box.data(), R.get(), R.size() * sizeof(int2), hipMemcpyDeviceToHost, s),
                          "download");
            }
            hip_check(hipStreamSynchronize(s), "sync");
            std::vector<int> tri;
            for (size_t i = 0; i < m; ++i) {
                if (i > 0 && keys[i - 1] == keys[i]) continue;
                int cx, cy, cz;
                grid_.cell_of(pos[i], cx, cy, cz);
                long long id = (long long)grid_.cps * ((long long)grid_.cps * cz + cy) + cx;
                int lx = cx - grid_.x0, lz = cz - grid_.z0;
                int2 e = make_int2(0, 0);
                if (lx >= 0 && lx < grid_.dev.bnx && lz >= 0 && lz < grid_.dev.bnz)
                    e = box[((size_t)lz * grid_.cps + cy) * grid_.dev.bnx + lx];
                tri.push_back((int)id);
                tri.push_back(e.x);
                tri.push_back(e.y);
            }
            long long cnt = (long long)tri.size() / 3;
            if (!out) return cnt;
            if ((long long)tri.size() > cap * 3) throw Error(SPH_ERR_RANGE, "output buffer too small");
            std::memcpy(out, tri.data(), tri.size() * sizeof(int));
            return cnt;
        }
        default: throw Error(SPH_ERR_INVALID, "unknown field id " + std::to_string(field));
    }
}

void SimulationBase::stats(sph_step_stats& o) {
    std::memset(&o, 0, sizeof o);
    DevState st{};
    hip_check(hipMemcpyAsync(&st, d_state.get(), sizeof st, hipMemcpyDeviceToHost, stream_), "download");
    unsigned long long acc[10] = {0};
    launch_stats(n_, d_nc.get(), d_bnc.get(), d_pos.get(), d_acc.get(), stream_);
    hip_check(hipMemcpyAsync(acc, d_acc.get(), sizeof acc, hipMemcpyDeviceToHost, stream_), "download");
    hip_check(hipStreamSynchronize(stream_), "sync");
    o.iterations = st.iterations;
    o.edge_reads = st.edge_reads;
    o.out_of_grid = st.out_of_grid;
    o.fluid_neighbours = (long long)acc[0];
    o.boundary_neighbours = (long long)acc[1];
    o.fluid_list_overflows = (int)acc[2];
    o.boundary_list_overflows = (int)acc[3];
    o.nan_positions = (int)acc[4];
    int it = st.iterations > 0 ? st.iterations : 1;
    uint32_t bits = 0;
    for (int l = 0; l < ERR_LANES; ++l) bits = std::max(bits, st.maxerr[it & (MAXERR_SLOTS - 1)][l]);
    float mx;
    std::memcpy(&mx, &bits, 4);
    o.max_density_error = s_.method == SPH_PCISPH ? mx : 0.0f;
}

// ----------------------------------------------------------------- PCISPH

PCISPHSimulation::PCISPHSimulation(const Settings& s) : SimulationBase(s) { initialize_params(); }

// _initialize_params + _deduce_density_scale_factor
// (pcisphsimluation.cpp:230-255, 429-469)
void PCISPHSimulation::initialize_params() {
    float r = s_.fluid_particle_radius, h = s_.fluid_support_radius;
    mass_ = s_.rest_density * powf(2.0f * r, 3.0f) / 1.15f;
    st_main_ = (float)(32.0f / (M_PI * (double)powf(h, 9.0f)));
    st_term_ = powf(h, 6.0f) / 64.0f;
    deduce_density_scale_factor();
    fill_params();
}

static float muller_default_eval(float h) { return (float)(365.0f / (64.0f * M_PI * ::pow((double)h, 9.0))); }
static float muller_default_grad(float h) { return (float)(-945.0f / (32.0f * M_PI * ::pow((double)h, 9.0))); }
static float muller_pressure_grad(float h) { return (float)(-45.0f / (M_PI * ::pow((double)h, 6.0))); }
static float muller_viscosity_lapl(float h) { return (float)(45.0f / (M_PI * ::pow((double)h, 6.0))); }

void PCISPHSimulation::deduce_density_scale_factor() {
    float h = s_.fluid_support_radius, r = s_.fluid_particle_radius;
    float spiky = muller_pressure_grad(h);
    float poly6g = muller_default_grad(h);
    float beta = 2.0f * powf((s_.time_step * mass_) / s_.rest_density, 2.0f);
    float h2 = h * h;
    float gsum[3] = {0.f, 0.f, 0.f}, ssum[3] = {0.f, 0.f, 0.f};
    float gdots = 0.f;
    for (float x = -h - r; x <= h + r; x += 2.0f * r)
        for (float y = -h - r; y <= h + r; y += 2.0f * r)
            for (float z = -h - r; z <= h + r; z += 2.0f * r) {
                float r2 = (x * x + y * y + z * z);
                float rn = sqrtf(r2);
                if (r2 < h2) {
                    float sfac = (float)((double)(spiky * (h - rn) * (h - rn)) * (1.0 / (double)rn));
                    float gfac = poly6g * (h2 - r2) * (h2 - r2);
                    float sv[3] = {sfac * x, sfac * y, sfac * z};
                    float gv[3] = {gfac * x, gfac * y, gfac * z};
                    for (int k = 0; k < 3; ++k) {
                        gsum[k] += gv[k];
                        ssum[k] += sv[k];
                        gdots += gv[k] * sv[k];
                    }
                }
            }
    float sdot = 0.f;
    for (int k = 0; k < 3; ++k) sdot += ssum[k] * gsum[k];
    delta_ = -1.0f / (beta * (-sdot - gdots));
}

void PCISPHSimulation::fill_params() {
    float h = s_.fluid_support_radius;
    PciParams& P = P_;
    P.H = define_literal(h);
    P.H2 = P.H * P.H;
    P.M = define_literal(mass_);
    P.rho0_d = define_literal(s_.rest_density);
    P.w_eval = muller_default_eval(h);
    P.w_grad = muller_default_grad(h);
    P.visc_lapl = muller_viscosity_lapl(h);
    P.press_grad = muller_pressure_grad(h);
    P.st_main = st_main_;
    P.st_term = st_term_;
    P.k_visc = s_.k_viscosity;
    P.sigma = s_.surface_tension;
    P.delta = delta_;
    P.dt = s_.time_step;
    P.max_vel = s_.max_vel;
    P.gx = 0.0f;
    P.gy = -s_.gravity;
    P.gz = 0.0f;
    P.lo_x = -container[0] / 2.0f;
    P.hi_x = container[0] / 2.0f;
    P.lo_y = -container[1] / 2.0f;
    P.lo_z = -container[2] / 2.0f;
    P.hi_z = container[2] / 2.0f;
    P.rho0_host = s_.rest_density;
    P.eta = s_.pcisph_error_ratio;
    P.min_iter = 3;
    P.max_iter = s_.pcisph_max_iterations;
    P.fast = (s_.flags & SPH_FAST_MATH) ? 1 : 0;
}

// relaxation step and velocity reset (pcisphsimluation.cpp:548-550)
void PCISPHSimulation::after_initialize() {
    if (n_ > 0) {
        simulate_step(3, 10000);
        hip_check(hipMemsetAsync(d_vel.get(), 0, n_ * sizeof(float4), stream_), "memset");
        hip_check(hipStreamSynchronize(stream_), "sync");
    }
}

// _simulate_pcisph_step (pcisphsimluation.cpp:72-165)
void PCISPHSimulation::simulate_step(int min_iter, int max_iter) {
    require_ready();
    ensure_ranges();
    fill_params();
    P_.min_iter = min_iter;
    P_.max_iter = max_iter;
    hipStream_t s = stream_;
    launch_step_begin(d_state.get(), s);
    boundary_.sync(false);
    if (n_ == 0) return;
    neighbour_search();
    const int nb = boundary_.count();
    const int has_b = nb > 0 ? 1 : 0;
    DensityParams D{};
    D.H2 = P_.H2;
    D.M = P_.M;
    D.w_eval = P_.w_eval;
    launch_lists_density(false, n_, d_pos_s.get(), d_ranges.get(), nb, boundary_.d_pos_s.get(),
                         boundary_.d_phi_s.get(), boundary_.d_ranges.get(), boundary_.d_near.get(), grid_.dev, D,
                         d_nl.get(), d_nc.get(), d_bnl.get(), d_bnc.get(), d_density.get(), nullptr, d_prho.get(),
                         d_state.get(), s, &prof_);
    launch_normals(n_, d_prho.get(), d_nl.get(), d_nc.get(), P_.H2, P_.w_grad, P_.M, P_.H, d_normal.get(), s,
                   &prof_, P_.fast != 0, shapes_[T_NORMALS]);
    if (max_iter <= 0) {
        launch_fill_u32(d_pressure.get(), n_, 0u, s);
        launch_fill_u32(d_f_press.get(), (size_t)n_ * 4, 0u, s);
        launch_forces_predict(n_, d_prho.get(), d_vel_s.get(), d_normal.get(), d_nl.get(), d_nc.get(), P_,
                              d_f_other.get(), d_pos_pred.get(), s, &prof_, shapes_[T_FORCES]);
        boundary_.apply_fluid_forces();
        launch_pci_integrate_only(n_, d_pos_s.get(), d_vel_s.get(), d_f_other.get(), P_, d_pos.get(), d_vel.get(), s,
                                  &prof_);
        return;
    }
    launch_forces_predict(n_, d_prho.get(), d_vel_s.get(), d_normal.get(), d_nl.get(), d_nc.get(), P_,
                          d_f_other.get(), d_pos_pred.get(), s, &prof_, shapes_[T_FORCES]);
    // The corrections that always run get one launch per phase on full grids;
    // the rest of the loop, up to max_iter, on small grids that return at once
    // after the device-side exit test ended the loop, so a step needs no host
    // round trip. A cap above TAIL_CHUNK further corrections (the relaxation
    // step) is enqueued TAIL_CHUNK corrections at a time, and the host reads
    // the exit flag between chunks.
    const int always = min_iter < max_iter ? min_iter : max_iter;
    for (int i = 1; i <= always; ++i) {
        launch_update_pressure(n_, i, d_pos_pred.get(), d_prho.get(), d_nl.get(), d_nc.get(), d_bnl.get(),
                               d_bnc.get(), boundary_.d_pos_s.get(), boundary_.d_phi_s.get(), has_b, P_, d_err.get(),
                               d_pressure.get(), d_pc.get(), d_state.get(), s, &prof_, shapes_[T_UPDATE]);
        launch_pressure_force(n_, i, d_pc.get(), d_vel_s.get(), d_f_other.get(), d_nl.get(), d_nc.get(), d_bnl.get(),
                              d_bnc.get(), boundary_.d_pos_s.get(), boundary_.d_phi_s.get(), has_b, P_,
                              d_f_press.get(), d_pos_pred.get(), d_pos.get(), d_vel.get(), d_state.get(), s, &prof_,
                              shapes_[T_PFORCE]);
    }
    constexpr int TAIL_CHUNK = 8;
    for (int first = always + 1; first <= max_iter; first += TAIL_CHUNK) {
        if (first > always + 1) {
            int done = 0;
            hip_check(hipMemcpyAsync(&done, &d_state.get()->done_iter, sizeof done, hipMemcpyDeviceToHost, s),
                      "download");
            hip_check(hipStreamSynchronize(s), "sync");
            if (done != 0) break;
        }
        const int last = std::min(max_iter, first + TAIL_CHUNK - 1);
        launch_corrections(n_, first, last, d_prho.get(), d_vel_s.get(), d_f_other.get(), d_nl.get(), d_nc.get(),
                           d_bnl.get(), d_bnc.get(), boundary_.d_pos_s.get(), boundary_.d_phi_s.get(), has_b, P_,
                           d_err.get(), d_pressure.get(), d_pc.get(), d_f_press.get(), d_pos_pred.get(), d_pos.get(),
                           d_vel.get(), d_state.get(), s, &prof_);
    }
    boundary_.apply_fluid_forces();
}

// ------------------------------------------------------------ autotuning
const char* PCISPHSimulation::tuned_name(int k) {
    static const char* names[T_COUNT] = {"normals", "forces_predict", "update_pressure", "pressure_force"};
    return (k >= 0 && k < T_COUNT) ? names[k] : "?";
}

void PCISPHSimulation::set_shape(int k, Shape sh) {
    if (k < 0 || k >= T_COUNT) throw Error(SPH_ERR_RANGE, "no such tuned kernel");
    if (sh.block != 64 && sh.block != 128 && sh.block != 256)
        throw Error(SPH_ERR_INVALID, "threads per block must be 64, 128 or 256");
    if (sh.grid < 0 || ((k == T_NORMALS || k == T_FORCES) && sh.grid != 0))
        throw Error(SPH_ERR_INVALID, "a fixed grid applies to update_pressure and pressure_force only");
    shapes_[k] = sh;
}

std::string PCISPHSimulation::tune_key() const {
    hipDeviceProp_t prop{};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::string arch = prop.gcnArchName;
    for (char& c : arch)
        if (c == ' ') c = '_';
    return "pcisph " + arch + " cu" + std::to_string(prop.multiProcessorCount) + " n" + std::to_string(n_) + " b" +
           std::to_string(boundary_.count());
}

int PCISPHSimulation::autotune(const char* cache_path, int trials) {
    require_ready();
    if (n_ == 0) return 0;
    const std::string key = tune_key();
    if (cache_path && *cache_path) {
        std::ifstream f(cache_path);
        std::string line;
        Shape found[T_COUNT];
        int seen = 0;
        while (std::getline(f, line)) {
            if (line.compare(0, key.size() + 1, key + " ") != 0) continue;
            std::istringstream ss(line.substr(key.size() + 1));
            std::string name;
            Shape sh;
            if (!(ss >> name >> sh.block >> sh.grid)) continue;
            for (int k = 0; k < T_COUNT; ++k)
                if (name == tuned_name(k)) {
                    found[k] = sh;
                    seen |= 1 << k;
                }
        }
        if (seen == (1 << T_COUNT) - 1) {
            for (int k = 0; k < T_COUNT; ++k) set_shape(k, found[k]);
            return 1;
        }
    }
    if (trials < 1) trials = 1;
    hipDeviceProp_t prop{};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    const int cus = std::max(prop.multiProcessorCount, 1);
    // the tuning steps advance the fluid; the caller's steps continue from
    // the state before them (a step depends on positions and velocities only)
    synchronize();
    std::vector<float4> save_p(n_), save_v(n_);
    get_state(0, &save_p[0].x, n_);
    get_state(1, &save_v[0].x, n_);
    const bool prof_was = prof_.enabled;
    const std::string filter_was = prof_.filter;
    const int blocks[3] = {64, 128, 256};
    for (int k = 0; k < T_COUNT; ++k) {
        std::vector<Shape> cands;
        for (int b : blocks) {
            cands.push_back(Shape{b, 0});
            if (k == T_UPDATE || k == T_PFORCE) {
                cands.push_back(Shape{b, 4 * cus});
                cands.push_back(Shape{b, 16 * cus});
            }
        }
        Shape best = shapes_[k];
        double best_ms = 1e30;
        for (const Shape& c : cands) {
            shapes_[k] = c;
            prof_.enabled = false;
            simulate_step(3, s_.pcisph_max_iterations);  // warm
            prof_.reset();
            prof_.filter = tuned_name(k);
            prof_.enabled = true;
            for (int t = 0; t < trials; ++t) simulate_step(3, s_.pcisph_max_iterations);
            synchronize();
            prof_.resolve();
            prof_.enabled = false;
            for (auto& e : prof_.entries)
                if (e.name == tuned_name(k) && e.launches > 0 && e.total_ms / e.launches < best_ms) {
                    best_ms = e.total_ms / e.launches;
                    best = c;
                }
        }
        shapes_[k] = best;
    }
    prof_.reset();
    prof_.filter = filter_was;
    prof_.enabled = prof_was;
    set_state(0, &save_p[0].x, n_);
    set_state(1, &save_v[0].x, n_);
    synchronize();
    if (cache_path && *cache_path) {
        std::ofstream f(cache_path, std::ios::app);
        if (!f.good()) throw Error(SPH_ERR_INVALID, "autotune: cannot write the cache " + std::string(cache_path));
        for (int k = 0; k < T_COUNT; ++k)
            f << key << " " << tuned_name(k) << " " << shapes_[k].block << " " << shapes_[k].grid << "\n";
    }
    return 0;
}

// ------------------------------------------------------------------ WCSPH
WCSPHSimulation::WCSPHSimulation(const Settings& s) : SimulationBase(s) { initialize_params(); }

// _initialize_params (wcsphsimluation.cpp:165-202)
void WCSPHSimulation::initialize_params() {
    float r = s_.fluid_particle_radius;
    mass_ = s_.rest_density * powf(2.0f * r, 3.0f) / 1.0f;
}

// _initialize_params (wcsphsimluation.cpp:165-202, 188-199) and the kernel
// arguments of simulate (wcsphsimluation.cpp:216-282), shared by the
// single-scene solver and the batched one so both pass the same bits.
WcStepParams wcsph_step_params(const Settings& s, const float container[3]) {
    WcStepParams o{};
    const float r = s.fluid_particle_radius, h = s.fluid_support_radius;
    const float mass = s.rest_density * powf(2.0f * r, 3.0f) / 1.0f;
    const float poly6_eval = (float)(315.0 / (64.0 * M_PI * std::pow((double)h, 9)));
    const float poly6_grad = (float)(-945.0 / (32.0 * M_PI * std::pow((double)h, 9)));
    const float spiky_grad = (float)(-45.0 / (M_PI * std::pow((double)h, 6)));
    const float visc_lapl = (float)(45.0 / (M_PI * std::pow((double)h, 6)));
    const float st_main = (float)(32.0f / (M_PI * (double)powf(h, 9.0f)));
    const float st_term = powf(h, 6.0f) / 64.0f;
    o.D.h = h;
    o.D.M = mass;
    o.D.w_eval = poly6_eval;
    o.D.rho0 = s.rest_density;
    o.D.k_gas = s.gas_stiffness;
    o.D.fix_density = (s.flags & SPH_FIX_WCSPH_DENSITY) ? 1 : 0;
    // compute_normals with the Q8 repair: PARTICLE_MASS is the WCSPH mass
    // printed like every other -D float constant
    const float H = define_literal(h);
    o.nH2 = H * H;
    o.nw_grad = poly6_grad;
    o.nM = define_literal(mass);
    o.nH = H;
    WcParams& P = o.P;
    P.h = h;
    P.M = mass;
    P.spiky_grad = spiky_grad;
    P.visc_lapl = visc_lapl;
    P.k_visc = s.k_viscosity;
    P.sigma = s.surface_tension;
    P.rho0 = s.rest_density;
    P.st_main = st_main;
    P.st_term = st_term;
    P.dt = s.time_step;
    P.max_vel = s.max_vel;
    P.gx = 0.0f;
    P.gy = -s.gravity;
    P.gz = 0.0f;
    P.lo_x = -container[0] / 2.0f;
    P.hi_x = container[0] / 2.0f;
    P.lo_y = -container[1] / 2.0f;
    P.lo_z = -container[2] / 2.0f;
    P.hi_z = container[2] / 2.0f;
    P.fast = (s.flags & SPH_FAST_MATH) ? 1 : 0;
    return o;
}

// WCSPHSimulation::simulate (wcsphsimluation.cpp:40-106)
void WCSPHSimulation::simulate() {
    require_ready();
    ++frames_done;
    sim_time += s_.time_step;
    ensure_ranges();
    hipStream_t s = stream_;
    launch_step_begin(d_state.get(), s);
    boundary_.sync(false);
    if (n_ == 0) return;
    neighbour_search();
    const int nb = boundary_.count();
    const WcStepParams W = wcsph_step_params(s_, container);
    launch_lists_density(true, n_, d_pos_s.get(), d_ranges.get(), nb, boundary_.d_pos_s.get(),
                         boundary_.d_phi_s.get(), boundary_.d_ranges.get(), boundary_.d_near.get(), grid_.dev, W.D,
                         d_nl.get(), d_nc.get(), d_bnl.get(), d_bnc.get(), d_density.get(), d_pressure.get(),
                         d_prho.get(), d_state.get(), s, &prof_);
    launch_normals(n_, d_prho.get(), d_nl.get(), d_nc.get(), W.nH2, W.nw_grad, W.nM, W.nH, d_normal.get(), s,
                   &prof_, W.P.fast != 0);
    boundary_.apply_fluid_forces();
    launch_accel_integrate(n_, d_prho.get(), d_vel_s.get(), d_vh_s.get(), d_pressure.get(), d_normal.get(),
                           d_nl.get(), d_nc.get(), d_bnl.get(), d_bnc.get(), boundary_.d_pos_s.get(),
                           boundary_.d_phi_s.get(), W.P, d_accel.get(), d_pos.get(), d_vel.get(), d_vh.get(), s,
                           &prof_);
}

std::unique_ptr<SimulationBase> build_simulation(const Settings& s) {
    switch (s.method) {
        case SPH_WCSPH: return std::unique_ptr<SimulationBase>(new WCSPHSimulation(s));
        case SPH_PCISPH: return std::unique_ptr<SimulationBase>(new PCISPHSimulation(s));
        default: throw Error(SPH_ERR_INVALID, "Unknown simulation method!");
    }
}

// ------------------------------------------------------------------ settings
// FileParser (settings/fileparser.cpp:7-50): for every line holding '=', the
// first whitespace-delimited token on each side; later keys overwrite earlier.
static std::string first_token(const std::string& s) {
    std::stringstream ss;
    ss << s;
    std::string out;
    ss >> out;
    return out;
}

Settings load_settings(const std::string& path) {
    std::ifstream f(path);
    if (!f.good()) throw Error(SPH_ERR_PARSE, "Configuration file '" + path + "' could not be parsed.");
    std::map<std::string, std::string> opt;
    while (f.good()) {
        std::string line;
        std::getline(f, line);
        auto eq = line.find_first_of('=');
        if (eq != std::string::npos) opt[first_token(line.substr(0, eq))] = first_token(line.substr(eq + 1));
    }
    auto get = [&](const char* k) -> const std::string& {
        auto it = opt.find(k);
        if (it == opt.end())
            throw Error(SPH_ERR_PARSE, std::string("Option '") + k + "' was not found in the configuration file");
        return it->second;
    };
    Settings s;
    // Settings::load (settings/settings.cpp:23-57): atof/atoi, then the enums
    s.time_step = (float)std::atof(get("time_step").c_str());
    s.max_vel = (float)std::atof(get("max_vel").c_str());
    s.fluid_particle_radius = (float)std::atof(get("fluid_particle_radius").c_str());
    s.fluid_support_radius = (float)std::atof(get("fluid_support_radius").c_str());
    s.pcisph_max_iterations = std::atoi(get("pcisph_max_iterations").c_str());
    s.pcisph_error_ratio = (float)std::atof(get("pcisph_error_ratio").c_str());
    const std::string& m = get("method");
    if (m == "wcsph") s.method = SPH_WCSPH;
    else if (m == "pcisph") s.method = SPH_PCISPH;
    else throw Error(SPH_ERR_PARSE, "Unknown simulation method '" + m + "'!");
    s.rest_density = (float)std::atof(get("rest_density").c_str());
    s.k_viscosity = (float)std::atof(get("k_viscosity").c_str());
    s.gravity = (float)std::atof(get("gravity").c_str());
    s.gas_stiffness = (float)std::atof(get("gas_stiffness").c_str());
    s.surface_tension = (float)std::atof(get("surface_tension").c_str());
    // render_method fills GraphicsSettings (settings.cpp:45-53), which the
    // solver does not use: validated when present, not required (config.wc
    // has no such key and would otherwise not load)
    auto rmi = opt.find("render_method");
    if (rmi != opt.end() && rmi->second != "particles" && rmi->second != "screenspace")
        throw Error(SPH_ERR_PARSE, "Unknown rendering method '" + rmi->second + "'!");
    s.flags = 0;
    return s;
}

}  // namespace sph
