// Host side of batched WCSPH (batch.h): scene packing, the step sequence,
// its capture into a HIP graph, and the launch-shape autotuner.
#include "batch_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace sph {

BatchWCSPH::BatchWCSPH() { hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate"); }

BatchWCSPH::~BatchWCSPH() {
    if (exec_) hipGraphExecDestroy(exec_);
    if (graph_) hipGraphDestroy(graph_);
    if (stream_) hipStreamDestroy(stream_);
}

int BatchWCSPH::add_scene(const Settings& s, const float container[3], const float* xyzw, int n) {
    if (finalized_) throw Error(SPH_ERR_STATE, "batch: scenes are added before the batch is finalised");
    if (s.method != SPH_WCSPH) throw Error(SPH_ERR_INVALID, "batch: only WCSPH scenes can be batched");
    if (n < 0 || (n > 0 && !xyzw)) throw Error(SPH_ERR_INVALID, "batch: bad particle array");
    for (int k = 0; k < 3; ++k)
        if (!(container[k] > 0.0f)) throw Error(SPH_ERR_INVALID, "batch: container sides must be > 0");
    if ((int)scenes_.size() >= 1024) throw Error(SPH_ERR_RANGE, "batch: at most 1024 scenes");
    long long total = 0;
    for (auto& h : host_) total += h.n;
    if (total + n > (long long)INT32_MAX / LIST_CAP) throw Error(SPH_ERR_RANGE, "batch: too many particles");
    HostScene h;
    h.s = s;
    for (int k = 0; k < 3; ++k) h.container[k] = container[k];
    h.n = n;
    h.pos.resize(n);
    for (int i = 0; i < n; ++i) {
        h.pos[i] = make_float4(xyzw[4 * i], xyzw[4 * i + 1], xyzw[4 * i + 2], 1.0f);  // w = 1 (Q21)
        if (!(std::isfinite(h.pos[i].x) && std::isfinite(h.pos[i].y) && std::isfinite(h.pos[i].z)))
            throw Error(SPH_ERR_INVALID, "batch: non-finite particle position");
    }
    host_.push_back(std::move(h));
    scenes_.push_back(BatchScene{});
    return (int)scenes_.size() - 1;
}

void BatchWCSPH::finalize() {
    if (finalized_) return;
    if (host_.empty()) throw Error(SPH_ERR_STATE, "batch: no scenes");
    long long roff = 0;
    int lo = 0;
    key_bits_ = 1;
    std::vector<float4> pos;
    std::vector<int> sid;
    for (size_t k = 0; k < host_.size(); ++k) {
        HostScene& h = host_[k];
        // the grid a single-scene run builds: cell size h, box over the
        // container's corners and the particles, 2 cells of margin
        Grid g;
        g.reset(h.s.fluid_support_radius);
        std::vector<float4> corners;
        for (int sx = -1; sx <= 1; sx += 2)
            for (int sz = -1; sz <= 1; sz += 2)
                corners.push_back(make_float4(sx * h.container[0] / 2, 0.0f, sz * h.container[2] / 2, 1.0f));
        g.include(corners, 2);
        g.include(h.pos, 2);
        key_bits_ = std::max(key_bits_, g.key_bits);
        BatchScene& S = scenes_[k];
        S.g = g.dev;
        const WcStepParams W = wcsph_step_params(h.s, h.container);
        S.D = W.D;
        S.P = W.P;
        S.nH2 = W.nH2;
        S.nw_grad = W.nw_grad;
        S.nM = W.nM;
        S.nH = W.nH;
        S.roff = roff;
        S.lo = lo;
        S.hi = lo + h.n;
        roff += (long long)g.box_cells();
        lo += h.n;
        pos.insert(pos.end(), h.pos.begin(), h.pos.end());
        sid.insert(sid.end(), (size_t)h.n, (int)k);
    }
    n_ = lo;
    sid_bits_ = 1;
    while ((1 << sid_bits_) < (int)scenes_.size()) ++sid_bits_;
    sid_bits_ = std::max(sid_bits_, 8);
    const size_t N = (size_t)std::max(n_, 1);
    d_pos.resize(N); d_vel.resize(N); d_vh.resize(N);
    d_pos_s.resize(N); d_vel_s.resize(N); d_vh_s.resize(N);
    d_prho.resize(N); d_normal.resize(N); d_accel.resize(N);
    d_density.resize(N); d_pressure.resize(N);
    d_keys_a.resize(N); d_keys_b.resize(N); d_vals_a.resize(N); d_vals_b.resize(N);
    d_hist.resize(sort_hist_words(n_));
    d_nl.resize(N * LIST_CAP); d_nc.resize(N); d_bnc.resize(N);
    d_sid.resize(N);
    d_ranges.resize((size_t)std::max(roff, 1LL));
    d_scenes.resize(scenes_.size());
    d_state.resize(1);
    d_resort.resize(resort_bytes(n_));
    resort_ = resort_carve(d_resort.get(), n_);
    resort_clear(resort_, n_, stream_);
    sorted_hint_ = false;
    hip_check(hipMemsetAsync(d_ranges.get(), 0, d_ranges.size() * sizeof(int2), stream_), "memset");
    hip_check(hipMemsetAsync(d_bnc.get(), 0, N * sizeof(int), stream_), "memset");
    hip_check(hipMemsetAsync(d_keys_a.get(), 0, N * sizeof(uint32_t), stream_), "memset");
    if (n_ > 0) {
        hip_check(hipMemcpyAsync(d_pos.get(), pos.data(), n_ * sizeof(float4), hipMemcpyHostToDevice, stream_),
                  "upload");
        hip_check(hipMemcpyAsync(d_sid.get(), sid.data(), n_ * sizeof(int), hipMemcpyHostToDevice, stream_), "upload");
        hip_check(hipMemsetAsync(d_vel.get(), 0, n_ * sizeof(float4), stream_), "memset");
        hip_check(hipMemsetAsync(d_vh.get(), 0, n_ * sizeof(float4), stream_), "memset");
    }
    hip_check(hipMemcpyAsync(d_scenes.get(), scenes_.data(), scenes_.size() * sizeof(BatchScene),
                             hipMemcpyHostToDevice, stream_),
              "upload");
    hip_check(hipStreamSynchronize(stream_), "sync");
    for (auto& h : host_) std::vector<float4>().swap(h.pos);
    fresh_ = true;
    finalized_ = true;
}

void BatchWCSPH::require_final() const {
    if (!finalized_) throw Error(SPH_ERR_STATE, "batch: finalise the batch first");
}

void BatchWCSPH::enqueue_step(bool clear) {
    hipStream_t s = stream_;
    Profiler* prof = &prof_;
    launch_step_begin(d_state.get(), s);
    if (sorted_hint_) {
        // every scene keeps its segment and its particles are stored in the
        // last step's order: only the keys that changed cell are placed
        launch_batch_hash_resort(n_, d_sid.get(), d_scenes.get(), d_pos.get(), d_keys_b.get(), d_keys_a.get(),
                                 d_vals_a.get(), clear ? d_ranges.get() : nullptr, d_state.get(), resort_, s, prof);
    } else {
        const int passes = sort_passes(key_bits_, sid_bits_);
        uint32_t* hash_out = (passes % 2) ? d_keys_b.get() : d_keys_a.get();
        launch_batch_hash(n_, d_sid.get(), d_scenes.get(), d_pos.get(), hash_out, d_keys_a.get(), d_ranges.get(),
                          clear, d_state.get(), cfg_[K_HASH], s, prof);
        launch_radix_sort(d_keys_a.get(), d_keys_b.get(), d_vals_a.get(), d_vals_b.get(), d_hist.get(), n_, key_bits_,
                          s, prof, d_sid.get(), sid_bits_);
    }
    launch_batch_gather(n_, d_sid.get(), d_scenes.get(), d_vals_a.get(), d_keys_a.get(), d_pos.get(), d_vel.get(),
                        d_vh.get(), d_pos_s.get(), d_vel_s.get(), d_vh_s.get(), d_ranges.get(), cfg_[K_GATHER], s,
                        prof);
    launch_batch_lists_density(n_, d_sid.get(), d_scenes.get(), d_pos_s.get(), d_ranges.get(), d_nl.get(),
                               d_nc.get(), d_bnc.get(), d_density.get(), d_pressure.get(), d_prho.get(),
                               d_state.get(), cfg_[K_LISTS], s, prof);
    launch_batch_normals(n_, d_sid.get(), d_scenes.get(), d_prho.get(), d_nl.get(), d_nc.get(), d_normal.get(),
                         cfg_[K_NORMALS], s, prof);
    launch_batch_accel_integrate(n_, d_sid.get(), d_scenes.get(), d_prho.get(), d_vel_s.get(), d_vh_s.get(),
                                 d_pressure.get(), d_normal.get(), d_nl.get(), d_nc.get(), d_bnc.get(),
                                 d_accel.get(), d_pos.get(), d_vel.get(), d_vh.get(), cfg_[K_ACCEL], s, prof);
}

void BatchWCSPH::drop_graph() {
    if (exec_) hipGraphExecDestroy(exec_);
    if (graph_) hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
}

void BatchWCSPH::step(int nsteps) {
    require_final();
    if (nsteps < 0) throw Error(SPH_ERR_RANGE, "batch: negative step count");
    for (int k = 0; k < nsteps; ++k) {
        if (n_ == 0) return;
        // the first step has no previous keys to clear ranges by; profiling
        // records events around each launch, which a graph cannot hold
        if (fresh_ || !sorted_hint_ || !use_graph_ || prof_.enabled) {
            enqueue_step(!fresh_);
            fresh_ = false;
            sorted_hint_ = true;
            continue;
        }
        if (!exec_) {
            hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                enqueue_step(true);
            } catch (...) {
                hipGraph_t g = nullptr;
                hipStreamEndCapture(stream_, &g);
                if (g) hipGraphDestroy(g);
                throw;
            }
            hip_check(hipStreamEndCapture(stream_, &graph_), "hipStreamEndCapture");
            hip_check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
        }
        hip_check(hipGraphLaunch(exec_, stream_), "hipGraphLaunch");
    }
}

void BatchWCSPH::synchronize() { hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

void BatchWCSPH::set_graph(bool on) {
    use_graph_ = on;
    if (!on) drop_graph();
}

void BatchWCSPH::set_config(int kernel, const LaunchCfg& c) {
    if (kernel < 0 || kernel >= K_COUNT) throw Error(SPH_ERR_RANGE, "batch: no such kernel");
    if (c.block != 64 && c.block != 128 && c.block != 256) throw Error(SPH_ERR_INVALID, "batch: block must be 64/128/256");
    if (c.grid < 0) throw Error(SPH_ERR_INVALID, "batch: grid must be >= 0");
    if (c.stage != 0 && c.stage != 24) throw Error(SPH_ERR_INVALID, "batch: stage must be 0 or 24");
    cfg_[kernel] = c;
    drop_graph();
}

const char* BatchWCSPH::kernel_name(int k) {
    static const char* names[K_COUNT] = {"batch_hash", "batch_gather_ranges", "batch_lists_density",
                                         "batch_normals", "batch_accel_integrate"};
    return (k >= 0 && k < K_COUNT) ? names[k] : "";
}

void BatchWCSPH::check_scene(int scene) const {
    if (scene < 0 || scene >= (int)scenes_.size()) throw Error(SPH_ERR_RANGE, "batch: no scene " + std::to_string(scene));
}

int BatchWCSPH::scene_particles(int scene) const {
    check_scene(scene);
    return finalized_ ? scenes_[scene].hi - scenes_[scene].lo : host_[scene].n;
}

void BatchWCSPH::copy_scene(int which, int scene, float* out, int n) {
    require_final();
    check_scene(scene);
    const BatchScene& S = scenes_[scene];
    if (n != S.hi - S.lo) throw Error(SPH_ERR_RANGE, "batch: size mismatch");
    if (n == 0) return;
    const float4* src = which == 0 ? d_pos.get() : which == 1 ? d_vel.get() : d_vh.get();
    if (which == 3) {  // density of the last step, sorted order
        hip_check(hipMemcpyAsync(out, d_density.get() + S.lo, n * sizeof(float), hipMemcpyDeviceToHost, stream_),
                  "download");
    } else if (which == 4) {  // neighbour counts, sorted order
        hip_check(hipMemcpyAsync(out, d_nc.get() + S.lo, n * sizeof(int), hipMemcpyDeviceToHost, stream_),
                  "download");
    } else {
        hip_check(hipMemcpyAsync(out, src + S.lo, n * sizeof(float4), hipMemcpyDeviceToHost, stream_), "download");
    }
    synchronize();
}

void BatchWCSPH::set_scene_state(int which, int scene, const float* in, int n) {
    require_final();
    check_scene(scene);
    const BatchScene& S = scenes_[scene];
    if (n != S.hi - S.lo) throw Error(SPH_ERR_RANGE, "batch: size mismatch");
    if (n == 0) return;
    std::vector<float4> v(n);
    for (int i = 0; i < n; ++i)
        v[i] = make_float4(in[4 * i], in[4 * i + 1], in[4 * i + 2], which == 0 ? 1.0f : 0.0f);
    float4* dst = which == 0 ? d_pos.get() : which == 1 ? d_vel.get() : d_vh.get();
    hip_check(hipMemcpyAsync(dst + S.lo, v.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream_), "upload");
    synchronize();
    if (which == 0) sorted_hint_ = false;  // new positions: the next step sorts every key
}

long long BatchWCSPH::range_cells() const { return (long long)d_ranges.size(); }

// ----------------------------------------------------------------- autotune
std::string BatchWCSPH::tune_key() const {
    hipDeviceProp_t prop{};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::ostringstream k;
    k << prop.gcnArchName << "/cu" << prop.multiProcessorCount << "/n" << n_ << "/s" << scenes_.size();
    std::string key = k.str();
    std::replace(key.begin(), key.end(), ' ', '_');
    return key;
}

bool BatchWCSPH::load_tuning(const std::string& path) {
    std::ifstream f(path);
    if (!f.good()) return false;
    const std::string key = tune_key();
    std::string line;
    LaunchCfg found[K_COUNT];
    int seen = 0;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string k, kernel;
        LaunchCfg c;
        if (!(in >> k >> kernel >> c.block >> c.grid >> c.stage)) continue;
        if (k != key) continue;
        for (int i = 0; i < K_COUNT; ++i)
            if (kernel == kernel_name(i)) {
                found[i] = c;
                seen |= 1 << i;
            }
    }
    if (seen != (1 << K_COUNT) - 1) return false;
    for (int i = 0; i < K_COUNT; ++i) set_config(i, found[i]);
    return true;
}

void BatchWCSPH::save_tuning(const std::string& path) const {
    std::ofstream f(path, std::ios::app);
    if (!f.good()) throw Error(SPH_ERR_INVALID, "batch: cannot write the tuning cache " + path);
    const std::string key = tune_key();
    for (int i = 0; i < K_COUNT; ++i)
        f << key << " " << kernel_name(i) << " " << cfg_[i].block << " " << cfg_[i].grid << " " << cfg_[i].stage
          << "\n";
}

int BatchWCSPH::autotune(const char* cache_path, int trials) {
    require_final();
    if (n_ == 0) return 0;
    if (cache_path && *cache_path && load_tuning(cache_path)) return 1;
    if (trials < 1) trials = 1;
    hipDeviceProp_t prop{};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    const int cus = std::max(prop.multiProcessorCount, 1);
    // keep the state: tuning steps advance the scenes, the caller's steps
    // must start from where they were
    DeviceArray<float4> save_p, save_v, save_h;
    save_p.resize(n_); save_v.resize(n_); save_h.resize(n_);
    hip_check(hipMemcpyAsync(save_p.get(), d_pos.get(), n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_), "copy");
    hip_check(hipMemcpyAsync(save_v.get(), d_vel.get(), n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_), "copy");
    hip_check(hipMemcpyAsync(save_h.get(), d_vh.get(), n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_), "copy");
    const bool prof_was = prof_.enabled;
    const std::string filter_was = prof_.filter;
    const int grids[3] = {0, 4 * cus, 16 * cus};
    const int blocks[3] = {64, 128, 256};
    for (int k = 0; k < K_COUNT; ++k) {
        // batch_hash runs on the first step only: later steps hash inside
        // the sort of the moved keys, whose launch shapes are fixed
        if (k == K_HASH) continue;
        LaunchCfg best = cfg_[k];
        double best_ms = 1e30;
        const int nstage = k == K_LISTS ? 2 : 1;
        for (int b : blocks)
            for (int g : grids)
                for (int st = 0; st < nstage; ++st) {
                    LaunchCfg c;
                    c.block = b;
                    c.grid = g;
                    c.stage = (k == K_LISTS && st == 0) ? 24 : (k == K_LISTS ? 0 : 24);
                    set_config(k, c);
                    prof_.enabled = false;
                    step(1);  // warm
                    prof_.reset();
                    prof_.filter = kernel_name(k);
                    prof_.enabled = true;
                    step(trials);
                    synchronize();
                    prof_.resolve();
                    prof_.enabled = false;
                    double ms = 1e30;
                    for (auto& e : prof_.entries)
                        if (e.name == kernel_name(k) && e.launches > 0) ms = e.total_ms / e.launches;
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = c;
                    }
                    tuned_ms_[k] = best_ms;
                }
        set_config(k, best);
    }
    prof_.reset();
    prof_.filter = filter_was;
    prof_.enabled = prof_was;
    hip_check(hipMemcpyAsync(d_pos.get(), save_p.get(), n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_), "copy");
    hip_check(hipMemcpyAsync(d_vel.get(), save_v.get(), n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_), "copy");
    hip_check(hipMemcpyAsync(d_vh.get(), save_h.get(), n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_), "copy");
    synchronize();
    sorted_hint_ = false;  // the restored positions belong to an earlier order
    if (cache_path && *cache_path) save_tuning(cache_path);
    return 0;
}

}  // namespace sph
