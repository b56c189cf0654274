// Scene-file front-end: Scene::load_scene (scene/scene.cpp:322-476) without
// the rendering parts. A scene file names the container, box-shaped fluid
// volumes and rigid bodies (cube, sphere, model, wall); loading it builds
// the fluid solver in the reference's init order (set_rect_limits,
// add_volume*, add_boundary*), the rigid-body world, and the pairs of
// (boundary body, rigid body) that the per-frame coupling exchanges.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "rigid.h"
#include "solver.h"

namespace sph {

struct SceneBody {
    std::string type, name, obj_filename;
    float mass = 0.0f;
    float size[3] = {0, 0, 0};     // cube: side; sphere: radius; wall: width, height
    float center[3] = {0, 0, 0};
    float rotation[4] = {0, 0, 0, 1};  // (x, y, z, w) as written in the file
};

struct SceneFile {
    float container[3] = {0, 0, 0};
    std::vector<std::pair<std::vector<float>, std::vector<float>>> volumes;  // (size, center)
    std::vector<SceneBody> bodies;
};

// Parses a scene file (json11 semantics for the keys the loader reads:
// missing numbers read as 0, unknown keys are ignored).
SceneFile parse_scene(const std::string& path);

// Body-local surface samples of a body at the given spacing.
std::vector<float4> body_samples(const SceneBody& b, float spacing, const std::string& base_dir);

class SceneRun {
public:
    // settings: the solver settings (config.pci / config.wc); base_dir
    // resolves relative model paths (the reference resolves them against its
    // working directory)
    SceneRun(const std::string& path, const Settings& s, const std::string& base_dir);
    SimulationBase& sim() { return *simp_; }
    RigidWorld& world() { return world_; }
    // one frame of Scene::render: body states to the fluid, fluid step,
    // wrenches to bodies with mass > 0, rigid step of the fluid's time step
    void step();
    struct Entry {
        SceneBody body;
        int sph_id = -1, rigid_id = -1, samples = 0;
    };
    const std::vector<Entry>& bodies() const { return entries_; }
    // ownership of the solver, for a caller that hands it out as a handle;
    // sim() stays valid while that owner lives
    std::unique_ptr<SimulationBase>& sim_owner() { return sim_; }

private:
    std::unique_ptr<SimulationBase> sim_;
    SimulationBase* simp_ = nullptr;
    RigidWorld world_;
    std::vector<Entry> entries_;
};

}  // namespace sph
