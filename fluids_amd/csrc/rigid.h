// Host rigid-body world that takes the place of Bullet in the two-way
// coupling loop (scene/scene.cpp:100-110, 304-316): the fluid steps, hands a
// force and torque to each body with mass > 0 (rigidbody.cpp:49-54:
// clearForces, applyCentralForce, applyTorque), then the world advances one
// variable step of dt (stepSimulation(dt, 0)) under gravity -9.8 m/s^2.
//
// The step follows the order of a discrete dynamics world:
//   1. integrate velocities: v += (F + m g) dt / m, w += I_world^-1 T dt
//      (|w| dt clamped to pi/2);
//   2. contacts from the current poses: box corners and sphere caps against
//      the container's floor and four walls (FishTank is an open-top box,
//      fishtank.cpp:143-182) and against other bodies;
//   3. sequential impulses (normal impulses accumulated >= 0 with a
//      Baumgarte bias, Coulomb friction clamped by the normal impulse);
//   4. integrate poses: x += v dt, q <- exp(w dt / 2) q, renormalised;
//   5. clear the applied forces.
// Bullet itself is not available to compare against, so the trajectories of
// this world are pinned by closed forms (free fall, torque-free rotation,
// resting contact), not by Bullet's own numbers.
#pragma once

#include <array>
#include <vector>

namespace sph {

struct RigidBody {
    enum Shape { STATIC = 0, BOX = 1, SPHERE = 2 };
    Shape shape = STATIC;
    float mass = 0.0f, inv_mass = 0.0f;
    float half[3] = {0, 0, 0};  // box half extents
    float radius = 0.0f;        // sphere radius
    float inv_inertia[3] = {0, 0, 0};  // body-frame diagonal
    float pos[3] = {0, 0, 0};   // centre of mass
    float q[4] = {0, 0, 0, 1};  // x, y, z, w (unit)
    float lin[3] = {0, 0, 0}, ang[3] = {0, 0, 0};
    float force[3] = {0, 0, 0}, torque[3] = {0, 0, 0};
    float friction = 0.5f;  // per body; a contact uses the product of both
    bool dynamic() const { return inv_mass > 0.0f; }
};

struct RigidContact {
    int a, b;          // body indices; b = -1 for the container
    float p[3];        // world contact point
    float n[3];        // unit normal, pushing body a out
    float depth;       // > 0 when overlapping
    float mu;          // friction coefficient
    float ln = 0, lt1 = 0, lt2 = 0;  // accumulated impulses
};

class RigidWorld {
public:
    explicit RigidWorld(float gravity_y = -9.8f) : gravity_{0.0f, gravity_y, 0.0f} {}

    // FishTank(w, h, d) centred at `center`: floor y = cy - h/2 and walls at
    // x = cx +- w/2, z = cz +- d/2; no lid.
    void set_container(float w, float h, float d, const float center[3]);
    bool has_container() const { return container_; }

    // quat is (x, y, z, w) and need not be normalised (a btQuaternion built
    // from scene-file numbers, scene.cpp:399-404)
    int add_box(float mass, const float half[3], const float pos[3], const float quat[4]);
    int add_sphere(float mass, float radius, const float pos[3]);
    int add_static(const float pos[3], const float quat[4]);

    int count() const { return (int)bodies_.size(); }
    RigidBody& body(int id);
    const RigidBody& body(int id) const;

    // RigidBody::apply_force_n_torque (rigidbody.cpp:49-54)
    void apply_force_n_torque(int id, const float f[3], const float t[3]);
    void set_velocity(int id, const float lin[3], const float ang[3]);
    void set_pose(int id, const float pos[3], const float quat[4]);

    // one stepSimulation(dt, 0)
    void step(float dt);

    // basis rows of the rotation and the centre of mass, as the boundary
    // handler reads them (boundaryhandler.cpp:259-270)
    void transform(int id, float R[9], float com[3]) const;
    const std::vector<RigidContact>& last_contacts() const { return contacts_; }

    int solver_iterations = 10;
    float erp = 0.2f;               // fraction of the overlap removed per step
    float contact_distance = 0.02f; // contacts are made below this separation

private:
    void collide();
    void contacts_container(int a);
    void contacts_pair(int a, int b);
    void add_contact(int a, int b, const float p[3], const float n[3], float depth);
    void solve(float dt);

    std::array<float, 3> gravity_;
    bool container_ = false;
    float lo_[3] = {0, 0, 0}, hi_[3] = {0, 0, 0};
    std::vector<RigidBody> bodies_;
    std::vector<RigidContact> contacts_;
};

}  // namespace sph
