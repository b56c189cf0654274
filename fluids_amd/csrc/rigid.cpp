// Host rigid-body world for the two-way coupling loop; the step order and the
// contact model are described in rigid.h.
#include "rigid.h"

#include <cmath>
#include <stdexcept>
#include <string>

namespace sph {
namespace {

using V = std::array<float, 3>;

V v3(const float* a) { return {a[0], a[1], a[2]}; }
void store(float* o, const V& a) { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; }
V add(const V& a, const V& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
V sub(const V& a, const V& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V mul(const V& a, float s) { return {a[0] * s, a[1] * s, a[2] * s}; }
float dot(const V& a, const V& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V cross(const V& a, const V& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
float norm(const V& a) { return std::sqrt(dot(a, a)); }

// rotation matrix (rows) of a unit quaternion (x, y, z, w)
void quat_rows(const float* q, float R[9]) {
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float d = x * x + y * y + z * z + w * w;
    const float s = 2.0f / d;
    const float xs = x * s, ys = y * s, zs = z * s;
    const float wx = w * xs, wy = w * ys, wz = w * zs;
    const float xx = x * xs, xy = x * ys, xz = x * zs;
    const float yy = y * ys, yz = y * zs, zz = z * zs;
    R[0] = 1.0f - (yy + zz); R[1] = xy - wz;          R[2] = xz + wy;
    R[3] = xy + wz;          R[4] = 1.0f - (xx + zz); R[5] = yz - wx;
    R[6] = xz - wy;          R[7] = yz + wx;          R[8] = 1.0f - (xx + yy);
}

V rot(const float R[9], const V& a) {  // R a
    return {R[0] * a[0] + R[1] * a[1] + R[2] * a[2], R[3] * a[0] + R[4] * a[1] + R[5] * a[2],
            R[6] * a[0] + R[7] * a[1] + R[8] * a[2]};
}
V rot_t(const float R[9], const V& a) {  // R^T a
    return {R[0] * a[0] + R[3] * a[1] + R[6] * a[2], R[1] * a[0] + R[4] * a[1] + R[7] * a[2],
            R[2] * a[0] + R[5] * a[1] + R[8] * a[2]};
}

// I_world^-1 a = R diag(inv) R^T a
V inv_inertia_world(const RigidBody& b, const V& a) {
    if (!b.dynamic()) return {0, 0, 0};
    float R[9];
    quat_rows(b.q, R);
    V l = rot_t(R, a);
    for (int k = 0; k < 3; ++k) l[k] *= b.inv_inertia[k];
    return rot(R, l);
}

void normalise(float* q) {
    float d = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (!(d > 0.0f)) throw std::invalid_argument("rigid body: zero quaternion");
    for (int k = 0; k < 4; ++k) q[k] /= d;
}

void check_vec(const float* a, const char* what) {
    if (!a) throw std::invalid_argument(std::string("rigid body: null ") + what);
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(a[k])) throw std::invalid_argument(std::string("rigid body: non-finite ") + what);
}

}  // namespace

void RigidWorld::set_container(float w, float h, float d, const float center[3]) {
    if (!(w > 0 && h > 0 && d > 0)) throw std::invalid_argument("rigid world: container sides must be > 0");
    float c[3] = {0, 0, 0};
    if (center) for (int k = 0; k < 3; ++k) c[k] = center[k];
    lo_[0] = c[0] - w / 2; hi_[0] = c[0] + w / 2;
    lo_[1] = c[1] - h / 2; hi_[1] = c[1] + h / 2;
    lo_[2] = c[2] - d / 2; hi_[2] = c[2] + d / 2;
    container_ = true;
}

int RigidWorld::add_box(float mass, const float half[3], const float pos[3], const float quat[4]) {
    check_vec(half, "half extents");
    check_vec(pos, "position");
    if (!(mass >= 0.0f)) throw std::invalid_argument("rigid body: mass must be >= 0");
    RigidBody b;
    b.shape = RigidBody::BOX;
    for (int k = 0; k < 3; ++k) {
        if (!(half[k] > 0)) throw std::invalid_argument("rigid body: box half extents must be > 0");
        b.half[k] = half[k];
        b.pos[k] = pos[k];
    }
    if (quat) for (int k = 0; k < 4; ++k) b.q[k] = quat[k];
    normalise(b.q);
    b.mass = mass;
    if (mass > 0.0f) {
        b.inv_mass = 1.0f / mass;
        // btBoxShape::calculateLocalInertia: m/12 (ly^2 + lz^2), l = 2 half
        const float lx = 2 * half[0], ly = 2 * half[1], lz = 2 * half[2];
        b.inv_inertia[0] = 12.0f / (mass * (ly * ly + lz * lz));
        b.inv_inertia[1] = 12.0f / (mass * (lx * lx + lz * lz));
        b.inv_inertia[2] = 12.0f / (mass * (lx * lx + ly * ly));
    }
    bodies_.push_back(b);
    return (int)bodies_.size() - 1;
}

int RigidWorld::add_sphere(float mass, float radius, const float pos[3]) {
    check_vec(pos, "position");
    if (!(mass >= 0.0f)) throw std::invalid_argument("rigid body: mass must be >= 0");
    if (!(radius > 0)) throw std::invalid_argument("rigid body: sphere radius must be > 0");
    RigidBody b;
    b.shape = RigidBody::SPHERE;
    b.radius = radius;
    for (int k = 0; k < 3; ++k) b.pos[k] = pos[k];
    b.mass = mass;
    if (mass > 0.0f) {
        b.inv_mass = 1.0f / mass;
        const float inv = 1.0f / (0.4f * mass * radius * radius);  // btSphereShape: 2/5 m r^2
        b.inv_inertia[0] = b.inv_inertia[1] = b.inv_inertia[2] = inv;
    }
    bodies_.push_back(b);
    return (int)bodies_.size() - 1;
}

int RigidWorld::add_static(const float pos[3], const float quat[4]) {
    check_vec(pos, "position");
    RigidBody b;
    for (int k = 0; k < 3; ++k) b.pos[k] = pos[k];
    if (quat) for (int k = 0; k < 4; ++k) b.q[k] = quat[k];
    normalise(b.q);
    bodies_.push_back(b);
    return (int)bodies_.size() - 1;
}

RigidBody& RigidWorld::body(int id) {
    if (id < 0 || id >= (int)bodies_.size()) throw std::out_of_range("rigid world: no body " + std::to_string(id));
    return bodies_[id];
}
const RigidBody& RigidWorld::body(int id) const {
    if (id < 0 || id >= (int)bodies_.size()) throw std::out_of_range("rigid world: no body " + std::to_string(id));
    return bodies_[id];
}

void RigidWorld::apply_force_n_torque(int id, const float f[3], const float t[3]) {
    check_vec(f, "force");
    check_vec(t, "torque");
    RigidBody& b = body(id);
    // clearForces(); applyCentralForce(f); applyTorque(t)
    for (int k = 0; k < 3; ++k) {
        b.force[k] = f[k];
        b.torque[k] = t[k];
    }
}

void RigidWorld::set_velocity(int id, const float lin[3], const float ang[3]) {
    RigidBody& b = body(id);
    if (lin) { check_vec(lin, "linear velocity"); for (int k = 0; k < 3; ++k) b.lin[k] = lin[k]; }
    if (ang) { check_vec(ang, "angular velocity"); for (int k = 0; k < 3; ++k) b.ang[k] = ang[k]; }
}

void RigidWorld::set_pose(int id, const float pos[3], const float quat[4]) {
    RigidBody& b = body(id);
    if (pos) { check_vec(pos, "position"); for (int k = 0; k < 3; ++k) b.pos[k] = pos[k]; }
    if (quat) { for (int k = 0; k < 4; ++k) b.q[k] = quat[k]; normalise(b.q); }
}

void RigidWorld::transform(int id, float R[9], float com[3]) const {
    const RigidBody& b = body(id);
    if (R) quat_rows(b.q, R);
    if (com) for (int k = 0; k < 3; ++k) com[k] = b.pos[k];
}

void RigidWorld::add_contact(int a, int b, const float p[3], const float n[3], float depth) {
    RigidContact c;
    c.a = a;
    c.b = b;
    for (int k = 0; k < 3; ++k) {
        c.p[k] = p[k];
        c.n[k] = n[k];
    }
    c.depth = depth;
    c.mu = bodies_[a].friction * (b >= 0 ? bodies_[b].friction : 0.5f);
    contacts_.push_back(c);
}

void RigidWorld::contacts_container(int a) {
    const RigidBody& B = bodies_[a];
    // inward normals and the plane each one bounds: floor and four walls
    const float normals[5][3] = {{0, 1, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, -1}};
    auto depth_of = [&](int f, const V& p) {
        switch (f) {
            case 0: return lo_[1] - p[1];
            case 1: return lo_[0] - p[0];
            case 2: return p[0] - hi_[0];
            case 3: return lo_[2] - p[2];
            default: return p[2] - hi_[2];
        }
    };
    if (B.shape == RigidBody::BOX) {
        float R[9];
        quat_rows(B.q, R);
        for (int c = 0; c < 8; ++c) {
            V l = {(c & 1) ? B.half[0] : -B.half[0], (c & 2) ? B.half[1] : -B.half[1],
                   (c & 4) ? B.half[2] : -B.half[2]};
            V p = add(v3(B.pos), rot(R, l));
            for (int f = 0; f < 5; ++f) {
                float d = depth_of(f, p);
                if (d > -contact_distance) add_contact(a, -1, p.data(), normals[f], d);
            }
        }
    } else if (B.shape == RigidBody::SPHERE) {
        for (int f = 0; f < 5; ++f) {
            V n = v3(normals[f]);
            V p = sub(v3(B.pos), mul(n, B.radius));
            float d = depth_of(f, p);
            if (d > -contact_distance) add_contact(a, -1, p.data(), normals[f], d);
        }
    }
}

void RigidWorld::contacts_pair(int a, int b) {
    RigidBody& A = bodies_[a];
    RigidBody& B = bodies_[b];
    if (A.shape == RigidBody::STATIC || B.shape == RigidBody::STATIC) return;
    if (A.shape == RigidBody::SPHERE && B.shape == RigidBody::SPHERE) {
        V d = sub(v3(A.pos), v3(B.pos));
        float dist = norm(d);
        float depth = A.radius + B.radius - dist;
        if (depth > -contact_distance && dist > 0.0f) {
            V n = mul(d, 1.0f / dist);
            V p = add(v3(B.pos), mul(n, B.radius));
            add_contact(a, b, p.data(), n.data(), depth);
        }
        return;
    }
    // a point of `pi` (a box corner or a sphere's nearest point) against box `bi`
    auto point_in_box = [&](int pi, int bi, const V& p, float extra) {
        const RigidBody& P = bodies_[pi];
        const RigidBody& X = bodies_[bi];
        float R[9];
        quat_rows(X.q, R);
        V l = rot_t(R, sub(p, v3(X.pos)));
        int axis = -1;
        float pen = 0.0f;
        for (int k = 0; k < 3; ++k) {
            float s = X.half[k] + extra - std::fabs(l[k]);
            if (s <= -contact_distance) return;
            if (axis < 0 || s < pen) {
                axis = k;
                pen = s;
            }
        }
        V e = {0, 0, 0};
        e[axis] = l[axis] >= 0 ? 1.0f : -1.0f;
        V n = rot(R, e);  // out of X, towards P
        (void)P;
        add_contact(pi, bi, p.data(), n.data(), pen);
    };
    auto box_corners = [&](int bi, int other) {
        const RigidBody& X = bodies_[bi];
        float R[9];
        quat_rows(X.q, R);
        for (int c = 0; c < 8; ++c) {
            V l = {(c & 1) ? X.half[0] : -X.half[0], (c & 2) ? X.half[1] : -X.half[1],
                   (c & 4) ? X.half[2] : -X.half[2]};
            point_in_box(bi, other, add(v3(X.pos), rot(R, l)), 0.0f);
        }
    };
    if (A.shape == RigidBody::BOX && B.shape == RigidBody::BOX) {
        box_corners(a, b);
        box_corners(b, a);
        return;
    }
    // sphere against box
    int s = A.shape == RigidBody::SPHERE ? a : b;
    int x = s == a ? b : a;
    const RigidBody& S = bodies_[s];
    const RigidBody& X = bodies_[x];
    float R[9];
    quat_rows(X.q, R);
    V l = rot_t(R, sub(v3(S.pos), v3(X.pos)));
    V c = l;
    for (int k = 0; k < 3; ++k) c[k] = std::fmax(-X.half[k], std::fmin(X.half[k], l[k]));
    V d = sub(l, c);
    float dist = norm(d);
    if (dist > 0.0f) {
        float depth = S.radius - dist;
        if (depth > -contact_distance) {
            V n = rot(R, mul(d, 1.0f / dist));
            V p = add(v3(X.pos), rot(R, c));
            add_contact(s, x, p.data(), n.data(), depth);
        }
    } else {
        point_in_box(s, x, v3(S.pos), S.radius);
    }
}

void RigidWorld::collide() {
    contacts_.clear();
    const int n = (int)bodies_.size();
    for (int a = 0; a < n; ++a) {
        if (!bodies_[a].dynamic()) continue;
        if (container_) contacts_container(a);
        for (int b = 0; b < n; ++b)
            if (b != a && (b > a || !bodies_[b].dynamic())) contacts_pair(a, b);
    }
}

void RigidWorld::solve(float dt) {
    auto vel_at = [&](int i, const V& r) {
        if (i < 0) return V{0, 0, 0};
        const RigidBody& b = bodies_[i];
        return add(v3(b.lin), cross(v3(b.ang), r));
    };
    auto apply = [&](int i, const V& r, const V& imp, float sign) {
        if (i < 0) return;
        RigidBody& b = bodies_[i];
        if (!b.dynamic()) return;
        V j = mul(imp, sign);
        for (int k = 0; k < 3; ++k) b.lin[k] += j[k] * b.inv_mass;
        V w = inv_inertia_world(b, cross(r, j));
        for (int k = 0; k < 3; ++k) b.ang[k] += w[k];
    };
    auto k_eff = [&](int i, const V& r, const V& n) {
        if (i < 0) return 0.0f;
        const RigidBody& b = bodies_[i];
        if (!b.dynamic()) return 0.0f;
        return b.inv_mass + dot(cross(inv_inertia_world(b, cross(r, n)), r), n);
    };
    struct Row { V ra, rb, n, t1, t2; float kn, kt1, kt2, bias; };
    std::vector<Row> rows(contacts_.size());
    for (size_t c = 0; c < contacts_.size(); ++c) {
        RigidContact& C = contacts_[c];
        Row& w = rows[c];
        w.n = v3(C.n);
        w.ra = sub(v3(C.p), v3(bodies_[C.a].pos));
        w.rb = C.b >= 0 ? sub(v3(C.p), v3(bodies_[C.b].pos)) : V{0, 0, 0};
        V vr = sub(vel_at(C.a, w.ra), vel_at(C.b, w.rb));
        V vt = sub(vr, mul(w.n, dot(vr, w.n)));
        float lt = norm(vt);
        if (lt > 1e-6f) {
            w.t1 = mul(vt, 1.0f / lt);
        } else {  // any direction in the tangent plane
            V ref = std::fabs(w.n[0]) < 0.9f ? V{1, 0, 0} : V{0, 1, 0};
            V t = cross(w.n, ref);
            w.t1 = mul(t, 1.0f / norm(t));
        }
        w.t2 = cross(w.n, w.t1);
        w.kn = k_eff(C.a, w.ra, w.n) + k_eff(C.b, w.rb, w.n);
        w.kt1 = k_eff(C.a, w.ra, w.t1) + k_eff(C.b, w.rb, w.t1);
        w.kt2 = k_eff(C.a, w.ra, w.t2) + k_eff(C.b, w.rb, w.t2);
        // overlapping: push apart a fraction erp of the overlap per step;
        // separated by s: allow closing at up to s / dt
        w.bias = C.depth > 0.0f ? erp * C.depth / dt : C.depth / dt;
    }
    for (int it = 0; it < solver_iterations; ++it) {
        for (size_t c = 0; c < contacts_.size(); ++c) {
            RigidContact& C = contacts_[c];
            Row& w = rows[c];
            if (!(w.kn > 0.0f)) continue;
            V vr = sub(vel_at(C.a, w.ra), vel_at(C.b, w.rb));
            float dl = (w.bias - dot(vr, w.n)) / w.kn;
            float nl = std::fmax(C.ln + dl, 0.0f);
            dl = nl - C.ln;
            C.ln = nl;
            V imp = mul(w.n, dl);
            apply(C.a, w.ra, imp, 1.0f);
            apply(C.b, w.rb, imp, -1.0f);
            const float lim = C.mu * C.ln;
            for (int t = 0; t < 2; ++t) {
                const V& dir = t == 0 ? w.t1 : w.t2;
                float kt = t == 0 ? w.kt1 : w.kt2;
                float& acc = t == 0 ? C.lt1 : C.lt2;
                if (!(kt > 0.0f)) continue;
                vr = sub(vel_at(C.a, w.ra), vel_at(C.b, w.rb));
                float d = -dot(vr, dir) / kt;
                float na = std::fmax(-lim, std::fmin(lim, acc + d));
                d = na - acc;
                acc = na;
                V ti = mul(dir, d);
                apply(C.a, w.ra, ti, 1.0f);
                apply(C.b, w.rb, ti, -1.0f);
            }
        }
    }
}

void RigidWorld::step(float dt) {
    if (!(dt > 0.0f) || !std::isfinite(dt)) throw std::invalid_argument("rigid world: dt must be > 0");
    const float max_turn = 1.5707963f;  // |w| dt is kept below pi/2
    for (RigidBody& b : bodies_) {
        if (!b.dynamic()) continue;
        for (int k = 0; k < 3; ++k) b.lin[k] += (b.force[k] * b.inv_mass + gravity_[k]) * dt;
        V dw = inv_inertia_world(b, v3(b.torque));
        for (int k = 0; k < 3; ++k) b.ang[k] += dw[k] * dt;
        float w = norm(v3(b.ang));
        if (w * dt > max_turn) {
            float s = max_turn / dt / w;
            for (int k = 0; k < 3; ++k) b.ang[k] *= s;
        }
    }
    collide();
    solve(dt);
    for (RigidBody& b : bodies_) {
        if (!b.dynamic()) continue;
        for (int k = 0; k < 3; ++k) b.pos[k] += b.lin[k] * dt;
        V w = v3(b.ang);
        float wl = norm(w);
        float half = 0.5f * wl * dt;
        float s = wl > 1e-6f ? std::sin(half) / wl : 0.5f * dt;
        float dq[4] = {w[0] * s, w[1] * s, w[2] * s, std::cos(half)};
        const float* q = b.q;
        float r[4] = {dq[3] * q[0] + q[3] * dq[0] + (dq[1] * q[2] - dq[2] * q[1]),
                      dq[3] * q[1] + q[3] * dq[1] + (dq[2] * q[0] - dq[0] * q[2]),
                      dq[3] * q[2] + q[3] * dq[2] + (dq[0] * q[1] - dq[1] * q[0]),
                      dq[3] * q[3] - (dq[0] * q[0] + dq[1] * q[1] + dq[2] * q[2])};
        normalise(r);
        for (int k = 0; k < 4; ++k) b.q[k] = r[k];
        for (int k = 0; k < 3; ++k) {
            b.force[k] = 0.0f;
            b.torque[k] = 0.0f;
        }
    }
}

}  // namespace sph
