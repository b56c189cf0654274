// Scene-file front-end (scene.h): a small JSON reader and the builder that
// follows Scene::load_scene (scene/scene.cpp:322-476).
#include "scene.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

namespace sph {

std::vector<float4> sample_sphere(float radius, float spacing);
std::vector<float4> sample_model(const std::string& obj_path, float spacing);

namespace {

// ------------------------------------------------------------------ JSON
struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    // json11's accessors: a missing key or index reads as null, a null or
    // non-number reads as 0 / "" / empty
    const Json& operator[](const std::string& k) const {
        static const Json null;
        if (kind != OBJ) return null;
        auto it = obj.find(k);
        return it == obj.end() ? null : it->second;
    }
    const Json& operator[](size_t i) const {
        static const Json null;
        return (kind == ARR && i < arr.size()) ? arr[i] : null;
    }
    double number() const { return kind == NUM ? num : 0.0; }
    std::string string() const { return kind == STR ? str : std::string(); }
};

class JsonReader {
public:
    explicit JsonReader(const std::string& text) : s_(text) {}
    Json parse() {
        Json v = value();
        ws();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    [[noreturn]] void fail(const std::string& what) {
        throw Error(SPH_ERR_PARSE, "scene file: " + what + " at offset " + std::to_string(p_));
    }
    void ws() {
        while (p_ < s_.size()) {
            char c = s_[p_];
            if (c == ' ' || c == '\t' || c == '\n' || c == '\r') {
                ++p_;
            } else if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '/') {  // json11 accepts comments
                while (p_ < s_.size() && s_[p_] != '\n') ++p_;
            } else {
                break;
            }
        }
    }
    Json value() {
        ws();
        if (p_ >= s_.size()) fail("unexpected end");
        char c = s_[p_];
        Json v;
        if (c == '{') {
            v.kind = Json::OBJ;
            ++p_;
            ws();
            if (peek('}')) return v;
            for (;;) {
                ws();
                if (!peek('"')) fail("expected a key");
                std::string k = string_body();
                ws();
                if (!peek(':')) fail("expected ':'");
                v.obj[k] = value();
                ws();
                if (peek('}')) return v;
                if (!peek(',')) fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Json::ARR;
            ++p_;
            ws();
            if (peek(']')) return v;
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (peek(']')) return v;
                if (!peek(',')) fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            ++p_;
            v.kind = Json::STR;
            v.str = string_body();
            return v;
        }
        if (s_.compare(p_, 4, "true") == 0) { p_ += 4; v.kind = Json::BOOL; v.b = true; return v; }
        if (s_.compare(p_, 5, "false") == 0) { p_ += 5; v.kind = Json::BOOL; return v; }
        if (s_.compare(p_, 4, "null") == 0) { p_ += 4; return v; }
        if (c == '-' || std::isdigit((unsigned char)c)) {
            const char* b = s_.c_str() + p_;
            char* e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) fail("bad number");
            p_ += (size_t)(e - b);
            v.kind = Json::NUM;
            return v;
        }
        fail("unexpected character");
    }
    bool peek(char c) {
        if (p_ < s_.size() && s_[p_] == c) {
            ++p_;
            return true;
        }
        return false;
    }
    // after the opening quote
    std::string string_body() {
        std::string out;
        while (p_ < s_.size() && s_[p_] != '"') {
            char c = s_[p_++];
            if (c == '\\' && p_ < s_.size()) {
                char e = s_[p_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u':
                        if (p_ + 4 > s_.size()) fail("bad escape");
                        out += (char)std::strtol(s_.substr(p_, 4).c_str(), nullptr, 16);
                        p_ += 4;
                        break;
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p_ >= s_.size()) fail("unterminated string");
        ++p_;
        return out;
    }
    const std::string& s_;
    size_t p_ = 0;
};

float f32(double v) { return (float)v; }

void rows_of(const float q[4], float R[9]) {
    // btMatrix3x3::setRotation: s = 2 / |q|^2 (the quaternion need not be unit)
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float d = x * x + y * y + z * z + w * w;
    const float s = 2.0f / d;
    const float xs = x * s, ys = y * s, zs = z * s;
    const float wx = w * xs, wy = w * ys, wz = w * zs, xx = x * xs, xy = x * ys, xz = x * zs;
    const float yy = y * ys, yz = y * zs, zz = z * zs;
    R[0] = 1.0f - (yy + zz); R[1] = xy - wz; R[2] = xz + wy;
    R[3] = xy + wz; R[4] = 1.0f - (xx + zz); R[5] = yz - wx;
    R[6] = xz - wy; R[7] = yz + wx; R[8] = 1.0f - (xx + yy);
}

std::string join(const std::string& dir, const std::string& p) {
    if (p.empty() || p[0] == '/' || dir.empty()) return p;
    return dir.back() == '/' ? dir + p : dir + "/" + p;
}

}  // namespace

SceneFile parse_scene(const std::string& path) {
    std::ifstream f(path);
    if (!f.good()) throw Error(SPH_ERR_PARSE, "Could not parse scene file: " + path);  // scene.cpp:328-330
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    Json root = JsonReader(text).parse();
    SceneFile out;
    const Json& c = root["container"];
    out.container[0] = f32(c["width"].number());
    out.container[1] = f32(c["height"].number());
    out.container[2] = f32(c["depth"].number());
    for (const Json& v : root["fluid_volumes"].arr) {
        if (v["type"].string() != "box") continue;  // the only volume type the loader reads
        std::vector<float> size = {f32(v["size"][0].number()), f32(v["size"][1].number()), f32(v["size"][2].number())};
        std::vector<float> center = {f32(v["center"][0].number()), f32(v["center"][1].number()),
                                     f32(v["center"][2].number())};
        out.volumes.push_back({size, center});
    }
    for (const Json& r : root["rigid_bodies"].arr) {
        SceneBody b;
        b.type = r["type"].string();
        b.name = r["name"].string();
        b.mass = f32(r["mass"].number());
        for (int k = 0; k < 3; ++k) b.center[k] = f32(r["center"][k].number());
        if (b.type == "cube" || b.type == "model" || b.type == "wall")
            for (int k = 0; k < 4; ++k) b.rotation[k] = f32(r["rotation"][k].number());
        if (b.type == "cube" || b.type == "sphere") {
            b.size[0] = f32(r["size"].number());
        } else if (b.type == "wall") {
            b.size[0] = f32(r["size"][0].number());
            b.size[1] = f32(r["size"][1].number());
        } else if (b.type == "model") {
            b.obj_filename = r["obj_filename"].string();
        } else {
            continue;  // unknown types create no body (scene.cpp:466-471)
        }
        out.bodies.push_back(b);
    }
    return out;
}

std::vector<float4> body_samples(const SceneBody& b, float spacing, const std::string& base_dir) {
    if (b.type == "cube") return sample_cube(b.size[0], spacing);
    if (b.type == "sphere") return sample_sphere(b.size[0], spacing);
    if (b.type == "wall") return sample_wall(b.size[0], b.size[1], spacing);
    if (b.type == "model") return sample_model(join(base_dir, b.obj_filename), spacing);
    throw Error(SPH_ERR_INVALID, "unknown body type '" + b.type + "'");
}

SceneRun::SceneRun(const std::string& path, const Settings& s, const std::string& base_dir)
    : sim_(build_simulation(s)), world_(-9.8f) {
    simp_ = sim_.get();
    SceneFile sf = parse_scene(path);
    // container: a static FishTank in the rigid world and the clamp box
    world_.set_container(sf.container[0], sf.container[1], sf.container[2], nullptr);
    simp_->set_rect_limits(sf.container[0], sf.container[1], sf.container[2]);
    for (auto& v : sf.volumes) simp_->add_volume(std::make_shared<BoxVolume>(v.first.data(), v.second.data()));
    const float spacing = 2.0f * s.fluid_particle_radius;  // boundaryhandler.cpp:19
    for (const SceneBody& b : sf.bodies) {
        Entry e;
        e.body = b;
        std::vector<float4> local = body_samples(b, spacing, base_dir);
        e.samples = (int)local.size();
        float R[9];
        rows_of(b.rotation, R);
        const bool moves = b.mass > 0.0f;  // scene.cpp:471
        e.sph_id = simp_->add_boundary(local, moves, b.mass, R, b.center);
        if (b.type == "cube") {
            const float h = b.size[0] / 2, half[3] = {h, h, h};
            e.rigid_id = world_.add_box(b.mass, half, b.center, b.rotation);
        } else if (b.type == "sphere") {
            e.rigid_id = world_.add_sphere(b.mass, b.size[0], b.center);
        } else if (b.type == "model" && moves) {
            // the reference builds a convex hull (model.cpp:130-148); here the
            // samples' bounding box stands in for it
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (const float4& p : local) {
                const float q[3] = {p.x, p.y, p.z};
                for (int k = 0; k < 3; ++k) {
                    lo[k] = std::min(lo[k], q[k]);
                    hi[k] = std::max(hi[k], q[k]);
                }
            }
            float half[3], off[3];
            for (int k = 0; k < 3; ++k) {
                half[k] = std::max((hi[k] - lo[k]) / 2, 1e-4f);
                off[k] = (hi[k] + lo[k]) / 2;
            }
            (void)off;
            e.rigid_id = world_.add_box(b.mass, half, b.center, b.rotation);
        } else {
            e.rigid_id = world_.add_static(b.center, b.rotation);
        }
        entries_.push_back(e);
    }
}

void SceneRun::step() {
    for (const Entry& e : entries_) {
        if (!(e.body.mass > 0.0f)) continue;
        float R[9], com[3];
        world_.transform(e.rigid_id, R, com);
        const RigidBody& b = world_.body(e.rigid_id);
        simp_->set_body_state(e.sph_id, R, com, b.lin, b.ang);
    }
    simp_->simulate();
    for (const Entry& e : entries_) {
        if (!(e.body.mass > 0.0f)) continue;
        float f[3], t[3];
        simp_->body_wrench(e.sph_id, f, t);
        world_.apply_force_n_torque(e.rigid_id, f, t);
    }
    world_.step(simp_->settings().time_step);
}

}  // namespace sph
