// axiom::dynamics::PhysicsWorld — C++20 host API over the C ABI (axiom/axm.h).
//
// The reference only forward-declares this namespace
// (include/axiom/debug/physics_debug_draw.hpp:14-22: collision::Shape,
// collision::ContactPoint, dynamics::RigidBody, dynamics::PhysicsWorld) and
// reserves src/collision and src/dynamics for it.  This header gives that
// surface, following the reference's conventions:
//   - no exceptions: every fallible call returns a Status whose code is the
//     numeric axiom::core::ErrorCode (include/axiom/core/error_code.hpp:21-57);
//     an engine build wraps it into core::Result<T> (see toResult below);
//   - factories return std::unique_ptr (VkContext::create,
//     include/axiom/gpu/vk_instance.hpp:34); handles are non-copyable, movable;
//   - camelCase methods, trailing-underscore members (.clang-tidy:48-86);
//   - stats fill any struct with gui::PhysicsWorldStats' field names
//     (include/axiom/gui/physics_panel.hpp:11-31) and bodies convert from any
//     struct with gui::RigidBodyData's field names
//     (include/axiom/gui/body_inspector.hpp:58-90), without including the GUI.
// Header-only; link libaxm.so.
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

#include "axiom/axm.h"

namespace axiom::dynamics {

// Outcome of a call: the axm_status code and its message.
class Status {
public:
    constexpr Status() noexcept = default;
    constexpr explicit Status(axm_status code) noexcept : code_(code) {}
    bool ok() const noexcept { return code_ == AXM_SUCCESS; }
    explicit operator bool() const noexcept { return ok(); }
    axm_status code() const noexcept { return code_; }
    const char* message() const noexcept { return axm_status_string(code_); }

private:
    axm_status code_ = AXM_SUCCESS;
};

// Maps a Status onto the engine's Result type (core::Result<void> in the
// reference, include/axiom/core/result.hpp:233-270): R needs success() and
// failure(ErrorCode, const char*).
template <class R, class ErrorCodeT> R toResult(Status s) {
    return s.ok() ? R::success() : R::failure(static_cast<ErrorCodeT>(s.code()), s.message());
}

enum class ShapeType : uint32_t { Sphere = AXM_SHAPE_SPHERE, Box = AXM_SHAPE_BOX, Capsule = AXM_SHAPE_CAPSULE };

// Host structure-of-arrays body set (the upload / download format).  Per-body
// material, damping, kinematic flag and collision filter travel too once any body
// sets them (a body added without them takes the world defaults given here).
class BodySet {
public:
    // defaults of the per-body extras for bodies added without them (a world
    // description's friction, restitution and damping)
    void setDefaults(float friction, float restitution, float linearDamping, float angularDamping) noexcept {
        defMat_[0] = friction; defMat_[1] = restitution; defMat_[2] = linearDamping; defMat_[3] = angularDamping;
    }
    uint32_t size() const noexcept { return static_cast<uint32_t>(invMass_.size()); }
    void reserve(uint32_t n) {
        pos_.reserve(3ull * n); quat_.reserve(4ull * n); linVel_.reserve(3ull * n); angVel_.reserve(3ull * n);
        invMass_.reserve(n); invInertia_.reserve(3ull * n); type_.reserve(n); params_.reserve(3ull * n);
        world_.reserve(n); material_.reserve(4ull * n); flags_.reserve(n); filter_.reserve(3ull * n);
    }

    // Adds a body of the given shape; density 0 makes it static.  Mass
    // properties come from axm_mass_properties (solid shapes of that density).
    Status add(ShapeType type, const float params[3], const float position[3], const float rotation[4],
               const float linearVelocity[3], const float angularVelocity[3], float density = 1000.0f,
               uint32_t world = 0) {
        float im = 0.0f, ii[3] = {0.0f, 0.0f, 0.0f};
        if (density > 0.0f) {
            const Status s(axm_mass_properties(static_cast<uint32_t>(type), params, density, &im, ii));
            if (!s) return s;
        }
        push(type, params, position, rotation, linearVelocity, angularVelocity, im, ii, world, defMat_, 0u, nullptr);
        return Status();
    }

    // Adds a body from any struct with gui::RigidBodyData's fields
    // (include/axiom/gui/body_inspector.hpp:58-90): position, rotation (x, y, z, w),
    // linearVelocity, angularVelocity, mass, inertiaTensor (diagonal), shapeType
    // (Sphere=0, Box=1, Capsule=2), shapeExtents (sphere: x = r; box: half extents;
    // capsule: x = r, y = half-height) and, when present: type (BodyType: Static = 0,
    // Dynamic = 1, Kinematic = 2), material.friction / material.restitution,
    // linearDamping / angularDamping, filter.categoryBits / maskBits / groupIndex.
    // A Static or Kinematic body gets zero inverse mass and inertia whatever its mass
    // field holds (RigidBodyData defaults mass to 1); a Dynamic body needs mass > 0.
    // Without a type field, mass 0 means static.
    template <class RigidBodyDataT> Status addRigidBody(const RigidBodyDataT& rb, uint32_t world = 0) {
        const uint32_t t = static_cast<uint32_t>(rb.shapeType);
        if (t > AXM_SHAPE_CAPSULE) return Status(AXM_INVALID_SHAPE);
        if (!(rb.mass >= 0.0f)) return Status(AXM_INVALID_RIGID_BODY);
        bool dyn = rb.mass > 0.0f;
        uint32_t flags = 0;
        if constexpr (requires { rb.type; }) {
            const uint32_t bt = static_cast<uint32_t>(rb.type);
            if (bt > 2u) return Status(AXM_INVALID_RIGID_BODY);
            if (bt == 1u && !(rb.mass > 0.0f)) return Status(AXM_INVALID_RIGID_BODY);
            dyn = bt == 1u;
            if (bt == 2u) flags |= 1u;  // kinematic
        }
        float mat[4] = {defMat_[0], defMat_[1], defMat_[2], defMat_[3]};
        if constexpr (requires { rb.material.friction; rb.material.restitution; }) {
            mat[0] = rb.material.friction;
            mat[1] = rb.material.restitution;
        }
        if constexpr (requires { rb.linearDamping; rb.angularDamping; }) {
            mat[2] = rb.linearDamping;
            mat[3] = rb.angularDamping;
        }
        uint32_t filt[3] = {1u, 0xFFFFFFFFu, 0u};
        const uint32_t* fp = nullptr;
        if constexpr (requires { rb.filter.categoryBits; rb.filter.maskBits; rb.filter.groupIndex; }) {
            filt[0] = static_cast<uint32_t>(rb.filter.categoryBits);
            filt[1] = static_cast<uint32_t>(rb.filter.maskBits);
            filt[2] = static_cast<uint32_t>(static_cast<int32_t>(rb.filter.groupIndex));
            fp = filt;
        }
        const float prm[3] = {rb.shapeExtents.x, rb.shapeExtents.y, rb.shapeExtents.z};
        const float pos[3] = {rb.position.x, rb.position.y, rb.position.z};
        const float rot[4] = {rb.rotation.x, rb.rotation.y, rb.rotation.z, rb.rotation.w};
        const float v[3] = {rb.linearVelocity.x, rb.linearVelocity.y, rb.linearVelocity.z};
        const float w[3] = {rb.angularVelocity.x, rb.angularVelocity.y, rb.angularVelocity.z};
        const float ii[3] = {dyn && rb.inertiaTensor.x > 0.0f ? 1.0f / rb.inertiaTensor.x : 0.0f,
                             dyn && rb.inertiaTensor.y > 0.0f ? 1.0f / rb.inertiaTensor.y : 0.0f,
                             dyn && rb.inertiaTensor.z > 0.0f ? 1.0f / rb.inertiaTensor.z : 0.0f};
        push(static_cast<ShapeType>(t), prm, pos, rot, v, w, dyn ? 1.0f / rb.mass : 0.0f, ii, world, mat, flags, fp);
        return Status();
    }

    axm_body_soa soa() noexcept {
        return axm_body_soa{pos_.data(), quat_.data(), linVel_.data(), angVel_.data(), invMass_.data(),
                            invInertia_.data(), type_.data(), params_.data(), world_.data(),
                            material_.data(), flags_.data(), anyFilter_ ? filter_.data() : nullptr};
    }
    void resize(uint32_t n) {
        pos_.resize(3ull * n); quat_.resize(4ull * n); linVel_.resize(3ull * n); angVel_.resize(3ull * n);
        invMass_.resize(n); invInertia_.resize(3ull * n); type_.resize(n); params_.resize(3ull * n);
        world_.resize(n); material_.resize(4ull * n); flags_.resize(n); filter_.resize(3ull * n);
    }

    const float* position(uint32_t i) const noexcept { return &pos_[3ull * i]; }
    const float* rotation(uint32_t i) const noexcept { return &quat_[4ull * i]; }
    const float* linearVelocity(uint32_t i) const noexcept { return &linVel_[3ull * i]; }
    const float* angularVelocity(uint32_t i) const noexcept { return &angVel_[3ull * i]; }
    float inverseMass(uint32_t i) const noexcept { return invMass_[i]; }
    bool kinematic(uint32_t i) const noexcept { return (flags_[i] & 1u) != 0u; }

private:
    void push(ShapeType type, const float* prm, const float* pos, const float* rot, const float* v, const float* w,
              float im, const float* ii, uint32_t world, const float* mat, uint32_t flags, const uint32_t* filt) {
        for (int k = 0; k < 3; ++k) {
            pos_.push_back(pos[k]); linVel_.push_back(v[k]); angVel_.push_back(w[k]);
            invInertia_.push_back(ii[k]); params_.push_back(prm[k]);
        }
        for (int k = 0; k < 4; ++k) quat_.push_back(rot[k]);
        for (int k = 0; k < 4; ++k) material_.push_back(mat[k]);
        invMass_.push_back(im);
        type_.push_back(static_cast<uint32_t>(type));
        world_.push_back(world);
        flags_.push_back(flags);
        filter_.push_back(filt ? filt[0] : 1u);
        filter_.push_back(filt ? filt[1] : 0xFFFFFFFFu);
        filter_.push_back(filt ? filt[2] : 0u);
        anyFilter_ = anyFilter_ || filt != nullptr;
    }

    std::vector<float> pos_, quat_, linVel_, angVel_, invMass_, invInertia_, params_, material_;
    std::vector<uint32_t> type_, world_, flags_, filter_;
    float defMat_[4] = {0.5f, 0.1f, 0.0f, 0.0f};
    bool anyFilter_ = false;
};

// One simulation world on one GPU.  Calls on one world are not thread-safe;
// steps run asynchronously until sync() or a read-back.
class PhysicsWorld {
public:
    static axm_world_desc defaultDesc() noexcept {
        axm_world_desc d;
        axm_world_desc_default(&d);
        return d;
    }

    static std::unique_ptr<PhysicsWorld> create(const axm_world_desc& desc, Status* status = nullptr) {
        axm_world w = nullptr;
        const Status s(axm_world_create(&desc, &w));
        if (status) *status = s;
        if (!s) return nullptr;
        return std::unique_ptr<PhysicsWorld>(new PhysicsWorld(w));
    }

    ~PhysicsWorld() {
        if (world_) axm_world_destroy(world_);
    }
    PhysicsWorld(const PhysicsWorld&) = delete;
    PhysicsWorld& operator=(const PhysicsWorld&) = delete;
    PhysicsWorld(PhysicsWorld&& o) noexcept
        : world_(std::exchange(o.world_, nullptr)), last_(o.last_), configDt_(o.configDt_), configIters_(o.configIters_) {}
    PhysicsWorld& operator=(PhysicsWorld&& o) noexcept {
        if (this != &o) {
            if (world_) axm_world_destroy(world_);
            world_ = std::exchange(o.world_, nullptr);
            last_ = o.last_;
            configDt_ = o.configDt_;
            configIters_ = o.configIters_;
        }
        return *this;
    }

    Status setBodies(BodySet& bodies) {
        axm_body_soa s = bodies.soa();
        return Status(axm_world_set_bodies(world_, &s, bodies.size()));
    }
    // Copies positions, rotations and velocities back into `bodies` (resized to the world's count).
    Status getBodies(BodySet& bodies) const {
        uint32_t n = 0;
        const Status c(axm_world_body_count(world_, &n));
        if (!c) return c;
        if (bodies.size() != n) bodies.resize(n);
        axm_body_soa s = bodies.soa();
        s.inv_mass = nullptr; s.inv_inertia = nullptr; s.shape_type = nullptr; s.shape_params = nullptr;
        s.world = nullptr;
        return Status(axm_world_get_bodies(world_, &s));
    }

    // PhysicsWorld::step(dt): broadphase, narrowphase, `iterations` solver sweeps, integration.
    Status step(float dt, uint32_t iterations, uint32_t flags = 0) {
        return Status(axm_world_step(world_, dt, iterations, flags, &last_));
    }
    // `steps` steps; graph = true replays a captured hipGraph of the step.
    Status run(float dt, uint32_t iterations, uint32_t steps, bool graph = true) {
        return Status(axm_world_run(world_, dt, iterations, steps, graph ? 1 : 0, &last_));
    }
    Status sync() { return Status(axm_world_sync(world_)); }

    const axm_step_stats& lastStats() const noexcept { return last_; }

    // Writes any struct with gui::PhysicsWorldConfig's fields (include/axiom/gui/physics_panel.hpp:34-43)
    // into a world description: gravity, allowSleep and the sleep thresholds; the config's
    // timeStep and velocityIterations are what step() takes (see configStep / configIterations).
    // positionIterations has no counterpart (positions are corrected through the Baumgarte bias).
    template <class ConfigT> static void applyConfig(const ConfigT& c, axm_world_desc& d) noexcept {
        d.gravity[0] = c.gravity.x;
        d.gravity[1] = c.gravity.y;
        d.gravity[2] = c.gravity.z;
        d.allow_sleep = c.allowSleep ? 1u : 0u;
        d.sleep_linear = c.sleepLinearThreshold;
        d.sleep_angular = c.sleepAngularThreshold;
        d.sleep_time = c.sleepTimeThreshold;
    }
    // Applies a PhysicsWorldConfig to this live world (axm_world_configure) and keeps its time step
    // and velocity iterations for stepConfigured().
    template <class ConfigT> Status applyConfig(const ConfigT& c) {
        axm_world_desc d{};
        Status s(axm_world_get_desc(world_, &d));
        if (!s) return s;
        applyConfig(c, d);
        s = Status(axm_world_configure(world_, &d));
        if (s) {
            configDt_ = c.timeStep;
            configIters_ = c.velocityIterations;
        }
        return s;
    }
    // One step with the applied config's time step and velocity iterations.
    Status stepConfigured(uint32_t flags = 0) { return step(configDt_, configIters_, flags); }

    // Contact points of the last step as any struct with debug::DebugContactPoint's fields
    // (include/axiom/debug/physics_debug_draw.hpp:128-132): position (witness point on A, world
    // space), normal (A to B), penetrationDepth.
    template <class DebugContactPointT> Status debugContacts(std::vector<DebugContactPointT>& out) const {
        Manifolds m;
        const Status s = contacts(m);
        if (!s) return s;
        out.clear();
        for (size_t k = 0; k < m.a.size(); ++k)
            for (uint32_t p = 0; p < m.npts[k]; ++p) {
                DebugContactPointT c{};
                const size_t o = (k * 4 + p) * 3;
                c.position.x = m.position[o];
                c.position.y = m.position[o + 1];
                c.position.z = m.position[o + 2];
                c.normal.x = m.normal[o];
                c.normal.y = m.normal[o + 1];
                c.normal.z = m.normal[o + 2];
                c.penetrationDepth = m.depth[k * 4 + p];
                out.push_back(c);
            }
        return Status();
    }

    // Fills any struct with gui::PhysicsWorldStats' field names from the last step.
    template <class StatsT> void fillStats(StatsT& s, uint32_t totalBodies, uint32_t staticBodies,
                                           uint32_t kinematicBodies = 0) const {
        s.totalBodies = totalBodies;
        s.staticBodies = staticBodies;
        s.kinematicBodies = kinematicBodies;
        s.dynamicBodies = totalBodies - staticBodies - kinematicBodies;
        s.sleepingBodies = last_.sleeping;
        s.activeBodies = s.dynamicBodies - last_.sleeping;
        s.contactPointCount = last_.contacts;
        s.constraintCount = 0;
        s.islandCount = last_.islands;
        s.totalStepTime = last_.total_ms;
        s.broadphaseTime = last_.broadphase_ms;
        s.narrowphaseTime = last_.narrowphase_ms;
        s.solverTime = last_.solver_ms;
        s.integrationTime = last_.integration_ms;
    }

    // Contact manifolds of the last step (collision::ContactPoint records, up to 4 per manifold).
    struct Manifolds {
        std::vector<uint32_t> a, b, npts, feature;
        std::vector<float> position, normal, depth;
    };
    Status contacts(Manifolds& m) const {
        uint32_t count = 0;
        Status s(axm_world_get_manifolds(world_, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                                         &count));
        if (!s && s.code() != AXM_OUT_OF_RANGE) return s;
        m.a.resize(count); m.b.resize(count); m.npts.resize(count); m.feature.resize(4ull * count);
        m.position.resize(12ull * count); m.normal.resize(12ull * count); m.depth.resize(4ull * count);
        return Status(axm_world_get_manifolds(world_, m.a.data(), m.b.data(), m.npts.data(), m.position.data(),
                                              m.normal.data(), m.depth.data(), m.feature.data(), count, &count));
    }

    axm_world handle() const noexcept { return world_; }

private:
    explicit PhysicsWorld(axm_world w) noexcept : world_(w) {}
    axm_world world_ = nullptr;
    axm_step_stats last_{};
    float configDt_ = 1.0f / 60.0f;  // PhysicsWorldConfig defaults
    uint32_t configIters_ = 8;
};

}  // namespace axiom::dynamics
