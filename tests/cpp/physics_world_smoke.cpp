// Compile-and-run check of the C++20 API (include/axiom/dynamics/physics_world.hpp).
// Without a GPU it checks the host-side pieces and that world creation fails
// with a status (never a CPU fallback); with a GPU it steps a sphere resting on
// the ground and checks the normal impulse carries the weight.
#include <cmath>
#include <cstdio>
#include <vector>

#include "axiom/dynamics/physics_world.hpp"

namespace {
// Field names of gui::RigidBodyData (include/axiom/gui/body_inspector.hpp:58-90).
struct V3 { float x, y, z; };
struct Q4 { float x, y, z, w; };
struct RigidBodyLike {
    V3 position{0, 0, 0};
    Q4 rotation{0, 0, 0, 1};
    V3 linearVelocity{0, 0, 0};
    V3 angularVelocity{0, 0, 0};
    float mass = 1.0f;
    V3 inertiaTensor{1, 1, 1};
    int shapeType = 0;
    V3 shapeExtents{1, 1, 1};
};
// The same with gui::RigidBodyData's body type (include/axiom/gui/body_inspector.hpp:13-17).
enum class BodyTypeLike : uint32_t { Static = 0, Dynamic = 1, Kinematic = 2 };
struct MaterialLike { float restitution = 0.5f, friction = 0.5f, rollingFriction = 0.0f, density = 1000.0f; };
struct FilterLike { uint32_t categoryBits = 1, maskBits = 0xFFFF; int16_t groupIndex = 0; };
struct TypedBodyLike : RigidBodyLike {
    BodyTypeLike type = BodyTypeLike::Dynamic;
    float linearDamping = 0.01f, angularDamping = 0.05f;
    MaterialLike material;
    FilterLike filter;
};
// Field names of gui::PhysicsWorldConfig (physics_panel.hpp:34-43) and
// debug::DebugContactPoint (physics_debug_draw.hpp:128-132).
struct ConfigLike {
    V3 gravity{0.0f, -9.81f, 0.0f};
    float timeStep = 1.0f / 120.0f;
    uint32_t velocityIterations = 10;
    uint32_t positionIterations = 3;
    bool allowSleep = true;
    float sleepLinearThreshold = 0.01f, sleepAngularThreshold = 0.01f, sleepTimeThreshold = 0.5f;
};
struct ContactLike {
    V3 position;
    V3 normal;
    float penetrationDepth = 0.0f;
};
struct StatsLike {
    uint32_t totalBodies, activeBodies, sleepingBodies, staticBodies, dynamicBodies, kinematicBodies;
    uint32_t contactPointCount, constraintCount, islandCount;
    float totalStepTime, broadphaseTime, narrowphaseTime, solverTime, integrationTime;
};
int fails = 0;
void check(bool c, const char* what) {
    if (!c) { std::printf("FAIL %s\n", what); ++fails; }
}
}  // namespace

int main() {
    using namespace axiom::dynamics;
    BodySet bodies;
    const float r[3] = {0.5f, 0, 0}, at[3] = {0, 0.5f, 0}, q[4] = {0, 0, 0, 1}, zero[3] = {0, 0, 0};
    check(bodies.add(ShapeType::Sphere, r, at, q, zero, zero).ok(), "add sphere");
    const float bad[3] = {-1, 0, 0};
    check(bodies.add(ShapeType::Sphere, bad, at, q, zero, zero).code() == AXM_INVALID_SHAPE, "invalid shape");
    RigidBodyLike rb;
    rb.position = {3, 1, 0};
    rb.mass = 0.0f;  // static
    rb.shapeType = 1;
    rb.shapeExtents = {1, 1, 1};
    check(bodies.addRigidBody(rb).ok(), "addRigidBody");
    check(bodies.size() == 2 && bodies.inverseMass(1) == 0.0f, "static body");
    // a Static body keeps zero inverse mass although its mass field holds the default 1 kg
    TypedBodyLike tb;
    tb.position = {-3, 0.4f, 0};
    tb.shapeType = 1;
    tb.shapeExtents = {0.5f, 0.4f, 0.5f};
    tb.type = BodyTypeLike::Static;
    check(tb.mass == 1.0f && bodies.addRigidBody(tb).ok(), "addRigidBody static");
    check(bodies.size() == 3 && bodies.inverseMass(2) == 0.0f, "static type overrides mass");
    // a Kinematic body moves at its own velocity, unaffected by gravity (away from everything here)
    TypedBodyLike kin = tb;
    kin.type = BodyTypeLike::Kinematic;
    kin.position = {0, 20, 10};
    kin.linearVelocity = {1, 0, 0};
    check(bodies.addRigidBody(kin).ok() && bodies.size() == 4 && bodies.kinematic(3) && bodies.inverseMass(3) == 0.0f,
          "kinematic body");
    TypedBodyLike dyn0 = tb;
    dyn0.type = BodyTypeLike::Dynamic;
    dyn0.mass = 0.0f;
    check(bodies.addRigidBody(dyn0).code() == AXM_INVALID_RIGID_BODY, "dynamic needs mass");
    const float m = 1000.0f * 4.0f / 3.0f * 3.14159265f * 0.125f;
    check(std::fabs(1.0f / bodies.inverseMass(0) - m) < 1e-3f * m, "sphere mass");

    Status s;
    axm_world_desc d = PhysicsWorld::defaultDesc();
    auto world = PhysicsWorld::create(d, &s);
    if (axm_device_count() == 0) {
        check(!world && (s.code() == AXM_GPU_OPERATION_FAILED || s.code() == AXM_GPU_INVALID_OPERATION),
              "no device: creation fails with a GPU status");
        std::printf("%s (no device: %s)\n", fails ? "FAILED" : "OK", s.message());
        return fails ? 1 : 0;
    }
    check(world != nullptr, "create");
    check(world->setBodies(bodies).ok(), "setBodies");
    for (int k = 0; k < 60 && world; ++k) check(world->step(1.0f / 120.0f, 10).ok(), "step");
    check(world->run(1.0f / 120.0f, 10, 240, true).ok(), "run graph");
    BodySet out;
    check(world->getBodies(out).ok() && out.size() == 4, "getBodies");
    check(std::fabs(out.position(3)[0] - 300.0f / 120.0f) < 1e-3f && out.position(3)[1] == 20.0f &&
              out.linearVelocity(3)[0] == 1.0f, "kinematic body moves at its velocity");
    check(out.position(2)[1] == 0.4f && out.linearVelocity(2)[1] == 0.0f, "static box stays put");
    check(std::fabs(out.linearVelocity(0)[1]) < 1e-3f, "sphere at rest");
    check(std::fabs(out.position(0)[1] - 0.5f) < 0.01f, "sphere on the ground");
    PhysicsWorld::Manifolds mf;
    check(world->contacts(mf).ok() && mf.a.size() == 1 && mf.npts[0] == 1, "one contact");
    std::vector<ContactLike> dc;
    check(world->debugContacts(dc).ok() && dc.size() == 1 && dc[0].normal.y < -0.99f &&
              std::fabs(dc[0].position.y) < 0.01f, "debug contact under the sphere");
    StatsLike st{};
    world->fillStats(st, 4, 2, 1);
    check(st.contactPointCount == 1 && st.dynamicBodies == 1 && st.kinematicBodies == 1 && st.islandCount == 1 &&
              st.sleepingBodies == 0 && st.activeBodies == 1, "stats");
    // PhysicsWorldConfig with allowSleep: the resting sphere falls asleep after sleepTimeThreshold
    ConfigLike cfg;
    check(world->applyConfig(cfg).ok(), "applyConfig");
    axm_world_desc got{};
    check(axm_world_get_desc(world->handle(), &got) == AXM_SUCCESS && got.allow_sleep == 1u && got.sleep_time == 0.5f,
          "config applied");
    for (int k = 0; k < 90; ++k) check(world->stepConfigured().ok(), "step with the applied config");
    world->fillStats(st, 4, 2, 1);
    check(st.sleepingBodies == 1 && st.activeBodies == 0 && st.islandCount == 1, "sleeping stats");
    PhysicsWorld moved(std::move(*world));
    check(moved.handle() != nullptr && world->handle() == nullptr, "move");
    std::printf("%s\n", fails ? "FAILED" : "OK");
    return fails ? 1 : 0;
}
