// Seeded synthetic scenes (host only) and the status-string table.
//
// The random stream restates the reference's DeterministicRNG
// (include/axiom/math/random.hpp:19-68): PCG-XSH-RR (O'Neill 2014) with
// state = seed | 1, ten warm-up draws, multiplier 6364136223846793005,
// increment 1442695040888963407, nextFloat = float(u32) / 2^32 (which rounds to
// exactly 1.0f for u32 >= 0xFFFFFF80) and nextFloat(min,max) = min + f*(max-min).
// Orientations use Shoemake's (1992) uniform quaternion; normal deviates use the
// Box-Muller cosine branch in double precision.
#include <cmath>
#include <cstring>
#include <vector>

#include "axiom/axm.h"

namespace {

class Pcg {
public:
    explicit Pcg(uint64_t seed) : state_(seed | 1ULL) {
        for (int i = 0; i < 10; ++i) next();
    }
    uint32_t next() {
        const uint64_t old = state_;
        state_ = old * 6364136223846793005ULL + 1442695040888963407ULL;
        const uint32_t xs = static_cast<uint32_t>(((old >> 18U) ^ old) >> 27U);
        const uint32_t rot = static_cast<uint32_t>(old >> 59U);
        return (xs >> rot) | (xs << ((~rot + 1U) & 31));
    }
    float nextFloat() { return static_cast<float>(next()) / static_cast<float>(0x100000000ULL); }
    float nextFloat(float lo, float hi) { return lo + nextFloat() * (hi - lo); }
    float normal(float sigma) {
        double u1 = nextFloat();
        const double u2 = nextFloat();
        if (u1 < 1e-7) u1 = 1e-7;
        return static_cast<float>(std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2) * sigma);
    }
    void quat(float* q) {
        const double u1 = nextFloat(), u2 = nextFloat(), u3 = nextFloat();
        const double s1 = std::sqrt(1.0 - u1), s2 = std::sqrt(u1);
        double x = s1 * std::sin(2.0 * M_PI * u2), y = s1 * std::cos(2.0 * M_PI * u2);
        double z = s2 * std::sin(2.0 * M_PI * u3), w = s2 * std::cos(2.0 * M_PI * u3);
        const double l = std::sqrt(x * x + y * y + z * z + w * w);
        q[0] = (float)(x / l); q[1] = (float)(y / l); q[2] = (float)(z / l); q[3] = (float)(w / l);
    }

private:
    uint64_t state_;
};

constexpr double kDensity = 1000.0;

struct Body {
    float pos[3], quat[4], v[3], w[3], im, ii[3], prm[3];
    uint32_t type, world;
};

// Mass and body-frame principal inertia (closed forms: solid sphere; box on
// half extents; capsule along Y = cylinder of height 2h + two hemispheres, each
// hemisphere moved from its centroid (3r/8 off the flat face) by the parallel
// axis theorem).  Returns false for an unknown type or non-positive size.
bool mass_inertia(uint32_t type, const float* prm, double density, double& m, double I[3]) {
    m = 0.0;
    I[0] = I[1] = I[2] = 0.0;
    if (!(density > 0.0)) return false;
    if (type == AXM_SHAPE_SPHERE) {
        const double r = prm[0];
        if (!(r > 0.0)) return false;
        m = density * 4.0 / 3.0 * M_PI * r * r * r;
        I[0] = I[1] = I[2] = 0.4 * m * r * r;
    } else if (type == AXM_SHAPE_BOX) {
        const double x = prm[0], y = prm[1], z = prm[2];
        if (!(x > 0.0 && y > 0.0 && z > 0.0)) return false;
        m = density * 8.0 * x * y * z;
        I[0] = m / 3.0 * (y * y + z * z);
        I[1] = m / 3.0 * (x * x + z * z);
        I[2] = m / 3.0 * (x * x + y * y);
    } else if (type == AXM_SHAPE_CAPSULE) {
        const double r = prm[0], h = prm[1];
        if (!(r > 0.0 && h >= 0.0)) return false;
        const double mc = density * M_PI * r * r * (2.0 * h);
        const double ms = density * 4.0 / 3.0 * M_PI * r * r * r;
        m = mc + ms;
        I[1] = mc * r * r / 2.0 + ms * 0.4 * r * r;
        I[0] = I[2] = mc * (r * r / 4.0 + (2.0 * h) * (2.0 * h) / 12.0) + ms * (0.4 * r * r + h * h + 0.75 * h * r);
    } else {
        return false;
    }
    return true;
}

void mass_props(Body& b) {
    double m, I[3];
    mass_inertia(b.type, b.prm, kDensity, m, I);
    b.im = (float)(1.0 / m);
    for (int k = 0; k < 3; ++k) b.ii[k] = (float)(1.0 / I[k]);
}

float bound_radius(const Body& b) {
    if (b.type == AXM_SHAPE_SPHERE) return b.prm[0];
    if (b.type == AXM_SHAPE_BOX) return std::sqrt(b.prm[0] * b.prm[0] + b.prm[1] * b.prm[1] + b.prm[2] * b.prm[2]);
    return b.prm[0] + b.prm[1];
}

void draw_shape(Pcg& r, Body& b, uint32_t type) {
    b.type = type;
    b.prm[0] = b.prm[1] = b.prm[2] = 0.0f;
    if (type == AXM_SHAPE_SPHERE) {
        b.prm[0] = r.nextFloat(0.25f, 0.75f);
    } else if (type == AXM_SHAPE_BOX) {
        b.prm[0] = r.nextFloat(0.25f, 0.75f);
        b.prm[1] = r.nextFloat(0.25f, 0.75f);
        b.prm[2] = r.nextFloat(0.25f, 0.75f);
    } else {
        b.prm[0] = r.nextFloat(0.2f, 0.5f);
        b.prm[1] = r.nextFloat(0.2f, 1.0f);
    }
    mass_props(b);
}

// Mixed spheres / boxes, uniform in a cube of side 30*(n/1000)^(1/3) above the ground.
void gen_mixed(uint64_t seed, uint32_t n, uint32_t world, std::vector<Body>& out) {
    Pcg r(seed);
    const float L = 30.0f * (float)std::cbrt((double)n / 1000.0);
    for (uint32_t i = 0; i < n; ++i) {
        Body b{};
        b.world = world;
        draw_shape(r, b, (i & 1u) ? AXM_SHAPE_BOX : AXM_SHAPE_SPHERE);
        b.pos[0] = r.nextFloat(-0.5f * L, 0.5f * L);
        b.pos[1] = r.nextFloat(0.0f, L) + 1.5f;
        b.pos[2] = r.nextFloat(-0.5f * L, 0.5f * L);
        r.quat(b.quat);
        for (int k = 0; k < 3; ++k) b.v[k] = r.normal(1.0f);
        for (int k = 0; k < 3; ++k) b.w[k] = r.normal(0.5f);
        out.push_back(b);
    }
}

// Dense stacking: thirds spheres / boxes / capsules on a 1.5 m lattice over a
// 20 x 20 m footprint inside four wall planes, random orientation, at rest.
void gen_pile(uint64_t seed, uint32_t n, std::vector<Body>& out) {
    Pcg r(seed);
    const float s = 1.5f;
    const int per_row = 13;
    for (uint32_t i = 0; i < n; ++i) {
        Body b{};
        b.world = 0;
        draw_shape(r, b, i % 3u);
        const uint32_t layer = i / (per_row * per_row), cell = i % (per_row * per_row);
        const float jx = r.nextFloat(-0.05f, 0.05f), jz = r.nextFloat(-0.05f, 0.05f);
        b.pos[0] = -9.0f + s * (float)(cell % per_row) + jx;
        b.pos[1] = 0.8f + s * (float)layer;
        b.pos[2] = -9.0f + s * (float)(cell / per_row) + jz;
        r.quat(b.quat);
        out.push_back(b);
    }
}

// Sparse large scene: 50/50 spheres / boxes on a jittered 4 m lattice filling a
// 400 m cube above the ground (about 2% of bodies start overlapping).
void gen_sparse(uint64_t seed, uint32_t n, std::vector<Body>& out) {
    Pcg r(seed);
    // 4 m lattice (1M bodies fill the 400 m cube); smaller n keep the density
    // and shrink the cube, so every sample has the same ~2 % initial overlap
    uint32_t side = (uint32_t)std::ceil(std::cbrt((double)n));
    if (side < 1) side = 1;
    const float s = 4.0f;
    const float J = 0.45f * s;
    const float half = 0.5f * s * (float)side;
    for (uint32_t i = 0; i < n; ++i) {
        Body b{};
        b.world = 0;
        draw_shape(r, b, (i & 1u) ? AXM_SHAPE_BOX : AXM_SHAPE_SPHERE);
        const uint32_t ix = i % side, iy = (i / side) % side, iz = i / (side * side);
        b.pos[0] = -half + s * ((float)ix + 0.5f) + r.nextFloat(-J, J);
        b.pos[1] = 1.5f + s * ((float)iy + 0.5f) + r.nextFloat(-J, J);
        b.pos[2] = -half + s * ((float)iz + 0.5f) + r.nextFloat(-J, J);
        r.quat(b.quat);
        for (int k = 0; k < 3; ++k) b.v[k] = r.normal(1.0f);
        for (int k = 0; k < 3; ++k) b.w[k] = r.normal(0.5f);
        out.push_back(b);
    }
}

}  // namespace

extern "C" {

const char* axm_status_string(axm_status s) {
    switch (s) {
    case AXM_SUCCESS: return "Success";
    case AXM_OUT_OF_MEMORY: return "Out of memory";
    case AXM_NULL_POINTER: return "Unexpected null pointer";
    case AXM_INVALID_SHAPE: return "Invalid collision shape";
    case AXM_GJK_FAILED_TO_CONVERGE: return "GJK algorithm failed to converge";
    case AXM_EPA_FAILED_TO_CONVERGE: return "EPA algorithm failed to converge";
    case AXM_INVALID_RIGID_BODY: return "Invalid rigid body properties";
    case AXM_CONSTRAINT_SOLVER_FAILURE: return "Constraint solver failed to converge";
    case AXM_BUFFER_ALLOCATION_FAILED: return "GPU buffer allocation failed";
    case AXM_GPU_INVALID_OPERATION: return "Invalid GPU operation or state";
    case AXM_GPU_TIMEOUT: return "GPU operation timed out";
    case AXM_GPU_OPERATION_FAILED: return "GPU operation failed";
    case AXM_INVALID_PARAMETER: return "Invalid parameter";
    case AXM_OUT_OF_RANGE: return "Value out of range";
    default: return "Unknown error";
    }
}

void axm_world_desc_default(axm_world_desc* d) {
    std::memset(d, 0, sizeof(*d));
    d->gravity[0] = 0.0f;
    d->gravity[1] = -9.81f;
    d->gravity[2] = 0.0f;
    d->friction = 0.5f;
    d->restitution = 0.1f;
    d->rest_threshold = 1.0f;
    d->baumgarte = 0.2f;
    d->slop = 0.005f;
    d->max_bias = 2.0f;
    d->margin = 0.02f;
    d->lin_damp = 0.0f;
    d->ang_damp = 0.0f;
    d->cell_size = 0.0f;
    d->num_planes = 1;
    d->planes[0][0] = 0.0f;
    d->planes[0][1] = 1.0f;
    d->planes[0][2] = 0.0f;
    d->planes[0][3] = 0.0f;
    d->warm_start = AXM_WARM_CACHE;
    d->record_residual = 0;
    d->softness = 0.1f;
    d->allow_sleep = 0;
    d->sleep_linear = 0.01f;
    d->sleep_angular = 0.01f;
    d->sleep_time = 0.5f;
    d->island_solve = 1;
    d->auto_grow = 1;
    d->residual_tol = 0.0f;
    d->residual_every = 1;
    d->max_iterations = 200;
    d->device = 0;
}

axm_status axm_mass_properties(uint32_t type, const float* params, float density, float* inv_mass,
                               float* inv_inertia) {
    if (!params || !inv_mass || !inv_inertia) return AXM_NULL_POINTER;
    double m, I[3];
    if (!mass_inertia(type, params, (double)density, m, I)) return AXM_INVALID_SHAPE;
    *inv_mass = (float)(1.0 / m);
    for (int k = 0; k < 3; ++k) inv_inertia[k] = (float)(1.0 / I[k]);
    return AXM_SUCCESS;
}

void axm_rng_stream(uint64_t seed, uint32_t n, uint32_t* u32_out, float lo, float hi, float* f_out) {
    if (u32_out) {
        Pcg r(seed);
        for (uint32_t i = 0; i < n; ++i) u32_out[i] = r.next();
    }
    if (f_out) {
        Pcg r(seed);
        for (uint32_t i = 0; i < n; ++i) f_out[i] = r.nextFloat(lo, hi);
    }
}

axm_status axm_scene_generate(uint32_t config, uint64_t seed, uint32_t n, uint32_t n_worlds, axm_body_soa* out,
                              uint32_t* count, axm_world_desc* desc) {
    std::vector<Body> bodies;
    axm_world_desc d;
    axm_world_desc_default(&d);
    switch (config) {
    case 0:
        gen_mixed(seed, n ? n : 1000u, 0, bodies);
        d.friction = 0.5f;
        break;
    case 1:
        gen_pile(seed, n ? n : 100000u, bodies);
        d.friction = 0.6f;
        d.num_planes = 5;
        {
            const float walls[4][4] = {{1, 0, 0, -10}, {-1, 0, 0, -10}, {0, 0, 1, -10}, {0, 0, -1, -10}};
            for (int k = 0; k < 4; ++k)
                for (int c = 0; c < 4; ++c) d.planes[1 + k][c] = walls[k][c];
        }
        break;
    case 2:
        gen_sparse(seed, n ? n : 1000000u, bodies);
        d.friction = 0.5f;
        break;
    case 3: {
        const uint32_t W = n_worlds ? n_worlds : 4096u;
        Pcg sizes(seed);
        const double lo = std::log(16.0), hi = std::log(4096.0);
        for (uint32_t w = 0; w < W; ++w) {
            uint32_t nw = (uint32_t)std::floor(std::exp(lo + (double)sizes.nextFloat() * (hi - lo)));
            if (nw < 16) nw = 16;
            if (nw > 4096) nw = 4096;
            // DeterministicRNG starts from seed | 1, so seeds that differ only in
            // bit 0 give one stream: world seeds step by 2
            gen_mixed(seed + 2ull * (w + 1), nw, w, bodies);
        }
        d.friction = 0.5f;
        break;
    }
    default:
        return AXM_INVALID_PARAMETER;
    }
    float rmax = 0.0f;
    for (const Body& b : bodies) rmax = std::fmax(rmax, bound_radius(b));
    d.cell_size = 2.0f * (rmax + d.margin);
    d.origin[0] = -512.0f * d.cell_size;
    d.origin[1] = -16.0f * d.cell_size;
    d.origin[2] = -512.0f * d.cell_size;
    if (count) *count = (uint32_t)bodies.size();
    if (desc) *desc = d;
    if (!out || !out->pos) return AXM_SUCCESS;
    for (size_t i = 0; i < bodies.size(); ++i) {
        const Body& b = bodies[i];
        std::memcpy(out->pos + 3 * i, b.pos, 12);
        std::memcpy(out->quat + 4 * i, b.quat, 16);
        std::memcpy(out->lin_vel + 3 * i, b.v, 12);
        std::memcpy(out->ang_vel + 3 * i, b.w, 12);
        out->inv_mass[i] = b.im;
        std::memcpy(out->inv_inertia + 3 * i, b.ii, 12);
        out->shape_type[i] = b.type;
        std::memcpy(out->shape_params + 3 * i, b.prm, 12);
        if (out->world) out->world[i] = b.world;
    }
    return AXM_SUCCESS;
}

}  // extern "C"
