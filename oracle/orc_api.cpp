// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// extern "C" entry points of the oracle shared library (oracle/liborc.so).
// Only tests/, the smoke check and the CPU-baseline leg of the benchmark load
// this library; the product never links or calls it.
#include <chrono>
#include <cstdio>
#include <memory>
#include <thread>

#include "orc_mlp.hpp"
#include "orc_world.hpp"

using namespace orc;

namespace {

struct Handle {
    int fp64;
    std::unique_ptr<World<float>> f;
    std::unique_ptr<World<double>> d;
};

template <class T> void set_bodies(World<T>& W, uint32_t n, const float* pos, const float* quat, const float* vel,
                                   const float* angvel, const float* inv_mass, const float* inv_inertia,
                                   const uint32_t* type, const float* prm, const uint32_t* world) {
    W.n = n;
    W.x.resize(n);
    W.q.resize(n);
    W.v.resize(n);
    W.w.resize(n);
    W.im.resize(n);
    W.ii.resize(n);
    W.type.resize(n);
    W.prm.resize(n);
    W.world.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        W.x[i] = {T(pos[3 * i]), T(pos[3 * i + 1]), T(pos[3 * i + 2])};
        W.q[i] = {T(quat[4 * i]), T(quat[4 * i + 1]), T(quat[4 * i + 2]), T(quat[4 * i + 3])};
        W.v[i] = {T(vel[3 * i]), T(vel[3 * i + 1]), T(vel[3 * i + 2])};
        W.w[i] = {T(angvel[3 * i]), T(angvel[3 * i + 1]), T(angvel[3 * i + 2])};
        W.im[i] = T(inv_mass[i]);
        // static bodies (inverse mass 0) carry no inverse inertia
        W.ii[i] = inv_mass[i] == 0.0f ? V3<T>{T(0), T(0), T(0)}
                                      : V3<T>{T(inv_inertia[3 * i]), T(inv_inertia[3 * i + 1]), T(inv_inertia[3 * i + 2])};
        W.type[i] = type[i];
        W.prm[i] = {T(prm[3 * i]), T(prm[3 * i + 1]), T(prm[3 * i + 2])};
        W.world[i] = world ? world[i] : 0u;
    }
    W.cache.clear();
    W.nsteps = 0;
    W.reset_sleep();
    W.reset_extras();
}

template <class T> void get_bodies(const World<T>& W, float* pos, float* quat, float* vel, float* angvel) {
    for (uint32_t i = 0; i < W.n; ++i) {
        if (pos) { pos[3 * i] = (float)W.x[i].x; pos[3 * i + 1] = (float)W.x[i].y; pos[3 * i + 2] = (float)W.x[i].z; }
        if (quat) {
            quat[4 * i] = (float)W.q[i].x; quat[4 * i + 1] = (float)W.q[i].y;
            quat[4 * i + 2] = (float)W.q[i].z; quat[4 * i + 3] = (float)W.q[i].w;
        }
        if (vel) { vel[3 * i] = (float)W.v[i].x; vel[3 * i + 1] = (float)W.v[i].y; vel[3 * i + 2] = (float)W.v[i].z; }
        if (angvel) {
            angvel[3 * i] = (float)W.w[i].x; angvel[3 * i + 1] = (float)W.w[i].y; angvel[3 * i + 2] = (float)W.w[i].z;
        }
    }
}

template <class T> uint32_t get_manifolds(const World<T>& W, uint32_t* a, uint32_t* b, uint32_t* np, float* p, float* nrm,
                                          float* depth, uint32_t* fid, uint32_t cap) {
    const uint32_t M = (uint32_t)W.man.size();
    for (uint32_t m = 0; m < M && m < cap; ++m) {
        if (a) a[m] = W.man_a[m];
        if (b) b[m] = W.man_b[m];
        if (np) np[m] = W.man[m].np;
        for (uint32_t k = 0; k < 4; ++k) {
            const bool on = k < W.man[m].np;
            const CPt<T>& c = W.man[m].pt[k];
            const size_t s = (size_t)m * 4 + k;
            if (p) { p[3 * s] = on ? (float)c.p.x : 0.f; p[3 * s + 1] = on ? (float)c.p.y : 0.f; p[3 * s + 2] = on ? (float)c.p.z : 0.f; }
            if (nrm) { nrm[3 * s] = on ? (float)c.n.x : 0.f; nrm[3 * s + 1] = on ? (float)c.n.y : 0.f; nrm[3 * s + 2] = on ? (float)c.n.z : 0.f; }
            if (depth) depth[s] = on ? (float)c.depth : 0.f;
            if (fid) fid[s] = on ? c.fid : 0u;
        }
    }
    return M;
}

template <class T> void load_manifolds(World<T>& W, uint32_t M, const uint32_t* a, const uint32_t* b, const uint32_t* np,
                                       const float* p, const float* nrm, const float* depth, const uint32_t* fid) {
    W.man_a.assign(a, a + M);
    W.man_b.assign(b, b + M);
    W.man.assign(M, Man<T>{});
    for (uint32_t m = 0; m < M; ++m) {
        W.man[m].np = np[m];
        for (uint32_t k = 0; k < np[m]; ++k) {
            const size_t s = (size_t)m * 4 + k;
            CPt<T>& c = W.man[m].pt[k];
            c.p = {T(p[3 * s]), T(p[3 * s + 1]), T(p[3 * s + 2])};
            c.n = {T(nrm[3 * s]), T(nrm[3 * s + 1]), T(nrm[3 * s + 2])};
            c.depth = T(depth[s]);
            c.fid = fid[s];
        }
    }
}

template <class T> Shape<T> mk_shape(uint32_t type, const float* prm, const float* pos, const float* quat) {
    Shape<T> s;
    s.type = type;
    s.prm = {T(prm[0]), T(prm[1]), T(prm[2])};
    s.c = {T(pos[0]), T(pos[1]), T(pos[2])};
    s.R = quat_to_mat(Q4<T>{T(quat[0]), T(quat[1]), T(quat[2]), T(quat[3])});
    return s;
}

template <class T> void out_man(const Man<T>& m, uint32_t* np, float* p, float* n, float* depth, uint32_t* fid) {
    *np = m.np;
    for (uint32_t k = 0; k < m.np; ++k) {
        p[3 * k] = (float)m.pt[k].p.x; p[3 * k + 1] = (float)m.pt[k].p.y; p[3 * k + 2] = (float)m.pt[k].p.z;
        n[3 * k] = (float)m.pt[k].n.x; n[3 * k + 1] = (float)m.pt[k].n.y; n[3 * k + 2] = (float)m.pt[k].n.z;
        depth[k] = (float)m.pt[k].depth;
        fid[k] = m.pt[k].fid;
    }
}

}  // namespace

#define DISPATCH(h, ...)             \
    do {                             \
        Handle* H = (Handle*)(h);    \
        if (H->fp64) {               \
            auto& W = *H->d;         \
            using T = double;        \
            (void)sizeof(T);         \
            __VA_ARGS__;             \
        } else {                     \
            auto& W = *H->f;         \
            using T = float;         \
            (void)sizeof(T);         \
            __VA_ARGS__;             \
        }                            \
    } while (0)

extern "C" {

uint32_t orc_params_size(void) { return (uint32_t)sizeof(Params); }
uint32_t orc_stats_size(void) { return (uint32_t)sizeof(Stats); }

// ---------------------------------------------------------------- math primitives (for known-answer tests)
// Hamilton product a (x) b of (x, y, z, w) quaternions (include/axiom/math/quat.hpp:29-227 convention).
void orc_quat_mul(const float* a, const float* b, float* out) {
    const Q4<float> r = qmul(Q4<float>{a[0], a[1], a[2], a[3]}, Q4<float>{b[0], b[1], b[2], b[3]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
// Rotates v by the unit quaternion q (through its rotation matrix, as every stage does).
void orc_quat_rotate(const float* q, const float* v, float* out) {
    const V3<float> r = mul(quat_to_mat(Q4<float>{q[0], q[1], q[2], q[3]}), V3<float>{v[0], v[1], v[2]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
// World inverse inertia R diag(d) R^T as (xx, yy, zz, xy, xz, yz).
void orc_world_inv_inertia(const float* q, const float* d, float* out) {
    const S3<float> s = world_inv_inertia(quat_to_mat(Q4<float>{q[0], q[1], q[2], q[3]}), V3<float>{d[0], d[1], d[2]});
    out[0] = s.xx; out[1] = s.yy; out[2] = s.zz; out[3] = s.xy; out[4] = s.xz; out[5] = s.yz;
}
void orc_tangents(const float* n, float* t1, float* t2) {
    V3<float> a, b;
    tangents(V3<float>{n[0], n[1], n[2]}, a, b);
    t1[0] = a.x; t1[1] = a.y; t1[2] = a.z;
    t2[0] = b.x; t2[1] = b.y; t2[2] = b.z;
}
uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }
uint32_t orc_pair_hash24(uint32_t a, uint32_t b) { return pair_hash24(a, b); }

void* orc_create(const Params* p, int fp64) {
    Handle* h = new Handle;
    h->fp64 = fp64;
    if (fp64) {
        h->d.reset(new World<double>);
        h->d->P = *p;
    } else {
        h->f.reset(new World<float>);
        h->f->P = *p;
    }
    return h;
}

void orc_destroy(void* h) { delete (Handle*)h; }

// Per-body extras after orc_set_bodies (NULL arrays keep the defaults): material (4n: friction,
// restitution, linear damping, angular damping), flags (n: bit 0 kinematic), filter (3n: category,
// mask, group as int32).
void orc_set_body_extras(void* h, const float* material, const uint32_t* flags, const uint32_t* filter) {
    DISPATCH(h, {
        W.reset_extras();
        for (uint32_t i = 0; i < W.n; ++i) {
            if (material) {
                W.mu_b[i] = T(material[4 * i]);
                W.e_b[i] = T(material[4 * i + 1]);
                W.lkeep[i] = T(1) - T(material[4 * i + 2]);
                W.akeep[i] = T(1) - T(material[4 * i + 3]);
            }
            if (flags) W.kin[i] = (flags[i] & 1u) ? 1 : 0;
            if (filter) {
                W.fcat[i] = filter[3 * i];
                W.fmask[i] = filter[3 * i + 1];
                W.fgroup[i] = (int32_t)filter[3 * i + 2];
            }
        }
        W.use_filter = filter != nullptr;
    });
}

// Friction bound of every contact point of the last step (M x 4, manifold order).
uint32_t orc_get_friction_bounds(void* h, float* out, uint32_t cap) {
    uint32_t M = 0;
    DISPATCH(h, {
        M = (uint32_t)W.man.size();
        if (out)
            for (size_t s = 0; s < (size_t)std::min(M, cap) * 4 && s < W.fbound.size(); ++s) out[s] = (float)W.fbound[s];
    });
    return M;
}

// Island label (0x7FFFFFFF: none), asleep flag and rest time of every body after the last step.
void orc_get_islands(void* h, uint32_t* label, uint32_t* asleep, float* rest) {
    DISPATCH(h, {
        for (uint32_t i = 0; i < W.n; ++i) {
            if (label) label[i] = W.label[i];
            if (asleep) asleep[i] = W.asleep[i];
            if (rest) rest[i] = (float)W.rest[i];
        }
    });
}

// Host threads of the oracle's parallel loops (results do not depend on the count).
void orc_set_threads(void* h, uint32_t threads) { DISPATCH(h, W.set_threads(threads)); }

// Residual-stopped sweeps (tol > 0): residual checked after every `every` sweeps, the
// solve stops once it is <= tol, at most max_iters sweeps (the step's iters argument is
// then ignored).  tol = 0 restores fixed sweep counts.
void orc_set_tolerance(void* h, float tol, uint32_t max_iters, uint32_t every) {
    DISPATCH(h, {
        W.residual_tol = T(tol);
        W.max_iters = max_iters;
        W.res_every = every ? every : 1u;
    });
}

void orc_set_params(void* h, const Params* p) { DISPATCH(h, W.P = *p); }

void orc_set_bodies(void* h, uint32_t n, const float* pos, const float* quat, const float* vel, const float* angvel,
                    const float* inv_mass, const float* inv_inertia, const uint32_t* type, const float* prm,
                    const uint32_t* world) {
    DISPATCH(h, set_bodies(W, n, pos, quat, vel, angvel, inv_mass, inv_inertia, type, prm, world));
}

void orc_get_bodies(void* h, float* pos, float* quat, float* vel, float* angvel) {
    DISPATCH(h, get_bodies(W, pos, quat, vel, angvel));
}

void orc_get_pre_solve_velocities(void* h, float* vel, float* angvel) {
    DISPATCH(h, {
        for (uint32_t i = 0; i < W.n && i < W.v_pre.size(); ++i) {
            vel[3 * i] = (float)W.v_pre[i].x; vel[3 * i + 1] = (float)W.v_pre[i].y; vel[3 * i + 2] = (float)W.v_pre[i].z;
            angvel[3 * i] = (float)W.w_pre[i].x; angvel[3 * i + 1] = (float)W.w_pre[i].y; angvel[3 * i + 2] = (float)W.w_pre[i].z;
        }
    });
}

void orc_step(void* h, float dt, uint32_t iters, Stats* st) {
    DISPATCH(h, {
        W.step(T(dt), iters);
        if (st) *st = W.st;
    });
}

// Steps n times; used by the CPU baseline timing.
double orc_run(void* h, float dt, uint32_t iters, uint32_t steps) {
    double t0 = 0;
    auto start = std::chrono::steady_clock::now();
    DISPATCH(h, {
        for (uint32_t s = 0; s < steps; ++s) W.step(T(dt), iters);
    });
    t0 = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    return t0;
}

// Runs several independent oracle worlds on `threads` host threads (CPU baseline, batched form).
double orc_run_many(void** hs, uint32_t count, float dt, uint32_t iters, uint32_t steps, uint32_t threads) {
    auto start = std::chrono::steady_clock::now();
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t) {
        pool.emplace_back([=]() {
            for (uint32_t k = t; k < count; k += threads) orc_run(hs[k], dt, iters, steps);
        });
    }
    for (auto& th : pool) th.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
}

// Runs only the broadphase on the current state.
uint32_t orc_broadphase(void* h) {
    uint32_t r = 0;
    DISPATCH(h, {
        std::vector<V3<T>> mn, mx;
        W.broadphase(mn, mx);
        r = (uint32_t)W.pair_a.size();
    });
    return r;
}

// Runs broadphase + narrowphase on the current state (no time advance).
uint32_t orc_collide_world(void* h) {
    uint32_t r = 0;
    DISPATCH(h, {
        std::vector<V3<T>> mn, mx;
        W.broadphase(mn, mx);
        W.narrowphase();
        r = (uint32_t)W.man.size();
    });
    return r;
}

uint32_t orc_get_pairs(void* h, uint32_t* a, uint32_t* b, uint32_t cap) {
    uint32_t r = 0;
    DISPATCH(h, {
        r = (uint32_t)W.pair_a.size();
        for (uint32_t k = 0; k < r && k < cap; ++k) { a[k] = W.pair_a[k]; b[k] = W.pair_b[k]; }
    });
    return r;
}

void orc_get_keys(void* h, uint64_t* keys, uint32_t* order) {
    DISPATCH(h, {
        for (uint32_t i = 0; i < W.n; ++i) { if (keys) keys[i] = W.keys[i]; if (order) order[i] = W.order[i]; }
    });
}

// O(n^2) reference pair set (same AABBs and plane test), in (a, b) sorted order.
uint32_t orc_brute_pairs(void* h, uint32_t* a, uint32_t* b, uint32_t cap) {
    uint32_t cnt = 0;
    DISPATCH(h, {
        std::vector<V3<T>> mn, mx;
        W.broadphase(mn, mx);
        for (uint32_t i = 0; i < W.n; ++i) {
            if (W.dynamic(i)) {
                for (uint32_t k = 0; k < W.P.num_planes; ++k) {
                    const V3<T> pn = {T(W.P.planes[k][0]), T(W.P.planes[k][1]), T(W.P.planes[k][2])};
                    const V3<T> c = W.x[i];
                    const V3<T> e = mx[i] - c;
                    const T s = dot(pn, c) - (std::fabs(pn.x) * e.x + std::fabs(pn.y) * e.y + std::fabs(pn.z) * e.z);
                    if (s <= T(W.P.planes[k][3])) {
                        if (cnt < cap) { a[cnt] = i; b[cnt] = PLANE_BIT | k; }
                        ++cnt;
                    }
                }
            }
            for (uint32_t j = i + 1; j < W.n; ++j) {
                if (W.world[i] != W.world[j]) continue;
                if (!W.dynamic(i) && !W.dynamic(j)) continue;
                if (mn[i].x <= mx[j].x && mx[i].x >= mn[j].x && mn[i].y <= mx[j].y && mx[i].y >= mn[j].y &&
                    mn[i].z <= mx[j].z && mx[i].z >= mn[j].z) {
                    if (cnt < cap) { a[cnt] = i; b[cnt] = j; }
                    ++cnt;
                }
            }
        }
    });
    return cnt;
}

uint32_t orc_get_manifolds(void* h, uint32_t* a, uint32_t* b, uint32_t* np, float* p, float* n, float* depth,
                           uint32_t* fid, uint32_t cap) {
    uint32_t r = 0;
    DISPATCH(h, r = get_manifolds(W, a, b, np, p, n, depth, fid, cap));
    return r;
}

uint32_t orc_get_colours(void* h, uint32_t* colour, uint32_t* solve_order, uint32_t cap) {
    uint32_t K = 0;
    DISPATCH(h, {
        for (uint32_t m = 0; m < W.colour.size() && m < cap; ++m) {
            if (colour) colour[m] = W.colour[m];
            if (solve_order) solve_order[m] = W.solve_order[m];
        }
        K = W.st.colours;
    });
    return K;
}

// lambda per manifold (pair order) x 4 points x (n, t1, t2) after the last solve
void orc_get_lambda(void* h, float* out, uint32_t cap_manifolds) {
    DISPATCH(h, {
        const size_t M = W.man.size() < cap_manifolds ? W.man.size() : cap_manifolds;
        for (size_t s = 0; s < M * 12; ++s) out[s] = (float)W.lambda[s];
    });
}

// Warm-start network inputs of the last step: manifold order x 4 points x 24.
uint32_t orc_get_features(void* h, float* out, uint32_t cap_manifolds) {
    uint32_t M = 0;
    DISPATCH(h, {
        M = (uint32_t)W.man.size();
        if (out && cap_manifolds >= M)
            for (size_t i = 0; i < (size_t)M * 96 && i < W.feat.size(); ++i) out[i] = (float)W.feat[i];
    });
    return M;
}

uint32_t orc_get_residuals(void* h, float* out, uint32_t cap) {
    uint32_t r = 0;
    DISPATCH(h, {
        r = (uint32_t)W.residual_hist.size();
        for (uint32_t k = 0; k < r && k < cap; ++k) out[k] = (float)W.residual_hist[k];
    });
    return r;
}

// One step that takes its manifolds (pair order) from the caller instead of
// running the broadphase and narrowphase: velocity integration, colouring,
// solve (warm start from `warm` when given, else from the oracle's cache) and
// position integration.
void orc_step_given(void* h, float dt, uint32_t iters, uint32_t M, const uint32_t* a, const uint32_t* b,
                    const uint32_t* np, const float* p, const float* n, const float* depth, const uint32_t* fid,
                    const float* warm, const uint32_t* colour_in, const float* bounds, Stats* st) {
    DISPATCH(h, {
        W.st = Stats{};
        load_manifolds(W, M, a, b, np, p, n, depth, fid);
        W.integrate_velocities(T(dt));
        W.v_pre = W.v;
        W.w_pre = W.w;
        W.colour_manifolds(colour_in);
        std::vector<T> wt;
        if (warm) {
            wt.resize((size_t)M * 12);
            for (size_t s = 0; s < wt.size(); ++s) wt[s] = T(warm[s]);
        }
        std::vector<T> bt;
        if (bounds) {
            bt.resize((size_t)M * 4);
            for (size_t s = 0; s < bt.size(); ++s) bt[s] = T(bounds[s]);
            W.fb_in = bt.data();
        }
        W.solve(T(dt), iters, warm ? wt.data() : nullptr);
        W.fb_in = nullptr;
        W.integrate_positions(T(dt));
        W.islands_and_sleep(T(dt));
        ++W.nsteps;
        W.st.manifolds = M;
        if (st) *st = W.st;
    });
}

// Pairwise geometry entry points.
int orc_collide(int fp64, uint32_t ta, const float* pa, const float* xa, const float* qa, uint32_t tb, const float* pb,
                const float* xb, const float* qb, float margin, uint32_t* np, float* p, float* n, float* depth,
                uint32_t* fid) {
    if (fp64) {
        Man<double> m;
        int s = collide(m, mk_shape<double>(ta, pa, xa, qa), mk_shape<double>(tb, pb, xb, qb), (double)margin);
        out_man(m, np, p, n, depth, fid);
        return s;
    }
    Man<float> m;
    int s = collide(m, mk_shape<float>(ta, pa, xa, qa), mk_shape<float>(tb, pb, xb, qb), margin);
    out_man(m, np, p, n, depth, fid);
    return s;
}

void orc_collide_plane(int fp64, uint32_t ta, const float* pa, const float* xa, const float* qa, const float* plane,
                       float margin, uint32_t* np, float* p, float* n, float* depth, uint32_t* fid) {
    if (fp64) {
        Man<double> m;
        collide_plane(m, mk_shape<double>(ta, pa, xa, qa), V3<double>{plane[0], plane[1], plane[2]}, (double)plane[3],
                      (double)margin);
        out_man(m, np, p, n, depth, fid);
        return;
    }
    Man<float> m;
    collide_plane(m, mk_shape<float>(ta, pa, xa, qa), V3<float>{plane[0], plane[1], plane[2]}, plane[3], margin);
    out_man(m, np, p, n, depth, fid);
}

// GJK between the cores (box polytope / capsule segment / sphere centre) of two
// shapes, followed by EPA when the cores intersect.  out: dist_or_depth,
// pA[3], pB[3], n[3]; returns 0 separated, 1 penetrating (EPA ok), 301 / 302 on failure.
int orc_gjk_epa_capped(int fp64, uint32_t ta, const float* pa, const float* xa, const float* qa, uint32_t tb,
                       const float* pb, const float* xb, const float* qb, int gjk_iters, int epa_iters, float* out) {
    auto run = [&](auto tag) -> int {
        using T = decltype(tag);
        const Shape<T> A = mk_shape<T>(ta, pa, xa, qa), B = mk_shape<T>(tb, pb, xb, qb);
        const GjkOut<T> g = gjk(A, B, gjk_iters);
        if (g.status == 1) {
            const EpaOut<T> e = epa(A, B, g, epa_iters);
            out[0] = (float)e.depth;
            out[1] = (float)e.pA.x; out[2] = (float)e.pA.y; out[3] = (float)e.pA.z;
            out[4] = (float)e.pB.x; out[5] = (float)e.pB.y; out[6] = (float)e.pB.z;
            out[7] = (float)e.n.x; out[8] = (float)e.n.y; out[9] = (float)e.n.z;
            return e.status == 0 ? 1 : GEOM_EPA_FAIL;
        }
        out[0] = (float)g.dist;
        out[1] = (float)g.pA.x; out[2] = (float)g.pA.y; out[3] = (float)g.pA.z;
        out[4] = (float)g.pB.x; out[5] = (float)g.pB.y; out[6] = (float)g.pB.z;
        out[7] = out[8] = out[9] = 0.f;
        return g.status == 0 ? 0 : GEOM_GJK_FAIL;
    };
    return fp64 ? run(0.0) : run(0.0f);
}
int orc_gjk_epa(int fp64, uint32_t ta, const float* pa, const float* xa, const float* qa, uint32_t tb, const float* pb,
                const float* xb, const float* qb, float* out) {
    return orc_gjk_epa_capped(fp64, ta, pa, xa, qa, tb, pb, xb, qb, 32, 48, out);
}

// ---------------------------------------------------------------- warm-start MLP
uint32_t orc_mlp_nparam(void) { return (uint32_t)MLP_NPARAM; }

void orc_bf16_round(const float* x, float* y, int n) {
    for (int i = 0; i < n; ++i) y[i] = bf16r(x[i]);
}

void orc_mlp_forward(const float* prm, const float* X, int B, int emu, float* Y) {
    MlpCache<float> c;
    mlp_forward<float>(prm, X, B, emu != 0, c);
    for (size_t i = 0; i < c.y.size(); ++i) Y[i] = c.y[i];
}

// Loss and gradient; fp64 != 0 runs the plain double path (no bf16 rounding).
double orc_mlp_grad(const float* prm, const float* X, const float* Y, int B, int emu, int fp64, double* grad) {
    if (fp64) return mlp_loss_grad<double>(prm, X, Y, B, false, grad);
    std::vector<float> g(MLP_NPARAM);
    const double l = mlp_loss_grad<float>(prm, X, Y, B, emu != 0, g.data());
    for (int i = 0; i < MLP_NPARAM; ++i) grad[i] = g[i];
    return l;
}

double orc_mlp_loss(const float* prm, const float* X, const float* Y, int B) {
    MlpCache<double> c;
    mlp_forward<double>(prm, X, B, false, c);
    double loss = 0.0;
    for (size_t i = 0; i < c.y.size(); ++i) {
        const double e = c.y[i] - (double)Y[i];
        loss += e * e;
    }
    return loss / (double)(B * MLP_OUT);
}

void orc_adam(float* prm, float* m, float* v, const float* g, int n, int t, float lr, float b1, float b2, float eps) {
    adam_step(prm, m, v, g, n, t, lr, b1, b2, eps);
}

}  // extern "C"
