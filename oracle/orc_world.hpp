// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// Sequential restatement of one world step: uniform-grid broadphase, primitive
// narrowphase, Jones-Plassmann colouring, colour-ordered PGS and semi-implicit
// Euler.  The reference only sketches this loop
// (examples/profiling_example.cpp:117-174: Broadphase -> Narrowphase -> Solver
// -> Integration; semi-implicit order v += g dt then x += v dt in
// examples/body_inspector_example.cpp:48-60; default config
// include/axiom/gui/physics_panel.hpp:34-43), so the algorithm below is the
// project's own specification; the HIP kernels implement the same sequence.
//
// Step(dt, I):
//   1. v += g dt; v *= (1 - linDamp); w *= (1 - angDamp)            (dynamic bodies)
//   2. AABB = shape bounds + margin; cell = clamp(floor((x - origin)/cell), 0, 1023);
//      key = world << 30 | morton(cell); stable sort of bodies by key
//   3. pairs, in sorted-body order t: planes first, then the 27 neighbour cells
//      (dz, dy, dx each -1..1), then ascending sorted position; pair (i, j) kept
//      iff j > i, not both static, AABBs overlap (inclusive)
//   4. narrowphase per pair -> manifold of <= 4 points; manifolds = pairs with
//      points, in pair order
//   5. colouring (rounds r = 0,1,..): tag(m) = (r+1)<<48 | hash24(a,b)<<24 | m;
//      every dynamic body keeps the max tag of its uncoloured manifolds; a
//      manifold that holds the max at all its dynamic bodies takes the first
//      colour free at both bodies (64-bit masks; 64+r when full)
//   6. stable order of manifolds by colour = solve order
//   7. prep per point, warm start (cache keyed by (a, b, feature id))
//   8. I PGS sweeps, colour by colour, manifold by manifold: friction rows of
//      all points, then normal rows (iteration 0 first applies the warm start).
//      Each friction row is clamped to |lam_t| <= mu * B (pyramid), with the bound B
//      fixed when the step starts: the point's normal impulse of the previous step
//      when the warm-start cache holds its feature, else the impulse m_N (target -
//      v_n)+ that stops the point's approach on its own.  With fixed bounds the step's problem is a
//      convex quadratic program over a box, so the sweeps cannot cycle (a bound that
//      follows lam_n during the sweeps can leave friction and normal rows trading
//      impulse in a period-2 cycle that never settles).
//      Every row is regularised with a compliance eps*k proportional to its own
//      diagonal k (soft constraint, constraint-force mixing): the row solves
//      v + eps*k*lam = target, so one visit is
//          lam' = proj(keep*lam + m*(target - v)),  m = 1/(k(1+eps)), keep = 1/(1+eps).
//      The regularised problem is strictly convex in the normal rows, so the
//      sweeps converge even where a manifold's rows are redundant (four
//      coplanar points, two friction rows per point) and their targets are not
//      mutually attainable, which otherwise leaves PGS drifting along the
//      null space with a constant residual.
//   9. x += v dt; q += (dt/2) (w,0) q; q /= |q|
//  Per-body extras (defaults from the world parameters): material {mu, e} combined
//  per contact as mu = sqrt(mu_A mu_B) (mu_A when equal) and e = max(e_A, e_B), a
//  plane taking the world's mu and e; per-body damping factors; kinematic bodies
//  (no mass, moved by their own velocity, pushing dynamic bodies as static ones
//  do); collision filters (category, mask, group: a pair with one nonzero group
//  collides iff that group is positive, else iff each category meets the other's
//  mask; planes collide with every dynamic body).
//  10. islands: connected components of the awake dynamic bodies over this step's
//      manifolds (a manifold with a static, sleeping or plane partner does not
//      connect); an island's label is its smallest body index.  Sleeping: an
//      awake body whose |v| and |w| are below the thresholds adds dt to its rest
//      time (else the time restarts at 0); an island whose bodies have all rested
//      sleep_time seconds falls asleep (velocities zeroed).  A sleeping body acts
//      as a static one in every stage; a manifold between an awake dynamic body and
//      a sleeping one wakes the sleeping body's island at the start of the next step.
#pragma once
#include <algorithm>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <vector>

#include "orc_geom.hpp"
#include "orc_par.hpp"

namespace orc {

constexpr uint32_t PLANE_BIT = 0x80000000u;
constexpr int MAX_PLANES = 8;
constexpr uint32_t CELL_MAX = 1023u;

struct Params {
    float gravity[3];
    float friction, restitution, rest_threshold;
    float baumgarte, slop, max_bias, margin;
    float lin_damp, ang_damp;
    float cell_size;
    float origin[3];
    uint32_t num_planes;
    float planes[MAX_PLANES][4];  // n.xyz, d
    uint32_t warm_start;          // 0 none, 1 cache
    uint32_t record_residual;     // compute residual after every sweep
    float softness;               // relative compliance eps of every row (0: rigid rows)
    uint32_t allow_sleep;         // islands whose bodies all rest for sleep_time seconds go to sleep
    float sleep_linear;           // m/s: |v| below this counts as resting
    float sleep_angular;          // rad/s: |w| below this counts as resting
    float sleep_time;             // s
    uint32_t island_solve;        // islands of at most LOCAL_MANIFOLDS manifolds solved on their own
};

struct Stats {
    uint32_t pairs, manifolds, contacts, colours, rounds, iterations;
    float residual;      // after the last sweep (velocity units)
    float residual_first;
    double kinetic, potential;
    uint32_t gjk_fail, epa_fail, warm_hits;
    uint32_t islands;   // islands of dynamic bodies (awake and asleep) after the step
    uint32_t sleeping;  // bodies asleep after the step
};

template <class T> struct World {
    Params P;
    uint32_t n = 0;
    std::vector<V3<T>> x, v, w, ii, prm;
    std::vector<Q4<T>> q;
    std::vector<T> im;
    std::vector<uint32_t> type, world;
    // stage outputs of the last step
    std::vector<uint64_t> keys;
    std::vector<uint32_t> order;
    std::vector<uint32_t> pair_a, pair_b;
    std::vector<uint32_t> man_a, man_b;  // pair order
    std::vector<Man<T>> man;
    std::vector<uint32_t> colour;      // per manifold (pair order)
    std::vector<uint32_t> solve_order;  // manifold ids by colour
    std::vector<uint32_t> colour_off;   // K+1
    std::vector<V3<T>> v_pre, w_pre;    // velocities entering the solver (after gravity)
    std::vector<T> residual_hist;
    // warm-start cache: (a,b) -> list of (fid, lambda)
    struct CacheEnt {
        uint32_t np;
        uint32_t fid[4];
        T lam[4][3];
        uint32_t colour;
    };
    uint32_t nsteps = 0;  // steps since the bodies were set (full recolouring every RECOLOUR_PERIOD)
    static constexpr uint32_t RECOLOUR_PERIOD = 64;
    std::map<uint64_t, CacheEnt> cache;
    std::vector<T> lambda;  // per manifold x 4 points x 3 (solve output)
    std::vector<T> feat;    // per manifold x 4 points x 24 warm-start network inputs
    // islands and sleeping (per body): island label (LABEL_NONE for static bodies and
    // bodies that never joined an island), rest time, asleep flag; wake stamps per label
    static constexpr uint32_t LABEL_NONE = 0x7FFFFFFFu;
    std::vector<uint32_t> label;
    std::vector<T> rest;
    std::vector<uint8_t> asleep;
    std::vector<uint32_t> wake;  // indexed by label: the step number at which that island wakes
    // per-body extras
    std::vector<T> mu_b, e_b, lkeep, akeep;
    std::vector<uint8_t> kin;
    std::vector<uint32_t> fcat, fmask;
    std::vector<int32_t> fgroup;
    bool use_filter = false;
    Stats st{};

    void reset_extras() {
        mu_b.assign(n, T(P.friction));
        e_b.assign(n, T(P.restitution));
        lkeep.assign(n, T(1) - T(P.lin_damp));
        akeep.assign(n, T(1) - T(P.ang_damp));
        kin.assign(n, 0);
        fcat.assign(n, 1u);
        fmask.assign(n, 0xFFFFFFFFu);
        fgroup.assign(n, 0);
        use_filter = false;
    }
    bool should_collide(uint32_t i, uint32_t j) const {
        if (!use_filter) return true;
        if (fgroup[i] == fgroup[j] && fgroup[i] != 0) return fgroup[i] > 0;
        return (fmask[i] & fcat[j]) != 0u && (fcat[i] & fmask[j]) != 0u;
    }
    T mix_mu(uint32_t a, uint32_t b) const {
        if (b & 0x80000000u) return T(P.friction) == mu_b[a] ? mu_b[a] : std::sqrt(mu_b[a] * T(P.friction));
        return mu_b[a] == mu_b[b] ? mu_b[a] : std::sqrt(mu_b[a] * mu_b[b]);
    }
    T mix_e(uint32_t a, uint32_t b) const {
        const T eb = (b & 0x80000000u) ? T(P.restitution) : e_b[b];
        return e_b[a] > eb ? e_b[a] : eb;
    }
    // residual-stopped sweeps: with tol > 0 the residual is checked after every sweep and
    // the solve stops once it is <= tol (at most max_iters sweeps)
    T residual_tol = T(0);
    uint32_t max_iters = 0;
    uint32_t res_every = 1;
    // host threads: every parallel loop below splits its range by a fixed grain and
    // combines partial results in range order, so results do not depend on the count
    std::unique_ptr<Pool> pool;
    void set_threads(unsigned t) {
        if (t <= 1) pool.reset();
        else pool.reset(new Pool(t));
    }
    // f(chunk, begin, end) over [0, n) in Pool::chunks_for(n, grain) ranges
    template <class F> void par(uint64_t n, uint64_t grain, F&& f) {
        if (n == 0) return;
        const std::function<void(uint64_t, uint64_t, uint64_t)> fn = f;
        if (pool) {
            pool->parallel_for(n, grain, fn);
        } else {
            const uint64_t c = Pool::chunks_for(n, grain);
            for (uint64_t k = 0; k < c; ++k) fn(k, n * k / c, n * (k + 1) / c);
        }
    }

    bool dynamic(uint32_t i) const { return im[i] > T(0) && !asleep[i]; }

    Shape<T> shape_of(uint32_t i) const {
        Shape<T> s;
        s.c = x[i];
        s.R = quat_to_mat(q[i]);
        s.type = type[i];
        s.prm = prm[i];
        return s;
    }

    void aabb_ext(uint32_t i, const M3<T>& R, V3<T>& e) const {
        const T mg = T(P.margin);
        const V3<T> p = prm[i];
        if (type[i] == SHAPE_SPHERE) {
            e = {p.x + mg, p.x + mg, p.x + mg};
        } else if (type[i] == SHAPE_BOX) {
            e.x = std::fabs(R.r0.x) * p.x + std::fabs(R.r0.y) * p.y + std::fabs(R.r0.z) * p.z + mg;
            e.y = std::fabs(R.r1.x) * p.x + std::fabs(R.r1.y) * p.y + std::fabs(R.r1.z) * p.z + mg;
            e.z = std::fabs(R.r2.x) * p.x + std::fabs(R.r2.y) * p.y + std::fabs(R.r2.z) * p.z + mg;
        } else {
            e.x = std::fabs(R.r0.y) * p.y + p.x + mg;
            e.y = std::fabs(R.r1.y) * p.y + p.x + mg;
            e.z = std::fabs(R.r2.y) * p.y + p.x + mg;
        }
    }

    static uint32_t cell_coord(T p, T o, T inv) {
        T f = (p - o) * inv;
        f = f > T(0) ? f : T(0);  // NaN -> 0
        f = f < T(CELL_MAX) ? f : T(CELL_MAX);
        return (uint32_t)std::floor(f);
    }

    // ---------------------------------------------------------------- broadphase
    void broadphase(std::vector<V3<T>>& mn, std::vector<V3<T>>& mx) {
        mn.resize(n);
        mx.resize(n);
        keys.resize(n);
        const T inv = T(1) / T(P.cell_size);
        const V3<T> o = {T(P.origin[0]), T(P.origin[1]), T(P.origin[2])};
        std::vector<uint32_t> cx(n), cy(n), cz(n);
        par(n, 4096, [&](uint64_t, uint64_t b, uint64_t e) {
            for (uint32_t i = (uint32_t)b; i < (uint32_t)e; ++i) {
                const M3<T> R = quat_to_mat(q[i]);
                V3<T> ext;
                aabb_ext(i, R, ext);
                mn[i] = x[i] - ext;
                mx[i] = x[i] + ext;
                cx[i] = cell_coord(x[i].x, o.x, inv);
                cy[i] = cell_coord(x[i].y, o.y, inv);
                cz[i] = cell_coord(x[i].z, o.z, inv);
                keys[i] = ((uint64_t)world[i] << 30) | morton3(cx[i], cy[i], cz[i]);
            }
        });
        // stable sort by key = sort by (key, index): sorted runs merged pairwise
        {
            std::vector<std::pair<uint64_t, uint32_t>> kv(n), tmp(n);
            for (uint32_t i = 0; i < n; ++i) kv[i] = {keys[i], i};
            const uint64_t RUN = 65536;
            const uint64_t runs = (n + RUN - 1) / RUN;
            par(runs, 1, [&](uint64_t, uint64_t b, uint64_t e) {
                for (uint64_t r = b; r < e; ++r)
                    std::sort(kv.begin() + r * RUN, kv.begin() + std::min<uint64_t>(n, (r + 1) * RUN));
            });
            for (uint64_t wd = RUN; wd < n; wd *= 2) {
                const uint64_t pairs = (n + 2 * wd - 1) / (2 * wd);
                par(pairs, 1, [&](uint64_t, uint64_t b, uint64_t e) {
                    for (uint64_t r = b; r < e; ++r) {
                        const uint64_t lo = r * 2 * wd, mid = std::min<uint64_t>(n, lo + wd),
                                       hi = std::min<uint64_t>(n, lo + 2 * wd);
                        std::merge(kv.begin() + lo, kv.begin() + mid, kv.begin() + mid, kv.begin() + hi,
                                   tmp.begin() + lo);
                    }
                });
                kv.swap(tmp);
            }
            order.resize(n);
            for (uint32_t t = 0; t < n; ++t) order[t] = kv[t].second;
        }
        std::vector<uint64_t> sk(n);
        for (uint32_t t = 0; t < n; ++t) sk[t] = keys[order[t]];
        pair_a.clear();
        pair_b.clear();
        const uint64_t PGRAIN = 2048;
        const uint64_t nchunk = Pool::chunks_for(n, PGRAIN);
        std::vector<std::vector<uint32_t>> ca(nchunk), cb(nchunk);
        par(n, PGRAIN, [&](uint64_t ch, uint64_t tb, uint64_t te) {
            std::vector<uint32_t>& pa = ca[ch];
            std::vector<uint32_t>& pb = cb[ch];
            for (uint32_t t = (uint32_t)tb; t < (uint32_t)te; ++t) {
                const uint32_t i = order[t];
                if (dynamic(i)) {
                    for (uint32_t k = 0; k < P.num_planes; ++k) {
                        const V3<T> pn = {T(P.planes[k][0]), T(P.planes[k][1]), T(P.planes[k][2])};
                        const V3<T> c = x[i];
                        const V3<T> e = mx[i] - c;
                        const T sd = dot(pn, c) - (std::fabs(pn.x) * e.x + std::fabs(pn.y) * e.y + std::fabs(pn.z) * e.z);
                        if (sd <= T(P.planes[k][3])) {
                            pa.push_back(i);
                            pb.push_back(PLANE_BIT | k);
                        }
                    }
                }
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int64_t X = (int64_t)cx[i] + dx, Y = (int64_t)cy[i] + dy, Z = (int64_t)cz[i] + dz;
                            if (X < 0 || Y < 0 || Z < 0 || X > CELL_MAX || Y > CELL_MAX || Z > CELL_MAX) continue;
                            const uint64_t key =
                                ((uint64_t)world[i] << 30) | morton3((uint32_t)X, (uint32_t)Y, (uint32_t)Z);
                            auto lo = std::lower_bound(sk.begin(), sk.end(), key);
                            auto hi = std::upper_bound(lo, sk.end(), key);
                            for (auto it = lo; it != hi; ++it) {
                                const uint32_t j = order[(size_t)(it - sk.begin())];
                                if (j <= i) continue;
                                if (!dynamic(i) && !dynamic(j)) continue;
                                if (!should_collide(i, j)) continue;
                                if (mn[i].x <= mx[j].x && mx[i].x >= mn[j].x && mn[i].y <= mx[j].y &&
                                    mx[i].y >= mn[j].y && mn[i].z <= mx[j].z && mx[i].z >= mn[j].z) {
                                    pa.push_back(i);
                                    pb.push_back(j);
                                }
                            }
                        }
            }
        });
        for (uint64_t c = 0; c < nchunk; ++c) {
            pair_a.insert(pair_a.end(), ca[c].begin(), ca[c].end());
            pair_b.insert(pair_b.end(), cb[c].begin(), cb[c].end());
        }
    }

    // ---------------------------------------------------------------- narrowphase
    void narrowphase() {
        man_a.clear();
        man_b.clear();
        man.clear();
        const T mg = T(P.margin);
        const size_t NP = pair_a.size();
        std::vector<Man<T>> pm(NP);
        std::vector<int> ps(NP, GEOM_OK);
        par(NP, 1024, [&](uint64_t, uint64_t b, uint64_t e) {
            for (size_t p = b; p < e; ++p) {
                const uint32_t a = pair_a[p], bb = pair_b[p];
                const Shape<T> A = shape_of(a);
                if (bb & PLANE_BIT) {
                    const uint32_t k = bb & ~PLANE_BIT;
                    collide_plane(pm[p], A, V3<T>{T(P.planes[k][0]), T(P.planes[k][1]), T(P.planes[k][2])},
                                  T(P.planes[k][3]), mg);
                } else {
                    ps[p] = collide(pm[p], A, shape_of(bb), mg);
                }
            }
        });
        for (size_t p = 0; p < NP; ++p) {
            if (ps[p] == GEOM_GJK_FAIL) ++st.gjk_fail;
            if (ps[p] == GEOM_EPA_FAIL) ++st.epa_fail;
            if (pm[p].np > 0) {
                man_a.push_back(pair_a[p]);
                man_b.push_back(pair_b[p]);
                man.push_back(pm[p]);
            }
        }
    }

    // ---------------------------------------------------------------- colouring
    bool dyn_partner(uint32_t b) const { return !(b & PLANE_BIT) && dynamic(b); }

    uint32_t local_id(uint32_t b, const std::vector<uint32_t>& world_first) const {
        if (b & PLANE_BIT) return b;
        return b - world_first[world[b]];
    }

    // Deterministic Jones-Plassmann colouring (hash priorities, first-fit colours).
    // Between full recolourings (every RECOLOUR_PERIOD steps) a manifold whose body
    // pair was in contact in the previous step keeps that step's colour, and only
    // the new manifolds are coloured around the kept ones.
    void colour_manifolds(const uint32_t* colour_in = nullptr) {
        const uint32_t M = (uint32_t)man.size();
        colour.assign(M, 0xFFFFFFFFu);
        if (colour_in) {
            uint32_t K = 0;
            for (uint32_t m = 0; m < M; ++m) {
                colour[m] = colour_in[m];
                K = std::max(K, colour[m] + 1);
            }
            st.rounds = 0;
            st.colours = K;
            order_by_colour(K);
            return;
        }
        // first body index of every world (bodies of a world are contiguous)
        uint32_t nw = 0;
        for (uint32_t i = 0; i < n; ++i) nw = std::max(nw, world[i] + 1);
        std::vector<uint32_t> wf(nw + 1, 0xFFFFFFFFu);
        for (uint32_t i = n; i-- > 0;) wf[world[i]] = i;
        std::vector<uint64_t> prio(M);
        for (uint32_t m = 0; m < M; ++m)
            prio[m] = ((uint64_t)pair_hash24(local_id(man_a[m], wf), local_id(man_b[m], wf)) << 24) | (m & 0xFFFFFFu);
        std::vector<uint64_t> bmax(n, 0);
        std::vector<uint64_t> mask(n, 0);
        uint32_t left = M, round = 0, K = 0;
        if (nsteps > 0 && nsteps % RECOLOUR_PERIOD != 0) {
            for (uint32_t m = 0; m < M; ++m) {
                auto it = cache.find(((uint64_t)man_a[m] << 32) | man_b[m]);
                if (it == cache.end() || it->second.colour >= 64) continue;
                const uint32_t c = it->second.colour;
                colour[m] = c;
                mask[man_a[m]] |= 1ull << c;
                if (dyn_partner(man_b[m])) mask[man_b[m]] |= 1ull << c;
                K = std::max(K, c + 1);
                --left;
            }
        }
        while (left > 0) {
            const uint64_t tagbase = (uint64_t)(round + 1) << 48;
            for (uint32_t m = 0; m < M; ++m) {
                if (colour[m] != 0xFFFFFFFFu) continue;
                const uint64_t t = tagbase | prio[m];
                bmax[man_a[m]] = std::max(bmax[man_a[m]], t);
                if (dyn_partner(man_b[m])) bmax[man_b[m]] = std::max(bmax[man_b[m]], t);
            }
            std::vector<uint32_t> win;
            for (uint32_t m = 0; m < M; ++m) {
                if (colour[m] != 0xFFFFFFFFu) continue;
                const uint64_t t = tagbase | prio[m];
                const bool bd = dyn_partner(man_b[m]);
                if (bmax[man_a[m]] == t && (!bd || bmax[man_b[m]] == t)) win.push_back(m);
            }
            for (uint32_t m : win) {
                const bool bd = dyn_partner(man_b[m]);
                const uint64_t used = mask[man_a[m]] | (bd ? mask[man_b[m]] : 0);
                uint32_t c;
                if (~used == 0) {
                    c = 64 + round;
                } else {
                    c = 0;
                    while (used & (1ull << c)) ++c;
                    mask[man_a[m]] |= 1ull << c;
                    if (bd) mask[man_b[m]] |= 1ull << c;
                }
                colour[m] = c;
                K = std::max(K, c + 1);
                --left;
            }
            ++round;
        }
        st.rounds = round;
        st.colours = K;
        order_by_colour(K);
    }

    // Solve order (P.island_solve, without record_residual, below 2^23 bodies): the manifolds
    // of the large islands first, by colour (then manifold index); then the small islands (at
    // most LOCAL_MANIFOLDS manifolds; an island is the bodies joined by manifolds between two
    // dynamic bodies), one after another in the order of their roots (smallest body index),
    // each by colour (then manifold index).  A manifold with no dynamic body belongs to the
    // large part.  colour_off covers the large islands' part [0, n_glob); local_off[i] is
    // where small island i starts (local_off.back() = M).  Without the split every manifold
    // is in the large part.
    static constexpr uint32_t LOCAL_MANIFOLDS = 4096;
    uint32_t n_glob = 0;
    std::vector<uint32_t> local_off;
    std::vector<uint32_t> step_root;  // island (union-find root) of every body in this step

    void step_islands() {
        const uint32_t M = (uint32_t)man.size();
        step_root.resize(n);
        for (uint32_t i = 0; i < n; ++i) step_root[i] = i;
        for (uint32_t m = 0; m < M; ++m) {
            const uint32_t a = man_a[m], b = man_b[m];
            if ((b & PLANE_BIT) || !dynamic(a) || !dynamic(b)) continue;
            uint32_t ra = uf_find(step_root, a), rb = uf_find(step_root, b);
            if (ra < rb) step_root[rb] = ra;
            else if (rb < ra) step_root[ra] = rb;
        }
        for (uint32_t i = 0; i < n; ++i) step_root[i] = uf_find(step_root, i);
    }
    // island of manifold m: that of its dynamic body (UINT32_MAX with none)
    uint32_t man_root(uint32_t m) const {
        const uint32_t a = man_a[m], b = man_b[m];
        if (dynamic(a)) return step_root[a];
        if (!(b & PLANE_BIT) && dynamic(b)) return step_root[b];
        return UINT32_MAX;
    }

    void order_by_colour(uint32_t K) {
        const uint32_t M = (uint32_t)man.size();
        std::vector<uint8_t> loc(M, 0);
        if (P.island_solve && !P.record_residual && n < (1u << 23)) {
            step_islands();
            std::vector<uint32_t> size(n, 0u);
            for (uint32_t m = 0; m < M; ++m)
                if (man_root(m) != UINT32_MAX) ++size[man_root(m)];
            for (uint32_t m = 0; m < M; ++m) {
                const uint32_t r = man_root(m);
                loc[m] = r != UINT32_MAX && size[r] <= LOCAL_MANIFOLDS ? 1 : 0;
            }
        }
        colour_off.assign(K + 1, 0);
        for (uint32_t m = 0; m < M; ++m)
            if (!loc[m]) colour_off[colour[m] + 1]++;
        for (uint32_t c = 0; c < K; ++c) colour_off[c + 1] += colour_off[c];
        n_glob = colour_off[K];
        solve_order.resize(M);
        std::vector<uint32_t> fill(colour_off.begin(), colour_off.end() - 1);
        std::vector<uint32_t> lm;
        for (uint32_t m = 0; m < M; ++m) {
            if (!loc[m]) solve_order[fill[colour[m]]++] = m;
            else lm.push_back(m);
        }
        std::stable_sort(lm.begin(), lm.end(), [&](uint32_t x, uint32_t y) {
            const uint32_t rx = man_root(x), ry = man_root(y);
            return rx != ry ? rx < ry : colour[x] < colour[y];
        });
        local_off.clear();
        for (size_t j = 0; j < lm.size(); ++j) {
            solve_order[n_glob + j] = lm[j];
            if (j == 0 || man_root(lm[j]) != man_root(lm[j - 1])) local_off.push_back(n_glob + (uint32_t)j);
        }
        local_off.push_back(M);
    }

    // ---------------------------------------------------------------- solver
    // Per contact point: the three constraint directions (n, t1, t2) with their
    // angular Jacobians a = rA x d, b = rB x d and inverse-inertia-weighted
    // terms ia = IA a, ib = IB b, effective masses m = 1 / k, the velocity target.
    struct Row {
        V3<T> rA, rB, n, t1, t2;
        V3<T> a[3], b[3], ia[3], ib[3];
        T mN, mT1, mT2, kN, kT1, kT2, target;
        T fb;  // friction bound mu * B of this step
        V3<T> dir(int d) const { return d == 0 ? n : (d == 1 ? t1 : t2); }
        T mass(int d) const { return d == 0 ? mN : (d == 1 ? mT1 : mT2); }
    };
    struct SBody {
        V3<T> v, w;
        T im;
        S3<T> I;
        V3<T> x;
        bool dyn;
    };

    // Row update primitives of the sweep (explicit fused multiply-adds, the same
    // sequence as the device kernels).
    static T fdot(V3<T> x, V3<T> y) { return std::fma(x.z, y.z, std::fma(x.y, y.y, x.x * y.x)); }
    // relative velocity along direction d: d.(vB - vA) + b.wB - a.wA
    static T jvel(const Row& r, int d, const SBody& A, const SBody& B) {
        const T x = fdot(r.dir(d), B.v - A.v) + fdot(r.b[d], B.w);
        return x - fdot(r.a[d], A.w);
    }
    // impulse lam along direction d: A gets -lam, B +lam
    static void japply(const Row& r, int d, T lam, SBody& A, SBody& B) {
        const V3<T> dd = r.dir(d);
        if (A.dyn) {
            const T sa = -(lam * A.im);
            A.v = {std::fma(sa, dd.x, A.v.x), std::fma(sa, dd.y, A.v.y), std::fma(sa, dd.z, A.v.z)};
            A.w = {std::fma(-lam, r.ia[d].x, A.w.x), std::fma(-lam, r.ia[d].y, A.w.y), std::fma(-lam, r.ia[d].z, A.w.z)};
        }
        if (B.dyn) {
            const T sb = lam * B.im;
            B.v = {std::fma(sb, dd.x, B.v.x), std::fma(sb, dd.y, B.v.y), std::fma(sb, dd.z, B.v.z)};
            B.w = {std::fma(lam, r.ib[d].x, B.w.x), std::fma(lam, r.ib[d].y, B.w.y), std::fma(lam, r.ib[d].z, B.w.z)};
        }
    }
    static V3<T> vrel(const SBody& A, const SBody& B, const Row& r) {
        return (B.v + cross(B.w, r.rB)) - (A.v + cross(A.w, r.rA));
    }

    std::vector<T> fbound;  // friction bound of every point of the last step (M x 4)
    const T* fb_in = nullptr;  // optional: bounds to use instead of computing them (M x 4)

    void solve(T dt, uint32_t iters, const T* warm_in /* optional: M x 4 x 3 in manifold order */) {
        const uint32_t M = (uint32_t)man.size();
        std::vector<SBody> sb(n);
        par(n, 4096, [&](uint64_t, uint64_t b0, uint64_t b1) {
            for (uint32_t i = (uint32_t)b0; i < (uint32_t)b1; ++i) {
                SBody& b = sb[i];
                b.v = v[i];
                b.w = w[i];
                b.dyn = dynamic(i);
                b.im = b.dyn ? im[i] : T(0);
                b.x = x[i];
                const M3<T> R = quat_to_mat(q[i]);
                b.I = b.dyn ? world_inv_inertia(R, ii[i]) : S3<T>{T(0), T(0), T(0), T(0), T(0), T(0)};
            }
        });
        SBody ground{};
        ground.v = ground.w = ground.x = {T(0), T(0), T(0)};
        ground.im = T(0);
        ground.I = {T(0), T(0), T(0), T(0), T(0), T(0)};
        ground.dyn = false;
        std::vector<Row> rows((size_t)M * 4);
        lambda.assign((size_t)M * 12, T(0));
        feat.assign((size_t)M * 4 * 24, T(0));
        fbound.assign((size_t)M * 4, T(0));
        const T inv_dt = T(1) / dt;
        const T soft1 = T(1.0f + P.softness);
        const T keep = T(1) / soft1;
        st.warm_hits = 0;
        std::vector<uint32_t> mhits(M, 0u);
        par(M, 256, [&](uint64_t, uint64_t m0, uint64_t m1) {
        for (uint32_t m = (uint32_t)m0; m < (uint32_t)m1; ++m) {
            const uint32_t a = man_a[m], b = man_b[m];
            const SBody& A = sb[a];
            const SBody& B = (b & PLANE_BIT) ? ground : sb[b];
            const uint64_t ck = ((uint64_t)a << 32) | b;
            const CacheEnt* ce = nullptr;
            if (P.warm_start != 0) {
                auto it = cache.find(ck);
                if (it != cache.end()) ce = &it->second;
            }
            const T mu = mix_mu(a, b), rest_e = mix_e(a, b);
            for (uint32_t k = 0; k < man[m].np; ++k) {
                const CPt<T>& c = man[m].pt[k];
                Row& r = rows[(size_t)m * 4 + k];
                r.n = c.n;
                r.rA = c.p - A.x;
                r.rB = (b & PLANE_BIT) ? V3<T>{T(0), T(0), T(0)} : c.p - B.x;
                tangents(c.n, r.t1, r.t2);
                const V3<T> dirs[3] = {r.n, r.t1, r.t2};
                T kk[3];
                for (int d = 0; d < 3; ++d) {
                    r.a[d] = cross(r.rA, dirs[d]);
                    r.b[d] = cross(r.rB, dirs[d]);
                    r.ia[d] = smul(A.I, r.a[d]);
                    r.ib[d] = smul(B.I, r.b[d]);
                    kk[d] = (A.im + B.im + dot(r.a[d], r.ia[d]) + dot(r.b[d], r.ib[d])) * soft1;
                }
                r.kN = kk[0];
                r.kT1 = kk[1];
                r.kT2 = kk[2];
                r.mN = kk[0] > T(0) ? T(1) / kk[0] : T(0);
                r.mT1 = kk[1] > T(0) ? T(1) / kk[1] : T(0);
                r.mT2 = kk[2] > T(0) ? T(1) / kk[2] : T(0);
                const T vn0 = dot(vrel(A, B, r), r.n);
                T bias = T(0);
                if (c.depth > T(P.slop)) {
                    bias = T(P.baumgarte) * (c.depth - T(P.slop)) * inv_dt;
                    if (bias > T(P.max_bias)) bias = T(P.max_bias);
                } else if (c.depth < T(0)) {
                    bias = c.depth * inv_dt;
                }
                if (vn0 < -T(P.rest_threshold)) {
                    const T rb = -rest_e * vn0;
                    if (rb > bias) bias = rb;
                }
                r.target = bias;
                // previous step's impulse of the same feature (the cached warm start)
                T Lp[3] = {T(0), T(0), T(0)};
                bool hit = false;
                if (ce) {
                    for (uint32_t j = 0; j < ce->np; ++j)
                        if (ce->fid[j] == c.fid) {
                            Lp[0] = ce->lam[j][0];
                            Lp[1] = ce->lam[j][1];
                            Lp[2] = ce->lam[j][2];
                            hit = true;
                            ++mhits[m];
                            break;
                        }
                }
                {
                    const T est = r.mN * (bias - vn0);
                    const T nb = hit ? Lp[0] : est;
                    r.fb = fb_in ? fb_in[(size_t)m * 4 + k] : mu * (nb > T(0) ? nb : T(0));
                    fbound[(size_t)m * 4 + k] = r.fb;
                }
                T* L = &lambda[((size_t)m * 4 + k) * 3];
                for (int d = 0; d < 3; ++d) L[d] = warm_in ? warm_in[((size_t)m * 4 + k) * 3 + d] : Lp[d];
                {
                    // warm-start network inputs: contact frame, impulses in units of u = mN g dt
                    const T gdt = T(9.81f) * dt;
                    const T inv_gdt = T(1) / gdt;
                    const T u = r.mN * gdt;
                    const T inv_u = u > T(0) ? T(1) / u : T(0);
                    const V3<T> vr = vrel(A, B, r);
                    T* f = &feat[((size_t)m * 4 + k) * 24];
                    f[0] = r.n.x; f[1] = r.n.y; f[2] = r.n.z; f[3] = c.depth * T(20);
                    f[4] = vn0 * inv_gdt; f[5] = dot(vr, r.t1) * inv_gdt; f[6] = dot(vr, r.t2) * inv_gdt;
                    f[7] = dot(r.rA, r.n); f[8] = dot(r.rA, r.t1); f[9] = dot(r.rA, r.t2);
                    f[10] = dot(r.rB, r.n); f[11] = dot(r.rB, r.t1); f[12] = dot(r.rB, r.t2);
                    f[13] = A.im * r.mN; f[14] = B.im * r.mN; f[15] = r.mT1 * kk[0]; f[16] = r.mT2 * kk[0];
                    f[17] = bias * inv_gdt; f[18] = Lp[0] * inv_u; f[19] = Lp[1] * inv_u; f[20] = Lp[2] * inv_u;
                    f[21] = hit ? T(1) : T(0); f[22] = T(man[m].np) * T(0.25f); f[23] = (b & PLANE_BIT) ? T(1) : T(0);
                }
            }
        }
        });
        for (uint32_t m = 0; m < M; ++m) st.warm_hits += mhits[m];
        residual_hist.clear();
        // one visit of solve slot s (manifold solve_order[s])
        auto visit = [&](uint32_t s, bool first) {
            const uint32_t m = solve_order[s];
            const uint32_t a = man_a[m], b = man_b[m];
            SBody& A = sb[a];
            SBody& B = (b & PLANE_BIT) ? ground : sb[b];
            const uint32_t np = man[m].np;
            if (first) {
                for (uint32_t k = 0; k < np; ++k) {
                    const Row& r = rows[(size_t)m * 4 + k];
                    const T* L = &lambda[((size_t)m * 4 + k) * 3];
                    japply(r, 0, L[0], A, B);
                    japply(r, 1, L[1], A, B);
                    japply(r, 2, L[2], A, B);
                }
            }
            for (uint32_t k = 0; k < np; ++k) {
                const Row& r = rows[(size_t)m * 4 + k];
                T* L = &lambda[((size_t)m * 4 + k) * 3];
                const T maxf = r.fb;
                for (int dd = 1; dd <= 2; ++dd) {
                    T nl = std::fma(-r.mass(dd), jvel(r, dd, A, B), keep * L[dd]);
                    nl = nl < -maxf ? -maxf : (nl > maxf ? maxf : nl);
                    const T d = nl - L[dd];
                    L[dd] = nl;
                    japply(r, dd, d, A, B);
                }
            }
            for (uint32_t k = 0; k < np; ++k) {
                const Row& r = rows[(size_t)m * 4 + k];
                T* L = &lambda[((size_t)m * 4 + k) * 3];
                T nl = std::fma(r.mN, r.target - jvel(r, 0, A, B), keep * L[0]);
                nl = nl > T(0) ? nl : T(0);
                const T d = nl - L[0];
                L[0] = nl;
                japply(r, 0, d, A, B);
            }
        };
        // residual: the largest velocity-space change one more visit would make to any row
        auto residual = [&]() {
            const uint64_t RG = 1024;
            std::vector<T> part(Pool::chunks_for(M, RG), T(0));
            par(M, RG, [&](uint64_t ch, uint64_t m0, uint64_t m1) {
                T res = T(0);
                for (uint32_t m = (uint32_t)m0; m < (uint32_t)m1; ++m) {
                    const uint32_t a = man_a[m], b = man_b[m];
                    const SBody& A = sb[a];
                    const SBody& B = (b & PLANE_BIT) ? ground : sb[b];
                    for (uint32_t k = 0; k < man[m].np; ++k) {
                        const Row& r = rows[(size_t)m * 4 + k];
                        const T* L = &lambda[((size_t)m * 4 + k) * 3];
                        T nl = std::fma(r.mN, r.target - jvel(r, 0, A, B), keep * L[0]);
                        nl = nl > T(0) ? nl : T(0);
                        T e = std::fabs(nl - L[0]) * r.kN;
                        const T maxf = r.fb;
                        T t1 = std::fma(-r.mT1, jvel(r, 1, A, B), keep * L[1]);
                        t1 = t1 < -maxf ? -maxf : (t1 > maxf ? maxf : t1);
                        const T e1 = std::fabs(t1 - L[1]) * r.kT1;
                        T t2 = std::fma(-r.mT2, jvel(r, 2, A, B), keep * L[2]);
                        t2 = t2 < -maxf ? -maxf : (t2 > maxf ? maxf : t2);
                        const T e2 = std::fabs(t2 - L[2]) * r.kT2;
                        if (e1 > e) e = e1;
                        if (e2 > e) e = e2;
                        if (e > res) res = e;
                    }
                }
                part[ch] = res;
            });
            T res = T(0);
            for (const T p : part)
                if (p > res) res = p;
            return res;
        };
        // colour batches: the manifolds of one colour share no dynamic body, so a batch's
        // visits are independent and run in parallel without changing any result
        const uint32_t K = colour_off.empty() ? 0u : (uint32_t)colour_off.size() - 1u;
        const bool tol_mode = residual_tol > T(0);
        const uint32_t cap = tol_mode ? std::max<uint32_t>(max_iters, 1u) : iters;
        const uint32_t every = std::max<uint32_t>(res_every, 1u);
        uint32_t ran = 0;
        if (cap == 0) {
            // no sweeps: apply the starting impulses (the warm start) and measure their residual
            for (uint32_t s = 0; s < M; ++s) {
                const uint32_t m = solve_order[s];
                const uint32_t a = man_a[m], b = man_b[m];
                SBody& A = sb[a];
                SBody& B = (b & PLANE_BIT) ? ground : sb[b];
                for (uint32_t k = 0; k < man[m].np; ++k) {
                    const Row& r = rows[(size_t)m * 4 + k];
                    const T* L = &lambda[((size_t)m * 4 + k) * 3];
                    japply(r, 0, L[0], A, B);
                    japply(r, 1, L[1], A, B);
                    japply(r, 2, L[2], A, B);
                }
            }
            residual_hist.push_back(residual());
        }
        // residual of the manifolds in solve slots [s0, s1)
        auto residual_slots = [&](uint32_t s0, uint32_t s1) {
            T res = T(0);
            for (uint32_t s = s0; s < s1; ++s) {
                const uint32_t m = solve_order[s];
                const uint32_t a = man_a[m], b = man_b[m];
                const SBody& A = sb[a];
                const SBody& B = (b & PLANE_BIT) ? ground : sb[b];
                for (uint32_t k = 0; k < man[m].np; ++k) {
                    const Row& r = rows[(size_t)m * 4 + k];
                    const T* L = &lambda[((size_t)m * 4 + k) * 3];
                    T nl = std::fma(r.mN, r.target - jvel(r, 0, A, B), keep * L[0]);
                    nl = nl > T(0) ? nl : T(0);
                    T e = std::fabs(nl - L[0]) * r.kN;
                    const T maxf = r.fb;
                    T t1 = std::fma(-r.mT1, jvel(r, 1, A, B), keep * L[1]);
                    t1 = t1 < -maxf ? -maxf : (t1 > maxf ? maxf : t1);
                    const T e1 = std::fabs(t1 - L[1]) * r.kT1;
                    T t2 = std::fma(-r.mT2, jvel(r, 2, A, B), keep * L[2]);
                    t2 = t2 < -maxf ? -maxf : (t2 > maxf ? maxf : t2);
                    const T e2 = std::fabs(t2 - L[2]) * r.kT2;
                    if (e1 > e) e = e1;
                    if (e2 > e) e = e2;
                    if (e > res) res = e;
                }
            }
            return res;
        };
        // small islands: each runs its own sweeps, visiting its slots in order, and (with a
        // tolerance) stops after the first sweep that leaves its own residual <= tol
        const uint32_t nloc = local_off.empty() ? 0u : (uint32_t)local_off.size() - 1u;
        std::vector<uint32_t> loc_ran(nloc, 0u);
        std::vector<T> loc_res(nloc, T(0));
        if (cap > 0)
            par(nloc, 64, [&](uint64_t, uint64_t i0, uint64_t i1) {
                for (uint32_t i = (uint32_t)i0; i < (uint32_t)i1; ++i) {
                    const uint32_t s0 = local_off[i], s1 = local_off[i + 1];
                    uint32_t it = 0;
                    T res = T(0);
                    for (; it < cap; ++it) {
                        for (uint32_t s = s0; s < s1; ++s) visit(s, it == 0);
                        if (tol_mode || it + 1 == cap) {
                            res = residual_slots(s0, s1);
                            if (tol_mode && res <= residual_tol) {
                                ++it;
                                break;
                            }
                        }
                    }
                    loc_ran[i] = it > cap ? cap : it;
                    loc_res[i] = res;
                }
            });
        uint32_t loc_max = 0;
        T loc_res_max = T(0);
        for (uint32_t i = 0; i < nloc; ++i) {
            loc_max = std::max(loc_max, loc_ran[i]);
            if (loc_res[i] > loc_res_max) loc_res_max = loc_res[i];
        }
        // large islands: together, colour batch by colour batch, over solve slots [0, n_glob)
        // (also with no manifold at all, so an empty step counts its sweeps as before)
        const bool glob = n_glob > 0 || nloc == 0;
        for (uint32_t it = 0; it < cap && glob; ++it) {
            for (uint32_t c = 0; c < K; ++c) {
                const uint32_t s0 = colour_off[c], s1 = colour_off[c + 1];
                if (s1 - s0 >= 512 && pool) {
                    par(s1 - s0, 256, [&](uint64_t, uint64_t b0, uint64_t b1) {
                        for (uint32_t s = s0 + (uint32_t)b0; s < s0 + (uint32_t)b1; ++s) visit(s, it == 0);
                    });
                } else {
                    for (uint32_t s = s0; s < s1; ++s) visit(s, it == 0);
                }
            }
            ran = it + 1;
            const bool check = tol_mode && ran % every == 0;
            if (P.record_residual || check || ran == cap) {
                // the large islands stop on their own residual; the step reports the largest
                const T gres = nloc ? residual_slots(0, n_glob) : residual();
                residual_hist.push_back(gres > loc_res_max ? gres : loc_res_max);
                if (check && gres <= residual_tol) break;
            }
        }
        if (!glob && cap > 0) residual_hist.push_back(loc_res_max);
        ran = std::max(ran, loc_max);
        st.iterations = ran;
        st.residual = residual_hist.empty() ? 0.0f : (float)residual_hist.back();
        st.residual_first = residual_hist.empty() ? 0.0f : (float)residual_hist.front();
        for (uint32_t i = 0; i < n; ++i) {
            if (!sb[i].dyn) continue;
            v[i] = sb[i].v;
            w[i] = sb[i].w;
        }
        // refresh the warm-start cache
        cache.clear();
        for (uint32_t m = 0; m < M; ++m) {
            CacheEnt e;
            e.np = man[m].np;
            e.colour = colour[m];
            for (uint32_t k = 0; k < e.np; ++k) {
                e.fid[k] = man[m].pt[k].fid;
                for (int d = 0; d < 3; ++d) e.lam[k][d] = lambda[((size_t)m * 4 + k) * 3 + d];
            }
            cache[((uint64_t)man_a[m] << 32) | man_b[m]] = e;
        }
    }

    void integrate_velocities(T dt) {
        const V3<T> g = {T(P.gravity[0]), T(P.gravity[1]), T(P.gravity[2])};
        par(n, 4096, [&](uint64_t, uint64_t b, uint64_t e) {
            for (uint32_t i = (uint32_t)b; i < (uint32_t)e; ++i) {
                if (!dynamic(i)) continue;
                v[i] = (v[i] + g * dt) * lkeep[i];
                w[i] = w[i] * akeep[i];
            }
        });
    }

    void integrate_positions(T dt) {
        const T hdt = T(0.5) * dt;
        par(n, 4096, [&](uint64_t, uint64_t b, uint64_t e) {
            for (uint32_t i = (uint32_t)b; i < (uint32_t)e; ++i) {
                if (!dynamic(i) && !kin[i]) continue;
                x[i] = x[i] + v[i] * dt;
                const Q4<T> wq = {w[i].x, w[i].y, w[i].z, T(0)};
                const Q4<T> dq = qmul(wq, q[i]);
                Q4<T> nq = {q[i].x + dq.x * hdt, q[i].y + dq.y * hdt, q[i].z + dq.z * hdt, q[i].w + dq.w * hdt};
                const T len = std::sqrt(nq.x * nq.x + nq.y * nq.y + nq.z * nq.z + nq.w * nq.w);
                const T inv = T(1) / len;
                q[i] = {nq.x * inv, nq.y * inv, nq.z * inv, nq.w * inv};
            }
        });
    }

    // Kinetic energy (translational + rotational about the centre, body-frame
    // principal inertia) and potential energy -m g.x of the dynamic bodies, in
    // double; summed per 4096-body range, ranges in order.
    void energy(double& ke, double& pe) {
        const uint64_t G = 4096;
        const uint64_t nch = Pool::chunks_for(n, G);
        std::vector<double> pk(nch, 0.0), pp(nch, 0.0);
        par(n, G, [&](uint64_t ch, uint64_t b, uint64_t e) {
            double k = 0.0, u = 0.0;
            for (uint32_t i = (uint32_t)b; i < (uint32_t)e; ++i) {
                if (!(im[i] > T(0))) continue;
                const double m = 1.0 / (double)im[i];
                const double vx = v[i].x, vy = v[i].y, vz = v[i].z;
                k += 0.5 * m * (vx * vx + vy * vy + vz * vz);
                const M3<T> R = quat_to_mat(q[i]);
                const V3<T> wl = mulT(R, w[i]);
                const double I[3] = {ii[i].x > 0 ? 1.0 / (double)ii[i].x : 0.0,
                                     ii[i].y > 0 ? 1.0 / (double)ii[i].y : 0.0,
                                     ii[i].z > 0 ? 1.0 / (double)ii[i].z : 0.0};
                k += 0.5 * (I[0] * (double)wl.x * (double)wl.x + I[1] * (double)wl.y * (double)wl.y +
                            I[2] * (double)wl.z * (double)wl.z);
                u -= m * ((double)P.gravity[0] * (double)x[i].x + (double)P.gravity[1] * (double)x[i].y +
                          (double)P.gravity[2] * (double)x[i].z);
            }
            pk[ch] = k;
            pp[ch] = u;
        });
        ke = pe = 0.0;
        for (uint64_t c = 0; c < pk.size(); ++c) {
            ke += pk[c];
            pe += pp[c];
        }
    }

    void reset_sleep() {
        label.assign(n, LABEL_NONE);
        rest.assign(n, T(0));
        asleep.assign(n, 0);
        wake.assign(n, 0xFFFFFFFFu);
    }

    // Start of a step: sleeping bodies whose island was touched by an awake body in the
    // previous step wake up.
    void wake_up() {
        par(n, 4096, [&](uint64_t, uint64_t b, uint64_t e) {
            for (uint32_t i = (uint32_t)b; i < (uint32_t)e; ++i)
                if (asleep[i] && label[i] < n && wake[label[i]] == nsteps) {
                    asleep[i] = 0;
                    rest[i] = T(0);
                }
        });
    }

    static uint32_t uf_find(std::vector<uint32_t>& p, uint32_t x) {
        while (p[x] != x) {
            p[x] = p[p[x]];
            x = p[x];
        }
        return x;
    }

    // End of a step (after integration): wake requests, island labels, rest times, sleep.
    void islands_and_sleep(T dt) {
        const uint32_t M = (uint32_t)man_a.size();
        std::vector<uint32_t> p(n);
        for (uint32_t i = 0; i < n; ++i) p[i] = i;
        for (uint32_t m = 0; m < M; ++m) {
            const uint32_t a = man_a[m], b = man_b[m];
            if (b & PLANE_BIT) continue;
            const bool da = dynamic(a), db = dynamic(b);
            if (da && db) {
                uint32_t ra = uf_find(p, a), rb = uf_find(p, b);
                if (ra != rb) {
                    if (ra < rb) p[rb] = ra;
                    else p[ra] = rb;
                }
            } else if (da && asleep[b] && label[b] < n) {
                wake[label[b]] = nsteps + 1;
            } else if (db && asleep[a] && label[a] < n) {
                wake[label[a]] = nsteps + 1;
            }
        }
        const T lin2 = T(P.sleep_linear) * T(P.sleep_linear), ang2 = T(P.sleep_angular) * T(P.sleep_angular);
        std::vector<T> imin(n, std::numeric_limits<T>::infinity());
        for (uint32_t i = 0; i < n; ++i) {
            if (!dynamic(i)) continue;
            const uint32_t r = uf_find(p, i);
            label[i] = r;
            const bool still = dot(v[i], v[i]) < lin2 && dot(w[i], w[i]) < ang2;
            rest[i] = still ? rest[i] + dt : T(0);
            if (rest[i] < imin[r]) imin[r] = rest[i];
        }
        uint32_t islands = 0, sleeping = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (!(im[i] > T(0))) continue;
            if (!asleep[i] && P.allow_sleep && imin[label[i]] >= T(P.sleep_time)) {
                asleep[i] = 1;
                v[i] = w[i] = V3<T>{T(0), T(0), T(0)};
            }
            if (label[i] == i) ++islands;
            if (asleep[i]) ++sleeping;
        }
        st.islands = islands;
        st.sleeping = sleeping;
    }

    void step(T dt, uint32_t iters) {
        st = Stats{};
        wake_up();
        std::vector<V3<T>> mn, mx;
        broadphase(mn, mx);
        narrowphase();
        integrate_velocities(dt);
        v_pre = v;
        w_pre = w;
        colour_manifolds();
        solve(dt, iters, nullptr);
        integrate_positions(dt);
        islands_and_sleep(dt);
        ++nsteps;
        st.pairs = (uint32_t)pair_a.size();
        st.manifolds = (uint32_t)man.size();
        uint32_t c = 0;
        for (auto& m : man) c += m.np;
        st.contacts = c;
        energy(st.kinetic, st.potential);
    }
};

}  // namespace orc
