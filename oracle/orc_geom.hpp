// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// Primitive narrowphase, GJK and EPA.  The reference reserves these stages
// (collision::Shape / collision::ContactPoint forward declarations,
// include/axiom/debug/physics_debug_draw.hpp:14-17; GJKFailedToConverge /
// EPAFailedToConverge, include/axiom/core/error_code.hpp:37-38) but contains no
// implementation, so this file restates the textbook algorithms:
//   * closed-form primitive tests     Ericson, Real-Time Collision Detection (2005) ch. 4-5
//   * box-box SAT + face clipping     Ericson 4.4 / Sutherland-Hodgman polygon clipping
//   * GJK distance                    Gilbert, Johnson & Keerthi (1988); closest-point-on-simplex
//                                     routines of Ericson 5.1.5 / 5.1.6
//   * EPA penetration depth           van den Bergen (2001)
// Contact convention (include/axiom/debug/physics_debug_draw.hpp:128-132 and
// src/debug/physics_debug_draw.cpp:111-115): the normal points from body A to
// body B, depth > 0 means penetration, `p` is the witness point on A and
// p - n*depth the witness point on B.  Capsules lie along local +Y with half
// height hh (src/debug/physics_debug_draw.cpp:254-266).
#pragma once
#include "orc_math.hpp"

namespace orc {

enum : uint32_t { SHAPE_SPHERE = 0, SHAPE_BOX = 1, SHAPE_CAPSULE = 2, SHAPE_PLANE = 3 };
enum : int { GEOM_OK = 0, GEOM_GJK_FAIL = 301, GEOM_EPA_FAIL = 302 };

template <class T> struct Shape {
    V3<T> c;        // centre
    M3<T> R;        // orientation
    uint32_t type;  // SHAPE_*
    V3<T> prm;      // sphere (r,-,-) box (hx,hy,hz) capsule (r,hh,-)
};

template <class T> struct CPt {
    V3<T> p;  // witness point on A
    V3<T> n;  // A -> B
    T depth;
    uint32_t fid;
};
template <class T> struct Man {
    uint32_t np = 0;
    CPt<T> pt[4];
};

template <class T> inline void man_add(Man<T>& m, V3<T> p, V3<T> n, T depth, uint32_t fid) {
    if (m.np < 4) m.pt[m.np++] = CPt<T>{p, n, depth, fid};
}

// ---------------------------------------------------------------- spheres
template <class T> inline void sphere_sphere(Man<T>& m, V3<T> cA, T rA, V3<T> cB, T rB, T margin, uint32_t fid) {
    const V3<T> d = cB - cA;
    const T d2 = dot(d, d);
    const T R = rA + rB;
    const T s = R + margin;
    if (d2 > s * s) return;
    const T dist = std::sqrt(d2);
    V3<T> n;
    if (dist > T(1e-6)) {
        const T inv = T(1) / dist;
        n = d * inv;
    } else {
        n = {T(0), T(1), T(0)};
    }
    man_add(m, cA + n * rA, n, R - dist, fid);
}

template <class T> inline V3<T> closest_on_segment(V3<T> p, V3<T> a, V3<T> b) {
    const V3<T> e = b - a;
    const T ee = dot(e, e);
    T t = T(0);
    if (ee > T(0)) t = dot(p - a, e) / ee;
    t = clampv(t, T(0), T(1));
    return a + e * t;
}

template <class T> inline void sphere_box(Man<T>& m, V3<T> c, T r, const Shape<T>& B, T margin) {
    const V3<T> h = B.prm;
    const V3<T> pl = mulT(B.R, c - B.c);
    const V3<T> q = {clampv(pl.x, -h.x, h.x), clampv(pl.y, -h.y, h.y), clampv(pl.z, -h.z, h.z)};
    const V3<T> diff = q - pl;
    const T d2 = dot(diff, diff);
    if (d2 > T(1e-12)) {
        const T s = r + margin;
        if (d2 > s * s) return;
        const T dist = std::sqrt(d2);
        const T inv = T(1) / dist;
        const V3<T> n = mul(B.R, diff * inv);
        man_add(m, c + n * r, n, r - dist, 0u);
        return;
    }
    // centre inside the box: leave through the nearest face
    const T fx = h.x - std::fabs(pl.x), fy = h.y - std::fabs(pl.y), fz = h.z - std::fabs(pl.z);
    int k = 0;
    T f = fx;
    if (fy < f) { k = 1; f = fy; }
    if (fz < f) { k = 2; f = fz; }
    const T comp = k == 0 ? pl.x : (k == 1 ? pl.y : pl.z);
    const V3<T> n = B.R.col(k) * (-sgn1(comp));
    man_add(m, c + n * r, n, r + f, 1u);
}

template <class T> inline void capsule_segment(const Shape<T>& S, V3<T>& p0, V3<T>& p1) {
    const V3<T> u = S.R.col(1);
    p0 = S.c - u * S.prm.y;
    p1 = S.c + u * S.prm.y;
}

// ---------------------------------------------------------------- planes
// Plane: dot(n, x) = d, solid where dot(n, x) < d.
template <class T> inline void sphere_plane(Man<T>& m, V3<T> c, T r, V3<T> pn, T pd, T margin, uint32_t fid) {
    const T dist = dot(pn, c) - pd;
    if (dist - r > margin) return;
    man_add(m, c - pn * r, -pn, r - dist, fid);
}

template <class T> inline void box_plane(Man<T>& m, const Shape<T>& A, V3<T> pn, T pd, T margin) {
    T dep[8];
    V3<T> vx[8];
    int cand[8];
    int nc = 0;
    for (int k = 0; k < 8; ++k) {
        const V3<T> loc = {(k & 1) ? A.prm.x : -A.prm.x, (k & 2) ? A.prm.y : -A.prm.y, (k & 4) ? A.prm.z : -A.prm.z};
        vx[k] = A.c + mul(A.R, loc);
        dep[k] = pd - dot(pn, vx[k]);
        if (dep[k] > -margin) cand[nc++] = k;
    }
    // keep up to four deepest (ties: lower vertex index), emitted in vertex order
    bool keep[8] = {false, false, false, false, false, false, false, false};
    const int want = nc < 4 ? nc : 4;
    for (int s = 0; s < want; ++s) {
        int best = -1;
        for (int i = 0; i < nc; ++i) {
            const int k = cand[i];
            if (keep[k]) continue;
            if (best < 0 || dep[k] > dep[best]) best = k;
        }
        keep[best] = true;
    }
    for (int k = 0; k < 8; ++k)
        if (keep[k]) man_add(m, vx[k], -pn, dep[k], (uint32_t)k);
}

// ---------------------------------------------------------------- capsules
// Closest points of segments p1q1 and p2q2 (Ericson 5.1.9).
template <class T> inline void seg_seg(V3<T> p1, V3<T> q1, V3<T> p2, V3<T> q2, V3<T>& c1, V3<T>& c2) {
    const V3<T> d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    const T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    T s, t;
    const T eps = T(1e-12);
    if (a <= eps && e <= eps) {
        s = t = T(0);
    } else if (a <= eps) {
        s = T(0);
        t = clampv(f / e, T(0), T(1));
    } else {
        const T c = dot(d1, r);
        if (e <= eps) {
            t = T(0);
            s = clampv(-c / a, T(0), T(1));
        } else {
            const T b = dot(d1, d2);
            const T den = a * e - b * b;
            s = den != T(0) ? clampv((b * f - c * e) / den, T(0), T(1)) : T(0);
            t = (b * s + f) / e;
            if (t < T(0)) {
                t = T(0);
                s = clampv(-c / a, T(0), T(1));
            } else if (t > T(1)) {
                t = T(1);
                s = clampv((b - c) / a, T(0), T(1));
            }
        }
    }
    c1 = p1 + d1 * s;
    c2 = p2 + d2 * t;
}

template <class T> inline void capsule_capsule(Man<T>& m, const Shape<T>& A, const Shape<T>& B, T margin) {
    V3<T> a0, a1, b0, b1;
    capsule_segment(A, a0, a1);
    capsule_segment(B, b0, b1);
    const T rA = A.prm.x, rB = B.prm.x;
    const V3<T> dA = a1 - a0, dB = b1 - b0;
    const V3<T> cr = cross(dA, dB);
    const T aa = dot(dA, dA), bb = dot(dB, dB);
    if (aa > T(1e-12) && bb > T(1e-12) && dot(cr, cr) <= T(1e-6) * aa * bb) {
        // near-parallel: two points at the ends of the overlap of B projected onto A
        const T tb0 = dot(b0 - a0, dA) / aa;
        const T tb1 = dot(b1 - a0, dA) / aa;
        const T lo = clampv(tb0 < tb1 ? tb0 : tb1, T(0), T(1));
        const T hi = clampv(tb0 < tb1 ? tb1 : tb0, T(0), T(1));
        if (hi - lo > T(1e-4)) {
            const V3<T> pa0 = a0 + dA * lo;
            const V3<T> pa1 = a0 + dA * hi;
            sphere_sphere(m, pa0, rA, closest_on_segment(pa0, b0, b1), rB, margin, 0u);
            sphere_sphere(m, pa1, rA, closest_on_segment(pa1, b0, b1), rB, margin, 1u);
            return;
        }
    }
    V3<T> c1, c2;
    seg_seg(a0, a1, b0, b1, c1, c2);
    sphere_sphere(m, c1, rA, c2, rB, margin, 2u);
}

// ---------------------------------------------------------------- GJK
template <class T> struct SV {
    V3<T> w, a, b;  // w = a - b
};
template <class T> struct GjkOut {
    int status;  // 0 separated, 1 cores intersect, -1 iteration cap
    T dist;
    V3<T> pA, pB;
    int n;
    SV<T> s[4];
    int iters;
};

// Support of the "core" shapes used by GJK: box polytope, capsule segment, sphere point.
template <class T> inline V3<T> core_support(const Shape<T>& S, V3<T> d) {
    if (S.type == SHAPE_BOX) {
        const V3<T> dl = mulT(S.R, d);
        const V3<T> l = {sgn1(dl.x) * S.prm.x, sgn1(dl.y) * S.prm.y, sgn1(dl.z) * S.prm.z};
        return S.c + mul(S.R, l);
    }
    if (S.type == SHAPE_CAPSULE) {
        V3<T> p0, p1;
        capsule_segment(S, p0, p1);
        return dot(d, p1 - p0) >= T(0) ? p1 : p0;
    }
    return S.c;
}

template <class T> inline SV<T> mink_support(const Shape<T>& A, const Shape<T>& B, V3<T> d) {
    SV<T> r;
    r.a = core_support(A, d);
    r.b = core_support(B, -d);
    r.w = r.a - r.b;
    return r;
}

// Closest point to the origin on the simplex s[0..n); reduces s to the supporting
// sub-simplex and writes barycentric weights. Returns false when the origin lies
// inside a tetrahedron.
template <class T> inline bool simplex_closest(SV<T>* s, int& n, T* lam, V3<T>& v) {
    if (n == 1) {
        lam[0] = T(1);
        v = s[0].w;
        return true;
    }
    if (n == 2) {
        const V3<T> a = s[0].w, ab = s[1].w - s[0].w;
        const T den = dot(ab, ab);
        const T t = den > T(0) ? dot(-a, ab) / den : T(0);
        if (t <= T(0)) { n = 1; lam[0] = T(1); v = a; return true; }
        if (t >= T(1)) { s[0] = s[1]; n = 1; lam[0] = T(1); v = s[0].w; return true; }
        lam[0] = T(1) - t;
        lam[1] = t;
        v = a + ab * t;
        return true;
    }
    if (n == 3) {
        const V3<T> a = s[0].w, b = s[1].w, c = s[2].w;
        const V3<T> ab = b - a, ac = c - a, ap = -a;
        const T d1 = dot(ab, ap), d2 = dot(ac, ap);
        if (d1 <= T(0) && d2 <= T(0)) { n = 1; lam[0] = T(1); v = a; return true; }
        const V3<T> bp = -b;
        const T d3 = dot(ab, bp), d4 = dot(ac, bp);
        if (d3 >= T(0) && d4 <= d3) { s[0] = s[1]; n = 1; lam[0] = T(1); v = b; return true; }
        const T vc = d1 * d4 - d3 * d2;
        if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
            const T t = d1 / (d1 - d3);
            n = 2; lam[0] = T(1) - t; lam[1] = t; v = a + ab * t;
            return true;
        }
        const V3<T> cp = -c;
        const T d5 = dot(ab, cp), d6 = dot(ac, cp);
        if (d6 >= T(0) && d5 <= d6) { s[0] = s[2]; n = 1; lam[0] = T(1); v = c; return true; }
        const T vb = d5 * d2 - d1 * d6;
        if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
            const T t = d2 / (d2 - d6);
            s[1] = s[2]; n = 2; lam[0] = T(1) - t; lam[1] = t; v = a + ac * t;
            return true;
        }
        const T va = d3 * d6 - d5 * d4;
        if (va <= T(0) && (d4 - d3) >= T(0) && (d5 - d6) >= T(0)) {
            const T t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            s[0] = s[1]; s[1] = s[2]; n = 2; lam[0] = T(1) - t; lam[1] = t; v = b + (c - b) * t;
            return true;
        }
        const T den = T(1) / (va + vb + vc);
        const T bv = vb * den, cw = vc * den;
        lam[0] = T(1) - bv - cw; lam[1] = bv; lam[2] = cw;
        v = a + ab * bv + ac * cw;
        return true;
    }
    // tetrahedron: examine the faces the origin lies outside of
    static const int F[4][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {1, 3, 2, 0}};
    bool inside = true;
    T best = T(0);
    int bestn = 0;
    SV<T> bs[3];
    T bl[3];
    V3<T> bv3 = {T(0), T(0), T(0)};
    for (int f = 0; f < 4; ++f) {
        const V3<T> a = s[F[f][0]].w, b = s[F[f][1]].w, c = s[F[f][2]].w, d = s[F[f][3]].w;
        const V3<T> nrm = cross(b - a, c - a);
        const T sp = dot(-a, nrm), sd = dot(d - a, nrm);
        if (sp * sd <= T(0)) {
            inside = false;
            SV<T> ts[4] = {s[F[f][0]], s[F[f][1]], s[F[f][2]], s[F[f][2]]};
            int tn = 3;
            T tl[4];
            V3<T> tv;
            simplex_closest(ts, tn, tl, tv);
            const T dd = dot(tv, tv);
            if (bestn == 0 || dd < best) {
                best = dd;
                bestn = tn;
                for (int i = 0; i < tn; ++i) { bs[i] = ts[i]; bl[i] = tl[i]; }
                bv3 = tv;
            }
        }
    }
    if (inside) {
        v = {T(0), T(0), T(0)};
        return false;
    }
    n = bestn;
    for (int i = 0; i < n; ++i) { s[i] = bs[i]; lam[i] = bl[i]; }
    v = bv3;
    return true;
}

// GJK and EPA run in a frame centred on A (coordinates of the size of the
// shapes, not of the scene) and translate their witness points back.
template <class T> inline void centre_on_a(const Shape<T>& A, const Shape<T>& B, Shape<T>& A0, Shape<T>& B0) {
    A0 = A;
    B0 = B;
    A0.c = {T(0), T(0), T(0)};
    B0.c = B.c - A.c;
}

template <class T> inline GjkOut<T> gjk(const Shape<T>& Aw, const Shape<T>& Bw, int max_iter = 32) {
    Shape<T> A, B;
    centre_on_a(Aw, Bw, A, B);
    GjkOut<T> o;
    o.status = -1;
    V3<T> d = B.c - A.c;
    if (dot(d, d) < T(1e-12)) d = {T(1), T(0), T(0)};
    o.s[0] = mink_support(A, B, d);
    o.n = 1;
    T lam[4] = {T(1), T(0), T(0), T(0)};
    V3<T> v = o.s[0].w;
    int it = 0;
    for (; it < max_iter; ++it) {
        const T vv = dot(v, v);
        if (vv <= T(1e-12)) { o.status = 1; break; }
        const SV<T> w = mink_support(A, B, -v);
        if (vv - dot(v, w.w) <= T(1e-5) * vv + T(1e-9)) { o.status = 0; break; }
        bool dup = false;
        for (int i = 0; i < o.n; ++i)
            if (o.s[i].w.x == w.w.x && o.s[i].w.y == w.w.y && o.s[i].w.z == w.w.z) dup = true;
        if (dup) { o.status = 0; break; }
        SV<T> ps[4];
        T pl[4];
        const int pn = o.n;
        for (int i = 0; i < pn; ++i) { ps[i] = o.s[i]; pl[i] = lam[i]; }
        const V3<T> pv = v;
        o.s[o.n++] = w;
        if (!simplex_closest(o.s, o.n, lam, v)) { o.status = 1; break; }
        if (!(dot(v, v) < vv)) {
            // no progress: keep the previous simplex (numerical limit reached)
            o.n = pn;
            for (int i = 0; i < pn; ++i) { o.s[i] = ps[i]; lam[i] = pl[i]; }
            v = pv;
            o.status = 0;
            break;
        }
    }
    o.iters = it;
    V3<T> pa = {T(0), T(0), T(0)}, pb = {T(0), T(0), T(0)};
    for (int i = 0; i < o.n; ++i) {
        pa = pa + o.s[i].a * lam[i];
        pb = pb + o.s[i].b * lam[i];
    }
    o.pA = pa + Aw.c;
    o.pB = pb + Aw.c;
    o.dist = o.status == 1 ? T(0) : std::sqrt(dot(v, v));
    return o;
}

// ---------------------------------------------------------------- EPA
template <class T> struct EpaOut {
    int status;  // 0 ok, -1 failed / degenerate
    V3<T> n;     // outward normal of A - B (= A -> B)
    T depth;
    V3<T> pA, pB;
    int iters;
};

template <class T> inline EpaOut<T> epa(const Shape<T>& Aw, const Shape<T>& Bw, const GjkOut<T>& g, int max_iter = 48) {
    Shape<T> A, B;
    centre_on_a(Aw, Bw, A, B);
    constexpr int MAXV = 64, MAXF = 128;
    SV<T> vt[MAXV];
    int fv[MAXF][3];
    V3<T> fn[MAXF];
    T fd[MAXF];
    int nv = 0, nf = 0;
    EpaOut<T> o;
    o.status = -1;
    o.depth = T(0);
    o.n = {T(0), T(1), T(0)};
    o.pA = o.pB = Aw.c;
    o.iters = 0;
    for (int i = 0; i < g.n; ++i) vt[nv++] = g.s[i];
    if (nv == 1) return o;
    if (nv == 2) {
        const V3<T> dd = vt[1].w - vt[0].w;
        const T ax = std::fabs(dd.x), ay = std::fabs(dd.y), az = std::fabs(dd.z);
        V3<T> e = ax <= ay && ax <= az ? V3<T>{T(1), T(0), T(0)}
                                       : (ay <= az ? V3<T>{T(0), T(1), T(0)} : V3<T>{T(0), T(0), T(1)});
        V3<T> p = cross(dd, e);
        const T pl = std::sqrt(dot(p, p));
        if (!(pl > T(0))) return o;
        p = p * (T(1) / pl);
        const T dl = std::sqrt(dot(dd, dd));
        const V3<T> u = dd * (T(1) / dl);
        const T s3 = T(0.86602540378443864676);
        V3<T> dir = p;
        for (int k = 0; k < 3; ++k) {
            vt[nv++] = mink_support(A, B, dir);
            dir = dir * T(-0.5) + cross(u, dir) * s3;
        }
        const int F6[6][3] = {{0, 2, 3}, {0, 3, 4}, {0, 4, 2}, {1, 3, 2}, {1, 4, 3}, {1, 2, 4}};
        for (int f = 0; f < 6; ++f) { fv[nf][0] = F6[f][0]; fv[nf][1] = F6[f][1]; fv[nf][2] = F6[f][2]; ++nf; }
    } else if (nv == 3) {
        V3<T> nn = cross(vt[1].w - vt[0].w, vt[2].w - vt[0].w);
        const T l = std::sqrt(dot(nn, nn));
        if (!(l > T(0))) return o;
        nn = nn * (T(1) / l);
        vt[nv++] = mink_support(A, B, nn);
        vt[nv++] = mink_support(A, B, -nn);
        const int F6[6][3] = {{0, 1, 3}, {1, 2, 3}, {2, 0, 3}, {1, 0, 4}, {2, 1, 4}, {0, 2, 4}};
        for (int f = 0; f < 6; ++f) { fv[nf][0] = F6[f][0]; fv[nf][1] = F6[f][1]; fv[nf][2] = F6[f][2]; ++nf; }
    } else {
        const int F4[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
        for (int f = 0; f < 4; ++f) { fv[nf][0] = F4[f][0]; fv[nf][1] = F4[f][1]; fv[nf][2] = F4[f][2]; ++nf; }
    }
    // orient every face away from the centroid; reject degenerate faces
    V3<T> cen = {T(0), T(0), T(0)};
    for (int i = 0; i < nv; ++i) cen = cen + vt[i].w;
    cen = cen * (T(1) / T(nv));
    auto make_face = [&](int f) -> bool {
        const V3<T> a = vt[fv[f][0]].w, b = vt[fv[f][1]].w, c = vt[fv[f][2]].w;
        V3<T> nn = cross(b - a, c - a);
        const T l = std::sqrt(dot(nn, nn));
        if (!(l > T(1e-12))) return false;
        nn = nn * (T(1) / l);
        fn[f] = nn;
        fd[f] = dot(nn, a);
        return true;
    };
    for (int f = 0; f < nf; ++f) {
        const V3<T> a = vt[fv[f][0]].w;
        const V3<T> nn = cross(vt[fv[f][1]].w - a, vt[fv[f][2]].w - a);
        if (dot(nn, cen - a) > T(0)) { const int t = fv[f][1]; fv[f][1] = fv[f][2]; fv[f][2] = t; }
        if (!make_face(f)) return o;
    }
    int best = 0;
    int it = 0;
    bool converged = false;
    for (; it < max_iter; ++it) {
        best = 0;
        for (int f = 1; f < nf; ++f)
            if (fd[f] < fd[best]) best = f;
        const SV<T> w = mink_support(A, B, fn[best]);
        const T dw = dot(w.w, fn[best]);
        if (dw - fd[best] <= T(1e-5) + T(1e-5) * std::fabs(fd[best])) { converged = true; break; }
        if (nv >= MAXV) break;
        const int wi = nv;
        vt[nv++] = w;
        // remove faces visible from w, collecting the horizon
        int edges[MAXF * 3][2];
        int ne = 0;
        int keep = 0;
        for (int f = 0; f < nf; ++f) {
            if (dot(fn[f], w.w - vt[fv[f][0]].w) > T(0)) {
                for (int k = 0; k < 3; ++k) {
                    const int e0 = fv[f][k], e1 = fv[f][(k + 1) % 3];
                    int found = -1;
                    for (int q = 0; q < ne; ++q)
                        if (edges[q][0] == e1 && edges[q][1] == e0) { found = q; break; }
                    if (found >= 0) {
                        edges[found][0] = edges[ne - 1][0];
                        edges[found][1] = edges[ne - 1][1];
                        --ne;
                    } else {
                        edges[ne][0] = e0;
                        edges[ne][1] = e1;
                        ++ne;
                    }
                }
            } else {
                fv[keep][0] = fv[f][0]; fv[keep][1] = fv[f][1]; fv[keep][2] = fv[f][2];
                fn[keep] = fn[f];
                fd[keep] = fd[f];
                ++keep;
            }
        }
        nf = keep;
        if (nf + ne > MAXF) break;
        bool bad = false;
        for (int q = 0; q < ne; ++q) {
            fv[nf][0] = edges[q][0];
            fv[nf][1] = edges[q][1];
            fv[nf][2] = wi;
            if (!make_face(nf)) bad = true;
            ++nf;
        }
        if (bad || nf == 0) break;
    }
    o.iters = it;
    best = 0;
    for (int f = 1; f < nf; ++f)
        if (fd[f] < fd[best]) best = f;
    // barycentric weights of the origin's projection onto the closest face
    const V3<T> pr = fn[best] * fd[best];
    const V3<T> a = vt[fv[best][0]].w, b = vt[fv[best][1]].w, c = vt[fv[best][2]].w;
    const V3<T> e0 = b - a, e1 = c - a, e2 = pr - a;
    const T d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1), d20 = dot(e2, e0), d21 = dot(e2, e1);
    const T den = d00 * d11 - d01 * d01;
    T lv = T(0), lw = T(0);
    if (den != T(0)) {
        lv = (d11 * d20 - d01 * d21) / den;
        lw = (d00 * d21 - d01 * d20) / den;
    }
    const T lu = T(1) - lv - lw;
    o.pA = vt[fv[best][0]].a * lu + vt[fv[best][1]].a * lv + vt[fv[best][2]].a * lw + Aw.c;
    o.pB = vt[fv[best][0]].b * lu + vt[fv[best][1]].b * lv + vt[fv[best][2]].b * lw + Aw.c;
    o.n = fn[best];
    o.depth = fd[best];
    o.status = converged ? 0 : -1;
    return o;
}

// ---------------------------------------------------------------- box-capsule (GJK / EPA)
template <class T> inline int box_capsule(Man<T>& m, const Shape<T>& A, const Shape<T>& B, T margin) {
    const T r = B.prm.x;
    const GjkOut<T> g = gjk(A, B);
    if (g.status == 1) {
        const EpaOut<T> e = epa(A, B, g);
        if (e.status != 0 && !(e.depth > T(0))) return GEOM_EPA_FAIL;
        man_add(m, e.pA, e.n, e.depth + r, 3u);
        return e.status == 0 ? GEOM_OK : GEOM_EPA_FAIL;
    }
    if (g.dist > r + margin) return g.status == 0 ? GEOM_OK : GEOM_GJK_FAIL;
    if (!(g.dist > T(1e-6))) return GEOM_OK;
    const V3<T> n = (g.pB - g.pA) * (T(1) / g.dist);
    const V3<T> u = B.R.col(1);
    if (std::fabs(dot(u, n)) < T(0.2)) {
        // capsule lying along a face: try one contact per segment end
        V3<T> e[2];
        capsule_segment(B, e[0], e[1]);
        Man<T> m2;
        for (int k = 0; k < 2; ++k) {
            const V3<T> pl = mulT(A.R, e[k] - A.c);
            const V3<T> q = {clampv(pl.x, -A.prm.x, A.prm.x), clampv(pl.y, -A.prm.y, A.prm.y),
                             clampv(pl.z, -A.prm.z, A.prm.z)};
            const V3<T> diff = pl - q;
            const T d2 = dot(diff, diff);
            const T s = r + margin;
            if (d2 > T(1e-12) && d2 <= s * s) {
                const T dist = std::sqrt(d2);
                const V3<T> nk = mul(A.R, diff * (T(1) / dist));
                man_add(m2, A.c + mul(A.R, q), nk, r - dist, (uint32_t)k);
            }
        }
        if (m2.np == 2) {
            m = m2;
            return GEOM_OK;
        }
    }
    man_add(m, g.pA, n, r - g.dist, 2u);
    return g.status == 0 ? GEOM_OK : GEOM_GJK_FAIL;
}

// ---------------------------------------------------------------- box-box (SAT + clipping)
template <class T> inline T box_proj(const Shape<T>& S, V3<T> L) {
    return S.prm.x * std::fabs(dot(S.R.col(0), L)) + S.prm.y * std::fabs(dot(S.R.col(1), L)) +
           S.prm.z * std::fabs(dot(S.R.col(2), L));
}

template <class T> inline T hcomp(V3<T> h, int k) { return k == 0 ? h.x : (k == 1 ? h.y : h.z); }

template <class T> inline void box_box(Man<T>& m, const Shape<T>& A, const Shape<T>& B, T margin) {
    const V3<T> Tv = B.c - A.c;
    T sFace[6];
    for (int i = 0; i < 3; ++i) {
        const V3<T> L = A.R.col(i);
        sFace[i] = std::fabs(dot(Tv, L)) - (hcomp(A.prm, i) + box_proj(B, L));
        if (sFace[i] > margin) return;
    }
    for (int i = 0; i < 3; ++i) {
        const V3<T> L = B.R.col(i);
        sFace[3 + i] = std::fabs(dot(Tv, L)) - (box_proj(A, L) + hcomp(B.prm, i));
        if (sFace[3 + i] > margin) return;
    }
    T sEdge = T(0);
    int eBest = -1;
    V3<T> LEdge = {T(0), T(0), T(0)};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            V3<T> L = cross(A.R.col(i), B.R.col(j));
            const T l2 = dot(L, L);
            if (l2 < T(1e-6)) continue;
            L = L * (T(1) / std::sqrt(l2));
            const T s = std::fabs(dot(Tv, L)) - (box_proj(A, L) + box_proj(B, L));
            if (s > margin) return;
            if (eBest < 0 || s > sEdge) {
                sEdge = s;
                eBest = i * 3 + j;
                LEdge = L;
            }
        }
    }
    int fA = 0;
    for (int i = 1; i < 3; ++i)
        if (sFace[i] > sFace[fA]) fA = i;
    int fB = 3;
    for (int i = 4; i < 6; ++i)
        if (sFace[i] > sFace[fB]) fB = i;
    int face = fA;
    if (sFace[fB] > T(0.95) * sFace[fA] + T(0.001)) face = fB;
    const T sF = sFace[face];
    if (eBest >= 0 && sEdge > T(0.95) * sF + T(0.001)) {
        // edge-edge contact
        V3<T> L = LEdge;
        if (dot(Tv, L) < T(0)) L = -L;
        const int i = eBest / 3, j = eBest % 3;
        V3<T> pa = A.c, pb = B.c;
        for (int k = 0; k < 3; ++k) {
            if (k != i) pa = pa + A.R.col(k) * (hcomp(A.prm, k) * sgn1(dot(A.R.col(k), L)));
            if (k != j) pb = pb + B.R.col(k) * (hcomp(B.prm, k) * (-sgn1(dot(B.R.col(k), L))));
        }
        const V3<T> ea = A.R.col(i) * hcomp(A.prm, i);
        const V3<T> eb = B.R.col(j) * hcomp(B.prm, j);
        V3<T> c1, c2;
        seg_seg(pa - ea, pa + ea, pb - eb, pb + eb, c1, c2);
        man_add(m, c1, L, -sEdge, 0xE000u | (uint32_t)eBest);
        return;
    }
    // face contact: reference face on R, incident face on I
    const bool refB = face >= 3;
    const Shape<T>& Rf = refB ? B : A;
    const Shape<T>& In = refB ? A : B;
    const int ri = refB ? face - 3 : face;
    V3<T> nRef = Rf.R.col(ri);
    if (dot(In.c - Rf.c, nRef) < T(0)) nRef = -nRef;
    // incident face: most anti-parallel to nRef
    int ij = 0;
    T best = std::fabs(dot(In.R.col(0), nRef));
    for (int k = 1; k < 3; ++k) {
        const T d = std::fabs(dot(In.R.col(k), nRef));
        if (d > best) { best = d; ij = k; }
    }
    const V3<T> ax = In.R.col(ij);
    const V3<T> fnI = dot(ax, nRef) > T(0) ? -ax : ax;
    const V3<T> fc = In.c + fnI * hcomp(In.prm, ij);
    const int iu = (ij + 1) % 3, iv = (ij + 2) % 3;
    const V3<T> u = In.R.col(iu) * hcomp(In.prm, iu);
    const V3<T> v = In.R.col(iv) * hcomp(In.prm, iv);
    V3<T> poly[8], tmp[8];
    int np = 4;
    poly[0] = fc + u + v;
    poly[1] = fc - u + v;
    poly[2] = fc - u - v;
    poly[3] = fc + u - v;
    // clip against the four side planes of the reference face
    const int ru = (ri + 1) % 3, rv = (ri + 2) % 3;
    for (int p = 0; p < 4; ++p) {
        const V3<T> axp = Rf.R.col(p < 2 ? ru : rv);
        const T hp = hcomp(Rf.prm, p < 2 ? ru : rv);
        const V3<T> pn = (p & 1) ? -axp : axp;
        const T pd = dot(pn, Rf.c) + hp;
        // at most eight vertices are kept: a convex face gains at most one per plane, and
        // more can only come from rounding on a nearly degenerate face
        int nt = 0;
        auto put = [&](const V3<T>& q) {
            if (nt < 8) tmp[nt] = q;
            ++nt;
        };
        for (int k = 0; k < np; ++k) {
            const V3<T> a = poly[(k + np - 1) % np], b = poly[k];
            const T da = dot(pn, a) - pd, db = dot(pn, b) - pd;
            if (da <= T(0)) {
                if (db <= T(0)) {
                    put(b);
                } else {
                    const T t = da / (da - db);
                    put(a + (b - a) * t);
                }
            } else if (db <= T(0)) {
                const T t = da / (da - db);
                put(a + (b - a) * t);
                put(b);
            }
        }
        np = nt < 8 ? nt : 8;
        for (int k = 0; k < np; ++k) poly[k] = tmp[k];
        if (np == 0) return;
    }
    const T refD = dot(nRef, Rf.c) + hcomp(Rf.prm, ri);
    V3<T> cp[8];
    T cd[8];
    int cid[8];
    int nc = 0;
    for (int k = 0; k < np; ++k) {
        const T d = refD - dot(nRef, poly[k]);
        if (d > -margin) {
            cp[nc] = poly[k];
            cd[nc] = d;
            cid[nc] = k;
            ++nc;
        }
    }
    if (nc == 0) return;
    int sel[4];
    int ns = 0;
    if (nc <= 4) {
        for (int k = 0; k < nc; ++k) sel[ns++] = k;
    } else {
        int i0 = 0;
        for (int k = 1; k < nc; ++k)
            if (cd[k] > cd[i0]) i0 = k;
        int i1 = -1;
        T dm = T(0);
        for (int k = 0; k < nc; ++k) {
            if (k == i0) continue;
            const V3<T> dd = cp[k] - cp[i0];
            const T d2 = dot(dd, dd);
            if (i1 < 0 || d2 > dm) { dm = d2; i1 = k; }
        }
        int i2 = -1, i3 = -1;
        T amax = T(0), amin = T(0);
        for (int k = 0; k < nc; ++k) {
            if (k == i0 || k == i1) continue;
            const T a = dot(cross(cp[i1] - cp[i0], cp[k] - cp[i0]), nRef);
            if (i2 < 0 || a > amax) { amax = a; i2 = k; }
            if (i3 < 0 || a < amin) { amin = a; i3 = k; }
        }
        if (i3 == i2) {
            for (int k = 0; k < nc; ++k)
                if (k != i0 && k != i1 && k != i2) { i3 = k; break; }
        }
        bool pick[8] = {false, false, false, false, false, false, false, false};
        pick[i0] = pick[i1] = pick[i2] = pick[i3] = true;
        for (int k = 0; k < nc; ++k)
            if (pick[k]) sel[ns++] = k;
    }
    const uint32_t fbase = ((refB ? 1u : 0u) << 12) | ((uint32_t)ri << 8) | ((uint32_t)ij << 4);
    for (int s = 0; s < ns; ++s) {
        const int k = sel[s];
        if (refB) {
            man_add(m, cp[k], -nRef, cd[k], fbase | (uint32_t)cid[k]);
        } else {
            man_add(m, cp[k] + nRef * cd[k], nRef, cd[k], fbase | (uint32_t)cid[k]);
        }
    }
}

// ---------------------------------------------------------------- dispatch
// Canonical ordering: type(A) <= type(B); otherwise the pair is swapped and the
// result converted back (normal negated, witness moved to the other body).
template <class T> inline int collide_canon(Man<T>& m, const Shape<T>& A, const Shape<T>& B, T margin) {
    const uint32_t ta = A.type, tb = B.type;
    if (ta == SHAPE_SPHERE && tb == SHAPE_SPHERE) { sphere_sphere(m, A.c, A.prm.x, B.c, B.prm.x, margin, 0u); return GEOM_OK; }
    if (ta == SHAPE_SPHERE && tb == SHAPE_BOX) { sphere_box(m, A.c, A.prm.x, B, margin); return GEOM_OK; }
    if (ta == SHAPE_SPHERE && tb == SHAPE_CAPSULE) {
        V3<T> p0, p1;
        capsule_segment(B, p0, p1);
        sphere_sphere(m, A.c, A.prm.x, closest_on_segment(A.c, p0, p1), B.prm.x, margin, 0u);
        return GEOM_OK;
    }
    if (ta == SHAPE_BOX && tb == SHAPE_BOX) { box_box(m, A, B, margin); return GEOM_OK; }
    if (ta == SHAPE_BOX && tb == SHAPE_CAPSULE) return box_capsule(m, A, B, margin);
    if (ta == SHAPE_CAPSULE && tb == SHAPE_CAPSULE) { capsule_capsule(m, A, B, margin); return GEOM_OK; }
    return GEOM_OK;
}

template <class T> inline int collide(Man<T>& m, const Shape<T>& A, const Shape<T>& B, T margin) {
    m.np = 0;
    if (A.type <= B.type) return collide_canon(m, A, B, margin);
    Man<T> t;
    const int st = collide_canon(t, B, A, margin);
    m.np = t.np;
    for (uint32_t k = 0; k < t.np; ++k) {
        m.pt[k].p = t.pt[k].p - t.pt[k].n * t.pt[k].depth;
        m.pt[k].n = -t.pt[k].n;
        m.pt[k].depth = t.pt[k].depth;
        m.pt[k].fid = t.pt[k].fid;
    }
    return st;
}

template <class T> inline void collide_plane(Man<T>& m, const Shape<T>& A, V3<T> pn, T pd, T margin) {
    m.np = 0;
    if (A.type == SHAPE_SPHERE) {
        sphere_plane(m, A.c, A.prm.x, pn, pd, margin, 0u);
    } else if (A.type == SHAPE_BOX) {
        box_plane(m, A, pn, pd, margin);
    } else if (A.type == SHAPE_CAPSULE) {
        V3<T> p0, p1;
        capsule_segment(A, p0, p1);
        sphere_plane(m, p0, A.prm.x, pn, pd, margin, 0u);
        sphere_plane(m, p1, A.prm.x, pn, pd, margin, 1u);
    }
}

}  // namespace orc
