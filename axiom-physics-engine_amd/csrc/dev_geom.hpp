// Device narrowphase: closed-form primitive tests, box-box SAT with reference-
// face clipping, and GJK / EPA on shape cores (box polytope, capsule segment).
//
// The reference reserves this stage (collision::Shape / collision::ContactPoint,
// include/axiom/debug/physics_debug_draw.hpp:14-17; GJKFailedToConverge /
// EPAFailedToConverge, include/axiom/core/error_code.hpp:37-38) without an
// implementation.  Algorithms: Ericson, Real-Time Collision Detection (2005)
// ch. 4-5 (closed forms, SAT, segment-segment closest points 5.1.9, closest
// point on triangle / tetrahedron 5.1.5-5.1.6); GJK (Gilbert, Johnson & Keerthi
// 1988); EPA (van den Bergen 2001).  Contact convention: normal from A to B,
// depth > 0 when penetrating, witness point on A, witness on B = p - n*depth
// (src/debug/physics_debug_draw.cpp:111-115).  Capsules along local +Y
// (src/debug/physics_debug_draw.cpp:254-266).
#pragma once
#include "dev_math.hpp"

namespace axm {

enum : uint32_t { SHAPE_SPHERE = 0, SHAPE_BOX = 1, SHAPE_CAPSULE = 2, SHAPE_PLANE = 3 };
enum : int { GEOM_OK = 0, GEOM_GJK_FAIL = 301, GEOM_EPA_FAIL = 302 };

struct Shp {
    f3 c;
    m3 R;
    uint32_t type;
    f3 prm;
};

struct CPt {
    f3 p;
    f3 n;
    float depth;
    uint32_t fid;
};
struct Man {
    uint32_t np;
    CPt pt[4];
};

AXM_D void man_add(Man& m, f3 p, f3 n, float depth, uint32_t fid) {
    // every slot is rewritten with per-component selects (constant indices, no
    // conditional stores), so that the manifold stays in registers: conditional stores
    // get merged into one indexed store, which places the manifold in scratch memory
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const bool w = k == m.np;
        CPt& q = m.pt[k];
        q.p.x = w ? p.x : q.p.x;
        q.p.y = w ? p.y : q.p.y;
        q.p.z = w ? p.z : q.p.z;
        q.n.x = w ? n.x : q.n.x;
        q.n.y = w ? n.y : q.n.y;
        q.n.z = w ? n.z : q.n.z;
        q.depth = w ? depth : q.depth;
        q.fid = w ? fid : q.fid;
    }
    m.np = m.np < 4u ? m.np + 1u : m.np;
}

AXM_D void sphere_sphere(Man& m, f3 cA, float rA, f3 cB, float rB, float margin, uint32_t fid) {
    const f3 d = cB - cA;
    const float d2 = dot(d, d);
    const float R = rA + rB;
    const float s = R + margin;
    if (d2 > s * s) return;
    const float dist = sqrtf(d2);
    f3 n;
    if (dist > 1e-6f) {
        const float inv = 1.0f / dist;
        n = d * inv;
    } else {
        n = {0.0f, 1.0f, 0.0f};
    }
    man_add(m, cA + n * rA, n, R - dist, fid);
}

AXM_D f3 closest_on_segment(f3 p, f3 a, f3 b) {
    const f3 e = b - a;
    const float ee = dot(e, e);
    float t = 0.0f;
    if (ee > 0.0f) t = dot(p - a, e) / ee;
    t = clampf(t, 0.0f, 1.0f);
    return a + e * t;
}

AXM_D void sphere_box(Man& m, f3 c, float r, const Shp& B, float margin) {
    const f3 h = B.prm;
    const f3 pl = mulT(B.R, c - B.c);
    const f3 q = {clampf(pl.x, -h.x, h.x), clampf(pl.y, -h.y, h.y), clampf(pl.z, -h.z, h.z)};
    const f3 diff = q - pl;
    const float d2 = dot(diff, diff);
    if (d2 > 1e-12f) {
        const float s = r + margin;
        if (d2 > s * s) return;
        const float dist = sqrtf(d2);
        const float inv = 1.0f / dist;
        const f3 n = mul(B.R, diff * inv);
        man_add(m, c + n * r, n, r - dist, 0u);
        return;
    }
    const float fx = h.x - fabsf(pl.x), fy = h.y - fabsf(pl.y), fz = h.z - fabsf(pl.z);
    int k = 0;
    float f = fx;
    if (fy < f) { k = 1; f = fy; }
    if (fz < f) { k = 2; f = fz; }
    const float comp = k == 0 ? pl.x : (k == 1 ? pl.y : pl.z);
    const f3 n = B.R.col(k) * (-sgn1(comp));
    man_add(m, c + n * r, n, r + f, 1u);
}

AXM_D void capsule_segment(const Shp& S, f3& p0, f3& p1) {
    const f3 u = S.R.col(1);
    p0 = S.c - u * S.prm.y;
    p1 = S.c + u * S.prm.y;
}

AXM_D void sphere_plane(Man& m, f3 c, float r, f3 pn, float pd, float margin, uint32_t fid) {
    const float dist = dot(pn, c) - pd;
    if (dist - r > margin) return;
    man_add(m, c - pn * r, -pn, r - dist, fid);
}

AXM_D void box_plane(Man& m, const Shp& A, f3 pn, float pd, float margin) {
    float dep[8];
    f3 vx[8];
    uint32_t candmask = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const f3 loc = {(k & 1) ? A.prm.x : -A.prm.x, (k & 2) ? A.prm.y : -A.prm.y, (k & 4) ? A.prm.z : -A.prm.z};
        vx[k] = A.c + mul(A.R, loc);
        dep[k] = pd - dot(pn, vx[k]);
        if (dep[k] > -margin) candmask |= 1u << k;
    }
    const int nc = __popc(candmask);
    uint32_t keep = 0;
    const int want = nc < 4 ? nc : 4;
    for (int s = 0; s < want; ++s) {
        int best = -1;
        float bd = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (!((candmask >> k) & 1u) || ((keep >> k) & 1u)) continue;
            if (best < 0 || dep[k] > bd) { best = k; bd = dep[k]; }
        }
        keep |= 1u << best;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if ((keep >> k) & 1u) man_add(m, vx[k], -pn, dep[k], (uint32_t)k);
}

AXM_D void seg_seg(f3 p1, f3 q1, f3 p2, f3 q2, f3& c1, f3& c2) {
    const f3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    float s, t;
    const float eps = 1e-12f;
    if (a <= eps && e <= eps) {
        s = t = 0.0f;
    } else if (a <= eps) {
        s = 0.0f;
        t = clampf(f / e, 0.0f, 1.0f);
    } else {
        const float c = dot(d1, r);
        if (e <= eps) {
            t = 0.0f;
            s = clampf(-c / a, 0.0f, 1.0f);
        } else {
            const float b = dot(d1, d2);
            const float den = a * e - b * b;
            s = den != 0.0f ? clampf((b * f - c * e) / den, 0.0f, 1.0f) : 0.0f;
            t = (b * s + f) / e;
            if (t < 0.0f) {
                t = 0.0f;
                s = clampf(-c / a, 0.0f, 1.0f);
            } else if (t > 1.0f) {
                t = 1.0f;
                s = clampf((b - c) / a, 0.0f, 1.0f);
            }
        }
    }
    c1 = p1 + d1 * s;
    c2 = p2 + d2 * t;
}

AXM_D void capsule_capsule(Man& m, const Shp& A, const Shp& B, float margin) {
    f3 a0, a1, b0, b1;
    capsule_segment(A, a0, a1);
    capsule_segment(B, b0, b1);
    const float rA = A.prm.x, rB = B.prm.x;
    const f3 dA = a1 - a0, dB = b1 - b0;
    const f3 cr = cross(dA, dB);
    const float aa = dot(dA, dA), bb = dot(dB, dB);
    if (aa > 1e-12f && bb > 1e-12f && dot(cr, cr) <= 1e-6f * aa * bb) {
        const float tb0 = dot(b0 - a0, dA) / aa;
        const float tb1 = dot(b1 - a0, dA) / aa;
        const float lo = clampf(tb0 < tb1 ? tb0 : tb1, 0.0f, 1.0f);
        const float hi = clampf(tb0 < tb1 ? tb1 : tb0, 0.0f, 1.0f);
        if (hi - lo > 1e-4f) {
            const f3 pa0 = a0 + dA * lo;
            const f3 pa1 = a0 + dA * hi;
            sphere_sphere(m, pa0, rA, closest_on_segment(pa0, b0, b1), rB, margin, 0u);
            sphere_sphere(m, pa1, rA, closest_on_segment(pa1, b0, b1), rB, margin, 1u);
            return;
        }
    }
    f3 c1, c2;
    seg_seg(a0, a1, b0, b1, c1, c2);
    sphere_sphere(m, c1, rA, c2, rB, margin, 2u);
}

// ---------------------------------------------------------------- GJK
struct SV {
    f3 w, a, b;
};
struct GjkOut {
    int status;  // 0 separated, 1 cores intersect, -1 iteration cap
    float dist;
    f3 pA, pB;
    int n;
    SV s[4];
};

AXM_D f3 core_support(const Shp& S, f3 d) {
    if (S.type == SHAPE_BOX) {
        const f3 dl = mulT(S.R, d);
        const f3 l = {sgn1(dl.x) * S.prm.x, sgn1(dl.y) * S.prm.y, sgn1(dl.z) * S.prm.z};
        return S.c + mul(S.R, l);
    }
    if (S.type == SHAPE_CAPSULE) {
        f3 p0, p1;
        capsule_segment(S, p0, p1);
        return dot(d, p1 - p0) >= 0.0f ? p1 : p0;
    }
    return S.c;
}

AXM_D SV mink_support(const Shp& A, const Shp& B, f3 d) {
    SV r;
    r.a = core_support(A, d);
    r.b = core_support(B, -d);
    r.w = r.a - r.b;
    return r;
}

// Closest point on a triangle simplex (Ericson 5.1.5), reducing s[0..3).
AXM_D void tri_closest(SV* s, int& n, float* lam, f3& v) {
    const f3 a = s[0].w, b = s[1].w, c = s[2].w;
    const f3 ab = b - a, ac = c - a, ap = -a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) { n = 1; lam[0] = 1.0f; v = a; return; }
    const f3 bp = -b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) { s[0] = s[1]; n = 1; lam[0] = 1.0f; v = b; return; }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float t = d1 / (d1 - d3);
        n = 2; lam[0] = 1.0f - t; lam[1] = t; v = a + ab * t;
        return;
    }
    const f3 cp = -c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) { s[0] = s[2]; n = 1; lam[0] = 1.0f; v = c; return; }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float t = d2 / (d2 - d6);
        s[1] = s[2]; n = 2; lam[0] = 1.0f - t; lam[1] = t; v = a + ac * t;
        return;
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        s[0] = s[1]; s[1] = s[2]; n = 2; lam[0] = 1.0f - t; lam[1] = t; v = b + (c - b) * t;
        return;
    }
    const float den = 1.0f / (va + vb + vc);
    const float bv = vb * den, cw = vc * den;
    lam[0] = 1.0f - bv - cw; lam[1] = bv; lam[2] = cw;
    v = a + ab * bv + ac * cw;
}

AXM_D bool simplex_closest(SV* s, int& n, float* lam, f3& v) {
    if (n == 1) {
        lam[0] = 1.0f;
        v = s[0].w;
        return true;
    }
    if (n == 2) {
        const f3 a = s[0].w, ab = s[1].w - s[0].w;
        const float den = dot(ab, ab);
        const float t = den > 0.0f ? dot(-a, ab) / den : 0.0f;
        if (t <= 0.0f) { n = 1; lam[0] = 1.0f; v = a; return true; }
        if (t >= 1.0f) { s[0] = s[1]; n = 1; lam[0] = 1.0f; v = s[0].w; return true; }
        lam[0] = 1.0f - t;
        lam[1] = t;
        v = a + ab * t;
        return true;
    }
    if (n == 3) {
        tri_closest(s, n, lam, v);
        return true;
    }
    const int F[4][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {1, 3, 2, 0}};
    bool inside = true;
    float best = 0.0f;
    int bestn = 0;
    SV bs[3];
    float bl[3] = {0.0f, 0.0f, 0.0f};
    f3 bv3 = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const f3 a = s[F[f][0]].w, b = s[F[f][1]].w, c = s[F[f][2]].w, d = s[F[f][3]].w;
        const f3 nrm = cross(b - a, c - a);
        const float sp = dot(-a, nrm), sd = dot(d - a, nrm);
        if (sp * sd <= 0.0f) {
            inside = false;
            SV ts[3] = {s[F[f][0]], s[F[f][1]], s[F[f][2]]};
            int tn = 3;
            float tl[3] = {0.0f, 0.0f, 0.0f};
            f3 tv;
            tri_closest(ts, tn, tl, tv);
            const float dd = dot(tv, tv);
            if (bestn == 0 || dd < best) {
                best = dd;
                bestn = tn;
                // all three copied (constant indices keep the simplex in registers);
                // entries from tn on are never read
#pragma unroll
                for (int i = 0; i < 3; ++i) { bs[i] = ts[i]; bl[i] = tl[i]; }
                bv3 = tv;
            }
        }
    }
    if (inside) {
        v = {0.0f, 0.0f, 0.0f};
        return false;
    }
    n = bestn;
#pragma unroll
    for (int i = 0; i < 3; ++i) { s[i] = bs[i]; lam[i] = bl[i]; }
    v = bv3;
    return true;
}

// GJK / EPA work in a frame centred on A and translate witnesses back.
AXM_D void centre_on_a(const Shp& A, const Shp& B, Shp& A0, Shp& B0) {
    A0 = A;
    B0 = B;
    A0.c = {0.0f, 0.0f, 0.0f};
    B0.c = B.c - A.c;
}

AXM_D GjkOut gjk(const Shp& Aw, const Shp& Bw, int max_iter = 32) {
    Shp A, B;
    centre_on_a(Aw, Bw, A, B);
    GjkOut o;
    o.status = -1;
    f3 d = B.c - A.c;
    if (dot(d, d) < 1e-12f) d = {1.0f, 0.0f, 0.0f};
    o.s[0] = mink_support(A, B, d);
    o.n = 1;
    float lam[4] = {1.0f, 0.0f, 0.0f, 0.0f};
    f3 v = o.s[0].w;
    for (int it = 0; it < max_iter; ++it) {
        const float vv = dot(v, v);
        if (vv <= 1e-12f) { o.status = 1; break; }
        const SV w = mink_support(A, B, -v);
        if (vv - dot(v, w.w) <= 1e-5f * vv + 1e-9f) { o.status = 0; break; }
        // loops over the four slots with constant indices keep the simplex in registers;
        // slots from o.n on are never read
        bool dup = false;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (i < o.n && o.s[i].w.x == w.w.x && o.s[i].w.y == w.w.y && o.s[i].w.z == w.w.z) dup = true;
        if (dup) { o.status = 0; break; }
        SV ps[4];
        float pl[4];
        const int pn = o.n;
#pragma unroll
        for (int i = 0; i < 4; ++i) { ps[i] = o.s[i]; pl[i] = lam[i]; }
        const f3 pv = v;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (i == o.n) o.s[i] = w;
        ++o.n;
        if (!simplex_closest(o.s, o.n, lam, v)) { o.status = 1; break; }
        if (!(dot(v, v) < vv)) {
            o.n = pn;
#pragma unroll
            for (int i = 0; i < 4; ++i) { o.s[i] = ps[i]; lam[i] = pl[i]; }
            v = pv;
            o.status = 0;
            break;
        }
    }
    f3 pa = {0.0f, 0.0f, 0.0f}, pb = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= o.n) break;
        pa = pa + o.s[i].a * lam[i];
        pb = pb + o.s[i].b * lam[i];
    }
    o.pA = pa + Aw.c;
    o.pB = pb + Aw.c;
    o.dist = o.status == 1 ? 0.0f : sqrtf(dot(v, v));
    return o;
}

// ---------------------------------------------------------------- EPA
struct EpaOut {
    int status;
    f3 n;
    float depth;
    f3 pA, pB;
};

constexpr int EPA_MAXV = 64, EPA_MAXF = 128;

struct EpaWork {
    SV vt[EPA_MAXV];
    int fv[EPA_MAXF][3];
    f3 fn[EPA_MAXF];
    float fd[EPA_MAXF];
    int edges[EPA_MAXF * 3][2];
};

AXM_D bool epa_make_face(EpaWork& W, int f) {
    const f3 a = W.vt[W.fv[f][0]].w, b = W.vt[W.fv[f][1]].w, c = W.vt[W.fv[f][2]].w;
    f3 nn = cross(b - a, c - a);
    const float l = sqrtf(dot(nn, nn));
    if (!(l > 1e-12f)) return false;
    nn = nn * (1.0f / l);
    W.fn[f] = nn;
    W.fd[f] = dot(nn, a);
    return true;
}

AXM_D EpaOut epa(const Shp& Aw, const Shp& Bw, const GjkOut& g, EpaWork& W, int max_iter = 48) {
    Shp A, B;
    centre_on_a(Aw, Bw, A, B);
    int nv = 0, nf = 0;
    EpaOut o;
    o.status = -1;
    o.depth = 0.0f;
    o.n = {0.0f, 1.0f, 0.0f};
    o.pA = o.pB = Aw.c;
    for (int i = 0; i < g.n; ++i) W.vt[nv++] = g.s[i];
    if (nv == 1) return o;
    if (nv == 2) {
        const f3 dd = W.vt[1].w - W.vt[0].w;
        const float ax = fabsf(dd.x), ay = fabsf(dd.y), az = fabsf(dd.z);
        const f3 e = ax <= ay && ax <= az ? f3{1.0f, 0.0f, 0.0f} : (ay <= az ? f3{0.0f, 1.0f, 0.0f} : f3{0.0f, 0.0f, 1.0f});
        f3 p = cross(dd, e);
        const float pl = sqrtf(dot(p, p));
        if (!(pl > 0.0f)) return o;
        p = p * (1.0f / pl);
        const float dl = sqrtf(dot(dd, dd));
        const f3 u = dd * (1.0f / dl);
        const float s3v = 0.86602540378443864676f;
        f3 dir = p;
        for (int k = 0; k < 3; ++k) {
            W.vt[nv++] = mink_support(A, B, dir);
            dir = dir * (-0.5f) + cross(u, dir) * s3v;
        }
        const int F6[6][3] = {{0, 2, 3}, {0, 3, 4}, {0, 4, 2}, {1, 3, 2}, {1, 4, 3}, {1, 2, 4}};
        for (int f = 0; f < 6; ++f) { W.fv[nf][0] = F6[f][0]; W.fv[nf][1] = F6[f][1]; W.fv[nf][2] = F6[f][2]; ++nf; }
    } else if (nv == 3) {
        f3 nn = cross(W.vt[1].w - W.vt[0].w, W.vt[2].w - W.vt[0].w);
        const float l = sqrtf(dot(nn, nn));
        if (!(l > 0.0f)) return o;
        nn = nn * (1.0f / l);
        W.vt[nv++] = mink_support(A, B, nn);
        W.vt[nv++] = mink_support(A, B, -nn);
        const int F6[6][3] = {{0, 1, 3}, {1, 2, 3}, {2, 0, 3}, {1, 0, 4}, {2, 1, 4}, {0, 2, 4}};
        for (int f = 0; f < 6; ++f) { W.fv[nf][0] = F6[f][0]; W.fv[nf][1] = F6[f][1]; W.fv[nf][2] = F6[f][2]; ++nf; }
    } else {
        const int F4[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
        for (int f = 0; f < 4; ++f) { W.fv[nf][0] = F4[f][0]; W.fv[nf][1] = F4[f][1]; W.fv[nf][2] = F4[f][2]; ++nf; }
    }
    f3 cen = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < nv; ++i) cen = cen + W.vt[i].w;
    cen = cen * (1.0f / (float)nv);
    for (int f = 0; f < nf; ++f) {
        const f3 a = W.vt[W.fv[f][0]].w;
        const f3 nn = cross(W.vt[W.fv[f][1]].w - a, W.vt[W.fv[f][2]].w - a);
        if (dot(nn, cen - a) > 0.0f) { const int t = W.fv[f][1]; W.fv[f][1] = W.fv[f][2]; W.fv[f][2] = t; }
        if (!epa_make_face(W, f)) return o;
    }
    int best = 0;
    bool converged = false;
    for (int it = 0; it < max_iter; ++it) {
        best = 0;
        for (int f = 1; f < nf; ++f)
            if (W.fd[f] < W.fd[best]) best = f;
        const SV w = mink_support(A, B, W.fn[best]);
        const float dw = dot(w.w, W.fn[best]);
        if (dw - W.fd[best] <= 1e-5f + 1e-5f * fabsf(W.fd[best])) { converged = true; break; }
        if (nv >= EPA_MAXV) break;
        const int wi = nv;
        W.vt[nv++] = w;
        int ne = 0;
        int keep = 0;
        for (int f = 0; f < nf; ++f) {
            if (dot(W.fn[f], w.w - W.vt[W.fv[f][0]].w) > 0.0f) {
                for (int k = 0; k < 3; ++k) {
                    const int e0 = W.fv[f][k], e1 = W.fv[f][(k + 1) % 3];
                    int found = -1;
                    for (int q = 0; q < ne; ++q)
                        if (W.edges[q][0] == e1 && W.edges[q][1] == e0) { found = q; break; }
                    if (found >= 0) {
                        W.edges[found][0] = W.edges[ne - 1][0];
                        W.edges[found][1] = W.edges[ne - 1][1];
                        --ne;
                    } else {
                        W.edges[ne][0] = e0;
                        W.edges[ne][1] = e1;
                        ++ne;
                    }
                }
            } else {
                W.fv[keep][0] = W.fv[f][0]; W.fv[keep][1] = W.fv[f][1]; W.fv[keep][2] = W.fv[f][2];
                W.fn[keep] = W.fn[f];
                W.fd[keep] = W.fd[f];
                ++keep;
            }
        }
        nf = keep;
        if (nf + ne > EPA_MAXF) break;
        bool bad = false;
        for (int q = 0; q < ne; ++q) {
            W.fv[nf][0] = W.edges[q][0];
            W.fv[nf][1] = W.edges[q][1];
            W.fv[nf][2] = wi;
            if (!epa_make_face(W, nf)) bad = true;
            ++nf;
        }
        if (bad || nf == 0) break;
    }
    best = 0;
    for (int f = 1; f < nf; ++f)
        if (W.fd[f] < W.fd[best]) best = f;
    const f3 pr = W.fn[best] * W.fd[best];
    const f3 a = W.vt[W.fv[best][0]].w, b = W.vt[W.fv[best][1]].w, c = W.vt[W.fv[best][2]].w;
    const f3 e0 = b - a, e1 = c - a, e2 = pr - a;
    const float d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1), d20 = dot(e2, e0), d21 = dot(e2, e1);
    const float den = d00 * d11 - d01 * d01;
    float lv = 0.0f, lw = 0.0f;
    if (den != 0.0f) {
        lv = (d11 * d20 - d01 * d21) / den;
        lw = (d00 * d21 - d01 * d20) / den;
    }
    const float lu = 1.0f - lv - lw;
    o.pA = W.vt[W.fv[best][0]].a * lu + W.vt[W.fv[best][1]].a * lv + W.vt[W.fv[best][2]].a * lw + Aw.c;
    o.pB = W.vt[W.fv[best][0]].b * lu + W.vt[W.fv[best][1]].b * lv + W.vt[W.fv[best][2]].b * lw + Aw.c;
    o.n = W.fn[best];
    o.depth = W.fd[best];
    o.status = converged ? 0 : -1;
    return o;
}

// ---------------------------------------------------------------- box-capsule
// Returns GEOM_OK, or GEOM_EPA_FAIL when the cores intersect (the caller runs
// EPA in a separate pass: `need_epa` set, manifold left empty).
AXM_D int box_capsule_gjk(Man& m, const Shp& A, const Shp& B, float margin, bool& need_epa) {
    need_epa = false;
    const float r = B.prm.x;
    const GjkOut g = gjk(A, B);
    if (g.status == 1) {
        need_epa = true;
        return GEOM_OK;
    }
    if (g.dist > r + margin) return g.status == 0 ? GEOM_OK : GEOM_GJK_FAIL;
    if (!(g.dist > 1e-6f)) return GEOM_OK;
    const f3 n = (g.pB - g.pA) * (1.0f / g.dist);
    const f3 u = B.R.col(1);
    if (fabsf(dot(u, n)) < 0.2f) {
        f3 e[2];
        capsule_segment(B, e[0], e[1]);
        Man m2;
        m2.np = 0;
        for (int k = 0; k < 2; ++k) {
            const f3 pl = mulT(A.R, e[k] - A.c);
            const f3 q = {clampf(pl.x, -A.prm.x, A.prm.x), clampf(pl.y, -A.prm.y, A.prm.y), clampf(pl.z, -A.prm.z, A.prm.z)};
            const f3 diff = pl - q;
            const float d2 = dot(diff, diff);
            const float s = r + margin;
            if (d2 > 1e-12f && d2 <= s * s) {
                const float dist = sqrtf(d2);
                const f3 nk = mul(A.R, diff * (1.0f / dist));
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

// Deep box-capsule: cores intersect, EPA gives the penetration.
AXM_D int box_capsule_epa(Man& m, const Shp& A, const Shp& B, EpaWork& W) {
    const GjkOut g = gjk(A, B);
    const EpaOut e = epa(A, B, g, W);
    if (e.status != 0 && !(e.depth > 0.0f)) return GEOM_EPA_FAIL;
    man_add(m, e.pA, e.n, e.depth + B.prm.x, 3u);
    return e.status == 0 ? GEOM_OK : GEOM_EPA_FAIL;
}

// ---------------------------------------------------------------- box-box
AXM_D float box_proj(const Shp& S, f3 L) {
    return S.prm.x * fabsf(dot(S.R.col(0), L)) + S.prm.y * fabsf(dot(S.R.col(1), L)) +
           S.prm.z * fabsf(dot(S.R.col(2), L));
}
AXM_D float hcomp(f3 h, int k) { return k == 0 ? h.x : (k == 1 ? h.y : h.z); }

// Up to eight points held in registers: 8-wide vectors, whose run-time element
// accesses the compiler expands into selects (an array indexed at run time would
// live in scratch memory).
typedef float vf8 __attribute__((ext_vector_type(8)));
typedef int vi8 __attribute__((ext_vector_type(8)));
struct Pts8 {
    vf8 x, y, z;
    AXM_D f3 get(int k) const { return f3{x[k], y[k], z[k]}; }
    AXM_D void set(int k, f3 v) {
        x[k] = v.x;
        y[k] = v.y;
        z[k] = v.z;
    }
};

// A shape chosen at run time, by per-field selects (a reference chosen at run time
// would keep both shapes in memory).
AXM_D Shp select_shape(bool first, const Shp& S0, const Shp& S1) {
    Shp r;
    r.c = f3{first ? S0.c.x : S1.c.x, first ? S0.c.y : S1.c.y, first ? S0.c.z : S1.c.z};
    r.R.r0 = f3{first ? S0.R.r0.x : S1.R.r0.x, first ? S0.R.r0.y : S1.R.r0.y, first ? S0.R.r0.z : S1.R.r0.z};
    r.R.r1 = f3{first ? S0.R.r1.x : S1.R.r1.x, first ? S0.R.r1.y : S1.R.r1.y, first ? S0.R.r1.z : S1.R.r1.z};
    r.R.r2 = f3{first ? S0.R.r2.x : S1.R.r2.x, first ? S0.R.r2.y : S1.R.r2.y, first ? S0.R.r2.z : S1.R.r2.z};
    r.type = first ? S0.type : S1.type;
    r.prm = f3{first ? S0.prm.x : S1.prm.x, first ? S0.prm.y : S1.prm.y, first ? S0.prm.z : S1.prm.z};
    return r;
}

AXM_D void box_box(Man& m, const Shp& A, const Shp& B, float margin) {
    // separating axes: the 15 tests are evaluated without a branch between them (a pair
    // separated along any axis has no contact, whichever axis is found first), so that
    // their arithmetic overlaps; the best edge axis is chosen in the same order
    const f3 Tv = B.c - A.c;
    vf8 sFace;
    bool apart = false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const f3 L = A.R.col(i);
        sFace[i] = fabsf(dot(Tv, L)) - (hcomp(A.prm, i) + box_proj(B, L));
        apart = apart || sFace[i] > margin;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const f3 L = B.R.col(i);
        sFace[3 + i] = fabsf(dot(Tv, L)) - (box_proj(A, L) + hcomp(B.prm, i));
        apart = apart || sFace[3 + i] > margin;
    }
    float sEdge = 0.0f;
    int eBest = -1;
    f3 LEdge = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            f3 L = cross(A.R.col(i), B.R.col(j));
            const float l2 = dot(L, L);
            const bool use = !(l2 < 1e-6f);
            L = L * (1.0f / sqrtf(l2));
            const float s = fabsf(dot(Tv, L)) - (box_proj(A, L) + box_proj(B, L));
            apart = apart || (use && s > margin);
            if (use && (eBest < 0 || s > sEdge)) {
                sEdge = s;
                eBest = i * 3 + j;
                LEdge = L;
            }
        }
    }
    if (apart) return;
    int fA = 0;
#pragma unroll
    for (int i = 1; i < 3; ++i)
        if (sFace[i] > sFace[fA]) fA = i;
    int fB = 3;
#pragma unroll
    for (int i = 4; i < 6; ++i)
        if (sFace[i] > sFace[fB]) fB = i;
    int face = fA;
    if (sFace[fB] > 0.95f * sFace[fA] + 0.001f) face = fB;
    const float sF = sFace[face];
    if (eBest >= 0 && sEdge > 0.95f * sF + 0.001f) {
        f3 L = LEdge;
        if (dot(Tv, L) < 0.0f) L = -L;
        const int i = eBest / 3, j = eBest % 3;
        f3 pa = A.c, pb = B.c;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k != i) pa = pa + A.R.col(k) * (hcomp(A.prm, k) * sgn1(dot(A.R.col(k), L)));
            if (k != j) pb = pb + B.R.col(k) * (hcomp(B.prm, k) * (-sgn1(dot(B.R.col(k), L))));
        }
        const f3 ea = A.R.col(i) * hcomp(A.prm, i);
        const f3 eb = B.R.col(j) * hcomp(B.prm, j);
        f3 c1, c2;
        seg_seg(pa - ea, pa + ea, pb - eb, pb + eb, c1, c2);
        man_add(m, c1, L, -sEdge, 0xE000u | (uint32_t)eBest);
        return;
    }
    const bool refB = face >= 3;
    const Shp Rf = select_shape(refB, B, A);
    const Shp In = select_shape(refB, A, B);
    const int ri = refB ? face - 3 : face;
    f3 nRef = Rf.R.col(ri);
    if (dot(In.c - Rf.c, nRef) < 0.0f) nRef = -nRef;
    int ij = 0;
    float best = fabsf(dot(In.R.col(0), nRef));
#pragma unroll
    for (int k = 1; k < 3; ++k) {
        const float d = fabsf(dot(In.R.col(k), nRef));
        if (d > best) { best = d; ij = k; }
    }
    const f3 ax = In.R.col(ij);
    const f3 fnI = dot(ax, nRef) > 0.0f ? -ax : ax;
    const f3 fc = In.c + fnI * hcomp(In.prm, ij);
    const int iu = (ij + 1) % 3, iv = (ij + 2) % 3;
    const f3 u = In.R.col(iu) * hcomp(In.prm, iu);
    const f3 v = In.R.col(iv) * hcomp(In.prm, iv);
    // incident face clipped against the four side planes of the reference face
    // (Sutherland-Hodgman)
    Pts8 poly, tmp;
    int np = 4;
    poly.set(0, fc + u + v);
    poly.set(1, fc - u + v);
    poly.set(2, fc - u - v);
    poly.set(3, fc + u - v);
    const int ru = (ri + 1) % 3, rv = (ri + 2) % 3;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const f3 axp = Rf.R.col(p < 2 ? ru : rv);
        const float hp = hcomp(Rf.prm, p < 2 ? ru : rv);
        const f3 pn = (p & 1) ? -axp : axp;
        const float pd = dot(pn, Rf.c) + hp;
        // at most eight vertices are kept: a convex face gains at most one per plane, and
        // more can only come from rounding on a nearly degenerate face
        int nt = 0;
        auto put = [&](f3 q) {
            if (nt < 8) tmp.set(nt, q);
            ++nt;
        };
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k >= np) break;
            const f3 a = poly.get(k == 0 ? np - 1 : k - 1), b = poly.get(k);
            const float da = dot(pn, a) - pd, db = dot(pn, b) - pd;
            if (da <= 0.0f) {
                if (db <= 0.0f) {
                    put(b);
                } else {
                    const float t = da / (da - db);
                    put(a + (b - a) * t);
                }
            } else if (db <= 0.0f) {
                const float t = da / (da - db);
                put(a + (b - a) * t);
                put(b);
            }
        }
        np = nt < 8 ? nt : 8;
        poly = tmp;
        if (np == 0) return;
    }
    const float refD = dot(nRef, Rf.c) + hcomp(Rf.prm, ri);
    Pts8 cp;
    vf8 cd;
    vi8 cid;
    int nc = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (k >= np) break;
        const f3 q = poly.get(k);
        const float d = refD - dot(nRef, q);
        if (d > -margin) {
            cp.set(nc, q);
            cd[nc] = d;
            cid[nc] = k;
            ++nc;
        }
    }
    if (nc == 0) return;
    uint32_t pick = 0;
    if (nc <= 4) {
        pick = (1u << nc) - 1u;
    } else {
        int i0 = 0;
#pragma unroll
        for (int k = 1; k < 8; ++k)
            if (k < nc && cd[k] > cd[i0]) i0 = k;
        const f3 c0 = cp.get(i0);
        int i1 = -1;
        float dm = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k >= nc || k == i0) continue;
            const f3 dd = cp.get(k) - c0;
            const float d2 = dot(dd, dd);
            if (i1 < 0 || d2 > dm) { dm = d2; i1 = k; }
        }
        const f3 e01 = cp.get(i1) - c0;
        int i2 = -1, i3 = -1;
        float amax = 0.0f, amin = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k >= nc || k == i0 || k == i1) continue;
            const float a = dot(cross(e01, cp.get(k) - c0), nRef);
            if (i2 < 0 || a > amax) { amax = a; i2 = k; }
            if (i3 < 0 || a < amin) { amin = a; i3 = k; }
        }
        if (i3 == i2) {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (k < nc && k != i0 && k != i1 && k != i2) { i3 = k; break; }
        }
        pick = (1u << i0) | (1u << i1) | (1u << i2) | (1u << i3);
    }
    const uint32_t fbase = ((refB ? 1u : 0u) << 12) | ((uint32_t)ri << 8) | ((uint32_t)ij << 4);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (k >= nc) break;
        if (!((pick >> k) & 1u)) continue;
        if (refB) {
            man_add(m, cp.get(k), -nRef, cd[k], fbase | (uint32_t)cid[k]);
        } else {
            man_add(m, cp.get(k) + nRef * cd[k], nRef, cd[k], fbase | (uint32_t)cid[k]);
        }
    }
}

// Canonical-order dispatch (type(A) <= type(B)).  TESTS selects which tests are compiled
// in: COLLIDE_LIGHT (sphere-sphere, sphere-box, sphere-capsule, capsule-capsule),
// COLLIDE_HEAVY (box-box face clipping, box-capsule GJK, which need several times the
// registers and local arrays of the others), or both.
constexpr int COLLIDE_LIGHT = 1, COLLIDE_HEAVY = 2, COLLIDE_ALL = 3;
template <int TESTS = COLLIDE_ALL>
AXM_D int collide_canon(Man& m, const Shp& A, const Shp& B, float margin, bool& need_epa) {
    need_epa = false;
    const uint32_t ta = A.type, tb = B.type;
    if constexpr ((TESTS & COLLIDE_LIGHT) != 0) {
        if (ta == SHAPE_SPHERE && tb == SHAPE_SPHERE) { sphere_sphere(m, A.c, A.prm.x, B.c, B.prm.x, margin, 0u); return GEOM_OK; }
        if (ta == SHAPE_SPHERE && tb == SHAPE_BOX) { sphere_box(m, A.c, A.prm.x, B, margin); return GEOM_OK; }
        if (ta == SHAPE_SPHERE && tb == SHAPE_CAPSULE) {
            f3 p0, p1;
            capsule_segment(B, p0, p1);
            sphere_sphere(m, A.c, A.prm.x, closest_on_segment(A.c, p0, p1), B.prm.x, margin, 0u);
            return GEOM_OK;
        }
        if (ta == SHAPE_CAPSULE && tb == SHAPE_CAPSULE) { capsule_capsule(m, A, B, margin); return GEOM_OK; }
    }
    if constexpr ((TESTS & COLLIDE_HEAVY) != 0) {
        if (ta == SHAPE_BOX && tb == SHAPE_BOX) { box_box(m, A, B, margin); return GEOM_OK; }
        if (ta == SHAPE_BOX && tb == SHAPE_CAPSULE) return box_capsule_gjk(m, A, B, margin, need_epa);
    }
    return GEOM_OK;
}

// Converts a manifold computed as (B, A) back to (A, B).
AXM_D void flip_manifold(Man& m) {
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k >= m.np) break;
        m.pt[k].p = m.pt[k].p - m.pt[k].n * m.pt[k].depth;
        m.pt[k].n = -m.pt[k].n;
    }
}

AXM_D void collide_plane(Man& m, const Shp& A, f3 pn, float pd, float margin) {
    if (A.type == SHAPE_SPHERE) {
        sphere_plane(m, A.c, A.prm.x, pn, pd, margin, 0u);
    } else if (A.type == SHAPE_BOX) {
        box_plane(m, A, pn, pd, margin);
    } else if (A.type == SHAPE_CAPSULE) {
        f3 p0, p1;
        capsule_segment(A, p0, p1);
        sphere_plane(m, p0, A.prm.x, pn, pd, margin, 0u);
        sphere_plane(m, p1, A.prm.x, pn, pd, margin, 1u);
    }
}

}  // namespace axm
