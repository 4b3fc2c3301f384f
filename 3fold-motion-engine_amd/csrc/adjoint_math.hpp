// Adjoint of the ColorCache contact step: the per-element forward updates written once over a
// scalar type S, and their hand-written reverses.
//
// The forward functions restate, expression for expression, the device updates of the solve
// (cc_warm / cc_vel<true> / cc_pos in tiled.hpp, the per-contact constants of k_t_prep, integrate_body,
// the world inverse inertia of k_body_pre): with S = float they round exactly like those, so a
// reverse that recomputes its forward intermediates takes the same branches (clamps, cone
// projection) as the step it differentiates. With S = double they are the reference the CPU checks
// differentiate by central differences.
//
// Notation: for y = f(x), the reverse adds xb += (df/dx)^T yb ("b" = adjoint).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace me {
namespace ad {

#define AD_HD __host__ __device__ __forceinline__

template <class S> struct V3 { S x, y, z; };
template <class S> struct Q4 { S x, y, z, w; };
template <class S> struct M3 { S m[3][3]; };

template <class S> AD_HD V3<S> v3(S x, S y, S z) { V3<S> r; r.x = x; r.y = y; r.z = z; return r; }
template <class S> AD_HD V3<S> zero3() { return v3<S>(S(0), S(0), S(0)); }
template <class S> AD_HD V3<S> operator+(V3<S> a, V3<S> b) { return v3<S>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class S> AD_HD V3<S> operator-(V3<S> a, V3<S> b) { return v3<S>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class S> AD_HD V3<S> operator-(V3<S> a) { return v3<S>(-a.x, -a.y, -a.z); }
template <class S> AD_HD V3<S> operator*(V3<S> a, S s) { return v3<S>(a.x * s, a.y * s, a.z * s); }
template <class S> AD_HD V3<S> operator*(S s, V3<S> a) { return v3<S>(s * a.x, s * a.y, s * a.z); }
template <class S> AD_HD V3<S> operator/(V3<S> a, S s) { return v3<S>(a.x / s, a.y / s, a.z / s); }
template <class S> AD_HD V3<S> &operator+=(V3<S> &a, V3<S> b) { a = a + b; return a; }
template <class S> AD_HD V3<S> &operator-=(V3<S> &a, V3<S> b) { a = a - b; return a; }
template <class S> AD_HD S dot(V3<S> a, V3<S> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class S> AD_HD V3<S> cross(V3<S> a, V3<S> b) {
    return v3<S>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class S> AD_HD V3<S> mv(const M3<S> &M, V3<S> b) {
    return v3<S>(M.m[0][0] * b.x + M.m[0][1] * b.y + M.m[0][2] * b.z,
                 M.m[1][0] * b.x + M.m[1][1] * b.y + M.m[1][2] * b.z,
                 M.m[2][0] * b.x + M.m[2][1] * b.y + M.m[2][2] * b.z);
}
template <class S> AD_HD V3<S> mtv(const M3<S> &M, V3<S> b) {  // M^T b
    return v3<S>(M.m[0][0] * b.x + M.m[1][0] * b.y + M.m[2][0] * b.z,
                 M.m[0][1] * b.x + M.m[1][1] * b.y + M.m[2][1] * b.z,
                 M.m[0][2] * b.x + M.m[1][2] * b.y + M.m[2][2] * b.z);
}
template <class S> AD_HD void add_outer(M3<S> &M, V3<S> a, V3<S> b, S s) {  // M += s * a b^T
    S av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M.m[i][j] += s * av[i] * bv[j];
}
template <class S> AD_HD M3<S> zero33() {
    M3<S> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = S(0);
    return r;
}
template <class S> AD_HD S sq(S a) { return a * a; }
template <class S> AD_HD S wmax(S a, S b) { return a > b ? a : b; }
AD_HD float ssqrt(float a) { return sqrtf(a); }
AD_HD double ssqrt(double a) { return sqrt(a); }

// reverse of c = cross(a, b): ab += cross(b, cb), bb += cross(cb, a)
template <class S> AD_HD void cross_rev(V3<S> a, V3<S> b, V3<S> cb, V3<S> &ab, V3<S> &bb) {
    ab += cross(b, cb);
    bb += cross(cb, a);
}
// reverse of y = M x: xb += M^T yb, Mb += yb x^T
template <class S> AD_HD void mv_rev(const M3<S> &M, V3<S> x, V3<S> yb, V3<S> &xb, M3<S> &Mb) {
    xb += mtv(M, yb);
    add_outer(Mb, yb, x, S(1));
}

// ---- per-contact constants (k_t_prep with the fixed frame) ----

template <class S> struct Con {
    V3<S> n, rA, rB, t1, t2;
    S meff, meft1, meft2, bias, mu;
};

// adjoint of Con, same layout (20 scalars)
template <class S> AD_HD Con<S> con_zero() {
    Con<S> c;
    c.n = c.rA = c.rB = c.t1 = c.t2 = zero3<S>();
    c.meff = c.meft1 = c.meft2 = c.bias = c.mu = S(0);
    return c;
}

// normal-derived tangent (fixed_t1): aa = ex unless |n.x| >= 0.9, t1 = normalize(aa - n (aa.n))
template <class S> AD_HD V3<S> fixed_t1(V3<S> n) {
    V3<S> aa = v3<S>(S(1), S(0), S(0));
    if ((n.x < S(0) ? -n.x : n.x) >= S(0.9f)) aa = v3<S>(S(0), S(1), S(0));
    V3<S> u = aa - n * dot(aa, n);
    S l = ssqrt(dot(u, u));
    if (l > S(0)) return u / l;
    return zero3<S>();
}
template <class S> AD_HD void fixed_t1_rev(V3<S> n, V3<S> tb, V3<S> &nb) {
    V3<S> aa = v3<S>(S(1), S(0), S(0));
    if ((n.x < S(0) ? -n.x : n.x) >= S(0.9f)) aa = v3<S>(S(0), S(1), S(0));
    S d = dot(aa, n);
    V3<S> u = aa - n * d;
    S l = ssqrt(dot(u, u));
    if (!(l > S(0))) return;
    V3<S> t = u / l;
    V3<S> ub = (tb - t * dot(t, tb)) / l;
    nb -= ub * d;
    S db = -dot(ub, n);
    nb += aa * db;
}

// quadratic form m = iM + r.(I r) with r = cross(rr, d)
template <class S> AD_HD S qform(V3<S> rr, V3<S> d, const M3<S> &I) {
    V3<S> r = cross(rr, d);
    return dot(r, mv(I, r));
}
template <class S> AD_HD void qform_rev(V3<S> rr, V3<S> d, const M3<S> &I, S mb, V3<S> &rrb, V3<S> &db, M3<S> &Ib) {
    V3<S> r = cross(rr, d);
    V3<S> y = mv(I, r);
    V3<S> rb = y * mb;
    V3<S> yb = r * mb;
    mv_rev(I, r, yb, rb, Ib);
    cross_rev(rr, d, rb, rrb, db);
}

// constants of one contact from the contact data and the bodies (k_t_prep, FIXED):
// pA/pB contact points, xA/xB centres, iMA/iMB inverse masses, IA/IB world inverse inertia
template <class S> AD_HD Con<S> prep_fwd(V3<S> n, V3<S> pA, V3<S> pB, V3<S> xA, V3<S> xB, bool hasB, S iMA, S iMB,
                                        const M3<S> &IA, const M3<S> &IB, S pen, S mu, S dt) {
    Con<S> c;
    c.n = n;
    c.rA = pA - xA;
    c.rB = hasB ? pB - xB : zero3<S>();
    c.meff = iMA + qform(c.rA, n, IA);
    if (hasB) c.meff = c.meff + iMB + qform(c.rB, n, IB);
    c.t1 = fixed_t1(n);
    c.t2 = cross(n, c.t1);
    S m1 = iMA + qform(c.rA, c.t1, IA);
    S m2 = iMA + qform(c.rA, c.t2, IA);
    if (hasB) {
        m1 = m1 + iMB + qform(c.rB, c.t1, IB);
        m2 = m2 + iMB + qform(c.rB, c.t2, IB);
    }
    c.meft1 = m1;
    c.meft2 = m2;
    c.bias = S(0.2f) * wmax(pen - S(1e-4f), S(0)) / dt;
    c.mu = mu;
    return c;
}

// reverse of prep_fwd: cb (adjoint of the constants) -> adjoints of n, pA, pB, xA, xB, IA, IB, pen, mu
template <class S> AD_HD void prep_rev(V3<S> n, V3<S> pA, V3<S> pB, V3<S> xA, V3<S> xB, bool hasB, const M3<S> &IA,
                                      const M3<S> &IB, S pen, S dt, const Con<S> &cb, V3<S> &nb, V3<S> &pAb,
                                      V3<S> &pBb, V3<S> &xAb, V3<S> &xBb, M3<S> &IAb, M3<S> &IBb, S &penb, S &mub) {
    V3<S> rA = pA - xA, rB = hasB ? pB - xB : zero3<S>();
    V3<S> t1 = fixed_t1(n), t2 = cross(n, t1);
    V3<S> rAb = cb.rA, rBb = cb.rB, t1b = cb.t1, t2b = cb.t2;
    nb += cb.n;
    qform_rev(rA, t1, IA, cb.meft1, rAb, t1b, IAb);
    qform_rev(rA, t2, IA, cb.meft2, rAb, t2b, IAb);
    if (hasB) {
        qform_rev(rB, t1, IB, cb.meft1, rBb, t1b, IBb);
        qform_rev(rB, t2, IB, cb.meft2, rBb, t2b, IBb);
    }
    qform_rev(rA, n, IA, cb.meff, rAb, nb, IAb);
    if (hasB) qform_rev(rB, n, IB, cb.meff, rBb, nb, IBb);
    cross_rev(n, t1, t2b, nb, t1b);
    fixed_t1_rev(n, t1b, nb);
    pAb += rAb;
    xAb -= rAb;
    if (hasB) {
        pBb += rBb;
        xBb -= rBb;
    }
    if (pen - S(1e-4f) > S(0)) penb += cb.bias * S(0.2f) / dt;
    mub += cb.mu;
}

// ---- body state of one pair in the solve ----

template <class S> struct PairBodies {
    V3<S> vA, wA, vB, wB;  // (v, w) in velocity passes, (pv, po) in position passes
};

struct PairFlags {
    bool hasB, dynA, dynB;
};

// one contact of apply_warm (cc_warm) with impulses j = (jn, jt1, jt2)
template <class S> AD_HD void warm_fwd(PairBodies<S> &B, const Con<S> &C, S jn, S jt1, S jt2, PairFlags f, S iMA,
                                      S iMB, const M3<S> &IA, const M3<S> &IB) {
    V3<S> imp = (jn * C.n + jt1 * C.t1) + jt2 * C.t2;
    if (f.dynA) {
        B.vA = B.vA + iMA * imp;
        B.wA = B.wA + mv(IA, cross(C.rA, imp));
    }
    if (f.hasB && f.dynB) {
        B.vB = B.vB - iMB * imp;
        B.wB = B.wB - mv(IB, cross(C.rB, imp));
    }
}
// the body adjoint passes through unchanged (the update adds a term independent of the state)
template <class S> AD_HD void warm_rev(const PairBodies<S> &Bb, const Con<S> &C, S jn, S jt1, S jt2, PairFlags f,
                                      S iMA, S iMB, const M3<S> &IA, const M3<S> &IB, S &jnb, S &jt1b, S &jt2b,
                                      Con<S> &cb, M3<S> &IAb, M3<S> &IBb) {
    V3<S> imp = (jn * C.n + jt1 * C.t1) + jt2 * C.t2;
    V3<S> impb = zero3<S>();
    if (f.dynA) {
        impb += Bb.vA * iMA;
        V3<S> k = cross(C.rA, imp);
        V3<S> kb = zero3<S>();
        mv_rev(IA, k, Bb.wA, kb, IAb);
        cross_rev(C.rA, imp, kb, cb.rA, impb);
    }
    if (f.hasB && f.dynB) {
        impb -= Bb.vB * iMB;
        V3<S> k = cross(C.rB, imp);
        V3<S> kb = zero3<S>();
        mv_rev(IB, k, -Bb.wB, kb, IBb);
        cross_rev(C.rB, imp, kb, cb.rB, impb);
    }
    jnb += dot(impb, C.n);
    cb.n += impb * jn;
    jt1b += dot(impb, C.t1);
    cb.t1 += impb * jt1;
    jt2b += dot(impb, C.t2);
    cb.t2 += impb * jt2;
}

// the impulse J applied to the pair: vA += J iMA, wA += IA (rA x J), vB -= J iMB, wB -= IB (rB x J)
template <class S> AD_HD void apply_fwd(PairBodies<S> &B, V3<S> J, V3<S> rA, V3<S> rB, PairFlags f, S iMA, S iMB,
                                       const M3<S> &IA, const M3<S> &IB) {
    if (f.dynA) { B.vA = B.vA + J * iMA; B.wA = B.wA + mv(IA, cross(rA, J)); }
    if (f.hasB && f.dynB) { B.vB = B.vB - J * iMB; B.wB = B.wB - mv(IB, cross(rB, J)); }
}
template <class S> AD_HD void apply_rev(const PairBodies<S> &Bb, V3<S> J, V3<S> rA, V3<S> rB, PairFlags f, S iMA,
                                       S iMB, const M3<S> &IA, const M3<S> &IB, V3<S> &Jb, V3<S> &rAb, V3<S> &rBb,
                                       M3<S> &IAb, M3<S> &IBb) {
    if (f.dynA) {
        Jb += Bb.vA * iMA;
        V3<S> k = cross(rA, J), kb = zero3<S>();
        mv_rev(IA, k, Bb.wA, kb, IAb);
        cross_rev(rA, J, kb, rAb, Jb);
    }
    if (f.hasB && f.dynB) {
        Jb -= Bb.vB * iMB;
        V3<S> k = cross(rB, J), kb = zero3<S>();
        mv_rev(IB, k, -Bb.wB, kb, IBb);
        cross_rev(rB, J, kb, rBb, Jb);
    }
}

// relative velocity vrel = (vA + wA x rA) - (vB + wB x rB)
template <class S> AD_HD V3<S> vrel_fwd(const PairBodies<S> &B, V3<S> rA, V3<S> rB, bool hasB) {
    V3<S> va = B.vA + cross(B.wA, rA);
    V3<S> vb = zero3<S>();
    if (hasB) vb = B.vB + cross(B.wB, rB);
    return va - vb;
}
template <class S> AD_HD void vrel_rev(const PairBodies<S> &B, V3<S> rA, V3<S> rB, bool hasB, V3<S> vrelb,
                                      PairBodies<S> &Bb, V3<S> &rAb, V3<S> &rBb) {
    Bb.vA += vrelb;
    cross_rev(B.wA, rA, vrelb, Bb.wA, rAb);
    if (hasB) {
        Bb.vB -= vrelb;
        cross_rev(B.wB, rB, -vrelb, Bb.wB, rBb);
    }
}

// one contact of the fixed-frame velocity sweep (cc_vel<true>, t2d = 1): updates B and (jn, jt1, jt2)
template <class S> AD_HD void vel_fwd(PairBodies<S> &B, const Con<S> &C, S &jn, S &jt1, S &jt2, PairFlags f, S iMA,
                                     S iMB, const M3<S> &IA, const M3<S> &IB) {
    V3<S> vrel = vrel_fwd(B, C.rA, C.rB, f.hasB);
    S vn = dot(vrel, C.n);
    if (!(C.meff >= S(1e-12f))) return;
    S jo = jn;
    S dj = -vn / C.meff;
    S nw = wmax(S(0), jo + dj);
    dj = nw - jo;
    jn = nw;
    V3<S> Jn = C.n * dj;
    apply_fwd(B, Jn, C.rA, C.rB, f, iMA, iMB, IA, IB);
    vrel = vrel_fwd(B, C.rA, C.rB, f.hasB);
    if (C.meft1 >= S(1e-12f) && C.meft2 >= S(1e-12f)) {
        S o1 = jt1, o2 = jt2;
        S a1 = o1 - dot(vrel, C.t1) / C.meft1;
        S a2 = o2 - dot(vrel, C.t2) / C.meft2;
        S mag = ssqrt(a1 * a1 + a2 * a2);
        S lim = C.mu * nw;
        if (mag > lim && mag > S(1e-12f)) { a1 = a1 * lim / mag; a2 = a2 * lim / mag; }
        jt1 = a1;
        jt2 = a2;
        V3<S> Jt = C.t1 * (a1 - o1) + C.t2 * (S(1) * (a2 - o2));
        apply_fwd(B, Jt, C.rA, C.rB, f, iMA, iMB, IA, IB);
    }
}

// reverse of vel_fwd at the input state B0 / impulses (jn0, jt10, jt20). On entry Bb and (jnb, jt1b,
// jt2b) hold the adjoints of the outputs; on exit, of the inputs. cb, IAb, IBb accumulate.
template <class S> AD_HD void vel_rev(const PairBodies<S> &B0, const Con<S> &C, S jn0, S jt10, S jt20, PairFlags f,
                                     S iMA, S iMB, const M3<S> &IA, const M3<S> &IB, PairBodies<S> &Bb, S &jnb,
                                     S &jt1b, S &jt2b, Con<S> &cb, M3<S> &IAb, M3<S> &IBb) {
    if (!(C.meff >= S(1e-12f))) return;
    // forward intermediates
    PairBodies<S> B = B0;
    V3<S> vrel0 = vrel_fwd(B, C.rA, C.rB, f.hasB);
    S vn = dot(vrel0, C.n);
    S jo = jn0;
    S dj0 = -vn / C.meff;
    S s = jo + dj0;
    S nw = wmax(S(0), s);
    S dj = nw - jo;
    V3<S> Jn = C.n * dj;
    apply_fwd(B, Jn, C.rA, C.rB, f, iMA, iMB, IA, IB);
    PairBodies<S> B1 = B;
    V3<S> vrel1 = vrel_fwd(B1, C.rA, C.rB, f.hasB);
    S nwb = jnb;  // jn (out) = nw
    if (C.meft1 >= S(1e-12f) && C.meft2 >= S(1e-12f)) {
        S o1 = jt10, o2 = jt20;
        S u1 = dot(vrel1, C.t1), u2 = dot(vrel1, C.t2);
        S a1 = o1 - u1 / C.meft1;
        S a2 = o2 - u2 / C.meft2;
        S mag = ssqrt(a1 * a1 + a2 * a2);
        S lim = C.mu * nw;
        bool clamp = mag > lim && mag > S(1e-12f);
        S b1 = a1, b2 = a2;
        if (clamp) { b1 = a1 * lim / mag; b2 = a2 * lim / mag; }
        V3<S> Jt = C.t1 * (b1 - o1) + C.t2 * (S(1) * (b2 - o2));
        // Jt applied to B1
        V3<S> Jtb = zero3<S>();
        apply_rev(Bb, Jt, C.rA, C.rB, f, iMA, iMB, IA, IB, Jtb, cb.rA, cb.rB, IAb, IBb);
        cb.t1 += Jtb * (b1 - o1);
        cb.t2 += Jtb * (S(1) * (b2 - o2));
        S d1b = dot(Jtb, C.t1), d2b = S(1) * dot(Jtb, C.t2);
        S b1b = jt1b + d1b, b2b = jt2b + d2b;
        S o1b = -d1b, o2b = -d2b;
        S a1b, a2b;
        if (clamp) {
            S limb = (b1b * a1 + b2b * a2) / mag;
            S magb = -(b1b * a1 + b2b * a2) * lim / (mag * mag);
            a1b = b1b * lim / mag + magb * a1 / mag;
            a2b = b2b * lim / mag + magb * a2 / mag;
            cb.mu += limb * nw;
            nwb += limb * C.mu;
        } else {
            a1b = b1b;
            a2b = b2b;
        }
        o1b += a1b;
        o2b += a2b;
        S u1b = -a1b / C.meft1, u2b = -a2b / C.meft2;
        cb.meft1 += a1b * u1 / (C.meft1 * C.meft1);
        cb.meft2 += a2b * u2 / (C.meft2 * C.meft2);
        V3<S> vrel1b = C.t1 * u1b + C.t2 * u2b;
        cb.t1 += vrel1 * u1b;
        cb.t2 += vrel1 * u2b;
        vrel_rev(B1, C.rA, C.rB, f.hasB, vrel1b, Bb, cb.rA, cb.rB);
        jt1b = o1b;
        jt2b = o2b;
    }
    // normal step
    V3<S> Jnb = zero3<S>();
    apply_rev(Bb, Jn, C.rA, C.rB, f, iMA, iMB, IA, IB, Jnb, cb.rA, cb.rB, IAb, IBb);
    cb.n += Jnb * dj;
    S djb = dot(Jnb, C.n);
    nwb += djb;
    S job = -djb;
    S sb = (S(0) > s) ? S(0) : nwb;
    job += sb;
    S dj0b = sb;
    S vnb = -dj0b / C.meff;
    cb.meff += -dj0b * dj0 / C.meff;
    jnb = job;
    V3<S> vrel0b = C.n * vnb;
    cb.n += vrel0 * vnb;
    vrel_rev(B0, C.rA, C.rB, f.hasB, vrel0b, Bb, cb.rA, cb.rB);
}

// one contact of the position sweep (cc_pos): pseudo-velocities in B, accumulated jp
template <class S> AD_HD void pos_fwd(PairBodies<S> &B, const Con<S> &C, S &jp, PairFlags f, S iMA, S iMB,
                                     const M3<S> &IA, const M3<S> &IB) {
    V3<S> rel = B.vA + cross(B.wA, C.rA);
    if (f.hasB) rel = rel - (B.vB + cross(B.wB, C.rB));
    if (!(C.meff >= S(1e-12f))) return;
    S dj = (C.bias - dot(rel, C.n)) / C.meff;
    S jo = jp;
    S nw = wmax(S(0), jo + dj);
    dj = nw - jo;
    jp = nw;
    apply_fwd(B, C.n * dj, C.rA, C.rB, f, iMA, iMB, IA, IB);
}
template <class S> AD_HD void pos_rev(const PairBodies<S> &B0, const Con<S> &C, S jp0, PairFlags f, S iMA, S iMB,
                                     const M3<S> &IA, const M3<S> &IB, PairBodies<S> &Bb, S &jpb, Con<S> &cb,
                                     M3<S> &IAb, M3<S> &IBb) {
    if (!(C.meff >= S(1e-12f))) return;
    V3<S> rel = vrel_fwd(B0, C.rA, C.rB, f.hasB);
    S dj0 = (C.bias - dot(rel, C.n)) / C.meff;
    S jo = jp0;
    S s = jo + dj0;
    S nw = wmax(S(0), s);
    S dj = nw - jo;
    V3<S> Jn = C.n * dj;
    V3<S> Jnb = zero3<S>();
    apply_rev(Bb, Jn, C.rA, C.rB, f, iMA, iMB, IA, IB, Jnb, cb.rA, cb.rB, IAb, IBb);
    cb.n += Jnb * dj;
    S djb = dot(Jnb, C.n);
    S nwb = jpb + djb;
    S job = -djb;
    S sb = (S(0) > s) ? S(0) : nwb;
    job += sb;
    S dj0b = sb;
    cb.bias += dj0b / C.meff;
    S relnb = -dj0b / C.meff;
    cb.meff += -dj0b * dj0 / C.meff;
    jpb = job;
    V3<S> relb = C.n * relnb;
    cb.n += rel * relnb;
    vrel_rev(B0, C.rA, C.rB, f.hasB, relb, Bb, cb.rA, cb.rB);
}

// ---- quaternions ----

template <class S> AD_HD Q4<S> q4(S x, S y, S z, S w) { Q4<S> r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
template <class S> AD_HD Q4<S> qmul(Q4<S> a, Q4<S> b) {
    return q4<S>(a.w * b.x + b.w * a.x + a.y * b.z - b.y * a.z, a.w * b.y + b.w * a.y + a.z * b.x - b.z * a.x,
                 a.w * b.z + b.w * a.z + a.x * b.y - b.x * a.y, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
template <class S> AD_HD void qmul_rev(Q4<S> a, Q4<S> b, Q4<S> r, Q4<S> &ab, Q4<S> &bb) {
    ab.x += r.x * b.w - r.y * b.z + r.z * b.y - r.w * b.x;
    ab.y += r.x * b.z + r.y * b.w - r.z * b.x - r.w * b.y;
    ab.z += -r.x * b.y + r.y * b.x + r.z * b.w - r.w * b.z;
    ab.w += r.x * b.x + r.y * b.y + r.z * b.z + r.w * b.w;
    bb.x += r.x * a.w + r.y * a.z - r.z * a.y - r.w * a.x;
    bb.y += -r.x * a.z + r.y * a.w + r.z * a.x - r.w * a.y;
    bb.z += r.x * a.y - r.y * a.x + r.z * a.w - r.w * a.z;
    bb.w += r.x * a.x + r.y * a.y + r.z * a.z + r.w * a.w;
}

// quat_rotate(q, e_k) columns: R(q) as quat_to_matrix
template <class S> AD_HD V3<S> quat_rotate(Q4<S> q, V3<S> x) {
    S c = S(2) * q.w * q.w - S(1);
    S d = S(2) * (q.x * x.x + q.y * x.y + q.z * x.z);
    V3<S> qv = v3<S>(q.x, q.y, q.z);
    return (x * c + qv * d) + (cross(qv, x) * q.w) * S(2);
}
template <class S> AD_HD M3<S> quat_to_matrix(Q4<S> q) {
    M3<S> R;
    for (int k = 0; k < 3; ++k) {
        V3<S> e = v3<S>(S(k == 0), S(k == 1), S(k == 2));
        V3<S> c = quat_rotate(q, e);
        R.m[0][k] = c.x; R.m[1][k] = c.y; R.m[2][k] = c.z;
    }
    return R;
}
// reverse of R = quat_to_matrix(q)
template <class S> AD_HD void quat_to_matrix_rev(Q4<S> q, const M3<S> &Rb, Q4<S> &qb) {
    V3<S> qv = v3<S>(q.x, q.y, q.z);
    for (int k = 0; k < 3; ++k) {
        V3<S> e = v3<S>(S(k == 0), S(k == 1), S(k == 2));
        V3<S> cb = v3<S>(Rb.m[0][k], Rb.m[1][k], Rb.m[2][k]);
        // c = e (2w^2 - 1) + qv (2 qv.e) + (qv x e) 2w
        S qk = k == 0 ? q.x : (k == 1 ? q.y : q.z);
        qb.w += dot(cb, e) * S(4) * q.w + dot(cb, cross(qv, e)) * S(2);
        V3<S> qvb = cb * (S(2) * qk) + e * (S(2) * dot(qv, cb)) + cross(e, cb * (S(2) * q.w));
        qb.x += qvb.x; qb.y += qvb.y; qb.z += qvb.z;
    }
}

// world inverse inertia W = (R B) R^T; reverse: Rb = W_b R B^T + W_b^T R B
template <class S> AD_HD M3<S> mm(const M3<S> &A, const M3<S> &B) {
    M3<S> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            S s = S(0);
            for (int k = 0; k < 3; ++k) s += A.m[i][k] * B.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}
template <class S> AD_HD M3<S> tr(const M3<S> &A) {
    M3<S> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = A.m[j][i];
    return r;
}
template <class S> AD_HD void inv_inertia_rev(const M3<S> &R, const M3<S> &B, const M3<S> &Wb, M3<S> &Rb) {
    M3<S> a = mm(mm(Wb, R), tr(B)), b = mm(mm(tr(Wb), R), B);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Rb.m[i][j] += a.m[i][j] + b.m[i][j];
}

// integrate one dynamic body (integrate_body): x1 = x + (v + pv) dt, q1 = normalize(q + (h q) dt),
// h = (0.5 (w + po), 0)
template <class S> AD_HD void integrate_fwd(V3<S> &x, Q4<S> &q, V3<S> v, V3<S> w, V3<S> pv, V3<S> po, S dt) {
    x = x + (v + pv) * dt;
    V3<S> ww = w + po;
    Q4<S> h = q4<S>(S(0.5f) * ww.x, S(0.5f) * ww.y, S(0.5f) * ww.z, S(0.5f) * S(0));
    Q4<S> m = qmul(h, q);
    Q4<S> qn = q4<S>(q.x + m.x * dt, q.y + m.y * dt, q.z + m.z * dt, q.w + m.w * dt);
    S l = ssqrt(qn.x * qn.x + qn.y * qn.y + qn.z * qn.z + qn.w * qn.w);
    if (l > S(0)) {
        S inv = S(1) / l;
        q = q4<S>(qn.x * inv, qn.y * inv, qn.z * inv, qn.w * inv);
    } else {
        q = q4<S>(S(0), S(0), S(0), S(1));
    }
}
// reverse at the input (x, q, v, w, pv, po): on entry xb, qb hold the adjoints of x1, q1; on exit of
// x, q; vb, wb, pvb, pob accumulate
template <class S> AD_HD void integrate_rev(Q4<S> q, V3<S> w, V3<S> po, S dt, V3<S> &xb, Q4<S> &qb, V3<S> &vb,
                                           V3<S> &wb, V3<S> &pvb, V3<S> &pob) {
    vb += xb * dt;
    pvb += xb * dt;
    V3<S> ww = w + po;
    Q4<S> h = q4<S>(S(0.5f) * ww.x, S(0.5f) * ww.y, S(0.5f) * ww.z, S(0));
    Q4<S> m = qmul(h, q);
    Q4<S> qn = q4<S>(q.x + m.x * dt, q.y + m.y * dt, q.z + m.z * dt, q.w + m.w * dt);
    S l = ssqrt(qn.x * qn.x + qn.y * qn.y + qn.z * qn.z + qn.w * qn.w);
    Q4<S> qnb = q4<S>(S(0), S(0), S(0), S(0));
    if (l > S(0)) {
        Q4<S> u = q4<S>(qn.x / l, qn.y / l, qn.z / l, qn.w / l);
        S d = u.x * qb.x + u.y * qb.y + u.z * qb.z + u.w * qb.w;
        qnb = q4<S>((qb.x - u.x * d) / l, (qb.y - u.y * d) / l, (qb.z - u.z * d) / l, (qb.w - u.w * d) / l);
    }
    Q4<S> qout = qnb;  // q contributes directly (qn = q + m dt)
    Q4<S> mb = q4<S>(qnb.x * dt, qnb.y * dt, qnb.z * dt, qnb.w * dt);
    Q4<S> hb = q4<S>(S(0), S(0), S(0), S(0));
    qmul_rev(h, q, mb, hb, qout);
    V3<S> wwb = v3<S>(S(0.5f) * hb.x, S(0.5f) * hb.y, S(0.5f) * hb.z);
    wb += wwb;
    pob += wwb;
    qb = qout;
}

}  // namespace ad
}  // namespace me
