// Per-contact and per-pair solver updates, expression for expression in the reference's order.
// Body state is reached through a BodyView whose pointers may be global memory or LDS; the
// arithmetic does not depend on where the state lives.
#pragma once
#include "vecmath.hpp"

namespace me {

struct BodyView {
    float *v, *w, *pv, *po;
    const float *x, *invM, *invIw;
};

struct ContactView {
    const float *pA, *pB, *n, *pen, *mu;
    float *jn, *jt1, *jt2, *jp;
};

// normal-derived tangent (contact_engine_gpu_fixed_frame.py:36-40)
__device__ __forceinline__ f3 fixed_t1(f3 nrm) {
    f3 aa = mk(1.0f, 0.0f, 0.0f);
    if (fabsf(nrm.x) >= 0.9f) aa = mk(0.0f, 1.0f, 0.0f);
    return normalize(aa - nrm * dot(aa, nrm));
}

// apply_warm of one pair (contact_engine_gpu_applied_warm.py:7-54); body a is BA[a], body b is BB[b]
// (b < 0: ground)
__device__ __forceinline__ void apply_warm_pair(const BodyView &BA, const BodyView &BB, const ContactView &Cv,
                                                int a, int b, int st, int cnt) {
    const BodyView &B = BA;
    f3 va = ld3(B.v, a), wa = ld3(B.w, a);
    f3 vb = mk(0.0f, 0.0f, 0.0f), wb = mk(0.0f, 0.0f, 0.0f);
    if (b >= 0) { vb = ld3(BB.v, b); wb = ld3(BB.w, b); }
    m3 IA = ldm(B.invIw, a), IB;
    if (b >= 0) IB = ldm(BB.invIw, b);
    float iMA = B.invM[a], iMB = b >= 0 ? BB.invM[b] : 0.0f;
    f3 xa = ld3(B.x, a), xb = b >= 0 ? ld3(BB.x, b) : mk(0.0f, 0.0f, 0.0f);
    for (int k = 0; k < cnt; ++k) {
        int c = st + k;
        f3 n = ld3(Cv.n, c);
        f3 t1 = cfixed_t1(n);
        f3 t2 = ccross(n, t1);
        f3 imp = cmadd(Cv.jt2[c], t2, cmadd(Cv.jn[c], n, Cv.jt1[c] * t1));
        if (iMA > 0.0f) {
            va = cmadd(iMA, imp, va);
            wa = wa + cmv(IA, ccross(ld3(Cv.pA, c) - xa, imp));
        }
        if (b >= 0 && iMB > 0.0f) {
            vb = cmadd(-iMB, imp, vb);
            wb = wb - cmv(IB, ccross(ld3(Cv.pB, c) - xb, imp));
        }
    }
    if (iMA > 0.0f) { st3(B.v, a, va); st3(B.w, a, wa); }
    if (b >= 0 && iMB > 0.0f) { st3(BB.v, b, vb); st3(BB.w, b, wb); }
}

// pair-chunk velocity sweep. FIXED: normal-derived frame (k_fixed_frame,
// contact_engine_gpu_fixed_frame.py:7-77 / unfused_solve.py:117-220); otherwise the slip frame of
// k_chunk_vel_r (contact_engine_gpu_colored.py:413-490)
template <bool FIXED>
__device__ __forceinline__ void chunk_vel(const BodyView &B, const BodyView &BB, const ContactView &Cv, int bi,
                                          int bj, int st, int cnt, float t2d) {
    float iMA = B.invM[bi];
    m3 IA = ldm(B.invIw, bi), IB = IA;
    f3 vA = ld3(B.v, bi), wA = ld3(B.w, bi);
    float iMB = 0.0f;
    f3 vB = mk(0.0f, 0.0f, 0.0f), wB = mk(0.0f, 0.0f, 0.0f);
    f3 xA = ld3(B.x, bi), xB = mk(0.0f, 0.0f, 0.0f);
    if (bj >= 0) { iMB = BB.invM[bj]; IB = ldm(BB.invIw, bj); vB = ld3(BB.v, bj); wB = ld3(BB.w, bj); xB = ld3(BB.x, bj); }
    for (int k = 0; k < cnt; ++k) {
        int c = st + k;
        f3 nrm = ld3(Cv.n, c), rA = ld3(Cv.pA, c) - xA;
        f3 va = vA + cross(wA, rA);
        f3 vb = mk(0.0f, 0.0f, 0.0f), rB = mk(0.0f, 0.0f, 0.0f);
        if (bj >= 0) { rB = ld3(Cv.pB, c) - xB; vb = vB + cross(wB, rB); }
        f3 vrel = va - vb;
        float vn = dot(vrel, nrm);
        f3 t1;
        if (FIXED) {
            t1 = fixed_t1(nrm);
        } else {
            f3 vt = vrel - nrm * vn;
            float mt = length(vt);
            t1 = vt / mt;
            if (mt <= 5e-3f) t1 = fixed_t1(nrm);
        }
        f3 t2 = cross(nrm, t1);
        f3 rnA = cross(rA, nrm);
        float meff = iMA + dot(rnA, mv(IA, rnA));
        if (bj >= 0) {
            f3 rnB = cross(rB, nrm);
            meff = meff + iMB + dot(rnB, mv(IB, rnB));
        }
        if (!(meff >= 1e-12f)) continue;
        float jo = Cv.jn[c];
        float dj = -vn / meff;
        float nw = wmax(0.0f, jo + dj);
        dj = nw - jo;
        Cv.jn[c] = nw;
        f3 Jn = nrm * dj;
        if (iMA > 0.0f) { vA = vA + Jn * iMA; wA = wA + mv(IA, cross(rA, Jn)); }
        if (bj >= 0 && iMB > 0.0f) { vB = vB - Jn * iMB; wB = wB - mv(IB, cross(rB, Jn)); }
        va = vA + cross(wA, rA);
        vb = mk(0.0f, 0.0f, 0.0f);
        if (bj >= 0) vb = vB + cross(wB, rB);
        vrel = va - vb;
        f3 rtA1 = cross(rA, t1);
        float meft1 = iMA + dot(rtA1, mv(IA, rtA1));
        f3 rtA2 = cross(rA, t2);
        float meft2 = iMA + dot(rtA2, mv(IA, rtA2));
        if (bj >= 0) {
            f3 rtB1 = cross(rB, t1);
            meft1 = meft1 + iMB + dot(rtB1, mv(IB, rtB1));
            f3 rtB2 = cross(rB, t2);
            meft2 = meft2 + iMB + dot(rtB2, mv(IB, rtB2));
        }
        if (meft1 >= 1e-12f && meft2 >= 1e-12f) {
            float o1 = Cv.jt1[c], o2 = Cv.jt2[c];
            float a1 = o1 - dot(vrel, t1) / meft1;
            float a2 = o2 - dot(vrel, t2) / meft2;
            float mag = sqrtf(a1 * a1 + a2 * a2);
            float lim = Cv.mu[c] * nw;
            if (mag > lim && mag > 1e-12f) { a1 = a1 * lim / mag; a2 = a2 * lim / mag; }
            Cv.jt1[c] = a1;
            Cv.jt2[c] = a2;
            f3 Jt = t1 * (a1 - o1) + t2 * (t2d * (a2 - o2));
            if (iMA > 0.0f) { vA = vA + Jt * iMA; wA = wA + mv(IA, cross(rA, Jt)); }
            if (bj >= 0 && iMB > 0.0f) { vB = vB - Jt * iMB; wB = wB - mv(IB, cross(rB, Jt)); }
        }
    }
    if (iMA > 0.0f) { st3(B.v, bi, vA); st3(B.w, bi, wA); }
    if (bj >= 0 && iMB > 0.0f) { st3(BB.v, bj, vB); st3(BB.w, bj, wB); }
}

// pair-chunk position sweep (k_chunk_pos_r, contact_engine_gpu_colored.py:511-554;
// unfused_position.py:7-65)
__device__ __forceinline__ void chunk_pos(const BodyView &B, const BodyView &BB, const ContactView &Cv, int bi,
                                          int bj, int st, int cnt, float dt) {
    float iMA = B.invM[bi];
    m3 IA = ldm(B.invIw, bi), IB = IA;
    f3 pvA = ld3(B.pv, bi), poA = ld3(B.po, bi);
    float iMB = 0.0f;
    f3 pvB = mk(0.0f, 0.0f, 0.0f), poB = mk(0.0f, 0.0f, 0.0f);
    f3 xA = ld3(B.x, bi), xB = mk(0.0f, 0.0f, 0.0f);
    if (bj >= 0) { iMB = BB.invM[bj]; IB = ldm(BB.invIw, bj); pvB = ld3(BB.pv, bj); poB = ld3(BB.po, bj); xB = ld3(BB.x, bj); }
    for (int k = 0; k < cnt; ++k) {
        int c = st + k;
        f3 nrm = ld3(Cv.n, c), rA = ld3(Cv.pA, c) - xA;
        f3 rel = pvA + cross(poA, rA);
        f3 rB = mk(0.0f, 0.0f, 0.0f);
        if (bj >= 0) { rB = ld3(Cv.pB, c) - xB; rel = rel - (pvB + cross(poB, rB)); }
        f3 rnA = cross(rA, nrm);
        float meff = iMA + dot(rnA, mv(IA, rnA));
        if (bj >= 0) {
            f3 rnB = cross(rB, nrm);
            meff = meff + iMB + dot(rnB, mv(IB, rnB));
        }
        if (!(meff >= 1e-12f)) continue;
        float bias = 0.2f * wmax(Cv.pen[c] - 1e-4f, 0.0f) / dt;
        float dj = (bias - dot(rel, nrm)) / meff;
        float jo = Cv.jp[c];
        float nw = wmax(0.0f, jo + dj);
        dj = nw - jo;
        Cv.jp[c] = nw;
        f3 Jn = nrm * dj;
        if (iMA > 0.0f) { pvA = pvA + Jn * iMA; poA = poA + mv(IA, cross(rA, Jn)); }
        if (bj >= 0 && iMB > 0.0f) { pvB = pvB - Jn * iMB; poB = poB - mv(IB, cross(rB, Jn)); }
    }
    if (iMA > 0.0f) { st3(B.pv, bi, pvA); st3(B.po, bi, poA); }
    if (bj >= 0 && iMB > 0.0f) { st3(BB.pv, bj, pvB); st3(BB.po, bj, poB); }
}

// integrate one body (k_integ, contact_engine_gpu.py:172-180)
__device__ __forceinline__ void integrate_body(float *x, float *q, f3 v, f3 w, f3 pv, f3 po, int i, float dt) {
    st3(x, i, cmadd(dt, v + pv, ld3(x, i)));
    f3 ww = w + po;
    f4 h;
    h.x = 0.5f * ww.x; h.y = 0.5f * ww.y; h.z = 0.5f * ww.z; h.w = 0.5f * 0.0f;
    f4 qi = ld4(q, i);
    f4 m = cqmul(h, qi);
    f4 qn;
    qn.x = fmaf(m.x, dt, qi.x); qn.y = fmaf(m.y, dt, qi.y); qn.z = fmaf(m.z, dt, qi.z); qn.w = fmaf(m.w, dt, qi.w);
    f4 r = cqnormalize(qn);
    q[4 * i] = r.x; q[4 * i + 1] = r.y; q[4 * i + 2] = r.z; q[4 * i + 3] = r.w;
}

}  // namespace me
