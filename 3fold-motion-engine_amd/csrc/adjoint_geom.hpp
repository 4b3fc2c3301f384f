// Derivatives of the box narrow phase for the adjoint: the narrow phase of geometry.hpp written
// over a scalar type S, evaluated with forward-mode dual numbers (value + K partials) to get the
// Jacobian of every kept contact's point, normal and penetration with respect to the two bodies'
// centres and orientation quaternions. Every discrete decision (separating axis, reference face,
// clipping side, duplicate removal, depth filter, reduction to 4 points, feature order) is taken on
// the values, which are computed by the same float operations in the same order as box_box /
// box_ground, so the dual evaluation follows exactly the branch the step took.
#pragma once
#include "adjoint_math.hpp"

namespace me {
namespace ad {

template <class T, int K> struct Dual {
    T v;
    T d[K];
    AD_HD Dual() {}
    AD_HD Dual(T x) : v(x) {
        for (int k = 0; k < K; ++k) d[k] = T(0);
    }
};
template <class T, int K> AD_HD Dual<T, K> operator+(const Dual<T, K> &a, const Dual<T, K> &b) {
    Dual<T, K> r; r.v = a.v + b.v;
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <class T, int K> AD_HD Dual<T, K> operator-(const Dual<T, K> &a, const Dual<T, K> &b) {
    Dual<T, K> r; r.v = a.v - b.v;
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <class T, int K> AD_HD Dual<T, K> operator-(const Dual<T, K> &a) {
    Dual<T, K> r; r.v = -a.v;
    for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
    return r;
}
template <class T, int K> AD_HD Dual<T, K> operator*(const Dual<T, K> &a, const Dual<T, K> &b) {
    Dual<T, K> r; r.v = a.v * b.v;
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <class T, int K> AD_HD Dual<T, K> operator/(const Dual<T, K> &a, const Dual<T, K> &b) {
    Dual<T, K> r; r.v = a.v / b.v;
    for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r;
}
template <class T, int K> AD_HD Dual<T, K> &operator+=(Dual<T, K> &a, const Dual<T, K> &b) { a = a + b; return a; }
template <class T, int K> AD_HD bool operator<(const Dual<T, K> &a, const Dual<T, K> &b) { return a.v < b.v; }
template <class T, int K> AD_HD bool operator>(const Dual<T, K> &a, const Dual<T, K> &b) { return a.v > b.v; }
template <class T, int K> AD_HD bool operator<=(const Dual<T, K> &a, const Dual<T, K> &b) { return a.v <= b.v; }
template <class T, int K> AD_HD bool operator>=(const Dual<T, K> &a, const Dual<T, K> &b) { return a.v >= b.v; }
template <class T, int K> AD_HD Dual<T, K> ssqrt(const Dual<T, K> &a) {
    Dual<T, K> r; r.v = ssqrt(a.v);
    for (int k = 0; k < K; ++k) r.d[k] = r.v > T(0) ? a.d[k] / (T(2) * r.v) : T(0);
    return r;
}

AD_HD float val(float a) { return a; }
AD_HD double val(double a) { return a; }
template <class T, int K> AD_HD T val(const Dual<T, K> &a) { return a.v; }
AD_HD float sabs(float a) { return fabsf(a); }
AD_HD double sabs(double a) { return fabs(a); }
template <class T, int K> AD_HD Dual<T, K> sabs(const Dual<T, K> &a) { return a.v < T(0) ? -a : a; }

template <class S> AD_HD V3<S> colm(const M3<S> &R, int i) { return v3<S>(R.m[0][i], R.m[1][i], R.m[2][i]); }
template <class S> AD_HD S hget(V3<S> h, int i) { return i == 0 ? h.x : (i == 1 ? h.y : h.z); }

template <class S> AD_HD V3<S> box_vertex(V3<S> c, const M3<S> &R, V3<S> h, int k) {
    V3<S> l = v3<S>((k & 1) ? h.x : -h.x, (k & 2) ? h.y : -h.y, (k & 4) ? h.z : -h.z);
    return c + mv(R, l);
}

template <class S> AD_HD void face_poly(V3<S> c, const M3<S> &R, V3<S> h, int a, float s, V3<S> out[4]) {
    int o0 = a == 0 ? 1 : 0, o1 = a == 2 ? 1 : 2;
    const float S0[4] = {-1.0f, -1.0f, 1.0f, 1.0f};
    const float S1[4] = {-1.0f, 1.0f, 1.0f, -1.0f};
    for (int k = 0; k < 4; ++k) {
        S l[3];
        l[a] = S(s) * hget(h, a);
        l[o0] = S(S0[k]) * hget(h, o0);
        l[o1] = S(S1[k]) * hget(h, o1);
        out[k] = c + mv(R, v3<S>(l[0], l[1], l[2]));
    }
}

#define AD_MAXPOLY 16

template <class S> struct PolyG {
    V3<S> p[AD_MAXPOLY];
    int code[AD_MAXPOLY];
    int edge[AD_MAXPOLY];
    int n;
};

template <class S> AD_HD void clip_poly(const PolyG<S> &in, V3<S> sn, S d, int plane, PolyG<S> &out) {
    const float eps = 1e-6f;
    out.n = 0;
    int m = in.n;
    for (int i = 0; i < m; ++i) {
        V3<S> a = in.p[i], b = in.p[(i + 1) % m];
        S da = dot(sn, a) - d;
        S db = dot(sn, b) - d;
        int E = in.edge[i];
        if (val(da) <= eps && out.n < AD_MAXPOLY) {
            out.p[out.n] = a;
            out.code[out.n] = in.code[i];
            out.edge[out.n] = (val(db) > eps && !(val(da) < -eps)) ? 4 + plane : E;
            out.n++;
        }
        bool enter = val(da) > eps && val(db) < -eps;
        bool exitp = val(da) < -eps && val(db) > eps;
        if ((enter || exitp) && out.n < AD_MAXPOLY) {
            S t = da / (da - db);
            out.p[out.n] = a + t * (b - a);
            out.code[out.n] = E < 4 ? 4 + 4 * E + plane : 20 + 4 * (E - 4) + plane;
            out.edge[out.n] = exitp ? 4 + plane : E;
            out.n++;
        }
    }
}

// reduce4 on values (same selection and order as geometry.hpp)
template <class S> AD_HD int reduce4_idx(int n, const V3<S> *p, const S *gap, const int *code, int *keep) {
    int ord[AD_MAXPOLY];
    for (int i = 0; i < n; ++i) ord[i] = i;
    for (int a = 1; a < n; ++a) {
        int t = ord[a], b = a - 1;
        while (b >= 0 && val(gap[ord[b]]) > val(gap[t])) { ord[b + 1] = ord[b]; --b; }
        ord[b + 1] = t;
    }
    int nk;
    if (n <= 4) {
        for (int i = 0; i < n; ++i) keep[i] = ord[i];
        nk = n;
    } else {
        unsigned used = 1u;
        keep[0] = ord[0];
        nk = 1;
        while (nk < 4) {
            int best = -1;
            float bd = -1.0f;
            for (int r = 1; r < n; ++r) {
                if (used & (1u << r)) continue;
                float md = 3.4e38f;
                for (int k = 0; k < nk; ++k) {
                    V3<S> e = p[ord[r]] - p[keep[k]];
                    float d2 = val(e.x) * val(e.x) + val(e.y) * val(e.y) + val(e.z) * val(e.z);
                    if (d2 < md) md = d2;
                }
                if (md > bd) { bd = md; best = r; }
            }
            used |= 1u << best;
            keep[nk++] = ord[best];
        }
    }
    for (int a = 1; a < nk; ++a) {
        int t = keep[a], b = a - 1;
        while (b >= 0 && code[keep[b]] > code[t]) { keep[b + 1] = keep[b]; --b; }
        keep[b + 1] = t;
    }
    return nk;
}

template <class S> AD_HD int box_ground_g(V3<S> c, const M3<S> &R, V3<S> h, float margin, V3<S> *pout, S *pen) {
    V3<S> p[8];
    S gap[8];
    int cd[8];
    int n = 0;
    for (int k = 0; k < 8; ++k) {
        V3<S> vk = box_vertex(c, R, h, k);
        if (val(vk.z) < margin) { p[n] = vk; gap[n] = vk.z; cd[n] = k; n++; }
    }
    int keep[AD_MAXPOLY];
    int nk = reduce4_idx(n, p, gap, cd, keep);
    for (int i = 0; i < nk; ++i) { pout[i] = p[keep[i]]; pen[i] = -gap[keep[i]]; }
    return nk;
}

template <class S> AD_HD float sat_overlap_v(V3<S> ax, const M3<S> &RA, V3<S> hA, const M3<S> &RB, V3<S> hB, V3<S> d) {
    auto vv = [](V3<S> a) { return v3<float>(val(a.x), val(a.y), val(a.z)); };
    auto dt = [](V3<float> a, V3<float> b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    V3<float> a = vv(ax), dd = vv(d), h1 = vv(hA), h2 = vv(hB);
    float ra = fabsf(dt(a, vv(colm(RA, 0)))) * h1.x + fabsf(dt(a, vv(colm(RA, 1)))) * h1.y +
               fabsf(dt(a, vv(colm(RA, 2)))) * h1.z;
    float rb = fabsf(dt(a, vv(colm(RB, 0)))) * h2.x + fabsf(dt(a, vv(colm(RB, 1)))) * h2.y +
               fabsf(dt(a, vv(colm(RB, 2)))) * h2.z;
    return ra + rb - fabsf(dt(a, dd));
}

// box_box over S: returns the kept contacts (point, normal from B to A, penetration), or 0
template <class S> AD_HD int box_box_g(V3<S> cA, const M3<S> &RA, V3<S> hA, V3<S> cB, const M3<S> &RB, V3<S> hB,
                                      float margin, V3<S> *pout, V3<S> *nout, S *pen) {
    V3<S> d = cA - cB;
    float best = 3.4e38f;
    int bfrom = -1;
    for (int i = 0; i < 6; ++i) {
        V3<S> ax = i < 3 ? colm(RA, i) : colm(RB, i - 3);
        float o = sat_overlap_v(ax, RA, hA, RB, hB, d);
        if (o < best) { best = o; bfrom = i; }
    }
    if (bfrom < 0 || best < -margin) return 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            V3<float> ca = v3<float>(val(RA.m[0][i]), val(RA.m[1][i]), val(RA.m[2][i]));
            V3<float> cb = v3<float>(val(RB.m[0][j]), val(RB.m[1][j]), val(RB.m[2][j]));
            V3<float> L = cross(ca, cb);
            float l2 = L.x * L.x + L.y * L.y + L.z * L.z;
            if (l2 < 1e-12f) continue;
            V3<float> ax = L / sqrtf(l2);
            M3<float> ra, rb;
            for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 3; ++s) { ra.m[r][s] = val(RA.m[r][s]); rb.m[r][s] = val(RB.m[r][s]); }
            V3<float> hAv = v3<float>(val(hA.x), val(hA.y), val(hA.z)), hBv = v3<float>(val(hB.x), val(hB.y), val(hB.z));
            V3<float> dv = v3<float>(val(d.x), val(d.y), val(d.z));
            if (sat_overlap_v(ax, ra, hAv, rb, hBv, dv) < -margin) return 0;
        }
    V3<S> n = bfrom < 3 ? colm(RA, bfrom) : colm(RB, bfrom - 3);
    if (val(dot(n, d)) < 0.0f) n = -n;
    int refIsB = bfrom >= 3, ai = refIsB ? bfrom - 3 : bfrom;
    V3<S> refc, incc, refh, inch, ref_n;
    M3<S> refR, incR;
    float ref_sign;
    if (!refIsB) {
        refc = cA; refR = RA; refh = hA; incc = cB; incR = RB; inch = hB;
        ref_n = -n;
        ref_sign = val(dot(colm(RA, ai), n)) > 0.0f ? -1.0f : 1.0f;
    } else {
        refc = cB; refR = RB; refh = hB; incc = cA; incR = RA; inch = hA;
        ref_n = n;
        ref_sign = val(dot(colm(RB, ai), n)) > 0.0f ? 1.0f : -1.0f;
    }
    int bi = 0;
    float bs = 1.0f, bv = 3.4e38f;
    for (int i = 0; i < 3; ++i)
        for (int si = 0; si < 2; ++si) {
            float s = si == 0 ? -1.0f : 1.0f;
            float vl = val(dot(colm(incR, i) * S(s), ref_n));
            if (vl < bv) { bv = vl; bi = i; bs = s; }
        }
    PolyG<S> A, B;
    face_poly(incc, incR, inch, bi, bs, A.p);
    A.n = 4;
    for (int k = 0; k < 4; ++k) { A.code[k] = k; A.edge[k] = k; }
    V3<S> rp[4];
    face_poly(refc, refR, refh, ai, ref_sign, rp);
    PolyG<S> *cur = &A, *nxt = &B;
    for (int i = 0; i < 4; ++i) {
        V3<S> e = rp[(i + 1) % 4] - rp[i];
        V3<S> sn = cross(e, ref_n);
        S nl = ssqrt(dot(sn, sn));
        if (val(nl) < 1e-12f) continue;
        sn = sn / nl;
        if (val(dot(sn, refc - rp[i])) > 0.0f) sn = -sn;
        clip_poly(*cur, sn, dot(sn, rp[i]), i, *nxt);
        PolyG<S> *t = cur; cur = nxt; nxt = t;
        if (cur->n == 0) return 0;
    }
    V3<S> p[AD_MAXPOLY];
    S gap[AD_MAXPOLY];
    int cd[AD_MAXPOLY];
    int n2 = 0;
    for (int i = 0; i < cur->n; ++i) {
        bool dup = false;
        for (int k = 0; k < i; ++k) {
            V3<S> e = cur->p[i] - cur->p[k];
            float d2 = val(e.x) * val(e.x) + val(e.y) * val(e.y) + val(e.z) * val(e.z);
            if (d2 <= 1e-12f) { dup = true; break; }
        }
        if (dup) continue;
        S g = dot(ref_n, cur->p[i] - rp[0]);
        if (val(g) < margin) {
            p[n2] = cur->p[i] - ref_n * (S(0.5f) * g);
            gap[n2] = g;
            cd[n2] = cur->code[i];
            n2++;
        }
    }
    int keep[AD_MAXPOLY];
    int nk = reduce4_idx(n2, p, gap, cd, keep);
    for (int i = 0; i < nk; ++i) {
        pout[i] = p[keep[i]];
        nout[i] = n;
        pen[i] = -gap[keep[i]];
    }
    return nk;
}

}  // namespace ad
}  // namespace me
