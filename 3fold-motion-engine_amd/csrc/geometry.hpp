// Narrow phase for the box body model (per-body extents, <= 4 contact points per body pair).
//
// The reference's GPU engines model bodies as bead clouds and cannot express the box workloads;
// its own box precedent is ncp/ncp_ref.py:1065-1211 (face-axis SAT, reference-face clipping,
// deepest-first farthest-point reduction to 4 points, normal from B to A). This is that algorithm
// in float32 for one thread per body pair, with three additions:
//   * the 9 edge-edge axes are separation tests (no edge-edge manifold, as in the precedent);
//   * every clipped vertex carries a feature code, so the warm start can match contacts by
//     (pair, reference face, incident face, feature) across steps;
//   * the kept points are ordered by feature code, so the contact list of a step is sorted by
//     feature key without a sort.
// Feature code (13 bits): refIsB<<11 | refFace<<8 | incFace<<5 | vertex code (0..35), where a
// face is 2*axis + (sign > 0) and the vertex code is an incident-face corner (0..3), an incident
// edge e crossing reference side plane i (4 + 4e + i), or a clip-line j crossing plane i
// (20 + 4j + i). Box-ground contacts use the box corner index (0..7).
#pragma once
#include "vecmath.hpp"

namespace me {

#define ME_CLIP_EPS 1e-6f
#define ME_MAXPOLY 16
#define ME_BP_SLACK 1e-4f

__device__ __forceinline__ float hget(f3 h, int i) { return i == 0 ? h.x : (i == 1 ? h.y : h.z); }

__device__ __forceinline__ f3 box_vertex(f3 c, const m3 &R, f3 h, int k) {
    f3 l = mk((k & 1) ? h.x : -h.x, (k & 2) ? h.y : -h.y, (k & 4) ? h.z : -h.z);
    return c + mv(R, l);
}

// face (axis a, sign s) corners in the order (-,-), (-,+), (+,+), (+,-) over the other two axes
__device__ __forceinline__ void face_poly(f3 c, const m3 &R, f3 h, int a, float s, f3 out[4]) {
    int o0 = a == 0 ? 1 : 0, o1 = a == 2 ? 1 : 2;
    const float S0[4] = {-1.0f, -1.0f, 1.0f, 1.0f};
    const float S1[4] = {-1.0f, 1.0f, 1.0f, -1.0f};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float l[3];
        l[a] = s * hget(h, a);
        l[o0] = S0[k] * hget(h, o0);
        l[o1] = S1[k] * hget(h, o1);
        out[k] = c + mv(R, mk(l[0], l[1], l[2]));
    }
}

struct Poly {
    f3 p[ME_MAXPOLY];
    unsigned char code[ME_MAXPOLY];  // feature codes < 40
    unsigned char edge[ME_MAXPOLY];  // edge ids < 8
    int n;
};

// Sutherland-Hodgman against sn.x <= d; a vertex within eps of the plane is inside and makes no
// intersection point
__device__ __forceinline__ void clip_poly(const Poly &in, f3 sn, float d, int plane, Poly &out) {
    out.n = 0;
    int m = in.n;
    for (int i = 0; i < m; ++i) {
        f3 a = in.p[i], b = in.p[(i + 1) % m];
        float da = dot(sn, a) - d;
        float db = dot(sn, b) - d;
        int E = in.edge[i];
        if (da <= ME_CLIP_EPS && out.n < ME_MAXPOLY) {
            out.p[out.n] = a;
            out.code[out.n] = in.code[i];
            out.edge[out.n] = (db > ME_CLIP_EPS && !(da < -ME_CLIP_EPS)) ? 4 + plane : E;
            out.n++;
        }
        bool enter = da > ME_CLIP_EPS && db < -ME_CLIP_EPS;
        bool exitp = da < -ME_CLIP_EPS && db > ME_CLIP_EPS;
        if ((enter || exitp) && out.n < ME_MAXPOLY) {
            float t = da / (da - db);
            out.p[out.n] = a + t * (b - a);
            out.code[out.n] = E < 4 ? 4 + 4 * E + plane : 20 + 4 * (E - 4) + plane;
            out.edge[out.n] = exitp ? 4 + plane : E;
            out.n++;
        }
    }
}

// deepest first (stable by gap), then greedily the point with the largest squared distance to
// the kept ones; returns the kept indices ordered by feature code
__device__ __forceinline__ int reduce4(int n, const f3 *p, const float *gap, const int *code, int *keep, int *ord) {
    for (int i = 0; i < n; ++i) ord[i] = i;
    for (int a = 1; a < n; ++a) {
        int t = ord[a], b = a - 1;
        while (b >= 0 && gap[ord[b]] > gap[t]) { ord[b + 1] = ord[b]; --b; }
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
                    f3 e = p[ord[r]] - p[keep[k]];
                    float d2 = dot(e, e);
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

// box vs the ground plane z = 0 (normal +z, ground -> box): corners with z < margin, at most 4
__device__ __forceinline__ int box_ground(f3 c, const m3 &R, f3 h, float margin, f3 *pout, float *pen,
                                          int *code) {
    f3 p[8];
    float gap[8];
    int cd[8];
    int n = 0;
    for (int k = 0; k < 8; ++k) {
        f3 vk = box_vertex(c, R, h, k);
        if (vk.z < margin) { p[n] = vk; gap[n] = vk.z; cd[n] = k; n++; }
    }
    int keep[ME_MAXPOLY], ord[ME_MAXPOLY];
    int nk = reduce4(n, p, gap, cd, keep, ord);
    for (int i = 0; i < nk; ++i) { pout[i] = p[keep[i]]; pen[i] = -gap[keep[i]]; code[i] = cd[keep[i]]; }
    return nk;
}

__device__ __forceinline__ float sat_overlap(f3 ax, const m3 &RA, f3 hA, const m3 &RB, f3 hB, f3 d) {
    float ra = fabsf(dot(ax, col(RA, 0))) * hA.x + fabsf(dot(ax, col(RA, 1))) * hA.y +
               fabsf(dot(ax, col(RA, 2))) * hA.z;
    float rb = fabsf(dot(ax, col(RB, 0))) * hB.x + fabsf(dot(ax, col(RB, 1))) * hB.y +
               fabsf(dot(ax, col(RB, 2))) * hB.z;
    return ra + rb - fabsf(dot(ax, d));
}

// box A vs box B, normal from B to A
// the dynamically indexed arrays of box_box: the clipping polygons and the clipped points (the
// caller's LDS: in private memory they would sit in scratch)
struct ClipWork {
    Poly A, B;
    f3 p[ME_MAXPOLY];
    float gap[ME_MAXPOLY];
    int cd[ME_MAXPOLY], keep[ME_MAXPOLY], ord[ME_MAXPOLY];
};

__device__ __forceinline__ int box_box(f3 cA, const m3 &RA, f3 hA, f3 cB, const m3 &RB, f3 hB, float margin,
                                       f3 *pout, f3 *nout, float *pen, int *code, ClipWork &W) {
    Poly &A = W.A, &B = W.B;
    f3 d = cA - cB;
    float best = 3.4e38f;
    int bfrom = -1;
    for (int i = 0; i < 6; ++i) {
        f3 ax = i < 3 ? col(RA, i) : col(RB, i - 3);
        float o = sat_overlap(ax, RA, hA, RB, hB, d);
        if (o < best) { best = o; bfrom = i; }
    }
    if (bfrom < 0 || best < -margin) return 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            f3 L = cross(col(RA, i), col(RB, j));
            float l2 = dot(L, L);
            if (l2 < 1e-12f) continue;
            f3 ax = L / sqrtf(l2);
            if (sat_overlap(ax, RA, hA, RB, hB, d) < -margin) return 0;
        }
    f3 n = bfrom < 3 ? col(RA, bfrom) : col(RB, bfrom - 3);
    if (dot(n, d) < 0.0f) n = -n;
    int refIsB = bfrom >= 3, ai = refIsB ? bfrom - 3 : bfrom;
    f3 refc, incc, refh, inch, ref_n;
    m3 refR, incR;
    float ref_sign;
    if (!refIsB) {
        refc = cA; refR = RA; refh = hA; incc = cB; incR = RB; inch = hB;
        ref_n = -n;
        ref_sign = dot(col(RA, ai), n) > 0.0f ? -1.0f : 1.0f;
    } else {
        refc = cB; refR = RB; refh = hB; incc = cA; incR = RA; inch = hA;
        ref_n = n;
        ref_sign = dot(col(RB, ai), n) > 0.0f ? 1.0f : -1.0f;
    }
    int bi = 0;
    float bs = 1.0f, bv = 3.4e38f;
    for (int i = 0; i < 3; ++i)
        for (int si = 0; si < 2; ++si) {
            float s = si == 0 ? -1.0f : 1.0f;
            float val = dot(col(incR, i) * s, ref_n);
            if (val < bv) { bv = val; bi = i; bs = s; }
        }
    face_poly(incc, incR, inch, bi, bs, A.p);
    A.n = 4;
    for (int k = 0; k < 4; ++k) { A.code[k] = k; A.edge[k] = k; }
    f3 rp[4];
    face_poly(refc, refR, refh, ai, ref_sign, rp);
    Poly *cur = &A, *nxt = &B;
    for (int i = 0; i < 4; ++i) {
        f3 e = rp[(i + 1) % 4] - rp[i];
        f3 sn = cross(e, ref_n);
        float nl = length(sn);
        if (nl < 1e-12f) continue;
        sn = sn / nl;
        if (dot(sn, refc - rp[i]) > 0.0f) sn = -sn;
        clip_poly(*cur, sn, dot(sn, rp[i]), i, *nxt);
        Poly *t = cur; cur = nxt; nxt = t;
        if (cur->n == 0) return 0;
    }
    f3 *p = W.p;
    float *gap = W.gap;
    int *cd = W.cd;
    int n2 = 0;
    for (int i = 0; i < cur->n; ++i) {
        bool dup = false;
        for (int k = 0; k < i; ++k) {
            f3 e = cur->p[i] - cur->p[k];
            if (dot(e, e) <= 1e-12f) { dup = true; break; }
        }
        if (dup) continue;
        float g = dot(ref_n, cur->p[i] - rp[0]);
        if (g < margin) {
            p[n2] = cur->p[i] - ref_n * (0.5f * g);
            gap[n2] = g;
            cd[n2] = cur->code[i];
            n2++;
        }
    }
    int *keep = W.keep;
    int nk = reduce4(n2, p, gap, cd, keep, W.ord);
    int face = (refIsB << 11) | ((ai * 2 + (ref_sign > 0.0f ? 1 : 0)) << 8) | ((bi * 2 + (bs > 0.0f ? 1 : 0)) << 5);
    for (int i = 0; i < nk; ++i) {
        pout[i] = p[keep[i]];
        nout[i] = n;
        pen[i] = -gap[keep[i]];
        code[i] = face | cd[keep[i]];
    }
    return nk;
}

}  // namespace me
