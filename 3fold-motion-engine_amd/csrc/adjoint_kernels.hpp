// Reverse pass of one recorded ColorCache substep (tiled solve, small tiles, box geometry).
//
// Order (the forward order reversed): integrate -> position sweeps -> velocity sweeps -> warm apply
// (k_adj_tiles, one workgroup per tile, replaying each sweep from its recorded start -- each pair
// keeping its contacts' pre-states in registers -- and reversing its colour steps) -> per-contact constants and the warm-start carry (k_adj_contacts) -> narrow
// phase (k_adj_narrow, dual numbers) -> world inverse inertia (k_adj_body). Gravity adds a constant,
// so its adjoint is the identity. Body-level sums from several contacts use float atomics.
#pragma once
#include "adjoint_geom.hpp"
#include "tiled.hpp"

namespace me {

using ad::Con;
using ad::M3;
using ad::PairBodies;
using ad::PairFlags;
using ad::Q4;
using ad::V3;

__device__ __forceinline__ V3<float> g3(const float *p, size_t i) { return ad::v3<float>(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ void s3(float *p, size_t i, V3<float> a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }
// Body-level adjoint contributions (several contacts, pairs and candidates per body) are summed in
// int64 fixed point, 2^-32 per unit: integer addition is associative, so the sums, and the gradients,
// are the same bits whatever order the atomics land in. Per body ADJ_ACC words: the centre (3),
// the orientation (4), the world inverse inertia (9) and the friction coefficient (1). A contribution
// of magnitude 2^30 or more flags OVF_ADJ and is dropped (the step then raises).
#define ADJ_ACC 17
#define ADJ_X 0
#define ADJ_Q 3
#define ADJ_I 7
#define ADJ_MU 16
__device__ __forceinline__ void fx_add(long long *acc, int f, size_t i, float v, int *istat) {
    if (v == 0.0f) return;
    if (!(fabsf(v) < 1073741824.0f)) {  // also NaN
        atomicOr(&istat[IS_OVF], OVF_ADJ);
        return;
    }
    const long long q = __double2ll_rn((double)v * 4294967296.0);
    atomicAdd(reinterpret_cast<unsigned long long *>(acc + ADJ_ACC * i + f), (unsigned long long)q);
}
__device__ __forceinline__ float fx_get(const long long *acc, int f, size_t i) {
    return (float)((double)acc[ADJ_ACC * i + f] * (1.0 / 4294967296.0));
}
__device__ __forceinline__ void a3(long long *acc, int f, size_t i, V3<float> a, int *istat) {
    fx_add(acc, f, i, a.x, istat); fx_add(acc, f + 1, i, a.y, istat); fx_add(acc, f + 2, i, a.z, istat);
}
__device__ __forceinline__ Q4<float> g4(const float *p, size_t i) {
    return ad::q4<float>(p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]);
}
__device__ __forceinline__ M3<float> g33(const float *p, size_t i) {
    M3<float> r;
    for (int k = 0; k < 9; ++k) r.m[k / 3][k % 3] = p[9 * i + k];
    return r;
}
__device__ __forceinline__ void a33(long long *acc, size_t i, const M3<float> &M, int *istat) {
    for (int k = 0; k < 9; ++k) fx_add(acc, ADJ_I + k, i, M.m[k / 3][k % 3], istat);
}

// zero fill of up to 8 arrays in one launch (the reverse pass's accumulators): a grid-stride loop
// over the concatenated index space
struct ZeroSpans {
    float *p[8];
    size_t n[8];
    int k;
};
__global__ void k_zero_spans(ZeroSpans Z) {
    size_t tot = 0;
    for (int s = 0; s < Z.k; ++s) tot += Z.n[s];
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < tot; g += (size_t)gridDim.x * blockDim.x) {
        size_t i = g;
        int s = 0;
        while (s < Z.k - 1 && i >= Z.n[s]) { i -= Z.n[s]; ++s; }
        Z.p[s][i] = 0.0f;
    }
}

// integrate reverse; vbar/wbar: adjoint of the step's output velocities in, of the solve's output out
__global__ void k_adj_integ(TiledArgs A, float *xbar, float *qbar, float *vbar, float *wbar) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.N) return;
    if (!(A.b.invM[i] > 0.0f)) return;
    V3<float> xb = g3(xbar, i), vb = g3(vbar, i), wb = g3(wbar, i);
    Q4<float> qb = g4(qbar, i);
    V3<float> pvb = ad::zero3<float>(), pob = ad::zero3<float>();
    ad::integrate_rev(g4(A.tape.qpre, i), g3(A.b.w, i), g3(A.po, i), A.dt, xb, qb, vb, wb, pvb, pob);
    s3(xbar, i, xb);
    qbar[4 * i] = qb.x; qbar[4 * i + 1] = qb.y; qbar[4 * i + 2] = qb.z; qbar[4 * i + 3] = qb.w;
    s3(vbar, i, vb);
    s3(wbar, i, wb);
    s3(A.tape.pvb, i, pvb);
    s3(A.tape.pob, i, pob);
}

// pair of the reverse tile kernel: constants, inertia, flags, slots and static-body state
struct AdjPair {
    int cnt, la, lb, a, b, c0;
    PairFlags f;
    float iMA, iMB;
    M3<float> IA, IB;
    Con<float> C[4];
    PairBodies<float> stat;  // static body state (velocities or pseudo-velocities of the phase)
};

__device__ __forceinline__ PairBodies<float> adj_load(const AdjPair &P, const float *s0, const float *s1) {
    PairBodies<float> B = P.stat;
    if (P.f.dynA) { B.vA = g3(s0, P.la); B.wA = g3(s1, P.la); }
    if (P.f.hasB && P.f.dynB) { B.vB = g3(s0, P.lb); B.wB = g3(s1, P.lb); }
    return B;
}
__device__ __forceinline__ void adj_store(const AdjPair &P, float *s0, float *s1, const PairBodies<float> &B) {
    if (P.f.dynA) { s3(s0, P.la, B.vA); s3(s1, P.la, B.wA); }
    if (P.f.hasB && P.f.dynB) { s3(s0, P.lb, B.vB); s3(s1, P.lb, B.wB); }
}
__device__ __forceinline__ PairBodies<float> adj_load_bar(const AdjPair &P, const float *s0, const float *s1) {
    PairBodies<float> B;
    B.vA = B.wA = B.vB = B.wB = ad::zero3<float>();
    if (P.f.dynA) { B.vA = g3(s0, P.la); B.wA = g3(s1, P.la); }
    if (P.f.hasB && P.f.dynB) { B.vB = g3(s0, P.lb); B.wB = g3(s1, P.lb); }
    return B;
}

// reverse of the solve of one small tile: position sweeps, velocity sweeps, warm apply
__global__ void __launch_bounds__(T_SMALL_P) k_adj_tiles(TiledArgs A, TiledBuffers T, float *vbar, float *wbar,
                                                         const float *jbar_end) {
    int t = blockIdx.x;
    if (T.tb0[t] == INT_MAX) return;
    int b0, nb, p0, np;
    tile_bounds(T, t, b0, nb, p0, np);
    if (!is_small_tile(nb, np)) return;
    __shared__ float sv[3 * T_SMALL_B], sw[3 * T_SMALL_B], bv[3 * T_SMALL_B], bw[3 * T_SMALL_B];
    __shared__ int cnt[ME_NCOLMAX + 1], fill[ME_NCOLMAX + 1], plist[T_SMALL_P], plcol[T_SMALL_P];
    __shared__ int maxc;
    const int tid = threadIdx.x;
    const int N = A.N, V = A.vit, Pn = A.pit;
    if (tid == 0) maxc = 0;
    int ncol = A.istat[IS_NCOL];
    if (ncol > ME_NCOLMAX) ncol = ME_NCOLMAX;
    for (int c = tid; c <= ncol; c += T_SMALL_P) { cnt[c] = 0; fill[c] = 0; }
    __syncthreads();
    // the forward's colour order within the tile
    int myc = -1;
    if (tid < np) {
        myc = A.Pp.col[T.psorted[p0 + tid]];
        if (myc < 0 || myc >= ncol) myc = ncol > 0 ? ncol - 1 : 0;
        myc = tile_class(T, myc, ncol);
        atomicAdd(&cnt[myc], 1);
        atomicMax(&maxc, myc + 1);
    }
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int c = 0; c < maxc; ++c) { int n = cnt[c]; cnt[c] = run; run += n; }
    }
    __syncthreads();
    if (tid < np) {
        int pos = cnt[myc] + atomicAdd(&fill[myc], 1);
        plist[pos] = T.psorted[p0 + tid];
        plcol[pos] = myc;
    }
    __syncthreads();
    const int nc = maxc;
    AdjPair P;
    int p = -1, c_own = -1;
    if (tid < np) {
        p = plist[tid];
        c_own = plcol[tid];
        int a = A.Pp.pbi[p], b = A.Pp.pbj[p];
        P.a = a; P.b = b;
        P.iMA = A.b.invM[a];
        P.f.hasB = b >= 0;
        P.iMB = P.f.hasB ? A.b.invM[b] : 0.0f;
        P.f.dynA = P.iMA > 0.0f;
        P.f.dynB = P.f.hasB && P.iMB > 0.0f;
        P.la = P.f.dynA ? T.bslot[a] - b0 : 0;
        P.lb = P.f.dynB ? T.bslot[b] - b0 : 0;
        P.IA = g33(A.invIw, a);
        P.IB = P.f.hasB ? g33(A.invIw, b) : P.IA;
        P.c0 = A.Pp.pstart[p];
        P.cnt = A.Pp.pcount[p];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k >= P.cnt) continue;
            const float *o = T.prep + 20 * (size_t)(P.c0 + k);
            Con<float> &C = P.C[k];
            C.n = ad::v3<float>(o[0], o[1], o[2]); C.rA = ad::v3<float>(o[3], o[4], o[5]);
            C.rB = ad::v3<float>(o[6], o[7], o[8]); C.t1 = ad::v3<float>(o[9], o[10], o[11]);
            C.t2 = ad::v3<float>(o[12], o[13], o[14]);
            C.meff = o[15]; C.meft1 = o[16]; C.meft2 = o[17]; C.bias = o[18]; C.mu = o[19];
        }
    }
    Con<float> cb[4];
    for (int k = 0; k < 4; ++k) cb[k] = ad::con_zero<float>();
    M3<float> IAb = ad::zero33<float>(), IBb = ad::zero33<float>();
    float jb[4][3];
    for (int k = 0; k < 4; ++k) { jb[k][0] = jb[k][1] = jb[k][2] = 0.0f; }

    // ---- position sweeps, last first ----
    if (p >= 0) {
        V3<float> z = ad::zero3<float>();
        P.stat.vA = P.f.dynA ? z : g3(A.pv, P.a); P.stat.wA = P.f.dynA ? z : g3(A.po, P.a);
        bool sB = P.f.hasB && !P.f.dynB;
        P.stat.vB = sB ? g3(A.pv, P.b) : z; P.stat.wB = sB ? g3(A.po, P.b) : z;
    }
    for (int j = tid; j < nb; j += T_SMALL_P) {
        int i = T.bsorted[b0 + j];
        s3(bv, j, g3(A.tape.pvb, i));
        s3(bw, j, g3(A.tape.pob, i));
    }
    float jpb[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int s = Pn - 1; s >= 0 && np > 0; --s) {
        const float *st = A.tape.sst + (size_t)(V + s) * 6 * N;
        for (int j = tid; j < nb; j += T_SMALL_P) {
            int i = T.bsorted[b0 + j];
            s3(sv, j, g3(st, 2 * (size_t)i));
            s3(sw, j, g3(st, 2 * (size_t)i + 1));
        }
        float jp0[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (p >= 0) {
            const float *si = A.tape.simp + (size_t)(V + s) * 4 * A.capC;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < P.cnt) jp0[k] = si[4 * (size_t)(P.c0 + k) + 3];
        }
        __syncthreads();
        // the forward replay keeps each contact's pre-state, which the reverse step then reads
        PairBodies<float> chain[4];
        for (int c = 0; c < nc; ++c) {
            if (c_own == c) {
                PairBodies<float> B = adj_load(P, sv, sw);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (k >= P.cnt) continue;
                    chain[k] = B;
                    float jp = jp0[k];
                    ad::pos_fwd(B, P.C[k], jp, P.f, P.iMA, P.iMB, P.IA, P.IB);
                }
                adj_store(P, sv, sw, B);
            }
            __syncthreads();
        }
        for (int c = nc - 1; c >= 0; --c) {
            if (c_own == c) {
                PairBodies<float> Bb = adj_load_bar(P, bv, bw);
#pragma unroll
                for (int k = 3; k >= 0; --k)
                    if (k < P.cnt)
                        ad::pos_rev(chain[k], P.C[k], jp0[k], P.f, P.iMA, P.iMB, P.IA, P.IB, Bb, jpb[k], cb[k], IAb,
                                    IBb);
                adj_store(P, bv, bw, Bb);
            }
            __syncthreads();
        }
    }

    // ---- velocity sweeps, last first ----
    if (p >= 0) {
        V3<float> z = ad::zero3<float>();
        P.stat.vA = P.f.dynA ? z : g3(A.b.v, P.a); P.stat.wA = P.f.dynA ? z : g3(A.b.w, P.a);
        bool sB = P.f.hasB && !P.f.dynB;
        P.stat.vB = sB ? g3(A.b.v, P.b) : z; P.stat.wB = sB ? g3(A.b.w, P.b) : z;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < P.cnt)
                for (int m = 0; m < 3; ++m) jb[k][m] = jbar_end[3 * (size_t)(P.c0 + k) + m];
    }
    __syncthreads();
    for (int j = tid; j < nb; j += T_SMALL_P) {
        int i = T.bsorted[b0 + j];
        s3(bv, j, g3(vbar, i));
        s3(bw, j, g3(wbar, i));
    }
    for (int s = V - 1; s >= 0 && np > 0; --s) {
        const float *st = A.tape.sst + (size_t)s * 6 * N;
        for (int j = tid; j < nb; j += T_SMALL_P) {
            int i = T.bsorted[b0 + j];
            s3(sv, j, g3(st, 2 * (size_t)i));
            s3(sw, j, g3(st, 2 * (size_t)i + 1));
        }
        float j0[4][3];
        if (p >= 0) {
            const float *si = A.tape.simp + (size_t)s * 4 * A.capC;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < P.cnt)
                    for (int m = 0; m < 3; ++m) j0[k][m] = si[4 * (size_t)(P.c0 + k) + m];
        }
        __syncthreads();
        PairBodies<float> chain[4];
        for (int c = 0; c < nc; ++c) {
            if (c_own == c) {
                PairBodies<float> B = adj_load(P, sv, sw);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (k >= P.cnt) continue;
                    chain[k] = B;
                    float jn = j0[k][0], jt1 = j0[k][1], jt2 = j0[k][2];
                    ad::vel_fwd(B, P.C[k], jn, jt1, jt2, P.f, P.iMA, P.iMB, P.IA, P.IB);
                }
                adj_store(P, sv, sw, B);
            }
            __syncthreads();
        }
        for (int c = nc - 1; c >= 0; --c) {
            if (c_own == c) {
                PairBodies<float> Bb = adj_load_bar(P, bv, bw);
#pragma unroll
                for (int k = 3; k >= 0; --k)
                    if (k < P.cnt)
                        ad::vel_rev(chain[k], P.C[k], j0[k][0], j0[k][1], j0[k][2], P.f, P.iMA, P.iMB, P.IA, P.IB, Bb,
                                    jb[k][0], jb[k][1], jb[k][2], cb[k], IAb, IBb);
                adj_store(P, bv, bw, Bb);
            }
            __syncthreads();
        }
    }

    // ---- warm apply (impulses of the first velocity sweep's start) ----
    if (p >= 0 && np > 0 && V > 0) {
        PairBodies<float> Bb = adj_load_bar(P, bv, bw);
        const float *si = A.tape.simp;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k >= P.cnt) continue;
            const float *o = si + 4 * (size_t)(P.c0 + k);
            ad::warm_rev(Bb, P.C[k], o[0], o[1], o[2], P.f, P.iMA, P.iMB, P.IA, P.IB, jb[k][0], jb[k][1], jb[k][2],
                         cb[k], IAb, IBb);
        }
    }
    __syncthreads();
    for (int j = tid; j < nb; j += T_SMALL_P) {
        int i = T.bsorted[b0 + j];
        s3(vbar, i, g3(bv, j));
        s3(wbar, i, g3(bw, j));
    }
    if (p >= 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k >= P.cnt) continue;
            size_t c = P.c0 + k;
            for (int m = 0; m < 3; ++m) A.tape.jbi[3 * c + m] = jb[k][m];
            float *o = A.tape.cbar + 20 * c;
            const Con<float> &C = cb[k];
            o[0] = C.n.x; o[1] = C.n.y; o[2] = C.n.z; o[3] = C.rA.x; o[4] = C.rA.y; o[5] = C.rA.z;
            o[6] = C.rB.x; o[7] = C.rB.y; o[8] = C.rB.z; o[9] = C.t1.x; o[10] = C.t1.y; o[11] = C.t1.z;
            o[12] = C.t2.x; o[13] = C.t2.y; o[14] = C.t2.z;
            o[15] = C.meff; o[16] = C.meft1; o[17] = C.meft2; o[18] = C.bias; o[19] = C.mu;
        }
        if (P.f.dynA) a33(A.tape.acc, P.a, IAb, A.istat);
        if (P.f.dynB) a33(A.tape.acc, P.b, IBb, A.istat);
    }
}

// per contact: constants -> contact data and bodies (centres and inverse inertias into the body
// sums); the warm-start carry; pair friction -> body mu
__global__ void k_adj_contacts(TiledArgs A, float *jbar_prev) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int C = A.istat[IS_C];
    if (C > A.capC) C = A.capC;
    if (c >= C) return;
    int a = A.Cc.bi[c], b = A.Cc.bj[c];
    bool hasB = b >= 0;
    const float *o = A.tape.cbar + 20 * (size_t)c;
    Con<float> cb;
    cb.n = ad::v3<float>(o[0], o[1], o[2]); cb.rA = ad::v3<float>(o[3], o[4], o[5]);
    cb.rB = ad::v3<float>(o[6], o[7], o[8]); cb.t1 = ad::v3<float>(o[9], o[10], o[11]);
    cb.t2 = ad::v3<float>(o[12], o[13], o[14]);
    cb.meff = o[15]; cb.meft1 = o[16]; cb.meft2 = o[17]; cb.bias = o[18]; cb.mu = o[19];
    M3<float> IA = g33(A.invIw, a), IB = hasB ? g33(A.invIw, b) : IA;
    V3<float> nb = ad::zero3<float>(), pAb = nb, pBb = nb, xAb = nb, xBb = nb;
    M3<float> IAb = ad::zero33<float>(), IBb = ad::zero33<float>();
    float penb = 0.0f, mub = 0.0f;
    ad::prep_rev(g3(A.Cc.n, c), g3(A.Cc.pA, c), g3(A.Cc.pB, c), g3(A.tape.xpre, a),
                 hasB ? g3(A.tape.xpre, b) : ad::zero3<float>(), hasB, IA, IB, A.Cc.pen[c], A.dt, cb, nb, pAb, pBb,
                 xAb, xBb, IAb, IBb, penb, mub);
    bool dA = A.b.invM[a] > 0.0f, dB = hasB && A.b.invM[b] > 0.0f;
    if (dA) { a3(A.tape.acc, ADJ_X, a, xAb, A.istat); a33(A.tape.acc, a, IAb, A.istat); }
    if (dB) { a3(A.tape.acc, ADJ_X, b, xBb, A.istat); a33(A.tape.acc, b, IBb, A.istat); }
    float *g = A.tape.gbar + 7 * (size_t)c;
    V3<float> pb = pAb + pBb;  // box contacts carry one point for both bodies
    g[0] = pb.x; g[1] = pb.y; g[2] = pb.z; g[3] = nb.x; g[4] = nb.y; g[5] = nb.z; g[6] = penb;
    // pair friction coefficient (pair_mu): mu_a (ground), mu_a == mu_b, or sqrt(mu_a mu_b)
    if (!hasB) {
        fx_add(A.tape.acc, ADJ_MU, a, mub, A.istat);
    } else {
        float ma = A.b.mu[a], mb = A.b.mu[b];
        if (ma == mb) {
            fx_add(A.tape.acc, ADJ_MU, a, 0.5f * mub, A.istat);
            fx_add(A.tape.acc, ADJ_MU, b, 0.5f * mub, A.istat);
        } else {
            float m = sqrtf(ma * mb);
            fx_add(A.tape.acc, ADJ_MU, a, mub * mb / (2.0f * m), A.istat);
            fx_add(A.tape.acc, ADJ_MU, b, mub * ma / (2.0f * m), A.istat);
        }
    }
    int w = A.tape.widx[c];
    if (w >= 0) {
        for (int m = 0; m < 3; ++m) jbar_prev[3 * (size_t)w + m] = A.tape.jbi[3 * (size_t)c + m];
    }
}

// narrow phase: contact point, normal and penetration of every kept contact -> centres and
// orientations, by forward-mode duals over one body's 7 coordinates at a time
typedef ad::Dual<float, 7> D7;

__device__ __forceinline__ void adj_narrow_body(int a, int bb, int k, const float *gb, float margin,
                                                const Bodies &b, const float *xpre, const float *qpre, bool seedA,
                                                float out[7], int *bad) {
    V3<D7> xA, xB;
    Q4<D7> qA, qB;
    V3<float> xa = g3(xpre, a);
    Q4<float> qa = g4(qpre, a);
    xA = ad::v3<D7>(D7(xa.x), D7(xa.y), D7(xa.z));
    qA = ad::q4<D7>(D7(qa.x), D7(qa.y), D7(qa.z), D7(qa.w));
    if (bb >= 0) {
        V3<float> xb = g3(xpre, bb);
        Q4<float> qbv = g4(qpre, bb);
        xB = ad::v3<D7>(D7(xb.x), D7(xb.y), D7(xb.z));
        qB = ad::q4<D7>(D7(qbv.x), D7(qbv.y), D7(qbv.z), D7(qbv.w));
    }
    V3<D7> &xs = seedA ? xA : xB;
    Q4<D7> &qs = seedA ? qA : qB;
    xs.x.d[0] = 1.0f; xs.y.d[1] = 1.0f; xs.z.d[2] = 1.0f;
    qs.x.d[3] = 1.0f; qs.y.d[4] = 1.0f; qs.z.d[5] = 1.0f; qs.w.d[6] = 1.0f;
    V3<float> ha = g3(b.half, a);
    V3<D7> hA = ad::v3<D7>(D7(ha.x), D7(ha.y), D7(ha.z));
    // values of the orientation as the forward step formed them (k_body_pre's contracted
    // quat_to_matrix), so the dual narrow phase takes the forward's branches; derivatives from ad::
    auto fwd_R = [](M3<D7> &R, Q4<float> q) {
        f4 qf;
        qf.x = q.x; qf.y = q.y; qf.z = q.z; qf.w = q.w;
        const m3 Rf = cquat_to_matrix(qf);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) R.m[i][j].v = Rf.m[i][j];
    };
    M3<D7> RA = ad::quat_to_matrix(qA);
    fwd_R(RA, qa);
    V3<D7> pp[4], nn[4];
    D7 pen[4];
    int nk;
    if (bb < 0) {
        nk = ad::box_ground_g(xA, RA, hA, margin, pp, pen);
        for (int i = 0; i < nk; ++i) nn[i] = ad::v3<D7>(D7(0.0f), D7(0.0f), D7(1.0f));
    } else {
        V3<float> hb = g3(b.half, bb);
        V3<D7> hB = ad::v3<D7>(D7(hb.x), D7(hb.y), D7(hb.z));
        M3<D7> RB = ad::quat_to_matrix(qB);
        fwd_R(RB, g4(qpre, bb));
        nk = ad::box_box_g(xA, RA, hA, xB, RB, hB, margin, pp, nn, pen);
    }
    if (nk != k) { atomicOr(bad, OVF_TAPE); nk = nk < k ? nk : k; }
    for (int d = 0; d < 7; ++d) out[d] = 0.0f;
    for (int i = 0; i < nk; ++i) {
        const float *g = gb + 7 * i;
        for (int d = 0; d < 7; ++d)
            out[d] += g[0] * pp[i].x.d[d] + g[1] * pp[i].y.d[d] + g[2] * pp[i].z.d[d] + g[3] * nn[i].x.d[d] +
                      g[4] * nn[i].y.d[d] + g[5] * nn[i].z.d[d] + g[6] * pen[i].d[d];
    }
}

// two threads per candidate: the even one differentiates with respect to body a, the odd one to body
// b (each re-runs the clipping in dual numbers seeded for its own body)
__global__ void k_adj_narrow(TiledArgs A, int cap, const int *cand_a, const int *cand_b, const int *slot_n,
                             const int *slot_off, float margin) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = t >> 1;
    const bool sideA = (t & 1) == 0;
    int n = A.istat[IS_CAND];
    if (n > cap) n = cap;
    if (p >= n) return;
    int k = slot_n[p];
    if (k <= 0) return;
    int a = cand_a[p], bb = cand_b[p], o = slot_off[p];
    if (o + k > A.capC) return;
    const int body = sideA ? a : bb;
    if (body < 0 || !(A.b.invM[body] > 0.0f)) return;
    const float *gb = A.tape.gbar + 7 * (size_t)o;
    float g[7];
    adj_narrow_body(a, bb, k, gb, margin, A.b, A.tape.xpre, A.tape.qpre, sideA, g, A.istat + IS_OVF);
    for (int m = 0; m < 7; ++m) fx_add(A.tape.acc, ADJ_X + m, body, g[m], A.istat);
}

// the body's fixed-point sums into its adjoints (centre, orientation, friction), then world inverse
// inertia -> orientation
__global__ void k_adj_body(TiledArgs A, float *xbar, float *qbar, float *mubar) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.N) return;
    const long long *acc = A.tape.acc;
    for (int k = 0; k < 3; ++k) xbar[3 * i + k] += fx_get(acc, ADJ_X + k, i);
    for (int k = 0; k < 4; ++k) qbar[4 * i + k] += fx_get(acc, ADJ_Q + k, i);
    mubar[i] += fx_get(acc, ADJ_MU, i);
    if (!(A.b.invM[i] > 0.0f) || A.b.kin[i] == 1) return;
    M3<float> Ib;
    for (int k = 0; k < 9; ++k) Ib.m[k / 3][k % 3] = fx_get(acc, ADJ_I + k, i);
    Q4<float> q = g4(A.tape.qpre, i);
    M3<float> R = ad::quat_to_matrix(q), Rb = ad::zero33<float>();
    ad::inv_inertia_rev(R, g33(A.b.IbInv, i), Ib, Rb);
    Q4<float> qb = ad::q4<float>(0.0f, 0.0f, 0.0f, 0.0f);
    ad::quat_to_matrix_rev(q, Rb, qb);
    qbar[4 * i] += qb.x; qbar[4 * i + 1] += qb.y; qbar[4 * i + 2] += qb.z; qbar[4 * i + 3] += qb.w;
}

// substep-start copy of centres and orientations
__global__ void k_tape_xq(int N, const float *x, const float *q, float *xpre, float *qpre) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int k = 0; k < 3; ++k) xpre[3 * i + k] = x[3 * i + k];
    for (int k = 0; k < 4; ++k) qpre[4 * i + k] = q[4 * i + k];
}

}  // namespace me
