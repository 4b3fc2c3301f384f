// Big tiles (more pairs than one small tile holds): one 512-thread workgroup per tile.
//
// LDS mode (tile within T_BIG_B bodies and T_BIG_P pairs): the tile's pairs are counting-sorted by
// colour; each sorted pair keeps a 32-bit word in LDS (body slots, contact count, flags) and its
// contact constants and impulses as 96-byte records in global memory in sorted order, so a colour
// step is one LDS read plus one round of 16-byte vector loads per thread. Body velocities
// (velocity passes) and then pseudo-velocities (position passes) sit in LDS.
// Global mode (larger tiles): the same sweeps with body state in global memory.
#pragma once

#define T_BIG_P 8192  // pairs of an LDS-mode big tile (32-bit word each)

// pair word: la (13 bits) | lb (13 bits) << 13 | (cnt - 1) (2 bits) << 26 | dynA << 28 | dynB << 29 | hasB << 30
__device__ __forceinline__ unsigned pw_make(int la, int lb, int cnt, bool dA, bool dB, bool hB) {
    return (unsigned)(la & 8191) | ((unsigned)(lb & 8191) << 13) | ((unsigned)(cnt - 1) << 26) |
           ((unsigned)dA << 28) | ((unsigned)dB << 29) | ((unsigned)hB << 30);
}

__device__ __forceinline__ void ld_sprep(const float *o, CCon &C) {
    float4 q0 = *reinterpret_cast<const float4 *>(o);
    float4 q1 = *reinterpret_cast<const float4 *>(o + 4);
    float4 q2 = *reinterpret_cast<const float4 *>(o + 8);
    float4 q3 = *reinterpret_cast<const float4 *>(o + 12);
    float4 q4 = *reinterpret_cast<const float4 *>(o + 16);
    float4 q5 = *reinterpret_cast<const float4 *>(o + 20);
    C.n = mk(q0.x, q0.y, q0.z); C.rA = mk(q0.w, q1.x, q1.y); C.rB = mk(q1.z, q1.w, q2.x);
    C.t1 = mk(q2.y, q2.z, q2.w); C.t2 = mk(q3.x, q3.y, q3.z);
    C.meff = q3.w; C.meft1 = q4.x; C.meft2 = q4.y; C.bias = q4.z; C.mu = q4.w;
    C.jn = q5.x; C.jt1 = q5.y; C.jt2 = q5.z; C.jp = q5.w;
}

__device__ __forceinline__ m3 ld_m12(const float *o) {
    float4 a = *reinterpret_cast<const float4 *>(o);
    float4 b = *reinterpret_cast<const float4 *>(o + 4);
    float4 c = *reinterpret_cast<const float4 *>(o + 8);
    m3 M;
    M.m[0][0] = a.x; M.m[0][1] = a.y; M.m[0][2] = a.z; M.m[1][0] = a.w; M.m[1][1] = b.x; M.m[1][2] = b.y;
    M.m[2][0] = b.z; M.m[2][1] = b.w; M.m[2][2] = c.x;
    return M;
}

// PairCache of the k-th colour-sorted pair of an LDS-mode big tile (q = p0 + k)
__device__ __forceinline__ void pc_from_word(PairCache &P, const TiledArgs &A, const TiledBuffers &T, unsigned w,
                                             int q, int b0) {
    P.la = w & 8191;
    P.lb = (w >> 13) & 8191;
    P.cnt = (int)((w >> 26) & 3) + 1;
    P.dynA = (w >> 28) & 1;
    P.dynB = (w >> 29) & 1;
    P.hasB = (w >> 30) & 1;
    f3 z = mk(0.0f, 0.0f, 0.0f);
    P.vAs = z; P.wAs = z; P.pvAs = z; P.poAs = z; P.vBs = z; P.wBs = z; P.pvBs = z; P.poBs = z;
    if (P.dynA) {
        P.iMA = T.invMs[b0 + P.la];
        P.IA = ld_m12(T.invIws + 12 * (size_t)(b0 + P.la));
    } else {
        int a = T.srec[4 * (size_t)q + 1];
        P.iMA = A.b.invM[a];
        P.IA = ldm(A.invIw, a);
        P.vAs = ld3(A.b.v, a); P.wAs = ld3(A.b.w, a); P.pvAs = ld3(A.pv, a); P.poAs = ld3(A.po, a);
    }
    P.iMB = 0.0f;
    P.IB = P.IA;
    if (P.dynB) {
        P.iMB = T.invMs[b0 + P.lb];
        P.IB = ld_m12(T.invIws + 12 * (size_t)(b0 + P.lb));
    } else if (P.hasB) {
        int b = T.srec[4 * (size_t)q + 2];
        P.iMB = A.b.invM[b];
        P.IB = ldm(A.invIw, b);
        P.vBs = ld3(A.b.v, b); P.wBs = ld3(A.b.w, b); P.pvBs = ld3(A.pv, b); P.poBs = ld3(A.po, b);
    }
    const float *o = T.sprep + 96 * (size_t)q;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= P.cnt) break;
        ld_sprep(o + 24 * k, P.c[k]);
    }
}

__device__ __forceinline__ void pc_store_sprep(const PairCache &P, const TiledBuffers &T, int q) {
    float *o = T.sprep + 96 * (size_t)q;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= P.cnt) break;
        *reinterpret_cast<float4 *>(o + 24 * k + 20) = make_float4(P.c[k].jn, P.c[k].jt1, P.c[k].jt2, P.c[k].jp);
    }
}

// PairCache of a colour-sorted global-mode record (8 ints: p, a, b, la, lb, st, cnt, flags)
__device__ __forceinline__ void pc_from_rec(PairCache &P, const TiledArgs &A, const TiledBuffers &T, const int *r) {
    int a = r[1], b = r[2], flags = r[7];
    P.la = r[3]; P.lb = r[4]; P.cnt = r[6];
    P.dynA = flags & 1; P.dynB = (flags & 2) != 0; P.hasB = b >= 0;
    const float *invM = A.b.invM;
    P.iMA = invM[a];
    P.iMB = P.hasB ? invM[b] : 0.0f;
    P.IA = ldm(A.invIw, a);
    P.IB = P.hasB ? ldm(A.invIw, b) : P.IA;
    f3 z = mk(0.0f, 0.0f, 0.0f);
    P.vAs = P.dynA ? z : ld3(A.b.v, a); P.wAs = P.dynA ? z : ld3(A.b.w, a);
    P.pvAs = P.dynA ? z : ld3(A.pv, a); P.poAs = P.dynA ? z : ld3(A.po, a);
    bool sB = P.hasB && !P.dynB;
    P.vBs = sB ? ld3(A.b.v, b) : z; P.wBs = sB ? ld3(A.b.w, b) : z;
    P.pvBs = sB ? ld3(A.pv, b) : z; P.poBs = sB ? ld3(A.po, b) : z;
    int st = r[5];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= P.cnt) break;
        int c = st + k;
        const float *o = T.prep + 20 * (size_t)c;
        CCon &C = P.c[k];
        C.n = mk(o[0], o[1], o[2]); C.rA = mk(o[3], o[4], o[5]); C.rB = mk(o[6], o[7], o[8]);
        C.t1 = mk(o[9], o[10], o[11]); C.t2 = mk(o[12], o[13], o[14]);
        C.meff = o[15]; C.meft1 = o[16]; C.meft2 = o[17]; C.bias = o[18]; C.mu = o[19];
        C.jn = A.Cc.jn[c]; C.jt1 = A.Cc.jt1[c]; C.jt2 = A.Cc.jt2[c]; C.jp = A.Cc.jp[c];
    }
}

__device__ __forceinline__ void ld_prep20(const float *o, CCon &C) {
    float4 q0 = *reinterpret_cast<const float4 *>(o);
    float4 q1 = *reinterpret_cast<const float4 *>(o + 4);
    float4 q2 = *reinterpret_cast<const float4 *>(o + 8);
    float4 q3 = *reinterpret_cast<const float4 *>(o + 12);
    float4 q4 = *reinterpret_cast<const float4 *>(o + 16);
    C.n = mk(q0.x, q0.y, q0.z); C.rA = mk(q0.w, q1.x, q1.y); C.rB = mk(q1.z, q1.w, q2.x);
    C.t1 = mk(q2.y, q2.z, q2.w); C.t2 = mk(q3.x, q3.y, q3.z);
    C.meff = q3.w; C.meft1 = q4.x; C.meft2 = q4.y; C.bias = q4.z; C.mu = q4.w;
}

// One pair of a big tile for one colour step, contacts streamed one at a time (few live registers).
// PASS 0 warm apply, 1 velocity, 2 position. LDS mode: body state s0/s1 in LDS by tile slot, the
// pair word and its colour-ordered records; global mode: record r (8 ints), body state in global
// memory by body id, constants from prep, impulses from the contact arrays.
template <int PASS, bool FIXED, bool LDSM>
__device__ __forceinline__ void big_pair(const TiledArgs &A, const TiledBuffers &T, unsigned w, const int *r, int q,
                                         int b0, float *s0, float *s1, float t2d) {
    int la, lb, cnt, a = -1, b = -1, st = 0;
    bool dynA, dynB, hasB;
    if (LDSM) {
        la = w & 8191; lb = (w >> 13) & 8191; cnt = (int)((w >> 26) & 3) + 1;
        dynA = (w >> 28) & 1; dynB = (w >> 29) & 1; hasB = (w >> 30) & 1;
    } else {
        a = r[1]; b = r[2]; la = a; lb = b; st = r[5]; cnt = r[6];
        dynA = r[7] & 1; dynB = (r[7] & 2) != 0; hasB = b >= 0;
    }
    const float *gs0 = PASS == 2 ? A.pv : A.b.v, *gs1 = PASS == 2 ? A.po : A.b.w;
    f3 z = mk(0.0f, 0.0f, 0.0f);
    float iMA, iMB = 0.0f;
    m3 IA, IB;
    f3 xA, yA, xB = z, yB = z;
    if (LDSM && dynA) {
        iMA = T.invMs[b0 + la];
        IA = ld_m12(T.invIws + 12 * (size_t)(b0 + la));
        xA = ld3(s0, la); yA = ld3(s1, la);
    } else {
        if (LDSM) a = T.srec[4 * (size_t)q + 1];
        iMA = A.b.invM[a];
        IA = ldm(A.invIw, a);
        xA = (!LDSM && dynA) ? ld3(s0, a) : ld3(gs0, a);
        yA = (!LDSM && dynA) ? ld3(s1, a) : ld3(gs1, a);
    }
    IB = IA;
    if (LDSM && dynB) {
        iMB = T.invMs[b0 + lb];
        IB = ld_m12(T.invIws + 12 * (size_t)(b0 + lb));
        xB = ld3(s0, lb); yB = ld3(s1, lb);
    } else if (hasB) {
        if (LDSM) b = T.srec[4 * (size_t)q + 2];
        iMB = A.b.invM[b];
        IB = ldm(A.invIw, b);
        xB = (!LDSM && dynB) ? ld3(s0, b) : ld3(gs0, b);
        yB = (!LDSM && dynB) ? ld3(s1, b) : ld3(gs1, b);
    }
    float *o = T.sprep + 96 * (size_t)q;
    // software pipeline: contact k+1 is loaded before contact k is solved and stored
    auto load = [&](int k, CCon &C) {
        if (LDSM) {
            ld_sprep(o + 24 * k, C);
        } else {
            int c = st + k;
            ld_prep20(T.prep + 20 * (size_t)c, C);
            C.jn = A.Cc.jn[c]; C.jt1 = A.Cc.jt1[c]; C.jt2 = A.Cc.jt2[c]; C.jp = A.Cc.jp[c];
        }
    };
    CCon Cnext;
    load(0, Cnext);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= cnt) break;
        CCon C = Cnext;
        if (k + 1 < cnt) load(k + 1, Cnext);
        if (PASS == 0) {
            cc_warm(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB);
        } else {
            if (PASS == 1) cc_vel<FIXED>(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB, t2d);
            else cc_pos(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB);
            if (LDSM) {
                *reinterpret_cast<float4 *>(o + 24 * k + 20) = make_float4(C.jn, C.jt1, C.jt2, C.jp);
            } else {
                int c = st + k;
                A.Cc.jn[c] = C.jn; A.Cc.jt1[c] = C.jt1; A.Cc.jt2[c] = C.jt2; A.Cc.jp[c] = C.jp;
            }
        }
    }
    if (iMA > 0.0f) { st3(s0, la, xA); st3(s1, la, yA); }
    if (hasB && iMB > 0.0f) { st3(s0, lb, xB); st3(s1, lb, yB); }
}

// LDS-mode pair data loaded ahead of its colour step: inverse mass and inertia of dynamic bodies and
// the contact records (constants + impulses). gfx9 counts stores in vmcnt, so loads issued after a
// step's impulse stores would wait for them; loading the next step's pair before the current pair
// is solved and stored keeps the stores off the critical path.
struct BigPre {
    unsigned w;
    int q;
    float iMA, iMB;
    m3 IA, IB;
    CCon c[4];
};

__device__ __forceinline__ void big_prefetch(BigPre &B, const TiledBuffers &T, unsigned w, int q, int b0) {
    B.w = w;
    B.q = q;
    int la = w & 8191, lb = (w >> 13) & 8191, cnt = (int)((w >> 26) & 3) + 1;
    bool dynA = (w >> 28) & 1, dynB = (w >> 29) & 1;
    if (dynA) {
        B.iMA = T.invMs[b0 + la];
        B.IA = ld_m12(T.invIws + 12 * (size_t)(b0 + la));
    }
    if (dynB) {
        B.iMB = T.invMs[b0 + lb];
        B.IB = ld_m12(T.invIws + 12 * (size_t)(b0 + lb));
    }
    const float *o = T.sprep + 96 * (size_t)q;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k < cnt) ld_sprep(o + 24 * k, B.c[k]);
}

// solve one prefetched pair (static bodies, which are rare, are read here)
template <int PASS, bool FIXED>
__device__ __forceinline__ void big_solve(BigPre &B, const TiledArgs &A, const TiledBuffers &T, float *s0, float *s1,
                                          float t2d) {
    unsigned w = B.w;
    int q = B.q;
    int la = w & 8191, lb = (w >> 13) & 8191, cnt = (int)((w >> 26) & 3) + 1;
    bool dynA = (w >> 28) & 1, dynB = (w >> 29) & 1, hasB = (w >> 30) & 1;
    const float *gs0 = PASS == 2 ? A.pv : A.b.v, *gs1 = PASS == 2 ? A.po : A.b.w;
    f3 z = mk(0.0f, 0.0f, 0.0f);
    float iMA, iMB = 0.0f;
    m3 IA, IB;
    f3 xA, yA, xB = z, yB = z;
    if (dynA) {
        iMA = B.iMA; IA = B.IA;
        xA = ld3(s0, la); yA = ld3(s1, la);
    } else {
        int a = T.srec[4 * (size_t)q + 1];
        iMA = A.b.invM[a]; IA = ldm(A.invIw, a);
        xA = ld3(gs0, a); yA = ld3(gs1, a);
    }
    IB = IA;
    if (dynB) {
        iMB = B.iMB; IB = B.IB;
        xB = ld3(s0, lb); yB = ld3(s1, lb);
    } else if (hasB) {
        int b = T.srec[4 * (size_t)q + 2];
        iMB = A.b.invM[b]; IB = ldm(A.invIw, b);
        xB = ld3(gs0, b); yB = ld3(gs1, b);
    }
    float *o = T.sprep + 96 * (size_t)q;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= cnt) break;
        CCon &C = B.c[k];
        if (PASS == 0) {
            cc_warm(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB);
        } else {
            if (PASS == 1) cc_vel<FIXED>(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB, t2d);
            else cc_pos(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB);
            *reinterpret_cast<float4 *>(o + 24 * k + 20) = make_float4(C.jn, C.jt1, C.jt2, C.jp);
        }
    }
    if (iMA > 0.0f) { st3(s0, la, xA); st3(s1, la, yA); }
    if (hasB && iMB > 0.0f) { st3(s0, lb, xB); st3(s1, lb, yB); }
}

#define T_BIG_WG 1024
#define T_BIG_DYN (6 * sizeof(float) * T_BIG_B + sizeof(unsigned) * T_BIG_P)

template <int SOLVER>
__global__ void __launch_bounds__(T_BIG_WG) k_t_big(TiledArgs A, TiledBuffers T) {
    constexpr bool FIXED = SOLVER == ME_SOLVER_COLOR_CACHE;
    int t = blockIdx.x;
    if (T.tb0[t] == INT_MAX) return;
    int b0, nb, p0, np;
    tile_bounds(T, t, b0, nb, p0, np);
    if (is_small_tile(nb, np)) return;
    extern __shared__ float dsm[];
    __shared__ int cst[ME_NCOLMAX + 1], cfl[ME_NCOLMAX + 1];
    __shared__ int maxc;
    const int tid = threadIdx.x;
    const bool lds = nb <= T_BIG_B && np <= T_BIG_P;
    if (tid == 0 && nb >= 1000) { T.dbg[0] = wall_clock64(); T.dbg[7] = nb; T.dbg[8] = np; }
    unsigned *pword = reinterpret_cast<unsigned *>(dsm + 6 * T_BIG_B);
    if (tid == 0) {
        atomicAdd(&A.istat[IS_TILES], 1);
        atomicMax(&A.istat[IS_MAXTB], nb);
        maxc = 0;
    }
    int ncol = A.istat[IS_NCOL];
    if (ncol > ME_NCOLMAX) ncol = ME_NCOLMAX;
    float *bv = lds ? dsm : A.b.v, *bw = lds ? dsm + 3 * T_BIG_B : A.b.w;
    float *bpv = lds ? dsm : A.pv, *bpo = lds ? dsm + 3 * T_BIG_B : A.po;
    if (lds)
        for (int j = tid; j < nb; j += T_BIG_WG) {
            int i = T.bsorted[b0 + j];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                bv[3 * j + k] = A.b.v[3 * i + k];
                bw[3 * j + k] = A.b.w[3 * i + k];
            }
        }
    for (int c = tid; c <= ncol; c += T_BIG_WG) { cst[c] = 0; cfl[c] = 0; }
    __syncthreads();
    if (tid == 0 && nb >= 1000) T.dbg[1] = wall_clock64();
    for (int k = tid; k < np; k += T_BIG_WG) {
        int c = A.Pp.col[T.psorted[p0 + k]];
        if (c < 0 || c >= ncol) c = ncol > 0 ? ncol - 1 : 0;
        atomicAdd(&cst[c], 1);
        atomicMax(&maxc, c + 1);
    }
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int c = 0; c < maxc; ++c) { int n = cst[c]; cst[c] = run; run += n; }
        cst[maxc] = run;
    }
    __syncthreads();
    if (tid == 0 && nb >= 1000) T.dbg[2] = wall_clock64();
    const float *invM = A.b.invM;
    for (int k = tid; k < np; k += T_BIG_WG) {
        int p = T.psorted[p0 + k];
        int c = A.Pp.col[p];
        if (c < 0 || c >= ncol) c = ncol > 0 ? ncol - 1 : 0;
        int pos = cst[c] + atomicAdd(&cfl[c], 1);
        int q = p0 + pos;
        int a = A.Pp.pbi[p], b = A.Pp.pbj[p];
        bool dA = invM[a] > 0.0f, dB = b >= 0 && invM[b] > 0.0f;
        int st = A.Pp.pstart[p], cnt = A.Pp.pcount[p];
        if (lds) {
            int la = dA ? T.bslot[a] - b0 : 0, lb = dB ? T.bslot[b] - b0 : 0;
            pword[pos] = pw_make(la, lb, cnt, dA, dB, b >= 0);
            int *sr = T.srec + 4 * (size_t)q;
            sr[0] = p; sr[1] = a; sr[2] = b; sr[3] = st;
            float4 *o = reinterpret_cast<float4 *>(T.sprep + 96 * (size_t)q);
            for (int i = 0; i < cnt; ++i) {
                const float4 *s4 = reinterpret_cast<const float4 *>(T.prep + 20 * (size_t)(st + i));
                float4 v0 = s4[0], v1 = s4[1], v2 = s4[2], v3 = s4[3], v4 = s4[4];
                float4 v5 = make_float4(A.Cc.jn[st + i], A.Cc.jt1[st + i], A.Cc.jt2[st + i], A.Cc.jp[st + i]);
                float4 *d = o + 6 * i;
                d[0] = v0; d[1] = v1; d[2] = v2; d[3] = v3; d[4] = v4; d[5] = v5;
            }
        } else {
            int *r = T.rec + 8 * (size_t)q;
            r[0] = p; r[1] = a; r[2] = b;
            r[3] = a;
            r[4] = b;
            r[5] = st; r[6] = cnt;
            r[7] = (dA ? 1 : 0) | (dB ? 2 : 0);
        }
    }
    __syncthreads();
    if (tid == 0 && nb >= 1000) T.dbg[3] = wall_clock64();
    const int nc = maxc;
    const float t2d = FIXED ? 1.0f : A.t2d;
    auto sweep = [&](int pass) {
        for (int c = 0; c < nc; ++c) {
            for (int k = cst[c] + tid; k < cst[c + 1]; k += T_BIG_WG) {
                int q = p0 + k;
                const int *r = T.rec + 8 * (size_t)q;
                unsigned wd = lds ? pword[k] : 0u;
                if (lds) {
                    if (pass == 0) big_pair<0, FIXED, true>(A, T, wd, r, q, b0, bv, bw, t2d);
                    else if (pass == 1) big_pair<1, FIXED, true>(A, T, wd, r, q, b0, bv, bw, t2d);
                    else big_pair<2, FIXED, true>(A, T, wd, r, q, b0, bpv, bpo, t2d);
                } else {
                    if (pass == 0) big_pair<0, FIXED, false>(A, T, wd, r, q, b0, bv, bw, t2d);
                    else if (pass == 1) big_pair<1, FIXED, false>(A, T, wd, r, q, b0, bv, bw, t2d);
                    else big_pair<2, FIXED, false>(A, T, wd, r, q, b0, bpv, bpo, t2d);
                }
            }
            // LDS mode: only LDS state crosses threads (each record is read and written by one thread),
            // so the colour step ends with an LDS-only barrier and leaves impulse stores in flight
            if (lds) lds_barrier();
            else __syncthreads();
        }
    };
    if (np > 0) {
        if (FIXED) sweep(0);
        for (int it = 0; it < A.vit; ++it) sweep(1);
    }
    __syncthreads();
    if (tid == 0 && nb >= 1000) T.dbg[4] = wall_clock64();
    // velocities back to global memory; the LDS region now holds the pseudo-velocities
    if (lds)
        for (int j = tid; j < nb; j += T_BIG_WG) {
            int i = T.bsorted[b0 + j];
            st3(A.b.v, i, ld3(bv, j));
            st3(A.b.w, i, ld3(bw, j));
        }
    __syncthreads();
    if (lds)
        for (int j = tid; j < nb; j += T_BIG_WG) {
            st3(bpv, j, mk(0.0f, 0.0f, 0.0f));
            st3(bpo, j, mk(0.0f, 0.0f, 0.0f));
        }
    __syncthreads();
    if (np > 0)
        for (int it = 0; it < A.pit; ++it) sweep(2);
    __syncthreads();
    if (tid == 0 && nb >= 1000) T.dbg[5] = wall_clock64();
    if (lds)  // impulses back to the contact arrays
        for (int k = tid; k < np; k += T_BIG_WG) {
            int q = p0 + k;
            int st = T.srec[4 * (size_t)q + 3];
            int cnt = (int)((pword[k] >> 26) & 3) + 1;
            const float *o = T.sprep + 96 * (size_t)q;
            for (int i = 0; i < cnt; ++i) {
                A.Cc.jn[st + i] = o[24 * i + 20];
                A.Cc.jt1[st + i] = o[24 * i + 21];
                A.Cc.jt2[st + i] = o[24 * i + 22];
                A.Cc.jp[st + i] = o[24 * i + 23];
            }
        }
    for (int j = tid; j < nb; j += T_BIG_WG) {
        int i = T.bsorted[b0 + j];
        f3 pv = lds ? ld3(bpv, j) : ld3(A.pv, i), po = lds ? ld3(bpo, j) : ld3(A.po, i);
        if (lds) { st3(A.pv, i, pv); st3(A.po, i, po); }
        if (A.b.invM[i] > 0.0f) integrate_body(A.b.x, A.b.q, ld3(A.b.v, i), ld3(A.b.w, i), pv, po, i, A.dt);
    }
    __syncthreads();
    if (tid == 0 && nb >= 1000) T.dbg[6] = wall_clock64();
}
