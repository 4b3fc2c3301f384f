// Big tiles (more pairs than one small tile holds): one 1024-thread workgroup per tile.
//
// LDS mode (tile within T_BIG_B bodies and T_BIG_P pairs): the tile's pairs are counting-sorted by
// colour; each sorted pair keeps a 32-bit word in LDS (body slots, contact count, flags), its contact
// constants as a lane-coalesced record in global memory in sorted order (T_REC4 float4 fields), and
// its impulses beside (simp), so a colour step is one LDS read plus one round of 16-byte vector loads
// per thread. Body velocities (velocity passes) and then pseudo-velocities (position passes) sit in
// LDS.
// Larger tiles (giant islands): k_t_giant, the same sweeps by a grid of workgroups with body state in
// global memory.
#pragma once

#define T_BIG_P 8192  // pairs of an LDS-mode big tile (32-bit word each)

// pair word: la (13 bits) | lb (13 bits) << 13 | (cnt - 1) (2 bits) << 26 | dynA << 28 | dynB << 29 | hasB << 30
__device__ __forceinline__ unsigned pw_make(int la, int lb, int cnt, bool dA, bool dB, bool hB) {
    return (unsigned)(la & 8191) | ((unsigned)(lb & 8191) << 13) | ((unsigned)(cnt - 1) << 26) |
           ((unsigned)dA << 28) | ((unsigned)dB << 29) | ((unsigned)hB << 30);
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

// a giant tile's pair record (8 ints: p, a, b, a, b, first contact, count, dynamic flags) as loaded
struct GRec {
    int a, b, st, cnt, fl;
};
__device__ __forceinline__ GRec grec_load(const int *r) {
    int4 x = *reinterpret_cast<const int4 *>(r), y = *reinterpret_cast<const int4 *>(r + 4);
    return GRec{x.y, x.z, y.y, y.z, y.w};
}

// A colour-sorted pair's record: a header of 10 float4 -- both sides' inverse mass (and two pad
// floats), inverse world inertia (3 float4 each, row-major with 3 pad), and the velocity and angular
// velocity of a static (not dynamic) side (3 float4: vA, wA, vB, wB) -- then per contact 5 float4 of
// k_t_prep's constants, regrouped so that the warm and position passes read only the first three:
// (n, meff), (rA, bias), (rB, mu), (t1, meft1), (t2, meft2). The records are lane-coalesced in blocks
// of T_RB consecutive sorted pairs: float4 field j of sorted pair q sits at
// sprep[(q / T_RB) * T_REC4 * T_RB + j * T_RB + q % T_RB], so the 64 lanes of a wave, which hold
// consecutive sorted pairs of a colour, read whole 128-byte lines per field (with one record after
// another, each field load touched 64 cache lines), and every field is an immediate offset (j * 128
// bytes) from one per-lane address (a field-major layout over all pairs kept 30 64-bit field
// addresses live, more than the scalar registers hold). The record is written once per substep and
// only read by the sweeps; the impulses (jn, jt1, jt2, jp), which every colour step rewrites, sit apart
// in simp, contact k of pair q at simp[(q / T_RB) * 4 * T_RB + k * T_RB + q % T_RB]. Everything a colour
// step reads from global memory sits at addresses known before the pair word is read, so the step
// issues all of its loads at once.
#define T_RB 8
__device__ __forceinline__ size_t rec_off(int q) { return (size_t)(q / T_RB) * (T_REC4 * T_RB) + q % T_RB; }
__device__ __forceinline__ size_t imp_off(int q) { return (size_t)(q / T_RB) * (4 * T_RB) + q % T_RB; }

// contact k of a sorted pair from its record (Rq = sprep + rec_off(q)) and impulses (Iq = simp +
// imp_off(q)); the friction frame and its effective masses only for the velocity pass
template <int PASS>
__device__ __forceinline__ void ld_srec(const float4 *Rq, const float4 *Iq, int k, CCon &C) {
    constexpr int S = T_RB;
    const float4 *c = Rq + (T_REC_HDR4 + 5 * k) * S;
    const float4 a = c[0], b = c[S], d = c[2 * S];
    C.n = mk(a.x, a.y, a.z); C.meff = a.w;
    C.rA = mk(b.x, b.y, b.z); C.bias = b.w;
    C.rB = mk(d.x, d.y, d.z); C.mu = d.w;
    if (PASS == 1) {
        const float4 e = c[3 * S], f = c[4 * S];
        C.t1 = mk(e.x, e.y, e.z); C.meft1 = e.w;
        C.t2 = mk(f.x, f.y, f.z); C.meft2 = f.w;
    } else {
        C.t1 = C.t2 = mk(0.0f, 0.0f, 0.0f);
        C.meft1 = C.meft2 = 0.0f;
    }
    const float4 i = Iq[k * S];
    C.jn = i.x; C.jt1 = i.y; C.jt2 = i.z; C.jp = i.w;
}

__device__ __forceinline__ m3 ld_m12s(const float4 *o) {
    constexpr int S = T_RB;
    const float4 a = o[0], b = o[S], c = o[2 * S];
    m3 M;
    M.m[0][0] = a.x; M.m[0][1] = a.y; M.m[0][2] = a.z; M.m[1][0] = a.w; M.m[1][1] = b.x; M.m[1][2] = b.y;
    M.m[2][0] = b.z; M.m[2][1] = b.w; M.m[2][2] = c.x;
    return M;
}

// One pair of a big tile for one colour step, contacts streamed one at a time (few live registers).
// PASS 0 warm apply, 1 velocity, 2 position. LDS mode: body state s0/s1 in LDS by tile slot, the
// pair word and its colour-ordered records; global mode: record g, body state in global memory by
// body id, constants from prep, impulses from the contact arrays. FAST: the fast-arithmetic contact
// updates (cc_*_c), with the same records (k_t_prep then stores the reciprocal effective masses).
template <int PASS, bool FIXED, bool LDSM, bool FAST = false>
__device__ __forceinline__ void big_pair(const TiledArgs &A, const TiledBuffers &T, unsigned w, GRec g, int q,
                                         int b0, float *s0, float *s1, float t2d) {
    int la, lb, cnt, a = -1, b = -1, st = 0;
    bool dynA, dynB, hasB;
    if (LDSM) {
        la = w & 8191; lb = (w >> 13) & 8191; cnt = (int)((w >> 26) & 3) + 1;
        dynA = (w >> 28) & 1; dynB = (w >> 29) & 1; hasB = (w >> 30) & 1;
    } else {
        a = g.a; b = g.b; la = a; lb = b; st = g.st; cnt = g.cnt;
        dynA = g.fl & 1; dynB = (g.fl & 2) != 0; hasB = b >= 0;
    }
    const float *gs0 = PASS == 2 ? A.pv : A.b.v, *gs1 = PASS == 2 ? A.po : A.b.w;
    f3 z = mk(0.0f, 0.0f, 0.0f);
    float iMA, iMB = 0.0f;
    m3 IA, IB;
    f3 xA, yA, xB = z, yB = z;
    constexpr int S = T_RB;
    const float4 *Rq = LDSM ? reinterpret_cast<const float4 *>(T.sprep) + rec_off(q) : nullptr;
    float4 *Iq = LDSM ? reinterpret_cast<float4 *>(T.simp) + imp_off(q) : nullptr;
    if (LDSM) {
        // the record header: masses, inertias, a static side's velocities (its pseudo-velocities are
        // zero: they are cleared every substep and a static body takes no impulse)
        const float4 *H = Rq;
        const float4 h0 = H[0];
        iMA = h0.x;
        iMB = h0.y;
        IA = ld_m12s(H + S);
        IB = ld_m12s(H + 4 * S);
        float4 h7 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), h8 = h7, h9 = h7;
        if (PASS != 2) { h7 = H[7 * S]; h8 = H[8 * S]; h9 = H[9 * S]; }
        if (dynA) {
            xA = ld3(s0, la); yA = ld3(s1, la);
        } else if (PASS == 2) {
            xA = z; yA = z;
        } else {
            xA = mk(h7.x, h7.y, h7.z); yA = mk(h7.w, h8.x, h8.y);
        }
        if (dynB) {
            xB = ld3(s0, lb); yB = ld3(s1, lb);
        } else if (hasB && PASS != 2) {
            xB = mk(h8.z, h8.w, h9.x); yB = mk(h9.y, h9.z, h9.w);
        }
    } else {
        iMA = A.b.invM[a];
        IA = ldm(A.invIw, a);
        xA = dynA ? ld3(s0, a) : ld3(gs0, a);
        yA = dynA ? ld3(s1, a) : ld3(gs1, a);
        IB = IA;
        if (hasB) {
            iMB = A.b.invM[b];
            IB = ldm(A.invIw, b);
            xB = dynB ? ld3(s0, b) : ld3(gs0, b);
            yB = dynB ? ld3(s1, b) : ld3(gs1, b);
        }
    }
    auto solve = [&](int k, CCon &C) {
        if (PASS == 0) {
            if (FAST) cc_warm_c(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB);
            else cc_warm(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB);
        } else {
            if (PASS == 1 && FAST) cc_vel_c(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB);
            else if (PASS == 1) cc_vel<FIXED>(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB, t2d);
            else if (FAST) cc_pos_c(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB);
            else cc_pos(C, xA, yA, xB, yB, hasB, iMA, iMB, IA, IB);
            if (LDSM) {
                Iq[k * S] = make_float4(C.jn, C.jt1, C.jt2, C.jp);
            } else {
                int c = st + k;
                A.Cc.jn[c] = C.jn; A.Cc.jt1[c] = C.jt1; A.Cc.jt2[c] = C.jt2; A.Cc.jp[c] = C.jp;
            }
        }
    };
    // software pipeline: contact k+1 is loaded before contact k is solved and stored (loading every
    // contact at once, in 512-thread workgroups for the registers, measured slower)
    auto load = [&](int k, CCon &C) {
        if (LDSM) {
            ld_srec<PASS>(Rq, Iq, k, C);
        } else {
            int c = st + k;
            ld_prep20(T.prep + 20 * (size_t)c, C);
            C.jn = A.Cc.jn[c]; C.jt1 = A.Cc.jt1[c]; C.jt2 = A.Cc.jt2[c]; C.jp = A.Cc.jp[c];
        }
    };
    // four named contacts rather than one rotated buffer: no register copies between them
    CCon c0, c1, c2, c3;
    load(0, c0);
    if (cnt > 1) load(1, c1);
    solve(0, c0);
    if (cnt > 2) load(2, c2);
    if (cnt > 1) solve(1, c1);
    if (cnt > 3) load(3, c3);
    if (cnt > 2) solve(2, c2);
    if (cnt > 3) solve(3, c3);
    if (iMA > 0.0f) { st3(s0, la, xA); st3(s1, la, yA); }
    if (hasB && iMB > 0.0f) { st3(s0, lb, xB); st3(s1, lb, yB); }
}

#define T_BIG_WG 1024   // threads of a giant-grid workgroup
#define T_BIGT_WG 1024  // threads of a k_t_big workgroup
#define T_BIG_DYN (6 * sizeof(float) * T_BIG_B + sizeof(unsigned) * T_BIG_P)

// `reps` sweeps of one pass over an LDS-mode big tile, contacts streamed (big_pair). A colour step
// of a big island is bound by the latency of one pair's dependent chain: a colour of up to 256 pairs
// leaves one wave on each SIMD, with nothing to hide the chain behind. Measured with cycle stamps on
// the settled 10 000-box pile (2.25 GHz): a velocity step takes about 10 000 cycles, of which about
// 1 800 wait for the pair's first records, and a position step about 5 000 (900). Prefetching the next
// pair across colour steps (to keep loads clear of the step's impulse stores, which gfx9 counts in
// vmcnt) costs more instructions than the latency it hides.
template <int PASS, bool FIXED, bool FAST>
__device__ __forceinline__ void big_sweeps_stream(const TiledArgs &A, const TiledBuffers &T, const int *cst,
                                                  const unsigned *pword, int nc, int reps, int p0, int b0, float *s0,
                                                  float *s1, float t2d) {
    const int tid = threadIdx.x;
    for (int it = 0; it < reps; ++it)
        for (int c = 0; c < nc; ++c) {
            for (int k = cst[c] + tid; k < cst[c + 1]; k += T_BIGT_WG)
                big_pair<PASS, FIXED, true, FAST>(A, T, pword[k], GRec{}, p0 + k, b0, s0, s1, t2d);
            lds_barrier();
        }
}

// The colour order and pair records of the big tiles (LDS mode), built before k_t_big by kernels that
// spread over the GPU, so that the one workgroup that sweeps a large island starts its sweeps at once.
// k_t_bigord: one workgroup per big tile counting-sorts its pairs by colour class (pairs with two
// dynamic bodies first within a class, one-sided pairs last, so most waves skip one side's updates as
// a whole; the order within a class is free, its pairs sharing no dynamic body) and stores each sorted
// pair's word, class and (p, a, b, first contact). k_t_bigfill: one thread per sorted pair writes its
// record (lane-coalesced blocks, above) and impulses.
__device__ __forceinline__ bool lds_big_tile(const TiledBuffers &T, int t, int &b0, int &nb, int &p0, int &np) {
    if (T.tb0[t] == INT_MAX) return false;
    tile_bounds(T, t, b0, nb, p0, np);
    return !is_small_tile(nb, np) && nb <= T_BIG_B && np <= T_BIG_P;
}

#define T_ORD_PER (T_BIG_P / 1024)  // pairs per thread of k_t_bigord (kept in registers between its passes)
__global__ void __launch_bounds__(1024) k_t_bigord(TiledArgs A, TiledBuffers T) {
    __shared__ int cst[ME_NCOLMAX + 1], cfl[ME_NCOLMAX + 1], cfe[ME_NCOLMAX + 1];
    __shared__ int maxc, wdyn, wcon;
    const int tid = threadIdx.x, nbig = A.istat[IS_NBIG];
    int ncol = A.istat[IS_NCOL];
    if (ncol > ME_NCOLMAX) ncol = ME_NCOLMAX;
    const float *invM = A.b.invM;
    for (int i = blockIdx.x; i < nbig; i += gridDim.x) {
        int b0, nb, p0, np;
        if (!lds_big_tile(T, T.bigq[i], b0, nb, p0, np)) continue;
        if (tid == 0) { maxc = 0; wdyn = 0; wcon = 0; }
        for (int c = tid; c <= ncol; c += 1024) { cst[c] = 0; cfl[c] = 0; cfe[c] = 0; }
        __syncthreads();
        // pass 1: each thread's pairs (k = tid + 1024 j), their class, word and record entries
        int pc[T_ORD_PER], pp[T_ORD_PER], pa[T_ORD_PER], pb[T_ORD_PER], pst[T_ORD_PER];
        unsigned pw[T_ORD_PER];
        int mdyn = 0, mcon = 0;
#pragma unroll
        for (int j = 0; j < T_ORD_PER; ++j) {
            const int k = tid + 1024 * j;
            pc[j] = -1;
            if (k >= np) continue;
            const int p = T.psorted[p0 + k];
            int c = A.Pp.col[p];
            if (c < 0 || c >= ncol) c = ncol > 0 ? ncol - 1 : 0;
            c = tile_class(T, c, ncol);
            const int a = A.Pp.pbi[p], b = A.Pp.pbj[p];
            const bool dA = invM[a] > 0.0f, dB = b >= 0 && invM[b] > 0.0f;
            const int cnt = A.Pp.pcount[p];
            const int la = dA ? T.bslot[a] - b0 : 0, lb = dB ? T.bslot[b] - b0 : 0;
            pc[j] = c; pp[j] = p; pa[j] = a; pb[j] = b; pst[j] = A.Pp.pstart[p];
            pw[j] = pw_make(la, lb, cnt, dA, dB, b >= 0);
            mdyn += (int)dA + (int)dB;
            mcon += cnt;
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
        // pass 2: the sorted places
#pragma unroll
        for (int j = 0; j < T_ORD_PER; ++j) {
            if (pc[j] < 0) continue;
            const int c = pc[j];
            const bool two = ((pw[j] >> 28) & 3u) == 3u;  // both bodies dynamic
            const int pos = two ? cst[c] + atomicAdd(&cfl[c], 1) : cst[c + 1] - 1 - atomicAdd(&cfe[c], 1);
            const int q = p0 + pos;
            T.pwordg[q] = pw[j];
            T.pclass[q] = c;
            *reinterpret_cast<int4 *>(T.srec + 4 * (size_t)q) = make_int4(pp[j], pa[j], pb[j], pst[j]);
        }
        if (mdyn) atomicAdd(&wdyn, mdyn);
        if (mcon) atomicAdd(&wcon, mcon);
        __syncthreads();
        if (tid == 0) {  // work of this tile for the algorithmic-byte count of k_t_big
            atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[0]), 1ull);
            atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[1]), (unsigned long long)np);
            atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[2]), (unsigned long long)wdyn);
            atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[3]), (unsigned long long)wcon);
        }
        __syncthreads();  // the next tile reuses the tables
    }
}

#define T_FILL_WG 256
#define T_FILL_CHUNKS (T_BIG_P / T_FILL_WG)
// grid (x, T_FILL_CHUNKS): workgroup (x, y) fills sorted pairs y * T_FILL_WG .. of big tiles x, x + gridDim.x, ...
__global__ void __launch_bounds__(T_FILL_WG) k_t_bigfill(TiledArgs A, TiledBuffers T) {
    const int nbig = A.istat[IS_NBIG];
    const float *invM = A.b.invM;
    for (int i = blockIdx.x; i < nbig; i += gridDim.x) {
        int b0, nb, p0, np;
        if (!lds_big_tile(T, T.bigq[i], b0, nb, p0, np)) continue;
        const int k = blockIdx.y * T_FILL_WG + threadIdx.x;
        if (k >= np) continue;
        const int q = p0 + k;
        const int4 sr = *reinterpret_cast<const int4 *>(T.srec + 4 * (size_t)q);
        const int a = sr.y, b = sr.z, st = sr.w;
        const unsigned w = T.pwordg[q];
        const int la = w & 8191, lb = (w >> 13) & 8191, cnt = (int)((w >> 26) & 3) + 1;
        const bool dA = (w >> 28) & 1, dB = (w >> 29) & 1;
        constexpr int S = T_RB;
        float4 *hd = reinterpret_cast<float4 *>(T.sprep) + rec_off(q);
        {
            float sA[12], sB[12];
            const float iA = dA ? T.invMs[b0 + la] : invM[a];
            const float iB = dB ? T.invMs[b0 + lb] : (b >= 0 ? invM[b] : 0.0f);
            for (int r = 0; r < 9; ++r) {
                sA[r] = dA ? T.invIws[12 * (size_t)(b0 + la) + r] : A.invIw[9 * (size_t)a + r];
                sB[r] = dB ? T.invIws[12 * (size_t)(b0 + lb) + r] : (b >= 0 ? A.invIw[9 * (size_t)b + r] : sA[r]);
            }
            sA[9] = sA[10] = sA[11] = 0.0f;
            sB[9] = sB[10] = sB[11] = 0.0f;
            f3 z3 = mk(0.0f, 0.0f, 0.0f);
            f3 vA = dA ? z3 : ld3(A.b.v, a), wA = dA ? z3 : ld3(A.b.w, a);
            f3 vB = (dB || b < 0) ? z3 : ld3(A.b.v, b), wB = (dB || b < 0) ? z3 : ld3(A.b.w, b);
            hd[0] = make_float4(iA, iB, 0.0f, 0.0f);
            hd[S] = make_float4(sA[0], sA[1], sA[2], sA[3]);
            hd[2 * S] = make_float4(sA[4], sA[5], sA[6], sA[7]);
            hd[3 * S] = make_float4(sA[8], sA[9], sA[10], sA[11]);
            hd[4 * S] = make_float4(sB[0], sB[1], sB[2], sB[3]);
            hd[5 * S] = make_float4(sB[4], sB[5], sB[6], sB[7]);
            hd[6 * S] = make_float4(sB[8], sB[9], sB[10], sB[11]);
            hd[7 * S] = make_float4(vA.x, vA.y, vA.z, wA.x);
            hd[8 * S] = make_float4(wA.y, wA.z, vB.x, vB.y);
            hd[9 * S] = make_float4(vB.z, wB.x, wB.y, wB.z);
        }
        float4 *imp = reinterpret_cast<float4 *>(T.simp) + imp_off(q);
        for (int c = 0; c < cnt; ++c) {
            // prep: n, rA, rB, t1, t2, meff, meft1, meft2, bias, mu (k_t_prep)
            const float4 *s4 = reinterpret_cast<const float4 *>(T.prep + 20 * (size_t)(st + c));
            const float4 v0 = s4[0], v1 = s4[1], v2 = s4[2], v3 = s4[3], v4 = s4[4];
            float4 *d = hd + (T_REC_HDR4 + 5 * c) * S;
            d[0] = make_float4(v0.x, v0.y, v0.z, v3.w);      // n, meff
            d[S] = make_float4(v0.w, v1.x, v1.y, v4.z);      // rA, bias
            d[2 * S] = make_float4(v1.z, v1.w, v2.x, v4.w);  // rB, mu
            d[3 * S] = make_float4(v2.y, v2.z, v2.w, v4.x);  // t1, meft1
            d[4 * S] = make_float4(v3.x, v3.y, v3.z, v4.y);  // t2, meft2
            imp[c * S] = make_float4(A.Cc.jn[st + c], A.Cc.jt1[st + c], A.Cc.jt2[st + c], A.Cc.jp[st + c]);
        }
    }
}

// one big tile t by the whole workgroup (LDS: dsm dynamic, the colour tables and scalars static), its
// colour order and records from k_t_bigord / k_t_bigfill; tiles beyond T_BIG_B bodies or T_BIG_P
// pairs are left to k_t_giant
template <int SOLVER, bool FAST>
__device__ __forceinline__ void big_tile(const TiledArgs &A, const TiledBuffers &T, int t, float *dsm, int *cst,
                                         int &maxc) {
    constexpr bool FIXED = SOLVER == ME_SOLVER_COLOR_CACHE;
    int b0, nb, p0, np;
    if (!lds_big_tile(T, t, b0, nb, p0, np)) return;
    if (A.taped && threadIdx.x == 0) atomicOr(&A.istat[IS_OVF], OVF_TAPE);  // no recording for big tiles
    const int tid = threadIdx.x;
    // phase stamps of the largest big tile (>= 1000 bodies) of the launch
    if (tid == 0 && nb >= 1000 && atomicMax(&T.dbg[7], (long long)nb) <= nb) { T.dbg[0] = wall_clock64(); T.dbg[8] = np; }
    auto stamp = [&](int k) {
        if (tid == 0 && nb >= 1000 && (long long)atomicAdd(reinterpret_cast<unsigned long long *>(&T.dbg[7]), 0ull) == nb) T.dbg[k] = wall_clock64();
    };
    unsigned *pword = reinterpret_cast<unsigned *>(dsm + 6 * T_BIG_B);
    if (tid == 0) {
        atomicAdd(&A.istat[IS_TILES], 1);
        atomicMax(&A.istat[IS_MAXTB], nb);
    }
    int ncol = A.istat[IS_NCOL];
    if (ncol > ME_NCOLMAX) ncol = ME_NCOLMAX;
    // body state (velocities, then pseudo-velocities) by tile slot; kept as LDS pointers so every
    // access compiles to ds_* instructions rather than generic (flat) ones
    float *lv = dsm, *lw = dsm + 3 * T_BIG_B;
    for (int j = tid; j < nb; j += T_BIGT_WG) {
        int i = T.bsorted[b0 + j];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lv[3 * j + k] = A.b.v[3 * i + k];
            lw[3 * j + k] = A.b.w[3 * i + k];
        }
    }
    for (int c = tid; c <= ncol; c += T_BIGT_WG) cst[c] = -1;
    __syncthreads();
    stamp(1);
    // the pair words, and the class starts from where the sorted classes change
    for (int k = tid; k < np; k += T_BIGT_WG) {
        pword[k] = T.pwordg[p0 + k];
        const int c = T.pclass[p0 + k];
        if (k == 0 || T.pclass[p0 + k - 1] != c) cst[c] = k;
        if (k == np - 1) maxc = c + 1;
    }
    __syncthreads();
    if (tid == 0) {  // empty classes start where the next one does
        if (np == 0) maxc = 0;
        cst[maxc] = np;
        for (int c = maxc - 1; c >= 0; --c)
            if (cst[c] < 0) cst[c] = cst[c + 1];
    }
    __syncthreads();
    stamp(2);
    stamp(3);
    const int nc = maxc;
    const float t2d = FIXED ? 1.0f : A.t2d;
    if (np > 0) {
        if (FIXED) big_sweeps_stream<0, FIXED, FAST>(A, T, cst, pword, nc, 1, p0, b0, lv, lw, t2d);
        big_sweeps_stream<1, FIXED, FAST>(A, T, cst, pword, nc, A.vit, p0, b0, lv, lw, t2d);
    }
    __syncthreads();
    stamp(4);
    // velocities back to global memory; the LDS region now holds the pseudo-velocities
    for (int j = tid; j < nb; j += T_BIGT_WG) {
        int i = T.bsorted[b0 + j];
        st3(A.b.v, i, ld3(lv, j));
        st3(A.b.w, i, ld3(lw, j));
    }
    __syncthreads();
    for (int j = tid; j < nb; j += T_BIGT_WG) {
        st3(lv, j, mk(0.0f, 0.0f, 0.0f));
        st3(lw, j, mk(0.0f, 0.0f, 0.0f));
    }
    __syncthreads();
    if (np > 0) big_sweeps_stream<2, FIXED, FAST>(A, T, cst, pword, nc, A.pit, p0, b0, lv, lw, t2d);
    __syncthreads();
    stamp(5);
    for (int k = tid; k < np; k += T_BIGT_WG) {  // impulses back to the contact arrays
        int q = p0 + k;
        int st = T.srec[4 * (size_t)q + 3];
        int cnt = (int)((pword[k] >> 26) & 3) + 1;
        const float4 *o = reinterpret_cast<const float4 *>(T.simp) + imp_off(q);
        for (int i = 0; i < cnt; ++i) {
            const float4 j = o[i * T_RB];
            A.Cc.jn[st + i] = j.x;
            A.Cc.jt1[st + i] = j.y;
            A.Cc.jt2[st + i] = j.z;
            A.Cc.jp[st + i] = j.w;
        }
    }
    for (int j = tid; j < nb; j += T_BIGT_WG) {
        int i = T.bsorted[b0 + j];
        f3 pv = ld3(lv, j), po = ld3(lw, j);
        st3(A.pv, i, pv);
        st3(A.po, i, po);
        if (A.b.invM[i] > 0.0f) integrate_body(A.b.x, A.b.q, ld3(A.b.v, i), ld3(A.b.w, i), pv, po, i, A.dt);
    }
    __syncthreads();
    stamp(6);
}

// the big tiles, taken from the largest-first queue (k_t_bigsort) by workgroups that each take the
// next entry when done, one workgroup per CU (a launch over every tile would schedule one 152 KB-LDS
// workgroup per tile, mostly to find a small tile and exit). Which workgroup solves a tile does not
// change its result.
#define T_BIG_GRID_MAX 256

template <int SOLVER, bool FAST = false>
__global__ void __launch_bounds__(T_BIGT_WG) k_t_big(TiledArgs A, TiledBuffers T) {
    extern __shared__ float dsm[];
    __shared__ int cst[ME_NCOLMAX + 1];
    __shared__ int maxc, qi;
    const int nbig = A.istat[IS_NBIG];
    for (;;) {
        if (threadIdx.x == 0) qi = atomicAdd(&A.istat[IS_BIGQ], 1);
        __syncthreads();
        const int i = qi;
        __syncthreads();
        if (i >= nbig) break;
        big_tile<SOLVER, FAST>(A, T, T.bigq[i], dsm, cst, maxc);
    }
}

// ── giant tiles: an island beyond the LDS limits of one workgroup, swept by a grid of workgroups ──
// Each colour step spreads its pairs over every workgroup of the grid; a grid-wide barrier separates
// the colour steps (a colour's pairs share no dynamic body, and the next colour reads what it wrote).
// Body state, impulses and pair records live in global memory, as in big_tile's global mode; the
// per-pair work (big_pair) and the colour order are the same, so the result is too. Workgroup 0
// builds the colour classes and records first.
#ifndef T_GIANT_GRID
#define T_GIANT_GRID 64
#endif

// Grid barrier. `target` = (barriers so far in this launch) x G arrivals, the count being cleared
// before the launch (by k_t_biglist, every substep). Every wave first waits for its own stores to be acknowledged by its XCD's L2
// (vmcnt 0); after the workgroup barrier, thread 0 writes that L2 back (agent-scope release), arrives,
// waits, and invalidates the CU's L1 and the L2's copies of other XCDs' lines (agent-scope acquire)
// before the workgroup reads anything. One release and one acquire per workgroup: a fence in every
// wave costs about 8x more per barrier (32 workgroups: 12 us against 1.5 us).
// The host sizes the grid so that all of its workgroups can be resident at once on an idle GPU
// (giant_grid_size); a workgroup that is not yet resident only delays the barrier while the kernels
// beside it on the other stream run to completion on their own. The wait is still bounded, by the
// wall clock (`wait_ticks` of the constant-rate counter, about 100 ms): past it the launch flags
// OVF_GIANT and winds down, and the engine's step raises.
__device__ __forceinline__ bool giant_barrier(unsigned *count, unsigned target, int *istat, int &stop,
                                              long long wait_ticks) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // gfx9 encoding: vmcnt 0, expcnt and lgkmcnt not waited
    __syncthreads();
    if (threadIdx.x == 0 && !stop) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const long long t0 = wall_clock64();
        while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (wall_clock64() - t0 > wait_ticks) {
                atomicOr(&istat[IS_OVF], OVF_GIANT);
                stop = 1;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    return !stop;
}

// JP colouring of the tiles beyond the LDS colouring limits by a grid of workgroups: the rounds of
// t_color_tile (claims by body id in the global claim array) with a grid barrier after each of a
// round's three passes (reset, claim, win). The colours are those of the one-workgroup rounds: a
// round's claims are the minimum key over each body's uncoloured pairs, whatever the order.
// Uncoloured pairs per round are counted in gbar[2 + round % 3]: the counter of round r + 1 is
// cleared in round r, after every workgroup has read that of round r - 2.
__global__ void __launch_bounds__(T_BIG_WG) k_t_color_giant(TiledArgs A, TiledBuffers T) {
    if (A.istat[IS_MISMATCH] != 1) return;
    __shared__ int stop, wleft, hist[ME_NCOLMAX];
    const int tid = threadIdx.x, G = gridDim.x, gt = blockIdx.x * T_BIG_WG + tid, stride = G * T_BIG_WG;
    if (tid == 0) stop = 0;
    __syncthreads();
    unsigned target = 0;
    unsigned *gb = T.gbar + 1, *gleft = T.gbar + 2;
    const float *invM = A.b.invM;
    int *gcol = A.Pp.col;
    unsigned long long *claim = A.claim;
    const int ng = A.istat[IS_NGCOL];
    for (int i = 0; i < ng; ++i) {
        int b0, nb, p0, np;
        tile_bounds(T, T.gcollist[i], b0, nb, p0, np);
        for (int k = gt; k < np; k += stride) gcol[T.psorted[p0 + k]] = -1;
        int nround = 0;
        for (int cur = 0; cur <= np; ++cur) {
            for (int j = gt; j < nb; j += stride) claim[T.bsorted[b0 + j]] = ~0ull;
            if (gt == 0) __hip_atomic_store(&gleft[(cur + 1) % 3], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (tid == 0) wleft = 0;
            target += G;
            if (!giant_barrier(gb, target, A.istat, stop, A.giant_wait_ticks)) return;
            for (int k = gt; k < np; k += stride) {
                int p = T.psorted[p0 + k];
                if (gcol[p] >= 0) continue;
                unsigned long long key = (unsigned long long)jp_key(p);
                int a = A.Pp.pbi[p], b = A.Pp.pbj[p];
                if (invM[a] > 0.0f) atomicMin(&claim[a], key);
                if (b >= 0 && invM[b] > 0.0f) atomicMin(&claim[b], key);
            }
            target += G;
            if (!giant_barrier(gb, target, A.istat, stop, A.giant_wait_ticks)) return;
            int nl = 0;
            for (int k = gt; k < np; k += stride) {
                int p = T.psorted[p0 + k];
                if (gcol[p] >= 0) continue;
                unsigned long long key = (unsigned long long)jp_key(p);
                int a = A.Pp.pbi[p], b = A.Pp.pbj[p];
                bool ok = true;
                if (invM[a] > 0.0f && claim[a] != key) ok = false;
                if (b >= 0 && invM[b] > 0.0f && claim[b] != key) ok = false;
                if (ok) gcol[p] = cur;
                else ++nl;
            }
            if (nl) atomicAdd(&wleft, nl);
            __syncthreads();
            if (tid == 0 && wleft) atomicAdd(&gleft[cur % 3], (unsigned)wleft);
            target += G;
            if (!giant_barrier(gb, target, A.istat, stop, A.giant_wait_ticks)) return;
            if (__hip_atomic_load(&gleft[cur % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                nround = cur + 1;
                break;
            }
        }
        target += G;  // every workgroup has read the last round's counter
        if (!giant_barrier(gb, target, A.istat, stop, A.giant_wait_ticks)) return;
        if (gt < 3) __hip_atomic_store(&gleft[gt], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // class sizes: a histogram in LDS, then one global add per class and workgroup
        for (int c = tid; c < nround && c < ME_NCOLMAX; c += T_BIG_WG) hist[c] = 0;
        __syncthreads();
        bool ovf = false;
        for (int k = gt; k < np; k += stride) {
            int c = gcol[T.psorted[p0 + k]];
            if (c >= 0 && c < ME_NCOLMAX) atomicAdd(&hist[c], 1);
            else ovf = true;
        }
        __syncthreads();
        for (int c = tid; c < nround && c < ME_NCOLMAX; c += T_BIG_WG)
            if (hist[c]) atomicAdd(&T.ccount_jp[c], hist[c]);
        if (ovf) atomicOr(&A.istat[IS_OVF], OVF_COLORS);
        if (gt == 0) atomicMax(&A.istat[IS_NCOLJP], nround);
        target += G;  // the next tile's round counters start from zero
        if (!giant_barrier(gb, target, A.istat, stop, A.giant_wait_ticks)) return;
    }
}

template <int SOLVER, bool FAST = false>
__global__ void __launch_bounds__(T_BIG_WG) k_t_giant(TiledArgs A, TiledBuffers T) {
    __shared__ int cst[ME_NCOLMAX + 1], cfl[ME_NCOLMAX + 1], cfe[ME_NCOLMAX + 1];
    __shared__ int maxc, wdyn, wcon, stop;
    constexpr bool FIXED = SOLVER == ME_SOLVER_COLOR_CACHE;
    const int tid = threadIdx.x, G = gridDim.x;
    if (tid == 0) stop = 0;
    __syncthreads();
    unsigned target = 0;
    const float t2d = FIXED ? 1.0f : A.t2d;
    const int ng = A.istat[IS_NGIANT];
    for (int i = 0; i < ng; ++i) {
        int b0, nb, p0, np;
        tile_bounds(T, T.giantlist[i], b0, nb, p0, np);
        // colour classes and colour-ordered pair records, as big_tile's global mode, built by the whole
        // grid: each workgroup counts its share of the pairs by class (two dynamic bodies or one) in LDS
        // and reserves its range of each class with one global add, whose old value is the range's offset
        // (one workgroup building them took an eighth of the sweep's time on an island of 130 000 pairs)
        if (blockIdx.x == 0 && tid == 0) {
            if (A.taped) atomicOr(&A.istat[IS_OVF], OVF_TAPE);
            if (atomicMax(&T.dbg[7], (long long)nb) <= nb) { T.dbg[0] = wall_clock64(); T.dbg[8] = np; }
            atomicAdd(&A.istat[IS_TILES], 1);
            atomicMax(&A.istat[IS_MAXTB], nb);
        }
        int ncol = A.istat[IS_NCOL];
        if (ncol > ME_NCOLMAX) ncol = ME_NCOLMAX;
        const float *invM = A.b.invM;
        const int gt = blockIdx.x * T_BIG_WG + tid, stride = G * T_BIG_WG;
        int *gdual = T.gcnt, *gsing = T.gcnt + (ME_NCOLMAX + 1);
        auto pair_class = [&](int p, bool &dA, bool &dB) {
            int c = A.Pp.col[p];
            if (c < 0 || c >= ncol) c = ncol > 0 ? ncol - 1 : 0;
            const int a = A.Pp.pbi[p], b = A.Pp.pbj[p];
            dA = invM[a] > 0.0f;
            dB = b >= 0 && invM[b] > 0.0f;
            return tile_class(T, c, ncol);
        };
        for (int c = tid; c < ncol; c += T_BIG_WG) { cfl[c] = 0; cfe[c] = 0; }
        if (tid == 0) { wdyn = 0; wcon = 0; }
        __syncthreads();
        for (int k = gt; k < np; k += stride) {
            bool dA, dB;
            const int c = pair_class(T.psorted[p0 + k], dA, dB);
            atomicAdd((dA && dB) ? &cfl[c] : &cfe[c], 1);
        }
        __syncthreads();
        for (int c = tid; c < ncol; c += T_BIG_WG) {
            if (cfl[c]) cfl[c] = atomicAdd(&gdual[c], cfl[c]);
            if (cfe[c]) cfe[c] = atomicAdd(&gsing[c], cfe[c]);
        }
        target += G;
        if (!giant_barrier(T.gbar, target, A.istat, stop, A.giant_wait_ticks)) return;
        // the class starts from the totals, by one wave of every workgroup (the same in each)
        if (tid < 64) {
            constexpr int PERL = (ME_NCOLMAX + 63) / 64;
            const int base = tid * PERL;
            int n[PERL], sum = 0, hi = 0;
#pragma unroll
            for (int j = 0; j < PERL; ++j) {
                const int c = base + j;
                n[j] = c < ncol ? __hip_atomic_load(&gdual[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) +
                                      __hip_atomic_load(&gsing[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                : 0;
                sum += n[j];
                if (n[j]) hi = c + 1;
            }
            int incl = sum;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o);
                if (tid >= o) incl += v;
            }
            int run = incl - sum;
#pragma unroll
            for (int j = 0; j < PERL; ++j)
                if (base + j < ncol) { cst[base + j] = run; run += n[j]; }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) hi = max(hi, __shfl_xor(hi, o));
            if (tid == 63) cst[ncol] = incl;  // classes from the highest non-empty one on start at the total
            if (tid == 0) maxc = hi;
        }
        __syncthreads();
        int mdyn = 0, mcon = 0;
        for (int k = gt; k < np; k += stride) {
            const int p = T.psorted[p0 + k];
            bool dA, dB;
            const int c = pair_class(p, dA, dB);
            const int pos = (dA && dB) ? cst[c] + atomicAdd(&cfl[c], 1) : cst[c + 1] - 1 - atomicAdd(&cfe[c], 1);
            int *r = T.rec + 8 * (size_t)(p0 + pos);
            const int a = A.Pp.pbi[p], b = A.Pp.pbj[p], st = A.Pp.pstart[p], cnt = A.Pp.pcount[p];
            mdyn += (int)dA + (int)dB;
            mcon += cnt;
            r[0] = p; r[1] = a; r[2] = b; r[3] = a; r[4] = b; r[5] = st; r[6] = cnt;
            r[7] = (dA ? 1 : 0) | (dB ? 2 : 0);
        }
        if (mdyn) atomicAdd(&wdyn, mdyn);
        if (mcon) atomicAdd(&wcon, mcon);
        __syncthreads();
        if (blockIdx.x == 0)
            for (int c = tid; c <= maxc; c += T_BIG_WG) T.gcst[c] = cst[c];
        if (tid == 0) {
            if (blockIdx.x == 0) {
                T.gcst[ME_NCOLMAX + 1] = maxc;
                atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[8]), 1ull);
                atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[9]), (unsigned long long)np);
            }
            if (wdyn) atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[10]), (unsigned long long)wdyn);
            if (wcon) atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[11]), (unsigned long long)wcon);
        }
        target += G;
        if (!giant_barrier(T.gbar, target, A.istat, stop, A.giant_wait_ticks)) return;
        if (blockIdx.x == 0 && tid == 0 &&
            (long long)atomicAdd(reinterpret_cast<unsigned long long *>(&T.dbg[7]), 0ull) == nb)
            T.dbg[3] = wall_clock64();  // the colour order and records built
        // the class totals start from zero for the next giant tile (its adds follow the tile's last barrier)
        if (blockIdx.x == 0)
            for (int c = tid; c < 2 * (ME_NCOLMAX + 1); c += T_BIG_WG)
                __hip_atomic_store(&T.gcnt[c], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int nc = T.gcst[ME_NCOLMAX + 1];
        // a colour's pairs in G contiguous runs of whole waves, one run per workgroup (loading the next
        // colour step's records before the barrier, which they do not depend on, measured 4 % slower)
        auto sweep = [&](int pass) -> bool {
            for (int c = 0; c < nc; ++c) {
                const int c0 = T.gcst[c], c1 = T.gcst[c + 1];
                const int run = ((c1 - c0 + G - 1) / G + 63) & ~63;
                const int r0 = c0 + blockIdx.x * run, r1 = min(r0 + run, c1);
                for (int k = r0 + tid; k < r1; k += T_BIG_WG) {
                    const int q = p0 + k;
                    const GRec g = grec_load(T.rec + 8 * (size_t)q);
                    if (pass == 0) big_pair<0, FIXED, false, FAST>(A, T, 0u, g, q, b0, A.b.v, A.b.w, t2d);
                    else if (pass == 1) big_pair<1, FIXED, false, FAST>(A, T, 0u, g, q, b0, A.b.v, A.b.w, t2d);
                    else big_pair<2, FIXED, false, FAST>(A, T, 0u, g, q, b0, A.pv, A.po, t2d);
                }
                target += G;
                if (!giant_barrier(T.gbar, target, A.istat, stop, A.giant_wait_ticks)) return false;
            }
            return true;
        };
        if (np > 0) {
            if (FIXED && !sweep(0)) return;
            for (int it = 0; it < A.vit; ++it)
                if (!sweep(1)) return;
            if (blockIdx.x == 0 && tid == 0 &&
                (long long)atomicAdd(reinterpret_cast<unsigned long long *>(&T.dbg[7]), 0ull) == nb)
                T.dbg[4] = wall_clock64();
            for (int it = 0; it < A.pit; ++it)  // pv, po start at zero (k_body_pre)
                if (!sweep(2)) return;
        }
        for (int j = blockIdx.x * T_BIG_WG + tid; j < nb; j += G * T_BIG_WG) {  // k_integ of the tile's bodies
            int b = T.bsorted[b0 + j];
            if (A.b.invM[b] > 0.0f)
                integrate_body(A.b.x, A.b.q, ld3(A.b.v, b), ld3(A.b.w, b), ld3(A.pv, b), ld3(A.po, b), b, A.dt);
        }
        target += G;  // the next giant tile's records reuse gcst
        if (!giant_barrier(T.gbar, target, A.istat, stop, A.giant_wait_ticks)) return;
        if (blockIdx.x == 0 && tid == 0 &&
            (long long)atomicAdd(reinterpret_cast<unsigned long long *>(&T.dbg[7]), 0ull) == nb) {
            T.dbg[1] = T.dbg[2] = T.dbg[0];
            T.dbg[5] = T.dbg[6] = wall_clock64();
        }
    }
}
