// Bead geometry for the coloured solvers: per-body-pair bead contacts (A7) and the manifold
// reduction to at most mx points per body pair (A15), both on the device.
//
// Layout. The broad phase lists candidate body pairs in ascending pair key a*(N+1) + b + 1 (the
// ground pair (a, -1) first for every body). One wave per candidate pair tests its P x P bead
// pairs (P ground tests for a ground pair) and a ballot gives one bit per hit in slot order
// la*P + lb, so the generated contacts come out pair-major and, within a pair, in the reference's
// feature-key order (fa = a*P + la, fb = b*P + lb). A contact's key is the pair-major composite
// key pairkey*P*(P+1) + la*(P+1) + lb + 1: it orders like (pair key, feature key) and names the same
// feature as the reference's fa*(N*P+1) + fb + 1, so warm-start matches by key equality are the
// reference's matches. The kept contacts stay pair-major, so every pair is a contiguous run of at
// most mx contacts and the solvers read them without an index list.
#pragma once
#include "kernels.hpp"

namespace me {

struct BeadRaw {
    int64_t *key;
    float *pA, *pB, *n, *pen;
    double *jnd, *jnr;  // warm normal impulse (float64, as the reference's jn_in) and the redistributed one
    float *jt1, *jt2;
    int *keep;
};

// L-infinity prefilter of the reference's pair-bitmask generator: the float32 successor of 0.1f,
// wider than any hit (contact_generation_pair_bitmask.py:30)
#define ME_BEAD_LINF 0.10000000894069672f
#define ME_BEAD_MAXSEG 1024
#define ME_BEAD_MAXKEEP 8

// bead half extents of every body: the AABB of its bead spheres in the body frame
__global__ void k_fill_half(int N, float hx, float hy, float hz, float *half) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    half[3 * i] = hx; half[3 * i + 1] = hy; half[3 * i + 2] = hz;
}

__device__ __forceinline__ bool bead_hit(const float *allp, int a, int b, int slot, int P, float R) {
    if (b < 0) {
        if (slot >= P) return false;
        return allp[3 * (a * P + slot) + 2] - R < 0.0f;
    }
    if (slot >= P * P) return false;
    f3 d = ld3(allp, a * P + slot / P) - ld3(allp, b * P + slot % P);
    if (!(fabsf(d.x) <= ME_BEAD_LINF && fabsf(d.y) <= ME_BEAD_LINF && fabsf(d.z) <= ME_BEAD_LINF)) return false;
    float dist = clength(d);
    return dist < 2.0f * R && dist > 1e-9f;
}

// A7 count (contact_generation_pair_bitmask.py:16-38): one wave per candidate pair, one 64-bit
// ballot per word of 64 slots; the word masks are kept for the emit pass
__global__ void __launch_bounds__(ME_THREADS) k_bp_count(int cap, const int *istat, const int *cand_a,
                                                         const int *cand_b, const float *allp, int P, float R,
                                                         int words, uint64_t *mask, int *cnt) {
    int pair = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    int lane = threadIdx.x & 63;
    int n = istat[IS_CAND];
    if (n > cap) n = cap;
    if (pair >= n) return;  // uniform across the wave
    int a = cand_a[pair], b = cand_b[pair];
    int total = 0;
    for (int word = 0; word < words; ++word) {
        uint64_t m = __ballot(bead_hit(allp, a, b, word * 64 + lane, P, R));
        if (lane == 0) mask[(size_t)pair * words + word] = m;
        total += __popcll(m);
    }
    if (lane == 0) cnt[pair] = total;
}

// A7 emit (contact_generation_pair_bitmask.py:40-64): every hit goes to offset + its rank among the
// pair's hits; ground contacts: normal +z, pen R - z; bead pairs: normal (xi - xj)/|xi - xj|,
// pen 2R - |xi - xj| (k_gen_fid, contact_engine_gpu_colored.py:103-122)
__global__ void __launch_bounds__(ME_THREADS) k_bp_emit(int cap, const int *istat, const int *cand_a,
                                                        const int *cand_b, const float *allp, int N, int P, float R,
                                                        int words, const uint64_t *mask, const int *off, int capR,
                                                        BeadRaw raw, int *istat_w) {
    int pair = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    int lane = threadIdx.x & 63;
    int n = istat[IS_CAND];
    if (n > cap) n = cap;
    if (pair >= n) return;
    int a = cand_a[pair], b = cand_b[pair];
    int base = off[pair];
    int64_t pk = (int64_t)a * (int64_t)(N + 1) + (int64_t)b + 1;
    int64_t kstride = (int64_t)P * (int64_t)(P + 1);
    uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    bool over = false;
    for (int word = 0; word < words; ++word) {
        uint64_t m = mask[(size_t)pair * words + word];
        if ((m >> lane) & 1ull) {
            int t = base + __popcll(m & below);
            if (t >= capR) {
                over = true;
            } else {
                int slot = word * 64 + lane;
                int la = b < 0 ? slot : slot / P, lb = b < 0 ? -1 : slot % P;
                f3 xi = ld3(allp, a * P + la);
                raw.key[t] = pk * kstride + (int64_t)la * (P + 1) + (int64_t)(lb + 1);
                st3(raw.pA, t, xi);
                if (b < 0) {
                    st3(raw.pB, t, mk(0.0f, 0.0f, 0.0f));
                    st3(raw.n, t, mk(0.0f, 0.0f, 1.0f));
                    raw.pen[t] = -(xi.z - R);
                } else {
                    f3 xj = ld3(allp, b * P + lb);
                    f3 d = xi - xj;
                    float dist = clength(d);
                    st3(raw.pB, t, xj);
                    st3(raw.n, t, d / dist);
                    raw.pen[t] = 2.0f * R - dist;
                }
            }
        }
        base += __popcll(m);
    }
    if (over) atomicOr(&istat_w[IS_OVF], OVF_CONTACTS);
}

// planar coordinates about ref in the (u, w) basis: float32 difference, then float64
// (contact_engine_gpu_colored.py:585-590)
__device__ __forceinline__ void bead_planar(f3 p, f3 ref, const double *u, const double *w, double &r0, double &r1) {
    float dx = p.x - ref.x, dy = p.y - ref.y, dz = p.z - ref.z;
    double ax = (double)dx, ay = (double)dy, az = (double)dz;
    r0 = ax * u[0] + ay * u[1] + az * u[2];
    r1 = ax * w[0] + ay * w[1] + az * w[2];
}

// A14 + A15 for one candidate pair (contact_engine_gpu_colored.py:595-620 k_warm, :647-720 k_select):
// warm start of each generated contact by binary search of its key in the previous step's kept keys,
// then, above mx contacts, the E-optimal choice of mx of them (deepest first, then the greedy max of
// min(k+1, lambda_min) of the planar second moment about body A's centre, first index on ties) and
// the float64 redistribution of every contact's normal impulse onto the kept point nearest to it in
// the contact plane. One thread per pair; returns the kept count, or -1 past the fixed sizes.
__device__ int bead_select(int s0, int L, int mx, f3 ref, BeadRaw raw) {
    if (L <= mx) {
        for (int t = 0; t < L; ++t) { raw.keep[s0 + t] = 1; raw.jnr[s0 + t] = raw.jnd[s0 + t]; }
        return L;
    }
    if (L > ME_BEAD_MAXSEG || mx > ME_BEAD_MAXKEEP) {
        for (int t = 0; t < L; ++t) { raw.keep[s0 + t] = 1; raw.jnr[s0 + t] = raw.jnd[s0 + t]; }
        return -1;
    }
    for (int t = 0; t < L; ++t) raw.keep[s0 + t] = 0;
    int deep = 0;
    float dbest = raw.pen[s0];
    for (int t = 1; t < L; ++t)
        if (raw.pen[s0 + t] > dbest) { dbest = raw.pen[s0 + t]; deep = t; }
    f3 nr = ld3(raw.n, s0 + deep);
    double ax = 1.0, ay = 0.0, az = 0.0;
    if (fabsf(nr.x) >= 0.9f) { ax = 0.0; ay = 1.0; az = 0.0; }
    double n0 = (double)nr.x, n1 = (double)nr.y, n2 = (double)nr.z;
    double dn = ax * n0 + ay * n1 + az * n2;
    double u0 = ax - dn * n0, u1 = ay - dn * n1, u2 = az - dn * n2;
    double ln = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    double u[3] = {u0 / ln, u1 / ln, u2 / ln};
    double w[3] = {n1 * u[2] - n2 * u[1], n2 * u[0] - n0 * u[2], n0 * u[1] - n1 * u[0]};
    double r0, r1;
    bead_planar(ld3(raw.pA, s0 + deep), ref, u, w, r0, r1);
    double S00 = r0 * r0, S01 = r0 * r1, S10 = r1 * r0, S11 = r1 * r1;
    int pick[ME_BEAD_MAXKEEP];
    pick[0] = deep;
    for (int k = 1; k < mx; ++k) {
        double sbest = -1.0e300;
        int tbest = -1;
        double c00 = 0.0, c01 = 0.0, c10 = 0.0, c11 = 0.0;
        for (int t = 0; t < L; ++t) {
            bool taken = false;
            for (int q = 0; q < k; ++q) taken = taken || pick[q] == t;
            if (taken) continue;
            double t0, t1;
            bead_planar(ld3(raw.pA, s0 + t), ref, u, w, t0, t1);
            double a00 = S00 + t0 * t0, a01 = S01 + t0 * t1;
            double a10 = S10 + t1 * t0, a11 = S11 + t1 * t1;
            double tr = a00 + a11;
            double det = a00 * a11 - a01 * a10;
            double disc = tr * tr / 4.0 - det;
            if (!(disc > 0.0)) disc = 0.0;
            double lam = tr / 2.0 - sqrt(disc);
            double kk = (double)(k + 1);
            double sc = lam < kk ? lam : kk;
            if (sc > sbest) { sbest = sc; tbest = t; c00 = a00; c01 = a01; c10 = a10; c11 = a11; }
        }
        if (tbest < 0) return -1;
        pick[k] = tbest;
        S00 = c00; S01 = c01; S10 = c10; S11 = c11;
    }
    for (int k = 0; k < mx; ++k) { raw.keep[s0 + pick[k]] = 1; raw.jnr[s0 + pick[k]] = 0.0; }
    for (int t = 0; t < L; ++t) {
        double t0, t1;
        bead_planar(ld3(raw.pA, s0 + t), ref, u, w, t0, t1);
        double dbest2 = 1.0e300;
        int kbest = 0;
        for (int k = 0; k < mx; ++k) {
            double k0, k1;
            bead_planar(ld3(raw.pA, s0 + pick[k]), ref, u, w, k0, k1);
            double e0 = t0 - k0, e1 = t1 - k1;
            double d2 = e0 * e0 + e1 * e1;
            if (d2 < dbest2) { dbest2 = d2; kbest = k; }
        }
        int gk = s0 + pick[kbest];
        raw.jnr[gk] = raw.jnr[gk] + raw.jnd[s0 + t];
    }
    return mx;
}

// warm start of one pair's generated contacts by one thread: the pair's keys ascend, and so do the
// previous step's kept keys, so one binary search for the first kept key >= the pair's smallest key
// and then a merge give the exact-match set of a search per key
__device__ void bead_warm_serial(int s0, int L, int wn, const int64_t *wkey, const float *wjn, const float *wjt1,
                                 const float *wjt2, BeadRaw raw) {
    int j = 0;
    if (wn > 0) {
        int64_t k0 = raw.key[s0];
        int lo = 0, hi = wn;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (wkey[mid] < k0) lo = mid + 1;
            else hi = mid;
        }
        j = lo;
    }
    for (int t = 0; t < L; ++t) {
        int c = s0 + t;
        raw.jnd[c] = 0.0; raw.jt1[c] = 0.0f; raw.jt2[c] = 0.0f;
        if (wn == 0) continue;
        int64_t k = raw.key[c];
        while (j < wn && wkey[j] < k) ++j;
        if (j < wn && wkey[j] == k) { raw.jnd[c] = (double)wjn[j]; raw.jt1[c] = wjt1[j]; raw.jt2[c] = wjt2[j]; }
    }
}

// A14 + A15 of one candidate pair by a group of BP_G lanes, one generated contact per lane (segments
// of at most BP_G contacts with finite inputs; otherwise the group's first lane runs the one-thread
// form, bead_warm_serial + bead_select). Every step evaluates bead_select's expressions on the
// lane's own contact; the arg-max / arg-min choices reduce over the group with bead_select's
// tie-break (first index), and the float64 redistribution sums each kept point's share in contact
// order, as bead_select does.
#define BP_G 32
__device__ __forceinline__ double bp_shfl(double v, int src) { return __shfl(v, src, BP_G); }

__device__ __forceinline__ void bp_select_pair(int pair, int *istat, const int *cand_a, const float *x,
                                               const int *cnt, const int *off, int capR, int mx, int warm,
                                               const int64_t *wkey, const float *wjn, const float *wjt1,
                                               const float *wjt2, BeadRaw raw, int *kcnt, int *pflag) {
    const int lane = threadIdx.x & (BP_G - 1);
    int L = cnt[pair], s0 = off[pair];
    if (L > 0 && s0 + L > capR) L = 0;  // overflow already flagged by the emit pass
    if (lane == 0) pflag[pair] = L > 0 ? 1 : 0;
    if (L == 0) {
        if (lane == 0) kcnt[pair] = 0;
        return;
    }
    const int wn = warm ? istat[IS_WN] : 0;
    const f3 ref = ld3(x, cand_a[pair]);
    const bool mine = lane < L;
    // inputs of this lane's contact, and whether the whole segment takes the group form
    float pen = 0.0f;
    f3 pa = mk(0.0f, 0.0f, 0.0f), nr = mk(0.0f, 0.0f, 1.0f);
    if (mine) {
        pen = raw.pen[s0 + lane];
        pa = ld3(raw.pA, s0 + lane);
        nr = ld3(raw.n, s0 + lane);
    }
    bool fin = isfinite(pen) && isfinite(pa.x) && isfinite(pa.y) && isfinite(pa.z) && isfinite(nr.x) &&
               isfinite(nr.y) && isfinite(nr.z);
    const unsigned long long gmask = (BP_G == 64 ? ~0ull : ((1ull << BP_G) - 1ull)) << (threadIdx.x & 63 & ~(BP_G - 1));
    const bool group = L <= BP_G && (__ballot(!fin) & gmask) == 0ull;
    if (!group) {
        if (lane == 0) {
            bead_warm_serial(s0, L, wn, wkey, wjn, wjt1, wjt2, raw);
            int k = bead_select(s0, L, mx, ref, raw);
            if (k < 0) {
                atomicOr(&istat[IS_OVF], OVF_CONTACTS);
                k = L;
            }
            kcnt[pair] = k;
        }
        return;
    }
    // A14 warm start: the first kept key >= the pair's smallest key (one search), then each lane
    // walks forward from it to its own key
    double jnd = 0.0;
    float jt1 = 0.0f, jt2 = 0.0f;
    if (wn > 0) {
        int j0 = 0;
        if (lane == 0) {
            int64_t k0 = raw.key[s0];
            int lo = 0, hi = wn;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (wkey[mid] < k0) lo = mid + 1;
                else hi = mid;
            }
            j0 = lo;
        }
        j0 = __shfl(j0, 0, BP_G);
        if (mine) {
            int64_t k = raw.key[s0 + lane];
            int j = j0;
            while (j < wn && wkey[j] < k) ++j;
            if (j < wn && wkey[j] == k) { jnd = (double)wjn[j]; jt1 = wjt1[j]; jt2 = wjt2[j]; }
        }
    }
    if (mine) { raw.jnd[s0 + lane] = jnd; raw.jt1[s0 + lane] = jt1; raw.jt2[s0 + lane] = jt2; }
    if (L <= mx) {
        if (mine) { raw.keep[s0 + lane] = 1; raw.jnr[s0 + lane] = jnd; }
        if (lane == 0) kcnt[pair] = L;
        return;
    }
    if (mx > ME_BEAD_MAXKEEP) {  // bead_select's fixed-size refusal
        if (mine) { raw.keep[s0 + lane] = 1; raw.jnr[s0 + lane] = jnd; }
        if (lane == 0) { atomicOr(&istat[IS_OVF], OVF_CONTACTS); kcnt[pair] = L; }
        return;
    }
    // deepest contact, first index on ties
    float bp = mine ? pen : -INFINITY;
    int bt = mine ? lane : BP_G;
#pragma unroll
    for (int o = BP_G / 2; o > 0; o >>= 1) {
        float op = __shfl_xor(bp, o, BP_G);
        int ot = __shfl_xor(bt, o, BP_G);
        if (op > bp || (op == bp && ot < bt)) { bp = op; bt = ot; }
    }
    const int deep = bt;
    // plane basis from the deepest contact's normal (float64), as bead_select
    f3 nd = mk(__shfl(nr.x, deep, BP_G), __shfl(nr.y, deep, BP_G), __shfl(nr.z, deep, BP_G));
    double ax = 1.0, ay = 0.0, az = 0.0;
    if (fabsf(nd.x) >= 0.9f) { ax = 0.0; ay = 1.0; az = 0.0; }
    double n0 = (double)nd.x, n1 = (double)nd.y, n2 = (double)nd.z;
    double dn = ax * n0 + ay * n1 + az * n2;
    double u0 = ax - dn * n0, u1 = ay - dn * n1, u2 = az - dn * n2;
    double ln = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    double u[3] = {u0 / ln, u1 / ln, u2 / ln};
    double w[3] = {n1 * u[2] - n2 * u[1], n2 * u[0] - n0 * u[2], n0 * u[1] - n1 * u[0]};
    double t0 = 0.0, t1 = 0.0;
    bead_planar(pa, ref, u, w, t0, t1);
    const bool pfin = !mine || (isfinite(t0) && isfinite(t1));
    if ((__ballot(!pfin) & gmask) != 0ull) {  // a degenerate basis: the one-thread form decides
        if (lane == 0) {
            int k = bead_select(s0, L, mx, ref, raw);
            if (k < 0) {
                atomicOr(&istat[IS_OVF], OVF_CONTACTS);
                k = L;
            }
            kcnt[pair] = k;
        }
        return;
    }
    double r0 = bp_shfl(t0, deep), r1 = bp_shfl(t1, deep);
    double S00 = r0 * r0, S01 = r0 * r1, S10 = r1 * r0, S11 = r1 * r1;
    unsigned long long taken = 1ull << deep;
    int pick[ME_BEAD_MAXKEEP];
    pick[0] = deep;
    for (int k = 1; k < mx; ++k) {
        const bool cand = mine && !((taken >> lane) & 1ull);
        double a00 = S00 + t0 * t0, a01 = S01 + t0 * t1;
        double a10 = S10 + t1 * t0, a11 = S11 + t1 * t1;
        double tr = a00 + a11;
        double det = a00 * a11 - a01 * a10;
        double disc = tr * tr / 4.0 - det;
        if (!(disc > 0.0)) disc = 0.0;
        double lam = tr / 2.0 - sqrt(disc);
        double kk = (double)(k + 1);
        double sc = lam < kk ? lam : kk;
        double bs = cand ? sc : -1.0e300;
        int bi = cand ? lane : BP_G;
#pragma unroll
        for (int o = BP_G / 2; o > 0; o >>= 1) {
            double os = __shfl_xor(bs, o, BP_G);
            int oi = __shfl_xor(bi, o, BP_G);
            if (os > bs || (os == bs && oi < bi)) { bs = os; bi = oi; }
        }
        pick[k] = bi;
        taken |= 1ull << bi;
        S00 = bp_shfl(a00, bi); S01 = bp_shfl(a01, bi); S10 = bp_shfl(a10, bi); S11 = bp_shfl(a11, bi);
    }
    if (mine) raw.keep[s0 + lane] = (int)((taken >> lane) & 1ull);
    // nearest kept point in the plane (first index on ties), then each kept point sums its share of
    // the warm normal impulses in contact order
    double kx[ME_BEAD_MAXKEEP], ky[ME_BEAD_MAXKEEP];
    for (int k = 0; k < mx; ++k) { kx[k] = bp_shfl(t0, pick[k]); ky[k] = bp_shfl(t1, pick[k]); }
    int kbest = 0;
    double dbest2 = 1.0e300;
    for (int k = 0; k < mx; ++k) {
        double e0 = t0 - kx[k], e1 = t1 - ky[k];
        double d2 = e0 * e0 + e1 * e1;
        if (d2 < dbest2) { dbest2 = d2; kbest = k; }
    }
    double acc = 0.0;
    for (int t = 0; t < L; ++t) {
        int kb = __shfl(kbest, t, BP_G);
        double jd = bp_shfl(jnd, t);
        if (lane < mx && kb == lane) acc = acc + jd;
    }
    if (lane < mx) raw.jnr[s0 + pick[lane]] = acc;
    if (lane == 0) kcnt[pair] = mx;
}

// groups of BP_G lanes take the candidate pairs in turn (the grid is sized from the capacity; the
// loop bound is the live count)
__global__ void __launch_bounds__(ME_THREADS) k_bp_select(int cap, int *istat, const int *cand_a, const float *x,
                                                          const int *cnt, const int *off, int capR, int mx, int warm,
                                                          const int64_t *wkey, const float *wjn, const float *wjt1,
                                                          const float *wjt2, BeadRaw raw, int *kcnt, int *pflag) {
    int n = istat[IS_CAND];
    if (n > cap) n = cap;
    const int groups = gridDim.x * (blockDim.x / BP_G);
    for (int pair = (blockIdx.x * blockDim.x + threadIdx.x) / BP_G; pair < n; pair += groups)  // uniform per group
        bp_select_pair(pair, istat, cand_a, x, cnt, off, capR, mx, warm, wkey, wjn, wjt1, wjt2, raw, kcnt, pflag);
}

// compaction of the kept contacts, pair-major, and one pair per candidate pair with contacts
// (k_compact + the pair build, contact_engine_gpu_colored.py:723-740, 1138-1157)
__global__ void k_bp_compact(int cap, const int *istat, const int *cand_a, const int *cand_b, const float *mu,
                             const int *cnt, const int *off, const int *kcnt, const int *koff, const int *pflag,
                             const int *pidx, int capC, int capP, BeadRaw raw, Contacts Cc, Pairs Pp,
                             int *istat_w) {
    int pair = blockIdx.x * blockDim.x + threadIdx.x;
    int n = istat[IS_CAND];
    if (n > cap) n = cap;
    if (pair >= n || !pflag[pair]) return;
    int a = cand_a[pair], b = cand_b[pair];
    int p = pidx[pair], o = koff[pair], kn = kcnt[pair], L = cnt[pair], s0 = off[pair];
    if (p >= capP || o + kn > capC) {
        atomicOr(&istat_w[IS_OVF], OVF_CONTACTS);
        return;
    }
    Pp.pbi[p] = a; Pp.pbj[p] = b; Pp.pstart[p] = o; Pp.pcount[p] = kn;
    float m = pair_mu(mu, a, b);
    int d = o;
    for (int t = 0; t < L; ++t) {
        int c = s0 + t;
        if (!raw.keep[c]) continue;
        Cc.key[d] = raw.key[c]; Cc.bi[d] = a; Cc.bj[d] = b;
        st3(Cc.pA, d, ld3(raw.pA, c)); st3(Cc.pB, d, ld3(raw.pB, c)); st3(Cc.n, d, ld3(raw.n, c));
        Cc.pen[d] = raw.pen[c]; Cc.mu[d] = m;
        Cc.jn[d] = (float)raw.jnr[c]; Cc.jt1[d] = raw.jt1[c]; Cc.jt2[d] = raw.jt2[c]; Cc.jp[d] = 0.0f;
        ++d;
    }
}

}  // namespace me
