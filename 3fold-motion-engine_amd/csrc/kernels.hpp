// Pipeline kernels of the contact step (gfx950). One thread per body, candidate pair, contact or
// pair as noted; every count a later kernel needs stays on the device (istat slots).
#pragma once
#include "geometry.hpp"
#include "solvers.hpp"

namespace me {

// int32 status slots shared by the kernels of one engine
enum {
    IS_C = 0, IS_P, IS_CAND, IS_NCOL, IS_HITS, IS_MISSES, IS_OVF, IS_ISL, IS_TILES, IS_MAXTB,
    IS_MAXEXT, IS_LEFT, IS_WN, IS_CACHED_P, IS_CACHED_NCOL, IS_MISMATCH, IS_CELL, IS_NBODYSORT,
    IS_NPAIRSORT, IS_NCOLJP, IS_NBIG, IS_CRAW, IS_JTOT, IS_NGIANT, IS_NGCOL,
    IS_BIGQ,  // next entry of the largest-first big-tile queue taken by k_t_big
    IS_COLQ,  // the same queue for the big tiles' colouring (k_t_color)
    IS_NSMALL,  // small tiles of the substep (k_t_biglist)
    IS_SMALLQ,  // next entry of the largest-first small-tile queue taken by k_t_small
    IS_CCDIFF,  // the tiled solve's pair list differs from the colour cache's (k_t_setup)
    IS_COUNT = 32
};
enum { OVF_CONTACTS = 1, OVF_PAIRS = 2, OVF_CAND = 4, OVF_COLORS = 8, OVF_TILE = 16,
       OVF_TAPE = 32 /* a recorded (differentiated) substep met a tile beyond one small tile */,
       OVF_GIANT = 64 /* a grid barrier of the giant-tile sweep waited past its bound */,
       OVF_ADJ = 128 /* an adjoint contribution beyond the fixed-point accumulators' range */,
       OVF_SCAN = 256 /* a one-pass scan's look-back waited past its bound */ };

#define ME_NCOLMAX 512
#define ME_MAXCAND 32
#define ME_BEADCAND 64
#define ME_THREADS 256

__global__ void k_fill_i32(int *p, int n, int val) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = val;
}
__global__ void k_fill_f32(float *p, int n, float val) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = val;
}
__global__ void k_fill_i64(int64_t *p, int n, int64_t val) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = val;
}
__global__ void k_copy_f32(float *dst, const float *src, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ── exclusive scan (3 phases, 2048 items per block, deterministic) ──
#define SCAN_ITEMS 8
#define SCAN_BLOCK (ME_THREADS * SCAN_ITEMS)

__device__ __forceinline__ int block_excl_scan(int v, int *sh, int *total) {
    // wave64 inclusive scan, then across the 4 waves of the block
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    int base = 0, tot = 0;
    for (int k = 0; k < ME_THREADS / 64; ++k) {
        if (k < wv) base += sh[k];
        tot += sh[k];
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

__device__ __forceinline__ int scan_len(const int *n_dev, int n_host, int cap) {
    int n = n_dev ? *n_dev : n_host;
    return n < cap ? n : cap;
}

__global__ void __launch_bounds__(ME_THREADS) k_scan_reduce(const int *in, const int *n_dev, int n_host, int cap, int *bsum) {
    __shared__ int sh[ME_THREADS / 64];
    int n = scan_len(n_dev, n_host, cap);
    int base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) s += in[base + k];
    int tot;
    block_excl_scan(s, sh, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_scan_top(int *bsum, int nb, int *total) {
    __shared__ int sh[16];
    // each of 1024 threads owns ceil(nb/1024) consecutive block sums
    int per = (nb + 1023) / 1024;
    int b0 = threadIdx.x * per;
    int s = 0;
    for (int k = 0; k < per; ++k)
        if (b0 + k < nb) s += bsum[b0 + k];
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    int base = 0, tot = 0;
    for (int k = 0; k < 16; ++k) {
        if (k < wv) base += sh[k];
        tot += sh[k];
    }
    int run = base + x - s;
    for (int k = 0; k < per; ++k)
        if (b0 + k < nb) {
            int t = bsum[b0 + k];
            bsum[b0 + k] = run;
            run += t;
        }
    if (threadIdx.x == 0 && total) *total = tot;
}

__global__ void __launch_bounds__(ME_THREADS) k_scan_down(const int *in, int *out, const int *n_dev, int n_host, int cap,
                                                          const int *bsum) {
    __shared__ int sh[ME_THREADS / 64];
    int n = scan_len(n_dev, n_host, cap);
    int base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    int vals[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        vals[k] = (base + k < n) ? in[base + k] : 0;
        s += vals[k];
    }
    int tot;
    int run = block_excl_scan(s, sh, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = run;
        run += vals[k];
    }
}

// ── exclusive scan in one pass (decoupled look-back) ──
// One launch instead of three. Each block takes a ticket from its call site's counter, which only
// grows: ticket / nb numbers the launch and ticket % nb is the block's place in the scan, so a block
// only ever waits on blocks that started before it. Every block publishes its sum (AGG) as soon as it
// has it and its inclusive prefix (INC) once known, in one 64-bit word per block and value: launch
// number (30 bits) | state (2) | value (32); a word of another launch reads as not yet published, so
// the words need no clearing between launches. The first wave of a block reads 64 predecessors' words
// at a time and adds back to the nearest INC. Integer sums in a fixed order of blocks: the result is
// the three-pass scan's. NV = 2 also scans the flags (in[i] > 0) into out2. The last block writes the
// totals; with capC > 0 also the clamped contact and pair counts (IS_CRAW, IS_C, IS_P) of a
// compaction, as k_clamp_counts does. The look-back's wait is bounded by the wall clock (OVF_SCAN).
#define SCAN_SITES 12  // call sites with their own ticket counter and block words
#define SC_TILED_B 10   // the island build's two scans of large scenes (tiled_scan)
#define SC_TILED_P 11
#define SC_AGG 1ull
#define SC_INC 2ull

__device__ __forceinline__ unsigned long long sc_word(unsigned launch, unsigned long long state, int v) {
    return ((unsigned long long)(launch & 0x3fffffffu) << 34) | (state << 32) | (unsigned)v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// the exclusive prefix of block b (> 0) of launch L from its predecessors' words (wave 0 of the block)
__device__ __forceinline__ int sc_lookback(unsigned long long *words, int b, unsigned L, int *istat,
                                           long long wait_ticks) {
    const int lane = threadIdx.x & 63;
    const unsigned ep = L & 0x3fffffffu;
    int acc = 0, j = b - 1;
    const long long t0 = wall_clock64();
    for (;;) {
        const int idx = j - lane;
        bool ready = true, inc = true;
        int v = 0;
        if (idx >= 0) {
            const unsigned long long w = __hip_atomic_load(words + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned st = (unsigned)(w >> 32) & 3u;
            ready = (unsigned)(w >> 34) == ep && st != 0;
            inc = ready && st == (unsigned)SC_INC;
            v = (int)(unsigned)w;
        }
        const unsigned long long im = __ballot(inc), rm = __ballot(ready);
        const int fi = im ? __ffsll((long long)im) - 1 : 63;  // the nearest inclusive word, else the window's end
        const unsigned long long need = fi >= 63 ? ~0ull : ((2ull << fi) - 1);
        if ((rm & need) == need) {
            acc += wave_sum(lane <= fi ? v : 0);
            if (im) return acc;
            j -= 64;
            continue;
        }
        __builtin_amdgcn_s_sleep(1);
        if (wall_clock64() - t0 > wait_ticks) {
            if (lane == 0) atomicOr(&istat[IS_OVF], OVF_SCAN);
            return acc;
        }
    }
}

template <int NV>
__global__ void __launch_bounds__(ME_THREADS) k_scan1(const int *in, int *out, int *out2, const int *n_dev, int n_host,
                                                      int cap, int *total, int *total2, unsigned long long *site,
                                                      int *istat, long long wait_ticks, int capC, int capP) {
    __shared__ int sh[ME_THREADS / 64];
    __shared__ unsigned long long tk;
    __shared__ int pre[2];
    const int tid = threadIdx.x, nb = (int)gridDim.x;
    if (tid == 0) tk = __hip_atomic_fetch_add(site, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int b = (int)(tk % (unsigned long long)nb);
    const unsigned L = (unsigned)(tk / (unsigned long long)nb);
    unsigned long long *words = site + 1;
    const int n = scan_len(n_dev, n_host, cap);
    const int base = b * SCAN_BLOCK + tid * SCAN_ITEMS;
    int vals[SCAN_ITEMS];
    int s = 0, f = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        vals[k] = (base + k < n) ? in[base + k] : 0;
        s += vals[k];
        f += vals[k] > 0 ? 1 : 0;
    }
    int agg, agg2 = 0;
    const int x = block_excl_scan(s, sh, &agg);
    const int x2 = NV == 2 ? block_excl_scan(f, sh, &agg2) : 0;
    if (tid < 64) {
        int p0 = 0, p1 = 0;
        if (b == 0) {
            if (tid == 0) {
                __hip_atomic_store(words, sc_word(L, SC_INC, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (NV == 2)
                    __hip_atomic_store(words + nb, sc_word(L, SC_INC, agg2), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else {
            if (tid == 0) {
                __hip_atomic_store(words + b, sc_word(L, SC_AGG, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (NV == 2)
                    __hip_atomic_store(words + nb + b, sc_word(L, SC_AGG, agg2), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
            p0 = sc_lookback(words, b, L, istat, wait_ticks);
            if (NV == 2) p1 = sc_lookback(words + nb, b, L, istat, wait_ticks);
            if (tid == 0) {
                __hip_atomic_store(words + b, sc_word(L, SC_INC, p0 + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (NV == 2)
                    __hip_atomic_store(words + nb + b, sc_word(L, SC_INC, p1 + agg2), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (tid == 0) { pre[0] = p0; pre[1] = p1; }
    }
    __syncthreads();
    int run = pre[0] + x, run2 = pre[1] + x2;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) {
            out[base + k] = run;
            if (NV == 2) out2[base + k] = run2;
        }
        run += vals[k];
        run2 += vals[k] > 0 ? 1 : 0;
    }
    if (b == nb - 1 && tid == 0) {
        const int t1 = pre[0] + agg, t2 = pre[1] + agg2;
        if (total) *total = t1;
        if (NV == 2 && total2) *total2 = t2;
        if (capC > 0) {  // the compaction's counts, clamped (k_clamp_counts with copy_raw)
            istat[IS_CRAW] = t1;
            if (t1 > capC) { istat[IS_C] = capC; atomicOr(&istat[IS_OVF], OVF_CONTACTS); }
            if (NV == 2 && t2 > capP) { istat[IS_P] = capP; atomicOr(&istat[IS_OVF], OVF_PAIRS); }
        }
    }
}

// ── the engine state block, copied by its used items ──
// The state block (the workspace's leading bytes that carry a substep's state to the next and to the
// adjoint of a recorded substep) is a sequence of arrays; a segment records one array's offset and
// size and which status word counts its used items (of `unit` bytes), -1 for all of it. Copying only
// the used items of the capacity-sized arrays is what makes a state snapshot cheap.
#define ME_STATE_SEGS 64
struct StateSeg {
    unsigned long long off, bytes, unit;
    int slot;
};
struct StateSegs {
    StateSeg s[ME_STATE_SEGS];
    int n = 0, overflow = 0;
};

// grid (blocks, segments): segment blockIdx.y, its used bytes (from the source's status words, rounded
// up to 16 within the array) in 16-byte units
__global__ void __launch_bounds__(ME_THREADS) k_state_copy(char *dst, const char *src, StateSegs segs) {
    const StateSeg g = segs.s[blockIdx.y];
    unsigned long long used = g.bytes;
    if (g.slot >= 0) {
        const int c = reinterpret_cast<const int *>(src)[g.slot];  // istat leads the block
        const unsigned long long u = (unsigned long long)(c > 0 ? c : 0) * g.unit;
        used = u < used ? u : used;
    }
    const unsigned long long n16 = (used + 15) / 16;
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src + g.off);
    uint4 *d4 = reinterpret_cast<uint4 *>(dst + g.off);
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n16;
         i += (unsigned long long)gridDim.x * blockDim.x)
        d4[i] = s4[i];
}

// ── A2 gravity, A3 world inverse inertia, box AABBs, velocity snapshot ──
struct Bodies {
    float *x, *q, *v, *w;
    const float *invM, *IbInv, *half, *mu;
    const int *kin, *scene;
    const float *rest;
};

// One global atomic per wave instead of one per lane on a single status word (every lane of a
// large launch hitting one address serialises at the memory-side atomic unit). All lanes of the
// wave must call these.
__device__ __forceinline__ void wave_atomic_max(int *addr, int val, bool pred) {
    int v = pred ? val : INT_MIN;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int w = __shfl_xor(v, o);
        v = v > w ? v : w;
    }
    if (__lane_id() == 0 && v != INT_MIN) atomicMax(addr, v);
}
__device__ __forceinline__ void wave_atomic_count(int *addr, bool pred) {
    unsigned long long m = __ballot(pred);
    if (__lane_id() == 0 && m) atomicAdd(addr, (int)__popcll(m));
}

// Also clears, for the grid built next, its T bucket counters and cursors (grid-stride), and for the
// coloured solvers the pseudo-velocities pv, po (zeroed every substep, contact_engine_gpu_colored.py:1330).
__global__ void k_body_pre(int N, Bodies b, float *invIw, float *Rw, float *aabb, float *vsnap, float *xsnap,
                           float *wsnap, int *istat, float dt, int box_mode, int T, int *bcount, int *bcursor,
                           float *pv, float *po) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    for (int k = i; k < T; k += gridDim.x * blockDim.x) { bcount[k] = 0; bcursor[k] = 0; }
    if (pv && i < N) { st3(pv, i, mk(0.0f, 0.0f, 0.0f)); st3(po, i, mk(0.0f, 0.0f, 0.0f)); }
    float mext = 0.0f;
    bool has = false;
    if (i < N) {
        f3 v = ld3(b.v, i);
        if (b.invM[i] > 0.0f) {  // k_grav (contact_engine_gpu.py:41-44)
            v = cmadd(dt, mk(0.0f, 0.0f, -9.81f), v);
            st3(b.v, i, v);
        }
        st3(vsnap, i, v);
        st3(xsnap, i, ld3(b.x, i));
        st3(wsnap, i, ld3(b.w, i));
        m3 R = cquat_to_matrix(ld4(b.q, i));  // k_invIw (contact_engine_gpu.py:47-53)
        stm(Rw, i, R);
        if (b.kin[i] == 1) {
            m3 z;
    #pragma unroll
            for (int k = 0; k < 9; ++k) z.m[k / 3][k % 3] = 0.0f;
            stm(invIw, i, z);
        } else {
            m3 Ib = ldm(b.IbInv, i);
            m3 t = cmm(R, Ib);
            m3 r = cmm(t, transpose(R));
            stm(invIw, i, r);
        }
        if (box_mode) {
            f3 h = ld3(b.half, i), c = ld3(b.x, i);
            float e0 = fabsf(R.m[0][0]) * h.x + fabsf(R.m[0][1]) * h.y + fabsf(R.m[0][2]) * h.z;
            float e1 = fabsf(R.m[1][0]) * h.x + fabsf(R.m[1][1]) * h.y + fabsf(R.m[1][2]) * h.z;
            float e2 = fabsf(R.m[2][0]) * h.x + fabsf(R.m[2][1]) * h.y + fabsf(R.m[2][2]) * h.z;
            aabb[6 * i + 0] = c.x - e0; aabb[6 * i + 1] = c.y - e1; aabb[6 * i + 2] = c.z - e2;
            aabb[6 * i + 3] = c.x + e0; aabb[6 * i + 4] = c.y + e1; aabb[6 * i + 5] = c.z + e2;
            // largest AABB half extent actually stored (hi - centre), as the oracle measures it
            mext = fmaxf(fmaxf((c.x + e0) - c.x, (c.y + e1) - c.y), (c.z + e2) - c.z);
            has = true;
        }
    }
    // non-negative floats order as their bit patterns
    wave_atomic_max(&istat[IS_MAXEXT], __float_as_int(mext), has);
}

// cell size of the body grid: (2 * max half extent + margin + slack) * 1.001, so that any two
// overlapping AABBs have centre cells at most one apart even after rounding of x / cell
// the box grid's cell: twice the largest AABB half extent of the substep plus the margin, padded
__device__ __forceinline__ float box_cell(const int *istat, float margin) {
    float hmax = __int_as_float(istat[IS_MAXEXT]);
    float cell = (2.0f * hmax + margin + ME_BP_SLACK) * 1.001f;
    return cell > 0.0f ? cell : 1.0f;
}

// cell -> bucket (mod the power-of-two table size): linear in cx, so the three cells (cx-1..cx+1) of a
// neighbour row take consecutive buckets -- their ranges share cache lines and their items lie back to
// back; cy, cz and the scene enter with large odd factors (a lookup tests the item's own cell)
__device__ __forceinline__ uint32_t cell_hash(int cx, int cy, int cz, int s) {
    return (uint32_t)cx + (uint32_t)cy * 2654435761u + (uint32_t)cz * 2246822519u + (uint32_t)s * 3266489917u;
}

// hash grid: bucket per item (box centre or bead point), bucket counts
__global__ void k_grid_count(int n, const float *pos, const int *scene_of, int div, const int *istat, float cell_fixed,
                             float margin, int T, int *cellc, int *bucket, int *bcount) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float cell = cell_fixed > 0.0f ? cell_fixed : box_cell(istat, margin);
    f3 p = ld3(pos, i);
    int cx = (int)floorf(p.x / cell), cy = (int)floorf(p.y / cell), cz = (int)floorf(p.z / cell);
    int s = scene_of ? scene_of[i / div] : 0;
    cellc[4 * i] = cx; cellc[4 * i + 1] = cy; cellc[4 * i + 2] = cz; cellc[4 * i + 3] = s;
    int bk = (int)(cell_hash(cx, cy, cz, s) & (uint32_t)(T - 1));
    bucket[i] = bk;
    atomicAdd(&bcount[bk], 1);
}
// items in bucket order; with spos, also each item's position and id in the same order (one 16-byte
// load per visited item for the bead grid); with scell / sbox (the body grid), each item's cell and
// scene, and its AABB with its inverse mass and id (lo.xyz, invM | hi.xyz, id), so a lookup reads an
// item's data at its bucket position in one round of independent loads
__global__ void k_grid_scatter(int n, const int *bucket, const int *bstart, int *bcursor, int *items, const float *pos,
                               float4 *spos, const int *cellc, const float *aabb, const float *invM, int4 *scell,
                               float4 *sbox) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int bk = bucket[i];
    int at = bstart[bk] + atomicAdd(&bcursor[bk], 1);
    items[at] = i;
    if (spos) {
        f3 p = ld3(pos, i);
        spos[at] = make_float4(p.x, p.y, p.z, __int_as_float(i));
    }
    if (scell) {
        scell[at] = make_int4(cellc[4 * i], cellc[4 * i + 1], cellc[4 * i + 2], cellc[4 * i + 3]);
        const float *bx = aabb + 6 * (size_t)i;
        sbox[2 * (size_t)at] = make_float4(bx[0], bx[1], bx[2], invM[i]);
        sbox[2 * (size_t)at + 1] = make_float4(bx[3], bx[4], bx[5], __int_as_float(i));
    }
}

__device__ __forceinline__ bool aabb_overlap(const float *a, const float *b, float m) {
    return a[0] <= b[3] + m && b[0] <= a[3] + m && a[1] <= b[4] + m && b[1] <= a[4] + m && a[2] <= b[5] + m &&
           b[2] <= a[5] + m;
}

// ── A6 body pairs: ground candidate + AABB overlaps with higher bodies of the same scene, in key order ──
// Half a wave per body a (32 lanes, two bodies per wave, eight per workgroup): lane l < 27 scans
// neighbour cell l's grid bucket for higher bodies of the same scene whose AABB overlaps a's (at most
// ME_CELLCAND per cell); a prefix sum over the half wave places the hits in LDS and the lanes rank
// them, so the list comes out ascending behind the ground candidate (the pair-key order
// a*(N+1) + b + 1). all_pairs (bead geometry): pairs of two static bodies and ground pairs of static
// bodies are kept, as the reference's bead contact generation emits their contacts too
// (contact_engine_gpu_colored.py:103-122).
#define ME_CELLCAND 16
#define ME_BP_PER_WG 8
__global__ void __launch_bounds__(32 * ME_BP_PER_WG) k_box_pairs(int N, Bodies b, const float *aabb, const int *cellc,
                                                                 const int *bstart, const int *bcount, const int4 *scell,
                                                                 const float4 *sbox, int T, float margin, int all_pairs,
                                                                 int *cand_n, int *cand_tmp, int *istat) {
    __shared__ int hits_all[ME_BP_PER_WG][ME_MAXCAND];
    const int half = threadIdx.x >> 5, lane = threadIdx.x & 31;
    const int a = blockIdx.x * ME_BP_PER_WG + half;
    const bool live = a < N;  // no early exit: the half waves shuffle and the block synchronises
    int *hits = hits_all[half];
    float m = margin + ME_BP_SLACK;
    const float *A = aabb + 6 * (live ? a : 0);
    bool dynA = live && b.invM[a] > 0.0f;
    int mine[ME_CELLCAND];
    int n = 0;
    bool over = false;
    if (live && lane < 27) {
        int X = cellc[4 * a] + lane % 3 - 1, Y = cellc[4 * a + 1] + (lane / 3) % 3 - 1, Z = cellc[4 * a + 2] + lane / 9 - 1;
        int s = cellc[4 * a + 3];
        int bk = (int)(cell_hash(X, Y, Z, s) & (uint32_t)(T - 1));
        int st = bstart[bk], cnt = bcount[bk];
        for (int t = 0; t < cnt; ++t) {
            int4 cj = scell[st + t];
            float4 lo = sbox[2 * (size_t)(st + t)], hi = sbox[2 * (size_t)(st + t) + 1];
            int j = __float_as_int(hi.w);
            if (j <= a) continue;
            if (cj.x != X || cj.y != Y || cj.z != Z || cj.w != s) continue;
            if (!all_pairs && !dynA && !(lo.w > 0.0f)) continue;
            const float B6[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
            if (!aabb_overlap(A, B6, m)) continue;
            if (n < ME_CELLCAND) mine[n++] = j;
            else over = true;
        }
    }
    int first = (live && (all_pairs || dynA) && A[2] < m) ? 1 : 0;
    int incl = n;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        int v = __shfl_up(incl, o, 32);
        if (lane >= o) incl += v;
    }
    int total = __shfl(incl, 31, 32);
    int off = first + incl - n;
    for (int k = 0; k < n; ++k)
        if (off + k < ME_MAXCAND) hits[off + k] = mine[k];
    over = over || first + total > ME_MAXCAND;
    __syncthreads();
    int nc = first + total < ME_MAXCAND ? first + total : ME_MAXCAND;
    for (int q = first + lane; live && q < nc; q += 32) {
        int c = hits[q], rank = first;
        for (int t = first; t < nc; ++t) rank += hits[t] < c;
        cand_tmp[a * ME_MAXCAND + rank] = c;
    }
    if (live && lane == 0) {
        if (first) cand_tmp[a * ME_MAXCAND] = -1;
        cand_n[a] = nc;
        if (a == 0) istat[IS_MAXEXT] = 0;  // read by this substep's grid only; the next k_body_pre maxes into it
    }
    if (__any(over) && threadIdx.x % 64 == 0) atomicOr(&istat[IS_OVF], OVF_CAND);
}

__global__ void k_box_cand_emit(int N, const int *cand_n, const int *cand_off, const int *cand_tmp, int cap,
                                int *cand_a, int *cand_b, int *istat) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    int o = cand_off[a], n = cand_n[a];
    for (int k = 0; k < n; ++k) {
        if (o + k < cap) {
            cand_a[o + k] = a;
            cand_b[o + k] = cand_tmp[a * ME_MAXCAND + k];
        } else {
            atomicOr(&istat[IS_OVF], OVF_PAIRS);
        }
    }
}

__device__ __forceinline__ float pair_mu(const float *mu, int a, int b) {
    if (b < 0) return mu[a];
    float ma = mu[a], mb = mu[b];
    return ma == mb ? ma : sqrtf(ma * mb);
}

struct Slots {
    int *n;
    int64_t *key;
    float *p, *nrm, *pen, *mu;
};

// ── narrow phase per candidate pair, up to 4 contacts in fixed slots ──
// LDSW: each thread's clipping work in LDS (64-thread workgroups, 58 KB each: at most two per CU),
// for scenes whose candidates fit about one round of such workgroups; else in private memory
// (scratch), 256 threads per workgroup and many workgroups per CU
#define ME_NARROW_WG 64
#define ME_NARROW_LDS_MAXN 12288  // bodies up to which k_narrow<true> runs (about 2.3 candidates per body)
template <bool LDSW>
__global__ void __launch_bounds__(LDSW ? ME_NARROW_WG : ME_THREADS) k_narrow(int cap, const int *istat, Bodies b,
                                                                             const float *Rw, const int *cand_a,
                                                                             const int *cand_b, float margin, int N,
                                                                             Slots S) {
    __shared__ ClipWork work[LDSW ? ME_NARROW_WG : 1];
    ClipWork own;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int n = istat[IS_CAND];
    if (n > cap) n = cap;
    if (p >= n) return;
    int a = cand_a[p], bb = cand_b[p];
    m3 RA = ldm(Rw, a);
    f3 cA = ld3(b.x, a), hA = ld3(b.half, a);
    f3 pp[4], nn[4];
    float pen[4];
    int code[4];
    int k;
    int64_t pk = (int64_t)a * (int64_t)(N + 1) + (int64_t)bb + 1;
    if (bb < 0) {
        k = box_ground(cA, RA, hA, margin, pp, pen, code);
        for (int i = 0; i < k; ++i) nn[i] = mk(0.0f, 0.0f, 1.0f);
    } else {
        m3 RB = ldm(Rw, bb);
        k = box_box(cA, RA, hA, ld3(b.x, bb), RB, ld3(b.half, bb), margin, pp, nn, pen, code,
                    LDSW ? work[threadIdx.x] : own);
    }
    float mu = pair_mu(b.mu, a, bb);
    S.n[p] = k;
    for (int i = 0; i < k; ++i) {
        int s = 4 * p + i;
        S.key[s] = pk * 8192 + code[i];
        st3(S.p, s, pp[i]);
        st3(S.nrm, s, nn[i]);
        S.pen[s] = pen[i];
        S.mu[s] = mu;
    }
}


struct Contacts {
    int64_t *key;
    int *bi, *bj;
    float *pA, *pB, *n, *pen, *mu, *jn, *jt1, *jt2, *jp;
};
struct Pairs {
    int *pbi, *pbj, *pstart, *pcount, *col, *porder;
};

__global__ void k_compact(int cap, int capC, int capP, int *istat, const int *cand_a, const int *cand_b, Slots S,
                          const int *slot_off, const int *pair_off, Contacts Cc, Pairs Pp) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int n = istat[IS_CAND];
    if (n > cap) n = cap;
    if (p >= n) return;
    int k = S.n[p];
    if (k == 0) return;
    int o = slot_off[p], g = pair_off[p];
    if (g >= capP || o + k > capC) {
        atomicOr(&istat[IS_OVF], OVF_CONTACTS);
        return;
    }
    int a = cand_a[p], bb = cand_b[p];
    Pp.pbi[g] = a; Pp.pbj[g] = bb; Pp.pstart[g] = o; Pp.pcount[g] = k;
    for (int i = 0; i < k; ++i) {
        int s = 4 * p + i, c = o + i;
        Cc.key[c] = S.key[s];
        Cc.bi[c] = a; Cc.bj[c] = bb;
        f3 pt = ld3(S.p, s);
        st3(Cc.pA, c, pt); st3(Cc.pB, c, pt);
        st3(Cc.n, c, ld3(S.nrm, s));
        Cc.pen[c] = S.pen[s];
        Cc.mu[c] = S.mu[s];
    }
}

// per-contact colouring (pair_chunks=False, contact_engine_gpu_colored.py:1105-1136): every
// contact is a chunk of its own
__global__ void k_pairs_per_contact(int capC, int *istat, const int *bi, const int *bj, Pairs Pp) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int C = istat[IS_C];
    if (C > capC) C = capC;
    if (c == 0) istat[IS_P] = C;
    if (c >= C) return;
    Pp.pbi[c] = bi[c]; Pp.pbj[c] = bj[c]; Pp.pstart[c] = c; Pp.pcount[c] = 1;
}

// clamp counts after compaction so later kernels stay in bounds
// generated = solved unless a reduction ran (copy_raw)
__global__ void k_clamp_counts(int *istat, int capC, int capP, int copy_raw) {
    if (copy_raw) istat[IS_CRAW] = istat[IS_C];
    if (istat[IS_C] > capC) { istat[IS_C] = capC; atomicOr(&istat[IS_OVF], OVF_CONTACTS); }
    if (istat[IS_P] > capP) { istat[IS_P] = capP; atomicOr(&istat[IS_OVF], OVF_PAIRS); }
}

// ── bead geometry (the reference's model): A4 world points, A5 ground + bead-pair contacts in feature-key order ──
__global__ void k_bead_points(int N, int P, Bodies b, float *allp) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N * P) return;
    int bi = g / P;
    st3(allp, g, ld3(b.x, bi) + cquat_rotate(ld4(b.q, bi), ld3(b.rest, g % P)));
}

// Every bead j > g of another body in the 27 grid cells around bead g (cell coordinates c, cell size
// h) with 1e-9 < |xi - xj| < 2R: fn(j), in an order no caller depends on. The grid holds each item's
// position and id in bucket order (spos); an item is in cell (X, Y, Z) iff floor(p / h) gives
// (X, Y, Z), the grid's own cell rule. The 27 cells form 9 rows along x whose three cells take
// consecutive buckets, so a row is one range of items (two range loads, usually one cache line);
// the 9 ranges and then the first item of each are loaded together. A row whose buckets wrap around
// the end of the table is walked bucket by bucket.
template <class F>
__device__ __forceinline__ void bead_neighbours(int g, int P, f3 xi, int4 c, const float4 *spos, const int *bstart,
                                                const int *bcount, int T, float R, float h, F &&fn) {
    const int bi = g / P;
    const uint32_t mask = (uint32_t)(T - 1);
    int lo[9], hi[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        uint32_t b0 = cell_hash(c.x - 1, c.y + r % 3 - 1, c.z + r / 3 - 1, 0) & mask;
        if (b0 + 2u <= mask) {
            lo[r] = bstart[b0];
            hi[r] = bstart[b0 + 2] + bcount[b0 + 2];
        } else {
            lo[r] = 0;
            hi[r] = -1;
        }
    }
    float4 first[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) first[r] = hi[r] > lo[r] ? spos[lo[r]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    auto test = [&](float4 q, int Y, int Z) {
        int j = __float_as_int(q.w);
        if (j <= g || j / P == bi) return;
        int fx = (int)floorf(q.x / h);
        if (fx < c.x - 1 || fx > c.x + 1 || (int)floorf(q.y / h) != Y || (int)floorf(q.z / h) != Z) return;
        f3 dv = xi - mk(q.x, q.y, q.z);
        float dist = clength(dv);
        if (!(dist < 2.0f * R && dist > 1e-9f)) return;
        fn(j);
    };
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int Y = c.y + r % 3 - 1, Z = c.z + r / 3 - 1;
        if (hi[r] >= 0) {
            for (int t = lo[r]; t < hi[r]; ++t) test(t == lo[r] ? first[r] : spos[t], Y, Z);
        } else {
            uint32_t b0 = cell_hash(c.x - 1, Y, Z, 0) & mask;
            for (uint32_t k = 0; k < 3u; ++k) {
                uint32_t b = (b0 + k) & mask;
                for (int t = bstart[b], e = bstart[b] + bcount[b]; t < e; ++t) test(spos[t], Y, Z);
            }
        }
    }
}

// A5 k_gen over the bead grid: ground contact (z - R < 0) first, then the bead pairs in ascending j
// (key g * (NP+1) + j + 1), at most ME_BEADCAND per bead. The count pass counts; the emit pass
// writes the keys of its matches into its slots, sorts them there (a handful per bead) and then
// fills the contacts in key order.
template <bool EMIT>
__global__ void k_bead_contacts(int NP, int P, const float *allp, const int *cellc, const int *bstart, const int *bcount,
                                const float4 *spos, int T, float R, const float *mu, int *cnt, const int *off, int capC,
                                Contacts Cc, int *istat) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= NP) return;
    f3 xi = ld3(allp, g);
    int bi = g / P;
    int4 cg = *reinterpret_cast<const int4 *>(cellc + 4 * (size_t)g);
    float s = xi.z - R;
    int ground = s < 0.0f ? 1 : 0;
    if (!EMIT) {
        int nc = 0;
        bead_neighbours(g, P, xi, cg, spos, bstart, bcount, T, R, 2.0f * R, [&](int) { ++nc; });
        if (nc > ME_BEADCAND) {
            atomicOr(&istat[IS_OVF], OVF_CAND);
            nc = ME_BEADCAND;
        }
        cnt[g] = ground + nc;
        return;
    }
    int o = off[g];
    int64_t stride = (int64_t)NP + 1, kg = (int64_t)g * stride;
    float m = mu[0];
    int c = o;
    if (ground) {
        if (c < capC) {
            Cc.key[c] = kg; Cc.bi[c] = bi; Cc.bj[c] = -1;
            st3(Cc.pA, c, xi); st3(Cc.pB, c, mk(0.0f, 0.0f, 0.0f)); st3(Cc.n, c, mk(0.0f, 0.0f, 1.0f));
            Cc.pen[c] = -s; Cc.mu[c] = m;
        }
        c++;
    }
    const int c0 = c;
    int nc = 0;
    bead_neighbours(g, P, xi, cg, spos, bstart, bcount, T, R, 2.0f * R, [&](int j) {
        if (nc < ME_BEADCAND) {
            if (c0 + nc < capC) Cc.key[c0 + nc] = kg + (int64_t)j + 1;
            ++nc;
        }
    });
    int nw = c0 + nc <= capC ? nc : (capC > c0 ? capC - c0 : 0);
    int64_t *key = Cc.key + c0;
    for (int a = 1; a < nw; ++a) {
        int64_t k = key[a];
        int b = a;
        while (b > 0 && key[b - 1] > k) { key[b] = key[b - 1]; --b; }
        key[b] = k;
    }
    for (int k = 0; k < nw; ++k) {
        int cc = c0 + k;
        int j = (int)(key[k] - kg - 1);
        f3 xj = ld3(allp, j);
        f3 dv = xi - xj;
        float dist = clength(dv);
        Cc.bi[cc] = bi; Cc.bj[cc] = j / P;
        st3(Cc.pA, cc, xi); st3(Cc.pB, cc, xj); st3(Cc.n, cc, dv / dist);
        Cc.pen[cc] = 2.0f * R - dist; Cc.mu[cc] = m;
    }
    if (o + ground + nc > capC) atomicOr(&istat[IS_OVF], OVF_CONTACTS);
}

// ── A8-A10 relaxed Jacobi with int64 fixed-point accumulation ──
// The reference adds every contact's quantised deltas into six int64 accumulators per body with
// atomics. Here each contact owns one six-int64 entry per body end, in a body-major list built once
// per substep, and stores its deltas there with plain stores; the apply pass sums a body's entries.
// Integer addition is associative, so the sums -- and everything after them -- are the same bits as
// the atomic form in any order, without a device-wide atomic in the iterations.

// _acc (contact_engine_gpu.py:32-38): int64(round(f64(d) * f64(sign * scale))), added to the entry
__device__ __forceinline__ void q3(int64_t *o, f3 d, float sign, float scale) {
    double sc = (double)(sign * scale);
    o[0] += (int64_t)round((double)d.x * sc);
    o[1] += (int64_t)round((double)d.y * sc);
    o[2] += (int64_t)round((double)d.z * sc);
}

// entry e = 6 int64 (linear xyz, angular xyz), 48 bytes: three 16-byte stores
__device__ __forceinline__ void st_entry(int64_t *d, int e, const int64_t *q) {
    longlong2 *p = reinterpret_cast<longlong2 *>(d + 6 * (size_t)e);
    p[0] = make_longlong2(q[0], q[1]);
    p[1] = make_longlong2(q[2], q[3]);
    p[2] = make_longlong2(q[4], q[5]);
}

// Slot of this lane in cnt[key] when lanes with equal keys sit in runs of consecutive lanes (contacts
// are grouped by body): one atomic per run. An inactive lane ends a run. Every lane of the wave must
// call it (no lane may have left the kernel).
__device__ __forceinline__ int wave_run_slot(int *cnt, int key, bool active) {
    const int lane = __lane_id();
    const int kk = active ? key : INT_MIN;
    const int prev = __shfl_up(kk, 1);
    const bool head = lane == 0 || prev != kk;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int h = 63 - __clzll(heads & upto);
    const unsigned long long above = heads & ~upto;
    const int next = above ? __ffsll((long long)above) - 1 : 64;
    int base = 0;
    if (head && active) base = atomicAdd(&cnt[key], next - lane);
    base = __shfl(base, h);
    return base + lane - h;
}

// entry counts and each contact's slot among its bodies' entries (A end, then B end unless ground)
__global__ void k_jac_slots(int capC, const int *istat, const int *bi, const int *bj, int *cnt, int *posA, int *posB) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int C = istat[IS_C] < capC ? istat[IS_C] : capC;
    bool live = c < C;  // no early exit: every lane takes part in the run ballots
    int a = live ? bi[c] : 0, b = live ? bj[c] : -1;
    int sa = wave_run_slot(cnt, a, live);
    int sb = wave_run_slot(cnt, b, live && b >= 0);
    if (live) { posA[c] = sa; posB[c] = b >= 0 ? sb : -1; }
}

// slots -> entry positions (off = exclusive scan of the counts)
__global__ void k_jac_place(int capC, const int *istat, const int *bi, const int *bj, const int *off, int *posA,
                            int *posB) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= istat[IS_C] || c >= capC) return;
    posA[c] += off[bi[c]];
    int b = bj[c];
    if (b >= 0) posB[c] += off[b];
}

struct JacArgs {
    const int *bi, *bj;
    const float *pA, *pB, *n, *pen, *mu;
    float *jn, *jt1, *jt2, *jp;
    const float *x, *invM, *invIw;
    const float *v, *w, *pv, *po;
    const int *posA, *posB;  // entry of each contact at its A and B body
    int64_t *d;              // entries, 6 int64 each
    int N;
};

// k_jac_vel (contact_engine_gpu.py:82-131), contracted forms throughout
__global__ void k_jac_vel(const int *istat, JacArgs A, float scale) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= istat[IS_C]) return;
    int bi = A.bi[c], bj = A.bj[c];
    int ea = A.posA[c], eb = A.posB[c];
    int64_t qa[6] = {0, 0, 0, 0, 0, 0}, qb[6] = {0, 0, 0, 0, 0, 0};
    f3 nrm = ld3(A.n, c), rA = ld3(A.pA, c) - ld3(A.x, bi);
    float iMA = A.invM[bi];
    m3 IA = ldm(A.invIw, bi), IB = IA;
    f3 va = ld3(A.v, bi) + ccross(ld3(A.w, bi), rA);
    f3 vb = mk(0.0f, 0.0f, 0.0f), rB = mk(0.0f, 0.0f, 0.0f);
    float iMB = 0.0f;
    if (bj >= 0) {
        rB = ld3(A.pB, c) - ld3(A.x, bj);
        vb = ld3(A.v, bj) + ccross(ld3(A.w, bj), rB);
        iMB = A.invM[bj];
        IB = ldm(A.invIw, bj);
    }
    f3 vrel = va - vb;
    float vn = cdot(vrel, nrm);
    f3 vt = cmadd(-vn, nrm, vrel);
    float mt = clength(vt);
    f3 t1 = vt / mt;
    if (mt <= 5e-3f) t1 = cfixed_t1(nrm);
    f3 t2 = ccross(nrm, t1);
    f3 rnA = ccross(rA, nrm);
    float meff = iMA + cdot(rnA, cmv(IA, rnA));
    if (bj >= 0) {
        f3 rnB = ccross(rB, nrm);
        meff = meff + iMB + cdot(rnB, cmv(IB, rnB));
    }
    if (meff < 1e-12f) {
        st_entry(A.d, ea, qa);
        if (bj >= 0) st_entry(A.d, eb, qb);
        return;
    }
    float jo = A.jn[c];
    float dj = -vn / meff;
    float nw = wmax(0.0f, jo + dj);
    dj = nw - jo;
    A.jn[c] = nw;
    f3 Jn = nrm * dj;
    q3(qa, Jn * iMA, 1.0f, scale);
    q3(qa + 3, cmv(IA, ccross(rA, Jn)), 1.0f, scale);
    if (bj >= 0) {
        q3(qb, Jn * iMB, -1.0f, scale);
        q3(qb + 3, cmv(IB, ccross(rB, Jn)), -1.0f, scale);
    }
    f3 rtA1 = ccross(rA, t1);
    float meft1 = iMA + cdot(rtA1, cmv(IA, rtA1));
    f3 rtA2 = ccross(rA, t2);
    float meft2 = iMA + cdot(rtA2, cmv(IA, rtA2));
    if (bj >= 0) {
        f3 rtB1 = ccross(rB, t1);
        meft1 = meft1 + iMB + cdot(rtB1, cmv(IB, rtB1));
        f3 rtB2 = ccross(rB, t2);
        meft2 = meft2 + iMB + cdot(rtB2, cmv(IB, rtB2));
    }
    float o1 = A.jt1[c], o2 = A.jt2[c];
    float a1 = o1 - cdot(vrel, t1) / meft1;
    float a2 = o2 - cdot(vrel, t2) / meft2;
    float mag = sqrtf(fmaf(a1, a1, a2 * a2));
    float lim = A.mu[c] * nw;
    if (mag > lim && mag > 1e-12f) { a1 = a1 * lim / mag; a2 = a2 * lim / mag; }
    A.jt1[c] = a1;
    A.jt2[c] = a2;
    f3 Jt = cmadd(a1 - o1, t1, t2 * (0.5f * (a2 - o2)));
    q3(qa, Jt * iMA, 1.0f, scale);
    q3(qa + 3, cmv(IA, ccross(rA, Jt)), 1.0f, scale);
    st_entry(A.d, ea, qa);
    if (bj >= 0) {
        q3(qb, Jt * iMB, -1.0f, scale);
        q3(qb + 3, cmv(IB, ccross(rB, Jt)), -1.0f, scale);
        st_entry(A.d, eb, qb);
    }
}

// k_jac_pos (contact_engine_gpu.py:146-169), contracted forms throughout
__global__ void k_jac_pos(const int *istat, JacArgs A, float dt, float scale) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= istat[IS_C]) return;
    int bi = A.bi[c], bj = A.bj[c];
    int ea = A.posA[c], eb = A.posB[c];
    int64_t qa[6] = {0, 0, 0, 0, 0, 0}, qb[6] = {0, 0, 0, 0, 0, 0};
    f3 nrm = ld3(A.n, c), rA = ld3(A.pA, c) - ld3(A.x, bi);
    float iMA = A.invM[bi];
    m3 IA = ldm(A.invIw, bi), IB = IA;
    f3 rel = ld3(A.pv, bi) + ccross(ld3(A.po, bi), rA);
    f3 rB = mk(0.0f, 0.0f, 0.0f);
    float iMB = 0.0f;
    if (bj >= 0) {
        rB = ld3(A.pB, c) - ld3(A.x, bj);
        rel = rel - (ld3(A.pv, bj) + ccross(ld3(A.po, bj), rB));
        iMB = A.invM[bj];
        IB = ldm(A.invIw, bj);
    }
    f3 rnA = ccross(rA, nrm);
    float meff = iMA + cdot(rnA, cmv(IA, rnA));
    if (bj >= 0) {
        f3 rnB = ccross(rB, nrm);
        meff = meff + iMB + cdot(rnB, cmv(IB, rnB));
    }
    if (meff < 1e-12f) {
        st_entry(A.d, ea, qa);
        if (bj >= 0) st_entry(A.d, eb, qb);
        return;
    }
    float bias = 0.2f * wmax(A.pen[c] - 1e-4f, 0.0f) / dt;
    float dj = (bias - cdot(rel, nrm)) / meff;
    float jo = A.jp[c];
    float nw = wmax(0.0f, jo + dj);
    dj = nw - jo;
    A.jp[c] = nw;
    f3 Jn = nrm * dj;
    q3(qa, Jn * iMA, 1.0f, scale);
    q3(qa + 3, cmv(IA, ccross(rA, Jn)), 1.0f, scale);
    st_entry(A.d, ea, qa);
    if (bj >= 0) {
        q3(qb, Jn * iMB, -1.0f, scale);
        q3(qb + 3, cmv(IB, ccross(rB, Jn)), -1.0f, scale);
        st_entry(A.d, eb, qb);
    }
}

// k_apply (contact_engine_gpu.py:134-143) over the entries: thread (body b, component k) sums b's
// entries in component k, dv = f32(f64(sum) * f64(f32(1/scale))), then v = fma(relax, dv, v). The
// reference zeroes its accumulators here; every entry is rewritten by the next pass instead.
__global__ void k_jac_gather(int N, const int *off, const int *total, const int64_t *d, float *v, float *w,
                             float relax, float scale) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * N) return;
    int b = t / 6, k = t - 6 * b;
    int e0 = off[b], e1 = b + 1 < N ? off[b + 1] : *total;
    const int64_t *p = d + k;
    int64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int e = e0;
    for (; e + 4 <= e1; e += 4) {
        s0 += p[6 * (size_t)e];
        s1 += p[6 * (size_t)(e + 1)];
        s2 += p[6 * (size_t)(e + 2)];
        s3 += p[6 * (size_t)(e + 3)];
    }
    for (; e < e1; ++e) s0 += p[6 * (size_t)e];
    int64_t sum = (s0 + s1) + (s2 + s3);
    float inv = 1.0f / scale;
    float dv = (float)((double)sum * (double)inv);
    float *dst = k < 3 ? v + 3 * (size_t)b + k : w + 3 * (size_t)b + (k - 3);
    *dst = fmaf(relax, dv, *dst);
}

// Jacobi substep set-up: impulses and entry counts to zero; pv, po to zero only when there are
// contacts (the reference zeroes pv/po inside `if C > 0`, contact_engine_gpu.py:266-269)
__global__ void k_jac_reset(int N, int capC, const int *istat, float *pv, float *po, int *cnt, float *jn, float *jt1,
                            float *jt2, float *jp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int C = istat[IS_C];
    if (C <= 0) return;
    if (i < N) {
        st3(pv, i, mk(0.0f, 0.0f, 0.0f));
        st3(po, i, mk(0.0f, 0.0f, 0.0f));
        cnt[i] = 0;
    }
    if (i < C && i < capC) { jn[i] = 0.0f; jt1[i] = 0.0f; jt2[i] = 0.0f; jp[i] = 0.0f; }
}

// ── coloured solvers ──
// A14 warm start by binary search in the previous step's kept keys (contact_engine_gpu_colored.py:602-626)
__device__ __forceinline__ void warm_lookup_at(int c, int capC, const int *istat, int warm, const int64_t *key,
                                               const int64_t *wkey, const float *wjn, const float *wjt1,
                                               const float *wjt2, float *jn, float *jt1, float *jt2, float *jp) {
    int C = istat[IS_C];
    if (c >= C || c >= capC) return;
    jn[c] = 0.0f; jt1[c] = 0.0f; jt2[c] = 0.0f; jp[c] = 0.0f;
    int wn = istat[IS_WN];
    if (!warm || wn == 0) return;
    int64_t k = key[c];
    int lo = 0, hi = wn - 1, hit = -1;
    while (lo <= hi) {
        int mid = (lo + hi) / 2;
        int64_t km = wkey[mid];
        if (km == k) { hit = mid; lo = hi + 1; }
        else if (km < k) lo = mid + 1;
        else hi = mid - 1;
    }
    if (hit >= 0) { jn[c] = wjn[hit]; jt1[c] = wjt1[hit]; jt2[c] = wjt2[hit]; }
}
__global__ void k_warm_lookup(int capC, const int *istat, int warm, const int64_t *key, const int64_t *wkey,
                              const float *wjn, const float *wjt1, const float *wjt2, float *jn, float *jt1, float *jt2,
                              float *jp) {
    warm_lookup_at(blockIdx.x * blockDim.x + threadIdx.x, capC, istat, warm, key, wkey, wjn, wjt1, wjt2, jn, jt1, jt2,
                   jp);
}
// the warm-start match of every contact (k_warm_lookup's binary search), kept for the adjoint
__device__ __forceinline__ void tape_widx_at(int c, int capC, const int *istat, int warm, const int64_t *key,
                                             const int64_t *wkey, int *widx) {
    int C = istat[IS_C];
    if (c >= C || c >= capC) return;
    int wn = istat[IS_WN], hit = -1;
    if (warm && wn > 0) {
        int64_t k = key[c];
        int lo = 0, hi = wn - 1;
        while (lo <= hi) {
            int mid = (lo + hi) / 2;
            int64_t km = wkey[mid];
            if (km == k) { hit = mid; lo = hi + 1; }
            else if (km < k) lo = mid + 1;
            else hi = mid - 1;
        }
    }
    widx[c] = hit;
}
__global__ void k_tape_widx(int capC, const int *istat, int warm, const int64_t *key, const int64_t *wkey, int *widx) {
    tape_widx_at(blockIdx.x * blockDim.x + threadIdx.x, capC, istat, warm, key, wkey, widx);
}

// A23 warm store (contact_engine_gpu_colored.py:1089-1093)
__global__ void k_warm_store(int capC, int *istat, const int64_t *key, const float *jn, const float *jt1,
                             const float *jt2, int64_t *wkey, float *wjn, float *wjt1, float *wjt2) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int C = istat[IS_C];
    if (C > capC) C = capC;
    if (c == 0) istat[IS_WN] = C;
    if (c >= C) return;
    wkey[c] = key[c]; wjn[c] = jn[c]; wjt1[c] = jt1[c]; wjt2[c] = jt2[c];
}

// ColorCache compare (contact_engine_gpu_color_cache.py:7-18, 37-54)
// status word `slot` := 1 when the pair list or a body's invM differs from the colour cache's (called
// by every thread of the grid: a wave ballot)
__device__ __forceinline__ void cc_compare_at(int i, int N, int *istat, int slot, const int *pbi, const int *pbj,
                                              const int *ca, const int *cb, const float *invM, const float *cinvM) {
    int P = istat[IS_P];
    if (P != istat[IS_CACHED_P]) {
        if (i == 0) istat[slot] = 1;
        return;
    }
    bool diff = (i < P && (pbi[i] != ca[i] || pbj[i] != cb[i])) || (i < N && invM[i] != cinvM[i]);
    if (__ballot(diff) && __lane_id() == 0) atomicMax(&istat[slot], 1);
}
__global__ void k_cc_compare(int N, int *istat, const int *pbi, const int *pbj, const int *ca, const int *cb,
                             const float *invM, const float *cinvM) {
    cc_compare_at(blockIdx.x * blockDim.x + threadIdx.x, N, istat, IS_MISMATCH, pbi, pbj, ca, cb, invM, cinvM);
}

// PCG hash (Jarzynski & Olano 2020), behind rand_init(seed, offset) / randi
__device__ __forceinline__ uint32_t pcg(uint32_t s) {
    uint32_t b = s * 747796405u + 2891336453u;
    uint32_t c = ((b >> ((b >> 28u) + 4u)) ^ b) * 277803737u;
    return (c >> 22u) ^ c;
}
// k_prio (contact_engine_gpu_colored.py:137-148)
__device__ __forceinline__ int64_t jp_key(int rank) {
    uint32_t st = pcg(pcg((uint32_t)7919 + pcg((uint32_t)rank)));
    int64_t p = (int64_t)(int32_t)st;
    if (p < 0) p = -p;
    return p * (int64_t)4294967296LL + (int64_t)rank;
}

// A17 JP rounds over the pair graph (k_claim / k_win, contact_engine_gpu_colored.py:151-185)
__global__ void k_jp_init(int capP, const int *istat, int *col) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < istat[IS_P] && p < capP) col[p] = -1;
}
__global__ void k_jp_claim(int capP, const int *istat, const int *col, const int *pbi, const int *pbj,
                           const float *invM, unsigned long long *claim) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= istat[IS_P] || p >= capP || col[p] >= 0) return;
    unsigned long long k = (unsigned long long)jp_key(p);
    int a = pbi[p], b = pbj[p];
    if (invM[a] > 0.0f) atomicMin(&claim[a], k);
    if (b >= 0 && invM[b] > 0.0f) atomicMin(&claim[b], k);
}
__global__ void k_jp_win(int capP, int *istat, int cur, int *col, const int *pbi, const int *pbj, const float *invM,
                         const unsigned long long *claim) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= istat[IS_P] || p >= capP || col[p] >= 0) return;
    unsigned long long k = (unsigned long long)jp_key(p);
    int a = pbi[p], b = pbj[p];
    bool ok = true;
    if (invM[a] > 0.0f && claim[a] != k) ok = false;
    if (b >= 0 && invM[b] > 0.0f && claim[b] != k) ok = false;
    if (ok) col[p] = cur;
    else atomicAdd(&istat[IS_LEFT], 1);
}

// colour class sizes; then (single thread) the split of classes into max(K, JP colours) classes
// colour class sizes: a histogram per workgroup in LDS, then one device atomic per non-empty class per
// workgroup (there are few classes, so per-pair device atomics on their counters would contend)
__global__ void __launch_bounds__(ME_THREADS) k_color_count(int capP, const int *istat, const int *col, int *ccount) {
    __shared__ int h[ME_NCOLMAX + 1];
    for (int c = threadIdx.x; c <= ME_NCOLMAX; c += blockDim.x) h[c] = 0;
    __syncthreads();
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < istat[IS_P] && p < capP) atomicAdd(&h[col[p]], 1);
    __syncthreads();
    for (int c = threadIdx.x; c <= ME_NCOLMAX; c += blockDim.x)
        if (h[c]) atomicAdd(&ccount[c], h[c]);
}
// sub[c] colours for JP class c: the extra K - ncol classes go one at a time to the class with the
// most pairs per sub-colour (lowest class first); base[c] = first final colour of class c
__global__ void k_color_split_plan(int *istat, int K, const int *ccount, int *sub, int *base, int maxcol) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int nc = istat[IS_NCOL];
    for (int c = 0; c < nc; ++c) sub[c] = 1;
    int tot = nc;
    if (K > nc && nc > 0) {
        for (int e = nc; e < K && e < maxcol; ++e) {
            int bc = 0;
            for (int c = 1; c < nc; ++c)
                if ((long long)ccount[c] * sub[bc] > (long long)ccount[bc] * sub[c]) bc = c;
            sub[bc]++;
        }
        tot = 0;
        for (int c = 0; c < nc; ++c) tot += sub[c];
    }
    base[0] = 0;
    for (int c = 0; c < nc; ++c) base[c + 1] = base[c] + sub[c];
    istat[IS_NCOL] = tot;
}
__global__ void k_color_split_apply(int capP, const int *istat, int *col, const int *sub, const int *base) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= istat[IS_P] || p >= capP) return;
    int c = col[p];
    col[p] = base[c] + p % sub[c];
}
// group pairs by colour (order inside a colour is free: its pairs touch disjoint dynamic bodies)
// pairs grouped by colour (any order within a colour: its pairs share no dynamic body); ranks within
// the workgroup from LDS atomics, one device atomic per class per workgroup for the group's base
__global__ void __launch_bounds__(ME_THREADS) k_color_scatter(int capP, const int *istat, const int *col,
                                                              const int *cstart, int *cfill, int *porder) {
    __shared__ int h[ME_NCOLMAX + 1], base[ME_NCOLMAX + 1];
    for (int c = threadIdx.x; c <= ME_NCOLMAX; c += blockDim.x) h[c] = 0;
    __syncthreads();
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    bool live = p < istat[IS_P] && p < capP;
    int c = 0, r = 0;
    if (live) {
        c = col[p];
        r = atomicAdd(&h[c], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k <= ME_NCOLMAX; k += blockDim.x)
        if (h[k]) base[k] = atomicAdd(&cfill[k], h[k]);
    __syncthreads();
    if (live) porder[cstart[c] + base[c] + r] = p;
}
__global__ void k_cc_store(int N, int capP, int *istat, const int *pbi, const int *pbj, const int *col, const float *invM,
                           int *ca, int *cb, int *ccol, float *cinvM, int hit) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int P = istat[IS_P];
    if (i == 0) {
        if (hit) istat[IS_HITS] += 1;
        else { istat[IS_MISSES] += 1; istat[IS_CACHED_P] = P; istat[IS_CACHED_NCOL] = istat[IS_NCOL]; }
    }
    if (hit) return;
    if (i < P && i < capP) { ca[i] = pbi[i]; cb[i] = pbj[i]; ccol[i] = col[i]; }
    if (i < N) cinvM[i] = invM[i];
}
__global__ void k_cc_restore(int capP, int *istat, const int *ccol, int *col) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) istat[IS_NCOL] = istat[IS_CACHED_NCOL];
    if (i < istat[IS_P] && i < capP) col[i] = ccol[i];
}

struct SweepArgs {
    BodyView B;
    ContactView Cv;
    const int *pbi, *pbj, *pstart, *pcount, *porder, *cstart, *ccount;
};

// one launch per colour: one thread per pair of colour ci
__global__ void k_sweep_warm(SweepArgs S, int ci) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.ccount[ci]) return;
    int p = S.porder[S.cstart[ci] + t];
    apply_warm_pair(S.B, S.B, S.Cv, S.pbi[p], S.pbj[p], S.pstart[p], S.pcount[p]);
}
template <bool FIXED>
__global__ void k_sweep_vel(SweepArgs S, int ci, float t2d) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.ccount[ci]) return;
    int p = S.porder[S.cstart[ci] + t];
    chunk_vel<FIXED>(S.B, S.B, S.Cv, S.pbi[p], S.pbj[p], S.pstart[p], S.pcount[p], t2d);
}
__global__ void k_sweep_pos(SweepArgs S, int ci, float dt) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.ccount[ci]) return;
    int p = S.porder[S.cstart[ci] + t];
    chunk_pos(S.B, S.B, S.Cv, S.pbi[p], S.pbj[p], S.pstart[p], S.pcount[p], dt);
}

// A11 integrate (contact_engine_gpu.py:172-180)
__global__ void k_integrate(int N, Bodies b, const float *pv, const float *po, float dt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (b.invM[i] <= 0.0f) return;
    integrate_body(b.x, b.q, ld3(b.v, i), ld3(b.w, i), ld3(pv, i), ld3(po, i), i, dt);
}

}  // namespace me
