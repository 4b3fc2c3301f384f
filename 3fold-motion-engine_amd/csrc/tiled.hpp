// Island-tiled coloured solve (placeholder until the LDS-resident kernel lands).
#pragma once
#include "kernels.hpp"

namespace me {

struct TiledBuffers {
    int *root;
};

struct TiledArgs {
    int N, capC, capP, solver, vit, pit, warm, n_colors, tile_bodies;
    float t2d, dt;
    Bodies b;
    float *invIw, *pv, *po;
    Contacts Cc;
    Pairs Pp;
    int *cache_a, *cache_b, *cache_col;
    float *cache_invM;
    int64_t *wkey;
    float *wjn, *wjt1, *wjt2;
    int *istat, *bsum;
    int nbsum;
};

template <class BumpT>
inline void tiled_layout(TiledBuffers &t, BumpT &B, int N, int capP, int capC) {
    (void)capP; (void)capC;
    t.root = B.template take<int>((size_t)N);
}

inline int tiled_solve(TiledBuffers &, const TiledArgs &, hipStream_t) { return ME_ERUNTIME; }

}  // namespace me

struct me_engine;
me::TiledArgs tiled_args_of(me_engine *e, float sdt);
