// Per-cage launches of the patch-local level step (cc_patch.h).
#include "cc_patch.h"

namespace cch {

// tiles [lb*hpt, lb*hpt + hpt) of the level, lb the XCD-contiguous block index
template <int GATHER, int CREASED, int TOPO, int UVMODE, int NT>
__global__ __launch_bounds__(NT) void k_level_patch(FusedArgs a, DCage c, LevelTab t, LevelPtrs lp, int d,
                                                    uint32_t tiles, uint32_t hpt) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, a.xcd_remap);
    for (uint32_t it = 0; it < hpt; ++it) {
        const uint32_t tile = lb * hpt + it;
        if (tile >= tiles) break;
        if (it) __syncthreads();   // the previous tile's LDS reads are done
        patch_tile<GATHER, CREASED, TOPO, UVMODE, NT>(a, c, t, lp, d, tile);
    }
}

// kDeferGroup mask words per wave
template <int GATHER, int CREASED>
__global__ __launch_bounds__(256) void k_level_deferred(FusedArgs a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    const int32_t words = (a.H + 63) >> 6;
    const int32_t wave = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    deferred_group<GATHER, CREASED>(a, c, t, lp, d, wave * kDeferGroup, words);
}

namespace {

template <int G, int C, int T, int U, int NT>
void launch_patch_nt(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d, uint32_t hpt,
                     hipStream_t st) {
    const uint32_t tiles = (uint32_t)((a.H + NT - 1) / NT);
    const uint32_t blocks = (tiles + hpt - 1) / hpt;
    hipLaunchKernelGGL((k_level_patch<G, C, T, U, NT>), dim3(blocks), dim3(NT), 0, st, a, c, t, lp, d, tiles, hpt);
}

template <int G, int C, int T, int U>
void launch_patch_u(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d, int nt,
                    uint32_t hpt, hipStream_t st) {
    if (nt == 128) launch_patch_nt<G, C, T, U, 128>(a, c, t, lp, d, hpt, st);
    else if (nt == 512) launch_patch_nt<G, C, T, U, 512>(a, c, t, lp, d, hpt, st);
    else launch_patch_nt<G, C, T, U, 256>(a, c, t, lp, d, hpt, st);
}

template <int G, int C>
void launch_patch_gc(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d, int topo,
                     int uvMode, int nt, uint32_t hpt, hipStream_t st) {
    if (topo) {
        if (uvMode == 1) launch_patch_u<G, C, 1, 1>(a, c, t, lp, d, nt, hpt, st);
        else if (uvMode == 2) launch_patch_u<G, C, 1, 2>(a, c, t, lp, d, nt, hpt, st);
        else launch_patch_u<G, C, 1, 0>(a, c, t, lp, d, nt, hpt, st);
    } else if (uvMode == 1) {
        launch_patch_u<G, C, 0, 1>(a, c, t, lp, d, nt, hpt, st);
    } else {
        launch_patch_u<G, C, 0, 0>(a, c, t, lp, d, nt, hpt, st);
    }
}

}  // namespace

void launch_level_patch(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d, int gather,
                        int creased, int topo, int uvMode, int nt, int hpt, hipStream_t st) {
    const uint32_t h = hpt < 1 ? 1u : (uint32_t)hpt;
    if (gather && creased) launch_patch_gc<1, 1>(a, c, t, lp, d, topo, uvMode, nt, h, st);
    else if (gather) launch_patch_gc<1, 0>(a, c, t, lp, d, topo, uvMode, nt, h, st);
    else if (creased) launch_patch_gc<0, 1>(a, c, t, lp, d, topo, uvMode, nt, h, st);
    else launch_patch_gc<0, 0>(a, c, t, lp, d, topo, uvMode, nt, h, st);
}

void launch_level_deferred(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d,
                           int gather, int creased, hipStream_t st) {
    const int64_t words = ((int64_t)a.H + 63) / 64;
    const int64_t waves = (words + kDeferGroup - 1) / kDeferGroup;
    const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    if (gather && creased) hipLaunchKernelGGL((k_level_deferred<1, 1>), grid, block, 0, st, a, c, t, lp, d);
    else if (gather) hipLaunchKernelGGL((k_level_deferred<1, 0>), grid, block, 0, st, a, c, t, lp, d);
    else if (creased) hipLaunchKernelGGL((k_level_deferred<0, 1>), grid, block, 0, st, a, c, t, lp, d);
    else hipLaunchKernelGGL((k_level_deferred<0, 0>), grid, block, 0, st, a, c, t, lp, d);
}

}  // namespace cch
