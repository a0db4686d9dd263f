// batch_plan.hpp -- device descriptors and the host plan object of the batched
// form (shared by the batched forward and VJP kernels).
#pragma once

#include "agg_tile.hpp"

namespace pbe {

constexpr int kBatchTiles[] = {1, 2, 4, 8};   // columns per lane of the tile groups
constexpr int kBatchGroups = 4;

struct BatchInstDev {
    int n, off, nb;
    int tile_u;          // columns per lane of this instance's pair tiles
    double log_x0, inv_log_r, b0, b_br, b_sh, S0, v_ref, alpha, v_min;
};

struct BatchItem {
    int inst;
    int unit;       // tile index (diagonal-major) or parent block index
    int slice;      // row slice of a diagonal-band tile, -1 for a whole tile
    int pad;
};

}  // namespace pbe

struct pbe_batch_plan_s {
    int count = 0;
    int total = 0;              // max(offset + n) over instances
    int kernel_type = -1;       // -1: no aggregation
    int selection = -1;         // -1: no breakage
    int n_quad = 0;
    int n_brk_items = 0;
    int grid = 0;
    // pair tiles of the instances whose tiles are 64 * kBatchTiles[g] wide,
    // longest first (one persistent launch per non-empty group)
    int n_agg_items[pbe::kBatchGroups] = {};
    double agg_cost[pbe::kBatchGroups] = {};
    pbe::BatchItem* d_agg_items[pbe::kBatchGroups] = {};
    pbe::BatchInstDev* d_inst = nullptr;
    pbe::BatchItem* d_brk_items = nullptr;
    pbe::BatchItem* d_row_items = nullptr;   // (instance, 64-row block) for the VJP
    int n_row_items = 0;
};
