// launch_shape.hpp -- the launch shapes the fused step kernels are built for.
// One definition shared by the solver, the group and the server, so that all
// three accept the same set.
#pragma once

// kernel_variant bits: 1 non-temporal loads, 2 non-temporal stores, 4 no XCD
// tile map, 8 / 16 two / four x-consecutive cells per work-item, 32 the XCD
// tile map in chunks (about kXcdChunkTiles consecutive tiles per XCD in turn;
// a group matches the count to its members' z planes) instead of one
// contiguous eighth per XCD.  Built: 0-3, 8, 11 (two cells
// with the non-temporal policy; the two-byte storages keep each pair packed
// in one 32-bit word) and 16, each optionally | 4 or | 32 (not both).
constexpr int kVarNoXcdMap = 4;
constexpr int kVarXcdChunks = 32;
constexpr int kXcdChunkTiles = 16;
inline int lbm_variant_map_bits(int variant) { return variant & (kVarNoXcdMap | kVarXcdChunks); }
inline bool lbm_variant_supported(int variant) {
    if (variant < 0 || variant > 63) return false;
    if (lbm_variant_map_bits(variant) == (kVarNoXcdMap | kVarXcdChunks)) return false;
    const int v = variant & ~(kVarNoXcdMap | kVarXcdChunks);
    return v == 0 || v == 1 || v == 2 || v == 3 || v == 8 || v == 11 || v == 16;
}

// work-items per work-group: whole wavefronts of 64, at most 512
inline bool lbm_work_group_supported(int wg) { return wg >= 64 && wg <= 512 && wg % 64 == 0; }
