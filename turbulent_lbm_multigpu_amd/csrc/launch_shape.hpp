// launch_shape.hpp -- the launch shapes the fused step kernels are built for.
// One definition shared by the solver, the group and the server, so that all
// three accept the same set.
#pragma once

// kernel_variant bits: 1 non-temporal loads, 2 non-temporal stores, 4 no XCD
// tile map, 8 / 16 two / four x-consecutive cells per work-item.  Built:
// 0-3, 8, 11 (two cells with the non-temporal policy; the two-byte storages
// keep each pair packed in one 32-bit word) and 16, each optionally | 4.
inline bool lbm_variant_supported(int variant) {
    if (variant < 0 || variant > 31) return false;
    const int v = variant & ~4;
    return v == 0 || v == 1 || v == 2 || v == 3 || v == 8 || v == 11 || v == 16;
}

// work-items per work-group: whole wavefronts of 64, at most 512
inline bool lbm_work_group_supported(int wg) { return wg >= 64 && wg <= 512 && wg % 64 == 0; }
