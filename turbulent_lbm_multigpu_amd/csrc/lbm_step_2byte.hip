// lbm_step_2byte.hip -- the step kernels of the two-byte storages (FP16 and
// BF16 holding f - w_i), in a translation unit of their own.
//
// They are compiled with floating-point contraction: the compiler may fuse a
// product and a sum into one FMA across statements.  These kernels move half
// the bytes of FP32 per cell with about the same per-cell instruction count,
// so their vector-instruction issue runs near the SIMDs' limit, and the
// fused collision (9-12 % fewer vector instructions) runs 2-3 % faster.  The FP32 and FP64 kernels (lbm_runtime.hip)
// keep every rounding of the reference's expressions (no contraction): their
// bit-for-bit parity with the CPU oracle depends on it.  The two-byte
// storages have no such counterpart; their results are bounded against the
// FP64 run, and every launch shape of them gives bit-identical results (the
// fused collision is the same code in every path), which the tests check.
#pragma clang fp contract(fast)

#include <hip/hip_runtime.h>

#include "../../include/lbm_capi.h"
#include "lbm_device.hpp"

namespace lbm {

hipError_t launch_step_2byte(int storage, bool beta, int variant, int lds, int tiles, int wg, const DomDev *descs,
                             const int *tile_table, hipStream_t stream, long long *ctr, const XChain *chains,
                             int xcd_chunk) {
    if (storage == LBM_STORAGE_F16_SHIFTED)
        return launch_step_storage<2>(beta, variant, lds, tiles, wg, descs, tile_table, stream, ctr, chains, xcd_chunk);
    return launch_step_storage<3>(beta, variant, lds, tiles, wg, descs, tile_table, stream, ctr, chains, xcd_chunk);
}

}  // namespace lbm
