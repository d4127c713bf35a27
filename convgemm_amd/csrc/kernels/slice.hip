// slice.hip -- one instantiation slice of the implicit-GEMM kernels: every tile
// configuration of one operand type for one operation and one kernel family
// (tile_tables.hpp).  The build compiles this file once per slice with
//   -DCG_SLICE_T=<float|uint16_t|uint8_t> -DCG_SLICE_TN=<f32|bf16|fp8>
//   -DCG_SLICE_OP=<cg_opkind> -DCG_SLICE_FAM=<0: general, 1: pipelined>
// so the slices compile in parallel; tiles.hip dispatches to them.  Without
// those definitions the file is empty.
#ifdef CG_SLICE_T
#include "igemm_kernel.hpp"
#include "lp_kernel.hpp"
#include "lp_nchw_kernel.hpp"
#include "ht_kernel.hpp"
#include "lpw_kernel.hpp"
#include "tile_tables.hpp"

using namespace cg;

namespace {

template <typename T, int OP, int FAM, int I = 0>
int launch_slice(int cfg, const cg_kparams *kp, const cg_group *g, int *occ, hipStream_t st) {
    if constexpr (I == num_tiles<T>()) {
        return (int)hipErrorInvalidValue;
    } else {
        if (cfg == I) {
            constexpr TileDef d = Tiles<T>::defs[I];
            if constexpr (FAM == 0 && d.fam == FAM_GEN)
                return launch_op<T, d.BM, d.BN, d.BK, d.WM, d.WN, d.nv, OP>(kp, g, occ, st);
            else if constexpr (FAM == 1 && d.fam == FAM_LP && OP == CG_FWD_NHWC)
                return launch_lp<T, d.BM, d.BN, d.WM, d.WN, d.NS>(kp, g, occ, st);
            else if constexpr (FAM == 1 && d.fam == FAM_LPN && OP == CG_FWD_NCHW && std::is_same<T, float>::value)
                return launch_lpn<d.BM, d.BN, d.WM, d.WN, d.NS>(kp, g, occ, st);
            else if constexpr (FAM == 1 && d.fam == FAM_HT && OP == CG_FWD_NHWC)
                return launch_ht<T, d.BM, d.BN, d.WM, d.WN, d.NS, d.HS>(kp, g, occ, st);
            else if constexpr (FAM == 1 && d.fam == FAM_LPW && OP == CG_WGRAD_NHWC && sizeof(T) != 1)
                return launch_lpw<T, d.BM, d.BN, d.WM, d.WN, d.NS>(kp, g, occ, st);
            else
                return (int)hipErrorInvalidValue;
        }
        return launch_slice<T, OP, FAM, I + 1>(cfg, kp, g, occ, st);
    }
}

}  // namespace

#define CG_SLICE_NAME3(tn, op, fam) cg_slice_##tn##_##op##_##fam
#define CG_SLICE_NAME2(tn, op, fam) CG_SLICE_NAME3(tn, op, fam)
#define CG_SLICE_NAME CG_SLICE_NAME2(CG_SLICE_TN, CG_SLICE_OP, CG_SLICE_FAM)

extern "C" int CG_SLICE_NAME(int cfg, const cg_kparams *kp, const cg_group *g, int *occ, void *stream) {
    return launch_slice<CG_SLICE_T, CG_SLICE_OP, CG_SLICE_FAM>(cfg, kp, g, occ, (hipStream_t)stream);
}
#endif
