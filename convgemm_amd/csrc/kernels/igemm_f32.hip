// igemm_f32.hip -- FP32 (v_mfma_f32_32x32x2f32) instantiations of the implicit-GEMM kernel and
// its tile-configuration table (BM, BN, BK, waves along M, waves along N, pipelined stages).
#include "igemm_kernel.hpp"
#include "lp_kernel.hpp"
#include "lp_nchw_kernel.hpp"

using namespace cg;

namespace {

const cg_tilecfg kCfg[] = {
    {128, 128, 32, 2, 2},  // 0
    {128, 64, 32, 2, 2},  // 1
    {64, 128, 32, 2, 2},  // 2
    {64, 64, 32, 2, 2},  // 3
    {256, 64, 32, 4, 1},  // 4
    {128, 128, 16, 2, 2},  // 5
    {64, 64, 16, 2, 2},  // 6
    {64, 256, 32, 1, 4},  // 7
    {256, 64, 16, 4, 1},  // 8
    {64, 256, 16, 1, 4},  // 9
    {256, 128, 16, 2, 2},  // 10
    {128, 64, 16, 2, 2},  // 11
    {128, 128, 32, 2, 2, 3},  // 12 pipelined, 3 LDS stages
    {128, 128, 32, 2, 2, 2},  // 13 pipelined, 2 LDS stages
    {256, 64, 32, 4, 1, 3},  // 14 pipelined, 3 LDS stages
    {128, 64, 32, 2, 2, 4},  // 15 pipelined, 4 LDS stages
    {64, 128, 32, 2, 2, 4},  // 16 pipelined, 4 LDS stages
    {256, 128, 32, 2, 2, 3},  // 17 pipelined, 3 LDS stages
    {128, 256, 32, 2, 2, 3},  // 18 pipelined, 3 LDS stages
    {64, 64, 32, 2, 2, 4},  // 19 pipelined, 4 LDS stages
    {256, 64, 32, 4, 1, 2},  // 20 pipelined, 2 LDS stages
    {128, 64, 32, 2, 2, 2},  // 21 pipelined, 2 LDS stages
    {128, 128, 32, 2, 2, 3},  // 22 pipelined NCHW, 3 LDS stages
    {128, 128, 32, 2, 2, 2},  // 23 pipelined NCHW, 2 LDS stages
    {128, 64, 32, 2, 2, 3},  // 24 pipelined NCHW, 3 LDS stages
    {64, 128, 32, 2, 2, 3},  // 25 pipelined NCHW, 3 LDS stages
    {64, 64, 32, 2, 2, 4},  // 26 pipelined NCHW, 4 LDS stages
    {256, 64, 32, 4, 1, 2},  // 27 pipelined NCHW, 2 LDS stages
    {256, 128, 32, 4, 2, 2},  // 28 pipelined, 8 waves, 2 LDS stages
    {128, 128, 32, 2, 4, 3},  // 29 pipelined, 8 waves, 3 LDS stages
    {256, 64, 32, 4, 2, 3},  // 30 pipelined, 8 waves, 3 LDS stages
    {64, 64, 64, 2, 2},  // 31 long K step (weight gradient: fewer barriers per MFMA)
};
constexpr int kNum = sizeof(kCfg) / sizeof(kCfg[0]);
// configurations that also have element-load (non-16-byte) instantiations
const bool kNoVec[kNum] = {true, false, false, true, false, false, true, false, false, false, false, false,
                           false, false, false, false, false, false, false, false, false, false,
                           false, false, false, false, false, false, false, false, false, false};
// configurations served by the pipelined forward kernel (lp_kernel.hpp)
const bool kLp[kNum] = {false, false, false, false, false, false, false, false, false, false, false, false,
                        true, true, true, true, true, true, true, true, true, true,
                        false, false, false, false, false, false, true, true, true, false};
// configurations served by the pipelined forward NCHW kernel (lp_nchw_kernel.hpp)
const bool kLpn[kNum] = {false, false, false, false, false, false, false, false, false, false, false, false,
                         false, false, false, false, false, false, false, false, false, false,
                         true, true, true, true, true, true, false, false, false, false};

}  // namespace

extern "C" int cg_f32_num_cfgs(void) { return kNum; }

extern "C" int cg_f32_get_cfg(int cfg, cg_tilecfg *out) {
    if (cfg < 0 || cfg >= kNum) return -1;
    *out = kCfg[cfg];
    return 0;
}

extern "C" int cg_f32_supports(int cfg, int op, int vec) { return supports<float>(kCfg, kNoVec, kNum, cfg, op, vec, kLp, kLpn); }

static int launch_any(int cfg, const cg_kparams *kp, const cg_group *g, int *occ, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    switch (cfg) {
    case 0: return launch_cfg<float, 128, 128, 32, 2, 2, true>(kp, g, occ, st);
    case 1: return launch_cfg<float, 128, 64, 32, 2, 2, false>(kp, g, occ, st);
    case 2: return launch_cfg<float, 64, 128, 32, 2, 2, false>(kp, g, occ, st);
    case 3: return launch_cfg<float, 64, 64, 32, 2, 2, true>(kp, g, occ, st);
    case 4: return launch_cfg<float, 256, 64, 32, 4, 1, false>(kp, g, occ, st);
    case 5: return launch_cfg<float, 128, 128, 16, 2, 2, false>(kp, g, occ, st);
    case 6: return launch_cfg<float, 64, 64, 16, 2, 2, true>(kp, g, occ, st);
    case 7: return launch_cfg<float, 64, 256, 32, 1, 4, false>(kp, g, occ, st);
    case 8: return launch_cfg<float, 256, 64, 16, 4, 1, false>(kp, g, occ, st);
    case 9: return launch_cfg<float, 64, 256, 16, 1, 4, false>(kp, g, occ, st);
    case 10: return launch_cfg<float, 256, 128, 16, 2, 2, false>(kp, g, occ, st);
    case 11: return launch_cfg<float, 128, 64, 16, 2, 2, false>(kp, g, occ, st);
    case 12: return launch_lp<float, 128, 128, 2, 2, 3>(kp, g, occ, st);
    case 13: return launch_lp<float, 128, 128, 2, 2, 2>(kp, g, occ, st);
    case 14: return launch_lp<float, 256, 64, 4, 1, 3>(kp, g, occ, st);
    case 15: return launch_lp<float, 128, 64, 2, 2, 4>(kp, g, occ, st);
    case 16: return launch_lp<float, 64, 128, 2, 2, 4>(kp, g, occ, st);
    case 17: return launch_lp<float, 256, 128, 2, 2, 3>(kp, g, occ, st);
    case 18: return launch_lp<float, 128, 256, 2, 2, 3>(kp, g, occ, st);
    case 19: return launch_lp<float, 64, 64, 2, 2, 4>(kp, g, occ, st);
    case 20: return launch_lp<float, 256, 64, 4, 1, 2>(kp, g, occ, st);
    case 21: return launch_lp<float, 128, 64, 2, 2, 2>(kp, g, occ, st);
    case 22: return launch_lpn<128, 128, 2, 2, 3>(kp, g, occ, st);
    case 23: return launch_lpn<128, 128, 2, 2, 2>(kp, g, occ, st);
    case 24: return launch_lpn<128, 64, 2, 2, 3>(kp, g, occ, st);
    case 25: return launch_lpn<64, 128, 2, 2, 3>(kp, g, occ, st);
    case 26: return launch_lpn<64, 64, 2, 2, 4>(kp, g, occ, st);
    case 27: return launch_lpn<256, 64, 4, 1, 2>(kp, g, occ, st);
    case 28: return launch_lp<float, 256, 128, 4, 2, 2>(kp, g, occ, st);
    case 29: return launch_lp<float, 128, 128, 2, 4, 3>(kp, g, occ, st);
    case 30: return launch_lp<float, 256, 64, 4, 2, 3>(kp, g, occ, st);
    case 31: return launch_cfg<float, 64, 64, 64, 2, 2, false>(kp, g, occ, st);
    }
    return (int)hipErrorInvalidValue;
}

extern "C" int cg_f32_launch(int cfg, const cg_kparams *kp, void *stream) {
    return launch_any(cfg, kp, nullptr, nullptr, stream);
}

extern "C" int cg_f32_launch_grouped(int cfg, const cg_kparams *kp0, const cg_group *g, void *stream) {
    return launch_any(cfg, kp0, g, nullptr, stream);
}

extern "C" int cg_f32_occupancy(int cfg, const cg_kparams *kp0) {
    int occ = 0;
    if (launch_any(cfg, kp0, nullptr, &occ, nullptr) != 0) return 0;
    return occ;
}
