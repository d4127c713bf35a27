// igemm_bf16.hip -- BF16 (v_mfma_f32_32x32x16_bf16) instantiations of the implicit-GEMM kernel and
// its tile-configuration table (BM, BN, BK, waves along M, waves along N, pipelined stages).
#include "igemm_kernel.hpp"
#include "lp_kernel.hpp"

using namespace cg;

namespace {

const cg_tilecfg kCfg[] = {
    {128, 128, 32, 2, 2},  // 0
    {128, 64, 32, 2, 2},  // 1
    {64, 128, 32, 2, 2},  // 2
    {64, 64, 32, 2, 2},  // 3
    {256, 64, 32, 4, 1},  // 4
    {256, 128, 32, 2, 2},  // 5
    {128, 128, 64, 2, 2},  // 6
    {64, 64, 64, 2, 2},  // 7
    {128, 256, 32, 2, 2},  // 8
    {128, 128, 64, 2, 2, 4},  // 9 pipelined, 4 LDS stages
    {256, 64, 64, 4, 1, 3},  // 10 pipelined, 3 LDS stages
    {128, 64, 64, 2, 2, 4},  // 11 pipelined, 4 LDS stages
    {256, 128, 64, 2, 2, 3},  // 12 pipelined, 3 LDS stages
    {64, 64, 64, 2, 2, 4},  // 13 pipelined, 4 LDS stages
    {128, 256, 64, 2, 2, 3},  // 14 pipelined, 3 LDS stages
    {128, 64, 64, 2, 2, 3},  // 15 pipelined, 3 LDS stages
    {64, 128, 64, 2, 2, 3},  // 16 pipelined, 3 LDS stages
    {128, 128, 64, 2, 2, 2},  // 17 pipelined, 2 LDS stages
    {256, 64, 64, 4, 1, 2},  // 18 pipelined, 2 LDS stages
    {128, 64, 64, 2, 2, 2},  // 19 pipelined, 2 LDS stages
    {64, 128, 64, 2, 2, 2},  // 20 pipelined, 2 LDS stages
    {256, 128, 64, 2, 2, 2},  // 21 pipelined, 2 LDS stages
    {64, 64, 64, 2, 2, 2},  // 22 pipelined, 2 LDS stages
    {256, 128, 64, 4, 2, 2},  // 23 pipelined, 8 waves, 2 LDS stages
    {256, 128, 64, 4, 2, 3},  // 24 pipelined, 8 waves, 3 LDS stages
    {256, 256, 64, 4, 2, 2},  // 25 pipelined, 8 waves, 2 LDS stages
    {128, 256, 64, 2, 4, 2},  // 26 pipelined, 8 waves, 2 LDS stages
    {512, 64, 64, 8, 1, 2},  // 27 pipelined, 8 waves, 2 LDS stages
    {256, 64, 64, 4, 2, 3},  // 28 pipelined, 8 waves, 3 LDS stages
    {128, 128, 64, 2, 4, 4},  // 29 pipelined, 8 waves, 4 LDS stages
    {64, 64, 128, 2, 2},  // 30 long K step (weight gradient: fewer barriers per MFMA)
};
constexpr int kNum = sizeof(kCfg) / sizeof(kCfg[0]);
// configurations that also have element-load (non-16-byte) instantiations
const bool kNoVec[kNum] = {true, false, false, true, false, false, false, true, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false};
// configurations served by the pipelined forward kernel (lp_kernel.hpp)
const bool kLp[kNum] = {false, false, false, false, false, false, false, false, false, true, true, true, true, true, true, true, true, true, true, true, true, true, true, true, true, true, true, true, true, true, false};

}  // namespace

extern "C" int cg_bf16_num_cfgs(void) { return kNum; }

extern "C" int cg_bf16_get_cfg(int cfg, cg_tilecfg *out) {
    if (cfg < 0 || cfg >= kNum) return -1;
    *out = kCfg[cfg];
    return 0;
}

extern "C" int cg_bf16_supports(int cfg, int op, int vec) { return supports<uint16_t>(kCfg, kNoVec, kNum, cfg, op, vec, kLp); }

static int launch_any(int cfg, const cg_kparams *kp, const cg_group *g, int *occ, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    switch (cfg) {
    case 0: return launch_cfg<uint16_t, 128, 128, 32, 2, 2, true>(kp, g, occ, st);
    case 1: return launch_cfg<uint16_t, 128, 64, 32, 2, 2, false>(kp, g, occ, st);
    case 2: return launch_cfg<uint16_t, 64, 128, 32, 2, 2, false>(kp, g, occ, st);
    case 3: return launch_cfg<uint16_t, 64, 64, 32, 2, 2, true>(kp, g, occ, st);
    case 4: return launch_cfg<uint16_t, 256, 64, 32, 4, 1, false>(kp, g, occ, st);
    case 5: return launch_cfg<uint16_t, 256, 128, 32, 2, 2, false>(kp, g, occ, st);
    case 6: return launch_cfg<uint16_t, 128, 128, 64, 2, 2, false>(kp, g, occ, st);
    case 7: return launch_cfg<uint16_t, 64, 64, 64, 2, 2, true>(kp, g, occ, st);
    case 8: return launch_cfg<uint16_t, 128, 256, 32, 2, 2, false>(kp, g, occ, st);
    case 9: return launch_lp<uint16_t, 128, 128, 2, 2, 4>(kp, g, occ, st);
    case 10: return launch_lp<uint16_t, 256, 64, 4, 1, 3>(kp, g, occ, st);
    case 11: return launch_lp<uint16_t, 128, 64, 2, 2, 4>(kp, g, occ, st);
    case 12: return launch_lp<uint16_t, 256, 128, 2, 2, 3>(kp, g, occ, st);
    case 13: return launch_lp<uint16_t, 64, 64, 2, 2, 4>(kp, g, occ, st);
    case 14: return launch_lp<uint16_t, 128, 256, 2, 2, 3>(kp, g, occ, st);
    case 15: return launch_lp<uint16_t, 128, 64, 2, 2, 3>(kp, g, occ, st);
    case 16: return launch_lp<uint16_t, 64, 128, 2, 2, 3>(kp, g, occ, st);
    case 17: return launch_lp<uint16_t, 128, 128, 2, 2, 2>(kp, g, occ, st);
    case 18: return launch_lp<uint16_t, 256, 64, 4, 1, 2>(kp, g, occ, st);
    case 19: return launch_lp<uint16_t, 128, 64, 2, 2, 2>(kp, g, occ, st);
    case 20: return launch_lp<uint16_t, 64, 128, 2, 2, 2>(kp, g, occ, st);
    case 21: return launch_lp<uint16_t, 256, 128, 2, 2, 2>(kp, g, occ, st);
    case 22: return launch_lp<uint16_t, 64, 64, 2, 2, 2>(kp, g, occ, st);
    case 23: return launch_lp<uint16_t, 256, 128, 4, 2, 2>(kp, g, occ, st);
    case 24: return launch_lp<uint16_t, 256, 128, 4, 2, 3>(kp, g, occ, st);
    case 25: return launch_lp<uint16_t, 256, 256, 4, 2, 2>(kp, g, occ, st);
    case 26: return launch_lp<uint16_t, 128, 256, 2, 4, 2>(kp, g, occ, st);
    case 27: return launch_lp<uint16_t, 512, 64, 8, 1, 2>(kp, g, occ, st);
    case 28: return launch_lp<uint16_t, 256, 64, 4, 2, 3>(kp, g, occ, st);
    case 29: return launch_lp<uint16_t, 128, 128, 2, 4, 4>(kp, g, occ, st);
    case 30: return launch_cfg<uint16_t, 64, 64, 128, 2, 2, false>(kp, g, occ, st);
    }
    return (int)hipErrorInvalidValue;
}

extern "C" int cg_bf16_launch(int cfg, const cg_kparams *kp, void *stream) {
    return launch_any(cfg, kp, nullptr, nullptr, stream);
}

extern "C" int cg_bf16_launch_grouped(int cfg, const cg_kparams *kp0, const cg_group *g, void *stream) {
    return launch_any(cfg, kp0, g, nullptr, stream);
}

extern "C" int cg_bf16_occupancy(int cfg, const cg_kparams *kp0) {
    int occ = 0;
    if (launch_any(cfg, kp0, nullptr, &occ, nullptr) != 0) return 0;
    return occ;
}
