// igemm_f32.hip -- FP32 (v_mfma_f32_32x32x2f32) instantiations of the implicit-GEMM kernel and
// its tile-configuration table (BM, BN, BK, waves along M, waves along N).
#include "igemm_kernel.hpp"

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
};
constexpr int kNum = sizeof(kCfg) / sizeof(kCfg[0]);
// configurations that also have element-load (non-16-byte) instantiations
const bool kNoVec[kNum] = {true, false, false, true, false, false, true, false, false, false, false, false};

}  // namespace

extern "C" int cg_f32_num_cfgs(void) { return kNum; }

extern "C" int cg_f32_get_cfg(int cfg, cg_tilecfg *out) {
    if (cfg < 0 || cfg >= kNum) return -1;
    *out = kCfg[cfg];
    return 0;
}

extern "C" int cg_f32_supports(int cfg, int op, int vec) { return supports<float>(kCfg, kNoVec, kNum, cfg, op, vec); }

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
