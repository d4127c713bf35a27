// single.hip -- the single-launch batch kernel: the forward NHWC problems of
// a batch (16-byte channel chunks) in ONE persistent launch.
//
// The workgroups walk a device table of problems: problem j owns work items
// [start[j], start[j+1]) (its tiles, padded to a multiple of 8) and runs on
// tile shape var[j] of SingleSet<T> (tile_tables.hpp), the pipelined tile of
// lp_kernel.hpp with its stages in one LDS buffer that every shape shares.
// The weights are read repacked ([kn][kh][kw][c], the batch's element-wise
// launch or its static repack): reading the caller's order in place made
// this kernel 1.5x slower on the 256-problem BF16 batch.
// Every shape holds 4 waves and at most SINGLE_LDS bytes of stages, so two
// workgroups are resident per CU whatever mix of shapes they run.
//
// The host orders the problems by the MFMA work of a tile, longest first, and
// launches as many workgroups as are resident at once.  Round r hands item
// r*G + w to workgroup slot w; odd rounds reverse the slots in groups of 8,
// so the longest-first order evens out across workgroups while an item keeps
// the XCD (blockIdx mod 8) that xcd_run assigned its tile to.
//
// The build compiles this file once per operand type with
//   -DCG_SINGLE_T=<float|uint16_t|uint8_t> -DCG_SINGLE_TN=<f32|bf16|fp8>
// Without those definitions the file is empty.
#ifdef CG_SINGLE_T
#include "lp_kernel.hpp"
#include "tile_tables.hpp"

using namespace cg;

namespace {

template <typename T, int V = 0>
__device__ __attribute__((always_inline)) inline void single_item(int v, const cg_kparams &p, int lin, char *lds) {
    if constexpr (V < num_single<T>()) {
        if (v == V) {
            constexpr TileDef d = Tiles<T>::defs[SingleSet<T>::ids[V]];
            static_assert(d.fam == FAM_LP && d.WM * d.WN == 4 && d.NS * (d.BM + d.BN) * 128 <= SINGLE_LDS,
                          "single-launch shapes: 4-wave pipelined tiles within the shared LDS buffer");
            lp_tile<T, d.BM, d.BN, d.WM, d.WN, d.NS, true>(p, lin, -1, 0, nullptr, lds);
            return;
        }
        single_item<T, V + 1>(v, p, lin, lds);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) single_kernel(const cg_kparams *__restrict__ plist,
                                                     const int *__restrict__ start, const int *__restrict__ var,
                                                     int nprob, int total) {
    __shared__ __attribute__((aligned(1024))) char lds[SINGLE_LDS];
    const int G = (int)gridDim.x, b = (int)blockIdx.x;  // G is a multiple of 8
    for (long long r = 0;; r++) {
        const int w = (r & 1) ? (G / 8 - 1 - b / 8) * 8 + (b & 7) : b;
        const long long item = r * G + w;
        if (item >= total) break;  // later rounds hold larger items only
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const cg_kparams &p = plist[lo];
        const int lin = xcd_run((int)(item - start[lo]), start[lo + 1] - start[lo]);
        __syncthreads();  // every LDS read and load of the previous item is complete
        if (lin < p.mtiles * p.ntiles) single_item<T>(var[lo], p, lin, lds);
    }
}

}  // namespace

#define CG_SINGLE_NAME2(tn) cg_single_##tn
#define CG_SINGLE_NAME1(tn) CG_SINGLE_NAME2(tn)
#define CG_SINGLE_NAME CG_SINGLE_NAME1(CG_SINGLE_TN)

// occ != nullptr: report resident workgroups per CU instead of launching.
extern "C" int CG_SINGLE_NAME(const cg_group *g, const int *var, int *occ, void *stream) {
    if (occ)
        return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            occ, reinterpret_cast<const void *>(&single_kernel<CG_SINGLE_T>), 256, 0);
    if (g->total <= 0) return 0;
    if (g->grid <= 0 || g->grid % 8) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(single_kernel<CG_SINGLE_T>, dim3((unsigned)g->grid), dim3(256), 0, (hipStream_t)stream,
                       g->plist, g->start, var, g->nprob, g->total);
    return (int)hipGetLastError();
}
#endif
