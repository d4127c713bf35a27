// tile_tables.hpp -- the tile configurations of the implicit-GEMM kernels, one
// table per operand type.  A configuration's index is its identity in plans,
// tuning databases and cost-model features, so entries are only appended.
//
//   BM x BN  output tile, BK reduction elements per K step (the general
//            kernel; the pipelined kernels move 128 bytes of each row per step)
//   WM x WN  waves along M and N (4 waves, or 8 for some pipelined tiles)
//   NS       LDS stages of a pipelined kernel (0: the general kernel)
//   fam      kernel family: GEN = igemm_tile (every operation), LP = the
//            pipelined forward NHWC kernel, LPN = the pipelined forward NCHW
//            kernel (FP32), HT = the halo-tile forward NHWC kernel
//            (ht_kernel.hpp; kernels larger than 1x1 whose halo fits HS),
//            LPW = the pipelined weight-gradient NHWC kernel (lpw_kernel.hpp;
//            BK is its K step in pixels)
//   nv       the general kernel also has element-load (non-16-byte)
//            instantiations of this tile
//   HS       halo slots (128 bytes each) of an HT tile
#pragma once
#include <stdint.h>

#include "../cg_internal.h"

namespace cg {

enum TileFam { FAM_GEN = 0, FAM_LP = 1, FAM_LPN = 2, FAM_HT = 3, FAM_LPW = 4 };

struct TileDef {
    int BM, BN, BK, WM, WN, NS;
    int fam;
    bool nv;
    int HS = 0;
};

template <typename T> struct Tiles;

template <> struct Tiles<float> {
    static constexpr TileDef defs[] = {
        {128, 128, 32, 2, 2, 0, FAM_GEN, true},   // 0
        {128, 64, 32, 2, 2, 0, FAM_GEN, false},   // 1
        {64, 128, 32, 2, 2, 0, FAM_GEN, false},   // 2
        {64, 64, 32, 2, 2, 0, FAM_GEN, true},     // 3
        {256, 64, 32, 4, 1, 0, FAM_GEN, false},   // 4
        {128, 128, 16, 2, 2, 0, FAM_GEN, false},  // 5
        {64, 64, 16, 2, 2, 0, FAM_GEN, true},     // 6
        {64, 256, 32, 1, 4, 0, FAM_GEN, false},   // 7
        {256, 64, 16, 4, 1, 0, FAM_GEN, false},   // 8
        {64, 256, 16, 1, 4, 0, FAM_GEN, false},   // 9
        {256, 128, 16, 2, 2, 0, FAM_GEN, false},  // 10
        {128, 64, 16, 2, 2, 0, FAM_GEN, false},   // 11
        {128, 128, 32, 2, 2, 3, FAM_LP, false},   // 12
        {128, 128, 32, 2, 2, 2, FAM_LP, false},   // 13
        {256, 64, 32, 4, 1, 3, FAM_LP, false},    // 14
        {128, 64, 32, 2, 2, 4, FAM_LP, false},    // 15
        {64, 128, 32, 2, 2, 4, FAM_LP, false},    // 16
        {256, 128, 32, 2, 2, 3, FAM_LP, false},   // 17
        {128, 256, 32, 2, 2, 3, FAM_LP, false},   // 18
        {64, 64, 32, 2, 2, 4, FAM_LP, false},     // 19
        {256, 64, 32, 4, 1, 2, FAM_LP, false},    // 20
        {128, 64, 32, 2, 2, 2, FAM_LP, false},    // 21
        {128, 128, 32, 2, 2, 3, FAM_LPN, false},  // 22
        {128, 128, 32, 2, 2, 2, FAM_LPN, false},  // 23
        {128, 64, 32, 2, 2, 3, FAM_LPN, false},   // 24
        {64, 128, 32, 2, 2, 3, FAM_LPN, false},   // 25
        {64, 64, 32, 2, 2, 4, FAM_LPN, false},    // 26
        {256, 64, 32, 4, 1, 2, FAM_LPN, false},   // 27
        {256, 128, 32, 4, 2, 2, FAM_LP, false},   // 28 8 waves
        {128, 128, 32, 2, 4, 3, FAM_LP, false},   // 29 8 waves
        {256, 64, 32, 4, 2, 3, FAM_LP, false},    // 30 8 waves
        {64, 64, 64, 2, 2, 0, FAM_GEN, false},    // 31 long K step (weight gradient: fewer barriers per MFMA)
        {256, 64, 32, 4, 1, 4, FAM_HT, false, 352},   // 32 halo tile
        {128, 128, 32, 2, 2, 3, FAM_HT, false, 256},  // 33 halo tile
        {128, 64, 32, 2, 2, 4, FAM_HT, false, 256},   // 34 halo tile
        {64, 128, 32, 2, 2, 4, FAM_HT, false, 128},   // 35 halo tile
        {64, 64, 32, 2, 2, 4, FAM_HT, false, 288},    // 36 halo tile
        {256, 128, 32, 4, 2, 3, FAM_HT, false, 384},  // 37 halo tile
        {128, 128, 32, 2, 2, 3, FAM_LPW, false},  // 38 pipelined weight gradient
        {128, 128, 32, 2, 2, 2, FAM_LPW, false},  // 39
        {128, 64, 32, 2, 2, 3, FAM_LPW, false},   // 40
        {64, 128, 32, 2, 2, 3, FAM_LPW, false},   // 41
        {64, 64, 32, 2, 2, 4, FAM_LPW, false},    // 42
        {256, 128, 32, 4, 2, 2, FAM_LPW, false},  // 43 8 waves
        {64, 64, 32, 2, 2, 2, FAM_LPW, false},    // 44 two stages: more workgroups per CU
    };
};

template <> struct Tiles<uint16_t> {
    static constexpr TileDef defs[] = {
        {128, 128, 32, 2, 2, 0, FAM_GEN, true},   // 0
        {128, 64, 32, 2, 2, 0, FAM_GEN, false},   // 1
        {64, 128, 32, 2, 2, 0, FAM_GEN, false},   // 2
        {64, 64, 32, 2, 2, 0, FAM_GEN, true},     // 3
        {256, 64, 32, 4, 1, 0, FAM_GEN, false},   // 4
        {256, 128, 32, 2, 2, 0, FAM_GEN, false},  // 5
        {128, 128, 64, 2, 2, 0, FAM_GEN, false},  // 6
        {64, 64, 64, 2, 2, 0, FAM_GEN, true},     // 7
        {128, 256, 32, 2, 2, 0, FAM_GEN, false},  // 8
        {128, 128, 64, 2, 2, 4, FAM_LP, false},   // 9
        {256, 64, 64, 4, 1, 3, FAM_LP, false},    // 10
        {128, 64, 64, 2, 2, 4, FAM_LP, false},    // 11
        {256, 128, 64, 2, 2, 3, FAM_LP, false},   // 12
        {64, 64, 64, 2, 2, 4, FAM_LP, false},     // 13
        {128, 256, 64, 2, 2, 3, FAM_LP, false},   // 14
        {128, 64, 64, 2, 2, 3, FAM_LP, false},    // 15
        {64, 128, 64, 2, 2, 3, FAM_LP, false},    // 16
        {128, 128, 64, 2, 2, 2, FAM_LP, false},   // 17
        {256, 64, 64, 4, 1, 2, FAM_LP, false},    // 18
        {128, 64, 64, 2, 2, 2, FAM_LP, false},    // 19
        {64, 128, 64, 2, 2, 2, FAM_LP, false},    // 20
        {256, 128, 64, 2, 2, 2, FAM_LP, false},   // 21
        {64, 64, 64, 2, 2, 2, FAM_LP, false},     // 22
        {256, 128, 64, 4, 2, 2, FAM_LP, false},   // 23 8 waves
        {256, 128, 64, 4, 2, 3, FAM_LP, false},   // 24 8 waves
        {256, 256, 64, 4, 2, 2, FAM_LP, false},   // 25 8 waves
        {128, 256, 64, 2, 4, 2, FAM_LP, false},   // 26 8 waves
        {512, 64, 64, 8, 1, 2, FAM_LP, false},    // 27 8 waves
        {256, 64, 64, 4, 2, 3, FAM_LP, false},    // 28 8 waves
        {128, 128, 64, 2, 4, 4, FAM_LP, false},   // 29 8 waves
        {64, 64, 128, 2, 2, 0, FAM_GEN, false},   // 30 long K step (weight gradient: fewer barriers per MFMA)
        {256, 64, 64, 4, 1, 4, FAM_HT, false, 352},   // 31 halo tile
        {128, 128, 64, 2, 2, 3, FAM_HT, false, 256},  // 32 halo tile
        {128, 64, 64, 2, 2, 4, FAM_HT, false, 256},   // 33 halo tile
        {64, 128, 64, 2, 2, 4, FAM_HT, false, 128},   // 34 halo tile
        {64, 64, 64, 2, 2, 4, FAM_HT, false, 288},    // 35 halo tile
        {256, 128, 64, 4, 2, 3, FAM_HT, false, 384},  // 36 halo tile
        {128, 128, 64, 2, 2, 3, FAM_LPW, false},  // 37 pipelined weight gradient
        {128, 128, 64, 2, 2, 2, FAM_LPW, false},  // 38
        {128, 64, 64, 2, 2, 3, FAM_LPW, false},   // 39
        {64, 128, 64, 2, 2, 3, FAM_LPW, false},   // 40
        {64, 64, 64, 2, 2, 4, FAM_LPW, false},    // 41
        {256, 128, 64, 4, 2, 2, FAM_LPW, false},  // 42 8 waves
        {128, 256, 64, 2, 4, 2, FAM_LPW, false},  // 43 8 waves
        {128, 64, 64, 2, 2, 2, FAM_LPW, false},   // 44
        {64, 64, 64, 2, 2, 2, FAM_LPW, false},    // 45 two stages: more workgroups per CU
        {64, 64, 64, 2, 2, 3, FAM_LPW, false},    // 46
    };
};

template <> struct Tiles<uint8_t> {
    static constexpr TileDef defs[] = {
        {128, 128, 64, 2, 2, 0, FAM_GEN, true},   // 0
        {128, 64, 64, 2, 2, 0, FAM_GEN, false},   // 1
        {64, 128, 64, 2, 2, 0, FAM_GEN, false},   // 2
        {64, 64, 64, 2, 2, 0, FAM_GEN, true},     // 3
        {256, 64, 64, 4, 1, 0, FAM_GEN, false},   // 4
        {256, 128, 64, 2, 2, 0, FAM_GEN, false},  // 5
        {128, 128, 128, 2, 2, 0, FAM_GEN, false}, // 6
        {128, 256, 64, 2, 2, 0, FAM_GEN, false},  // 7
        {128, 128, 128, 2, 2, 4, FAM_LP, false},  // 8
        {256, 64, 128, 4, 1, 3, FAM_LP, false},   // 9
        {128, 64, 128, 2, 2, 4, FAM_LP, false},   // 10
        {256, 128, 128, 2, 2, 3, FAM_LP, false},  // 11
        {64, 64, 128, 2, 2, 4, FAM_LP, false},    // 12
        {128, 256, 128, 2, 2, 3, FAM_LP, false},  // 13
        {128, 64, 128, 2, 2, 3, FAM_LP, false},   // 14
        {64, 128, 128, 2, 2, 3, FAM_LP, false},   // 15
        {128, 128, 128, 2, 2, 2, FAM_LP, false},  // 16
        {256, 64, 128, 4, 1, 2, FAM_LP, false},   // 17
        {128, 64, 128, 2, 2, 2, FAM_LP, false},   // 18
        {64, 128, 128, 2, 2, 2, FAM_LP, false},   // 19
        {256, 128, 128, 2, 2, 2, FAM_LP, false},  // 20
        {64, 64, 128, 2, 2, 2, FAM_LP, false},    // 21
        {256, 128, 128, 4, 2, 2, FAM_LP, false},  // 22 8 waves
        {256, 128, 128, 4, 2, 3, FAM_LP, false},  // 23 8 waves
        {256, 256, 128, 4, 2, 2, FAM_LP, false},  // 24 8 waves
        {128, 256, 128, 2, 4, 2, FAM_LP, false},  // 25 8 waves
        {512, 64, 128, 8, 1, 2, FAM_LP, false},   // 26 8 waves
        {256, 64, 128, 4, 2, 3, FAM_LP, false},   // 27 8 waves
        {128, 128, 128, 2, 4, 4, FAM_LP, false},  // 28 8 waves
        {256, 64, 128, 4, 1, 4, FAM_HT, false, 352},  // 29 halo tile
        {128, 128, 128, 2, 2, 3, FAM_HT, false, 256}, // 30 halo tile
        {128, 64, 128, 2, 2, 4, FAM_HT, false, 256},  // 31 halo tile
        {64, 128, 128, 2, 2, 4, FAM_HT, false, 128},  // 32 halo tile
        {64, 64, 128, 2, 2, 4, FAM_HT, false, 288},   // 33 halo tile
        {256, 128, 128, 4, 2, 3, FAM_HT, false, 384}, // 34 halo tile
    };
};

template <typename T> constexpr int num_tiles() { return (int)(sizeof(Tiles<T>::defs) / sizeof(TileDef)); }

// Tile shapes of the single-launch batch kernel (single.hip): 4-wave
// pipelined tiles whose NS stages fit SINGLE_LDS, so that two workgroups
// share a CU.  Entries are tile-shape ids of Tiles<T>; the kernel switches
// over them per problem.
constexpr int SINGLE_LDS = 80 * 1024;
template <typename T> struct SingleSet;
template <> struct SingleSet<float> {
    static constexpr int ids[] = {19, 21, 13, 20};  // 64x64, 128x64, 128x128, 256x64
};
template <> struct SingleSet<uint16_t> {
    static constexpr int ids[] = {13, 15, 16, 17, 18};  // 64x64, 128x64, 64x128, 128x128, 256x64
};
template <> struct SingleSet<uint8_t> {
    static constexpr int ids[] = {12, 14, 15, 16, 17};  // 64x64, 128x64, 64x128, 128x128, 256x64
};
template <typename T> constexpr int num_single() { return (int)(sizeof(SingleSet<T>::ids) / sizeof(int)); }

// Signature of one instantiation slice: the kernels of one (operand type,
// operation, kernel family), compiled as a translation unit of its own.
//   g == nullptr: one problem (kp); g != nullptr: the grouped kernel over g's
//   problems; occ != nullptr: report resident workgroups per CU instead of
//   launching.  Returns 0 or a hipError_t.
typedef int (*slice_fn)(int cfg, const cg_kparams *kp, const cg_group *g, int *occ, void *stream);

}  // namespace cg
