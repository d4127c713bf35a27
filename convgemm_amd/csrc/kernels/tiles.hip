// tiles.hip -- the tile-configuration tables (tile_tables.hpp) behind
// cg_internal.h's configuration queries, and the dispatch of a launch to the
// instantiation slice (slice.hip) holding its kernels.  A configuration is a
// tile shape and a raster order (cg_tilecfg); the order is a run-time field
// of the kernel parameters, so the slices hold the tile shapes only.
//
// Slices are referenced weakly: the build decides which (type, operation,
// family) slices exist, and a slice that was not built resolves to null here
// and reports hipErrorInvalidValue (FP8 has forward slices only; the
// pipelined family exists for the forward operations and the NHWC weight
// gradient).
#include "igemm_kernel.hpp"
#include "tile_tables.hpp"

using namespace cg;

#define CG_SLICE_DECL(tn, op, fam) \
    extern "C" __attribute__((weak)) int cg_slice_##tn##_##op##_##fam(int, const cg_kparams *, const cg_group *, int *, void *);
#define CG_SLICE_DECL_T(tn)                                                                                        \
    CG_SLICE_DECL(tn, 0, 0) CG_SLICE_DECL(tn, 1, 0) CG_SLICE_DECL(tn, 2, 0) CG_SLICE_DECL(tn, 3, 0)               \
    CG_SLICE_DECL(tn, 4, 0) CG_SLICE_DECL(tn, 5, 0) CG_SLICE_DECL(tn, 0, 1) CG_SLICE_DECL(tn, 1, 1)               \
    CG_SLICE_DECL(tn, 2, 1)
CG_SLICE_DECL_T(f32)
CG_SLICE_DECL_T(bf16)
CG_SLICE_DECL_T(fp8)
// the single-launch batch kernels (single.hip), one per operand type
extern "C" __attribute__((weak)) int cg_single_f32(const cg_group *, const int *, int *, void *);
extern "C" __attribute__((weak)) int cg_single_bf16(const cg_group *, const int *, int *, void *);
extern "C" __attribute__((weak)) int cg_single_fp8(const cg_group *, const int *, int *, void *);

namespace {

#define CG_SLICE_ROW(tn)                                                                                           \
    {{cg_slice_##tn##_0_0, cg_slice_##tn##_0_1}, {cg_slice_##tn##_1_0, cg_slice_##tn##_1_1},                       \
     {cg_slice_##tn##_2_0, cg_slice_##tn##_2_1}, {cg_slice_##tn##_3_0, nullptr},                                               \
     {cg_slice_##tn##_4_0, nullptr}, {cg_slice_##tn##_5_0, nullptr}}
// [elem][op][family: 0 general, 1 pipelined]
const slice_fn kSlices[3][CG_NUM_OPKINDS][2] = {CG_SLICE_ROW(f32), CG_SLICE_ROW(bf16), CG_SLICE_ROW(fp8)};

int num_shapes(int elem) {
    switch (elem) {
    case CG_ELEM_F32: return num_tiles<float>();
    case CG_ELEM_BF16: return num_tiles<uint16_t>();
    case CG_ELEM_FP8: return num_tiles<uint8_t>();
    }
    return 0;
}

// Configuration cfg = tile shape base + raster order r * (number of shapes).
const TileDef *def_of(int elem, int cfg, int *base = nullptr, int *raster = nullptr) {
    const int ns = num_shapes(elem);
    if (ns == 0 || cfg < 0 || cfg >= ns * CG_NUM_RASTERS) return nullptr;
    const int b = cfg % ns;
    if (base) *base = b;
    if (raster) *raster = cfg / ns;
    switch (elem) {
    case CG_ELEM_F32: return &Tiles<float>::defs[b];
    case CG_ELEM_BF16: return &Tiles<uint16_t>::defs[b];
    case CG_ELEM_FP8: return &Tiles<uint8_t>::defs[b];
    }
    return nullptr;
}

template <typename T> int supports_t(const TileDef &d, int op, int vec) {
    if (d.fam == FAM_LP || d.fam == FAM_HT) return op == CG_FWD_NHWC && vec;
    if (d.fam == FAM_LPN) return op == CG_FWD_NCHW && vec;
    if (d.fam == FAM_LPW) return op == CG_WGRAD_NHWC && vec && sizeof(T) != 1;
    if (!op_enabled<T>(op)) return 0;
    if (!vec && vec_relevant<T>(op) && !d.nv) return 0;
    // tiles of 256x128 and up spill registers in the backward-data gather
    if ((op == CG_DGRAD_NHWC || op == CG_DGRAD_NCHW) && d.BM * d.BN >= 256 * 128) return 0;
    return 1;
}

int launch_any(int elem, int cfg, const cg_kparams *kp, const cg_group *g, int *occ, void *stream) {
    int base, raster;
    const TileDef *d = def_of(elem, cfg, &base, &raster);
    if (!d || kp->op < 0 || kp->op >= CG_NUM_OPKINDS) return (int)hipErrorInvalidValue;
    if (!occ && !g && kp->raster != raster) return (int)hipErrorInvalidValue;  // tiles filled for another order
    const slice_fn f = kSlices[elem][kp->op][d->fam == FAM_GEN ? 0 : 1];
    if (!f) return (int)hipErrorInvalidValue;
    return f(base, kp, g, occ, stream);
}

template <typename T> int single_index(int base) {
    for (int v = 0; v < num_single<T>(); v++)
        if (SingleSet<T>::ids[v] == base) return v;
    return -1;
}

typedef int (*single_fn)(const cg_group *, const int *, int *, void *);
single_fn single_of(int elem) {
    switch (elem) {
    case CG_ELEM_F32: return cg_single_f32;
    case CG_ELEM_BF16: return cg_single_bf16;
    case CG_ELEM_FP8: return cg_single_fp8;
    }
    return nullptr;
}

}  // namespace

extern "C" int cg_single_variant(int elem, int cfg) {
    int base;
    if (!def_of(elem, cfg, &base)) return -1;
    switch (elem) {
    case CG_ELEM_F32: return single_index<float>(base);
    case CG_ELEM_BF16: return single_index<uint16_t>(base);
    case CG_ELEM_FP8: return single_index<uint8_t>(base);
    }
    return -1;
}

extern "C" int cg_launch_single(int elem, const cg_group *g, const int *var, void *stream) {
    const single_fn f = single_of(elem);
    return f ? f(g, var, nullptr, stream) : (int)hipErrorInvalidValue;
}

extern "C" int cg_single_occupancy(int elem) {
    const single_fn f = single_of(elem);
    int occ = 0;
    if (!f || f(nullptr, nullptr, &occ, nullptr) != 0) return 0;
    return occ;
}

extern "C" int cg_num_tilecfgs(int elem) { return num_shapes(elem) * CG_NUM_RASTERS; }

extern "C" int cg_num_tile_shapes(int elem) { return num_shapes(elem); }

extern "C" int cg_get_tilecfg(int elem, int cfg, cg_tilecfg *out) {
    int raster;
    const TileDef *d = def_of(elem, cfg, nullptr, &raster);
    if (!d) return -1;
    *out = cg_tilecfg{d->BM, d->BN, d->BK, d->WM, d->WN, d->NS, raster, d->HS};
    return 0;
}

extern "C" int cg_tilecfg_supports(int elem, int cfg, int op, int vec) {
    const TileDef *d = def_of(elem, cfg);
    if (!d || op < 0 || op >= CG_NUM_OPKINDS) return 0;
    switch (elem) {
    case CG_ELEM_F32: return supports_t<float>(*d, op, vec);
    case CG_ELEM_BF16: return supports_t<uint16_t>(*d, op, vec);
    case CG_ELEM_FP8: return supports_t<uint8_t>(*d, op, vec);
    }
    return 0;
}

extern "C" int cg_launch_igemm(int elem, int cfg, const cg_kparams *kp, void *stream) {
    return launch_any(elem, cfg, kp, nullptr, nullptr, stream);
}

extern "C" int cg_launch_igemm_grouped(int elem, int cfg, const cg_kparams *kp0, const cg_group *g, void *stream) {
    return launch_any(elem, cfg, kp0, g, nullptr, stream);
}

extern "C" int cg_igemm_occupancy(int elem, int cfg, const cg_kparams *kp0) {
    int occ = 0;
    if (launch_any(elem, cfg, kp0, nullptr, &occ, nullptr) != 0) return 0;
    return occ;
}
