// The slab pass in the line factorisations other than each edge's default (lfts_slab.h),
// for plain chains: the alternatives the kernel-shape autotuner times against the defaults
// of lfts_fused.hip. Kept in a translation unit of its own so that the two compile in
// parallel.
#include "lfts_slab.h"

namespace lfts {

hipError_t dispatch_edge_alt(int n, int shape, int v, const Rep* reps, int nact, int i, const SlabIO* io,
                             const double2* tw, hipStream_t st) {
    // must list the alternatives of the shape table in lfts_fused.hip, in its order, with
    // the same variant masks (slab_shape_variants)
    if (n == 64 && shape == 1) return dispatch_plain<64, 16, 4, 0xA>(v, reps, nact, i, io, tw, st);
    if (n == 64 && shape == 2) return dispatch_plain<64, 4, 16, 0x3>(v, reps, nact, i, io, tw, st);
    if (n == 56 && shape == 1) return dispatch_plain<56, 14, 4, 0xF>(v, reps, nact, i, io, tw, st);
    if (n == 48 && shape == 1) return dispatch_plain<48, 12, 4, 0xF>(v, reps, nact, i, io, tw, st);
    if (n == 40 && shape == 1) return dispatch_plain<40, 10, 4, 0xF>(v, reps, nact, i, io, tw, st);
    if (n == 32 && shape == 1) return dispatch_plain<32, 8, 4, 0xF>(v, reps, nact, i, io, tw, st);
    if (n == 24 && shape == 1) return dispatch_plain<24, 6, 4, 0xF>(v, reps, nact, i, io, tw, st);
    if (n == 16 && shape == 1) return dispatch_plain<16, 2, 8, 0xF>(v, reps, nact, i, io, tw, st);
    return hipErrorInvalidValue;
}

}  // namespace lfts
