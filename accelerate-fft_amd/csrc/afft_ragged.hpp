// afft_ragged.hpp -- descriptors of the ragged-batch persistent kernel (afft_ragged.hip) and its
// plan (afft_ragged_plan.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "afft_internal.hpp"

namespace afft {

constexpr int kRaggedBlock = 256;

enum RaggedKind { kDirect = 0, kBlue = 1 };

// one distinct length of the batch run by the persistent kernel
struct RaggedClass {
    int kind;
    int L;          // true length
    int E, C;       // elements per thread, transforms per item
    int M;          // Bluestein padded length
    PassTable pt;   // kDirect: length L; kBlue: length M
    const void *tw;
    const void *chirp;
    const void *filt;
};

struct RaggedItem {
    int cls;
    int first;  // first index into `moffs` (the transforms of the class, in class order)
    int count;  // transforms in the item
};

struct RaggedArgs {
    const void *in;
    void *out;
    const RaggedClass *classes;
    const RaggedItem *items;
    const int64_t *moffs;   // element offset of every transform, in class-member order
    int nitems;
    int conj;               // Reverse / Inverse
    int inverse;            // divide every transform by its own length
};

hipError_t launch_ragged(int dtype, const RaggedArgs &a, int64_t grid, size_t lds, hipStream_t s);
int ragged_occupancy(int dtype, size_t lds);

} // namespace afft
