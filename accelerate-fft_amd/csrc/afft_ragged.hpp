// afft_ragged.hpp -- descriptors of the ragged-batch persistent kernel (afft_ragged.hip) and its
// plan (afft_ragged_plan.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "afft_internal.hpp"

namespace afft {

constexpr int kRaggedBlock = 256;

enum RaggedKind { kDirect = 0, kBlue = 1, kTwoPhase = 2, kBlueTwo = 3 };

// one column phase of an in-workgroup four-step transform
struct RaggedSub {
    PassTable pt;
    const void *tw;
    int E;
    int C; // columns per LDS tile
};

// one distinct length of the batch
struct RaggedClass {
    int kind;
    int L;          // true length
    int E, C;       // kDirect / kBlue: elements per thread, transforms per item
    int M;          // Bluestein padded length
    PassTable pt;   // kDirect: length L; kBlue: length M
    const void *tw;
    const void *chirp;
    const void *filt;
    int64_t Lx;     // kTwoPhase: L; kBlueTwo: M;  Lx = s1.pt.n * s2.pt.n
    RaggedSub s1, s2;
    const void *hi; // W_Lx^e = hi[e >> shift] * lo[e & (2^shift - 1)]
    const void *lo;
    int shift;
};

struct RaggedItem {
    int cls;
    int first;  // first index into `moffs` (the transforms of the class, in class order)
    int count;  // short transforms: transforms in the item; long ones: column tiles in the item
    int stage;  // long transforms: the stage (= phase) of the item
    int tile;   // long transforms: the first column tile of the item
    int big;    // four-step Bluestein transforms: scratch slot
};

// stages of a long transform: A = column transforms of length L1 (twiddled) into scratch,
// B = column transforms of length L2 out of it.  Four-step Bluestein: A1 (chirp on load),
// B1 (filter on store), A2, B2 (chirp on store).
constexpr int kStagesTwo = 2;
constexpr int kStagesBlue = 4;

struct RaggedArgs {
    const void *in;
    void *out;
    const RaggedClass *classes;
    const RaggedItem *items;
    const int64_t *moffs;   // element offset of every transform, in class-member order
    int nitems;
    const int64_t *big_off; // scratch offset (elements) of every four-step Bluestein transform
    void *scratch;
    int conj;               // Reverse / Inverse
    int inverse;            // divide every transform by its own length
};

hipError_t launch_ragged(int dtype, const RaggedArgs &a, int64_t grid, size_t lds, hipStream_t s);
int ragged_occupancy(int dtype, size_t lds);

} // namespace afft
