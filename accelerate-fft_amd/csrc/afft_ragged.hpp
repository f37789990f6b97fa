// afft_ragged.hpp -- descriptors of the ragged-batch persistent kernel (afft_ragged.hip) and its
// plan (afft_ragged_plan.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "afft_internal.hpp"

namespace afft {

constexpr int kRaggedBlock = 256;

enum RaggedKind { kDirect = 0, kBlue = 1 };

// Compiled layouts of the persistent kernel (Complex Float): transform length N on T = N / E
// threads (a power of two, so C = kRaggedBlock / T whole transforms fill a workgroup), Stockham
// radices in pass order.  Every pass but the last has stride Ns <= T, so one twiddle base per
// pass and thread serves all its butterflies; the odd radix goes last.  The lengths are the
// 5-smooth sizes {1, 3, 5} * 2^k up to 4096 and so include the padded lengths 640 and 2048 of
// the chirp-z classes 257 and 1009.
#define AFFT_RAGGED_CT_LAYOUTS(X)                                                                \
    X(16, 16, 16)                                                                                \
    X(32, 16, 2, 16)                                                                             \
    X(48, 12, 4, 4, 3)                                                                           \
    X(64, 16, 4, 16)                                                                             \
    X(80, 20, 4, 4, 5)                                                                           \
    X(96, 24, 4, 8, 3)                                                                           \
    X(128, 16, 8, 16)                                                                            \
    X(160, 20, 2, 4, 4, 5)                                                                       \
    X(192, 24, 8, 8, 3)                                                                          \
    X(256, 16, 16, 16)                                                                           \
    X(320, 20, 4, 4, 4, 5)                                                                       \
    X(384, 24, 4, 4, 8, 3)                                                                       \
    X(512, 16, 2, 16, 16)                                                                        \
    X(640, 20, 4, 4, 2, 4, 5)                                                                    \
    X(768, 24, 4, 8, 8, 3)                                                                       \
    X(1024, 16, 16, 16, 4)                                                                       \
    X(1280, 20, 4, 4, 4, 4, 5)                                                                   \
    X(1536, 24, 8, 8, 8, 3)                                                                      \
    X(2048, 16, 16, 16, 8)                                                                       \
    X(2560, 20, 2, 4, 4, 4, 4, 5)                                                                \
    X(3072, 24, 8, 4, 4, 8, 3)                                                                   \
    X(4096, 16, 16, 16, 16)

// elements per thread of the compiled layout of length n, 0 if none
constexpr int ragged_ct_elems(int n)
{
#define AFFT_RAGGED_CT_E(N, E, ...) if (n == N) return E;
    AFFT_RAGGED_CT_LAYOUTS(AFFT_RAGGED_CT_E)
#undef AFFT_RAGGED_CT_E
    return 0;
}

// one distinct length of the batch run by the persistent kernel
struct RaggedClass {
    int kind;
    int L;          // true length
    int E, C;       // elements per thread, transforms per item
    int M;          // Bluestein padded length
    int ctn;        // transform length of a compiled layout (L, or M for kBlue), 0: pass table
    PassTable pt;   // kDirect: length L; kBlue: length M
    const void *tw; // pass-table twiddles, or with ctn the table W_ctn^m, m < ctn
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
