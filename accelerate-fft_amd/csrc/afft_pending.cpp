// afft_pending.cpp -- entry points declared in afft.h whose engines are not built yet.  Each
// returns AFFT_ERROR_INTERNAL so a caller gets a clean status instead of a missing symbol.
#include "../../include/afft.h"

extern "C" int afft_autotune(int, int, const int64_t *, int, int) { return AFFT_ERROR_INTERNAL; }
extern "C" int afft_tune_save(const char *) { return AFFT_ERROR_INTERNAL; }
extern "C" int afft_tune_load(const char *) { return AFFT_ERROR_INTERNAL; }
