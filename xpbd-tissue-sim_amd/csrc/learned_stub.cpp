// Placeholder for the batch autotuner entry point.
#include "capi_internal.h"

extern "C" int xpbd_autotune(xpbd_batch*, const char*) { return XPBD_E_UNSUPPORTED; }
