// abi_core.hip — error reporting, version and argument checks shared by all entry points.
#include <cstdarg>
#include <cstdio>
#include <initializer_list>

#include "waam_common.h"

namespace waam {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(WAAM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return WAAM_OK;
}

int check_dims(const char *what, int nx, int ny, int nz, int layout)
{
    if (nx <= 0 || ny <= 0 || nz <= 0)
        return set_error(WAAM_ERR_INVALID, "%s: invalid dims %d x %d x %d", what, nx, ny, nz);
    if (layout != kLayoutXYZ && layout != kLayoutZYX)
        return set_error(WAAM_ERR_INVALID, "%s: unknown field layout %d", what, layout);
    if ((int64_t)nx * ny * nz > ((int64_t)1 << 40))
        return set_error(WAAM_ERR_INVALID, "%s: grid too large", what);
    return WAAM_OK;
}

int check_ptrs(const char *what, std::initializer_list<const void *> ptrs)
{
    int i = 0;
    for (const void *p : ptrs) {
        if (p == nullptr) return set_error(WAAM_ERR_INVALID, "%s: pointer argument %d is NULL", what, i);
        ++i;
    }
    return WAAM_OK;
}

}  // namespace waam

extern "C" {

int waam_abi_version(void) { return WAAM_ABI_VERSION; }

const char *waam_last_error(void) { return waam::g_err; }

int64_t waam_scratch_bytes(void) { return (int64_t)waam::kScratchBytes; }

}  // extern "C"
