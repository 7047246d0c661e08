// Error plumbing of the C ABI: every entry point runs its body under guarded(), which
// turns exceptions into an SM_ERR_* code plus a thread-local message, so no C++ exception
// crosses the ABI (the reference throws std::runtime_error from its CUDA macros,
// include/cuda/errors.h:7-30, and ints from Blob/Script).
#pragma once
#include "../../include/softmold_hip.h"

#include <new>
#include <stdexcept>
#include <string>

namespace smapi {

extern thread_local std::string g_err;

struct ApiError : std::runtime_error {
    int code;
    ApiError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

template <class F>
int guarded(F &&f)
{
    try {
        f();
        return SM_OK;
    } catch (const ApiError &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = "host allocation failed";
        return SM_ERR_INVALID;
    } catch (const std::exception &e) {
        g_err = e.what();
        return SM_ERR_INVALID;
    }
}

[[noreturn]] inline void invalid(const std::string &m) { throw ApiError(SM_ERR_INVALID, m); }
[[noreturn]] inline void unsupported(const std::string &m) { throw ApiError(SM_ERR_UNSUPPORTED, m); }

}  // namespace smapi
