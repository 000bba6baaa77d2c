// filt_guard.hpp -- C ABI entry points that use C++ containers or threads run
// their body through guarded(), so that no C++ exception crosses the
// extern "C" boundary: an allocation failure becomes FILT_ERR_NOMEM, any
// other exception FILT_ERR_INVALID, with the text in filt_last_error().
#pragma once

#include <exception>
#include <new>

#include "../../include/filtration_hip.h"

namespace filt {

int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

template <class F>
int guarded(F&& body) noexcept
{
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return set_error(FILT_ERR_NOMEM, "host memory exhausted");
    } catch (const std::exception& e) {
        return set_error(FILT_ERR_INVALID, "%s", e.what());
    } catch (...) {
        return set_error(FILT_ERR_INVALID, "unexpected C++ exception");
    }
}

}  // namespace filt
