// Exception barrier of the C ABI: host code behind an extern "C" entry point
// uses std containers, whose allocation failures (std::bad_alloc,
// std::length_error) must become status codes instead of unwinding into a C
// or ctypes caller (which aborts the process).
#pragma once

#include <new>
#include <stdexcept>

#include "../../include/mesh_bem_amd.h"

// Runs fn() (returning an MBA_* status) and maps any C++ exception to a
// status: allocation failures to MBA_ERR_ALLOC, a container size beyond its
// limit to MBA_ERR_SIZE, anything else to `other`.
template <class Fn>
inline int mba_guard(Fn&& fn, int other = MBA_ERR_ARG) noexcept
{
    try {
        return fn();
    } catch (const std::bad_alloc&) {
        return MBA_ERR_ALLOC;
    } catch (const std::length_error&) {
        return MBA_ERR_SIZE;
    } catch (...) {
        return other;
    }
}
