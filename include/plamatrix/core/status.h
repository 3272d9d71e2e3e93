// Maps the C ABI's pm_status codes onto the exception classes the reference
// API throws (invalid_argument, out_of_range, overflow_error, logic_error,
// runtime_error).
#pragma once

#include <stdexcept>
#include <string>

#include <hip/hip_runtime_api.h>

#include "plamatrix_amd.h"

namespace plamatrix::detail
{

[[noreturn]] inline void throwStatus(int status)
{
    const std::string message = pm_last_error();
    switch (status)
    {
    case PM_INVALID_ARGUMENT: throw std::invalid_argument(message);
    case PM_OUT_OF_RANGE: throw std::out_of_range(message);
    case PM_OVERFLOW: throw std::overflow_error(message);
    case PM_LOGIC_ERROR: throw std::logic_error(message);
    default: throw std::runtime_error(message);
    }
}

inline void check(int status)
{
    if (status != PM_OK) throwStatus(status);
}

inline void checkHip(hipError_t err, const char* what)
{
    if (err != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(err));
}

} // namespace plamatrix::detail
