// Core types of the MI355X drop-in for PlaMatrix's sparse solve path
// (core/types.h:9-15 of the reference).
#pragma once

#include <cstdint>

namespace plamatrix
{

using Index = std::int64_t;

enum class Device
{
    CPU,
    GPU,
};

} // namespace plamatrix
