// status.hpp — C-ABI status codes -> the exception types the reference throws.
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "ozablas_hip.h"

namespace ozablas {
namespace detail {

inline void check(int status) {
    if (status == OZA_SUCCESS) return;
    const std::string msg = oza_last_error();
    switch (status) {
        case OZA_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case OZA_ALLOC_FAILED: throw std::bad_alloc();
        default: throw std::runtime_error(msg.empty() ? std::string(oza_status_string(status)) : msg);
    }
}

}  // namespace detail
}  // namespace ozablas
