// Error convention of the C ABI: entry points return 0 or a BN_ERR_* code and keep the
// message for bn_last_error(); the C++ layer turns a nonzero code back into
// std::runtime_error("<what>: <detail>") like the reference's cuda_check
// (cpp_src/gpu_core/cuda_utils.h:10-20).
#pragma once

#include <stdexcept>
#include <string>

#include "buryn_amd.h"

namespace bn {

// Stores the message for bn_last_error() and returns code.
int set_error(int code, const std::string& msg);

// Thrown by the host engine; carries the C-ABI code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

}  // namespace bn

// Inside a function returning an int code: return the error of a failing HIP call.
#define BN_HIP_TRY(expr, what)                                                                \
    do {                                                                                      \
        const hipError_t bn_e_ = (expr);                                                      \
        if (bn_e_ != hipSuccess)                                                              \
            return ::bn::set_error(BN_ERR_HIP, std::string(what) + ": " + hipGetErrorString(bn_e_)); \
    } while (0)

// Inside host engine code: throw on a failing HIP call (reference cuda_check).
#define BN_HIP_CHECK(expr, what)                                                              \
    do {                                                                                      \
        const hipError_t bn_e_ = (expr);                                                      \
        if (bn_e_ != hipSuccess)                                                              \
            throw ::bn::Error(BN_ERR_HIP, std::string(what) + ": " + hipGetErrorString(bn_e_)); \
    } while (0)

// Inside host engine code: throw if a C-ABI launcher returned an error.
#define BN_CALL(expr)                                                                         \
    do {                                                                                      \
        const int bn_rc_ = (expr);                                                            \
        if (bn_rc_ != 0) throw ::bn::Error(bn_rc_, bn_last_error());                          \
    } while (0)
