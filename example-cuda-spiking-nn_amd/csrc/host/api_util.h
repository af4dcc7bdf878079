// Error plumbing of the extern "C" API: exceptions become return codes + bn_last_error().
#pragma once

#include <stdexcept>
#include <string>

#include "../bn_errors.h"

namespace bn {

int set_error(int code, const std::string& msg);

// Runs f, mapping the exceptions of the host engine to the C ABI's codes (the reference
// throws std::runtime_error / std::invalid_argument / std::out_of_range, cuda_utils.h:10-20).
template <typename F>
inline int guarded(F&& f) {
    try {
        f();
        return BN_OK;
    } catch (const Error& e) {
        return set_error(e.code, e.what());
    } catch (const std::invalid_argument& e) {
        return set_error(BN_ERR_INVALID_ARGUMENT, e.what());
    } catch (const std::out_of_range& e) {
        return set_error(BN_ERR_OUT_OF_RANGE, e.what());
    } catch (const std::exception& e) {
        return set_error(BN_ERR_RUNTIME, e.what());
    }
}

}  // namespace bn
