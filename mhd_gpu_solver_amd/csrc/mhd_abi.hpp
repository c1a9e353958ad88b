// Error mapping shared by the C entry points: C++ exceptions become the status codes of
// include/mhd_hip.h, with the message kept per thread for mhd_hip_last_error().
#pragma once

#include "mhd_hip.h"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace mhdhip {

struct InvalidArgument : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

struct NoDevice : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// capacity (HBM budget or instance slots) exhausted: the request may be retried later
struct Full : std::runtime_error {
    using std::runtime_error::runtime_error;
};

std::string& last_error_slot();  // thread-local, defined in mhd_solver.hip

inline void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        throw std::runtime_error(std::string("HIP error in ") + what + ": " +
                                 hipGetErrorString(e));
    }
}

template <class F>
int guarded(F&& f)
{
    try {
        f();
        last_error_slot().clear();
        return MHD_HIP_OK;
    } catch (const NoDevice& e) {
        last_error_slot() = e.what();
        return MHD_HIP_ENODEV;
    } catch (const Full& e) {
        last_error_slot() = e.what();
        return MHD_HIP_EFULL;
    } catch (const std::invalid_argument& e) {
        last_error_slot() = e.what();
        return MHD_HIP_EINVAL;
    } catch (const std::exception& e) {
        last_error_slot() = e.what();
        return MHD_HIP_ERUNTIME;
    } catch (...) {
        last_error_slot() = "unknown error";
        return MHD_HIP_ERUNTIME;
    }
}

}  // namespace mhdhip
