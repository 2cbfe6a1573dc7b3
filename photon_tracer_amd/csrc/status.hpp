// Error plumbing for the C ABI: exceptions stop at the boundary and become a
// status code plus a thread-local message (pt_last_error).
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "../../include/photon_tracer.h"

namespace pt {

pt_status fail(pt_status s, const std::string& msg);

struct status_error : std::runtime_error {
    pt_status code;
    status_error(pt_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

}  // namespace pt

// Runs `body`; maps exceptions to status codes.
#define PT_GUARD(...)                                                   \
    try {                                                               \
        __VA_ARGS__;                                                         \
        return PT_OK;                                                   \
    } catch (const pt::status_error& e) {                               \
        return pt::fail(e.code, e.what());                              \
    } catch (const std::bad_alloc&) {                                   \
        return pt::fail(PT_ERR_NOMEM, "host allocation failed");        \
    } catch (const std::exception& e) {                                 \
        return pt::fail(PT_ERR_INVALID, e.what());                      \
    }
