// errors.hpp -- status codes, the exception that carries them and the
// thread's last error message.  Plain C++ (no HIP): shared by the device
// layer and the host-only scheduling code.
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "../../include/tiplhip.h"

namespace tiplhip {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);
void clear_last_error();

#define TIPLHIP_REQUIRE(cond, msg)                                                               \
    do {                                                                                         \
        if (!(cond)) throw ::tiplhip::Error(TIPLHIP_ERR_INVALID, (msg));                         \
    } while (0)

// Runs f, converting exceptions into status codes and the thread's last error.
template <class F>
int guarded(F&& f)
{
    try {
        clear_last_error();
        f();
        return TIPLHIP_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return TIPLHIP_ERR_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return TIPLHIP_ERR_INVALID;
    } catch (...) {
        set_last_error("unknown error");
        return TIPLHIP_ERR_INVALID;
    }
}

}  // namespace tiplhip
