#pragma once

// Exception-to-status translation for the C ABI: every entry point runs its body through
// guarded(), which catches the C++ exception vecdex would throw and records its message.

#include <new>
#include <stdexcept>
#include <string>

#include "../../../include/vdxgpu.h"
#include "common.hpp"

namespace vdx {

void set_last_error(const std::string& message);

template <typename Fn>
int guarded(Fn&& fn) noexcept {
    try {
        fn();
        return VDX_OK;
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return VDX_INVALID_ARGUMENT;
    } catch (const std::length_error& e) {
        set_last_error(e.what());
        return VDX_LENGTH_ERROR;
    } catch (const device_error& e) {
        set_last_error(e.what());
        return VDX_DEVICE_ERROR;
    } catch (const std::bad_alloc&) {
        set_last_error("out of memory");
        return VDX_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return VDX_RUNTIME_ERROR;
    } catch (...) {
        set_last_error("unknown error");
        return VDX_RUNTIME_ERROR;
    }
}

// The inverse of guarded(): rethrows a status (with the message it recorded) as its exception.
[[noreturn]] inline void throw_status(int status) {
    const std::string msg = vdx_last_error();
    switch (status) {
        case VDX_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case VDX_LENGTH_ERROR: throw std::length_error(msg);
        case VDX_DEVICE_ERROR: throw device_error(msg);
        case VDX_OUT_OF_MEMORY: throw std::bad_alloc();
        default: throw std::runtime_error(msg);
    }
}

inline void require(bool ok, const char* what) {
    if (!ok) {
        throw std::invalid_argument(what);
    }
}

}  // namespace vdx
