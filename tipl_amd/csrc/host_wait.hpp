// host_wait.hpp -- the host side of results that kernels store straight into
// coherent host memory.  Plain C++ (no HIP), so the logic also builds and is
// tested on a CPU-only host.
#pragma once

#include <chrono>
#include <cstddef>
#include <exception>

namespace tiplhip {

// Waits until every one of the n slots differs from `pending` (the pattern
// the host wrote into them before queueing the kernels that overwrite them).
// Spins for at most `spin`, then calls fallback() once -- a stream
// synchronisation, after which every store of the queued work has landed or
// the failed launch has been reported -- and returns.  Returns the number of
// slots seen written before the fallback (n when it was not needed).
template <class Slot, class Fallback>
size_t wait_for_slots(const volatile Slot* slots, size_t n, Slot pending,
                      std::chrono::microseconds spin, Fallback&& fallback)
{
    size_t ready = 0;
    const auto t0 = std::chrono::steady_clock::now();
    while (ready < n) {
        // acquire: the results read after the wait (the slots themselves and
        // whatever the kernel stored before them) are not read early
        if (__atomic_load_n(&slots[ready], __ATOMIC_ACQUIRE) != pending) {
            ++ready;
            continue;
        }
        if (std::chrono::steady_clock::now() - t0 > spin) {
            fallback();
            return ready;
        }
#if defined(__x86_64__) || defined(__i386__)
        __builtin_ia32_pause();
#endif
    }
    return n;
}

// Calls drain() when the scope is left by an exception.  Placed after a
// kernel that stores into host-visible memory has been queued: if a later
// step throws, the stream is drained before the caller can reuse (or free)
// what that kernel still writes to, so a stale store never lands in the
// next call's slots.
template <class Drain>
struct DrainOnUnwind {
    Drain drain;
    int uncaught = std::uncaught_exceptions();
    explicit DrainOnUnwind(Drain d) : drain(d) {}
    DrainOnUnwind(const DrainOnUnwind&) = delete;
    DrainOnUnwind& operator=(const DrainOnUnwind&) = delete;
    ~DrainOnUnwind()
    {
        if (std::uncaught_exceptions() > uncaught) drain();
    }
};

}  // namespace tiplhip
