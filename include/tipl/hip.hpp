// hip.hpp -- C++ ownership wrappers over the C ABI (include/tiplhip.h).
//
// Replaces the reference's device_vector / device_image RAII (cu.hpp:41-267)
// and its error convention: a failing call throws std::runtime_error with the
// library's message (cu.hpp:288-322).
#ifndef TIPL_HIP_HIP_HPP
#define TIPL_HIP_HIP_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../tiplhip.h"
#include "image.hpp"

namespace tipl {
namespace hip {

inline void check(int rc)
{
    if (rc != TIPLHIP_OK) throw std::runtime_error(std::string("tiplhip: ") + tiplhip_last_error());
}

inline int device_count()
{
    int n = 0;
    return tiplhip_device_count(&n) == TIPLHIP_OK ? n : 0;
}

class context {
public:
    explicit context(int device = 0) { check(tiplhip_ctx_create(device, &c_)); }
    // wraps a context owned by the caller (not destroyed here)
    explicit context(tiplhip_ctx* borrowed) : c_(borrowed), owned_(false) {}
    ~context()
    {
        if (c_ && owned_) tiplhip_ctx_destroy(c_);
    }
    context(const context&) = delete;
    context& operator=(const context&) = delete;
    tiplhip_ctx* get() const { return c_; }

private:
    tiplhip_ctx* c_ = nullptr;
    bool owned_ = true;
};

// One context per calling thread: the reference re-enters its cost functors
// from several threads (reg/linear.hpp:431-436, mt.hpp:115-222); contexts are
// serialised internally, so a thread-local one keeps callers concurrent.
inline context& thread_context()
{
    thread_local std::unique_ptr<context> c(new context(0));
    return *c;
}

template <class T>
constexpr int dtype_of()
{
    if constexpr (std::is_same_v<T, uint8_t>) return TIPLHIP_U8;
    else if constexpr (std::is_same_v<T, int16_t>) return TIPLHIP_I16;
    else return TIPLHIP_F32;
}

// A device-resident uint8 volume.
class volume {
public:
    volume(context& ctx, const uint8_t* data, const shape<3>& s) : ctx_(&ctx), sp_(s)
    {
        const int d[3] = {int(s[0]), int(s[1]), int(s[2])};
        check(tiplhip_volume_create(ctx.get(), data, TIPLHIP_U8, d, 255.99f, &v_));
    }
    template <class T>
    volume(context& ctx, const T* data, const shape<3>& s, float upper_limit) : ctx_(&ctx), sp_(s)
    {
        const int d[3] = {int(s[0]), int(s[1]), int(s[2])};
        check(tiplhip_volume_create(ctx.get(), data, dtype_of<T>(), d, upper_limit, &v_));
    }
    volume(context& ctx, tiplhip_volume* v) : ctx_(&ctx), v_(v)
    {
        int d[3];
        check(tiplhip_volume_dims(v, d));
        sp_ = tipl::shape<3>(unsigned(d[0]), unsigned(d[1]), unsigned(d[2]));
    }
    ~volume()
    {
        if (v_) tiplhip_volume_destroy(v_);
    }
    volume(const volume&) = delete;
    volume& operator=(const volume&) = delete;
    volume(volume&& r) noexcept : ctx_(r.ctx_), v_(r.v_), sp_(r.sp_) { r.v_ = nullptr; }
    tiplhip_volume* get() const { return v_; }
    const tipl::shape<3>& shape() const { return sp_; }
    volume downsample() const
    {
        tiplhip_volume* o = nullptr;
        check(tiplhip_volume_downsample(ctx_->get(), v_, &o));
        return volume(*ctx_, o);
    }
    image<3, uint8_t> read() const
    {
        image<3, uint8_t> out(sp_);
        check(tiplhip_volume_read(ctx_->get(), v_, out.data()));
        return out;
    }

private:
    context* ctx_;
    tiplhip_volume* v_ = nullptr;
    tipl::shape<3> sp_;
};

}  // namespace hip
}  // namespace tipl

#endif
