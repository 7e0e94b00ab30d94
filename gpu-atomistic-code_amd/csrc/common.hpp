// common.hpp - error reporting, HIP checks and small device buffers shared by the
// sources of libsllg.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/llg_capi.h"

namespace sllg {

// Thread-local message of the last failure, returned by llg_last_error().
void set_error(const char* fmt, ...);
const char* last_error();

#define SLLG_FAIL(code, ...)                 \
    do {                                     \
        ::sllg::set_error(__VA_ARGS__);      \
        return (code);                       \
    } while (0)

#define SLLG_HIP(call)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            SLLG_FAIL(LLG_E_DEVICE, "%s:%d %s: %s", __FILE__, __LINE__, #call,              \
                      hipGetErrorString(e_));                                               \
    } while (0)

#define SLLG_TRY(expr)          \
    do {                        \
        int rc_ = (expr);       \
        if (rc_ != 0) return rc_; \
    } while (0)

// Owning device allocation (hipMalloc / hipFree); moves only.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int alloc(size_t nbytes) {
        release();
        if (nbytes == 0) return 0;
        hipError_t e = hipMalloc(&p, nbytes);
        if (e != hipSuccess) {
            p = nullptr;
            SLLG_FAIL(LLG_E_NOMEM, "hipMalloc(%zu) failed: %s", nbytes, hipGetErrorString(e));
        }
        bytes = nbytes;
        return 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

inline unsigned blocks_for(size_t n, unsigned tpb) { return (unsigned)((n + tpb - 1) / tpb); }

}  // namespace sllg
