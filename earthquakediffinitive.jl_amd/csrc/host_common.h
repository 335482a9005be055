// Host-side helpers shared by the libeqd translation units: error reporting through
// eqd_status + a thread-local message, owning device allocations, device check.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/eqd.h"

namespace eqd {
namespace host {

inline thread_local std::string g_err;

struct Error : std::runtime_error {
    eqd_status code;
    Error(eqd_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            throw ::eqd::host::Error(e_ == hipErrorOutOfMemory ? EQD_ERR_NOMEM : EQD_ERR_DEVICE,     \
                                     std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

template <typename F>
eqd_status guard(F&& f) {
    try {
        f();
        return EQD_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return EQD_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return EQD_ERR_ARG;
    }
}

// Device memory: an owning allocation, or (borrow) a view of memory owned elsewhere, which
// release() leaves alone.
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    bool owned = false;
    DevMem() = default;
    explicit DevMem(size_t n) { alloc(n); }
    void alloc(size_t n) {
        release();
        if (n) HIPCHK(hipMalloc(&p, n));
        bytes = n;
        owned = n != 0;
    }
    void borrow(void* q, size_t n) {
        release();
        p = q;
        bytes = n;
        owned = false;
    }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        owned = false;
    }
    ~DevMem() { release(); }
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

inline void check_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw Error(EQD_ERR_DEVICE, "no HIP device is visible");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        throw Error(EQD_ERR_DEVICE, std::string("libeqd is built for gfx950, device is ") + prop.gcnArchName);
}

// Omega_f extents of a solver's operator (Omega_f is square: n2f == n3f).
void solver_fault_dims(const eqd_solver* s, int& nf, int& n2f, int& n3f);

}  // namespace host
}  // namespace eqd
