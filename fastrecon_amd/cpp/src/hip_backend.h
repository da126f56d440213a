// Shared state of the C++ API's MI355X backend: one fr_ctx (device 0, one
// HIP stream) for the process, and the reference's error convention: a
// failure prints a message to stderr and exits with status 1
// (e.g. ReconData.cpp:36-40), never a silent CPU fallback.
#ifndef FASTRECON_HIP_BACKEND_H
#define FASTRECON_HIP_BACKEND_H

#include <cstdlib>
#include <iostream>
#include <memory>

#include "fastrecon_hip.h"

namespace fastrecon {

inline fr_ctx *ctx()
{
    static std::shared_ptr<fr_ctx> c = [] {
        fr_ctx *h = nullptr;
        const char *dev = std::getenv("FASTRECON_DEVICE");
        fr_status st = fr_ctx_create(dev ? std::atoi(dev) : 0, &h);
        if (st != FR_OK) {
            std::cerr << "Error: no usable MI355X device (fr_ctx_create status " << st << ")" << std::endl;
            std::exit(1);
        }
        return std::shared_ptr<fr_ctx>(h, fr_ctx_destroy);
    }();
    return c.get();
}

inline void check(fr_status st, const char *what)
{
    if (st != FR_OK) {
        std::cerr << "Error: " << what << ": " << fr_last_error(ctx()) << std::endl;
        std::exit(1);
    }
}

}  // namespace fastrecon

#endif
