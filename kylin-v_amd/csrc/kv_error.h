// Error type shared by host-only and device code of the engine.
#pragma once

#include <stdexcept>
#include <string>

namespace kv {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

constexpr int ERR_ARG = -1, ERR_DIRS = -2, ERR_HIP = -3, ERR_SOLVER = -4, ERR_FORMAT = -5,
              ERR_SYM = -6, ERR_INTERNAL = -7;

#define KV_REQUIRE(cond, code, msg)                                                    \
    do {                                                                               \
        if (!(cond)) throw ::kv::Error((code), (msg));                                 \
    } while (0)

}  // namespace kv
