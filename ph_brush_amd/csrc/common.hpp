// common.hpp -- host-side error machinery shared by every translation unit.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "../../include/phb.h"

namespace phb {

struct Error : std::exception {
    int code;
    std::string msg;
    Error(int c, std::string m) : code(c), msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

void set_last_error(const std::string& m);

// Runs f, converting exceptions into status codes (no exception crosses the ABI).
template <class F>
int guard(F&& f) {
    try {
        return f();
    } catch (const Error& e) {
        set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return PHB_ERR_ALLOC;
    } catch (const std::exception& e) {
        set_last_error(std::string("internal error: ") + e.what());
        return PHB_ERR_INTERNAL;
    }
}

inline void require(bool cond, const char* msg, int code = PHB_ERR_ARG) {
    if (!cond) throw Error(code, msg);
}

}  // namespace phb
