// Error convention of the host layer: C-ABI status codes become exceptions
// (the reference exits the process from CUDA_CHECK,
// include/sjb/utils/ErrorChecking.cuh:7-14, and throws std::runtime_error
// from the factory, src/sjb/windowing/WindowingFactory.cpp:86).
#pragma once
#include <stdexcept>
#include <string>

#include "sjb.h"

class SjbError : public std::runtime_error {
public:
    SjbError(int code, const std::string &what)
        : std::runtime_error(what + ": " + sjb_strerror(code)), code(code) {}
    int code;
};

inline int sjbCheck(int rc, const char *what) {
    if (rc != SJB_OK) throw SjbError(rc, what);
    return rc;
}
