#pragma once

// Shared host helpers: thread-local error text and the status-returning
// wrapper every extern "C" entry point uses so no C++ exception crosses the
// C boundary.

#include <cstdint>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "numi_cloth_hip.h"

namespace numi_hip {

void setLastError(const std::string& message);

struct StatusError : std::runtime_error {
    int status;
    StatusError(int code, const std::string& message)
        : std::runtime_error(message), status(code) {}
};

[[noreturn]] inline void fail(int status, const std::string& message) {
    throw StatusError(status, message);
}

// Particle, distance and bend counts of one specimen of a batch.
void batchSpecimenCounts(const NumiClothHipBatch& batch, uint32_t specimen, uint32_t counts[3]);

template <typename Body>
int guarded(Body&& body) {
    try {
        body();
        setLastError("");
        return NUMI_CLOTH_HIP_OK;
    } catch (const StatusError& error) {
        setLastError(error.what());
        return error.status;
    } catch (const std::bad_alloc&) {
        setLastError("host allocation failed");
        return NUMI_CLOTH_HIP_ERROR_ALLOCATION;
    } catch (const std::exception& error) {
        setLastError(error.what());
        return NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT;
    }
}

}  // namespace numi_hip
