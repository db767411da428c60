// HIP error checking for the host layer: failures throw mek::Error (the CLI
// prints the message and exits 1, as the reference's CHECK_CUDA does at
// main.cu:44-52; the C ABI turns it into a status code).
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace mek {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* expr, const char* file, int line) {
  if (e != hipSuccess) {
    throw Error(std::string("HIP error ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) +
                ") at " + file + ":" + std::to_string(line) + ": " + expr);
  }
}

}  // namespace mek

#define MEK_HIP(expr) ::mek::hip_check((expr), #expr, __FILE__, __LINE__)
// kernel launches: report launch errors at the call site
#define MEK_HIP_LAUNCHED() ::mek::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)
