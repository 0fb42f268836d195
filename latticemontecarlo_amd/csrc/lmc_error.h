// Status codes and the thread-local message of the C ABI, shared by its source files.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "lmc_capi.h"

namespace lmc {

inline thread_local std::string g_last_error;

inline int Fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ModelError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void Check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// Runs f and maps its exceptions to status codes: HIP errors, model errors, invalid arguments,
// everything else as I/O.
template <typename F>
int Guard(F &&f) {
  try {
    return f();
  } catch (const HipError &e) {
    return Fail(LMC_ERR_HIP, e.what());
  } catch (const ModelError &e) {
    return Fail(LMC_ERR_MODEL, e.what());
  } catch (const std::invalid_argument &e) {
    return Fail(LMC_ERR_ARG, e.what());
  } catch (const std::exception &e) {
    return Fail(LMC_ERR_IO, e.what());
  }
}

}  // namespace lmc
