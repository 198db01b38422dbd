// Shared between the host layer (phx_host.cpp) and the device runtime (phx_device.hip).
#pragma once
#include <stdexcept>
#include <string>

#include "../../include/phx.h"

namespace phx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string& msg);

// the factors of the flux sum (output.hpp:43-53): pixel solid angle dOm and the Jy scale
void flux_scale(const phx_model* m, double* dOm, double* jy);

// Runs f, turning exceptions into an error code + thread-local message.
template <class F>
int guarded(F&& f) {
  try {
    f();
    return PHX_OK;
  } catch (const Error& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_error(e.what());
    return PHX_E_ARG;
  }
}

}  // namespace phx
