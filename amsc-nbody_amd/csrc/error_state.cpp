// error_state.cpp -- the per-thread last error of the C ABI (message and
// device code), set by every entry point's guard and read through
// nbh_last_error_message / nbh_last_device_code.  Host code only, so the
// host-side sources (particle files, frames, generators) also build and run
// without the HIP runtime, as in the sanitizer builds of their tests.
#include "nbody_amd.h"

#include <string>

namespace nbh {

namespace {
thread_local std::string t_message;
thread_local int t_device_code = 0;
} // namespace

void set_last_error(int code, const std::string& message)
{
    t_device_code = code;
    t_message = message;
}

} // namespace nbh

extern "C" {

const char* nbh_last_error_message(void) { return nbh::t_message.c_str(); }
int nbh_last_device_code(void) { return nbh::t_device_code; }
const char* nbh_version(void) { return "nbody_amd 0.1 gfx950"; }

} // extern "C"
