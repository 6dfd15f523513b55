// Host build of libfdtd_hip (test infrastructure): the product source
// fdtd_amd/csrc/fdtd_hip.hip compiled unchanged as C++ against the host
// stand-in hip/hip_runtime.h of this directory. See build.py.
#include "fdtd_hip.hip"

// Kernel launches run so far (lets a test check that work really ran here).
extern "C" int64_t fdtd_shim_launches(void) { return shim::g_launches; }
