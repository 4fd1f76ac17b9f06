// phx_flux_weno.hip -- the WENO5-Z flux sweeps of phx_flux_sweeps.hpp in the exact
// arithmetic, built in a translation unit of their own (phx_flux.hip dispatches to them),
// so that the kernel builds of the reconstructions run in parallel.
#include <hip/hip_runtime.h>

#include "phx.h"
#include "phx_device.hpp"
#include "phx_internal.hpp"

#define PHX_FLUX_NS flux_exact
#include "phx_flux_sweeps.hpp"
