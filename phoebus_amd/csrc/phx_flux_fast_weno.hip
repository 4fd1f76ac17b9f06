// phx_flux_fast_weno.hip -- the WENO5-Z flux sweeps of phx_flux_sweeps.hpp in the fast
// flux arithmetic (as phx_flux_fast.hip, which dispatches to them), built in a translation
// unit of their own so that the kernel builds of the reconstructions run in parallel.
#define PHX_FLUX_FAST 1
#include <hip/hip_runtime.h>

#include "phx.h"
#include "phx_device.hpp"
#include "phx_internal.hpp"

#define PHX_FLUX_NS flux_fast
#include "phx_flux_sweeps.hpp"
