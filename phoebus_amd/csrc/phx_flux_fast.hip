// phx_flux_fast.hip -- the flux sweeps of phx_flux_sweeps.hpp with the fast flux
// arithmetic (phx_fluid_params::flux_arith = PHX_FLUX_ARITH_FAST): built with contraction
// on (-ffp-contract=fast), and the reconstruction, prim_to_flux and Riemann divisions and
// square roots through Newton-refined hardware reciprocals (phx_device.hpp, frcp / frsqrt /
// fsqrt).  Fluxes agree with the exact build to about 1e-14 relative; the sweeps issue
// fewer VALU instructions.
#define PHX_FLUX_FAST 1
#include <hip/hip_runtime.h>

#include "phx.h"
#include "phx_device.hpp"
#include "phx_internal.hpp"

#define PHX_FLUX_NS flux_fast
#define PHX_FLUX_DISPATCH
#include "phx_flux_sweeps.hpp"
