/*
 * phx_driver.h -- one-block RK2 driver over the fluid task interface of libphx.so.
 *
 * Replaces, for one mesh block, the fluid-only part of PhoebusDriver::RungeKuttaStage
 * (phoebus_driver.cpp:204-738) and parthenon's EvolutionDriver loop around it: per stage
 * fluid::CalculateFluidSourceTerms (curved geometry only, phoebus_driver.cpp:323),
 * fluid::CalculateFluxes + fluid::FluxCT (:344-345), FluxDivergence + SumData +
 * AverageIndependentData + UpdateIndependentData (:560-576), FillDerived con2prim over
 * the entire block (:589), physical boundaries (:662) and FillDerived again (:667), each
 * FillDerived followed by fixup::ConservedToPrimitiveFixup and fixup::ApplyFloors (:592-597,
 * :671-676; no-ops unless phx_driver_set_fixup enabled them), and EstimateTimestep after
 * the last stage (:719).  Stage containers follow parthenon's
 * LowStorageIntegrator for rk2: "base" and "1", beta = {1, 1/2}.  The timestep grows by at
 * most 2x per step and the last step is clipped to tlim.
 *
 * The driver owns the device fields of its block (two conserved containers, shared
 * primitives and derived fields).  Primitives are set on the entire block (ghosts too)
 * before phx_driver_initialize, as a problem generator does (shock_tube.cpp:47-124).
 */
#ifndef PHX_DRIVER_H_
#define PHX_DRIVER_H_

#include <stdint.h>

#include "phx.h"
#include "phx_fixup.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phx_driver phx_driver;

/* physical boundary per face, bc[2*d + side] (side 0 = inner) */
enum { PHX_BC_OUTFLOW = 0, PHX_BC_PERIODIC = 1 };
/* fields readable with phx_driver_get */
enum { PHX_DRIVER_PRIM = 0, PHX_DRIVER_CONS = 1, PHX_DRIVER_FAIL = 2, PHX_DRIVER_C2P_MU = 3 };

/* Allocates the block's fields on the current device.  bc kinds as above; cfl is the
 * fluid package's "cfl" parameter (fluid.cpp:899-944); the first timestep is
 * dt_init_fact times the estimate (parthenon/time/dt_init_fact, shocktube.pin:36). */
int phx_driver_create(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                      const int bc[6], double cfl, double dt_init_fact, phx_driver **out);
int phx_driver_destroy(phx_driver *d);
/* the fixup package's switches (default: all off, the fixup.cpp:63-74 defaults) */
int phx_driver_set_fixup(phx_driver *d, const phx_fixup_params *fx);
/* host primitives [nprim][ncell] over the entire block */
int phx_driver_set_prim(phx_driver *d, const double *host_prim);
/* p2c, floors, boundaries, con2prim over the entire block, first timestep
 * (PostInitializationCommunication, phoebus_driver.cpp:147-202) */
int phx_driver_initialize(phx_driver *d);
/* one RK2 step of size dt (clipped to the tlim of the last execute call) */
int phx_driver_step(phx_driver *d);
/* steps until time >= tlim or max_steps steps were taken */
int phx_driver_execute(phx_driver *d, double tlim, int max_steps);
int phx_driver_time(const phx_driver *d, double *time, double *dt, int64_t *ncycle);
/* copies a field of the base container / shared fields to host (synchronizes) */
int phx_driver_get(const phx_driver *d, int what, double *host);

#ifdef __cplusplus
}
#endif
#endif /* PHX_DRIVER_H_ */
