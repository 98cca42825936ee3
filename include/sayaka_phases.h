/* sayaka_phases.h -- phase timing of the coarsest-level solve (no reference counterpart).
 *
 * Exported by lib/libsayaka_dev_phases.so only: the device library built with
 * SAYAKA_BOTTOM_PHASES, identical to libsayaka_dev.so (same C ABI, include/sayaka_amd.h) except
 * that the persistent coarsest-level BiCGSTAB kernel reads the wall clock at its phase
 * boundaries (workgroup 0) and accumulates the phase durations on the device.  A caller loads
 * this library instead of libsayaka_dev.so, runs solves, and reads the totals here
 * (sayakaamr_amd/bottom_phases.py prints the per-iteration breakdown). */
#ifndef SAYAKA_PHASES_H
#define SAYAKA_PHASES_H

#ifdef __cplusplus
extern "C" {
#endif

/* Totals since the last reset on GPU `device`: out32[k] wall-clock ticks of phase k --
 * 0 operator p -> v, 1-3 reduction A (ordered per-block chains, grid barrier, totals chain over
 * blocks), 4 operator s -> t, 5-7 reduction B, 8 x / r update, 9-11 reduction C, 12 r ghosts,
 * 13-15 the launch's first reduction, 16-19 the staging wait before each reduction's chains
 * (A, B, C, the launch's first) -- out32[30] iterations, out32[31] launches; *clock_khz the wall-clock
 * rate.  reset != 0 zeroes the totals after reading them. */
int sayaka_dev_bottom_phases(int device, unsigned long long* out32, int reset, int* clock_khz);

#ifdef __cplusplus
}
#endif

#endif
