/* Calibration of the chip's ceilings for the join path's access patterns.
 *
 * No counterpart in the reference (its joins run on CPU cores).  These entry
 * points serve measurement: they time, on the library's device, the two
 * patterns that bound the path -- a partition pass's scatter of short runs,
 * and a no-partitioning probe's random reads and atomics -- so that a kernel
 * can be rated against the rate its own access pattern allows as well as
 * against the HBM peak.  Call tbj_init first.  Each call runs one launch,
 * waits for it and stores its device time in *ms.
 */
#ifndef TBJ_CALIB_H
#define TBJ_CALIB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reads n 8-byte tuples from `in` and writes them to `out` (device buffers of
 * n tuples each) as a partition pass would: workgroups of bins * run tuples,
 * each writing one run of `run` tuples into each of `bins` regions of its
 * XCD.  n must be a multiple of 8 * bins * run, bins * run <= 2^20. */
int tbj_calibrate_runs(const void* in, void* out, uint64_t n, uint32_t bins, uint32_t run,
                       void* stream, float* ms);

/* nprobe (a multiple of 2048) random accesses to a device table of `bytes`
 * bytes (a power of two, >= 64): kind 0 = 8-byte reads, 1 = 32-byte reads,
 * 2 = u32 atomicAdd.  sink: 8 device bytes (written only if the reads sum
 * to one unlikely value, which keeps the reads alive). */
int tbj_calibrate_random(void* table, uint64_t bytes, uint64_t nprobe, int kind, void* sink,
                         void* stream, float* ms);

#ifdef __cplusplus
}
#endif

#endif
