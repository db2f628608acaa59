/* Measurement probe shipped beside libsortgpu (libsortgpu_probe.so). It has no counterpart in
 * SORT: it is not part of the drop-in boundary, and bench.py uses it to measure the ceiling the
 * traversal kernels run against (the rate at which a compute unit returns 16-byte chunks to lanes
 * that each read their own record). */
#ifndef SORTGPU_PROBE_H
#define SORTGPU_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Launches `blocks` workgroups of 256 lanes on `stream` (a hipStream_t, NULL = default stream).
 * Every lane reads, `iters` times, `chunks` (1, 3, 7, 8, 14 or 16) consecutive 16-byte chunks of a
 * pseudo-random 256-byte record of `buf` (a device buffer of n_rec * 256 bytes), all chunks of a
 * record issued back to back, and writes one float to out[lane] (blocks * 256 floats). Returns 0
 * or a hipError_t code (other chunk counts are refused); asynchronous. */
int sortgpu_probe_chunks(const void* buf, uint32_t n_rec, int chunks, int iters, int blocks, float* out,
                         void* stream);

#ifdef __cplusplus
}
#endif

#endif
