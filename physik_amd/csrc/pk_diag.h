/*
 * pk_diag.h - the diagnostics library libphysik_diag.so (built beside the
 * product library by physik_amd/build.py, loaded only by physik_amd/diag.py).
 * It has no counterpart in the reference and is not part of the product ABI.
 * Status: 0 ok, 1 bad argument, 2 HIP failure (message in pk_diag_last_error).
 */
#ifndef PK_DIAG_H
#define PK_DIAG_H

#ifdef __cplusplus
extern "C" {
#endif

#pragma GCC visibility push(default)

const char* pk_diag_last_error(void);

/* Copies `bytes` (a multiple of 32) between two fresh device buffers `reps`
 * times with a plain streaming kernel and returns the average time per copy. */
int pk_diag_copy(long long bytes, int reps, double* avg_ms);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* PK_DIAG_H */
