/* exoblas_amd: library-level utilities of the MI355X (gfx950) BLAS.
 *
 * The drop-in exo_* entry points are declared in exo_<kernel>.h (one header
 * per reference library target) and their stream-ordered device-pointer forms
 * in exoblas_amd_async.h.  None of the functions below has a counterpart in
 * the reference, whose functions report no errors (every exo_* returns void).
 */
#ifndef EXOBLAS_AMD_H
#define EXOBLAS_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes of the last call on this host thread. */
#define EXOBLAS_STATUS_OK 0
#define EXOBLAS_STATUS_INVALID_ENUM 1
#define EXOBLAS_STATUS_INVALID_SIZE 2
#define EXOBLAS_STATUS_INVALID_STRIDE 3
#define EXOBLAS_STATUS_INVALID_LD 4
#define EXOBLAS_STATUS_HIP_ERROR 5
#define EXOBLAS_STATUS_ALLOC_ERROR 6
#define EXOBLAS_STATUS_NOT_SUPPORTED 7
#define EXOBLAS_STATUS_INVALID_POINTER 8
#define EXOBLAS_STATUS_TIMEOUT 9
#define EXOBLAS_STATUS_CANCELLED 10

/* Creates the per-device state (stream, counters) for the current HIP device. */
int exoblas_init(void);
/* Status and message of the most recent failing call on this host thread. */
int exoblas_last_status(void);
const char *exoblas_last_error(void);
void exoblas_clear_error(void);
/* Waits for the library's own stream on the current device. */
int exoblas_synchronize(void);
/* Waits for `stream` (a hipStream_t) and reports whether a bounded wait
 * between workgroups expired in the library's work on it since the last
 * check: EXOBLAS_STATUS_TIMEOUT means the results of those calls are not
 * valid.  The synchronous exo_* calls make this check themselves. */
int exoblas_stream_check(void *stream);
/* Releases what the library holds for `stream` (a hipStream_t): waits for
 * its queued work, then frees the stream's scratch buffers, counters and
 * timeout word.  Call it before destroying a stream the library has been
 * used on (HIP may hand the same handle to a later stream, which then starts
 * fresh); graphs captured on the stream must not be replayed afterwards.
 * Returns 0 (also for a stream the library has not seen) or a status;
 * EXOBLAS_STATUS_NOT_SUPPORTED while the stream is being captured. */
int exoblas_stream_release(void *stream);
/* Bound of every wait between workgroups, in polls (< 0 restores the
 * default of 2^22, about a second).  A test hook: a small bound makes waits
 * give up early, so the timeout path can be exercised while every kernel
 * still ends normally.  Applies to the current device. */
int exoblas_set_spin_limit(long long polls);

#ifdef __cplusplus
}
#endif

#endif /* EXOBLAS_AMD_H */
