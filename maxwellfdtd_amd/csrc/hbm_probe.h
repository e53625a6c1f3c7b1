/* HBM bandwidth probes (libmfd_hbm_probe.so): measurement tools for the benchmark and the
 * byte-counter calibration, not part of the solver's interface.
 *
 * mfd_probe_copy: a float4 device copy of `bytes` bytes with plain loads and stores (the
 * calibration run for the FETCH_SIZE / WRITE_SIZE counters).  mfd_probe_copy_fast: the
 * fastest copy form measured on MI355X (4 float4 per lane in flight, nontemporal loads
 * and stores, 16 workgroups per CU: the practical HBM ceiling).  mfd_probe_multi: a
 * grid-stride kernel with nin reads and nout writes (nin == nout, 1..6) of
 * `bytes_per_array` each, blocks_per_cu (>= 1) workgroups per CU.  Each runs one warm-up
 * launch and `iters` timed ones and returns the average milliseconds per launch; status
 * 0 ok, -1 bad argument, -2 out of memory, -3 HIP error (mfd_probe_last_error). */
#ifndef MFD_HBM_PROBE_H
#define MFD_HBM_PROBE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
const char *mfd_probe_last_error(void);
int32_t mfd_probe_copy(int64_t bytes, int32_t iters, double *ms_per_iter);
int32_t mfd_probe_copy_fast(int64_t bytes, int32_t iters, double *ms_per_iter);
int32_t mfd_probe_multi(int64_t bytes_per_array, int32_t nin, int32_t nout, int32_t blocks_per_cu,
                        int32_t iters, double *ms_per_iter);
#ifdef __cplusplus
}
#endif
#endif
