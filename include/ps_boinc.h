/* ps_boinc -- the BOINC-facing pieces of the drop-in period_search binary,
 * without the BOINC client library.
 *
 * The reference links the BOINC API (boinc_api.h) and uses, from main()
 * (period_search_BOINC.cpp:171-907): boinc_get_init_data (the GPU number in
 * APP_INIT_DATA::gpu_device_num, :598-614), boinc_resolve_filename (logical
 * file names, :255-281), boinc_fraction_done (:652, :877-905;
 * start_CUDA.cu:1150-1191), boinc_time_to_checkpoint / boinc_is_standalone
 * (:757, start_CUDA.cu:1185) and boinc_finish (:907).  This host-only shim
 * reads what a BOINC client leaves in the slot directory instead:
 *
 *   init_data.xml         <gpu_device_num>, <checkpoint_period>, <gpu_usage>
 *   soft-link files       "<soft_link>path</soft_link>" names the real file
 *   boinc_finish_called   written by ps_boinc_finish when under a client
 *
 * Without init_data.xml the application is standalone: no device override,
 * a checkpoint after every batch (as the reference does when
 * boinc_is_standalone()), fraction done only reported to the caller.
 * Functions return 0 on success and never exit the process. */
#ifndef PS_BOINC_H
#define PS_BOINC_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int under_client;         /* init_data.xml was found and read */
  int gpu_device_num;       /* -1 if not given (then --device decides) */
  double checkpoint_period; /* seconds between checkpoints; 0 = every batch */
  double gpu_usage;         /* GPUs per task, 0 if not given */
} ps_boinc_init_data;

/* Reads init_data.xml from the directory `slot_dir` (NULL: current
 * directory) into *out; a missing file gives the standalone defaults
 * (under_client 0, gpu_device_num -1, checkpoint_period 0). */
int ps_boinc_read_init_data(const char* slot_dir, ps_boinc_init_data* out);

/* The physical file of a logical name: if the file `logical` holds a BOINC
 * soft link ("<soft_link>path</soft_link>") the linked path, else
 * `logical` itself.  Writes at most cap bytes (with the terminating 0);
 * returns 0, or -1 if the result does not fit. */
int ps_boinc_resolve_filename(const char* logical, char* physical, int cap);

/* Whether a checkpoint is due: always when standalone or the period is 0,
 * else when `period` seconds have passed since the last call of
 * ps_boinc_checkpoint_completed (or since process start). */
int ps_boinc_time_to_checkpoint(const ps_boinc_init_data* d);
void ps_boinc_checkpoint_completed(void);

/* Records the fraction of the work done (clamped to [0, 1]); returns the
 * value kept.  ps_boinc_fraction_done_get reads it back. */
double ps_boinc_fraction_done(double f);
double ps_boinc_fraction_done_get(void);

/* End of the task: under a client, writes boinc_finish_called (the exit
 * status on one line) into the slot directory.  Returns status. */
int ps_boinc_finish(const ps_boinc_init_data* d, const char* slot_dir, int status);

#ifdef __cplusplus
}
#endif
#endif
