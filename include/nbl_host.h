/* nbl_host.h — C entry of the host layer (lib/libnabla_host.so). */
#ifndef NBL_HOST_H
#define NBL_HOST_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One `nabla_solve run <case_json> [--restart <restart>]`
 * (core/apps/nabla_solve/main.cpp:131-175, core/src/flow/case.cpp:136-401)
 * on the calling thread: the same case keys, diagnostics.jsonl, snapshots and
 * checkpoints.  Several calls may run at once on different threads; each run
 * has its own device state and stream.  Returns 0 (msg: the CLI's summary),
 * 1 (msg: "error: ..."), or -15 when *cancel became non-zero between two
 * steps (the subprocess runner's SIGTERM; no final files are written).
 * restart and cancel may be NULL; msg is truncated to msg_len bytes. */
int nbl_case_run(const char* case_json, const char* restart, const volatile int* cancel, char* msg,
                 size_t msg_len);

#ifdef __cplusplus
}
#endif

#endif
