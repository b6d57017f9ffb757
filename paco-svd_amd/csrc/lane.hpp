// Internal interface between the batched engine (engine.hip) and the serving loop
// (serve.hip): a lane owns preallocated device scratch, a pinned plan buffer and a
// forked stream, so enqueueing a batch on it allocates nothing.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/paco_svd.h"

namespace paco {

struct Lane;

// Device bytes one job needs in a lane (scratch + its share of the plan and states).
size_t job_footprint(int M, int N);
// True when the job goes through the chip-wide path instead of a batch.
bool job_is_wide(int M, int N);
// A lane able to hold `capacity` bytes of job footprints and `max_jobs` jobs.
Lane *lane_create(size_t capacity, int max_jobs);
void lane_destroy(Lane *l);
// Enqueue the batched pipeline of the jobs on stream s (stream-ordered, no host sync
// beyond the lane's previous plan upload).  Returns 0 or an error code.
int lane_enqueue(Lane *l, const paco_svd_job *jobs, int count, hipStream_t s);
// Info words (LAPACK-style, one per job) of the lane's last enqueued batch, copied to the
// host; call after the lane's stream has passed the batch.  0 or an error code.
int lane_info(Lane *l, int *out, int count);
// Whether the engine takes an M x N job (sizes within the kernels' limits).
bool job_supported(int M, int N);
// Scratch for one wide job through the two entry points.
size_t wide_scratch_bytes(int M, int N);
// grid: tile grid of the chip-wide bidiagonalization (0: default); several wide lanes
// running at once each take a smaller grid; no_wait: without kernels whose workgroups wait
// on each other (the re-run of a solve whose waits timed out beside other work)
int wide_enqueue(const paco_svd_job &job, void *scratch, hipStream_t s, int grid, bool no_wait = false);
// Info word of the wide job last run in scratch (after its stream has passed it).
int wide_info(const paco_svd_job &job, const void *scratch, int *out);

} // namespace paco
