// tpsc_queue.hpp -- a whole solve queued without host waits (tpsc_capi.cpp), for the batch driver.
//
// tpsc_queue_solve re-targets a context to a point of its grid and precision and queues mu0,
// chi0, the Usp / Uch / mu root searches (a fixed number of Halley passes each, read from
// device memory by the stages that follow), the vertex, the self-energy and the Lambda_jj
// bubble + MT on the context's main stream, then the copies of the results into `slot` (pinned
// host memory of tpsc_queue_slot_doubles doubles).  Nothing waits: points of several contexts
// run on the GPU back to back while the host queues the next ones.  After the stream has
// drained, tpsc_queue_result reads the slot: TPSC_OK with the point's results (bit for bit those
// of tpsc_solve); kQueuePromote for a mixed-precision point enhanced enough for the FP64 chi0
// stage (1/(1 - Usp chi0max/2) > 8), to be queued again with chi0_fp64, as tpsc_solve redoes it;
// TPSC_E_NOROOT when the point must be solved again with tpsc_solve (a root search unsettled
// after its passes, or a mu0 without a root).
#pragma once
#include <cstddef>

#include "../../include/tpsc_capi.h"

// Halley passes queued per search: Usp settles in 3-5 passes, up to 15 next to the
// antiferromagnetic instability; Uch in 4-6; mu in 3-4.  A pass queued after the search has
// settled exits at once (~3 us of launches).
constexpr int kQueuedPassesUsp = 16;
constexpr int kQueuedPassesUch = 8;
constexpr int kQueuedPassesMu = 6;
constexpr int kQueuePromote = 1;

size_t tpsc_queue_slot_doubles(const tpsc_ctx* ctx);
int tpsc_queue_solve(tpsc_ctx* ctx, const tpsc_params* p, double* slot, bool chi0_fp64);
int tpsc_queue_sync(tpsc_ctx* ctx);
int tpsc_queue_result(const tpsc_params* p, const double* slot, bool chi0_fp64, tpsc_batch_result* r,
                      double* chijj, int ldc);
