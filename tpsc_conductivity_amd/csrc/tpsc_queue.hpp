// tpsc_queue.hpp -- a whole solve queued without host waits (tpsc_capi.cpp), for the batch driver.
//
// tpsc_queue_solve re-targets a context to a point of its grid and precision and queues mu0,
// chi0, the Usp / Uch / mu root searches (a fixed number of Halley passes each, one launch per
// pass, settling into device memory that the stages after them read), the vertex, the
// self-energy and the Lambda_jj bubble + MT on the context's main stream, then one kernel that
// writes the results into `slot_dev` (the device address of tpsc_queue_slot_doubles doubles of
// pinned, mapped host memory).  Nothing waits: points of several contexts run on the GPU back to
// back while the host queues the next ones.  After the stream has drained, tpsc_queue_result
// reads the slot (its host address):
//   TPSC_OK        the point's results, bit for bit those of tpsc_solve;
//   kQueueLonger   a search still running when its passes ran out: queue it again with
//                  long_searches;
//   kQueuePromote  a mixed-precision point enhanced enough for the FP64 chi0 stage
//                  (1/(1 - Usp chi0max/2) > 8): queue it again with chi0_fp64, as tpsc_solve
//                  redoes it;
//   TPSC_E_NOROOT  a failed search or mu0 without a root: solve it with tpsc_solve, whose
//                  bracketed fallbacks take over.
#pragma once
#include <cstddef>

#include "../../include/tpsc_capi.h"

// Halley passes queued per search {Usp, Uch, mu}.  Usp settles in 3-5 passes, up to 15 next to
// the antiferromagnetic instability; Uch in 4-6; mu in 3-4.  A pass queued after its search has
// settled returns at once, but its launch still costs the host ~10 us, the batch's bottleneck.
constexpr int kQueuedPasses[3] = {7, 6, 5};
constexpr int kQueuedPassesLong[3] = {20, 10, 8};
constexpr int kQueuePromote = 1;
constexpr int kQueueLonger = 2;

size_t tpsc_queue_slot_doubles(const tpsc_ctx* ctx);
int tpsc_queue_solve(tpsc_ctx* ctx, const tpsc_params* p, double* slot_dev, bool chi0_fp64, bool long_searches);
int tpsc_queue_sync(tpsc_ctx* ctx);
int tpsc_queue_result(const tpsc_params* p, const double* slot, bool chi0_fp64, tpsc_batch_result* r,
                      double* chijj, int ldc);
