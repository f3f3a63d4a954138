// XCD-local cooperative Householder column steps (tdcoop.hip).
//
// The one-stage tridiagonalisation needs a global product A22 v per column,
// which as one launch per column costs a kernel launch and a round trip to
// HBM per step.  Here one cooperative launch runs every column step of a
// batch: the rows of each matrix stay in the registers of the workgroups that
// own them for the whole reduction, and all workgroups of one matrix run on
// the same XCD (each workgroup reads the id of the XCD it was dispatched to
// and takes the next slot of that XCD; the dispatcher gives every XCD the same
// number of workgroups, checked on the device once per process), so the
// per-step exchange — p = A22 v, the next pivot row, one counter — goes
// through that XCD's shared L2 with loads that bypass the CU-local L1, without
// the device-wide L2 write-back a cross-XCD barrier would need.  Every wait is
// bounded: a wait that exceeds its limit marks the job failed, the workgroups
// leave, and the host raises.
#pragma once
#include <cstdint>
#include <vector>

#include "device.hpp"

namespace tk {

// One symmetric n x n matrix (full row-major storage at a_off of A); its d, e,
// tau and pending-update vectors at w_off.  Column steps k < k0 run here;
// afterwards A holds, as the one-launch-per-column kernel leaves it, the
// reflector of step k in column k (rows k+2..), the trailing block (rows and
// columns >= k0) with the update of step k0-2 applied, and v, A22 v of step
// k0-1 in the VP pair (k0-1) & 1.
struct CoopJob {
  int64_t a_off, w_off;
  int32_t n, k0;
};

// Whether the cooperative path can run on the context's device (the XCD
// dispatch order was confirmed and TK_TRIDIAG_COOP is not 0).
bool tridiag_coop_available(Context* ctx);

// Runs the column steps of the jobs that fit one XCD's registers and returns
// their indices; the caller runs the others one launch per column.  VP is the
// 4 * w_n pending-update buffer (v parity 0, 1; A22 v parity 0, 1), rowbuf
// 2 * w_n doubles of scratch for the published pivot rows.
std::vector<int> tridiag_columns_coop(Context* ctx, const std::vector<CoopJob>& jobs, double* A, double* D,
                                      double* E, double* TAU, double* VP, double* rowbuf, int64_t w_n);

}  // namespace tk
