// Batched symmetric eigensolver (eig.hip).
#pragma once
#include <cstdint>

#include "device.hpp"

namespace tk {

// One n x n symmetric matrix at element offset a_off of the input buffer
// (overwritten); eigenvalues at w_off, eigenvectors (columns, row-major n x n)
// at v_off of the output buffers.  The global workspace for matrices larger
// than the LDS limit uses the same a_off.
struct EigJob {
  int64_t a_off;
  int64_t w_off;
  int64_t v_off;
  int32_t n;
  int32_t pad;
};

int eig_max_n();
void launch_eig(Context* ctx, const EigJob* jobs_dev, int n_jobs, double* A, double* Vws, double* W, double* Vout,
                bool descending, int max_sweeps = 40);

}  // namespace tk
