// Particle-pool helpers of wake.hip shared with the panel wake (panwake.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace dust {

// Insertion of the first min(*kdev, kmax) new particles (all kmax when kdev
// is NULL) into the lowest free slots, in order, as dust_wake_shed_dev;
// `status` (device int) is or-ed with 1 when the free slots run out.  No host
// synchronisation.
int wake_insert_dev(int nmax, unsigned char *free_flag, double *cen, double *dir, double *mag,
                    double *r_vortex, double *vel, int kmax, const int *kdev,
                    const double *new_cen, const double *new_partvec, const double *new_rv,
                    const double *new_vel, const dust_wake_params *prm, int *n_prt, int *status,
                    hipStream_t st);

}  // namespace dust
