/*
 * svm_oracle_par.c -- TEST INFRASTRUCTURE ONLY (CPU baseline leg).
 *
 * Shot-parallel driver over the oracle: OpenMP over shots, one state per
 * thread, per-shot streams (the device executor's seeding convention).
 * The reference itself runs shots serially on one stream
 * (src/clifft/svm/svm.cc:218-252) and only threads amplitude sweeps at
 * k >= 18 (src/clifft/svm/svm_internal.h:60); this driver is the
 * shot-parallel "mode B" CPU baseline.
 */
#include "clifft_svm.h"

#include <omp.h>
#include <stdint.h>
#include <stdlib.h>

#define ORC_EXPORT __attribute__((visibility("default")))

int orc_sample(const clifft_program_view* p, uint64_t shots, uint64_t seed, int mode, uint64_t first_shot,
               uint8_t* meas, uint8_t* det, uint8_t* obs, double* ev, uint8_t* discarded);

ORC_EXPORT int orc_sample_parallel(const clifft_program_view* p, uint64_t shots, uint64_t seed, int threads,
                                   uint8_t* meas, uint8_t* det, uint8_t* obs, double* ev) {
    if (threads <= 0) threads = omp_get_max_threads();
    int rc = 0;
#pragma omp parallel num_threads(threads) reduction(| : rc)
    {
        int t = omp_get_thread_num(), nt = omp_get_num_threads();
        uint64_t lo = shots * (uint64_t)t / (uint64_t)nt, hi = shots * (uint64_t)(t + 1) / (uint64_t)nt;
        if (hi > lo)
            rc |= orc_sample(p, hi - lo, seed, 1, lo, meas ? meas + lo * p->num_measurements : NULL,
                             det ? det + lo * p->num_detectors : NULL, obs ? obs + lo * p->num_observables : NULL,
                             ev ? ev + lo * p->num_exp_vals : NULL, NULL);
    }
    return rc;
}

ORC_EXPORT int orc_max_threads(void) { return omp_get_max_threads(); }
