// svm_internal.h -- host-side declarations shared by the launchers and the
// C-ABI implementation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "clifft_svm.h"
#include "svm_device.h"

namespace csvm {

// One launch shape of the team interpreter.
struct LaunchShape {
    int nw = 1;                // frame kind: 1 or 2 register words, 0 = LDS frame (<= 16 words)
    bool block_team = false;   // team = whole workgroup (else = one wavefront)
    bool global_amps = false;  // amplitudes in global scratch (else LDS)
    uint32_t block = 256;      // threads per workgroup
    uint32_t grid = 1;         // workgroups
    size_t lds_bytes = 0;      // dynamic LDS per workgroup
    uint32_t teams_per_block() const { return block_team ? 1u : block / 64u; }
};

// P and A are device memory.
hipError_t launch_shots(const LaunchShape& L, const DevProgram* P, const ShotArgs* A, hipStream_t st);

// Batched persistent executor (B and counter are device memory; the counter
// is cleared by a fill kernel on `st` right before the launch).
hipError_t batch_occupancy(int nw, uint32_t block, size_t lds_bytes, int* blocks_per_cu);
hipError_t launch_batch(int nw, uint32_t grid, uint32_t block, size_t lds_bytes, const BatchArgs* B,
                        uint32_t* counter, hipStream_t st);

}  // namespace csvm
