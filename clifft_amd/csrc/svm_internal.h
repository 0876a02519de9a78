// svm_internal.h -- host-side declarations shared by the launchers and the
// C-ABI implementation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "clifft_svm.h"
#include "svm_device.h"
#include "svm_adjoint_args.h"

namespace csvm {

// One launch shape of the team interpreter.
struct LaunchShape {
    int nw = 1;                // frame kind: 1 or 2 register words, 0 = LDS frame (<= 16 words)
    bool block_team = false;   // team = whole workgroup (else = one wavefront)
    bool global_amps = false;  // amplitudes in global scratch (else LDS)
    bool lane = false;         // one shot per lane (small rank, <= 64 qubits)
    bool lane_dense = false;   // lane kernel built for 6 resident workgroups (small LDS columns)
    bool lane_narrow = false;  // lane kernel with 32-bit frame words (at most 32 qubits)
    uint32_t amp_slices = 1;   // global-amplitude slices per team (2: fused permutation/phase runs)
    uint32_t rec_stride = 0;   // LDS bytes of one wavefront's record block (team tiers)
    uint32_t block = 256;      // threads per workgroup
    uint32_t grid = 1;         // workgroups
    size_t lds_bytes = 0;      // dynamic LDS per workgroup
    uint32_t teams_per_block() const { return lane ? block : (block_team ? 1u : block / 64u); }
};

// Resident workgroups per compute unit for the shape's kernel.
hipError_t shot_occupancy(const LaunchShape& L, int* blocks_per_cu);

// P and A are device memory.
hipError_t launch_shots(const LaunchShape& L, const DevProgram* P, const ShotArgs* A, hipStream_t st);

// Dynamic LDS limit of the batched kernel (160 KiB less its static queue head).
constexpr size_t kBatchMaxDynLds = 160 * 1024 - 256;

// Batched persistent executor (B and counter are device memory; the counter
// is cleared by a fill kernel on `st` right before the launch).
hipError_t batch_occupancy(int nw, uint32_t block, size_t lds_bytes, int* blocks_per_cu);
hipError_t launch_batch(int nw, uint32_t grid, uint32_t block, size_t lds_bytes, const BatchArgs* B,
                        uint32_t* counter, hipStream_t st);

// Adjoint-gradient executor (P and A are device memory).
hipError_t adjoint_occupancy(bool global_amps, uint32_t block, size_t lds_bytes, int* blocks_per_cu);
hipError_t launch_adjoint(bool global_amps, uint32_t grid, uint32_t block, size_t lds_bytes, const DevProgram* P,
                          const AdjArgs* A, hipStream_t st);

}  // namespace csvm
