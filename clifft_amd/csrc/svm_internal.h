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
    uint32_t amp_slices = 1;   // global-amplitude slices per team (2: fused permutation/phase runs, 3: and candidate folds)
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

// LDS of one workgroup on gfx950.
constexpr size_t kLdsBytes = 160 * 1024;

// Largest peak rank the executor accepts: amplitude indices are 32-bit.
constexpr uint32_t kMaxPeakRank = 30;

// Dynamic LDS limit of the batched kernel (160 KiB less its static queue head).
constexpr size_t kBatchMaxDynLds = 160 * 1024 - 256;

// Batched execution: one persistent launch per launch group (B is device
// memory; L.grid, L.block and L.lds_bytes are the group's launch shape).
hipError_t batch_occupancy(const LaunchShape& L, int* blocks_per_cu);
hipError_t launch_batch_group(const LaunchShape& L, const BatchArgs* B, hipStream_t st);
// Zero n queue heads with a fill kernel (never a memset: graph-capturable).
hipError_t clear_counters(uint32_t* counters, uint32_t n, hipStream_t st);

// Adjoint-gradient executor (P and A are device memory).
hipError_t adjoint_occupancy(bool global_amps, uint32_t block, size_t lds_bytes, int* blocks_per_cu);
hipError_t launch_adjoint(bool global_amps, uint32_t grid, uint32_t block, size_t lds_bytes, const DevProgram* P,
                          const AdjArgs* A, hipStream_t st);

}  // namespace csvm
