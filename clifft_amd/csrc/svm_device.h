// svm_device.h -- device-side data layout and the per-shot interpreter core
// for gfx950 (CDNA4, 64-wide wavefronts).
//
// Execution model ("team interpreter"):
//   * One shot is run by one team: either a single wavefront (several teams
//     share a workgroup, no workgroup barriers) or a whole workgroup of up to
//     16 wavefronts.  Every lane walks the same bytecode, so opcode dispatch,
//     the Pauli frame (p_x, p_z), gamma, the active rank k and the
//     xoshiro256++ state are wavefront-uniform values (scalar registers for
//     the integer parts).
//   * The 2^k complex128 amplitudes live in LDS (k <= 13) or in a per-team
//     slice of a global scratch buffer (larger k); array opcodes sweep them
//     with the team's lanes striding over pair/quad indices.
//   * Measurement probabilities and EXP_VAL sums are team reductions with a
//     fixed association order (lane butterfly, then waves in index order), so
//     a given launch shape is bit-reproducible.
//   * Record bytes (measurement/detector/observable) live in LDS, one private
//     copy per wavefront, so read-modify-write opcodes need no barrier.
//
// Semantics restate the reference VM opcode handlers
// (src/clifft/svm/svm_kernels.inl:67-2173, svm_forced_kernels.cc:23-306,
// svm.h:40-90,167-248, svm_internal.h:46-79); each handler cites its source.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "clifft_svm.h"

#define CSVM_INLINE __device__ __forceinline__

namespace csvm {

constexpr double kInvSqrt2 = 0.70710678118654752440;
constexpr double kDustEpsilon = 1e-18;
constexpr double kUnreachable = -1.7976931348623157e308;
constexpr uint32_t kNoNoise = 0xFFFFFFFFu;

// Complex amplitude: a native two-lane double vector (loads/stores are single
// 16-byte memory operations in any address space).
typedef double dv2 __attribute__((ext_vector_type(2)));

// Device copy of a program (all pointers are device memory).
// Marker (payload bytes 8..9) of a lane-bytecode record parity held inline.
constexpr uint32_t kLaneInlineParity = 1;
// Lane-bytecode opcode (outside the reference's range) of a packed run of
// consecutive frame opcodes: axis_1 = the ops it applies (0..12; the run's
// opcodes that are the identity on every shot from the zero frame are left
// out), axis_2 = the bytecode instructions it replaces minus one (skipped
// after it); payload words 1..3 hold 4 ops each, 16 bits per op: opcode
// (bits 0-2), first axis (3-8), second axis (9-14).
constexpr uint32_t kLaneFrameRun = 0xF0;
constexpr uint32_t kLaneFrameRunMax = 12;
// Lane-bytecode opcode of a dormant measurement (MEAS_DORMANT_STATIC or
// MEAS_DORMANT_RANDOM, kept in the reserved byte) followed by an
// APPLY_PAULI conditioned on that measurement's record: payload words 2 and
// 3 hold the APPLY_PAULI masks, axis_2 = 1 (the APPLY_PAULI is skipped).
// Lane superinstructions are the opcodes from kLaneFrameRun up; each names
// in axis_2 how many instructions after it it covers.
constexpr uint32_t kLaneMeasPauli = 0xF1;
// Lane-bytecode opcode of a run of 2 or more DETECTORs whose parities are
// all inline: the first instruction of the run, axis_2 = run length - 1; the
// handler reads each detector's own instruction words.
constexpr uint32_t kLaneDetectorRun = 0xF2;
// Lane- and team-bytecode opcode of a frame fan: two or more consecutive
// frame CZs that share qubit p, or CNOTs that share their control or their
// target p (axis_1 = p, reserved byte = the kind below, payload word 1 = the
// mask m of the ops' other qubits, axis_2 = the instructions covered minus
// one).  On the one-word frame (x, z):
//   kFanCz:      z ^= x_p ? m : 0, z_p ^= parity(x & m), gamma negated when
//                x_p and parity(x & m) are both 1;
//   kFanControl: x ^= x_p ? m : 0, z_p ^= parity(z & m);
//   kFanTarget:  x_p ^= parity(x & m), z ^= z_p ? m : 0.
// No op of a fan changes a bit another op of it reads, so these equal the
// ops applied one at a time.
constexpr uint32_t kLaneFrameFan = 0xF3;
// Team-bytecode opcode of consecutive ARRAY_CNOTs that share their control
// (axis_1), payload word 1 = the mask of their targets (a target listed
// twice cancels), axis_2 = the instructions covered minus one.  The CNOTs
// commute; the amplitude permutation is x ^= mask where bit axis_1 of x is
// set, and the frame update a kFanControl fan.  Reserved byte 1: the
// ARRAY_MULTI_CZ from the same control (or ARRAY_CZ on it) that follows them
// is covered too, its target mask in payload word 2 (two or more CNOTs, or
// one and the CZ).
constexpr uint32_t kArrayFanout = 0xF4;
enum : uint32_t { kFanCz = 0, kFanControl = 1, kFanTarget = 2 };

struct DevProgram {
    const clifft_instruction* bc;
    const clifft_instruction* lane_bc;  // lane interpreter copy, operands inlined
    const clifft_instruction* team_bc;  // team interpreter copy: frame runs packed (one-word frames, zero initial frame)
    const uint64_t* pauli_x;
    const uint64_t* pauli_z;
    const uint8_t* pauli_sign;
    const uint64_t* ev_x;
    const uint64_t* ev_z;
    const uint8_t* ev_sign;
    const uint32_t* site_chan_off;
    const uint64_t* chan_x;
    const uint64_t* chan_z;
    const double* chan_prob;
    const double* chan_cum;   // running sum of chan_prob within each site
    const double* hazards;
    const uint32_t* hz_guide;  // hz_buckets + 1 entries (see hazard_upper_bound)
    const uint32_t* ro_meas_idx;
    const double* ro_prob;
    const uint32_t* det_off;
    const uint32_t* det_tgt;
    const uint32_t* obs_off;
    const uint32_t* obs_tgt;
    const clifft_u2_node* u2;
    const clifft_u4_node* u4;
    const uint8_t* expected_obs;
    uint32_t n_instr;
    uint32_t words;
    uint32_t peak_rank;
    uint32_t num_meas;
    uint32_t total_meas_slots;
    uint32_t num_det;
    uint32_t num_obs;
    uint32_t num_ev;
    uint32_t num_sites;
    uint32_t n_expected_obs;
    uint32_t hz_buckets;  // 0: no guide (no sites or zero total hazard)
    uint32_t pad0;
    double hz_scale;      // hz_buckets / hazards[num_sites - 1]
};

// Host-visible state blob for single-shot execution from a preset state.
struct PresetState {
    double gamma[2];
    uint64_t rng[4];
    uint32_t active_k;
    uint32_t discarded;
    uint32_t next_noise_idx;
    uint32_t forced_reachable;
    double forced_log_probability;
    uint64_t dust_clamps;
    uint64_t px[16];
    uint64_t pz[16];
};

enum : uint32_t { MODE_SAMPLE = 0, MODE_FORCED = 1, MODE_PRESET = 2, MODE_KFAULT = 3 };

// Importance sampling with exactly k faults per shot: the fault-set law over
// the n = num_sites + num_readout sites (svm.cc:356-546).  Uniform mode
// draws a partial Fisher-Yates prefix; otherwise a walk over the table
// dp[i * (k + 1) + j] = e_j(w[i..n)) with odds ratios w.
constexpr uint32_t kMaxForcedFaults = 64;
struct KfTables {
    const double* w;   // [n]
    const double* dp;  // [(n + 1) x (k + 1)]
    uint32_t n;        // sites + readout entries
    uint32_t n_q;      // quantum noise sites (indices below n_q)
    uint32_t k;
    uint32_t uniform;
};

struct ShotArgs {
    uint64_t shot_begin;  // global index of the first shot of this launch
    uint64_t shot_count;
    uint64_t seed;
    uint32_t mode;
    uint32_t flags;           // CLIFFT_REC_PACKED
    uint32_t rec_stride;      // LDS bytes of one wavefront's record block
    uint32_t amp_elems;       // amplitudes reserved per team (2^peak_rank)
    uint8_t* out_meas;
    uint8_t* out_det;
    uint8_t* out_obs;
    double* out_ev;
    uint8_t* out_disc;
    double* out_logp;                 // MODE_FORCED
    const uint8_t* forced_records;    // MODE_FORCED [shot_count x num_meas]
    dv2* gscratch;                    // global-resident amplitudes
    PresetState* preset;              // MODE_PRESET
    dv2* preset_amps;                 // MODE_PRESET amplitudes (2^peak_rank)
    uint8_t* preset_records;          // MODE_PRESET meas|det|obs bytes
    double* preset_ev;                // MODE_PRESET exp_vals
    unsigned long long* dust_counter; // optional telemetry
    const KfTables* kf;               // MODE_KFAULT
    uint32_t amp_slices;              // global tier: amplitude slices per team (2: fused runs; 3: and candidate folds)
    uint32_t pend_defer;              // with amp_slices >= 2: measurement passes read through pending expansions
};

// Opcodes that never read or write amplitudes nor depend on them, so
// pending expansions may stay pending across them.
constexpr uint64_t kNoAmpOps =
    (1ull << CLIFFT_OP_FRAME_CNOT) | (1ull << CLIFFT_OP_FRAME_CZ) | (1ull << CLIFFT_OP_FRAME_H) |
    (1ull << CLIFFT_OP_FRAME_S) | (1ull << CLIFFT_OP_FRAME_S_DAG) | (1ull << CLIFFT_OP_FRAME_SWAP) |
    (1ull << CLIFFT_OP_EXPAND) | (1ull << CLIFFT_OP_EXPAND_T) | (1ull << CLIFFT_OP_EXPAND_T_DAG) |
    (1ull << CLIFFT_OP_EXPAND_ROT) | (1ull << CLIFFT_OP_MEAS_DORMANT_STATIC) |
    (1ull << CLIFFT_OP_MEAS_DORMANT_RANDOM) | (1ull << CLIFFT_OP_APPLY_PAULI) | (1ull << CLIFFT_OP_NOISE) |
    (1ull << CLIFFT_OP_NOISE_BLOCK) | (1ull << CLIFFT_OP_READOUT_NOISE) | (1ull << CLIFFT_OP_DETECTOR) |
    (1ull << CLIFFT_OP_OBSERVABLE);
constexpr uint64_t kExpandOps = (1ull << CLIFFT_OP_EXPAND) | (1ull << CLIFFT_OP_EXPAND_T) |
                                (1ull << CLIFFT_OP_EXPAND_T_DAG) | (1ull << CLIFFT_OP_EXPAND_ROT);
// Permutation/phase opcodes that a fused run may hold, and what may sit
// between them (the amplitude-free opcodes other than EXPAND*).
constexpr uint64_t kMonoOps =
    (1ull << CLIFFT_OP_ARRAY_CNOT) | (1ull << CLIFFT_OP_ARRAY_CZ) | (1ull << CLIFFT_OP_ARRAY_SWAP) |
    (1ull << CLIFFT_OP_ARRAY_MULTI_CNOT) | (1ull << CLIFFT_OP_ARRAY_MULTI_CZ) | (1ull << CLIFFT_OP_ARRAY_S) |
    (1ull << CLIFFT_OP_ARRAY_S_DAG) | (1ull << CLIFFT_OP_ARRAY_T) | (1ull << CLIFFT_OP_ARRAY_T_DAG) |
    (1ull << CLIFFT_OP_ARRAY_ROT);
constexpr uint64_t kMonoKeepOps = kMonoOps | (kNoAmpOps & ~kExpandOps);
// Active measurements, which apply a deferred run in their probability pass.
constexpr uint64_t kMonoMeasOps =
    (1ull << CLIFFT_OP_MEAS_ACTIVE_DIAGONAL) | (1ull << CLIFFT_OP_MEAS_ACTIVE_INTERFERE) |
    (1ull << CLIFFT_OP_SWAP_MEAS_INTERFERE) | (1ull << CLIFFT_OP_MEAS_ACTIVE_DIAGONAL_FORCED) |
    (1ull << CLIFFT_OP_MEAS_ACTIVE_INTERFERE_FORCED) | (1ull << CLIFFT_OP_SWAP_MEAS_INTERFERE_FORCED);

// One deferred permutation/phase opcode of a fused run (global tier): the
// run's opcodes are applied to every amplitude in one out-of-place pass.
// The pass applies each in one branch-free form, to an entry at index x:
//   phase    when x & qa is nonzero and popc(x & qm) is odd, pk 1 negates
//            the entry and pk 2 multiplies it by (pr, pi) (pk 0: none);
//   permute  x ^= px when popc(x & pm) is odd (pm 0: none).
// CNOT(c, t): pm = c, px = t.  SWAP(a, b): pm = px = a | b (flips both bits
// when they differ).  MULTI_CNOT(t, m): pm = m, px = t.  CZ(a, b): qa = a,
// qm = b, pk 1.  MULTI_CZ(a, m): qa = a, qm = m, pk 1.  S/S_DAG/T/T_DAG/ROT
// on a: qa = qm = a, pk 2 (masks as single bits).  kind, a and b name the
// opcode for the single-opcode sweeps.
enum : uint32_t {
    MONO_CNOT = 0,
    MONO_CZ = 1,
    MONO_SWAP = 2,
    MONO_MULTI_CNOT = 3,
    MONO_MULTI_CZ = 4,
    MONO_PHASE = 5,
    MONO_FANOUT = 6,    // CNOTs from control a to every target in mask m
    MONO_FANOUT_CZ = 7  // the same, then a MULTI_CZ from a (phase part qa, qm)
};
struct MonoOp {
    uint32_t kind;
    uint32_t a;  // control / first axis / MULTI axis / phase axis
    uint32_t b;  // target / second axis
    uint32_t pm, px, qa, qm, pk;
    double pr, pi;  // MONO_PHASE factor
};
constexpr uint32_t kMaxMono = 32;
constexpr uint32_t kMonoBytes = kMaxMono * sizeof(MonoOp);

// Team-tier LDS scratch of a workgroup: the reduction area (16 waves x 4
// doubles) and, per wavefront, the phases of pending fused expansions
// (kMaxFusedExpand (re, im) pairs, padded to 16 doubles).
constexpr uint32_t kMaxFusedExpand = 5;
constexpr uint32_t kTeamScratchBytes = 16 * 4 * 8 + 16 * 16 * 8;

// Lane tier (one shot per lane): words of bit-packed records per lane, and
// the LDS bytes of one wavefront: 2^peak amplitude rows of 64 x 16 B, then
// the measurement, detector and observable words, 64 x 4 B each.
__host__ __device__ inline uint32_t lane_words(uint32_t bits) { return (bits + 31u) >> 5; }
__host__ __device__ inline uint32_t lane_wave_bytes(uint32_t peak_rank, uint32_t slots, uint32_t ndet, uint32_t nobs) {
    return (64u * 16u << peak_rank) + 256u * (lane_words(slots) + lane_words(ndet) + lane_words(nobs));
}

// Batched executor: one work-queue entry is a contiguous range of shots of
// one program, run by one team of its launch group.
struct BatchItem {
    uint32_t prog;  // index into BatchArgs::progs / args / seeds
    uint32_t shot_lo, shot_hi;
    uint32_t pad;
};

// Arguments of one launch group of a batch.
struct BatchArgs {
    const BatchItem* items;          // this group's items, by decreasing estimated cost
    const DevProgram* const* progs;  // device copy of each program's DevProgram
    const ShotArgs* args;            // per program (outputs; the group's amp_elems and rec_stride)
    const uint64_t* seeds;           // per program
    uint32_t* counter;               // this group's queue head, cleared by a fill kernel before each launch
    dv2* gscratch;                   // global-amplitude groups: team_elems per team
    uint64_t team_elems;             // amp_elems x amp_slices of the group
    uint32_t n_items;
    uint32_t wave_stride;            // lane groups: LDS bytes per wavefront
};

// Read-only program tables are dereferenced through constant-address-space
// pointers so uniform reads become scalar (s_load) instructions.
template <class T>
using cptr = const __attribute__((address_space(4))) T*;
template <class T>
CSVM_INLINE cptr<T> cp(const T* p) {
    return (cptr<T>)(p);
}

// Byte read from a constant-space table through the aligned dword that
// contains it.  Scalar loads require a dword-aligned base address; forming
// base + byte_index and reusing it as the base of a later scalar load would
// fetch from a misaligned address, so byte reads never create such a base.
// Tables are padded to at least 8 bytes, so the dword stays in bounds.
CSVM_INLINE uint32_t load_u8(cptr<uint8_t> base, uint32_t idx) {
    uint64_t a = (uint64_t)(base) + idx;
    cptr<uint32_t> w = (cptr<uint32_t>)(a & ~3ull);
    return (*w >> (8u * (uint32_t)(a & 3ull))) & 0xFFu;
}

CSVM_INLINE uint64_t uni64_(uint64_t x) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}
// Hide a table pointer's provenance from loop-invariant code motion so table
// fields are loaded where they are used instead of pinning registers across
// the interpreter loop.
template <class T>
CSVM_INLINE cptr<T> opaque(cptr<T> p) {
    uint64_t v = uni64_((uint64_t)p);
    __asm__ volatile("" : "+s"(v));
    return (cptr<T>)v;
}

// ---------------------------------------------------------------------------
// Uniform helpers
// ---------------------------------------------------------------------------
CSVM_INLINE uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
CSVM_INLINE uint64_t uni64(uint64_t x) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}
CSVM_INLINE double unid(double x) { return __longlong_as_double((long long)uni64((uint64_t)__double_as_longlong(x))); }


CSVM_INLINE dv2 mk(double re, double im) {
    dv2 r;
    r.x = re;
    r.y = im;
    return r;
}
CSVM_INLINE dv2 cmul(dv2 a, dv2 b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
CSVM_INLINE dv2 cadd(dv2 a, dv2 b) { return a + b; }
CSVM_INLINE dv2 csub(dv2 a, dv2 b) { return a - b; }
CSVM_INLINE dv2 cscale(dv2 a, double s) { return a * s; }
CSVM_INLINE dv2 cneg(dv2 a) { return -a; }

// ---------------------------------------------------------------------------
// Noise-site search (svm.h:229-243, svm_kernels.inl:1886-1920)
// ---------------------------------------------------------------------------
// First i in [lo, hi) with hz[i] > target, hi if none (hz non-decreasing).
template <class PD>
CSVM_INLINE uint32_t upper_bound_in(PD hz, uint32_t lo, uint32_t hi, double target) {
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (hz[mid] <= target)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// std::upper_bound(hz, hz + n, target) for n >= 1 (the gap sampler's next
// firing site).  The guide holds, for G buckets of width H / G over
// [0, H = hz[n-1]), the upper bound of each bucket's lower edge
// (guide[G] = n).  First try: that bound or the site after it, each
// accepted only if its neighbours bracket the target (with G about the site
// count most buckets hold one site).  Then the sites within one bucket
// either side of the target's bucket, the answer checked against its
// neighbours; should the bucket index have rounded past that window, a
// search of the whole table.  The result is upper_bound's in every case.
template <class PD, class PU>
CSVM_INLINE uint32_t hazard_upper_bound(PD hz, uint32_t n, PU guide, uint32_t G, double scale, double target) {
    if (!(target < hz[n - 1])) return n;
    if (G == 0) return upper_bound_in(hz, 0, n, target);
    const double fb = target * scale;
    const uint32_t b = fb < (double)G ? (uint32_t)fb : G - 1u;
    {
        const uint32_t g0 = guide[b];
        // three loads at once, indices clamped into the table (hazards are
        // >= 0, so -1 stands for "no site before", +inf for "none after")
        const double hp = hz[g0 > 0 ? g0 - 1u : 0u], h0 = hz[g0 < n ? g0 : n - 1u],
                     h1 = hz[g0 + 1u < n ? g0 + 1u : n - 1u];
        if ((g0 == 0 ? -1.0 : hp) <= target) {
            if (g0 >= n || h0 > target) return g0;
            if (g0 + 1u >= n || h1 > target) return g0 + 1u;
        }
    }
    uint32_t lo = upper_bound_in(hz, guide[b > 0 ? b - 1u : 0u], guide[b + 2u < G ? b + 2u : G], target);
    if ((lo > 0 && hz[lo - 1] > target) || (lo < n && hz[lo] <= target)) lo = upper_bound_in(hz, 0, n, target);
    return lo;
}

// First channel c in [c0, c1) with r < cum[c], c1 if none: the reference's
// running-sum scan, on running sums precomputed in the same order (so the
// same doubles).  r = u * cum[c1 - 1]; the channel u would pick if the
// channels were equally likely is tried first (accepted only if its running
// sums bracket r), then a binary search, as the sums never decrease.
template <class PD>
CSVM_INLINE uint32_t pick_channel(PD cum, uint32_t c0, uint32_t c1, double r, double u) {
    {
        const uint32_t cnt = c1 - c0, q = (uint32_t)(u * (double)cnt);
        const uint32_t g = c0 + (q < cnt ? q : cnt - 1u);
        const double below = cum[g > c0 ? g - 1u : g], at = cum[g];
        if ((g == c0 || below <= r) && r < at) return g;
    }
    uint32_t lo = c0, hi = c1;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (r < cum[mid])
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// svm_math.h:58-61
CSVM_INLINE uint32_t insert_zero_bit(uint32_t val, uint32_t pos) {
    uint32_t mask = (1u << pos) - 1u;
    return (val & mask) | ((val & ~mask) << 1);
}
CSVM_INLINE uint32_t insert_two_zero_bits(uint32_t val, uint32_t lo, uint32_t hi) {
    return insert_zero_bit(insert_zero_bit(val, lo), hi);
}

// ---------------------------------------------------------------------------
// PRNG: xoshiro256++ seeded by SplitMix64 (svm.h:40-90), uniform per team.
// ---------------------------------------------------------------------------
struct Rng {
    uint64_t s0, s1, s2, s3;
};

CSVM_INLINE uint64_t splitmix64_next(uint64_t& z) {
    z += 0x9e3779b97f4a7c15ULL;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

// Shot s takes SplitMix64 outputs 4s..4s+3 of the stream seeded by `seed`.
CSVM_INLINE Rng rng_for_shot(uint64_t seed, uint64_t shot) {
    uint64_t z = seed + shot * 4ULL * 0x9e3779b97f4a7c15ULL;
    Rng r;
    r.s0 = splitmix64_next(z);
    r.s1 = splitmix64_next(z);
    r.s2 = splitmix64_next(z);
    r.s3 = splitmix64_next(z);
    return r;
}

CSVM_INLINE uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

CSVM_INLINE uint64_t rng_next(Rng& r) {
    const uint64_t result = rotl64(r.s0 + r.s3, 23) + r.s0;
    const uint64_t t = r.s1 << 17;
    r.s2 ^= r.s0;
    r.s3 ^= r.s1;
    r.s1 ^= r.s2;
    r.s0 ^= r.s3;
    r.s2 ^= t;
    r.s3 = rotl64(r.s3, 45);
    return result;
}

// svm.h:150
CSVM_INLINE double rng_double(Rng& r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

}  // namespace csvm
