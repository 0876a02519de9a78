/*
 * clifft_svm.h -- C ABI of the MI355X (gfx950) Schrodinger-VM executor.
 *
 * The executor runs compiled near-Clifford programs: 32-byte bytecode
 * instructions plus a constant pool, over a dense active statevector of 2^k
 * complex128 amplitudes and a per-shot Pauli frame.  Every type below mirrors
 * the reference Program layout so a caller holding a reference
 * `clifft::CompiledModule` can fill a `clifft_program_view` without
 * re-encoding the bytecode:
 *
 *   clifft_instruction  <- clifft::Instruction        (src/clifft/backend/backend.h:95-163)
 *   CLIFFT_OP_*         <- clifft::Opcode enum order  (src/clifft/backend/backend.h:25-86)
 *   clifft_u2_node      <- clifft::FusedU2Node        (src/clifft/backend/backend.h:224-234)
 *   clifft_u4_node      <- clifft::FusedU4Node        (src/clifft/backend/backend.h:240-253)
 *   clifft_program_view <- clifft::CompiledModule + ConstantPool
 *                                                     (src/clifft/backend/backend.h:255-344)
 *
 * Conventions: every entry point returns 0 on success and a negative
 * CLIFFT_E_* code on failure; the message is thread-local and read with
 * clifft_svm_last_error().  No C++ exception crosses this boundary.  All
 * memory passed in is owned by the caller.
 */
#ifndef CLIFFT_SVM_H
#define CLIFFT_SVM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLIFFT_SVM_ABI_VERSION 2u

/* ---------------------------------------------------------------------------
 * Opcodes: numbering is the reference enum order (backend.h:25-86).
 * ------------------------------------------------------------------------- */
enum {
    CLIFFT_OP_FRAME_CNOT = 0,
    CLIFFT_OP_FRAME_CZ = 1,
    CLIFFT_OP_FRAME_H = 2,
    CLIFFT_OP_FRAME_S = 3,
    CLIFFT_OP_FRAME_S_DAG = 4,
    CLIFFT_OP_FRAME_SWAP = 5,
    CLIFFT_OP_ARRAY_CNOT = 6,
    CLIFFT_OP_ARRAY_CZ = 7,
    CLIFFT_OP_ARRAY_SWAP = 8,
    CLIFFT_OP_ARRAY_MULTI_CNOT = 9,
    CLIFFT_OP_ARRAY_MULTI_CZ = 10,
    CLIFFT_OP_ARRAY_H = 11,
    CLIFFT_OP_ARRAY_S = 12,
    CLIFFT_OP_ARRAY_S_DAG = 13,
    CLIFFT_OP_ARRAY_T = 14,
    CLIFFT_OP_ARRAY_T_DAG = 15,
    CLIFFT_OP_ARRAY_ROT = 16,
    CLIFFT_OP_ARRAY_U2 = 17,
    CLIFFT_OP_ARRAY_U4 = 18,
    CLIFFT_OP_EXPAND = 19,
    CLIFFT_OP_EXPAND_T = 20,
    CLIFFT_OP_EXPAND_T_DAG = 21,
    CLIFFT_OP_EXPAND_ROT = 22,
    CLIFFT_OP_MEAS_DORMANT_STATIC = 23,
    CLIFFT_OP_MEAS_DORMANT_RANDOM = 24,
    CLIFFT_OP_MEAS_ACTIVE_DIAGONAL = 25,
    CLIFFT_OP_MEAS_ACTIVE_INTERFERE = 26,
    CLIFFT_OP_SWAP_MEAS_INTERFERE = 27,
    CLIFFT_OP_MEAS_DORMANT_STATIC_FORCED = 28,
    CLIFFT_OP_MEAS_DORMANT_RANDOM_FORCED = 29,
    CLIFFT_OP_MEAS_ACTIVE_DIAGONAL_FORCED = 30,
    CLIFFT_OP_MEAS_ACTIVE_INTERFERE_FORCED = 31,
    CLIFFT_OP_SWAP_MEAS_INTERFERE_FORCED = 32,
    CLIFFT_OP_APPLY_PAULI = 33,
    CLIFFT_OP_NOISE = 34,
    CLIFFT_OP_NOISE_BLOCK = 35,
    CLIFFT_OP_READOUT_NOISE = 36,
    CLIFFT_OP_DETECTOR = 37,
    CLIFFT_OP_POSTSELECT = 38,
    CLIFFT_OP_OBSERVABLE = 39,
    CLIFFT_OP_EXP_VAL = 40,
    CLIFFT_NUM_OPCODES = 41
};

/* Instruction flag bits (backend.h:97-101). */
#define CLIFFT_FLAG_SIGN 1u
#define CLIFFT_FLAG_HIDDEN 2u
#define CLIFFT_FLAG_IDENTITY 4u
#define CLIFFT_FLAG_EXPECTED_ONE 8u

/* 32-byte instruction, byte-compatible with clifft::Instruction.
 * Payload views (offsets from the start of the record):
 *   math       : f64 weight_re @8, f64 weight_im @16        (ROT, EXPAND_ROT)
 *   classical  : u32 classical_idx @8, u32 expected_val @12  (measurements)
 *   pauli      : u32 cp_mask_idx @8, u32 condition_idx @12   (APPLY_PAULI, NOISE*,
 *                READOUT_NOISE, DETECTOR, POSTSELECT, OBSERVABLE)
 *   multi_gate : u64 mask @8                                 (MULTI_CNOT/CZ)
 *   u2 / u4    : u32 cp_idx @8
 *   exp_val    : u32 cp_exp_val_idx @8, u32 exp_val_idx @12 */
typedef struct clifft_instruction {
    uint8_t opcode;
    uint8_t reserved;
    uint8_t flags;
    uint8_t pad;
    uint16_t axis_1;
    uint16_t axis_2;
    union {
        struct { double weight_re, weight_im; } math;
        struct { uint32_t classical_idx, expected_val; } classical;
        struct { uint32_t cp_mask_idx, condition_idx; } pauli;
        struct { uint64_t mask; } multi_gate;
        struct { uint32_t cp_idx; } u;
        struct { uint32_t cp_exp_val_idx, exp_val_idx; } exp_val;
        uint8_t raw[24];
    } p;
} clifft_instruction;

/* Fused single-axis 2x2 node (328 bytes, same layout as clifft::FusedU2Node).
 * matrices[s] = {a, b, c, d} row-major for incoming frame state
 * s = (pz << 1) | px; complex numbers are (re, im) pairs. */
typedef struct clifft_u2_node {
    double matrices[4][4][2];
    double gamma[4][2];
    uint8_t out_states[4];
    uint8_t pad[4];
} clifft_u2_node;

/* Fused two-axis 4x4 node entry (280 bytes, clifft::FusedU4Node::Entry).
 * State index s = (pz_hi << 3) | (px_hi << 2) | (pz_lo << 1) | px_lo. */
typedef struct clifft_u4_entry {
    double matrix[4][4][2];
    double gamma[2];
    uint8_t out_state;
    uint8_t pad[7];
} clifft_u4_entry;

typedef struct clifft_u4_node {
    clifft_u4_entry entries[16];
} clifft_u4_node;

/* Read-only view of one compiled program.  Pauli masks are word-packed
 * little-endian by axis, `words = (num_qubits + 63) / 64` words per mask
 * (the PauliMaskArena layout, src/clifft/util/pauli_arena.h:63-110).
 * Noise channels are flattened in site order: site s owns channels
 * [site_chan_off[s], site_chan_off[s+1]). */
typedef struct clifft_program_view {
    uint32_t abi_version; /* CLIFFT_SVM_ABI_VERSION */
    uint32_t num_qubits;
    uint32_t peak_rank;
    uint32_t num_measurements; /* visible slots (output columns) */
    uint32_t total_meas_slots; /* visible + hidden slots (VM record size) */
    uint32_t num_detectors;
    uint32_t num_observables;
    uint32_t num_exp_vals;
    uint32_t has_postselection;
    uint32_t words;

    const clifft_instruction* bytecode;
    uint64_t num_instructions;

    /* OP_APPLY_PAULI masks */
    const uint64_t* pauli_x;
    const uint64_t* pauli_z;
    const uint8_t* pauli_sign;
    uint32_t num_pauli_masks;
    /* OP_EXP_VAL masks */
    const uint64_t* ev_x;
    const uint64_t* ev_z;
    const uint8_t* ev_sign;
    uint32_t num_ev_masks;
    /* OP_NOISE sites (CSR) and the cumulative hazard table (one per site) */
    const uint32_t* site_chan_off; /* num_sites + 1 entries */
    uint32_t num_sites;
    const uint64_t* chan_x;
    const uint64_t* chan_z;
    const double* chan_prob;
    uint32_t num_channels;
    const double* hazards; /* num_sites entries, or NULL when num_sites == 0 */
    /* OP_READOUT_NOISE entries */
    const uint32_t* ro_meas_idx;
    const double* ro_prob;
    uint32_t num_readout;
    /* detector and observable target lists (CSR) */
    const uint32_t* det_off; /* num_det_lists + 1 */
    const uint32_t* det_tgt;
    uint32_t num_det_lists;
    const uint32_t* obs_off; /* num_obs_lists + 1 */
    const uint32_t* obs_tgt;
    uint32_t num_obs_lists;
    /* fused nodes */
    const clifft_u2_node* u2_nodes;
    uint32_t num_u2;
    const clifft_u4_node* u4_nodes;
    uint32_t num_u4;
    /* expected noiseless observable parities (may be NULL) */
    const uint8_t* expected_observables;
    uint32_t num_expected_observables;
} clifft_program_view;

/* ---------------------------------------------------------------------------
 * Error codes
 * ------------------------------------------------------------------------- */
#define CLIFFT_OK 0
#define CLIFFT_E_INVALID (-1)     /* bad argument / malformed program   */
#define CLIFFT_E_UNSUPPORTED (-2) /* program outside the device limits  */
#define CLIFFT_E_HIP (-3)         /* HIP runtime error                  */
#define CLIFFT_E_NOMEM (-4)       /* allocation failure                 */
#define CLIFFT_E_NODEVICE (-5)    /* no gfx950 device visible           */

/* ---------------------------------------------------------------------------
 * Records.  Row-major [shots x n] uint8 (one byte per bit, the reference
 * SampleResult layout, svm.h:319-324) unless CLIFFT_REC_PACKED is set, in
 * which case each row is ceil(n / 8) bytes, bit j of a row at byte j / 8,
 * bit position j % 8 (numpy packbits bitorder="little").  exp_vals is
 * always float64 [shots x num_exp_vals].  Any pointer may be NULL to skip
 * that output.
 * ------------------------------------------------------------------------- */
#define CLIFFT_REC_DEVICE 1u /* output pointers are device memory */
#define CLIFFT_REC_PACKED 2u /* bit-pack meas/det/obs rows           */

typedef struct clifft_svm_records {
    uint8_t* measurements;
    uint8_t* detectors;
    uint8_t* observables;
    double* exp_vals;
    uint8_t* discarded; /* [shots] 1 if OP_POSTSELECT discarded the shot */
} clifft_svm_records;

/* Survivor summary (clifft::SurvivorResult counts, svm.h:338-356). */
typedef struct clifft_svm_survivors {
    uint64_t total_shots;
    uint64_t passed_shots;
    uint64_t logical_errors;
    uint64_t* observable_ones; /* caller array [num_observables] */
} clifft_svm_survivors;

/* Opaque device-resident program (bytecode + constant pool uploaded once,
 * launch plan and per-program scratch cached). */
typedef struct clifft_svm_program clifft_svm_program;

/* Host-side state for single-shot execution (the SchrodingerState fields the
 * reference VM tests preset, svm.h:115-289).  v holds 2^peak_rank complex128
 * as (re, im) pairs; px/pz hold `words` words; rng is the xoshiro256++ state.
 * All arrays are caller-owned host memory, read before and written after the
 * shot. */
typedef struct clifft_svm_state {
    double* v;
    uint64_t* px;
    uint64_t* pz;
    double gamma_re, gamma_im;
    uint32_t active_k;
    uint32_t discarded;
    uint8_t* meas_record; /* total_meas_slots */
    uint8_t* det_record;  /* num_detectors */
    uint8_t* obs_record;  /* num_observables */
    double* exp_vals;     /* num_exp_vals */
    uint64_t rng[4];
    uint32_t next_noise_idx;
    uint32_t forced_reachable;
    double forced_log_probability;
    const uint8_t* forced_record; /* total_meas_slots, or NULL */
    uint64_t dust_clamps;
} clifft_svm_state;

/* --- library --------------------------------------------------------------*/

/* Version of this ABI (CLIFFT_SVM_ABI_VERSION). */
uint32_t clifft_svm_abi_version(void);

/* Name of the executing backend, e.g. "hip:gfx950".  Replaces
 * clifft::svm_backend() (src/clifft/svm/svm.h:323; svm.cc:166-185). */
const char* clifft_svm_backend(void);

/* Message of the last failing call on this thread ("" if none). */
const char* clifft_svm_last_error(void);

/* --- programs -------------------------------------------------------------*/

/* Validate the view and upload it to the current HIP device.  The view's
 * arrays may be freed after this call returns. */
int clifft_svm_program_create(const clifft_program_view* view, clifft_svm_program** out);
int clifft_svm_program_destroy(clifft_svm_program* prog);

/* Static per-program facts the launch planner derived: peak_rank, the tier
 * chosen (0 = one shot per wavefront/workgroup team with LDS amplitudes,
 * 1 = one shot per workgroup team with global amplitudes, 2 = one shot per
 * lane), the number of array instructions and the algorithmic bytes of one
 * shot (sum over array instructions of their byte model at the static k). */
typedef struct clifft_svm_program_info {
    uint32_t peak_rank;
    uint32_t tier;
    uint32_t team_waves;
    uint32_t shots_per_block;
    uint64_t num_array_ops;
    double bytes_per_shot;
} clifft_svm_program_info;
int clifft_svm_program_info_get(const clifft_svm_program* prog, clifft_svm_program_info* info);

/* --- sampling --------------------------------------------------------------*/

/* Run `shots` independent shots.  Replaces clifft::sample
 * (src/clifft/svm/svm.h:331, svm.cc:191-255).  Shot s uses its own
 * xoshiro256++ stream: the SplitMix64 sequence seeded with `seed`, outputs
 * 4s .. 4s+3 (so shot 0 draws exactly what the reference's single stream
 * draws for its first shot).  Observables are XORed with the program's
 * expected_observables.  `stream` is a hipStream_t (NULL = the library's own
 * stream); the call returns after the records are written.  Programs whose
 * amplitudes live in global memory share one scratch arena per stream, so
 * launches on one stream never overlap in it; one program must not be run
 * concurrently on two streams. */
int clifft_svm_sample(clifft_svm_program* prog, uint64_t shots, uint64_t seed, uint32_t flags,
                      const clifft_svm_records* out, void* stream);

/* clifft_svm_sample that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Outputs must be device memory (flags must hold
 * CLIFFT_REC_DEVICE); records equal clifft_svm_sample's.  Programs whose
 * amplitudes live in global memory use a fixed 64 MiB device array of the
 * library instead of the per-stream arena: peak_rank <= 22 (else
 * CLIFFT_E_UNSUPPORTED), with as many teams in flight as the array holds.
 * Every call on one GPU shares that array, so calls that overlap on one GPU
 * (on different streams or threads) race; the call returns after the
 * records are written. */
int clifft_svm_sample_noalloc(clifft_svm_program* prog, uint64_t shots, uint64_t seed, uint32_t flags,
                              const clifft_svm_records* out, void* stream);

/* Survivor-only sampling.  Replaces clifft::sample_survivors
 * (src/clifft/svm/svm.h:358, svm.cc:266-345).  When `out` is non-NULL the
 * surviving rows are compacted into its arrays in shot order and
 * `summary->passed_shots` rows are valid. */
int clifft_svm_sample_survivors(clifft_svm_program* prog, uint64_t shots, uint64_t seed,
                                uint32_t flags, clifft_svm_survivors* summary,
                                const clifft_svm_records* out, void* stream);

/* Importance sampling: every shot has exactly k faults, the fault set drawn
 * from the conditional Poisson-binomial law over the quantum noise sites and
 * readout entries (or uniformly when all their probabilities are equal).
 * Replaces clifft::sample_k / sample_k_survivors (src/clifft/svm/svm.h:366-373,
 * svm.cc:347-711).  Shot s uses the per-shot stream of (seed, s) as in
 * clifft_svm_sample; shot 0 equals the reference's first shot (the uniform
 * sampler restarts its pool from the identity every shot, which the
 * reference does only for its first shot).  Infeasible k returns
 * CLIFFT_E_INVALID with the reference's message; k above 64 is
 * CLIFFT_E_UNSUPPORTED. */
int clifft_svm_sample_k(clifft_svm_program* prog, uint64_t shots, uint32_t k, uint64_t seed,
                        uint32_t flags, const clifft_svm_records* out, void* stream);
int clifft_svm_sample_k_survivors(clifft_svm_program* prog, uint64_t shots, uint32_t k,
                                  uint64_t seed, uint32_t flags, clifft_svm_survivors* summary,
                                  const clifft_svm_records* out, void* stream);
/* Per-site fault probabilities: quantum sites (sum of channel probabilities)
 * then readout entries (clifft::noise_site_probabilities, svm.h:378,
 * svm.cc:356-369).  `out` may be NULL to query *count. */
int clifft_svm_noise_site_probabilities(const clifft_svm_program* prog, double* out, uint32_t* count);

/* Exact log-probabilities of measurement records (forced execution).
 * Replaces clifft::record_probabilities (src/clifft/svm/svm.h:405,
 * src/clifft/api/record_probabilities.cc:96-191): measurement opcodes are
 * rewritten to their *_FORCED variants and one shot runs per record.
 * `records` is host memory [num_records x num_measurements] of 0/1 bytes;
 * unreachable records return -DBL_MAX. */
int clifft_svm_record_log_probs(clifft_svm_program* prog, const uint8_t* records,
                                uint64_t num_records, double* log_probs, void* stream);

/* Execute one shot from a caller-provided state (host memory in and out).
 * Replaces clifft::execute (src/clifft/svm/svm.h:296, svm.cc:158-164). */
int clifft_svm_execute(clifft_svm_program* prog, clifft_svm_state* state);

/* --- batching ---------------------------------------------------------------*/
/* No counterpart in the reference (its VM runs one program per call,
 * src/clifft/svm/svm.h:302-306).  Many programs of uneven peak rank and
 * bytecode length run as a few persistent launches, one per launch group:
 * the programs that share a kernel shape -- one shot per lane (peak_rank <=
 * lane_rank_max, at most 64 qubits; split by frame-word width and by
 * whether a workgroup holds one wavefront's columns or more), one shot
 * per wavefront (peak_rank <= wave_rank_max), one shot per workgroup with
 * LDS-resident amplitudes (<= lds_rank_max; split by team size), one
 * shot per wavefront with amplitudes in a global slice (<= gwave_rank_max),
 * or one shot per workgroup with amplitudes in a global slice
 * (global_threads per team).  A group's work items (a shot range of one program) are sorted by
 * descending estimated cost and drained by its teams from a device queue.
 * The groups run one after another on the caller's stream, or side by side
 * on the batch's own streams (concurrent = 1).
 *
 * Shot s of program i draws from the per-shot stream of (seeds[i], s), as in
 * clifft_svm_sample, so the records equal that call's; the team shape (and
 * so the order of floating-point reductions) may differ from the plain
 * launch, which moves expectation values by rounding only.
 * Outputs of a batch are device memory (CLIFFT_REC_DEVICE is implied). */

typedef struct clifft_svm_batch clifft_svm_batch;

typedef struct {
    uint32_t global_threads; /* 256, 512 or 1024: team size of the global-amplitude group */
    uint32_t wave_rank_max;  /* <= 10 (four LDS states per workgroup) */
    uint32_t lds_rank_max;   /* <= 13, >= wave_rank_max */
    uint32_t wave_chunk;     /* shots per work item, wavefront teams */
    uint32_t block_chunk;    /* shots per work item, workgroup teams */
    uint32_t concurrent;     /* 1: launch groups on parallel streams; 0: in sequence */
    int32_t lane_rank_max;   /* <= 64-qubit programs up to this rank run one shot
                                per lane (-1 = never), <= 7 */
    uint32_t lane_chunk;     /* shots per work item, lane tier (a multiple of 64) */
    uint32_t gwave_rank_max; /* global-amplitude programs up to this rank run one shot
                                per wavefront (<= 30; at or below lds_rank_max: never) */
} clifft_svm_batch_tuning;

/* One launch group of a batch's current plan. */
typedef struct {
    uint32_t kind;       /* 0 wavefront teams, 1 LDS workgroup teams, 2 global workgroup teams, 3 lanes,
                            4 global wavefront teams */
    uint32_t programs;   /* member programs */
    uint32_t items;      /* work items in its queue */
    uint32_t grid;       /* persistent workgroups */
    uint32_t block;      /* threads per workgroup */
    uint32_t lds_bytes;  /* dynamic LDS per workgroup */
    uint32_t amp_elems;  /* amplitudes per team slice (team kinds) */
    uint32_t amp_slices; /* slices per team (global kind) */
    double algorithmic_bytes; /* per launch: the members' per-shot state-traffic
                                 model times their shots, plus record bytes written */
} clifft_svm_batch_group_info;

/* outs[i] are device pointers; CLIFFT_REC_PACKED is honoured. */
int clifft_svm_batch_create(clifft_svm_program* const* programs, const uint64_t* shots,
                            uint32_t num_programs, uint32_t flags, const clifft_svm_records* outs,
                            clifft_svm_batch** out);
int clifft_svm_batch_destroy(clifft_svm_batch* batch);
int clifft_svm_batch_set_tuning(clifft_svm_batch* batch, const clifft_svm_batch_tuning* tuning);
int clifft_svm_batch_get_tuning(const clifft_svm_batch* batch, clifft_svm_batch_tuning* tuning);
/* The current plan: *num_groups groups, the first max_groups of them
 * described in info (may be NULL). */
int clifft_svm_batch_info(const clifft_svm_batch* batch, uint32_t max_groups, clifft_svm_batch_group_info* info,
                          uint32_t* num_groups);
/* Per-group kernel timing of eager runs: with timing enabled, each group's
 * launch in clifft_svm_batch_run is bracketed by timing events on the stream
 * it goes to; clifft_svm_batch_group_times waits for the last such run and
 * returns each group's kernel time in ms (graph replays are not timed). */
int clifft_svm_batch_time_groups(clifft_svm_batch* batch, uint32_t enable);
int clifft_svm_batch_group_times(clifft_svm_batch* batch, uint32_t max_groups, float* ms, uint32_t* num_groups);
/* Autotuner: times candidate tunings on the batch itself (each program
 * capped at `sample_shots` shots, outputs overwritten) with HIP events on
 * `stream`, keeps the fastest and returns it in `chosen` (may be NULL). */
int clifft_svm_batch_tune(clifft_svm_batch* batch, uint64_t sample_shots, void* stream,
                          clifft_svm_batch_tuning* chosen);
/* Every group's launch with per-program seeds (host array).  Asynchronous
 * on `stream` (NULL = the library stream). */
int clifft_svm_batch_run(clifft_svm_batch* batch, const uint64_t* seeds, void* stream);
/* Same work through a captured HIP graph (queue resets + the group
 * launches), captured on first use per stream and plan and replayed
 * afterwards. */
int clifft_svm_batch_run_graph(clifft_svm_batch* batch, const uint64_t* seeds, void* stream);

/* --- differentiation ---------------------------------------------------------*/
/* No counterpart in the reference (it has no gradients; its EXP_VAL probe is
 * src/clifft/svm/svm_kernels.inl:2031-2173).  For a noiseless,
 * measurement-free program whose parameterised Z rotations compiled to
 * ARRAY_ROT / EXPAND_ROT instructions (param_instr[i], with phase
 * exp(i param_sign[i] theta_i)), compute for a batch of theta vectors the
 * EXP_VAL results and the gradient of F = sum_j coeffs[j] * exp_vals[j] with
 * respect to theta, by one forward and one backward (adjoint) sweep per
 * theta vector on the device. */
typedef struct clifft_svm_adjoint clifft_svm_adjoint;

int clifft_svm_adjoint_create(clifft_svm_program* prog, const uint32_t* param_instr,
                              const int8_t* param_sign, uint32_t num_params,
                              clifft_svm_adjoint** out);
int clifft_svm_adjoint_destroy(clifft_svm_adjoint* adj);
/* thetas [batch x num_params]; coeffs [num_exp_vals] (needed with grads);
 * exp_vals [batch x num_exp_vals] or NULL; grads [batch x num_params] or NULL
 * (forward only).  flags: CLIFFT_REC_DEVICE = every pointer is device memory
 * and the call is asynchronous on `stream`; otherwise host memory and the
 * call returns with the results written. */
int clifft_svm_adjoint_run(clifft_svm_adjoint* adj, const double* thetas, uint64_t batch,
                           const double* coeffs, double* exp_vals, double* grads, uint32_t flags,
                           void* stream);

/* One-call form.  Without CLIFFT_BATCH_PER_PROGRAM it builds a batch (with
 * the default tuning), runs it once and destroys it; outputs may be host
 * memory (staged) or device memory.  With CLIFFT_BATCH_PER_PROGRAM it issues
 * one plain clifft_svm_sample launch per program (the unbatched baseline). */
#define CLIFFT_BATCH_PER_PROGRAM 0x100u
int clifft_svm_batch_sample(clifft_svm_program* const* programs, const uint64_t* shots,
                            const uint64_t* seeds, uint32_t num_programs, uint32_t flags,
                            const clifft_svm_records* outs, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CLIFFT_SVM_H */
