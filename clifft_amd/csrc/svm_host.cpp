// svm_host.cpp -- C-ABI implementation: program validation and upload,
// launch planning, and the sampling entry points (include/clifft_svm.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "svm_internal.h"

using namespace csvm;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) return fail(CLIFFT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

struct DeviceInfo {
    int ok = 0;
    int device = 0;
    int num_cus = 0;
    size_t lds_max = 0;
    std::string arch;
};

DeviceInfo& device_info() {
    static DeviceInfo info;
    static std::once_flag once;
    std::call_once(once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return;
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return;
        info.device = dev;
        info.num_cus = prop.multiProcessorCount;
        info.lds_max = 160 * 1024;
        info.arch = prop.gcnArchName;
        info.ok = info.arch.find("gfx950") != std::string::npos ? 1 : -1;
    });
    return info;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// The library's own stream, used when the caller passes none.
hipStream_t library_stream() {
    static hipStream_t st = nullptr;
    static std::once_flag once;
    std::call_once(once, [] { (void)hipStreamCreateWithFlags(&st, hipStreamNonBlocking); });
    return st;
}

// Global-memory amplitude scratch, one arena per stream: launches on one
// stream run in order, so every program launched there can share it.
struct Arena {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_arena_mu;
std::vector<std::pair<hipStream_t, Arena>> g_arenas;

Arena& arena_of(hipStream_t st) {
    for (auto& a : g_arenas)
        if (a.first == st) return a.second;
    g_arenas.push_back({st, Arena{}});
    return g_arenas.back().second;
}

// ---------------------------------------------------------------------------
// Static validation: walk the bytecode with the static active-rank
// trajectory and check every operand the kernels index with.
// ---------------------------------------------------------------------------
int validate(const clifft_program_view* v, std::vector<int32_t>& khist, bool& has_forced) {
    if (!v) return fail(CLIFFT_E_INVALID, "null program view");
    if (v->abi_version != CLIFFT_SVM_ABI_VERSION) return fail(CLIFFT_E_INVALID, "ABI version mismatch");
    if (v->peak_rank >= 63) return fail(CLIFFT_E_INVALID, "peak_rank >= 63 would cause undefined behavior");
    // Amplitude indices are 32-bit in every sweep (2^k and 2 * 2^(k-1) stay
    // below 2^31); a 2^30-amplitude state is 16 GiB.
    if (v->peak_rank > kMaxPeakRank)
        return fail(CLIFFT_E_UNSUPPORTED, "device executor supports peak_rank <= " + std::to_string(kMaxPeakRank) +
                                              " (32-bit amplitude indices)");
    const uint32_t nq = v->num_qubits > 0 ? v->num_qubits : std::max<uint32_t>(v->peak_rank, 1u);
    const uint32_t words = (nq + 63) / 64;
    if (v->words != words) return fail(CLIFFT_E_INVALID, "words != ceil(num_qubits / 64)");
    if (words > 16) return fail(CLIFFT_E_UNSUPPORTED, "device executor supports num_qubits <= 1024");
    if (v->total_meas_slots < v->num_measurements) return fail(CLIFFT_E_INVALID, "total_meas_slots < num_measurements");
    if (v->num_instructions > 0xFFFFFFFFull) return fail(CLIFFT_E_UNSUPPORTED, "too many instructions");
    if (v->num_sites > 0 && (!v->site_chan_off || !v->hazards)) return fail(CLIFFT_E_INVALID, "noise sites without CSR/hazards");
    if (v->num_sites > 0 && v->site_chan_off[v->num_sites] > v->num_channels)
        return fail(CLIFFT_E_INVALID, "noise channel CSR out of range");
    // The gap sampler's upper_bound needs a finite, non-decreasing hazard
    // table; otherwise a NOISE_BLOCK could revisit a site forever.
    for (uint32_t s = 0; s < v->num_sites; ++s) {
        double h = v->hazards[s];
        if (!(h >= 0.0) || !std::isfinite(h) || (s > 0 && h < v->hazards[s - 1]))
            return fail(CLIFFT_E_INVALID, "noise hazard table must be finite and non-decreasing");
    }
    // Channel probabilities are probabilities: the channel search relies on
    // their running sums never decreasing.
    for (uint32_t c = 0; c < v->num_channels; ++c)
        if (!(v->chan_prob[c] >= 0.0) || !std::isfinite(v->chan_prob[c]))
            return fail(CLIFFT_E_INVALID, "noise channel probabilities must be finite and non-negative");
    for (uint32_t i = 0; i < v->num_readout; ++i)
        if (v->ro_meas_idx[i] >= v->total_meas_slots) return fail(CLIFFT_E_INVALID, "readout meas_idx out of range");
    if (v->num_det_lists > 0 && v->det_off[v->num_det_lists] > 0)
        for (uint32_t j = 0; j < v->det_off[v->num_det_lists]; ++j)
            if (v->det_tgt[j] >= v->total_meas_slots) return fail(CLIFFT_E_INVALID, "detector target out of range");
    if (v->num_obs_lists > 0 && v->obs_off[v->num_obs_lists] > 0)
        for (uint32_t j = 0; j < v->obs_off[v->num_obs_lists]; ++j)
            if (v->obs_tgt[j] >= v->total_meas_slots) return fail(CLIFFT_E_INVALID, "observable target out of range");

    const uint32_t fbits = words * 64;
    has_forced = false;
    khist.assign(v->num_instructions, 0);
    int32_t k = 0, kmax = 0;
    auto bad = [&](uint64_t pc, const char* what) {
        return fail(CLIFFT_E_INVALID, "instruction " + std::to_string(pc) + ": " + what);
    };
    for (uint64_t pc = 0; pc < v->num_instructions; ++pc) {
        const clifft_instruction& in = v->bytecode[pc];
        khist[pc] = k;
        if (in.opcode >= CLIFFT_OP_MEAS_DORMANT_STATIC_FORCED && in.opcode <= CLIFFT_OP_SWAP_MEAS_INTERFERE_FORCED)
            has_forced = true;
        const uint32_t a1 = in.axis_1, a2 = in.axis_2, i0 = in.p.pauli.cp_mask_idx, i1 = in.p.pauli.condition_idx;
        switch (in.opcode) {
            case CLIFFT_OP_FRAME_CNOT:
            case CLIFFT_OP_FRAME_CZ:
            case CLIFFT_OP_FRAME_SWAP:
                if (a1 >= fbits || a2 >= fbits) return bad(pc, "frame axis out of range");
                break;
            case CLIFFT_OP_FRAME_H:
            case CLIFFT_OP_FRAME_S:
            case CLIFFT_OP_FRAME_S_DAG:
                if (a1 >= fbits) return bad(pc, "frame axis out of range");
                break;
            case CLIFFT_OP_ARRAY_CNOT:
            case CLIFFT_OP_ARRAY_CZ:
            case CLIFFT_OP_ARRAY_SWAP:
                if (a1 >= (uint32_t)k || a2 >= (uint32_t)k || a1 == a2) return bad(pc, "2-axis array op needs distinct active axes");
                break;
            case CLIFFT_OP_ARRAY_MULTI_CNOT:
            case CLIFFT_OP_ARRAY_MULTI_CZ:
                if (a1 >= (uint32_t)k) return bad(pc, "multi-gate axis not active");
                if (in.p.multi_gate.mask & (1ULL << a1)) return bad(pc, "multi-gate mask contains its own axis");
                break;
            case CLIFFT_OP_ARRAY_H:
            case CLIFFT_OP_ARRAY_S:
            case CLIFFT_OP_ARRAY_S_DAG:
                if (a1 >= (uint32_t)k) return bad(pc, "array axis not active");
                break;
            case CLIFFT_OP_ARRAY_T:
            case CLIFFT_OP_ARRAY_T_DAG:
            case CLIFFT_OP_ARRAY_ROT:
                if (a1 >= fbits) return bad(pc, "axis out of range");
                break;
            case CLIFFT_OP_ARRAY_U2:
                if (a1 >= fbits) return bad(pc, "axis out of range");
                if (in.p.u.cp_idx >= v->num_u2) return bad(pc, "U2 node index out of range");
                break;
            case CLIFFT_OP_ARRAY_U4:
                if (!(a1 < a2 && a2 < (uint32_t)k)) return bad(pc, "U4 needs lo < hi < k");
                if (in.p.u.cp_idx >= v->num_u4) return bad(pc, "U4 node index out of range");
                break;
            case CLIFFT_OP_EXPAND:
            case CLIFFT_OP_EXPAND_T:
            case CLIFFT_OP_EXPAND_T_DAG:
            case CLIFFT_OP_EXPAND_ROT:
                if (a1 != (uint32_t)k) return bad(pc, "EXPAND must target axis k");
                ++k;
                if ((uint32_t)k > v->peak_rank) return bad(pc, "EXPAND exceeds peak_rank");
                break;
            case CLIFFT_OP_MEAS_DORMANT_STATIC:
            case CLIFFT_OP_MEAS_DORMANT_RANDOM:
            case CLIFFT_OP_MEAS_DORMANT_STATIC_FORCED:
            case CLIFFT_OP_MEAS_DORMANT_RANDOM_FORCED:
                if (a1 >= fbits) return bad(pc, "axis out of range");
                if (in.p.classical.classical_idx >= v->total_meas_slots) return bad(pc, "record index out of range");
                break;
            case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL:
            case CLIFFT_OP_MEAS_ACTIVE_INTERFERE:
            case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL_FORCED:
            case CLIFFT_OP_MEAS_ACTIVE_INTERFERE_FORCED:
                if (k < 1 || a1 != (uint32_t)(k - 1)) return bad(pc, "active measurement must target axis k-1");
                if (in.p.classical.classical_idx >= v->total_meas_slots) return bad(pc, "record index out of range");
                --k;
                break;
            case CLIFFT_OP_SWAP_MEAS_INTERFERE:
            case CLIFFT_OP_SWAP_MEAS_INTERFERE_FORCED:
                if (k < 1 || a2 != (uint32_t)(k - 1) || a1 >= (uint32_t)k) return bad(pc, "swap-measure axes invalid");
                if (in.p.classical.classical_idx >= v->total_meas_slots) return bad(pc, "record index out of range");
                --k;
                break;
            case CLIFFT_OP_APPLY_PAULI:
                if (i0 >= v->num_pauli_masks || i1 >= v->total_meas_slots) return bad(pc, "APPLY_PAULI index out of range");
                break;
            case CLIFFT_OP_NOISE:
                if (i0 >= v->num_sites) return bad(pc, "noise site out of range");
                break;
            case CLIFFT_OP_NOISE_BLOCK:
                if ((uint64_t)i0 + i1 > v->num_sites) return bad(pc, "noise block out of range");
                break;
            case CLIFFT_OP_READOUT_NOISE:
                if (i0 >= v->num_readout) return bad(pc, "readout entry out of range");
                break;
            case CLIFFT_OP_DETECTOR:
            case CLIFFT_OP_POSTSELECT:
                if (i0 >= v->num_det_lists || i1 >= v->num_detectors) return bad(pc, "detector index out of range");
                break;
            case CLIFFT_OP_OBSERVABLE:
                if (i0 >= v->num_obs_lists || i1 >= v->num_observables) return bad(pc, "observable index out of range");
                break;
            case CLIFFT_OP_EXP_VAL:
                if (in.p.exp_val.cp_exp_val_idx >= v->num_ev_masks || in.p.exp_val.exp_val_idx >= v->num_exp_vals)
                    return bad(pc, "EXP_VAL index out of range");
                break;
            default:
                return bad(pc, "unknown opcode");
        }
        kmax = std::max(kmax, k);
    }
    if ((uint32_t)kmax > v->peak_rank) return fail(CLIFFT_E_INVALID, "active rank exceeds peak_rank");
    return CLIFFT_OK;
}

// Algorithmic bytes of one array instruction at pre-op rank k (16 B per
// amplitude; read + write of what the opcode must touch).
double op_bytes(const clifft_instruction& in, int32_t k, uint32_t fbits) {
    double K = std::ldexp(1.0, k);
    switch (in.opcode) {
        case CLIFFT_OP_ARRAY_H:
        case CLIFFT_OP_ARRAY_U4:
        case CLIFFT_OP_EXPAND:
        case CLIFFT_OP_EXPAND_T:
        case CLIFFT_OP_EXPAND_T_DAG:
        case CLIFFT_OP_EXPAND_ROT:
            return 32.0 * K;
        case CLIFFT_OP_ARRAY_U2:
            return in.axis_1 < (uint32_t)k ? 32.0 * K : 0.0;
        case CLIFFT_OP_ARRAY_CNOT:
        case CLIFFT_OP_ARRAY_SWAP:
        case CLIFFT_OP_ARRAY_MULTI_CNOT:
        case CLIFFT_OP_ARRAY_S:
        case CLIFFT_OP_ARRAY_S_DAG:
            return 16.0 * K;
        case CLIFFT_OP_ARRAY_T:
        case CLIFFT_OP_ARRAY_T_DAG:
        case CLIFFT_OP_ARRAY_ROT:
            return in.axis_1 < (uint32_t)k ? 16.0 * K : 0.0;
        case CLIFFT_OP_ARRAY_CZ:
        case CLIFFT_OP_ARRAY_MULTI_CZ:
            return 8.0 * K;
        case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL:
        case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL_FORCED:
            return 16.0 * K;
        case CLIFFT_OP_MEAS_ACTIVE_INTERFERE:
        case CLIFFT_OP_SWAP_MEAS_INTERFERE:
        case CLIFFT_OP_MEAS_ACTIVE_INTERFERE_FORCED:
        case CLIFFT_OP_SWAP_MEAS_INTERFERE_FORCED:
            return 40.0 * K;
        case CLIFFT_OP_EXP_VAL:
            return 16.0 * K;
        default:
            (void)fbits;
            return 0.0;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Program object
// ---------------------------------------------------------------------------
struct clifft_svm_program {
    void* dev_blob = nullptr;
    DevProgram dp{};                          // host copy (device pointers)
    DevProgram* dp_dev = nullptr;             // device copy, sampling bytecode
    DevProgram* dp_forced_dev = nullptr;      // device copy, *_FORCED bytecode
    clifft_instruction* forced_bc = nullptr;  // device bytecode with *_FORCED opcodes
    std::vector<clifft_instruction> host_bc;
    std::vector<int32_t> khist;
    std::vector<uint8_t> expected_obs;
    uint32_t num_qubits = 0, peak_rank = 0, words = 1, nw = 1;
    uint32_t num_meas = 0, total_slots = 0, num_det = 0, num_obs = 0, num_ev = 0;
    bool has_postselection = false;
    bool has_forced_ops = false;  // bytecode already contains *_FORCED opcodes
    bool has_mono_runs = false;   // some run of >= 2 permutation/phase opcodes could be fused
    bool has_mono_meas = false;   // such a run directly precedes an active measurement
    bool has_exp_meas = false;    // an EXPAND reaches an active measurement across such runs
    bool frame32 = false;         // every frame axis and frame mask fits 32 bits
    uint64_t num_array_ops = 0;
    double bytes_per_shot = 0.0;
    // importance sampling: per-site fault probabilities (quantum sites, then
    // readout entries) and the device tables of the last k used
    std::vector<double> fault_probs;
    uint32_t num_sites = 0;
    void* kf_dev = nullptr;
    size_t kf_dev_bytes = 0;
    int64_t kf_k = -1;
    // launch planning
    LaunchShape base_shape;
    // staging for host outputs (grown on demand)
    void* out_dev = nullptr;
    size_t out_dev_bytes = 0;
    unsigned long long* dust_dev = nullptr;
    ShotArgs* args_dev = nullptr;  // ring of launch-argument slots
    uint32_t args_slot = 0;
};

static constexpr uint32_t kArgSlots = 64;

namespace {

int frame_nw(uint32_t words) {
    return words <= 1 ? 1 : 0;  // 0 = frame spread over the wavefront's lanes
}

// LDS bytes of one wavefront's record block (+ the k-fault scratch of
// 5 * kMaxForcedFaults words at its end when sampling with forced faults).
size_t rec_stride_of(const clifft_svm_program* p, bool kfault = false) {
    return align_up(std::max<size_t>(p->total_slots + p->num_det + p->num_obs, 4), 16) +
           (kfault ? 20u * kMaxForcedFaults : 0u);
}

// The lane interpreter's copy of the bytecode (single frame word programs).
// Operands it reads from the instruction itself go into payload bytes 8..23,
// which these opcodes do not use:
//   APPLY_PAULI: the X and Z masks of the frame correction;
//   DETECTOR, POSTSELECT, OBSERVABLE: the record parity as (word, mask)
//     pairs over the 32-bit words of the measurement record -- bytes 8..9
//     hold kLaneInlineParity, 10..11 word a, 12..15 mask a, 16..17 word b,
//     20..23 mask b.  Targets listed twice cancel, as in the XOR they feed.
//     Lists spanning more than two words keep the CSR tables (marker 0).
// Runs of frame opcodes become kLaneFrameRun and runs of inline-parity
// DETECTORs kLaneDetectorRun (svm_device.h); the instructions they cover
// stay in place.
// CLIFFT_FRAME_PRUNE=0 keeps every frame opcode in the packed copies.
bool frame_prune_enabled() {
    static bool v = [] {
        const char* e = std::getenv("CLIFFT_FRAME_PRUNE");
        return !(e && e[0] == '0');
    }();
    return v;
}

// FRAME_* instructions that are the identity on every shot that starts from
// the zero frame (every sampled, forced and k-fault shot does).  A forward
// pass carries, per axis, whether p_x and p_z may be nonzero before each
// instruction: frame and array Clifford opcodes move and combine those bits
// as their frame updates do, a measurement sets p_x and clears p_z of its
// axis, APPLY_PAULI, noise sites and fused U2/U4 nodes may set the bits
// they name, and any other opcode that writes the frame marks every bit.
// A FRAME_* opcode whose input bits are all known zero changes neither the
// frame nor gamma (its handler acts only through those bits: CNOT reads
// x_c and z_t, CZ x_c and x_t, H x and z, S x, SWAP all four), so it is
// marked here and left out of its packed run.  No reference counterpart:
// the reference dispatches every frame opcode.  One-word frames only; a
// shot that starts from a preset frame runs the unpacked bytecode.
std::vector<uint8_t> identity_frame_ops(const clifft_program_view* v) {
    const uint64_t n = v->num_instructions;
    std::vector<uint8_t> idle(n, 0);
    if (v->words != 1 || !frame_prune_enabled()) return idle;
    uint64_t mx = 0, mz = 0;  // bits that may be nonzero before instruction i
    auto bit = [](uint32_t q) { return q < 64 ? 1ull << q : ~0ull; };
    auto has = [](uint64_t m, uint32_t q) { return q >= 64 || ((m >> q) & 1ull) != 0; };
    auto swap_bits = [&](uint64_t& m, uint32_t a, uint32_t b) {
        const uint64_t ba = bit(a), bb = bit(b);
        const bool xa = has(m, a), xb = has(m, b);
        m &= ~(ba | bb);
        m |= (xb ? ba : 0ull) | (xa ? bb : 0ull);
    };
    for (uint64_t i = 0; i < n; ++i) {
        const clifft_instruction& in = v->bytecode[i];
        const uint32_t op = in.opcode, a = in.axis_1, b = in.axis_2;
        const bool frame = op <= CLIFFT_OP_FRAME_SWAP;
        uint32_t i0, i1;
        std::memcpy(&i0, in.p.raw, 4);
        std::memcpy(&i1, in.p.raw + 4, 4);
        switch (op) {
            case CLIFFT_OP_FRAME_CNOT:
            case CLIFFT_OP_ARRAY_CNOT: {  // x_b ^= x_a, z_a ^= z_b
                const bool xa = has(mx, a), zb = has(mz, b);
                if (frame && !xa && !zb) idle[i] = 1;
                if (xa) mx |= bit(b);
                if (zb) mz |= bit(a);
                break;
            }
            case CLIFFT_OP_FRAME_CZ:
            case CLIFFT_OP_ARRAY_CZ: {  // z_b ^= x_a, z_a ^= x_b, gamma from x_a x_b
                const bool xa = has(mx, a), xb = has(mx, b);
                if (frame && !xa && !xb) idle[i] = 1;
                if (xa) mz |= bit(b);
                if (xb) mz |= bit(a);
                break;
            }
            case CLIFFT_OP_FRAME_H:
            case CLIFFT_OP_ARRAY_H:  // exchanges x_a and z_a, gamma from x_a z_a
                if (has(mx, a) || has(mz, a)) {
                    mx |= bit(a);
                    mz |= bit(a);
                } else if (frame) {
                    idle[i] = 1;
                }
                break;
            case CLIFFT_OP_FRAME_S:
            case CLIFFT_OP_FRAME_S_DAG:
            case CLIFFT_OP_ARRAY_S:
            case CLIFFT_OP_ARRAY_S_DAG:  // z_a ^= x_a, gamma from x_a
                if (has(mx, a))
                    mz |= bit(a);
                else if (frame)
                    idle[i] = 1;
                break;
            case CLIFFT_OP_FRAME_SWAP:
            case CLIFFT_OP_ARRAY_SWAP:
                if (frame && !has(mx, a) && !has(mx, b) && !has(mz, a) && !has(mz, b)) idle[i] = 1;
                swap_bits(mx, a, b);
                swap_bits(mz, a, b);
                break;
            case CLIFFT_OP_ARRAY_MULTI_CNOT: {  // target a, controls c: x_a ^= x_c, z_c ^= z_a
                const uint64_t m = in.p.multi_gate.mask;
                if (mx & m) mx |= bit(a);
                if (has(mz, a)) mz |= m;
                break;
            }
            case CLIFFT_OP_ARRAY_MULTI_CZ: {  // z_t ^= x_a, z_a ^= x_t
                const uint64_t m = in.p.multi_gate.mask;
                if (has(mx, a)) mz |= m;
                if (mx & m) mz |= bit(a);
                break;
            }
            case CLIFFT_OP_ARRAY_T:
            case CLIFFT_OP_ARRAY_T_DAG:
            case CLIFFT_OP_ARRAY_ROT:
            case CLIFFT_OP_EXPAND:
            case CLIFFT_OP_EXPAND_T:
            case CLIFFT_OP_EXPAND_T_DAG:
            case CLIFFT_OP_EXPAND_ROT:
            case CLIFFT_OP_MEAS_DORMANT_STATIC:
            case CLIFFT_OP_MEAS_DORMANT_STATIC_FORCED:
            case CLIFFT_OP_READOUT_NOISE:
            case CLIFFT_OP_DETECTOR:
            case CLIFFT_OP_POSTSELECT:
            case CLIFFT_OP_OBSERVABLE:
            case CLIFFT_OP_EXP_VAL:
                break;  // read the frame at most
            case CLIFFT_OP_MEAS_DORMANT_RANDOM:
            case CLIFFT_OP_MEAS_DORMANT_RANDOM_FORCED:
            case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL:
            case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL_FORCED:
            case CLIFFT_OP_MEAS_ACTIVE_INTERFERE:
            case CLIFFT_OP_MEAS_ACTIVE_INTERFERE_FORCED:  // x_a = outcome, z_a = 0
                mx |= bit(a);
                mz &= ~bit(a);
                break;
            case CLIFFT_OP_SWAP_MEAS_INTERFERE:
            case CLIFFT_OP_SWAP_MEAS_INTERFERE_FORCED:  // frame swap (a, b), then x_b = outcome, z_b = 0
                if (a != b) {
                    swap_bits(mx, a, b);
                    swap_bits(mz, a, b);
                }
                mx |= bit(b);
                mz &= ~bit(b);
                break;
            case CLIFFT_OP_ARRAY_U2:
                mx |= bit(a);
                mz |= bit(a);
                break;
            case CLIFFT_OP_ARRAY_U4:
                mx |= bit(a) | bit(b);
                mz |= bit(a) | bit(b);
                break;
            case CLIFFT_OP_APPLY_PAULI:
                if (i0 < v->num_pauli_masks) {
                    mx |= v->pauli_x[i0];
                    mz |= v->pauli_z[i0];
                } else {
                    mx = mz = ~0ull;
                }
                break;
            case CLIFFT_OP_NOISE:
            case CLIFFT_OP_NOISE_BLOCK: {
                const uint64_t s1 = op == CLIFFT_OP_NOISE ? (uint64_t)i0 + 1 : (uint64_t)i0 + i1;
                if (s1 > v->num_sites) {
                    mx = mz = ~0ull;
                    break;
                }
                for (uint64_t s = i0; s < s1; ++s)
                    for (uint32_t c = v->site_chan_off[s]; c < v->site_chan_off[s + 1]; ++c) {
                        mx |= v->chan_x[c];
                        mz |= v->chan_z[c];
                    }
                break;
            }
            default:
                mx = mz = ~0ull;
                break;
        }
    }
    return idle;
}

// Runs of frame opcodes on axes below 64 become superinstructions that
// cover the run and skip its identity ops (`idle`, from identity_frame_ops):
// * a fan (kLaneFrameFan) for two or more consecutive non-identity CZs that
//   share one qubit p, or CNOTs that share their control p or their target
//   p, held as p and the mask m of their other qubits (an op listed twice
//   cancels, as the ops do);
// * a packed run (kLaneFrameRun) for the other ops, at most
//   kLaneFrameRunMax of them as 16-bit fields (opcode, axis 1, axis 2), the
//   count in axis_1.
// Each names the instructions it covers beyond itself in axis_2; together
// they tile the run.  A single non-identity op covering only itself stays
// as it is.  Used by the lane copy and by the team interpreter's copy of
// one-word-frame programs.
void pack_frame_runs(clifft_instruction* code, uint64_t n, const std::vector<uint8_t>& idle) {
    auto frame_op = [&](uint64_t j) {
        return code[j].opcode <= CLIFFT_OP_FRAME_SWAP && code[j].axis_1 < 64 && code[j].axis_2 < 64;
    };
    // live ops of the fan of `kind` around `pivot` from live[u] on
    auto fan_len = [&](const std::vector<uint64_t>& live, size_t u, uint32_t kind, uint32_t pivot) {
        size_t c = 0;
        for (size_t v = u; v < live.size(); ++v, ++c) {
            const clifft_instruction& in = code[live[v]];
            const uint32_t a = in.axis_1, b = in.axis_2;
            if (a == b) break;
            if (kind == kFanCz) {
                if (in.opcode != CLIFFT_OP_FRAME_CZ || (a != pivot && b != pivot)) break;
            } else if (in.opcode != CLIFFT_OP_FRAME_CNOT || (kind == kFanControl ? a : b) != pivot) {
                break;
            }
        }
        return c;
    };
    auto emit_run = [&](uint64_t s, uint64_t e, const std::vector<uint64_t>& ops) {
        if (e <= s || (e - s == 1 && ops.size() == 1)) return;
        uint64_t w[3] = {0, 0, 0};
        for (size_t t = 0; t < ops.size(); ++t) {
            const clifft_instruction& in = code[ops[t]];
            const uint64_t f = (uint64_t)in.opcode | ((uint64_t)in.axis_1 << 3) | ((uint64_t)in.axis_2 << 9);
            w[t / 4] |= f << (16 * (t % 4));
        }
        clifft_instruction& h = code[s];
        h.opcode = (uint8_t)kLaneFrameRun;
        h.axis_1 = (uint16_t)ops.size();
        h.axis_2 = (uint16_t)(e - s - 1);
        std::memcpy(h.p.raw, w, 24);
    };
    std::vector<uint64_t> live, ops;
    for (uint64_t i = 0; i < n;) {
        uint64_t j = i;
        while (j < n && j - i < 0xFFFF && frame_op(j)) ++j;
        if (j == i) {
            ++i;
            continue;
        }
        live.clear();
        for (uint64_t t = i; t < j; ++t)
            if (!idle[t]) live.push_back(t);
        uint64_t start = i;  // first instruction of the unit being collected
        ops.clear();
        for (size_t u = 0; u < live.size();) {
            const clifft_instruction& in = code[live[u]];
            uint32_t kind = 0, pivot = 0;
            size_t len = 0;
            if (in.opcode == CLIFFT_OP_FRAME_CZ) {
                const size_t la = fan_len(live, u, kFanCz, in.axis_1), lb = fan_len(live, u, kFanCz, in.axis_2);
                kind = kFanCz;
                pivot = la >= lb ? in.axis_1 : in.axis_2;
                len = std::max(la, lb);
            } else if (in.opcode == CLIFFT_OP_FRAME_CNOT) {
                const size_t lc = fan_len(live, u, kFanControl, in.axis_1), lt = fan_len(live, u, kFanTarget, in.axis_2);
                kind = lc >= lt ? kFanControl : kFanTarget;
                pivot = lc >= lt ? in.axis_1 : in.axis_2;
                len = std::max(lc, lt);
            }
            if (len >= 2) {
                emit_run(start, live[u], ops);
                ops.clear();
                uint64_t m = 0;
                for (size_t v = u; v < u + len; ++v) {
                    const clifft_instruction& f = code[live[v]];
                    const uint32_t other = kind == kFanCz ? (f.axis_1 == pivot ? f.axis_2 : f.axis_1)
                                                          : (kind == kFanControl ? f.axis_2 : f.axis_1);
                    m ^= 1ull << other;
                }
                const uint64_t s = live[u], e = live[u + len - 1] + 1;
                clifft_instruction& h = code[s];
                h.opcode = (uint8_t)kLaneFrameFan;
                h.reserved = (uint8_t)kind;
                h.flags = 0;
                h.axis_1 = (uint16_t)pivot;
                h.axis_2 = (uint16_t)(e - s - 1);
                std::memset(h.p.raw, 0, 24);
                std::memcpy(h.p.raw, &m, 8);
                start = e;
                u += len;
                continue;
            }
            if (ops.size() == kLaneFrameRunMax) {
                emit_run(start, live[u], ops);
                ops.clear();
                start = live[u];
            }
            ops.push_back(live[u]);
            ++u;
        }
        emit_run(start, j, ops);
        i = j;
    }
}

// CLIFFT_ARRAY_FANOUT=0 keeps consecutive ARRAY_CNOTs one instruction each.
bool array_fanout_enabled() {
    static bool v = [] {
        const char* e = std::getenv("CLIFFT_ARRAY_FANOUT");
        return !(e && e[0] == '0');
    }();
    return v;
}

// Two or more consecutive ARRAY_CNOTs with the same control and targets other
// than it become one kArrayFanout instruction (control, target mask) that
// covers them: one deferred permutation opcode instead of one per CNOT.  An
// ARRAY_MULTI_CZ from the same control right after them (not targeting the
// control), or an ARRAY_CZ on the control, is covered too, also after a
// single CNOT.  Team copy of one-word-frame programs only.
void pack_array_fanouts(clifft_instruction* code, uint64_t n) {
    for (uint64_t i = 0; i < n;) {
        const clifft_instruction& h0 = code[i];
        if (h0.opcode != CLIFFT_OP_ARRAY_CNOT || h0.axis_1 == h0.axis_2 || h0.axis_1 >= 32 || h0.axis_2 >= 32) {
            ++i;
            continue;
        }
        const uint32_t c = h0.axis_1;
        uint64_t j = i;
        uint32_t m = 0;
        while (j < n && j - i < 0xFFFF && code[j].opcode == CLIFFT_OP_ARRAY_CNOT && code[j].axis_1 == c &&
               code[j].axis_2 != c && code[j].axis_2 < 32) {
            m ^= 1u << code[j].axis_2;
            ++j;
        }
        uint64_t cz = 0;
        bool with_cz = false;
        if (j < n && j - i < 0xFFFF) {
            const clifft_instruction& z = code[j];
            if (z.opcode == CLIFFT_OP_ARRAY_MULTI_CZ && z.axis_1 == c && ((z.p.multi_gate.mask >> c) & 1ull) == 0) {
                with_cz = true;
                cz = z.p.multi_gate.mask;
            } else if (z.opcode == CLIFFT_OP_ARRAY_CZ && z.axis_1 != z.axis_2 && (z.axis_1 == c || z.axis_2 == c) &&
                       z.axis_1 < 32 && z.axis_2 < 32) {
                with_cz = true;  // a CZ is a MULTI_CZ with one target
                cz = 1ull << (z.axis_1 == c ? z.axis_2 : z.axis_1);
            }
        }
        if (j - i >= 2 || with_cz) {
            if (with_cz) ++j;
            clifft_instruction& h = code[i];
            h.opcode = (uint8_t)kArrayFanout;
            h.reserved = with_cz ? 1u : 0u;
            h.axis_2 = (uint16_t)(j - i - 1);
            std::memset(h.p.raw, 0, 24);
            const uint64_t m64 = m;
            std::memcpy(h.p.raw, &m64, 8);
            std::memcpy(h.p.raw + 8, &cz, 8);
        }
        i = std::max(j, i + 1);
    }
}

void lane_inline_operands(const clifft_program_view* v, clifft_instruction* code, const std::vector<uint8_t>& idle) {
    const uint64_t n = v->num_instructions;
    pack_frame_runs(code, n, idle);  // first, so the runs never hold the opcodes below
    for (uint64_t i = 0; i < v->num_instructions; ++i) {
        clifft_instruction& in = code[i];
        uint64_t w2 = 0, w3 = 0;
        if (in.opcode == CLIFFT_OP_APPLY_PAULI) {
            w2 = v->pauli_x[in.p.pauli.cp_mask_idx];
            w3 = v->pauli_z[in.p.pauli.cp_mask_idx];
        } else if (in.opcode == CLIFFT_OP_DETECTOR || in.opcode == CLIFFT_OP_POSTSELECT ||
                   in.opcode == CLIFFT_OP_OBSERVABLE) {
            const bool obs = in.opcode == CLIFFT_OP_OBSERVABLE;
            const uint32_t* off = obs ? v->obs_off : v->det_off;
            const uint32_t* tgt = obs ? v->obs_tgt : v->det_tgt;
            const uint32_t l = in.p.classical.classical_idx;
            std::vector<std::pair<uint32_t, uint32_t>> words;  // (word, mask)
            for (uint32_t j = off[l]; j < off[l + 1]; ++j) {
                const uint32_t w = tgt[j] >> 5, bit = 1u << (tgt[j] & 31u);
                auto it = std::find_if(words.begin(), words.end(), [&](const auto& e) { return e.first == w; });
                if (it == words.end())
                    words.push_back({w, bit});
                else
                    it->second ^= bit;
            }
            words.erase(std::remove_if(words.begin(), words.end(), [](const auto& e) { return e.second == 0; }),
                        words.end());
            if (words.size() > 2 ||
                std::any_of(words.begin(), words.end(), [](const auto& e) { return e.first > 0xFFFF; }))
                continue;  // marker stays 0: CSR tables
            uint32_t wa = 0, ma = 0, wb = 0, mb = 0;
            if (words.size() > 0) wa = words[0].first, ma = words[0].second;
            if (words.size() > 1) wb = words[1].first, mb = words[1].second;
            w2 = (uint64_t)kLaneInlineParity | ((uint64_t)wa << 16) | ((uint64_t)ma << 32);
            w3 = (uint64_t)wb | ((uint64_t)mb << 32);
        } else {
            continue;
        }
        std::memcpy(in.p.raw + 8, &w2, 8);
        std::memcpy(in.p.raw + 16, &w3, 8);
    }
    // runs of DETECTORs with inline parities: one dispatch for the run
    auto inline_det = [&](uint64_t i) {
        uint16_t marker;
        std::memcpy(&marker, code[i].p.raw + 8, 2);
        return code[i].opcode == CLIFFT_OP_DETECTOR && marker == kLaneInlineParity;
    };
    for (uint64_t i = 0; i < n;) {
        uint64_t j = i;
        while (j < n && j - i < 0xFFFF && inline_det(j)) ++j;
        if (j - i >= 2) {
            code[i].opcode = (uint8_t)kLaneDetectorRun;
            code[i].axis_2 = (uint16_t)(j - i - 1);
        }
        i = std::max(j, i + 1);
    }
}

// A dormant measurement immediately followed by the APPLY_PAULI that its
// record conditions (the reset after a measurement) becomes one lane
// instruction (kLaneMeasPauli); the APPLY_PAULI's masks are already inline.
void lane_fuse_meas_pauli(const clifft_program_view* v, clifft_instruction* code) {
    for (uint64_t i = 0; i + 1 < v->num_instructions; ++i) {
        clifft_instruction& m = code[i];
        const clifft_instruction& a = code[i + 1];
        if ((m.opcode != CLIFFT_OP_MEAS_DORMANT_STATIC && m.opcode != CLIFFT_OP_MEAS_DORMANT_RANDOM) ||
            a.opcode != CLIFFT_OP_APPLY_PAULI || a.p.pauli.condition_idx != m.p.classical.classical_idx)
            continue;
        m.reserved = m.opcode;
        m.opcode = (uint8_t)kLaneMeasPauli;
        m.axis_2 = 1;
        std::memcpy(m.p.raw + 8, a.p.raw + 8, 16);
        ++i;
    }
}

// Highest peak rank run one shot per lane (CLIFFT_LANE_MAX_RANK overrides;
// a negative value turns the lane tier off).  Measured on the 20-qubit,
// peak-6 Clifford+T workload: lanes 1.5e8 shots/s, wavefront teams 1.3e7.
int lane_max_rank() {
    static int v = [] {
        const char* e = std::getenv("CLIFFT_LANE_MAX_RANK");
        return e ? std::atoi(e) : 6;
    }();
    return v;
}

size_t global_workset_bytes() {
    static size_t v = [] {
        const char* e = std::getenv("CLIFFT_GLOBAL_WORKSET_MIB");
        return (size_t)(e ? std::atoll(e) : 0) << 20;
    }();
    return v;
}

// Fused permutation/phase runs in the global tier (CLIFFT_FUSE_MONO=0 turns
// them off).
// Candidate folds: a third slice for programs whose fused runs precede
// active measurements (CLIFFT_FUSE_MEAS_FOLD=0 keeps two slices: the run's
// pass, then the measurement's fold in place).
bool fuse_meas_fold_enabled() {
    static bool v = [] {
        const char* e = std::getenv("CLIFFT_FUSE_MEAS_FOLD");
        return !(e && e[0] == '0');
    }();
    return v;
}

// Pending expansions read through by the next active measurement's pass
// (CLIFFT_DEFER_EXPAND=0: every expansion run is written by its own pass).
bool defer_expand_enabled() {
    static bool v = [] {
        const char* e = std::getenv("CLIFFT_DEFER_EXPAND");
        return !(e && e[0] == '0');
    }();
    return v;
}

// Whether a program gains from a second global slice: a fusable run, or an
// expansion read through by a measurement pass.
bool has_fusable(const clifft_svm_program* p) {
    return p->has_mono_runs || (p->has_exp_meas && defer_expand_enabled());
}

// Global-amplitude slices per team for a program with fusable runs.
uint32_t mono_slices(const clifft_svm_program* p) {
    const bool meas = p->has_mono_meas || (p->has_exp_meas && defer_expand_enabled());
    return meas && p->peak_rank <= 23 && fuse_meas_fold_enabled() ? 3u : 2u;
}

// Packed frame runs in the team interpreter's bytecode (CLIFFT_TEAM_FRAME_RUNS=0
// keeps one dispatch per frame opcode).
bool team_frame_runs_enabled() {
    static bool v = [] {
        const char* e = std::getenv("CLIFFT_TEAM_FRAME_RUNS");
        return !(e && e[0] == '0');
    }();
    return v;
}

bool fuse_mono_enabled() {
    static bool v = [] {
        const char* e = std::getenv("CLIFFT_FUSE_MONO");
        return !(e && std::atoi(e) == 0);
    }();
    return v;
}

// Highest peak rank whose amplitudes stay in LDS in a workgroup team
// (CLIFFT_LDS_MAX_RANK, 7..13; above it the global tier).  Measured on one
// MI355X, per shot of random Clifford+T with 20000 shots: at peak 8 and 10
// the LDS teams are faster (0.12 vs 0.28 us, 0.28 vs 0.35 us); at peak 11,
// 12 and 13 the global tier with fused runs and small teams is 1.6x, 1.6x
// and 3.1x faster (0.39 vs 0.64 us, 0.44 vs 0.70 us, 0.53 vs 1.66 us).
int lds_max_rank() {
    static int v = [] {
        const char* e = std::getenv("CLIFFT_LDS_MAX_RANK");
        const int r = e ? std::atoi(e) : 10;
        return std::min(std::max(r, 7), 13);
    }();
    return v;
}

// Threads of a workgroup team on global amplitudes: CLIFFT_GLOBAL_TEAM_THREADS
// (256, 512 or 1024) fixes it; 0 (unset) lets finalize_shape pick it from
// the shot count.
uint32_t global_team_threads() {
    static uint32_t v = [] {
        const char* e = std::getenv("CLIFFT_GLOBAL_TEAM_THREADS");
        const int t = e ? std::atoi(e) : 0;
        return (uint32_t)(t == 256 || t == 512 || t == 1024 ? t : 0);
    }();
    return v;
}

// Highest peak rank whose global-amplitude shots may run one per wavefront
// (four wavefront teams per 256-thread workgroup) instead of one per
// workgroup team, when a launch has the shots to fill every wavefront slot;
// CLIFFT_GWAVE_MAX_RANK overrides (0: workgroup teams only).
int gwave_max_rank() {
    static int v = [] {
        const char* e = std::getenv("CLIFFT_GWAVE_MAX_RANK");
        return e ? std::atoi(e) : 16;
    }();
    return v;
}

// LDS of the deferred-run opcode lists of a global-amplitude launch: one
// per team (a wavefront team or a workgroup).
size_t mono_list_bytes(const LaunchShape& L) {
    return (size_t)(L.block_team ? 1u : L.teams_per_block()) * kMonoBytes;
}

// Choose the team shape and storage tier for a program's peak rank.
LaunchShape plan_shape(const clifft_svm_program* p, bool allow_lane = true, bool kfault = false) {
    LaunchShape L;
    L.nw = (int)p->nw;
    const uint32_t k = p->peak_rank;
    if (allow_lane && p->words == 1 && (int)k <= lane_max_rank()) {
        // up to 4 wavefronts per workgroup, as many as the LDS holds
        const size_t wb = lane_wave_bytes(k, p->total_slots, p->num_det, p->num_obs);
        const size_t waves = std::min<size_t>(4, (160 * 1024) / wb);
        if (waves > 0) {
            L.lane = true;
            L.block = (uint32_t)(64 * waves);
            L.lds_bytes = waves * wb;
            // Measured on one MI355X: the 6-workgroup build is 11% faster on
            // the 17-qubit surface-code memory (peak rank 1); where LDS caps
            // residency anyway, fewer registers only add spills (an
            // 8-workgroup build was 9% slower on 20-qubit Clifford+T, peak 6).
            L.lane_dense = 6 * L.lds_bytes <= 160 * 1024;
            L.lane_narrow = p->frame32;
            return L;
        }
    }
    const size_t rec_stride = rec_stride_of(p, kfault);
    const size_t amp_bytes = (size_t)16 << k;
    const size_t red = kTeamScratchBytes;
    if (k <= 7) {
        L.block_team = false;
        L.block = 256;
        L.global_amps = false;
    } else if ((int)k <= lds_max_rank()) {
        L.block_team = true;
        uint32_t waves = k <= 9 ? 2 : (k <= 10 ? 4 : (k <= 12 ? 8 : 16));
        L.block = 64 * waves;
        L.global_amps = false;
    } else {
        L.block_team = true;
        L.block = global_team_threads() ? global_team_threads() : 1024;
        L.global_amps = true;
    }
    L.rec_stride = (uint32_t)rec_stride;
    size_t lds = red + (L.block / 64) * rec_stride;
    if (!L.global_amps) lds += L.teams_per_block() * amp_bytes;
    if (lds > kLdsBytes) {  // records too large next to LDS amplitudes
        L.global_amps = true;
        L.block_team = true;
        L.block = global_team_threads() ? global_team_threads() : 1024;
        lds = red + (L.block / 64) * rec_stride;
    }
    // every wavefront keeps its own record copy: fewer wavefronts per team
    // when the copies do not fit (a shape with lds_bytes > kLdsBytes is
    // refused at program creation)
    while (L.global_amps && lds > kLdsBytes && L.block > 256) {
        L.block /= 2;
        lds = red + (L.block / 64) * rec_stride;
    }
    // Workgroup teams on global amplitudes get a second slice and an LDS
    // list for fused permutation/phase runs (one read and one write of the
    // state per run instead of one pass per opcode).
    // Only programs with a run of two or more, and states of at most 2^24
    // amplitudes (a second 512 MiB slice per team at most).
    // A run that precedes an active measurement writes both folded states
    // of the measurement in its pass, into a second and a third slice (states
    // of at most 2^23 amplitudes).
    const size_t mono = mono_list_bytes(L);
    if (L.global_amps && fuse_mono_enabled() && has_fusable(p) && k <= 24 && lds + mono <= kLdsBytes) {
        L.amp_slices = mono_slices(p);
        lds += mono;
    }
    L.lds_bytes = lds;
    return L;
}

int ensure_scratch(hipStream_t st, size_t bytes, void** ptr) {
    std::lock_guard<std::mutex> lock(g_arena_mu);
    Arena& a = arena_of(st);
    if (bytes > a.bytes) {
        if (a.ptr) {
            HIP_TRY(hipStreamSynchronize(st));
            (void)hipFree(a.ptr);
        }
        a.ptr = nullptr;
        a.bytes = 0;
        HIP_TRY(hipMalloc(&a.ptr, bytes));
        a.bytes = bytes;
    }
    *ptr = a.ptr;
    return CLIFFT_OK;
}

size_t arena_bytes(hipStream_t st) {
    std::lock_guard<std::mutex> lock(g_arena_mu);
    return arena_of(st).bytes;
}

int ensure_out(clifft_svm_program* p, size_t bytes) {
    if (bytes <= p->out_dev_bytes) return CLIFFT_OK;
    if (p->out_dev) (void)hipFree(p->out_dev);
    p->out_dev = nullptr;
    p->out_dev_bytes = 0;
    HIP_TRY(hipMalloc(&p->out_dev, bytes));
    p->out_dev_bytes = bytes;
    return CLIFFT_OK;
}

// Size the grid: persistent teams up to the resident capacity of the chip.
// Workgroup teams on global amplitudes: the smallest team (256, 512, 1024
// threads) whose resident teams the launch still fills to 3/4.  Smaller
// teams keep more states in flight; measured on one MI355X (random
// Clifford+T, per-shot time): peak 14 and 15 at 256 threads 3.0x and 2.3x
// faster than at 1024, peak 16 1.5x, peak 18 best at 512 with 500 shots,
// peak 20 and 22 (125 and 64 shots) best at 1024.
int choose_global_team(LaunchShape& L, uint64_t shots, uint32_t peak_rank) {
    if (!L.global_amps || !L.block_team || global_team_threads()) return CLIFFT_OK;
    DeviceInfo& di = device_info();
    // One shot per wavefront once the shots fill every resident wavefront
    // team: a wavefront walks the bytecode once per shot where a workgroup
    // team's wavefronts each walk it.  Measured on one MI355X, 24-qubit
    // Clifford+T: peak 13 x 20000 shots 9.77 -> 3.92 ms; peak 16 x 2000
    // shots (half the slots) 2.87 -> 3.95 ms, so fewer shots keep workgroups.
    if ((int)peak_rank <= gwave_max_rank()) {
        LaunchShape c = L;
        c.block_team = false;
        c.block = 256;
        c.lds_bytes = kTeamScratchBytes + 4 * (size_t)L.rec_stride + (L.amp_slices >= 2 ? mono_list_bytes(c) : 0);
        if (c.lds_bytes <= kLdsBytes) {
            int per_cu = 0;
            hipError_t e = shot_occupancy(c, &per_cu);
            if (e != hipSuccess) return fail(CLIFFT_E_HIP, std::string("occupancy query: ") + hipGetErrorString(e));
            if (per_cu > 0 && shots >= (uint64_t)di.num_cus * (uint64_t)per_cu * c.teams_per_block()) {
                L = c;
                return CLIFFT_OK;
            }
        }
    }
    for (uint32_t t : {256u, 512u, 1024u}) {
        LaunchShape c = L;
        c.block = t;
        c.lds_bytes = kTeamScratchBytes + (t / 64) * (size_t)L.rec_stride + (L.amp_slices >= 2 ? kMonoBytes : 0);
        if (c.lds_bytes > kLdsBytes) break;  // larger teams hold more record copies
        int per_cu = 0;
        hipError_t e = shot_occupancy(c, &per_cu);
        if (e != hipSuccess) return fail(CLIFFT_E_HIP, std::string("occupancy query: ") + hipGetErrorString(e));
        const uint64_t slots = (uint64_t)di.num_cus * (uint64_t)std::max(per_cu, 1);
        if (t == 1024u || 4 * shots >= 3 * slots) {
            L = c;
            return CLIFFT_OK;
        }
    }
    return CLIFFT_OK;
}

int finalize_shape(clifft_svm_program* p, LaunchShape& L, uint64_t shots, hipStream_t st, void** gscratch) {
    DeviceInfo& di = device_info();
    int trc = choose_global_team(L, shots, p->peak_rank);
    if (trc) return trc;
    const uint32_t tpb = L.teams_per_block();
    uint64_t blocks_needed = (shots + tpb - 1) / tpb;
    int per_cu_i = 0;
    hipError_t oe = shot_occupancy(L, &per_cu_i);
    if (oe != hipSuccess) return fail(CLIFFT_E_HIP, std::string("occupancy query: ") + hipGetErrorString(oe));
    uint32_t per_cu = (uint32_t)std::max(per_cu_i, 1);
    uint64_t cap = (uint64_t)di.num_cus * per_cu;
    if (L.global_amps) {
        // bound the global scratch by what the device has free (keep 40% or
        // 16 GiB headroom, whichever is larger), at most 192 GiB
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 8ull << 30;
        size_t budget = free_b + arena_bytes(st);
        size_t headroom = std::max<size_t>(budget * 2 / 5, 16ull << 30);
        budget = budget > headroom ? budget - headroom : (size_t)1 << 30;
        budget = std::min<size_t>(budget, 192ull << 30);
        size_t per_team = ((size_t)16 << p->peak_rank) * L.amp_slices;
        uint64_t max_teams = std::max<uint64_t>(1, budget / per_team);
        cap = std::min<uint64_t>(cap, std::max<uint64_t>(1, max_teams / tpb));
        // Optionally keep the states in flight within a cache budget
        // (CLIFFT_GLOBAL_WORKSET_MIB; default 0 = no cap), never below one
        // team per CU.  Measured on one MI355X, a 256 MiB cap made no
        // difference with 1024-thread teams and cost 2.2x at peak 16 with
        // 256-thread teams, so it is off.
        const size_t ws = global_workset_bytes();
        if (ws > 0) {
            const uint64_t by_ws = std::max<uint64_t>((uint64_t)di.num_cus, ws / per_team);
            cap = std::min<uint64_t>(cap, std::max<uint64_t>(1, by_ws / tpb));
        }
    }
    L.grid = (uint32_t)std::max<uint64_t>(1, std::min(blocks_needed, cap));
    *gscratch = nullptr;
    if (L.global_amps) {
        size_t need = (size_t)L.grid * tpb * ((size_t)16 << p->peak_rank) * L.amp_slices;
        int rc = ensure_scratch(st, need, gscratch);
        if (rc) return rc;
    }
    return CLIFFT_OK;
}

void fill_common_args(clifft_svm_program* p, ShotArgs& A, void* gscratch, bool kfault = false) {
    std::memset(&A, 0, sizeof(A));
    A.rec_stride = (uint32_t)rec_stride_of(p, kfault);
    A.amp_elems = 1u << p->peak_rank;
    A.gscratch = reinterpret_cast<dv2*>(gscratch);
    A.dust_counter = p->dust_dev;
    A.amp_slices = 1;
    A.pend_defer = defer_expand_enabled() ? 1u : 0u;
}

size_t row_bytes(uint32_t n, bool packed) { return packed ? (n + 7) / 8 : n; }

// Copy the launch arguments into the next device slot and launch.  The
// host copy is staged by the runtime before hipMemcpyAsync returns, so the
// caller's ShotArgs may be reused immediately.
hipError_t launch_with_args(clifft_svm_program* p, const LaunchShape& L, const DevProgram* dp, const ShotArgs& A,
                            hipStream_t st) {
    ShotArgs* slot = p->args_dev + (p->args_slot++ % kArgSlots);
    hipError_t e = hipMemcpyAsync(slot, &A, sizeof(ShotArgs), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    return launch_shots(L, dp, slot, st);
}

}  // namespace

extern "C" {

uint32_t clifft_svm_abi_version(void) { return CLIFFT_SVM_ABI_VERSION; }

const char* clifft_svm_backend(void) {
    DeviceInfo& di = device_info();
    if (di.ok == 1) return "hip:gfx950";
    if (di.ok == -1) return "hip:unsupported-device";
    return "hip:no-device";
}

const char* clifft_svm_last_error(void) { return g_last_error.c_str(); }

int clifft_svm_program_create(const clifft_program_view* view, clifft_svm_program** out) {
    if (!out) return fail(CLIFFT_E_INVALID, "null output handle");
    *out = nullptr;
    std::vector<int32_t> khist;
    bool has_forced = false;
    int rc = validate(view, khist, has_forced);
    if (rc) return rc;
    DeviceInfo& di = device_info();
    if (di.ok == 0) return fail(CLIFFT_E_NODEVICE, "no HIP device visible");
    if (di.ok < 0) return fail(CLIFFT_E_NODEVICE, "device is " + di.arch + ", this build targets gfx950");

    auto* p = new clifft_svm_program();
    p->num_qubits = view->num_qubits;
    p->peak_rank = view->peak_rank;
    p->words = view->words;
    p->nw = frame_nw(view->words);
    p->num_meas = view->num_measurements;
    p->total_slots = view->total_meas_slots;
    p->num_det = view->num_detectors;
    p->num_obs = view->num_observables;
    p->num_ev = view->num_exp_vals;
    p->has_postselection = view->has_postselection != 0;
    p->has_forced_ops = has_forced;
    p->host_bc.assign(view->bytecode, view->bytecode + view->num_instructions);
    p->khist = khist;
    p->num_sites = view->num_sites;
    for (uint32_t s = 0; s < view->num_sites; ++s) {  // svm.cc:356-369
        double acc = 0.0;
        for (uint32_t c = view->site_chan_off[s]; c < view->site_chan_off[s + 1]; ++c) acc += view->chan_prob[c];
        p->fault_probs.push_back(acc);
    }
    for (uint32_t e = 0; e < view->num_readout; ++e) p->fault_probs.push_back(view->ro_prob[e]);
    if (view->expected_observables)
        p->expected_obs.assign(view->expected_observables, view->expected_observables + view->num_expected_observables);
    // 32-bit frame words suffice when no operand axis and no Pauli/noise
    // mask reaches bit 32: the frame's upper bits then stay zero.
    {
        bool fits = view->words == 1;
        for (uint64_t i = 0; i < view->num_instructions && fits; ++i)
            fits = view->bytecode[i].axis_1 < 32 && view->bytecode[i].axis_2 < 32;
        uint64_t any = 0;
        for (uint32_t i = 0; fits && i < view->num_pauli_masks; ++i) any |= view->pauli_x[i] | view->pauli_z[i];
        for (uint32_t i = 0; fits && i < view->num_channels; ++i) any |= view->chan_x[i] | view->chan_z[i];
        p->frame32 = fits && (any >> 32) == 0;
    }
    const uint32_t fbits = view->words * 64;
    uint32_t run = 0;
    bool expand_pending = false;
    for (uint64_t i = 0; i < view->num_instructions; ++i) {
        const clifft_instruction& in = p->host_bc[i];
        double b = op_bytes(in, khist[i], fbits);
        if (b > 0) {
            p->num_array_ops++;
            p->bytes_per_shot += b;
        }
        // the interpreter's deferral rule: T/ROT on a dormant axis sweep nothing
        const uint64_t bit = in.opcode < 64 ? 1ull << in.opcode : 0ull;
        const bool dormant_phase = (in.opcode == CLIFFT_OP_ARRAY_T || in.opcode == CLIFFT_OP_ARRAY_T_DAG ||
                                    in.opcode == CLIFFT_OP_ARRAY_ROT) && in.axis_1 >= (uint32_t)khist[i];
        if ((kMonoOps & bit) && !dormant_phase) {
            if (++run >= 2) p->has_mono_runs = true;
        } else if (!(kMonoKeepOps & bit)) {
            if (run >= 2 && (kMonoMeasOps & bit)) p->has_mono_meas = true;
            run = 0;
        }
        // the interpreter's rule for pending expansions (with a second slice)
        if (in.opcode == CLIFFT_OP_EXPAND || in.opcode == CLIFFT_OP_EXPAND_T || in.opcode == CLIFFT_OP_EXPAND_T_DAG ||
            in.opcode == CLIFFT_OP_EXPAND_ROT) {
            expand_pending = true;
        } else if (kMonoMeasOps & bit) {
            if (expand_pending) p->has_exp_meas = true;
            expand_pending = false;
        } else if (!((kNoAmpOps | kMonoOps) & bit)) {
            expand_pending = false;
        }
    }

    // Pack every array into one device allocation (256-B aligned pieces).
    struct Piece {
        const void* src;
        size_t bytes;
        size_t off;
    };
    std::vector<Piece> pieces;
    size_t total = 0;
    auto add = [&](const void* src, size_t bytes) -> size_t {
        size_t off = total;
        pieces.push_back({src, bytes, off});
        total = align_up(total + std::max<size_t>(bytes, 8), 256);
        return off;
    };
    const size_t W = view->words;
    size_t o_bc = add(view->bytecode, view->num_instructions * sizeof(clifft_instruction));
    size_t o_fbc = add(nullptr, view->num_instructions * sizeof(clifft_instruction));
    size_t o_lbc = add(nullptr, view->num_instructions * sizeof(clifft_instruction));
    size_t o_lfbc = add(nullptr, view->num_instructions * sizeof(clifft_instruction));
    size_t o_tbc = add(nullptr, view->num_instructions * sizeof(clifft_instruction));
    size_t o_tfbc = add(nullptr, view->num_instructions * sizeof(clifft_instruction));
    size_t o_px = add(view->pauli_x, view->num_pauli_masks * W * 8);
    size_t o_pz = add(view->pauli_z, view->num_pauli_masks * W * 8);
    size_t o_ps = add(view->pauli_sign, view->num_pauli_masks);
    size_t o_ex = add(view->ev_x, view->num_ev_masks * W * 8);
    size_t o_ez = add(view->ev_z, view->num_ev_masks * W * 8);
    size_t o_es = add(view->ev_sign, view->num_ev_masks);
    size_t o_so = add(view->site_chan_off, view->num_sites > 0 ? (view->num_sites + 1) * 4 : 0);
    size_t o_cx = add(view->chan_x, view->num_channels * W * 8);
    size_t o_cz = add(view->chan_z, view->num_channels * W * 8);
    size_t o_cp = add(view->chan_prob, view->num_channels * 8);
    size_t o_hz = add(view->hazards, view->num_sites * 8);
    // Running channel sums per site, in the reference's summation order
    // (svm_kernels.inl:1894-1909), and the hazard guide: G buckets of width
    // H / G over [0, H), H the total hazard; entry b is the upper bound of
    // the bucket's lower edge, entry G is the site count.
    std::vector<double> chan_cum(view->num_channels);
    for (uint32_t s = 0; s < view->num_sites; ++s) {
        double acc = 0.0;
        for (uint32_t c = view->site_chan_off[s]; c < view->site_chan_off[s + 1]; ++c) {
            acc += view->chan_prob[c];
            chan_cum[c] = acc;
        }
    }
    const uint32_t n_sites = view->num_sites;
    const double h_total = n_sites ? view->hazards[n_sites - 1] : 0.0;
    const uint32_t hz_g = h_total > 0.0 ? std::min<uint32_t>(n_sites, 1u << 16) : 0u;
    const double hz_scale = hz_g ? (double)hz_g / h_total : 0.0;
    std::vector<uint32_t> hz_guide(hz_g + 1, n_sites);
    for (uint32_t b = 0; b < hz_g; ++b)
        hz_guide[b] = (uint32_t)(std::upper_bound(view->hazards, view->hazards + n_sites, (double)b / hz_scale) -
                                 view->hazards);
    size_t o_cc = add(chan_cum.data(), chan_cum.size() * 8);
    size_t o_hg = add(hz_guide.data(), hz_guide.size() * 4);
    size_t o_rm = add(view->ro_meas_idx, view->num_readout * 4);
    size_t o_rp = add(view->ro_prob, view->num_readout * 8);
    size_t o_do = add(view->det_off, view->num_det_lists > 0 ? (view->num_det_lists + 1) * 4 : 0);
    size_t o_dt = add(view->det_tgt, view->num_det_lists > 0 ? view->det_off[view->num_det_lists] * 4 : 0);
    size_t o_oo = add(view->obs_off, view->num_obs_lists > 0 ? (view->num_obs_lists + 1) * 4 : 0);
    size_t o_ot = add(view->obs_tgt, view->num_obs_lists > 0 ? view->obs_off[view->num_obs_lists] * 4 : 0);
    size_t o_u2 = add(view->u2_nodes, view->num_u2 * sizeof(clifft_u2_node));
    size_t o_u4 = add(view->u4_nodes, view->num_u4 * sizeof(clifft_u4_node));
    size_t o_eo = add(view->expected_observables, view->num_expected_observables);
    size_t o_dp = add(nullptr, 2 * sizeof(DevProgram));

    std::vector<uint8_t> staging(total, 0);
    for (const Piece& pc : pieces)
        if (pc.src && pc.bytes) std::memcpy(staging.data() + pc.off, pc.src, pc.bytes);
    // forced copy of the bytecode (api/record_probabilities.cc:96-118)
    {
        auto* fbc = reinterpret_cast<clifft_instruction*>(staging.data() + o_fbc);
        std::memcpy(fbc, view->bytecode, view->num_instructions * sizeof(clifft_instruction));
        for (uint64_t i = 0; i < view->num_instructions; ++i) {
            uint8_t& op = fbc[i].opcode;
            if (op >= CLIFFT_OP_MEAS_DORMANT_STATIC && op <= CLIFFT_OP_SWAP_MEAS_INTERFERE)
                op = (uint8_t)(op + (CLIFFT_OP_MEAS_DORMANT_STATIC_FORCED - CLIFFT_OP_MEAS_DORMANT_STATIC));
        }
        // the lane interpreter's copies (sampling and forced), with operands
        // inlined into the unused payload words
        auto* lbc = reinterpret_cast<clifft_instruction*>(staging.data() + o_lbc);
        std::memcpy(lbc, view->bytecode, view->num_instructions * sizeof(clifft_instruction));
        const std::vector<uint8_t> idle = identity_frame_ops(view);
        if (view->words == 1) lane_inline_operands(view, lbc, idle);
        auto* lfbc = reinterpret_cast<clifft_instruction*>(staging.data() + o_lfbc);
        std::memcpy(lfbc, lbc, view->num_instructions * sizeof(clifft_instruction));
        for (uint64_t i = 0; i < view->num_instructions; ++i) {
            uint8_t& op = lfbc[i].opcode;
            if (op >= CLIFFT_OP_MEAS_DORMANT_STATIC && op <= CLIFFT_OP_SWAP_MEAS_INTERFERE)
                op = (uint8_t)(op + (CLIFFT_OP_MEAS_DORMANT_STATIC_FORCED - CLIFFT_OP_MEAS_DORMANT_STATIC));
        }
        if (view->words == 1) lane_fuse_meas_pauli(view, lbc);  // sampling copy only
        // the team interpreter's copies: frame runs packed (one-word frames)
        auto* tbc = reinterpret_cast<clifft_instruction*>(staging.data() + o_tbc);
        auto* tfbc = reinterpret_cast<clifft_instruction*>(staging.data() + o_tfbc);
        std::memcpy(tbc, view->bytecode, view->num_instructions * sizeof(clifft_instruction));
        std::memcpy(tfbc, fbc, view->num_instructions * sizeof(clifft_instruction));
        if (view->words == 1 && team_frame_runs_enabled()) {
            pack_frame_runs(tbc, view->num_instructions, idle);
            pack_frame_runs(tfbc, view->num_instructions, idle);
            if (array_fanout_enabled()) {
                pack_array_fanouts(tbc, view->num_instructions);
                pack_array_fanouts(tfbc, view->num_instructions);
            }
        }
    }
    hipError_t e = hipMalloc(&p->dev_blob, total);
    if (e == hipSuccess) e = hipMemcpy(p->dev_blob, staging.data(), total, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&p->dust_dev), sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(p->dust_dev, 0, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&p->args_dev), kArgSlots * sizeof(ShotArgs));
    if (e != hipSuccess) {
        std::string msg = std::string("program upload: ") + hipGetErrorString(e);
        clifft_svm_program_destroy(p);
        return fail(CLIFFT_E_HIP, msg);
    }
    auto* base = static_cast<uint8_t*>(p->dev_blob);
    DevProgram& d = p->dp;
    d.bc = reinterpret_cast<const clifft_instruction*>(base + o_bc);
    p->forced_bc = reinterpret_cast<clifft_instruction*>(base + o_fbc);
    d.lane_bc = reinterpret_cast<const clifft_instruction*>(base + o_lbc);
    d.team_bc = reinterpret_cast<const clifft_instruction*>(base + o_tbc);
    d.pauli_x = reinterpret_cast<const uint64_t*>(base + o_px);
    d.pauli_z = reinterpret_cast<const uint64_t*>(base + o_pz);
    d.pauli_sign = base + o_ps;
    d.ev_x = reinterpret_cast<const uint64_t*>(base + o_ex);
    d.ev_z = reinterpret_cast<const uint64_t*>(base + o_ez);
    d.ev_sign = base + o_es;
    d.site_chan_off = reinterpret_cast<const uint32_t*>(base + o_so);
    d.chan_x = reinterpret_cast<const uint64_t*>(base + o_cx);
    d.chan_z = reinterpret_cast<const uint64_t*>(base + o_cz);
    d.chan_prob = reinterpret_cast<const double*>(base + o_cp);
    d.chan_cum = reinterpret_cast<const double*>(base + o_cc);
    d.hazards = reinterpret_cast<const double*>(base + o_hz);
    d.hz_guide = reinterpret_cast<const uint32_t*>(base + o_hg);
    d.hz_buckets = hz_g;
    d.hz_scale = hz_scale;
    d.ro_meas_idx = reinterpret_cast<const uint32_t*>(base + o_rm);
    d.ro_prob = reinterpret_cast<const double*>(base + o_rp);
    d.det_off = reinterpret_cast<const uint32_t*>(base + o_do);
    d.det_tgt = reinterpret_cast<const uint32_t*>(base + o_dt);
    d.obs_off = reinterpret_cast<const uint32_t*>(base + o_oo);
    d.obs_tgt = reinterpret_cast<const uint32_t*>(base + o_ot);
    d.u2 = reinterpret_cast<const clifft_u2_node*>(base + o_u2);
    d.u4 = reinterpret_cast<const clifft_u4_node*>(base + o_u4);
    d.expected_obs = base + o_eo;
    d.n_instr = (uint32_t)view->num_instructions;
    d.words = view->words;
    d.peak_rank = view->peak_rank;
    d.num_meas = view->num_measurements;
    d.total_meas_slots = view->total_meas_slots;
    d.num_det = view->num_detectors;
    d.num_obs = view->num_observables;
    d.num_ev = view->num_exp_vals;
    d.num_sites = view->num_sites;
    d.n_expected_obs = view->num_expected_observables;
    {
        DevProgram both[2] = {d, d};
        both[1].bc = p->forced_bc;
        both[1].lane_bc = reinterpret_cast<const clifft_instruction*>(base + o_lfbc);
        both[1].team_bc = reinterpret_cast<const clifft_instruction*>(base + o_tfbc);
        e = hipMemcpy(base + o_dp, both, sizeof(both), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            std::string msg = std::string("program upload: ") + hipGetErrorString(e);
            clifft_svm_program_destroy(p);
            return fail(CLIFFT_E_HIP, msg);
        }
        p->dp_dev = reinterpret_cast<DevProgram*>(base + o_dp);
        p->dp_forced_dev = p->dp_dev + 1;
    }
    p->base_shape = plan_shape(p);
    if (p->base_shape.lds_bytes > kLdsBytes) {
        const std::string msg = "records of " + std::to_string(p->total_slots + p->num_det + p->num_obs) +
                                " bytes per shot do not fit the LDS of a 256-thread team (one copy per "
                                "wavefront; at most " +
                                std::to_string((kLdsBytes - kTeamScratchBytes - kMonoBytes) / 4) + " bytes)";
        clifft_svm_program_destroy(p);
        return fail(CLIFFT_E_UNSUPPORTED, msg);
    }
    *out = p;
    return CLIFFT_OK;
}

int clifft_svm_program_destroy(clifft_svm_program* p) {
    if (!p) return CLIFFT_OK;
    // launches of this program may still be queued on any stream
    (void)hipDeviceSynchronize();
    if (p->dev_blob) (void)hipFree(p->dev_blob);
    if (p->out_dev) (void)hipFree(p->out_dev);
    if (p->dust_dev) (void)hipFree(p->dust_dev);
    if (p->kf_dev) (void)hipFree(p->kf_dev);
    if (p->args_dev) (void)hipFree(p->args_dev);
    delete p;
    return CLIFFT_OK;
}

int clifft_svm_program_info_get(const clifft_svm_program* p, clifft_svm_program_info* info) {
    if (!p || !info) return fail(CLIFFT_E_INVALID, "null argument");
    info->peak_rank = p->peak_rank;
    info->tier = p->base_shape.lane ? 2u : (p->base_shape.global_amps ? 1u : 0u);
    info->team_waves = p->base_shape.block_team ? p->base_shape.block / 64 : 1;
    info->shots_per_block = p->base_shape.teams_per_block();
    info->num_array_ops = p->num_array_ops;
    info->bytes_per_shot = p->bytes_per_shot;
    return CLIFFT_OK;
}

// Shared implementation of sample / forced runs.  Outputs go to device
// memory (either the caller's, or the program's staging buffer which is then
// copied to the caller's host arrays).
static int run_shots(clifft_svm_program* p, uint64_t shots, uint64_t seed, uint32_t flags, uint32_t mode,
                     const clifft_svm_records* out, const uint8_t* forced_host, double* logp_host,
                     hipStream_t st, const KfTables* kf = nullptr) {
    const bool dev_out = (flags & CLIFFT_REC_DEVICE) != 0;
    const bool packed = (flags & CLIFFT_REC_PACKED) != 0;
    if (mode == MODE_SAMPLE && p->has_forced_ops)
        return fail(CLIFFT_E_INVALID, "program contains *_FORCED opcodes; use clifft_svm_record_log_probs");
    const uint64_t chunk_max = dev_out ? shots : (uint64_t)1 << 20;
    for (uint64_t base = 0; base < shots; base += chunk_max) {
        const uint64_t n = std::min(chunk_max, shots - base);
        const bool kfault = mode == MODE_KFAULT;
        LaunchShape L = kfault ? plan_shape(p, false, true) : p->base_shape;
        if (L.lds_bytes > kLdsBytes)
            return fail(CLIFFT_E_UNSUPPORTED, "records and k-fault scratch do not fit the LDS of a team");
        void* gs = nullptr;
        int rc = finalize_shape(p, L, n, st, &gs);
        if (rc) return rc;
        ShotArgs A;
        fill_common_args(p, A, gs, kfault);
        A.amp_slices = L.global_amps ? L.amp_slices : 1u;
        A.kf = kf;
        A.shot_begin = base;
        A.shot_count = n;
        A.seed = seed;
        A.mode = mode;
        A.flags = flags;
        size_t bm = row_bytes(p->num_meas, packed), bd = row_bytes(p->num_det, packed), bo = row_bytes(p->num_obs, packed);
        if (dev_out) {
            if (out) {
                A.out_meas = out->measurements ? out->measurements + base * bm : nullptr;
                A.out_det = out->detectors ? out->detectors + base * bd : nullptr;
                A.out_obs = out->observables ? out->observables + base * bo : nullptr;
                A.out_ev = out->exp_vals ? out->exp_vals + base * p->num_ev : nullptr;
                A.out_disc = out->discarded ? out->discarded + base : nullptr;
            }
        } else {
            // device staging: [meas | det | obs | disc | ev | logp | forced]
            size_t s_m = align_up(n * bm, 256), s_d = align_up(n * bd, 256), s_o = align_up(n * bo, 256);
            size_t s_x = align_up(n, 256), s_e = align_up(n * p->num_ev * 8, 256), s_l = align_up(n * 8, 256);
            size_t s_f = forced_host ? align_up(n * p->num_meas, 256) : 0;
            rc = ensure_out(p, s_m + s_d + s_o + s_x + s_e + s_l + s_f + 256);
            if (rc) return rc;
            uint8_t* o = static_cast<uint8_t*>(p->out_dev);
            A.out_meas = (out && out->measurements) ? o : nullptr;
            A.out_det = (out && out->detectors) ? o + s_m : nullptr;
            A.out_obs = (out && out->observables) ? o + s_m + s_d : nullptr;
            A.out_disc = (out && out->discarded) ? o + s_m + s_d + s_o : nullptr;
            A.out_ev = (out && out->exp_vals) ? reinterpret_cast<double*>(o + s_m + s_d + s_o + s_x) : nullptr;
            A.out_logp = reinterpret_cast<double*>(o + s_m + s_d + s_o + s_x + s_e);
            if (forced_host) {
                uint8_t* fdev = o + s_m + s_d + s_o + s_x + s_e + s_l;
                HIP_TRY(hipMemcpyAsync(fdev, forced_host + base * p->num_meas, n * p->num_meas, hipMemcpyHostToDevice, st));
                A.forced_records = fdev;
            }
        }
        HIP_TRY(launch_with_args(p, L, mode == MODE_FORCED ? p->dp_forced_dev : p->dp_dev, A, st));
        if (!dev_out) {
            if (out && out->measurements && n * bm)
                HIP_TRY(hipMemcpyAsync(out->measurements + base * bm, A.out_meas, n * bm, hipMemcpyDeviceToHost, st));
            if (out && out->detectors && n * bd)
                HIP_TRY(hipMemcpyAsync(out->detectors + base * bd, A.out_det, n * bd, hipMemcpyDeviceToHost, st));
            if (out && out->observables && n * bo)
                HIP_TRY(hipMemcpyAsync(out->observables + base * bo, A.out_obs, n * bo, hipMemcpyDeviceToHost, st));
            if (out && out->discarded)
                HIP_TRY(hipMemcpyAsync(out->discarded + base, A.out_disc, n, hipMemcpyDeviceToHost, st));
            if (out && out->exp_vals && p->num_ev)
                HIP_TRY(hipMemcpyAsync(out->exp_vals + base * p->num_ev, A.out_ev, n * p->num_ev * 8, hipMemcpyDeviceToHost, st));
            if (logp_host)
                HIP_TRY(hipMemcpyAsync(logp_host + base, A.out_logp, n * 8, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
        }
    }
    return CLIFFT_OK;
}

int clifft_svm_sample(clifft_svm_program* p, uint64_t shots, uint64_t seed, uint32_t flags,
                      const clifft_svm_records* out, void* stream) {
    if (!p) return fail(CLIFFT_E_INVALID, "null program");
    if (shots == 0) return CLIFFT_OK;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : library_stream();
    int rc = run_shots(p, shots, seed, flags, MODE_SAMPLE, out, nullptr, nullptr, st);
    if (rc) return rc;
    if (flags & CLIFFT_REC_DEVICE) HIP_TRY(hipStreamSynchronize(st));
    return CLIFFT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Sampling that allocates nothing
// ---------------------------------------------------------------------------
// The global-memory amplitudes of clifft_svm_sample_noalloc live in this
// array: 2^22 complex128 = 64 MiB, one copy per device, held until the
// process ends.  In host code its name is only a stand-in; the device
// address comes from hipGetSymbolAddress with the device current, once per
// device, and only that address is passed to kernels.
static constexpr size_t kNoAllocScratchElems = (size_t)1 << 22;
static __device__ __align__(256) dv2 scratch[kNoAllocScratchElems];

namespace {

std::mutex g_noalloc_mu;
void* g_noalloc_ptr[64] = {};

int noalloc_scratch(void** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(CLIFFT_E_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(CLIFFT_E_HIP, "hipGetSymbolAddress returned a null scratch address");
        g_noalloc_ptr[dev] = ptr;
    }
    *out = g_noalloc_ptr[dev];
    return CLIFFT_OK;
}

}  // namespace

extern "C" {

int clifft_svm_sample_noalloc(clifft_svm_program* p, uint64_t shots, uint64_t seed, uint32_t flags,
                              const clifft_svm_records* out, void* stream) {
    if (!p) return fail(CLIFFT_E_INVALID, "null program");
    if (!(flags & CLIFFT_REC_DEVICE))
        return fail(CLIFFT_E_INVALID, "sample_noalloc writes device outputs only (CLIFFT_REC_DEVICE)");
    if (p->has_forced_ops)
        return fail(CLIFFT_E_INVALID, "program contains *_FORCED opcodes; use clifft_svm_record_log_probs");
    if (shots == 0) return CLIFFT_OK;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : library_stream();
    LaunchShape L = p->base_shape;
    void* gs = nullptr;
    uint64_t max_teams = ~0ull;
    if (L.global_amps) {
        const size_t cap = kNoAllocScratchElems * sizeof(dv2);
        const size_t one = (size_t)16 << p->peak_rank;
        while (L.amp_slices > 1 && L.amp_slices * one > cap) --L.amp_slices;  // fused runs: 2 or 3 slices
        if (one * L.amp_slices > cap)
            return fail(CLIFFT_E_UNSUPPORTED, "sample_noalloc: a peak_rank " + std::to_string(p->peak_rank) +
                                                  " state needs " + std::to_string(one) +
                                                  " bytes; its scratch holds " + std::to_string(cap) +
                                                  " (peak_rank <= 22)");
        if (L.amp_slices == 1 && p->base_shape.amp_slices >= 2) L.lds_bytes -= mono_list_bytes(L);
        max_teams = cap / (one * L.amp_slices);
        int rc = noalloc_scratch(&gs);
        if (rc) return rc;
    }
    int rc = choose_global_team(L, shots, p->peak_rank);
    if (rc) return rc;
    const uint32_t tpb = L.teams_per_block();
    int per_cu = 0;
    HIP_TRY(shot_occupancy(L, &per_cu));
    const uint64_t blocks_needed = (shots + tpb - 1) / tpb;
    uint64_t grid = std::min<uint64_t>(blocks_needed, (uint64_t)device_info().num_cus * (uint64_t)std::max(per_cu, 1));
    if (L.global_amps) grid = std::min<uint64_t>(grid, std::max<uint64_t>(1, max_teams / tpb));
    L.grid = (uint32_t)std::max<uint64_t>(1, grid);
    if (L.global_amps && (uint64_t)L.grid * tpb > max_teams)
        return fail(CLIFFT_E_INVALID, "sample_noalloc: launch exceeds the scratch");
    ShotArgs A;
    fill_common_args(p, A, gs);
    A.amp_slices = L.global_amps ? L.amp_slices : 1u;
    A.shot_begin = 0;
    A.shot_count = shots;
    A.seed = seed;
    A.mode = MODE_SAMPLE;
    A.flags = flags;
    if (out) {
        A.out_meas = out->measurements;
        A.out_det = out->detectors;
        A.out_obs = out->observables;
        A.out_ev = out->exp_vals;
        A.out_disc = out->discarded;
    }
    HIP_TRY(launch_with_args(p, L, p->dp_dev, A, st));
    HIP_TRY(hipStreamSynchronize(st));
    return CLIFFT_OK;
}

}  // extern "C"

static int survivors_impl(clifft_svm_program* p, uint64_t shots, uint64_t seed, uint32_t flags,
                          clifft_svm_survivors* summary, const clifft_svm_records* out, void* stream, uint32_t mode,
                          const KfTables* kf) {
    if (!p || !summary) return fail(CLIFFT_E_INVALID, "null argument");
    if (flags & CLIFFT_REC_DEVICE) return fail(CLIFFT_E_UNSUPPORTED, "survivor compaction writes host memory");
    summary->total_shots = shots;
    summary->passed_shots = 0;
    summary->logical_errors = 0;
    if (summary->observable_ones)
        for (uint32_t i = 0; i < p->num_obs; ++i) summary->observable_ones[i] = 0;
    if (shots == 0) return CLIFFT_OK;
    const bool packed = (flags & CLIFFT_REC_PACKED) != 0;
    std::vector<uint8_t> disc(shots), obs((size_t)shots * p->num_obs);
    std::vector<uint8_t> meas, det;
    std::vector<double> ev;
    clifft_svm_records r{};
    r.discarded = disc.data();
    r.observables = p->num_obs ? obs.data() : nullptr;
    if (out) {
        meas.resize((size_t)shots * p->num_meas);
        det.resize((size_t)shots * p->num_det);
        ev.resize((size_t)shots * p->num_ev);
        r.measurements = out->measurements ? meas.data() : nullptr;
        r.detectors = out->detectors ? det.data() : nullptr;
        r.exp_vals = out->exp_vals ? ev.data() : nullptr;
    }
    // observables are always fetched unpacked for counting
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : library_stream();
    int rc = run_shots(p, shots, seed, flags & ~CLIFFT_REC_PACKED, mode, &r, nullptr, nullptr, st, kf);
    if (rc) return rc;
    const size_t bm = row_bytes(p->num_meas, false), bd = row_bytes(p->num_det, false);
    const size_t pm = row_bytes(p->num_meas, packed), pd = row_bytes(p->num_det, packed), po = row_bytes(p->num_obs, packed);
    auto put_row = [&](uint8_t* dst, const uint8_t* src, uint32_t n) {
        if (!packed) {
            std::memcpy(dst, src, n);
            return;
        }
        for (uint32_t j = 0; j < (n + 7) / 8; ++j) {
            uint8_t b = 0;
            for (uint32_t i = 8 * j; i < std::min(n, 8 * j + 8); ++i) b |= (uint8_t)((src[i] & 1u) << (i - 8 * j));
            dst[j] = b;
        }
    };
    uint64_t row = 0;
    for (uint64_t s = 0; s < shots; ++s) {
        if (disc[s]) continue;
        bool any = false;
        for (uint32_t i = 0; i < p->num_obs; ++i) {
            uint8_t v = obs[s * p->num_obs + i];
            if (v) {
                any = true;
                if (summary->observable_ones) summary->observable_ones[i]++;
            }
        }
        if (any) summary->logical_errors++;
        if (out) {
            if (out->observables && p->num_obs) put_row(out->observables + row * po, obs.data() + s * p->num_obs, p->num_obs);
            if (out->measurements && bm) put_row(out->measurements + row * pm, meas.data() + s * bm, p->num_meas);
            if (out->detectors && bd) put_row(out->detectors + row * pd, det.data() + s * bd, p->num_det);
            if (out->exp_vals && p->num_ev)
                std::memcpy(out->exp_vals + row * p->num_ev, ev.data() + s * p->num_ev, p->num_ev * 8);
        }
        row++;
    }
    summary->passed_shots = row;
    return CLIFFT_OK;
}

extern "C" {

int clifft_svm_sample_survivors(clifft_svm_program* p, uint64_t shots, uint64_t seed, uint32_t flags,
                                clifft_svm_survivors* summary, const clifft_svm_records* out, void* stream) {
    return survivors_impl(p, shots, seed, flags, summary, out, stream, MODE_SAMPLE, nullptr);
}

// Device tables of the k-fault law (svm.cc:376-431, 525-546) for this k.
static int kfault_tables(clifft_svm_program* p, uint32_t k, hipStream_t st, const KfTables** out) {
    const std::vector<double>& probs = p->fault_probs;
    const uint32_t n = (uint32_t)probs.size();
    if (k > n)
        return fail(CLIFFT_E_INVALID, "k (" + std::to_string(k) + ") exceeds total fault sites (" + std::to_string(n) + ")");
    uint32_t certain = 0, impossible = 0;
    for (double q : probs) {
        if (q <= 0.0)
            impossible++;
        else if (q >= 1.0)
            certain++;
    }
    if (k < certain || k > n - impossible)
        return fail(CLIFFT_E_INVALID, "k-fault stratum k=" + std::to_string(k) + " has zero probability mass (" +
                                          std::to_string(certain) + " sites have p=1, " + std::to_string(impossible) +
                                          " sites have p=0)");
    if (k > kMaxForcedFaults)
        return fail(CLIFFT_E_UNSUPPORTED, "device k-fault sampling supports k <= " + std::to_string(kMaxForcedFaults));
    bool uniform = true;
    for (uint32_t i = 1; i < n; ++i)
        if (probs[i] != probs[0]) uniform = false;
    const size_t o_w = align_up(sizeof(KfTables), 256), o_dp = o_w + align_up((size_t)(n + 1) * 8, 256);
    const size_t total = o_dp + (size_t)(n + 1) * (k + 1) * 8;
    if ((int64_t)k != p->kf_k || total > p->kf_dev_bytes) {
        if (total > p->kf_dev_bytes) {
            if (p->kf_dev) {
                HIP_TRY(hipDeviceSynchronize());
                (void)hipFree(p->kf_dev);
            }
            p->kf_dev = nullptr;
            p->kf_dev_bytes = 0;
            HIP_TRY(hipMalloc(&p->kf_dev, total));
            p->kf_dev_bytes = total;
        }
        std::vector<uint8_t> staging(total, 0);
        auto* w = reinterpret_cast<double*>(staging.data() + o_w);
        auto* dp = reinterpret_cast<double*>(staging.data() + o_dp);
        if (!uniform) {
            double sum = 0.0;  // odds ratios rescaled to mean 1
            for (uint32_t i = 0; i < n; ++i) {
                const double q = std::clamp(probs[i], 0.0, 1.0 - 1e-15);
                w[i] = q / (1.0 - q);
                sum += w[i];
            }
            if (sum > 0.0)
                for (uint32_t i = 0; i < n; ++i) w[i] *= (double)n / sum;
            const uint32_t stride = k + 1;  // dp[i][j] = e_j(w[i..n))
            for (uint32_t i = 0; i <= n; ++i) dp[(size_t)i * stride] = 1.0;
            for (int i = (int)n - 1; i >= 0; --i) {
                const uint32_t max_j = std::min(n - (uint32_t)i, k);
                for (uint32_t j = 1; j <= max_j; ++j)
                    dp[(size_t)i * stride + j] =
                        dp[(size_t)(i + 1) * stride + j] + w[i] * dp[(size_t)(i + 1) * stride + j - 1];
            }
        }
        auto* base = static_cast<uint8_t*>(p->kf_dev);
        KfTables T{};
        T.w = reinterpret_cast<const double*>(base + o_w);
        T.dp = reinterpret_cast<const double*>(base + o_dp);
        T.n = n;
        T.n_q = p->num_sites;
        T.k = k;
        T.uniform = uniform ? 1u : 0u;
        std::memcpy(staging.data(), &T, sizeof(T));
        HIP_TRY(hipMemcpyAsync(p->kf_dev, staging.data(), total, hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));  // staging is released on return
        p->kf_k = k;
    }
    *out = static_cast<const KfTables*>(p->kf_dev);
    return CLIFFT_OK;
}

int clifft_svm_sample_k(clifft_svm_program* p, uint64_t shots, uint32_t k, uint64_t seed, uint32_t flags,
                        const clifft_svm_records* out, void* stream) {
    if (!p) return fail(CLIFFT_E_INVALID, "null program");
    if (shots == 0) return CLIFFT_OK;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : library_stream();
    const KfTables* kf = nullptr;
    int rc = kfault_tables(p, k, st, &kf);
    if (rc) return rc;
    rc = run_shots(p, shots, seed, flags, MODE_KFAULT, out, nullptr, nullptr, st, kf);
    if (rc) return rc;
    if (flags & CLIFFT_REC_DEVICE) HIP_TRY(hipStreamSynchronize(st));
    return CLIFFT_OK;
}

int clifft_svm_sample_k_survivors(clifft_svm_program* p, uint64_t shots, uint32_t k, uint64_t seed, uint32_t flags,
                                  clifft_svm_survivors* summary, const clifft_svm_records* out, void* stream) {
    if (!p || !summary) return fail(CLIFFT_E_INVALID, "null argument");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : library_stream();
    const KfTables* kf = nullptr;
    int rc = kfault_tables(p, k, st, &kf);
    if (rc) return rc;
    return survivors_impl(p, shots, seed, flags, summary, out, stream, MODE_KFAULT, kf);
}

int clifft_svm_noise_site_probabilities(const clifft_svm_program* p, double* out, uint32_t* count) {
    if (!p || !count) return fail(CLIFFT_E_INVALID, "null argument");
    if (out) std::memcpy(out, p->fault_probs.data(), p->fault_probs.size() * sizeof(double));
    *count = (uint32_t)p->fault_probs.size();
    return CLIFFT_OK;
}

int clifft_svm_record_log_probs(clifft_svm_program* p, const uint8_t* records, uint64_t num_records,
                                double* log_probs, void* stream) {
    if (!p || (!records && num_records) || (!log_probs && num_records)) return fail(CLIFFT_E_INVALID, "null argument");
    if (p->total_slots != p->num_meas)
        return fail(CLIFFT_E_INVALID, "record_log_probs: program has hidden measurement slots");
    if (num_records == 0) return CLIFFT_OK;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : library_stream();
    return run_shots(p, num_records, 0, 0, MODE_FORCED, nullptr, records, log_probs, st);
}

int clifft_svm_execute(clifft_svm_program* p, clifft_svm_state* s) {
    if (!p || !s || !s->v || !s->px || !s->pz) return fail(CLIFFT_E_INVALID, "null argument");
    if (s->active_k > p->peak_rank) return fail(CLIFFT_E_INVALID, "state active_k exceeds peak_rank");
    if (p->has_forced_ops && !s->forced_record)
        return fail(CLIFFT_E_INVALID, "program contains *_FORCED opcodes but the state has no forced_record");
    hipStream_t st = library_stream();
    const size_t namp = (size_t)1 << p->peak_rank;
    const size_t nrec = (size_t)p->total_slots + p->num_det + p->num_obs;
    PresetState ps{};
    ps.gamma[0] = s->gamma_re;
    ps.gamma[1] = s->gamma_im;
    for (int i = 0; i < 4; ++i) ps.rng[i] = s->rng[i];
    ps.active_k = s->active_k;
    ps.discarded = s->discarded;
    ps.next_noise_idx = s->next_noise_idx;
    ps.forced_reachable = s->forced_reachable;
    ps.forced_log_probability = s->forced_log_probability;
    ps.dust_clamps = s->dust_clamps;
    for (uint32_t w = 0; w < p->words; ++w) {
        ps.px[w] = s->px[w];
        ps.pz[w] = s->pz[w];
    }
    std::vector<uint8_t> recs(nrec + 16, 0);
    if (s->meas_record) std::memcpy(recs.data(), s->meas_record, p->total_slots);
    if (s->det_record) std::memcpy(recs.data() + p->total_slots, s->det_record, p->num_det);
    if (s->obs_record) std::memcpy(recs.data() + p->total_slots + p->num_det, s->obs_record, p->num_obs);
    size_t s_ps = align_up(sizeof(PresetState), 256), s_a = align_up(namp * 16, 256), s_r = align_up(nrec + 16, 256);
    size_t s_e = align_up((p->num_ev + 1) * 8, 256), s_f = align_up(p->total_slots + 16, 256);
    int rc = ensure_out(p, s_ps + s_a + s_r + s_e + s_f);
    if (rc) return rc;
    uint8_t* o = static_cast<uint8_t*>(p->out_dev);
    auto* d_ps = reinterpret_cast<PresetState*>(o);
    auto* d_amp = reinterpret_cast<dv2*>(o + s_ps);
    uint8_t* d_rec = o + s_ps + s_a;
    auto* d_ev = reinterpret_cast<double*>(o + s_ps + s_a + s_r);
    uint8_t* d_forced = o + s_ps + s_a + s_r + s_e;
    HIP_TRY(hipMemcpyAsync(d_ps, &ps, sizeof(ps), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_amp, s->v, namp * 16, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_rec, recs.data(), nrec, hipMemcpyHostToDevice, st));
    if (s->exp_vals && p->num_ev) HIP_TRY(hipMemcpyAsync(d_ev, s->exp_vals, p->num_ev * 8, hipMemcpyHostToDevice, st));
    if (s->forced_record) HIP_TRY(hipMemcpyAsync(d_forced, s->forced_record, p->total_slots, hipMemcpyHostToDevice, st));
    LaunchShape L = p->base_shape.lane ? plan_shape(p, false) : p->base_shape;
    if (L.lds_bytes > kLdsBytes) return fail(CLIFFT_E_UNSUPPORTED, "records do not fit the LDS of a team");
    void* gs = nullptr;
    rc = finalize_shape(p, L, 1, st, &gs);
    if (rc) return rc;
    ShotArgs A;
    fill_common_args(p, A, gs);
    A.amp_slices = L.global_amps ? L.amp_slices : 1u;
    A.shot_begin = 0;
    A.shot_count = 1;
    A.mode = MODE_PRESET;
    A.preset = d_ps;
    A.preset_amps = d_amp;
    A.preset_records = d_rec;
    A.preset_ev = d_ev;
    A.forced_records = s->forced_record ? d_forced : nullptr;
    HIP_TRY(launch_with_args(p, L, p->dp_dev, A, st));
    // the kernel writes the final amplitudes back to d_amp
    HIP_TRY(hipMemcpyAsync(&ps, d_ps, sizeof(ps), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(recs.data(), d_rec, nrec, hipMemcpyDeviceToHost, st));
    if (s->exp_vals && p->num_ev) HIP_TRY(hipMemcpyAsync(s->exp_vals, d_ev, p->num_ev * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->v, d_amp, namp * 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    s->gamma_re = ps.gamma[0];
    s->gamma_im = ps.gamma[1];
    for (int i = 0; i < 4; ++i) s->rng[i] = ps.rng[i];
    s->active_k = ps.active_k;
    s->discarded = ps.discarded;
    s->next_noise_idx = ps.next_noise_idx;
    s->forced_reachable = ps.forced_reachable;
    s->forced_log_probability = ps.forced_log_probability;
    s->dust_clamps = ps.dust_clamps;
    for (uint32_t w = 0; w < p->words; ++w) {
        s->px[w] = ps.px[w];
        s->pz[w] = ps.pz[w];
    }
    if (s->meas_record) std::memcpy(s->meas_record, recs.data(), p->total_slots);
    if (s->det_record) std::memcpy(s->det_record, recs.data() + p->total_slots, p->num_det);
    if (s->obs_record) std::memcpy(s->obs_record, recs.data() + p->total_slots + p->num_det, p->num_obs);
    return CLIFFT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched execution
// ---------------------------------------------------------------------------
// A batch is cut into launch groups, one per kernel shape: the lane kernel
// (by frame-word width), wavefront teams, LDS workgroup teams (by team size)
// and global-amplitude workgroup teams.  Each group is one persistent launch
// whose teams drain the group's device queue of work items; the groups run
// one after another on the caller's stream, or side by side on the batch's
// own streams (tuning.concurrent).  Every program belongs to one group, and
// its ShotArgs carry the group's largest amplitude count and record stride,
// so one LDS carve-up serves every item of the launch.  batch_audit checks
// every item against its launch's allocations on the host before anything
// is uploaded.

struct BatchGroup {
    LaunchShape L;               // kernel shape; grid, block, lds_bytes of the launch
    std::vector<uint32_t> progs; // member program indices
    uint32_t amp_elems = 1;      // team tiers: largest 2^peak_rank of the members
    uint32_t rec_stride = 0;     // team tiers: largest record stride
    uint32_t wave_stride = 0;    // lane tier: largest lane_wave_bytes
    uint32_t amp_slices = 1;     // global tier: 2 or 3 when a member runs fused permutation/phase runs
    uint32_t teams_resident = 0; // teams in flight (grid x teams per workgroup)
    std::vector<BatchItem> items;
    size_t items_off = 0;        // offset of this group's items in the plan allocation
    void* gscratch = nullptr;
    size_t gscratch_bytes = 0;
};

struct clifft_svm_batch {
    std::vector<clifft_svm_program*> progs;
    std::vector<uint64_t> shots;
    std::vector<clifft_svm_records> outs;
    uint32_t flags = 0;
    clifft_svm_batch_tuning tuning{};
    std::vector<BatchGroup> groups;
    std::vector<uint32_t> group_of;  // per program
    std::vector<ShotArgs> args;      // per program, for the current plan
    // device: [program pointers | ShotArgs[n] | seeds[n]]
    void* dev = nullptr;
    const DevProgram** progs_dev = nullptr;
    ShotArgs* args_dev = nullptr;
    uint64_t* seeds_dev = nullptr;
    // device, per plan: [counters[G] | BatchArgs[G] | items of every group]
    void* plan_dev = nullptr;
    size_t plan_dev_bytes = 0;
    uint32_t* counters_dev = nullptr;
    BatchArgs* bargs_dev = nullptr;
    // side streams of concurrent groups and their join events
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_done;
    hipEvent_t fork = nullptr;
    // optional per-group kernel timing of eager runs (timing events on the
    // stream each group's launch goes to)
    bool time_groups = false;
    std::vector<hipEvent_t> t0, t1;
    bool timed_valid = false;
    // captured graph
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    hipStream_t graph_stream = nullptr;
    uint64_t graph_epoch = 0, epoch = 0;
    // pinned seed staging: a ring of slots, each reused only after the copy
    // that last read it has completed
    static constexpr int kSeedSlots = 4;
    uint64_t* seeds_host = nullptr;
    hipEvent_t seed_done[kSeedSlots] = {};
    int seed_slot = 0;
};

namespace {

clifft_svm_batch_tuning default_tuning() {
    clifft_svm_batch_tuning t{};
    // The autotuner's pick on one MI355X for 4096 random Clifford+T programs
    // (peak 3-16, 1000 shots each): wavefront teams with LDS amplitudes up to
    // peak 8, every program of peak 9-16 one shot per wavefront on global
    // amplitudes (no workgroup teams), lanes up to peak 6 -- 5.99e6 shots/s
    // against 4.64e6 for the previous default (workgroup teams with LDS
    // amplitudes for peak 8-10 and with global ones above 14).  Lanes up to
    // peak 7 (one wavefront per workgroup for the peak-7 columns) measured
    // later on the same batch: 441.9 against 452.3 ms per step.  A single
    // program's launch keeps the lane tier to CLIFFT_LANE_MAX_RANK (6): its
    // shots alone rarely fill the device at one wavefront per workgroup.
    t.global_threads = 256;
    t.wave_rank_max = 8;
    t.lds_rank_max = 8;
    t.wave_chunk = 4;
    t.block_chunk = 16;
    t.concurrent = 1;
    t.lane_rank_max = std::getenv("CLIFFT_LANE_MAX_RANK") ? std::min(lane_max_rank(), 7) : 7;
    t.lane_chunk = 128;
    t.gwave_rank_max = 16;
    return t;
}

int batch_check_tuning(const clifft_svm_batch_tuning& t) {
    if (t.global_threads != 256 && t.global_threads != 512 && t.global_threads != 1024)
        return fail(CLIFFT_E_INVALID, "batch tuning: global_threads must be 256, 512 or 1024");
    if (t.wave_rank_max > 10 || t.lds_rank_max > 13 || t.lds_rank_max < t.wave_rank_max)
        return fail(CLIFFT_E_INVALID, "batch tuning: need wave_rank_max <= lds_rank_max, wave_rank_max <= 10, lds_rank_max <= 13");
    if (t.wave_chunk == 0 || t.block_chunk == 0 || t.lane_chunk == 0)
        return fail(CLIFFT_E_INVALID, "batch tuning: zero chunk");
    if (t.lane_rank_max > 7) return fail(CLIFFT_E_INVALID, "batch tuning: lane_rank_max > 7");
    if (t.concurrent > 1) return fail(CLIFFT_E_INVALID, "batch tuning: concurrent must be 0 or 1");
    if (t.gwave_rank_max > kMaxPeakRank) return fail(CLIFFT_E_INVALID, "batch tuning: gwave_rank_max > 30");
    return CLIFFT_OK;
}

// Kernel shape of one program in a batch (the plain planner's choice with
// the tuning's rank thresholds).  Returns false if its records do not fit.
bool batch_shape(const clifft_svm_program* p, const clifft_svm_batch_tuning& t, LaunchShape& L) {
    L = LaunchShape();
    L.nw = (int)p->nw;
    const uint32_t k = p->peak_rank;
    const size_t red = kTeamScratchBytes, rec = rec_stride_of(p);
    const size_t lane_bytes = lane_wave_bytes(k, p->total_slots, p->num_det, p->num_obs);
    if (p->words == 1 && (int32_t)k <= t.lane_rank_max && lane_bytes <= kBatchMaxDynLds) {
        L.lane = true;
        L.lane_narrow = p->frame32;
        // columns too large for two wavefronts per workgroup (peak 7): a
        // group of its own, so the smaller programs keep their residency
        if (2 * lane_bytes > kBatchMaxDynLds) L.block = 64;
        return true;
    }
    if (k <= t.wave_rank_max && red + 4 * (rec + ((size_t)16 << k)) <= kBatchMaxDynLds) {
        L.block = 256;
        return true;
    }
    if (k <= t.lds_rank_max) {
        const uint32_t waves = k <= 9 ? 2 : (k <= 10 ? 4 : (k <= 12 ? 8 : 16));
        if (red + waves * rec + ((size_t)16 << k) <= kBatchMaxDynLds) {
            L.block_team = true;
            L.block = 64 * waves;
            return true;
        }
    }
    L.global_amps = true;
    if (k <= t.gwave_rank_max && red + 4 * (rec + kMonoBytes) <= kBatchMaxDynLds) {
        L.block = 256;  // four wavefront teams per workgroup
        return true;
    }
    L.block_team = true;
    L.block = t.global_threads;
    while (red + (L.block / 64) * rec > kBatchMaxDynLds && L.block > 256) L.block /= 2;
    return red + (L.block / 64) * rec <= kBatchMaxDynLds;
}

bool same_kernel(const LaunchShape& a, const LaunchShape& b) {
    return a.lane == b.lane && a.lane_narrow == b.lane_narrow && a.nw == b.nw && a.block_team == b.block_team &&
           a.global_amps == b.global_amps && a.block == b.block;
}

bool mono_eligible(const clifft_svm_program* p) {
    return fuse_mono_enabled() && has_fusable(p) && p->peak_rank <= 24;
}

// Dynamic LDS of a group's launch from its largest members.
size_t group_lds(const BatchGroup& g) {
    const LaunchShape& L = g.L;
    if (L.lane) return (size_t)(L.block / 64) * g.wave_stride;
    size_t lds = kTeamScratchBytes + (size_t)(L.block / 64) * g.rec_stride;
    if (!L.global_amps) lds += (size_t)L.teams_per_block() * g.amp_elems * 16;
    if (L.global_amps && g.amp_slices >= 2) lds += (size_t)L.teams_per_block() * kMonoBytes;  // a list per team
    return lds;
}

void group_sizes(BatchGroup& g, const std::vector<clifft_svm_program*>& progs) {
    g.amp_elems = 1;
    g.rec_stride = 0;
    g.wave_stride = 0;
    g.amp_slices = 1;
    for (uint32_t i : g.progs) {
        const clifft_svm_program* p = progs[i];
        g.amp_elems = std::max(g.amp_elems, 1u << p->peak_rank);
        g.rec_stride = std::max(g.rec_stride, (uint32_t)rec_stride_of(p));
        g.wave_stride = std::max(g.wave_stride, lane_wave_bytes(p->peak_rank, p->total_slots, p->num_det, p->num_obs));
    }
    if (g.L.lane) {
        const uint32_t waves = (uint32_t)std::min<size_t>(4, kBatchMaxDynLds / std::max<uint32_t>(g.wave_stride, 1));
        g.L.block = 64 * std::max<uint32_t>(waves, 1);
        g.L.lds_bytes = group_lds(g);
        g.L.lane_dense = 6 * g.L.lds_bytes <= kLdsBytes;
        return;
    }
    if (g.L.global_amps) {
        bool any = false;
        for (uint32_t i : g.progs) any |= mono_eligible(progs[i]);
        uint32_t slices = 1;
        for (uint32_t i : g.progs)
            if (mono_eligible(progs[i])) slices = std::max(slices, mono_slices(progs[i]));
        g.amp_slices = slices;
        if (any && group_lds(g) > kBatchMaxDynLds) g.amp_slices = 1;
    }
    g.L.amp_slices = g.amp_slices;
    g.L.rec_stride = g.rec_stride;
    g.L.lds_bytes = group_lds(g);
}

// Estimated team time of one shot (arbitrary units): the interpreter's
// per-instruction issue cost plus the state traffic spread over the team.
double shot_cost(const clifft_svm_program* p, const LaunchShape& L) {
    if (L.lane) return ((double)p->host_bc.size() + p->bytes_per_shot / 64.0) / 64.0;
    const double lanes = L.block_team ? (double)L.block : 64.0;
    const double per_byte = L.global_amps ? 1.0 / 8.0 : 1.0 / 32.0;
    return (double)p->host_bc.size() + p->bytes_per_shot / lanes * per_byte;
}

// Group the programs and cut their shots (at most `cap` per program) into
// items, each group's sorted by decreasing estimated cost.
int batch_group(clifft_svm_batch* b, uint64_t cap) {
    const clifft_svm_batch_tuning& t = b->tuning;
    std::vector<BatchGroup> groups;
    std::vector<uint32_t> group_of(b->progs.size());
    std::vector<LaunchShape> shape(b->progs.size());
    for (size_t i = 0; i < b->progs.size(); ++i) {
        if (!batch_shape(b->progs[i], t, shape[i]))
            return fail(CLIFFT_E_UNSUPPORTED, "batch: the records of program " + std::to_string(i) +
                                                  " do not fit the LDS of a 256-thread team");
        size_t gi = 0;
        while (gi < groups.size() && !same_kernel(groups[gi].L, shape[i])) ++gi;
        if (gi == groups.size()) {
            groups.emplace_back();
            groups.back().L = shape[i];
        }
        groups[gi].progs.push_back((uint32_t)i);
    }
    // A team group whose largest amplitudes and largest records do not fit
    // together hands its largest-record members to the global group.
    std::vector<uint32_t> evicted;
    for (BatchGroup& g : groups) {
        if (g.L.lane || g.L.global_amps) continue;
        std::stable_sort(g.progs.begin(), g.progs.end(), [&](uint32_t x, uint32_t y) {
            return rec_stride_of(b->progs[x]) > rec_stride_of(b->progs[y]);
        });
        group_sizes(g, b->progs);
        while (g.L.lds_bytes > kBatchMaxDynLds && g.progs.size() > 1) {
            evicted.push_back(g.progs.front());
            g.progs.erase(g.progs.begin());
            group_sizes(g, b->progs);
        }
    }
    for (uint32_t i : evicted) {
        LaunchShape L;
        L.nw = (int)b->progs[i]->nw;
        L.block_team = true;
        L.global_amps = true;
        L.block = t.global_threads;
        const size_t rec = rec_stride_of(b->progs[i]);
        while (kTeamScratchBytes + (L.block / 64) * rec > kBatchMaxDynLds && L.block > 256) L.block /= 2;
        size_t gi = 0;
        while (gi < groups.size() && !same_kernel(groups[gi].L, L)) ++gi;
        if (gi == groups.size()) {
            groups.emplace_back();
            groups.back().L = L;
        }
        groups[gi].progs.push_back(i);
    }
    std::vector<BatchGroup> kept;
    for (BatchGroup& g : groups) {
        if (g.progs.empty()) continue;
        group_sizes(g, b->progs);
        if (g.L.lds_bytes > kBatchMaxDynLds)
            return fail(CLIFFT_E_UNSUPPORTED, "batch: a launch group's records do not fit the LDS");
        kept.push_back(std::move(g));
    }
    groups.swap(kept);
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        BatchGroup& g = groups[gi];
        struct Pending {
            BatchItem it;
            double cost;
        };
        std::vector<Pending> work;
        const uint32_t chunk = g.L.lane ? std::max<uint32_t>(64, t.lane_chunk / 64 * 64)
                                        : (g.L.block_team ? t.block_chunk : t.wave_chunk);
        for (uint32_t i : g.progs) {
            group_of[i] = (uint32_t)gi;
            const uint64_t n = std::min<uint64_t>(b->shots[i], cap);
            const double c = shot_cost(b->progs[i], g.L);
            for (uint64_t lo = 0; lo < n; lo += chunk) {
                const uint64_t hi = std::min<uint64_t>(n, lo + chunk);
                Pending w;
                w.it.prog = i;
                w.it.shot_lo = (uint32_t)lo;
                w.it.shot_hi = (uint32_t)hi;
                w.it.pad = 0;
                w.cost = c * (double)(hi - lo);
                work.push_back(w);
            }
        }
        std::stable_sort(work.begin(), work.end(), [](const Pending& x, const Pending& y) { return x.cost > y.cost; });
        g.items.resize(work.size());
        for (size_t j = 0; j < work.size(); ++j) g.items[j] = work[j].it;
    }
    b->groups.swap(groups);
    b->group_of.swap(group_of);
    return CLIFFT_OK;
}

// Host-side bounds audit of a plan: every item's program, shot range, LDS
// carve-up and global slice against the allocations its launch gets.
int batch_audit(const clifft_svm_batch* b) {
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        const BatchGroup& g = b->groups[gi];
        const LaunchShape& L = g.L;
        const std::string where = "batch plan, group " + std::to_string(gi) + ": ";
        if (L.lds_bytes > kBatchMaxDynLds) return fail(CLIFFT_E_INVALID, where + "LDS over the limit");
        if (L.block == 0 || L.block > 1024 || L.block % 64) return fail(CLIFFT_E_INVALID, where + "bad block");
        if (!L.block_team && !L.lane && L.block > 256) return fail(CLIFFT_E_INVALID, where + "bad wave-team block");
        if (L.global_amps && !g.gscratch) return fail(CLIFFT_E_INVALID, where + "global group without scratch");
        if (L.global_amps && (size_t)g.teams_resident * g.amp_elems * g.amp_slices * 16 > g.gscratch_bytes)
            return fail(CLIFFT_E_INVALID, where + "global slices exceed the scratch");
        for (const BatchItem& it : g.items) {
            if (it.prog >= b->progs.size() || b->group_of[it.prog] != gi)
                return fail(CLIFFT_E_INVALID, where + "item names a program outside the group");
            const clifft_svm_program* p = b->progs[it.prog];
            const ShotArgs& A = b->args[it.prog];
            if (it.shot_lo >= it.shot_hi || it.shot_hi > b->shots[it.prog])
                return fail(CLIFFT_E_INVALID, where + "item shot range outside the program's shots");
            if (L.lane) {
                const size_t need = lane_wave_bytes(p->peak_rank, p->total_slots, p->num_det, p->num_obs);
                if (need > g.wave_stride || (size_t)(L.block / 64) * g.wave_stride > L.lds_bytes)
                    return fail(CLIFFT_E_INVALID, where + "lane columns exceed the wave's LDS slice");
                continue;
            }
            if (A.amp_elems != g.amp_elems || A.rec_stride != g.rec_stride)
                return fail(CLIFFT_E_INVALID, where + "ShotArgs carve-up differs from the group's");
            if ((1u << p->peak_rank) > A.amp_elems || rec_stride_of(p) > A.rec_stride)
                return fail(CLIFFT_E_INVALID, where + "program exceeds the group's amplitude or record room");
            if (p->total_slots + p->num_det + p->num_obs > A.rec_stride)
                return fail(CLIFFT_E_INVALID, where + "records exceed the record stride");
            size_t carve = kTeamScratchBytes + (size_t)(L.block / 64) * A.rec_stride;
            if (!L.global_amps) carve += (size_t)L.teams_per_block() * A.amp_elems * 16;
            if (L.global_amps && A.amp_slices >= 2) carve += (size_t)L.teams_per_block() * kMonoBytes;
            if (carve > L.lds_bytes) return fail(CLIFFT_E_INVALID, where + "LDS carve-up exceeds the launch's LDS");
            if (L.global_amps && A.amp_slices > g.amp_slices)
                return fail(CLIFFT_E_INVALID, where + "program uses more slices than a team holds");
            if (!L.global_amps && A.amp_slices != 1)
                return fail(CLIFFT_E_INVALID, where + "LDS tier with a second slice");
        }
    }
    return CLIFFT_OK;
}

// Teams of a launch group that stay resident at once, from the kernel's
// occupancy; the global group is also held to the device memory budget.
int batch_grids(clifft_svm_batch* b) {
    DeviceInfo& di = device_info();
    for (BatchGroup& g : b->groups) {
        int per_cu = 0;
        HIP_TRY(batch_occupancy(g.L, &per_cu));
        if (per_cu <= 0) return fail(CLIFFT_E_UNSUPPORTED, "batch: a launch group does not fit a compute unit");
        uint64_t grid = (uint64_t)per_cu * (uint64_t)di.num_cus;
        const uint64_t tpb = g.L.lane ? g.L.block / 64 : g.L.teams_per_block();  // queue consumers per workgroup
        grid = std::min<uint64_t>(grid, std::max<uint64_t>(1, (g.items.size() + tpb - 1) / tpb));
        if (g.L.global_amps) {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 8ull << 30;
            size_t budget = free_b + g.gscratch_bytes;
            const size_t headroom = std::max<size_t>(budget * 2 / 5, 16ull << 30);
            budget = budget > headroom ? budget - headroom : (size_t)1 << 30;
            budget = std::min<size_t>(budget, 192ull << 30);
            const size_t per_team = (size_t)g.amp_elems * g.amp_slices * 16;
            const uint64_t teams = g.L.teams_per_block();
            grid = std::min<uint64_t>(grid, std::max<uint64_t>(1, budget / (per_team * teams)));
            const size_t need = (size_t)grid * teams * per_team;
            if (need > g.gscratch_bytes) {
                if (g.gscratch) {
                    HIP_TRY(hipDeviceSynchronize());
                    (void)hipFree(g.gscratch);
                }
                g.gscratch = nullptr;
                g.gscratch_bytes = 0;
                HIP_TRY(hipMalloc(&g.gscratch, need));
                g.gscratch_bytes = need;
            }
        }
        g.L.grid = (uint32_t)grid;
        g.teams_resident = (uint32_t)(grid * (g.L.lane ? 1 : g.L.teams_per_block()));
    }
    return CLIFFT_OK;
}

void free_group_scratch(std::vector<BatchGroup>& groups) {
    for (BatchGroup& g : groups)
        if (g.gscratch) (void)hipFree(g.gscratch);
}

// Plan the batch for its tuning (each program capped at `cap` shots, the
// autotuner's sample), audit the plan and upload it.
int batch_plan(clifft_svm_batch* b, uint64_t cap, hipStream_t st) {
    int rc = batch_check_tuning(b->tuning);
    if (rc) return rc;
    // launches of the previous plan may still read its items and arguments
    HIP_TRY(hipDeviceSynchronize());
    std::vector<BatchGroup> old;
    old.swap(b->groups);
    rc = batch_group(b, cap);
    if (rc) {
        b->groups.swap(old);
        return rc;
    }
    // keep the global scratch of the previous plan when its shape matches
    for (BatchGroup& g : b->groups)
        for (BatchGroup& o : old)
            if (o.gscratch && same_kernel(o.L, g.L) && g.L.global_amps) {
                g.gscratch = o.gscratch;
                g.gscratch_bytes = o.gscratch_bytes;
                o.gscratch = nullptr;
            }
    free_group_scratch(old);
    rc = batch_grids(b);
    if (rc) return rc;
    // per-program launch arguments with the group's carve-up
    for (size_t i = 0; i < b->progs.size(); ++i) {
        clifft_svm_program* p = b->progs[i];
        const BatchGroup& g = b->groups[b->group_of[i]];
        ShotArgs& A = b->args[i];
        fill_common_args(p, A, nullptr);
        A.shot_begin = 0;
        A.shot_count = b->shots[i];
        A.mode = MODE_SAMPLE;
        A.flags = b->flags;
        A.out_meas = b->outs[i].measurements;
        A.out_det = b->outs[i].detectors;
        A.out_obs = b->outs[i].observables;
        A.out_ev = b->outs[i].exp_vals;
        A.out_disc = b->outs[i].discarded;
        if (!g.L.lane) {
            A.amp_elems = g.amp_elems;
            A.rec_stride = g.rec_stride;
            A.amp_slices = g.L.global_amps && g.amp_slices >= 2 && mono_eligible(p)
                               ? std::min(g.amp_slices, mono_slices(p)) : 1u;
        }
    }
    rc = batch_audit(b);
    if (rc) return rc;
    const size_t G = b->groups.size();
    const size_t o_bargs = align_up(G * 4, 256);
    size_t o_items = o_bargs + align_up(G * sizeof(BatchArgs), 256);
    size_t total = o_items;
    for (BatchGroup& g : b->groups) {
        g.items_off = total;
        total += align_up(std::max<size_t>(g.items.size(), 1) * sizeof(BatchItem), 256);
    }
    if (total > b->plan_dev_bytes) {
        if (b->plan_dev) {
            HIP_TRY(hipDeviceSynchronize());
            (void)hipFree(b->plan_dev);
        }
        b->plan_dev = nullptr;
        b->plan_dev_bytes = 0;
        HIP_TRY(hipMalloc(&b->plan_dev, total));
        b->plan_dev_bytes = total;
    }
    auto* base = static_cast<uint8_t*>(b->plan_dev);
    b->counters_dev = reinterpret_cast<uint32_t*>(base);
    b->bargs_dev = reinterpret_cast<BatchArgs*>(base + o_bargs);
    std::vector<uint8_t> staging(total, 0);
    for (size_t gi = 0; gi < G; ++gi) {
        const BatchGroup& g = b->groups[gi];
        BatchArgs B{};
        B.items = reinterpret_cast<const BatchItem*>(base + g.items_off);
        B.progs = b->progs_dev;
        B.args = b->args_dev;
        B.seeds = b->seeds_dev;
        B.counter = b->counters_dev + gi;
        B.gscratch = reinterpret_cast<dv2*>(g.gscratch);
        B.team_elems = g.L.global_amps ? (uint64_t)g.amp_elems * g.amp_slices : 0;
        B.n_items = (uint32_t)g.items.size();
        B.wave_stride = g.wave_stride;
        std::memcpy(staging.data() + o_bargs + gi * sizeof(BatchArgs), &B, sizeof(B));
        if (!g.items.empty())
            std::memcpy(staging.data() + g.items_off, g.items.data(), g.items.size() * sizeof(BatchItem));
    }
    HIP_TRY(hipMemcpyAsync(b->plan_dev, staging.data(), total, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->args_dev, b->args.data(), b->args.size() * sizeof(ShotArgs), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));  // host staging above is freed on return
    // side streams for concurrent groups (default priority)
    while (b->side.size() < G) {
        hipStream_t s2;
        hipEvent_t ev;
        HIP_TRY(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        b->side.push_back(s2);
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b->side_done.push_back(ev);
    }
    if (!b->fork) HIP_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    while (b->t0.size() < G) {
        hipEvent_t a, c;
        HIP_TRY(hipEventCreate(&a));
        b->t0.push_back(a);
        HIP_TRY(hipEventCreate(&c));
        b->t1.push_back(c);
    }
    b->timed_valid = false;
    b->epoch++;
    return CLIFFT_OK;
}

int batch_upload_seeds(clifft_svm_batch* b, const uint64_t* seeds, hipStream_t st) {
    if (!seeds) return fail(CLIFFT_E_INVALID, "null seeds");
    const size_t n = b->progs.size();
    const int slot = b->seed_slot;
    b->seed_slot = (slot + 1) % clifft_svm_batch::kSeedSlots;
    HIP_TRY(hipEventSynchronize(b->seed_done[slot]));
    uint64_t* host = b->seeds_host + (size_t)slot * n;
    std::memcpy(host, seeds, n * sizeof(uint64_t));
    HIP_TRY(hipMemcpyAsync(b->seeds_dev, host, n * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(b->seed_done[slot], st));
    return CLIFFT_OK;
}

// Queue resets and every group's launch, ordered after prior work on `st`
// and before later work on it.
int batch_enqueue(clifft_svm_batch* b, hipStream_t st, bool capturing = false) {
    const uint32_t G = (uint32_t)b->groups.size();
    if (G == 0) return CLIFFT_OK;
    b->timed_valid = b->time_groups && !capturing;
    HIP_TRY(clear_counters(b->counters_dev, G, st));
    const bool conc = b->tuning.concurrent && G > 1;
    if (conc) HIP_TRY(hipEventRecord(b->fork, st));
    for (uint32_t gi = 0; gi < G; ++gi) {
        const BatchGroup& g = b->groups[gi];
        if (g.items.empty()) continue;
        hipStream_t s = st;
        if (conc) {
            s = b->side[gi];
            HIP_TRY(hipStreamWaitEvent(s, b->fork, 0));
        }
        const bool timed = b->time_groups && !capturing;
        if (timed) HIP_TRY(hipEventRecord(b->t0[gi], s));
        HIP_TRY(launch_batch_group(g.L, b->bargs_dev + gi, s));
        if (timed) HIP_TRY(hipEventRecord(b->t1[gi], s));
        if (conc) {
            HIP_TRY(hipEventRecord(b->side_done[gi], s));
            HIP_TRY(hipStreamWaitEvent(st, b->side_done[gi], 0));
        }
    }
    return CLIFFT_OK;
}

}  // namespace

extern "C" {

int clifft_svm_batch_create(clifft_svm_program* const* programs, const uint64_t* shots, uint32_t num_programs,
                            uint32_t flags, const clifft_svm_records* outs, clifft_svm_batch** out) {
    if (!out) return fail(CLIFFT_E_INVALID, "null output handle");
    *out = nullptr;
    if (!programs || !shots || !outs || num_programs == 0) return fail(CLIFFT_E_INVALID, "null argument or empty batch");
    auto* b = new clifft_svm_batch();
    b->flags = flags | CLIFFT_REC_DEVICE;
    for (uint32_t i = 0; i < num_programs; ++i) {
        clifft_svm_program* p = programs[i];
        if (!p) {
            delete b;
            return fail(CLIFFT_E_INVALID, "null program in batch");
        }
        if (p->has_forced_ops) {
            delete b;
            return fail(CLIFFT_E_INVALID, "batch: program contains *_FORCED opcodes");
        }
        if (shots[i] > 0xFFFFFFFFull) {
            delete b;
            return fail(CLIFFT_E_UNSUPPORTED, "batch: more than 2^32-1 shots for one program");
        }
        b->progs.push_back(p);
        b->shots.push_back(shots[i]);
        b->outs.push_back(outs[i]);
    }
    const size_t n = num_programs;
    b->args.resize(n);
    const size_t o_args = align_up(n * 8, 256);
    const size_t o_seeds = o_args + align_up(n * sizeof(ShotArgs), 256);
    const size_t total = o_seeds + align_up(n * 8, 256);
    hipError_t e = hipMalloc(&b->dev, total);
    if (e != hipSuccess) {
        delete b;
        return fail(CLIFFT_E_HIP, std::string("batch upload: ") + hipGetErrorString(e));
    }
    auto* base = static_cast<uint8_t*>(b->dev);
    b->progs_dev = reinterpret_cast<const DevProgram**>(base);
    b->args_dev = reinterpret_cast<ShotArgs*>(base + o_args);
    b->seeds_dev = reinterpret_cast<uint64_t*>(base + o_seeds);
    std::vector<uint8_t> staging(total, 0);
    for (size_t i = 0; i < n; ++i) {
        const DevProgram* d = b->progs[i]->dp_dev;
        std::memcpy(staging.data() + i * 8, &d, 8);
    }
    e = hipMemcpy(base, staging.data(), total, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&b->seeds_host), clifft_svm_batch::kSeedSlots * n * sizeof(uint64_t),
                          hipHostMallocDefault);
    for (int i = 0; i < clifft_svm_batch::kSeedSlots && e == hipSuccess; ++i)
        e = hipEventCreateWithFlags(&b->seed_done[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        clifft_svm_batch_destroy(b);
        return fail(CLIFFT_E_HIP, std::string("batch upload: ") + hipGetErrorString(e));
    }
    b->tuning = default_tuning();
    int rc = batch_plan(b, ~0ull, library_stream());
    if (rc) {
        std::string msg = g_last_error;
        clifft_svm_batch_destroy(b);
        return fail(rc, msg);
    }
    *out = b;
    return CLIFFT_OK;
}

int clifft_svm_batch_destroy(clifft_svm_batch* b) {
    if (!b) return CLIFFT_OK;
    (void)hipDeviceSynchronize();
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->dev) (void)hipFree(b->dev);
    if (b->plan_dev) (void)hipFree(b->plan_dev);
    free_group_scratch(b->groups);
    for (hipStream_t s : b->side) (void)hipStreamDestroy(s);
    for (hipEvent_t ev : b->side_done) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : b->t0) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : b->t1) (void)hipEventDestroy(ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->seeds_host) (void)hipHostFree(b->seeds_host);
    for (auto& ev : b->seed_done)
        if (ev) (void)hipEventDestroy(ev);
    delete b;
    return CLIFFT_OK;
}

int clifft_svm_batch_set_tuning(clifft_svm_batch* b, const clifft_svm_batch_tuning* t) {
    if (!b || !t) return fail(CLIFFT_E_INVALID, "null argument");
    int rc = batch_check_tuning(*t);
    if (rc) return rc;
    clifft_svm_batch_tuning old = b->tuning;
    b->tuning = *t;
    rc = batch_plan(b, ~0ull, library_stream());
    if (rc) {
        std::string msg = g_last_error;
        b->tuning = old;
        (void)batch_plan(b, ~0ull, library_stream());
        return fail(rc, msg);
    }
    return CLIFFT_OK;
}

int clifft_svm_batch_get_tuning(const clifft_svm_batch* b, clifft_svm_batch_tuning* t) {
    if (!b || !t) return fail(CLIFFT_E_INVALID, "null argument");
    *t = b->tuning;
    return CLIFFT_OK;
}

int clifft_svm_batch_info(const clifft_svm_batch* b, uint32_t max_groups, clifft_svm_batch_group_info* info,
                          uint32_t* num_groups) {
    if (!b || !num_groups) return fail(CLIFFT_E_INVALID, "null argument");
    *num_groups = (uint32_t)b->groups.size();
    for (uint32_t gi = 0; gi < max_groups && gi < b->groups.size() && info; ++gi) {
        const BatchGroup& g = b->groups[gi];
        clifft_svm_batch_group_info& o = info[gi];
        o.kind = g.L.lane ? 3u : (g.L.global_amps ? (g.L.block_team ? 2u : 4u) : (g.L.block_team ? 1u : 0u));
        o.programs = (uint32_t)g.progs.size();
        o.items = (uint32_t)g.items.size();
        o.grid = g.L.grid;
        o.block = g.L.block;
        o.lds_bytes = (uint32_t)g.L.lds_bytes;
        o.amp_elems = g.L.lane ? 0u : g.amp_elems;
        o.amp_slices = g.amp_slices;
        const bool packed = (b->flags & CLIFFT_REC_PACKED) != 0;
        double bytes = 0.0;
        for (uint32_t i : g.progs) {
            const clifft_svm_program* p = b->progs[i];
            const clifft_svm_records& r = b->outs[i];
            const double rec = (r.measurements ? row_bytes(p->num_meas, packed) : 0) +
                               (r.detectors ? row_bytes(p->num_det, packed) : 0) +
                               (r.observables ? row_bytes(p->num_obs, packed) : 0);
            bytes += (p->bytes_per_shot + rec) * (double)b->shots[i];
        }
        o.algorithmic_bytes = bytes;
    }
    return CLIFFT_OK;
}

int clifft_svm_batch_time_groups(clifft_svm_batch* b, uint32_t enable) {
    if (!b) return fail(CLIFFT_E_INVALID, "null batch");
    b->time_groups = enable != 0;
    b->timed_valid = false;
    return CLIFFT_OK;
}

int clifft_svm_batch_group_times(clifft_svm_batch* b, uint32_t max_groups, float* ms, uint32_t* num_groups) {
    if (!b || !num_groups) return fail(CLIFFT_E_INVALID, "null argument");
    if (!b->timed_valid) return fail(CLIFFT_E_INVALID, "no timed eager run since timing was enabled or the plan changed");
    *num_groups = (uint32_t)b->groups.size();
    for (uint32_t gi = 0; gi < max_groups && gi < b->groups.size() && ms; ++gi) {
        ms[gi] = 0.f;
        if (b->groups[gi].items.empty()) continue;
        HIP_TRY(hipEventSynchronize(b->t1[gi]));
        HIP_TRY(hipEventElapsedTime(&ms[gi], b->t0[gi], b->t1[gi]));
    }
    return CLIFFT_OK;
}

int clifft_svm_batch_run(clifft_svm_batch* b, const uint64_t* seeds, void* stream) {
    if (!b) return fail(CLIFFT_E_INVALID, "null batch");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : library_stream();
    int rc = batch_upload_seeds(b, seeds, st);
    if (rc) return rc;
    return batch_enqueue(b, st);
}

int clifft_svm_batch_run_graph(clifft_svm_batch* b, const uint64_t* seeds, void* stream) {
    if (!b) return fail(CLIFFT_E_INVALID, "null batch");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : library_stream();
    int rc = batch_upload_seeds(b, seeds, st);
    if (rc) return rc;
    if (b->groups.empty()) return CLIFFT_OK;
    if (!b->graph_exec || b->graph_stream != st || b->graph_epoch != b->epoch) {
        if (b->graph_exec) {
            HIP_TRY(hipStreamSynchronize(st));
            (void)hipGraphExecDestroy(b->graph_exec);
            b->graph_exec = nullptr;
        }
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
        hipStream_t cap;
        HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            const int qrc = batch_enqueue(b, cap, true);
            e = hipStreamEndCapture(cap, &b->graph);
            if (qrc != CLIFFT_OK && e == hipSuccess) e = hipErrorUnknown;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cap);
        if (e != hipSuccess) return fail(CLIFFT_E_HIP, std::string("batch graph capture: ") + hipGetErrorString(e));
        b->graph_stream = st;
        b->graph_epoch = b->epoch;
    }
    HIP_TRY(hipGraphLaunch(b->graph_exec, st));
    return CLIFFT_OK;
}

int clifft_svm_batch_tune(clifft_svm_batch* b, uint64_t sample_shots, void* stream, clifft_svm_batch_tuning* chosen) {
    if (!b) return fail(CLIFFT_E_INVALID, "null batch");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : library_stream();
    if (sample_shots == 0) sample_shots = ~0ull;
    std::vector<uint64_t> seeds(b->progs.size());
    for (size_t i = 0; i < seeds.size(); ++i) seeds[i] = 0x5eed0000ull + i;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // a plan's time: the faster of two timed runs after a warm-up (one run's
    // time varies by a few per cent, as much as the knobs' differences)
    auto time_of = [&](const clifft_svm_batch_tuning& t, float* ms) -> int {
        b->tuning = t;
        int rc = batch_plan(b, sample_shots, st);
        if (rc) return rc;
        rc = clifft_svm_batch_run(b, seeds.data(), st);  // warm-up
        if (rc) return rc;
        for (int rep = 0; rep < 2; ++rep) {
            HIP_TRY(hipEventRecord(e0, st));
            rc = clifft_svm_batch_run(b, seeds.data(), st);
            if (rc) return rc;
            HIP_TRY(hipEventRecord(e1, st));
            HIP_TRY(hipEventSynchronize(e1));
            float one = 0.f;
            HIP_TRY(hipEventElapsedTime(&one, e0, e1));
            if (rep == 0 || one < *ms) *ms = one;
        }
        return CLIFFT_OK;
    };
    // Coordinate descent over one knob at a time, starting from the default.
    clifft_svm_batch_tuning best = default_tuning();
    float best_ms = 0.f;
    int rc = time_of(best, &best_ms);
    auto try_u32 = [&](uint32_t clifft_svm_batch_tuning::*field, std::initializer_list<uint32_t> values) {
        for (uint32_t v : values) {
            if (rc) return;
            clifft_svm_batch_tuning t = best;
            t.*field = v;
            if (batch_check_tuning(t) != CLIFFT_OK) continue;
            float ms = 0.f;
            int r = time_of(t, &ms);
            if (r == CLIFFT_E_HIP) {
                rc = r;
                return;
            }
            if (r == CLIFFT_OK && ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
    };
    auto try_lane = [&](std::initializer_list<int32_t> values) {
        for (int32_t v : values) {
            if (rc) return;
            clifft_svm_batch_tuning t = best;
            t.lane_rank_max = v;
            float ms = 0.f;
            int r = time_of(t, &ms);
            if (r == CLIFFT_E_HIP) {
                rc = r;
                return;
            }
            if (r == CLIFFT_OK && ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
    };
    if (!rc) try_u32(&clifft_svm_batch_tuning::concurrent, {0});
    if (!rc) try_lane({-1, 4, 5, 6});
    if (!rc) try_u32(&clifft_svm_batch_tuning::lane_chunk, {64, 256, 512, 1024});
    // the LDS thresholds together: which peaks run on wavefront teams in LDS,
    // on workgroup teams in LDS, and (above both) on global amplitudes
    auto try_lds = [&](std::initializer_list<std::pair<uint32_t, uint32_t>> pairs) {
        for (const auto& wl : pairs) {
            if (rc) return;
            clifft_svm_batch_tuning t = best;
            t.wave_rank_max = wl.first;
            t.lds_rank_max = wl.second;
            if (batch_check_tuning(t) != CLIFFT_OK) continue;
            float ms = 0.f;
            int r = time_of(t, &ms);
            if (r == CLIFFT_E_HIP) {
                rc = r;
                return;
            }
            if (r == CLIFFT_OK && ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
    };
    if (!rc) try_lds({{6, 6}, {7, 7}, {9, 9}, {10, 10}, {7, 8}, {8, 10}, {8, 12}});
    if (!rc) try_u32(&clifft_svm_batch_tuning::gwave_rank_max, {12, 13, 14, 15, 16, 18});
    if (!rc) try_u32(&clifft_svm_batch_tuning::global_threads, {512, 1024});
    if (!rc) try_u32(&clifft_svm_batch_tuning::wave_chunk, {2, 8, 16, 32});
    if (!rc) try_u32(&clifft_svm_batch_tuning::block_chunk, {2, 4, 8});
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    b->tuning = best;
    rc = batch_plan(b, ~0ull, st);
    if (rc) return rc;
    if (chosen) *chosen = best;
    return CLIFFT_OK;
}

int clifft_svm_batch_sample(clifft_svm_program* const* programs, const uint64_t* shots, const uint64_t* seeds,
                            uint32_t num_programs, uint32_t flags, const clifft_svm_records* outs, void* stream) {
    if (!programs || !shots || !seeds || !outs) return fail(CLIFFT_E_INVALID, "null argument");
    if (num_programs == 0) return CLIFFT_OK;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : library_stream();
    if (flags & CLIFFT_BATCH_PER_PROGRAM) {
        const uint32_t f = flags & ~CLIFFT_BATCH_PER_PROGRAM;
        for (uint32_t i = 0; i < num_programs; ++i) {
            int rc = run_shots(programs[i], shots[i], seeds[i], f, MODE_SAMPLE, &outs[i], nullptr, nullptr, st);
            if (rc) return rc;
        }
        if (f & CLIFFT_REC_DEVICE) HIP_TRY(hipStreamSynchronize(st));
        return CLIFFT_OK;
    }
    const bool dev_out = (flags & CLIFFT_REC_DEVICE) != 0;
    const bool packed = (flags & CLIFFT_REC_PACKED) != 0;
    std::vector<clifft_svm_records> douts(outs, outs + num_programs);
    void* staging = nullptr;
    std::vector<size_t> off;
    if (!dev_out) {
        // one device staging buffer for every program's rows
        size_t total = 0;
        for (uint32_t i = 0; i < num_programs; ++i) {
            clifft_svm_program* p = programs[i];
            if (!p) return fail(CLIFFT_E_INVALID, "null program in batch");
            const uint64_t n = shots[i];
            size_t parts[5] = {outs[i].measurements ? n * row_bytes(p->num_meas, packed) : 0,
                               outs[i].detectors ? n * row_bytes(p->num_det, packed) : 0,
                               outs[i].observables ? n * row_bytes(p->num_obs, packed) : 0,
                               outs[i].exp_vals ? n * p->num_ev * 8 : 0, outs[i].discarded ? n : 0};
            for (size_t part : parts) {
                off.push_back(total);
                total += align_up(part, 256);
            }
        }
        HIP_TRY(hipMalloc(&staging, std::max<size_t>(total, 256)));
        auto* s = static_cast<uint8_t*>(staging);
        for (uint32_t i = 0; i < num_programs; ++i) {
            clifft_svm_records& r = douts[i];
            r.measurements = outs[i].measurements ? s + off[5 * i + 0] : nullptr;
            r.detectors = outs[i].detectors ? s + off[5 * i + 1] : nullptr;
            r.observables = outs[i].observables ? s + off[5 * i + 2] : nullptr;
            r.exp_vals = outs[i].exp_vals ? reinterpret_cast<double*>(s + off[5 * i + 3]) : nullptr;
            r.discarded = outs[i].discarded ? s + off[5 * i + 4] : nullptr;
        }
    }
    clifft_svm_batch* b = nullptr;
    int rc = clifft_svm_batch_create(programs, shots, num_programs, flags, douts.data(), &b);
    if (!rc) rc = clifft_svm_batch_run(b, seeds, st);
    if (!rc && !dev_out) {
        for (uint32_t i = 0; i < num_programs && !rc; ++i) {
            clifft_svm_program* p = programs[i];
            const uint64_t n = shots[i];
            auto cp_out = [&](void* dst, const void* src, size_t bytes) {
                if (rc || !dst || !bytes) return;
                hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st);
                if (e != hipSuccess) rc = fail(CLIFFT_E_HIP, std::string("batch copy-out: ") + hipGetErrorString(e));
            };
            cp_out(outs[i].measurements, douts[i].measurements, n * row_bytes(p->num_meas, packed));
            cp_out(outs[i].detectors, douts[i].detectors, n * row_bytes(p->num_det, packed));
            cp_out(outs[i].observables, douts[i].observables, n * row_bytes(p->num_obs, packed));
            cp_out(outs[i].exp_vals, douts[i].exp_vals, n * p->num_ev * 8);
            cp_out(outs[i].discarded, douts[i].discarded, n);
        }
    }
    if (!rc) {
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fail(CLIFFT_E_HIP, std::string("batch: ") + hipGetErrorString(e));
    }
    std::string msg = g_last_error;
    if (b) clifft_svm_batch_destroy(b);
    if (staging) (void)hipFree(staging);
    if (rc) return fail(rc, msg);
    return CLIFFT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Adjoint gradients
// ---------------------------------------------------------------------------
struct clifft_svm_adjoint {
    clifft_svm_program* prog = nullptr;
    uint32_t num_params = 0;
    void* dev = nullptr;  // [slot_of (n_instr x i32) | psign (np x f64) | AdjArgs ring]
    int32_t* slot_of = nullptr;
    double* psign = nullptr;
    AdjArgs* args_dev = nullptr;
    uint32_t args_slot = 0;
    bool global_amps = false;
    uint32_t block = 256;
    size_t lds_bytes = 0;
    int per_cu = 1;
    void* gscratch = nullptr;
    size_t gscratch_bytes = 0;
    void* staging = nullptr;
    size_t staging_bytes = 0;
};

namespace {

constexpr uint32_t kAdjArgSlots = 64;

bool adjoint_op_ok(uint8_t op) {
    switch (op) {
        case CLIFFT_OP_FRAME_CNOT: case CLIFFT_OP_FRAME_CZ: case CLIFFT_OP_FRAME_H: case CLIFFT_OP_FRAME_S:
        case CLIFFT_OP_FRAME_S_DAG: case CLIFFT_OP_FRAME_SWAP: case CLIFFT_OP_ARRAY_CNOT: case CLIFFT_OP_ARRAY_CZ:
        case CLIFFT_OP_ARRAY_SWAP: case CLIFFT_OP_ARRAY_MULTI_CNOT: case CLIFFT_OP_ARRAY_MULTI_CZ:
        case CLIFFT_OP_ARRAY_H: case CLIFFT_OP_ARRAY_S: case CLIFFT_OP_ARRAY_S_DAG: case CLIFFT_OP_ARRAY_T:
        case CLIFFT_OP_ARRAY_T_DAG: case CLIFFT_OP_ARRAY_ROT: case CLIFFT_OP_ARRAY_U2: case CLIFFT_OP_ARRAY_U4:
        case CLIFFT_OP_EXPAND: case CLIFFT_OP_EXPAND_T: case CLIFFT_OP_EXPAND_T_DAG: case CLIFFT_OP_EXPAND_ROT:
        case CLIFFT_OP_EXP_VAL:
            return true;
        default:
            return false;
    }
}

int adjoint_ensure(void** ptr, size_t* cur, size_t bytes) {
    if (bytes <= *cur) return CLIFFT_OK;
    if (*ptr) {
        HIP_TRY(hipDeviceSynchronize());
        (void)hipFree(*ptr);
    }
    *ptr = nullptr;
    *cur = 0;
    HIP_TRY(hipMalloc(ptr, bytes));
    *cur = bytes;
    return CLIFFT_OK;
}

}  // namespace

extern "C" {

int clifft_svm_adjoint_create(clifft_svm_program* p, const uint32_t* param_instr, const int8_t* param_sign,
                              uint32_t num_params, clifft_svm_adjoint** out) {
    if (!out) return fail(CLIFFT_E_INVALID, "null output handle");
    *out = nullptr;
    if (!p || (num_params && (!param_instr || !param_sign))) return fail(CLIFFT_E_INVALID, "null argument");
    const size_t n = p->host_bc.size();
    for (size_t i = 0; i < n; ++i)
        if (!adjoint_op_ok(p->host_bc[i].opcode))
            return fail(CLIFFT_E_UNSUPPORTED, "adjoint: program has measurement, noise or record opcode at " +
                                                  std::to_string(i) + "; only noiseless measurement-free programs");
    std::vector<int32_t> slot_of(n, -1);
    std::vector<double> psign(std::max<uint32_t>(num_params, 1), 1.0);
    for (uint32_t i = 0; i < num_params; ++i) {
        const uint32_t pc = param_instr[i];
        if (pc >= n) return fail(CLIFFT_E_INVALID, "adjoint: param_instr out of range");
        const uint8_t op = p->host_bc[pc].opcode;
        if (op != CLIFFT_OP_ARRAY_ROT && op != CLIFFT_OP_EXPAND_ROT)
            return fail(CLIFFT_E_INVALID, "adjoint: parameter " + std::to_string(i) +
                                              " is not an ARRAY_ROT / EXPAND_ROT instruction");
        if (slot_of[pc] >= 0) return fail(CLIFFT_E_INVALID, "adjoint: two parameters on one instruction");
        if (param_sign[i] != 1 && param_sign[i] != -1) return fail(CLIFFT_E_INVALID, "adjoint: param_sign must be +-1");
        slot_of[pc] = (int32_t)i;
        psign[i] = (double)param_sign[i];
    }
    auto* a = new clifft_svm_adjoint();
    a->prog = p;
    a->num_params = num_params;
    const size_t o_ps = align_up(std::max<size_t>(n, 1) * 4, 256);
    const size_t o_args = o_ps + align_up(psign.size() * 8, 256);
    const size_t total = o_args + kAdjArgSlots * sizeof(AdjArgs);
    hipError_t e = hipMalloc(&a->dev, total);
    if (e == hipSuccess && n) e = hipMemcpy(a->dev, slot_of.data(), n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(static_cast<uint8_t*>(a->dev) + o_ps, psign.data(), psign.size() * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        clifft_svm_adjoint_destroy(a);
        return fail(CLIFFT_E_HIP, std::string("adjoint upload: ") + hipGetErrorString(e));
    }
    a->slot_of = static_cast<int32_t*>(a->dev);
    a->psign = reinterpret_cast<double*>(static_cast<uint8_t*>(a->dev) + o_ps);
    a->args_dev = reinterpret_cast<AdjArgs*>(static_cast<uint8_t*>(a->dev) + o_args);
    // team shape: phi and lambda in LDS up to rank 12, else global slices
    const uint32_t k = p->peak_rank;
    const size_t red_dens = 16 * 4 * sizeof(double) + kProbeBytes + align_up((size_t)std::max<uint32_t>(p->num_ev, 1) * 8, 16);
    a->global_amps = k > 12 || red_dens + ((size_t)32 << k) > 160 * 1024;
    a->block = k <= 6 ? 64 : (k <= 8 ? 256 : (k <= 10 ? 512 : 1024));
    if (a->global_amps) {  // CLIFFT_ADJOINT_TEAM_THREADS: 256, 512 or 1024
        const char* e = std::getenv("CLIFFT_ADJOINT_TEAM_THREADS");
        const int t = e ? std::atoi(e) : 0;
        if (t == 256 || t == 512 || t == 1024) a->block = (uint32_t)t;
    }
    a->lds_bytes = red_dens + (a->global_amps ? 0 : ((size_t)32 << k));
    if (a->lds_bytes > 160 * 1024) {
        clifft_svm_adjoint_destroy(a);
        return fail(CLIFFT_E_UNSUPPORTED, "adjoint: too many EXP_VAL probes for the LDS norm table");
    }
    e = adjoint_occupancy(a->global_amps, a->block, a->lds_bytes, &a->per_cu);
    if (e != hipSuccess) {
        clifft_svm_adjoint_destroy(a);
        return fail(CLIFFT_E_HIP, std::string("adjoint occupancy: ") + hipGetErrorString(e));
    }
    a->per_cu = std::max(a->per_cu, 1);
    *out = a;
    return CLIFFT_OK;
}

int clifft_svm_adjoint_destroy(clifft_svm_adjoint* a) {
    if (!a) return CLIFFT_OK;
    (void)hipDeviceSynchronize();
    if (a->dev) (void)hipFree(a->dev);
    if (a->gscratch) (void)hipFree(a->gscratch);
    if (a->staging) (void)hipFree(a->staging);
    delete a;
    return CLIFFT_OK;
}

int clifft_svm_adjoint_run(clifft_svm_adjoint* a, const double* thetas, uint64_t batch, const double* coeffs,
                           double* exp_vals, double* grads, uint32_t flags, void* stream) {
    if (!a) return fail(CLIFFT_E_INVALID, "null adjoint");
    if (batch == 0) return CLIFFT_OK;
    if (a->num_params && !thetas) return fail(CLIFFT_E_INVALID, "null thetas");
    if (grads && !coeffs && a->prog->num_ev) return fail(CLIFFT_E_INVALID, "gradients need coeffs");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : library_stream();
    clifft_svm_program* p = a->prog;
    const bool dev = (flags & CLIFFT_REC_DEVICE) != 0;
    const size_t np = a->num_params, nev = p->num_ev;
    const double *d_th = thetas, *d_c = coeffs;
    double *d_ev = exp_vals, *d_g = grads;
    if (!dev) {
        const size_t s_th = align_up(batch * np * 8 + 8, 256), s_c = align_up(nev * 8 + 8, 256);
        const size_t s_ev = align_up(batch * nev * 8 + 8, 256), s_g = align_up(batch * np * 8 + 8, 256);
        int rc = adjoint_ensure(&a->staging, &a->staging_bytes, s_th + s_c + s_ev + s_g);
        if (rc) return rc;
        auto* b = static_cast<uint8_t*>(a->staging);
        double* th = reinterpret_cast<double*>(b);
        double* c = reinterpret_cast<double*>(b + s_th);
        d_ev = exp_vals ? reinterpret_cast<double*>(b + s_th + s_c) : nullptr;
        d_g = grads ? reinterpret_cast<double*>(b + s_th + s_c + s_ev) : nullptr;
        if (np) HIP_TRY(hipMemcpyAsync(th, thetas, batch * np * 8, hipMemcpyHostToDevice, st));
        if (coeffs && nev) HIP_TRY(hipMemcpyAsync(c, coeffs, nev * 8, hipMemcpyHostToDevice, st));
        d_th = th;
        d_c = coeffs ? c : nullptr;
    }
    DeviceInfo& di = device_info();
    uint64_t grid = std::min<uint64_t>(batch, (uint64_t)a->per_cu * di.num_cus);
    if (a->global_amps) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 8ull << 30;
        const size_t per_team = (size_t)32 << p->peak_rank;
        const size_t budget = std::min<size_t>((free_b + a->gscratch_bytes) / 2, 64ull << 30);
        grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, budget / per_team));
        int rc = adjoint_ensure(&a->gscratch, &a->gscratch_bytes, grid * per_team);
        if (rc) return rc;
    }
    AdjArgs A{};
    A.thetas = d_th;
    A.psign = a->psign;
    A.slot_of = a->slot_of;
    A.coeffs = d_c;
    A.ev_out = d_ev;
    A.grad_out = d_g;
    A.gscratch = reinterpret_cast<dv2*>(a->gscratch);
    A.batch = batch;
    A.num_params = (uint32_t)np;
    A.amp_elems = 1u << p->peak_rank;
    AdjArgs* slot = a->args_dev + (a->args_slot++ % kAdjArgSlots);
    HIP_TRY(hipMemcpyAsync(slot, &A, sizeof(A), hipMemcpyHostToDevice, st));
    HIP_TRY(launch_adjoint(a->global_amps, (uint32_t)grid, a->block, a->lds_bytes, p->dp_dev, slot, st));
    if (!dev) {
        if (exp_vals && nev) HIP_TRY(hipMemcpyAsync(exp_vals, d_ev, batch * nev * 8, hipMemcpyDeviceToHost, st));
        if (grads && np) HIP_TRY(hipMemcpyAsync(grads, d_g, batch * np * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return CLIFFT_OK;
}

}  // extern "C"
