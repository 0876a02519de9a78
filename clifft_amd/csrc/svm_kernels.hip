// svm_kernels.hip -- team-interpreter kernels for gfx950 and their launchers.
#include <atomic>

#include "svm_internal.h"
#include "svm_interp.h"
#include "svm_lane.h"
#include "svm_adjoint.h"

namespace csvm {

// Frame kinds: one word in scalar registers (<= 64 qubits), or up to 64
// words spread over the wavefront's lanes (<= 4096 qubits).
enum : int { FRAME_W1 = 1, FRAME_LANES = 0 };

template <int FK>
struct FrameOf;
template <>
struct FrameOf<FRAME_W1> {
    typedef FrameReg<1> type;
};
template <>
struct FrameOf<FRAME_LANES> {
    typedef FrameLane type;
};

// LDS carve-up of one workgroup:
//   [teams_per_block x amp_elems x 16 B]  amplitudes (LDS tier only)
//   [512 B]                               team reduction scratch (16 x 4 f64)
//   [waves_per_block x rec_stride]        per-wavefront record copies
//
// The calling team runs shots s_first, s_first + s_step, ... < s_end of the
// program; `gamp` is the team's global amplitude slice (GA only) and `seed`
// the program's sampling seed.
template <int FK, bool BT, bool GA>
__device__ __forceinline__ void shot_range(cptr<DevProgram> Pg, cptr<ShotArgs> Ag, uint64_t s_first, uint64_t s_end,
                                           uint64_t s_step, dv2* gamp, uint64_t seed) {
    typedef typename FrameOf<FK>::type F;
    typedef typename std::conditional<GA, glb_d2, lds_d2>::type AmpT;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    cptr<ShotArgs> A = opaque(Ag);
    const uint32_t waves_block = blockDim.x >> 6;
    const uint32_t wave_in_block = uni(threadIdx.x >> 6);
    const uint32_t wl = threadIdx.x & 63u;
    Interp<F, BT, AmpT> S;
    uint32_t teams_per_block, team_in_block;
    if (BT) {
        team_in_block = 0;
        teams_per_block = 1;
        S.g.lane = threadIdx.x;
        S.g.size = blockDim.x;
        S.g.wave = wave_in_block;
        S.g.nwaves = waves_block;
    } else {
        team_in_block = wave_in_block;
        teams_per_block = waves_block;
        S.g.lane = wl;
        S.g.size = 64;
        S.g.wave = 0;
        S.g.nwaves = 1;
    }
    const uint32_t amp_elems = A->amp_elems;
    const uint32_t rec_stride = A->rec_stride;
    unsigned char* p = smem;
    unsigned char* lds_amp = p;
    if (!GA) p += (size_t)teams_per_block * amp_elems * sizeof(dv2);
    S.red = (lds_f64*)(p);
    S.pend = S.red + 16 * 4 + 16 * wave_in_block;
    p += kTeamScratchBytes;
    unsigned char* wblock = p + (size_t)wave_in_block * rec_stride;
    unsigned char* rec = wblock;
    typedef __attribute__((address_space(3))) uint8_t lds_u8;
    // MODE_KFAULT: the last 5 * kMaxForcedFaults words of the wave's block
    __attribute__((address_space(3))) uint32_t* kf_block =
        (__attribute__((address_space(3))) uint32_t*)(wblock + rec_stride - 20u * kMaxForcedFaults);
    {
        cptr<DevProgram> P = opaque(Pg);
        S.meas = (lds_u8*)(rec);
        S.det = S.meas + P->total_meas_slots;
        S.obs = S.det + P->num_det;
    }
    if (GA)
        S.amp = (AmpT*)(gamp);
    else
        S.amp = (AmpT*)((dv2*)lds_amp + (size_t)team_in_block * amp_elems);
    // fused permutation/phase runs: teams on global amplitudes with a second
    // slice; the opcode list (one per team) follows the wavefronts' record
    // blocks
    S.amp_alt = nullptr;
    S.amp_alt2 = nullptr;
    S.mops = nullptr;
    S.nmono = 0;
    S.cands = 0;
    S.pend_defer = 0;
    if (GA && A->amp_slices >= 2) {  // wavefront teams: one list per wavefront
        S.pend_defer = A->pend_defer ? 1u : 0u;
        S.amp_alt = (AmpT*)(gamp + amp_elems);
        if (A->amp_slices == 3) S.amp_alt2 = (AmpT*)(gamp + 2 * (size_t)amp_elems);
        S.mops = (lds_mono*)(p + (size_t)waves_block * rec_stride + (BT ? 0u : wave_in_block * kMonoBytes));
    }

    for (uint64_t s = s_first; s < s_end; s += s_step) {
        A = opaque(Ag);
        cptr<DevProgram> P = opaque(Pg);
        const uint64_t shot = A->shot_begin + s;
        const uint32_t mode = A->mode;
        // ---- per-shot reset (svm_state.cc:220-267) ----
        for (uint32_t i = wl; i < rec_stride / 4; i += 64) reinterpret_cast<uint32_t*>(wblock)[i] = 0u;
        wave_fence();
        S.f.init(wblock);
        S.kf = 0;
        S.npend = 0;
        S.nmono = 0;  // a shot that ended early leaves no deferred run to the next one
        S.cands = 0;
        S.pend_mask = 0;
        S.dust = 0;
        S.discarded = 0;
        S.reachable = 1;
        S.logp = 0.0;
        S.forced = nullptr;
        if (mode == MODE_PRESET) {
            cptr<PresetState> ps = cp(A->preset);
            S.g_re = ps->gamma[0];
            S.g_im = ps->gamma[1];
            S.rng.s0 = ps->rng[0];
            S.rng.s1 = ps->rng[1];
            S.rng.s2 = ps->rng[2];
            S.rng.s3 = ps->rng[3];
            S.k = ps->active_k;
            S.next_noise = ps->next_noise_idx;
            S.logp = ps->forced_log_probability;
            S.reachable = ps->forced_reachable;
            for (uint32_t w = 0; w < P->words; ++w) S.f.set_words(w, ps->px[w], ps->pz[w]);
            uint32_t n = 1u << P->peak_rank;
            for (uint32_t i = S.g.lane; i < n; i += S.g.size) S.amp[i] = cp(A->preset_amps)[i];
            uint32_t nrec = P->total_meas_slots + P->num_det + P->num_obs;
            for (uint32_t i = wl; i < nrec; i += 64) rec[i] = (uint8_t)load_u8(cp(A->preset_records), i);
            wave_fence();
            S.forced = cp(A->forced_records);
        } else {
            S.g_re = 1.0;
            S.g_im = 0.0;
            S.k = 0;
            S.rng = rng_for_shot(seed, shot);
            if (S.g.lane == 0) S.amp[0] = mk(1.0, 0.0);
            if (mode == MODE_KFAULT) {  // svm.cc:602-606: no gap draw, fault set first
                S.kfault_prepare(cp(A->kf), kf_block);
            } else {
                S.next_noise = 0;
                S.draw_next_noise(P);
            }
            if (mode == MODE_FORCED) S.forced = cp(A->forced_records) + s * P->num_meas;
        }
        team_sync<BT>();

        double* ev_row = nullptr;
        if (mode == MODE_PRESET) {
            ev_row = A->preset_ev;
        } else if (A->out_ev) {
            ev_row = A->out_ev + s * P->num_ev;
            if (S.g.lane == 0)
                for (uint32_t i = 0; i < P->num_ev; ++i) ev_row[i] = 0.0;
        }

        S.run(Pg, ev_row, mode != MODE_PRESET);

        A = opaque(Ag);
        P = opaque(Pg);
        if (mode == MODE_PRESET) {
            uint32_t n = 1u << P->peak_rank;
            for (uint32_t i = S.g.lane; i < n; i += S.g.size) A->preset_amps[i] = S.amp[i];
        }
        // ---- outputs (wave 0 of the team) ----
        if (S.g.wave == 0) {
            if (mode == MODE_PRESET) {
                PresetState* ps = A->preset;
                uint32_t nrec = P->total_meas_slots + P->num_det + P->num_obs;
                for (uint32_t i = wl; i < nrec; i += 64) A->preset_records[i] = rec[i];
                uint32_t W = P->words;
                uint64_t fx[16], fz[16];
                for (uint32_t w = 0; w < 16; ++w) {
                    fx[w] = w < W ? S.f.wx(w) : 0;
                    fz[w] = w < W ? S.f.wz(w) : 0;
                }
                if (wl == 0) {
                    ps->gamma[0] = S.g_re;
                    ps->gamma[1] = S.g_im;
                    ps->rng[0] = S.rng.s0;
                    ps->rng[1] = S.rng.s1;
                    ps->rng[2] = S.rng.s2;
                    ps->rng[3] = S.rng.s3;
                    ps->active_k = S.k;
                    ps->discarded = S.discarded;
                    ps->next_noise_idx = S.next_noise;
                    ps->forced_reachable = S.reachable;
                    ps->forced_log_probability = S.logp;
                    ps->dust_clamps += S.dust;
                    for (uint32_t w = 0; w < W; ++w) {
                        ps->px[w] = fx[w];
                        ps->pz[w] = fz[w];
                    }
                }
            } else {
                const bool packed = (A->flags & CLIFFT_REC_PACKED) != 0;
                const uint32_t n_exp = P->n_expected_obs;
                cptr<uint8_t> expo = cp(P->expected_obs);
                auto emit = [&](uint8_t* out, const lds_u8* src, uint32_t n, bool xor_expected) {
                    if (!out || n == 0) return;
                    if (packed) {
                        uint32_t nb = (n + 7) / 8;
                        for (uint32_t j = wl; j < nb; j += 64) {
                            uint32_t byte = 0;
                            for (uint32_t b = 0; b < 8; ++b) {
                                uint32_t i = 8 * j + b;
                                if (i < n) {
                                    uint32_t v = src[i];
                                    if (xor_expected && i < n_exp) v ^= load_u8(expo, i) ? 1u : 0u;
                                    byte |= (v & 1u) << b;
                                }
                            }
                            out[s * nb + j] = (uint8_t)byte;
                        }
                    } else {
                        for (uint32_t i = wl; i < n; i += 64) {
                            uint32_t v = src[i];
                            if (xor_expected && i < n_exp) v ^= load_u8(expo, i) ? 1u : 0u;
                            out[s * n + i] = (uint8_t)v;
                        }
                    }
                };
                emit(A->out_meas, S.meas, P->num_meas, false);
                emit(A->out_det, S.det, P->num_det, false);
                emit(A->out_obs, S.obs, P->num_obs, true);
                if (wl == 0) {
                    if (A->out_disc) A->out_disc[s] = (uint8_t)S.discarded;
                    if (mode == MODE_FORCED && A->out_logp) A->out_logp[s] = S.reachable ? S.logp : kUnreachable;
                    if (S.dust && A->dust_counter) atomicAdd(A->dust_counter, (unsigned long long)S.dust);
                }
            }
        }
        team_sync<BT>();
    }
}

// Lane tier: one shot per lane.  Wavefront `wave` (0 <= wave < waves in the
// workgroup) of this workgroup owns the LDS slice [wave * wave_stride, +
// lane_wave_bytes()) (wave_stride = lane_wave_bytes() of the program, or
// the largest of a batched launch's programs); the calling wavefront runs
// shot groups g = g_first, g_first + g_step, ..., lane l taking shot
// s_lo + 64 g + l while that is below s_hi.  FW is the frame word
// (uint32_t for programs of at most 32 qubits).
template <bool kFrameFirst, class FW>
__device__ __forceinline__ void lane_range(cptr<DevProgram> Pg, cptr<ShotArgs> Ag, uint64_t g_first, uint64_t g_step,
                                           uint64_t s_lo, uint64_t s_hi, uint64_t seed, uint32_t wave_stride) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = uni(threadIdx.x >> 6);
    cptr<DevProgram> P = opaque(Pg);
    const uint32_t peak = P->peak_rank, slots = P->total_meas_slots, ndet = P->num_det, nobs = P->num_obs;
    const uint32_t wm = lane_words(slots), wd = lane_words(ndet), nwords = wm + wd + lane_words(nobs);
    unsigned char* base = smem + (size_t)wave * wave_stride;
    LaneInterp<FW> S;
    S.amp = (lds_col_d2*)(base) + lane;
    lds_col_u32* recs = (lds_col_u32*)(base + (64u * 16u << peak)) + lane;
    S.rm = recs;
    S.rd = recs + (wm << 6);
    S.ro = S.rd + (wd << 6);
    for (uint64_t g = g_first;; g += g_step) {
        const uint64_t s0 = s_lo + g * 64u;
        if (s0 >= s_hi) break;
        const uint64_t s = s0 + lane;  // shot index within the launch
        if (s < s_hi) {
            cptr<ShotArgs> A = opaque(Ag);
            P = opaque(Pg);
            const uint32_t mode = A->mode;
            for (uint32_t w = 0; w < nwords; ++w) recs[w << 6] = 0u;
            S.px = S.pz = 0;
            S.g_mag = 1.0;
            S.k = 0;
            S.dust = 0;
            S.discarded = 0;
            S.reachable = 1;
            S.logp = 0.0;
            S.forced = nullptr;
            S.rng = rng_for_shot(seed, A->shot_begin + s);
            S.S(0, mk(1.0, 0.0));
            S.next_noise = 0;
            S.draw_next_noise(P, 0.0);
            if (mode == MODE_FORCED) S.forced = A->forced_records + s * P->num_meas;
            const uint32_t nev = P->num_ev;
            double* ev_row = A->out_ev ? A->out_ev + s * nev : nullptr;
            if (ev_row)
                for (uint32_t i = 0; i < nev; ++i) ev_row[i] = 0.0;

            S.template run<kFrameFirst>(Pg, ev_row);

            A = opaque(Ag);
            P = opaque(Pg);
            const bool packed = (A->flags & CLIFFT_REC_PACKED) != 0;
            const uint32_t n_exp = P->n_expected_obs;
            cptr<uint8_t> expo = cp(P->expected_obs);
            // record bits -> rows (bytes of 0/1, or little-endian bit-packed)
            auto emit = [&](uint8_t* out, const lds_col_u32* src, uint32_t n, bool xor_expected) {
                if (!out || n == 0) return;
                if (packed) {
                    const uint32_t nb = (n + 7u) / 8u;
                    for (uint32_t j = 0; j < nb; ++j) {
                        uint32_t byte = (src[(j >> 2) << 6] >> ((j & 3u) * 8u)) & 0xFFu;
                        if (xor_expected) {
                            uint32_t em = 0;
                            for (uint32_t b = 0; b < 8; ++b)
                                if (8 * j + b < n_exp) em |= (load_u8(expo, 8 * j + b) ? 1u : 0u) << b;
                            byte ^= em;
                        }
                        if (8 * j + 8 > n) byte &= (1u << (n - 8 * j)) - 1u;
                        out[s * nb + j] = (uint8_t)byte;
                    }
                } else {
                    for (uint32_t i = 0; i < n; ++i) {
                        uint32_t v = LaneInterp<FW>::rget(src, i);
                        if (xor_expected && i < n_exp) v ^= load_u8(expo, i) ? 1u : 0u;
                        out[s * n + i] = (uint8_t)v;
                    }
                }
            };
            emit(A->out_meas, S.rm, P->num_meas, false);
            emit(A->out_det, S.rd, ndet, false);
            emit(A->out_obs, S.ro, nobs, true);
            if (A->out_disc) A->out_disc[s] = (uint8_t)S.discarded;
            if (mode == MODE_FORCED && A->out_logp) A->out_logp[s] = S.reachable ? S.logp : kUnreachable;
            if (S.dust && A->dust_counter) atomicAdd(A->dust_counter, (unsigned long long)S.dust);
        }
    }
}

// MINB = resident workgroups of 256 the register budget is sized for: 4
// (128 VGPRs, no spills) for programs whose LDS columns limit occupancy
// anyway, 6 (80 VGPRs, a few spills) when LDS would allow more waves.
// FW: 32-bit frame words for programs of at most 32 qubits, else 64-bit.
template <int MINB, class FW>
__global__ void __launch_bounds__(256, MINB) lane_kernel(const DevProgram* P, const ShotArgs* A) {
    cptr<ShotArgs> Ag = cp(A);
    const uint64_t wpb = blockDim.x >> 6;
    const uint64_t wave_global = (uint64_t)blockIdx.x * wpb + uni(threadIdx.x >> 6);
    cptr<DevProgram> Pc = cp(P);
    const uint32_t stride = lane_wave_bytes(Pc->peak_rank, Pc->total_meas_slots, Pc->num_det, Pc->num_obs);
    lane_range<MINB == 4, FW>(Pc, Ag, wave_global, (uint64_t)gridDim.x * wpb, 0, Ag->shot_count, Ag->seed, stride);
}

// One program: teams stride over the launch's shots.
template <int FK, bool BT, bool GA>
__global__ void __launch_bounds__(BT ? 1024 : 256, 4) shot_kernel(const DevProgram* P, const ShotArgs* A) {
    cptr<ShotArgs> Ag = cp(A);
    const uint32_t tpb = BT ? 1u : (blockDim.x >> 6);
    const uint32_t team_in_block = BT ? 0u : uni(threadIdx.x >> 6);
    const uint64_t team_global = (uint64_t)blockIdx.x * tpb + team_in_block;
    const uint64_t nteams = (uint64_t)gridDim.x * tpb;
    dv2* gamp = GA ? Ag->gscratch + team_global * Ag->amp_elems * Ag->amp_slices : nullptr;
    shot_range<FK, BT, GA>(cp(P), Ag, team_global, Ag->shot_count, nteams, gamp, Ag->seed);
}

// ---------------------------------------------------------------------------
// Batched execution: many programs per launch.  A batch is split into
// launch groups, one per kernel shape (the shapes of the per-program
// launches); each group's launch is persistent and its teams pop work items
// (a shot range of one program, in decreasing order of estimated cost) from
// the group's device queue until it is empty.  The per-program ShotArgs of a
// group carry the group's largest amplitude count and record stride, so
// every team of a workgroup carves the group's LDS the same way whichever
// program it runs.
// ---------------------------------------------------------------------------

// One wavefront pops the next item index (lane 0's atomic, broadcast).
CSVM_INLINE uint32_t pop_wave(uint32_t* counter) {
    uint32_t v = 0;
    if ((threadIdx.x & 63u) == 0) v = atomicAdd(counter, 1u);
    return uni((uint32_t)__shfl((int)v, 0, 64));
}

template <int MINB, class FW>
__global__ void __launch_bounds__(256, MINB) batch_lane_kernel(const BatchArgs* B) {
    cptr<BatchArgs> Bg = cp(B);
    for (;;) {
        const uint32_t it = pop_wave(Bg->counter);
        cptr<BatchArgs> Bc = opaque(Bg);
        if (it >= Bc->n_items) break;
        cptr<BatchItem> I = cp(Bc->items) + it;
        const uint32_t pid = I->prog;
        lane_range<MINB == 4, FW>(cp(cp(Bc->progs)[pid]), cp(Bc->args) + pid, 0, 1, I->shot_lo, I->shot_hi,
                                  cp(Bc->seeds)[pid], Bc->wave_stride);
    }
}

template <int FK, bool BT, bool GA>
__global__ void __launch_bounds__(BT ? 1024 : 256, 4) batch_team_kernel(const BatchArgs* B) {
    __shared__ uint32_t head;
    cptr<BatchArgs> Bg = cp(B);
    const uint32_t tpb = BT ? 1u : (blockDim.x >> 6);
    const uint32_t team_in_block = BT ? 0u : uni(threadIdx.x >> 6);
    const uint64_t team_global = (uint64_t)blockIdx.x * tpb + team_in_block;
    dv2* gamp = GA ? Bg->gscratch + team_global * Bg->team_elems : nullptr;
    for (;;) {
        uint32_t it;
        if (BT) {
            if (threadIdx.x == 0) head = atomicAdd(Bg->counter, 1u);
            __syncthreads();
            it = uni(head);
            __syncthreads();
        } else {
            it = pop_wave(Bg->counter);
        }
        cptr<BatchArgs> Bc = opaque(Bg);
        if (it >= Bc->n_items) break;
        cptr<BatchItem> I = cp(Bc->items) + it;
        const uint32_t pid = I->prog;
        shot_range<FK, BT, GA>(cp(cp(Bc->progs)[pid]), cp(Bc->args) + pid, I->shot_lo, I->shot_hi, 1, gamp,
                               cp(Bc->seeds)[pid]);
    }
}

// Expectation values and adjoint gradients: one workgroup team per theta
// vector, phi and lambda in LDS (LDS tier) or in the workgroup's slice of
// global scratch; LDS also holds the reduction scratch and the probe norms.
template <bool GA>
__global__ void __launch_bounds__(1024) adjoint_kernel(const DevProgram* P, const AdjArgs* A) {
    typedef typename std::conditional<GA, glb_d2, lds_d2>::type AmpT;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    cptr<AdjArgs> Ag = cp(A);
    AdjInterp<AmpT> S;
    S.g.lane = threadIdx.x;
    S.g.size = blockDim.x;
    S.g.wave = uni(threadIdx.x >> 6);
    S.g.nwaves = blockDim.x >> 6;
    const uint32_t n = Ag->amp_elems;
    unsigned char* p = smem;
    if (GA) {
        S.phi = (AmpT*)(Ag->gscratch + (size_t)blockIdx.x * 2u * n);
    } else {
        S.phi = (AmpT*)(p);
        p += 2u * n * sizeof(dv2);
    }
    S.lam = S.phi + n;
    S.red = (lds_f64*)(p);
    p += 16 * 4 * sizeof(double);
    S.probes = (lds_probe*)(p);
    p += kProbeBytes;
    lds_f64* dens = (lds_f64*)(p);
    const uint64_t batch = Ag->batch;
    for (uint64_t b = blockIdx.x; b < batch; b += gridDim.x) {
        cptr<AdjArgs> A2 = opaque(Ag);
        cptr<DevProgram> P0 = cp(P);
        const uint32_t np = A2->num_params, nev = P0->num_ev;
        const double* th = A2->thetas + b * np;
        double* ev_row = A2->ev_out ? A2->ev_out + b * nev : nullptr;
        double* grad_row = A2->grad_out ? A2->grad_out + b * np : nullptr;
        if (grad_row && S.g.lane == 0)
            for (uint32_t i = 0; i < np; ++i) grad_row[i] = 0.0;
        S.forward(cp(P), Ag, th, ev_row, dens);
        if (grad_row) S.backward(cp(P), Ag, th, dens, grad_row);
        team_sync<true>();
    }
}

hipError_t adjoint_occupancy(bool global_amps, uint32_t block, size_t lds_bytes, int* blocks_per_cu) {
    const void* fn = global_amps ? reinterpret_cast<const void*>(adjoint_kernel<true>)
                                 : reinterpret_cast<const void*>(adjoint_kernel<false>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, (int)block, lds_bytes);
}

hipError_t launch_adjoint(bool global_amps, uint32_t grid, uint32_t block, size_t lds_bytes, const DevProgram* P,
                          const AdjArgs* A, hipStream_t st) {
    if (global_amps)
        hipLaunchKernelGGL(adjoint_kernel<true>, dim3(grid), dim3(block), lds_bytes, st, P, A);
    else
        hipLaunchKernelGGL(adjoint_kernel<false>, dim3(grid), dim3(block), lds_bytes, st, P, A);
    return hipGetLastError();
}

__global__ void clear_u32_kernel(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Launch dispatch over the template parameters.
// ---------------------------------------------------------------------------
template <int FK, bool BT, bool GA>
static hipError_t launch_one(const LaunchShape& L, const DevProgram* P, const ShotArgs* A, hipStream_t st) {
    auto kern = shot_kernel<FK, BT, GA>;
    static std::atomic<bool> attr_done{false};
    if (!attr_done.load(std::memory_order_acquire)) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_done.store(true, std::memory_order_release);
    }
    hipLaunchKernelGGL(kern, dim3(L.grid), dim3(L.block), L.lds_bytes, st, P, A);
    return hipGetLastError();
}

template <class FW>
static const void* lane_fn_fw(const LaunchShape& L) {
    return L.lane_dense ? reinterpret_cast<const void*>(lane_kernel<6, FW>)
                        : reinterpret_cast<const void*>(lane_kernel<4, FW>);
}

static const void* lane_fn(const LaunchShape& L) {
    return L.lane_narrow ? lane_fn_fw<uint32_t>(L) : lane_fn_fw<uint64_t>(L);
}

static hipError_t launch_lane(const LaunchShape& L, const DevProgram* P, const ShotArgs* A, hipStream_t st) {
    static std::atomic<bool> attr_done[4] = {{false}, {false}, {false}, {false}};
    const int v = (L.lane_dense ? 1 : 0) + (L.lane_narrow ? 2 : 0);
    if (!attr_done[v].load(std::memory_order_acquire)) {
        hipError_t e = hipFuncSetAttribute(lane_fn(L), hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_done[v].store(true, std::memory_order_release);
    }
    switch (v) {
        case 0: hipLaunchKernelGGL((lane_kernel<4, uint64_t>), dim3(L.grid), dim3(L.block), L.lds_bytes, st, P, A); break;
        case 1: hipLaunchKernelGGL((lane_kernel<6, uint64_t>), dim3(L.grid), dim3(L.block), L.lds_bytes, st, P, A); break;
        case 2: hipLaunchKernelGGL((lane_kernel<4, uint32_t>), dim3(L.grid), dim3(L.block), L.lds_bytes, st, P, A); break;
        default: hipLaunchKernelGGL((lane_kernel<6, uint32_t>), dim3(L.grid), dim3(L.block), L.lds_bytes, st, P, A); break;
    }
    return hipGetLastError();
}

template <int FK>
static hipError_t launch_fk(const LaunchShape& L, const DevProgram* P, const ShotArgs* A, hipStream_t st) {
    if (L.block_team)
        return L.global_amps ? launch_one<FK, true, true>(L, P, A, st) : launch_one<FK, true, false>(L, P, A, st);
    return L.global_amps ? launch_one<FK, false, true>(L, P, A, st) : launch_one<FK, false, false>(L, P, A, st);
}

template <int FK>
static const void* shot_fn_fk(const LaunchShape& L) {
    if (L.block_team)
        return L.global_amps ? reinterpret_cast<const void*>(shot_kernel<FK, true, true>)
                             : reinterpret_cast<const void*>(shot_kernel<FK, true, false>);
    return L.global_amps ? reinterpret_cast<const void*>(shot_kernel<FK, false, true>)
                         : reinterpret_cast<const void*>(shot_kernel<FK, false, false>);
}

hipError_t shot_occupancy(const LaunchShape& L, int* blocks_per_cu) {
    const void* fn = L.lane ? lane_fn(L)
                            : (L.nw == 1 ? shot_fn_fk<FRAME_W1>(L) : shot_fn_fk<FRAME_LANES>(L));
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, (int)L.block, L.lds_bytes);
}

hipError_t launch_shots(const LaunchShape& L, const DevProgram* P, const ShotArgs* A, hipStream_t st) {
    if (L.lane) return launch_lane(L, P, A, st);
    switch (L.nw) {
        case 1: return launch_fk<FRAME_W1>(L, P, A, st);
        default: return launch_fk<FRAME_LANES>(L, P, A, st);
    }
}

template <int FK>
static const void* batch_team_fn(bool block_team, bool global_amps) {
    if (block_team)
        return global_amps ? reinterpret_cast<const void*>(batch_team_kernel<FK, true, true>)
                           : reinterpret_cast<const void*>(batch_team_kernel<FK, true, false>);
    return global_amps ? reinterpret_cast<const void*>(batch_team_kernel<FK, false, true>)
                       : reinterpret_cast<const void*>(batch_team_kernel<FK, false, false>);
}

static const void* batch_fn(const LaunchShape& L) {
    if (L.lane) {
        if (L.lane_narrow)
            return L.lane_dense ? reinterpret_cast<const void*>(batch_lane_kernel<6, uint32_t>)
                                : reinterpret_cast<const void*>(batch_lane_kernel<4, uint32_t>);
        return L.lane_dense ? reinterpret_cast<const void*>(batch_lane_kernel<6, uint64_t>)
                            : reinterpret_cast<const void*>(batch_lane_kernel<4, uint64_t>);
    }
    return L.nw == 1 ? batch_team_fn<FRAME_W1>(L.block_team, L.global_amps)
                     : batch_team_fn<FRAME_LANES>(L.block_team, L.global_amps);
}

hipError_t batch_occupancy(const LaunchShape& L, int* blocks_per_cu) {
    const void* fn = batch_fn(L);
    // dynamic LDS may use what the static queue head leaves
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kBatchMaxDynLds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, (int)L.block, L.lds_bytes);
}

hipError_t clear_counters(uint32_t* counters, uint32_t n, hipStream_t st) {
    hipLaunchKernelGGL(clear_u32_kernel, dim3((n + 63) / 64), dim3(64), 0, st, counters, n);
    return hipGetLastError();
}

// The dynamic-LDS attribute was set by batch_occupancy when the plan was
// made, so a launch is a stream operation only (graph-capturable).
hipError_t launch_batch_group(const LaunchShape& L, const BatchArgs* B, hipStream_t st) {
    const void* fn = batch_fn(L);
    void* args[] = {&B};
    return hipLaunchKernel(fn, dim3(L.grid), dim3(L.block), args, L.lds_bytes, st);
}

}  // namespace csvm
