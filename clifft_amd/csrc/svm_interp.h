// svm_interp.h -- the per-shot bytecode interpreter run by one team.
//
// Every handler restates one reference opcode handler (citation on each,
// paths relative to the reference repository).  Frame, gamma, rank and PRNG
// updates are wavefront-uniform and done redundantly by every lane; the
// amplitude work is delegated to the sweeps in svm_sweeps.h.
#pragma once

#include "svm_sweeps.h"

namespace csvm {

// Pauli frame of NW words held in scalar registers (used with NW = 1).
template <int NW>
struct FrameReg {
    uint64_t x[NW];
    uint64_t z[NW];
    CSVM_INLINE void init(void*) {
#pragma unroll
        for (int i = 0; i < NW; ++i) x[i] = z[i] = 0;
    }
    CSVM_INLINE static uint64_t sel(const uint64_t (&m)[NW], uint32_t w) {
        if (NW == 1) return m[0];
        return w == 0 ? m[0] : m[NW - 1];
    }
    CSVM_INLINE static void xorw(uint64_t (&m)[NW], uint32_t w, uint64_t v) {
        if (NW == 1 || w == 0)
            m[0] ^= v;
        else
            m[NW - 1] ^= v;
    }
    CSVM_INLINE uint32_t gx(uint32_t q) const { return (uint32_t)(sel(x, q >> 6) >> (q & 63)) & 1u; }
    CSVM_INLINE uint32_t gz(uint32_t q) const { return (uint32_t)(sel(z, q >> 6) >> (q & 63)) & 1u; }
    CSVM_INLINE void fx(uint32_t q, uint32_t b) { xorw(x, q >> 6, (uint64_t)(b & 1u) << (q & 63)); }
    CSVM_INLINE void fz(uint32_t q, uint32_t b) { xorw(z, q >> 6, (uint64_t)(b & 1u) << (q & 63)); }
    CSVM_INLINE uint64_t wx(uint32_t w) const { return sel(x, w); }
    CSVM_INLINE uint64_t wz(uint32_t w) const { return sel(z, w); }
    CSVM_INLINE void xor_words(uint32_t w, uint64_t ex, uint64_t ez) {
        xorw(x, w, ex);
        xorw(z, w, ez);
    }
    CSVM_INLINE void set_words(uint32_t w, uint64_t vx, uint64_t vz) {
        xorw(x, w, sel(x, w) ^ vx);
        xorw(z, w, sel(z, w) ^ vz);
    }
};

// Pauli frame of up to 64 words (4096 qubits) distributed over the
// wavefront: lane w holds word w of p_x and p_z in vector registers.  A
// uniform word is read with v_readlane (uniform lane index); a word update is
// a masked XOR in the owning lane.  Every wavefront of a team keeps its own
// identical copy, so no memory or barrier is involved.
struct FrameLane {
    uint64_t x, z;
    CSVM_INLINE static uint32_t my_lane() { return threadIdx.x & 63u; }
    CSVM_INLINE static uint64_t readlane64(uint64_t v, uint32_t w) {
        uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, w);
        uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), w);
        return ((uint64_t)hi << 32) | lo;
    }
    CSVM_INLINE void init(void*) { x = z = 0; }
    CSVM_INLINE uint64_t wx(uint32_t w) const { return readlane64(x, w); }
    CSVM_INLINE uint64_t wz(uint32_t w) const { return readlane64(z, w); }
    CSVM_INLINE uint32_t gx(uint32_t q) const { return (uint32_t)(wx(q >> 6) >> (q & 63)) & 1u; }
    CSVM_INLINE uint32_t gz(uint32_t q) const { return (uint32_t)(wz(q >> 6) >> (q & 63)) & 1u; }
    CSVM_INLINE void xor_words(uint32_t w, uint64_t ex, uint64_t ez) {
        bool mine = my_lane() == w;
        x ^= mine ? ex : 0ull;
        z ^= mine ? ez : 0ull;
    }
    CSVM_INLINE void fx(uint32_t q, uint32_t b) { xor_words(q >> 6, (uint64_t)(b & 1u) << (q & 63), 0); }
    CSVM_INLINE void fz(uint32_t q, uint32_t b) { xor_words(q >> 6, 0, (uint64_t)(b & 1u) << (q & 63)); }
    CSVM_INLINE void set_words(uint32_t w, uint64_t vx, uint64_t vz) {
        bool mine = my_lane() == w;
        x = mine ? vx : x;
        z = mine ? vz : z;
    }
};

// Record-byte fence: lane 0 of a wavefront writes a record byte that every
// lane then reads back.
CSVM_INLINE void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class F, bool BT, class AmpT>
struct Interp {
    // amplitudes in global memory: the only teams with deferred runs
    static constexpr bool kGlobalAmps = std::is_same<AmpT, glb_d2>::value;
    TeamGeom g;
    AmpT* amp;
    lds_f64* red;
    lds_f64* pend;       // this wavefront's pending-expansion phases
    uint32_t npend;      // expansions recorded but not yet written
    uint32_t pend_mask;  // which of them carry a phase
    // 1: pending expansions stay pending under deferred permutation/phase
    // runs and are read through by the next active measurement's pass
    // (teams with a second slice only)
    uint32_t pend_defer;
    __attribute__((address_space(3))) uint8_t* meas;
    __attribute__((address_space(3))) uint8_t* det;
    __attribute__((address_space(3))) uint8_t* obs;
    F f;
    double g_re, g_im;
    uint32_t k;
    Rng rng;
    uint32_t next_noise;
    uint32_t discarded;
    uint32_t reachable;
    double logp;
    uint32_t dust;
    cptr<uint8_t> forced;
    // MODE_KFAULT: the shot's forced noise sites then readout entries, each
    // ascending, in this wavefront's LDS block (identical in every wave)
    uint32_t kf;
    const __attribute__((address_space(3))) uint32_t* kf_list;
    uint32_t kf_nn, kf_npos, kf_nr, kf_rpos;
    // Global-tier workgroup teams: the team's second amplitude slice and the
    // LDS list of deferred permutation/phase opcodes (null: no fusion).
    AmpT* amp_alt;
    lds_mono* mops;
    uint32_t nmono;
    // A third slice (null: none): the run before an active measurement then
    // writes both folded states, amp_alt (outcome 0) and amp_alt2 (outcome
    // 1), and `cands` is set until the measurement takes one of them.
    AmpT* amp_alt2;
    uint32_t cands;

    // ----------------------------------------------------------------- gamma
    CSVM_INLINE void mul_phase(double re, double im) {
        double a = g_re, b = g_im;
        g_re = a * re - b * im;
        g_im = a * im + b * re;
    }
    CSVM_INLINE void negate_gamma() {
        g_re = -g_re;
        g_im = -g_im;
    }
    // Write the oldest n pending expansions, kMaxFusedExpand per pass; the
    // rest stay pending (their phases move to the front of the list).
    CSVM_INLINE void flush_expand_oldest(uint32_t n) {
        if constexpr (!kGlobalAmps) {  // LDS amplitudes: n is always all of them (at most kMaxFusedExpand)
            sweep_expand_multi<BT>(amp, k - npend, npend, pend, pend_mask, g);
            npend = 0;
            pend_mask = 0;
            return;
        }
        while (n > 0) {
            const uint32_t m = n < kMaxFusedExpand ? n : kMaxFusedExpand;
            sweep_expand_multi<BT>(amp, k - npend, m, pend, pend_mask & ((1u << m) - 1u), g);
            npend -= m;
            pend_mask >>= m;
            n -= m;
            if ((threadIdx.x & 63u) == 0)
                for (uint32_t j = 0; j < 2 * npend; ++j) pend[j] = pend[j + 2 * m];
            wave_fence();
        }
    }
    // Record one permutation/phase opcode of the current rank k instead of
    // sweeping now (false: fusion is off for this team; sweep as usual).
    CSVM_INLINE bool defer_mono(uint32_t kind, uint32_t a, uint32_t b, uint32_t m, double pr, double pi) {
        if (!kGlobalAmps || !amp_alt) return false;
        if (nmono == kMaxMono) {
            if (npend) flush_expand_oldest(npend);
            flush_mono();
        }
        if (g.lane == 0) {
            MonoOp o;
            o.kind = kind;
            o.a = a;
            o.b = b;
            o.pm = o.px = o.qa = o.qm = o.pk = 0u;
            o.pr = pr;
            o.pi = pi;
            if (kind == MONO_CNOT) {
                o.pm = 1u << a;
                o.px = 1u << b;
            } else if (kind == MONO_FANOUT || kind == MONO_FANOUT_CZ) {
                o.pm = 1u << a;
                o.px = m;
                if (kind == MONO_FANOUT_CZ) {  // the phase, as seen before the permutation
                    o.qa = 1u << a;
                    o.qm = b;
                    o.pk = 1u;
                }
            } else if (kind == MONO_SWAP) {
                o.pm = o.px = (1u << a) | (1u << b);
            } else if (kind == MONO_MULTI_CNOT) {
                o.pm = m;
                o.px = 1u << a;
            } else if (kind == MONO_CZ || kind == MONO_MULTI_CZ) {
                o.qa = 1u << a;
                o.qm = kind == MONO_CZ ? 1u << b : m;
                o.pk = 1u;
            } else {  // MONO_PHASE
                o.qa = o.qm = 1u << a;
                o.pk = 2u;
            }
            store_mono(mops + nmono, o);
        }
        ++nmono;
        return true;
    }
    // Apply the deferred run: a single opcode in place with its own sweep
    // (a fan-out through the out-of-place pass), two or more in one
    // out-of-place pass into the other slice.  The run
    // was recorded over the expanded state: no expansion may be pending.
    CSVM_INLINE void flush_mono() {
        if (!kGlobalAmps || nmono == 0) return;
        team_sync<BT>();  // the list written by lane 0 is visible to every wave
        if (nmono == 1 && uni(mops->kind) < MONO_FANOUT) {
            const uint32_t kind = uni(mops->kind), a = uni(mops->a), b = uni(mops->b);
            if (kind == MONO_CNOT) sweep_two_axis<BT>(amp, k, a, b, PERM_CNOT, g);
            else if (kind == MONO_CZ) sweep_two_axis<BT>(amp, k, a, b, PERM_CZ, g);
            else if (kind == MONO_SWAP) sweep_two_axis<BT>(amp, k, a, b, PERM_SWAP, g);
            else if (kind == MONO_MULTI_CNOT) sweep_multi<BT>(amp, k, a, uni(mops->pm), 0u, g);
            else if (kind == MONO_MULTI_CZ) sweep_multi<BT>(amp, k, a, uni(mops->qm), 1u, g);
            else sweep_phase<BT>(amp, k, a, unid(mops->pr), unid(mops->pi), g);
        } else {
            sweep_mono<BT>(amp, amp_alt, k, mops, nmono, g);
            AmpT* t = amp;
            amp = amp_alt;
            amp_alt = t;
        }
        nmono = 0;
    }

    // Outcome probabilities of an active measurement.  A deferred run of two
    // or more permutation/phase opcodes (any run, with a third slice) is
    // applied in the same pass (into the other slice) instead of a gather of
    // its own followed by a reading pass; pending expansions under it (or
    // under no run) are read through, so the pass reads the state from
    // before them.
    template <class View>
    CSVM_INLINE double2 fused_probs(const View& src, uint32_t kind, uint32_t fa) {
        double2 pr;
        if (amp_alt2) {
            pr = sweep_mono_cands<BT>(src, amp_alt, amp_alt2, k, mops, nmono, kind, fa, g, red);
            cands = 1;
        } else {
            pr = sweep_mono_probs<BT>(src, amp_alt, k, mops, nmono, kind, fa, g, red);
            AmpT* t = amp;
            amp = amp_alt;
            amp_alt = t;
        }
        nmono = 0;
        npend = 0;
        pend_mask = 0;
        return pr;
    }
    CSVM_INLINE double2 probs(uint32_t kind, uint32_t fa) {
        // (teams with LDS amplitudes never defer: not compiled into them)
        if constexpr (kGlobalAmps) {
            if (npend) {
                team_sync<BT>();  // the list written by lane 0 is visible to every wave
                return fused_probs(ExpandView<AmpT>(amp, k - npend, npend, pend, pend_mask), kind, fa);
            }
            // with a third slice every active measurement folds in its pass
            // (a run of zero or one opcode too: one read and one write of
            // the state instead of the opcode's sweep, the sums and the fold)
            if (nmono >= 2 || amp_alt2) {
                team_sync<BT>();
                StoredView<AmpT> src;
                src.a = amp;
                return fused_probs(src, kind, fa);
            }
        }
        return sweep_probs<BT>(amp, k, kind, fa, g, red);
    }

    // Keep the folded state of outcome b written by sweep_mono_cands; the
    // run's source and the other candidate become the free slices.
    CSVM_INLINE void take_cand(uint32_t b) {
        AmpT* src = amp;
        AmpT* other = b ? amp_alt : amp_alt2;
        amp = b ? amp_alt2 : amp_alt;
        amp_alt = src;
        amp_alt2 = other;
        cands = 0;
    }

    // svm.h:177-186 (pending sweeps first: expansions, then the run over them)
    CSVM_INLINE void scale_magnitude(double s) {
        g_re *= s;
        g_im *= s;
        double gm = hypot(g_re, g_im);
        if (gm > 1e100 || (gm < 1e-100 && gm > 0.0)) {
            if (npend) flush_expand_oldest(npend);
            flush_mono();
            sweep_scale<BT>(amp, k, gm, g);
            g_re /= gm;
            g_im /= gm;
        }
    }
    // svm_internal.h:65-79
    CSVM_INLINE uint32_t sample_branch(double p0, double p1, double total) {
        double eps = kDustEpsilon * total;
        if (p1 <= eps) {
            if (p1 > 0.0) dust++;
            return 0;
        }
        if (p0 <= eps) {
            if (p0 > 0.0) dust++;
            return 1;
        }
        double r = rng_double(rng);
        return (r * total < p0) ? 0u : 1u;
    }
    // svm_forced_kernels.cc:43-61
    CSVM_INLINE bool forced_increment(double p0, double p1, double total, uint32_t b, double& inc) {
        double eps = kDustEpsilon * total;
        inc = 0.0;
        if (p1 <= eps) {
            if (p1 > 0.0) dust++;
            return b == 0;
        }
        if (p0 <= eps) {
            if (p0 > 0.0) dust++;
            return b == 1;
        }
        inc = log((b ? p1 : p0) / total);
        return true;
    }

    // ------------------------------------------------------------ records
    CSVM_INLINE void rec_write(__attribute__((address_space(3))) uint8_t* r, uint32_t idx, uint32_t val) {
        if ((threadIdx.x & 63u) == 0) r[idx] = (uint8_t)val;
        wave_fence();
    }
    CSVM_INLINE uint32_t rec_read(const __attribute__((address_space(3))) uint8_t* r, uint32_t idx) {
        return uni(r[idx]);
    }

    // ------------------------------------------------------------ frame ops
    // svm_kernels.inl:67-116, 749-755
    CSVM_INLINE void frame_cnot(uint32_t c, uint32_t t) {
        uint32_t pxc = f.gx(c), pzt = f.gz(t);
        f.fx(t, pxc);
        f.fz(c, pzt);
    }
    CSVM_INLINE void frame_cz(uint32_t c, uint32_t t) {
        uint32_t pxc = f.gx(c), pxt = f.gx(t);
        if (pxc & pxt) negate_gamma();
        f.fz(t, pxc);
        f.fz(c, pxt);
    }
    CSVM_INLINE void frame_h(uint32_t v) {
        uint32_t px = f.gx(v), pz = f.gz(v);
        if (px & pz) negate_gamma();
        if (px != pz) {
            f.fx(v, 1);
            f.fz(v, 1);
        }
    }
    CSVM_INLINE void frame_s(uint32_t v, double sgn) {
        uint32_t px = f.gx(v);
        if (px) mul_phase(0.0, sgn);
        f.fz(v, px);
    }
    CSVM_INLINE void frame_swap(uint32_t a, uint32_t b) {
        uint32_t dx = f.gx(a) ^ f.gx(b), dz = f.gz(a) ^ f.gz(b);
        f.fx(a, dx);
        f.fx(b, dx);
        f.fz(a, dz);
        f.fz(b, dz);
    }
    // A fan of frame CZs or CNOTs sharing qubit p (kLaneFrameFan,
    // svm_device.h): the ops one at a time, on the frame's first word.
    CSVM_INLINE void frame_fan(uint32_t kind, uint32_t p, uint64_t m) {
        const uint64_t x = f.wx(0), z = f.wz(0), pb = 1ull << p;
        if (kind == kFanCz) {
            const uint64_t xp = (x >> p) & 1ull;
            const uint32_t par = (uint32_t)__popcll(x & m) & 1u;
            if (xp && par) negate_gamma();
            f.xor_words(0, 0ull, (xp ? m : 0ull) ^ (par ? pb : 0ull));
        } else if (kind == kFanControl) {
            const uint64_t xp = (x >> p) & 1ull;
            const uint32_t par = (uint32_t)__popcll(z & m) & 1u;
            f.xor_words(0, xp ? m : 0ull, par ? pb : 0ull);
        } else {
            const uint64_t zp = (z >> p) & 1ull;
            const uint32_t par = (uint32_t)__popcll(x & m) & 1u;
            f.xor_words(0, par ? pb : 0ull, zp ? m : 0ull);
        }
    }
    CSVM_INLINE void set_x(uint32_t q, uint32_t b) { f.fx(q, f.gx(q) ^ b); }
    CSVM_INLINE void set_z(uint32_t q, uint32_t b) { f.fz(q, f.gz(q) ^ b); }

    // :1812-1866
    CSVM_INLINE void apply_pauli(cptr<uint64_t> ex, cptr<uint64_t> ez, uint32_t words, uint32_t sign) {
        uint32_t parity = 0;
        for (uint32_t w = 0; w < words; ++w) parity += __popcll(f.wx(w) & ez[w]);
        if (parity & 1u) negate_gamma();
        for (uint32_t w = 0; w < words; ++w) f.xor_words(w, ex[w], ez[w]);
        if (sign) negate_gamma();
    }
    // svm.h:239-248
    CSVM_INLINE void draw_next_noise(cptr<DevProgram> P) {
        uint32_t n = P->num_sites;
        if (n == 0 || next_noise >= n) {
            next_noise = kNoNoise;
            return;
        }
        cptr<double> hz = cp(P->hazards);
        double cur = next_noise == 0 ? 0.0 : hz[next_noise - 1];
        double gap = -log(1.0 - rng_double(rng));
        next_noise = uni(hazard_upper_bound(hz, n, cp(P->hz_guide), P->hz_buckets, P->hz_scale, cur + gap));
    }
    // :1886-1920
    CSVM_INLINE void noise(cptr<DevProgram> P, uint32_t site) {
        if (site != next_noise) return;
        cptr<uint32_t> off = cp(P->site_chan_off);
        cptr<double> cum = cp(P->chan_cum);
        const uint32_t c0 = off[site], c1 = off[site + 1];
        const double u = rng_double(rng);
        if (c1 > c0) {
            const uint32_t c = uni(pick_channel(cum, c0, c1, u * cum[c1 - 1], u));
            if (c < c1) {
                uint32_t W = P->words;
                apply_pauli(cp(P->chan_x) + (size_t)c * W, cp(P->chan_z) + (size_t)c * W, W, 0);
            }
        }
        if (kf) {  // svm.h:218-225
            next_noise = kf_npos < kf_nn ? uni(kf_list[kf_npos]) : kNoNoise;
            kf_npos++;
        } else {
            next_noise++;
            draw_next_noise(P);
        }
    }
    CSVM_INLINE uint32_t list_parity(cptr<uint32_t> off, cptr<uint32_t> tgt, uint32_t l) {
        uint32_t par = 0;
        uint32_t j1 = off[l + 1];
        for (uint32_t j = off[l]; j < j1; ++j) par ^= meas[tgt[j]];
        return uni(par);
    }

    // ------------------------------------------------------------ measurement
    // Shared tail of the active measurements: outcome choice (sampled or
    // forced), then the caller's compaction, rank decrement, renormalisation
    // and frame reset.  Returns false if a forced outcome is unreachable.
    CSVM_INLINE bool choose(double p0, double p1, double total, uint32_t frame_bit, uint32_t sign, uint32_t idx,
                            bool is_forced, uint32_t& b, uint32_t& m_abs, uint32_t& rec) {
        if (is_forced) {
            rec = uni(load_u8(forced, idx));
            m_abs = rec ^ sign;
            b = m_abs ^ frame_bit;
            double inc;
            if (!forced_increment(p0, p1, total, b, inc)) {
                reachable = 0;
                return false;
            }
            logp += inc;
        } else {
            b = sample_branch(p0, p1, total);
            m_abs = b ^ frame_bit;
            rec = m_abs ^ sign;
        }
        return true;
    }
    // :1573-1627; svm_forced_kernels.cc:118-171
    CSVM_INLINE bool meas_diagonal(uint32_t v, uint32_t idx, uint32_t sign, bool is_forced) {
        uint32_t px = f.gx(v), pz = f.gz(v);
        double2 pr = probs(PROB_DIAG, 0u);
        double total = pr.x + pr.y;
        uint32_t b, m_abs, rec;
        if (!choose(pr.x, pr.y, total, px, sign, idx, is_forced, b, m_abs, rec)) return false;
        if (b == 1 && pz) negate_gamma();
        if (cands) take_cand(b);
        else if (b == 1) sweep_take_upper<BT>(amp, k, g);
        k--;
        double pb = b ? pr.y : pr.x;
        if (is_forced || pb > 0.0) scale_magnitude(sqrt(total / pb));
        set_x(v, m_abs);
        set_z(v, 0);
        rec_write(meas, idx, rec);
        return true;
    }
    // :1634-1708; svm_forced_kernels.cc:173-229
    CSVM_INLINE bool meas_interfere(uint32_t v, uint32_t idx, uint32_t sign, bool is_forced) {
        uint32_t px = f.gx(v), pz = f.gz(v);
        double2 pr = probs(PROB_INTERFERE, 0u);
        double total = pr.x + pr.y;
        uint32_t bx, m_abs, rec;
        if (!choose(pr.x, pr.y, total, pz, sign, idx, is_forced, bx, m_abs, rec)) return false;
        if (cands) take_cand(bx);
        else sweep_fold<BT>(amp, k, bx ? -1.0 : 1.0, g);
        if (px && m_abs) negate_gamma();
        k--;
        double pb = bx ? pr.y : pr.x;
        if (is_forced || pb > 0.0) scale_magnitude(sqrt(total / pb));
        set_x(v, m_abs);
        set_z(v, 0);
        rec_write(meas, idx, rec);
        return true;
    }
    // :1728-1804; svm_forced_kernels.cc:231-304
    CSVM_INLINE bool swap_meas_interfere(uint32_t fa, uint32_t t, uint32_t idx, uint32_t sign, bool is_forced) {
        if (fa == t) return meas_interfere(t, idx, sign, is_forced);
        frame_swap(fa, t);
        uint32_t px = f.gx(t), pz = f.gz(t);
        double2 pr = probs(PROB_SWAP_INTERFERE, fa);
        double total = pr.x + pr.y;
        uint32_t bx, m_abs, rec;
        if (!choose(pr.x, pr.y, total, pz, sign, idx, is_forced, bx, m_abs, rec)) return false;
        if (cands) take_cand(bx);
        else sweep_swap_fold<BT>(amp, k, fa, bx ? -1.0 : 1.0, g);
        k--;
        double pb = bx ? pr.y : pr.x;
        if (is_forced || pb > 0.0) scale_magnitude(sqrt(total / pb));
        if (px && m_abs) negate_gamma();
        set_x(t, m_abs);
        set_z(t, 0);
        rec_write(meas, idx, rec);
        return true;
    }

    // :2031-2173
    CSVM_INLINE double exp_val(cptr<DevProgram> P, uint32_t ci) {
        uint32_t W = P->words;
        cptr<uint64_t> mx = cp(P->ev_x) + (size_t)ci * W;
        cptr<uint64_t> mz = cp(P->ev_z) + (size_t)ci * W;
        uint32_t fsign = load_u8(cp(P->ev_sign), ci) ? 1u : 0u;
        uint32_t parity = 0;
        for (uint32_t w = 0; w < W; ++w) parity += __popcll(mx[w] & f.wz(w)) + __popcll(mz[w] & f.wx(w));
        if (parity & 1u) fsign ^= 1u;
        uint32_t full = k / 64, rem = k % 64;
        bool dormant_x = false;
        if (rem > 0 && full < W && (mx[full] & ~((1ULL << rem) - 1ULL))) dormant_x = true;
        for (uint32_t w = (rem == 0 ? full : full + 1); w < W; ++w)
            if (mx[w]) dormant_x = true;
        if (dormant_x) return 0.0;
        uint32_t xa = 0, za = 0;
        if (k > 0) {
            uint64_t am = (1ULL << k) - 1ULL;
            xa = (uint32_t)(mx[0] & am);
            za = (uint32_t)(mz[0] & am);
        }
        uint32_t yc = __popc(xa & za) & 3u;
        double ph_r = (yc == 0) ? 1.0 : (yc == 2 ? -1.0 : 0.0);
        double ph_i = (yc == 1) ? 1.0 : (yc == 3 ? -1.0 : 0.0);
        if (fsign) {
            ph_r = -ph_r;
            ph_i = -ph_i;
        }
        double3 sums = sweep_exp_val<BT>(amp, k, xa, za, g, red);
        double num_re = sums.x * ph_r - sums.y * ph_i;
        double res = sums.z > 0.0 ? num_re / sums.z : 0.0;
        if (res > 1.0) res = 1.0;
        if (res < -1.0) res = -1.0;
        return res;
    }

    // ------------------------------------------------- k-fault fault sets
    // Choose this shot's fault set (svm.cc:435-520) into `list` (per-wave
    // LDS, room for 5 * kMaxForcedFaults words: picks, then the sparse
    // Fisher-Yates map of at most 2k positions) and arm the noise cursor.
    // Every lane runs the same uniform loop and draws; lane 0 stores.
    CSVM_INLINE void kfault_prepare(cptr<KfTables> T, __attribute__((address_space(3))) uint32_t* list) {
        const uint32_t n = T->n, k = T->k, n_q = T->n_q;
        const bool lead = (threadIdx.x & 63u) == 0;
        uint32_t m = 0;
        if (T->uniform) {
            __attribute__((address_space(3))) uint32_t* keys = list + kMaxForcedFaults;
            __attribute__((address_space(3))) uint32_t* vals = keys + 2 * kMaxForcedFaults;
            uint32_t nk = 0;
            auto find = [&](uint32_t pos) -> uint32_t {  // index in the map, or nk
                for (uint32_t e = 0; e < nk; ++e)
                    if (uni(keys[e]) == pos) return e;
                return nk;
            };
            for (uint32_t j = 0; j < k; ++j) {
                const uint32_t remaining = n - j;
                const uint32_t pick = j + (uint32_t)(rng_double(rng) * (double)remaining);
                if (pick == j) continue;  // swap with itself
                const uint32_t nk0 = nk;
                const uint32_t ej = find(j), ep = find(pick);
                const uint32_t vj = ej < nk0 ? uni(vals[ej]) : j;
                const uint32_t vp = ep < nk0 ? uni(vals[ep]) : pick;
                // pool[j] <- pool[pick], pool[pick] <- pool[j]
                const uint32_t sj = ej < nk0 ? ej : nk++;
                const uint32_t sp = ep < nk0 ? ep : nk++;
                if (lead) {
                    keys[sj] = j;
                    vals[sj] = vp;
                    keys[sp] = pick;
                    vals[sp] = vj;
                }
                wave_fence();
            }
            for (uint32_t j = 0; j < k; ++j) {  // prefix pool[0..k), then sort it
                const uint32_t e = find(j);
                const uint32_t v = e < nk ? uni(vals[e]) : j;
                if (lead) list[j] = v;
            }
            wave_fence();
            for (uint32_t a = 1; a < k; ++a) {
                const uint32_t v = uni(list[a]);
                uint32_t b = a;
                while (b > 0 && uni(list[b - 1]) > v) {
                    const uint32_t prev = uni(list[b - 1]);
                    if (lead) list[b] = prev;
                    wave_fence();
                    --b;
                }
                if (lead) list[b] = v;
                wave_fence();
            }
            m = k;
        } else {
            cptr<double> w = cp(T->w), dp = cp(T->dp);
            const uint32_t stride = k + 1;
            uint32_t needed = k;
            for (uint32_t i = 0; i < n && needed > 0; ++i) {
                const uint32_t remaining = n - i;
                double inc;
                if (needed == remaining) {
                    inc = 1.0;
                } else {
                    const double denom = dp[(size_t)i * stride + needed];
                    inc = denom > 0.0 ? (w[i] * dp[(size_t)(i + 1) * stride + needed - 1]) / denom : 0.0;
                }
                if (rng_double(rng) < inc) {
                    if (lead) list[m] = i;
                    ++m;
                    --needed;
                }
            }
            wave_fence();
        }
        // picks are ascending: quantum sites first, then readout entries
        uint32_t nn = 0;
        while (nn < m && uni(list[nn]) < n_q) ++nn;
        for (uint32_t j = nn; j < m; ++j) {
            const uint32_t v = uni(list[j]);
            if (lead) list[j] = v - n_q;
        }
        wave_fence();
        kf = 1;
        kf_list = list;
        kf_nn = nn;
        kf_nr = m - nn;
        kf_rpos = 0;
        if (nn > 0) {
            next_noise = uni(list[0]);
            kf_npos = 1;
        } else {
            next_noise = kNoNoise;
            kf_npos = 0;
        }
    }

    // ------------------------------------------------------------- dispatch
    // svm_kernels.inl:2183-2646.  Returns when the bytecode ends, a
    // POSTSELECT discards the shot, or a forced outcome is unreachable.
    // One-word frames that start at zero (zero_frame) read the team copy of
    // the bytecode, where runs of frame opcodes are packed into kLaneFrameRun
    // and kLaneFrameFan instructions (one dispatch per unit; the covered
    // instructions stay in place and are skipped) without the ops that are
    // the identity from the zero frame; a preset frame reads the full
    // bytecode.
    CSVM_INLINE void run(cptr<DevProgram> Pg, double* ev_out, bool zero_frame) {
        constexpr bool kPacked = std::is_same<F, FrameReg<1>>::value;
        cptr<DevProgram> P0 = opaque(Pg);
        const bool packed = kPacked && zero_frame;
        cptr<uint64_t> code = cp(reinterpret_cast<const uint64_t*>(packed ? P0->team_bc : P0->bc));
        const uint32_t n_instr = P0->n_instr;
        uint32_t skip = 0;
        for (uint32_t pc = 0; pc < n_instr; pc += 1 + skip) {
            const uint64_t w0 = code[4 * pc], w1 = code[4 * pc + 1];
            const uint32_t op = (uint32_t)(w0 & 0xFF);
            const uint32_t flags = (uint32_t)((w0 >> 16) & 0xFF);
            const uint32_t a1 = (uint32_t)((w0 >> 32) & 0xFFFF);
            const uint32_t a2 = (uint32_t)(w0 >> 48);
            const uint32_t i0 = (uint32_t)w1, i1 = (uint32_t)(w1 >> 32);
            const uint32_t sign = flags & CLIFFT_FLAG_SIGN;
            skip = 0;
            if (kPacked && op == kLaneFrameRun) {  // never in the full bytecode
                // frame opcodes keep deferred runs and pending expansions
                const uint64_t w2 = code[4 * pc + 2], w3 = code[4 * pc + 3];
                for (uint32_t j = 0; j < a1; ++j) {
                    const uint64_t w = j < 4 ? w1 : (j < 8 ? w2 : w3);
                    const uint32_t fr = (uint32_t)(w >> (16 * (j & 3u)));
                    const uint32_t fop = fr & 7u, fa = (fr >> 3) & 63u, fb = (fr >> 9) & 63u;
                    if (fop == CLIFFT_OP_FRAME_CNOT) frame_cnot(fa, fb);
                    else if (fop == CLIFFT_OP_FRAME_CZ) frame_cz(fa, fb);
                    else if (fop == CLIFFT_OP_FRAME_H) frame_h(fa);
                    else if (fop == CLIFFT_OP_FRAME_S) frame_s(fa, 1.0);
                    else if (fop == CLIFFT_OP_FRAME_S_DAG) frame_s(fa, -1.0);
                    else frame_swap(fa, fb);
                }
                skip = a2;
                continue;
            }
            if (kPacked && op == kLaneFrameFan) {  // never in the full bytecode
                frame_fan((uint32_t)(w0 >> 8) & 0xFFu, a1, w1);
                skip = a2;
                continue;
            }
            // An active measurement takes a run of two or more itself
            // (probs), and any run over pending expansions; with pend_defer
            // the expansions stay pending under runs and up to that
            // measurement.  Expansions are written before the run recorded
            // over them.
            // a fan-out of ARRAY_CNOTs flushes what one ARRAY_CNOT would
            const uint32_t opk = (kPacked && op == kArrayFanout) ? (uint32_t)CLIFFT_OP_ARRAY_CNOT : op;
            const bool mflush = nmono && !((kMonoKeepOps >> opk) & 1ull) &&
                                !((nmono >= 2 || npend || amp_alt2) && ((kMonoMeasOps >> opk) & 1ull));
            if (npend && (mflush || (!((kNoAmpOps >> opk) & 1ull) && !(pend_defer && (((kMonoOps | kMonoMeasOps) >> opk) & 1ull)))))
                flush_expand_oldest(npend);
            if (mflush) flush_mono();
            switch (op) {
                case CLIFFT_OP_FRAME_CNOT: frame_cnot(a1, a2); break;
                case CLIFFT_OP_FRAME_CZ: frame_cz(a1, a2); break;
                case CLIFFT_OP_FRAME_H: frame_h(a1); break;
                case CLIFFT_OP_FRAME_S: frame_s(a1, 1.0); break;
                case CLIFFT_OP_FRAME_S_DAG: frame_s(a1, -1.0); break;
                case CLIFFT_OP_FRAME_SWAP: frame_swap(a1, a2); break;
                case CLIFFT_OP_ARRAY_CNOT:  // :124-252
                    if (!defer_mono(MONO_CNOT, a1, a2, 0u, 0.0, 0.0)) sweep_two_axis<BT>(amp, k, a1, a2, PERM_CNOT, g);
                    frame_cnot(a1, a2);
                    break;
                case kArrayFanout: {  // ARRAY_CNOTs from control a1 to the targets in w1
                    const uint32_t t = (uint32_t)w1;
                    // then the covered MULTI_CZ from a1 (:559-635), on the active axes
                    const bool cz = ((w0 >> 8) & 0xFFu) != 0u;
                    const uint64_t mz = cz ? (code[4 * pc + 2] & ((1ULL << k) - 1ULL)) : 0ull;
                    // its sign condition x_a1 & parity(y & mz) at y = the permuted
                    // index, as a condition on the index before the permutation
                    const uint32_t qm = (uint32_t)mz ^ ((__popcll((uint64_t)t & mz) & 1) ? (1u << a1) : 0u);
                    if (!defer_mono(cz ? MONO_FANOUT_CZ : MONO_FANOUT, a1, qm, t, 0.0, 0.0)) {
                        for (uint32_t r = t; r; r &= r - 1u)
                            sweep_two_axis<BT>(amp, k, a1, (uint32_t)__builtin_ctz(r), PERM_CNOT, g);
                        if (cz) sweep_multi<BT>(amp, k, a1, (uint32_t)mz, 1u, g);
                    }
                    frame_fan(kFanControl, a1, (uint64_t)t);
                    if (cz) frame_fan(kFanCz, a1, mz);
                    skip = a2;
                    break;
                }
                case CLIFFT_OP_ARRAY_CZ:  // :255-365
                    if (!defer_mono(MONO_CZ, a1, a2, 0u, 0.0, 0.0)) sweep_two_axis<BT>(amp, k, a1, a2, PERM_CZ, g);
                    frame_cz(a1, a2);
                    break;
                case CLIFFT_OP_ARRAY_SWAP:  // :369-458
                    if (!defer_mono(MONO_SWAP, a1, a2, 0u, 0.0, 0.0)) sweep_two_axis<BT>(amp, k, a1, a2, PERM_SWAP, g);
                    frame_swap(a1, a2);
                    break;
                case CLIFFT_OP_ARRAY_MULTI_CNOT: {  // :464-550
                    uint64_t m = w1 & ((1ULL << k) - 1ULL);
                    if (!defer_mono(MONO_MULTI_CNOT, a1, 0u, (uint32_t)m, 0.0, 0.0)) sweep_multi<BT>(amp, k, a1, (uint32_t)m, 0u, g);
                    while (m) {
                        uint32_t c = (uint32_t)__builtin_ctzll(m);
                        m &= m - 1;
                        frame_cnot(c, a1);
                    }
                    break;
                }
                case CLIFFT_OP_ARRAY_MULTI_CZ: {  // :559-635
                    uint64_t m = w1 & ((1ULL << k) - 1ULL);
                    if (!defer_mono(MONO_MULTI_CZ, a1, 0u, (uint32_t)m, 0.0, 0.0)) sweep_multi<BT>(amp, k, a1, (uint32_t)m, 1u, g);
                    while (m) {
                        uint32_t t = (uint32_t)__builtin_ctzll(m);
                        m &= m - 1;
                        frame_cz(a1, t);
                    }
                    break;
                }
                case CLIFFT_OP_ARRAY_H:  // :639-744
                    sweep_h<BT>(amp, k, a1, g);
                    frame_h(a1);
                    break;
                case CLIFFT_OP_ARRAY_S:  // :854-858
                    if (!defer_mono(MONO_PHASE, a1, 0u, 0u, 0.0, 1.0)) sweep_phase<BT>(amp, k, a1, 0.0, 1.0, g);
                    frame_s(a1, 1.0);
                    break;
                case CLIFFT_OP_ARRAY_S_DAG:  // :862-866
                    if (!defer_mono(MONO_PHASE, a1, 0u, 0u, 0.0, -1.0)) sweep_phase<BT>(amp, k, a1, 0.0, -1.0, g);
                    frame_s(a1, -1.0);
                    break;
                case CLIFFT_OP_ARRAY_T:
                case CLIFFT_OP_ARRAY_T_DAG:
                case CLIFFT_OP_ARRAY_ROT: {  // :951-988, 1044-1061
                    double zr = kInvSqrt2, zi = (op == CLIFFT_OP_ARRAY_T_DAG) ? -kInvSqrt2 : kInvSqrt2;
                    if (op == CLIFFT_OP_ARRAY_ROT) {
                        zr = __longlong_as_double((long long)w1);
                        zi = __longlong_as_double((long long)code[4 * pc + 2]);
                    }
                    uint32_t px = f.gx(a1);
                    if (a1 < k && !defer_mono(MONO_PHASE, a1, 0u, 0u, zr, px ? -zi : zi))
                        sweep_phase<BT>(amp, k, a1, zr, px ? -zi : zi, g);
                    if (px) mul_phase(zr, zi);
                    break;
                }
                case CLIFFT_OP_ARRAY_U2: {  // :1091-1297
                    cptr<DevProgram> P = opaque(Pg);
                    cptr<clifft_u2_node> node = cp(P->u2) + i0;
                    uint32_t s = (f.gz(a1) << 1) | f.gx(a1);
                    mul_phase(node->gamma[s][0], node->gamma[s][1]);
                    uint32_t out = load_u8(&node->out_states[0], s);
                    set_x(a1, out & 1u);
                    set_z(a1, (out >> 1) & 1u);
                    if (a1 < k) sweep_u2<BT>(amp, k, a1, &node->matrices[s][0][0], g);
                    break;
                }
                case CLIFFT_OP_ARRAY_U4: {  // :1310-1531
                    cptr<DevProgram> P = opaque(Pg);
                    uint32_t s = (f.gz(a2) << 3) | (f.gx(a2) << 2) | (f.gz(a1) << 1) | f.gx(a1);
                    cptr<clifft_u4_entry> e = &cp(P->u4)[i0].entries[s];
                    mul_phase(e->gamma[0], e->gamma[1]);
                    uint32_t out = load_u8(&e->out_state, 0);
                    set_x(a1, out & 1u);
                    set_z(a1, (out >> 1) & 1u);
                    set_x(a2, (out >> 2) & 1u);
                    set_z(a2, (out >> 3) & 1u);
                    sweep_u4<BT>(amp, k, a1, a2, &e->matrix[0][0][0], g);
                    break;
                }
                case CLIFFT_OP_EXPAND:
                case CLIFFT_OP_EXPAND_T:
                case CLIFFT_OP_EXPAND_T_DAG:
                case CLIFFT_OP_EXPAND_ROT: {  // :876-945, 995-1034, 1065-1081
                    double zr = kInvSqrt2, zi = (op == CLIFFT_OP_EXPAND_T_DAG) ? -kInvSqrt2 : kInvSqrt2;
                    if (op == CLIFFT_OP_EXPAND_ROT) {
                        zr = __longlong_as_double((long long)w1);
                        zi = __longlong_as_double((long long)code[4 * pc + 2]);
                    }
                    const bool has_phase = op != CLIFFT_OP_EXPAND;
                    uint32_t px = f.gx(a1);
                    // recorded now, written by flush_expand_oldest() together
                    // with the expansions that follow it (or read through by
                    // a measurement pass)
                    if (npend == kMaxFusedExpand) flush_expand_oldest(pend_defer ? 2u : kMaxFusedExpand);
                    if ((threadIdx.x & 63u) == 0) {
                        pend[2 * npend] = has_phase ? zr : 1.0;
                        pend[2 * npend + 1] = has_phase ? (px ? -zi : zi) : 0.0;
                    }
                    wave_fence();
                    pend_mask |= (has_phase ? 1u : 0u) << npend;
                    npend++;
                    k++;
                    scale_magnitude(1.0 / sqrt(2.0));
                    if (has_phase && px) mul_phase(zr, zi);
                    break;
                }
                case CLIFFT_OP_MEAS_DORMANT_STATIC:  // :1540-1545, :2361-2369
                    rec_write(meas, i0, (flags & CLIFFT_FLAG_IDENTITY) ? sign : (f.gx(a1) ^ sign));
                    break;
                case CLIFFT_OP_MEAS_DORMANT_RANDOM: {  // :1551-1566
                    uint32_t m = rng_double(rng) < 0.5 ? 0u : 1u;
                    if (f.gx(a1) && m) negate_gamma();
                    set_x(a1, m);
                    set_z(a1, 0);
                    rec_write(meas, i0, m ^ sign);
                    break;
                }
                case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL: meas_diagonal(a1, i0, sign, false); break;
                case CLIFFT_OP_MEAS_ACTIVE_INTERFERE: meas_interfere(a1, i0, sign, false); break;
                case CLIFFT_OP_SWAP_MEAS_INTERFERE: swap_meas_interfere(a1, a2, i0, sign, false); break;
                case CLIFFT_OP_MEAS_DORMANT_STATIC_FORCED: {  // svm_forced_kernels.cc:63-89
                    uint32_t det_v = (flags & CLIFFT_FLAG_IDENTITY) ? sign : (f.gx(a1) ^ sign);
                    if (uni(load_u8(forced, i0)) != det_v) {
                        reachable = 0;
                        return;
                    }
                    rec_write(meas, i0, det_v);
                    break;
                }
                case CLIFFT_OP_MEAS_DORMANT_RANDOM_FORCED: {  // svm_forced_kernels.cc:91-116
                    uint32_t req = uni(load_u8(forced, i0));
                    uint32_t m = req ^ sign;
                    logp += log(0.5);
                    if (f.gx(a1) && m) negate_gamma();
                    set_x(a1, m);
                    set_z(a1, 0);
                    rec_write(meas, i0, req);
                    break;
                }
                case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL_FORCED:
                    if (!meas_diagonal(a1, i0, sign, true)) return;
                    break;
                case CLIFFT_OP_MEAS_ACTIVE_INTERFERE_FORCED:
                    if (!meas_interfere(a1, i0, sign, true)) return;
                    break;
                case CLIFFT_OP_SWAP_MEAS_INTERFERE_FORCED:
                    if (!swap_meas_interfere(a1, a2, i0, sign, true)) return;
                    break;
                case CLIFFT_OP_APPLY_PAULI:  // :1869-1880
                    if (rec_read(meas, i1)) {
                        cptr<DevProgram> P = opaque(Pg);
                        uint32_t W = P->words;
                        apply_pauli(cp(P->pauli_x) + (size_t)i0 * W, cp(P->pauli_z) + (size_t)i0 * W, W, load_u8(cp(P->pauli_sign), i0));
                    }
                    break;
                case CLIFFT_OP_NOISE:
                    if (i0 == next_noise) noise(opaque(Pg), i0);
                    break;
                case CLIFFT_OP_NOISE_BLOCK: {  // :1926-1932
                    uint32_t end = i0 + i1;
                    if (next_noise >= i0 && next_noise < end) {
                        cptr<DevProgram> P = opaque(Pg);
                        while (next_noise >= i0 && next_noise < end) noise(P, next_noise);
                    }
                    break;
                }
                case CLIFFT_OP_READOUT_NOISE: {  // :1937-1959
                    cptr<DevProgram> P = opaque(Pg);
                    bool fire;
                    if (kf) {
                        fire = kf_rpos < kf_nr && uni(kf_list[kf_nn + kf_rpos]) == i0;
                        if (fire) kf_rpos++;
                    } else {
                        fire = rng_double(rng) < cp(P->ro_prob)[i0];
                    }
                    if (fire) {
                        uint32_t mi = cp(P->ro_meas_idx)[i0];
                        rec_write(meas, mi, rec_read(meas, mi) ^ 1u);
                    }
                    break;
                }
                case CLIFFT_OP_DETECTOR: {  // :1964-1979
                    cptr<DevProgram> P = opaque(Pg);
                    uint32_t par = ((flags & CLIFFT_FLAG_EXPECTED_ONE) ? 1u : 0u) ^ list_parity(cp(P->det_off), cp(P->det_tgt), i0);
                    rec_write(det, i1, par);
                    break;
                }
                case CLIFFT_OP_POSTSELECT: {  // :1984-2004
                    cptr<DevProgram> P = opaque(Pg);
                    uint32_t par = ((flags & CLIFFT_FLAG_EXPECTED_ONE) ? 1u : 0u) ^ list_parity(cp(P->det_off), cp(P->det_tgt), i0);
                    rec_write(det, i1, 0);
                    if (par) {
                        discarded = 1;
                        return;
                    }
                    break;
                }
                case CLIFFT_OP_OBSERVABLE: {  // :2007-2020
                    cptr<DevProgram> P = opaque(Pg);
                    rec_write(obs, i1, rec_read(obs, i1) ^ list_parity(cp(P->obs_off), cp(P->obs_tgt), i0));
                    break;
                }
                case CLIFFT_OP_EXP_VAL: {  // :2031-2173
                    double r = exp_val(opaque(Pg), i0);
                    if (ev_out && g.lane == 0) ev_out[i1] = r;
                    break;
                }
                default: break;
            }
        }
        if (npend) flush_expand_oldest(npend);
        flush_mono();
    }
};

}  // namespace csvm
