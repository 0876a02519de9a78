// svm_lane.h -- lane-per-shot interpreter for small active rank.
//
// One wavefront runs 64 shots of the same program at once, one per lane.
// The bytecode walk (opcode fetch, operands, table addresses, the static
// active rank k) is wavefront-uniform and lives in scalar registers; every
// piece of per-shot state is a per-lane value:
//   * Pauli frame (p_x, p_z): two 32-bit values per lane for programs of at
//     most 32 qubits, two 64-bit values up to 64 qubits;
//   * |gamma|, the xoshiro256++ state, the next-noise index, discard/forced
//     bookkeeping;
//   * the 2^k amplitudes in LDS, column-major over the wavefront
//     (amplitude i of lane l at [i * 64 + l]), so the uniform loops of the
//     array opcodes read and write 64 consecutive 16-byte elements;
//   * measurement/detector/observable records bit-packed in LDS words, also
//     column-major ([w * 64 + l]).
// No lane ever reads another lane's data, so no reductions, shuffles or
// barriers are needed, and probabilities are summed in the same sequential
// order as the reference VM.  Lanes diverge only in data-dependent control
// (noise events, forced-outcome checks, post-selection exits).
//
// Handler semantics are those of the team interpreter (svm_interp.h), which
// cites the reference opcode handlers (src/clifft/svm/svm_kernels.inl,
// svm_forced_kernels.cc, svm.h:177-248, svm_internal.h:65-79), with one
// difference: this tier writes records, detector/observable rows,
// expectation values and log-probabilities, none of which depends on the
// phase of gamma (expectation values are normalised by the array's own
// norm, svm_kernels.inl:2031-2173).  Only |gamma| is kept: it decides when
// the array is rescaled (svm.h:177-186).  Unit phases (+-1, +-i, the T
// phases) leave it as it is; ROT weights and fused U2/U4 gamma multipliers
// multiply it by their modulus.
#pragma once

#include "svm_device.h"

namespace csvm {

typedef __attribute__((address_space(3))) dv2 lds_col_d2;
typedef __attribute__((address_space(3))) uint32_t lds_col_u32;

template <class FW>  // frame word: uint32_t (<= 32 qubits) or uint64_t
struct LaneInterp {
    lds_col_d2* amp;   // this lane's column: amplitude i at amp[i << 6]
    lds_col_u32* rm;   // measurement bits (word w at rm[w << 6])
    lds_col_u32* rd;   // detector bits
    lds_col_u32* ro;   // observable bits
    FW px, pz;
    double g_mag;  // |gamma|
    uint32_t k;  // wavefront-uniform
    Rng rng;
    uint32_t next_noise;
    uint32_t discarded;
    uint32_t reachable;
    uint32_t dust;
    double logp;
    const uint8_t* forced;  // this lane's forced record row (MODE_FORCED)

    // k is the same on every live lane (static rank trajectory)
    CSVM_INLINE uint32_t ku() const { return uni(k); }
    CSVM_INLINE dv2 A(uint32_t i) const { return amp[i << 6]; }
    CSVM_INLINE void S(uint32_t i, dv2 v) { amp[i << 6] = v; }

    // ------------------------------------------------------------ records
    CSVM_INLINE static uint32_t rget(const lds_col_u32* r, uint32_t idx) { return (r[(idx >> 5) << 6] >> (idx & 31u)) & 1u; }
    // parity of the measurement-record bits under two (word, mask) pairs
    // held in an instruction's payload words 2 and 3
    CSVM_INLINE uint32_t inline_parity(uint64_t w2, uint64_t w3) const {
        const uint32_t wa = (uint32_t)(w2 >> 16) & 0xFFFFu, wb = (uint32_t)w3 & 0xFFFFu;
        return (uint32_t)(__popc(rm[wa << 6] & (uint32_t)(w2 >> 32)) + __popc(rm[wb << 6] & (uint32_t)(w3 >> 32))) & 1u;
    }
    CSVM_INLINE static void rset(lds_col_u32* r, uint32_t idx, uint32_t v) {
        lds_col_u32* w = r + ((idx >> 5) << 6);
        uint32_t bit = 1u << (idx & 31u);
        *w = (*w & ~bit) | ((v & 1u) ? bit : 0u);
    }

    // ----------------------------------------------------------------- frame
    CSVM_INLINE uint32_t gx(uint32_t q) const { return (uint32_t)(px >> q) & 1u; }
    CSVM_INLINE uint32_t gz(uint32_t q) const { return (uint32_t)(pz >> q) & 1u; }
    CSVM_INLINE void set_x(uint32_t q, uint32_t b) { px = (px & ~((FW)1 << q)) | ((FW)(b & 1u) << q); }
    CSVM_INLINE void set_z(uint32_t q, uint32_t b) { pz = (pz & ~((FW)1 << q)) | ((FW)(b & 1u) << q); }
    // gamma *= (re + i im): only the modulus is kept
    CSVM_INLINE void mul_gamma(double re, double im) { g_mag *= hypot(re, im); }
    // Heisenberg frame updates (the phase they give gamma is not kept)
    CSVM_INLINE void frame_cnot(uint32_t c, uint32_t t) {
        const FW xc = (px >> c) & 1u, zt = (pz >> t) & 1u;
        px ^= xc << t;
        pz ^= zt << c;
    }
    CSVM_INLINE void frame_cz(uint32_t c, uint32_t t) {
        const FW xc = (px >> c) & 1u, xt = (px >> t) & 1u;
        pz ^= (xc << t) | (xt << c);
    }
    CSVM_INLINE void frame_h(uint32_t v) {
        const FW d = (((px ^ pz) >> v) & 1u) << v;
        px ^= d;
        pz ^= d;
    }
    CSVM_INLINE void frame_s(uint32_t v) { pz ^= ((px >> v) & 1u) << v; }
    CSVM_INLINE void frame_swap(uint32_t a, uint32_t b) {
        const FW dx = ((px >> a) ^ (px >> b)) & 1u, dz = ((pz >> a) ^ (pz >> b)) & 1u;
        px ^= (dx << a) | (dx << b);
        pz ^= (dz << a) | (dz << b);
    }
    // masks fit FW: every qubit index is below the program's qubit count
    CSVM_INLINE void apply_pauli(uint64_t ex, uint64_t ez) {
        px ^= (FW)ex;
        pz ^= (FW)ez;
    }

    // ------------------------------------------------------------ magnitude
    CSVM_INLINE void scale_magnitude(double s) {
        g_mag *= s;
        const double gm = g_mag;
        if (gm > 1e100 || (gm < 1e-100 && gm > 0.0)) {
            const uint32_t n = 1u << ku();
            for (uint32_t i = 0; i < n; ++i) S(i, cscale(A(i), gm));
            g_mag = 1.0;
        }
    }
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

    // ----------------------------------------------------------------- noise
    // Gap sampler (svm.h:229-243): the next firing site is the first whose
    // cumulative hazard exceeds cur (the hazard through the last site
    // reached) plus an Exp(1) gap.
    CSVM_INLINE void draw_next_noise(cptr<DevProgram> P, double cur) {
        const uint32_t n = P->num_sites;
        if (n == 0 || next_noise >= n) {
            next_noise = kNoNoise;
            return;
        }
        const double gap = -log(1.0 - rng_double(rng));
        next_noise = hazard_upper_bound(P->hazards, n, P->hz_guide, P->hz_buckets, P->hz_scale, cur + gap);
    }
    // Fire the noise site `site` (== next_noise) and draw the next one
    // (svm_kernels.inl:1886-1920).
    CSVM_INLINE void noise(cptr<DevProgram> P, uint32_t site) {
        const uint32_t c0 = P->site_chan_off[site], c1 = P->site_chan_off[site + 1];
        const double cur = P->hazards[site];
        const double u = rng_double(rng);
        if (c1 > c0) {
            const double* cum = P->chan_cum;
            const uint32_t c = pick_channel(cum, c0, c1, u * cum[c1 - 1], u);
            if (c < c1) apply_pauli(P->chan_x[c], P->chan_z[c]);
        }
        next_noise++;
        draw_next_noise(P, cur);
    }

    // ---------------------------------------------------------- measurement
    CSVM_INLINE bool choose(double p0, double p1, double total, uint32_t frame_bit, uint32_t sign, uint32_t idx,
                            bool is_forced, uint32_t& b, uint32_t& m_abs, uint32_t& rec) {
        if (is_forced) {
            rec = forced[idx] ? 1u : 0u;
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
    CSVM_INLINE bool meas_diagonal(uint32_t v, uint32_t idx, uint32_t sign, bool is_forced) {
        const uint32_t x = gx(v), half = 1u << (ku() - 1);
        double p0 = 0.0, p1 = 0.0;
        for (uint32_t i = 0; i < half; ++i) {
            dv2 a = A(i), b = A(i + half);
            p0 += a.x * a.x + a.y * a.y;
            p1 += b.x * b.x + b.y * b.y;
        }
        double total = p0 + p1;
        uint32_t b, m_abs, rec;
        if (!choose(p0, p1, total, x, sign, idx, is_forced, b, m_abs, rec)) return false;
        if (b)
            for (uint32_t i = 0; i < half; ++i) S(i, A(i + half));
        k = ku() - 1;
        double pb = b ? p1 : p0;
        if (is_forced || pb > 0.0) scale_magnitude(sqrt(total / pb));
        set_x(v, m_abs);
        set_z(v, 0);
        rset(rm, idx, rec);
        return true;
    }
    CSVM_INLINE bool meas_interfere(uint32_t v, uint32_t idx, uint32_t sign, bool is_forced) {
        const uint32_t z = gz(v), half = 1u << (ku() - 1);
        double p0 = 0.0, p1 = 0.0;
        for (uint32_t i = 0; i < half; ++i) {
            dv2 a = A(i), b = A(i + half);
            dv2 s = cadd(a, b), d = csub(a, b);
            p0 += s.x * s.x + s.y * s.y;
            p1 += d.x * d.x + d.y * d.y;
        }
        double total = p0 + p1;
        uint32_t bx, m_abs, rec;
        if (!choose(p0, p1, total, z, sign, idx, is_forced, bx, m_abs, rec)) return false;
        const double s = bx ? -1.0 : 1.0;
        for (uint32_t i = 0; i < half; ++i) {
            dv2 a = A(i), b = A(i + half);
            S(i, mk((a.x + s * b.x) * kInvSqrt2, (a.y + s * b.y) * kInvSqrt2));
        }
        k = ku() - 1;
        double pb = bx ? p1 : p0;
        if (is_forced || pb > 0.0) scale_magnitude(sqrt(total / pb));
        set_x(v, m_abs);
        set_z(v, 0);
        rset(rm, idx, rec);
        return true;
    }
    CSVM_INLINE bool swap_meas_interfere(uint32_t fa, uint32_t t, uint32_t idx, uint32_t sign, bool is_forced) {
        if (fa == t) return meas_interfere(t, idx, sign, is_forced);
        frame_swap(fa, t);
        const uint32_t z = gz(t);
        const uint32_t half = 1u << (ku() - 1), top = ku() - 1, fb = 1u << fa, tb = 1u << top;
        double p0 = 0.0, p1 = 0.0;
        for (uint32_t j = 0; j < half; ++j) {
            uint32_t bf = (j >> fa) & 1u, base = (j & ~fb) | (bf << top);
            dv2 a = A(base), b = A(base | fb);
            dv2 s = cadd(a, b), d = csub(a, b);
            p0 += s.x * s.x + s.y * s.y;
            p1 += d.x * d.x + d.y * d.y;
        }
        double total = p0 + p1;
        uint32_t bx, m_abs, rec;
        if (!choose(p0, p1, total, z, sign, idx, is_forced, bx, m_abs, rec)) return false;
        const double s = bx ? -1.0 : 1.0;
        const uint32_t quarter = 1u << (top - 1);
        for (uint32_t j = 0; j < quarter; ++j) {
            uint32_t i0 = insert_zero_bit(j, fa), i1 = i0 | fb;
            dv2 a0 = A(i0), a1 = A(i1), b0 = A(i0 | tb), b1 = A(i1 | tb);
            S(i0, mk((a0.x + s * a1.x) * kInvSqrt2, (a0.y + s * a1.y) * kInvSqrt2));
            S(i1, mk((b0.x + s * b1.x) * kInvSqrt2, (b0.y + s * b1.y) * kInvSqrt2));
        }
        k = ku() - 1;
        double pb = bx ? p1 : p0;
        if (is_forced || pb > 0.0) scale_magnitude(sqrt(total / pb));
        set_x(t, m_abs);
        set_z(t, 0);
        rset(rm, idx, rec);
        return true;
    }

    CSVM_INLINE double exp_val(uint64_t mx, uint64_t mz, uint32_t msign) {
        uint32_t fsign = msign ^ ((uint32_t)(__popcll(mx & (uint64_t)pz) + __popcll(mz & (uint64_t)px)) & 1u);
        const uint64_t am = (1ull << ku()) - 1ull;
        if (mx & ~am) return 0.0;  // X/Y support on a dormant qubit
        const uint32_t xa = (uint32_t)(mx & am), za = (uint32_t)(mz & am);
        const uint32_t yc = __popc(xa & za) & 3u;
        double ph_r = (yc == 0) ? 1.0 : (yc == 2 ? -1.0 : 0.0);
        double ph_i = (yc == 1) ? 1.0 : (yc == 3 ? -1.0 : 0.0);
        if (fsign) {
            ph_r = -ph_r;
            ph_i = -ph_i;
        }
        double nr = 0.0, ni = 0.0, den = 0.0;
        const uint32_t dim = 1u << ku();
        for (uint32_t j = 0; j < dim; ++j) {
            dv2 a = A(j), b = A(j ^ xa);
            den += a.x * a.x + a.y * a.y;
            double zs = (__popc(j & za) & 1u) ? -1.0 : 1.0;
            double cr = zs * a.x, ci = zs * a.y;
            nr += b.x * cr + b.y * ci;
            ni += b.x * ci - b.y * cr;
        }
        double num_re = nr * ph_r - ni * ph_i;
        double res = den > 0.0 ? num_re / den : 0.0;
        if (res > 1.0) res = 1.0;
        if (res < -1.0) res = -1.0;
        return res;
    }

    // ------------------------------------------------------------- dispatch
    // Runs this lane's shot over the lane copy of the bytecode (operands
    // inlined, superinstructions for runs of frame opcodes, detector runs and
    // measurement + reset pairs; see DevProgram::lane_bc).  A shot that ends
    // early (post-selection discard, unreachable forced outcome) stops
    // executing opcodes but stays in the loop, so pc remains uniform.
    // kFrameFirst: dispatch the frame opcodes (0..5) before the general
    // switch.  Measured on one MI355X: 11% faster on the 20-qubit Clifford+T
    // (peak 6, the 4-workgroup build), 2% slower on the surface-code memory
    // (the 6-workgroup build), so only the 4-workgroup build uses it.
    template <bool kFrameFirst>
    CSVM_INLINE void run(cptr<DevProgram> Pg, double* ev_row) {
        cptr<DevProgram> P0 = opaque(Pg);
        cptr<uint64_t> code = cp(reinterpret_cast<const uint64_t*>(P0->lane_bc));
        const uint32_t n_instr = P0->n_instr;
        // A shot that ends early (post-selection discard, unreachable
        // forced record) clears `live` and skips the remaining opcodes
        // instead of leaving the loop: the loop then has no per-lane exit,
        // so the compiler keeps pc, the opcode fetch and the dispatch on the
        // scalar unit (a per-lane exit would make them per-lane values,
        // fetched with vector loads and dispatched under exec masks).
        uint32_t live = 1;
        uint32_t skip = 0;  // instructions a lane superinstruction covers beyond its own
        for (uint32_t pc = 0; pc < n_instr; pc += 1 + skip) {
            const uint64_t w0 = code[4 * pc], w1 = code[4 * pc + 1], w2 = code[4 * pc + 2], w3 = code[4 * pc + 3];
            const uint32_t op = (uint32_t)(w0 & 0xFF);
            const uint32_t flags = (uint32_t)((w0 >> 16) & 0xFF);
            const uint32_t a1 = (uint32_t)((w0 >> 32) & 0xFFFF);
            const uint32_t a2 = (uint32_t)(w0 >> 48);
            const uint32_t i0 = (uint32_t)w1, i1 = (uint32_t)(w1 >> 32);
            const uint32_t sign = flags & CLIFFT_FLAG_SIGN;
            // decided before the per-lane `live` test, so that pc stays uniform
            skip = op >= kLaneFrameRun ? a2 : 0u;
            if (!live) continue;
            if (kFrameFirst && op <= CLIFFT_OP_FRAME_SWAP) {
                if (op == CLIFFT_OP_FRAME_CNOT) frame_cnot(a1, a2);
                else if (op == CLIFFT_OP_FRAME_CZ) frame_cz(a1, a2);
                else if (op == CLIFFT_OP_FRAME_SWAP) frame_swap(a1, a2);
                else if (op == CLIFFT_OP_FRAME_H) frame_h(a1);
                else frame_s(a1);
                continue;
            }
            switch (op) {
                case kLaneFrameRun:  // a packed run of frame opcodes
                    for (uint32_t j = 0; j < a1; ++j) {
                        const uint64_t w = j < 4 ? w1 : (j < 8 ? w2 : w3);
                        const uint32_t f = (uint32_t)(w >> (16 * (j & 3u)));
                        const uint32_t fop = f & 7u, fa = (f >> 3) & 63u, fb = (f >> 9) & 63u;
                        if (fop == CLIFFT_OP_FRAME_CNOT) frame_cnot(fa, fb);
                        else if (fop == CLIFFT_OP_FRAME_CZ) frame_cz(fa, fb);
                        else if (fop == CLIFFT_OP_FRAME_SWAP) frame_swap(fa, fb);
                        else if (fop == CLIFFT_OP_FRAME_H) frame_h(fa);
                        else frame_s(fa);
                    }
                    break;
                case kLaneFrameFan: {  // frame CZs or CNOTs sharing qubit a1 (svm_device.h)
                    const FW m = (FW)w1, pb = (FW)1 << a1;
                    const uint32_t kind = (uint32_t)(w0 >> 8) & 0xFFu;
                    if (kind == kFanCz) {
                        const FW xp = (px >> a1) & 1u;
                        const FW par = (FW)(__popcll((uint64_t)(px & m)) & 1);
                        pz ^= (xp ? m : (FW)0) ^ (par ? pb : (FW)0);
                    } else if (kind == kFanControl) {
                        const FW xp = (px >> a1) & 1u;
                        const FW par = (FW)(__popcll((uint64_t)(pz & m)) & 1);
                        px ^= xp ? m : (FW)0;
                        pz ^= par ? pb : (FW)0;
                    } else {
                        const FW zp = (pz >> a1) & 1u;
                        const FW par = (FW)(__popcll((uint64_t)(px & m)) & 1);
                        px ^= par ? pb : (FW)0;
                        pz ^= zp ? m : (FW)0;
                    }
                    break;
                }
                case kLaneDetectorRun:  // DETECTORs with inline parities, each from its own words
                    for (uint32_t j = 0; j <= a2; ++j) {
                        const uint64_t x0 = code[4 * (pc + j)], x1 = code[4 * (pc + j) + 1];
                        const uint64_t x2 = code[4 * (pc + j) + 2], x3 = code[4 * (pc + j) + 3];
                        const uint32_t e = (((uint32_t)(x0 >> 16) & 0xFF) & CLIFFT_FLAG_EXPECTED_ONE) ? 1u : 0u;
                        rset(rd, (uint32_t)(x1 >> 32), e ^ inline_parity(x2, x3));
                    }
                    break;
                case kLaneMeasPauli: {  // dormant measurement, then the APPLY_PAULI on its record
                    uint32_t rec;
                    if (((w0 >> 8) & 0xFF) == CLIFFT_OP_MEAS_DORMANT_STATIC) {
                        rec = (flags & CLIFFT_FLAG_IDENTITY) ? sign : (gx(a1) ^ sign);
                    } else {
                        const uint32_t m = rng_double(rng) < 0.5 ? 0u : 1u;
                        set_x(a1, m);
                        set_z(a1, 0);
                        rec = m ^ sign;
                    }
                    rset(rm, i0, rec);
                    if (rec) apply_pauli(w2, w3);
                    break;
                }
                case CLIFFT_OP_FRAME_CNOT: frame_cnot(a1, a2); break;
                case CLIFFT_OP_FRAME_CZ: frame_cz(a1, a2); break;
                case CLIFFT_OP_FRAME_H: frame_h(a1); break;
                case CLIFFT_OP_FRAME_S:
                case CLIFFT_OP_FRAME_S_DAG: frame_s(a1); break;
                case CLIFFT_OP_FRAME_SWAP: frame_swap(a1, a2); break;
                case CLIFFT_OP_ARRAY_CNOT: {
                    const uint32_t n = 1u << (ku() - 2), ab = 1u << a1, bb = 1u << a2;
                    const uint32_t lo = a1 < a2 ? a1 : a2, hi = a1 < a2 ? a2 : a1;
                    for (uint32_t i = 0; i < n; ++i) {
                        uint32_t b0 = insert_two_zero_bits(i, lo, hi) | ab, b1 = b0 | bb;
                        dv2 x = A(b0), y = A(b1);
                        S(b0, y);
                        S(b1, x);
                    }
                    frame_cnot(a1, a2);
                    break;
                }
                case CLIFFT_OP_ARRAY_CZ: {
                    const uint32_t n = 1u << (ku() - 2), ab = 1u << a1, bb = 1u << a2;
                    const uint32_t lo = a1 < a2 ? a1 : a2, hi = a1 < a2 ? a2 : a1;
                    for (uint32_t i = 0; i < n; ++i) {
                        uint32_t idx = insert_two_zero_bits(i, lo, hi) | ab | bb;
                        S(idx, cneg(A(idx)));
                    }
                    frame_cz(a1, a2);
                    break;
                }
                case CLIFFT_OP_ARRAY_SWAP: {
                    const uint32_t n = 1u << (ku() - 2), ab = 1u << a1, bb = 1u << a2;
                    const uint32_t lo = a1 < a2 ? a1 : a2, hi = a1 < a2 ? a2 : a1;
                    for (uint32_t i = 0; i < n; ++i) {
                        uint32_t base = insert_two_zero_bits(i, lo, hi);
                        dv2 x = A(base | ab), y = A(base | bb);
                        S(base | ab, y);
                        S(base | bb, x);
                    }
                    frame_swap(a1, a2);
                    break;
                }
                case CLIFFT_OP_ARRAY_MULTI_CNOT: {
                    const uint32_t m = (uint32_t)(w1 & ((1ULL << ku()) - 1ULL));
                    const uint32_t half = 1u << (ku() - 1), tb = 1u << a1;
                    for (uint32_t i = 0; i < half; ++i) {
                        uint32_t a = insert_zero_bit(i, a1);
                        if (__popc(a & m) & 1u) {
                            dv2 x = A(a), y = A(a | tb);
                            S(a, y);
                            S(a | tb, x);
                        }
                    }
                    uint32_t mm = m;
                    while (mm) {
                        uint32_t c = (uint32_t)__builtin_ctz(mm);
                        mm &= mm - 1;
                        frame_cnot(c, a1);
                    }
                    break;
                }
                case CLIFFT_OP_ARRAY_MULTI_CZ: {
                    const uint32_t m = (uint32_t)(w1 & ((1ULL << ku()) - 1ULL));
                    const uint32_t half = 1u << (ku() - 1), cb = 1u << a1;
                    for (uint32_t i = 0; i < half; ++i) {
                        uint32_t a = insert_zero_bit(i, a1) | cb;
                        if (__popc(a & m) & 1u) S(a, cneg(A(a)));
                    }
                    uint32_t mm = m;
                    while (mm) {
                        uint32_t t = (uint32_t)__builtin_ctz(mm);
                        mm &= mm - 1;
                        frame_cz(a1, t);
                    }
                    break;
                }
                case CLIFFT_OP_ARRAY_H: {
                    const uint32_t half = 1u << (ku() - 1), vb = 1u << a1;
                    for (uint32_t j = 0; j < half; ++j) {
                        uint32_t x0 = insert_zero_bit(j, a1);
                        dv2 a = A(x0), b = A(x0 | vb);
                        S(x0, cscale(cadd(a, b), kInvSqrt2));
                        S(x0 | vb, cscale(csub(a, b), kInvSqrt2));
                    }
                    frame_h(a1);
                    break;
                }
                case CLIFFT_OP_ARRAY_S:
                case CLIFFT_OP_ARRAY_S_DAG: {
                    const double sg = op == CLIFFT_OP_ARRAY_S ? 1.0 : -1.0;
                    const uint32_t half = 1u << (ku() - 1), vb = 1u << a1;
                    const dv2 ph = mk(0.0, sg);
                    for (uint32_t j = 0; j < half; ++j) {
                        uint32_t x1 = insert_zero_bit(j, a1) | vb;
                        S(x1, cmul(A(x1), ph));
                    }
                    frame_s(a1);
                    break;
                }
                case CLIFFT_OP_ARRAY_T:
                case CLIFFT_OP_ARRAY_T_DAG:
                case CLIFFT_OP_ARRAY_ROT: {
                    double zr = kInvSqrt2, zi = (op == CLIFFT_OP_ARRAY_T_DAG) ? -kInvSqrt2 : kInvSqrt2;
                    if (op == CLIFFT_OP_ARRAY_ROT) {
                        zr = __longlong_as_double((long long)w1);
                        zi = __longlong_as_double((long long)w2);
                    }
                    const uint32_t x = gx(a1);
                    if (a1 < ku()) {
                        const uint32_t half = 1u << (ku() - 1), vb = 1u << a1;
                        const dv2 ph = mk(zr, x ? -zi : zi);
                        for (uint32_t j = 0; j < half; ++j) {
                            uint32_t x1 = insert_zero_bit(j, a1) | vb;
                            S(x1, cmul(A(x1), ph));
                        }
                    }
                    if (x && op == CLIFFT_OP_ARRAY_ROT) mul_gamma(zr, zi);
                    break;
                }
                case CLIFFT_OP_ARRAY_U2: {
                    cptr<DevProgram> P = opaque(Pg);
                    const clifft_u2_node* node = P->u2 + i0;
                    const uint32_t s = (gz(a1) << 1) | gx(a1);
                    mul_gamma(node->gamma[s][0], node->gamma[s][1]);
                    const uint32_t out = node->out_states[s];
                    set_x(a1, out & 1u);
                    set_z(a1, (out >> 1) & 1u);
                    if (a1 < ku()) {
                        const double* m = &node->matrices[s][0][0];
                        const dv2 m00 = mk(m[0], m[1]), m01 = mk(m[2], m[3]), m10 = mk(m[4], m[5]),
                                  m11 = mk(m[6], m[7]);
                        const uint32_t half = 1u << (ku() - 1), vb = 1u << a1;
                        for (uint32_t j = 0; j < half; ++j) {
                            uint32_t x0 = insert_zero_bit(j, a1);
                            dv2 a = A(x0), b = A(x0 | vb);
                            S(x0, cadd(cmul(m00, a), cmul(m01, b)));
                            S(x0 | vb, cadd(cmul(m10, a), cmul(m11, b)));
                        }
                    }
                    break;
                }
                case CLIFFT_OP_ARRAY_U4: {
                    cptr<DevProgram> P = opaque(Pg);
                    const uint32_t s = (gz(a2) << 3) | (gx(a2) << 2) | (gz(a1) << 1) | gx(a1);
                    const clifft_u4_entry* e = &P->u4[i0].entries[s];
                    mul_gamma(e->gamma[0], e->gamma[1]);
                    const uint32_t out = e->out_state;
                    set_x(a1, out & 1u);
                    set_z(a1, (out >> 1) & 1u);
                    set_x(a2, (out >> 2) & 1u);
                    set_z(a2, (out >> 3) & 1u);
                    const uint32_t n = 1u << (ku() - 2), lb = 1u << a1, hb = 1u << a2;
                    const double* m0 = &e->matrix[0][0][0];
                    for (uint32_t i = 0; i < n; ++i) {
                        uint32_t base = insert_two_zero_bits(i, a1, a2);
                        uint32_t ix[4] = {base, base | lb, base | hb, base | lb | hb};
                        dv2 in[4] = {A(ix[0]), A(ix[1]), A(ix[2]), A(ix[3])};
                        // re-read the (per-lane) matrix rows each trip instead
                        // of holding 32 doubles in registers across the loop
                        const double* m = m0;
                        __asm__ volatile("" : "+v"(m));
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const double* row = m + 8 * r;
                            dv2 acc = cmul(mk(row[0], row[1]), in[0]);
                            acc = cadd(acc, cmul(mk(row[2], row[3]), in[1]));
                            acc = cadd(acc, cmul(mk(row[4], row[5]), in[2]));
                            acc = cadd(acc, cmul(mk(row[6], row[7]), in[3]));
                            S(ix[r], acc);
                        }
                    }
                    break;
                }
                case CLIFFT_OP_EXPAND:
                case CLIFFT_OP_EXPAND_T:
                case CLIFFT_OP_EXPAND_T_DAG:
                case CLIFFT_OP_EXPAND_ROT: {
                    double zr = kInvSqrt2, zi = (op == CLIFFT_OP_EXPAND_T_DAG) ? -kInvSqrt2 : kInvSqrt2;
                    if (op == CLIFFT_OP_EXPAND_ROT) {
                        zr = __longlong_as_double((long long)w1);
                        zi = __longlong_as_double((long long)w2);
                    }
                    const bool has_phase = op != CLIFFT_OP_EXPAND;
                    const uint32_t x = gx(a1);
                    const uint32_t half = 1u << ku();
                    const dv2 ph = has_phase ? mk(zr, x ? -zi : zi) : mk(1.0, 0.0);
                    for (uint32_t i = 0; i < half; ++i) S(i + half, has_phase ? cmul(A(i), ph) : A(i));
                    k = ku() + 1;
                    scale_magnitude(1.0 / sqrt(2.0));
                    if (x && op == CLIFFT_OP_EXPAND_ROT) mul_gamma(zr, zi);
                    break;
                }
                case CLIFFT_OP_MEAS_DORMANT_STATIC:
                    rset(rm, i0, (flags & CLIFFT_FLAG_IDENTITY) ? sign : (gx(a1) ^ sign));
                    break;
                case CLIFFT_OP_MEAS_DORMANT_RANDOM: {
                    const uint32_t m = rng_double(rng) < 0.5 ? 0u : 1u;
                    set_x(a1, m);
                    set_z(a1, 0);
                    rset(rm, i0, m ^ sign);
                    break;
                }
                case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL: meas_diagonal(a1, i0, sign, false); break;
                case CLIFFT_OP_MEAS_ACTIVE_INTERFERE: meas_interfere(a1, i0, sign, false); break;
                case CLIFFT_OP_SWAP_MEAS_INTERFERE: swap_meas_interfere(a1, a2, i0, sign, false); break;
                case CLIFFT_OP_MEAS_DORMANT_STATIC_FORCED: {
                    const uint32_t det_v = (flags & CLIFFT_FLAG_IDENTITY) ? sign : (gx(a1) ^ sign);
                    if ((forced[i0] ? 1u : 0u) != det_v) {
                        reachable = 0;
                        live = 0;
                        break;
                    }
                    rset(rm, i0, det_v);
                    break;
                }
                case CLIFFT_OP_MEAS_DORMANT_RANDOM_FORCED: {
                    const uint32_t req = forced[i0] ? 1u : 0u;
                    const uint32_t m = req ^ sign;
                    logp += log(0.5);
                    set_x(a1, m);
                    set_z(a1, 0);
                    rset(rm, i0, req);
                    break;
                }
                case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL_FORCED:
                    if (!meas_diagonal(a1, i0, sign, true)) live = 0;
                    break;
                case CLIFFT_OP_MEAS_ACTIVE_INTERFERE_FORCED:
                    if (!meas_interfere(a1, i0, sign, true)) live = 0;
                    break;
                case CLIFFT_OP_SWAP_MEAS_INTERFERE_FORCED:
                    if (!swap_meas_interfere(a1, a2, i0, sign, true)) live = 0;
                    break;
                case CLIFFT_OP_APPLY_PAULI:  // masks inlined in payload words 2 and 3
                    if (rget(rm, i1)) apply_pauli(w2, w3);
                    break;
                case CLIFFT_OP_NOISE:
                    if (i0 == next_noise) noise(opaque(Pg), i0);
                    break;
                case CLIFFT_OP_NOISE_BLOCK: {
                    const uint32_t end = i0 + i1;
                    if (next_noise >= i0 && next_noise < end) {
                        cptr<DevProgram> P = opaque(Pg);
                        while (next_noise >= i0 && next_noise < end) noise(P, next_noise);
                    }
                    break;
                }
                case CLIFFT_OP_READOUT_NOISE: {
                    cptr<DevProgram> P = opaque(Pg);
                    if (rng_double(rng) < cp(P->ro_prob)[i0]) {
                        const uint32_t mi = cp(P->ro_meas_idx)[i0];
                        rset(rm, mi, rget(rm, mi) ^ 1u);
                    }
                    break;
                }
                case CLIFFT_OP_DETECTOR:
                case CLIFFT_OP_POSTSELECT: {
                    uint32_t par = (flags & CLIFFT_FLAG_EXPECTED_ONE) ? 1u : 0u;
                    if ((uint32_t)(w2 & 0xFFFF) == kLaneInlineParity) {
                        par ^= inline_parity(w2, w3);
                    } else {
                        cptr<DevProgram> P = opaque(Pg);
                        cptr<uint32_t> off = cp(P->det_off), tgt = cp(P->det_tgt);
                        const uint32_t j1 = off[i0 + 1];
                        for (uint32_t j = off[i0]; j < j1; ++j) par ^= rget(rm, tgt[j]);
                    }
                    if (op == CLIFFT_OP_DETECTOR) {
                        rset(rd, i1, par);
                    } else {
                        rset(rd, i1, 0);
                        if (par) {
                            discarded = 1;
                            live = 0;
                        }
                    }
                    break;
                }
                case CLIFFT_OP_OBSERVABLE: {
                    uint32_t par = 0;
                    if ((uint32_t)(w2 & 0xFFFF) == kLaneInlineParity) {
                        par = inline_parity(w2, w3);
                    } else {
                        cptr<DevProgram> P = opaque(Pg);
                        cptr<uint32_t> off = cp(P->obs_off), tgt = cp(P->obs_tgt);
                        const uint32_t j1 = off[i0 + 1];
                        for (uint32_t j = off[i0]; j < j1; ++j) par ^= rget(rm, tgt[j]);
                    }
                    rset(ro, i1, rget(ro, i1) ^ par);
                    break;
                }
                case CLIFFT_OP_EXP_VAL: {
                    cptr<DevProgram> P = opaque(Pg);
                    const double r = exp_val(cp(P->ev_x)[i0], cp(P->ev_z)[i0], load_u8(cp(P->ev_sign), i0) ? 1u : 0u);
                    if (ev_row) ev_row[i1] = r;
                    break;
                }
                default: break;
            }
        }
    }
};

}  // namespace csvm
