// svm_sweeps.h -- amplitude sweeps and reductions of the team interpreter.
//
// Each sweep takes only what it touches (amplitude pointer in its own
// address space, rank, axes, coefficients) and is inlined into the
// interpreter.  AmpPtr is either an LDS pointer (address space 3:
// ds_read/ds_write) or a global pointer (address space 1:
// global_load/global_store).
#pragma once

#include "svm_device.h"

namespace csvm {

typedef __attribute__((address_space(3))) dv2 lds_d2;
typedef __attribute__((address_space(1))) dv2 glb_d2;
typedef __attribute__((address_space(3))) double lds_f64;

#define CSVM_SWEEP __device__ __forceinline__

// Team synchronisation: wavefront-level (one-wave teams) or workgroup barrier.
template <bool BT>
CSVM_INLINE void team_sync() {
    if (BT) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

struct TeamGeom {
    uint32_t lane;    // lane within team
    uint32_t size;    // threads in team
    uint32_t wave;    // wavefront within team
    uint32_t nwaves;  // wavefronts in team
};

// Global-memory sweeps keep several independent 16-byte loads in flight per
// lane (loads of a whole group are issued before its stores); LDS sweeps do
// one element per trip.
template <class AmpT>
struct Unroll {
    static constexpr uint32_t value = std::is_same<AmpT, glb_d2>::value ? 4u : 1u;
};

// Team sum of N doubles, identical bits on every lane: lane butterfly, then
// (workgroup teams) waves summed in index order through LDS scratch `red`
// (16 x 4 doubles), fenced by a barrier on both sides.
template <bool BT, int N>
CSVM_INLINE void team_reduce(double (&v)[N], const TeamGeom& g, lds_f64* red) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] += __shfl_xor(v[i], off, 64);
    }
    if (BT) {
        if (g.nwaves > 1) {
            __syncthreads();
            if ((g.lane & 63u) == 0) {
#pragma unroll
                for (int i = 0; i < N; ++i) red[g.wave * 4 + i] = v[i];
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double acc = red[i];
                for (uint32_t w = 1; w < g.nwaves; ++w) acc += red[w * 4 + i];
                v[i] = acc;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = unid(v[i]);
}

// Sweep arguments are re-asserted uniform (readfirstlane) so that loop
// control and coefficient loads stay scalar whichever registers the caller
// held them in.
CSVM_INLINE TeamGeom uni_geom(TeamGeom g) {
    TeamGeom r;
    r.lane = g.lane;
    r.size = uni(g.size);
    r.wave = uni(g.wave);
    r.nwaves = uni(g.nwaves);
    return r;
}

// --------------------------------------------------------------------------
// Permutation / sign sweeps (svm_kernels.inl:124-635)
// --------------------------------------------------------------------------
enum : uint32_t { PERM_CNOT = 0, PERM_CZ = 1, PERM_SWAP = 2 };

template <bool BT, class AmpT>
CSVM_SWEEP void sweep_two_axis(AmpT* amp, uint32_t k, uint32_t a, uint32_t b, uint32_t kind, TeamGeom g) {
    k = uni(k); a = uni(a); b = uni(b); kind = uni(kind); g = uni_geom(g);
    uint32_t n = 1u << (k - 2), ab = 1u << a, bb = 1u << b;
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    if (kind == PERM_CNOT) {  // a = control, b = target
        for (uint32_t i = g.lane; i < n; i += g.size) {
            uint32_t b0 = insert_two_zero_bits(i, lo, hi) | ab, b1 = b0 | bb;
            dv2 x = amp[b0], y = amp[b1];
            amp[b0] = y;
            amp[b1] = x;
        }
    } else if (kind == PERM_CZ) {
        for (uint32_t i = g.lane; i < n; i += g.size) {
            uint32_t idx = insert_two_zero_bits(i, lo, hi) | ab | bb;
            amp[idx] = cneg(amp[idx]);
        }
    } else {
        for (uint32_t i = g.lane; i < n; i += g.size) {
            uint32_t base = insert_two_zero_bits(i, lo, hi);
            dv2 x = amp[base | ab], y = amp[base | bb];
            amp[base | ab] = y;
            amp[base | bb] = x;
        }
    }
    team_sync<BT>();
}

// MULTI_CNOT (is_cz = 0): swap the target pair where parity(idx & mask) is
// odd.  MULTI_CZ (is_cz = 1): negate entries with the control set and odd
// parity(idx & mask).  `mask` is already restricted to active axes.
template <bool BT, class AmpT>
CSVM_SWEEP void sweep_multi(AmpT* amp, uint32_t k, uint32_t axis, uint32_t mask, uint32_t is_cz, TeamGeom g) {
    k = uni(k); axis = uni(axis); mask = uni(mask); is_cz = uni(is_cz); g = uni_geom(g);
    uint32_t half = 1u << (k - 1), ab = 1u << axis;
    if (!is_cz) {
        for (uint32_t i = g.lane; i < half; i += g.size) {
            uint32_t a = insert_zero_bit(i, axis);
            if (__popc(a & mask) & 1u) {
                dv2 x = amp[a], y = amp[a | ab];
                amp[a] = y;
                amp[a | ab] = x;
            }
        }
    } else {
        for (uint32_t i = g.lane; i < half; i += g.size) {
            uint32_t a = insert_zero_bit(i, axis) | ab;
            if (__popc(a & mask) & 1u) amp[a] = cneg(amp[a]);
        }
    }
    team_sync<BT>();
}

// --------------------------------------------------------------------------
// Single-axis sweeps: H butterfly (:639-744), diag(1, ph) (:764-850),
// dense 2x2 (:1091-1297)
// --------------------------------------------------------------------------
template <bool BT, class AmpT>
CSVM_SWEEP void sweep_h(AmpT* amp, uint32_t k, uint32_t v, TeamGeom g) {
    k = uni(k); v = uni(v); g = uni_geom(g);
    uint32_t half = 1u << (k - 1), vb = 1u << v;
    constexpr uint32_t U = Unroll<AmpT>::value;
    const uint32_t step = U * g.size;
    for (uint32_t j0 = g.lane; j0 < half; j0 += step) {
        dv2 a[U], b[U];
        uint32_t i0[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            uint32_t i = j0 + u * g.size;
            i0[u] = insert_zero_bit(i, v);
            if (i < half) {
                a[u] = amp[i0[u]];
                b[u] = amp[i0[u] | vb];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (j0 + u * g.size < half) {
                amp[i0[u]] = cscale(cadd(a[u], b[u]), kInvSqrt2);
                amp[i0[u] | vb] = cscale(csub(a[u], b[u]), kInvSqrt2);
            }
        }
    }
    team_sync<BT>();
}

template <bool BT, class AmpT>
CSVM_SWEEP void sweep_phase(AmpT* amp, uint32_t k, uint32_t v, double pr, double pi, TeamGeom g) {
    k = uni(k); v = uni(v); pr = unid(pr); pi = unid(pi); g = uni_geom(g);
    uint32_t half = 1u << (k - 1), vb = 1u << v;
    dv2 ph = mk(pr, pi);
    constexpr uint32_t U = Unroll<AmpT>::value;
    const uint32_t step = U * g.size;
    for (uint32_t i0 = g.lane; i0 < half; i0 += step) {
        dv2 x[U];
        uint32_t idx[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            uint32_t i = i0 + u * g.size;
            idx[u] = insert_zero_bit(i, v) | vb;
            if (i < half) x[u] = amp[idx[u]];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (i0 + u * g.size < half) amp[idx[u]] = cmul(x[u], ph);
    }
    team_sync<BT>();
}

// m: row-major {a, b, c, d} as (re, im) pairs
template <bool BT, class AmpT>
CSVM_SWEEP void sweep_u2(AmpT* amp, uint32_t k, uint32_t v, cptr<double> m, TeamGeom g) {
    k = uni(k); v = uni(v); g = uni_geom(g);
    dv2 m00 = mk(m[0], m[1]), m01 = mk(m[2], m[3]);
    dv2 m10 = mk(m[4], m[5]), m11 = mk(m[6], m[7]);
    uint32_t half = 1u << (k - 1), vb = 1u << v;
    constexpr uint32_t U = Unroll<AmpT>::value;
    const uint32_t step = U * g.size;
    for (uint32_t j0 = g.lane; j0 < half; j0 += step) {
        dv2 a[U], b[U];
        uint32_t i0[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            uint32_t i = j0 + u * g.size;
            i0[u] = insert_zero_bit(i, v);
            if (i < half) {
                a[u] = amp[i0[u]];
                b[u] = amp[i0[u] | vb];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (j0 + u * g.size < half) {
                amp[i0[u]] = cadd(cmul(m00, a[u]), cmul(m01, b[u]));
                amp[i0[u] | vb] = cadd(cmul(m10, a[u]), cmul(m11, b[u]));
            }
        }
    }
    team_sync<BT>();
}

// Dense 4x4 on quads {base, |lo, |hi, |lo|hi} (:1310-1531).  The matrix is
// re-read (scalar loads) row by row inside the loop rather than held in 64
// registers.
template <bool BT, class AmpT>
CSVM_SWEEP void sweep_u4(AmpT* amp, uint32_t k, uint32_t lo, uint32_t hi, cptr<double> m, TeamGeom g) {
    k = uni(k); lo = uni(lo); hi = uni(hi); g = uni_geom(g);
    uint32_t n = 1u << (k - 2), lb = 1u << lo, hb = 1u << hi;
    for (uint32_t i = g.lane; i < n; i += g.size) {
        uint32_t base = insert_two_zero_bits(i, lo, hi);
        uint32_t ix0 = base, ix1 = base | lb, ix2 = base | hb, ix3 = base | lb | hb;
        dv2 in0 = amp[ix0], in1 = amp[ix1], in2 = amp[ix2], in3 = amp[ix3];
        dv2 o[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cptr<double> row = opaque(m) + 8 * r;
            dv2 acc = cmul(mk(row[0], row[1]), in0);
            acc = cadd(acc, cmul(mk(row[2], row[3]), in1));
            acc = cadd(acc, cmul(mk(row[4], row[5]), in2));
            acc = cadd(acc, cmul(mk(row[6], row[7]), in3));
            o[r] = acc;
        }
        amp[ix0] = o[0];
        amp[ix1] = o[1];
        amp[ix2] = o[2];
        amp[ix3] = o[3];
    }
    team_sync<BT>();
}

// A run of permutation/phase opcodes in one out-of-place pass: dst[y] is the
// entry that the run carries to index y.  Its source x is found by undoing
// the permutations (each is its own inverse) in reverse order; the entry is
// then walked forward through the run, multiplied by each phase at the index
// it holds at that point, exactly as the one-at-a-time sweeps would (same
// products in the same order, so the values are bit-identical).  Each index
// stays below 2^k: the opcodes only exchange or flip bits below k.
typedef __attribute__((address_space(3))) MonoOp lds_mono;
CSVM_INLINE void store_mono(lds_mono* p, const MonoOp& o) {
    p->kind = o.kind;
    p->a = o.a;
    p->b = o.b;
    p->pm = o.pm;
    p->px = o.px;
    p->qa = o.qa;
    p->qm = o.qm;
    p->pk = o.pk;
    p->pr = o.pr;
    p->pi = o.pi;
}
// The run's permutations undone in reverse order (each is its own inverse):
// the source index of each of the N entries.  One uniform test per opcode;
// the entries' work is branch-free.
template <uint32_t N>
CSVM_INLINE void mono_sources(const lds_mono* ops, uint32_t n, uint32_t (&x)[N]) {
    for (uint32_t j = n; j-- > 0;) {
        const uint32_t pm = uni(ops[j].pm), px = uni(ops[j].px);
        if (pm) {
#pragma unroll
            for (uint32_t e = 0; e < N; ++e) x[e] = (__popc(x[e] & pm) & 1u) ? x[e] ^ px : x[e];
        }
    }
}
// The entries walked forward through the run: each opcode's phase at the
// index the entry holds at that point, then its permutation.
template <uint32_t N>
CSVM_INLINE void mono_walk(const lds_mono* ops, uint32_t n, uint32_t (&x)[N], dv2 (&v)[N]) {
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t pk = uni(ops[j].pk), pm = uni(ops[j].pm), px = uni(ops[j].px);
        if (pk) {
            const uint32_t qa = uni(ops[j].qa), qm = uni(ops[j].qm);
            if (pk == 1) {
#pragma unroll
                for (uint32_t e = 0; e < N; ++e)
                    if ((x[e] & qa) && (__popc(x[e] & qm) & 1u)) v[e] = cneg(v[e]);
            } else {
                const dv2 ph = mk(unid(ops[j].pr), unid(ops[j].pi));
#pragma unroll
                for (uint32_t e = 0; e < N; ++e)
                    if ((x[e] & qa) && (__popc(x[e] & qm) & 1u)) v[e] = cmul(v[e], ph);
            }
        }
        if (pm) {
#pragma unroll
            for (uint32_t e = 0; e < N; ++e) x[e] = (__popc(x[e] & pm) & 1u) ? x[e] ^ px : x[e];
        }
    }
}
template <bool BT, class AmpT>
CSVM_SWEEP void sweep_mono(const AmpT* src, AmpT* dst, uint32_t k, const lds_mono* ops, uint32_t n, TeamGeom g) {
    k = uni(k); n = uni(n); g = uni_geom(g);
    const uint32_t K = 1u << k;
    constexpr uint32_t U = Unroll<AmpT>::value;
    const uint32_t step = U * g.size;
    for (uint32_t y0 = g.lane; y0 < K; y0 += step) {
        uint32_t x[U];
        dv2 v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) x[u] = y0 + u * g.size;
        mono_sources(ops, n, x);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (y0 + u * g.size < K) v[u] = src[x[u]];
        mono_walk(ops, n, x, v);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (y0 + u * g.size < K) dst[y0 + u * g.size] = v[u];
    }
    team_sync<BT>();
}

// Pair layouts of the active measurements' probability sums.
enum : uint32_t { PROB_DIAG = 0, PROB_INTERFERE = 1, PROB_SWAP_INTERFERE = 2 };

// Entry t 2^k0 + i of a state after m expansions b = 0 .. m-1 (phase z_b
// when bit b of `mask` is set, else a plain copy) from entry i of the state
// before them (t < 2^m): a times z_b for each set bit b of t, lowest first --
// the products the one-at-a-time EXPAND sweeps form.
CSVM_INLINE dv2 expand_entry(dv2 a, uint32_t t, const dv2 (&z)[kMaxFusedExpand], uint32_t mask) {
#pragma unroll
    for (uint32_t b = 0; b < kMaxFusedExpand; ++b)
        if (((t >> b) & 1u) && ((mask >> b) & 1u)) a = cmul(a, z[b]);
    return a;
}

// Read views of the state a fused measurement pass gathers from: the stored
// entries, or the entries that m pending expansions of a stored state of
// 2^k0 entries would write, computed on the fly by expand_entry, so the pass
// reads 2^k0 entries instead of 2^(k0+m) and the expansions are never
// written.
template <class AmpT>
struct StoredView {
    const AmpT* a;
    CSVM_INLINE dv2 operator()(uint32_t x) const { return a[x]; }
};
template <class AmpT>
struct ExpandView {
    const AmpT* a;
    uint32_t k0, lo, m, mask;
    dv2 z[kMaxFusedExpand];
    CSVM_INLINE ExpandView(const AmpT* src, uint32_t k0_, uint32_t m_, const lds_f64* ph, uint32_t mask_)
        : a(src), k0(uni(k0_)), lo((1u << uni(k0_)) - 1u), m(uni(m_)), mask(uni(mask_)) {
#pragma unroll
        for (uint32_t b = 0; b < kMaxFusedExpand; ++b)
            z[b] = b < m ? mk(unid(ph[2 * b]), unid(ph[2 * b + 1])) : mk(1.0, 0.0);
    }
    CSVM_INLINE dv2 operator()(uint32_t x) const { return expand_entry(a[x & lo], x >> k0, z, mask); }
};

// A deferred permutation/phase run and the outcome probabilities of the
// active measurement that follows it, in one pass.  dst receives the run's
// result, entry for entry as sweep_mono writes it; the sums are sweep_probs'
// over that result, each lane visiting the same pair indices in the same
// order, so the measurement reads the state once instead of twice.  `src`
// is a StoredView or an ExpandView (pending expansions read through).
template <bool BT, class View, class AmpT>
CSVM_SWEEP double2 sweep_mono_probs(const View src, AmpT* dst, uint32_t k, const lds_mono* ops, uint32_t n,
                                    uint32_t kind, uint32_t fa, TeamGeom g, lds_f64* red) {
    k = uni(k); n = uni(n); kind = uni(kind); fa = uni(fa); g = uni_geom(g);
    const uint32_t half = 1u << (k - 1), t = k - 1;
    // pair axis: fa for SWAP_MEAS_INTERFERE, else the top axis t
    const uint32_t fs = kind == PROB_SWAP_INTERFERE ? fa : t, fb = 1u << fs;
    // pairs in flight per lane: as many loads as sweep_mono keeps
    constexpr uint32_t U = Unroll<AmpT>::value > 1 ? Unroll<AmpT>::value / 2 : 1;
    double v[2] = {0.0, 0.0};
    for (uint32_t i0 = g.lane; i0 < half; i0 += U * g.size) {
        uint32_t y[2 * U], x[2 * U];
        dv2 a[2 * U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            // the pair sweep_probs reads for index i (in range or not,
            // y stays below 2^k); with fs = k-1 this is (i, i + half)
            const uint32_t i = (i0 + u * g.size) & (half - 1u);
            y[2 * u] = (i & ~fb) | (((i >> fs) & 1u) << t);
            y[2 * u + 1] = y[2 * u] | fb;
            x[2 * u] = y[2 * u];
            x[2 * u + 1] = y[2 * u + 1];
        }
        mono_sources(ops, n, x);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (i0 + u * g.size < half) {
                a[2 * u] = src(x[2 * u]);
                a[2 * u + 1] = src(x[2 * u + 1]);
            }
        mono_walk(ops, n, x, a);
        if (kind == PROB_DIAG) {
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                if (i0 + u * g.size < half) {
                    const dv2 p = a[2 * u], q = a[2 * u + 1];
                    dst[y[2 * u]] = p;
                    dst[y[2 * u + 1]] = q;
                    v[0] += p.x * p.x + p.y * p.y;
                    v[1] += q.x * q.x + q.y * q.y;
                }
        } else {
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                if (i0 + u * g.size < half) {
                    const dv2 p = a[2 * u], q = a[2 * u + 1];
                    dst[y[2 * u]] = p;
                    dst[y[2 * u + 1]] = q;
                    const dv2 s = cadd(p, q), d = csub(p, q);
                    v[0] += s.x * s.x + s.y * s.y;
                    v[1] += d.x * d.x + d.y * d.y;
                }
        }
    }
    team_sync<BT>();
    team_reduce<BT, 2>(v, g, red);
    return make_double2(v[0], v[1]);
}

// A deferred run, the outcome probabilities of the active measurement that
// follows it, and both of that measurement's folded states, in one pass.
// For each pair the probability sum visits (output index i < 2^(k-1)),
// c0[i] and c1[i] receive the entry that the measurement leaves at i for
// outcome 0 and for outcome 1: the run's result entries p (lower) and q
// (upper) for the diagonal kind (keep the lower half or take the upper);
// (p + q) / sqrt 2 and (p - q) / sqrt 2 for the interfere kinds, with the
// index mapping of sweep_swap_fold for SWAP_MEAS_INTERFERE.  These are the
// products sweep_take_upper, sweep_fold and sweep_swap_fold form from the
// stored run result, so the state kept is bit-identical to gathering,
// measuring and folding in place; the sums are sweep_mono_probs'.  One read
// and one write of 2^k entries replace the gather's read and write plus the
// fold's read of 2^k and write of 2^(k-1); through an ExpandView the read
// is of the 2^(k-m) entries before m pending expansions.
template <bool BT, class View, class AmpT>
CSVM_SWEEP double2 sweep_mono_cands(const View src, AmpT* c0, AmpT* c1, uint32_t k, const lds_mono* ops, uint32_t n,
                                    uint32_t kind, uint32_t fa, TeamGeom g, lds_f64* red) {
    k = uni(k); n = uni(n); kind = uni(kind); fa = uni(fa); g = uni_geom(g);
    const uint32_t half = 1u << (k - 1), t = k - 1;
    // pair axis: fa for SWAP_MEAS_INTERFERE, else the top axis t
    const uint32_t fs = kind == PROB_SWAP_INTERFERE ? fa : t, fb = 1u << fs;
    constexpr uint32_t U = Unroll<AmpT>::value > 1 ? Unroll<AmpT>::value / 2 : 1;
    double v[2] = {0.0, 0.0};
    for (uint32_t i0 = g.lane; i0 < half; i0 += U * g.size) {
        uint32_t y[2 * U], x[2 * U];
        dv2 a[2 * U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t i = (i0 + u * g.size) & (half - 1u);
            y[2 * u] = (i & ~fb) | (((i >> fs) & 1u) << t);
            y[2 * u + 1] = y[2 * u] | fb;
            x[2 * u] = y[2 * u];
            x[2 * u + 1] = y[2 * u + 1];
        }
        mono_sources(ops, n, x);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (i0 + u * g.size < half) {
                a[2 * u] = src(x[2 * u]);
                a[2 * u + 1] = src(x[2 * u + 1]);
            }
        mono_walk(ops, n, x, a);
        if (kind == PROB_DIAG) {
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t i = i0 + u * g.size;
                if (i < half) {
                    const dv2 p = a[2 * u], q = a[2 * u + 1];
                    c0[i] = p;
                    c1[i] = q;
                    v[0] += p.x * p.x + p.y * p.y;
                    v[1] += q.x * q.x + q.y * q.y;
                }
            }
        } else {
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t i = i0 + u * g.size;
                if (i < half) {
                    const dv2 p = a[2 * u], q = a[2 * u + 1];
                    c0[i] = mk((p.x + 1.0 * q.x) * kInvSqrt2, (p.y + 1.0 * q.y) * kInvSqrt2);
                    c1[i] = mk((p.x + -1.0 * q.x) * kInvSqrt2, (p.y + -1.0 * q.y) * kInvSqrt2);
                    const dv2 s = cadd(p, q), d = csub(p, q);
                    v[0] += s.x * s.x + s.y * s.y;
                    v[1] += d.x * d.x + d.y * d.y;
                }
            }
        }
    }
    team_sync<BT>();
    team_reduce<BT, 2>(v, g, red);
    return make_double2(v[0], v[1]);
}

// EXPAND: amp[i + 2^k] = amp[i] * ph for i < 2^k (:876-945)
template <bool BT, class AmpT>
CSVM_SWEEP void sweep_expand(AmpT* amp, uint32_t k, double pr, double pi, uint32_t has_phase, TeamGeom g) {
    k = uni(k); pr = unid(pr); pi = unid(pi); has_phase = uni(has_phase); g = uni_geom(g);
    uint32_t half = 1u << k;
    dv2 ph = has_phase ? mk(pr, pi) : mk(1.0, 0.0);
    constexpr uint32_t U = Unroll<AmpT>::value;
    const uint32_t step = U * g.size;
    for (uint32_t i0 = g.lane; i0 < half; i0 += step) {
        dv2 v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            uint32_t i = i0 + u * g.size;
            if (i < half) v[u] = amp[i];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            uint32_t i = i0 + u * g.size;
            if (i < half) amp[i + half] = has_phase ? cmul(v[u], ph) : v[u];
        }
    }
    team_sync<BT>();
}

// Several consecutive EXPANDs in one pass (the opcodes between them touch
// only the frame, gamma, the PRNG or records).  Expansion b (0-based,
// phase z_b when bit b of `mask` is set, else a plain copy) maps a state of
// 2^(k0+b) entries to twice that, so
//     amp[i + t 2^k0] = (((amp[i] z_b1) z_b2) ... z_bj)   for bits b1 < ... < bj of t,
// the very products the one-at-a-time sweeps form: one read and 2^m - 1
// writes per source entry instead of 2^(m+1) - 2 of each.
template <bool BT, class AmpT>
CSVM_SWEEP void sweep_expand_multi(AmpT* amp, uint32_t k0, uint32_t m, const lds_f64* ph, uint32_t mask, TeamGeom g) {
    k0 = uni(k0); m = uni(m); mask = uni(mask); g = uni_geom(g);
    dv2 z[kMaxFusedExpand];
#pragma unroll
    for (uint32_t b = 0; b < kMaxFusedExpand; ++b) z[b] = b < m ? mk(unid(ph[2 * b]), unid(ph[2 * b + 1])) : mk(1.0, 0.0);
    const uint32_t K = 1u << k0, T = 1u << m;
    for (uint32_t i = g.lane; i < K; i += g.size) {
        const dv2 a = amp[i];
        for (uint32_t t = 1; t < T; ++t) amp[i + t * K] = expand_entry(a, t, z, mask);
    }
    team_sync<BT>();
}

template <bool BT, class AmpT>
CSVM_SWEEP void sweep_scale(AmpT* amp, uint32_t k, double s, TeamGeom g) {
    k = uni(k); s = unid(s); g = uni_geom(g);
    uint32_t n = 1u << k;
    constexpr uint32_t U = Unroll<AmpT>::value;
    for (uint32_t i0 = g.lane; i0 < n; i0 += U * g.size) {
        dv2 x[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (i0 + u * g.size < n) x[u] = amp[i0 + u * g.size];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (i0 + u * g.size < n) amp[i0 + u * g.size] = cscale(x[u], s);
    }
    team_sync<BT>();
}

// --------------------------------------------------------------------------
// Measurement reductions and compaction (:1573-1804)
// --------------------------------------------------------------------------

// Returns (p0, p1) of the branch pair on axis k-1 (or the swapped mapping
// for SWAP_MEAS_INTERFERE with from-axis fa).
template <bool BT, class AmpT>
CSVM_SWEEP double2 sweep_probs(AmpT* amp, uint32_t k, uint32_t kind, uint32_t fa, TeamGeom g, lds_f64* red) {
    k = uni(k); kind = uni(kind); fa = uni(fa); g = uni_geom(g);
    uint32_t half = 1u << (k - 1);
    double v[2] = {0.0, 0.0};
    if (kind == PROB_DIAG) {
        for (uint32_t i = g.lane; i < half; i += g.size) {
            dv2 a = amp[i], b = amp[i + half];
            v[0] += a.x * a.x + a.y * a.y;
            v[1] += b.x * b.x + b.y * b.y;
        }
    } else if (kind == PROB_INTERFERE) {
        for (uint32_t i = g.lane; i < half; i += g.size) {
            dv2 a = amp[i], b = amp[i + half];
            dv2 s = cadd(a, b), d = csub(a, b);
            v[0] += s.x * s.x + s.y * s.y;
            v[1] += d.x * d.x + d.y * d.y;
        }
    } else {
        uint32_t t = k - 1, fb = 1u << fa;
        for (uint32_t idx = g.lane; idx < half; idx += g.size) {
            uint32_t bf = (idx >> fa) & 1u, base = (idx & ~fb) | (bf << t);
            dv2 a = amp[base], b = amp[base | fb];
            dv2 s = cadd(a, b), d = csub(a, b);
            v[0] += s.x * s.x + s.y * s.y;
            v[1] += d.x * d.x + d.y * d.y;
        }
    }
    team_reduce<BT, 2>(v, g, red);
    return make_double2(v[0], v[1]);
}

// Diagonal outcome b=1: keep the upper half (copy into the lower half).
template <bool BT, class AmpT>
CSVM_SWEEP void sweep_take_upper(AmpT* amp, uint32_t k, TeamGeom g) {
    k = uni(k); g = uni_geom(g);
    uint32_t half = 1u << (k - 1);
    constexpr uint32_t U = Unroll<AmpT>::value;
    for (uint32_t i0 = g.lane; i0 < half; i0 += U * g.size) {
        dv2 x[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (i0 + u * g.size < half) x[u] = amp[i0 + u * g.size + half];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (i0 + u * g.size < half) amp[i0 + u * g.size] = x[u];
    }
    team_sync<BT>();
}

// Interfere fold v'[i] = (v[i] +/- v[i+half]) / sqrt(2) (:1672-1687)
template <bool BT, class AmpT>
CSVM_SWEEP void sweep_fold(AmpT* amp, uint32_t k, double s, TeamGeom g) {
    k = uni(k); s = unid(s); g = uni_geom(g);
    uint32_t half = 1u << (k - 1);
    constexpr uint32_t U = Unroll<AmpT>::value;
    for (uint32_t i0 = g.lane; i0 < half; i0 += U * g.size) {
        dv2 a[U], b[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            uint32_t i = i0 + u * g.size;
            if (i < half) {
                a[u] = amp[i];
                b[u] = amp[i + half];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            uint32_t i = i0 + u * g.size;
            if (i < half) amp[i] = mk((a[u].x + s * b[u].x) * kInvSqrt2, (a[u].y + s * b[u].y) * kInvSqrt2);
        }
    }
    team_sync<BT>();
}

// SWAP_MEAS_INTERFERE fold with the swapped index mapping (:1768-1780).
// Output idx < 2^t reads (base, base|f), base = (idx & ~f) | (bit_f(idx) << t).
// Each lane owns the pair of outputs {i0, i0|f} (bit f of i0 clear); their
// sources {i0, i0|f, i0|2^t, i0|2^t|f} belong to no other lane, so reading
// all four before writing reproduces the serial in-place result.
template <bool BT, class AmpT>
CSVM_SWEEP void sweep_swap_fold(AmpT* amp, uint32_t k, uint32_t fa, double s, TeamGeom g) {
    k = uni(k); fa = uni(fa); s = unid(s); g = uni_geom(g);
    uint32_t t = k - 1, quarter = 1u << (t - 1), fb = 1u << fa, tb = 1u << t;
    for (uint32_t j = g.lane; j < quarter; j += g.size) {
        uint32_t i0 = insert_zero_bit(j, fa), i1 = i0 | fb;
        dv2 a0 = amp[i0], a1 = amp[i1], b0 = amp[i0 | tb], b1 = amp[i1 | tb];
        amp[i0] = mk((a0.x + s * a1.x) * kInvSqrt2, (a0.y + s * a1.y) * kInvSqrt2);
        amp[i1] = mk((b0.x + s * b1.x) * kInvSqrt2, (b0.y + s * b1.y) * kInvSqrt2);
    }
    team_sync<BT>();
}

// EXP_VAL sums (:2139-2152): returns (Re num, Im num, denominator).
template <bool BT, class AmpT>
CSVM_SWEEP double3 sweep_exp_val(AmpT* amp, uint32_t k, uint32_t xa, uint32_t za, TeamGeom g, lds_f64* red) {
    k = uni(k); xa = uni(xa); za = uni(za); g = uni_geom(g);
    uint32_t dim = 1u << k;
    double v[3] = {0.0, 0.0, 0.0};
    for (uint32_t j = g.lane; j < dim; j += g.size) {
        dv2 a = amp[j], b = amp[j ^ xa];
        v[2] += a.x * a.x + a.y * a.y;
        double zs = (__popc(j & za) & 1u) ? -1.0 : 1.0;
        double cr = zs * a.x, ci = zs * a.y;
        v[0] += b.x * cr + b.y * ci;
        v[1] += b.x * ci - b.y * cr;
    }
    team_reduce<BT, 3>(v, g, red);
    return make_double3(v[0], v[1], v[2]);
}

}  // namespace csvm
