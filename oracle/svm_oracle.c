/*
 * svm_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A scalar CPU restatement of the reference Schrodinger VM, used as the
 * parity checker for the HIP executor.  Only tests, the smoke check and the
 * CPU-baseline leg of the benchmark load this file; the product path never
 * does.  Each function names the reference file:line whose behaviour it
 * restates (paths relative to the reference repository root).
 *
 * Pinning (what checks each part, see tests/test_oracle_*.py):
 *   xoshiro256++/SplitMix64 ... published reference outputs (prng.di.unimi.it
 *                               algorithm) + a pure-Python re-derivation
 *   frame ops / APPLY_PAULI ... reference KATs tests/test_svm.cc:93-272,
 *                               839-1000 restated + brute-force P.U = U.P'
 *   array ops, EXPAND*, U2/U4 . dense 2^n matrices on random states
 *   measurements / forced .... closed-form Born probabilities, norm and
 *                               gamma-scale invariants, KATs
 *                               tests/test_svm.cc:516-788, 1097-1571,
 *                               tests/test_forced_kernels.cc
 *   noise / readout .......... NOISE_BLOCK == NOISE sequence, binomial
 *                               5-sigma bounds on firing rates
 *   EXP_VAL .................. dense Pauli-matrix expectation
 */
#include "clifft_svm.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

static const double kInvSqrt2 = 0.70710678118654752440; /* util/constants.h:8 */
static const double kDustEpsilon = 1e-18;               /* svm/svm_internal.h:46 */

/* ------------------------------------------------------------------------
 * PRNG: xoshiro256++ 1.0 seeded by SplitMix64 (svm/svm.h:40-90)
 * ---------------------------------------------------------------------- */
typedef struct orc_rng {
    uint64_t s[4];
} orc_rng;

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* svm/svm.h:84-89 */
static inline uint64_t splitmix64_next(uint64_t* state) {
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* svm/svm.h:46-52 */
static void rng_seed(orc_rng* r, uint64_t seed) {
    uint64_t z = seed;
    r->s[0] = splitmix64_next(&z);
    r->s[1] = splitmix64_next(&z);
    r->s[2] = splitmix64_next(&z);
    r->s[3] = splitmix64_next(&z);
}

/* Per-shot stream used by the device executor: shot s takes SplitMix64
 * outputs 4s..4s+3 of the stream seeded by `seed` (shot 0 == rng_seed). */
static void rng_seed_shot(orc_rng* r, uint64_t seed, uint64_t shot) {
    uint64_t z = seed + shot * 4ULL * 0x9e3779b97f4a7c15ULL;
    r->s[0] = splitmix64_next(&z);
    r->s[1] = splitmix64_next(&z);
    r->s[2] = splitmix64_next(&z);
    r->s[3] = splitmix64_next(&z);
}

/* svm/svm.h:66-79 */
static inline uint64_t rng_next(orc_rng* r) {
    uint64_t* s = r->s;
    const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

/* svm/svm.h:150 */
static inline double rng_double(orc_rng* r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

/* ------------------------------------------------------------------------
 * State (svm/svm.h:105-289, svm/svm_state.cc:36-267)
 * ---------------------------------------------------------------------- */
typedef struct orc_state {
    uint32_t peak_rank, num_qubits, words;
    uint32_t n_meas, n_det, n_obs, n_ev;
    double* v; /* 2 * 2^peak_rank doubles */
    uint64_t* px;
    uint64_t* pz;
    double g_re, g_im;
    uint32_t active_k;
    uint32_t discarded;
    uint8_t* meas;
    uint8_t* det;
    uint8_t* obs;
    double* ev;
    uint32_t next_noise_idx;
    uint64_t dust_clamps;
    orc_rng rng;
    /* forced faults (svm/svm.h:208-227) */
    uint32_t ff_active;
    uint32_t* ff_noise;
    uint32_t ff_n_noise, ff_noise_pos;
    uint32_t* ff_ro;
    uint32_t ff_n_ro, ff_ro_pos;
    /* forced records (svm/svm.h:271-288) */
    const uint8_t* forced_record;
    double forced_log_prob;
    uint32_t forced_reachable;
} orc_state;

typedef clifft_program_view prog_t;

ORC_EXPORT orc_state* orc_state_new(uint32_t peak_rank, uint32_t n_meas, uint32_t num_qubits,
                                    uint32_t n_det, uint32_t n_obs, uint32_t n_ev, uint64_t seed) {
    if (peak_rank >= 63) return NULL; /* svm_state.cc:38-41 */
    orc_state* st = (orc_state*)calloc(1, sizeof(orc_state));
    st->peak_rank = peak_rank;
    /* svm_state.cc:57: frame width falls back to peak_rank */
    st->num_qubits = num_qubits > 0 ? num_qubits : (peak_rank > 1 ? peak_rank : 1);
    st->words = (st->num_qubits + 63) / 64;
    st->n_meas = n_meas;
    st->n_det = n_det;
    st->n_obs = n_obs;
    st->n_ev = n_ev;
    st->v = (double*)calloc((size_t)2 << peak_rank, sizeof(double));
    st->px = (uint64_t*)calloc(st->words, 8);
    st->pz = (uint64_t*)calloc(st->words, 8);
    st->meas = (uint8_t*)calloc(n_meas + 1, 1);
    st->det = (uint8_t*)calloc(n_det + 1, 1);
    st->obs = (uint8_t*)calloc(n_obs + 1, 1);
    st->ev = (double*)calloc(n_ev + 1, sizeof(double));
    st->v[0] = 1.0;
    st->g_re = 1.0;
    st->g_im = 0.0;
    st->forced_reachable = 1;
    rng_seed(&st->rng, seed);
    return st;
}

ORC_EXPORT void orc_state_free(orc_state* st) {
    if (!st) return;
    free(st->v);
    free(st->px);
    free(st->pz);
    free(st->meas);
    free(st->det);
    free(st->obs);
    free(st->ev);
    free(st->ff_noise);
    free(st->ff_ro);
    free(st);
}

/* svm_state.cc:220-267 */
ORC_EXPORT void orc_state_reset(orc_state* st) {
    uint64_t active_size = st->active_k > 0 ? (1ULL << st->active_k) : 1;
    memset(st->v, 0, active_size * 16);
    st->v[0] = 1.0;
    st->v[1] = 0.0;
    memset(st->px, 0, st->words * 8);
    memset(st->pz, 0, st->words * 8);
    st->g_re = 1.0;
    st->g_im = 0.0;
    st->active_k = 0;
    if (st->discarded) {
        memset(st->meas, 0, st->n_meas);
        memset(st->det, 0, st->n_det);
    }
    st->discarded = 0;
    memset(st->obs, 0, st->n_obs);
    memset(st->ev, 0, st->n_ev * sizeof(double));
    st->ff_noise_pos = 0;
    st->ff_ro_pos = 0;
    st->forced_record = NULL;
    st->forced_log_prob = 0.0;
    st->forced_reachable = 1;
}

/* Field accessors for the ctypes test harness. */
ORC_EXPORT double* orc_state_v(orc_state* st) { return st->v; }
ORC_EXPORT uint64_t* orc_state_px(orc_state* st) { return st->px; }
ORC_EXPORT uint64_t* orc_state_pz(orc_state* st) { return st->pz; }
ORC_EXPORT uint8_t* orc_state_meas(orc_state* st) { return st->meas; }
ORC_EXPORT uint8_t* orc_state_det(orc_state* st) { return st->det; }
ORC_EXPORT uint8_t* orc_state_obs(orc_state* st) { return st->obs; }
ORC_EXPORT double* orc_state_ev(orc_state* st) { return st->ev; }
ORC_EXPORT uint64_t* orc_state_rng(orc_state* st) { return st->rng.s; }
ORC_EXPORT void orc_state_get_gamma(const orc_state* st, double* out) {
    out[0] = st->g_re;
    out[1] = st->g_im;
}
ORC_EXPORT void orc_state_set_gamma(orc_state* st, double re, double im) {
    st->g_re = re;
    st->g_im = im;
}
ORC_EXPORT uint32_t orc_state_get_k(const orc_state* st) { return st->active_k; }
ORC_EXPORT void orc_state_set_k(orc_state* st, uint32_t k) { st->active_k = k; }
ORC_EXPORT uint32_t orc_state_discarded(const orc_state* st) { return st->discarded; }
ORC_EXPORT uint64_t orc_state_dust_clamps(const orc_state* st) { return st->dust_clamps; }
ORC_EXPORT uint32_t orc_state_next_noise(const orc_state* st) { return st->next_noise_idx; }
ORC_EXPORT void orc_state_set_next_noise(orc_state* st, uint32_t v) { st->next_noise_idx = v; }
ORC_EXPORT void orc_state_reseed(orc_state* st, uint64_t seed) { rng_seed(&st->rng, seed); }
ORC_EXPORT double orc_state_random_double(orc_state* st) { return rng_double(&st->rng); }
ORC_EXPORT double orc_state_forced_log_prob(const orc_state* st) { return st->forced_log_prob; }
ORC_EXPORT uint32_t orc_state_forced_reachable(const orc_state* st) { return st->forced_reachable; }
ORC_EXPORT void orc_state_set_forced_record(orc_state* st, const uint8_t* rec) {
    st->forced_record = rec;
}

/* ------------------------------------------------------------------------
 * Scalar helpers
 * ---------------------------------------------------------------------- */
static inline int bit_get(const uint64_t* m, uint32_t i) { return (int)((m[i >> 6] >> (i & 63)) & 1ULL); }
static inline void bit_set(uint64_t* m, uint32_t i, int v) {
    uint64_t b = 1ULL << (i & 63);
    if (v)
        m[i >> 6] |= b;
    else
        m[i >> 6] &= ~b;
}
static inline void bit_xor(uint64_t* m, uint32_t i, int v) {
    if (v) m[i >> 6] ^= 1ULL << (i & 63);
}

/* gamma *= (re, im)  (svm/svm.h:167-170) */
static inline void mul_phase(orc_state* st, double re, double im) {
    double a = st->g_re, b = st->g_im;
    st->g_re = a * re - b * im;
    st->g_im = a * im + b * re;
}
static inline void negate_gamma(orc_state* st) {
    st->g_re = -st->g_re;
    st->g_im = -st->g_im;
}

/* svm/svm.h:177-186 */
static void scale_magnitude(orc_state* st, double scale) {
    st->g_re *= scale;
    st->g_im *= scale;
    double g = hypot(st->g_re, st->g_im);
    if (g > 1e100 || (g < 1e-100 && g > 0.0)) {
        uint64_t sz = 1ULL << st->active_k;
        for (uint64_t i = 0; i < 2 * sz; ++i) st->v[i] *= g;
        st->g_re /= g;
        st->g_im /= g;
    }
}

/* svm/svm_internal.h:65-79 */
static int sample_branch(orc_state* st, double p0, double p1, double total) {
    double eps = kDustEpsilon * total;
    if (p1 <= eps) {
        if (p1 > 0.0) st->dust_clamps++;
        return 0;
    }
    if (p0 <= eps) {
        if (p0 > 0.0) st->dust_clamps++;
        return 1;
    }
    double r = rng_double(&st->rng);
    return (r * total < p0) ? 0 : 1;
}

/* svm/svm_math.h:58-61 */
static inline uint64_t insert_zero_bit(uint64_t val, uint32_t pos) {
    uint64_t mask = (1ULL << pos) - 1;
    return (val & mask) | ((val & ~mask) << 1);
}
static inline uint64_t insert_two_zero_bits(uint64_t val, uint32_t a, uint32_t b) {
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return insert_zero_bit(insert_zero_bit(val, lo), hi);
}

/* complex helpers over the interleaved array */
#define RE(i) st->v[2 * (i)]
#define IM(i) st->v[2 * (i) + 1]

/* ------------------------------------------------------------------------
 * Frame opcodes (svm/svm_kernels.inl:67-116, 749-755)
 * ---------------------------------------------------------------------- */
static void frame_cnot(orc_state* st, uint32_t c, uint32_t t) {
    int pxc = bit_get(st->px, c), pzt = bit_get(st->pz, t);
    bit_xor(st->px, t, pxc);
    bit_xor(st->pz, c, pzt);
}
static void frame_cz(orc_state* st, uint32_t c, uint32_t t) {
    int pxc = bit_get(st->px, c), pxt = bit_get(st->px, t);
    if (pxc && pxt) negate_gamma(st);
    bit_xor(st->pz, t, pxc);
    bit_xor(st->pz, c, pxt);
}
static void frame_h(orc_state* st, uint32_t v) {
    int px = bit_get(st->px, v), pz = bit_get(st->pz, v);
    if (px && pz) negate_gamma(st);
    bit_set(st->px, v, pz);
    bit_set(st->pz, v, px);
}
static void frame_s(orc_state* st, uint32_t v) {
    int px = bit_get(st->px, v);
    if (px) mul_phase(st, 0.0, 1.0);
    bit_xor(st->pz, v, px);
}
static void frame_s_dag(orc_state* st, uint32_t v) {
    int px = bit_get(st->px, v);
    if (px) mul_phase(st, 0.0, -1.0);
    bit_xor(st->pz, v, px);
}
static void frame_swap(orc_state* st, uint32_t a, uint32_t b) {
    int xa = bit_get(st->px, a), xb = bit_get(st->px, b);
    if (xa != xb) {
        bit_xor(st->px, a, 1);
        bit_xor(st->px, b, 1);
    }
    int za = bit_get(st->pz, a), zb = bit_get(st->pz, b);
    if (za != zb) {
        bit_xor(st->pz, a, 1);
        bit_xor(st->pz, b, 1);
    }
}

/* ------------------------------------------------------------------------
 * Array opcodes (scalar paths of svm/svm_kernels.inl:124-1531)
 * ---------------------------------------------------------------------- */
static void array_cnot(orc_state* st, uint32_t c, uint32_t t) { /* :238-252 */
    uint64_t iters = 1ULL << (st->active_k - 2), cb = 1ULL << c, tb = 1ULL << t;
    for (uint64_t i = 0; i < iters; ++i) {
        uint64_t b0 = insert_two_zero_bits(i, c, t) | cb, b1 = b0 | tb;
        double r = RE(b0), m = IM(b0);
        RE(b0) = RE(b1);
        IM(b0) = IM(b1);
        RE(b1) = r;
        IM(b1) = m;
    }
    frame_cnot(st, c, t);
}
static void array_cz(orc_state* st, uint32_t c, uint32_t t) { /* :352-365 */
    uint64_t iters = 1ULL << (st->active_k - 2), both = (1ULL << c) | (1ULL << t);
    for (uint64_t i = 0; i < iters; ++i) {
        uint64_t idx = insert_two_zero_bits(i, c, t) | both;
        RE(idx) = -RE(idx);
        IM(idx) = -IM(idx);
    }
    frame_cz(st, c, t);
}
static void array_swap(orc_state* st, uint32_t a, uint32_t b) { /* :444-458 */
    uint64_t iters = 1ULL << (st->active_k - 2), ab = 1ULL << a, bb = 1ULL << b;
    for (uint64_t i = 0; i < iters; ++i) {
        uint64_t base = insert_two_zero_bits(i, a, b);
        uint64_t i0 = base | ab, i1 = base | bb;
        double r = RE(i0), m = IM(i0);
        RE(i0) = RE(i1);
        IM(i0) = IM(i1);
        RE(i1) = r;
        IM(i1) = m;
    }
    frame_swap(st, a, b);
}
static void array_multi_cnot(orc_state* st, uint32_t target, uint64_t cmask) { /* :464-550 */
    uint64_t tb = 1ULL << target, half = 1ULL << (st->active_k - 1);
    for (uint64_t i = 0; i < half; ++i) {
        uint64_t a = insert_zero_bit(i, target);
        if (__builtin_popcountll(a & cmask) & 1) {
            uint64_t b = a | tb;
            double r = RE(a), m = IM(a);
            RE(a) = RE(b);
            IM(a) = IM(b);
            RE(b) = r;
            IM(b) = m;
        }
    }
    for (uint32_t c = 0; c < st->active_k; ++c)
        if (cmask & (1ULL << c)) frame_cnot(st, c, target);
}
static void array_multi_cz(orc_state* st, uint32_t control, uint64_t tmask) { /* :559-635 */
    uint64_t cb = 1ULL << control, half = 1ULL << (st->active_k - 1);
    for (uint64_t i = 0; i < half; ++i) {
        uint64_t a = insert_zero_bit(i, control) | cb;
        if (__builtin_popcountll(a & tmask) & 1) {
            RE(a) = -RE(a);
            IM(a) = -IM(a);
        }
    }
    for (uint32_t t = 0; t < st->active_k; ++t)
        if (tmask & (1ULL << t)) frame_cz(st, control, t);
}
static void array_h(orc_state* st, uint32_t v) { /* :729-744 */
    uint64_t half = 1ULL << (st->active_k - 1), vb = 1ULL << v;
    for (uint64_t i = 0; i < half; ++i) {
        uint64_t i0 = insert_zero_bit(i, v), i1 = i0 | vb;
        double ar = RE(i0), ai = IM(i0), br = RE(i1), bi = IM(i1);
        RE(i0) = (ar + br) * kInvSqrt2;
        IM(i0) = (ai + bi) * kInvSqrt2;
        RE(i1) = (ar - br) * kInvSqrt2;
        IM(i1) = (ai - bi) * kInvSqrt2;
    }
    frame_h(st, v);
}
/* apply_phase_waterfall scalar path (:843-850) */
static void phase_waterfall(orc_state* st, uint32_t v, double pr, double pi) {
    uint64_t half = 1ULL << (st->active_k - 1), vb = 1ULL << v;
    for (uint64_t i = 0; i < half; ++i) {
        uint64_t idx = insert_zero_bit(i, v) | vb;
        double a = RE(idx), b = IM(idx);
        RE(idx) = a * pr - b * pi;
        IM(idx) = a * pi + b * pr;
    }
}
static void array_s(orc_state* st, uint32_t v) { phase_waterfall(st, v, 0.0, 1.0); frame_s(st, v); }
static void array_s_dag(orc_state* st, uint32_t v) { phase_waterfall(st, v, 0.0, -1.0); frame_s_dag(st, v); }

/* :951-966, 972-988, 1044-1061: diag(1, z) with the p_x conjugation rule */
static void array_phase(orc_state* st, uint32_t v, double zr, double zi) {
    int px = bit_get(st->px, v);
    if (v >= st->active_k) {
        if (px) mul_phase(st, zr, zi);
        return;
    }
    if (px) {
        phase_waterfall(st, v, zr, -zi);
        mul_phase(st, zr, zi);
    } else {
        phase_waterfall(st, v, zr, zi);
    }
}

/* :876-897 (EXPAND) and :906-945, 995-1034, 1065-1081 (EXPAND + phase) */
static void expand_phase(orc_state* st, uint32_t v, double zr, double zi, int has_phase) {
    uint64_t half = 1ULL << st->active_k;
    int px = bit_get(st->px, v);
    (void)v;
    if (!has_phase) {
        memcpy(st->v + 2 * half, st->v, half * 16);
    } else {
        double pi = px ? -zi : zi;
        for (uint64_t i = 0; i < half; ++i) {
            double a = RE(i), b = IM(i);
            RE(i + half) = a * zr - b * pi;
            IM(i + half) = a * pi + b * zr;
        }
    }
    st->active_k++;
    scale_magnitude(st, 1.0 / sqrt(2.0));
    if (has_phase && px) mul_phase(st, zr, zi);
}

/* :1091-1297 */
static void array_u2(orc_state* st, const prog_t* p, uint32_t axis, uint32_t cp) {
    const clifft_u2_node* node = &p->u2_nodes[cp];
    int s = (bit_get(st->pz, axis) ? 2 : 0) | (bit_get(st->px, axis) ? 1 : 0);
    const double(*m)[2] = node->matrices[s];
    mul_phase(st, node->gamma[s][0], node->gamma[s][1]);
    uint8_t out = node->out_states[s];
    bit_set(st->px, axis, out & 1);
    bit_set(st->pz, axis, (out >> 1) & 1);
    if (axis >= st->active_k) return;
    uint64_t half = 1ULL << (st->active_k - 1), vb = 1ULL << axis;
    for (uint64_t i = 0; i < half; ++i) {
        uint64_t i0 = insert_zero_bit(i, axis), i1 = i0 | vb;
        double ar = RE(i0), ai = IM(i0), br = RE(i1), bi = IM(i1);
        RE(i0) = m[0][0] * ar - m[0][1] * ai + m[1][0] * br - m[1][1] * bi;
        IM(i0) = m[0][0] * ai + m[0][1] * ar + m[1][0] * bi + m[1][1] * br;
        RE(i1) = m[2][0] * ar - m[2][1] * ai + m[3][0] * br - m[3][1] * bi;
        IM(i1) = m[2][0] * ai + m[2][1] * ar + m[3][0] * bi + m[3][1] * br;
    }
}

/* :1310-1531 */
static void array_u4(orc_state* st, const prog_t* p, uint32_t lo, uint32_t hi, uint32_t cp) {
    const clifft_u4_node* node = &p->u4_nodes[cp];
    int s = (bit_get(st->pz, hi) << 3) | (bit_get(st->px, hi) << 2) | (bit_get(st->pz, lo) << 1) |
            bit_get(st->px, lo);
    const clifft_u4_entry* e = &node->entries[s];
    mul_phase(st, e->gamma[0], e->gamma[1]);
    uint8_t out = e->out_state;
    bit_set(st->px, lo, out & 1);
    bit_set(st->pz, lo, (out >> 1) & 1);
    bit_set(st->px, hi, (out >> 2) & 1);
    bit_set(st->pz, hi, (out >> 3) & 1);
    uint64_t lb = 1ULL << lo, hb = 1ULL << hi, iters = 1ULL << (st->active_k - 2);
    for (uint64_t i = 0; i < iters; ++i) {
        uint64_t base = insert_two_zero_bits(i, lo, hi);
        uint64_t ix[4] = {base, base | lb, base | hb, base | lb | hb};
        double in_r[4], in_i[4];
        for (int j = 0; j < 4; ++j) {
            in_r[j] = RE(ix[j]);
            in_i[j] = IM(ix[j]);
        }
        for (int r = 0; r < 4; ++r) {
            double acc_r = 0.0, acc_i = 0.0;
            for (int c = 0; c < 4; ++c) {
                double mr = e->matrix[r][c][0], mi = e->matrix[r][c][1];
                acc_r += mr * in_r[c] - mi * in_i[c];
                acc_i += mr * in_i[c] + mi * in_r[c];
            }
            RE(ix[r]) = acc_r;
            IM(ix[r]) = acc_i;
        }
    }
}

/* ------------------------------------------------------------------------
 * Measurements (svm/svm_kernels.inl:1540-1804)
 * ---------------------------------------------------------------------- */
static void meas_dormant_static(orc_state* st, uint32_t v, uint32_t idx, int sign) {
    st->meas[idx] = (uint8_t)(bit_get(st->px, v) ^ sign);
}
static void meas_dormant_random(orc_state* st, uint32_t v, uint32_t idx, int sign) {
    int m = rng_double(&st->rng) < 0.5 ? 0 : 1;
    if (bit_get(st->px, v) && m) negate_gamma(st);
    bit_set(st->px, v, m);
    bit_set(st->pz, v, 0);
    st->meas[idx] = (uint8_t)(m ^ sign);
}

static void diag_probs(orc_state* st, double* p0, double* p1) {
    uint64_t half = 1ULL << (st->active_k - 1);
    double a = 0.0, b = 0.0;
    for (uint64_t i = 0; i < half; ++i) {
        a += RE(i) * RE(i) + IM(i) * IM(i);
        b += RE(i + half) * RE(i + half) + IM(i + half) * IM(i + half);
    }
    *p0 = a;
    *p1 = b;
}
static void interfere_probs(orc_state* st, double* pp, double* pm) {
    uint64_t half = 1ULL << (st->active_k - 1);
    double a = 0.0, b = 0.0;
    for (uint64_t i = 0; i < half; ++i) {
        double sr = RE(i) + RE(i + half), si = IM(i) + IM(i + half);
        double dr = RE(i) - RE(i + half), di = IM(i) - IM(i + half);
        a += sr * sr + si * si;
        b += dr * dr + di * di;
    }
    *pp = a;
    *pm = b;
}
static void swap_interfere_probs(orc_state* st, uint32_t f, uint32_t t, double* pp, double* pm) {
    uint64_t half = 1ULL << t, fb = 1ULL << f;
    double a = 0.0, b = 0.0;
    for (uint64_t idx = 0; idx < half; ++idx) {
        uint64_t bf = (idx >> f) & 1, base = (idx & ~fb) | (bf << t);
        double xr = RE(base), xi = IM(base), yr = RE(base | fb), yi = IM(base | fb);
        double sr = xr + yr, si = xi + yi, dr = xr - yr, di = xi - yi;
        a += sr * sr + si * si;
        b += dr * dr + di * di;
    }
    *pp = a;
    *pm = b;
}
static void diag_compact(orc_state* st, int b) {
    uint64_t half = 1ULL << (st->active_k - 1);
    if (b == 1) memmove(st->v, st->v + 2 * half, half * 16);
}
static void interfere_fold(orc_state* st, int bx) {
    uint64_t half = 1ULL << (st->active_k - 1);
    double s = bx ? -1.0 : 1.0;
    for (uint64_t i = 0; i < half; ++i) {
        RE(i) = (RE(i) + s * RE(i + half)) * kInvSqrt2;
        IM(i) = (IM(i) + s * IM(i + half)) * kInvSqrt2;
    }
}
static void swap_interfere_fold(orc_state* st, uint32_t f, uint32_t t, int bx) {
    uint64_t half = 1ULL << t, fb = 1ULL << f;
    double s = bx ? -1.0 : 1.0;
    for (uint64_t idx = 0; idx < half; ++idx) {
        uint64_t bf = (idx >> f) & 1, base = (idx & ~fb) | (bf << t);
        double xr = RE(base), xi = IM(base), yr = RE(base | fb), yi = IM(base | fb);
        RE(idx) = (xr + s * yr) * kInvSqrt2;
        IM(idx) = (xi + s * yi) * kInvSqrt2;
    }
}

/* :1573-1627 */
static void meas_active_diagonal(orc_state* st, uint32_t v, uint32_t idx, int sign) {
    int px = bit_get(st->px, v), pz = bit_get(st->pz, v);
    double p0, p1;
    diag_probs(st, &p0, &p1);
    double total = p0 + p1;
    int b = sample_branch(st, p0, p1, total);
    int m_abs = b ^ px;
    if (b == 1 && pz) negate_gamma(st);
    diag_compact(st, b);
    st->active_k--;
    double pb = b ? p1 : p0;
    if (pb > 0.0) scale_magnitude(st, sqrt(total / pb));
    bit_set(st->px, v, m_abs);
    bit_set(st->pz, v, 0);
    st->meas[idx] = (uint8_t)(m_abs ^ sign);
}
/* :1634-1708 */
static void meas_active_interfere(orc_state* st, uint32_t v, uint32_t idx, int sign) {
    int px = bit_get(st->px, v), pz = bit_get(st->pz, v);
    double pp, pm;
    interfere_probs(st, &pp, &pm);
    double total = pp + pm;
    int bx = sample_branch(st, pp, pm, total);
    int m_abs = bx ^ pz;
    interfere_fold(st, bx);
    if (px && m_abs) negate_gamma(st);
    st->active_k--;
    double pb = bx ? pm : pp;
    if (pb > 0.0) scale_magnitude(st, sqrt(total / pb));
    bit_set(st->px, v, m_abs);
    bit_set(st->pz, v, 0);
    st->meas[idx] = (uint8_t)(m_abs ^ sign);
}
/* :1728-1804 */
static void swap_meas_interfere(orc_state* st, uint32_t f, uint32_t t, uint32_t idx, int sign) {
    if (f == t) {
        meas_active_interfere(st, t, idx, sign);
        return;
    }
    frame_swap(st, f, t);
    int px = bit_get(st->px, t), pz = bit_get(st->pz, t);
    double pp, pm;
    swap_interfere_probs(st, f, t, &pp, &pm);
    double total = pp + pm;
    int bx = sample_branch(st, pp, pm, total);
    swap_interfere_fold(st, f, t, bx);
    st->active_k--;
    double pb = bx ? pm : pp;
    if (pb > 0.0) scale_magnitude(st, sqrt(total / pb));
    int m_abs = bx ^ pz;
    if (px && m_abs) negate_gamma(st);
    bit_set(st->px, t, m_abs);
    bit_set(st->pz, t, 0);
    st->meas[idx] = (uint8_t)(m_abs ^ sign);
}

/* ------------------------------------------------------------------------
 * Forced measurement variants (svm/svm_forced_kernels.cc:23-306)
 * ---------------------------------------------------------------------- */
/* :43-61; returns reachable, writes the log increment */
static int forced_log_increment(orc_state* st, double p0, double p1, double total, int b,
                                double* inc) {
    double eps = kDustEpsilon * total;
    *inc = 0.0;
    if (p1 <= eps) {
        if (p1 > 0.0) st->dust_clamps++;
        return b == 0;
    }
    if (p0 <= eps) {
        if (p0 > 0.0) st->dust_clamps++;
        return b == 1;
    }
    *inc = log((b ? p1 : p0) / total);
    return 1;
}
static int unreachable(orc_state* st) {
    st->forced_reachable = 0;
    return 0;
}
static int forced_static(orc_state* st, int identity, uint32_t v, uint32_t idx, int sign) {
    uint8_t det = identity ? (uint8_t)sign : (uint8_t)(bit_get(st->px, v) ^ sign);
    if (st->forced_record[idx] != det) return unreachable(st);
    st->meas[idx] = det;
    return 1;
}
static int forced_dormant_random(orc_state* st, uint32_t v, uint32_t idx, int sign) {
    uint8_t req = st->forced_record[idx];
    int m = req ^ sign;
    st->forced_log_prob += log(0.5);
    if (bit_get(st->px, v) && m) negate_gamma(st);
    bit_set(st->px, v, m);
    bit_set(st->pz, v, 0);
    st->meas[idx] = req;
    return 1;
}
static int forced_diagonal(orc_state* st, uint32_t v, uint32_t idx, int sign) {
    int px = bit_get(st->px, v), pz = bit_get(st->pz, v);
    double p0, p1, inc;
    diag_probs(st, &p0, &p1);
    double total = p0 + p1;
    uint8_t req = st->forced_record[idx];
    int m_abs = req ^ sign, b = m_abs ^ px;
    if (!forced_log_increment(st, p0, p1, total, b, &inc)) return unreachable(st);
    st->forced_log_prob += inc;
    double pb = b ? p1 : p0;
    if (b == 1 && pz) negate_gamma(st);
    diag_compact(st, b);
    st->active_k--;
    scale_magnitude(st, sqrt(total / pb));
    bit_set(st->px, v, m_abs);
    bit_set(st->pz, v, 0);
    st->meas[idx] = req;
    return 1;
}
static int forced_interfere(orc_state* st, uint32_t v, uint32_t idx, int sign) {
    int px = bit_get(st->px, v), pz = bit_get(st->pz, v);
    double pp, pm, inc;
    interfere_probs(st, &pp, &pm);
    double total = pp + pm;
    uint8_t req = st->forced_record[idx];
    int m_abs = req ^ sign, bx = m_abs ^ pz;
    if (!forced_log_increment(st, pp, pm, total, bx, &inc)) return unreachable(st);
    st->forced_log_prob += inc;
    double pb = bx ? pm : pp;
    interfere_fold(st, bx);
    if (px && m_abs) negate_gamma(st);
    st->active_k--;
    scale_magnitude(st, sqrt(total / pb));
    bit_set(st->px, v, m_abs);
    bit_set(st->pz, v, 0);
    st->meas[idx] = req;
    return 1;
}
static int forced_swap_interfere(orc_state* st, uint32_t f, uint32_t t, uint32_t idx, int sign) {
    if (f == t) return forced_interfere(st, t, idx, sign);
    frame_swap(st, f, t);
    int px = bit_get(st->px, t), pz = bit_get(st->pz, t);
    double pp, pm, inc;
    swap_interfere_probs(st, f, t, &pp, &pm);
    double total = pp + pm;
    uint8_t req = st->forced_record[idx];
    int m_abs = req ^ sign, bx = m_abs ^ pz;
    if (!forced_log_increment(st, pp, pm, total, bx, &inc)) return unreachable(st);
    st->forced_log_prob += inc;
    double pb = bx ? pm : pp;
    swap_interfere_fold(st, f, t, bx);
    st->active_k--;
    scale_magnitude(st, sqrt(total / pb));
    if (px && m_abs) negate_gamma(st);
    bit_set(st->px, t, m_abs);
    bit_set(st->pz, t, 0);
    st->meas[idx] = req;
    return 1;
}

/* ------------------------------------------------------------------------
 * Classical / noise (svm/svm_kernels.inl:1812-2020; svm/svm.h:231-248)
 * ---------------------------------------------------------------------- */
/* :1829-1866 */
static void apply_pauli_frame(orc_state* st, const uint64_t* ex, const uint64_t* ez, int sign) {
    int parity = 0;
    for (uint32_t w = 0; w < st->words; ++w) parity += __builtin_popcountll(st->px[w] & ez[w]);
    if (parity & 1) negate_gamma(st);
    for (uint32_t w = 0; w < st->words; ++w) {
        st->px[w] ^= ex[w];
        st->pz[w] ^= ez[w];
    }
    if (sign) negate_gamma(st);
}

/* svm/svm.h:239-248: upper_bound over the cumulative hazard table */
static void draw_next_noise(orc_state* st, const prog_t* p) {
    uint32_t n = p->num_sites;
    if (n == 0 || p->hazards == NULL || st->next_noise_idx >= n) {
        st->next_noise_idx = 0xFFFFFFFFu;
        return;
    }
    double cur = st->next_noise_idx == 0 ? 0.0 : p->hazards[st->next_noise_idx - 1];
    double gap = -log(1.0 - rng_double(&st->rng));
    double target = cur + gap;
    uint32_t lo = 0, hi = n; /* first index with hazards[i] > target */
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (p->hazards[mid] <= target)
            lo = mid + 1;
        else
            hi = mid;
    }
    st->next_noise_idx = lo;
}

/* svm/svm.h:218-225 */
static void advance_forced_noise(orc_state* st) {
    if (st->ff_noise_pos < st->ff_n_noise)
        st->next_noise_idx = st->ff_noise[st->ff_noise_pos++];
    else
        st->next_noise_idx = 0xFFFFFFFFu;
}

/* :1886-1920 */
static void exec_noise(orc_state* st, const prog_t* p, uint32_t site) {
    if (site != st->next_noise_idx) return;
    uint32_t c0 = p->site_chan_off[site], c1 = p->site_chan_off[site + 1];
    double psum = 0.0;
    for (uint32_t c = c0; c < c1; ++c) psum += p->chan_prob[c];
    double r = rng_double(&st->rng) * psum;
    double cum = 0.0;
    for (uint32_t c = c0; c < c1; ++c) {
        cum += p->chan_prob[c];
        if (r < cum) {
            apply_pauli_frame(st, p->chan_x + (size_t)c * p->words, p->chan_z + (size_t)c * p->words, 0);
            break;
        }
    }
    if (st->ff_active) {
        advance_forced_noise(st);
    } else {
        st->next_noise_idx++;
        draw_next_noise(st, p);
    }
}
/* :1926-1932 */
static void exec_noise_block(orc_state* st, const prog_t* p, uint32_t start, uint32_t count) {
    uint32_t end = start + count;
    while (st->next_noise_idx >= start && st->next_noise_idx < end) exec_noise(st, p, st->next_noise_idx);
}
/* :1937-1959 */
static void exec_readout_noise(orc_state* st, const prog_t* p, uint32_t e) {
    int fire = 0;
    if (st->ff_active) {
        if (st->ff_ro_pos < st->ff_n_ro && st->ff_ro[st->ff_ro_pos] == e) {
            fire = 1;
            st->ff_ro_pos++;
        }
    } else {
        fire = rng_double(&st->rng) < p->ro_prob[e];
    }
    if (fire) st->meas[p->ro_meas_idx[e]] ^= 1;
}
static uint8_t list_parity(const orc_state* st, const uint32_t* off, const uint32_t* tgt, uint32_t l) {
    uint8_t par = 0;
    for (uint32_t j = off[l]; j < off[l + 1]; ++j) par ^= st->meas[tgt[j]];
    return par;
}

/* :2031-2173 */
static void exec_exp_val(orc_state* st, const prog_t* p, uint32_t cp, uint32_t out) {
    const uint64_t* mx = p->ev_x + (size_t)cp * p->words;
    const uint64_t* mz = p->ev_z + (size_t)cp * p->words;
    int fsign = p->ev_sign[cp] ? 1 : 0;
    int parity = 0;
    for (uint32_t w = 0; w < p->words; ++w)
        parity += __builtin_popcountll(mx[w] & st->pz[w]) + __builtin_popcountll(mz[w] & st->px[w]);
    if (parity & 1) fsign = !fsign;
    uint32_t k = st->active_k;
    uint32_t full = k / 64, rem = k % 64;
    int dormant_x = 0;
    if (rem > 0 && full < p->words && (mx[full] & ~((1ULL << rem) - 1))) dormant_x = 1;
    for (uint32_t w = (rem == 0 ? full : full + 1); w < p->words && !dormant_x; ++w)
        if (mx[w]) dormant_x = 1;
    if (dormant_x) {
        st->ev[out] = 0.0;
        return;
    }
    uint64_t xa = 0, za = 0;
    if (k > 0 && p->words > 0) {
        uint64_t am = (k == 64) ? ~0ULL : ((1ULL << k) - 1);
        xa = mx[0] & am;
        za = mz[0] & am;
    }
    uint32_t yc = __builtin_popcountll(xa & za) & 3;
    double ph_r = (yc == 0) ? 1.0 : (yc == 2 ? -1.0 : 0.0);
    double ph_i = (yc == 1) ? 1.0 : (yc == 3 ? -1.0 : 0.0);
    if (fsign) {
        ph_r = -ph_r;
        ph_i = -ph_i;
    }
    uint64_t dim = 1ULL << k;
    double nre = 0.0, nim = 0.0, den = 0.0;
    for (uint64_t j = 0; j < dim; ++j) {
        double ar = RE(j), ai = IM(j);
        den += ar * ar + ai * ai;
        uint64_t jx = j ^ xa;
        double zs = (__builtin_popcountll(j & za) & 1) ? -1.0 : 1.0;
        double br = RE(jx), bi = -IM(jx); /* conj(a[j ^ x]) */
        double cr = zs * ar, ci = zs * ai;
        nre += br * cr - bi * ci;
        nim += br * ci + bi * cr;
    }
    double num_re = nre * ph_r - nim * ph_i;
    double res = den > 0.0 ? num_re / den : 0.0;
    if (res > 1.0) res = 1.0;
    if (res < -1.0) res = -1.0;
    st->ev[out] = res;
}

/* ------------------------------------------------------------------------
 * Dispatch loop (svm/svm_kernels.inl:2183-2646)
 * ---------------------------------------------------------------------- */
ORC_EXPORT void orc_execute(const prog_t* p, orc_state* st) {
    for (uint64_t pc = 0; pc < p->num_instructions; ++pc) {
        const clifft_instruction* in = &p->bytecode[pc];
        uint32_t a1 = in->axis_1, a2 = in->axis_2;
        int sign = (in->flags & CLIFFT_FLAG_SIGN) != 0;
        switch (in->opcode) {
            case CLIFFT_OP_FRAME_CNOT: frame_cnot(st, a1, a2); break;
            case CLIFFT_OP_FRAME_CZ: frame_cz(st, a1, a2); break;
            case CLIFFT_OP_FRAME_H: frame_h(st, a1); break;
            case CLIFFT_OP_FRAME_S: frame_s(st, a1); break;
            case CLIFFT_OP_FRAME_S_DAG: frame_s_dag(st, a1); break;
            case CLIFFT_OP_FRAME_SWAP: frame_swap(st, a1, a2); break;
            case CLIFFT_OP_ARRAY_CNOT: array_cnot(st, a1, a2); break;
            case CLIFFT_OP_ARRAY_CZ: array_cz(st, a1, a2); break;
            case CLIFFT_OP_ARRAY_SWAP: array_swap(st, a1, a2); break;
            case CLIFFT_OP_ARRAY_MULTI_CNOT: array_multi_cnot(st, a1, in->p.multi_gate.mask); break;
            case CLIFFT_OP_ARRAY_MULTI_CZ: array_multi_cz(st, a1, in->p.multi_gate.mask); break;
            case CLIFFT_OP_ARRAY_H: array_h(st, a1); break;
            case CLIFFT_OP_ARRAY_S: array_s(st, a1); break;
            case CLIFFT_OP_ARRAY_S_DAG: array_s_dag(st, a1); break;
            case CLIFFT_OP_ARRAY_T: array_phase(st, a1, kInvSqrt2, kInvSqrt2); break;
            case CLIFFT_OP_ARRAY_T_DAG: array_phase(st, a1, kInvSqrt2, -kInvSqrt2); break;
            case CLIFFT_OP_ARRAY_ROT: array_phase(st, a1, in->p.math.weight_re, in->p.math.weight_im); break;
            case CLIFFT_OP_ARRAY_U2: array_u2(st, p, a1, in->p.u.cp_idx); break;
            case CLIFFT_OP_ARRAY_U4: array_u4(st, p, a1, a2, in->p.u.cp_idx); break;
            case CLIFFT_OP_EXPAND: expand_phase(st, a1, 1.0, 0.0, 0); break;
            case CLIFFT_OP_EXPAND_T: expand_phase(st, a1, kInvSqrt2, kInvSqrt2, 1); break;
            case CLIFFT_OP_EXPAND_T_DAG: expand_phase(st, a1, kInvSqrt2, -kInvSqrt2, 1); break;
            case CLIFFT_OP_EXPAND_ROT: expand_phase(st, a1, in->p.math.weight_re, in->p.math.weight_im, 1); break;
            case CLIFFT_OP_MEAS_DORMANT_STATIC:
                if (in->flags & CLIFFT_FLAG_IDENTITY)
                    st->meas[in->p.classical.classical_idx] = (uint8_t)sign;
                else
                    meas_dormant_static(st, a1, in->p.classical.classical_idx, sign);
                break;
            case CLIFFT_OP_MEAS_DORMANT_RANDOM: meas_dormant_random(st, a1, in->p.classical.classical_idx, sign); break;
            case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL: meas_active_diagonal(st, a1, in->p.classical.classical_idx, sign); break;
            case CLIFFT_OP_MEAS_ACTIVE_INTERFERE: meas_active_interfere(st, a1, in->p.classical.classical_idx, sign); break;
            case CLIFFT_OP_SWAP_MEAS_INTERFERE: swap_meas_interfere(st, a1, a2, in->p.classical.classical_idx, sign); break;
            case CLIFFT_OP_MEAS_DORMANT_STATIC_FORCED:
                if (!forced_static(st, (in->flags & CLIFFT_FLAG_IDENTITY) != 0, a1, in->p.classical.classical_idx, sign)) return;
                break;
            case CLIFFT_OP_MEAS_DORMANT_RANDOM_FORCED:
                if (!forced_dormant_random(st, a1, in->p.classical.classical_idx, sign)) return;
                break;
            case CLIFFT_OP_MEAS_ACTIVE_DIAGONAL_FORCED:
                if (!forced_diagonal(st, a1, in->p.classical.classical_idx, sign)) return;
                break;
            case CLIFFT_OP_MEAS_ACTIVE_INTERFERE_FORCED:
                if (!forced_interfere(st, a1, in->p.classical.classical_idx, sign)) return;
                break;
            case CLIFFT_OP_SWAP_MEAS_INTERFERE_FORCED:
                if (!forced_swap_interfere(st, a1, a2, in->p.classical.classical_idx, sign)) return;
                break;
            case CLIFFT_OP_APPLY_PAULI: {
                uint32_t cp = in->p.pauli.cp_mask_idx, cond = in->p.pauli.condition_idx;
                if (st->meas[cond] == 0) break;
                apply_pauli_frame(st, p->pauli_x + (size_t)cp * p->words, p->pauli_z + (size_t)cp * p->words,
                                  p->pauli_sign[cp] != 0);
                break;
            }
            case CLIFFT_OP_NOISE: exec_noise(st, p, in->p.pauli.cp_mask_idx); break;
            case CLIFFT_OP_NOISE_BLOCK: exec_noise_block(st, p, in->p.pauli.cp_mask_idx, in->p.pauli.condition_idx); break;
            case CLIFFT_OP_READOUT_NOISE: exec_readout_noise(st, p, in->p.pauli.cp_mask_idx); break;
            case CLIFFT_OP_DETECTOR: {
                uint8_t par = (in->flags & CLIFFT_FLAG_EXPECTED_ONE) ? 1 : 0;
                par ^= list_parity(st, p->det_off, p->det_tgt, in->p.pauli.cp_mask_idx);
                st->det[in->p.pauli.condition_idx] = par;
                break;
            }
            case CLIFFT_OP_POSTSELECT: {
                uint8_t par = (in->flags & CLIFFT_FLAG_EXPECTED_ONE) ? 1 : 0;
                par ^= list_parity(st, p->det_off, p->det_tgt, in->p.pauli.cp_mask_idx);
                st->det[in->p.pauli.condition_idx] = 0;
                if (par) {
                    st->discarded = 1;
                    return;
                }
                break;
            }
            case CLIFFT_OP_OBSERVABLE:
                st->obs[in->p.pauli.condition_idx] ^= list_parity(st, p->obs_off, p->obs_tgt, in->p.pauli.cp_mask_idx);
                break;
            case CLIFFT_OP_EXP_VAL: exec_exp_val(st, p, in->p.exp_val.cp_exp_val_idx, in->p.exp_val.exp_val_idx); break;
            default: break;
        }
    }
}

/* ------------------------------------------------------------------------
 * Shot drivers (svm/svm.cc:191-345)
 * ---------------------------------------------------------------------- */
static void copy_out(const prog_t* p, const orc_state* st, uint64_t row, uint8_t* meas, uint8_t* det,
                     uint8_t* obs, double* ev) {
    if (meas && p->num_measurements) memcpy(meas + row * p->num_measurements, st->meas, p->num_measurements);
    if (det && p->num_detectors) memcpy(det + row * p->num_detectors, st->det, p->num_detectors);
    if (obs)
        for (uint32_t i = 0; i < p->num_observables; ++i) {
            uint8_t val = st->obs[i];
            if (i < p->num_expected_observables && p->expected_observables[i]) val ^= 1;
            obs[row * p->num_observables + i] = val;
        }
    if (ev && p->num_exp_vals) memcpy(ev + row * p->num_exp_vals, st->ev, p->num_exp_vals * sizeof(double));
}

static orc_state* state_for(const prog_t* p, uint64_t seed) {
    return orc_state_new(p->peak_rank, p->total_meas_slots, p->num_qubits, p->num_detectors,
                         p->num_observables, p->num_exp_vals, seed);
}

/* mode 0: one stream seeded once (svm.cc:191-255).  mode 1: per-shot
 * stream (rng_seed_shot), the device executor's convention.  `first_shot`
 * offsets the shot index in mode 1 so a bounded sample can reproduce rows
 * of a larger run. */
ORC_EXPORT int orc_sample(const prog_t* p, uint64_t shots, uint64_t seed, int mode, uint64_t first_shot,
                          uint8_t* meas, uint8_t* det, uint8_t* obs, double* ev, uint8_t* discarded) {
    orc_state* st = state_for(p, seed);
    if (!st) return -1;
    for (uint64_t s = 0; s < shots; ++s) {
        if (s > 0) orc_state_reset(st);
        if (mode == 1) rng_seed_shot(&st->rng, seed, first_shot + s);
        st->next_noise_idx = 0;
        draw_next_noise(st, p);
        orc_execute(p, st);
        copy_out(p, st, s, meas, det, obs, ev);
        if (discarded) discarded[s] = (uint8_t)st->discarded;
    }
    orc_state_free(st);
    return 0;
}

/* svm.cc:266-345.  counts[0] = passed, counts[1] = logical_errors,
 * obs_ones[num_observables]; records of survivors are compacted. */
ORC_EXPORT int orc_sample_survivors(const prog_t* p, uint64_t shots, uint64_t seed, int mode, uint64_t* counts,
                                    uint64_t* obs_ones, uint8_t* meas, uint8_t* det, uint8_t* obs, double* ev) {
    orc_state* st = state_for(p, seed);
    if (!st) return -1;
    counts[0] = counts[1] = 0;
    for (uint32_t i = 0; i < p->num_observables; ++i) obs_ones[i] = 0;
    for (uint64_t s = 0; s < shots; ++s) {
        if (s > 0) orc_state_reset(st);
        if (mode == 1) rng_seed_shot(&st->rng, seed, s);
        st->next_noise_idx = 0;
        draw_next_noise(st, p);
        orc_execute(p, st);
        if (st->discarded) continue;
        int any = 0;
        for (uint32_t i = 0; i < p->num_observables; ++i) {
            uint8_t val = st->obs[i];
            if (i < p->num_expected_observables && p->expected_observables[i]) val ^= 1;
            if (val) {
                obs_ones[i]++;
                any = 1;
            }
        }
        if (any) counts[1]++;
        copy_out(p, st, counts[0], meas, det, obs, ev);
        counts[0]++;
    }
    orc_state_free(st);
    return 0;
}

/* record_probabilities (api/record_probabilities.cc:96-191): the caller
 * passes a program whose measurement opcodes are already the *_FORCED
 * variants; records are [n x num_measurements]; hidden slots (beyond
 * num_measurements) are rejected by the caller.  Seed 0 as in the
 * reference (:172). */
ORC_EXPORT int orc_record_log_probs(const prog_t* p, const uint8_t* records, uint64_t n, double* out) {
    orc_state* st = state_for(p, 0);
    if (!st) return -1;
    uint8_t* rec = (uint8_t*)calloc(p->total_meas_slots + 1, 1);
    for (uint64_t r = 0; r < n; ++r) {
        orc_state_reset(st);
        memcpy(rec, records + r * p->num_measurements, p->num_measurements);
        st->forced_record = rec;
        st->next_noise_idx = 0;
        draw_next_noise(st, p);
        orc_execute(p, st);
        out[r] = st->forced_reachable ? st->forced_log_prob : -1.7976931348623157e308;
    }
    free(rec);
    orc_state_free(st);
    return 0;
}

/* Run one shot with a caller-chosen fault set (ForcedFaults, svm.h:208-227;
 * the per-shot preparation of svm.cc:497-520 with the indices supplied). */
ORC_EXPORT int orc_execute_forced_faults(const prog_t* p, orc_state* st, const uint32_t* noise_idx,
                                         uint32_t n_noise, const uint32_t* ro_idx, uint32_t n_ro) {
    free(st->ff_noise);
    free(st->ff_ro);
    st->ff_noise = (uint32_t*)malloc((n_noise + 1) * 4);
    st->ff_ro = (uint32_t*)malloc((n_ro + 1) * 4);
    memcpy(st->ff_noise, noise_idx, n_noise * 4);
    memcpy(st->ff_ro, ro_idx, n_ro * 4);
    st->ff_n_noise = n_noise;
    st->ff_n_ro = n_ro;
    st->ff_noise_pos = 0;
    st->ff_ro_pos = 0;
    st->ff_active = 1;
    if (n_noise > 0) {
        st->next_noise_idx = noise_idx[0];
        st->ff_noise_pos = 1;
    } else {
        st->next_noise_idx = 0xFFFFFFFFu;
    }
    orc_execute(p, st);
    st->ff_active = 0;
    return 0;
}

/* ---------------------------------------------------------------------
 * Importance sampling with exactly k forced faults per shot
 * (svm.cc:347-622).  Sites are the quantum noise sites followed by the
 * readout entries.
 * ------------------------------------------------------------------- */

/* svm.cc:356-369 */
static double* site_probabilities(const prog_t* p, uint32_t* n_out) {
    uint32_t n = p->num_sites + p->num_readout;
    double* probs = (double*)malloc((n + 1) * sizeof(double));
    for (uint32_t s = 0; s < p->num_sites; ++s) {
        double acc = 0.0;
        for (uint32_t c = p->site_chan_off[s]; c < p->site_chan_off[s + 1]; ++c) acc += p->chan_prob[c];
        probs[s] = acc;
    }
    for (uint32_t e = 0; e < p->num_readout; ++e) probs[p->num_sites + e] = p->ro_prob[e];
    *n_out = n;
    return probs;
}

ORC_EXPORT uint32_t orc_noise_site_probabilities(const prog_t* p, double* out) {
    uint32_t n;
    double* probs = site_probabilities(p, &n);
    if (out) memcpy(out, probs, n * sizeof(double));
    free(probs);
    return n;
}

/* svm.cc:377-392: odds ratios w = p/(1-p), p clamped to [0, 1-1e-15],
 * rescaled to mean 1 */
static void odds_ratios(const double* probs, uint32_t n, double* w) {
    double sum = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double q = probs[i] < 0.0 ? 0.0 : (probs[i] > 1.0 - 1e-15 ? 1.0 - 1e-15 : probs[i]);
        w[i] = q / (1.0 - q);
        sum += w[i];
    }
    if (sum > 0.0) {
        double scale = (double)n / sum;
        for (uint32_t i = 0; i < n; ++i) w[i] *= scale;
    }
}

/* svm.cc:411-431: dp[i*(k+1)+j] = e_j(w[i..n)) */
static void dp_table(const double* w, uint32_t n, uint32_t k, double* dp) {
    const uint32_t stride = k + 1;
    memset(dp, 0, (size_t)(n + 1) * stride * sizeof(double));
    for (uint32_t i = 0; i <= n; ++i) dp[(size_t)i * stride] = 1.0;
    for (int i = (int)n - 1; i >= 0; --i) {
        uint32_t remaining = n - (uint32_t)i, max_j = remaining < k ? remaining : k;
        for (uint32_t j = 1; j <= max_j; ++j)
            dp[(size_t)i * stride + j] = dp[(size_t)(i + 1) * stride + j] + w[i] * dp[(size_t)(i + 1) * stride + j - 1];
    }
}

static int all_equal(const double* probs, uint32_t n) {
    for (uint32_t i = 1; i < n; ++i)
        if (probs[i] != probs[0]) return 0;
    return 1;
}

/* svm.cc:525-546; returns 0 or -2 (k above the site count) / -3 (zero mass) */
static int stratum_ok(const double* probs, uint32_t n, uint32_t k) {
    if (k > n) return -2;
    uint32_t certain = 0, impossible = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (probs[i] <= 0.0)
            impossible++;
        else if (probs[i] >= 1.0)
            certain++;
    }
    if (k < certain || k > n - impossible) return -3;
    return 0;
}

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : (x > y);
}

/* svm.cc:435-520: choose the fault set of one shot (DP walk, or partial
 * Fisher-Yates on `pool` with the chosen prefix sorted) and arm the state */
static void prepare_k_shot(orc_state* st, const prog_t* p, uint32_t k, uint32_t n, int uniform, const double* w,
                           const double* dp, uint32_t* pool, uint32_t* picks) {
    uint32_t m = 0;
    if (uniform) {
        for (uint32_t j = 0; j < k; ++j) {
            uint32_t remaining = n - j;
            uint32_t pick = j + (uint32_t)(rng_double(&st->rng) * remaining);
            uint32_t t = pool[j];
            pool[j] = pool[pick];
            pool[pick] = t;
        }
        qsort(pool, k, sizeof(uint32_t), cmp_u32);
        for (uint32_t j = 0; j < k; ++j) picks[m++] = pool[j];
    } else {
        const uint32_t stride = k + 1;
        uint32_t needed = k;
        for (uint32_t i = 0; i < n && needed > 0; ++i) {
            double inc;
            uint32_t remaining = n - i;
            if (needed == remaining) {
                inc = 1.0;
            } else {
                double denom = dp[(size_t)i * stride + needed];
                inc = denom > 0.0 ? (w[i] * dp[(size_t)(i + 1) * stride + needed - 1]) / denom : 0.0;
            }
            if (rng_double(&st->rng) < inc) {
                picks[m++] = i;
                needed--;
            }
        }
    }
    uint32_t nn = 0, nr = 0;
    free(st->ff_noise);
    free(st->ff_ro);
    st->ff_noise = (uint32_t*)malloc((m + 1) * 4);
    st->ff_ro = (uint32_t*)malloc((m + 1) * 4);
    for (uint32_t j = 0; j < m; ++j) {
        if (picks[j] < p->num_sites)
            st->ff_noise[nn++] = picks[j];
        else
            st->ff_ro[nr++] = picks[j] - p->num_sites;
    }
    st->ff_n_noise = nn;
    st->ff_n_ro = nr;
    st->ff_ro_pos = 0;
    st->ff_active = 1;
    if (nn > 0) {
        st->next_noise_idx = st->ff_noise[0];
        st->ff_noise_pos = 1;
    } else {
        st->next_noise_idx = 0xFFFFFFFFu;
        st->ff_noise_pos = 0;
    }
}

/* svm.cc:549-622 (sample_k).  mode 0: one stream and a Fisher-Yates pool
 * carried across shots (the reference); mode 1: per-shot streams and a
 * fresh identity pool per shot (the device convention; shot 0 agrees with
 * mode 0).  Returns 0, -2 (k exceeds the number of sites) or -3 (the
 * k-fault stratum has zero probability mass). */
ORC_EXPORT int orc_sample_k(const prog_t* p, uint64_t shots, uint32_t k, uint64_t seed, int mode, uint64_t first_shot,
                            uint8_t* meas, uint8_t* det, uint8_t* obs, double* ev, uint8_t* discarded) {
    uint32_t n;
    double* probs = site_probabilities(p, &n);
    int rc = stratum_ok(probs, n, k);
    if (rc) {
        free(probs);
        return rc;
    }
    const int uniform = all_equal(probs, n);
    double* w = (double*)calloc(n + 1, sizeof(double));
    double* dp = (double*)calloc((size_t)(n + 1) * (k + 1), sizeof(double));
    uint32_t* pool = (uint32_t*)malloc((n + 1) * 4);
    uint32_t* picks = (uint32_t*)malloc((k + 1) * 4);
    if (!uniform) {
        odds_ratios(probs, n, w);
        dp_table(w, n, k, dp);
    }
    for (uint32_t i = 0; i < n; ++i) pool[i] = i;
    orc_state* st = state_for(p, seed);
    for (uint64_t s = 0; s < shots; ++s) {
        if (s > 0) orc_state_reset(st);
        if (mode == 1) {
            rng_seed_shot(&st->rng, seed, first_shot + s);
            for (uint32_t i = 0; i < n; ++i) pool[i] = i;
        }
        prepare_k_shot(st, p, k, n, uniform, w, dp, pool, picks);
        orc_execute(p, st);
        st->ff_active = 0;
        copy_out(p, st, s, meas, det, obs, ev);
        if (discarded) discarded[s] = (uint8_t)st->discarded;
    }
    orc_state_free(st);
    free(probs);
    free(w);
    free(dp);
    free(pool);
    free(picks);
    return 0;
}

/* Seed the state's rng with the device per-shot convention. */
ORC_EXPORT void orc_state_seed_shot(orc_state* st, uint64_t seed, uint64_t shot) {
    rng_seed_shot(&st->rng, seed, shot);
}
ORC_EXPORT void orc_draw_next_noise(const prog_t* p, orc_state* st) { draw_next_noise(st, p); }
