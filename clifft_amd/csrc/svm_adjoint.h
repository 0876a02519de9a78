// svm_adjoint.h -- expectation values and their gradients with respect to
// parameterised Z rotations, by one forward and one backward (adjoint) sweep.
//
// Scope: noiseless, measurement-free programs (Clifford array ops, T/ROT
// phases, U2/U4 nodes, EXPAND*, EXP_VAL probes).  The Pauli frame of such a
// program stays zero, so every frame-conditioned choice takes its frame-0
// branch (svm_kernels.inl:951-988, 1091-1297, 1310-1531, 876-945) and the
// state is phi_t = V_t ... V_1 |0>, with V_t unitary or, for EXPAND*, the
// isometry phi -> [phi ; z phi].  EXP_VAL j reports
//     <P_j> = Re(ph_j <phi|X^xa Z^za|phi>) / <phi|phi>      (:2031-2173)
// and the objective is F = sum_j c_j <P_j>.
//
// Parameter i sets the phase z = exp(i s_i theta_i) of one ARRAY_ROT or
// EXPAND_ROT on axis a (s_i = +-1 from compilation), so
//     d phi_t / d theta_i = i s_i Pi_a phi_t    (Pi_a: bit a of the index set)
// and dF/dtheta_i = -2 s_i Im <lambda_t | Pi_a phi_t>, where lambda is swept
// backward from lambda = 0: at EXP_VAL j, lambda += (c_j / D_j) Q_j phi
// (Q_j = ph_j X^xa Z^za, D_j = <phi|phi> at the probe); at each op,
// lambda <- V_t^dagger lambda and phi <- V_t^-1 phi (EXPAND: phi keeps its
// lower half, lambda <- lambda_lo + conj(z) lambda_hi).
#pragma once

#include "svm_adjoint_args.h"
#include "svm_sweeps.h"

namespace csvm {

// --------------------------------------------------------------------------
// Sweeps used only by the adjoint pass (workgroup teams)
// --------------------------------------------------------------------------
template <class AmpT>
CSVM_SWEEP void sweep_u2m(AmpT* amp, uint32_t k, uint32_t v, dv2 m00, dv2 m01, dv2 m10, dv2 m11, TeamGeom g) {
    k = uni(k); v = uni(v); g = uni_geom(g);
    const uint32_t half = 1u << (k - 1), vb = 1u << v;
    for (uint32_t j = g.lane; j < half; j += g.size) {
        uint32_t i0 = insert_zero_bit(j, v);
        dv2 a = amp[i0], b = amp[i0 | vb];
        amp[i0] = cadd(cmul(m00, a), cmul(m01, b));
        amp[i0 | vb] = cadd(cmul(m10, a), cmul(m11, b));
    }
    team_sync<true>();
}

// amp <- M^dagger amp on quads {base, |lo, |hi, |lo|hi}; row r of M^dagger
// is the conjugated column r of M.
template <class AmpT>
CSVM_SWEEP void sweep_u4_adj(AmpT* amp, uint32_t k, uint32_t lo, uint32_t hi, cptr<double> m, TeamGeom g) {
    k = uni(k); lo = uni(lo); hi = uni(hi); g = uni_geom(g);
    const uint32_t n = 1u << (k - 2), lb = 1u << lo, hb = 1u << hi;
    for (uint32_t i = g.lane; i < n; i += g.size) {
        uint32_t base = insert_two_zero_bits(i, lo, hi);
        uint32_t ix0 = base, ix1 = base | lb, ix2 = base | hb, ix3 = base | lb | hb;
        dv2 in0 = amp[ix0], in1 = amp[ix1], in2 = amp[ix2], in3 = amp[ix3];
        dv2 o[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cptr<double> col = opaque(m) + 2 * r;  // M[c][r] at col + 8 c
            dv2 acc = cmul(mk(col[0], -col[1]), in0);
            acc = cadd(acc, cmul(mk(col[8], -col[9]), in1));
            acc = cadd(acc, cmul(mk(col[16], -col[17]), in2));
            acc = cadd(acc, cmul(mk(col[24], -col[25]), in3));
            o[r] = acc;
        }
        amp[ix0] = o[0];
        amp[ix1] = o[1];
        amp[ix2] = o[2];
        amp[ix3] = o[3];
    }
    team_sync<true>();
}

// lambda[i] <- lambda[i] + conj(z) lambda[i + half], i < half = 2^(k-1)
template <class AmpT>
CSVM_SWEEP void sweep_unexpand_adj(AmpT* lam, uint32_t k, double zr, double zi, TeamGeom g) {
    k = uni(k); zr = unid(zr); zi = unid(zi); g = uni_geom(g);
    const uint32_t half = 1u << (k - 1);
    const dv2 zc = mk(zr, -zi);
    for (uint32_t i = g.lane; i < half; i += g.size) lam[i] = cadd(lam[i], cmul(zc, lam[i + half]));
    team_sync<true>();
}

// lambda[j] += w (-1)^popc((j ^ xa) & za) phi[j ^ xa]
template <class AmpT>
CSVM_SWEEP void sweep_pauli_acc(AmpT* lam, const AmpT* phi, uint32_t k, uint32_t xa, uint32_t za, double wr, double wi,
                                TeamGeom g) {
    k = uni(k); xa = uni(xa); za = uni(za); wr = unid(wr); wi = unid(wi); g = uni_geom(g);
    const uint32_t dim = 1u << k;
    const dv2 w = mk(wr, wi);
    for (uint32_t j = g.lane; j < dim; j += g.size) {
        uint32_t src = j ^ xa;
        dv2 t = cmul(w, phi[src]);
        if (__popc(src & za) & 1u) t = cneg(t);
        lam[j] = cadd(lam[j], t);
    }
    team_sync<true>();
}

// sum over indices with bit `axis` set of conj(lambda_i) phi_i
template <class AmpT>
CSVM_SWEEP double2 sweep_inner_upper(const AmpT* lam, const AmpT* phi, uint32_t k, uint32_t axis, TeamGeom g,
                                     lds_f64* red) {
    k = uni(k); axis = uni(axis); g = uni_geom(g);
    const uint32_t half = 1u << (k - 1), ab = 1u << axis;
    double v[2] = {0.0, 0.0};
    for (uint32_t j = g.lane; j < half; j += g.size) {
        uint32_t i = insert_zero_bit(j, axis) | ab;
        dv2 l = lam[i], p = phi[i];
        v[0] += l.x * p.x + l.y * p.y;
        v[1] += l.x * p.y - l.y * p.x;
    }
    team_reduce<true, 2>(v, g, red);
    return make_double2(v[0], v[1]);
}

template <class AmpT>
CSVM_SWEEP void sweep_zero(AmpT* amp, uint32_t n, TeamGeom g) {
    n = uni(n); g = uni_geom(g);
    for (uint32_t i = g.lane; i < n; i += g.size) amp[i] = mk(0.0, 0.0);
    team_sync<true>();
}

// --------------------------------------------------------------------------
// The forward/backward walk of one theta vector by one workgroup team
// --------------------------------------------------------------------------
// Consecutive EXP_VAL probes share one pass over phi (forward) and one pass
// over lambda (backward), at most kProbeGroup at a time.
typedef __attribute__((address_space(3))) ProbeRec lds_probe;
CSVM_INLINE void store_probe(lds_probe* p, uint32_t xa, uint32_t za, uint32_t out, double wr, double wi) {
    p->xa = xa;
    p->za = za;
    p->out = out;
    p->pad = 0;
    p->wr = wr;
    p->wi = wi;
}

CSVM_INLINE ProbeRec load_probe(const lds_probe* p) {
    ProbeRec r;
    r.xa = p->xa;
    r.za = p->za;
    r.out = p->out;
    r.pad = 0;
    r.wr = p->wr;
    r.wi = p->wi;
    return r;
}

// The sums of sweep_exp_val for n probes in one pass: out[2p], out[2p + 1]
// the numerator of probe p, out[2 kProbeGroup] the shared norm.
template <class AmpT>
CSVM_SWEEP void sweep_exp_val_group(const AmpT* amp, uint32_t k, const lds_probe* pr, uint32_t n, TeamGeom g,
                                    lds_f64* red, double (&out)[2 * kProbeGroup + 4]) {
    k = uni(k); n = uni(n); g = uni_geom(g);
    uint32_t xa[kProbeGroup], za[kProbeGroup];
#pragma unroll
    for (uint32_t p = 0; p < kProbeGroup; ++p) {
        xa[p] = p < n ? uni(pr[p].xa) : 0u;
        za[p] = p < n ? uni(pr[p].za) : 0u;
    }
    const uint32_t dim = 1u << k;
#pragma unroll
    for (uint32_t i = 0; i < 2 * kProbeGroup + 4; ++i) out[i] = 0.0;
    for (uint32_t j = g.lane; j < dim; j += g.size) {
        const dv2 a = amp[j];
        out[2 * kProbeGroup] += a.x * a.x + a.y * a.y;
#pragma unroll
        for (uint32_t p = 0; p < kProbeGroup; ++p) {
            if (p < n) {
                const dv2 b = amp[j ^ xa[p]];
                const double zs = (__popc(j & za[p]) & 1u) ? -1.0 : 1.0;
                const double cr = zs * a.x, ci = zs * a.y;
                out[2 * p] += b.x * cr + b.y * ci;
                out[2 * p + 1] += b.x * ci - b.y * cr;
            }
        }
    }
#pragma unroll
    for (uint32_t c = 0; c < 2 * kProbeGroup + 4; c += 4) {
        double t[4] = {out[c], out[c + 1], out[c + 2], out[c + 3]};
        team_reduce<true, 4>(t, g, red);
        out[c] = t[0];
        out[c + 1] = t[1];
        out[c + 2] = t[2];
        out[c + 3] = t[3];
    }
}

// lambda[j] += sum_p w_p (-1)^popc((j ^ xa_p) & za_p) phi[j ^ xa_p], the
// terms added in list order (the order of the one-probe sweeps).
template <class AmpT>
CSVM_SWEEP void sweep_pauli_acc_group(AmpT* lam, const AmpT* phi, uint32_t k, const lds_probe* pr, uint32_t n,
                                      TeamGeom g) {
    k = uni(k); n = uni(n); g = uni_geom(g);
    uint32_t xa[kProbeGroup], za[kProbeGroup];
    dv2 w[kProbeGroup];
#pragma unroll
    for (uint32_t p = 0; p < kProbeGroup; ++p) {
        xa[p] = p < n ? uni(pr[p].xa) : 0u;
        za[p] = p < n ? uni(pr[p].za) : 0u;
        w[p] = p < n ? mk(unid(pr[p].wr), unid(pr[p].wi)) : mk(0.0, 0.0);
    }
    const uint32_t dim = 1u << k;
    for (uint32_t j = g.lane; j < dim; j += g.size) {
        dv2 acc = lam[j];
#pragma unroll
        for (uint32_t p = 0; p < kProbeGroup; ++p) {
            if (p < n) {
                const uint32_t src = j ^ xa[p];
                dv2 t = cmul(w[p], phi[src]);
                if (__popc(src & za[p]) & 1u) t = cneg(t);
                acc = cadd(acc, t);
            }
        }
        lam[j] = acc;
    }
    team_sync<true>();
}

template <class AmpT>
struct AdjInterp {
    TeamGeom g;
    AmpT* phi;
    AmpT* lam;
    lds_f64* red;
    lds_probe* probes;  // kProbeGroup records in LDS
    uint32_t k;

    // phase of instruction pc: the theta-vector's parameter or the payload
    CSVM_INLINE void rot_phase(cptr<AdjArgs> A, const double* th, uint32_t pc, uint64_t w1, uint64_t w2, double& zr,
                               double& zi) {
        int32_t slot = cp(A->slot_of)[pc];
        if (slot >= 0) {
            double a = cp(A->psign)[slot] * th[slot];
            sincos(a, &zi, &zr);
            zr = unid(zr);
            zi = unid(zi);
        } else {
            zr = __longlong_as_double((long long)w1);
            zi = __longlong_as_double((long long)w2);
        }
    }

    CSVM_INLINE void fixed_phase(uint32_t op, double& zr, double& zi) {
        zr = kInvSqrt2;
        zi = (op == CLIFFT_OP_ARRAY_T_DAG || op == CLIFFT_OP_EXPAND_T_DAG) ? -kInvSqrt2 : kInvSqrt2;
    }

    // Pauli probe of EXP_VAL ci on the zero frame: returns false when it has
    // X support on a dormant axis (value 0); else (xa, za) on active axes and
    // the phase ph = i^popc(xa & za) * (-1)^sign.
    CSVM_INLINE bool probe(cptr<DevProgram> P, uint32_t ci, uint32_t& xa, uint32_t& za, double& ph_r, double& ph_i) {
        const uint32_t W = P->words;
        cptr<uint64_t> mx = cp(P->ev_x) + (size_t)ci * W;
        cptr<uint64_t> mz = cp(P->ev_z) + (size_t)ci * W;
        const uint32_t full = k / 64, rem = k % 64;
        if (rem > 0 && full < W && (mx[full] & ~((1ULL << rem) - 1ULL))) return false;
        for (uint32_t w = (rem == 0 ? full : full + 1); w < W; ++w)
            if (mx[w]) return false;
        xa = 0;
        za = 0;
        if (k > 0) {
            uint64_t am = (1ULL << k) - 1ULL;
            xa = (uint32_t)(mx[0] & am);
            za = (uint32_t)(mz[0] & am);
        }
        const uint32_t yc = __popc(xa & za) & 3u;
        ph_r = (yc == 0) ? 1.0 : (yc == 2 ? -1.0 : 0.0);
        ph_i = (yc == 1) ? 1.0 : (yc == 3 ? -1.0 : 0.0);
        if (load_u8(cp(P->ev_sign), ci)) {
            ph_r = -ph_r;
            ph_i = -ph_i;
        }
        return true;
    }

    CSVM_INLINE void forward(cptr<DevProgram> Pg, cptr<AdjArgs> Ag, const double* th, double* ev_row, lds_f64* den_row) {
        cptr<DevProgram> P0 = opaque(Pg);
        cptr<uint64_t> code = cp(reinterpret_cast<const uint64_t*>(P0->bc));
        const uint32_t n_instr = P0->n_instr;
        k = 0;
        if (g.lane == 0) phi[0] = mk(1.0, 0.0);
        team_sync<true>();
        for (uint32_t pc = 0; pc < n_instr; ++pc) {
            const uint64_t w0 = code[4 * pc], w1 = code[4 * pc + 1];
            const uint32_t op = (uint32_t)(w0 & 0xFF);
            const uint32_t a1 = (uint32_t)((w0 >> 32) & 0xFFFF), a2 = (uint32_t)(w0 >> 48);
            const uint32_t i0 = (uint32_t)w1, i1 = (uint32_t)(w1 >> 32);
            switch (op) {
                case CLIFFT_OP_ARRAY_CNOT: sweep_two_axis<true>(phi, k, a1, a2, PERM_CNOT, g); break;
                case CLIFFT_OP_ARRAY_CZ: sweep_two_axis<true>(phi, k, a1, a2, PERM_CZ, g); break;
                case CLIFFT_OP_ARRAY_SWAP: sweep_two_axis<true>(phi, k, a1, a2, PERM_SWAP, g); break;
                case CLIFFT_OP_ARRAY_MULTI_CNOT:
                    sweep_multi<true>(phi, k, a1, (uint32_t)(w1 & ((1ULL << k) - 1ULL)), 0u, g);
                    break;
                case CLIFFT_OP_ARRAY_MULTI_CZ:
                    sweep_multi<true>(phi, k, a1, (uint32_t)(w1 & ((1ULL << k) - 1ULL)), 1u, g);
                    break;
                case CLIFFT_OP_ARRAY_H: sweep_h<true>(phi, k, a1, g); break;
                case CLIFFT_OP_ARRAY_S: sweep_phase<true>(phi, k, a1, 0.0, 1.0, g); break;
                case CLIFFT_OP_ARRAY_S_DAG: sweep_phase<true>(phi, k, a1, 0.0, -1.0, g); break;
                case CLIFFT_OP_ARRAY_T:
                case CLIFFT_OP_ARRAY_T_DAG:
                case CLIFFT_OP_ARRAY_ROT: {
                    double zr, zi;
                    if (op == CLIFFT_OP_ARRAY_ROT)
                        rot_phase(opaque(Ag), th, pc, w1, code[4 * pc + 2], zr, zi);
                    else
                        fixed_phase(op, zr, zi);
                    if (a1 < k) sweep_phase<true>(phi, k, a1, zr, zi, g);
                    break;
                }
                case CLIFFT_OP_ARRAY_U2:
                    if (a1 < k) {
                        cptr<DevProgram> P = opaque(Pg);
                        sweep_u2<true>(phi, k, a1, &cp(P->u2)[i0].matrices[0][0][0], g);
                    }
                    break;
                case CLIFFT_OP_ARRAY_U4: {
                    cptr<DevProgram> P = opaque(Pg);
                    sweep_u4<true>(phi, k, a1, a2, &cp(P->u4)[i0].entries[0].matrix[0][0][0], g);
                    break;
                }
                case CLIFFT_OP_EXPAND:
                case CLIFFT_OP_EXPAND_T:
                case CLIFFT_OP_EXPAND_T_DAG:
                case CLIFFT_OP_EXPAND_ROT: {
                    double zr = 1.0, zi = 0.0;
                    if (op == CLIFFT_OP_EXPAND_ROT)
                        rot_phase(opaque(Ag), th, pc, w1, code[4 * pc + 2], zr, zi);
                    else if (op != CLIFFT_OP_EXPAND)
                        fixed_phase(op, zr, zi);
                    sweep_expand<true>(phi, k, zr, zi, op != CLIFFT_OP_EXPAND ? 1u : 0u, g);
                    k++;
                    break;
                }
                case CLIFFT_OP_EXP_VAL: {
                    // this probe and the EXP_VALs that follow it (frame
                    // opcodes between them act on the zero frame only)
                    cptr<DevProgram> P = opaque(Pg);
                    uint32_t ng = 0, last = pc;
                    for (uint32_t t = pc; t < n_instr; ++t) {
                        const uint64_t u0 = code[4 * t], u1 = code[4 * t + 1];
                        const uint32_t uop = (uint32_t)(u0 & 0xFF);
                        if (uop <= CLIFFT_OP_FRAME_SWAP) continue;
                        if (uop != CLIFFT_OP_EXP_VAL || ng == kProbeGroup) break;
                        last = t;
                        const uint32_t ci = (uint32_t)u1, oi = (uint32_t)(u1 >> 32);
                        uint32_t xa, za;
                        double ph_r, ph_i;
                        if (probe(P, ci, xa, za, ph_r, ph_i)) {
                            if (g.lane == 0) store_probe(probes + ng, xa, za, oi, ph_r, ph_i);
                            ++ng;
                        } else if (g.lane == 0) {  // X on a dormant axis: <P> = 0
                            if (ev_row) ev_row[oi] = 0.0;
                            den_row[oi] = 1.0;
                        }
                    }
                    pc = last;
                    if (ng) {
                        team_sync<true>();
                        double sums[2 * kProbeGroup + 4];
                        sweep_exp_val_group(phi, k, probes, ng, g, red, sums);
                        const double den = sums[2 * kProbeGroup];
                        if (g.lane == 0) {
#pragma unroll
                            for (uint32_t p = 0; p < kProbeGroup; ++p) {
                                if (p >= ng) break;
                                const ProbeRec r = load_probe(probes + p);
                                const double num_re = sums[2 * p] * r.wr - sums[2 * p + 1] * r.wi;
                                double res = den > 0.0 ? num_re / den : 0.0;
                                if (res > 1.0) res = 1.0;
                                if (res < -1.0) res = -1.0;
                                if (ev_row) ev_row[r.out] = res;
                                den_row[r.out] = den;
                            }
                        }
                        team_sync<true>();  // the list is rewritten by the next group
                    }
                    break;
                }
                default: break;  // frame ops act on the zero frame only
            }
        }
    }

    CSVM_INLINE void backward(cptr<DevProgram> Pg, cptr<AdjArgs> Ag, const double* th, const lds_f64* den_row,
                              double* grad_row) {
        cptr<DevProgram> P0 = opaque(Pg);
        cptr<uint64_t> code = cp(reinterpret_cast<const uint64_t*>(P0->bc));
        const uint32_t n_instr = P0->n_instr;
        sweep_zero(lam, 1u << k, g);
        for (uint32_t pc = n_instr; pc-- > 0;) {
            const uint64_t w0 = code[4 * pc], w1 = code[4 * pc + 1];
            const uint32_t op = (uint32_t)(w0 & 0xFF);
            const uint32_t a1 = (uint32_t)((w0 >> 32) & 0xFFFF), a2 = (uint32_t)(w0 >> 48);
            const uint32_t i0 = (uint32_t)w1, i1 = (uint32_t)(w1 >> 32);
            switch (op) {
                case CLIFFT_OP_ARRAY_CNOT:
                case CLIFFT_OP_ARRAY_CZ:
                case CLIFFT_OP_ARRAY_SWAP: {
                    const uint32_t kind = op == CLIFFT_OP_ARRAY_CNOT ? PERM_CNOT : (op == CLIFFT_OP_ARRAY_CZ ? PERM_CZ : PERM_SWAP);
                    sweep_two_axis<true>(phi, k, a1, a2, kind, g);
                    sweep_two_axis<true>(lam, k, a1, a2, kind, g);
                    break;
                }
                case CLIFFT_OP_ARRAY_MULTI_CNOT:
                case CLIFFT_OP_ARRAY_MULTI_CZ: {
                    const uint32_t m = (uint32_t)(w1 & ((1ULL << k) - 1ULL)), cz = op == CLIFFT_OP_ARRAY_MULTI_CZ;
                    sweep_multi<true>(phi, k, a1, m, cz, g);
                    sweep_multi<true>(lam, k, a1, m, cz, g);
                    break;
                }
                case CLIFFT_OP_ARRAY_H:
                    sweep_h<true>(phi, k, a1, g);
                    sweep_h<true>(lam, k, a1, g);
                    break;
                case CLIFFT_OP_ARRAY_S:
                case CLIFFT_OP_ARRAY_S_DAG: {
                    const double s = op == CLIFFT_OP_ARRAY_S ? -1.0 : 1.0;  // inverse phase
                    sweep_phase<true>(phi, k, a1, 0.0, s, g);
                    sweep_phase<true>(lam, k, a1, 0.0, s, g);
                    break;
                }
                case CLIFFT_OP_ARRAY_T:
                case CLIFFT_OP_ARRAY_T_DAG:
                case CLIFFT_OP_ARRAY_ROT: {
                    double zr, zi;
                    if (op == CLIFFT_OP_ARRAY_ROT)
                        rot_phase(opaque(Ag), th, pc, w1, code[4 * pc + 2], zr, zi);
                    else
                        fixed_phase(op, zr, zi);
                    if (a1 < k) {
                        const int32_t slot = cp(opaque(Ag)->slot_of)[pc];
                        if (slot >= 0 && grad_row) {
                            double2 w = sweep_inner_upper(lam, phi, k, a1, g, red);
                            if (g.lane == 0) grad_row[slot] = -2.0 * cp(opaque(Ag)->psign)[slot] * w.y;
                        }
                        sweep_phase<true>(phi, k, a1, zr, -zi, g);
                        sweep_phase<true>(lam, k, a1, zr, -zi, g);
                    } else if (g.lane == 0 && grad_row) {
                        const int32_t slot = cp(opaque(Ag)->slot_of)[pc];
                        if (slot >= 0) grad_row[slot] = 0.0;  // phase on a dormant |0> axis
                    }
                    break;
                }
                case CLIFFT_OP_ARRAY_U2:
                    if (a1 < k) {
                        cptr<DevProgram> P = opaque(Pg);
                        cptr<double> m = &cp(P->u2)[i0].matrices[0][0][0];
                        const dv2 c00 = mk(m[0], -m[1]), c01 = mk(m[4], -m[5]), c10 = mk(m[2], -m[3]),
                                  c11 = mk(m[6], -m[7]);
                        sweep_u2m(phi, k, a1, c00, c01, c10, c11, g);
                        sweep_u2m(lam, k, a1, c00, c01, c10, c11, g);
                    }
                    break;
                case CLIFFT_OP_ARRAY_U4: {
                    cptr<DevProgram> P = opaque(Pg);
                    cptr<double> m = &cp(P->u4)[i0].entries[0].matrix[0][0][0];
                    sweep_u4_adj(phi, k, a1, a2, m, g);
                    sweep_u4_adj(lam, k, a1, a2, m, g);
                    break;
                }
                case CLIFFT_OP_EXPAND:
                case CLIFFT_OP_EXPAND_T:
                case CLIFFT_OP_EXPAND_T_DAG:
                case CLIFFT_OP_EXPAND_ROT: {
                    double zr = 1.0, zi = 0.0;
                    if (op == CLIFFT_OP_EXPAND_ROT)
                        rot_phase(opaque(Ag), th, pc, w1, code[4 * pc + 2], zr, zi);
                    else if (op != CLIFFT_OP_EXPAND)
                        fixed_phase(op, zr, zi);
                    if (op == CLIFFT_OP_EXPAND_ROT && grad_row) {
                        const int32_t slot = cp(opaque(Ag)->slot_of)[pc];
                        if (slot >= 0) {
                            double2 w = sweep_inner_upper(lam, phi, k, k - 1, g, red);
                            if (g.lane == 0) grad_row[slot] = -2.0 * cp(opaque(Ag)->psign)[slot] * w.y;
                        }
                    }
                    sweep_unexpand_adj(lam, k, zr, zi, g);
                    k--;  // phi keeps its lower half
                    break;
                }
                case CLIFFT_OP_EXP_VAL: {
                    // this probe and the EXP_VALs before it (walking down;
                    // frame opcodes between them are no-ops here)
                    cptr<AdjArgs> A = opaque(Ag);
                    cptr<DevProgram> P = opaque(Pg);
                    uint32_t ng = 0, last = pc;
                    for (uint32_t t = pc + 1; t-- > 0;) {
                        const uint64_t u0 = code[4 * t], u1 = code[4 * t + 1];
                        const uint32_t uop = (uint32_t)(u0 & 0xFF);
                        if (uop <= CLIFFT_OP_FRAME_SWAP) continue;
                        if (uop != CLIFFT_OP_EXP_VAL || ng == kProbeGroup) break;
                        last = t;
                        const uint32_t ci = (uint32_t)u1, oi = (uint32_t)(u1 >> 32);
                        const double c = cp(A->coeffs)[oi];
                        const double den = den_row[oi];
                        uint32_t xa, za;
                        double ph_r, ph_i;
                        if (c != 0.0 && den > 0.0 && probe(P, ci, xa, za, ph_r, ph_i)) {
                            if (g.lane == 0) store_probe(probes + ng, xa, za, oi, c * ph_r / den, c * ph_i / den);
                            ++ng;
                        }
                    }
                    pc = last;
                    if (ng) {
                        team_sync<true>();
                        sweep_pauli_acc_group(lam, (const AmpT*)phi, k, probes, ng, g);
                    }
                    break;
                }
                default: break;
            }
        }
    }
};

}  // namespace csvm
