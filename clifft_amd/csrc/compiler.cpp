// compiler.cpp -- circuit text -> Program (include/clifft_compile.h).
//
// Representation.  The physical state is |Psi> = w * U * P * (phi (x) |0>_D)
// where U is a Clifford known at compile time, P the runtime Pauli frame, phi
// the runtime active array and w a scalar.  The compiler keeps the Heisenberg
// images T(X_q) = U^dag X_q U and T(Z_q) = U^dag Z_q U of every physical
// generator as signed Pauli strings on the virtual axes.
//
//  * A physical Clifford G updates U <- G U, i.e. T(P) <- T(G^dag P G), a
//    product of at most two images (no bytecode).
//  * A virtual Clifford W applied to the VM state (frame or array opcode)
//    updates U <- U W^dag, i.e. every image is conjugated by W.
//  * A non-Clifford rotation or a measurement of physical Z_q acts on the
//    virtual Pauli Q = T(Z_q).  Q is localized to a single axis (+-X_p or
//    +-Z_p) by CNOT/CZ/S virtual gates, then realized by one VM opcode.
//
// Localization picks, for X-support, a dormant pivot when one exists (the
// clearing CNOTs/CZs then have a dormant control and are frame-only), and for
// pure Z-support an active pivot (so a measurement can compact the array).
// This is the constructive localization lemma the reference implements in
// src/clifft/backend/backend.cc:420-523; the code here is written from the
// stabilizer algebra.  VM measurement conventions (checked against dense
// projectors in tests/test_oracle_kats.py): diagonal = project Z; interfere
// and dormant-random = project X, then H on that axis.
#include "clifft_compile.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

typedef std::complex<double> cd;
const double kInvSqrt2 = 0.70710678118654752440;
const double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Signed Pauli strings: i^e * X^x * Z^z (Z applied first)
// ---------------------------------------------------------------------------
struct Pauli {
    std::vector<uint64_t> x, z;
    int e = 0;
    Pauli() {}
    explicit Pauli(uint32_t words) : x(words, 0), z(words, 0) {}
    bool gx(uint32_t q) const { return (x[q >> 6] >> (q & 63)) & 1; }
    bool gz(uint32_t q) const { return (z[q >> 6] >> (q & 63)) & 1; }
    void fx(uint32_t q) { x[q >> 6] ^= 1ull << (q & 63); }
    void fz(uint32_t q) { z[q >> 6] ^= 1ull << (q & 63); }
    bool x_empty() const {
        for (uint64_t w : x)
            if (w) return false;
        return true;
    }
    bool z_empty() const {
        for (uint64_t w : z)
            if (w) return false;
        return true;
    }
    int xz_overlap() const {
        int c = 0;
        for (size_t w = 0; w < x.size(); ++w) c += __builtin_popcountll(x[w] & z[w]);
        return c;
    }
};

// (i^ea X^xa Z^za)(i^eb X^xb Z^zb) = i^(ea+eb) (-1)^|za.xb| X^(xa^xb) Z^(za^zb)
Pauli mul(const Pauli& a, const Pauli& b) {
    Pauli r(a.x.size());
    int par = 0;
    for (size_t w = 0; w < a.x.size(); ++w) {
        par += __builtin_popcountll(a.z[w] & b.x[w]);
        r.x[w] = a.x[w] ^ b.x[w];
        r.z[w] = a.z[w] ^ b.z[w];
    }
    r.e = (a.e + b.e + 2 * (par & 1)) & 3;
    return r;
}

// Conjugation W P W^dag by virtual gates (the same rules the VM applies to
// its frame, src/clifft/svm/svm_kernels.inl:67-116).
void conj_h(Pauli& p, uint32_t v) {
    bool x = p.gx(v), z = p.gz(v);
    if (x && z) p.e = (p.e + 2) & 3;
    if (x != z) {
        p.fx(v);
        p.fz(v);
    }
}
void conj_s(Pauli& p, uint32_t v, bool dag) {
    if (p.gx(v)) {
        p.e = (p.e + (dag ? 3 : 1)) & 3;
        p.fz(v);
    }
}
void conj_cnot(Pauli& p, uint32_t c, uint32_t t) {
    if (p.gx(c)) p.fx(t);
    if (p.gz(t)) p.fz(c);
}
void conj_cz(Pauli& p, uint32_t a, uint32_t b) {
    bool xa = p.gx(a), xb = p.gx(b);
    if (xa && xb) p.e = (p.e + 2) & 3;
    if (xa) p.fz(b);
    if (xb) p.fz(a);
}
void conj_swap(Pauli& p, uint32_t a, uint32_t b) {
    if (p.gx(a) != p.gx(b)) {
        p.fx(a);
        p.fx(b);
    }
    if (p.gz(a) != p.gz(b)) {
        p.fz(a);
        p.fz(b);
    }
}

// ---------------------------------------------------------------------------
// Circuit parsing (Stim text subset)
// ---------------------------------------------------------------------------
struct Target {
    enum Kind { QUBIT, REC, PAULI } kind;
    int64_t v = 0;                                    // qubit or rec offset
    std::vector<std::pair<char, uint32_t>> product;   // Pauli product terms
};
struct Op {
    std::string name;
    std::vector<double> args;
    std::vector<Target> targets;
};

std::string upper(std::string s) {
    for (char& c : s) c = (char)toupper((unsigned char)c);
    return s;
}

void parse_lines(const std::vector<std::string>& lines, size_t& i, std::vector<Op>& out, int depth) {
    if (depth > 16) throw std::runtime_error("REPEAT nesting too deep");
    while (i < lines.size()) {
        std::string line = lines[i++];
        size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        // trim
        size_t a = line.find_first_not_of(" \t\r");
        if (a == std::string::npos) continue;
        line = line.substr(a);
        size_t b = line.find_last_not_of(" \t\r");
        line = line.substr(0, b + 1);
        if (line == "}") return;
        // name[(args)] targets...
        size_t p = 0;
        while (p < line.size() && (isalnum((unsigned char)line[p]) || line[p] == '_')) ++p;
        Op op;
        op.name = upper(line.substr(0, p));
        std::string rest = line.substr(p);
        if (!rest.empty() && rest[0] == '(') {
            size_t close = rest.find(')');
            if (close == std::string::npos) throw std::runtime_error("unterminated '(' in: " + line);
            std::string args = rest.substr(1, close - 1);
            std::stringstream ss(args);
            std::string tok;
            while (std::getline(ss, tok, ',')) op.args.push_back(std::strtod(tok.c_str(), nullptr));
            rest = rest.substr(close + 1);
        }
        if (op.name == "REPEAT") {
            std::stringstream ss(rest);
            long long count = 0;
            std::string brace;
            ss >> count >> brace;
            if (brace != "{" || count < 0) throw std::runtime_error("malformed REPEAT: " + line);
            std::vector<Op> body;
            parse_lines(lines, i, body, depth + 1);
            for (long long r = 0; r < count; ++r) out.insert(out.end(), body.begin(), body.end());
            if (out.size() > 10000000) throw std::runtime_error("circuit too large after REPEAT unrolling");
            continue;
        }
        std::stringstream ss(rest);
        std::string tok;
        while (ss >> tok) {
            Target t;
            if (tok.rfind("rec[", 0) == 0) {
                t.kind = Target::REC;
                t.v = std::strtoll(tok.c_str() + 4, nullptr, 10);
            } else if (isdigit((unsigned char)tok[0])) {
                t.kind = Target::QUBIT;
                t.v = std::strtoll(tok.c_str(), nullptr, 10);
            } else {
                t.kind = Target::PAULI;
                std::stringstream ps(tok);
                std::string term;
                while (std::getline(ps, term, '*')) {
                    if (term.size() < 2) throw std::runtime_error("bad Pauli term: " + tok);
                    char c = (char)toupper((unsigned char)term[0]);
                    if (c != 'X' && c != 'Y' && c != 'Z') throw std::runtime_error("bad Pauli term: " + tok);
                    t.product.push_back({c, (uint32_t)std::strtoul(term.c_str() + 1, nullptr, 10)});
                }
            }
            op.targets.push_back(t);
        }
        out.push_back(op);
    }
}

// ---------------------------------------------------------------------------
// Lowering context
// ---------------------------------------------------------------------------
struct Mask {
    std::vector<uint64_t> x, z;
    uint8_t sign;
};

struct Lowerer {
    uint32_t n = 0, W = 1, k = 0, kmax = 0;
    std::vector<Pauli> img;  // 2q: T(X_q), 2q+1: T(Z_q)
    std::vector<clifft_instruction> bc;
    std::vector<int32_t> tag;  // parameter slot of each instruction, -1 if none
    std::vector<Mask> pauli_masks, ev_masks;
    std::vector<uint32_t> site_off{0};
    std::vector<Mask> chans;
    std::vector<double> chan_prob;
    std::vector<std::vector<uint32_t>> det_lists, obs_lists;
    uint32_t num_visible = 0, next_visible = 0, next_hidden = 0, num_det = 0, num_obs = 0, num_ev = 0;
    std::vector<int8_t> param_sign;
    cd global_weight{1.0, 0.0};

    clifft_instruction& emit(uint8_t op, uint32_t a1 = 0, uint32_t a2 = 0) {
        clifft_instruction in;
        std::memset(&in, 0, sizeof(in));
        in.opcode = op;
        in.axis_1 = (uint16_t)a1;
        in.axis_2 = (uint16_t)a2;
        bc.push_back(in);
        tag.push_back(-1);
        return bc.back();
    }

    void init(uint32_t nq) {
        n = nq;
        W = (n + 63) / 64;
        if (W == 0) W = 1;
        img.assign(2 * n, Pauli(W));
        for (uint32_t q = 0; q < n; ++q) {
            img[2 * q].fx(q);
            img[2 * q + 1].fz(q);
        }
    }

    // --- virtual gates: conjugate every image, emit the VM opcode ----------
    void v_cnot(uint32_t c, uint32_t t, Pauli* extra) {
        for (auto& p : img) conj_cnot(p, c, t);
        if (extra) conj_cnot(*extra, c, t);
        if (c >= k)
            emit(CLIFFT_OP_FRAME_CNOT, c, t);
        else
            emit(CLIFFT_OP_ARRAY_CNOT, c, t);
    }
    void v_cz(uint32_t a, uint32_t b, Pauli* extra) {
        for (auto& p : img) conj_cz(p, a, b);
        if (extra) conj_cz(*extra, a, b);
        emit((a >= k || b >= k) ? CLIFFT_OP_FRAME_CZ : CLIFFT_OP_ARRAY_CZ, a, b);
    }
    void v_s(uint32_t v, Pauli* extra) {
        for (auto& p : img) conj_s(p, v, false);
        if (extra) conj_s(*extra, v, false);
        emit(v >= k ? CLIFFT_OP_FRAME_S : CLIFFT_OP_ARRAY_S, v);
    }
    void v_swap(uint32_t a, uint32_t b, Pauli* extra) {
        if (a == b) return;
        for (auto& p : img) conj_swap(p, a, b);
        if (extra) conj_swap(*extra, a, b);
        emit(a >= k ? CLIFFT_OP_FRAME_SWAP : CLIFFT_OP_ARRAY_SWAP, a, b);
    }
    void v_h_active(uint32_t v, Pauli* extra) {
        for (auto& p : img) conj_h(p, v);
        if (extra) conj_h(*extra, v);
        emit(CLIFFT_OP_ARRAY_H, v);
    }
    // H on dormant axis v (swapped to axis k first): FRAME_H + EXPAND.
    uint32_t v_h_expand(uint32_t v, Pauli* extra) {
        if (v != k) v_swap(v, k, extra);
        v = k;
        for (auto& p : img) conj_h(p, v);
        if (extra) conj_h(*extra, v);
        emit(CLIFFT_OP_FRAME_H, v);
        emit(CLIFFT_OP_EXPAND, v);
        ++k;
        kmax = std::max(kmax, k);
        return v;
    }
    void conj_all_h(uint32_t v) {
        for (auto& p : img) conj_h(p, v);
    }

    static uint32_t lowest(const std::vector<uint64_t>& m, uint32_t from) {
        for (uint32_t w = from / 64; w < m.size(); ++w) {
            uint64_t bits = m[w];
            if (w == from / 64 && (from & 63)) bits &= ~((1ull << (from & 63)) - 1);
            if (bits) return w * 64 + (uint32_t)__builtin_ctzll(bits);
        }
        return UINT32_MAX;
    }

    // Localize Q (modified in place) to +-X_p (basis 'X') or +-Z_p ('Z').
    void localize(Pauli& Q, uint32_t& p, char& basis) {
        if (!Q.x_empty()) {
            p = lowest(Q.x, k);  // dormant X-support first (axes >= k)
            if (p == UINT32_MAX) p = lowest(Q.x, 0);
            for (uint32_t q = lowest(Q.x, 0); q != UINT32_MAX; q = lowest(Q.x, q + 1))
                if (q != p) v_cnot(p, q, &Q);
            for (uint32_t q = lowest(Q.z, 0); q != UINT32_MAX; q = lowest(Q.z, q + 1))
                if (q != p) v_cz(p, q, &Q);
            if (Q.gz(p)) v_s(p, &Q);
            basis = 'X';
        } else {
            if (Q.z_empty()) throw std::runtime_error("internal: identity Pauli in localization");
            // active axes are [0, k): the lowest Z bit is active whenever any is
            p = lowest(Q.z, 0);
            for (uint32_t q = lowest(Q.z, 0); q != UINT32_MAX; q = lowest(Q.z, q + 1))
                if (q != p) v_cnot(q, p, &Q);
            basis = 'Z';
        }
        if (Q.e & 1) throw std::runtime_error("internal: non-Hermitian localized Pauli");
    }

    // --- physical Cliffords (U <- G U) -------------------------------------
    void phys_h(uint32_t q) { std::swap(img[2 * q], img[2 * q + 1]); }
    void phys_s(uint32_t q, bool dag) {  // S^dag X S = -iXZ ; S X S^dag = iXZ
        Pauli t = mul(img[2 * q], img[2 * q + 1]);
        t.e = (t.e + (dag ? 1 : 3)) & 3;
        img[2 * q] = t;
    }
    void phys_pauli(uint32_t q, char c) {
        if (c == 'X' || c == 'Y') img[2 * q + 1].e = (img[2 * q + 1].e + 2) & 3;
        if (c == 'Z' || c == 'Y') img[2 * q].e = (img[2 * q].e + 2) & 3;
    }
    void phys_cx(uint32_t c, uint32_t t) {
        img[2 * c] = mul(img[2 * c], img[2 * t]);
        img[2 * t + 1] = mul(img[2 * c + 1], img[2 * t + 1]);
    }
    void phys_cz(uint32_t a, uint32_t b) {
        Pauli xa = mul(img[2 * a], img[2 * b + 1]);
        Pauli xb = mul(img[2 * a + 1], img[2 * b]);
        img[2 * a] = xa;
        img[2 * b] = xb;
    }
    void phys_swap(uint32_t a, uint32_t b) {
        std::swap(img[2 * a], img[2 * b]);
        std::swap(img[2 * a + 1], img[2 * b + 1]);
    }

    Pauli image_y(uint32_t q) {  // Y = i X Z
        Pauli t = mul(img[2 * q], img[2 * q + 1]);
        t.e = (t.e + 1) & 3;
        return t;
    }
    Pauli image_of(char c, uint32_t q) {
        if (c == 'X') return img[2 * q];
        if (c == 'Z') return img[2 * q + 1];
        return image_y(q);
    }
    static Mask frame_mask(const Pauli& p) { return Mask{p.x, p.z, (uint8_t)((p.e >> 1) & 1)}; }
    static Mask hermitian_mask(const Pauli& p) {
        int s = ((p.e - p.xz_overlap()) % 4 + 4) % 4;
        if (s & 1) throw std::runtime_error("internal: non-Hermitian probe");
        return Mask{p.x, p.z, (uint8_t)(s == 2)};
    }

    // --- rotation exp(-i alpha pi/2 Z_q) (T: alpha=1/4, T_DAG: -1/4) --------
    void rotation(uint32_t q, double alpha, int kind /*0 T, 1 T_DAG, 2 R_Z*/, int32_t param) {
        Pauli Q = img[2 * q + 1];
        uint32_t p;
        char basis;
        localize(Q, p, basis);
        if (basis == 'X') p = (p >= k) ? v_h_expand(p, &Q) : (v_h_active(p, &Q), p);
        int s = (Q.e == 2) ? -1 : 1;  // Q = s * Z_p
        if (kind == 0 || kind == 1) {
            bool t_dag = (kind == 1) != (s < 0);
            emit(t_dag ? CLIFFT_OP_ARRAY_T_DAG : CLIFFT_OP_ARRAY_T, p);
            // T = e^{i pi/8} R_Z(1/4); localized to -Z the array op is T^dag
            // up to e^{+-i pi/4}
            if (s < 0) global_weight *= std::polar(1.0, (kind == 0 ? 1 : -1) * kPi / 4);
        } else {
            double a = kPi * alpha * s;
            clifft_instruction& in = emit(CLIFFT_OP_ARRAY_ROT, p);
            in.p.math.weight_re = std::cos(a);
            in.p.math.weight_im = std::sin(a);
            tag.back() = param;
            if (param >= 0) {
                if (param_sign.size() <= (size_t)param) param_sign.resize(param + 1, 1);
                param_sign[param] = (int8_t)s;
            }
            global_weight *= std::polar(1.0, -a / 2);
        }
    }

    // --- measurement of physical Z_q into record slot `slot` ---------------
    void measure(uint32_t q, uint32_t slot) {
        Pauli Q = img[2 * q + 1];
        uint32_t p;
        char basis;
        localize(Q, p, basis);
        uint8_t flags = (Q.e == 2) ? CLIFFT_FLAG_SIGN : 0;
        if (basis == 'Z') {
            if (p >= k) {
                clifft_instruction& in = emit(CLIFFT_OP_MEAS_DORMANT_STATIC, p);
                in.flags = flags;
                in.p.classical.classical_idx = slot;
            } else {
                if (p != k - 1) v_swap(p, k - 1, &Q);
                clifft_instruction& in = emit(CLIFFT_OP_MEAS_ACTIVE_DIAGONAL, k - 1);
                in.flags = flags;
                in.p.classical.classical_idx = slot;
                --k;
            }
        } else {
            if (p >= k) {
                clifft_instruction& in = emit(CLIFFT_OP_MEAS_DORMANT_RANDOM, p);
                in.flags = flags;
                in.p.classical.classical_idx = slot;
                conj_all_h(p);
            } else {
                if (p != k - 1) v_swap(p, k - 1, &Q);
                clifft_instruction& in = emit(CLIFFT_OP_MEAS_ACTIVE_INTERFERE, k - 1);
                in.flags = flags;
                in.p.classical.classical_idx = slot;
                conj_all_h(k - 1);
                --k;
            }
        }
    }

    void conditional_x(uint32_t q, uint32_t slot) {
        pauli_masks.push_back(frame_mask(img[2 * q]));
        clifft_instruction& in = emit(CLIFFT_OP_APPLY_PAULI);
        in.p.pauli.cp_mask_idx = (uint32_t)pauli_masks.size() - 1;
        in.p.pauli.condition_idx = slot;
    }

    void noise_site(const std::vector<std::pair<Pauli, double>>& channels) {
        for (auto& c : channels) {
            chans.push_back(frame_mask(c.first));
            chan_prob.push_back(c.second);
        }
        site_off.push_back((uint32_t)chans.size());
        clifft_instruction& in = emit(CLIFFT_OP_NOISE);
        in.p.pauli.cp_mask_idx = (uint32_t)site_off.size() - 2;
    }
};

uint32_t qubit_of(const Target& t, uint32_t n) {
    if (t.kind != Target::QUBIT || t.v < 0 || (uint64_t)t.v >= n) throw std::runtime_error("bad qubit target");
    return (uint32_t)t.v;
}

// ---------------------------------------------------------------------------
// Bytecode passes (semantics of src/clifft/optimizer/*_pass.cc; own code)
// ---------------------------------------------------------------------------
typedef std::vector<clifft_instruction> Code;

void pass_noise_block(Code& bc, std::vector<int32_t>& tag) {
    Code out;
    std::vector<int32_t> ot;
    for (size_t i = 0; i < bc.size();) {
        if (bc[i].opcode == CLIFFT_OP_NOISE) {
            size_t j = i + 1;
            uint32_t start = bc[i].p.pauli.cp_mask_idx;
            while (j < bc.size() && bc[j].opcode == CLIFFT_OP_NOISE &&
                   bc[j].p.pauli.cp_mask_idx == start + (uint32_t)(j - i))
                ++j;
            if (j - i >= 2) {
                clifft_instruction in = bc[i];
                in.opcode = CLIFFT_OP_NOISE_BLOCK;
                in.p.pauli.cp_mask_idx = start;
                in.p.pauli.condition_idx = (uint32_t)(j - i);
                out.push_back(in);
                ot.push_back(-1);
                i = j;
                continue;
            }
        }
        out.push_back(bc[i]);
        ot.push_back(tag[i]);
        ++i;
    }
    bc.swap(out);
    tag.swap(ot);
}

void pass_multi_gate(Code& bc, std::vector<int32_t>& tag) {
    Code out;
    std::vector<int32_t> ot;
    for (size_t i = 0; i < bc.size();) {
        const clifft_instruction& a = bc[i];
        if (a.opcode == CLIFFT_OP_ARRAY_CNOT && a.axis_2 < 64) {
            uint32_t t = a.axis_2;
            uint64_t mask = 0;
            size_t j = i;
            while (j < bc.size() && bc[j].opcode == CLIFFT_OP_ARRAY_CNOT && bc[j].axis_2 == t && bc[j].axis_1 < 64 &&
                   !(mask & (1ull << bc[j].axis_1))) {
                mask |= 1ull << bc[j].axis_1;
                ++j;
            }
            if (j - i >= 2) {
                clifft_instruction in = a;
                in.opcode = CLIFFT_OP_ARRAY_MULTI_CNOT;
                in.axis_1 = (uint16_t)t;
                in.axis_2 = 0;
                std::memset(&in.p, 0, sizeof(in.p));
                in.p.multi_gate.mask = mask;
                out.push_back(in);
                ot.push_back(-1);
                i = j;
                continue;
            }
        } else if (a.opcode == CLIFFT_OP_ARRAY_CZ && i + 1 < bc.size() && bc[i + 1].opcode == CLIFFT_OP_ARRAY_CZ) {
            // common axis of the first two CZs
            const clifft_instruction& b = bc[i + 1];
            int common = -1;
            if (a.axis_1 == b.axis_1 || a.axis_1 == b.axis_2) common = a.axis_1;
            else if (a.axis_2 == b.axis_1 || a.axis_2 == b.axis_2) common = a.axis_2;
            if (common >= 0 && common < 64) {
                uint64_t mask = 0;
                size_t j = i;
                while (j < bc.size() && bc[j].opcode == CLIFFT_OP_ARRAY_CZ &&
                       (bc[j].axis_1 == common || bc[j].axis_2 == common)) {
                    uint32_t other = bc[j].axis_1 == common ? bc[j].axis_2 : bc[j].axis_1;
                    if (other >= 64 || (mask & (1ull << other))) break;
                    mask |= 1ull << other;
                    ++j;
                }
                if (j - i >= 2) {
                    clifft_instruction in = a;
                    in.opcode = CLIFFT_OP_ARRAY_MULTI_CZ;
                    in.axis_1 = (uint16_t)common;
                    in.axis_2 = 0;
                    std::memset(&in.p, 0, sizeof(in.p));
                    in.p.multi_gate.mask = mask;
                    out.push_back(in);
                    ot.push_back(-1);
                    i = j;
                    continue;
                }
            }
        }
        out.push_back(bc[i]);
        ot.push_back(tag[i]);
        ++i;
    }
    bc.swap(out);
    tag.swap(ot);
}

void pass_expand_phase(Code& bc, std::vector<int32_t>& tag) {
    Code out;
    std::vector<int32_t> ot;
    for (size_t i = 0; i < bc.size(); ++i) {
        if (bc[i].opcode == CLIFFT_OP_EXPAND && i + 1 < bc.size() && bc[i + 1].axis_1 == bc[i].axis_1 &&
            (bc[i + 1].opcode == CLIFFT_OP_ARRAY_T || bc[i + 1].opcode == CLIFFT_OP_ARRAY_T_DAG ||
             bc[i + 1].opcode == CLIFFT_OP_ARRAY_ROT)) {
            clifft_instruction in = bc[i + 1];
            in.opcode = bc[i + 1].opcode == CLIFFT_OP_ARRAY_T
                            ? CLIFFT_OP_EXPAND_T
                            : (bc[i + 1].opcode == CLIFFT_OP_ARRAY_T_DAG ? CLIFFT_OP_EXPAND_T_DAG : CLIFFT_OP_EXPAND_ROT);
            out.push_back(in);
            ot.push_back(tag[i + 1]);
            ++i;
            continue;
        }
        out.push_back(bc[i]);
        ot.push_back(tag[i]);
    }
    bc.swap(out);
    tag.swap(ot);
}

void pass_swap_meas(Code& bc, std::vector<int32_t>& tag) {
    Code out;
    std::vector<int32_t> ot;
    for (size_t i = 0; i < bc.size(); ++i) {
        if (bc[i].opcode == CLIFFT_OP_ARRAY_SWAP && i + 1 < bc.size() &&
            bc[i + 1].opcode == CLIFFT_OP_MEAS_ACTIVE_INTERFERE &&
            (bc[i + 1].axis_1 == bc[i].axis_1 || bc[i + 1].axis_1 == bc[i].axis_2)) {
            clifft_instruction in = bc[i + 1];
            uint32_t t = bc[i + 1].axis_1;
            uint32_t f = bc[i].axis_1 == t ? bc[i].axis_2 : bc[i].axis_1;
            in.opcode = CLIFFT_OP_SWAP_MEAS_INTERFERE;
            in.axis_1 = (uint16_t)f;
            in.axis_2 = (uint16_t)t;
            out.push_back(in);
            ot.push_back(-1);
            ++i;
            continue;
        }
        out.push_back(bc[i]);
        ot.push_back(tag[i]);
    }
    bc.swap(out);
    tag.swap(ot);
}

// Single-axis fusion into U2 nodes (behaviour of
// src/clifft/optimizer/single_axis_fusion_pass.cc:65-233): a run of
// single-axis ops on one axis with >= 3 array ops (or >= 2 with a rotation)
// becomes one ARRAY_U2 whose node holds, for each incoming frame state, the
// product matrix, the accumulated gamma factor and the outgoing frame state.
bool u2_fusible(uint8_t op, bool keep_rot) {
    switch (op) {
        case CLIFFT_OP_ARRAY_H:
        case CLIFFT_OP_ARRAY_S:
        case CLIFFT_OP_ARRAY_S_DAG:
        case CLIFFT_OP_ARRAY_T:
        case CLIFFT_OP_ARRAY_T_DAG:
        case CLIFFT_OP_FRAME_H:
        case CLIFFT_OP_FRAME_S:
        case CLIFFT_OP_FRAME_S_DAG:
            return true;
        case CLIFFT_OP_ARRAY_ROT:
            return !keep_rot;
        default:
            return false;
    }
}
bool u2_array(uint8_t op) { return op >= CLIFFT_OP_ARRAY_H && op <= CLIFFT_OP_ARRAY_ROT; }

void simulate_u2(const clifft_instruction& in, int& px, int& pz, cd M[4], cd& g) {
    cd U[4] = {1.0, 0.0, 0.0, 1.0};
    const cd I(0, 1), w(kInvSqrt2, kInvSqrt2), wc(kInvSqrt2, -kInvSqrt2);
    bool array_op = true;
    switch (in.opcode) {
        case CLIFFT_OP_ARRAY_H:
        case CLIFFT_OP_FRAME_H:
            U[0] = kInvSqrt2, U[1] = kInvSqrt2, U[2] = kInvSqrt2, U[3] = -kInvSqrt2;
            if (px && pz) g *= -1.0;
            std::swap(px, pz);
            array_op = in.opcode == CLIFFT_OP_ARRAY_H;
            break;
        case CLIFFT_OP_ARRAY_S:
        case CLIFFT_OP_FRAME_S:
            U[3] = I;
            if (px) g *= I;
            pz ^= px;
            array_op = in.opcode == CLIFFT_OP_ARRAY_S;
            break;
        case CLIFFT_OP_ARRAY_S_DAG:
        case CLIFFT_OP_FRAME_S_DAG:
            U[3] = -I;
            if (px) g *= -I;
            pz ^= px;
            array_op = in.opcode == CLIFFT_OP_ARRAY_S_DAG;
            break;
        case CLIFFT_OP_ARRAY_T:
            U[3] = px ? wc : w;
            if (px) g *= w;
            break;
        case CLIFFT_OP_ARRAY_T_DAG:
            U[3] = px ? w : wc;
            if (px) g *= wc;
            break;
        case CLIFFT_OP_ARRAY_ROT: {
            cd z(in.p.math.weight_re, in.p.math.weight_im);
            U[3] = px ? std::conj(z) : z;
            if (px) g *= z;
            break;
        }
        default:
            return;
    }
    if (!array_op) return;
    cd o[4] = {U[0] * M[0] + U[1] * M[2], U[0] * M[1] + U[1] * M[3], U[2] * M[0] + U[3] * M[2],
               U[2] * M[1] + U[3] * M[3]};
    for (int i = 0; i < 4; ++i) M[i] = o[i];
}

void pass_u2(Code& bc, std::vector<int32_t>& tag, std::vector<clifft_u2_node>& nodes, bool keep_rot) {
    Code out;
    std::vector<int32_t> ot;
    for (size_t i = 0; i < bc.size();) {
        if (!u2_fusible(bc[i].opcode, keep_rot)) {
            out.push_back(bc[i]);
            ot.push_back(tag[i]);
            ++i;
            continue;
        }
        uint16_t axis = bc[i].axis_1;
        size_t j = i;
        int n_array = 0;
        bool has_rot = false;
        while (j < bc.size() && u2_fusible(bc[j].opcode, keep_rot) && bc[j].axis_1 == axis) {
            if (u2_array(bc[j].opcode)) ++n_array;
            if (bc[j].opcode == CLIFFT_OP_ARRAY_ROT) has_rot = true;
            ++j;
        }
        if (n_array >= 3 || (n_array >= 2 && has_rot)) {
            clifft_u2_node node;
            std::memset(&node, 0, sizeof(node));
            for (int s = 0; s < 4; ++s) {
                cd M[4] = {1.0, 0.0, 0.0, 1.0};
                cd g(1.0, 0.0);
                int px = s & 1, pz = (s >> 1) & 1;
                for (size_t r = i; r < j; ++r) simulate_u2(bc[r], px, pz, M, g);
                for (int e = 0; e < 4; ++e) {
                    node.matrices[s][e][0] = M[e].real();
                    node.matrices[s][e][1] = M[e].imag();
                }
                node.gamma[s][0] = g.real();
                node.gamma[s][1] = g.imag();
                node.out_states[s] = (uint8_t)((pz << 1) | px);
            }
            clifft_instruction in;
            std::memset(&in, 0, sizeof(in));
            in.opcode = CLIFFT_OP_ARRAY_U2;
            in.axis_1 = axis;
            in.p.u.cp_idx = (uint32_t)nodes.size();
            nodes.push_back(node);
            out.push_back(in);
            ot.push_back(-1);
        } else {
            for (size_t r = i; r < j; ++r) {
                out.push_back(bc[r]);
                ot.push_back(tag[r]);
            }
        }
        i = j;
    }
    bc.swap(out);
    tag.swap(ot);
}

// Two-axis (tile) fusion into U4 nodes (behaviour of
// src/clifft/optimizer/tile_axis_fusion_pass.cc:393-458, node semantics
// :208-391).  A run starts at a two-axis array op on axes {lo, hi} and
// extends while each instruction acts on lo and/or hi only (single-axis
// array or frame ops on either axis, two-axis ops on exactly {lo, hi}); a
// run holding >= 3 array ops becomes one ARRAY_U4.  Its node stores, for
// each of the 16 incoming frame states s = (pz_hi, px_hi, pz_lo, px_lo), the
// 4x4 product acting on the tile index (b_hi << 1) | b_lo, the gamma factor
// and the outgoing frame state.  Parameterised rotations stay unfused when
// `keep_rot` is set.
struct TileFrame {
    int px[2], pz[2];  // [0] = lo axis, [1] = hi axis
};

int tile_kind(const clifft_instruction& in, bool keep_rot) {  // 1: one axis, 2: two axes, 0: neither
    switch (in.opcode) {
        case CLIFFT_OP_ARRAY_H: case CLIFFT_OP_ARRAY_S: case CLIFFT_OP_ARRAY_S_DAG: case CLIFFT_OP_ARRAY_T:
        case CLIFFT_OP_ARRAY_T_DAG: case CLIFFT_OP_FRAME_H: case CLIFFT_OP_FRAME_S: case CLIFFT_OP_FRAME_S_DAG:
            return 1;
        case CLIFFT_OP_ARRAY_ROT:
            return keep_rot ? 0 : 1;
        case CLIFFT_OP_ARRAY_CNOT: case CLIFFT_OP_ARRAY_CZ: case CLIFFT_OP_ARRAY_SWAP: case CLIFFT_OP_FRAME_CNOT:
        case CLIFFT_OP_FRAME_CZ: case CLIFFT_OP_FRAME_SWAP:
            return 2;
        default:
            return 0;
    }
}

bool tile_array_op(uint8_t op) {
    return (op >= CLIFFT_OP_ARRAY_CNOT && op <= CLIFFT_OP_ARRAY_SWAP) ||
           (op >= CLIFFT_OP_ARRAY_H && op <= CLIFFT_OP_ARRAY_ROT);
}

// Left-multiply the 4x4 M (row-major) by a 2x2 u acting on tile bit `pos`.
void tile_apply_1q(cd M[16], const cd u[4], int pos) {
    const int bit = 1 << pos;
    for (int r0 = 0; r0 < 4; ++r0) {
        if (r0 & bit) continue;
        const int r1 = r0 | bit;
        for (int col = 0; col < 4; ++col) {
            cd a = M[r0 * 4 + col], b = M[r1 * 4 + col];
            M[r0 * 4 + col] = u[0] * a + u[1] * b;
            M[r1 * 4 + col] = u[2] * a + u[3] * b;
        }
    }
}

void tile_swap_rows(cd M[16], int r0, int r1) {
    for (int col = 0; col < 4; ++col) std::swap(M[r0 * 4 + col], M[r1 * 4 + col]);
}

void simulate_tile(const clifft_instruction& in, uint16_t lo, TileFrame& f, cd M[16], cd& g) {
    const cd I(0, 1), w(kInvSqrt2, kInvSqrt2), wc(kInvSqrt2, -kInvSqrt2);
    const int pos = in.axis_1 == lo ? 0 : 1;  // single-axis ops: which tile bit
    int& px = f.px[pos];
    int& pz = f.pz[pos];
    const bool array = tile_array_op(in.opcode);
    switch (in.opcode) {
        case CLIFFT_OP_ARRAY_H:
        case CLIFFT_OP_FRAME_H: {
            if (px && pz) g *= -1.0;
            std::swap(px, pz);
            if (array) {
                const cd u[4] = {kInvSqrt2, kInvSqrt2, kInvSqrt2, -kInvSqrt2};
                tile_apply_1q(M, u, pos);
            }
            return;
        }
        case CLIFFT_OP_ARRAY_S:
        case CLIFFT_OP_ARRAY_S_DAG:
        case CLIFFT_OP_FRAME_S:
        case CLIFFT_OP_FRAME_S_DAG: {
            const bool dag = in.opcode == CLIFFT_OP_ARRAY_S_DAG || in.opcode == CLIFFT_OP_FRAME_S_DAG;
            if (px) g *= dag ? -I : I;
            pz ^= px;
            if (array) {
                const cd u[4] = {1.0, 0.0, 0.0, dag ? -I : I};
                tile_apply_1q(M, u, pos);
            }
            return;
        }
        case CLIFFT_OP_ARRAY_T:
        case CLIFFT_OP_ARRAY_T_DAG:
        case CLIFFT_OP_ARRAY_ROT: {
            cd z = in.opcode == CLIFFT_OP_ARRAY_T ? w : (in.opcode == CLIFFT_OP_ARRAY_T_DAG ? wc : cd(in.p.math.weight_re, in.p.math.weight_im));
            const cd u[4] = {1.0, 0.0, 0.0, px ? std::conj(z) : z};
            if (px) g *= z;
            tile_apply_1q(M, u, pos);
            return;
        }
        default:
            break;
    }
    // two-axis ops on exactly {lo, hi}
    const int cpos = in.axis_1 == lo ? 0 : 1, tpos = 1 - cpos;
    switch (in.opcode) {
        case CLIFFT_OP_ARRAY_CNOT:
        case CLIFFT_OP_FRAME_CNOT:
            f.px[tpos] ^= f.px[cpos];
            f.pz[cpos] ^= f.pz[tpos];
            if (in.opcode == CLIFFT_OP_ARRAY_CNOT) {
                // rows with the control bit set swap their target bit
                const int cb = 1 << cpos, tb = 1 << tpos;
                tile_swap_rows(M, cb, cb | tb);
            }
            return;
        case CLIFFT_OP_ARRAY_CZ:
        case CLIFFT_OP_FRAME_CZ:
            if (f.px[0] && f.px[1]) g *= -1.0;
            f.pz[1] ^= f.px[0];
            f.pz[0] ^= f.px[1];
            if (in.opcode == CLIFFT_OP_ARRAY_CZ)
                for (int col = 0; col < 4; ++col) M[12 + col] = -M[12 + col];
            return;
        case CLIFFT_OP_ARRAY_SWAP:
        case CLIFFT_OP_FRAME_SWAP:
            std::swap(f.px[0], f.px[1]);
            std::swap(f.pz[0], f.pz[1]);
            if (in.opcode == CLIFFT_OP_ARRAY_SWAP) tile_swap_rows(M, 1, 2);
            return;
        default:
            return;
    }
}

void pass_u4(Code& bc, std::vector<int32_t>& tag, std::vector<clifft_u4_node>& nodes, bool keep_rot) {
    Code out;
    std::vector<int32_t> ot;
    for (size_t i = 0; i < bc.size();) {
        const clifft_instruction& head = bc[i];
        const bool starts = head.opcode == CLIFFT_OP_ARRAY_CNOT || head.opcode == CLIFFT_OP_ARRAY_CZ ||
                            head.opcode == CLIFFT_OP_ARRAY_SWAP;
        if (!starts) {
            out.push_back(head);
            ot.push_back(tag[i]);
            ++i;
            continue;
        }
        const uint16_t lo = std::min(head.axis_1, head.axis_2), hi = std::max(head.axis_1, head.axis_2);
        size_t j = i + 1;
        int n_array = 1;
        while (j < bc.size()) {
            const clifft_instruction& in = bc[j];
            const int kind = tile_kind(in, keep_rot);
            bool fits = false;
            if (kind == 1) fits = in.axis_1 == lo || in.axis_1 == hi;
            if (kind == 2) fits = std::min(in.axis_1, in.axis_2) == lo && std::max(in.axis_1, in.axis_2) == hi;
            if (!fits) break;
            if (tile_array_op(in.opcode)) ++n_array;
            ++j;
        }
        if (n_array >= 3) {
            clifft_u4_node node;
            std::memset(&node, 0, sizeof(node));
            for (int s = 0; s < 16; ++s) {
                TileFrame f{{s & 1, (s >> 2) & 1}, {(s >> 1) & 1, (s >> 3) & 1}};
                cd M[16] = {};
                for (int d = 0; d < 4; ++d) M[d * 5] = 1.0;
                cd g(1.0, 0.0);
                for (size_t r = i; r < j; ++r) simulate_tile(bc[r], lo, f, M, g);
                clifft_u4_entry& e = node.entries[s];
                for (int r = 0; r < 4; ++r)
                    for (int col = 0; col < 4; ++col) {
                        e.matrix[r][col][0] = M[r * 4 + col].real();
                        e.matrix[r][col][1] = M[r * 4 + col].imag();
                    }
                e.gamma[0] = g.real();
                e.gamma[1] = g.imag();
                e.out_state = (uint8_t)((f.pz[1] << 3) | (f.px[1] << 2) | (f.pz[0] << 1) | f.px[0]);
            }
            clifft_instruction in;
            std::memset(&in, 0, sizeof(in));
            in.opcode = CLIFFT_OP_ARRAY_U4;
            in.axis_1 = lo;
            in.axis_2 = hi;
            in.p.u.cp_idx = (uint32_t)nodes.size();
            nodes.push_back(node);
            out.push_back(in);
            ot.push_back(-1);
        } else {
            for (size_t r = i; r < j; ++r) {
                out.push_back(bc[r]);
                ot.push_back(tag[r]);
            }
        }
        i = j;
    }
    bc.swap(out);
    tag.swap(ot);
}

}  // namespace

// ---------------------------------------------------------------------------
// Compiled module
// ---------------------------------------------------------------------------
struct clifft_compiled {
    uint32_t num_qubits = 0, peak_rank = 0, num_meas = 0, total_slots = 0, num_det = 0, num_obs = 0, num_ev = 0;
    uint32_t words = 1;
    std::vector<clifft_instruction> bc;
    std::vector<uint64_t> pauli_x, pauli_z, ev_x, ev_z, chan_x, chan_z;
    std::vector<uint8_t> pauli_sign, ev_sign;
    std::vector<uint32_t> site_off, det_off, det_tgt, obs_off, obs_tgt;
    std::vector<double> chan_prob, hazards;
    std::vector<clifft_u2_node> u2;
    std::vector<clifft_u4_node> u4;
    std::vector<uint32_t> param_instr;
    std::vector<int8_t> param_sign;
    cd global_weight{1.0, 0.0};
};

namespace {

void flatten(const std::vector<Mask>& ms, uint32_t W, std::vector<uint64_t>& x, std::vector<uint64_t>& z,
             std::vector<uint8_t>* sign) {
    for (const Mask& m : ms) {
        for (uint32_t w = 0; w < W; ++w) {
            x.push_back(m.x[w]);
            z.push_back(m.z[w]);
        }
        if (sign) sign->push_back(m.sign);
    }
}

void csr(const std::vector<std::vector<uint32_t>>& lists, std::vector<uint32_t>& off, std::vector<uint32_t>& tgt) {
    off.assign(1, 0);
    for (auto& l : lists) {
        tgt.insert(tgt.end(), l.begin(), l.end());
        off.push_back((uint32_t)tgt.size());
    }
}

clifft_compiled* compile_impl(const char* text, uint32_t options) {
    std::vector<std::string> lines;
    {
        std::stringstream ss(text);
        std::string l;
        while (std::getline(ss, l)) lines.push_back(l);
    }
    std::vector<Op> ops;
    size_t li = 0;
    parse_lines(lines, li, ops, 0);

    // qubit count and visible measurement count
    uint32_t n = 0, num_visible = 0;
    for (const Op& op : ops) {
        for (const Target& t : op.targets) {
            if (t.kind == Target::QUBIT && op.name != "DETECTOR" && op.name != "OBSERVABLE_INCLUDE")
                n = std::max<uint32_t>(n, (uint32_t)t.v + 1);
            for (auto& term : t.product) n = std::max<uint32_t>(n, term.second + 1);
        }
        if (op.name == "M" || op.name == "MR" || op.name == "MZ" || op.name == "MRZ")
            num_visible += (uint32_t)op.targets.size();
    }
    if (n == 0) n = 1;
    if (n > 65535) throw std::runtime_error("too many qubits");

    Lowerer L;
    L.init(n);
    L.num_visible = num_visible;
    uint32_t hidden = 0;
    uint32_t n_params = 0;

    for (const Op& op : ops) {
        const std::string& g = op.name;
        auto q_at = [&](size_t i) { return qubit_of(op.targets[i], n); };
        if (g == "TICK" || g == "QUBIT_COORDS" || g == "SHIFT_COORDS" || g == "I") continue;
        if (g == "H") {
            for (size_t i = 0; i < op.targets.size(); ++i) L.phys_h(q_at(i));
        } else if (g == "S" || g == "SQRT_Z") {
            for (size_t i = 0; i < op.targets.size(); ++i) L.phys_s(q_at(i), false);
        } else if (g == "S_DAG" || g == "SQRT_Z_DAG") {
            for (size_t i = 0; i < op.targets.size(); ++i) L.phys_s(q_at(i), true);
        } else if (g == "X" || g == "Y" || g == "Z") {
            for (size_t i = 0; i < op.targets.size(); ++i) L.phys_pauli(q_at(i), g[0]);
        } else if (g == "CX" || g == "CNOT" || g == "ZCX" || g == "CZ" || g == "SWAP") {
            if (op.targets.size() % 2) throw std::runtime_error(g + " needs qubit pairs");
            for (size_t i = 0; i < op.targets.size(); i += 2) {
                uint32_t a = q_at(i), b = q_at(i + 1);
                if (a == b) throw std::runtime_error(g + " on identical qubits");
                if (g == "CZ")
                    L.phys_cz(a, b);
                else if (g == "SWAP")
                    L.phys_swap(a, b);
                else
                    L.phys_cx(a, b);
            }
        } else if (g == "T" || g == "T_DAG") {
            for (size_t i = 0; i < op.targets.size(); ++i) L.rotation(q_at(i), g == "T" ? 0.25 : -0.25, g == "T" ? 0 : 1, -1);
        } else if (g == "R_Z") {
            if (op.args.size() != 1) throw std::runtime_error("R_Z needs one angle");
            for (size_t i = 0; i < op.targets.size(); ++i) L.rotation(q_at(i), op.args[0], 2, (int32_t)n_params++);
        } else if (g == "M" || g == "MZ" || g == "MR" || g == "MRZ") {
            if (!op.args.empty()) throw std::runtime_error(g + "(p) readout noise is not supported");
            for (size_t i = 0; i < op.targets.size(); ++i) {
                uint32_t q = q_at(i), slot = L.next_visible++;
                L.measure(q, slot);
                if (g == "MR" || g == "MRZ") L.conditional_x(q, slot);
            }
        } else if (g == "R" || g == "RZ") {
            for (size_t i = 0; i < op.targets.size(); ++i) {
                uint32_t q = q_at(i), slot = num_visible + hidden++;
                L.measure(q, slot);
                L.conditional_x(q, slot);
            }
        } else if (g == "X_ERROR" || g == "Y_ERROR" || g == "Z_ERROR" || g == "DEPOLARIZE1") {
            if (op.args.size() != 1) throw std::runtime_error(g + " needs one probability");
            double p = op.args[0];
            if (p < 0 || p > 1) throw std::runtime_error(g + " probability out of range");
            for (size_t i = 0; i < op.targets.size(); ++i) {
                uint32_t q = q_at(i);
                std::vector<std::pair<Pauli, double>> ch;
                if (g == "DEPOLARIZE1") {
                    for (char c : {'X', 'Y', 'Z'}) ch.push_back({L.image_of(c, q), p / 3});
                } else {
                    ch.push_back({L.image_of(g[0], q), p});
                }
                if (p > 0) L.noise_site(ch);
            }
        } else if (g == "DEPOLARIZE2") {
            if (op.args.size() != 1 || op.targets.size() % 2) throw std::runtime_error("DEPOLARIZE2(p) needs pairs");
            double p = op.args[0];
            const char* P4 = "IXYZ";
            for (size_t i = 0; i < op.targets.size(); i += 2) {
                uint32_t a = q_at(i), b = q_at(i + 1);
                std::vector<std::pair<Pauli, double>> ch;
                for (int u = 0; u < 4; ++u)
                    for (int v = 0; v < 4; ++v) {
                        if (!u && !v) continue;
                        Pauli pa(L.W);
                        if (u) pa = L.image_of(P4[u], a);
                        if (v) pa = mul(pa, L.image_of(P4[v], b));
                        ch.push_back({pa, p / 15});
                    }
                if (p > 0) L.noise_site(ch);
            }
        } else if (g == "DETECTOR" || g == "OBSERVABLE_INCLUDE") {
            std::vector<uint32_t> tg;
            for (const Target& t : op.targets) {
                if (t.kind != Target::REC || t.v >= 0 || -t.v > (int64_t)L.next_visible)
                    throw std::runtime_error(g + " needs rec[-k] targets within the record");
                tg.push_back((uint32_t)((int64_t)L.next_visible + t.v));
            }
            if (g == "DETECTOR") {
                L.det_lists.push_back(tg);
                clifft_instruction& in = L.emit(CLIFFT_OP_DETECTOR);
                in.p.pauli.cp_mask_idx = (uint32_t)L.det_lists.size() - 1;
                in.p.pauli.condition_idx = L.num_det++;
            } else {
                uint32_t idx = op.args.empty() ? 0 : (uint32_t)op.args[0];
                L.obs_lists.push_back(tg);
                clifft_instruction& in = L.emit(CLIFFT_OP_OBSERVABLE);
                in.p.pauli.cp_mask_idx = (uint32_t)L.obs_lists.size() - 1;
                in.p.pauli.condition_idx = idx;
                L.num_obs = std::max(L.num_obs, idx + 1);
            }
        } else if (g == "EXP_VAL") {
            for (const Target& t : op.targets) {
                if (t.kind != Target::PAULI) throw std::runtime_error("EXP_VAL needs Pauli products like X0*Z1");
                Pauli acc(L.W);
                for (auto& term : t.product) acc = mul(acc, L.image_of(term.first, term.second));
                L.ev_masks.push_back(Lowerer::hermitian_mask(acc));
                clifft_instruction& in = L.emit(CLIFFT_OP_EXP_VAL);
                in.p.exp_val.cp_exp_val_idx = (uint32_t)L.ev_masks.size() - 1;
                in.p.exp_val.exp_val_idx = L.num_ev++;
            }
        } else {
            throw std::runtime_error("unsupported instruction: " + g);
        }
    }

    std::vector<double> alphas;  // one per R_Z target, circuit order
    for (const Op& op : ops)
        if (op.name == "R_Z")
            for (size_t i = 0; i < op.targets.size(); ++i) alphas.push_back(op.args[0]);

    Code bc = L.bc;
    std::vector<int32_t> tag = L.tag;
    auto* c = new clifft_compiled();
    if (!(options & CLIFFT_COMPILE_NO_PASSES)) {
        pass_noise_block(bc, tag);
        pass_multi_gate(bc, tag);
        pass_expand_phase(bc, tag);
        pass_swap_meas(bc, tag);
        if (!(options & CLIFFT_COMPILE_NO_FUSION)) {
            const bool keep_rot = (options & CLIFFT_COMPILE_KEEP_ROTATIONS) != 0;
            pass_u4(bc, tag, c->u4, keep_rot);
            pass_u2(bc, tag, c->u2, keep_rot);
        }
    }
    c->num_qubits = n;
    c->words = L.W;
    c->peak_rank = L.kmax;
    c->num_meas = num_visible;
    c->total_slots = num_visible + hidden;
    c->num_det = L.num_det;
    c->num_obs = L.num_obs;
    c->num_ev = L.num_ev;
    c->bc = bc;
    flatten(L.pauli_masks, L.W, c->pauli_x, c->pauli_z, &c->pauli_sign);
    flatten(L.ev_masks, L.W, c->ev_x, c->ev_z, &c->ev_sign);
    flatten(L.chans, L.W, c->chan_x, c->chan_z, nullptr);
    c->chan_prob = L.chan_prob;
    c->site_off = L.site_off;
    double acc = 0.0;
    for (size_t s = 0; s + 1 < c->site_off.size(); ++s) {
        double psum = 0.0;
        for (uint32_t ch = c->site_off[s]; ch < c->site_off[s + 1]; ++ch) psum += c->chan_prob[ch];
        acc += -std::log1p(-std::min(psum, 1.0 - 0x1.0p-53));  // backend.cc:861-868
        c->hazards.push_back(acc);
    }
    csr(L.det_lists, c->det_off, c->det_tgt);
    csr(L.obs_lists, c->obs_off, c->obs_tgt);
    c->param_instr.assign(alphas.size(), UINT32_MAX);
    c->param_sign.assign(alphas.size(), 1);
    for (size_t i = 0; i < L.param_sign.size() && i < alphas.size(); ++i) c->param_sign[i] = L.param_sign[i];
    for (size_t i = 0; i < bc.size(); ++i)
        if (tag[i] >= 0) c->param_instr[(uint32_t)tag[i]] = (uint32_t)i;
    c->global_weight = L.global_weight;
    return c;
}

}  // namespace

extern "C" {

int clifft_compile(const char* text, uint32_t options, clifft_compiled** out) {
    if (!text || !out) {
        g_err = "null argument";
        return CLIFFT_E_INVALID;
    }
    *out = nullptr;
    try {
        *out = compile_impl(text, options);
    } catch (const std::exception& e) {
        g_err = e.what();
        return CLIFFT_E_INVALID;
    }
    return CLIFFT_OK;
}

int clifft_compiled_view(const clifft_compiled* c, clifft_program_view* v) {
    if (!c || !v) {
        g_err = "null argument";
        return CLIFFT_E_INVALID;
    }
    std::memset(v, 0, sizeof(*v));
    v->abi_version = CLIFFT_SVM_ABI_VERSION;
    v->num_qubits = c->num_qubits;
    v->peak_rank = c->peak_rank;
    v->num_measurements = c->num_meas;
    v->total_meas_slots = c->total_slots;
    v->num_detectors = c->num_det;
    v->num_observables = c->num_obs;
    v->num_exp_vals = c->num_ev;
    v->words = c->words;
    v->bytecode = c->bc.data();
    v->num_instructions = c->bc.size();
    v->pauli_x = c->pauli_x.data();
    v->pauli_z = c->pauli_z.data();
    v->pauli_sign = c->pauli_sign.data();
    v->num_pauli_masks = (uint32_t)c->pauli_sign.size();
    v->ev_x = c->ev_x.data();
    v->ev_z = c->ev_z.data();
    v->ev_sign = c->ev_sign.data();
    v->num_ev_masks = (uint32_t)c->ev_sign.size();
    v->site_chan_off = c->site_off.data();
    v->num_sites = (uint32_t)c->site_off.size() - 1;
    v->chan_x = c->chan_x.data();
    v->chan_z = c->chan_z.data();
    v->chan_prob = c->chan_prob.data();
    v->num_channels = (uint32_t)c->chan_prob.size();
    v->hazards = c->hazards.empty() ? nullptr : c->hazards.data();
    v->det_off = c->det_off.data();
    v->det_tgt = c->det_tgt.data();
    v->num_det_lists = (uint32_t)c->det_off.size() - 1;
    v->obs_off = c->obs_off.data();
    v->obs_tgt = c->obs_tgt.data();
    v->num_obs_lists = (uint32_t)c->obs_off.size() - 1;
    v->u2_nodes = c->u2.data();
    v->num_u2 = (uint32_t)c->u2.size();
    v->u4_nodes = c->u4.data();
    v->num_u4 = (uint32_t)c->u4.size();
    return CLIFFT_OK;
}

uint32_t clifft_compiled_num_params(const clifft_compiled* c) { return c ? (uint32_t)c->param_instr.size() : 0; }
const uint32_t* clifft_compiled_param_instr(const clifft_compiled* c) { return c ? c->param_instr.data() : nullptr; }
const int8_t* clifft_compiled_param_sign(const clifft_compiled* c) { return c ? c->param_sign.data() : nullptr; }

void clifft_compiled_global_weight(const clifft_compiled* c, double* re, double* im) {
    if (!c) return;
    if (re) *re = c->global_weight.real();
    if (im) *im = c->global_weight.imag();
}

void clifft_compiled_destroy(clifft_compiled* c) { delete c; }

const char* clifft_compile_last_error(void) { return g_err.c_str(); }

}  // extern "C"
