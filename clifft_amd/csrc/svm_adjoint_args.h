// svm_adjoint_args.h -- launch arguments of the adjoint-gradient kernel
// (shared by the kernel and the host code).
#pragma once

#include "svm_device.h"

namespace csvm {

struct AdjArgs {
    const double* thetas;     // [batch x num_params]
    const double* psign;      // [num_params]: +-1, the sign s_i of the phase exp(i s_i theta_i)
    const int32_t* slot_of;   // [n_instr]: parameter slot of the instruction, or -1
    const double* coeffs;     // [num_ev] (adjoint mode)
    double* ev_out;           // [batch x num_ev] or null
    double* grad_out;         // [batch x num_params] or null (forward only)
    dv2* gscratch;            // global tier: 2 x amp_elems per team
    uint64_t batch;
    uint32_t num_params;
    uint32_t amp_elems;       // 2^peak_rank
};

// Consecutive EXP_VAL probes share one pass over the state, at most
// kProbeGroup at a time; their records sit in the team's LDS.
constexpr uint32_t kProbeGroup = 8;
struct ProbeRec {
    uint32_t xa, za, out, pad;
    double wr, wi;  // forward: the probe phase; backward: the lambda weight
};
constexpr uint32_t kProbeBytes = kProbeGroup * sizeof(ProbeRec);

}  // namespace csvm
