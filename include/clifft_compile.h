/*
 * clifft_compile.h -- C ABI of the host-side Program producer.
 *
 * Compiles a Stim-format circuit (the subset listed below) into the
 * bytecode + constant pool layout of include/clifft_svm.h.  It stands in
 * for the reference compile pipeline (parse -> trace -> lower -> bytecode
 * passes; src/python/clifft/__init__.py:324-374, src/clifft/backend/
 * backend.cc:640-1005, src/clifft/optimizer/pass_registry.h:67-94), which
 * depends on Stim and cannot be built here.  The emitted programs obey the
 * same VM contract (opcode semantics, static active-rank trajectory), but
 * their instruction sequences are this compiler's own.
 *
 * Supported instructions: H S S_DAG SQRT_Z SQRT_Z_DAG X Y Z I CX CNOT ZCX
 * CZ SWAP T T_DAG R_Z(a) M MR R DEPOLARIZE1(p) DEPOLARIZE2(p) X_ERROR(p)
 * Y_ERROR(p) Z_ERROR(p) DETECTOR OBSERVABLE_INCLUDE(i) EXP_VAL TICK
 * QUBIT_COORDS SHIFT_COORDS REPEAT n { ... }.  Angles are in half-turns
 * (R_Z(a) = exp(-i a pi/2 Z)).
 */
#ifndef CLIFFT_COMPILE_H
#define CLIFFT_COMPILE_H

#include <stdint.h>

#include "clifft_svm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Option bits */
#define CLIFFT_COMPILE_NO_FUSION 1u       /* skip the U2/U4 fusion passes            */
#define CLIFFT_COMPILE_KEEP_ROTATIONS 2u  /* never fuse R_Z rotations (parameter slots) */
#define CLIFFT_COMPILE_NO_PASSES 4u       /* emit the lowered bytecode only          */

typedef struct clifft_compiled clifft_compiled;

/* Compile `text`; on success *out owns the program arrays. */
int clifft_compile(const char* text, uint32_t options, clifft_compiled** out);

/* Fill a view whose pointers reference arrays owned by `c` (valid until
 * clifft_compiled_destroy). */
int clifft_compiled_view(const clifft_compiled* c, clifft_program_view* view);

/* Parameter slots: for each R_Z in circuit order, the index of the
 * instruction that applies it and the sign s in its weight
 * z = exp(i * pi * s * alpha).  Arrays owned by `c`. */
uint32_t clifft_compiled_num_params(const clifft_compiled* c);
const uint32_t* clifft_compiled_param_instr(const clifft_compiled* c);
const int8_t* clifft_compiled_param_sign(const clifft_compiled* c);

/* Accumulated global weight of the compiled program (cosmetic phase). */
void clifft_compiled_global_weight(const clifft_compiled* c, double* re, double* im);

void clifft_compiled_destroy(clifft_compiled* c);

/* Message of the last failing compile call on this thread. */
const char* clifft_compile_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CLIFFT_COMPILE_H */
