/*
 * insem_oracle.h -- CPU restatement of insem's explicit time-stepping hot
 * path.  TEST INFRASTRUCTURE ONLY: the parity checker of the HIP backend and
 * the CPU baseline leg of bench.py.  The product never links or calls it.
 *
 * It consumes the same insem_hip_setup the HIP backend takes and mirrors the
 * insem_hip_* entry points one for one (same arguments, same layouts, same
 * status codes), so a test can drive both with the same calls.
 */
#ifndef INSEM_ORACLE_H
#define INSEM_ORACLE_H

#include "../include/insem_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

const char* oracle_last_error(void);
int oracle_create(const insem_hip_setup* setup, oracle_ctx** out);
int oracle_destroy(oracle_ctx* ctx);
int oracle_set_time(oracle_ctx* ctx, double dt, int32_t n_steps, const float* amplitudes);
int oracle_lambda_max(oracle_ctx* ctx, const float* start, double* lambda);
int oracle_set_state(oracle_ctx* ctx, int32_t domain, const float* u, const float* v);
int oracle_get_field(oracle_ctx* ctx, int32_t domain, int32_t field, float* out);
int oracle_advance(oracle_ctx* ctx, int32_t n);
int oracle_step(oracle_ctx* ctx, int32_t* step);
int oracle_n_samples(oracle_ctx* ctx, int32_t* n);
int oracle_read_traces(oracle_ctx* ctx, double* out);
int oracle_n_energy(oracle_ctx* ctx, int32_t* n);
int oracle_read_energy(oracle_ctx* ctx, double* out);
int oracle_internal_forces(oracle_ctx* ctx, int32_t domain, const float* u, float* f, double* dot);
int oracle_block_forces(oracle_ctx* ctx, int32_t domain, int32_t block, const float* u_local, float* f_local);
/* DofMap gather (u_local = C u) and assemble (f = C^T f_local) of one domain. */
int oracle_gather(oracle_ctx* ctx, int32_t domain, const float* global, float* local);
int oracle_assemble(oracle_ctx* ctx, int32_t domain, const float* local, float* global);
/* Threads used by the OpenMP loops (0 = runtime default). */
int oracle_set_threads(int32_t n);

#ifdef __cplusplus
}
#endif

#endif
