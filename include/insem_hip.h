/*
 * insem_hip.h -- C-ABI of the MI355X (gfx950) backend for the explicit
 * time-stepping hot path of insem.
 *
 * The boundary is plain C: pointers, sizes and status codes; no C++ or torch
 * types cross it.  Every entry point returns INSEM_HIP_OK (0) or a negative
 * status; the message of the last failure on the calling thread is returned
 * by insem_hip_last_error().  A C++ caller rethrows it as insem::Error
 * (reference: cpp/include/insem/core/error.hpp:9-25), Python as InsemError.
 *
 * Data handed to insem_hip_create() is copied; the caller keeps ownership of
 * every host buffer.  The context owns all device memory and releases it in
 * insem_hip_destroy(): its blocks go to a per-device cache that later
 * contexts draw from (hipFree would wait for every stream of the device,
 * stalling a host that destroys one context while others have work queued);
 * insem_hip_release_cache() hands the cache back to the device.  A context
 * is not thread-safe (one host thread per context, as the reference's
 * Solver); contexts may be created and destroyed on several threads.
 *
 * Layouts follow the reference exactly (cpp/include/insem/ops/dof_map.hpp:15-18):
 *   global field:  [component][global point]
 *   local field:   [element][component][j * ngll + i]
 * The backend may store them differently on the device; every read/write
 * through this ABI uses the layouts above.
 */
#ifndef INSEM_HIP_H
#define INSEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define INSEM_HIP_OK 0
#define INSEM_HIP_EINVAL (-1)   /* invalid argument or unsupported configuration */
#define INSEM_HIP_EDEVICE (-2)  /* HIP runtime failure or no device */
#define INSEM_HIP_ESTATE (-3)   /* call not allowed in the current state */
#define INSEM_HIP_ENOMEM (-4)   /* device allocation failed */

/* Geometry kinds (reference: cpp/include/insem/geometry/geometry.hpp:17-22). */
#define INSEM_KIND_RECT 0
#define INSEM_KIND_AFFINE 1
#define INSEM_KIND_CURVED 2
#define INSEM_KIND_AXISYM 3

/* Media (reference: cpp/include/insem/sim/description.hpp:11). */
#define INSEM_MEDIUM_ACOUSTIC 0
#define INSEM_MEDIUM_ELASTIC 1

/* Recorded quantities (reference: description.hpp:225-244). */
#define INSEM_Q_FIELD 0             /* displacement (elastic) or pressure (acoustic): u */
#define INSEM_Q_VELOCITY 1          /* v */
#define INSEM_Q_ACCELERATION 2      /* a */
#define INSEM_Q_PARTICLE_VELOCITY 3 /* acoustic: trapezoid of -grad(p)/rho */

/* Fields for insem_hip_get_field / set_field. */
#define INSEM_FIELD_U 0
#define INSEM_FIELD_V 1
#define INSEM_FIELD_A 2
#define INSEM_FIELD_F 3

/*
 * One homogeneous element group (reference: ops::ElementBlock,
 * cpp/include/insem/ops/element_block.hpp:31-46).
 *   metric: Rect 3 per element (xi_x, eta_z, jac); Affine 5 per element;
 *           Curved and Axisym 5 per point (xi_x, xi_z, eta_x, eta_z, jac),
 *           [block element][k][5].
 *   coef0/coef1: per point [block element][k]; elastic lambda, mu;
 *           acoustic 1/rho and unused.
 *   wr, hoop, r: Axisym only, per point [block element][k].
 *   axial: Axisym only, per block element.
 */
typedef struct insem_hip_block {
  int32_t kind;
  int32_t azimuthal_order;
  int32_t n_elements;
  const int32_t* elements; /* domain-local element index of each block element */
  const float* metric;
  const float* coef0;
  const float* coef1;
  const float* wr;
  const float* hoop;
  const float* r;
  const uint8_t* axial;
} insem_hip_block;

/*
 * One medium (reference: solver::Domain, cpp/include/insem/solver/solver.hpp:25-66,
 * with its ops::DofMap, dof_map.hpp:30-92).
 *   l2g: [element][k]; >= 0 a global point, < 0 the constrained slot of row -1 - l2g.
 *   rows: constraint rows (row_offsets has n_rows + 1 entries), weights double.
 *   t_offsets/t_slots(/t_weights): transpose CSR, slots ascending per point;
 *     t_weights NULL when the map has no constraints (all weights 1).
 *   mass, inv_mass: lumped diagonal mass and its inverse, [n_global].
 *   absorbing: Stacey points and their ncomp x ncomp damping matrices (row-major).
 *   dirichlet: points whose components are all zero; axis points zero the listed
 *     axis components.
 */
typedef struct insem_hip_domain {
  int32_t medium;
  int32_t ncomp;
  int32_t azimuthal_order;
  int32_t n_elements;
  int32_t n_global;
  const int32_t* l2g;
  int32_t n_rows;
  const int32_t* row_offsets;
  const int32_t* row_dofs;
  const double* row_weights;
  const int32_t* t_offsets;
  const int32_t* t_slots;
  const double* t_weights;
  const float* mass;
  const float* inv_mass;
  int32_t n_absorbing;
  const int32_t* absorbing_points;
  const float* absorbing_coef;
  int32_t n_dirichlet;
  const int32_t* dirichlet_points;
  int32_t n_axis;
  const int32_t* axis_points;
  int32_t n_axis_components;
  const int32_t* axis_components;
  int32_t n_blocks;
  const insem_hip_block* blocks;
} insem_hip_domain;

/* Fluid-solid interface (reference: Solver::Coupling, solver.hpp:208-211). */
typedef struct insem_hip_coupling {
  int32_t n_points;
  const int32_t* fluid_points;
  const int32_t* solid_points;
  const float* cx;
  const float* cz;
} insem_hip_coupling;

/* Point source (reference: Solver::SourceTerm, solver.hpp:155-160); coef is [comp][point]. */
typedef struct insem_hip_source {
  int32_t domain;
  int32_t n_points;
  const int32_t* points;
  const float* coef;
} insem_hip_source;

/* Receiver (reference: Solver::ReceiverTerm, solver.hpp:161-172). */
typedef struct insem_hip_receiver {
  int32_t domain;
  int32_t n_points;
  const int32_t* points;
  const float* value;
  const float* d_dx;
  const float* d_dz;
  double rho;
  int32_t n_quantities;
  const int32_t* quantities; /* INSEM_Q_* */
} insem_hip_receiver;

/*
 * Everything the time loop needs (reference: the state Solver::Solver builds,
 * cpp/src/solver/solver.cpp:91-145).  Domains are in the reference's order
 * (acoustic first, then elastic, solver.cpp:156).
 */
typedef struct insem_hip_setup {
  int32_t ngll;
  const float* D;  /* [i * ngll + l] = l_l'(x_i) on GLL */
  const float* Dj; /* same on GLJ(0,1) */
  const float* w2; /* [j * ngll + i] = w_i w_j */
  int32_t n_domains;
  const insem_hip_domain* domains;
  insem_hip_coupling coupling;
  int32_t n_sources;
  const insem_hip_source* sources;
  int32_t n_receivers;
  const insem_hip_receiver* receivers;
  int32_t record_every;
  int32_t energy_every;
} insem_hip_setup;

typedef struct insem_hip_ctx insem_hip_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char* insem_hip_last_error(void);

/* Number of visible devices (0 without a GPU; never fails). */
int insem_hip_device_count(void);
/* The calling thread's current device, and setting it (contexts are created
 * on the current device).  No counterpart in the reference, which has no
 * device. */
int insem_hip_get_device(int32_t* device);
int insem_hip_set_device(int32_t device);

/*
 * Uploads a setup and allocates the fields (zero).  Replaces the CPU state
 * the reference's Solver constructor builds (solver.cpp:91-145) on the device.
 */
int insem_hip_create(const insem_hip_setup* setup, insem_hip_ctx** out);
int insem_hip_destroy(insem_hip_ctx* ctx);

/*
 * insem_hip_create in two halves, so that a host thread that must not stall
 * (a request loop) makes only the short device calls and the set-up's host
 * work runs elsewhere.  No counterpart in the reference (its Solver
 * constructor builds host vectors, solver.cpp:91-145).
 *   insem_hip_stage_size   host only: the device bytes a context of the
 *                          setup takes (its arrays laid end to end).
 *   insem_hip_reserve      device: a block of at least that many bytes on
 *                          the current device (from the block cache).
 *   insem_hip_stage        host only, any thread: the context laid out for
 *                          the block, its arrays in a host image; no device
 *                          call is made.  The block stays reserved if it
 *                          fails (insem_hip_unreserve gives it back).
 *   insem_hip_realize      device: copies the image into the block; the
 *                          context is then what insem_hip_create returns
 *                          (bitwise the same arrays).  A staged context
 *                          refuses every other call until then, and
 *                          insem_hip_destroy of a staged context gives its
 *                          block back.
 */
int insem_hip_stage_size(const insem_hip_setup* setup, int64_t* bytes);
int insem_hip_reserve(int64_t bytes, void** block);
int insem_hip_unreserve(void* block);
int insem_hip_stage(const insem_hip_setup* setup, void* block, insem_hip_ctx** out);
int insem_hip_realize(insem_hip_ctx* ctx);

/*
 * Time step, step count and source amplitudes.  amplitudes is
 * [n_sources][n_steps + 1], amplitudes[s][n] = float(stf_s(n * dt)), the value
 * the reference adds at step n (solver.cpp:681).  Must be called before the
 * first advance; replaces the dt/n_steps logic of solver.cpp:131-144.
 */
int insem_hip_set_time(insem_hip_ctx* ctx, double dt, int32_t n_steps, const float* amplitudes);

/*
 * lambda_max(M^-1 K) by power iteration on the device (FP64 reductions):
 * reference Solver::estimate_stable_dt and coupled_lambda_max
 * (solver.cpp:583-667).  start holds the start vectors in the reference's
 * draw order: one vector per domain ([ncomp][n_global], domain order), then,
 * for coupled models, solid u then fluid u.  The fields are restored.
 */
int insem_hip_lambda_max(insem_hip_ctx* ctx, const float* start, double* lambda);
/* The same in two halves: _start queues the iterations on the context's
 * stream and returns (the start vectors are copied), _result waits and
 * returns lambda.  Power iterations of several contexts then overlap on the
 * device (a batch sets up its models this way). */
int insem_hip_lambda_start(insem_hip_ctx* ctx, const float* start);
int insem_hip_lambda_result(insem_hip_ctx* ctx, double* lambda);

/* Initial conditions of one domain ([ncomp][n_global]); before the first step
 * only (reference: Solver::set_state, solver.cpp:764-778). */
int insem_hip_set_state(insem_hip_ctx* ctx, int32_t domain, const float* u, const float* v);

/* Reads a global field of one domain ([ncomp][n_global]). */
int insem_hip_get_field(insem_hip_ctx* ctx, int32_t domain, int32_t field, float* out);

/* Advances n steps, clipped to n_steps (reference: Solver::advance,
 * solver.cpp:917-963, with initialize 878-898 on first use). */
int insem_hip_advance(insem_hip_ctx* ctx, int32_t n);

/* Captures and instantiates, without running them, the step-chunk graphs that
 * insem_hip_advance(ctx, n) would replay from the current step, so that a
 * later advance of that length only launches.  No counterpart in the
 * reference (which has no graphs); a timing harness calls it before its
 * timed region, and a serving loop before its first chunk. */
int insem_hip_prepare(insem_hip_ctx* ctx, int32_t n);

/* Autotuner variants of the paired element-force launch of a time step (no
 * counterpart in the reference, whose kernels have one shape).  variant -1
 * (the default) takes the winner that the committed tuning table records for
 * the model's order and element kinds; 0 .. 5 force one variant: level
 * (variant % 3) budgets the kernel's registers for the default waves per SIMD,
 * at most 6 or at most 4, and variant / 3 picks LLVM's max-ilp (0) or
 * default (1) scheduling strategy.  Orders 4 to 6 have all six variants,
 * other orders run variant 0.  Every variant gives the same bits; a change
 * drops the captured step chunks.  The environment variable
 * INSEM_HIP_FORCE_VARIANT sets the initial value. */
int insem_hip_set_force_variant(insem_hip_ctx* ctx, int32_t variant);
/* The variant the context's paired launch runs, and whether the model has a
 * paired launch at all (paired = 0: single-domain or batch-less models whose
 * first blocks have no paired kernel; variant is then 0). */
int insem_hip_force_variant(insem_hip_ctx* ctx, int32_t* variant, int32_t* paired);
/* Steps per captured hipGraph chunk of insem_hip_advance (no counterpart in
 * the reference): 1 .. 1024 fixes it, 0 returns to the tuned table's value
 * for the model's order and element kinds (32 where it records none).  The
 * chunk length changes no result bit; a change drops the captured chunks.
 * The environment variable INSEM_HIP_CHUNK sets the initial value. */
int insem_hip_set_chunk(insem_hip_ctx* ctx, int32_t steps);
int insem_hip_chunk(insem_hip_ctx* ctx, int32_t* steps);

/* Steps done. */
int insem_hip_step(insem_hip_ctx* ctx, int32_t* step);

/* Steps the device has completed, without waiting: the step count reached by
 * the latest insem_hip_advance call whose launches have all finished (no
 * counterpart in the reference, whose advance returns when its steps are
 * done).  A caller that keeps several advance calls queued (a serving loop)
 * reads results of models whose last step is at most this. */
int insem_hip_progress(insem_hip_ctx* ctx, int32_t* step);
/* The trace rows of the steps the device has completed (insem_hip_progress),
 * laid out as insem_hip_read_traces lays out all rows, copied without waiting
 * for the advance calls still queued (no counterpart in the reference); the
 * row count goes to n_samples.  A serving loop reads a batch model's traces
 * this way as soon as the model's last step is done. */
int insem_hip_read_traces_done(insem_hip_ctx* ctx, double* out, int32_t* n_samples);

/*
 * Receiver traces.  n_samples = samples recorded so far.  For receiver r and
 * quantity q the trace is [n_samples][ncomp] doubles, concatenated in
 * receiver/quantity order.  Elastic azimuthal-order-1 traces hold
 * (U_r, U_theta, U_z) (reference: Solver::record, solver.cpp:780-811).
 */
int insem_hip_n_samples(insem_hip_ctx* ctx, int32_t* n_samples);
int insem_hip_read_traces(insem_hip_ctx* ctx, double* out);

/* Energy rows [n_rows][3] = (time, kinetic, potential) (solver.cpp:900-915). */
int insem_hip_n_energy(insem_hip_ctx* ctx, int32_t* n_rows);
int insem_hip_read_energy(insem_hip_ctx* ctx, double* out);

/*
 * Operator-level seams for unit parity.
 * insem_hip_internal_forces: f = gather -> all blocks -> assemble on the
 *   given global field, returns u.f in double when dot != NULL
 *   (reference: Solver::internal_forces, solver.cpp:570-581).
 * insem_hip_block_forces: one block on an element-local field, writing only the
 *   block's elements (reference: physics::elastic_internal_forces /
 *   acoustic_internal_forces, cpp/include/insem/physics/kernels.hpp:32-35).
 *   u_local and f_local are [n_elements][ncomp][npe] of the domain.
 */
int insem_hip_internal_forces(insem_hip_ctx* ctx, int32_t domain, const float* u, float* f, double* dot);
int insem_hip_block_forces(insem_hip_ctx* ctx, int32_t domain, int32_t block, const float* u_local,
                           float* f_local);
/*
 * insem_hip_block_forces without any device allocation: no counterpart in
 * the reference.  It is for callers that cannot allocate (hipMalloc) at call
 * time: code under graph capture, tight serving loops, harnesses that forbid
 * allocation.  Same arguments, results and errors as insem_hip_block_forces
 * (bitwise equal f_local), but its device output staging is a fixed 32 MiB
 * array per GPU (held from the library's load until the process ends)
 * instead of a buffer allocated per call.  A domain whose element-local
 * field exceeds 8,388,608 values (n_elements * ncomp * npe) is refused with
 * INSEM_HIP_EINVAL.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (from several host threads or contexts) race and
 * must be serialised by the caller; calls in sequence are independent (the
 * caller's f_local is copied in first, so nothing of an earlier call
 * remains in the result).
 */
int insem_hip_block_forces_noalloc(insem_hip_ctx* ctx, int32_t domain, int32_t block, const float* u_local,
                                   float* f_local);

/* Blocks until all work of the context has finished. */
int insem_hip_synchronize(insem_hip_ctx* ctx);
/* *idle = 1 when all queued work of the context has finished, else 0; does
 * not block (a request loop polls it to keep several contexts busy). */
int insem_hip_query(insem_hip_ctx* ctx, int32_t* idle);
/* Free and total bytes of the current device (admission control); free
 * bytes include the blocks cached for new contexts. */
int insem_hip_device_memory(int64_t* free_bytes, int64_t* total_bytes);
/* Frees the current device's cached blocks (those of destroyed contexts),
 * waiting for the device's streams; *bytes = bytes released (may be NULL).
 * No counterpart in the reference, whose buffers are host vectors. */
int insem_hip_release_cache(int64_t* bytes);

/*
 * Kernel timing with HIP events on the context's stream.  When enabled,
 * every element-force launch is bracketed by events; insem_hip_kernel_time
 * returns the summed milliseconds and launch count of the element-force
 * kernels and of the assembly/update kernels since the last reset.
 */
int insem_hip_set_timing(insem_hip_ctx* ctx, int32_t enabled);
int insem_hip_kernel_time(insem_hip_ctx* ctx, double* force_ms, int32_t* force_launches, double* update_ms,
                          int32_t* update_launches);
/* Milliseconds and launch count of the element-force launches of one block
 * accumulated by the last insem_hip_kernel_time call. */
int insem_hip_block_time(insem_hip_ctx* ctx, int32_t domain, int32_t block, double* ms, int32_t* launches);
/* Point counts of one block's force launch in a time step: owned_slots =
 * element slots that update their point inside the launch (the one slot of
 * a point with a single unit-weight slot, the first slot of a batch-local
 * point; points with no boundary, coupling, source or receiver work; 0 when
 * that fusion is off), points_read = distinct global points the launch
 * gathers.  Used to count the launch's algorithmic bytes; no counterpart in
 * the reference. */
int insem_hip_block_points(insem_hip_ctx* ctx, int32_t domain, int32_t block, int64_t* owned_slots,
                           int64_t* points_read);
/* Slots of one block's force launch in a time step whose force stays in the
 * work-group: the slots of batch-local points (all of whose slots lie in one
 * work-group's elements) other than the point's first, which sums them and
 * is counted in owned_slots.  0 when that fusion is off.  Used to count the
 * launch's algorithmic bytes; no counterpart in the reference. */
int insem_hip_block_local_slots(insem_hip_ctx* ctx, int32_t domain, int32_t block, int64_t* local_slots);

/* Device bytes held by the context. */
int insem_hip_footprint(insem_hip_ctx* ctx, int64_t* bytes);

/*
 * Batched models (no counterpart in the reference, which runs one model per
 * Solver): the setup is the disjoint union of n_models independent models
 * that share the polynomial order, coordinate system, azimuthal order,
 * media and record_every.  The union steps as one model, so each element
 * launch and tail covers every model, while each model keeps its own time
 * step and step count: every per-point computation uses its own model's
 * coefficients, so a model's traces and fields are bitwise those of running
 * it alone.
 *   point_model[d]: model of each global point of union domain d;
 *   receiver_model: model of each receiver.
 * Energy rows are not kept and the power iteration is not available on a
 * batch (each model's Δt comes from its own run of insem_hip_lambda_max or
 * from the caller).
 */
typedef struct insem_hip_batch {
  int32_t n_models;
  const int32_t* point_model[2];
  const int32_t* receiver_model;
} insem_hip_batch;

int insem_hip_create_batch(const insem_hip_setup* setup, const insem_hip_batch* batch, insem_hip_ctx** out);
/* dt[m], n_steps[m] per model; amplitudes [n_sources][max_m n_steps[m] + 1],
 * each source's values at its own model's dt.  insem_hip_advance then steps
 * the union: a model stops at its own n_steps (its last step stores a and
 * leaves u, v as a run of that length alone does) while longer ones go on;
 * traces of a model beyond its n_steps / record_every + 1 samples are 0. */
int insem_hip_set_time_batch(insem_hip_ctx* ctx, const double* dt, const int32_t* n_steps, const float* amplitudes);

#ifdef __cplusplus
}
#endif

#endif /* INSEM_HIP_H */
