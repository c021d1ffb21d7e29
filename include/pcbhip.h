/*
 * pcbhip -- C ABI of the MI355X (gfx950) matrix-free hexahedral Q1
 * conduction solver: the thermal steady-conduction and hex DC-conduction
 * operator, its Jacobi diagonal and its mixed-precision iterative-refinement
 * (MPIR) PCG solve, Jacobi or two-level preconditioned.
 *
 * Everything is extern "C" with plain pointers and sizes.  Handles are
 * opaque.  Every function returns PCBHIP_OK or an error code; the message of
 * the last error on the calling thread is pcbhip_last_error().  The caller
 * owns every host buffer for the duration of a call only; a handle owns its
 * device copies.  Functions ending in _dev take device pointers and a HIP
 * stream (NULL: the handle's own stream) and do not synchronise.
 *
 * Node layout is the reference's flat C order ((z * (rows + 1) + y) *
 * (cols + 1) + x), elements ((ez * rows + ey) * cols + ex), coefficients
 * (3, slabs, rows, cols) = (a_x, a_y, a_z) with a_x = k_in h_y h_z / h_x,
 * a_y = k_in h_x h_z / h_y, a_z = k_z h_x h_y / h_z
 * (reference: src/thermal/matrix_free_mpir_fem/mesh.py:276-293).
 */
#ifndef PCBHIP_H
#define PCBHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pcbhip_status {
    PCBHIP_OK = 0,
    PCBHIP_INVALID_INPUT = 1, /* ValueError in the Python layer */
    PCBHIP_RUNTIME = 2,       /* RuntimeError: HIP failure, not SPD */
    PCBHIP_NON_FINITE = 3,    /* FloatingPointError */
    PCBHIP_NO_MEMORY = 4,     /* MemoryError */
    PCBHIP_FULL = 5           /* a server has no free slot (the caller queues the request) */
};

typedef struct pcbhip_operator pcbhip_operator;
typedef struct pcbhip_batch pcbhip_batch;
typedef struct pcbhip_server pcbhip_server;

/* Replaces cpp/include/pcbcore/fem/mpir.hpp:11-17 (MpirConfig). */
typedef struct pcbhip_mpir_config {
    double relative_tolerance;       /* 1e-10 */
    double absolute_tolerance;       /* 0 */
    double inner_relative_tolerance; /* 2e-3 */
    int max_outer_iterations;        /* 8 (16 for the thermal solve) */
    int max_inner_iterations;        /* 200 */
} pcbhip_mpir_config;

/* Replaces cpp/include/pcbcore/fem/mpir.hpp:27-35 (MpirResult); the history
 * rows go to a separate caller buffer of 4 doubles per outer step:
 * (outer step, FP64 relative residual before it, inner iterations, inner
 * relative residual), as cpp/bindings/mpir_convert.hpp:28-44 returns them. */
typedef struct pcbhip_mpir_result {
    int converged;
    int outer_iterations;
    int inner_iterations;
    int high_operator_applications;
    int low_operator_applications;
    int not_spd;
    double relative_residual;
} pcbhip_mpir_result;

/* Message of the last failed call on this thread ("" if none). */
const char *pcbhip_last_error(void);

/* Number of visible HIP devices (replaces cpp/bindings/cuda_module.cpp:90-93,
 * device_count). */
int pcbhip_device_count(int *count);

/* A prepared operator resident on ``device``: replaces the reference's
 * ThermalSystem construction (cpp/src/thermal/thermal_system.cpp:252-363)
 * followed by the device upload of DeviceSystem::thermal
 * (cpp/bindings/cuda_module.cpp:52-62, cpp/cuda/inner_pcg.cu:619-640).
 *   coef        (3, slabs, rows, cols) FP64 element coefficients
 *   robin       (nodes) FP64 lumped Robin conductance plus C/dt
 *   free_nodes  (nodes) 1 free, 0 fixed (fixed rows are the identity)
 * The Jacobi diagonal is formed on the device; a free node without positive
 * conductance is PCBHIP_INVALID_INPUT, and so is a node plane of 2^25 nodes
 * or more. */
int pcbhip_operator_create(int device, int slabs, int rows, int cols, const double *coef, const double *robin,
                           const uint8_t *free_nodes, pcbhip_operator **out);
/* The layered-sheet DC operator of copper layers joined by resistive vias;
 * replaces MatrixFreePCBOperator's arrays and LayeredDCSystem
 * (src/electrical/matrix_free_mpir_fem/pcb.py:297-418, 575-624;
 * cpp/include/pcbcore/fem/layered_dc_system.hpp:66-172).  Nodes are
 * (layers, rows + 1, cols + 1) in flat C order; coef is (2, layers, rows,
 * cols) = (c_x, c_y) with c_x = sigma t h_y / h_x, c_y = sigma t h_x / h_y
 * (pcb.py:173-186); free_nodes as for the hexahedral operator; via k joins
 * flat nodes via_a[k] and via_b[k] with conductance via_g[k] (pcb.py:40-64).
 * The returned handle works with every apply, solve, batch and server call.
 * An operator of 2^25 nodes or more is PCBHIP_INVALID_INPUT. */
int pcbhip_sheet_operator_create(int device, int layers, int rows, int cols, const double *coef,
                                 const uint8_t *free_nodes, int64_t via_count, const int64_t *via_a,
                                 const int64_t *via_b, const double *via_g, pcbhip_operator **out);
int pcbhip_operator_destroy(pcbhip_operator *op);
/* Node count of the operator. */
int64_t pcbhip_operator_size(const pcbhip_operator *op);
/* FP64 Jacobi diagonal (replaces operator.py:289-306 / thermal_system.cpp:329-349). */
int pcbhip_operator_diagonal(const pcbhip_operator *op, double *diag);
/* Two-level preconditioner M^-1 = D^-1 + Z C^-1 Z^T (replaces the coarse data
 * of AggregationCoarseCorrection, src/electrical/matrix_free_mpir_fem/
 * two_level.py:54-229, which the native solve takes as block and cinv32,
 * thermal_hex.hpp:35-39 and hex_q1.cpp:388-432).  Z is constant over
 * block x block node patches of one node layer (node planes of a hexahedral
 * operator, copper layers of a sheet operator), zero on fixed nodes; cinv is
 * the nc x nc row-major inverse of Z^T A Z, nc = layers * ceil((R + 1) /
 * block) * ceil((C + 1) / block) <= 16384, stored in FP32.  nc = 0 removes
 * it.  Batches made and server submissions made afterwards run two-level
 * PCG on this operator; slots made before keep what they were made with.
 * Not to be called while another thread uses the same operator. */
int pcbhip_operator_set_coarse(pcbhip_operator *op, int block, int nc, const double *cinv);
/* The same from the coarse matrix itself: Z^T A Z (nc x nc FP64, symmetric)
 * is factored and inverted on the device (Cholesky and its inverse, the
 * LLT inverse of hex_q1.cpp:819-834 and two_level.py:108-122); a matrix that
 * is not positive definite is PCBHIP_INVALID_INPUT.  inverse (NULL or nc x nc
 * doubles) receives the FP64 inverse. */
int pcbhip_operator_set_coarse_matrix(pcbhip_operator *op, int block, int nc, const double *matrix,
                                      double *inverse);
/* out[k] = Z^T A Z v[k] in FP64 for count coarse vectors of nc values each
 * (row-major count x nc), Z the free-node patch indicator of block x block
 * patches as above: the Galerkin assembly's applications on the device
 * (replaces the coloured apply loop of two_level.py:147-195 and
 * hex_q1.cpp:703-818, which run the FP64 action on host vectors). */
int pcbhip_operator_galerkin_f64(const pcbhip_operator *op, int block, int count, const double *v, double *out);
/* z = M^-1 r in FP32 with the operator's preconditioner, Jacobi or two-level
 * (replaces MatrixFreeThermalOperator.precondition_low, operator.py:366-371). */
int pcbhip_operator_precondition_f32(const pcbhip_operator *op, const float *r, float *z);

/* The same z = M^-1 r as pcbhip_operator_precondition_f32, on device vectors
 * r and z of pcbhip_operator_size() floats, in stream order on ``stream``
 * (NULL: the operator's stream), allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops and harnesses that forbid
 * allocation).  The two-level coarse vectors live in a static device array
 * of 2 x 16384 doubles (256 KiB) that each device holds until the process
 * ends; a coarse space of more than 16384 unknowns is PCBHIP_INVALID_INPUT.
 * Every call on one device shares that array: calls that overlap on one
 * device (different streams or host threads) race; calls in one stream's
 * order do not. */
int pcbhip_operator_precondition_f32_noalloc(const pcbhip_operator *op, const float *r, float *z, void *stream);

/* y = A x with fixed rows the identity and Robin on the diagonal, FP32
 * (replaces hex_q1.cpp:131-193 apply_lines<float>, pcbcore::fem::thermal_hex::
 * apply_low, thermal_hex.hpp:33) and FP64 (apply_high, thermal_hex.hpp:32). */
int pcbhip_apply_f32(const pcbhip_operator *op, const float *x, float *y);
int pcbhip_apply_f64(const pcbhip_operator *op, const double *x, double *y);
/* y = K x, the unconstrained stiffness with no Robin term
 * (replaces ThermalSystem::stiffness, thermal_system.cpp:381-391). */
int pcbhip_stiffness_f64(const pcbhip_operator *op, const double *x, double *y);
int pcbhip_apply_f32_dev(const pcbhip_operator *op, const float *x, float *y, void *stream);
int pcbhip_apply_f64_dev(const pcbhip_operator *op, const double *x, double *y, void *stream);
int pcbhip_stiffness_f64_dev(const pcbhip_operator *op, const double *x, double *y, void *stream);

/* One MPIR solve: replaces thermal_hex::solve_mpir (hex_q1.cpp:837-925,
 * thermal_hex.hpp:37-39) and DeviceSystem.solve (cuda_module.cpp:64-86).
 * x holds the initial guess on entry (zeros for none) and the solution on
 * return.  inner_fp64 = 0 is the FP32 inner Jacobi PCG; 1 runs the inner PCG
 * in FP64 (a plain FP64 Jacobi PCG when inner_relative_tolerance equals the
 * outer tolerance).  history: NULL or 4 * max_outer_iterations doubles.
 * A non-SPD operator returns PCBHIP_RUNTIME with result->not_spd = 1. */
int pcbhip_solve_mpir(const pcbhip_operator *op, const double *rhs, double *x, const pcbhip_mpir_config *config,
                      int inner_fp64, pcbhip_mpir_result *result, double *history);

/* Many right-hand sides of possibly different operators (of possibly
 * different sizes) solved together: every kernel launch covers every slot
 * still iterating and each slot's arithmetic does not depend on the others.
 * No reference counterpart beyond the per-solve contract above: the
 * reference solves these one at a time (cpp/src/fem/dc_ports.cpp:26-49).
 * ops[i] may repeat (many right-hand sides of one operator). */
int pcbhip_batch_create(int count, const pcbhip_operator *const *ops, int inner_fp64, pcbhip_batch **out);
int pcbhip_batch_destroy(pcbhip_batch *batch);
/* Host rhs and initial guess (NULL: zero) and the slot's limits. */
int pcbhip_batch_set(pcbhip_batch *batch, int slot, const double *rhs, const double *x0,
                     const pcbhip_mpir_config *config);
/* Tile shape of the batch's kernels: node rows per wave strip of the FP32
 * inner stencil (4 or 8) and the number of stencil workgroups a launch should
 * reach before each board's z sweep is split into blocks.  The autotuner
 * times candidates through this call.  No reference counterpart (the
 * reference's CPU team splits node lines statically, hex_q1.cpp:195-205). */
int pcbhip_batch_set_shape(pcbhip_batch *batch, int rows_inner, int target_groups);
/* Schedule of pcbhip_batch_run: MPIR steps per captured replay (1..1024,
 * default 64; the host checks for finished slots between replays), inner
 * iterations between outer steps (default 32) and the share of the tiled
 * slots, in percent, that must have ended before their tiles leave the
 * launches (default 25).  Results do not depend on it.  No reference
 * counterpart. */
int pcbhip_batch_set_schedule(pcbhip_batch *batch, int replay_steps, int outer_every, int compact_percent);
/* Tile shape of the standalone applies: rows per wave strip (FP32: 4 or 8,
 * FP64: 2 or 4) and node planes per z block. */
int pcbhip_operator_set_shape(pcbhip_operator *op, int rows32, int zb32, int rows64, int zb64);
/* Element action of the standalone FP32 apply: 0 = FP32 vector arithmetic;
 * 2 or 3 = matrix cores: the 24 x 8 stack of unit element matrices (exact
 * in bf16) against the corner values of 32 elements per product, the values
 * split into two bf16 terms (hi + lo, 16 significant bits) or three (exact
 * for FP32), FP32 accumulation.  No reference counterpart (the reference's
 * action is the scalar gather, hex_q1.cpp:90-193). */
int pcbhip_operator_set_engine(pcbhip_operator *op, int engine);
/* Multi-right-hand-side block apply: slots of the batch that share an
 * operator are applied in groups of at most group_size (1..64; 0 turns it
 * off), one tile of a group per workgroup, so the group reads the operator's
 * coefficients, Robin terms and diagonal once; engine is the element action
 * as in pcbhip_operator_set_engine.  Each slot keeps its own scalars and FP64 refinement.  The
 * reference solves such right-hand sides one at a time
 * (cpp/src/fem/dc_ports.cpp:26-49).  FP32 inner solves only. */
int pcbhip_batch_set_block(pcbhip_batch *batch, int group_size, int engine);
/* Device pointers to a slot's FP64 right-hand side (write it before run). */
int pcbhip_batch_rhs_dev(pcbhip_batch *batch, int slot, double **rhs_dev);
/* Reset every slot to its initial guess and iterate until every slot has
 * converged, reached its outer limit or failed. */
int pcbhip_batch_run(pcbhip_batch *batch);
/* Solution (host, may be NULL), result and history (NULL or 4 * max_outer). */
int pcbhip_batch_result(pcbhip_batch *batch, int slot, double *x, pcbhip_mpir_result *result, double *history);
/* Average device time, in microseconds, of one fused Jacobi-PCG pass over
 * every slot of the batch at its current tile shape: `launches` passes back
 * to back between two HIP events on the batch's stream (alpha = beta = 0, so
 * the vectors stay bounded).  Slot states and results are kept; call it
 * after pcbhip_batch_run.  A measurement aid with no reference counterpart. */
int pcbhip_batch_time_pass(pcbhip_batch *batch, int launches, double *avg_us);

/* Device bytes an operator handle holds (coefficients, Robin terms and
 * Jacobi inverses in both precisions): the admission footprint of a served
 * request's operator. */
int pcbhip_operator_bytes(const pcbhip_operator *op, int64_t *bytes);

/* Continuous batching of MPIR solves (the "serving submit/poll" export of the
 * drop-in boundary; the reference has no serving loop -- its callers solve one
 * right-hand side at a time, src/multiphysics/staggered_coupling/
 * electro_thermal.py:403-423, cpp/src/fem/dc_ports.cpp:26-49).
 *
 * A server owns up to max_slots slots.  submit() uploads a right-hand side of
 * a resident operator into a free slot (its vectors come from a pool of
 * released blocks); step() retires the oldest replay in flight, admits the
 * submitted slots into the tile tables and launches the next replay of a
 * captured graph of steps_per_replay MPIR steps, keeping at most two replays
 * in flight, so admission and retirement happen between replays while the
 * GPU keeps iterating.  finished() lists slots whose solve ended and
 * collect() copies one result out and frees its slot.  Each slot's
 * arithmetic is that of pcbhip_solve_mpir: a result does not depend on the
 * requests it shared the device with.  Calls on one server are not
 * thread-safe; distinct servers are independent. */
typedef struct pcbhip_server_info {
    int queued;         /* submitted, waiting for the next step */
    int running;
    int finished;       /* ended, not yet collected */
    int free_slots;
    int64_t bytes_in_use;  /* slot vectors of queued, running and finished slots */
    int64_t bytes_pooled;  /* released blocks kept for reuse */
    int64_t replays;
    int64_t graph_builds;
    int steps_per_replay;
    double wait_ms;        /* host time step() spent blocked on a replay in flight */
    double replay_ms;      /* device time of the retired replays (event to event) */
} pcbhip_server_info;

int pcbhip_server_create(int device, int max_slots, int inner_fp64, int steps_per_replay, pcbhip_server **out);
int pcbhip_server_destroy(pcbhip_server *server);
/* Inner stencil rows (4 or 8) and target workgroups, as pcbhip_batch_set_shape. */
int pcbhip_server_set_shape(pcbhip_server *server, int rows_inner, int target_groups);
/* Device bytes one slot of this operator takes (admission footprint). */
int pcbhip_server_slot_bytes(const pcbhip_server *server, const pcbhip_operator *op, int max_outer, int64_t *bytes);
/* Queue one solve; PCBHIP_FULL when every slot is in use.  The operator must
 * stay alive until the slot is collected.  x0: NULL for zero. */
int pcbhip_server_submit(pcbhip_server *server, const pcbhip_operator *op, const double *rhs, const double *x0,
                         const pcbhip_mpir_config *config, int64_t tag, int *slot);
/* One replay boundary (see above); running/finished may be NULL. */
int pcbhip_server_step(pcbhip_server *server, int *running, int *finished);
/* Up to capacity finished slots and their tags. */
int pcbhip_server_finished(const pcbhip_server *server, int *slots, int64_t *tags, int capacity, int *count);
/* Solution (NULL to skip), result and history (NULL or 4 * max_outer) of a
 * finished slot; frees the slot.  A non-SPD solve returns PCBHIP_RUNTIME
 * after filling result. */
int pcbhip_server_collect(pcbhip_server *server, int slot, double *x, pcbhip_mpir_result *result, double *history);
int pcbhip_server_stats(const pcbhip_server *server, pcbhip_server_info *info);
/* Free the pooled blocks. */
int pcbhip_server_trim(pcbhip_server *server);

/* Bandwidth reference: copy n_floats (multiple of 4) device floats. */
int pcbhip_copy_dev(const float *src, float *dst, int64_t n_floats, void *stream);
/* Bandwidth reference for kernels that stream many arrays at once: reads
 * (1..16) device arrays and writes (0..16) device arrays of n_floats
 * (multiple of 4) each, one 16-byte element of every array per thread,
 * the inputs summed into every output.  No counterpart in the reference:
 * the ceiling against which the fused PCG pass's bytes are judged. */
int pcbhip_streams_dev(const float *const *in, int reads, float *const *out, int writes, int64_t n_floats,
                       void *stream);

/* FP32 vectors of the inner-solve runtime on one device: the vector
 * operations of the LowPrecisionRuntime contract
 * (src/electrical/matrix_free_mpir_fem/runtime.py:20-44; the device runtime
 * CupyFloat32Runtime, runtime.py:121-170) for callers that drive an inner
 * PCG loop from the host.  Each call returns when its work is done (it runs
 * on a stream of the vector runtime's own and waits for it).  n is the
 * length in floats; n = 0 is a no-op.  Outputs must not overlap inputs.
 *   create   a zeroed vector of n floats (at least one is allocated);
 *   destroy  release one (NULL is a no-op);
 *   write    host -> device, read  device -> host, copy  device -> device;
 *   dot      sum of x[i] * y[i] accumulated in FP64 in an order fixed by n
 *            (the same bits on every call; the reference accumulates in FP32);
 *   axpy     out = y + alpha * x with FP32 rounding of the product and the sum,
 *            as runtime.py:74-82 computes it;
 *   divide   out = num / den, correctly rounded (runtime.py:84-90). */
int pcbhip_vec_create(int device, int64_t n, float **v);
int pcbhip_vec_destroy(int device, float *v);
int pcbhip_vec_write(int device, float *dst, const float *src, int64_t n);
int pcbhip_vec_read(int device, const float *src, float *dst, int64_t n);
int pcbhip_vec_copy(int device, const float *src, float *dst, int64_t n);
int pcbhip_vec_dot(int device, const float *x, const float *y, int64_t n, double *out);
int pcbhip_vec_axpy(int device, float alpha, const float *x, const float *y, float *out, int64_t n);
int pcbhip_vec_divide(int device, const float *num, const float *den, float *out, int64_t n);
/* Gram-Schmidt products of the restarted GMRES inner solve on basis rows
 * held as device vectors (replaces the device basis products of
 * src/electrical/matrix_free_mpir_fem/solver.py:400-577, which project a
 * candidate on the stored basis with GEMV, and the dot / axpy loop of the
 * host GMRES, solver.py:235-397).  count is 1..64, every row has n floats,
 * and each call returns when done:
 *   project  out[k] = rows[k] . w for k < count (host array of count
 *            doubles), each in FP64 with the bits pcbhip_vec_dot gives;
 *   combine  out = y + sum_k coef[k] rows[k] (coef a host array of count
 *            finite doubles; y NULL: zero), accumulated in FP64 in row order
 *            and rounded once to FP32.  out may be y or one of the rows. */
int pcbhip_vec_project(int device, const float *const *rows, int count, const float *w, int64_t n, double *out);
int pcbhip_vec_combine(int device, const float *const *rows, const double *coef, int count, const float *y,
                       float *out, int64_t n);
/* z = M^-1 r in FP32 on device vectors of pcbhip_operator_size() floats,
 * returning when done (replaces MatrixFreeThermalOperator.precondition_low on
 * runtime arrays, operator.py:366-371). */
int pcbhip_operator_precondition_f32_dev(const pcbhip_operator *op, const float *r, float *z);
/* Wait for the work queued on the operator's own stream (pcbhip_apply_f32_dev
 * with a NULL stream runs there). */
int pcbhip_operator_synchronize(const pcbhip_operator *op);

#ifdef __cplusplus
}
#endif
#endif
