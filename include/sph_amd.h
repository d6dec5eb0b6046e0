/*
 * sph_amd.h -- C ABI of the MI355X (gfx950) SPH step library, libsph_amd.so.
 *
 * The reference (iwoplaza/cuda-sph) has no FFI: its step is a Python template method,
 * AbstractSPHStrategy.compute_next_state (sim/src/sph/strategies/abstract_sph_strategy.py:31-46),
 * whose stages launch numba kernels.  The entry points below are what a binding of that
 * path needs: the strategy constructor becomes sph_create + sph_bind_workspace, the
 * per-step host->device copies become sph_upload_state, the stage launches become
 * sph_step (or sph_run_stage one stage at a time) and the device->host copies become
 * sph_download_state.  Each declaration cites the reference code it replaces.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative SPH_E_* code; sph_last_error
 *     returns the message of the last failure on the calling thread.
 *   - Pointers named dev_* are device pointers (for example torch tensor data_ptr()).
 *     Pointers without the prefix are host pointers.
 *   - stream is a hipStream_t passed as void*; NULL means the default stream.  All work
 *     is enqueued on that stream; the calls that wait for it are sph_bind_workspace,
 *     sph_set_batch_lanes, sph_check_finite, sph_step_timed and sph_step on a context
 *     created with SPH_FLAG_CHECK_FINITE.
 *   - The library never allocates device memory: the caller allocates the workspace
 *     (sph_workspace_bytes) and keeps it alive while the context is used.
 *   - A context is not re-entrant: calls on one context must not overlap.
 */
#ifndef SPH_AMD_H
#define SPH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_ABI_VERSION 1

#define SPH_OK 0
#define SPH_E_INVALID (-1)   /* bad argument or size                    */
#define SPH_E_STATE (-2)     /* call out of order (no workspace, no state) */
#define SPH_E_HIP (-3)       /* HIP runtime error                        */
#define SPH_E_NONFINITE (-4) /* a particle position is NaN or infinite   */

/* Physics constants.  In the reference these are module globals frozen into the numba
 * kernels at JIT time (config.py:18-23, derived forms config.py:24-29). */
typedef struct {
    double mass; /* MASS  */
    double rho0; /* RHO_0 */
    double h;    /* INF_R, the smoothing radius */
    double visc; /* VISC  */
    double k;    /* K, gas constant of p = K (rho - rho0) */
    double damp; /* DAMP, box-wall velocity damping */
} sph_constants_t;

#define SPH_MODE_BOX 0  /* config.py:13 SIM_MODE == 'BOX'  -> collision_kernel_box */
#define SPH_MODE_PIPE 1 /* config.py:13 SIM_MODE == 'PIPE' -> collision_kernel     */

#define SPH_FORM_NAIVE 0 /* config.py:14 'NAIVE' (naive_kernels.py) */
#define SPH_FORM_VOXEL 1 /* config.py:14 'VOXEL' (voxel_kernels.py) */

/* SimulationParameters (common/data_classes.py:70-78) plus the reference globals the
 * step reads.  dt = 1 / fps (abstract_sph_strategy.py:20). */
typedef struct {
    int64_t particle_count;   /* data_classes.py:72 */
    double external_force[3]; /* data_classes.py:73, a force density (divided by rho) */
    double dt;                /* 1 / fps */
    double space_size[3];     /* data_classes.py:77 */
    double voxel_size[3];     /* data_classes.py:78 */
    int32_t mode;             /* SPH_MODE_* */
    int32_t formulation;      /* SPH_FORM_* */
    int32_t pipe_rows;        /* rows of Pipe.to_numpy() (data_classes.py:34-44); 0 in BOX mode */
    int32_t flags;            /* SPH_FLAG_* */
    const double* pipe;       /* host, pipe_rows x 5 row-major [x, y, z, start_radius, length] */
    uint64_t seed;            /* collision RNG seed, abstract_sph_strategy.py:27 uses 16435234 */
} sph_params_t;

/* Keep the per-particle pressure term, viscosity term and result force of the last step
 * (readable with sph_read_buffer).  Off by default: they cost 48 bytes per particle. */
#define SPH_FLAG_KEEP_TERMS 1
/* Allocate the adjoint buffers (sph_adjoint_*): 64 bytes per particle more.  Plain
 * NAIVE-formulation contexts in BOX mode only. */
#define SPH_FLAG_ADJOINT 2
/* sph_step waits for its steps and fails with SPH_E_NONFINITE when a position became NaN
 * or infinite (it calls sph_check_finite; skipped while the stream is being captured into
 * a graph).  The reference has no such check: its state silently turns non-finite, e.g.
 * the VOXEL formulation's density leaves the particle itself out (voxel_kernels.py:127-130),
 * so an isolated particle has rho = 0 and base_kernels.py:47-52 divides by it. */
#define SPH_FLAG_CHECK_FINITE 4

typedef struct sph_ctx sph_ctx;

/* Replaces AbstractSPHStrategy.__init__ (abstract_sph_strategy.py:18-29) together with
 * thread_layout.organize (thread_layout.py:27-33): validates the parameters, plans the
 * voxel grid and launch shapes and, in PIPE mode, seeds one xoroshiro128+ stream per
 * particle exactly as create_xoroshiro128p_states(n, seed) does.  Host only. */
int sph_create(const sph_params_t* params, const sph_constants_t* consts, sph_ctx** out);

/* Batched form: count independent simulations (each with its own particle count, voxel
 * grid, h, mu, box and dt) advanced together, one launch per stage for all of them.
 * params[b] / consts[b] describe instance b; all instances must use SPH_FORM_NAIVE and
 * SPH_MODE_BOX.  State arrays of the other entry points are the instances' arrays
 * concatenated in order (instance b starts at particle first[b], see sph_batch_info).
 * The reference has no counterpart: it advances one simulation per process
 * (state_generator.py:14-37); this replaces running compute_next_state once per
 * simulation. */
int sph_create_batch(const sph_params_t* params, const sph_constants_t* consts, int32_t count, sph_ctx** out);

/* Instance count and the count + 1 particle offsets of a context (1 and {0, N} for a plain
 * step).  first may be NULL.  Host only. */
int sph_batch_info(const sph_ctx* ctx, int32_t* count, int32_t* first);

/* Releases the host context.  The workspace belongs to the caller. */
int sph_destroy(sph_ctx* ctx);

/* Device bytes the context needs (state, sort buffers, cell table, RNG streams). */
size_t sph_workspace_bytes(const sph_ctx* ctx);

/* Binds caller-owned device memory of at least sph_workspace_bytes bytes (256-byte
 * aligned) and uploads the pipe table and RNG streams.  Replaces the per-step
 * allocations of _send_arrays_to_gpu (abstract_sph_strategy.py:64-74): after this call
 * nothing is allocated per step. */
int sph_bind_workspace(sph_ctx* ctx, void* dev_workspace, size_t bytes, void* stream);

/* Loads a SimulationState (data_classes.py:81-85): dev_pos and dev_vel are float64
 * (N, 3) row-major in particle order, as cuda.to_device(old_state.position) holds them
 * (abstract_sph_strategy.py:65-66).  The state is converted to the internal FP32
 * cell-sorted layout and stays on the device across steps. */
int sph_upload_state(sph_ctx* ctx, const double* dev_pos, const double* dev_vel, void* stream);

/* Same as sph_upload_state from float32 structure-of-arrays inputs (x, y, z, vx, vy, vz
 * each of N floats). */
int sph_upload_state_f32(sph_ctx* ctx, const float* dev_x, const float* dev_y, const float* dev_z,
                         const float* dev_vx, const float* dev_vy, const float* dev_vz, void* stream);

/* Advances the resident state by nsteps steps.  One step is the body of
 * compute_next_state (abstract_sph_strategy.py:31-46): voxel hash, sort, cell table,
 * density, pressure and viscosity terms, integration (base_kernels.py:30-53) and
 * collision (base_kernels.py:56-98).  Returns SPH_E_STATE before the first upload. */
int sph_step(sph_ctx* ctx, int32_t nsteps, void* stream);

/* sph_step(ctx, nsteps) with HIP events recorded on the stream between its stages; waits
 * for them and writes to stage_ms[0..2] the mean milliseconds per step of the sort (with
 * the gather into cell order and the cell table), the density pass and the force pass
 * (1 <= nsteps <= 100000; not on a stream under graph capture).  No reference
 * counterpart: it is for measurement, the kernels and results are sph_step's. */
int sph_step_timed(sph_ctx* ctx, int32_t nsteps, double* stage_ms, void* stream);

/* Waits for the stream and reads the number of particle positions found NaN or infinite
 * when hashed since the last read (every step hashes every particle once, as does an upload),
 * then clears it.  *count (may be NULL) receives the number; returns SPH_E_NONFINITE if it is
 * not zero.  No reference counterpart: the reference returns a non-finite state without
 * notice (abstract_sph_strategy.py:76-84). */
int sph_check_finite(sph_ctx* ctx, int64_t* count, void* stream);

/* Replaces __finalize_computation (abstract_sph_strategy.py:76-84): writes the state in
 * particle order as float64 (N, 3) / (N,), i.e. the new SimulationState's position,
 * velocity and density, and the result_force of the last step.  Any pointer may be
 * NULL; dev_force requires SPH_FLAG_KEEP_TERMS. */
int sph_download_state(sph_ctx* ctx, double* dev_pos, double* dev_vel, double* dev_rho, double* dev_force,
                       void* stream);

/* Same as sph_download_state into float32 structure-of-arrays outputs (any may be NULL). */
int sph_download_state_f32(sph_ctx* ctx, float* dev_x, float* dev_y, float* dev_z, float* dev_vx, float* dev_vy,
                           float* dev_vz, float* dev_rho, void* stream);

/* Stage-by-stage execution of one step, for parity tests and profiling.  The stages in
 * order; sph_step(ctx, 1) runs exactly this sequence.
 *   SPH_STAGE_SORT     voxel_sph_strategy.py:81-89 (sort of the voxel map, ties by id)
 *   SPH_STAGE_TABLE    voxel_sph_strategy.py:92-107 (cell start table)
 *   SPH_STAGE_REORDER  gather of the state into cell order (no reference counterpart)
 *   SPH_STAGE_DENSITY  naive_kernels.py:7-21 / voxel_kernels.py:108-132
 *   SPH_STAGE_FORCE    naive_kernels.py:24-80 / voxel_kernels.py:135-211, then
 *                      base_kernels.py:30-53 and :56-98, then the voxel hash of the
 *                      new positions for the next step (voxel_kernels.py:88-105)
 * The hash of the uploaded state is part of sph_upload_state. */
#define SPH_STAGE_SORT 1
#define SPH_STAGE_TABLE 2
#define SPH_STAGE_REORDER 3
#define SPH_STAGE_DENSITY 4
#define SPH_STAGE_FORCE 5
int sph_run_stage(sph_ctx* ctx, int32_t stage, void* stream);

/* Copies an internal buffer to dev_out.  Buffers in particle order unless noted:
 *   SPH_BUF_KEYS        int32[N]   voxel key of each particle (voxel_kernels.py:88-105)
 *   SPH_BUF_PERM        int32[N]   particle id at each sorted position (voxel_particle_map
 *                                  ["particle_id"] after the sort, voxel_sph_strategy.py:88)
 *   SPH_BUF_SORTED_KEYS int32[N]   key at each sorted position (voxel_particle_map["voxel_id"])
 *   SPH_BUF_CELL_BEGIN  int32[C]   voxel_begin (voxel_sph_strategy.py:92-107): first sorted
 *                                  position of each voxel, -1 if empty; C = product of
 *                                  ceil(space/voxel)
 *   SPH_BUF_DENSITY     float32[N] density of the current step
 *   SPH_BUF_PRESSURE    float32[N*3] pressure term  (needs SPH_FLAG_KEEP_TERMS)
 *   SPH_BUF_VISCOSITY   float32[N*3] viscosity term (needs SPH_FLAG_KEEP_TERMS)
 *   SPH_BUF_FORCE       float32[N*3] result_force   (needs SPH_FLAG_KEEP_TERMS)
 *   SPH_BUF_RNG         uint64[N*2] xoroshiro128+ stream of each particle (s0, s1)
 *   SPH_BUF_NEIGHBOURS  int32[N*33] VOXEL formulation: count then up to 32 particle ids
 *                                  (get_neighbours, voxel_kernels.py:29-85)
 * sph_buffer_elems gives the element count of a buffer. */
#define SPH_BUF_KEYS 1
#define SPH_BUF_PERM 2
#define SPH_BUF_SORTED_KEYS 3
#define SPH_BUF_CELL_BEGIN 4
#define SPH_BUF_DENSITY 5
#define SPH_BUF_PRESSURE 6
#define SPH_BUF_VISCOSITY 7
#define SPH_BUF_FORCE 8
#define SPH_BUF_RNG 9
#define SPH_BUF_NEIGHBOURS 10
int64_t sph_buffer_elems(const sph_ctx* ctx, int32_t which);
int sph_read_buffer(sph_ctx* ctx, int32_t which, void* dev_out, void* stream);

/* FP64 global diagnostics of the resident state (no reference counterpart; the reference
 * only logs particle 0, abstract_sph_strategy.py:120-125).  dev_out8 receives eight
 * doubles: total mass, momentum x/y/z, kinetic energy, count of non-finite particles,
 * maximum speed, mean density. */
int sph_diagnostics(sph_ctx* ctx, double* dev_out8, void* stream);

/* Grid plan of the context: dims[3] (voxels per axis used for keys), stencil[3] (cells
 * searched on each side), key_bits and sort passes.  Host only. */
int sph_grid_info(const sph_ctx* ctx, int32_t* dims3, int32_t* stencil3, int32_t* key_bits, int32_t* passes);

/* Launch shapes of the neighbour kernels (threads per block), for the autotuner. */
int sph_set_launch(sph_ctx* ctx, int32_t density_block, int32_t force_block);

/* Lanes of a wavefront that share one particle in the NAIVE density and force passes
 * (1, 2, 4, 8, 16, 32 or 64; default 1), for the autotuner: small or dense problems (few particles,
 * many neighbours each) get that many times the parallelism.  The neighbour lists are the
 * same for every value; sums are added in a different order (FP32 rounding only).  The
 * reference launches one thread per particle (thread_layout.py:27-33).  On a batched
 * context every instance gets `lanes`. */
int sph_set_lanes(sph_ctx* ctx, int32_t lanes);

/* NAIVE formulation, plain context without SPH_FLAG_ADJOINT: whether the fused step cuts
 * each cell into four x slabs (enable = 1, the default), so that a row of cells is searched
 * slab by slab, or searches whole cells (enable = 0), for the autotuner: cells of many
 * particles gain little from the slabs and pay for ordering them.  The pairs found are the
 * same either way; sums are added in a different order (FP32 rounding only).  With a state
 * resident, its keys are rewritten on `stream` for the next sort, and the last step's cell
 * table can no longer be read.  HIP graphs of sph_step recorded before the call keep the
 * previous layout in their kernel arguments: record them again.  No counterpart in the
 * reference (an all-pairs search). */
int sph_set_slabs(sph_ctx* ctx, int32_t enable, void* stream);

/* Batched context: lanes per particle of each instance (count = the number of instances, each
 * 1, 2, 4, 8, 16, 32 or 64), so that instances with many neighbours per particle get more lanes and
 * sparse ones fewer.  Each instance computes what a plain context with its lane count
 * computes.  The instance table is re-uploaded on `stream` (synchronised) when the workspace
 * is bound.  No counterpart in the reference, which advances one simulation per process. */
int sph_set_batch_lanes(sph_ctx* ctx, const int32_t* lanes, int32_t count, void* stream);

/* Replaces the physics constants of a plain context between steps (the reference freezes
 * them at JIT time, config.py:18-23).  mass, rho0, visc, k and damp may change; h must be
 * the one the context was created with (it fixes the voxel grid).  Host only. */
int sph_set_constants(sph_ctx* ctx, const sph_constants_t* consts);

/* ---------------------------------------------------------------------------------------
 * Reverse mode (no reference counterpart: the reference has no gradients).  For a context
 * created with SPH_FLAG_ADJOINT the library holds one cotangent (gx, gv) of a state, in
 * particle order, and pulls it back through steps one at a time:
 *   sph_adjoint_seed   sets the cotangent of the state the rollout ended in (dL/dx_T,
 *                      dL/dv_T) and clears the viscosity gradient
 *   sph_adjoint_step   with the state x_t resident (uploaded, e.g. from a checkpoint),
 *                      advances it to x_{t+1} exactly as sph_step does and, in the same
 *                      pass, maps the cotangent of x_{t+1} to that of x_t and adds the
 *                      step's dL/dmu
 *   sph_adjoint_read   the current cotangent and the accumulated dL/dmu
 * A rollout of T steps is differentiated by visiting t = T-1 .. 0 with x_t re-uploaded
 * before each call (states recomputed from checkpoints).  The pullback is the exact
 * derivative of the FP32 step, wall bounces included (a bounced component's position is
 * constant and its velocity is scaled by -damp).
 * --------------------------------------------------------------------------------------- */
/* dev_gx, dev_gv: float32 (N, 3) row-major in particle order; NULL means zero. */
int sph_adjoint_seed(sph_ctx* ctx, const float* dev_gx, const float* dev_gv, void* stream);
int sph_adjoint_step(sph_ctx* ctx, void* stream);
/* dev_gx, dev_gv: float32 (N, 3) row-major in particle order; dev_gmu: one float32.  Any
 * pointer may be NULL. */
int sph_adjoint_read(sph_ctx* ctx, float* dev_gx, float* dev_gv, float* dev_gmu, void* stream);

/* Step snapshots (no reference counterpart): what sph_adjoint_step recomputes before its
 * pullback -- the step's cell-ordered state, densities, neighbour lists, cell table and
 * force sums --
 * saved right after the forward step that produced them, so that a checkpointed backward
 * can reverse each step without sorting and recomputing it a second time.
 *   sph_adjoint_snapshot_bytes  bytes of one snapshot of this context (0 if it has no
 *                               reverse-mode buffers)
 *   sph_adjoint_save            after sph_step (or sph_run_stage through SPH_STAGE_FORCE):
 *                               copies that step's snapshot to dev_snapshot
 *   sph_step_save               sph_step(ctx, 1) writing the step's sorted state,
 *                               densities, lists and cell table into dev_snapshot instead
 *                               of the context (no copy); the state it produces is the
 *                               same, but until the next step sph_read_buffer refuses the
 *                               buffers the snapshot holds and sph_adjoint_save refuses
 *   sph_adjoint_step_from       sph_adjoint_step for the step whose snapshot dev_snapshot
 *                               holds, read in place (the resident state is not read); the
 *                               cotangent and dL/dmu come out bitwise as sph_adjoint_step
 *                               gives them
 * dev_snapshot: 256-byte aligned device memory of `bytes` bytes, at least
 * sph_adjoint_snapshot_bytes (SPH_E_INVALID otherwise). */
int64_t sph_adjoint_snapshot_bytes(const sph_ctx* ctx);
int sph_adjoint_save(sph_ctx* ctx, void* dev_snapshot, size_t bytes, void* stream);
int sph_step_save(sph_ctx* ctx, void* dev_snapshot, size_t bytes, void* stream);
int sph_adjoint_step_from(sph_ctx* ctx, const void* dev_snapshot, size_t bytes, void* stream);

/* ---------------------------------------------------------------------------------------
 * The step with no device allocation.  No counterpart in the reference: this is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  sph_step_noalloc plans a plain context on the stack,
 * takes its workspace from a static device array of sph_noalloc_scratch_bytes() bytes
 * (64 MiB) that the library holds on every GPU it runs on until the process ends, and runs
 * sph_upload_state, sph_step(nsteps) and sph_download_state on it: dev_pos_in / dev_vel_in
 * and dev_pos_out / dev_vel_out are float64 (N, 3) in particle order, dev_rho_out float64
 * (N,) or NULL.  Results are bitwise those of that sequence on a context of its own.  BOX
 * mode only, without SPH_FLAG_ADJOINT or SPH_FLAG_KEEP_TERMS (SPH_FLAG_CHECK_FINITE is
 * honoured); a problem whose workspace (sph_workspace_bytes of the same parameters) exceeds
 * the array fails with SPH_E_INVALID.  Nothing persists between calls: every call on a GPU
 * shares that GPU's one array, so calls that overlap on one GPU (other streams, other
 * threads) race and must be serialised by the caller.  It calls no hipMalloc, hipFree or
 * other allocating routine.
 * --------------------------------------------------------------------------------------- */
size_t sph_noalloc_scratch_bytes(void);
int sph_step_noalloc(const sph_params_t* params, const sph_constants_t* consts, const double* dev_pos_in,
                     const double* dev_vel_in, double* dev_pos_out, double* dev_vel_out, double* dev_rho_out,
                     int32_t nsteps, void* stream);

/* Message of the last error on this thread (always NUL-terminated). */
int sph_last_error(char* buf, size_t len);

int sph_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SPH_AMD_H */
