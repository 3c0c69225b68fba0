/*
 * swfd -- MI355X (gfx950) kernels for the 2D ring-array FDTD hot path.
 *
 * C ABI of libswfd.so. Every pointer is a device pointer unless stated
 * otherwise; every field is float32, row-major with x (columns) fastest;
 * batched fields are (n_shots, ny, nx). Streams are hipStream_t passed as
 * void* (0 = the null stream). Functions return SWFD_OK or an error code;
 * swfd_last_error() gives the message of the last failure on the calling
 * thread. The library allocates no device memory: callers own every buffer.
 *
 * The reference (scjoshi/SoundWaveSimulation) has no native layer: its hot
 * path is a sequence of PyTorch ops. Each entry point below names the
 * reference function(s) whose op sequence it replaces; the Python host layer
 * (soundwavesimulation_amd.torch_backend) keeps the reference call surface
 * and calls these entry points through ctypes.
 */
#ifndef SWFD_H
#define SWFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWFD_OK 0
#define SWFD_EINVAL 1 /* bad argument (shape, null pointer, range) */
#define SWFD_EHIP 2   /* HIP runtime error */

/* Grid and scheme constants, all as the float32 values the kernels use.
 * Laplacian taps: inv_dx2 = 1/dx^2, inv_dy2 = 1/dy^2, center = -2 (inv_dx2 +
 * inv_dy2), each rounded from double (torch_backend.py:122-132); dt2 is dt*dt
 * rounded from double (torch_backend.py:242). */
typedef struct swfd_grid {
    int32_t ny, nx; /* rows (y) and columns (x); both >= 3 */
    float inv_dx2, inv_dy2, center;
    float dt2;
} swfd_grid;

/* Medium coefficients produced by swfd_prepare_medium: c^2, dt^2 c^2, and the
 * Mur coefficients k_left[ny], k_right[ny], k_bottom[nx], k_top[nx],
 * corner kx[4], corner ky[4] (corners (0,0), (0,-1), (-1,0), (-1,-1)). */
typedef struct swfd_medium {
    const float *c2;    /* (ny, nx) */
    const float *c2dt2; /* (ny, nx) */
    const float *mur;   /* 2 ny + 2 nx + 8 */
} swfd_medium;

/* Bilinear receivers (array.py:60-76) and, for adjoint injection, the
 * scatter table: cell_map[y*nx+x] = entry id or -1; entry e covers pairs
 * [entry_start[e], entry_start[e+1]) of (pair_rec, pair_w), summed in that
 * order; entry_cell[e] = y*nx+x. The table may be absent (cell_map = NULL)
 * when no adjoint injection or scatter is requested. */
typedef struct swfd_receivers {
    int32_t n;               /* R */
    const int32_t *row0;     /* (R) footprint rows row0, row0+1 */
    const int32_t *col0;     /* (R) footprint cols col0, col0+1 */
    const float *weights;    /* (4, R): w00, w10, w01, w11 */
    int32_t n_entries;       /* E */
    const int32_t *cell_map; /* (ny*nx) */
    const int32_t *entry_start; /* (E+1) */
    const int32_t *entry_cell;  /* (E) */
    const int32_t *pair_rec;    /* (P) */
    const float *pair_w;        /* (P) */
    /* optional (ny), from swfd_footprint_rows; NULL: sampled forward runs
     * store every step's field in full */
    const uint32_t *footprint_rows;
} swfd_receivers;

/* One time step for n_shots shots (torch_backend.py:177-233 single shot,
 * :493-598 batched): interior leapfrog, sources, Mur edges then corners.
 * adjoint = 0: u' = 2u - u_old + dt^2 c^2 Lap(u); adjoint = 1: Lap(c^2 u).
 * Sources, all optional and added before the boundary update:
 *   point: u'[s, src_rows[s], src_cols[s]] += src_values[src_index];
 *   dense: u' += dense[s];
 *   residual (S, R): bilinear scatter of the residual (needs the table).
 * Imaging (optional, image_field != NULL):
 *   grad_out = grad_in + image_scale * sum_s u'[s] * Lap(image_field[s]).
 * samples (optional): gather of u' after the boundary, (S, R).
 * u_next must not alias u_old or u; u_next_copy (optional) receives u' too.
 * src_weights (optional, (S, 4), no counterpart in the reference's solver; the
 * reference's suggested reciprocity fix, desmond/docs/upstream_suggestions.md:
 * 65-71): a bilinear source instead of the point source: u'[s, r + dy, c + dx]
 * += src_values[src_index] * src_weights[s][2 dy + dx] for dy, dx in {0, 1},
 * (r, c) = (src_rows[s], src_cols[s]) the transmitter's footprint corner and
 * the weights the receivers' bilinear weights (order w00, w10, w01, w11);
 * the footprint must lie inside the grid. */
typedef struct swfd_step_args {
    int32_t n_shots;
    int32_t adjoint;
    const float *u_old, *u;
    float *u_next, *u_next_copy;
    const int32_t *src_rows, *src_cols;
    const float *src_values;
    int32_t src_index;
    const float *dense;
    const float *residual;
    const float *image_field;
    const float *grad_in;
    float *grad_out;
    float image_scale;
    float *samples;
    const float *src_weights; /* optional (S, 4): bilinear source (see above) */
} swfd_step_args;

/* A run of forward steps [step_begin, step_begin + n_steps) with point
 * sources (simulate_shot torch_backend.py:443-474, simulate_shots_batch
 * :736-770, forward_store(_batch) invert_ring.py:395-455).
 *   init_old/init_u: state before step_begin (read only).
 *   work0..2: scratch fields; work3 (optional) a fourth one: with it the run
 *   advances several steps per launch (same results, fewer field reads);
 *   with samples, two more (work4, work5) let full-width runs take more than
 *   two steps per launch (the intermediate steps' fields are kept there for
 *   the gather).
 *   final_old/final_u (optional) receive the state after the last step (they
 *   may be work0/work1 or the init buffers).
 *   src_values: per global step, length >= step_begin + n_steps.
 *   samples (optional): (N_total, S, R), row n = gather after step n.
 *   hist (optional): output of step n is stored in slot (n - hist_first) /
 *   hist_every when n >= hist_first and aligned; hist_slots bounds the slot.
 *   ckpt (optional): the state before step j*ckpt_every (j >= 1) is stored as
 *   ckpt slots 2(j-1) (older) and 2(j-1)+1; ckpt_slots bounds the slot.
 *   hist_lap = 1 (no counterpart in the reference; for the adjoint's
 *   imaging): the slot of step n receives the 5-point Laplacian of the
 *   output of step n (zero outside the grid) instead of the output, for
 *   n = step_begin - 1 (the Laplacian of init_u) up to step_begin + n_steps
 *   - 2 (the last step's output gets none); needs work3, full-width launches
 *   (swfd_lap_history) and no samples or checkpoints. */
typedef struct swfd_forward_args {
    int32_t n_shots;
    int32_t step_begin, n_steps;
    const float *init_old, *init_u;
    float *work0, *work1, *work2, *work3;
    float *final_old, *final_u;
    const int32_t *src_rows, *src_cols;
    const float *src_values;
    float *samples;
    float *hist;
    int32_t hist_first, hist_every, hist_slots;
    float *ckpt;
    int32_t ckpt_every, ckpt_slots;
    int32_t hist_lap; /* 1: history slots receive Laplacians (see above) */
    float *work4, *work5; /* optional: with samples, up to 4 steps per launch */
    const float *src_weights; /* optional (S, 4): bilinear sources, as in swfd_step_args */
    /* optional boundary ring (no counterpart in the reference): the boundary
     * cells of the field after step j, for j in [bnd_first, bnd_first +
     * bnd_slots), go to slot j - bnd_first, each slot (S, swfd_boundary_floats)
     * floats per shot: row 0, row ny-1, column 0 (rows 1..ny-2), column nx-1
     * (rows 1..ny-2). Lets swfd_adjoint reconstruct the forward fields. In
     * batched runs one batched copy per tick takes the ring of every run. */
    float *bnd;
    int32_t bnd_first, bnd_slots;
} swfd_forward_args;

/* A run of adjoint steps n = step_hi - 1 down to step_lo with residual
 * injection and imaging (TorchLeastSquaresFWI._adjoint_accumulate_batch,
 * invert_ring.py:683-701; gradient.py:124-141).
 *   residual: (N_total, S, R), row n drives step n.
 *   hist: forward field after step j lives in slot (j - hist_first) /
 *   hist_every; step n images with the field after step n-1 when n >= 1,
 *   n-1 >= hist_first and (n-1-hist_first) % hist_every == 0, adding
 *   image_scale * sum_s lam'[s] * Lap(field) into grad (in place). Shot groups
 *   accumulate their partial images in `workspace` across the run, summed into
 *   grad in a fixed order at its end, so the result is deterministic. */
typedef struct swfd_adjoint_args {
    int32_t n_shots;
    int32_t step_hi, step_lo;
    const float *init_old, *init_u;
    float *work0, *work1, *work2, *work3; /* work3 optional, as for swfd_forward */
    float *final_old, *final_u;
    const float *residual;
    const float *hist;
    int32_t hist_first, hist_every, hist_slots;
    float image_scale;
    float *grad;
    float *workspace; /* swfd_adjoint_workspace_floats() floats, needed with hist */
    int32_t hist_lap; /* 1: slots hold Lap(field): the imaging adds lam' * slot */
    /* Optional, instead of hist (no counterpart in the reference: replaces its
     * stored wavefield, invert_ring.py:442-455, with a reconstruction): the
     * forward fields are rebuilt backwards in time inside the adjoint launches,
     * U(j-2) = 2 U(j-1) - U(j) + dt2 c^2 Lap(U(j-1)) + source of step j in the
     * interior and the boundary ring (see swfd_forward_args) on the boundary,
     * starting from recon_next / recon_cur = the forward fields after steps
     * step_hi-1 / step_hi-2 (read only); step n >= 1 images against the
     * reconstructed field after step n-1. recon_work0..3: four fields for the
     * reconstructed pairs; src_*: the forward run's sources. Needs full-width
     * launches (see swfd_steps_per_launch) and work3; in batched runs the
     * reconstructing launches of a tick go out as one batched launch per step
     * count and row width. */
    const float *recon_next, *recon_cur;
    float *recon_work0, *recon_work1, *recon_work2, *recon_work3;
    const float *bnd;
    int32_t bnd_first, bnd_slots;
    const int32_t *src_rows, *src_cols;
    const float *src_values;
    const float *src_weights;
} swfd_adjoint_args;

/* Floats per shot in one slot of the boundary ring of an ny x nx grid:
 * 2 nx + 2 (ny - 2) (no counterpart in the reference). */
size_t swfd_boundary_floats(int32_t ny, int32_t nx);

/* 1 when swfd_adjoint can reconstruct the forward fields on an ny x nx grid
 * (full-width launches with work3), else 0 (no counterpart in the
 * reference). */
int32_t swfd_adjoint_recon(int32_t ny, int32_t nx);

/* 1 when runs on an ny x nx grid (with work3) support Laplacian histories
 * (hist_lap), else 0 (no counterpart in the reference). */
int32_t swfd_lap_history(int32_t ny, int32_t nx);

/* Steps per launch that swfd_forward (kind 0; samples != 0: with receiver
 * samples) or swfd_adjoint (kind 1) takes on an ny x nx grid given work3,
 * under the current tuning table and environment; 0 for bad arguments (no
 * counterpart in the reference; for profiling). */
int32_t swfd_steps_per_launch(int32_t kind, int32_t ny, int32_t nx, int32_t samples);

/* Launch shape of the imaging adjoint runs of n_shots shots on an ny x nx
 * grid given work3 (no counterpart in the reference; for profiling), for
 * runs that rebuild the forward field (recon = 1) or read a history (0):
 * out[0] steps per launch, out[1] 1 for full-width launches, and for those
 * out[2] rows per segment, out[3] segments, out[4] shot groups, out[5] shots
 * per group (zeros otherwise). Host only, but the shape depends on the
 * current device's compute-unit count. */
int swfd_adjoint_launch_shape(int32_t ny, int32_t nx, int32_t n_shots, int32_t recon,
                              int32_t *out);

/* Rows and full-width wavefronts that hold receiver footprint cells
 * (host-only, no launch; host pointers; no counterpart in the reference):
 * out[y] (ny entries, overwritten) gets bit w when row y has a footprint cell
 * (row0[i] + {0,1}, col0[i] + {0,1}) among the columns wavefront w of a
 * full-width launch covers; all zero when nx > 1023 (no full-width launches).
 * A device copy in swfd_receivers.footprint_rows lets sampled forward runs
 * that keep intermediate steps only for the receiver gather (work4/work5,
 * no history slot) store those steps' fields only there; samples, states,
 * checkpoints and histories are unchanged. */
int swfd_footprint_rows(int32_t ny, int32_t nx, int32_t n, const int32_t *row0,
                        const int32_t *col0, uint32_t *out);

/* Thread-local message of the last failure. */
const char *swfd_last_error(void);

/* Library version string. */
const char *swfd_version(void);

/* Floats in the medium-coefficient buffer for an (ny, nx) grid. */
size_t swfd_medium_floats(int32_t ny, int32_t nx);

/* c (ny, nx) -> coef = [c^2 | dt2 c^2 | Mur coefficients]
 * (TorchScalarWave2D.set_c torch_backend.py:158-168, _refresh_boundaries
 * :264-275). dt, dx, dy are rounded to float32 as the reference's tensor ops
 * do. */
int swfd_prepare_medium(const swfd_grid *grid, const float *c, double dt, double dx,
                        double dy, float *coef, void *stream);

/* Fill a swfd_medium view of a coefficient buffer (host-only, no launch). */
void swfd_medium_view(const swfd_grid *grid, const float *coef, swfd_medium *out);

/* One step; see swfd_step_args (TorchFDTD2D._step torch_backend.py:300-329,
 * TorchFDTD2DBatch.step/.step_record/.adjoint_step_residual/
 * .adjoint_step_image/.adjoint_step :665-733). */
int swfd_step(const swfd_grid *grid, const swfd_medium *medium, const swfd_receivers *rec,
              const swfd_step_args *args, void *stream);

/* Bilinear gather (S, ny, nx) -> (S, R) (TorchRingSampler.record/.record_batch
 * torch_backend.py:368-387, _gather_receivers :540-548). */
int swfd_gather(const swfd_grid *grid, const swfd_receivers *rec, int32_t n_shots,
                const float *fields, float *samples, void *stream);

/* Bilinear scatter (S, R) -> (S, ny, nx), output overwritten
 * (TorchRingSampler.record_adjoint(_batch) torch_backend.py:389-440). */
int swfd_scatter(const swfd_grid *grid, const swfd_receivers *rec, int32_t n_shots,
                 const float *residual, float *fields, void *stream);

/* Zero-padded 5-point Laplacian of n fields (TorchScalarWave2D.laplacian
 * torch_backend.py:170-174). */
int swfd_laplacian(const swfd_grid *grid, int32_t n_fields, const float *fields, float *out,
                   void *stream);

/* Reduced-precision wavefield histories (store_every > 1 with wave_dtype
 * float16 / bfloat16, invert_ring.py:443-453): the reference casts each
 * stored field into wave_dtype (wave[k].copy_(fields), :450-451) and back to
 * float32 before imaging (u_fwd.to(dtype), torch_backend.py:595).
 * direction 0 packs n float32 values of src into n 16-bit codes at dst,
 * direction 1 unpacks codes into float32; type SWFD_HISTORY_F16 (IEEE
 * binary16) or SWFD_HISTORY_BF16, both rounded to nearest even. */
#define SWFD_HISTORY_F16 1
#define SWFD_HISTORY_BF16 2
int swfd_history_convert(size_t n, int32_t direction, int32_t type, const void *src, void *dst,
                         void *stream);

/* Multi-step forward run; see swfd_forward_args. */
int swfd_forward(const swfd_grid *grid, const swfd_medium *medium, const swfd_receivers *rec,
                 const swfd_forward_args *args, void *stream);

/* Multi-step adjoint run; see swfd_adjoint_args. */
int swfd_adjoint(const swfd_grid *grid, const swfd_medium *medium, const swfd_receivers *rec,
                 const swfd_adjoint_args *args, void *stream);

/* Workspace floats needed by swfd_adjoint with imaging: one ny x nx partial
 * image per shot group (at most n_shots groups and at least 8; up to 32 on
 * grids of 1024^2 cells and more, up to 2^25 / (ny nx) -- 512 at 256^2 --
 * on smaller ones). */
size_t swfd_adjoint_workspace_floats(int32_t ny, int32_t nx, int32_t n_shots);

/* Workspace floats needed by swfd_residual. */
size_t swfd_residual_workspace_floats(int32_t n_steps, int32_t n_shots, int32_t n_rec);

/* Residual and misfit (TorchLeastSquaresFWI._residual_batch and
 * 0.5 * sum(residual^2), invert_ring.py:603-609,630,651):
 *   residual[n, s, r] = samples[n, s, r] - observed[obs_index[s], r, n],
 *   set to 0 where r == mute[s] (mute[s] < 0: no muting);
 *   *misfit (device float) = 0.5 * sum residual^2, summed in a fixed order.
 * observed element (e, r, n) is at observed[e*obs_stride_e + r*obs_stride_r + n]. */
int swfd_residual(int32_t n_steps, int32_t n_shots, int32_t n_rec, const float *samples,
                  const float *observed, int64_t obs_stride_e, int64_t obs_stride_r,
                  const int32_t *obs_index, const int32_t *mute, float *residual,
                  float *workspace, float *misfit, void *stream);

/* Partial sums the built-in scratch of swfd_residual_noalloc holds (4 MiB). */
#define SWFD_RESIDUAL_SCRATCH_FLOATS (1u << 20)

/* swfd_residual without a workspace argument: no counterpart in the reference.
 * It is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The per-tile
 * partial sums live in a device array built into the library (one per GPU,
 * held until the process ends); the call allocates nothing. Same kernels,
 * arguments and results as swfd_residual, bit for bit. Limit:
 * swfd_residual_workspace_floats(n_steps, n_shots, n_rec) must not exceed
 * SWFD_RESIDUAL_SCRATCH_FLOATS (else SWFD_EINVAL), e.g. 256 shots x 256
 * receivers x 64 000 steps. Calls that overlap on one GPU (two host threads,
 * or two streams not ordered by the caller) share that array and race. */
int swfd_residual_noalloc(int32_t n_steps, int32_t n_shots, int32_t n_rec, const float *samples,
                          const float *observed, int64_t obs_stride_e, int64_t obs_stride_r,
                          const int32_t *obs_index, const int32_t *mute, float *residual,
                          float *misfit, void *stream);

/* Batched runs (no counterpart in the reference, whose batching stops at the
 * shots of one problem, torch_backend.py:601-770): n independent problems,
 * each with its own grid, medium, receivers and step range, advance in
 * lockstep, and at every tick the multi-step launches of all problems that
 * have one go out as one batched kernel launch per kernel variant and row
 * width (a per-problem parameter table; a trailing single step runs as a
 * single-problem launch). Forward runs with samples and work4 / work5 take
 * up to four steps per launch, followed by one batched launch that gathers
 * the samples of every level. Results equal n separate swfd_forward /
 * swfd_adjoint calls bit for bit.
 *
 * Three phases, so that the launch phase can be captured in a HIP graph:
 *   *_table_bytes: size of the device table the run needs (host only);
 *   *_prepare:     writes the tables into `table` (device memory of that
 *                  size) with a blocking copy; call again when any argument
 *                  (pointer, step range, tuning) changes;
 *   swfd_*_batch:  launches, reading the prepared tables; asynchronous on
 *                  `stream`, and contains no host synchronisation.
 * The argument arrays hold one entry per problem. */
int swfd_forward_batch_table_bytes(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                                   const swfd_receivers *recs, const swfd_forward_args *args,
                                   size_t *bytes);
int swfd_forward_batch_prepare(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                               const swfd_receivers *recs, const swfd_forward_args *args,
                               void *table, size_t table_bytes);
int swfd_forward_batch(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                       const swfd_receivers *recs, const swfd_forward_args *args,
                       const void *table, size_t table_bytes, void *stream);
int swfd_adjoint_batch_table_bytes(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                                   const swfd_receivers *recs, const swfd_adjoint_args *args,
                                   size_t *bytes);
int swfd_adjoint_batch_prepare(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                               const swfd_receivers *recs, const swfd_adjoint_args *args,
                               void *table, size_t table_bytes);
int swfd_adjoint_batch(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                       const swfd_receivers *recs, const swfd_adjoint_args *args,
                       const void *table, size_t table_bytes, void *stream);

/* Launch-shape tuning for one grid size (no counterpart in the reference):
 *   kind 0: multi-step forward; seg_rows = rows per march, param = target
 *           work items per CU (64-column blocks);
 *   kind 1: multi-step adjoint; seg_rows = rows per march, param = shot
 *           groups (<= 8);
 *   kind 2: forward steps per launch and form: depth 1 = single-step
 *           kernels (adjoint too), 2..4 = steps per launch; param 1 =
 *           full-width workgroups (rows of up to 1023 columns; required for
 *           more than two steps), 0 = 64-column blocks (two steps; with
 *           depth 2 the adjoint uses blocks too);
 *   kind 3: adjoint steps per launch and form, as kind 2.
 * Without entries: full width where a row fits, 4 forward steps per launch
 * (2 when receiver samples are taken), 4 adjoint steps; batched launches
 * always use two-step blocks. Values <= 0 keep the defaults. swfd_tune_clear forgets every entry. The
 * table is process-wide host state: set it before running (not while another
 * host thread launches). */
int swfd_tune(int32_t kind, int32_t ny, int32_t nx, int32_t seg_rows, int32_t param,
              int32_t depth);
void swfd_tune_clear(void);

#ifdef __cplusplus
}
#endif

#endif /* SWFD_H */
