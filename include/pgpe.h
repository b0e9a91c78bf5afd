/*
 * pgpe.h -- C ABI of libpgpe_hip.so, the gfx950 kernels of the
 * driven-dissipative Gross-Pitaevskii time stepper.
 *
 * Conventions
 *   - Every function returns 0 on success and a nonzero pgpe_status code on
 *     failure; pgpe_last_error() returns a thread-local message for the last
 *     failure on the calling thread.
 *   - Pointers named *_dev are device addresses obtained from pgpe_malloc.
 *   - Complex fields are interleaved (re, im) pairs; precision 0 = single
 *     (complex64 / float32), 1 = double (complex128 / float64).
 *   - Fields are packed member by member; inside a member a field is
 *     (components, ny, nx) with x fastest, matching the reference's canonical
 *     (B, C, ny, nx) layout (polarism/simulation_state.py:84-151).
 *   - Every launch takes an explicit stream (a hipStream_t passed as void*;
 *     NULL = the default stream).  Nothing on a launch path allocates.
 *
 * The Python side binds these with ctypes (polarism_amd/_native.py).  Each
 * entry point names the reference interface it replaces.
 */
#ifndef PGPE_H
#define PGPE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pgpe_status {
    PGPE_OK = 0,
    PGPE_ERR_ARG = 1,     /* invalid argument, shape or size */
    PGPE_ERR_HIP = 2,     /* a HIP runtime call failed */
    PGPE_ERR_UNSUPPORTED = 3
};

enum pgpe_reservoir_kind {
    PGPE_RES_NONE = 0,    /* reservoir/none_reservoir.py */
    PGPE_RES_SINGLE = 1,  /* reservoir/single_reservoir.py:81-98 */
    PGPE_RES_DOUBLE = 2,  /* reservoir/double_reservoir.py:255-279 */
    PGPE_RES_QDOUBLE = 3  /* reservoir/quadratic_double_reservoir.py:131-187 */
};

/* ------------------------------------------------------------------ */
/* Runtime: device, memory, streams, events, graphs                     */
/* (replaces compute_engine.py:33-97 and CuPy's allocator/streams)      */
/* ------------------------------------------------------------------ */
const char* pgpe_last_error(void);
int pgpe_version(void);
int pgpe_device_count(int* count);
int pgpe_device_init(int device);
int pgpe_device_name(int device, char* buf, size_t len);
int pgpe_device_cu_count(int device, int* cus);
int pgpe_malloc(void** ptr_dev, size_t nbytes);
int pgpe_free(void* ptr_dev);
int pgpe_memcpy_h2d(void* dst_dev, const void* src_host, size_t nbytes, void* stream);
int pgpe_memcpy_d2h(void* dst_host, const void* src_dev, size_t nbytes, void* stream);
int pgpe_memcpy_d2d(void* dst_dev, const void* src_dev, size_t nbytes, void* stream);
/* Pinned host memory (hipHostMalloc) and a device-to-host copy that returns without waiting
 * (into pinned memory; wait on an event recorded after it). */
int pgpe_host_alloc(void** ptr_host, size_t nbytes);
int pgpe_host_free(void* ptr_host);
int pgpe_memcpy_d2h_async(void* dst_host, const void* src_dev, size_t nbytes, void* stream);
/* Fills with a small kernel (not hipMemsetAsync) so the call may sit in a captured graph. */
int pgpe_fill_zero(void* ptr_dev, size_t nbytes, void* stream);
int pgpe_stream_create(void** stream);
int pgpe_stream_destroy(void* stream);
int pgpe_stream_sync(void* stream);
int pgpe_device_sync(void);
int pgpe_event_create(void** event);
int pgpe_event_destroy(void* event);
int pgpe_event_record(void* event, void* stream);
int pgpe_event_elapsed_ms(void* start, void* stop, float* ms);
int pgpe_event_sync(void* event);
/* The stream waits for the event's work (replaces the storage writer's                    */
/* stream.wait_event, results/visitors/storage_visitor.py:188-189).                         */
int pgpe_stream_wait_event(void* stream, void* event);
/* Graph capture of a chunk of launches on one stream
 * (replaces rk4_cuda_solver.py:3396-3449 CUDA-graph chunks). */
int pgpe_graph_begin(void* stream);
int pgpe_graph_end(void* stream, void** graph_exec);
int pgpe_graph_launch(void* graph_exec, void* stream);
int pgpe_graph_destroy(void* graph_exec);

/* ------------------------------------------------------------------ */
/* Member table: one row per sweep member, device resident.  Launches  */
/* cover a tile map of (member, local tile) int32 pairs; a member of   */
/* nx x ny points has tiles_x * tiles_y tiles of 64 x tile_h points.   */
/* Per-member physics replaces BatchPhysics columns                    */
/* (config/batch_physics.py:17-64; rk4_cuda_solver.py:983-1021).       */
/* ------------------------------------------------------------------ */
typedef struct pgpe_member {
    int64_t off;            /* offset of this member in packed per-point arrays (points) */
    int64_t pump_off;       /* offset of its pump envelope (points); shared pumps use 0 */
    int32_t nx, ny;
    int32_t tiles_x, tiles_y;
    int32_t tile_begin;     /* first tile index of this member */
    int32_t components;     /* 1 = scalar, 2 = spinor (psi+, psi-) */
    int32_t tile_h;         /* rows per tile (tiles are 64 points wide) */
    int32_t res_components; /* reservoir components: 1 shared, 2 per spin */
    double inv_dx2, inv_dy2;
    double hbar_over_2m;    /* hbar / (2 m_eff) */
    double inv_hbar;
    double kin;             /* -hbar^2 / (2 m_eff) */
    double hbar, m_eff;
    double g_C, g_R, g_I, R, gamma_C, eta;
    double res_a, res_b, res_c, res_d; /* single: gamma_R; double: gamma_I, gamma_A, R_IA, R_AI;
                                          quadratic-double: kappa, gamma_R, gamma_I */
    double alpha_2, zeeman, te_tm, inv_4dxdy;
    double noise_gamma, noise_R, noise_strength, noise_dt_over_4v;
    uint64_t seed;          /* Philox key of this member */
    double pump_w_plus, pump_w_minus; /* spinor pump weights (1 +/- c)/2 */
    int64_t pot_off;        /* offset of its potential in the packed complex potential array (points) */
    int32_t laplacian;      /* pgpe_laplacian_kind */
    int32_t boundary;       /* pgpe_boundary_kind */
} pgpe_member;

/* Finite-difference Laplacian of the rk4 kernels: the periodic or mirrored 5-point
 * stencil (model.py:210-248) or the isotropic 9-point stencil of rk4-cuda
 * (rk4_cuda_solver.py:719-770; needs dx == dy). */
enum pgpe_laplacian_kind { PGPE_LAP_FIVE = 0, PGPE_LAP_NINE = 1 };
/* Grid boundary of the rk4 kernels: periodic wrap, or the closed interval with
 * mirrored (Neumann) neighbours psi[-1] = psi[1] (model.py:227-248). */
enum pgpe_boundary_kind { PGPE_BC_PERIODIC = 0, PGPE_BC_CLOSED = 1 };

/* ------------------------------------------------------------------ */
/* Philox4x32-10 and the truncated-Wigner increment                    */
/* (replaces noise/philox.py:59-130 and noise/stochastic.py:224-253)    */
/* ------------------------------------------------------------------ */
/* out[4j..4j+3] = Philox4x32-10(counters[4j..4j+3], key) for j < n. */
int pgpe_philox_words(const uint32_t* counters_dev, uint32_t* out_dev, uint64_t key,
                      int64_t n, void* stream);
/* Box-Muller pair of the canonical counter (cell = j, step, stream), j < n:
 * out is interleaved (re, im) in the given precision (philox.py:41-56). */
int pgpe_normal_pairs(void* out_dev, int precision, uint64_t key, uint64_t step,
                      uint32_t stream_word, int64_t n, void* stream);
/* psi += sigma (N1 + i N2) over all members of the table, with
 * sigma = sqrt(strength (gamma + R n_active) dt / (4 dx dy)).  active_dev may be
 * NULL (zero density); active_components is 1 (shared) or equal to components. */
int pgpe_wigner_add(void* psi_dev, const void* active_dev, int active_components,
                    const pgpe_member* table_dev, const int32_t* tile_map_dev,
                    int num_tiles, int members, int precision, uint64_t step, void* stream);

/* ------------------------------------------------------------------ */
/* Fused RK4-FDM stage (replaces RK4FDMSolver.step rk4_fdm_solver.py:36-77,  */
/* Model.rhs / _spinor_rhs / _te_tm_energy model.py:89-208 and the      */
/* rk4_cuda_solver.py kernels :43-419, :1556-2194).  Spinor fields are  */
/* (members, 2, ny, nx); spinor reservoirs (members, res_components,    */
/* ny, nx).                                                             */
/* ------------------------------------------------------------------ */
typedef struct pgpe_rk4_args {
    int32_t precision;      /* 0 single, 1 double */
    int32_t reservoir;      /* pgpe_reservoir_kind */
    int32_t stage;          /* 0..3; stage 3 also combines and writes the new state */
    int32_t num_tiles;      /* total tiles over all members */
    int32_t tile_y;         /* rows per tile, equal to the table's tile_h: 8, 16 or 32 */
    int32_t noise;          /* stage 3 only: add the Wigner increment of this step */
    int32_t members;
    int32_t components;     /* 1 scalar; 2 spinor (psi+, psi-) with TE-TM, Zeeman and alpha_2 */
    int32_t res_components; /* spinor reservoir: 1 shared (rho = |psi+|^2 + |psi-|^2), 2 per spin */
    int32_t five_point;     /* pgpe_rk4_step with strips only: 1 promises that every member of the
                               table uses the 5-point Laplacian, so the kernel leaves the 9-point
                               branch out; 0 = general */
    double dt;
    const pgpe_member* table;        /* device */
    const int32_t* tile_map;         /* device, num_tiles (member, local tile) int32 pairs */
    const void* psi0;                /* state at the step start */
    const void* kprev;               /* stages 1..3: derivative of the previous stage */
    const void* k1; const void* k2;  /* stage 3: derivatives of stages 0..2 */
    void* kout;                      /* stages 0..2: this stage's derivative */
    void* psi_out;                   /* stage 3: new state (must not alias psi0) */
    const void* res0[2];             /* reservoir fields at the step start */
    const void* kres_prev[2];
    const void* kres1[2]; const void* kres2[2];
    void* kres_out[2];
    void* res_out[2];                /* stage 3: new reservoir (may alias res0) */
    const void* pump;                /* packed pump envelopes (real) */
    const void* coef;                /* real, members entries: pump coefficient of this stage */
    uint64_t step;                   /* noise counter step = (*step_base if set) + step */
    const uint64_t* step_base;       /* device or NULL */
    const void* potential;           /* complex V packed per member at pot_off, or NULL for V = 0;
                                        enters as (V + gC|psi|^2 + gR nA + gI nI) psi (model.py:89-127) */
    const int32_t* active_tiles;     /* device or NULL: how many leading tile_map entries run; the
                                        launch is then sized for num_tiles entries (the map's
                                        capacity), so a captured graph stays valid while members
                                        retire and the map is rewritten in place */
    int32_t strips;                  /* pgpe_rk4_step only: 0 = 64 x tile_y tiles of the member
                                        table; 1 = row strips of 56 output columns x tile_y rows
                                        (8, 16, 24, or 32 and 40 in single precision), local index
                                        (row << 16) | column, from the strip map; every
                                        member must hold at most 2^28 points (the kernel
                                        addresses a member with 32-bit byte offsets) */
    int32_t no_eta;                  /* pgpe_rk4_step with strips only: 1 promises that no member
                                        of the table has a kinetic-relaxation term (eta = 0), so
                                        the kernel leaves it out; 0 = general */
} pgpe_rk4_args;

int pgpe_rk4_stage(const pgpe_rk4_args* args, void* stream);

/* One whole RK4 step (all four stages, the combination, the clamp and the optional
 * Wigner increment) in a single launch over the same member table and tile map; the
 * same result as the four pgpe_rk4_stage launches (bit for bit on periodic grids, to
 * rounding on closed intervals, whose mirrored halo points are recomputed).  Replaces the
 * reference's one-launch tiled step (rk4_cuda_solver.py:1835-2194).  Scalar fields
 * (components 1): tile_y is the table's tile_h, 16, 32, or 64 (single precision).  Spinor
 * fields (components 2, res_components 1 or 2): single precision, tile_y 16, periodic
 * grids only (closed intervals take pgpe_rk4_stage: the TE-TM cross term changes sign under
 * a reflection, so the whole step cannot recompute a mirrored halo; not checked here, the
 * table is device memory).
 * Uses psi0, res0, pump, psi_out, res_out, noise, step, step_base and coef, which here
 * points at 4 x members coefficients (stage s of member m at coef[s * members + m]).
 * psi_out must not alias psi0 and res_out must not alias res0 (tiles read their
 * neighbours' step-start state). */
int pgpe_rk4_step(const pgpe_rk4_args* args, void* stream);

/* ------------------------------------------------------------------ */
/* Steady-state reduction (replaces simulation_controller.py:710-734)   */
/* For every member: sums[3m+0] = sum rho, sums[3m+1] = sum (rho-old)^2, */
/* sums[3m+2] = sum old^2 in double, then old <- rho when update != 0.   */
/* partials_dev holds 3 doubles per tile of the whole table; the tile  */
/* map may list a subset of the tiles (members still running).         */
/* ------------------------------------------------------------------ */
int pgpe_density_delta(const void* psi_dev, void* old_density_dev, int precision,
                       int components, const pgpe_member* table_dev,
                       const int32_t* tile_map_dev, int num_tiles, int members,
                       double* partials_dev, double* sums_dev, int update, void* stream);

/* Allocation-free variant of pgpe_density_delta.  No counterpart in the    */
/* reference: it is for callers that cannot allocate device memory (code    */
/* under HIP graph capture, tight serving loops).  The per-tile partial      */
/* sums live in a file-scope device array of the library (1.5 MiB per       */
/* device, held until the process ends) instead of partials_dev, so no      */
/* workspace argument is taken and nothing is allocated.  total_tiles is    */
/* the member table's tile count (the partial slots it addresses); more     */
/* than 65536 is rejected with PGPE_ERR_ARG.  The partials are cleared at   */
/* the start of every call (members without tiles in the map sum to 0).     */
/* Every call on one device shares that array: calls that overlap on one    */
/* device (other streams or threads) race and must be serialised.           */
int pgpe_density_delta_noalloc(const void* psi_dev, void* old_density_dev, int precision,
                               int components, const pgpe_member* table_dev,
                               const int32_t* tile_map_dev, int num_tiles, int members,
                               int total_tiles, double* sums_dev, int update, void* stream);

/* ------------------------------------------------------------------ */
/* In-run accumulators (replace the device sums of                     */
/* results/accumulators.py:94-188).  n points of a complex field; the   */
/* sums are double.  pgpe_accumulate_power: sum[i] += |f[i]|^2 (the      */
/* momentum distribution, f = FFT of the field).                        */
/* pgpe_accumulate_moments: rho = |f[i]|^2; sum_rho[i] += rho,           */
/* sum_rho2[i] += rho^2 (g2).  Components are widened to double first.  */
/* ------------------------------------------------------------------ */
int pgpe_accumulate_power(const void* field_dev, double* sum_dev, int64_t n, int precision, void* stream);
int pgpe_accumulate_moments(const void* field_dev, double* sum_rho_dev, double* sum_rho2_dev, int64_t n,
                            int precision, void* stream);
/* Time spectra (accumulators.py:173-270).  pgpe_spectrum_line: planes x nx line cut into a   */
/* ring slot; mode 0 = row `row` of the (planes, ny, nx) field, mode 1 = row ky = 0 of its     */
/* 2-D transform times scale, in fftshift order along x.  pgpe_spectrum_block: the ring of     */
/* segment slots of width values, oldest (slot `position`) first, times the FP64 window, into */
/* a complex128 block.  pgpe_spectrum_power: psd[(line segment + f) nx + x] += |B[f][line nx   */
/* + x]|^2 for the transformed block B (width = lines x nx).                                  */
int pgpe_spectrum_line(const void* field_dev, void* slot_dev, int planes, int nx, int ny, int row, int mode,
                       double scale, int precision, void* stream);
int pgpe_spectrum_block(const void* ring_dev, void* block_dev, const double* window_dev, int segment, int64_t width,
                        int position, int precision, void* stream);
int pgpe_spectrum_power(const void* block_k_dev, double* psd_dev, int segment, int64_t width, int nx, void* stream);
/* Result-frame observables (replaces _compute_density / _compute_stokes /                  */
/* _compute_degree_of_polarization, simulation_controller.py:63-89) of a (members, components,  */
/* plane) field: density (members, components, plane), and for two components the Stokes       */
/* vector (members, 4, plane) and degree of polarisation (members, plane); null outputs skipped. */
int pgpe_frame_fields(const void* psi_dev, int precision, int components, int64_t plane, int64_t members,
                      void* density_dev, void* stokes_dev, void* dop_dev, void* stream);

/* ------------------------------------------------------------------ */
/* Batched 2-D FFT (LDS Stockham, radix 2-16) and the fused split-step  */
/* step (replaces cuFFT plans split_step_fft_cuda_solver.py:729-735,    */
/* numpy.fft in split_step_fft_solver.py:70-79, and the local kernels   */
/* split_step_fft_cuda_solver.py:106-184).                              */
/* ------------------------------------------------------------------ */
typedef struct pgpe_fft_plan pgpe_fft_plan;  /* opaque; twiddles and wave numbers */

/* Sizes 2^a or 3 * 2^a (2 <= n <= 3072); allocates its tables once. */
int pgpe_fft_plan_create(int nx, int ny, double dx, double dy, int precision, pgpe_fft_plan** plan);
int pgpe_fft_plan_destroy(pgpe_fft_plan* plan);
/* planes x (ny, nx) complex: forward (unnormalised) or inverse (scaled by
 * 1/(nx ny)), the numpy.fft.fft2 / ifft2 conventions.  in may equal out. */
int pgpe_fft2(const pgpe_fft_plan* plan, const void* in_dev, void* out_dev, int planes, int inverse,
              void* stream);
/* Forward transform along y only of planes x (ny, width) complex (unnormalised, any width;
 * the plan's ny and precision are used, its nx is not): numpy.fft.fft(a, axis=-2). */
int pgpe_fft_columns(const pgpe_fft_plan* plan, const void* in_dev, void* out_dev, int width, int planes,
                     void* stream);

typedef struct pgpe_split_args {
    int32_t precision;
    int32_t reservoir;      /* pgpe_reservoir_kind */
    int32_t ordering;       /* 0: R K N noise K R (CPU solver order); 1: K [R N R noise] K */
    int32_t noise;          /* add the Wigner increment */
    int32_t members;        /* members of equal grids, one table row each */
    int32_t components;     /* 1 (scalar) or 2 (spinor psi+, psi-; 0 reads as 1) */
    double dt;
    const pgpe_member* table;
    const pgpe_fft_plan* plan;
    void* psi;              /* in/out, members x components x (ny, nx) complex */
    void* tmp;              /* scratch of the same size */
    void* res[2];           /* in/out reservoir fields, members x res_components x (ny, nx) */
    const void* pump;       /* envelopes (real) */
    const void* coef;       /* 2 * members reals: pump coefficient at t+dt/4, then at t+3dt/4;
                               3 * members with PGPE_CHAIN_OUT (row 2: the next step's t+dt/4) */
    uint64_t step;          /* noise counter step = (*step_base if set) + step */
    const uint64_t* step_base;
    int32_t res_components; /* spinor: 2 = one reservoir per component driven by its density and
                               its weighted pump, 1 = one shared reservoir driven by the total
                               density (split_step_fft_solver.py:81-94); 0 reads as 1 */
    int32_t chain;          /* 0, or bits of pgpe_split_chain for consecutive steps of the scalar
                               CPU order (ordering 0): the field stays in tmp between them, and
                               psi holds the state only after a step without PGPE_CHAIN_OUT */
} pgpe_split_args;

/* Chained split steps: one row pass instead of two between consecutive steps (the last pass of
 * a step and the first of the next, same operations in the same order: bit-identical). */
enum pgpe_split_chain {
    PGPE_CHAIN_IN = 1,   /* the field enters in tmp, as a PGPE_CHAIN_OUT step left it */
    PGPE_CHAIN_OUT = 2   /* leave the field in tmp, row-transformed after the next step's first
                            reservoir half-step (coefficient row 2) */
};

int pgpe_split_step(const pgpe_split_args* args, void* stream);

/* ------------------------------------------------------------------ */
/* Lawson integrating-factor RK (replaces IFRKFFTCudaSolver._generic_step */
/* and _adaptive_error, ifrk_fft_cuda_solver.py:148-293, and the fused   */
/* kernels of ifrk_cuda_fused.py:10-374).                                */
/* Stage values are kept in the interaction picture: the stage call with */
/* frac0 = c_i writes f_out = g_i = L(-c_i h) FFT(N_i), so that           */
/* L(c h) (psi0^k + sum_j hc[j] g_j) = L(c h) psi0^k                      */
/*   + sum_j hc[j] L((c - c_j) h) FFT(N_j), the reference's combination   */
/* with one propagator per k point instead of one per term (fr[] is not  */
/* read).                                                                 */
/* op STAGE: psi_i^k = L(frac0 h) (psi0^k + sum_j hc[j] f_j), inverse     */
/*   transform, N(psi_i, n_i) with n_i = n0 + sum_j rc[j] dres_j,          */
/*   writes dres_out and f_out = L(-frac0 h) FFT(N).  stage_real = 1 takes */
/*   psi_real (real space) as psi_i instead of the combination.           */
/* op FINAL: psi_new^k = L(h) (psi0^k + sum_j hc[j] f_j), inverse         */
/*   into psi_out, res_out = max(n0 + sum_j rc[j] dres_j, 0), optional     */
/*   noise; with nerr > 0 also the embedded error fields (he5, he3) and    */
/*   per-member sums of (|e| / (atol + rtol max(|start|, |new|)))^2 into   */
/*   error_sums[2 m], [2 m + 1] (double).                                  */
/* ------------------------------------------------------------------ */
#define PGPE_MAX_STAGES 12
enum pgpe_lawson_op { PGPE_LAWSON_STAGE = 0, PGPE_LAWSON_FINAL = 1 };

typedef struct pgpe_lawson_args {
    int32_t precision, reservoir, op, components, res_components, members;
    int32_t nterms, lap, nerr, stage_real, noise, pad0;
    double h, frac0, atol, rtol;
    double hc[PGPE_MAX_STAGES], fr[PGPE_MAX_STAGES], rc[PGPE_MAX_STAGES];
    double he5[PGPE_MAX_STAGES], he3[PGPE_MAX_STAGES];
    const pgpe_member* table;
    const pgpe_fft_plan* plan;
    const void* psi0k;                       /* members x comps x (ny, nx) complex */
    const void* f[PGPE_MAX_STAGES];          /* stage transforms */
    const void* dres[PGPE_MAX_STAGES][2];    /* stage reservoir derivatives */
    void* work;                              /* scratch, one field */
    void* aux;                               /* scratch, two fields (lap or error fields) */
    void* f_out;
    const void* psi_real;
    const void* res0[2];
    void* dres_out[2];
    const void* pump;
    const void* coef;                        /* members reals: pump coefficient at the stage time */
    const void* potential;                   /* complex (ny, nx) potential or NULL */
    void* psi_out;
    void* res_out[2];
    void* psik_out;                          /* optional: psi_new^k */
    const void* psi_start;                   /* real-space field at the trial start */
    double* partials;                        /* 2 * members * ny doubles */
    double* error_sums;                      /* 2 * members doubles */
    uint64_t step;
    const uint64_t* step_base;
    /* Device-driven adaptive trials (both NULL otherwise): the trial's h is read from
     * *h_dev and hc, rc, he5, he3 then hold the bare tableau weights a_ij, b_j, b_j - b_hat_j
     * (the kernels form h * weight); every kernel returns at once while *skip_dev != 0. */
    const double* h_dev;
    const int32_t* skip_dev;
} pgpe_lawson_args;

int pgpe_lawson(const pgpe_lawson_args* args, void* stream);

/* ------------------------------------------------------------------ */
/* On-device control of adaptive trials: replaces the host loop of      */
/* abstract_solver.py:86-96 with solver/step_controller.py:48-81 (PI     */
/* update, tick landing, underflow guard) and the host-side member max  */
/* of ifrk_fft_cuda_solver.py:293.  The state lives in device memory.   */
/* propose: h from h_next, clamped onto target.  decide: member norms   */
/* from the (err5, err3) sums of pgpe_lawson's final op (DP5            */
/* sqrt(err5 / count); blend != 0: DOP853 err5 / sqrt((err5 + 0.01 err3) */
/* count)), max over members, accept when <= 1, PI update; done when t  */
/* reaches target.  commit: copies each src region over dst when the     */
/* last trial was accepted.  All three do nothing once done is set.     */
/* ------------------------------------------------------------------ */
enum { PGPE_ADAPTIVE_OK = 0, PGPE_ADAPTIVE_UNDERFLOW = 1 };
#define PGPE_COMMIT_MAX 8

typedef struct pgpe_adaptive_state {
    double t, target;          /* current time; the tick this advance lands on */
    double h_next, err_prev;   /* controller memory */
    double dt_max;             /* <= 0: no limit */
    double k;                  /* min(order, embedded order) + 1 */
    double h, prop, error;     /* the last trial: step, proposal, norm */
    double dt_min, dt_max_used, dt_last;
    double count;              /* points per member in the error norm */
    int32_t members, blend;
    int32_t clamped, no_growth, accepted, rejected, done, commit, status, pad0;
} pgpe_adaptive_state;

typedef struct pgpe_commit_list {
    void* dst[PGPE_COMMIT_MAX];
    const void* src[PGPE_COMMIT_MAX];
    uint64_t bytes[PGPE_COMMIT_MAX];   /* multiples of 4; regions 16-byte aligned */
    int32_t count, pad0;
} pgpe_commit_list;

int pgpe_adaptive_propose(pgpe_adaptive_state* state_dev, void* stream);
int pgpe_adaptive_decide(pgpe_adaptive_state* state_dev, const double* error_sums_dev, void* stream);
int pgpe_adaptive_commit(const pgpe_adaptive_state* state_dev, const pgpe_commit_list* list, void* stream);

/* ------------------------------------------------------------------ */
/* Surrogate (Fourier neural operator) kernels: BF16 MFMA GEMM with fp32 */
/* accumulation and fused epilogue, plus the elementwise pieces of the   */
/* forward and backward passes.  No counterpart in the reference.        */
/* C[b](m, n) = alpha sum_k A[b](m, k) B[b](k, n) + beta C + bias + res,  */
/* then optional GELU; element (m, k) of A is at A + b sab + m sam + k sak */
/* (A needs sak == 1 or sam == 1; B needs sbk == 1 or sbn == 1).          */
/* ------------------------------------------------------------------ */
typedef struct pgpe_gemm_args {
    const void* A;            /* bf16 */
    const void* B;            /* bf16 */
    void* C;                  /* fp32, or bf16 when out_bf16 */
    const float* bias;        /* M (bias_mode 1) or N (bias_mode 2) floats */
    const float* residual;    /* fp32 added before GELU, strides of C, batch stride srb */
    int32_t M, N, K, batch;
    int64_t sam, sak, sab;
    int64_t sbk, sbn, sbb;
    int64_t scm, scn, scb;
    int64_t srb;
    float alpha, beta;
    int32_t bias_mode, gelu, out_bf16, pad0;
} pgpe_gemm_args;

int pgpe_gemm_bf16(const pgpe_gemm_args* args, void* stream);
int pgpe_f32_to_bf16(const float* x_dev, void* y_dev, int64_t n, void* stream);
/* y = GELU(x) (erf form) into fp32 y and / or bf16 yb. */
int pgpe_gelu_fwd(const float* x_dev, float* y_dev, void* yb_dev, int64_t n, void* stream);
/* dx = dy GELU'(x) into fp32 dx and / or bf16 dxb. */
int pgpe_gelu_bwd(const float* x_dev, const float* dy_dev, float* dx_dev, void* dxb_dev, int64_t n, void* stream);
/* Planar spectra laid out [ky][c][b][kx] (my x C x B x mx): u = d[c, ky, kx] v, with
 * conj(d) when conj; fp32 (ur, ui) and / or bf16 (urb, uib) outputs. */
int pgpe_mode_scale(const float* vr_dev, const float* vi_dev, const float* dr_dev, const float* di_dev,
                    float* ur_dev, float* ui_dev, void* urb_dev, void* uib_dev, int my, int C, int B, int mx,
                    int conj, void* stream);
/* gd[c, ky, kx] = sum_b conj(v) du over the same layout (gradient of the per-mode scale). */
int pgpe_mode_scale_grad(const float* vr_dev, const float* vi_dev, const float* dur_dev, const float* dui_dev,
                         float* gdr_dev, float* gdi_dev, int my, int C, int B, int mx, void* stream);
/* out[c] = sum_r x[r row_stride + c col_stride], fp32. */
int pgpe_colsum(const float* x_dev, float* out_dev, int64_t rows, int64_t cols, int64_t row_stride,
                int64_t col_stride, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PGPE_H */
