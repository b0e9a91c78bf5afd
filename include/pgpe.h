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
/* Graph capture of a chunk of launches on one stream
 * (replaces rk4_cuda_solver.py:3396-3449 CUDA-graph chunks). */
int pgpe_graph_begin(void* stream);
int pgpe_graph_end(void* stream, void** graph_exec);
int pgpe_graph_launch(void* graph_exec, void* stream);
int pgpe_graph_destroy(void* graph_exec);

/* ------------------------------------------------------------------ */
/* Member table: one row per sweep member, device resident.            */
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
    double g_C, g_R, g_I, R, gamma_C, eta;
    double res_a, res_b, res_c, res_d; /* single: gamma_R; double: gamma_I, gamma_A, R_IA, R_AI;
                                          quadratic-double: kappa, gamma_R, gamma_I */
    double alpha_2, zeeman, te_tm, inv_4dxdy;
    double noise_gamma, noise_R, noise_strength, noise_dt_over_4v;
    uint64_t seed;          /* Philox key of this member */
    double pump_w_plus, pump_w_minus; /* spinor pump weights (1 +/- c)/2 */
} pgpe_member;

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
                    const pgpe_member* table_dev, const int32_t* tile_member_dev,
                    int num_tiles, int members, int precision, uint64_t step, void* stream);

/* ------------------------------------------------------------------ */
/* Fused RK4-FDM stage (replaces RK4FDMSolver.step rk4_fdm_solver.py:36-77,  */
/* Model.rhs model.py:89-127 and rk4_cuda_solver.py:43-419)             */
/* ------------------------------------------------------------------ */
typedef struct pgpe_rk4_args {
    int32_t precision;      /* 0 single, 1 double */
    int32_t reservoir;      /* pgpe_reservoir_kind */
    int32_t stage;          /* 0..3; stage 3 also combines and writes the new state */
    int32_t num_tiles;      /* total tiles over all members */
    int32_t tile_y;         /* rows per tile, equal to the table's tile_h: 8, 16 or 32 */
    int32_t noise;          /* stage 3 only: add the Wigner increment of this step */
    int32_t members;
    int32_t pad0;
    double dt;
    const pgpe_member* table;        /* device */
    const int32_t* tile_member;      /* device, num_tiles entries */
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
} pgpe_rk4_args;

int pgpe_rk4_stage(const pgpe_rk4_args* args, void* stream);

/* ------------------------------------------------------------------ */
/* Steady-state reduction (replaces simulation_controller.py:710-734)   */
/* For every member: sums[3m+0] = sum rho, sums[3m+1] = sum (rho-old)^2, */
/* sums[3m+2] = sum old^2 in double, then old <- rho when update != 0.   */
/* partials_dev holds 3 * num_tiles doubles.                            */
/* ------------------------------------------------------------------ */
int pgpe_density_delta(const void* psi_dev, void* old_density_dev, int precision,
                       int components, const pgpe_member* table_dev,
                       const int32_t* tile_member_dev, int num_tiles, int members,
                       double* partials_dev, double* sums_dev, int update, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PGPE_H */
