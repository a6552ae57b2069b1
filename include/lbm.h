/*
 * lbm.h -- C ABI of the gfx950 D2Q9 lattice-Boltzmann engine (liblbm.so).
 *
 * This is the drop-in boundary for the reference's kernel entry points.  The reference launches CUDA
 * kernels through CuPy (cp.RawModule(...).get_function(name)(grid, block, args), lbm_smemaa_main.py:84-86,
 * 141-142); here each kernel becomes an extern "C" function taking plain device pointers, sizes and a
 * HIP stream, and the Python driver (lbm-cuda_amd/lbm) binds them with ctypes.
 *
 * Conventions (shared by every entry point):
 *   - Populations are SoA FP32 planes f[q][y][x], x fastest, plane stride W*H (lbm_core4.cu:26-27,40;
 *     lbm_smemaa_main.py:35,43).  q follows the reference's D2Q9 order: c_q = (0,0) (1,0) (0,1) (-1,0)
 *     (0,-1) (1,1) (-1,1) (-1,-1) (1,-1) (lbm_core4.cu:1-3).
 *   - mask[y][x] is one byte per cell, non-zero = solid (the reference's bool mask, lbm_smemaa_main.py:68).
 *     In channel mode rows 0 and H-1 must be solid and columns 0 and W-1 open on rows 1..H-2
 *     (lbm_smemaa_main.py:61-66); the Python driver enforces this.
 *   - inlet_u[y] is the Zou-He inlet speed per row (the reference computes u_in * ramp(y) inline,
 *     lbm_core4.cu:46-51; the driver builds the ramp or a parabola on the host in FP64).
 *   - AA parity: 0 = canonical (slot q at x holds the incoming f_q(x)); 1 = reversed (slot opp(q) at x
 *     holds the post-collision f*_q(x)).  This is the A/B convention of lbm_core_smemaa.cu:92-106,274-292
 *     applied to the HBM state.
 *   - Pointers are device addresses owned by the caller; only lbm_ensemble_create (its descriptor table)
 *     and lbm_graph_end (the instantiated graph) allocate.  Calls are asynchronous
 *     on `stream` (NULL = default stream).  Every function returns 0 on success, otherwise a negative
 *     library code or a positive hipError_t, and lbm_last_error() describes the last failure of the
 *     calling thread.
 */
#ifndef LBM_H
#define LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *lbm_stream_t; /* == hipStream_t */

enum lbm_model {
    LBM_MRT_LES = 0, /* MRT + Smagorinsky with the reference clamps (lbm_core4.cu:124-210) */
    LBM_BGK = 1      /* BGK of the temporally blocked family (lbm_core_smemaa.cu:194-220)  */
};

enum lbm_bc {
    LBM_BC_CHANNEL = 0, /* wall rows, Zou-He inlet at x=0, zero-gradient outlet ghost at x=W-1 */
    LBM_BC_PERIODIC = 1 /* both directions periodic; mask cells are bounce-back obstacles      */
};

/* Storage format of the populations.  LBM_FMT_F32 stores f; the 16-bit formats store deviations
 * h = f - w_q from the rest state (w_q the FP32 weights 0.444444444f, 0.1111111111f, 0.0277777778f of
 * lbm_core_smemaa.cu:4-5), rounded to nearest even; the collision always computes in FP32.  With
 * lbm_params.row_base set they store deviations from a per-row base state instead (see lbm_row_base). */
enum lbm_format {
    LBM_FMT_F32 = 0,
    LBM_FMT_F16 = 1, /* IEEE half */
    LBM_FMT_BF16 = 2 /* bfloat16 */
};

enum lbm_status {
    LBM_OK = 0,
    LBM_ERR_ARG = -1,      /* invalid argument (see lbm_last_error) */
    LBM_ERR_SIZE = -2,     /* lattice too large / too small          */
    LBM_ERR_UNSUPPORTED = -3
};

/* Physical and numerical parameters.  The reference hard-codes Cs = 0.2 (lbm_core4.cu:123), the MRT
 * rates si = 1.1 (lbm_core4.cu:6) and the clamps (lbm_core4.cu:142-144); they are fields here. */
typedef struct lbm_params {
    int32_t W, H;   /* lattice size (totwidth, totheight) */
    int32_t model;  /* enum lbm_model */
    int32_t bc;     /* enum lbm_bc */
    int32_t clamp;  /* 1: |u| <= 0.4 and rho in [0,1.5] before MRT collision (populations are always
                       clamped to [0,10], lbm_core4.cu:203, lbm_core_smemaa.cu:220) */
    float tau_inv;  /* relaxation rate of the stress moments (MRT) / BGK rate */
    float cs;       /* Smagorinsky constant; 0 = plain MRT (no LES) */
    float s_e, s_eps, s_q; /* MRT rates of e, eps and q_x/q_y (reference: 1.1) */
    int32_t sums;   /* 0: plain FP32 sums of rho - 1 and j in the MRT collision; 1: compensated (Neumaier) sums of
                       both, for the precision study.  Every stepping entry point takes 1 (single, pull, fused,
                       ensemble; each temporally blocked form bit-identical to its single steps); lbm_neq and the
                       closure step return LBM_ERR_UNSUPPORTED */
    int32_t fast;   /* 0: the collision's reciprocals and square roots are correctly rounded (the CPU FP32 mirror of the
                       oracle reproduces every kernel bit for bit); 1: the hardware's approximate v_rcp_f32 / v_sqrt_f32
                       (within 1 ulp), the arithmetic of the reference's -use_fast_math build (lbm_smemaa_main.py:84):
                       MRT only, taken by lbm_step_aa(_h16)(_n), the tile kernels of lbm_step_fused(_h16)(_n) (strip 0)
                       and the ensembles (all members alike), each blocked form still bit-identical to its single
                       steps; other entry points return LBM_ERR_UNSUPPORTED, and so does fast = 1 with sums = 1 */
    const float *row_base; /* NULL, or the per-row base of the 16-bit formats: a device table of LBM_ROW_BASE_STRIDE * H
                              floats, 16-byte aligned (lbm_row_base).  Used by lbm_step_aa_h16(_n), lbm_step_fused_h16(_n)
                              (k = 2, two-step tile kernel), lbm_macro_h16 and lbm_convert_state; FP32 entry points ignore
                              it; lbm_convert and lbm_ensemble_create reject it */
    const struct lbm_base_window *window; /* NULL, or a window of per-cell base velocities (needs row_base; see
                                             lbm_base_window); taken by the entry points that take row_base */
} lbm_params;

/*
 * A window of per-cell base states for row-based 16-bit storage (no reference counterpart).  Cells in
 * [x0, x0 + w) x [y0, y0 + h) take the base values of their slots from a velocity of their own instead of from
 * their row: `vel` is a device array of two IEEE halves (ux, uy) per window cell, row-major over the window,
 * 4-byte aligned (cell (x, y) at vel[2 ((y - y0) w + (x - x0))]), and the base value of direction q is the
 * equilibrium deviation at rho = 1, d_q = w_q ((3 cu + 4.5 cu^2) - 1.5 (ux^2 + uy^2)) with cu = c_q.u, evaluated in
 * FP32 from the half values (the equilibrium lbm_row_base tabulates per row in FP64, here per cell).  Where the flow departs from every per-row
 * profile -- behind a body -- the stored deviations, and their rounding, stay small only with a local base;
 * lbm_rebase_window re-derives it from the current flow.  The window must lie in
 * 1 <= x <= W - 3, 1 <= y <= H - 2: the edges, and the channel's last fluid column with its outlet ghost (a copy
 * stored against the ghost's own base), stay on their rows' base.
 */
typedef struct lbm_base_window {
    int32_t x0, y0, w, h;
    const uint16_t *vel;
} lbm_base_window;

/*
 * Per-row base state of the 16-bit formats (no reference counterpart: the reference stores FP32 only).  A 16-bit
 * word holds about 3 significant decimal digits of the stored deviation, so the smaller the deviation the smaller its
 * rounding error.  With lbm_params.row_base = table, a slot on row y that holds a population moving in direction q
 * stores half(clamp(h) - d) instead of half(clamp(h)), where h = f - w_q is the deviation the collision works on,
 * clamp the population clamp [-w_q, 10 - w_q] and d = table[LBM_ROW_BASE_STRIDE * y + q]; a load returns
 * (float)half + d.  Both FP32 operations round once.  A table built from the flow's mean velocity profile (a channel's
 * inlet profile) stores mostly the departure from that profile: on the 4096 x 1024 cylinder flow it lowers the FP16
 * velocity error against FP64 about tenfold.  Which direction a slot holds follows the AA parity: canonical slot q
 * holds direction q, reversed slot q holds opp(q).
 *
 * lbm_row_base fills a host table of LBM_ROW_BASE_STRIDE * H floats: entry q < 9 of row y is the deviation from w_q of
 * the equilibrium at rho = 1, u = (ux[y], uy[y]) (uy may be NULL: 0), w_q (3 c.u + 4.5 (c.u)^2 - 1.5 u.u) evaluated
 * in FP64 with the exact weights and rounded once; entries 9..15 are 0.  Any other table is valid too.
 */
#define LBM_ROW_BASE_STRIDE 16
int lbm_row_base(int H, const float *ux, const float *uy, float *table);

/* ABI version of this header (bumped on any signature change). */
int lbm_abi_version(void);

/* Last error message of the calling thread ("" if none). */
const char *lbm_last_error(void);

/* Wait for all work on `stream`; returns the first asynchronous error, if any. */
int lbm_sync(lbm_stream_t stream);

/*
 * One in-place AA collide-stream step of the whole lattice.
 * Replaces the reference's per-step kernel (fused_lbmkernel of lbm_core4.cu:7-13) with a single buffer:
 * parity 0 collides locally and stores reversed; parity 1 gathers from neighbours, collides and
 * scatters back to canonical (bounce-back on solid sources/targets, x=0 Zou-He, W-1 zero-gradient ghost).
 * ux/uy/rho may all be NULL (no macro output) or all non-NULL (written for every updated cell, as
 * lbm_core4.cu:59-61,145-147 does).
 */
int lbm_step_aa(const lbm_params *p, const uint8_t *mask, float *f, const float *inlet_u, int parity,
                float *ux, float *uy, float *rho, lbm_stream_t stream);

/*
 * nsteps consecutive lbm_step_aa launches on `stream` (parities parity, parity^1, ...), issued from C so
 * that a long run costs no host-language overhead per step.  The state ends at parity (parity+nsteps)&1.
 * Replaces the reference's replay loop of a captured 2-launch graph (lbm_mine2.py:73-94).
 */
int lbm_step_aa_n(const lbm_params *p, const uint8_t *mask, float *f, const float *inlet_u, int parity, int nsteps,
                  lbm_stream_t stream);

/*
 * Reference-compatible two-buffer pull step: fused_lbmkernel of lbm_core4.cu:7-13
 * (mask, f_now, f_out, ux, uy, rho, totwidth, totheight, tau_inv, u_in, tau_inv_bnd).  f_out receives
 * post-collision populations, the W-1 column a copy of W-2, and ux/uy/rho are always written.
 * `tau_inv_bnd` is unused by the reference (lbm_core4.cu:12) and is therefore not a parameter;
 * u_in becomes the per-row inlet_u array.
 */
int lbm_step_pull(const lbm_params *p, const uint8_t *mask, const float *f_now, float *f_out, float *ux,
                  float *uy, float *rho, const float *inlet_u, lbm_stream_t stream);

/*
 * k collide-stream steps, temporally blocked in LDS (the fused_lbmkernel of lbm_core_smemaa.cu:12-15 /
 * lbm_core_smemaa_endfused.cu:12-14, with T = k chosen at run time): one tile launch for k = 2 and 4 from an even
 * state, otherwise a tile launch or a copy followed by single steps on f_out (lbm_fused_default_cfg).
 * f_now (parity `parity`) is read, f_out receives the state after k steps (parity (parity + k) & 1);
 * f_now and f_out must not overlap.  The result is bit-identical to k calls of lbm_step_aa.
 * 1 <= k <= lbm_fused_max_k().  ux/uy/rho: NULL or the macros of the last step's collision.
 */
int lbm_step_fused(const lbm_params *p, const uint8_t *mask, const float *f_now, float *f_out,
                   const float *inlet_u, int parity, int k, float *ux, float *uy, float *rho,
                   lbm_stream_t stream);

/* Largest k accepted by lbm_step_fused. */
int lbm_fused_max_k(void);

/* Launch configuration of the k-step kernel (no reference counterpart: the reference fixes its tile and
 * T at compile time, lbm_core_smemaa.cu:8-9).  strip = LDS columns per workgroup including the halo
 * (64 or 128); chunk = rows per workgroup.  strip = 0 selects the two-step tile kernel instead (k = 2 from
 * parity 0 only; chunk = 8, 12, 16 or 24 core rows per 128-column tile). */
typedef struct lbm_fused_cfg {
    int32_t strip;
    int32_t chunk;
} lbm_fused_cfg;

/* A strip configuration for k (128-column strips, the tallest chunk that leaves at least 512 workgroups), for callers
 * that select the strip sweep.  lbm_step_fused without a configuration does not run the strip sweep: from an even
 * state with k = 2 (any storage) or k = 4 (FP32 populations) it runs the tile kernels, {strip 0, chunk 12}; for
 * other k and states one tile launch (even states, k >= 2) or a device copy of f_now to f_out (odd states, k = 1),
 * then the remaining steps as single steps in place on f_out, bit-identical like every form and faster than the
 * strip sweep at 4096 x 1024.  Such a call issues up to k launches (and a copy) on `stream`. */
int lbm_fused_default_cfg(const lbm_params *p, int k, lbm_fused_cfg *cfg);

/* lbm_step_fused with an explicit configuration (NULL = default); used by the autotuner. */
int lbm_step_fused_cfg(const lbm_params *p, const uint8_t *mask, const float *f_now, float *f_out,
                       const float *inlet_u, int parity, int k, const lbm_fused_cfg *cfg, float *ux, float *uy,
                       float *rho, lbm_stream_t stream);

/*
 * nlaunch consecutive lbm_step_fused_cfg launches on `stream`, alternating between the two buffers: launch i
 * reads f_a (i even) or f_b (i odd), writes the other one, and starts at parity (parity + i k) & 1.  The state
 * ends in f_b when nlaunch is odd, in f_a otherwise; the other buffer's contents are then unspecified (without a
 * configuration a launch from an odd state with k >= 3 takes its first step in place on the buffer it reads, then
 * a tile launch, instead of copying that buffer).  Issued from C so that a long run costs no
 * host-language overhead per launch; replaces the replay loop of the reference's captured graph of two fused
 * launches (lbm_smemaa_main.py:136-151,173-174).
 */
int lbm_step_fused_n(const lbm_params *p, const uint8_t *mask, float *f_a, float *f_b, const float *inlet_u,
                     int parity, int k, int nlaunch, const lbm_fused_cfg *cfg, lbm_stream_t stream);

/*
 * 16-bit deviation storage (no reference counterpart: the reference stores FP32 only).  Same semantics as
 * lbm_step_aa / lbm_step_aa_n / lbm_step_fused_cfg / lbm_macro on h = f - w stored as `fmt`
 * (LBM_FMT_F16 or LBM_FMT_BF16); stores clamp h to [-w, 10 - w], the population clamp of lbm_core4.cu:203.
 */
int lbm_step_aa_h16(const lbm_params *p, const uint8_t *mask, uint16_t *h, int fmt, const float *inlet_u,
                    int parity, float *ux, float *uy, float *rho, lbm_stream_t stream);
int lbm_step_aa_h16_n(const lbm_params *p, const uint8_t *mask, uint16_t *h, int fmt, const float *inlet_u,
                      int parity, int nsteps, lbm_stream_t stream);
int lbm_step_fused_h16(const lbm_params *p, const uint8_t *mask, const uint16_t *h_now, uint16_t *h_out, int fmt,
                       const float *inlet_u, int parity, int k, const lbm_fused_cfg *cfg, float *ux, float *uy,
                       float *rho, lbm_stream_t stream);
int lbm_step_fused_h16_n(const lbm_params *p, const uint8_t *mask, uint16_t *h_a, uint16_t *h_b, int fmt,
                         const float *inlet_u, int parity, int k, int nlaunch, const lbm_fused_cfg *cfg,
                         lbm_stream_t stream);
int lbm_macro_h16(const lbm_params *p, const uint8_t *mask, const uint16_t *h, int fmt, const float *inlet_u,
                  int parity, float *ux, float *uy, float *rho, lbm_stream_t stream);

/*
 * Convert a whole 9-plane state between formats (FP32 f <-> 16-bit h = f - w; 16-bit -> 16-bit through
 * FP32).  Round to nearest even, no clamping.  src and dst must not overlap.
 */
int lbm_convert(const lbm_params *p, const void *src, int src_fmt, void *dst, int dst_fmt, lbm_stream_t stream);

/*
 * lbm_convert for a state of AA parity `parity`, with p->row_base applied to the 16-bit side(s): 16-bit -> FP32 gives
 * w_q + ((float)half + d), FP32 -> 16-bit half((f - w_q) - d), d the base value of the direction the slot holds at that
 * parity.  Without p->row_base it is lbm_convert.  Same-format calls copy the stored words.
 */
int lbm_convert_state(const lbm_params *p, const void *src, int src_fmt, void *dst, int dst_fmt, int parity,
                      lbm_stream_t stream);

/*
 * Re-derive the base window (p->window) from the current flow: for every window cell, the new base velocity is the
 * cell's u = j / rho (FP32, from its stored populations decoded with the current base; 0 where rho is outside
 * (0.5, 2) or a component is not below 1 in magnitude) rounded to IEEE half, written to new_vel (the layout of
 * lbm_base_window.vel; it must not overlap it), and the cell's 16-bit words are re-encoded against the base values
 * it gives (decoded with the current ones, as lbm_convert_state does).  The caller then points p->window->vel at
 * new_vel.  parity: the state's AA parity.  No reference counterpart.
 */
int lbm_rebase_window(const lbm_params *p, uint16_t *h, int fmt, int parity, uint16_t *new_vel, lbm_stream_t stream);

/*
 * rho, ux, uy of a stored state (macro_kernel of lbm_core_smemaa_endfused.cu:377-379, extended to both
 * parities): canonical states get Zou-He at x=0; reversed states are read through opp().  Only updated
 * cells are written.
 */
int lbm_macro(const lbm_params *p, const uint8_t *mask, const float *f, const float *inlet_u, int parity,
              float *ux, float *uy, float *rho, lbm_stream_t stream);

/*
 * Vorticity colour map into RGBA8 (visualizekernel of lbm_core4.cu:214-217): interior cells only,
 * solids black, vort = ((uy[x+1]-uy[x-1]) - (ux[y+1]-ux[y-1])) * vort_scale.  rho is unused, as in the
 * reference, and therefore not a parameter.
 */
int lbm_visualize(const lbm_params *p, const float *ux, const float *uy, const uint8_t *mask, uint8_t *rgba,
                  float vort_scale, lbm_stream_t stream);

/*
 * FP64 diagnostics of a stored state, reduced deterministically on the GPU (no reference counterpart:
 * the reference keeps no diagnostics; the build needs them for drag/lift and conservation checks).
 * out[0] = mass, out[1..2] = momentum (same cells and Zou-He completion as lbm_macro), out[3..4] =
 * momentum-exchange force 2 sum c_q f*_q over links from updated cells to interior solid cells
 * (wall rows excluded).  `work` is a device buffer of lbm_diagnostics_work_size(p) bytes; `out` is a
 * device buffer of 5 doubles.
 */
size_t lbm_diagnostics_work_size(const lbm_params *p);
int lbm_diagnostics(const lbm_params *p, const uint8_t *mask, const float *f, const float *inlet_u, int parity,
                    double *work, double *out, lbm_stream_t stream);

/*
 * lbm_diagnostics without a workspace argument and without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under graph capture, tight serving or health-check loops).
 * The partial sums live in a static device array of lbm_diagnostics_scratch_bytes() bytes (40 KiB) that each
 * GPU holds until the process ends; its size does not depend on the lattice, so every lattice the library
 * accepts fits.  All calls on one GPU share that array: calls that overlap on one GPU (different streams, or
 * threads) race and must be serialised by the caller.  Results equal lbm_diagnostics' bit for bit.
 */
size_t lbm_diagnostics_scratch_bytes(void);
int lbm_diagnostics_noalloc(const lbm_params *p, const uint8_t *mask, const float *f, const float *inlet_u,
                            int parity, double *out, lbm_stream_t stream);

/*
 * Batched ensembles (no reference counterpart: the reference advances one lattice per process,
 * lbm_smemaa_main.py:8-10).  An ensemble is a set of independent lattices of any sizes that share model,
 * bc and storage format; lbm_ensemble_step advances every member by one in-place AA step (the update of
 * lbm_step_aa / lbm_step_aa_h16, bit for bit) in ONE persistent launch whose workgroups walk a tile list
 * spanning all members.  The instance table is copied at creation; the buffers it names must stay alive
 * and in place while the ensemble exists.  All members share one parity: step them together.
 * `blocks` <= 0 picks the default grid (as many workgroups as are resident at once); other values are
 * rounded up to a multiple of 8.
 */
typedef struct lbm_instance {
    lbm_params p;
    const uint8_t *mask; /* W*H bytes */
    void *f;             /* 9*W*H FP32 populations, or 16-bit deviations for LBM_FMT_F16 / LBM_FMT_BF16 */
    const float *inlet_u; /* H floats (channel bc; may be NULL when periodic) */
    void *f_alt;         /* a second state buffer of the same size for the two-step path (NULL: k = 1 only) */
} lbm_instance;
typedef struct lbm_ensemble lbm_ensemble;
int lbm_ensemble_create(const lbm_instance *inst, int n, int fmt, lbm_ensemble **out);
/* Tiles of 128 x 2 cells over all members, cells over all members, and the default grid. */
int lbm_ensemble_info(const lbm_ensemble *e, int *total_tiles, long long *total_cells, int *default_blocks);
int lbm_ensemble_step(lbm_ensemble *e, int parity, int blocks, lbm_stream_t stream);
/*
 * nsteps steps of every member, grouped for the last-level cache: members are taken in order in groups
 * whose states total at most group_bytes (a larger member is a group of its own), and each group is
 * advanced all nsteps steps (one launch per step) before the next group starts.  Members are independent,
 * so every member ends bit-identical to nsteps calls of lbm_ensemble_step; a group that stays resident in
 * the 256 MB Infinity Cache is read from HBM once instead of nsteps times.  group_bytes <= 0: one group.
 */
int lbm_ensemble_step_n(lbm_ensemble *e, int parity, int nsteps, long long group_bytes, int blocks,
                        lbm_stream_t stream);
/*
 * npairs two-step launches per cache group (even then odd step, the update of the two-step tile kernel of
 * lbm_step_fused with strip 0, bit for bit) from an even state; each launch moves every member from one of its
 * buffers into the other.  chunk = tile rows (8, 12, 16 or 24).  Every member needs f_alt.
 * lbm_ensemble_side reports which buffer holds the state (0: f, 1: f_alt); lbm_ensemble_step(_n) update that
 * buffer in place.
 */
int lbm_ensemble_pair_n(lbm_ensemble *e, int npairs, long long group_bytes, int chunk, int blocks,
                        lbm_stream_t stream);
int lbm_ensemble_side(const lbm_ensemble *e);
/* tiles of the two-step list with `chunk` rows, and its grid (one workgroup per tile) */
int lbm_ensemble_pair_info(lbm_ensemble *e, int chunk, int *total_tiles, int *default_blocks);
/*
 * Members [m0, m1) (in creation order) advanced nsteps steps with k steps per launch, so that each size class of
 * an ensemble can run with its own kernel and tile shape:
 *   k = 1  single steps in place from `parity` (chunk ignored);
 *   k = 2  the two-step kernel (chunk = tile rows 8, 12, 16 or 24) from an even state, nsteps a multiple of 2;
 *   k = 4  the four-step kernel of lbm_step_fused with strip 0 (chunk 8: 128 x 8 tiles, 12, 20 or 24: 64-wide tiles;
 *          FP32 populations) from an even state, nsteps a multiple of 4.
 * The members are grouped for the cache as in lbm_ensemble_step_n (group_bytes) and every group runs all its
 * launches before the next.  `blocks` sets the persistent grid of the k = 1 kernel (<= 0: as many workgroups as
 * are resident at once); the k = 2 and 4 kernels launch one workgroup per tile and ignore it.  The range's
 * members must hold their states on the same side (the buffer
 * lbm_ensemble_member_side reports); k = 2 and 4 launches move them between f and f_alt.  Every member ends
 * bit-identical to its solo run.  lbm_ensemble_step, _step_n and _pair_n are this on every member.
 */
int lbm_ensemble_run(lbm_ensemble *e, int m0, int m1, int k, int chunk, int nsteps, int parity,
                     long long group_bytes, int blocks, lbm_stream_t stream);
/* tiles of the k-step list with `chunk` rows (k = 1: the 128 x 2 list), and its default grid (k = 2, 4: the
 * number of tiles, one workgroup each) */
int lbm_ensemble_run_info(lbm_ensemble *e, int k, int chunk, int *total_tiles, int *default_blocks);
/* which buffer of member i holds its state (0: f, 1: f_alt); lbm_ensemble_side fails when they differ */
int lbm_ensemble_member_side(const lbm_ensemble *e, int i);
int lbm_ensemble_destroy(lbm_ensemble *e);

/*
 * Stream-capture graphs: the C-ABI form of the reference driver's CuPy capture of two fused launches
 * (lbm_smemaa_main.py:136-151, replayed at :173-174).  lbm_graph_begin starts capturing `stream` (not the default
 * stream); every liblbm call issued on it until lbm_graph_end is recorded, not run; lbm_graph_end instantiates
 * the graph and uploads it to the device (so that its first launch carries no setup) before it returns.
 * lbm_graph_launch replays the recorded sequence on any stream.  No liblbm entry point records a memset, so the graph
 * may be replayed any number of times.
 */
typedef struct lbm_graph lbm_graph;
int lbm_graph_begin(lbm_stream_t stream);
int lbm_graph_end(lbm_stream_t stream, lbm_graph **out);
int lbm_graph_launch(lbm_graph *g, lbm_stream_t stream);
int lbm_graph_destroy(lbm_graph *g);

/*
 * Learned subgrid closure (no reference counterpart: the reference's only closure is the Smagorinsky model
 * of lbm_core4.cu:171-177).  An MLP 18 -> 64 -> 64 -> 1 with GELU, evaluated and trained on the matrix
 * cores with BF16 operands and FP32 accumulation.  `params` holds the 5441 FP32 parameters in the order
 *   W1[18][64] (input-major), b1[64], W2[64][64], b2[64], W3[64], b3;
 * `packed` holds lbm_mlp_npack() BF16 values derived from them by lbm_mlp_pack (the kernels read both).
 * Samples are rows of x[n][18] (FP32, rounded to BF16 on load).
 */
int lbm_mlp_nparam(void); /* 5441 */
int lbm_mlp_npack(void);  /* 11264 */
int lbm_mlp_pack(const float *params, uint16_t *packed, lbm_stream_t stream);
int lbm_mlp_forward(const uint16_t *packed, const float *params, const float *x, int n, float *y, lbm_stream_t stream);
/* Bytes of the `work` buffer lbm_mlp_grad / lbm_mlp_train_step need for n samples. */
size_t lbm_mlp_work_size(int n);
/* grad = dL/dparams of L = sum_s wgt[s] (y_s - t[s])^2; loss (may be NULL) = L.  Deterministic. */
int lbm_mlp_grad(const uint16_t *packed, const float *params, const float *x, const float *t, const float *wgt,
                 int n, float *work, float *grad, float *loss, lbm_stream_t stream);
/* Adam: m = beta1 m + (1 - beta1) g; v = beta2 v + (1 - beta2) g^2; p -= lr (m / bc1) / (sqrt(v / bc2) + eps),
 * with the bias corrections bc1 = 1 - beta1^t, bc2 = 1 - beta2^t supplied by the caller. */
typedef struct lbm_adam {
    float lr, beta1, beta2, eps, bc1, bc2;
} lbm_adam;
/* One training step: the gradient of lbm_mlp_grad, an Adam update of params / m / v, and a repack.
 * grad and loss may be NULL. */
int lbm_mlp_train_step(uint16_t *packed, float *params, float *m, float *v, const float *x, const float *t,
                       const float *wgt, int n, const lbm_adam *adam, float *work, float *grad, float *loss,
                       lbm_stream_t stream);

/* m7neq and m8neq (lbm_core4.cu:166-173) of the pre-collision state of every cell of a periodic MRT lattice
 * into neq[2][H][W] (solid cells: 0). */
int lbm_neq(const lbm_params *p, const uint8_t *mask, const float *f, int parity, float *neq, lbm_stream_t stream);

/* Feature scaling and trust region of the closure. */
typedef struct lbm_closure {
    float in_scale[18]; /* feature 2 s + c, s = 3 (dy + 1) + (dx + 1), c = 0: m7neq, 1: m8neq, is multiplied by
                           in_scale[2 s + c] */
    float nu_scale;     /* nu_t = nu_scale * network output (lattice units) */
    float nu_max;       /* nu_t above nu_max (or NaN) falls back to Smagorinsky; negative nu_t clamps to 0 */
    float feat_max;     /* a scaled feature with |value| > feat_max falls back to Smagorinsky */
} lbm_closure;
/*
 * One in-place AA step of a periodic MRT lattice (FP32 storage) whose stress moments relax at
 * 1 / (tau_0 + 3 nu_t), nu_t from the network on the 3 x 3 stencil of (m7neq, m8neq) around each cell;
 * fallback cells use Smagorinsky with p->cs (lbm_core4.cu:171-177; plain MRT when cs = 0).  `neq` is a work
 * buffer of 2 W H floats (left holding the step's features); nut (may be NULL) receives nu_t per cell (-1 where
 * the fallback was taken); fallbacks (may be NULL) is incremented by the number of fluid fallback cells.
 */
int lbm_step_closure(const lbm_params *p, const uint8_t *mask, float *f, int parity, const uint16_t *packed,
                     const float *params, const lbm_closure *cl, float *neq, float *nut, unsigned *fallbacks,
                     lbm_stream_t stream);

/*
 * Exhaustive self-check of the short arithmetic sequences of the collision over all 2^32 FP32 inputs (no
 * reference counterpart).  what = 0: the reciprocal against the general correctly rounded 1/x; what = 1:
 * the square root against the correctly rounded sqrtf on its domain (normal x >= 1e-8).  `bad` and `first`
 * are device arrays of lbm_selftest_threads() entries: per thread, the number of mismatching inputs and
 * the first one (0xFFFFFFFF if none).
 */
int lbm_selftest(int what, uint32_t *bad, uint32_t *first, lbm_stream_t stream);
int lbm_selftest_threads(void);

/*
 * Bandwidth ceiling of the in-place single step's memory pattern (no reference counterpart: a measurement
 * kernel).  `buf` holds 9 planes of W * H FP32 values.  In the single step's tiles and XCD order, every thread
 * reads `width` (1, 2 or 4) consecutive values of each plane through one buffer resource, multiplies them by s
 * and writes them back in place.  shift != 0 moves plane q's window by the lattice link (c_qx + c_qy * W), the
 * odd step's neighbour pattern (unaligned for width > 1).  W must be a multiple of width; 36 * W * H < 2^32.
 */
int lbm_rmw_probe(float *buf, int W, int H, int width, int shift, float s, lbm_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* LBM_H */
