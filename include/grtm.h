/*
 * grtm.h -- C ABI of libgrtm.so, the MI355X (gfx950) implementation of the
 * layered-medium Green's-function path of cangyeone/grtm_cuda.
 *
 * Plain C types only: host pointers, sizes, and an opaque `void *stream`
 * (a hipStream_t, or NULL for the library's own stream).
 */
#ifndef GRTM_H_
#define GRTM_H_

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define GRTM_OK 0
#define GRTM_E_IO -1          /* cannot open / write a file */
#define GRTM_E_PARSE -2       /* malformed input.txt */
#define GRTM_E_INPUT -3       /* invalid values (sizes, NaN, depths, ...) */
#define GRTM_E_UNSUPPORTED -4 /* a configuration the selected mode cannot run */
#define GRTM_E_DEVICE -5      /* HIP runtime error or no GPU */
#define GRTM_E_NOMEM -6       /* device allocation failed */
#define GRTM_E_TOO_LARGE -7   /* run exceeds the fixed scratch of grtm_green_mem_noalloc */

#define GRTM_PREC_FP64 0 /* FP64 recursion, FP64 contraction */
#define GRTM_PREC_FP32 1 /* mixed: FP32 recursion with FP64 vertical wavenumbers,
                            FP64 low-frequency rows and FP64 re-solve of
                            non-finite cells; FP32 contraction */
#define GRTM_PREC_BF16X3 2 /* mixed recursion as GRTM_PREC_FP32; contraction on
                              split-BF16 matrix-core products, FP32 accumulation:
                              each FP32 operand as two bf16 terms, the three
                              products above 2^-16 of the leading one */
#define GRTM_PREC_BF16X6 3 /* as GRTM_PREC_BF16X3 with three bf16 terms per
                              operand and the six products above 2^-24 (FP32
                              rounding level).  Both split forms apply to trace
                              plans and batches whose wavenumber and receiver
                              counts are multiples of 4 (stream-K over the
                              pair of contractions); other shapes, and column
                              plans, contract as GRTM_PREC_FP32 (BF16X6) or on
                              the streamed BF16x3 kernel (BF16X3) */
#define GRTM_PREC_FP16X3 4 /* mixed recursion as GRTM_PREC_FP32; contraction on
                              split-FP16 matrix-core products: each frequency
                              row of the operands scaled by a power of two set
                              by its peak, two fp16 terms per value, the three
                              products above 2^-22.  Trace plans and batches
                              of the pair shape only (GRTM_E_UNSUPPORTED
                              otherwise, and for column plans) */

/* ---------------------------------------------------------- drop-in entry */
/* Replaces the reference's only exported symbol:
 *   void green(char *iname, char *oname_w, char *oname_t);
 * declared at cusrc/popm/popm.h:159-161, defined at
 * cusrc/popm/func/green.cu:1975-2281.  Same input.txt format, same output
 * files (oname_w: N rows of 30 "re,im," pairs; oname_t: N rows of 30 reals),
 * same column meaning, computed on the GPU.  Reproduces the reference's
 * behaviour including its quirks (compat mode).  Prints nothing; errors are
 * reported through grtm_last_status(). */
void green(char *iname, char *oname_w, char *oname_t);

/* Status-returning form of green(); compat = 1 reproduces the reference,
 * compat = 0 runs the corrected physics (receiver layer located from the
 * receiver depth, attenuated mu, fixed d/dr columns, ...). */
int grtm_green(const char *iname, const char *oname_w, const char *oname_t, int compat);

/* In-memory form used by bindings and tests.  spec: 30 x nw complex
 * (interleaved re,im; the post-processed spectrum rows 0..nw-1 of oname_w),
 * traces: 30 x N reals (column c of oname_t at [c*N + j]).  Either may be
 * NULL.  *nw_out receives nw (call first with both NULL to size buffers). */
int grtm_green_mem(const char *iname, int compat, double *spec, double *traces, int *nw_out);

/* grtm_green_mem without any device allocation, for callers that cannot
 * allocate (serving loops, harnesses that forbid hipMalloc).  No counterpart
 * in the reference.  All device memory of the run -- layer tables, per-row
 * k counts, the k-sum partials, the spectra and traces before their copy to
 * the host, the post-processing tables -- lives in one fixed 32 MiB array per
 * GPU (held until the process ends); a run needing more returns
 * GRTM_E_TOO_LARGE before any device work.  The partials dominate:
 * 480 nw ceil(nk_max / 256) bytes for nw frequencies and at most nk_max
 * wavenumbers per frequency (the 2048-sample demo, nw = 1024 and
 * nk_max = 10929, needs 23.8 MB in all).  Host staging uses bounded
 * stack buffers.  Runs on `stream` (NULL: the null stream) and returns when the
 * outputs are on the host.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU race.  Each call writes every scratch value
 * it reads, so nothing carries over between calls.  Outputs as
 * grtm_green_mem, bit for bit. */
int grtm_green_mem_noalloc(const char *iname, int compat, double *spec, double *traces, int *nw_out,
                           void *stream);

/* Status of the last green() call on this thread. */
int grtm_last_status(void);
const char *grtm_status_str(int status);
const char *grtm_version(void);

/* -------------------------------------------------- multi-receiver traces */
/* One layered model, one receiver depth, many epicentral distances, two
 * source kinds (explosion = isotropic moment tensor, vertical force) from
 * one recursion.  Replaces the reference's per-distance green() call loop
 * (the reference handles one distance per call, green.cu:2052) and its host
 * k-sum (green.cu:1878-1903) with a k-sum contraction on the GPU. */
typedef struct grtm_ctx grtm_ctx;
typedef struct grtm_plan grtm_plan;

typedef struct {
    int nw;          /* frequencies; N = 2 nw time samples */
    double dt;       /* sample interval (s) */
    int nk;          /* uniform wavenumbers k_i = i dk, i < nk */
    double kmax;     /* dk = kmax / nk */
    double omega_i;  /* imaginary frequency; <= 0 selects pi / T */
    double t0;       /* reduction time (s) */
    double taper;    /* fraction of the band cosine-tapered at the top */
} grtm_grid;

typedef struct {
    int n_layers;          /* rows including the half-space */
    const double *layers;  /* n_layers x 6: thickness km, vp, vs, rho, qp, qs
                              (thickness of the last row is ignored) */
    double src_depth;      /* km */
    double rcv_depth;      /* km */
    int nr;                /* number of distances */
    const double *dists;   /* km */
} grtm_job;

int grtm_create(grtm_ctx **ctx, int device);
void grtm_destroy(grtm_ctx *ctx);

/* Prepare a job: layer tables, Bessel tables and all device buffers. */
int grtm_plan_traces(grtm_ctx *ctx, const grtm_job *job, const grtm_grid *grid, int precision, int compat,
                     grtm_plan **plan);
/* Displacement columns for many receivers from one recursion: replaces the
 * per-distance integrand columns of green() (cusrc/popm/func/green.cu
 * 1558-1621 single force, 1626-1716 tensor; one distance per call at
 * green.cu:2052) on a dense uniform wavenumber grid.  grtm_plan_traces_out
 * then gives GRTM_N_COLUMNS traces per receiver, traces[(c*nr + ir)*N + j]:
 * c = 0..9 the tensor columns 0-9 (R n=2, R n=1, R n=0 from s2, R n=0 from
 * s0, T n=2, T n=1, Z n=2, Z n=1, Z n=0 from s0, Z n=0 from s2) and
 * c = 10..14 the single-force columns 0-4 (horizontal force R, vertical
 * force R, horizontal force T, horizontal force Z, vertical force Z).
 * compat = 1 keeps the reference's column contents (index slips in tensor
 * columns 2, 3, 5; its tensor SH factors; reference Bessel forms).
 * derivatives = 1 (physics mode only; compat returns GRTM_E_UNSUPPORTED)
 * adds the strain columns of the reference's physics layout
 * (green.cu:1665-1691, 1589-1600): GRTM_N_STRAIN_COLUMNS traces per receiver,
 * tensor columns 0-29 (0-9 displacement, 10-19 d/dz, 20-29 d/dr) then
 * single-force columns 0-14 (0-4 displacement, 5-9 d/dz, 10-14 d/dr).
 * grtm_plan_set_adaptive works when nk and nr are multiples of 4 (FP64,
 * FP32); grtm_plan_kernels_out returns GRTM_E_UNSUPPORTED. */
#define GRTM_N_COLUMNS 15
#define GRTM_N_STRAIN_COLUMNS 45
int grtm_plan_columns(grtm_ctx *ctx, const grtm_job *job, const grtm_grid *grid, int precision, int compat,
                      int derivatives, grtm_plan **plan);
/* Run recursion + contraction + IFFT on resident data (asynchronous). */
int grtm_plan_execute(grtm_plan *plan, void *stream);
/* Wait for the plan's work: its own stream and the stream its last
 * grtm_plan_execute ran on (other streams and plans are not waited for). */
int grtm_plan_sync(grtm_plan *plan);
/* Copy the traces out: traces[((q*2 + c)*nr + ir)*N + j], q: 0 explosion,
 * 1 vertical force; c: 0 Z, 1 R (radial).  Column plans: the layout above
 * (GRTM_N_COLUMNS or GRTM_N_STRAIN_COLUMNS, times nr * N values). */
int grtm_plan_traces_out(grtm_plan *plan, double *traces);
/* Raw k-kernels (tests): kern[((iw*nk + ik)*4 + {zr,zi,rr,ri})] for source q,
 * unweighted, for frequency rows [w_lo, w_hi). */
int grtm_plan_kernels_out(grtm_plan *plan, int q, int w_lo, int w_hi, double *kern);
/* Timing on the plan's stream with HIP events: reps executions;
 * ms[0] total per execution, ms[1] recursion, ms[2] contraction,
 * ms[3] IFFT, ms[4] Bessel tables (one-off, part of planning), ms[5] the
 * main recursion launch alone (FP32 rows of a mixed plan, all rows of an
 * FP64 plan; its cells are work[9]), ms[6] the formation of a learned grid
 * (grtm_plan_set_learned: network, cuts, item list; part of ms[1]; 0
 * otherwise).  7 doubles. */
int grtm_plan_timing(grtm_plan *plan, int reps, double *ms);
/* Work of one execution: [0] cells, [1] cell-layers, [2] contraction flops,
 * [3] IFFT algorithmic bytes, [4] traces (column plans: column traces),
 * [5] recursion flops per cell with the full upward sweep,
 * [6] transcendental calls per cell, [7] frequency rows solved in FP64,
 * [8] FP32 cells re-solved in FP64 in the last execution, [9] cells of the
 * main recursion launch, [10] their mean flops per cell with the upward
 * sweep cut where the stack below no longer reaches the source (11 doubles). */
int grtm_plan_work(grtm_plan *plan, double *work);
void grtm_plan_destroy(grtm_plan *plan);

/* Mixed-precision plans: solve frequency rows [0, rows64) of the recursion in
 * FP64 (the rest in FP32).  The plan's default is the rows with
 * |w| < 2.5 Vs_min / |z_s - z_r| (at least 4); rows64 = nw solves every row
 * in FP64 and keeps FP32 operands for the contraction.  No reference
 * counterpart (precision control of the new path). */
int grtm_plan_set_rows64(grtm_plan *plan, int rows64);

/* Peak-trough averaging (PTAM) of every k-sum's tail on a dense Z/R plan.
 * The reference reads n_ptam from input.txt (green.cu:2059) and leaves the
 * method as an unused sketch (integral.cu:38-88).  At every execution the
 * recursion also solves `tail` wavenumbers beyond the grid (same dk; 0: enough
 * for n_ptam / 2 + 2 periods of J(k r) at the nearest receiver, at most 4 nk),
 * and each sum over the grid is continued over min(tail, ceil((n_ptam/2 + 2)
 * 2 pi / (r dk))) of them and replaced by the binomially weighted mean of the
 * partial sums at its first n_ptam extrema (parabola vertices), or by the
 * extended sum where fewer appear.  2 <= n_ptam <= GRTM_PTAM_MAX; n_ptam = 0
 * removes it.  Not with adaptive grids (GRTM_E_UNSUPPORTED). */
#define GRTM_PTAM_MAX 32
int grtm_plan_set_ptam(grtm_plan *plan, int n_ptam, int tail);
/* out[0] n_ptam, out[1] tail wavenumbers, out[2] sums of the last execution
 * that took the extended sum, out[3] sums per execution (4 doubles). */
int grtm_plan_ptam_stats(grtm_plan *plan, double *out);

/* Test hook (no reference counterpart): at every execution of a
 * mixed-precision Z/R plan, give the FP32 cells (cells[2i] = frequency row in
 * [rows64, nw), cells[2i+1] = wavenumber index) non-finite operands and put
 * them on the list of cells re-solved in FP64, as an FP32 cell that overflowed
 * is.  n <= GRTM_MAX_PLANT; n = 0 removes them.  For the batch: cells[3i] =
 * model, then row and wavenumber (recaptures the batch graph). */
#define GRTM_MAX_PLANT 4096
int grtm_plan_plant_nonfinite(grtm_plan *plan, const int *cells, int n);

/* FP64 cells next to branch points and surface-wave poles (mixed-precision
 * Z/R plans).  At every execution, cells of the FP32 rows within
 * branch_cells wavenumber cells of Re(w/alpha) and Re(w/beta) of every layer
 * (the branch points of the vertical wavenumbers, green.cu:184-194), and
 * within pole_widths half widths of each pole cluster the plan's network
 * predicts for the row (learned plans only; at most 64 cells either side),
 * below the row's cut, are re-solved in FP64 after the FP32 launch, as
 * non-finite FP32 cells are.  branch_cells < 0 and pole_widths = 0 turn the
 * two parts off (the default).  branch_cells <= GRTM_MAX_FLAG_CELLS,
 * 0 <= pole_widths <= 8.  The reference solves every cell in FP64; it has no
 * counterpart. */
#define GRTM_MAX_FLAG_CELLS 64
int grtm_plan_set_flag64(grtm_plan *plan, int branch_cells, double pole_widths);
/* Cells (row * nk + k) re-solved in FP64 in the last execution, up to cap
 * of them into cells; returns their number (a status < 0 on error). */
int grtm_plan_flag_list(grtm_plan *plan, unsigned *cells, int cap);

/* Adaptive wavenumber grid (no counterpart in the reference, whose ragged
 * grid uses the cal_kmax heuristic, green.cu:108-162).  kcut[iw] (km^-1) is
 * the truncation wavenumber of each frequency row (from the learned
 * predictor or any other estimate): the recursion solves k_i < kcut only.  A
 * residual check solves every `stride`-th wavenumber in a window of `window`
 * km^-1 beyond the cut; a row whose checked tail exceeds tol times the row's
 * peak is completed on the dense grid in FP64 (fallback).  kcut = NULL
 * returns the plan to the dense grid. */
int grtm_plan_set_adaptive(grtm_plan *plan, const double *kcut, double window, int stride, double tol);
/* Training labels of every frequency row after a dense execution: out[26*iw
 * + 0] peak |kernel|, [1] 1 + last wavenumber index whose |kernel| exceeds
 * tau x peak, [2+3i, 3+3i, 4+3i] index, half width (samples) and relative
 * height of the i-th largest local maximum (index -1 when absent), i < 8. */
int grtm_plan_row_labels(grtm_plan *plan, double tau, float *out);
/* After an execution: out[0] cells of the predicted grid, out[1] residual-check
 * cells, out[2] cells after fallback, out[3] rows that fell back. */
int grtm_plan_adaptive_stats(grtm_plan *plan, double *out);
/* The current grid per frequency row (a learned grid: as of the last
 * execution): cut[iw] wavenumbers solved before any fallback, checks[iw]
 * residual-check cells; the dense grid gives nk and 0. */
int grtm_plan_adaptive_rows(grtm_plan *plan, int *cut, int *checks);

/* ------------------------------------------------------------ batches
 * Many layered models (uneven layer counts and source depths) on one grid
 * and one set of surface-receiver distances: the production workload of
 * Green's-function databases.  No counterpart in the reference, which runs
 * one model and one distance per green() call.  Models are processed in
 * groups sharing device buffers: one recursion launch per group over
 * (model, frequency row, k tile) work items, ordered heaviest model first,
 * one contraction over the group's stacked k-kernels, one inverse FFT; the
 * whole sequence is captured once in a HIP graph and replayed by
 * grtm_batch_execute.  With GRTM_BATCH_AUTOTUNE the recursion tile (64, 128
 * or 256 wavenumbers per block) and the contraction's K slice are chosen by
 * timing the candidates on the first group. */
typedef struct grtm_batch grtm_batch;
typedef struct {
    int n_layers;          /* rows including the half-space */
    const double *layers;  /* n_layers x 6, as in grtm_job */
    double src_depth;      /* km */
    double rcv_depth;      /* km */
} grtm_model;
#define GRTM_BATCH_AUTOTUNE 1
int grtm_batch_create(grtm_ctx *ctx, const grtm_model *models, int n_models, const double *dists, int nr,
                      const grtm_grid *grid, int precision, int compat, int flags, grtm_batch **batch);
/* Adaptive wavenumber grids for every model (as grtm_plan_set_adaptive):
 * kcut[m * nw + iw] km^-1, window[m] km^-1, one stride and tolerance for
 * all; kcut = NULL returns to the dense grid.  The graph is captured again. */
int grtm_batch_set_adaptive(grtm_batch *batch, const double *kcut, const double *window, int stride, double tol);
/* Summed over the models after an execution: out[0] cells of the predicted
 * grids, out[1] residual-check cells, out[2] cells after fallback, out[3]
 * rows that fell back. */
int grtm_batch_adaptive_stats(grtm_batch *batch, double *out);
/* Replay the captured graph on `stream` (NULL: the batch's stream). */
int grtm_batch_execute(grtm_batch *batch, void *stream);
/* The test hook of grtm_plan_plant_nonfinite for a mixed-precision batch. */
int grtm_batch_plant_nonfinite(grtm_batch *batch, const int *cells, int n);
/* The same launches issued one by one on the batch's stream, no graph. */
int grtm_batch_execute_eager(grtm_batch *batch);
int grtm_batch_sync(grtm_batch *batch);
/* Traces of models [m0, m1): traces[(((m - m0)*2 + q)*2 + c)*nr + ir)*N + j]. */
int grtm_batch_traces_out(grtm_batch *batch, int m0, int m1, double *traces);
/* ms[0]: mean of reps graph replays (HIP events on the batch's stream);
 * ms[1..3]: recursion, contraction and inverse FFT of one eager pass (sums
 * over the groups); ms[4]: the learned grids' formation in that pass
 * (grtm_batch_set_learned; part of ms[1], 0 otherwise).  5 doubles. */
int grtm_batch_timing(grtm_batch *batch, int reps, double *ms);
/* out[0] models per group, [1] recursion tile (wavenumbers per block),
 * [2] contraction K slice (0: split-K plan of the FP64 / BF16x3 kernels),
 * [3] groups, [4] graph nodes, [5] autotune candidates timed, [6] autotune ms,
 * [7] items a persistent recursion block takes at a time (learned grids
 * formed on the device), [8] 1 for the resident layout, [9] rows whose extent
 * it had to cut since it was set up (0 unless a later pass fell back on rows
 * the set-up pass did not).  10 doubles. */
int grtm_batch_config(grtm_batch *batch, double *out);
/* out[0] cells, [1] cell-layers, [2] contraction flops, [3] traces,
 * [4] recursion flops of the dense grid with full upward sweeps (operation
 * count of the cell arithmetic), [5] recursion flops of the cells the
 * current grid solves with the sweeps cut by depth (estimated on every 8th
 * frequency row and 16th wavenumber) (6 doubles). */
int grtm_batch_work(grtm_batch *batch, double *out);
void grtm_batch_destroy(grtm_batch *batch);

/* ------------------------------------------------------------ self-tests */
/* Host-compiled copies of the device arithmetic, for unit checks without a
 * GPU.  Not used by any entry point above. */
/* lay: n_layer rows of (h, alpha re, im, beta re, im, mu re, im); out as
 * the oracle's cell: a0[2],a1[2],a2[2],b1,b2 and d/dz versions, 18 complex. */
/* C[M x N] = A[M x K] B[K x N] (row-major host arrays) on the GPU with the
 * contraction kernels: mode -1 VALU FP64, 0 FP64 MFMA, 1 FP32 MFMA,
 * 2 BF16x3 MFMA, 3 streamed FP32 MFMA pair kernel, 4 streamed FP64 MFMA
 * pair kernel (modes 3, 4: both of the pair's problems solved on (A, B) and
 * required to agree bit for bit); operands rounded to FP32 first for modes
 * 1-3. */
int grtm_selftest_gemm(int mode, int M, int N, int K, const double *A, const double *B, double *C);
/* The learned-grid contraction pair (prec 1: FP32 stream-K, 0: FP64 split-K,
 * 3 / 6: split-BF16 stream-K with that many products, on FP32 operands)
 * on (A, B) twice: row r of A is read up to column nk_w[r / 4] (M / 4
 * entries, four operand rows per frequency row), the tile extents formed on
 * the GPU as the plans form them; the halves must agree bit for bit. */
int grtm_selftest_gemm_cut(int prec, int M, int N, int K, const double *A, const double *B, const int *nk_w,
                           double *C);

/* Lane-layout probe of the FP64 matrix-core instruction: lane l0 feeds a = 1,
 * lane l feeds b = 1000 + l; out[64*4] receives every lane's accumulators. */
int grtm_selftest_mfma_probe(int l0, double *out);
/* One plain device copy of n_int32 4-byte words (k_copy_i32): the known-size
 * kernel that calibrates hardware byte counters against algorithmic bytes.
 * No reference counterpart (measurement hook). */
int grtm_selftest_copy(int n_int32);

int grtm_selftest_cell(int compat, int src_type, double k, double wr, double wi, int n_layer,
                       const double *lay, int s_layer, double zs, int r_layer, double zr,
                       const double *par0_rv, double *out);
/* The same cell solved in double-double arithmetic (the precision of the
 * drop-in path's lowest frequency rows).  No reference counterpart (test
 * hook). */
int grtm_selftest_cell_dd(int compat, int src_type, double k, double wr, double wi, int n_layer,
                          const double *lay, int s_layer, double zs, int r_layer, double zr,
                          const double *par0_rv, double *out);

/* Exact operation count of one recursion cell of the trace path with the
 * full upward sweep (the device template run on a counting scalar type on
 * the host):
 * out[0] = add/sub/mul/div/sqrt, out[1] = transcendental calls
 * (exp, sincos, log, atan2, hypot).  branch/up as the geometry selects
 * (1: receiver layer above the source layer, 2/3: same layer above/below,
 * 4: below). */
int grtm_opcount_cell(int n_layers, int s_layer, int r_layer, int branch, int up, double *out);
/* The same for a cell of the column plans (recursion + the coefficient rows
 * of both sources; derivatives: with the strain rows). */
int grtm_opcount_columns_cell(int n_layers, int s_layer, int r_layer, int branch, int up, int derivatives,
                              double *out);

/* ------------------------------------------- learned wavenumber predictor */
/* Multilayer perceptron n_in -> depth x width (ReLU) -> n_out, trained and
 * evaluated on the GPU: BF16 weights and activations, FP32 accumulation,
 * FP32 master weights and Adam.  No counterpart in the reference: it predicts
 * the per-frequency truncation wavenumber and pole clusters that drive the
 * adaptive grid (grtm_plan_set_adaptive).  Parameters are laid out layer by
 * layer as W_l [fan_in x fan_out] row-major followed by b_l. */
typedef struct grtm_mlp grtm_mlp;
int grtm_mlp_create(int n_in, int width, int depth, int n_out, int max_batch, grtm_mlp **mlp);
void grtm_mlp_destroy(grtm_mlp *mlp);
long grtm_mlp_nparams(grtm_mlp *mlp);
int grtm_mlp_set_params(grtm_mlp *mlp, const float *params);
int grtm_mlp_get_params(grtm_mlp *mlp, float *params);
/* Loss: the first n_quantile outputs use the pinball loss of quantile q
 * (a per-row conservative bound), the others the masked squared error. */
int grtm_mlp_set_loss(grtm_mlp *mlp, int n_quantile, float q);
/* Resident training set: X [n x n_in], targets T and mask M [n x n_out]. */
int grtm_mlp_set_data(grtm_mlp *mlp, const float *X, const float *T, const float *M, int n);
/* `steps` Adam steps on random batches (rows from a splitmix64 stream seeded
 * by `seed`); loss of the masked mean squared error per step in losses. */
int grtm_mlp_train(grtm_mlp *mlp, int steps, int batch, float lr, unsigned long long seed, float *losses);
/* Gradient of the loss on n <= max_batch samples (params layout). */
int grtm_mlp_grads(grtm_mlp *mlp, const float *X, const float *T, const float *M, int n, float *grads, float *loss);
int grtm_mlp_predict(grtm_mlp *mlp, const float *X, int n, float *Y);
/* The same outputs from the training forward (VALU tiles) instead of the
 * fused matrix-core forward grtm_mlp_predict uses: test hook. */
int grtm_mlp_predict_train_forward(grtm_mlp *mlp, const float *X, int n, float *Y);

/* Learned adaptive grid formed on the GPU at every grtm_plan_execute, inside
 * the execution (no counterpart in the reference, whose ragged grid uses the
 * cal_kmax heuristic, green.cu:108-162): the network's features of the plan's
 * model at every frequency row, one fused matrix-core forward, the cut
 * k = exp(y0 + lmargin) (|w| / Vs_min + 1 / max(|z_s - z_r|, 0.5)) (at least
 * the analytic cut 1.15 w / Vs_min + ln(1e7) / dz on the rows solved in
 * FP64, at most kmax), the residual check of grtm_plan_set_adaptive (window,
 * stride, tol) extended to reach every predicted pole cluster (outputs 1-16:
 * centre / s, log(half width * 100)) that starts beyond the cut plus window,
 * and the recursion's item list, taken by persistent launches.  A network of
 * 17 outputs; Z / R plans with the streamed contraction (FP64, FP32; nk and nr
 * multiples of 4), models of at most 30 layers.  The plan keeps `net`, which must
 * outlive it or be detached with net = NULL (which returns the plan to the
 * dense grid). */
int grtm_plan_set_learned(grtm_plan *plan, grtm_mlp *net, double lmargin, double window, int stride, double tol);
/* The same for every model of a batch, inside its captured graph: per group,
 * one forward of the network over the group's (model, frequency) rows, the
 * cuts and check counts, the group's item list, persistent recursion
 * launches.  window[m] per model; batches with the streamed contraction
 * (FP64, FP32).  The graph is captured again.  An FP32 batch then takes the
 * resident layout when it fits the device: one group of every model, each
 * frequency row's operands stored up to the extent the set-up pass reached
 * (its cut, or nk where it fell back), one recursion launch per precision pass
 * for every receiver branch, one contraction, one inverse FFT. */
int grtm_batch_set_learned(grtm_batch *batch, grtm_mlp *net, double lmargin, const double *window, int stride,
                           double tol);

#ifdef __cplusplus
}
#endif
#endif
