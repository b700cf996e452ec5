// kernels.h -- launch interface between the host driver and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "cell.h"

struct grtm_mlp;  // the predictor network (mlp.hip; C ABI in grtm.h)

namespace grtm {

// Operation count of one cell whose upward sweep ends at layer bot (the depth
// truncation's choice; L - 1: none; -1: the criterion's choice), as
// grtm_opcount_cell (kind 0) and grtm_opcount_columns_cell (kind 1; 2 with
// the strain rows); top: the receiver on its layer's top, as a receiver on
// the free surface (the public counters take a buried receiver).
int opcount_cell_bot(int kind, int L, int s, int rl, int branch, int up, int bot, double *out, bool top = false);

constexpr int kMaxLayers = 64;

// Spectral post-processing (cusrc/popm/func/green.cu:1905-1921) and time
// scaling (green.cu:2236-2237) parameters.
struct PostParams {
    int nw;          // number of frequencies (N = 2 nw)
    int wc;          // taper start index (int)(nw (1 - taper))
    double df;       // frequency step
    double dt;       // time step
    double t0;       // reduction time
    double omega_i;  // imaginary frequency
    // device tables in the trace precision (FP64 or FP32), computed in FP64:
    // pv[iw] = e^{2 pi i df iw t0} taper(iw) df nw / pi (complex, nw),
    // tw[m] = e^{2 pi i m / N} (complex, nw), ts[j] = e^{w_I (j dt + t0)} / N (N)
    const void *pv, *tw, *ts;
};

// Displacement columns (tensor columns 0-9, single-force columns 0-4) and,
// in physics mode, their d/dz and d/dr (tensor 10-29, single force 5-14):
// operands of the contractions against J0..J4 (rows of mt_rows_from_cell),
// and each column's terms over their outputs.
constexpr int kMtColsMax = 45;
constexpr int kMtTermsMax = 4;
struct MtOps {
    void *A[kMtSetsMax];  // A_s: [2 rows[s] nw][lda], FP64 or FP32 as the launch
    int rows[kMtSetsMax]; // complex rows per frequency
};
struct MtTerm {
    int set, row;   // contraction and complex row within a frequency
    int rpow, pad;  // divided by the receiver distance to this power
    double scale;
};
// Lives in device memory (larger than a kernel's argument block).
struct MtCols {
    const void *C[kMtSetsMax];  // contraction outputs [2 rows[s] nw][ldc]
    int rows[kMtSetsMax];
    int nterm[kMtColsMax];
    MtTerm term[kMtColsMax][kMtTermsMax];
};
// prec: 0 = FP64 cells / FP64 operands, 1 = FP32 cells / FP32 operands
// (non-finite cells appended to flag_list), 2 = FP64 cells / FP32 operands.
// Work items are (row, tile of 256 wavenumbers), one block each: the list
// `tiles` (ntiles entries cells_tile(row, tile)), or, when tiles is null,
// every tile of rows [w0, w0 + nwr).  An item cells_tile(row, c) |
// kCheckItem solves the residual-check cells 256 c .. 256 c + 255 of the row
// (wavenumbers nk_w[row] + stride j + stride / 2, j < nchk[row]; largest
// |value| to tail_max[row], 3e38 if not finite) instead of storing operands.
int cells_tiles_per_row(int nk);
unsigned cells_tile(int iw, int kb);
constexpr unsigned kCheckItem = 1u << 31;
int cells_tile_limit(int nw, int nk);  // 1 if (nw, nk) fit the item encoding
// dev_count (optional): the item list was built on the device and its length
// is only known there -- items [dev_count[which], dev_count[which + 1]) of
// `tiles`, taken by a persistent grid from the counter *ctr (zeroed by the
// launcher); ntiles is then an upper bound on the item count.
// k_base: the rows' cells are the wavenumbers k_base .. k_base + nk - 1, stored
// from column 0 of A (a tail beyond a grid of k_base wavenumbers; dense
// launches without flags only).
hipError_t launch_cells_zr(int prec, int compat, const void *lw, int L, const Geo &g, int nk, double dk,
                           double df, double omega_i, int w0, int nwr, const unsigned *tiles, int ntiles, void *A0,
                           void *A1, long lda, unsigned *flag_count, unsigned *flag_list, const int *nk_w,
                           unsigned *row_peak, hipStream_t st, const int *nchk = nullptr, int stride = 0,
                           unsigned *tail_max = nullptr, const unsigned *dev_count = nullptr, int which = 0,
                           unsigned *ctr = nullptr, int k_base = 0, bool ctr_zeroed = false);

// ---- learned wavenumber grid built on the device
// The predictor's inputs for one model and the decoding of its cut output
// (the rows of learned.py: features, decode_kcut, the low-frequency floor).
constexpr int kLearnSlots = 30;
struct LearnModel {
    double rows[kLearnSlots][6];  // thickness km, vp, vs, rho, qp, qs (first n rows)
    int n;                        // layer rows, <= kLearnSlots
    int floor_rows;               // rows [0, floor_rows) keep at least the analytic cut
    double src, rcv;              // source / receiver depth km
    double vs_min;                // slowest shear velocity of the model
    double window;                // residual-check window km^-1
};
struct LearnGrid {
    int nw, nk, stride;
    // 1: rows [0, floor_rows) of every model keep the cut and check count the
    // host stored in nk_pred / nchk (the network's outputs for them unused)
    int low_fixed;
    double dt, dk, kmax, lmargin;
};
// Per-row cut nk_pred[m nw + iw] and residual-check count nchk[...] of
// n_models models from one execution of the network (learned.hip / mlp.hip).
// poles (optional): each row's 8 predicted pole clusters as (centre, half
// width) in km^-1, [n_models nw][16] floats.
int learned_grid(grtm_mlp *net, const LearnModel *models, int n_models, const LearnGrid &lg, int *nk_pred,
                 int *nchk, hipStream_t st, float *poles = nullptr);
// The state an execution of a learned grid starts from, set by the item
// kernels instead of launches of their own: up to kGridZero scalar counters
// cleared, and per row nk_w = nk_pred, row peak and checked-tail maximum zero
// (null pointers are skipped).
constexpr int kGridZero = 4;
struct GridInit {
    unsigned *zero[kGridZero];
    int *nk_w;
    unsigned *row_peak, *tail_max;
};
// Recursion items of one model's learned grid from nk_pred / nchk: tiles of
// rows [0, r64), their residual-check items, tiles of rows [r64, nw), theirs;
// counts = {0, items of rows [0, r64), all items}.  off: 2 nw scratch words
// (each row's first tile and first check item).  low_apart: rows [0, r64)
// get no items and gi leaves their state alone (their launch has a list and
// state of its own), counts = {0, 0, items of rows [r64, nw)}.
hipError_t launch_grid_items(int nw, int r64, const int *nk_pred, const int *nchk, unsigned *items, unsigned *off,
                             unsigned *counts, hipStream_t st, const GridInit &gi, bool low_apart = false);
// Rows [0, n) start an execution: nk_w = nk_src, row peak and checked-tail
// maximum zero.
hipError_t launch_rows_init(int n, const int *nk_src, int *nk_w, unsigned *row_peak, unsigned *tail_max,
                            hipStream_t st);
// adaptive grid: dense fallback, tile extents
// Rows whose checked tail exceeds tol x row peak are completed on the dense
// grid: FP64 plans in FP64; mixed plans in FP64 below row r64 and in FP32
// (lw32) above it, non-finite FP32 cells appended to the flag list as
// (row_base + row) nk + k.  With kend, the same launch that selects the rows
// also forms the K extents of contraction tiles of rows_per_tile (<= 64) rows
// from the final counts, as launch_kend_tiles.
hipError_t launch_fallback(int out_f64, int compat, const void *lw64, const void *lw32, int r64, int L, const Geo &g,
                           int nk, int nw, double dk, double df, double omega_i, double tol,
                           unsigned *row_peak, const unsigned *tail_max, int *nk_w, unsigned *fb_count,
                           unsigned *fb_list, void *A0, void *A1, long lda, unsigned *flag_count,
                           unsigned *flag_list, int row_base, hipStream_t st, const MtOps *mto = nullptr,
                           int deriv = 0, int *kend = nullptr, int rows_per_tile = 0);
hipError_t launch_kend_tiles(int nw, int rows_per_tile, int nk, const int *nk_w, int *kend, hipStream_t st);
// extents of the bm-row tiles of n contractions with rpw[q] rows per frequency
// row, laid end to end in kend
hipError_t launch_kend_sets(int nw, int nk, int n, const int *rpw, int bm, const int *nk_w, int *kend, hipStream_t st);
hipError_t launch_copy_i32(const int *a, int *b, int n, hipStream_t st);
// per-row training labels (peak, truncation index, 8 local maxima); out: nw x 26 floats
hipError_t launch_row_labels(int prec, int nw, int nk, long lda, const void *A0, const void *A1, double tau,
                             float *out, hipStream_t st);
// Test hook: non-finite FP32 operands in the listed cells (row * nk + k), each
// appended to the flag list.
hipError_t launch_plant_nonfinite(const unsigned *cells, int n, int nk, void *A0, void *A1, long lda,
                                  unsigned *flag_count, unsigned *flag_list, hipStream_t st);
// FP64 cells next to branch points and predicted surface-wave poles
// (grtm_plan_set_flag64): per row of [r64, nw), the cells within hb of
// Re(w/alpha), Re(w/beta) of every layer (hb < 0: none) and within
// pole_widths half widths of each cluster in poles ([nw][8][centre, half
// width] km^-1; null: none), below nk_w[iw], join the flag list once each.
hipError_t launch_flag_cells(const void *lw64, int L, int r64, int nw, int nk, double dk, int hb, const float *poles,
                             float pole_widths, const int *nk_w, unsigned *flag_count, unsigned *flag_list,
                             hipStream_t st);
// row_peak (optional): per-row peaks raised to cover the re-solved values
hipError_t launch_cells_fix(int compat, const void *lw64, int L, const Geo &g, int nk, double dk, double df,
                            double omega_i, void *A0, void *A1, long lda, const unsigned *flag_count,
                            const unsigned *flag_list, hipStream_t st, unsigned *row_peak = nullptr);
hipError_t launch_zero_u32(unsigned *p, int n, hipStream_t st);
// drop-in cell kernel (green.hip) over rows [w_base, w_base + nrows); dd_rows:
// solve them in double-double arithmetic
hipError_t launch_cells_full(int compat, int src, int dd_rows, const void *lw, const int *nk_w, int L,
                             const Geo &g, int w_base, int nrows, int nblk, double dk, double dist, double df,
                             double omega_i, const double p0[4], double *partial, hipStream_t st);
hipError_t launch_green_finish(int nw, int nblk, const double *partial, double *raw, hipStream_t st);
// B0 = J_n0(k r), B1 = J_n0+1(k r)
hipError_t launch_bessel_tab(int prec, int mode, int nk, double dk, const double *dists, int nr, void *B0,
                             void *B1, long ldb, hipStream_t st, int n0 = 0, int k0 = 0);
// Peak-trough averaging of the k-sums' tails (Z/R plans): for every frequency
// row, source, component and receiver, the contraction's sum C (FP64 or FP32
// per prec, [4 nw][nr]) is continued with the tail terms At[row][j] Bt[j][r]
// (At: [4 nw][nt] weighted kernels of the wavenumbers beyond the grid, Bt:
// [nt][nr] J0 / J1 there) over min(nt, ceil((n/2 + 2) 2 pi / (r dk))) terms,
// and replaced by the binomially weighted mean of its first n extrema
// (parabola vertices), or by the extended sum if fewer appear; misses counts
// the sums of the latter kind.
hipError_t launch_ptam(int prec, int nw, int nr, int nt, int n, double dk, const double *dists, const void *At0,
                       const void *At1, const void *Bt0, const void *Bt1, void *C0, void *C1, unsigned *misses,
                       hipStream_t st);
hipError_t launch_gemm(int prec, int M, int N, int K, const void *A, long lda, const void *B, long ldb, void *Cm,
                       long ldc, hipStream_t st);
size_t post_ifft_lds_bytes(int prec, int nw, int R);
// k-sum contraction on the matrix cores; mode 0 FP64, 1 FP32, 2 BF16x3
size_t gemm_workspace_bytes(int mode, int M, int N, int K);
hipError_t launch_probe_f64(int l0, double *out, hipStream_t st);
// nk_w (optional): A row r is zero at and beyond column nk_w[r / 4]
hipError_t launch_gemm_mfma(int mode, int M, int N, int K, const void *A, long lda, const void *B, long ldb,
                            void *Cm, void *P, const int *kend, const int *nk_w, hipStream_t st);
// Both contractions C_p = A_p B_p (p = 0, 1) in FP32 on the matrix cores,
// streamed kernel: K slices of kc (a multiple of 32; gemm_pair_slice(1, ...)
// picks one), workspace P of gemm_pair_workspace_bytes, kend per 128 operand
// rows.  mode: 1 FP32, 0 FP64.
int gemm_pair_slice(int mode, int M, int N, int K, double k_avg);
size_t gemm_pair_workspace_bytes(int M, int N, int K, int kc);
bool gemm_pair_ok(long lda, long ldb, const void *A0, const void *A1, const void *B0, const void *B1);
// foff / fext (optional, with nk_w): A rows in the resident batch layout
// (RowLayout: row r at foff[r / 4] + (r % 4) fext[r / 4]) instead of r lda.
hipError_t launch_gemm_pair_f32(int M, int N, int K, int kc, const float *A0, const float *A1, long lda,
                                const float *B0, const float *B1, long ldb, float *C0, float *C1, float *P,
                                const int *kend, const int *nk_w, hipStream_t st, const long long *foff = nullptr,
                                const int *fext = nullptr);
// Split-precision pair: the contractions of launch_gemm_pair_f32 on 16-bit
// matrix cores, A split while it is staged, B given as split planes S0 / S1
// (launch_split_table: planes of N ldk values, ldk a multiple of 32 >= K);
// stream-K for any kend / nk_w (null: dense).  nprod (the split code):
// kSplitBf16x3 two bf16 terms, three products (about 2^-16 per product);
// kSplitBf16x6 three bf16 terms, six products (about 2^-24); kSplitF16x3 two
// fp16 terms, three products (about 2^-22), each frequency row of A scaled by
// a power of two set by its peak |A| (rpk: per frequency row, float bits; the
// rows' largest |A| over every value they hold) and B (|B| < 8) by 2^13.  Workspace:
// gemm_bsplit_workspace_bytes(M, N, K).
constexpr int kSplitBf16x3 = 3, kSplitBf16x6 = 6, kSplitF16x3 = 16;
inline int split_planes(int nprod) { return nprod == kSplitBf16x6 ? 3 : 2; }
size_t gemm_bsplit_workspace_bytes(int M, int N, int K);
hipError_t launch_split_table(int nprod, int K, int N, const float *B, long ldb, void *S, long ldk, hipStream_t st);
hipError_t launch_gemm_pair_bsplit(int nprod, int M, int N, int K, const float *A0, const float *A1, long lda,
                                   const void *S0, const void *S1, long ldk, float *C0, float *C1, float *P,
                                   const int *kend, const int *nk_w, hipStream_t st, const long long *foff = nullptr,
                                   const int *fext = nullptr, const unsigned *rpk = nullptr);
// The same in FP64 (128 x 128 tiles, kend per 128-row tile; workspace
// 2 x gemm_pair_workspace_bytes(M, N, K, kc)).
hipError_t launch_gemm_pair_f64(int M, int N, int K, int kc, const double *A0, const double *A1, long lda,
                                const double *B0, const double *B1, long ldb, double *C0, double *C1, double *P,
                                const int *kend, const int *nk_w, hipStream_t st);
// n <= 5 contractions C_q = A_q B_q (M[q] rows, rpw[q] rows per frequency
// row of nk_w; shared N, K, strides) in one streamed launch; prec 0 FP64,
// 1 FP32; the same alignment rules as the pairs; workspace P of
// gemm_set_workspace_bytes; kend (optional) per contraction and 128-row tile.
size_t gemm_set_workspace_bytes(int n, const int *M, int N, int K, int kc, size_t es);
// FP32 sets and pairs with kend run stream-K: workspace of at least this size
// (the pair: n = 2, M = {M, M}).
size_t gemm_streamk_workspace_bytes(int n, const int *M, int N, int K);
int gemm_set_slice(int mode, int n, const int *M, int N, int K);
hipError_t launch_gemm_set(int prec, int n, const int *M, const int *rpw, int N, int K, int kc, const void *const *A,
                           long lda, const void *const *B, long ldb, void *const *C, void *P,
                           const int *const *kend, const int *nk_w, hipStream_t st);
// layout 1 with several models: trace t reads model t / (4 nr), whose
// contraction outputs start cms elements further per model
hipError_t launch_post_ifft(int prec, const PostParams &pp, int layout, int ntr, const void *raw, const void *C0,
                            const void *C1, long ldc, int nr, void *traces, void *spec_post, hipStream_t st,
                            long cms = 0);

// prec and work items as launch_cells_zr; deriv: the strain rows as well
// (physics mode only)
hipError_t launch_cells_mt(int prec, int compat, int deriv, const void *lw, int L, const Geo &g, int nk, double dk,
                           double df, double omega_i, int w0, int nwr, const unsigned *tiles, int ntiles,
                           const MtOps &o, long lda, unsigned *flag_count, unsigned *flag_list, const int *nk_w,
                           unsigned *row_peak, hipStream_t st, const int *nchk = nullptr, int stride = 0,
                           unsigned *tail_max = nullptr);
hipError_t launch_cells_mt_fix(int compat, int deriv, const void *lw64, int L, const Geo &g, int nk, double dk,
                               double df, double omega_i, const MtOps &o, long lda, const unsigned *flag_count,
                               const unsigned *flag_list, hipStream_t st);
// traces [ncol][nr][2 nw]; mc in device memory
hipError_t launch_post_ifft_cols(int prec, const PostParams &pp, const MtCols *mc, int ncol, long ldc, int nr,
                                 const double *dists, void *traces, hipStream_t st);

// One model of a batch on the device: geometry and the offset (in LayerW
// entries) of its [nw x L] layer table in the batch table.
struct ModelDesc {
    Geo g;
    long lw_off;
    long r64;  // frequency rows [0, r64) solved in FP64 (all rows of an FP64 batch)
};
// Where the A operand rows of a batch group live.  Dense: group frequency row
// g has its four component rows at (4 g + i) lda.  Resident (foff set): row g
// has them at foff[g] + i fext[g], fext[g] a multiple of 4 wavenumbers; cells
// at or past fext[g] are not stored.
struct RowLayout {
    const long long *foff;
    const int *fext;
    long lda;
};
// Flag list of a batch's non-finite FP32 cells: *count entries (those below
// cap stored) of (group row) nk + k.
struct FlagList {
    unsigned *count;
    unsigned long long *list;
    unsigned cap;
};
// Recursion of a group of models; items (j << 32) | cells_tile(row, tile)
// with tiles of bs in {64, 128, 256} wavenumbers; prec as launch_cells_zr;
// every item's model has receiver branch br (1-4), or br = 0: each item's own
// (the kernel holds every branch's solve).
// dev_count: persistent form over items [dev_count[which], dev_count[which + 1])
// of a list built on the device (nitems: an upper bound), taken `chunk` at a
// time from counter *ctr.
hipError_t launch_cells_zr_batch(int prec, int compat, int bs, int br, const ModelDesc *md, int m0, const void *lw_all,
                                 int nw, int nk, double dk, double df, double omega_i,
                                 const unsigned long long *items, int nitems, void *A0, void *A1, RowLayout rl,
                                 FlagList fl, const int *nk_w, unsigned *row_peak, hipStream_t st,
                                 const int *nchk = nullptr, int stride = 0, unsigned *tail_max = nullptr,
                                 const unsigned *dev_count = nullptr, int which = 0, unsigned *ctr = nullptr,
                                 int chunk = 1);
// A group's learned-grid items from its slot order (batch.hip k_batch_count,
// k_batch_scan_sums, k_batch_offsets, k_batch_fill): slots[q] = group row |
// kCheckItem for its check items; segment s covers slots [seg_slots[s],
// seg_slots[s + 1]) and gets items [seg_items[s], seg_items[s + 1]);
// slot_off: batch_items_words(nslots) scratch words.
size_t batch_items_words(int nslots);
hipError_t launch_batch_items(int nslots, const unsigned *slots, int nw, int bs, const int *nk_pred, const int *nchk,
                              unsigned *slot_off, int nseg, const unsigned *seg_slots, unsigned *seg_items,
                              unsigned long long *items, hipStream_t st);
// Dense fallback of a group's rows whose residual check failed (learned
// grids): one selection over the group's nrow rows (list: group rows, then
// their first wavenumbers; count: the group's counter; model_count: per
// model), then the completion of the listed rows, FP64 below each model's r64
// and in prec (0 FP64, else FP32) above.
// raise_peaks: the completed cells raise their rows' peaks (the FP16 split's
// row scales)
hipError_t launch_fallback_batch(int prec, int compat, const ModelDesc *md, int m0, const void *lw64, const void *lw32,
                                 int nrow, int nw, int nk, double dk, double df, double omega_i, double tol,
                                 unsigned *row_peak, const unsigned *tail_max, int *nk_w, unsigned *count,
                                 unsigned *list, unsigned *model_count, void *A0, void *A1, RowLayout rl,
                                 FlagList fl, hipStream_t st, bool raise_peaks = false);
// row_peak (optional, group rows): raised to cover the re-solved values
hipError_t launch_cells_fix_batch(int compat, const ModelDesc *md, int m0, const void *lw64, int nw, int nk, double dk,
                                  double df, double omega_i, void *A0, void *A1, RowLayout rl, FlagList fl,
                                  hipStream_t st, unsigned *row_peak = nullptr);
// Test hook: non-finite FP32 operands in the listed cells (group row nk + k),
// each appended to the flag list.
hipError_t launch_plant_nonfinite_batch(const unsigned long long *cells, int n, int nk, void *A0, void *A1,
                                        RowLayout rl, FlagList fl, hipStream_t st);
// Resident layout: a row's final extent beyond its stored length (a fallback
// the layout was not sized for) is cut to the stored length and counted in
// *over.
hipError_t launch_clamp_extent(int nrow, int *nk_w, const int *fext, unsigned *over, hipStream_t st);

}  // namespace grtm
