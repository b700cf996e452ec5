/*
 * grtm_oracle.h -- FP64 CPU restatement of the reference's Green's-function
 * path (cusrc/popm/func/green.cu).  TEST INFRASTRUCTURE ONLY: this library is
 * the checker for the HIP implementation.  Only tests/, the smoke check and the
 * CPU-baseline leg of the benchmark may load it; no product code links it.
 *
 * Two modes:
 *   compat  = 1  reproduce the reference's observable behaviour, including its
 *                quirks (receiver search keyed on the source depth,
 *                real unattenuated mu, "J''" formula, index slips in the
 *                tensor columns, ...), so that results can be compared with the
 *                reference's own output files.
 *   compat  = 0  "physics" mode: the same algorithm with those quirks fixed.
 */
#ifndef GRTM_ORACLE_H_
#define GRTM_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_LAYERS 64

/* status codes (shared meaning with include/grtm.h) */
#define OR_OK 0
#define OR_E_IO -1
#define OR_E_PARSE -2
#define OR_E_INPUT -3
#define OR_E_UNSUPPORTED -4

/* A parsed drop-in run (input.txt, green.cu:1989-2101). */
typedef struct {
    int n_layer;                 /* after source-layer insertion */
    double h[OR_MAX_LAYERS];     /* thickness (km); last entry keeps its top depth */
    double a0[OR_MAX_LAYERS], b0[OR_MAX_LAYERS], rho[OR_MAX_LAYERS];
    double qp[OR_MAX_LAYERS], qs[OR_MAX_LAYERS];
    double hs, hr;               /* source / receiver depth as read */
    int s_layer, r_layer;
    double zs, zr;               /* depths inside the source / receiver layer */
    double dt, dist, twin, t0, taper;
    int nf1, nf2, log2n, n_ptam, src_type;
    int split_done;              /* 1 if a layer was inserted for the source */
} or_run;

/* Parse input.txt.  compat selects the reference's receiver-layer search. */
int or_parse_input(const char *path, int compat, or_run *run);

/* Ragged reference grid: number of k samples for frequency index nw
 * (green.cu:2145-2194 with cal_kmax green.cu:108-162), and dk. */
int or_ref_grid(const or_run *run, int compat, double *dk, double *omega_i,
                int *nk_per_w /* [n_half] or NULL */, long long *nkw);

/* Post-processed spectrum rows (oname_w rows 0..Nw-1, 30 complex columns
 * each, interleaved re/im) for the frequency indices listed in rows[].
 * out has nrows*60 doubles. */
int or_green_rows(const char *iname, int compat, const int *rows, int nrows,
                  double *out, int nthreads);

/* Complete drop-in run: writes oname_w and oname_t like green.cu:2208-2258. */
int or_green_files(const char *iname, const char *oname_w, const char *oname_t,
                   int compat, int nthreads);

/* ---- cell-level entry points (for unit checks) ---- */

/* Dispersed complex velocities (green.cu:177-180 / 2148-2153). */
void or_dispersion(double wr, double wi, double v0, double q, double *out_re_im);

/* Interface R/T of the P-SV system between two layers at (k, w): solves the
 * 4x4 system E1 M = E2 (green.cu:235-262, 359-387).  Each layer is
 * (alpha re,im, beta re,im, mu re,im).  out: 4 blocks of 2x2 complex
 * (tdd, rud, rdu, tuu), row-major, interleaved re/im = 32 doubles; then the SH
 * scalars tdd0, rud0, rdu0, tuu0 (8 doubles).  Total 40 doubles. */
void or_interface_rt(double k, double wr, double wi, const double *lay_up,
                     const double *lay_dn, double *out);

/* Kernel coefficients of one (w,k) cell.  lay = n_layer rows of
 * (h, alpha re,im, beta re,im, mu re,im) = 7 doubles.  par0_rv = (r,v) of
 * layer r_layer in the first cell (k=0, w_0), complex, 4 doubles (compat only).
 * out (complex, interleaved): a0[2], a1[2], a2[2], b1, b2, a0t[2], a1t[2],
 * a2t[2], b1t, b2t  = 18 complex = 36 doubles. */
int or_cell(int compat, int src_type, double k, double wr, double wi,
            int n_layer, const double *lay, int s_layer, double zs,
            int r_layer, double zr, const double *par0_rv, double *out);

/* Test switch: 1 = the reference's SH factors of the tensor source
 * (green.cu:1408-1411) in physics mode as well. */
void or_set_tensor_sh_reference(int on);

/* Bessel family as used by the integrand: mode 0 = reference formulas
 * (green.cu:1418-1554), mode 1 = libm jn.  which: 0 J_n, 1 J_n', 2 "J_n''". */
double or_bessel(int mode, int which, int n, double x);

/* ---- multi-receiver traces on a dense uniform k grid ---- */
/* src_kind: 0 explosion, 1 vertical force.  Outputs (Z then R) per receiver:
 * traces[(ir*2 + c)*N + j], N = 2*nw.  spec (optional) gets the
 * post-processed half spectrum [(ir*2+c)*nw + iw] complex interleaved.
 * lay rows: (thickness, vp, vs, rho, qp, qs); last row = half-space.
 * w_lo..w_hi restricts the frequency rows evaluated (others left 0);
 * pass 0, nw for all. */
int or_traces(int compat, const double *lay, int n_in, double src_depth,
              double rcv_depth, const double *dists, int nr, int src_kind,
              int nw, double dt, int nk, double kmax, double omega_i,
              double t0, double taper, int w_lo, int w_hi, double *traces,
              double *spec, int nthreads);

/* The same with the tail of every k-sum beyond the grid taken by peak-trough
 * averaging (PTAM; the reference's n_ptam, green.cu:2059, integral.cu:38-88):
 * nt more wavenumbers on the same dk, of which receiver r uses
 * min(nt, ceil((n_ptam/2 + 2) 2 pi / (r dk))); n_ptam extrema per real
 * sequence, 2 <= n_ptam <= OR_PTAM_MAX.  nt = 0: or_traces. */
#define OR_PTAM_MAX 32
int or_traces_ptam(int compat, const double *lay, int n_in, double src_depth,
                   double rcv_depth, const double *dists, int nr, int src_kind,
                   int nw, double dt, int nk, double kmax, double omega_i,
                   double t0, double taper, int w_lo, int w_hi, int n_ptam,
                   int nt, double *traces, double *spec, int nthreads);
/* PTAM of one complex k-sum: s0 = the sum so far, terms[2 nt] the further
 * terms (re, im interleaved); out[2] the estimate, *found the extrema seen
 * (the smaller count of the two parts, at most n). */
int or_ptam(double s0_re, double s0_im, const double *terms, int nt, int n,
            double *out, int *found);

/* Displacement columns of both sources on the dense grid (all frequency
 * rows): tensor columns 0-9, then single-force columns 0-4, of the
 * reference's integrand, each summed over k and post-processed like
 * or_traces.  traces[(c*nr + ir)*N + j], c < OR_N_COLUMNS. */
#define OR_N_COLUMNS 15
int or_columns(int compat, const double *lay, int n_in, double src_depth,
               double rcv_depth, const double *dists, int nr, int nw,
               double dt, int nk, double kmax, double omega_i, double t0,
               double taper, double *traces, int nthreads);
/* all30 = 1: all 30 integrand columns of each source instead,
 * traces[((src*30 + c)*nr + ir)*N + j] with src 0 tensor, 1 single force
 * (physics mode: d/dz in columns 5-9 / 10-19, d/dr in 10-14 / 20-29). */
int or_columns_all(int compat, int all30, const double *lay, int n_in,
                   double src_depth, double rcv_depth, const double *dists,
                   int nr, int nw, double dt, int nk, double kmax,
                   double omega_i, double t0, double taper, double *traces,
                   int nthreads);

/* The same dense-grid integration but returning the raw k-kernels
 * Kz, Kr (complex) for every (iw, ik) in [w_lo,w_hi) x [0,nk):
 * kern[((iw-w_lo)*nk + ik)*4 + {0,1,2,3}] = Kz re, im, Kr re, im, where the
 * integrand of Z is Kz*J0(k r) and of R is Kr*J1(k r) (no dk weight). */
int or_kernels(int compat, const double *lay, int n_in, double src_depth,
               double rcv_depth, int src_kind, int nw, double dt, int nk,
               double kmax, double omega_i, int w_lo, int w_hi, double *kern,
               int nthreads);

/* Post-processing + inverse FFT of one half spectrum (nw complex) into a
 * trace of 2*nw reals (green.cu:1905-1938, 2222-2237). */
void or_post_ifft(int nw, const double *spec_raw, double df, double omega_i,
                  double dt, double t0, double taper, double *spec_post,
                  double *trace);

#ifdef __cplusplus
}
#endif
#endif
