/*
 * rbc_oracle.h -- FP64 CPU restatement of gpu_flow's Boussinesq Rayleigh-Benard time
 * stepper (jceyy/sdc), used ONLY as the parity checker and as the timed CPU baseline.
 *
 * TEST INFRASTRUCTURE. Nothing in the product library (sdc_amd/) may link, load or call
 * this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Layout (the reference's matrix layout, matrix/matrix_device.h:13-63): a spectral field is
 * (mox = M/2+1) x N x L complex doubles, x (kx column) fastest, then y (row), then the
 * vertical mode m.  F and G are L complex doubles.  Complex = interleaved (re, im).
 */
#ifndef RBC_ORACLE_H
#define RBC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Galerkin tables (util/coeff.cu) ------------------------------------------------- */
double orc_lambda(int k); /* k >= 1, coeff.cu:1232-1257 */

/* Index names of the 3-factor tables returned by orc_coeff_three (same order as the
 * reference accessors, coeff.h:80-92). */
enum {
    ORC_IAC1 = 0, ORC_IBB1, ORC_IAAB, ORC_IAAD, ORC_IABC, ORC_IA11, ORC_IA12, ORC_I111,
    ORC_IBBB, ORC_IB11, ORC_IAA1, ORC_IAB2, ORC_NTHREE
};
enum { ORC_IAC = 0, ORC_IAB, ORC_IA1, ORC_IA2, ORC_NTWO };

/* two[t][i + j*L], three[t][i + j*L + k*L*L]; returns 0 on success. */
int orc_coeff_tables(int L, double* two /* ORC_NTWO*L*L */, double* three /* ORC_NTHREE*L^3 */);

/* Gauss-Legendre rule on [-1/2, 1/2] with n nodes (ascending). */
int orc_gauss_legendre(int n, double* z, double* w);

/* ---- FFT (cuFFT boundary: C2R unnormalised inverse, R2C forward) ----------------------- */
/* real phys[N][M] (x fastest) <-> complex spec[N][M/2+1]; M, N powers of two. */
int orc_fft_r2c(int M, int N, const double* phys, double* spec);
int orc_fft_c2r(int M, int N, const double* spec, double* phys);

/* ---- Stepper context -------------------------------------------------------------------- */
typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(int M, int N, int L, double Lx, double Ly, double Ra, double Pr, double dt,
                    int nthreads);
void orc_destroy(orc_ctx* c);

/* theta, f, g: mox*N*L complex; F, G: L complex.  NULL skips a field. */
int orc_set_state(orc_ctx* c, const double* theta, const double* f, const double* g,
                  const double* F, const double* G);
int orc_get_state(orc_ctx* c, double* theta, double* f, double* g, double* F, double* G);

/* Operator-level entry points (each mirrors one reference call; in/out in full layout). */
int orc_nonlinear(orc_ctx* c, const double* th, const double* f, const double* g,
                  const double* F, const double* G, double* oth, double* of, double* og,
                  double* oF, double* oG);
int orc_inverse_B(orc_ctx* c, const double* th, const double* f, const double* g,
                  const double* F, const double* G, double* oth, double* of, double* og,
                  double* oF, double* oG);
int orc_implicit(orc_ctx* c, const double* th, const double* f, const double* g,
                 const double* F, const double* G, double* oth, double* of, double* og,
                 double* oF, double* oG);

void orc_dealias(int mox, int N, int L, double* field);   /* implicit_explicit_stepping.cu:163-182 */
void orc_make_real(int mox, int N, int L, double* field); /* util.cu make_real */

/* Penalization of a PENALIZED build (nonlinear_operator_rayleigh_noslip.cu:1097-1343) with
 * the mask shapes of :1652-1740 (rectangle: the default build, circle: -DCIRCLE, disk array:
 * -DDISKARRAY, everywhere: the commented alternative); eta = 0 turns it off. */
enum { ORC_MASK_RECTANGLE = 0, ORC_MASK_CIRCLE, ORC_MASK_DISKARRAY, ORC_MASK_EVERYWHERE };
int orc_set_penalization(orc_ctx* c, double eta, int shape);

/* One reference time step (step_time + make_real_signal), nsteps times. */
int orc_step(orc_ctx* c, int nsteps);
void orc_reset_history(orc_ctx* c);

/* Diagnostics defined by this project (the reference computes no Nusselt number):
 * nu[0] = plate Nu at z=-1/2, nu[1] = plate Nu at z=+1/2, nu[2] = 1 + <w theta>. */
int orc_nusselt(orc_ctx* c, double* nu3);

/* CPU-baseline timing of one step with every pair_stride-th (i,j) pair of the nonlinear term;
 * out[4] = {extrapolated seconds per step, physical-field phase, sampled pair phase, rest}. */
int orc_time_step_sampled(orc_ctx* c, int pair_stride, double* out);

/* Per-wavenumber implicit factor block for (kx col, row): out 4*L*L doubles (TT, Tf, fT, ff). */
int orc_implicit_factor(orc_ctx* c, int col, int row, double* out);

#ifdef __cplusplus
}
#endif
#endif
