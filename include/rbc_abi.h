/*
 * rbc_abi.h -- C-ABI of the MI355X-native Rayleigh-Benard time stepper (libsdc_rbc.so).
 *
 * Drop-in boundary for gpu_flow's hot path (jceyy/sdc, gpu_flow/code): the Boussinesq
 * RBC IMEX time stepper with its nonlinear-term evaluation, inverse-B, implicit linear
 * step, and the Nusselt diagnostics.  Plain C types only; all calls return an int status
 * (0 = OK, < 0 = error; rbc_last_error() gives the text).  Nothing calls exit().
 *
 * Field layout at the boundary is the reference's (matrix/matrix_device.h:13-63):
 *   theta, f, g : (M/2+1) x N x L complex, kx column fastest, then row (ky), then mode m;
 *   F, G        : L complex.
 * Complex values are interleaved (re, im) doubles at the boundary for every precision.
 *
 * Threading: one HIP stream per context; calls on one context are not thread-safe;
 * distinct contexts may be driven from distinct host threads.  All device memory is
 * allocated in rbc_create; rbc_step and the operator calls allocate nothing.
 */
#ifndef RBC_ABI_H
#define RBC_ABI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBC_ABI_VERSION 1

/* Status codes */
#define RBC_OK 0
#define RBC_ERR_ARG (-1)
#define RBC_ERR_HIP (-2)
#define RBC_ERR_NOMEM (-3)
#define RBC_ERR_NUMERIC (-4)
#define RBC_ERR_STATE (-5)
#define RBC_ERR_IO (-6)

/* Arithmetic of the GPU path.
 *   RBC_F64    : every GPU stage in FP64 (parity form; PRECISION_DOUBLE, cuda_defines.h:9-27)
 *   RBC_F32    : FFTs, products, vertical transforms and implicit apply in FP32
 *   RBC_BF16X3 : as RBC_F32, but the vertical transforms run as BF16 hi/lo split MFMA
 *                products accumulated in FP32
 * Per-wavenumber factors are always built in FP64 on the host. */
typedef enum { RBC_F64 = 0, RBC_F32 = 1, RBC_BF16X3 = 2 } rbc_precision;

typedef struct {
    int M, N, L;     /* physical grid M x N, L vertical modes (main.cu: -nmodes M N L) */
    double Lx, Ly;   /* horizontal box lengths, height 1 (main.cu: -clength) */
    double Ra, Pr;   /* Rayleigh and Prandtl numbers (-ra, -pr) */
    double dt;       /* time step, baked into the implicit factors (-dt) */
    int precision;   /* rbc_precision */
    int nq;          /* vertical collocation levels; 0 = smallest exact rule for L */
    int device;      /* HIP device ordinal (-device) */
} rbc_params;

typedef struct rbc_ctx rbc_ctx;

/* ≙ Coeff<double>(L) + linear_implicit_operator::init_operator + inverse_B_matrix::
 *   init_operator + nonlinear_operator_rayleigh_noslip::init_operator +
 *   implicit_explicit_step::init_timestepping (main.cu:431-454). */
int rbc_create(const rbc_params* p, rbc_ctx** out);
void rbc_destroy(rbc_ctx* c);
const char* rbc_last_error(const rbc_ctx* c); /* c may be NULL: last create error */
int rbc_get_info(const rbc_ctx* c, int* nq, int* kx_kept, int* ky_kept, int* n_unique_k);

/* Host state exchange in the reference layout (see above).  NULL skips a field.
 * The stepper dealiases on entry (implicit_explicit_stepping.cu:40), so set_state keeps
 * only the modes the dealias mask keeps; get_state returns zeros in the removed band. */
int rbc_set_state(rbc_ctx* c, const double* theta, const double* f, const double* g,
                  const double* F, const double* G);
int rbc_get_state(rbc_ctx* c, double* theta, double* f, double* g, double* F, double* G);

/* ≙ implicit_explicit_step::step_time + make_real_signal, nsteps times (main.cu:612-616).
 * The context owns the AB2 history; the first step after create/reset is Euler
 * (implicit_explicit_stepping.cu:43-52). */
int rbc_step(rbc_ctx* c, int nsteps);
int rbc_reset_history(rbc_ctx* c);

/* Operator-level calls on host buffers in the reference layout (inputs are dealiased):
 *   rbc_nonlinear  ≙ nonlinear_operator_rayleigh_noslip::calculate_operator
 *                    (nonlinear_operator_rayleigh_noslip.h:56)
 *   rbc_inverse_B  ≙ inverse_B_matrix::calculate_inverse (inverse_B_matrix.h:18)
 *   rbc_implicit   ≙ linear_implicit_operator::calculate_operator
 *                    (linear_implicit_operator.h:65)
 * Outputs are written for the kept modes and zero elsewhere. */
int rbc_nonlinear(rbc_ctx* c, const double* theta, const double* f, const double* g,
                  const double* F, const double* G, double* o_theta, double* o_f, double* o_g,
                  double* o_F, double* o_G);
int rbc_inverse_B(rbc_ctx* c, const double* theta, const double* f, const double* g,
                  const double* F, const double* G, double* o_theta, double* o_f, double* o_g,
                  double* o_F, double* o_G);
int rbc_implicit(rbc_ctx* c, const double* theta, const double* f, const double* g,
                 const double* F, const double* G, double* o_theta, double* o_f, double* o_g,
                 double* o_F, double* o_G);

/* Diagnostics (FP64 reductions on the device; the reference computes none):
 *   nu[0] plate Nusselt number at z=-1/2, nu[1] at z=+1/2, nu[2] = 1 + <w theta>. */
int rbc_nusselt(rbc_ctx* c, double* nu3);

/* Per-stage timing with HIP events on the context's stream: runs nsteps steps and writes
 * the mean milliseconds per call of each stage into ms[RBC_NSTAGES] (stage names from
 * rbc_stage_name). */
#define RBC_NSTAGES 8
int rbc_profile_steps(rbc_ctx* c, int nsteps, double* ms);
const char* rbc_stage_name(int i);

/* Blocks until all work queued on the context's stream has finished. */
int rbc_synchronize(rbc_ctx* c);

#ifdef __cplusplus
}
#endif
#endif /* RBC_ABI_H */
