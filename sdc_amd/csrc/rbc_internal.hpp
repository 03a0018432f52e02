// rbc_internal.hpp -- internal declarations shared by the host setup, the kernels and the
// C-ABI of libsdc_rbc.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "rbc_abi.h"

namespace rbc {

// ---------------------------------------------------------------------------------------
// FP64 host setup (rbc_setup.cpp)
// ---------------------------------------------------------------------------------------

// lambda_k of the Chandrasekhar functions as the reference tabulates them
// (util/coeff.cu:1232-1257): 12-digit roots for k <= 8, pi (k + 1/2) beyond.
double lambda_k(int k);

// Gauss-Legendre nodes/weights on [-1/2, 1/2], ascending.
void gauss_legendre(int n, std::vector<double>& z, std::vector<double>& w);

// Basis values at nodes z: out[(t*L + m)*nz + q] for t = 0..5 = C, C', C'', C''', S, S'.
void basis_values(int L, const std::vector<double>& z, std::vector<double>& out);

// Closed-form 2-factor integrals used by the linear operators, restated from
// util/coeff.cu:77-116 (Iac, including its even-even off-diagonal zeros) and :146-178 (Ia1).
// t[i + j*L].
void closed_form_iac_ia1(int L, std::vector<double>& Iac, std::vector<double>& Ia1);

// Smallest number of collocation levels (multiple of 8) whose Gauss-Legendre rule
// reproduces the Galerkin triple products of L modes to relative accuracy tol.
int exact_level_count(int L, double tol, double* achieved);

// Kept-mode geometry of the reference dealias mask (implicit_explicit_stepping.cu:163-182).
struct Geometry {
    int M, N, L, mox;
    int KX;                 // kept kx columns 0..KX-1
    std::vector<int> rows;  // kept rows (ascending), NYc entries
    int NYc() const { return (int)rows.size(); }
    int P() const { return KX * NYc(); }  // compact mode count; p = kx*NYc + kyc
};
Geometry make_geometry(int M, int N, int L);

// Per-|k|^2 factors, split by vertical parity class (class 0: 0-based even m, class 1: odd).
// For each group: [Cinv class0 (n0 x n0, column-major), Cinv class1, Binv_f class0 (L0 x L0),
// Binv_f class1], n_c = 2*L_c, unknown order within a class: theta modes then f modes.
struct Factors {
    int L0, L1, stride;                 // class sizes and doubles per group
    std::vector<double> mats;           // ngroups * stride
    std::vector<int> group_start;       // ngroups + 1
    std::vector<int> members;           // P, compact mode indices grouped
    std::vector<double> k2;             // P, |k|^2 per compact mode
    std::vector<double> kx, ky;         // P, wavenumbers per compact mode
    int ngroups() const { return (int)group_start.size() - 1; }
};
// Follows linear_implicit_operator.cu:47-214 and inverse_B_matrix.cu:98-164.
// Returns 0 or a negative status (singular matrix).
int build_factors(const Geometry& g, double Lx, double Ly, double Ra, double Pr, double dt,
                  Factors& out, std::string& err);

// ---------------------------------------------------------------------------------------
// Device-side launchers (rbc_kernels.hip), templated on the real type T.
// ---------------------------------------------------------------------------------------
struct DevGeom {
    int M, N, L, nq, KX, KXp, NYc, P;
    int logM, logN;
    double Lx, Ly;
};

struct Buffers {
    // mode space, [p][m] complex (m fastest)
    void* X[3];
    void* XF;
    void* XG;  // L complex
    void* Nn[3];
    void* NF;
    void* NG;
    void* Rp[3];
    void* RpF;
    void* RpG;
    // operator-level staging (same shapes as X / Nn)
    void* Xo[3];
    void* XoF;
    void* XoG;
    void* Oo[3];
    void* OoF;
    void* OoG;
    // level space
    void* lev;    // 6 * nq * P complex  [field][l][p]
    void* hin;    // 6 * nq * N * KXp complex  [field][l][y][kx]
    void* hout;   // 7 * nq * N * KXp complex  [prod][l][y][kx]
    void* prod;   // 7 * nq * P complex  [prod][l][p] (generic path only)
    // the projection's five level combinations live in `lev` (free once the inverse y-FFT has
    // consumed it): [5][l][p] complex; fgw[l] = (P_uw, P_vw) at k = 0 for the mean flow
    void* fgw;    // nq * 2 complex
    void* refstage;  // (M/2+1) * N * L complex FP64: one field of the host boundary in the reference layout
    // tables
    void* evtab;  // nq * 4 * L : C', C, S, S' at level l   [l][t][m]
    void* prtab;  // nq * 5 * L : w*C', w*C'', w*C, w*S, w*S'  [l][t][m]
    void* twM;    // M complex: exp(-2 pi i k / M), k < M
    void* twN;    // N complex
    void* twP = nullptr;  // at M = N = 512 (FP32 storage): 576 complex, the 512-point twiddles in the
                          // per-pass order of the 512-point kernels (TwByPass)
    void* kx;     // P real
    void* ky;     // P real
    void* k2;     // P real
    int* rows;    // NYc kept rows
    int* gstart;  // ngroups + 1
    int* gmem;    // P
    // work chunks of at most kChunk members of one group (a group's chunks are consecutive):
    // chunk c covers members [cbeg[c], cbeg[c+1]) of group cgrp[c]
    int* cgrp;
    int* cbeg;
    int nchunk;
    void* gmat;   // ngroups * stride real
    double* red;  // reduction scratch (doubles)
    int nred;
    // step health: health[0] = 1 + the step count at which the end-of-step check first saw a
    // non-finite value (0 = none), health[1] = steps finished since create (both written by the
    // last block of the end-of-step kernel only)
    int* health;
    // matrix-core vertical transforms (FP32 storage): -1 = off, 0 = FP32 MFMA, 1 = BF16 x3
    int mma = -1;
    void* evA = nullptr;  // [4][ev_nlt][ev_nks][64][KPL] floats (C', C, S, S')
    void* prA = nullptr;  // [5][pr_nks][64][KPL] floats (wC', wC'', wC, wS, wS')
    int ev_nlt = 0, ev_nks = 0, pr_nks = 0;
    // launch shapes of the generic (non-512) FFT kernels; 0 = built-in choice.  Set by the
    // ensemble autotuner; the arithmetic per element does not depend on them.
    int tune_ytx = 0;           // kx columns per block of the y-FFTs
    int tune_rows_threads = 0;  // threads per block of the row kernel (64, 128 or 256)
};

// Penalization of a PENALIZED build (nonlinear_operator_rayleigh_noslip.cu:1097-1343): the
// Brinkman terms of a z-independent mask chi(x, y) added to the nonlinear term.  Device arrays
// in the context's real type; the four work arrays may alias the context's level-space buffers
// (the penalty runs after the projection has consumed them).
struct Penalty {
    double eta = 0.0;        // 0 = off
    void* mask = nullptr;    // M * N reals, x fastest
    void* iac = nullptr;     // L * L reals, Iac(j, k) at j + k L (coeff.cu:77-116, quirk kept)
    void* ia2 = nullptr;     // L * L reals, Ia2(i, j) = <C_i S'_j> at i + j L
    void* spec = nullptr;    // [10][L][P] complex: the mode-space fields before the transforms
    void* hin = nullptr;     // [10][L][N][KXp] complex (row-major hybrid rows)
    void* hout = nullptr;    // same, after the mask products
    void* spec_out = nullptr;  // [10][L][P] complex: the transforms of the mask products
};

// Ia2(i, j) = <C_i S'_j> by a 4 L + 64 point Gauss-Legendre rule (coeff.cu's ra2), i + j L.
void quadrature_ia2(int L, std::vector<double>& Ia2);

// The mask shapes of nonlinear_operator_rayleigh_noslip.cu:1652-1740 on the M x N grid.
void penalty_mask(int shape, int M, int N, std::vector<double>& mask);

struct StepCoef {
    double a_cur, a_prev;  // X += a_cur*R ; X += a_prev*R_prev
    int use_prev;
};

enum UpdateMode { UPD_STEP = 0, UPD_INVB = 1, UPD_IMPLICIT = 2 };

// plane operators on the state's mode planes (freeze_operator.cu, masktemp.cu, masking.cu)
enum PlaneOp { PLANE_CAPTURE = 0, PLANE_FREEZE = 1, PLANE_HEAT = 2, PLANE_SCALE = 3, PLANE_SET = 4 };

// one field between an FP64 reference-layout device buffer and the compact field of the
// context's real type (rsz = 4 or 8); enqueued on s
void field_from_ref64(const DevGeom& g, int rsz, const void* ref, void* dst, hipStream_t s);
void field_to_ref64(const DevGeom& g, int rsz, const void* src, void* ref, hipStream_t s);

template <typename T>
struct Launch {
    // stage 1: vertical evaluation -> six level fields (spectral)
    static void eval(const DevGeom& g, const Buffers& b, void* const X[3], const void* XF,
                     const void* XG, hipStream_t s);
    // stage 2: inverse y-FFT of level fields into hybrid (kx spectral, y physical) rows
    static void fft_y_inv(const DevGeom& g, const Buffers& b, hipStream_t s);
    // stage 3: x C2R, pointwise products, x R2C per (level, y) row
    static void rows(const DevGeom& g, const Buffers& b, hipStream_t s);
    // stage 4: forward y-FFT of the products back to compact spectral layout
    static void fft_y_fwd(const DevGeom& g, const Buffers& b, hipStream_t s);
    // stage 5: vertical projection -> N (theta, f, g, F, G), scaled as the reference does
    static void project(const DevGeom& g, const Buffers& b, double Pr, void* const Nout[3],
                        void* NFo, void* NGo, hipStream_t s);
    // stage 6: per-|k|^2 group update: inverse B, AB2 combine, implicit Euler
    static void update(const DevGeom& g, const Buffers& b, int ngroups, int L0, int L1,
                       int stride, double Pr, double dt, StepCoef sc, int mode,
                       void* const Xio[3], void* const Nin[3], void* const Rp[3], hipStream_t s);
    // F and G (mean flow) update
    static void update_FG(const DevGeom& g, double Pr, double dt, StepCoef sc, int mode,
                          void* XF, void* XG, const void* NF, const void* NG, void* RpF,
                          void* RpG, hipStream_t s);
    // make_real of X and the F, G update of a time step (mode UPD_STEP) in one launch
    static void finish(const DevGeom& g, const Buffers& b, double Pr, double dt, StepCoef sc, hipStream_t s);
    // make_real on the kx = 0 column (util.cu make_real)
    static void make_real(const DevGeom& g, const Buffers& b, void* const X[3], hipStream_t s);
    static void nusselt(const DevGeom& g, const Buffers& b, const void* ia1, hipStream_t s);
    // device fields in the reference layout ((M/2+1) x N x L, T complex) -> compact [p][m]
    // (kept modes); null source fields are skipped
    static void from_ref(const DevGeom& g, const void* const src[3], void* const dst[3], hipStream_t s);
    // compact [p][m] -> reference layout with zeros in the removed band; null targets skipped
    static void to_ref(const DevGeom& g, const void* const src[3], void* const dst[3], hipStream_t s);
    // initial state of the context's X, XF, XG (kind: 0 zeros, 1 random theta, 2 init_test)
    static void init_state(const DevGeom& g, const Buffers& b, int kind, uint64_t seed, double amp, hipStream_t s);
    // the real plane of an initial pattern (util/init.cu), [row][col] of M x N, in the real type
    static void init_pattern(const DevGeom& g, int kind, int mode, double amp, uint64_t seed, void* plane,
                             hipStream_t s);
    // F_0 = v (force_mean)
    static void force_mean(const Buffers& b, double v, hipStream_t s);
    // plane operator `op` on the first nf fields of X (theta, f, g) in place (PLANE_CAPTURE
    // writes the physical planes to `frozen` and leaves X); spec / spec_out hold nf L P complex,
    // hin / hout nf L N KXp complex
    static void plane_op(const DevGeom& g, const Buffers& b, int op, int nf, const int* imask, double factor,
                         void* frozen, void* const X[3], void* spec, void* hin, void* hout, void* spec_out,
                         hipStream_t s);
    // out[x + y M] = sum_i S[i] planes[i][y][x] (float), planes: L physical mode planes of theta
    static void temperature_sum(const DevGeom& g, const void* planes, const float* S, float* out, hipStream_t s);
    // adds the penalization terms of state X to N (already scaled as the projection scales)
    static void penalty(const DevGeom& g, const Buffers& b, const Penalty& pen, double Pr, void* const X[3],
                        const void* XF, const void* XG, void* const Nout[3], void* NF, void* NG, hipStream_t s);
};

// Batched steps of the members of one ensemble size class (rbc_kernels.hip).  A member is
// described by its geometry, buffers and step constants; batch_create returns nullptr (reason
// in why) when the members are not on the path the batched kernels cover (FP32 storage on the
// matrix cores, group transforms of 64..256 points, L <= 32) or differ in launch shape.
struct BatchMember {
    const DevGeom* g;
    const Buffers* b;
    double Pr, dt;
    int L0, L1, stride;
};
struct StepBatch;
StepBatch* batch_create(const std::vector<BatchMember>& m, std::string& why);
// one AB2 step of every member (all of them must have their AB2 history)
void batch_step(const StepBatch* sb, hipStream_t s);
void batch_destroy(StepBatch* sb);

}  // namespace rbc
