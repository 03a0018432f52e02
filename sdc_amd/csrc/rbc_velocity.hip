// rbc_velocity.hip -- the velocity operator and what is built on it: the velocity on a
// grid of heights (calculate_velocity_operator), passive tracers in the reference's mode
// (RK4 with trilinear interpolation on a 128-level grid, particle_tracer.cu) and in a
// spectral mode (midpoint RK2 with the velocity summed exactly at each particle), and the
// kinetic-energy spectra.
//
// Velocity from the state (calculate_velocity_operator.cu:172-245, 426-505):
//   u = sum_m C'_m(z) d_x f_m + S_m(z) (d_y g_m + F_m at k = 0)
//   v = sum_m C'_m(z) d_y f_m + S_m(z) (-d_x g_m + G_m at k = 0)
//   w = sum_m C_m(z) (-Laplace_h f_m)
// with the unnormalised C2R as the horizontal synthesis.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rbc_ctx.hpp"
#include "rbc_device.hpp"

using namespace rbc;

namespace {

constexpr int kZNum = 128;                    // PARTICLE_Z_NUM (cuda_defines.h:44)
constexpr double kZMax = 0.5f - 1e-3;         // PARTICLE_Z_MAX (cuda_defines.h:45), a double
constexpr int kTile = 64;                     // modes per LDS tile of the spectral evaluation

// ---- counter-based random numbers (SplitMix64, Steele, Lea & Flood 2014) ----------------
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// uniform in (0, 1] from (seed, counter), 24 bits
__device__ inline float rng_u01(uint64_t seed, uint64_t ctr) {
    const uint64_t h = splitmix64(seed * 4294967296ull + ctr);
    return (float)((h >> 40) + 1) * (1.0f / 16777216.0f);
}
// standard normal by Box-Muller from the uniforms at 2 ctr and 2 ctr + 1
__device__ inline float rng_normal(uint64_t seed, uint64_t ctr) {
    const float u1 = rng_u01(seed, 2 * ctr), u2 = rng_u01(seed, 2 * ctr + 1);
    return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
}

// ---- basis values at one height, FP64 (coeff.h:17-36; the clamp of get_z_coeff_at,
//      calculate_velocity_operator.cu:462-466) --------------------------------------------
// The mode's constants: cosh and cos of lam / 2 (odd k = m + 1), sinh and sin (even k).
__host__ __device__ inline void basis_consts(int m, double lam, double& c1, double& c2) {
    if ((m + 1) % 2) {
        c1 = cosh(0.5 * lam);
        c2 = cos(0.5 * lam);
    } else {
        c1 = sinh(0.5 * lam);
        c2 = sin(0.5 * lam);
    }
}
// The height-dependent part: cosh and sinh of lam z from one exponential (|lam z| <= lam / 2, no
// overflow), cos and sin from one sincos.
__host__ __device__ inline void basis_at_c(int m, double lam, double c1, double c2, double z, double& C, double& Cp,
                                           double& S) {
    if (z < -0.5) z = -0.5;
    else if (z > 0.5) z = 0.5;
    const int k = m + 1;
    const double lz = lam * z;
    const double ex = exp(lz), iex = 1.0 / ex;
    const double chz = 0.5 * (ex + iex), shz = 0.5 * (ex - iex);
    double sz, cz;
    sincos(lz, &sz, &cz);
    if (k % 2) {
        C = chz / c1 - cz / c2;
        Cp = lam * (shz / c1 + sz / c2);
    } else {
        C = shz / c1 - sz / c2;
        Cp = lam * (chz / c1 - cz / c2);
    }
    S = sin(M_PI * k * (z + 0.5));
}
__host__ __device__ inline void basis_at(int m, double lam, double z, double& C, double& Cp, double& S) {
    double c1, c2;
    basis_consts(m, lam, c1, c2);
    basis_at_c(m, lam, c1, c2, z, C, Cp, S);
}

// Basis values at the points' heights for the FP32 point evaluations, once per evaluation:
// zb[i][m] = (C'_m, C_m, S_m, 0) at z_i, basis_at in FP64 rounded to FP32 -- the values the
// point kernels would otherwise compute in every mode slice.
__global__ void __launch_bounds__(256) k_point_basis(int np, int L, const float* __restrict__ pz,
                                                     const double* __restrict__ lam, float4* __restrict__ zb) {
    // thread: mode m = e % L of the points g * 8 .. g * 8 + 7, g = e / L (the mode's constants once)
    constexpr int R = 8;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t ng = ((size_t)np + R - 1) / R;
    if (e >= ng * L) return;
    const int m = (int)(e % L);
    const size_t g0 = e / L * R;
    const double lm = lam[m];
    double c1, c2;
    basis_consts(m, lm, c1, c2);
    for (int r = 0; r < R; ++r) {
        const size_t i = g0 + r;
        if (i >= (size_t)np) break;
        double C, Cp, S;
        basis_at_c(m, lm, c1, c2, (double)pz[i], C, Cp, S);
        zb[i * L + m] = make_float4((float)Cp, (float)C, (float)S, 0.f);
    }
}

// ---------------------------------------------------------------------------------------
// Spectral velocity at nz heights: spec[c][l][p] for c = u, v, w (complex), thread per mode.
// zb[l][3][L] = C'_m, C_m, S_m at height l.
// ---------------------------------------------------------------------------------------
template <typename T, typename O>
__global__ void __launch_bounds__(256) k_vel_levels(DevGeom g, int nz, const double* __restrict__ zb,
                                                    const cpx<T>* __restrict__ Xf, const cpx<T>* __restrict__ Xg,
                                                    const cpx<T>* __restrict__ XF, const cpx<T>* __restrict__ XG,
                                                    const T* __restrict__ kxa, const T* __restrict__ kya,
                                                    cpx<O>* __restrict__ spec) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x, l = blockIdx.y;
    if (p >= g.P) return;
    const int L = g.L;
    const double* tb = zb + (size_t)l * 3 * L;
    cpx<double> A = mk<double>(0, 0), B = A, Cc = A;
    for (int m = 0; m < L; ++m) {
        const cpx<T> f = Xf[(size_t)p * L + m], gg = Xg[(size_t)p * L + m];
        const cpx<double> fd = mk<double>(f.x, f.y), gd = mk<double>(gg.x, gg.y);
        A = A + cscale(fd, tb[m]);
        Cc = Cc + cscale(fd, tb[L + m]);
        B = B + cscale(gd, tb[2 * L + m]);
    }
    const double kx = kxa[p], ky = kya[p];
    cpx<double> u = imul(A, kx) + imul(B, ky);
    cpx<double> v = imul(A, ky) - imul(B, kx);
    const cpx<double> w = cscale(Cc, kx * kx + ky * ky);
    if (p == 0) {  // add_dc_part (calculate_velocity_operator.cu:479-494)
        for (int m = 0; m < L; ++m) {
            u = u + cscale(mk<double>(XF[m].x, XF[m].y), tb[2 * L + m]);
            v = v + cscale(mk<double>(XG[m].x, XG[m].y), tb[2 * L + m]);
        }
    }
    const size_t plane = (size_t)nz * g.P;
    cpx<O>* o = spec + (size_t)l * g.P + p;
    o[0] = mk<O>((O)u.x, (O)u.y);
    o[plane] = mk<O>((O)v.x, (O)v.y);
    o[2 * plane] = mk<O>((O)w.x, (O)w.y);
}

// ---------------------------------------------------------------------------------------
// x C2R of the hybrid rows [c][l][y][kx] of the three components into the real grid
// [c][l][y][x] (the reference's u_i[x + y M + l M N] layout).  Block per (y, l): u + i v and
// w go through two complex inverse FFTs of length M.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_vel_rows(DevGeom g, const cpx<float>* __restrict__ hyb,
                                                  const cpx<float>* __restrict__ twM, float* __restrict__ grid) {
    extern __shared__ unsigned char smem_raw[];
    const int M = g.M, N = g.N, KX = g.KX, KXp = g.KXp, nz = g.nq;
    cpx<float>* a = reinterpret_cast<cpx<float>*>(smem_raw);  // 2 M
    cpx<float>* b = a + 2 * M;
    const int y = blockIdx.x, l = blockIdx.y;
    const size_t fstride = (size_t)nz * N * KXp;
    const size_t roff = ((size_t)l * N + y) * KXp;
    for (int e = threadIdx.x; e < 2 * M; e += blockDim.x) a[e] = mk<float>(0, 0);
    __syncthreads();
    for (int e = threadIdx.x; e < 2 * KX; e += blockDim.x) {
        const int pr = e / KX, k = e - pr * KX;
        cpx<float> x1 = hyb[(size_t)(2 * pr) * fstride + roff + k];
        cpx<float> x2 = pr == 0 ? hyb[fstride + roff + k] : mk<float>(0, 0);
        if (k == 0) {  // a C2R ignores the imaginary part at kx = 0
            x1.y = 0;
            x2.y = 0;
        }
        cpx<float>* z = a + pr * M;
        z[k] = mk(x1.x - x2.y, x1.y + x2.x);
        if (k > 0) z[M - k] = mk(x1.x + x2.y, -x1.y + x2.x);
    }
    __syncthreads();
    const cpx<float>* r = lds_fft(a, b, g.logM, 2, twM, true);
    const size_t cs = (size_t)nz * N * M;
    float* o = grid + ((size_t)l * N + y) * M;
    for (int x = threadIdx.x; x < M; x += blockDim.x) {
        o[x] = r[x].x;
        o[cs + x] = r[x].y;
        o[2 * cs + x] = r[M + x].x;
    }
}

// ---------------------------------------------------------------------------------------
// Reference tracer step: RK4 in the frozen velocity grid with trilinear interpolation and
// periodic lookup (particle_tracer.cu:508-687, PERIODIC build), optional diffusion
// (:692-717) and reflection at the lids (:722-764).  Mixed float/double expressions are
// kept in the reference's types.
// ---------------------------------------------------------------------------------------
__device__ inline void vel_trilinear(const float* __restrict__ ux, const float* __restrict__ uy,
                                     const float* __restrict__ uz, int xs, int ys, float clx, float cly, float x,
                                     float y, float z, float& ru, float& rv, float& rw) {
    const int xy = xs * ys;
    float xg = x * xs / clx;
    float yg = y * ys / cly;
    float zg = (kZMax + z) / (2 * kZMax) * (kZNum - 1);
    xg -= floorf(xg / xs) * xs;
    yg -= floorf(yg / ys) * ys;
    int ix0 = (int)floorf(xg), iy0 = (int)floorf(yg), iz0 = (int)floorf(zg);
    if (ix0 < 0 || ix0 >= xs) ix0 = 0;
    if (iy0 < 0 || iy0 >= ys) iy0 = 0;
    if (iz0 < 0) iz0 = 0;
    else if (iz0 >= kZNum) iz0 = kZNum - 1;
    const int ix1 = (ix0 + 1 >= xs) ? 0 : ix0 + 1;
    const int iy1 = (iy0 + 1 >= ys) ? 0 : iy0 + 1;
    const int iz1 = (iz0 + 1 >= kZNum) ? iz0 : iz0 + 1;
    float wR = xg - float(ix0), wL = 1 - wR;
    auto lerpx = [&](int iy, int iz, float& a, float& b, float& c) {
        const int i0 = ix0 + iy * xs + iz * xy, i1 = ix1 + iy * xs + iz * xy;
        a = wL * ux[i0] + wR * ux[i1];
        b = wL * uy[i0] + wR * uy[i1];
        c = wL * uz[i0] + wR * uz[i1];
    };
    float x00, y00, z00, x10, y10, z10, x01, y01, z01, x11, y11, z11;
    lerpx(iy0, iz0, x00, y00, z00);
    lerpx(iy1, iz0, x10, y10, z10);
    lerpx(iy0, iz1, x01, y01, z01);
    lerpx(iy1, iz1, x11, y11, z11);
    wR = yg - float(iy0);
    wL = 1 - wR;
    const float cx0 = wL * x00 + wR * x10, cy0 = wL * y00 + wR * y10, cz0 = wL * z00 + wR * z10;
    const float cx1 = wL * x01 + wR * x11, cy1 = wL * y01 + wR * y11, cz1 = wL * z01 + wR * z11;
    wR = zg - float(iz0);
    wL = 1 - wR;
    ru = wL * cx0 + wR * cx1;
    rv = wL * cy0 + wR * cy1;
    rw = wL * cz0 + wR * cz1;
}

// diffusion with spread sigma (counter-based normals) and the lid reflection
__device__ inline void finish_particle(float& x, float& y, float& z, int i, int np, float sigma, uint64_t seed,
                                       uint64_t ctr0) {
    if (sigma > 0.f) {
        x += sigma * rng_normal(seed, ctr0 + i);
        y += sigma * rng_normal(seed, ctr0 + (uint64_t)np + i);
        z += sigma * rng_normal(seed, ctr0 + 2 * (uint64_t)np + i);
    }
    if (!(z <= kZMax)) {  // NaN goes to the top lid (particle_tracer.cu:741-749)
        z = (z < 2 * kZMax) ? (float)(2 * kZMax - z) : (float)kZMax;
    } else if (z < -kZMax) {
        z = (z > -2 * kZMax) ? (float)(-2 * kZMax - z) : (float)(-kZMax);
    }
}

__global__ void __launch_bounds__(256) k_tracer_rk4(int np, const float* __restrict__ grid, int M, int N, float Lx,
                                                    float Ly, const float* __restrict__ x0, const float* __restrict__ y0,
                                                    const float* __restrict__ z0, float dt, float* __restrict__ x1,
                                                    float* __restrict__ y1, float* __restrict__ z1, float sigma,
                                                    uint64_t seed, uint64_t ctr0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const size_t cs = (size_t)kZNum * N * M;
    const float *ux = grid, *uy = grid + cs, *uz = grid + 2 * cs;
    float x = x0[i], y = y0[i], z = z0[i];
    float vx, vy, vz;
    vel_trilinear(ux, uy, uz, M, N, Lx, Ly, x, y, z, vx, vy, vz);
    const float dx1 = dt * vx, dy1 = dt * vy, dz1 = dt * vz;
    float xt = x + 0.5 * dx1, yt = y + 0.5 * dy1, zt = z + 0.5 * dz1;
    vel_trilinear(ux, uy, uz, M, N, Lx, Ly, xt, yt, zt, vx, vy, vz);
    const float dx2 = dt * vx, dy2 = dt * vy, dz2 = dt * vz;
    xt = x + 0.5 * dx2;
    yt = y + 0.5 * dy2;
    zt = z + 0.5 * dz2;
    vel_trilinear(ux, uy, uz, M, N, Lx, Ly, xt, yt, zt, vx, vy, vz);
    const float dx3 = dt * vx, dy3 = dt * vy, dz3 = dt * vz;
    xt = x + dx3;
    yt = y + dy3;
    zt = z + dz3;
    vel_trilinear(ux, uy, uz, M, N, Lx, Ly, xt, yt, zt, vx, vy, vz);
    const float dx4 = dt * vx, dy4 = dt * vy, dz4 = dt * vz;
    x += (dx1 + 2 * (dx2 + dx3) + dx4) / 6.;
    y += (dy1 + 2 * (dy2 + dy3) + dy4) / 6.;
    z += (dz1 + 2 * (dz2 + dz3) + dz4) / 6.;
    finish_particle(x, y, z, i, np, sigma, seed, ctr0);
    x1[i] = x;
    y1[i] = y;
    z1[i] = z;
}

// ---------------------------------------------------------------------------------------
// Spectral velocity at particle positions: u(x) = sum over kept modes of
// w_k Re(exp(i k.x) u_k(z)), w_k = 1 at kx = 0 (both +-ky rows are stored) and 2 beyond.
// Block = 256 particles; the modes stream through LDS in tiles of kTile; blockIdx.y takes a
// slice of the tiles (partial sums part[slice][3][np], summed in order by k_sum_slices).
// Basis values at the particle's height are held in registers (LMAX >= L, zero-padded).
// ---------------------------------------------------------------------------------------
template <typename T, int LMAX>
__global__ void __launch_bounds__(256) k_vel_points(DevGeom g, int np, const float* __restrict__ px,
                                                    const float* __restrict__ py, const float* __restrict__ pz,
                                                    const cpx<T>* __restrict__ Xf, const cpx<T>* __restrict__ Xg,
                                                    const cpx<T>* __restrict__ XF, const cpx<T>* __restrict__ XG,
                                                    const double* __restrict__ lam, const float4* __restrict__ modes,
                                                    int tiles_per_slice, float* __restrict__ part) {
    __shared__ float4 sc[kTile][LMAX];  // (f.re, f.im, g.re, g.im)
    __shared__ float4 sm[kTile];        // (kx index, ky index, kx, ky)
    const int L = g.L, P = g.P;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int slice = blockIdx.y;
    const bool live = i < np;
    const float x = live ? px[i] : 0.f, y = live ? py[i] : 0.f, z = live ? pz[i] : 0.f;
    float cp[LMAX], cc[LMAX], ss[LMAX];
#pragma unroll
    for (int m = 0; m < LMAX; ++m) {
        double C = 0, Cp = 0, S = 0;
        if (m < L) basis_at(m, lam[m], (double)z, C, Cp, S);
        cp[m] = (float)Cp;
        cc[m] = (float)C;
        ss[m] = (float)S;
    }
    // periodic coordinates in units of the box
    const float xr = x / (float)g.Lx - floorf(x / (float)g.Lx);
    const float yr = y / (float)g.Ly - floorf(y / (float)g.Ly);
    float u = 0.f, v = 0.f, w = 0.f;
    if (slice == 0) {  // mean flow (k = 0): u += sum S_m Re F_m, v += sum S_m Re G_m
#pragma unroll
        for (int m = 0; m < LMAX; ++m)
            if (m < L) {
                u += ss[m] * (float)XF[m].x;
                v += ss[m] * (float)XG[m].x;
            }
    }
    const int ntile = (P + kTile - 1) / kTile;
    const int t0 = slice * tiles_per_slice, t1 = min(ntile, t0 + tiles_per_slice);
    for (int t = t0; t < t1; ++t) {
        const int pbase = t * kTile, nm = min(kTile, P - pbase);
        __syncthreads();
        for (int e = threadIdx.x; e < kTile * LMAX; e += blockDim.x) {
            const int j = e / LMAX, m = e - j * LMAX;
            float4 val = make_float4(0.f, 0.f, 0.f, 0.f);
            if (j < nm && m < L) {
                const cpx<T> f = Xf[(size_t)(pbase + j) * L + m], gg = Xg[(size_t)(pbase + j) * L + m];
                val = make_float4((float)f.x, (float)f.y, (float)gg.x, (float)gg.y);
            }
            sc[j][m] = val;
        }
        for (int j = threadIdx.x; j < kTile; j += blockDim.x)
            sm[j] = j < nm ? modes[pbase + j] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        for (int j = 0; j < nm; ++j) {
            float ar = 0.f, ai = 0.f, br = 0.f, bi = 0.f, cr = 0.f, ci = 0.f;
#pragma unroll
            for (int m = 0; m < LMAX; ++m) {
                const float4 c = sc[j][m];
                ar += c.x * cp[m];
                ai += c.y * cp[m];
                cr += c.x * cc[m];
                ci += c.y * cc[m];
                br += c.z * ss[m];
                bi += c.w * ss[m];
            }
            const float4 md = sm[j];
            float ph = md.x * xr + md.y * yr;
            ph -= rintf(ph);
            float sn, cs;
            sincospif(2.0f * ph, &sn, &cs);
            const float wk = md.x == 0.f ? 1.f : 2.f;
            const float kx = md.z, ky = md.w;
            // u_k = i kx A + i ky B, v_k = i ky A - i kx B, w_k = k^2 Cc
            const float ur = -kx * ai - ky * bi, ui = kx * ar + ky * br;
            const float vr = -ky * ai + kx * bi, vi = ky * ar - kx * br;
            const float k2 = kx * kx + ky * ky;
            u += wk * (cs * ur - sn * ui);
            v += wk * (cs * vr - sn * vi);
            w += wk * k2 * (cs * cr - sn * ci);
        }
    }
    if (live) {
        float* o = part + (size_t)slice * 3 * np;
        o[i] = u;
        o[np + i] = v;
        o[2 * (size_t)np + i] = w;
    }
}

// The same sums on the FP32 matrix cores (FP32 state, L <= 32).  Block = 8 waves x 16
// particles.  For a tile of 16 modes the six vertical sums of every (mode, particle) pair,
// A = sum f_m C'_m, Cc = sum f_m C_m and B = sum g_m S_m with their real and imaginary parts,
// are products D[mode][particle] = sum_m coef[mode][m] basis[m][particle] with k = m:
// v_mfma_f32_16x16x4_f32 (K = 4), A operand = staged coefficients (row = mode = lane % 16,
// k = lane / 16), B operand = basis values at the lane's particle (column = particle = lane % 16,
// k = lane / 16).  The phase and the velocity combination then run on the VALU in the MFMA output
// layout: lane l holds column l % 16 (its particle) and the mode rows 4 (l / 16) + r, r < 4, so the
// four 16-lane groups hold complementary rows of the same particles and are added at the end.
// The weight w_k (1 at kx = 0, else 2) is folded into the staged coefficients.  The modes come
// from the padded list (padded_mode_table), where the 4 rows of a lane group share kx and have
// consecutive ky: one sincospi per group gives exp(2 pi i k.x) of its first row, and the others
// follow by a complex product with exp(2 pi i r y / Ly), r = 1, 2, 3, taken once per particle.
// The coefficients of the next 64-mode tile are loaded into registers while the current tile is
// computed, across LDS-only barriers.
typedef float vpf32x4 __attribute__((ext_vector_type(4)));
constexpr int kVelMmaThreads = 512;  // 8 waves

// phases and velocity combination of one 16-mode sub-tile in the MFMA output layout of
// v_mfma_f32_16x16x*: lane l holds its particle's (column l % 16) mode rows b0 + r, r < 4, with
// b0 = 16 sub + 4 (l / 16); acc = A.re, A.im, Cc.re, Cc.im, B.re, B.im
__device__ __forceinline__ void vel_combine(const float2* sg, const float4* sm, int b0, float xr, float yr,
                                            const float (&dc)[4], const float (&ds)[4], const vpf32x4 (&acc)[6],
                                            float& u, float& v, float& w) {
    float s0, c0;
    {
        const float2 m0 = sg[b0 >> 2];
        float ph = m0.x * xr + m0.y * yr;
        ph -= rintf(ph);
        sincospif(2.0f * ph, &s0, &c0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float4 md = sm[b0 + r];
        const float cs = r == 0 ? c0 : c0 * dc[r] - s0 * ds[r];
        const float sn = r == 0 ? s0 : s0 * dc[r] + c0 * ds[r];
        const float kx = md.x, ky = md.y, k2 = md.z;
        // u_k = i kx A + i ky B, v_k = i ky A - i kx B, w_k = k^2 Cc; with
        // PA = Im(e A) = cs A.im + sn A.re and PB likewise, Re(e u_k) = -kx PA - ky PB and
        // Re(e v_k) = -ky PA + kx PB (e = cs + i sn)
        const float pa = cs * acc[1][r] + sn * acc[0][r], pb = cs * acc[5][r] + sn * acc[4][r];
        u -= kx * pa + ky * pb;
        v += kx * pb - ky * pa;
        w += k2 * (cs * acc[2][r] - sn * acc[3][r]);
    }
}
template <int LMAX>
__global__ void __launch_bounds__(kVelMmaThreads) k_vel_points_mma(
    DevGeom g, int np, const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const cpx<float>* __restrict__ Xf, const cpx<float>* __restrict__ Xg, const cpx<float>* __restrict__ XF,
    const cpx<float>* __restrict__ XG, const double* __restrict__ lam, const float4* __restrict__ modes,
    const int* __restrict__ pmap, int npad, int tiles_per_slice, float* __restrict__ part, const float4* __restrict__ zb) {
    static_assert(LMAX % 4 == 0 && LMAX <= 32, "k-steps of 4 modes");
    constexpr int KS = LMAX / 4;
    constexpr int PER = (kTile * LMAX + kVelMmaThreads - 1) / kVelMmaThreads;  // staged values per thread
    __shared__ float4 sc[kTile][LMAX + 1];  // w_k (f.re, f.im, g.re, g.im); odd row length spreads the banks
    __shared__ float4 sm[kTile];            // (kx, ky, k^2, 0)
    __shared__ float2 sg[kTile / 4];        // (kx index, ky index) of the first mode of each group of 4
    const int L = g.L, P = g.P, NYc = g.NYc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = lane & 15, q = lane >> 4;
    const int i = blockIdx.x * (kVelMmaThreads / 4) + wave * 16 + j;
    const int slice = blockIdx.y;
    const bool live = i < np;
    const float x = live ? px[i] : 0.f, y = live ? py[i] : 0.f, z = live ? pz[i] : 0.f;
    // basis values of the lane's k index m = 4 s + q at the particle's height
    float bcp[KS], bcc[KS], bss[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int m = 4 * s + q;
        bcp[s] = bcc[s] = bss[s] = 0.f;
        if (m < L) {
            if (zb) {  // computed once per evaluation by k_point_basis
                if (live) {
                    const float4 t = zb[(size_t)i * L + m];
                    bcp[s] = t.x;
                    bcc[s] = t.y;
                    bss[s] = t.z;
                }
            } else {
                double C = 0, Cp = 0, S = 0;
                basis_at(m, lam[m], (double)z, C, Cp, S);
                bcp[s] = (float)Cp;
                bcc[s] = (float)C;
                bss[s] = (float)S;
            }
        }
    }
    const float xr = x / (float)g.Lx - floorf(x / (float)g.Lx);
    const float yr = y / (float)g.Ly - floorf(y / (float)g.Ly);
    // exp(2 pi i r yr), r = 1, 2, 3
    float ds[4], dc[4];
    sincospif(2.0f * yr, &ds[1], &dc[1]);
    dc[2] = dc[1] * dc[1] - ds[1] * ds[1];
    ds[2] = 2.f * ds[1] * dc[1];
    dc[3] = dc[2] * dc[1] - ds[2] * ds[1];
    ds[3] = ds[2] * dc[1] + dc[2] * ds[1];
    float u = 0.f, v = 0.f, w = 0.f;
    if (slice == 0) {  // mean flow (k = 0), each lane group its residue of m modulo 4
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int m = 4 * s + q;
            if (m < L) {
                u += bss[s] * XF[m].x;
                v += bss[s] * XG[m].x;
            }
        }
    }
    const int ntile = (npad + kTile - 1) / kTile;
    const int t0 = slice * tiles_per_slice, t1 = min(ntile, t0 + tiles_per_slice);
    // the coefficients of tile t into registers: value e = threadIdx.x + u * threads of the tile
    float4 pre[PER], pmd = make_float4(0.f, 0.f, 0.f, 0.f);
    auto fetch = [&](int t) {
        const int pbase = t * kTile, nm = min(kTile, npad - pbase);
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = threadIdx.x + u * kVelMmaThreads, jj = e / LMAX, m = e - jj * LMAX;
            float4 val = make_float4(0.f, 0.f, 0.f, 0.f);
            const int p = (e < kTile * LMAX && jj < nm) ? pmap[pbase + jj] : -1;
            if (p >= 0 && p < P && m < L) {
                const cpx<float> f = Xf[(size_t)p * L + m], gg = Xg[(size_t)p * L + m];
                const float wk = p < NYc ? 1.f : 2.f;  // compact p < NYc <=> kx = 0
                val = make_float4(wk * f.x, wk * f.y, wk * gg.x, wk * gg.y);
            }
            pre[u] = val;
        }
        if (threadIdx.x < kTile)
            pmd = (int)threadIdx.x < nm ? modes[pbase + threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    if (t0 < t1) fetch(t0);
    for (int t = t0; t < t1; ++t) {
        const int nm = min(kTile, npad - t * kTile);
        lds_barrier();  // the previous tile is consumed
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = threadIdx.x + u * kVelMmaThreads;
            if (e < kTile * LMAX) sc[e / LMAX][e % LMAX] = pre[u];
        }
        if (threadIdx.x < kTile) {
            sm[threadIdx.x] = make_float4(pmd.z, pmd.w, pmd.z * pmd.z + pmd.w * pmd.w, 0.f);
            if ((threadIdx.x & 3) == 0) sg[threadIdx.x >> 2] = make_float2(pmd.x, pmd.y);
        }
        lds_barrier();
        if (t + 1 < t1) fetch(t + 1);  // in flight while tile t is computed
        const int nsub = (nm + 15) / 16;
        // the six products of a 16-mode sub-tile on the matrix cores
        auto products = [&](int sub, vpf32x4 (&acc)[6]) {
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[k] = vpf32x4{0.f, 0.f, 0.f, 0.f};
            const float4* row = sc[sub * 16 + j];
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const float4 c4 = row[4 * s + q];
                acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(c4.x, bcp[s], acc[0], 0, 0, 0);  // A.re
                acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(c4.y, bcp[s], acc[1], 0, 0, 0);  // A.im
                acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(c4.x, bcc[s], acc[2], 0, 0, 0);  // Cc.re
                acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(c4.y, bcc[s], acc[3], 0, 0, 0);  // Cc.im
                acc[4] = __builtin_amdgcn_mfma_f32_16x16x4f32(c4.z, bss[s], acc[4], 0, 0, 0);  // B.re
                acc[5] = __builtin_amdgcn_mfma_f32_16x16x4f32(c4.w, bss[s], acc[5], 0, 0, 0);  // B.im
            }
        };
        auto combine = [&](int sub, const vpf32x4 (&acc)[6]) {
            vel_combine(sg, sm, sub * 16 + 4 * q, xr, yr, dc, ds, acc, u, v, w);
        };
        // software pipeline: the products of sub-tile sub + 1 are issued before the combination of
        // sub-tile sub, so that the matrix cores and the VALU work at the same time
        vpf32x4 acc0[6], acc1[6];
        products(0, acc0);
#pragma unroll
        for (int sub = 0; sub < kTile / 16; sub += 2) {
            if (sub + 1 < nsub) products(sub + 1, acc1);
            if (sub < nsub) combine(sub, acc0);
            if (sub + 2 < nsub) products(sub + 2, acc0);
            if (sub + 1 < nsub) combine(sub + 1, acc1);
        }
    }
    u += __shfl_xor(u, 16);
    v += __shfl_xor(v, 16);
    w += __shfl_xor(w, 16);
    u += __shfl_xor(u, 32);
    v += __shfl_xor(v, 32);
    w += __shfl_xor(w, 32);
    if (live && q == 0) {
        float* o = part + (size_t)slice * 3 * np;
        o[i] = u;
        o[np + i] = v;
        o[2 * (size_t)np + i] = w;
    }
}

// BF16X3 form of k_vel_points_mma (contexts of precision RBC_BF16X3): the same six vertical sums
// with every FP32 operand split into BF16 hi + lo halves and a.b ~ ah.bh + ah.bl + al.bh
// accumulated in FP32 (v_mfma_f32_16x16x32_bf16, K = 32 in one instruction).  Lane (j, q) takes the
// KPL = LMAX / 4 modes m = KPL q + e, e < KPL, at k = 8 q + e and zeros at the other k of its
// slice, so that the basis values per lane stay balanced for every L.  The coefficients are split
// once, when a tile is staged ([component][hi, lo][mode][m] BF16 in LDS); the basis values once per
// particle.  Phases, combination, slices and prefetch as in k_vel_points_mma.
typedef __bf16 vpbf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int vpu32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ unsigned short bf16_bits(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
__device__ __forceinline__ float bf16_value(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
template <int LMAX>
__global__ void __launch_bounds__(kVelMmaThreads) k_vel_points_bf3(
    DevGeom g, int np, const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const cpx<float>* __restrict__ Xf, const cpx<float>* __restrict__ Xg, const cpx<float>* __restrict__ XF,
    const cpx<float>* __restrict__ XG, const double* __restrict__ lam, const float4* __restrict__ modes,
    const int* __restrict__ pmap, int npad, int tiles_per_slice, float* __restrict__ part, const float4* __restrict__ zb) {
    static_assert(LMAX % 8 == 0 && LMAX <= 32, "KPL = LMAX / 4 even, at most 8");
    constexpr int KPL = LMAX / 4;
    constexpr int RS = LMAX == 24 ? 40 : LMAX + 8;  // LDS row (BF16 values): 8-byte aligned, spreads the banks
    constexpr int NPAIR = kTile * LMAX / 2;         // (mode, pair of m) entries of a tile
    constexpr int PER = (NPAIR + kVelMmaThreads - 1) / kVelMmaThreads;
    __shared__ unsigned int sa[4][2][kTile][RS / 2];  // [f.re, f.im, g.re, g.im][hi, lo][mode][m pair]
    __shared__ float4 sm[kTile];                      // (kx, ky, k^2, 0)
    __shared__ float2 sg[kTile / 4];                  // (kx index, ky index) of the first mode of each group of 4
    const int L = g.L, P = g.P, NYc = g.NYc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = lane & 15, q = lane >> 4;
    const int i = blockIdx.x * (kVelMmaThreads / 4) + wave * 16 + j;
    const int slice = blockIdx.y;
    const bool live = i < np;
    const float x = live ? px[i] : 0.f, y = live ? py[i] : 0.f, z = live ? pz[i] : 0.f;
    // basis values of m = KPL q + e at the particle's height, split into BF16 hi and lo
    vpbf16x8 bh[3], bl[3];  // C', C, S
    float ssf[KPL];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int m = KPL * q + e;
        float val[3] = {0.f, 0.f, 0.f};
        if (e < KPL && m < L) {
            if (zb) {  // computed once per evaluation by k_point_basis
                if (live) {
                    const float4 t = zb[(size_t)i * L + m];
                    val[0] = t.x;
                    val[1] = t.y;
                    val[2] = t.z;
                }
            } else {
                double C = 0, Cp = 0, S = 0;
                basis_at(m, lam[m], (double)z, C, Cp, S);
                val[0] = (float)Cp;
                val[1] = (float)C;
                val[2] = (float)S;
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const unsigned short h = bf16_bits(val[k]);
            bh[k][e] = __builtin_bit_cast(__bf16, h);
            bl[k][e] = (__bf16)(val[k] - bf16_value(h));
        }
        if (e < KPL) ssf[e] = val[2];
    }
    const float xr = x / (float)g.Lx - floorf(x / (float)g.Lx);
    const float yr = y / (float)g.Ly - floorf(y / (float)g.Ly);
    float ds[4], dc[4];
    sincospif(2.0f * yr, &ds[1], &dc[1]);
    dc[2] = dc[1] * dc[1] - ds[1] * ds[1];
    ds[2] = 2.f * ds[1] * dc[1];
    dc[3] = dc[2] * dc[1] - ds[2] * ds[1];
    ds[3] = ds[2] * dc[1] + dc[2] * ds[1];
    float u = 0.f, v = 0.f, w = 0.f;
    if (slice == 0) {  // mean flow (k = 0), each lane group its KPL modes
#pragma unroll
        for (int e = 0; e < KPL; ++e) {
            const int m = KPL * q + e;
            if (m < L) {
                u += ssf[e] * XF[m].x;
                v += ssf[e] * XG[m].x;
            }
        }
    }
    const int ntile = (npad + kTile - 1) / kTile;
    const int t0 = slice * tiles_per_slice, t1 = min(ntile, t0 + tiles_per_slice);
    // pairs (m, m + 1) of the next tile's coefficients, times w_k, into registers
    float4 pf[PER], pg[PER];
    float4 pmd = make_float4(0.f, 0.f, 0.f, 0.f);
    auto fetch = [&](int t) {
        const int pbase = t * kTile, nm = min(kTile, npad - pbase);
#pragma unroll
        for (int uu = 0; uu < PER; ++uu) {
            const int e = threadIdx.x + uu * kVelMmaThreads, jj = e / (LMAX / 2), m = 2 * (e - jj * (LMAX / 2));
            float4 fv = make_float4(0.f, 0.f, 0.f, 0.f), gv = fv;
            const int p = (e < NPAIR && jj < nm) ? pmap[pbase + jj] : -1;
            if (p >= 0 && p < P && m < L) {
                const float wk = p < NYc ? 1.f : 2.f;  // compact p < NYc <=> kx = 0
                const cpx<float> f0 = Xf[(size_t)p * L + m], g0 = Xg[(size_t)p * L + m];
                const cpx<float> f1 = m + 1 < L ? Xf[(size_t)p * L + m + 1] : mk(0.f, 0.f);
                const cpx<float> g1 = m + 1 < L ? Xg[(size_t)p * L + m + 1] : mk(0.f, 0.f);
                fv = make_float4(wk * f0.x, wk * f0.y, wk * f1.x, wk * f1.y);
                gv = make_float4(wk * g0.x, wk * g0.y, wk * g1.x, wk * g1.y);
            }
            pf[uu] = fv;
            pg[uu] = gv;
        }
        if (threadIdx.x < kTile)
            pmd = (int)threadIdx.x < nm ? modes[pbase + threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    // hi and lo BF16 of (a, b) packed into two 32-bit words (a in the low half)
    auto split2 = [](float a, float b, unsigned int& hi, unsigned int& lo) {
        const unsigned short ha = bf16_bits(a), hb = bf16_bits(b);
        const unsigned short la = bf16_bits(a - bf16_value(ha)), lb = bf16_bits(b - bf16_value(hb));
        hi = (unsigned int)ha | ((unsigned int)hb << 16);
        lo = (unsigned int)la | ((unsigned int)lb << 16);
    };
    if (t0 < t1) fetch(t0);
    for (int t = t0; t < t1; ++t) {
        const int nm = min(kTile, npad - t * kTile);
        lds_barrier();  // the previous tile is consumed
#pragma unroll
        for (int uu = 0; uu < PER; ++uu) {
            const int e = threadIdx.x + uu * kVelMmaThreads;
            if (e < NPAIR) {
                const int jj = e / (LMAX / 2), mp = e - jj * (LMAX / 2);
                unsigned int hi, lo;
                split2(pf[uu].x, pf[uu].z, hi, lo);  // f.re at m, m + 1
                sa[0][0][jj][mp] = hi;
                sa[0][1][jj][mp] = lo;
                split2(pf[uu].y, pf[uu].w, hi, lo);  // f.im
                sa[1][0][jj][mp] = hi;
                sa[1][1][jj][mp] = lo;
                split2(pg[uu].x, pg[uu].z, hi, lo);  // g.re
                sa[2][0][jj][mp] = hi;
                sa[2][1][jj][mp] = lo;
                split2(pg[uu].y, pg[uu].w, hi, lo);  // g.im
                sa[3][0][jj][mp] = hi;
                sa[3][1][jj][mp] = lo;
            }
        }
        if (threadIdx.x < kTile) {
            sm[threadIdx.x] = make_float4(pmd.z, pmd.w, pmd.z * pmd.z + pmd.w * pmd.w, 0.f);
            if ((threadIdx.x & 3) == 0) sg[threadIdx.x >> 2] = make_float2(pmd.x, pmd.y);
        }
        lds_barrier();
        if (t + 1 < t1) fetch(t + 1);  // in flight while tile t is computed
        const int nsub = (nm + 15) / 16;
        auto products = [&](int sub, vpf32x4 (&acc)[6]) {
            const int row = sub * 16 + j;
            vpbf16x8 ah[4], al[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                vpu32x4 wh = {0u, 0u, 0u, 0u}, wl = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int e2 = 0; e2 < KPL / 2; ++e2) {
                    wh[e2] = sa[c][0][row][(KPL * q) / 2 + e2];
                    wl[e2] = sa[c][1][row][(KPL * q) / 2 + e2];
                }
                ah[c] = __builtin_bit_cast(vpbf16x8, wh);
                al[c] = __builtin_bit_cast(vpbf16x8, wl);
            }
            static constexpr int comp[6] = {0, 1, 0, 1, 2, 3}, kind[6] = {0, 0, 1, 1, 2, 2};
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                vpf32x4 a4 = {0.f, 0.f, 0.f, 0.f};
                a4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[comp[k]], bh[kind[k]], a4, 0, 0, 0);
                a4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[comp[k]], bl[kind[k]], a4, 0, 0, 0);
                a4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[comp[k]], bh[kind[k]], a4, 0, 0, 0);
                acc[k] = a4;
            }
        };
        auto combine = [&](int sub, const vpf32x4 (&acc)[6]) {
            vel_combine(sg, sm, sub * 16 + 4 * q, xr, yr, dc, ds, acc, u, v, w);
        };
        vpf32x4 acc0[6], acc1[6];
        products(0, acc0);
#pragma unroll
        for (int sub = 0; sub < kTile / 16; sub += 2) {
            if (sub + 1 < nsub) products(sub + 1, acc1);
            if (sub < nsub) combine(sub, acc0);
            if (sub + 2 < nsub) products(sub + 2, acc0);
            if (sub + 1 < nsub) combine(sub + 1, acc1);
        }
    }
    u += __shfl_xor(u, 16);
    v += __shfl_xor(v, 16);
    w += __shfl_xor(w, 16);
    u += __shfl_xor(u, 32);
    v += __shfl_xor(v, 32);
    w += __shfl_xor(w, 32);
    if (live && q == 0) {
        float* o = part + (size_t)slice * 3 * np;
        o[i] = u;
        o[np + i] = v;
        o[2 * (size_t)np + i] = w;
    }
}

__global__ void k_sum_slices(int np, int nslice, const float* __restrict__ part, float* __restrict__ u,
                             float* __restrict__ v, float* __restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float a = 0.f, b = 0.f, c = 0.f;
    for (int s = 0; s < nslice; ++s) {
        const float* o = part + (size_t)s * 3 * np;
        a += o[i];
        b += o[np + i];
        c += o[2 * (size_t)np + i];
    }
    u[i] = a;
    v[i] = b;
    w[i] = c;
}

// midpoint position x + (dt/2) u (the spectral RK2 half step)
__global__ void k_tracer_half(int np, const float* __restrict__ x0, const float* __restrict__ y0,
                              const float* __restrict__ z0, const float* __restrict__ u, const float* __restrict__ v,
                              const float* __restrict__ w, float hdt, float* __restrict__ xm, float* __restrict__ ym,
                              float* __restrict__ zm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    xm[i] = x0[i] + hdt * u[i];
    ym[i] = y0[i] + hdt * v[i];
    zm[i] = z0[i] + hdt * w[i];
}

// x1 = x0 + dt u(midpoint), then diffusion and the lid reflection
__global__ void k_tracer_full(int np, const float* __restrict__ x0, const float* __restrict__ y0,
                              const float* __restrict__ z0, const float* __restrict__ u, const float* __restrict__ v,
                              const float* __restrict__ w, float dt, float* __restrict__ x1, float* __restrict__ y1,
                              float* __restrict__ z1, float sigma, uint64_t seed, uint64_t ctr0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float x = x0[i] + dt * u[i], y = y0[i] + dt * v[i], z = z0[i] + dt * w[i];
    finish_particle(x, y, z, i, np, sigma, seed, ctr0);
    x1[i] = x;
    y1[i] = y;
    z1[i] = z;
}

// initial positions (particle_tracer.cu:81-135, 390-453) with counter-based random numbers
__global__ void k_tracer_init(int type, int np, float Lx, float Ly, float cx, float cy, float cz, float sigma,
                              uint64_t seed, float* __restrict__ x, float* __restrict__ y, float* __restrict__ z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float a, b, c;
    switch (type) {
        case RBC_PART_RANDOM:
            a = Lx * rng_u01(seed, 3 * (uint64_t)i);
            b = Ly * rng_u01(seed, 3 * (uint64_t)i + 1);
            c = rng_u01(seed, 3 * (uint64_t)i + 2) - 0.5f;
            break;
        case RBC_PART_DIAGONAL:
            a = (Lx * i) / float(np);
            b = (Ly * i) / float(np);
            c = 0.f;
            break;
        case RBC_PART_PAIRS: {
            const uint64_t pi = (uint64_t)(i / 2);
            a = Lx * rng_u01(seed, 2 * pi);
            b = Ly * rng_u01(seed, 2 * pi + 1);
            c = 0.f;
            break;
        }
        default:  // center, center_sigma, coord_sigma
            a = cx;
            b = cy;
            c = cz;
            if (sigma > 0.f) {
                a += sigma * rng_normal(seed, (uint64_t)i);
                b += sigma * rng_normal(seed, (uint64_t)np + i);
                c += sigma * rng_normal(seed, 2 * (uint64_t)np + i);
            }
    }
    x[i] = a;
    y[i] = b;
    z[i] = c;
}

// moments of the current positions in FP64, one block, fixed order (particle_tracer.cu:771-821
// computes the same quantities in the device real type):
// out = {<x>, <y>, <x^2+y^2> - <x>^2 - <y>^2, <|x_2i - x_2i+1|^2> (minimum image)}
__global__ void __launch_bounds__(1024) k_tracer_moments(int np, const float* __restrict__ x,
                                                         const float* __restrict__ y, float Lx, float Ly,
                                                         double* __restrict__ out) {
    __shared__ double s[4][1024];
    double a = 0, b = 0, c = 0, d = 0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        const double xi = x[i], yi = y[i];
        a += xi;
        b += yi;
        c += xi * xi + yi * yi;
    }
    for (int j = threadIdx.x; 2 * j + 1 < np; j += blockDim.x) {
        float dx = fabsf(x[2 * j + 1] - x[2 * j]), dy = fabsf(y[2 * j + 1] - y[2 * j]);
        if (dx > 0.5f * Lx) dx = Lx - dx;
        if (dy > 0.5f * Ly) dy = Ly - dy;
        d += (double)dx * dx + (double)dy * dy;
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    s[2][threadIdx.x] = c;
    s[3][threadIdx.x] = d;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 4; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double mx = s[0][0] / np, my = s[1][0] / np;
        out[0] = mx;
        out[1] = my;
        out[2] = s[2][0] / np - mx * mx - my * my;
        out[3] = np >= 2 ? s[3][0] / (np / 2) : 0.0;
    }
}

// ---------------------------------------------------------------------------------------
// Kinetic energy per mode, FP64: out[c][p] = sum_l wz_l |c_k(z_l)|^2 for c = u, v, w.
// ---------------------------------------------------------------------------------------
// Characteristic velocity <|u|> over a velocity grid (main_extract_vel.cu:182-218): per block
// the sum of sqrt(u^2 + v^2 + w^2) over its points (the float components squared and summed in
// FP64, points whose sum is NaN skipped) and the count of points summed; the host adds the
// blocks in order.
constexpr int kCharvelBlocks = 1024;
__global__ void __launch_bounds__(256) k_charvel(size_t n, const float* __restrict__ grid, double* __restrict__ part) {
    __shared__ double ss[256], sc[256];
    double s = 0.0, cnt = 0.0;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        double val = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double v = grid[(size_t)c * n + e];
            val += v * v;
        }
        if (val == val) {
            cnt += 1.0;
            s += sqrt(val);
        }
    }
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ss[threadIdx.x] += ss[threadIdx.x + w];
            sc[threadIdx.x] += sc[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ss[0];
        part[2 * blockIdx.x + 1] = sc[0];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_energy_modes(DevGeom g, int nz, const double* __restrict__ zb,
                                                      const double* __restrict__ wz, const cpx<T>* __restrict__ Xf,
                                                      const cpx<T>* __restrict__ Xg, const cpx<T>* __restrict__ XF,
                                                      const cpx<T>* __restrict__ XG, const T* __restrict__ kxa,
                                                      const T* __restrict__ kya, double* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const int L = g.L;
    const double kx = kxa[p], ky = kya[p], k2 = kx * kx + ky * ky;
    double eu = 0, ev = 0, ew = 0;
    for (int l = 0; l < nz; ++l) {
        const double* tb = zb + (size_t)l * 3 * L;
        cpx<double> A = mk<double>(0, 0), B = A, Cc = A;
        for (int m = 0; m < L; ++m) {
            const cpx<T> f = Xf[(size_t)p * L + m], gg = Xg[(size_t)p * L + m];
            A = A + cscale(mk<double>(f.x, f.y), tb[m]);
            Cc = Cc + cscale(mk<double>(f.x, f.y), tb[L + m]);
            B = B + cscale(mk<double>(gg.x, gg.y), tb[2 * L + m]);
        }
        cpx<double> u = imul(A, kx) + imul(B, ky), v = imul(A, ky) - imul(B, kx);
        if (p == 0)
            for (int m = 0; m < L; ++m) {
                u = u + cscale(mk<double>(XF[m].x, XF[m].y), tb[2 * L + m]);
                v = v + cscale(mk<double>(XG[m].x, XG[m].y), tb[2 * L + m]);
            }
        const cpx<double> w = cscale(Cc, k2);
        eu += wz[l] * (u.x * u.x + u.y * u.y);
        ev += wz[l] * (v.x * v.x + v.y * v.y);
        ew += wz[l] * (w.x * w.x + w.y * w.y);
    }
    out[p] = eu;
    out[g.P + p] = ev;
    out[2 * (size_t)g.P + p] = ew;
}

// ---- host helpers ------------------------------------------------------------------------

// basis table zb[l][3][L] = C'_m, C_m, S_m at heights z (FP64)
std::vector<double> basis_table(int L, const std::vector<double>& z) {
    std::vector<double> t((size_t)z.size() * 3 * L);
    for (size_t l = 0; l < z.size(); ++l)
        for (int m = 0; m < L; ++m) {
            double C, Cp, S;
            basis_at(m, lambda_k(m + 1), z[l], C, Cp, S);
            t[(l * 3 + 0) * L + m] = Cp;
            t[(l * 3 + 1) * L + m] = C;
            t[(l * 3 + 2) * L + m] = S;
        }
    return t;
}

std::vector<float> twiddles_f(int n) {
    std::vector<float> t(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        const double a = -2.0 * M_PI * k / n;
        t[2 * k] = (float)std::cos(a);
        t[2 * k + 1] = (float)std::sin(a);
    }
    return t;
}

// mode table (kx index, ky index, kx, ky) per compact mode
std::vector<float> mode_table(const rbc_ctx* c) {
    const Geometry& g = c->geo;
    const int NYc = g.NYc();
    std::vector<float> t(4 * (size_t)g.P());
    for (int p = 0; p < g.P(); ++p) {
        const int kx = p / NYc, row = g.rows[p % NYc];
        const int ky = row <= g.N / 2 ? row : row - g.N;
        t[4 * p] = (float)kx;
        t[4 * p + 1] = (float)ky;
        t[4 * p + 2] = (float)c->fac.kx[p];
        t[4 * p + 3] = (float)c->fac.ky[p];
    }
    return t;
}

// Mode list of k_vel_points_mma: per kx column, every run of consecutive ky (the ky >= 0 and
// ky < 0 halves of the kept rows) padded to a multiple of 4 with slots of zero coefficient whose
// ky continues the run, so that every aligned group of 4 slots has one kx and consecutive ky.
// t: (kx index, ky index, kx, ky) per slot; map: compact mode of the slot or -1 (padding).
void padded_mode_table(const rbc_ctx* c, std::vector<float>& t, std::vector<int>& map) {
    const Geometry& g = c->geo;
    const int NYc = g.NYc();
    const double Lx = c->p.Lx, Ly = c->p.Ly;
    t.clear();
    map.clear();
    auto ky_of = [&](int kyc) { const int row = g.rows[kyc]; return row <= g.N / 2 ? row : row - g.N; };
    auto push = [&](int kx, int ky, int p) {
        t.push_back((float)kx);
        t.push_back((float)ky);
        t.push_back((float)((M_PI + M_PI) * (kx / Lx)));
        t.push_back((float)((M_PI + M_PI) * (ky / Ly)));
        map.push_back(p);
    };
    for (int kx = 0; kx < g.KX; ++kx) {
        int kyc = 0;
        while (kyc < NYc) {
            int len = 1;
            while (kyc + len < NYc && ky_of(kyc + len) == ky_of(kyc) + len) ++len;
            for (int e = 0; e < ((len + 3) & ~3); ++e)
                push(kx, ky_of(kyc) + e, e < len ? kx * NYc + kyc + e : -1);
            kyc += len;
        }
    }
}

template <typename T>
void launch_vel_levels(rbc_ctx* c, const DevGeom& gt, int nz, const double* zb, void* spec) {
    const Buffers& b = c->b;
    dim3 grid((c->dg.P + 255) / 256, nz);
    hipLaunchKernelGGL((k_vel_levels<T, float>), grid, dim3(256), 0, c->stream, gt, nz, zb, (const cpx<T>*)b.X[1],
                       (const cpx<T>*)b.X[2], (const cpx<T>*)b.XF, (const cpx<T>*)b.XG, (const T*)b.kx,
                       (const T*)b.ky, (cpx<float>*)spec);
}

// scratch of one velocity-grid evaluation at nz heights
struct GridWork {
    double* zb = nullptr;       // [nz][3][L]
    cpx<float>* spec = nullptr; // [3][nz][P]
    cpx<float>* hyb = nullptr;  // [3][nz][N][KXp]
    cpx<float>* twM = nullptr;
    cpx<float>* twN = nullptr;
    float* grid = nullptr;      // [3][nz][N][M]
    int nz = 0;
};

void alloc_grid_work(rbc_ctx* c, int nz, GridWork& w, std::vector<void*>& owned) {
    const DevGeom& g = c->dg;
    auto al = [&](size_t bytes) {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, bytes));
        owned.push_back(p);
        return p;
    };
    w.nz = nz;
    w.zb = (double*)al((size_t)nz * 3 * g.L * sizeof(double));
    w.spec = (cpx<float>*)al((size_t)3 * nz * g.P * sizeof(cpx<float>));
    w.hyb = (cpx<float>*)al((size_t)3 * nz * g.N * g.KXp * sizeof(cpx<float>));
    w.grid = (float*)al((size_t)3 * nz * g.N * g.M * sizeof(float));
    w.twM = (cpx<float>*)al((size_t)g.M * sizeof(cpx<float>));
    w.twN = (cpx<float>*)al((size_t)g.N * sizeof(cpx<float>));
    const std::vector<float> tm = twiddles_f(g.M), tn = twiddles_f(g.N);
    HIPCHK(hipMemcpy(w.twM, tm.data(), tm.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(w.twN, tn.data(), tn.size() * sizeof(float), hipMemcpyHostToDevice));
}

void set_heights(rbc_ctx* c, GridWork& w, const std::vector<double>& z) {
    const std::vector<double> t = basis_table(c->dg.L, z);
    HIPCHK(hipMemcpy(w.zb, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
}

// velocity grid at the heights of w (spectral -> y-FFT -> x C2R), on the context's stream
void velocity_grid(rbc_ctx* c, const GridWork& w) {
    DevGeom gt = c->dg;
    gt.nq = w.nz;
    if (c->rsz == 8) launch_vel_levels<double>(c, gt, w.nz, w.zb, w.spec);
    else launch_vel_levels<float>(c, gt, w.nz, w.zb, w.spec);
    if (!launch_y_inv_group(gt, 3, (const cpx<float>*)w.spec, (const cpx<float>*)w.twN, w.hyb, c->stream)) {
        int TX = (int)(65536 / (2 * (size_t)gt.N * sizeof(cpx<float>)));
        TX = std::max(1, std::min(16, TX));
        dim3 gy((gt.KX + TX - 1) / TX, w.nz, 3);
        hipLaunchKernelGGL(k_fft_y_inv<float>, gy, dim3(256), (size_t)2 * TX * gt.N * sizeof(cpx<float>), c->stream,
                           gt, TX, (const cpx<float>*)w.spec, c->b.rows, (const cpx<float>*)w.twN, w.hyb);
    }
    hipLaunchKernelGGL(k_vel_rows, dim3(gt.N, w.nz), dim3(256), (size_t)4 * gt.M * sizeof(cpx<float>), c->stream, gt,
                       (const cpx<float>*)w.hyb, (const cpx<float>*)w.twM, w.grid);
    HIPCHK(hipGetLastError());
}

// spectral velocity at np points (device arrays); part holds nslice*3*np floats; lam and
// modes are the context's constants (VelWork)
struct PointWork {
    const double* lam = nullptr;
    const float4* modes = nullptr;   // compact modes (FP64 kernel)
    const float4* pmodes = nullptr;  // padded mode list (matrix-core kernel)
    const int* pmap = nullptr;
    int npad = 0;
    float* part = nullptr;
    float4* zb = nullptr;  // FP32 states: np x L basis values at the points (null: in the kernel)
    int nslice = 1, tiles_per_slice = 1;
};

int lmax_for(int L) { return L <= 8 ? 8 : L <= 16 ? 16 : L <= 24 ? 24 : L <= 32 ? 32 : 0; }

int padded_modes(rbc_ctx* c);

void plan_points(rbc_ctx* c, int np, PointWork& pw) {
    const int nmodes = c->rsz == 4 ? padded_modes(c) : c->dg.P;
    const int ntile = (nmodes + kTile - 1) / kTile;
    // particles per block: 128 for an FP32 state (k_vel_points_mma), 256 for FP64 (k_vel_points)
    const int per = c->rsz == 4 ? kVelMmaThreads / 4 : 256;
    const int nblk = (np + per - 1) / per;
    // enough blocks to give every SIMD of the chip several waves: 4096 blocks for an FP32 state
    // (1e5 tracers at 128^2 x 16, basis values computed once per evaluation: 2048 blocks 585,
    // 4096 610, 8192 599 tracer steps/s), 2048 for FP64
    const int want = c->rsz == 4 ? 4096 : 2048;
    int ns = (want + nblk - 1) / nblk;
    ns = std::max(1, std::min(std::min(16, ns), ntile));
    pw.tiles_per_slice = (ntile + ns - 1) / ns;
    pw.nslice = (ntile + pw.tiles_per_slice - 1) / pw.tiles_per_slice;
}

// floats of partial sums an evaluation at np points needs
size_t part_floats(rbc_ctx* c, int np) {
    PointWork pw;
    plan_points(c, np, pw);
    return (size_t)pw.nslice * 3 * np;
}

// ---- scratch of rbc_velocity_at_device_noalloc ------------------------------------------------
// The partial sums of the spectral point evaluation (part[slice][3][n]) for callers that cannot
// allocate: a file-scope device array, one copy on every GPU the library runs on, held until
// the process ends.  Its address on a device comes from hipGetSymbolAddress, taken once per
// device by rbc_create with that device current; host code uses that address, never the
// array's name.  16 MiB holds the partial sums of every n <= 1,398,101 points.
constexpr size_t kNoAllocScratchFloats = (size_t)4 << 20;
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// address of this device's copy of `scratch` (the device is current); nullptr when the runtime
// cannot give it (the no-allocation entry point then fails with RBC_ERR_HIP, nothing else does)
void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

// Per-context workspace of the velocity diagnostics.  lam, modes and the spectrum scratch are
// made by velocity_setup (rbc_create); the grid work and the point partial sums of the
// host-buffer calls grow on demand and are kept for later calls.
struct VelWork {
    double* lam = nullptr;    // L
    float4* modes = nullptr;  // P
    float4* pmodes = nullptr; // padded mode list of the matrix-core point evaluation
    int* pmap = nullptr;
    int npad = 0;
    double* ezb = nullptr;    // spectrum: basis values [nz][3][L], nz <= max(nq, 1)
    double* ew = nullptr;     // spectrum: level weights
    double* ee = nullptr;     // spectrum: per-mode energies [3][P]
    int ez_cap = 0;
    GridWork gw;              // rbc_velocity_grid, capacity gw_cap heights
    int gw_cap = 0;
    std::vector<void*> gw_owned;
    float* part = nullptr;    // rbc_velocity_at partial sums
    size_t part_cap = 0;
    float4* zbas = nullptr;   // rbc_velocity_at basis values at the points (FP32 states)
    size_t zbas_cap = 0;
    float* io = nullptr;      // rbc_velocity_at points and results, 6 n floats
    size_t io_cap = 0;
    double* cpart = nullptr;  // characteristic velocity: per block (sum, count)
    float* noalloc = nullptr; // this device's copy of `scratch` (rbc_velocity_at_device_noalloc)
    std::vector<void*> owned;
    ~VelWork() {
        for (void* p : gw_owned) hipFree(p);
        if (part) hipFree(part);
        if (zbas) hipFree(zbas);
        if (io) hipFree(io);
        for (void* p : owned) hipFree(p);
    }
};

void free_vel_work(void* p) { delete (VelWork*)p; }

VelWork* vel_work(rbc_ctx* c) { return (VelWork*)c->vel_ws; }

int padded_modes(rbc_ctx* c) { return vel_work(c)->npad; }

// the context's constants of a point evaluation
void point_constants(rbc_ctx* c, PointWork& pw) {
    VelWork* vw = vel_work(c);
    pw.lam = vw->lam;
    pw.modes = vw->modes;
    pw.pmodes = vw->pmodes;
    pw.pmap = vw->pmap;
    pw.npad = vw->npad;
}

void alloc_point_work(rbc_ctx* c, int np, PointWork& pw, std::vector<void*>& owned) {
    plan_points(c, np, pw);
    point_constants(c, pw);
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(16, (size_t)pw.nslice * 3 * np * sizeof(float))));
    owned.push_back(p);
    pw.part = (float*)p;
    if (c->rsz == 4) {
        void* q = nullptr;
        HIPCHK(hipMalloc(&q, std::max<size_t>(16, (size_t)np * c->dg.L * sizeof(float4))));
        owned.push_back(q);
        pw.zb = (float4*)q;
    }
}

// grows a device buffer to at least `bytes` (contents not kept)
template <typename P>
void grow(P*& buf, size_t& cap, size_t need) {
    if (need <= cap && buf) return;
    if (buf) HIPCHK(hipFree(buf));
    buf = nullptr;
    cap = 0;
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(16, need * sizeof(P))));
    buf = (P*)p;
    cap = need;
}

template <typename T, int LM>
void launch_points_t(rbc_ctx* c, int np, const float* x, const float* y, const float* z, const PointWork& pw,
                     float* u, float* v, float* w) {
    const Buffers& b = c->b;
    if constexpr (sizeof(T) == 4) {  // FP32 state: the vertical sums on the matrix cores
        const float4* zb = pw.zb;
        if (zb) {
            const size_t ne = ((size_t)np + 7) / 8 * c->dg.L;  // threads: (8 points, mode)
            hipLaunchKernelGGL(k_point_basis, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, c->stream, np, c->dg.L,
                               z, (const double*)pw.lam, pw.zb);
        }
        dim3 grid((np + kVelMmaThreads / 4 - 1) / (kVelMmaThreads / 4), pw.nslice);
        if (c->p.precision == RBC_BF16X3)  // BF16 hi/lo split products, as the step's vertical transforms
            hipLaunchKernelGGL((k_vel_points_bf3<LM>), grid, dim3(kVelMmaThreads), 0, c->stream, c->dg, np, x, y, z,
                               (const cpx<float>*)b.X[1], (const cpx<float>*)b.X[2], (const cpx<float>*)b.XF,
                               (const cpx<float>*)b.XG, (const double*)pw.lam, pw.pmodes, pw.pmap, pw.npad,
                               pw.tiles_per_slice, pw.part, zb);
        else
            hipLaunchKernelGGL((k_vel_points_mma<LM>), grid, dim3(kVelMmaThreads), 0, c->stream, c->dg, np, x, y, z,
                               (const cpx<float>*)b.X[1], (const cpx<float>*)b.X[2], (const cpx<float>*)b.XF,
                               (const cpx<float>*)b.XG, (const double*)pw.lam, pw.pmodes, pw.pmap, pw.npad,
                               pw.tiles_per_slice, pw.part, zb);
    } else {
        dim3 grid((np + 255) / 256, pw.nslice);
        hipLaunchKernelGGL((k_vel_points<T, LM>), grid, dim3(256), 0, c->stream, c->dg, np, x, y, z,
                           (const cpx<T>*)b.X[1], (const cpx<T>*)b.X[2], (const cpx<T>*)b.XF, (const cpx<T>*)b.XG,
                           (const double*)pw.lam, (const float4*)pw.modes, pw.tiles_per_slice, pw.part);
    }
    hipLaunchKernelGGL(k_sum_slices, dim3((np + 255) / 256), dim3(256), 0, c->stream, np, pw.nslice,
                       (const float*)pw.part, u, v, w);
}

template <typename T>
void launch_points(rbc_ctx* c, int np, const float* x, const float* y, const float* z, const PointWork& pw, float* u,
                   float* v, float* w) {
    switch (lmax_for(c->dg.L)) {
        case 8: launch_points_t<T, 8>(c, np, x, y, z, pw, u, v, w); break;
        case 16: launch_points_t<T, 16>(c, np, x, y, z, pw, u, v, w); break;
        case 24: launch_points_t<T, 24>(c, np, x, y, z, pw, u, v, w); break;
        default: launch_points_t<T, 32>(c, np, x, y, z, pw, u, v, w); break;
    }
}

void velocity_points(rbc_ctx* c, int np, const float* x, const float* y, const float* z, const PointWork& pw, float* u,
                     float* v, float* w) {
    if (np <= 0) return;
    if (c->rsz == 8) launch_points<double>(c, np, x, y, z, pw, u, v, w);
    else launch_points<float>(c, np, x, y, z, pw, u, v, w);
    HIPCHK(hipGetLastError());
}

}  // namespace

namespace rbc {

void velocity_setup(rbc_ctx* c) {
    VelWork* vw = new VelWork;
    c->vel_ws = vw;
    c->vel_ws_free = free_vel_work;
    auto al = [&](size_t bytes) {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, std::max<size_t>(16, bytes)));
        vw->owned.push_back(p);
        return p;
    };
    const int L = c->dg.L, P = c->dg.P;
    vw->lam = (double*)al((size_t)L * sizeof(double));
    vw->modes = (float4*)al((size_t)P * sizeof(float4));
    std::vector<double> lam(L);
    for (int m = 0; m < L; ++m) lam[m] = lambda_k(m + 1);
    const std::vector<float> mt = mode_table(c);
    HIPCHK(hipMemcpy(vw->lam, lam.data(), lam.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(vw->modes, mt.data(), mt.size() * sizeof(float), hipMemcpyHostToDevice));
    {
        std::vector<float> pt;
        std::vector<int> pm;
        padded_mode_table(c, pt, pm);
        vw->npad = (int)pm.size();
        vw->pmodes = (float4*)al(pm.size() * sizeof(float4));
        vw->pmap = (int*)al(pm.size() * sizeof(int));
        HIPCHK(hipMemcpy(vw->pmodes, pt.data(), pt.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(vw->pmap, pm.data(), pm.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    vw->ez_cap = std::max(1, c->dg.nq);
    vw->ezb = (double*)al((size_t)vw->ez_cap * 3 * L * sizeof(double));
    vw->ew = (double*)al((size_t)vw->ez_cap * sizeof(double));
    vw->ee = (double*)al((size_t)3 * P * sizeof(double));
    vw->cpart = (double*)al((size_t)2 * kCharvelBlocks * sizeof(double));
    vw->noalloc = (float*)noalloc_scratch(c->p.device);
}

}  // namespace rbc

// ---- tracer object -----------------------------------------------------------------------
struct rbc_tracer {
    rbc_ctx* c = nullptr;
    int np = 0, buf = 0, mode = 0;
    double diff = 0.0;
    uint64_t seed = 0;
    float *x = nullptr, *y = nullptr, *z = nullptr;  // [buf][np]
    int iter = 0, ptime = 0;
    GridWork gw;               // reference mode
    PointWork pw;              // spectral mode
    float *vel = nullptr;      // 3 * np (spectral mode)
    float *mid = nullptr;      // 3 * np
    double* red = nullptr;     // 4 doubles
    std::vector<void*> owned;
    std::string err;
    ~rbc_tracer() {
        if (c && c->stream) hipStreamSynchronize(c->stream);
        for (void* p : owned) hipFree(p);
    }
};

namespace {

int tfail(rbc_tracer* t, int code, const std::string& msg) {
    t->err = msg;
    t->c->err = msg;
    return code;
}

int check_ctx_for_velocity(rbc_ctx* c) {
    if (c->dg.M > 4096 || c->dg.N > 4096) return fail(c, RBC_ERR_ARG, "velocity grid supports M, N <= 4096");
    return RBC_OK;
}

}  // namespace

extern "C" {

int rbc_tracer_create(rbc_ctx* c, int num_particles, int buffer_size, int mode, double diff_coeff,
                      unsigned long long seed, rbc_tracer** out) {
    if (!c || !out) return RBC_ERR_ARG;
    *out = nullptr;
    if (num_particles < 1 || buffer_size < 2) return fail(c, RBC_ERR_ARG, "need num_particles >= 1 and buffer_size >= 2");
    if (mode != RBC_TRACER_RK4_TRILINEAR && mode != RBC_TRACER_SPECTRAL_RK2)
        return fail(c, RBC_ERR_ARG, "unknown tracer mode");
    if (mode == RBC_TRACER_SPECTRAL_RK2 && lmax_for(c->dg.L) == 0)
        return fail(c, RBC_ERR_ARG, "the spectral tracer supports L <= 32");
    if (diff_coeff < 0) return fail(c, RBC_ERR_ARG, "diffusion coefficient must be >= 0");
    if (int rc = check_ctx_for_velocity(c)) return rc;
    rbc_tracer* t = new rbc_tracer;
    t->c = c;
    t->np = num_particles;
    t->buf = buffer_size;
    t->mode = mode;
    t->diff = diff_coeff;
    t->seed = seed;
    try {
        hipSetDevice(c->p.device);
        auto al = [&](size_t bytes) {
            void* p = nullptr;
            HIPCHK(hipMalloc(&p, bytes));
            t->owned.push_back(p);
            return p;
        };
        const size_t pos = (size_t)num_particles * buffer_size * sizeof(float);
        t->x = (float*)al(pos);
        t->y = (float*)al(pos);
        t->z = (float*)al(pos);
        t->red = (double*)al(4 * sizeof(double));
        if (mode == RBC_TRACER_RK4_TRILINEAR) {
            alloc_grid_work(c, kZNum, t->gw, t->owned);
            // linspace(-PARTICLE_Z_MAX, PARTICLE_Z_MAX) in the device real type (particle_tracer.cu:494-501)
            const float zmin = (float)(-kZMax), zmax = (float)kZMax;
            std::vector<double> zl(kZNum);
            for (int l = 0; l < kZNum; ++l) zl[l] = (double)(zmin + l * (zmax - zmin) / (kZNum - 1));
            set_heights(c, t->gw, zl);
        } else {
            alloc_point_work(c, num_particles, t->pw, t->owned);
            t->vel = (float*)al((size_t)3 * num_particles * sizeof(float));
            t->mid = (float*)al((size_t)3 * num_particles * sizeof(float));
        }
        // the constructor places every particle at the centre (particle_tracer.cu:43)
        hipLaunchKernelGGL(k_tracer_init, dim3((num_particles + 255) / 256), dim3(256), 0, c->stream,
                           (int)RBC_PART_CENTER, num_particles, (float)c->p.Lx, (float)c->p.Ly,
                           (float)(0.5 * c->p.Lx), (float)(0.5 * c->p.Ly), 0.f, 0.f, (uint64_t)seed, t->x, t->y, t->z);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        const int rc = hip_fail(c, e);
        delete t;
        return rc;
    }
    *out = t;
    return RBC_OK;
}

void rbc_tracer_destroy(rbc_tracer* t) { delete t; }

const char* rbc_tracer_last_error(const rbc_tracer* t) { return t ? t->err.c_str() : ""; }

int rbc_tracer_init(rbc_tracer* t, int type, double cx, double cy, double cz, double sigma) {
    if (!t) return RBC_ERR_ARG;
    if (type < RBC_PART_RANDOM || type > RBC_PART_COORD_SIGMA) return tfail(t, RBC_ERR_ARG, "unknown init type");
    if (type == RBC_PART_PAIRS && (t->np & 1))
        return tfail(t, RBC_ERR_ARG, "Position pairs only possible with even number of particles.");
    rbc_ctx* c = t->c;
    float ax = (float)cx, ay = (float)cy, az = (float)cz, sg = (float)sigma;
    if (type == RBC_PART_CENTER || type == RBC_PART_CENTER_SIGMA) {
        ax = (float)(0.5 * c->p.Lx);
        ay = (float)(0.5 * c->p.Ly);
        az = 0.f;
        if (type == RBC_PART_CENTER) sg = 0.f;
    }
    try {
        hipSetDevice(c->p.device);
        const size_t off = (size_t)t->iter * t->np;
        hipLaunchKernelGGL(k_tracer_init, dim3((t->np + 255) / 256), dim3(256), 0, c->stream, type, t->np,
                           (float)c->p.Lx, (float)c->p.Ly, ax, ay, az, sg, t->seed, t->x + off, t->y + off, t->z + off);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_tracer_set_positions(rbc_tracer* t, const float* x, const float* y, const float* z) {
    if (!t || !x || !y || !z) return RBC_ERR_ARG;
    rbc_ctx* c = t->c;
    try {
        hipSetDevice(c->p.device);
        const size_t off = (size_t)t->iter * t->np, n = (size_t)t->np * sizeof(float);
        HIPCHK(hipMemcpyAsync(t->x + off, x, n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(t->y + off, y, n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(t->z + off, z, n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_tracer_step(rbc_tracer* t, double dt) {
    if (!t) return RBC_ERR_ARG;
    if (t->iter >= t->buf - 1)  // particle_tracer.cu:145-147, is_buffer_full :263-265
        return tfail(t, RBC_ERR_STATE, "buffer of particle tracer is filled! read out the positions first");
    rbc_ctx* c = t->c;
    const int np = t->np;
    const size_t in = (size_t)t->iter * np, outo = in + np;
    const float sigma = t->diff > 0 ? (float)std::sqrt(2.0 * t->diff * dt) : 0.f;
    const uint64_t ctr0 = (uint64_t)(t->ptime + 1) * 3 * (uint64_t)np;
    const dim3 pg((np + 255) / 256);
    try {
        hipSetDevice(c->p.device);
        if (t->mode == RBC_TRACER_RK4_TRILINEAR) {
            velocity_grid(c, t->gw);
            hipLaunchKernelGGL(k_tracer_rk4, pg, dim3(256), 0, c->stream, np, (const float*)t->gw.grid, c->dg.M,
                               c->dg.N, (float)c->p.Lx, (float)c->p.Ly, (const float*)t->x + in,
                               (const float*)t->y + in, (const float*)t->z + in, (float)dt, t->x + outo, t->y + outo,
                               t->z + outo, sigma, t->seed, ctr0);
        } else {
            float *u = t->vel, *v = t->vel + np, *w = t->vel + 2 * (size_t)np;
            float *xm = t->mid, *ym = t->mid + np, *zm = t->mid + 2 * (size_t)np;
            velocity_points(c, np, t->x + in, t->y + in, t->z + in, t->pw, u, v, w);
            hipLaunchKernelGGL(k_tracer_half, pg, dim3(256), 0, c->stream, np, (const float*)t->x + in,
                               (const float*)t->y + in, (const float*)t->z + in, (const float*)u, (const float*)v,
                               (const float*)w, (float)(0.5 * dt), xm, ym, zm);
            velocity_points(c, np, xm, ym, zm, t->pw, u, v, w);
            hipLaunchKernelGGL(k_tracer_full, pg, dim3(256), 0, c->stream, np, (const float*)t->x + in,
                               (const float*)t->y + in, (const float*)t->z + in, (const float*)u, (const float*)v,
                               (const float*)w, (float)dt, t->x + outo, t->y + outo, t->z + outo, sigma, t->seed,
                               ctr0);
        }
        HIPCHK(hipGetLastError());
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    t->iter++;
    t->ptime++;
    return RBC_OK;
}

int rbc_tracer_get_positions(rbc_tracer* t, int clear, float* x, float* y, float* z, int* num_timesteps) {
    if (!t || !x || !y || !z) return RBC_ERR_ARG;
    rbc_ctx* c = t->c;
    // particle_tracer.cu:272-311: the current slot is left out unless nothing was stepped
    int nt = t->iter;
    if (t->iter >= t->buf) nt = t->buf - 1;
    else if (t->iter < 1) nt = 1;
    try {
        hipSetDevice(c->p.device);
        const size_t n = (size_t)t->np * nt * sizeof(float);
        HIPCHK(hipMemcpyAsync(x, t->x, n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(y, t->y, n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(z, t->z, n, hipMemcpyDeviceToHost, c->stream));
        if (clear && t->iter > 0) {  // reset_particle_positions: slot nt becomes slot 0
            const size_t off = (size_t)nt * t->np, m = (size_t)t->np * sizeof(float);
            HIPCHK(hipMemcpyAsync(t->x, t->x + off, m, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(t->y, t->y + off, m, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(t->z, t->z + off, m, hipMemcpyDeviceToDevice, c->stream));
            t->iter = 0;
        }
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    if (num_timesteps) *num_timesteps = nt;
    return RBC_OK;
}

int rbc_tracer_current(rbc_tracer* t, float* x, float* y, float* z) {
    if (!t || !x || !y || !z) return RBC_ERR_ARG;
    rbc_ctx* c = t->c;
    try {
        hipSetDevice(c->p.device);
        const size_t off = (size_t)t->iter * t->np, n = (size_t)t->np * sizeof(float);
        HIPCHK(hipMemcpyAsync(x, t->x + off, n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(y, t->y + off, n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(z, t->z + off, n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_tracer_info(const rbc_tracer* t, int* iteration_index, int* particle_time) {
    if (!t) return RBC_ERR_ARG;
    if (iteration_index) *iteration_index = t->iter;
    if (particle_time) *particle_time = t->ptime;
    return RBC_OK;
}

int rbc_tracer_moments(rbc_tracer* t, double* out4) {
    if (!t || !out4) return RBC_ERR_ARG;
    rbc_ctx* c = t->c;
    try {
        hipSetDevice(c->p.device);
        const size_t off = (size_t)t->iter * t->np;
        hipLaunchKernelGGL(k_tracer_moments, dim3(1), dim3(1024), 0, c->stream, t->np, (const float*)t->x + off,
                           (const float*)t->y + off, (float)c->p.Lx, (float)c->p.Ly, t->red);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out4, t->red, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_velocity_grid(rbc_ctx* c, int nz, const double* z, float* ux, float* uy, float* uz) {
    if (!c || nz < 1 || !z || !ux || !uy || !uz) return RBC_ERR_ARG;
    if (int rc = check_ctx_for_velocity(c)) return rc;
    try {
        hipSetDevice(c->p.device);
        VelWork* vw = vel_work(c);
        if (nz > vw->gw_cap) {  // grow the grid work (kept for later calls)
            HIPCHK(hipStreamSynchronize(c->stream));
            for (void* p : vw->gw_owned) HIPCHK(hipFree(p));
            vw->gw_owned.clear();
            vw->gw = GridWork();
            vw->gw_cap = 0;
            alloc_grid_work(c, nz, vw->gw, vw->gw_owned);
            vw->gw_cap = nz;
        }
        GridWork& w = vw->gw;
        w.nz = nz;
        set_heights(c, w, std::vector<double>(z, z + nz));
        velocity_grid(c, w);
        const size_t n = (size_t)nz * c->dg.N * c->dg.M;
        HIPCHK(hipMemcpyAsync(ux, w.grid, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(uy, w.grid + n, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(uz, w.grid + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_characteristic_velocity(rbc_ctx* c, int nz, double* U) {
    if (!c || nz < 2 || !U) return RBC_ERR_ARG;
    if (int rc = check_ctx_for_velocity(c)) return rc;
    try {
        hipSetDevice(c->p.device);
        VelWork* vw = vel_work(c);
        if (nz > vw->gw_cap) {
            HIPCHK(hipStreamSynchronize(c->stream));
            for (void* p : vw->gw_owned) HIPCHK(hipFree(p));
            vw->gw_owned.clear();
            vw->gw = GridWork();
            vw->gw_cap = 0;
            alloc_grid_work(c, nz, vw->gw, vw->gw_owned);
            vw->gw_cap = nz;
        }
        GridWork& w = vw->gw;
        w.nz = nz;
        // linspace(-0.5, 0.5, nz) in the device real type (calculate_velocity_operator.cu:56-58, 320-327)
        std::vector<double> z(nz);
        for (int l = 0; l < nz; ++l) z[l] = (double)(-0.5f + l * (0.5f - -0.5f) / (nz - 1));
        HIPCHK(hipStreamSynchronize(c->stream));  // the heights table may still be read
        set_heights(c, w, z);
        velocity_grid(c, w);
        const size_t n = (size_t)nz * c->dg.N * c->dg.M;
        hipLaunchKernelGGL(k_charvel, dim3(kCharvelBlocks), dim3(256), 0, c->stream, n, (const float*)w.grid,
                           vw->cpart);
        HIPCHK(hipGetLastError());
        std::vector<double> part(2 * (size_t)kCharvelBlocks);
        HIPCHK(hipMemcpyAsync(part.data(), vw->cpart, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double s = 0.0, cnt = 0.0;
        for (int b = 0; b < kCharvelBlocks; ++b) {
            s += part[2 * b];
            cnt += part[2 * b + 1];
        }
        *U = s / cnt;
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_velocity_at(rbc_ctx* c, int n, const float* x, const float* y, const float* z, float* u, float* v, float* w) {
    if (!c || n < 0 || (n > 0 && (!x || !y || !z || !u || !v || !w))) return RBC_ERR_ARG;
    if (lmax_for(c->dg.L) == 0) return fail(c, RBC_ERR_ARG, "spectral evaluation supports L <= 32");
    if (n == 0) return RBC_OK;
    try {
        hipSetDevice(c->p.device);
        VelWork* vw = vel_work(c);
        HIPCHK(hipStreamSynchronize(c->stream));  // earlier calls may still use the buffers
        grow(vw->io, vw->io_cap, (size_t)6 * n);
        grow(vw->part, vw->part_cap, part_floats(c, n));
        if (c->rsz == 4) grow(vw->zbas, vw->zbas_cap, (size_t)n * c->dg.L);
        PointWork pw;
        plan_points(c, n, pw);
        point_constants(c, pw);
        pw.part = vw->part;
        pw.zb = vw->zbas;
        float* d = vw->io;
        const size_t b = (size_t)n * sizeof(float);
        HIPCHK(hipMemcpyAsync(d, x, b, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(d + n, y, b, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(d + 2 * (size_t)n, z, b, hipMemcpyHostToDevice, c->stream));
        float* o = d + 3 * (size_t)n;
        velocity_points(c, n, d, d + n, d + 2 * (size_t)n, pw, o, o + n, o + 2 * (size_t)n);
        HIPCHK(hipMemcpyAsync(u, o, b, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(v, o + n, b, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(w, o + 2 * (size_t)n, b, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_velocity_at_device_noalloc(rbc_ctx* c, int n, const float* x, const float* y, const float* z, float* u,
                                   float* v, float* w) {
    if (!c || n < 0 || (n > 0 && (!x || !y || !z || !u || !v || !w))) return RBC_ERR_ARG;
    if (lmax_for(c->dg.L) == 0) return fail(c, RBC_ERR_ARG, "spectral evaluation supports L <= 32");
    if (n == 0) return RBC_OK;
    VelWork* vw = vel_work(c);
    if (!vw->noalloc) return fail(c, RBC_ERR_HIP, "no-allocation scratch not available on this device");
    PointWork pw;
    plan_points(c, n, pw);
    const size_t need = (size_t)pw.nslice * 3 * n;
    if (need > kNoAllocScratchFloats)
        return fail(c, RBC_ERR_ARG,
                    std::to_string(n) + " points need " + std::to_string(need) +
                        " floats of partial sums; the no-allocation scratch holds " +
                        std::to_string(kNoAllocScratchFloats) + " (every n <= 1398101 fits)");
    point_constants(c, pw);
    pw.part = vw->noalloc;  // every slot the sum reads is written first in the same call
    try {
        hipSetDevice(c->p.device);
        velocity_points(c, n, x, y, z, pw, u, v, w);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_velocity_at_device(rbc_ctx* c, int n, const float* x, const float* y, const float* z, float* u, float* v,
                           float* w) {
    if (!c || n < 0 || (n > 0 && (!x || !y || !z || !u || !v || !w))) return RBC_ERR_ARG;
    if (lmax_for(c->dg.L) == 0) return fail(c, RBC_ERR_ARG, "spectral evaluation supports L <= 32");
    if (n == 0) return RBC_OK;
    try {
        hipSetDevice(c->p.device);
        VelWork* vw = vel_work(c);
        const size_t need = part_floats(c, n);
        const size_t nzb = c->rsz == 4 ? (size_t)n * c->dg.L : 0;
        if (need > vw->part_cap || nzb > vw->zbas_cap) {
            HIPCHK(hipStreamSynchronize(c->stream));  // the old buffers may be in use
            grow(vw->part, vw->part_cap, need);
            if (nzb) grow(vw->zbas, vw->zbas_cap, nzb);
        }
        PointWork pw;
        plan_points(c, n, pw);
        point_constants(c, pw);
        pw.part = vw->part;
        pw.zb = vw->zbas;
        velocity_points(c, n, x, y, z, pw, u, v, w);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

// per-mode energies at heights z with weights wz (FP64, host result [3][P]); nz <= the
// context's level count (the spectrum scratch is made at create)
static int mode_energies(rbc_ctx* c, const std::vector<double>& z, const std::vector<double>& wz,
                         std::vector<double>& e) {
    try {
        hipSetDevice(c->p.device);
        const DevGeom& g = c->dg;
        VelWork* vw = vel_work(c);
        const int nz = (int)z.size();
        if (nz > vw->ez_cap) return fail(c, RBC_ERR_ARG, "too many spectrum levels");
        const std::vector<double> tb = basis_table(g.L, z);
        HIPCHK(hipStreamSynchronize(c->stream));  // the scratch may still be read by an earlier call
        HIPCHK(hipMemcpy(vw->ezb, tb.data(), tb.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(vw->ew, wz.data(), wz.size() * sizeof(double), hipMemcpyHostToDevice));
        const Buffers& b = c->b;
        if (c->rsz == 8)
            hipLaunchKernelGGL(k_energy_modes<double>, dim3((g.P + 255) / 256), dim3(256), 0, c->stream, g, nz,
                               (const double*)vw->ezb, (const double*)vw->ew, (const cpx<double>*)b.X[1],
                               (const cpx<double>*)b.X[2], (const cpx<double>*)b.XF, (const cpx<double>*)b.XG,
                               (const double*)b.kx, (const double*)b.ky, vw->ee);
        else
            hipLaunchKernelGGL(k_energy_modes<float>, dim3((g.P + 255) / 256), dim3(256), 0, c->stream, g, nz,
                               (const double*)vw->ezb, (const double*)vw->ew, (const cpx<float>*)b.X[1],
                               (const cpx<float>*)b.X[2], (const cpx<float>*)b.XF, (const cpx<float>*)b.XG,
                               (const float*)b.kx, (const float*)b.ky, vw->ee);
        HIPCHK(hipGetLastError());
        e.resize((size_t)3 * g.P);
        HIPCHK(hipMemcpyAsync(e.data(), vw->ee, e.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& err) {
        return hip_fail(c, err);
    }
    return RBC_OK;
}

int rbc_energy_spectrum(rbc_ctx* c, int cap, double* E, int* nbins, double* totals) {
    if (!c || !nbins) return RBC_ERR_ARG;
    const Geometry& g = c->geo;
    const double dk = std::min(2 * M_PI / c->p.Lx, 2 * M_PI / c->p.Ly);
    std::vector<double> zq, wq;
    gauss_legendre(c->dg.nq, zq, wq);
    std::vector<double> e;
    if (int rc = mode_energies(c, zq, wq, e)) return rc;
    const int P = g.P(), NYc = g.NYc();
    std::vector<double> spec;
    double eh = 0, ez = 0;
    for (int p = 0; p < P; ++p) {
        const double wk = (p / NYc == 0) ? 1.0 : 2.0;
        const double kk = std::sqrt(c->fac.kx[p] * c->fac.kx[p] + c->fac.ky[p] * c->fac.ky[p]);
        const int s = (int)std::floor(kk / dk + 0.5);
        if ((int)spec.size() <= s) spec.resize(s + 1, 0.0);
        const double h = 0.5 * wk * (e[p] + e[P + p]);
        spec[s] += h;
        eh += h;
        ez += 0.5 * wk * e[2 * (size_t)P + p];
    }
    *nbins = (int)spec.size();
    if (E) {
        if (cap < (int)spec.size()) return fail(c, RBC_ERR_ARG, "spectrum buffer too small");
        std::copy(spec.begin(), spec.end(), E);
    }
    if (totals) {
        totals[0] = eh;
        totals[1] = ez;
    }
    return RBC_OK;
}

int rbc_energy_spectrum_ref(rbc_ctx* c, double z, int cap, double* E3, int* nbins) {
    if (!c || !nbins) return RBC_ERR_ARG;
    const Geometry& g = c->geo;
    std::vector<double> e;
    if (int rc = mode_energies(c, std::vector<double>{z}, std::vector<double>{1.0}, e)) return rc;
    // energy_spectrum_2d.cu:53-56, 85-129: size = max(M/2+1, N/2+1) bins of width
    // |k_max| / size, |k_max| from (2 pi (M/2) / Lx, 2 pi (N/2) / Ly); each bin sums |c_k|^2
    // over the stored half spectrum (no factor 2 for kx > 0).
    const int mox = g.mox, size = std::max(mox, g.N / 2 + 1);
    const double kmx = 2.0 * M_PI * (mox - 1) / c->p.Lx, kmy = 2.0 * M_PI * (g.N / 2) / c->p.Ly;
    const double inc = std::sqrt(kmx * kmx + kmy * kmy) / size;
    *nbins = size;
    if (!E3) return RBC_OK;
    if (cap < size) return fail(c, RBC_ERR_ARG, "spectrum buffer too small");
    for (int i = 0; i < 3 * size; ++i) E3[i] = 0.0;
    const int P = g.P(), NYc = g.NYc();
    for (int p = 0; p < P; ++p) {
        const int col = p / NYc, row = g.rows[p % NYc];
        const double mx = 2.0 * M_PI * col / c->p.Lx;
        const double my = 2.0 * M_PI * (row <= g.N / 2 ? row : g.N - row) / c->p.Ly;
        const double a = std::sqrt(mx * mx + my * my);
        const int b0 = std::max(0, (int)std::floor(a / inc) - 1);
        for (int b = b0; b < std::min(size, b0 + 3); ++b)
            if (b * inc <= a && (b + 1) * inc > a) {
                for (int k = 0; k < 3; ++k) E3[k * size + b] += e[(size_t)k * P + p];
                break;
            }
    }
    return RBC_OK;
}

}  // extern "C"

// ---- packets of the visual stream (output/server_writer.cu:79-252, output/datastruct.h) -----
// The frames the reference's server_writer sends over UDP to its viewer, built in caller
// memory; sending them is left to the caller.  Layout (packed as the reference's structs):
// every packet is 2048 bytes, a 16-byte header {uint16 Type, uint16 NumPackets, uint32 Nt,
// uint32 Nx, uint32 Ny} and a payload; packet 0 (Type 0x491C) carries the simulation header,
// packets 1.. (Type 0x08B1) {int32 iStart, int32 iEnd, float Data[506]}: the image sampled at
// Nx x Ny points, x outer, y inner.
namespace {
constexpr int kUPack = 2048, kUData = kUPack / 4 - 6, kUDescr = 32;
constexpr uint16_t kUTypeH = 0x491C, kUTypeP = 0x08B1;
struct UHeader {
    uint16_t Type, NumPackets;
    uint32_t Nt, Nx, Ny;
};
struct USimHeader {
    float uVersion, cVersion, tFinal, dt, Ra, Pr, eta, clx, cly, clz;
    uint16_t nmx, nmy, nmz;
    char device[kUDescr], descr[kUDescr], Err1[kUDescr], Err2[kUDescr], Err3[kUDescr], Err4[kUDescr];
};
struct USimPayload {
    int32_t iStart, iEnd;
    float Data[kUData];
};
struct UPacket {
    UHeader H;
    USimPayload P;
};
static_assert(sizeof(UPacket) == kUPack, "visual-stream packet size");
static_assert(sizeof(USimHeader) <= sizeof(USimPayload), "simulation header fits the payload");
}  // namespace

extern "C" int rbc_stream_packets(rbc_ctx* c, rbc_tracer* t, int img_type, int res_req, int Nt, double eta,
                                  double t_final, void* out, size_t cap, size_t* nbytes) {
    if (!c || !nbytes || img_type < RBC_IMG_TEMPERATURE || img_type > RBC_IMG_TEMP_PARTICLES) return RBC_ERR_ARG;
    const bool want_t = img_type != RBC_IMG_PARTICLES, want_p = img_type != RBC_IMG_TEMPERATURE;
    if (want_p && !t) return fail(c, RBC_ERR_ARG, "particle images need a tracer");
    const int M = c->dg.M, N = c->dg.N;
    // the image (x fastest): temperature at z = 0, or particle counts on a res x res grid
    int xs = M, ys = N;
    std::vector<float> img;
    std::string err;
    if (want_t) {
        img.resize((size_t)M * N);
        const double z0 = 0.0;
        if (int rc = rbc_temperature_grid(c, 1, &z0, img.data())) return rc;
    } else {
        xs = ys = res_req <= 0 ? 128 : res_req;
        img.assign((size_t)xs * ys, 0.f);
    }
    if (want_p) {
        std::vector<float> px((size_t)(t->buf - 1) * t->np), py(px.size()), pz(px.size());
        int nt = 0;
        if (int rc = rbc_tracer_get_positions(t, 0, px.data(), py.data(), pz.data(), &nt)) return rc;
        const size_t off = (size_t)(nt - 1) * t->np;  // the last slot read out (noClear)
        const float clx = (float)c->p.Lx, cly = (float)c->p.Ly;
        for (int i = t->np - 1; i >= 0; --i) {  // reversed so that particle 0 is drawn last
            const float x = px[off + i], y = py[off + i];
            const float xg = (xs * x) / clx, yg = (ys * y) / cly;
            const int ix = (int)(xg - std::floor(xg / xs) * xs), iy = (int)(yg - std::floor(yg / ys) * ys);
            if (ix < 0 || ix >= xs || iy < 0 || iy >= ys) {
                char b[96];
                snprintf(b, sizeof b, "Particle %d at (%f,%f) out of bounds.", i, x, y);
                err = b;
                continue;
            }
            float& e = img[(size_t)iy * xs + ix];
            e = img_type == RBC_IMG_TEMP_PARTICLES ? 0.f : e + 1.f;
        }
    }
    int Nx = res_req, Ny = res_req;
    if (Nx <= 0 || Nx > xs) Nx = xs;
    if (Ny <= 0 || Ny > ys) Ny = ys;
    const int npack = 1 + (Nx * Ny + kUData - 1) / kUData;
    *nbytes = (size_t)npack * kUPack;
    if (!out) return RBC_OK;
    if (cap < *nbytes) return fail(c, RBC_ERR_ARG, "packet buffer too small");
    UPacket* pk = (UPacket*)out;
    std::memset(out, 0, *nbytes);
    for (int i = 0; i < npack; ++i) {
        pk[i].H.Type = i == 0 ? kUTypeH : kUTypeP;
        pk[i].H.NumPackets = (uint16_t)npack;
        pk[i].H.Nt = (uint32_t)Nt;
        pk[i].H.Nx = (uint32_t)Nx;
        pk[i].H.Ny = (uint32_t)Ny;
    }
    USimHeader h;
    std::memset(&h, 0, sizeof h);
    h.uVersion = 1.1f;
    h.cVersion = 1.0f;
    h.tFinal = (float)t_final;
    h.dt = (float)c->p.dt;
    h.Ra = (float)c->p.Ra;
    h.Pr = (float)c->p.Pr;
    h.eta = (float)eta;
    h.clx = (float)c->p.Lx;
    h.cly = (float)c->p.Ly;
    h.clz = 1.f;
    h.nmx = (uint16_t)M;
    h.nmy = (uint16_t)N;
    h.nmz = (uint16_t)c->dg.L;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->p.device) == hipSuccess) strncpy(h.device, prop.name, kUDescr);
    const char* descr = img_type == RBC_IMG_TEMPERATURE ? "Temperature field at z = 0"
                        : img_type == RBC_IMG_TEMP_PARTICLES ? "Temperature at z = 0 w/ Part" : "Particles";
    strncpy(h.descr, descr, kUDescr);
    strncpy(h.Err1, err.c_str(), kUDescr);
    std::memcpy(&pk[0].P, &h, sizeof h);
    int tot = 0, ip = 1, in = 0;
    USimPayload* P = &pk[ip].P;
    P->iStart = 0;
    for (int ix = 0; ix < Nx; ++ix)
        for (int iy = 0; iy < Ny; ++iy, ++in, ++tot) {
            if (in >= kUData) {
                P->iEnd = tot;
                P = &pk[++ip].P;
                P->iStart = tot;
                in = 0;
            }
            const int sx = ix * xs / Nx, sy = iy * ys / Ny;
            P->Data[in] = img[(size_t)sy * xs + sx];
        }
    P->iEnd = tot;
    return RBC_OK;
}
