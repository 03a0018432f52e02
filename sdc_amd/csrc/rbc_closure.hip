// rbc_closure.hip -- learned closure for coarse runs: a pointwise MLP (8 -> 64 -> 64 -> 3,
// tanh, BF16 weights, FP32 arithmetic) trained on the GPU to predict, at the points of a
// coarse grid, the part of the nonlinear term that the coarse grid misses.
//
// Training data (rbc_closure_sample): from a fine state X_f,
//   R = P_c N_f(X_f) - N_c(P_c X_f)
// where P_c truncates to the coarse kept modes and vertical modes (the Galerkin bases are
// nested), N_f / N_c are the fine / coarse nonlinear terms (scaled as the stepper uses them).
// R is synthesised at the coarse points (x, y, z_l) as r_theta = sum S_m R_theta,m,
// r_f = sum C_m R_f,m, r_g = sum S_m R_g,m; the inputs are the local coarse fields
// (u, v, w, theta, d_x theta, d_y theta, d_z theta, z).
// Correction (rbc_closure_enable): inside every coarse step, the predicted r at the points
// is projected back onto the coarse modes (R2C, forward y-FFT, Gauss-Legendre projection
// on S_m / C_m) and added to N before the inverse-B / AB2 / implicit update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "rbc_ctx.hpp"
#include "rbc_device.hpp"

using namespace rbc;

namespace {

constexpr int NIN = 8, NH = 64, NOUT = 3, NFEAT = 7;
constexpr int OFF_W1 = 0;                   // [NH][NIN]
constexpr int OFF_B1 = OFF_W1 + NH * NIN;   // [NH]
constexpr int OFF_W2 = OFF_B1 + NH;         // [NH][NH]  (row j = hidden unit j of layer 2)
constexpr int OFF_B2 = OFF_W2 + NH * NH;    // [NH]
constexpr int OFF_W3 = OFF_B2 + NH;         // [NOUT][NH]
constexpr int OFF_B3 = OFF_W3 + NOUT * NH;  // [NOUT]
constexpr int NW = OFF_B3 + NOUT;           // 4931 parameters
constexpr int kTrainWaves = 512;

struct Norm {
    float in_mean[NIN], in_inv[NIN];  // x_n = (x - mean) * inv
    float out_scale[NOUT];            // r = y_n * scale
};

__device__ __forceinline__ float bf16_bits_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__host__ __device__ inline uint16_t f32_to_bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));  // inf / NaN
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline uint64_t mix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// ---- point synthesis ---------------------------------------------------------------------
// features at the levels: spec[f][l][p], f = u, v, w, theta, theta_x, theta_y, theta_z;
// zb[l][4][L] = C', C, S, S' (FP64).  Block = MB modes (MB = 64, 32 or 16, so that LDS holds
// them): their [MB][L] slabs of theta, f, g are staged in LDS as [field][m][mode] (coalesced
// reads, each mode read once for all levels), the mean flow at every level is staged by the
// block holding p = 0, and lane (mode, sub-level) of wave w then walks levels
// (w 64 / MB + sub) + 256 / MB i.
template <typename T>
__global__ void __launch_bounds__(256) k_feat_levels(DevGeom g, int nz, int MB, const T* __restrict__ zb,
                                                     const cpx<T>* __restrict__ Xt, const cpx<T>* __restrict__ Xf,
                                                     const cpx<T>* __restrict__ Xg, const cpx<T>* __restrict__ XF,
                                                     const cpx<T>* __restrict__ XG, const T* __restrict__ kxa,
                                                     const T* __restrict__ kya, cpx<float>* __restrict__ spec) {
    extern __shared__ unsigned char smem_raw[];
    const int L = g.L, P = g.P;
    cpx<T>* sm = reinterpret_cast<cpx<T>*>(smem_raw);  // [3][L][MB]
    cpx<T>* smf = sm + 3 * L * MB;                      // [2][nz] mean flow (block 0)
    const int p0 = blockIdx.x * MB, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int e = threadIdx.x; e < MB * L; e += blockDim.x) {
        const int mode = e / L, m = e - mode * L, pp = p0 + mode;
        cpx<T> zt = mk<T>(0, 0), zf = zt, zg = zt;
        if (pp < P) {
            const size_t gi = (size_t)pp * L + m;
            zt = Xt[gi];
            zf = Xf[gi];
            zg = Xg[gi];
        }
        sm[(0 * L + m) * MB + mode] = zt;
        sm[(1 * L + m) * MB + mode] = zf;
        sm[(2 * L + m) * MB + mode] = zg;
    }
    if (p0 == 0)
        for (int t = threadIdx.x; t < 2 * nz; t += blockDim.x) {
            const int which = t / nz, l = t - which * nz;
            const T* tb = zb + (size_t)l * 4 * L + 2 * L;
            const cpx<T>* Xm = which ? XG : XF;
            cpx<T> sum = mk<T>(0, 0);
            for (int m = 0; m < L; ++m) sum = sum + cscale(Xm[m], tb[m]);
            smf[t] = sum;
        }
    __syncthreads();
    const int mode = lane % MB, sub = lane / MB, nsub = 64 / MB;
    const int p = p0 + mode;
    if (p >= P) return;
    const T kx = kxa[p], ky = kya[p], k2 = kx * kx + ky * ky;
    const size_t plane = (size_t)nz * P;
    for (int l = wave * nsub + sub; l < nz; l += 4 * nsub) {
        const T* tb = zb + (size_t)l * 4 * L;
        cpx<T> A = mk<T>(0, 0), B = A, Cc = A, Ts = A, Tz = A;
        for (int m = 0; m < L; ++m) {
            const cpx<T> t = sm[(0 * L + m) * MB + mode], f = sm[(1 * L + m) * MB + mode],
                         gg = sm[(2 * L + m) * MB + mode];
            A = A + cscale(f, tb[m]);
            Cc = Cc + cscale(f, tb[L + m]);
            B = B + cscale(gg, tb[2 * L + m]);
            Ts = Ts + cscale(t, tb[2 * L + m]);
            Tz = Tz + cscale(t, tb[3 * L + m]);
        }
        cpx<T> u = imul(A, kx) + imul(B, ky), v = imul(A, ky) - imul(B, kx);
        if (p == 0) {
            u = u + smf[l];
            v = v + smf[nz + l];
        }
        const cpx<T> vals[NFEAT] = {u, v, cscale(Cc, k2), Ts, imul(Ts, kx), imul(Ts, ky), Tz};
#pragma unroll
        for (int f = 0; f < NFEAT; ++f)
            spec[f * plane + (size_t)l * P + p] = mk<float>((float)vals[f].x, (float)vals[f].y);
    }
}

// residual fields at the levels: spec[f][l][p] for (Nt - No) of theta (S), f (C), g (S)
template <typename T>
__global__ void __launch_bounds__(256) k_res_levels(DevGeom g, int nz, const T* __restrict__ zb,
                                                    const cpx<T>* __restrict__ Nt0, const cpx<T>* __restrict__ Nt1,
                                                    const cpx<T>* __restrict__ Nt2, const cpx<T>* __restrict__ No0,
                                                    const cpx<T>* __restrict__ No1, const cpx<T>* __restrict__ No2,
                                                    cpx<float>* __restrict__ spec) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x, l = blockIdx.y;
    if (p >= g.P) return;
    const int L = g.L;
    const T* tb = zb + (size_t)l * 4 * L;
    cpx<T> a = mk<T>(0, 0), b = a, c = a;
    for (int m = 0; m < L; ++m) {
        const size_t i = (size_t)p * L + m;
        a = a + cscale(Nt0[i] - No0[i], tb[2 * L + m]);
        b = b + cscale(Nt1[i] - No1[i], tb[L + m]);
        c = c + cscale(Nt2[i] - No2[i], tb[2 * L + m]);
    }
    const size_t plane = (size_t)nz * g.P;
    spec[(size_t)l * g.P + p] = mk<float>((float)a.x, (float)a.y);
    spec[plane + (size_t)l * g.P + p] = mk<float>((float)b.x, (float)b.y);
    spec[2 * plane + (size_t)l * g.P + p] = mk<float>((float)c.x, (float)c.y);
}

// x C2R of F hybrid fields per (y, l) row into point-major values out[pt * os + f]
// (pt = (l N + y) M + x); with zcol >= 0 the level's height goes to column zcol.  Each field
// has its own complex transform: the fields differ by orders of magnitude (velocity vs theta,
// the f residual vs the theta residual), and packing two into one transform would leave the
// smaller one with the rounding error of the larger.
__global__ void __launch_bounds__(256) k_synth_rows(DevGeom g, int F, int os, int zcol, const float* __restrict__ zl,
                                                    const cpx<float>* __restrict__ hyb,
                                                    const cpx<float>* __restrict__ twM, float* __restrict__ out) {
    extern __shared__ unsigned char smem_raw[];
    const int M = g.M, N = g.N, KX = g.KX, KXp = g.KXp, nz = g.nq;
    cpx<float>* a = reinterpret_cast<cpx<float>*>(smem_raw);
    cpx<float>* b = a + (size_t)F * M;
    const int y = blockIdx.x, l = blockIdx.y;
    const size_t fstride = (size_t)nz * N * KXp, roff = ((size_t)l * N + y) * KXp;
    for (int e = threadIdx.x; e < F * M; e += blockDim.x) a[e] = mk<float>(0, 0);
    __syncthreads();
    for (int e = threadIdx.x; e < F * KX; e += blockDim.x) {
        const int f = e / KX, k = e - f * KX;
        cpx<float> x1 = hyb[(size_t)f * fstride + roff + k];
        if (k == 0) x1.y = 0;  // a C2R ignores the imaginary part at kx = 0
        cpx<float>* z = a + f * M;
        z[k] = x1;
        if (k > 0) z[M - k] = conjg(x1);
    }
    __syncthreads();
    const cpx<float>* r = lds_fft(a, b, g.logM, F, twM, true);
    const float zv = zcol >= 0 ? zl[l] : 0.f;
    for (int x = threadIdx.x; x < M; x += blockDim.x) {
        float* o = out + (((size_t)l * N + y) * M + x) * os;
        for (int f = 0; f < F; ++f) o[f] = r[f * M + x].x;
        if (zcol >= 0) o[zcol] = zv;
    }
}

// R2C of the three predicted fields per (y, l) row into hybrid rows (scaled 1/(M N)), one
// complex transform per field (see k_synth_rows)
__global__ void __launch_bounds__(256) k_anal_rows(DevGeom g, const float* __restrict__ in,
                                                   const cpx<float>* __restrict__ twM, cpx<float>* __restrict__ hyb) {
    extern __shared__ unsigned char smem_raw[];
    const int M = g.M, N = g.N, KX = g.KX, KXp = g.KXp, nz = g.nq;
    cpx<float>* a = reinterpret_cast<cpx<float>*>(smem_raw);  // NOUT M
    cpx<float>* b = a + NOUT * M;
    const int y = blockIdx.x, l = blockIdx.y;
    for (int x = threadIdx.x; x < M; x += blockDim.x) {
        const float* v = in + (((size_t)l * N + y) * M + x) * NOUT;
#pragma unroll
        for (int f = 0; f < NOUT; ++f) a[f * M + x] = mk(v[f], 0.f);
    }
    __syncthreads();
    const cpx<float>* R = lds_fft(a, b, g.logM, NOUT, twM, false);
    const float s = 1.0f / ((float)M * (float)N);
    const size_t fstride = (size_t)nz * N * KXp, roff = ((size_t)l * N + y) * KXp;
    for (int e = threadIdx.x; e < NOUT * KX; e += blockDim.x) {
        const int f = e / KX, k = e - f * KX;
        hyb[(size_t)f * fstride + roff + k] = cscale(R[f * M + k], s);
    }
}

// Group-transform versions (FP32, M = 64, 128, 256; see fft_group): one group of W = M / 8
// lanes per (y, l) row, WPB = 4 waves per block.  The real fields go two at a time through one
// complex transform with the second balanced to the first by a power of two (pair_scale over
// the row), so each keeps the rounding of its own size.
template <int LOGM>
__global__ void __launch_bounds__(256) k_synth_rows_w(DevGeom g, int F, int os, int zcol,
                                                      const float* __restrict__ zl, const cpx<float>* __restrict__ hyb,
                                                      const cpx<float>* __restrict__ twM, float* __restrict__ out) {
    constexpr int M = 1 << LOGM, W = M / 8, G = 64 / W, RPB = 4 * G, SP = M + M / 8, KX = kept_cols(M);
    __shared__ cpx<float> tw[M];
    __shared__ cpx<float> scr[RPB][SP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, grp = lane / W, j = lane - grp * W;
    const int N = g.N, KXp = g.KXp, nz = g.nq;
    const int y = blockIdx.x * RPB + wave * G + grp, l = blockIdx.y;
    const size_t fstride = (size_t)nz * N * KXp, roff = ((size_t)l * N + y) * KXp;
    for (int e = threadIdx.x; e < M; e += blockDim.x) tw[e] = twM[e];
    __syncthreads();
    cpx<float>* sb = scr[wave * G + grp];
    auto gmax = [](float x) { return group_max<W>(x); };
    float* orow = out + ((size_t)l * N + y) * M * os;
    for (int f0 = 0; f0 < F; f0 += 2) {
        const bool two = f0 + 1 < F;
        const cpx<float>* x1 = hyb + (size_t)f0 * fstride + roff;
        const cpx<float>* x2 = hyb + (size_t)(two ? f0 + 1 : f0) * fstride + roff;
        cpx<float> a[8], b[8];
        float ma = 0.f, mb = 0.f;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int k = j + W * r;
            const int src = k < KX ? k : (k > M - KX ? M - k : -1);
            a[r] = b[r] = mk(0.f, 0.f);
            if (src >= 0) {
                a[r] = x1[src];
                if (two) b[r] = x2[src];
                if (src == 0) {  // a C2R ignores the imaginary part at kx = 0
                    a[r].y = 0.f;
                    b[r].y = 0.f;
                }
            }
            ma = fmaxf(ma, fmaxf(fabsf(a[r].x), fabsf(a[r].y)));
            mb = fmaxf(mb, fmaxf(fabsf(b[r].x), fabsf(b[r].y)));
        }
        const float sg = pair_scale(gmax(ma), gmax(mb)), isg = inv_pow2(sg);
        cpx<float> v[1][8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int k = j + W * r;
            const cpx<float> bs = mk(sg * b[r].x, sg * b[r].y);
            if (k < KX) v[0][r] = mk(a[r].x - bs.y, a[r].y + bs.x);            // X1 + i s X2
            else if (k > M - KX) v[0][r] = mk(a[r].x + bs.y, -a[r].y + bs.x);  // conj(X1) + i s conj(X2)
            else v[0][r] = mk(0.f, 0.f);
        }
        fft_group<LOGM, true, 1>(v, sb, tw, j);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            float* o = orow + (size_t)(j + W * r) * os + f0;
            o[0] = v[0][r].x;
            if (two) o[1] = isg * v[0][r].y;
        }
        wave_sync();  // the scratch serves the next pair
    }
    if (zcol >= 0) {
        const float zv = zl[l];
#pragma unroll
        for (int r = 0; r < 8; ++r) orow[(size_t)(j + W * r) * os + zcol] = zv;
    }
}

// R2C of the three predicted fields per (y, l) row: (r_theta + i s r_f) and r_g as two
// complex transforms, unpacked and scaled by 1 / (M N)
template <int LOGM>
__global__ void __launch_bounds__(256) k_anal_rows_w(DevGeom g, const float* __restrict__ in,
                                                     const cpx<float>* __restrict__ twM, cpx<float>* __restrict__ hyb) {
    constexpr int M = 1 << LOGM, W = M / 8, G = 64 / W, RPB = 4 * G, SP = M + M / 8, KX = kept_cols(M);
    __shared__ cpx<float> tw[M];
    __shared__ cpx<float> scr[RPB][SP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, grp = lane / W, j = lane - grp * W;
    const int N = g.N, KXp = g.KXp, nz = g.nq;
    const int y = blockIdx.x * RPB + wave * G + grp, l = blockIdx.y;
    const size_t fstride = (size_t)nz * N * KXp, roff = ((size_t)l * N + y) * KXp;
    for (int e = threadIdx.x; e < M; e += blockDim.x) tw[e] = twM[e];
    __syncthreads();
    cpx<float>* sb = scr[wave * G + grp];
    auto gmax = [](float x) { return group_max<W>(x); };
    const float* irow = in + ((size_t)l * N + y) * M * NOUT;
    float r0[8], r1[8], r2[8], m0 = 0.f, m1 = 0.f;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const float* p = irow + (size_t)(j + W * r) * NOUT;
        r0[r] = p[0];
        r1[r] = p[1];
        r2[r] = p[2];
        m0 = fmaxf(m0, fabsf(r0[r]));
        m1 = fmaxf(m1, fabsf(r1[r]));
    }
    const float sg = pair_scale(gmax(m0), gmax(m1)), isg = inv_pow2(sg);
    const float s = 1.0f / ((float)M * (float)N);
    cpx<float> v[2][8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        v[0][r] = mk(r0[r], sg * r1[r]);
        v[1][r] = mk(r2[r], 0.f);
    }
    fft_group<LOGM, false, 2>(v, sb, tw, j);
    // field 2 directly; fields 0, 1 from Z_k and Z_{M-k} (another lane: through the scratch)
    cpx<float>* o2 = hyb + (size_t)2 * fstride + roff;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int k = j + W * r;
        if (k < KX) o2[k] = cscale(v[1][r], s);
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) sb[pad8(j + W * r)] = v[0][r];
    wave_sync();
    cpx<float>* o0 = hyb + roff;
    cpx<float>* o1 = hyb + fstride + roff;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int k = j + W * r;
        if (k < KX) {
            const cpx<float> zk = v[0][r], zm = sb[pad8((M - k) & (M - 1))];
            o0[k] = mk(0.5f * s * (zk.x + zm.x), 0.5f * s * (zk.y - zm.y));
            o1[k] = mk(0.5f * s * isg * (zk.y + zm.y), -0.5f * s * isg * (zk.x - zm.x));
        }
    }
}

// Projection of the level spectra of the three predicted fields onto the modes, added to
// N: N_t += sum_l pw[0][m][l] s0_l, N_f += sum_l pw[1][m][l] s1_l, N_g += ...; pw holds
// w_l psi_m(z_l) / <psi_m, psi_m>.  One thread per (field, mode, m); the correction is also
// kept in dN (reused by steps that do not re-evaluate the network).  Per-block sums of
// |dN|^2 and |N|^2 go to guard.
template <typename T>
__global__ void __launch_bounds__(256) k_closure_project(DevGeom g, int nz, const T* __restrict__ pw,
                                                         const cpx<float>* __restrict__ spec, cpx<T>* __restrict__ N0,
                                                         cpx<T>* __restrict__ N1, cpx<T>* __restrict__ N2,
                                                         cpx<T>* __restrict__ dN, double* __restrict__ guard,
                                                         T gain, T k2cut, const T* __restrict__ kxa,
                                                         const T* __restrict__ kya) {
    __shared__ double sd[256], sn[256];
    const int L = g.L, P = g.P;
    const size_t PL = (size_t)P * L;
    const size_t e3 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double dd = 0.0, nn = 0.0;
    if (e3 < 3 * PL) {
        const int f = (int)(e3 / PL);
        const size_t e = e3 - f * PL;
        const int p = (int)(e / L), m = (int)(e - (size_t)p * L);
        const size_t plane = (size_t)nz * P;
        cpx<T>* out = f == 0 ? N0 : (f == 1 ? N1 : N2);
        const T* w = pw + ((size_t)f * L + m) * nz;
        T ax = 0, ay = 0;
        // eight levels' loads in flight at a time; the sum runs in level order
        for (int l0 = 0; l0 < nz; l0 += 8) {
            cpx<float> v[8];
            T wl[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int l = l0 + u;
                v[u] = l < nz ? spec[f * plane + (size_t)l * P + p] : mk<float>(0, 0);
                wl[u] = l < nz ? w[l] : T(0);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (l0 + u < nz) {
                    ax += wl[u] * (T)v[u].x;
                    ay += wl[u] * (T)v[u].y;
                }
        }
        // damping: the correction times gain on the modes with |k|^2 <= k2cut, zero above
        const T k2 = kxa[p] * kxa[p] + kya[p] * kya[p];
        const T fac = k2 <= k2cut ? gain : T(0);
        ax *= fac;
        ay *= fac;
        cpx<T> n = out[e];
        nn = (double)n.x * n.x + (double)n.y * n.y;
        dd = (double)ax * ax + (double)ay * ay;
        dN[e3] = mk<T>(ax, ay);
        n.x += ax;
        n.y += ay;
        out[e] = n;
    }
    sd[threadIdx.x] = dd;
    sn[threadIdx.x] = nn;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            sd[threadIdx.x] += sd[threadIdx.x + h];
            sn[threadIdx.x] += sn[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        guard[2 * blockIdx.x] = sd[0];
        guard[2 * blockIdx.x + 1] = sn[0];
    }
}

// N += dN (the last evaluated correction) on the steps between evaluations
template <typename T>
__global__ void __launch_bounds__(256) k_closure_reuse(size_t PL, const cpx<T>* __restrict__ dN,
                                                       cpx<T>* __restrict__ N0, cpx<T>* __restrict__ N1,
                                                       cpx<T>* __restrict__ N2) {
    const size_t e3 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e3 >= 3 * PL) return;
    const int f = (int)(e3 / PL);
    const size_t e = e3 - f * PL;
    cpx<T>* out = f == 0 ? N0 : (f == 1 ? N1 : N2);
    out[e] = out[e] + dN[e3];
}

// truncation of one field from a finer context: b[pc][m] = a[pmap[pc]][m], m < Lc
template <typename Tf, typename Tc>
__global__ void k_restrict(int Pc, int Lc, int Lf, const int* __restrict__ pmap, const cpx<Tf>* __restrict__ a,
                           cpx<Tc>* __restrict__ b) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Pc * Lc) return;
    const int pc = e / Lc, m = e - pc * Lc;
    const cpx<Tf> v = a[(size_t)pmap[pc] * Lf + m];
    b[e] = mk<Tc>((Tc)v.x, (Tc)v.y);
}

// ---- the network ---------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// tanh(x) = 1 - 2 / (e^{2x} + 1) with the hardware exponential and reciprocal (absolute error
// ~1e-7; saturates to +-1 and propagates NaN)
__device__ __forceinline__ float tanh_fast(float x) {
    const float e = __expf(2.f * x);
    return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}
// FP32 -> BF16, round to nearest even (one v_cvt_pk_bf16_f32 per pair on gfx950)
__device__ __forceinline__ __bf16 bf16_of(float x) { return (__bf16)x; }
__device__ __forceinline__ float f32_of(__bf16 h) {
    return __uint_as_float((unsigned)__builtin_bit_cast(unsigned short, h) << 16);
}
__device__ __forceinline__ __bf16 bf16_bits(uint16_t u) { return __builtin_bit_cast(__bf16, u); }

// Inference on the matrix cores, 32 points per wave.  Both hidden layers are computed
// transposed, H = W X^T (rows = hidden units, columns = points), with v_mfma_f32_32x32x16_bf16:
// the weights are BF16 already and each FP32 activation goes in as a BF16 hi + lo pair (two
// MFMAs), so the products keep FP32 accuracy.  The 32x32 accumulator of a lane holds column
// l % 32 (its point) and rows 8 (r / 4) + 4 (l / 32) + r % 4, r < 16; those registers are
// directly the next layer's B operand (lane half h supplies K elements 8 h .. 8 h + 7) if the
// K order of the next A operand is permuted to match (the a2 fragments below).  The 3 outputs
// are dot products over each lane's 32 hidden values plus one exchange between lane halves.
__global__ void __launch_bounds__(256) k_mlp_infer(int npts, const uint16_t* __restrict__ wbf, Norm nm,
                                                   const float* __restrict__ X, float* __restrict__ Y) {
    __shared__ bf16x8 a1s[2][64];
    __shared__ bf16x8 a2s[2][4][64];
    __shared__ float b1s[NH], b2s[NH], w3s[NOUT * NH], b3s[NOUT];
    const int tid = threadIdx.x;
    for (int i = tid; i < 2 * 64; i += blockDim.x) {
        const int t = i >> 6, l = i & 63, h = l >> 5, row = 32 * t + (l & 31);
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = 8 * h + e;
            v[e] = k < NIN ? bf16_bits(wbf[OFF_W1 + row * NIN + k]) : bf16_of(0.f);
        }
        a1s[t][l] = v;
    }
    for (int i = tid; i < 2 * 4 * 64; i += blockDim.x) {
        const int t2 = i >> 8, sblk = (i >> 6) & 3, l = i & 63, h = l >> 5, row = 32 * t2 + (l & 31);
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int r = 8 * (sblk & 1) + e;
            const int feat = 32 * (sblk >> 1) + 8 * (r >> 2) + 4 * h + (r & 3);
            v[e] = bf16_bits(wbf[OFF_W2 + row * NH + feat]);
        }
        a2s[t2][sblk][l] = v;
    }
    for (int i = tid; i < NH; i += blockDim.x) {
        b1s[i] = bf16_bits_to_f32(wbf[OFF_B1 + i]);
        b2s[i] = bf16_bits_to_f32(wbf[OFF_B2 + i]);
    }
    for (int i = tid; i < NOUT * NH; i += blockDim.x) w3s[i] = bf16_bits_to_f32(wbf[OFF_W3 + i]);
    if (tid < NOUT) b3s[tid] = bf16_bits_to_f32(wbf[OFF_B3 + tid]);
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6, h = lane >> 5;
    const int ngroups = (npts + 31) / 32;
    for (int gi = blockIdx.x * 4 + wave; gi < ngroups; gi += gridDim.x * 4) {  // wave-uniform loop
        const int pt = gi * 32 + (lane & 31);
        const bool valid = pt < npts;
        bf16x8 xh, xl;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float v = (h == 0 && valid) ? (X[(size_t)pt * NIN + e] - nm.in_mean[e]) * nm.in_inv[e] : 0.f;
            xh[e] = bf16_of(v);
            xl[e] = bf16_of(v - f32_of(xh[e]));
        }
        f32x16 d1[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            f32x16 c = {};
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1s[t][lane], xh, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1s[t][lane], xl, c, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) c[r] = tanh_fast(c[r] + b1s[32 * t + 8 * (r >> 2) + 4 * h + (r & 3)]);
            d1[t] = c;
        }
        f32x16 d2[2] = {};
#pragma unroll
        for (int sblk = 0; sblk < 4; ++sblk) {
            bf16x8 hh, hl;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float v = d1[sblk >> 1][8 * (sblk & 1) + e];
                hh[e] = bf16_of(v);
                hl[e] = bf16_of(v - f32_of(hh[e]));
            }
#pragma unroll
            for (int t2 = 0; t2 < 2; ++t2) {
                d2[t2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2s[t2][sblk][lane], hh, d2[t2], 0, 0, 0);
                d2[t2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2s[t2][sblk][lane], hl, d2[t2], 0, 0, 0);
            }
        }
        float y0 = 0.f, y1 = 0.f, y2 = 0.f;
#pragma unroll
        for (int t2 = 0; t2 < 2; ++t2)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int feat = 32 * t2 + 8 * (r >> 2) + 4 * h + (r & 3);
                const float hv = tanh_fast(d2[t2][r] + b2s[feat]);
                y0 += w3s[feat] * hv;
                y1 += w3s[NH + feat] * hv;
                y2 += w3s[2 * NH + feat] * hv;
            }
        y0 += __shfl_xor(y0, 32);
        y1 += __shfl_xor(y1, 32);
        y2 += __shfl_xor(y2, 32);
        if (h == 0 && valid) {
            Y[(size_t)pt * NOUT] = (y0 + b3s[0]) * nm.out_scale[0];
            Y[(size_t)pt * NOUT + 1] = (y1 + b3s[1]) * nm.out_scale[1];
            Y[(size_t)pt * NOUT + 2] = (y2 + b3s[2]) * nm.out_scale[2];
        }
    }
}

int infer_blocks(int npts) {
    const int groups = (npts + 31) / 32;
    return std::max(1, std::min(2048, (groups + 3) / 4));
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Training: gradient of the mean squared error (normalised targets, mean over points and
// outputs) over `batch` points.  One wave per block; lane j is hidden unit j of both layers.
// Wave w takes batch points [w per, (w+1) per); the point index is drawn uniformly from the
// dataset (random != 0) or is the batch position itself.  Partial gradients per wave go to
// gpart[w][NW], partial losses to lpart[w].
__global__ void __launch_bounds__(64) k_mlp_train(int count, int batch, int per, int random, uint64_t seed,
                                                  uint64_t step, const uint16_t* __restrict__ wbf, Norm nm,
                                                  const float* __restrict__ X, const float* __restrict__ Y,
                                                  float* __restrict__ gpart, double* __restrict__ lpart) {
    __shared__ float sW2[NH * NH];
    __shared__ float sh1[NH], sd2[NH];
    const int j = threadIdx.x, wv = blockIdx.x;
    for (int i = j; i < NH * NH; i += 64) sW2[i] = bf16_bits_to_f32(wbf[OFF_W2 + i]);
    float w1[NIN], w2[NH], w3[NOUT], b3[NOUT];
#pragma unroll
    for (int i = 0; i < NIN; ++i) w1[i] = bf16_bits_to_f32(wbf[OFF_W1 + j * NIN + i]);
    const float b1 = bf16_bits_to_f32(wbf[OFF_B1 + j]), b2 = bf16_bits_to_f32(wbf[OFF_B2 + j]);
#pragma unroll
    for (int o = 0; o < NOUT; ++o) {
        w3[o] = bf16_bits_to_f32(wbf[OFF_W3 + o * NH + j]);
        b3[o] = bf16_bits_to_f32(wbf[OFF_B3 + o]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NH; ++i) w2[i] = sW2[j * NH + i];
    float g1[NIN] = {}, g2[NH] = {}, g3[NOUT] = {}, gb3[NOUT] = {};
    float gb1 = 0.f, gb2 = 0.f;
    double loss = 0.0;
    const float inv_outs = 1.0f / ((float)batch * NOUT);
    const int b0 = wv * per, b1e = min(batch, b0 + per);
    for (int bi = b0; bi < b1e; ++bi) {
        size_t idx = (size_t)bi;
        if (random) idx = (size_t)(mix64(seed * 0x100000001B3ull + step * 0x9E3779B1ull + (uint64_t)bi) % (uint64_t)count);
        float x[NIN];
#pragma unroll
        for (int i = 0; i < NIN; ++i) x[i] = (X[idx * NIN + i] - nm.in_mean[i]) * nm.in_inv[i];
        float a1 = b1;
#pragma unroll
        for (int i = 0; i < NIN; ++i) a1 += w1[i] * x[i];
        const float h1 = tanh_fast(a1);
        sh1[j] = h1;
        wave_sync();
        float a2 = b2;
#pragma unroll
        for (int i = 0; i < NH; ++i) a2 += w2[i] * sh1[i];
        const float h2 = tanh_fast(a2);
        float dy[NOUT];
        float d2s = 0.f;
#pragma unroll
        for (int o = 0; o < NOUT; ++o) {
            const float y = wave_sum(w3[o] * h2) + b3[o];
            const float t = Y[idx * NOUT + o] / nm.out_scale[o];
            const float e = y - t;
            loss += (double)e * e;
            dy[o] = 2.f * e * inv_outs;
            g3[o] += dy[o] * h2;
            gb3[o] += dy[o];
            d2s += w3[o] * dy[o];
        }
        const float d2 = d2s * (1.f - h2 * h2);
        gb2 += d2;
#pragma unroll
        for (int i = 0; i < NH; ++i) g2[i] += d2 * sh1[i];
        sd2[j] = d2;
        wave_sync();
        float d1s = 0.f;
#pragma unroll
        for (int k = 0; k < NH; ++k) d1s += sW2[k * NH + j] * sd2[k];
        const float d1 = d1s * (1.f - h1 * h1);
        gb1 += d1;
#pragma unroll
        for (int i = 0; i < NIN; ++i) g1[i] += d1 * x[i];
        wave_sync();  // sh1 / sd2 are rewritten by the next point
    }
    float* gp = gpart + (size_t)wv * NW;
#pragma unroll
    for (int i = 0; i < NIN; ++i) gp[OFF_W1 + j * NIN + i] = g1[i];
    gp[OFF_B1 + j] = gb1;
#pragma unroll
    for (int i = 0; i < NH; ++i) gp[OFF_W2 + j * NH + i] = g2[i];
    gp[OFF_B2 + j] = gb2;
#pragma unroll
    for (int o = 0; o < NOUT; ++o) gp[OFF_W3 + o * NH + j] = g3[o];
    if (j == 0) {
#pragma unroll
        for (int o = 0; o < NOUT; ++o) gp[OFF_B3 + o] = gb3[o];
        lpart[wv] = loss / ((double)batch * NOUT);
    }
}

// sum of the per-wave partials (fixed order) and the batch loss into lossbuf[slot]
__global__ void k_grad_reduce(int nw, const float* __restrict__ gpart, const double* __restrict__ lpart,
                              float* __restrict__ grad, double* __restrict__ lossbuf, int slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < NW) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += gpart[(size_t)w * NW + i];
        grad[i] = (float)s;
    }
    if (i == 0) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += lpart[w];
        lossbuf[slot] = s;
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; writes the BF16 copies
__global__ void k_adam(const float* __restrict__ grad, float* __restrict__ wm, float* __restrict__ m1,
                       float* __restrict__ m2, uint16_t* __restrict__ wbf, float lr, float beta1, float beta2,
                       float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NW) return;
    const float gr = grad[i];
    const float a = beta1 * m1[i] + (1.f - beta1) * gr;
    const float b = beta2 * m2[i] + (1.f - beta2) * gr * gr;
    m1[i] = a;
    m2[i] = b;
    const float w = wm[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    wm[i] = w;
    wbf[i] = f32_to_bf16_bits(w);
}

// per-column sums and sums of squares (FP64) of a row-major [n][nc] array, per block
__global__ void k_colstats(int n, int nc, const float* __restrict__ A, double* __restrict__ part) {
    __shared__ double s[2 * NIN][256];  // nc <= NIN
    double acc[2 * NIN];
    for (int c = 0; c < 2 * nc; ++c) acc[c] = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
        for (int c = 0; c < nc; ++c) {
            const double v = A[(size_t)r * nc + c];
            acc[c] += v;
            acc[nc + c] += v * v;
        }
    for (int c = 0; c < 2 * nc; ++c) s[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int c = 0; c < 2 * nc; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 2 * nc; ++c) part[(size_t)blockIdx.x * 2 * nc + c] = s[c][0];
}

// ---- host side -----------------------------------------------------------------------------
std::vector<float> twiddles(int n) {
    std::vector<float> t(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        t[2 * k] = (float)std::cos(-2.0 * M_PI * k / n);
        t[2 * k + 1] = (float)std::sin(-2.0 * M_PI * k / n);
    }
    return t;
}

}  // namespace

struct rbc_closure {
    rbc_ctx* c = nullptr;
    int nz = 0, npts = 0, cap = 0, count = 0;
    DevGeom gz{};                     // coarse geometry with nq = nz (the closure's levels)
    std::vector<void*> owned;
    std::string err;
    // tables
    void* zb = nullptr;               // [nz][4][L] in the context's precision
    float* zl = nullptr;              // [nz]
    void* pw = nullptr;               // [3][L][nz] in the context's precision
    cpx<float>* twM = nullptr;
    cpx<float>* twN = nullptr;
    // scratch
    cpx<float>* spec = nullptr;       // [7][nz][P]
    cpx<float>* hyb = nullptr;        // [7][nz][N][KXp]
    float* feat = nullptr;            // [npts][8]
    float* pred = nullptr;            // [npts][3]
    double* guard = nullptr;          // [nblk][2]
    int nguard = 0;
    void* dN = nullptr;               // [3][P][L] last correction, the context's precision
    int every = 1;                    // network evaluated on every `every`-th corrected step
    int phase = 0;                    // corrected steps since enable
    // sampling
    int* pmap = nullptr;              // coarse compact mode -> fine compact mode (for the fine geometry below)
    int pmap_M = 0, pmap_N = 0, pmap_L = 0;
    void* Nt[3] = {};                 // restricted fine N (coarse precision)
    // dataset
    float* DX = nullptr;              // [cap][8]
    float* DY = nullptr;              // [cap][3]
    // network
    float* wm = nullptr;              // FP32 master
    uint16_t* wbf = nullptr;          // BF16 copy used by the kernels
    float* grad = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    float* gpart = nullptr;           // [kTrainWaves][NW]
    double* lpart = nullptr;          // [kTrainWaves]
    double* lossbuf = nullptr;        // [lossbuf_n]
    int lossbuf_n = 0;
    int adam_t = 0;
    Norm nm{};
    // run-time state
    bool enabled = false, fallback = false;
    double max_ratio = 1.0, last_ratio = 0.0;
    double gain = 1.0, k2cut = HUGE_VAL;  // damping (rbc_closure_damping): all modes at full gain
    std::vector<void*> snap;          // state + history copy for the fallback
    ~rbc_closure() {
        if (c) {
            if (c->post_user == this) {
                c->post_nonlinear = nullptr;
                c->post_user = nullptr;
            }
            if (c->stream) hipStreamSynchronize(c->stream);
        }
        for (void* p : owned) hipFree(p);
    }
    void* al(size_t bytes) {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, bytes > 0 ? bytes : 16));
        owned.push_back(p);
        return p;
    }
};

namespace {

int cfail(rbc_closure* cl, int code, const std::string& msg) {
    cl->err = msg;
    return code;
}

int chip(rbc_closure* cl, const HipError& e) {
    return cfail(cl, RBC_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
}

template <typename T>
void launch_feat_levels(rbc_closure* cl, void* const X[3], const void* XF, const void* XG) {
    rbc_ctx* c = cl->c;
    const DevGeom& g = cl->gz;
    int MB = 64;
    while (MB > 16 && ((size_t)3 * g.L * MB + 2 * (size_t)cl->nz) * sizeof(cpx<T>) > 65536) MB /= 2;
    const size_t sm = ((size_t)3 * g.L * MB + 2 * (size_t)cl->nz) * sizeof(cpx<T>);
    hipLaunchKernelGGL(k_feat_levels<T>, dim3((g.P + MB - 1) / MB), dim3(256), sm, c->stream, g, cl->nz, MB,
                       (const T*)cl->zb,
                       (const cpx<T>*)X[0], (const cpx<T>*)X[1], (const cpx<T>*)X[2], (const cpx<T>*)XF,
                       (const cpx<T>*)XG, (const T*)c->b.kx, (const T*)c->b.ky, cl->spec);
}

int y_tile(int N) {
    int tx = (int)(65536 / (2 * (size_t)N * sizeof(cpx<float>)));
    return std::max(1, std::min(16, tx));
}

// the F level-spectrum fields in cl->spec -> point values (os floats per point)
void synth_points(rbc_closure* cl, int F, int os, int zcol, float* out) {
    rbc_ctx* c = cl->c;
    const DevGeom& g = cl->gz;
    if (!launch_y_inv_group(g, F, cl->spec, cl->twN, cl->hyb, c->stream)) {
        const int TX = y_tile(g.N);
        hipLaunchKernelGGL(k_fft_y_inv<float>, dim3((g.KX + TX - 1) / TX, cl->nz, F), dim3(256),
                           (size_t)2 * TX * g.N * sizeof(cpx<float>), c->stream, g, TX, (const cpx<float>*)cl->spec,
                           c->b.rows, (const cpx<float>*)cl->twN, cl->hyb);
    }
    if (g.M >= 64 && g.M <= 256 && g.N % (2048 / g.M) == 0 && g.KX == kept_cols(g.M)) {
        const dim3 gr(g.N / (2048 / g.M), cl->nz);
#define RBC_SYNTHW(LG)                                                                                    \
    hipLaunchKernelGGL(k_synth_rows_w<LG>, gr, dim3(256), 0, c->stream, g, F, os, zcol, (const float*)cl->zl, \
                       (const cpx<float>*)cl->hyb, (const cpx<float>*)cl->twM, out)
        if (g.M == 64) RBC_SYNTHW(6);
        else if (g.M == 128) RBC_SYNTHW(7);
        else RBC_SYNTHW(8);
#undef RBC_SYNTHW
        return;
    }
    hipLaunchKernelGGL(k_synth_rows, dim3(g.N, cl->nz), dim3(256), (size_t)2 * F * g.M * sizeof(cpx<float>),
                       c->stream, g, F, os, zcol, (const float*)cl->zl, (const cpx<float>*)cl->hyb,
                       (const cpx<float>*)cl->twM, out);
}

void features_of(rbc_closure* cl, void* const X[3], const void* XF, const void* XG, float* out) {
    if (cl->c->rsz == 8) launch_feat_levels<double>(cl, X, XF, XG);
    else launch_feat_levels<float>(cl, X, XF, XG);
    synth_points(cl, NFEAT, NIN, NFEAT, out);
}

template <typename T>
void launch_project(rbc_closure* cl) {
    rbc_ctx* c = cl->c;
    const DevGeom& g = cl->gz;
    hipLaunchKernelGGL(k_closure_project<T>, dim3(cl->nguard), dim3(256), 0, c->stream, g, cl->nz,
                       (const T*)cl->pw, (const cpx<float>*)cl->spec, (cpx<T>*)c->b.Nn[0], (cpx<T>*)c->b.Nn[1],
                       (cpx<T>*)c->b.Nn[2], (cpx<T>*)cl->dN, cl->guard, (T)cl->gain, (T)cl->k2cut,
                       (const T*)c->b.kx, (const T*)c->b.ky);
}

template <typename T>
void launch_reuse(rbc_closure* cl) {
    rbc_ctx* c = cl->c;
    const size_t PL = (size_t)cl->gz.P * cl->gz.L;
    hipLaunchKernelGGL(k_closure_reuse<T>, dim3((unsigned)((3 * PL + 255) / 256)), dim3(256), 0, c->stream, PL,
                       (const cpx<T>*)cl->dN, (cpx<T>*)c->b.Nn[0], (cpx<T>*)c->b.Nn[1], (cpx<T>*)c->b.Nn[2]);
}

// the correction inside a coarse step (registered as the context's post-nonlinear hook)
void apply_closure(rbc_ctx* c, void* user) {
    rbc_closure* cl = (rbc_closure*)user;
    if (!cl->enabled) return;
    if (cl->phase++ % cl->every != 0) {  // between evaluations: the last correction again
        if (c->rsz == 8) launch_reuse<double>(cl);
        else launch_reuse<float>(cl);
        return;
    }
    features_of(cl, c->b.X, c->b.XF, c->b.XG, cl->feat);
    hipLaunchKernelGGL(k_mlp_infer, dim3(infer_blocks(cl->npts)), dim3(256), 0, c->stream, cl->npts,
                       (const uint16_t*)cl->wbf, cl->nm, (const float*)cl->feat, cl->pred);
    const DevGeom& g = cl->gz;
    if (g.M >= 64 && g.M <= 256 && g.N % (2048 / g.M) == 0 && g.KX == kept_cols(g.M)) {
        const dim3 gr(g.N / (2048 / g.M), cl->nz);
        if (g.M == 64)
            hipLaunchKernelGGL(k_anal_rows_w<6>, gr, dim3(256), 0, c->stream, g, (const float*)cl->pred,
                               (const cpx<float>*)cl->twM, cl->hyb);
        else if (g.M == 128)
            hipLaunchKernelGGL(k_anal_rows_w<7>, gr, dim3(256), 0, c->stream, g, (const float*)cl->pred,
                               (const cpx<float>*)cl->twM, cl->hyb);
        else
            hipLaunchKernelGGL(k_anal_rows_w<8>, gr, dim3(256), 0, c->stream, g, (const float*)cl->pred,
                               (const cpx<float>*)cl->twM, cl->hyb);
    } else {
        hipLaunchKernelGGL(k_anal_rows, dim3(g.N, cl->nz), dim3(256), (size_t)2 * NOUT * g.M * sizeof(cpx<float>),
                           c->stream, g, (const float*)cl->pred, (const cpx<float>*)cl->twM, cl->hyb);
    }
    if (!launch_y_fwd_group(g, 3, cl->hyb, cl->twN, cl->spec, c->stream)) {
        const int TX = y_tile(g.N);
        hipLaunchKernelGGL(k_fft_y_fwd<float>, dim3((g.KX + TX - 1) / TX, cl->nz, 3), dim3(256),
                           (size_t)2 * TX * g.N * sizeof(cpx<float>), c->stream, g, TX, (const cpx<float>*)cl->hyb,
                           c->b.rows, (const cpx<float>*)cl->twN, cl->spec);
    }
    if (c->rsz == 8) launch_project<double>(cl);
    else launch_project<float>(cl);
}

void upload_weights_bf16(rbc_closure* cl) {
    std::vector<float> w(NW);
    HIPCHK(hipMemcpy(w.data(), cl->wm, NW * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<uint16_t> h(NW);
    for (int i = 0; i < NW; ++i) h[i] = f32_to_bf16_bits(w[i]);
    HIPCHK(hipMemcpy(cl->wbf, h.data(), NW * sizeof(uint16_t), hipMemcpyHostToDevice));
}

// restrict map from a fine context's compact modes; false if a coarse mode has no fine twin
bool build_pmap(rbc_closure* cl, const rbc_ctx* f) {
    const Geometry& gc = cl->c->geo;
    const Geometry& gf = f->geo;
    std::vector<int> rowpos(gf.N, -1);
    for (int i = 0; i < gf.NYc(); ++i) rowpos[gf.rows[i]] = i;
    std::vector<int> map(gc.P());
    for (int kx = 0; kx < gc.KX; ++kx)
        for (int kc = 0; kc < gc.NYc(); ++kc) {
            const int row = gc.rows[kc];
            const int ky = row <= gc.N / 2 ? row : row - gc.N;
            const int frow = ky >= 0 ? ky : gf.N + ky;
            if (kx >= gf.KX || frow < 0 || frow >= gf.N || rowpos[frow] < 0) return false;
            map[kx * gc.NYc() + kc] = kx * gf.NYc() + rowpos[frow];
        }
    if (!cl->pmap) cl->pmap = (int*)cl->al(map.size() * sizeof(int));
    HIPCHK(hipMemcpy(cl->pmap, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice));
    cl->pmap_M = gf.M;
    cl->pmap_N = gf.N;
    cl->pmap_L = gf.L;
    return true;
}

template <typename Tf, typename Tc>
void restrict_fields(rbc_closure* cl, const rbc_ctx* f, const void* const src[3], const void* sF, const void* sG,
                     void* const dst[3], void* dF, void* dG) {
    rbc_ctx* c = cl->c;
    const int Pc = c->dg.P, Lc = c->dg.L, Lf = f->dg.L;
    const dim3 grid((Pc * Lc + 255) / 256);
    for (int i = 0; i < 3; ++i)
        hipLaunchKernelGGL((k_restrict<Tf, Tc>), grid, dim3(256), 0, c->stream, Pc, Lc, Lf, (const int*)cl->pmap,
                           (const cpx<Tf>*)src[i], (cpx<Tc>*)dst[i]);
    // the mean-flow vectors: one "mode" (index 0) with Lc entries
    if (sF && dF)
        hipLaunchKernelGGL((k_restrict<Tf, Tc>), dim3(1), dim3(256), 0, c->stream, 1, Lc, Lf, (const int*)cl->pmap,
                           (const cpx<Tf>*)sF, (cpx<Tc>*)dF);
    if (sG && dG)
        hipLaunchKernelGGL((k_restrict<Tf, Tc>), dim3(1), dim3(256), 0, c->stream, 1, Lc, Lf, (const int*)cl->pmap,
                           (const cpx<Tf>*)sG, (cpx<Tc>*)dG);
}

void restrict_any(rbc_closure* cl, const rbc_ctx* f, const void* const src[3], const void* sF, const void* sG,
                  void* const dst[3], void* dF, void* dG) {
    const bool fd = f->rsz == 8, cd = cl->c->rsz == 8;
    if (fd && cd) restrict_fields<double, double>(cl, f, src, sF, sG, dst, dF, dG);
    else if (fd) restrict_fields<double, float>(cl, f, src, sF, sG, dst, dF, dG);
    else if (cd) restrict_fields<float, double>(cl, f, src, sF, sG, dst, dF, dG);
    else restrict_fields<float, float>(cl, f, src, sF, sG, dst, dF, dG);
}

// pmap's entry 0 must be the k = 0 mode for the vector restriction (compact mode 0 is
// (kx, ky) = (0, 0) in both geometries)
int check_nesting(rbc_closure* cl, const rbc_ctx* f) {
    rbc_ctx* c = cl->c;
    if (f->p.device != c->p.device) return cfail(cl, RBC_ERR_ARG, "fine and coarse contexts must share a device");
    if (f->p.Lx != c->p.Lx || f->p.Ly != c->p.Ly) return cfail(cl, RBC_ERR_ARG, "fine and coarse boxes differ");
    if (f->dg.L < c->dg.L || f->dg.KX < c->dg.KX || f->dg.NYc < c->dg.NYc)
        return cfail(cl, RBC_ERR_ARG, "the coarse kept modes must be a subset of the fine ones");
    return RBC_OK;
}

template <typename T>
void launch_res_levels(rbc_closure* cl) {
    rbc_ctx* c = cl->c;
    dim3 grid((c->dg.P + 255) / 256, cl->nz);
    hipLaunchKernelGGL(k_res_levels<T>, grid, dim3(256), 0, c->stream, cl->gz, cl->nz, (const T*)cl->zb,
                       (const cpx<T>*)cl->Nt[0], (const cpx<T>*)cl->Nt[1], (const cpx<T>*)cl->Nt[2],
                       (const cpx<T>*)c->b.Oo[0], (const cpx<T>*)c->b.Oo[1], (const cpx<T>*)c->b.Oo[2], cl->spec);
}

// bytes of snapshot entry i: three fields, F, G, three history fields, their F, G, and the
// step-health word pair (a chunk redone from the snapshot must not keep a non-finite flag)
size_t snap_bytes(const rbc_ctx* c, int i) {
    if (i == 10) return 2 * sizeof(int);
    return (i % 5) < 3 ? (size_t)c->dg.P * c->dg.L * c->csz() : (size_t)c->dg.L * c->csz();
}

void take_snapshot(rbc_closure* cl) {
    rbc_ctx* c = cl->c;
    void* src[11] = {c->b.X[0], c->b.X[1], c->b.X[2], c->b.XF, c->b.XG, c->b.Rp[0], c->b.Rp[1], c->b.Rp[2],
                     c->b.RpF, c->b.RpG, c->b.health};
    for (int i = 0; i < 11; ++i)
        HIPCHK(hipMemcpyAsync(cl->snap[i], src[i], snap_bytes(c, i), hipMemcpyDeviceToDevice, c->stream));
}

void restore_snapshot(rbc_closure* cl) {
    rbc_ctx* c = cl->c;
    void* dst[11] = {c->b.X[0], c->b.X[1], c->b.X[2], c->b.XF, c->b.XG, c->b.Rp[0], c->b.Rp[1], c->b.Rp[2],
                     c->b.RpF, c->b.RpG, c->b.health};
    for (int i = 0; i < 11; ++i)
        HIPCHK(hipMemcpyAsync(dst[i], cl->snap[i], snap_bytes(c, i), hipMemcpyDeviceToDevice, c->stream));
}

// |dN| / |N| over the last correction (all three fields)
double read_ratio(rbc_closure* cl) {
    std::vector<double> g(2 * (size_t)cl->nguard);
    HIPCHK(hipMemcpyAsync(g.data(), cl->guard, g.size() * sizeof(double), hipMemcpyDeviceToHost, cl->c->stream));
    HIPCHK(hipStreamSynchronize(cl->c->stream));
    double d = 0, n = 0;
    for (int i = 0; i < cl->nguard; ++i) {
        d += g[2 * i];
        n += g[2 * i + 1];
    }
    if (!std::isfinite(d) || !std::isfinite(n)) return NAN;  // a non-finite N or correction
    return (n > 0) ? std::sqrt(d / n) : (d > 0 ? INFINITY : 0.0);
}

}  // namespace

extern "C" {

int rbc_closure_create(rbc_ctx* c, int nz, int capacity_points, rbc_closure** out) {
    if (!c || !out) return RBC_ERR_ARG;
    *out = nullptr;
    if (nz <= 0) nz = 2 * c->dg.L;
    if (nz > 512) return fail(c, RBC_ERR_ARG, "nz must be <= 512");
    if (capacity_points < 0) return fail(c, RBC_ERR_ARG, "capacity_points must be >= 0");
    // one LDS transform per input field and row: 2 * 7 * M complex floats within 64 KiB
    if ((size_t)2 * NFEAT * c->dg.M * sizeof(cpx<float>) > 65536)
        return fail(c, RBC_ERR_ARG, "the closure supports coarse grids with M <= 512");
    const size_t npts = (size_t)c->dg.M * c->dg.N * nz;
    if (npts > (size_t)1 << 30) return fail(c, RBC_ERR_ARG, "too many closure points");
    rbc_closure* cl = new rbc_closure;
    cl->c = c;
    cl->nz = nz;
    cl->npts = (int)npts;
    cl->cap = capacity_points;
    cl->gz = c->dg;
    cl->gz.nq = nz;
    try {
        hipSetDevice(c->p.device);
        const DevGeom& g = cl->gz;
        const int L = g.L;
        // Gauss-Legendre levels of the closure and the basis there
        std::vector<double> z, w;
        gauss_legendre(nz, z, w);
        std::vector<double> bv;
        basis_values(L, z, bv);  // [t][m][l], t = C, C', C'', C''', S, S'
        auto B = [&](int t, int m, int l) { return bv[((size_t)t * L + m) * nz + l]; };
        std::vector<double> zb((size_t)nz * 4 * L);
        for (int l = 0; l < nz; ++l)
            for (int m = 0; m < L; ++m) {
                zb[((size_t)l * 4 + 0) * L + m] = B(1, m, l);
                zb[((size_t)l * 4 + 1) * L + m] = B(0, m, l);
                zb[((size_t)l * 4 + 2) * L + m] = B(4, m, l);
                zb[((size_t)l * 4 + 3) * L + m] = B(5, m, l);
            }
        // projection weights w_l psi_m(z_l) / <psi_m, psi_m>, norms from a finer rule
        std::vector<double> zf, wf, bf;
        gauss_legendre(4 * L + 64, zf, wf);
        basis_values(L, zf, bf);
        const int nf = (int)zf.size();
        auto nrm = [&](int t, int m) {
            double s = 0;
            for (int q = 0; q < nf; ++q) s += wf[q] * bf[((size_t)t * L + m) * nf + q] * bf[((size_t)t * L + m) * nf + q];
            return s;
        };
        std::vector<double> pw((size_t)3 * L * nz);
        const int kinds[3] = {4, 0, 4};  // theta: S, f: C, g: S
        for (int f = 0; f < 3; ++f)
            for (int m = 0; m < L; ++m) {
                const double inv = 1.0 / nrm(kinds[f], m);
                for (int l = 0; l < nz; ++l) pw[((size_t)f * L + m) * nz + l] = w[l] * B(kinds[f], m, l) * inv;
            }
        std::vector<float> zlf(z.begin(), z.end());
        auto upload = [&](const std::vector<double>& v) {
            void* d = cl->al(v.size() * c->rsz);
            if (c->rsz == 8) {
                HIPCHK(hipMemcpy(d, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
            } else {
                std::vector<float> f(v.begin(), v.end());
                HIPCHK(hipMemcpy(d, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
            }
            return d;
        };
        cl->zb = upload(zb);
        cl->pw = upload(pw);
        cl->zl = (float*)cl->al(zlf.size() * sizeof(float));
        HIPCHK(hipMemcpy(cl->zl, zlf.data(), zlf.size() * sizeof(float), hipMemcpyHostToDevice));
        const std::vector<float> tm = twiddles(g.M), tn = twiddles(g.N);
        cl->twM = (cpx<float>*)cl->al(tm.size() * sizeof(float));
        cl->twN = (cpx<float>*)cl->al(tn.size() * sizeof(float));
        HIPCHK(hipMemcpy(cl->twM, tm.data(), tm.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(cl->twN, tn.data(), tn.size() * sizeof(float), hipMemcpyHostToDevice));
        cl->spec = (cpx<float>*)cl->al((size_t)NFEAT * nz * g.P * sizeof(cpx<float>));
        cl->hyb = (cpx<float>*)cl->al((size_t)NFEAT * nz * g.N * g.KXp * sizeof(cpx<float>));
        cl->feat = (float*)cl->al(npts * NIN * sizeof(float));
        cl->pred = (float*)cl->al(npts * NOUT * sizeof(float));
        cl->nguard = (int)((3 * (size_t)g.P * L + 255) / 256);
        cl->guard = (double*)cl->al((size_t)2 * cl->nguard * sizeof(double));
        cl->dN = cl->al((size_t)3 * g.P * L * c->csz());
        for (int i = 0; i < 3; ++i) cl->Nt[i] = cl->al((size_t)g.P * L * c->csz());
        cl->DX = (float*)cl->al((size_t)capacity_points * NIN * sizeof(float));
        cl->DY = (float*)cl->al((size_t)capacity_points * NOUT * sizeof(float));
        cl->wm = (float*)cl->al(NW * sizeof(float));
        cl->wbf = (uint16_t*)cl->al(NW * sizeof(uint16_t));
        cl->grad = (float*)cl->al(NW * sizeof(float));
        cl->m1 = (float*)cl->al(NW * sizeof(float));
        cl->m2 = (float*)cl->al(NW * sizeof(float));
        cl->gpart = (float*)cl->al((size_t)kTrainWaves * NW * sizeof(float));
        cl->lpart = (double*)cl->al(kTrainWaves * sizeof(double));
        for (int i = 0; i < 11; ++i) cl->snap.push_back(cl->al(snap_bytes(c, i)));
        // Glorot-uniform initial weights from SplitMix64, zero biases
        std::vector<float> w0(NW, 0.f);
        uint64_t ctr = 0;
        auto uni = [&]() { return (double)(mix64(0xC105EULL * 1000003ull + ctr++) >> 11) / 9007199254740992.0; };
        auto fill = [&](int off, int fan_out, int fan_in) {
            const double a = std::sqrt(6.0 / (fan_in + fan_out));
            for (int i = 0; i < fan_out * fan_in; ++i) w0[off + i] = (float)((2 * uni() - 1) * a);
        };
        fill(OFF_W1, NH, NIN);
        fill(OFF_W2, NH, NH);
        fill(OFF_W3, NOUT, NH);
        HIPCHK(hipMemcpy(cl->wm, w0.data(), NW * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(cl->m1, 0, NW * sizeof(float)));
        HIPCHK(hipMemset(cl->m2, 0, NW * sizeof(float)));
        upload_weights_bf16(cl);
        for (int i = 0; i < NIN; ++i) {
            cl->nm.in_mean[i] = 0.f;
            cl->nm.in_inv[i] = 1.f;
        }
        for (int o = 0; o < NOUT; ++o) cl->nm.out_scale[o] = 1.f;
    } catch (const HipError& e) {
        const int rc = hip_fail(c, e);
        delete cl;
        return rc;
    }
    *out = cl;
    return RBC_OK;
}

void rbc_closure_destroy(rbc_closure* cl) { delete cl; }

const char* rbc_closure_last_error(const rbc_closure* cl) { return cl ? cl->err.c_str() : ""; }

int rbc_closure_info(const rbc_closure* cl, int* npts, int* nweights, int* count, int* nz) {
    if (!cl) return RBC_ERR_ARG;
    if (npts) *npts = cl->npts;
    if (nweights) *nweights = NW;
    if (count) *count = cl->count;
    if (nz) *nz = cl->nz;
    return RBC_OK;
}

int rbc_closure_restrict_state(rbc_closure* cl, rbc_ctx* fine) {
    if (!cl || !fine) return RBC_ERR_ARG;
    if (int rc = check_nesting(cl, fine)) return rc;
    rbc_ctx* c = cl->c;
    try {
        hipSetDevice(c->p.device);
        HIPCHK(hipStreamSynchronize(fine->stream));
        if (!build_pmap(cl, fine)) return cfail(cl, RBC_ERR_ARG, "a coarse mode is missing from the fine context");
        restrict_any(cl, fine, fine->b.X, fine->b.XF, fine->b.XG, c->b.X, c->b.XF, c->b.XG);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        c->have_hist = false;
    } catch (const HipError& e) {
        return chip(cl, e);
    }
    return RBC_OK;
}

int rbc_closure_sample(rbc_closure* cl, rbc_ctx* fine, int* count_out) {
    if (!cl || !fine) return RBC_ERR_ARG;
    if (int rc = check_nesting(cl, fine)) return rc;
    if (cl->count + cl->npts > cl->cap) return cfail(cl, RBC_ERR_ARG, "closure dataset is full");
    rbc_ctx* c = cl->c;
    try {
        hipSetDevice(c->p.device);
        HIPCHK(hipStreamSynchronize(c->stream));
        if (!build_pmap(cl, fine)) return cfail(cl, RBC_ERR_ARG, "a coarse mode is missing from the fine context");
        // fine nonlinear term of the fine state (into the fine context's scratch N)
        nonlinear_op(fine, fine->b.X, fine->b.XF, fine->b.XG, fine->b.Oo, fine->b.OoF, fine->b.OoG);
        HIPCHK(hipStreamSynchronize(fine->stream));
        // truncations: state into the coarse operator staging, fine N into Nt
        restrict_any(cl, fine, fine->b.X, fine->b.XF, fine->b.XG, c->b.Xo, c->b.XoF, c->b.XoG);
        restrict_any(cl, fine, fine->b.Oo, nullptr, nullptr, cl->Nt, nullptr, nullptr);
        // coarse nonlinear term of the truncated state
        nonlinear_op(c, c->b.Xo, c->b.XoF, c->b.XoG, c->b.Oo, c->b.OoF, c->b.OoG);
        // inputs and targets at the points, appended to the dataset
        features_of(cl, c->b.Xo, c->b.XoF, c->b.XoG, cl->DX + (size_t)cl->count * NIN);
        if (c->rsz == 8) launch_res_levels<double>(cl);
        else launch_res_levels<float>(cl);
        synth_points(cl, NOUT, NOUT, -1, cl->DY + (size_t)cl->count * NOUT);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        return chip(cl, e);
    }
    cl->count += cl->npts;
    if (count_out) *count_out = cl->count;
    return RBC_OK;
}

int rbc_closure_dataset(rbc_closure* cl, int first, int n, float* X, float* Y) {
    if (!cl || first < 0 || n < 0 || first + n > cl->count) return RBC_ERR_ARG;
    try {
        hipSetDevice(cl->c->p.device);
        if (X) HIPCHK(hipMemcpy(X, cl->DX + (size_t)first * NIN, (size_t)n * NIN * sizeof(float), hipMemcpyDeviceToHost));
        if (Y) HIPCHK(hipMemcpy(Y, cl->DY + (size_t)first * NOUT, (size_t)n * NOUT * sizeof(float), hipMemcpyDeviceToHost));
    } catch (const HipError& e) {
        return chip(cl, e);
    }
    return RBC_OK;
}

int rbc_closure_load_dataset(rbc_closure* cl, int n, const float* X, const float* Y) {
    if (!cl || n < 0 || (n > 0 && (!X || !Y))) return RBC_ERR_ARG;
    if (n > cl->cap) return cfail(cl, RBC_ERR_ARG, "more points than the dataset capacity");
    try {
        hipSetDevice(cl->c->p.device);
        HIPCHK(hipMemcpy(cl->DX, X, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(cl->DY, Y, (size_t)n * NOUT * sizeof(float), hipMemcpyHostToDevice));
    } catch (const HipError& e) {
        return chip(cl, e);
    }
    cl->count = n;
    return RBC_OK;
}

int rbc_closure_clear_dataset(rbc_closure* cl) {
    if (!cl) return RBC_ERR_ARG;
    cl->count = 0;
    return RBC_OK;
}

int rbc_closure_fit_normalization(rbc_closure* cl, int n) {
    if (!cl) return RBC_ERR_ARG;
    if (n <= 0 || n > cl->count) n = cl->count;
    if (n < 2) return cfail(cl, RBC_ERR_STATE, "no training data");
    try {
        hipSetDevice(cl->c->p.device);
        const int nb = 256;
        double* part = nullptr;
        HIPCHK(hipMalloc(&part, (size_t)nb * 2 * 16 * sizeof(double)));
        std::vector<double> h((size_t)nb * 2 * 16);
        auto stats = [&](const float* A, int nc, std::vector<double>& mean, std::vector<double>& sd) {
            hipLaunchKernelGGL(k_colstats, dim3(nb), dim3(256), 0, cl->c->stream, n, nc, A, part);
            HIPCHK(hipMemcpyAsync(h.data(), part, (size_t)nb * 2 * nc * sizeof(double), hipMemcpyDeviceToHost,
                                  cl->c->stream));
            HIPCHK(hipStreamSynchronize(cl->c->stream));
            mean.assign(nc, 0.0);
            sd.assign(nc, 0.0);
            for (int cix = 0; cix < nc; ++cix) {
                double s = 0, q = 0;
                for (int b = 0; b < nb; ++b) {
                    s += h[(size_t)b * 2 * nc + cix];
                    q += h[(size_t)b * 2 * nc + nc + cix];
                }
                mean[cix] = s / n;
                sd[cix] = std::sqrt(std::max(0.0, q / n - mean[cix] * mean[cix]));
            }
        };
        std::vector<double> mx, sx, my, sy;
        stats(cl->DX, NIN, mx, sx);
        stats(cl->DY, NOUT, my, sy);
        hipFree(part);
        for (int i = 0; i < NIN; ++i) {
            cl->nm.in_mean[i] = (float)mx[i];
            cl->nm.in_inv[i] = sx[i] > 0 ? (float)(1.0 / sx[i]) : 1.f;
        }
        for (int o = 0; o < NOUT; ++o) {
            const double rms = std::sqrt(sy[o] * sy[o] + my[o] * my[o]);
            cl->nm.out_scale[o] = rms > 0 ? (float)rms : 1.f;
        }
    } catch (const HipError& e) {
        return chip(cl, e);
    }
    return RBC_OK;
}

int rbc_closure_get_normalization(const rbc_closure* cl, float* in_mean, float* in_inv, float* out_scale) {
    if (!cl) return RBC_ERR_ARG;
    for (int i = 0; i < NIN; ++i) {
        if (in_mean) in_mean[i] = cl->nm.in_mean[i];
        if (in_inv) in_inv[i] = cl->nm.in_inv[i];
    }
    for (int o = 0; o < NOUT; ++o)
        if (out_scale) out_scale[o] = cl->nm.out_scale[o];
    return RBC_OK;
}

int rbc_closure_set_normalization(rbc_closure* cl, const float* in_mean, const float* in_inv, const float* out_scale) {
    if (!cl || !in_mean || !in_inv || !out_scale) return RBC_ERR_ARG;
    for (int i = 0; i < NIN; ++i) {
        cl->nm.in_mean[i] = in_mean[i];
        cl->nm.in_inv[i] = in_inv[i];
    }
    for (int o = 0; o < NOUT; ++o) cl->nm.out_scale[o] = out_scale[o];
    return RBC_OK;
}

int rbc_closure_get_weights(rbc_closure* cl, float* w) {
    if (!cl || !w) return RBC_ERR_ARG;
    try {
        HIPCHK(hipMemcpy(w, cl->wm, NW * sizeof(float), hipMemcpyDeviceToHost));
    } catch (const HipError& e) {
        return chip(cl, e);
    }
    return RBC_OK;
}

int rbc_closure_set_weights(rbc_closure* cl, const float* w) {
    if (!cl || !w) return RBC_ERR_ARG;
    try {
        hipSetDevice(cl->c->p.device);
        HIPCHK(hipMemcpy(cl->wm, w, NW * sizeof(float), hipMemcpyHostToDevice));
        upload_weights_bf16(cl);
        HIPCHK(hipMemset(cl->m1, 0, NW * sizeof(float)));
        HIPCHK(hipMemset(cl->m2, 0, NW * sizeof(float)));
        cl->adam_t = 0;
    } catch (const HipError& e) {
        return chip(cl, e);
    }
    return RBC_OK;
}

int rbc_closure_gradient(rbc_closure* cl, int n, const float* X, const float* Y, float* grad, double* loss) {
    if (!cl || n < 1 || !X || !Y || !grad) return RBC_ERR_ARG;
    rbc_ctx* c = cl->c;
    float *dx = nullptr, *dy = nullptr;
    double* lb = nullptr;
    int rc = RBC_OK;
    try {
        hipSetDevice(c->p.device);
        HIPCHK(hipMalloc(&dx, (size_t)n * NIN * sizeof(float)));
        HIPCHK(hipMalloc(&dy, (size_t)n * NOUT * sizeof(float)));
        HIPCHK(hipMalloc(&lb, sizeof(double)));
        HIPCHK(hipMemcpy(dx, X, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dy, Y, (size_t)n * NOUT * sizeof(float), hipMemcpyHostToDevice));
        const int per = (n + kTrainWaves - 1) / kTrainWaves;
        const int nwv = (n + per - 1) / per;
        hipLaunchKernelGGL(k_mlp_train, dim3(nwv), dim3(64), 0, c->stream, n, n, per, 0, (uint64_t)0, (uint64_t)0,
                           (const uint16_t*)cl->wbf, cl->nm, (const float*)dx, (const float*)dy, cl->gpart, cl->lpart);
        hipLaunchKernelGGL(k_grad_reduce, dim3((NW + 255) / 256), dim3(256), 0, c->stream, nwv,
                           (const float*)cl->gpart, (const double*)cl->lpart, cl->grad, lb, 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(grad, cl->grad, NW * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        double l = 0;
        HIPCHK(hipMemcpyAsync(&l, lb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (loss) *loss = l;
    } catch (const HipError& e) {
        rc = chip(cl, e);
    }
    if (dx) hipFree(dx);
    if (dy) hipFree(dy);
    if (lb) hipFree(lb);
    return rc;
}

int rbc_closure_train(rbc_closure* cl, int steps, int batch, double lr, unsigned long long seed, int first, int n,
                      double* loss) {
    if (!cl || steps < 0 || batch < 1) return RBC_ERR_ARG;
    if (n <= 0) n = cl->count - first;
    if (first < 0 || first + n > cl->count || n < 1) return cfail(cl, RBC_ERR_ARG, "training range outside the dataset");
    rbc_ctx* c = cl->c;
    try {
        hipSetDevice(c->p.device);
        if (cl->lossbuf_n < steps) {
            cl->lossbuf = (double*)cl->al((size_t)std::max(1, steps) * sizeof(double));
            cl->lossbuf_n = steps;
        }
        const int per = (batch + kTrainWaves - 1) / kTrainWaves;
        const int nwv = (batch + per - 1) / per;
        const float* X = cl->DX + (size_t)first * NIN;
        const float* Y = cl->DY + (size_t)first * NOUT;
        for (int s = 0; s < steps; ++s) {
            hipLaunchKernelGGL(k_mlp_train, dim3(nwv), dim3(64), 0, c->stream, n, batch, per, 1, (uint64_t)seed,
                               (uint64_t)(cl->adam_t + 1), (const uint16_t*)cl->wbf, cl->nm, X, Y, cl->gpart,
                               cl->lpart);
            hipLaunchKernelGGL(k_grad_reduce, dim3((NW + 255) / 256), dim3(256), 0, c->stream, nwv,
                               (const float*)cl->gpart, (const double*)cl->lpart, cl->grad, cl->lossbuf, s);
            cl->adam_t++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - std::pow(b1, (float)cl->adam_t), c2 = 1.f - std::pow(b2, (float)cl->adam_t);
            hipLaunchKernelGGL(k_adam, dim3((NW + 255) / 256), dim3(256), 0, c->stream, (const float*)cl->grad, cl->wm,
                               cl->m1, cl->m2, cl->wbf, (float)lr, b1, b2, 1e-8f, c1, c2);
        }
        HIPCHK(hipGetLastError());
        if (loss && steps > 0)
            HIPCHK(hipMemcpyAsync(loss, cl->lossbuf, (size_t)steps * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        return chip(cl, e);
    }
    return RBC_OK;
}

int rbc_closure_predict(rbc_closure* cl, int n, const float* X, float* Y) {
    if (!cl || n < 1 || !X || !Y) return RBC_ERR_ARG;
    rbc_ctx* c = cl->c;
    float *dx = nullptr, *dy = nullptr;
    int rc = RBC_OK;
    try {
        hipSetDevice(c->p.device);
        HIPCHK(hipMalloc(&dx, (size_t)n * NIN * sizeof(float)));
        HIPCHK(hipMalloc(&dy, (size_t)n * NOUT * sizeof(float)));
        HIPCHK(hipMemcpy(dx, X, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_mlp_infer, dim3(infer_blocks(n)), dim3(256), 0, c->stream, n, (const uint16_t*)cl->wbf,
                           cl->nm, (const float*)dx, dy);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(Y, dy, (size_t)n * NOUT * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        rc = chip(cl, e);
    }
    if (dx) hipFree(dx);
    if (dy) hipFree(dy);
    return rc;
}

// mean squared error (normalised units, as trained) of the model and of the zero prediction
// over dataset points [first, first + n)
int rbc_closure_eval(rbc_closure* cl, int first, int n, double* mse_model, double* mse_zero) {
    if (!cl || first < 0 || n < 1 || first + n > cl->count) return RBC_ERR_ARG;
    std::vector<float> X((size_t)n * NIN), Y((size_t)n * NOUT), P((size_t)n * NOUT);
    if (int rc = rbc_closure_dataset(cl, first, n, X.data(), Y.data())) return rc;
    if (int rc = rbc_closure_predict(cl, n, X.data(), P.data())) return rc;
    double a = 0, b = 0;
    for (size_t i = 0; i < (size_t)n; ++i)
        for (int o = 0; o < NOUT; ++o) {
            const double s = cl->nm.out_scale[o];
            const double e = (P[i * NOUT + o] - Y[i * NOUT + o]) / s, t = Y[i * NOUT + o] / s;
            a += e * e;
            b += t * t;
        }
    if (mse_model) *mse_model = a / ((double)n * NOUT);
    if (mse_zero) *mse_zero = b / ((double)n * NOUT);
    return RBC_OK;
}

int rbc_closure_enable(rbc_closure* cl, int on, double max_ratio) {
    if (!cl) return RBC_ERR_ARG;
    rbc_ctx* c = cl->c;
    if (on) {
        if (c->post_nonlinear && c->post_user != cl)
            return cfail(cl, RBC_ERR_STATE, "another closure is attached to this context");
        c->post_nonlinear = apply_closure;
        c->post_user = cl;
        cl->enabled = true;
        cl->fallback = false;
        cl->max_ratio = max_ratio > 0 ? max_ratio : 1.0;
        cl->phase = 0;
    } else {
        if (c->post_user == cl) {
            c->post_nonlinear = nullptr;
            c->post_user = nullptr;
        }
        cl->enabled = false;
    }
    return RBC_OK;
}

int rbc_closure_damping(rbc_closure* cl, double gain, double kfrac) {
    if (!cl) return RBC_ERR_ARG;
    if (!std::isfinite(gain) || gain < 0 || !(kfrac > 0)) return cfail(cl, RBC_ERR_ARG, "gain >= 0 and kfrac > 0 required");
    const rbc::Factors& f = cl->c->fac;
    double k2max = 0;
    for (double k2 : f.k2) k2max = std::max(k2max, k2);
    cl->gain = gain;
    cl->k2cut = kfrac >= 1 ? HUGE_VAL : kfrac * kfrac * k2max;
    return RBC_OK;
}

int rbc_closure_interval(rbc_closure* cl, int every) {
    if (!cl) return RBC_ERR_ARG;
    if (every < 1) return cfail(cl, RBC_ERR_ARG, "every must be >= 1");
    cl->every = every;
    cl->phase = 0;
    return RBC_OK;
}

// steps the coarse context with the closure in chunks of check_every steps; after each
// chunk the size of the last correction relative to N is checked, and when it is not
// finite or exceeds max_ratio the chunk is undone and redone with the plain method, which
// then continues for the rest of the call (fallback).
int rbc_closure_step(rbc_closure* cl, int nsteps, int check_every) {
    if (!cl || nsteps < 0) return RBC_ERR_ARG;
    if (check_every <= 0) check_every = nsteps > 0 ? nsteps : 1;
    rbc_ctx* c = cl->c;
    try {
        hipSetDevice(c->p.device);
        for (int done = 0; done < nsteps;) {
            const int k = std::min(check_every, nsteps - done);
            if (!cl->enabled) {
                run_steps(c, k);
                done += k;
                continue;
            }
            const bool hist = c->have_hist;
            take_snapshot(cl);
            run_steps(c, k);
            const double r = read_ratio(cl);
            cl->last_ratio = r;
            if (!std::isfinite(r) || r > cl->max_ratio) {
                restore_snapshot(cl);
                c->have_hist = hist;
                rbc_closure_enable(cl, 0, 0);
                cl->fallback = true;
                run_steps(c, k);
            }
            done += k;
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        return chip(cl, e);
    }
    return RBC_OK;
}

int rbc_closure_status(const rbc_closure* cl, int* active, int* fallback, double* last_ratio) {
    if (!cl) return RBC_ERR_ARG;
    if (active) *active = cl->enabled ? 1 : 0;
    if (fallback) *fallback = cl->fallback ? 1 : 0;
    if (last_ratio) *last_ratio = cl->last_ratio;
    return RBC_OK;
}

// features and residual targets of the coarse context's current state against a fine
// context, at all closure points, without touching the dataset (for checks)
int rbc_closure_points(rbc_closure* cl, rbc_ctx* fine, float* X, float* Y) {
    if (!cl || !fine || !X) return RBC_ERR_ARG;
    // sample into the free dataset slot after the last point, read it back, forget it
    const int save = cl->count;
    if (cl->count + cl->npts > cl->cap) return cfail(cl, RBC_ERR_ARG, "closure dataset is full");
    int rc = rbc_closure_sample(cl, fine, nullptr);
    if (rc) return rc;
    rc = rbc_closure_dataset(cl, save, cl->npts, X, Y);
    cl->count = save;
    return rc;
}

}  // extern "C"
