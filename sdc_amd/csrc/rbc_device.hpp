// rbc_device.hpp -- device-side helpers shared by the kernel translation units of
// libsdc_rbc.so: complex arithmetic, XCD-aware block ids, wavefront LDS ordering, the
// LDS Stockham FFT with the generic y-FFTs, the natural-layout group FFT with the group y-FFTs, and
// the lane reductions.
#pragma once

#include <hip/hip_runtime.h>

#include "rbc_internal.hpp"

namespace rbc {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> mk(T a, T b) {
    cpx<T> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T>
__device__ __forceinline__ cpx<T> conjg(cpx<T> a) { return mk(a.x, -a.y); }
// i * s * a
template <typename T>
__device__ __forceinline__ cpx<T> imul(cpx<T> a, T s) { return mk(-s * a.y, s * a.x); }

// Kept columns (kx = 0 .. kept_cols - 1) and kept rows of the dealias rule of make_geometry,
// as compile-time functions of the grid size (same double expressions), so that the transform
// kernels of a fixed size resolve the zero band and the mirrored half at compile time.
constexpr int kept_cols(int M) {
    const int mox = M / 2 + 1;
    int kx = 0;
    for (int col = 0; col < mox; ++col)
        if (!(mox - col <= (1.0 / 3.0) * mox)) kx = col + 1;
    return kx;
}
constexpr int kept_rows(int N) {
    int n = 0;
    for (int r = 0; r < N; ++r) {
        const double d = 0.5 * N - r;
        if (!((d < 0 ? -d : d) <= (1.0 / 3.0) * 0.5 * N)) ++n;
    }
    return n;
}

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2).  This maps the
// hardware block id to a logical id so that logical neighbours (which share cache lines) run
// on the same XCD: XCD x takes the contiguous logical range [x*per, (x+1)*per).  Grids using
// it are launched with 8*per blocks; blocks with logical id >= the real count return at once.
__device__ __forceinline__ int xcd_logical_block(int b, int per) { return (b & 7) * per + (b >> 3); }

// Orders LDS traffic between the lanes of one wavefront (no block-wide barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Block-wide barrier that orders LDS traffic only: unlike __syncthreads it does not wait for
// the wave's outstanding global loads and stores, so prefetches stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One LDS pad slot every 8 complex values spreads the strided accesses of the FFT passes over
// the banks.
__device__ __forceinline__ int pad8(int i) { return i + (i >> 3); }

// 8-point DFT in registers (decimation in frequency, outputs reordered to natural order).
template <typename T, bool INV>
__device__ __forceinline__ void dft8(cpx<T> v[8]) {
    const T h = T(0.70710678118654752440);
    // stage 1 (span 4) with twiddles w8^i, w8 = exp(-+ i pi / 4)
    cpx<T> a[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        a[i] = v[i] + v[i + 4];
        a[i + 4] = v[i] - v[i + 4];
    }
    // a5 *= w8, a6 *= w8^2, a7 *= w8^3
    if (!INV) {
        a[5] = mk(h * (a[5].x + a[5].y), h * (a[5].y - a[5].x));
        a[6] = mk(a[6].y, -a[6].x);
        a[7] = mk(h * (a[7].y - a[7].x), -h * (a[7].x + a[7].y));
    } else {
        a[5] = mk(h * (a[5].x - a[5].y), h * (a[5].y + a[5].x));
        a[6] = mk(-a[6].y, a[6].x);
        a[7] = mk(-h * (a[7].x + a[7].y), h * (a[7].x - a[7].y));
    }
    // stage 2 (span 2) with w4 = -+ i on the second element of each pair
    cpx<T> b[8];
#pragma unroll
    for (int hh = 0; hh < 8; hh += 4) {
        b[hh + 0] = a[hh + 0] + a[hh + 2];
        b[hh + 2] = a[hh + 0] - a[hh + 2];
        b[hh + 1] = a[hh + 1] + a[hh + 3];
        const cpx<T> d = a[hh + 1] - a[hh + 3];
        b[hh + 3] = INV ? mk(-d.y, d.x) : mk(d.y, -d.x);
    }
    // stage 3 (span 1); outputs in bit-reversed order
    const cpx<T> c0 = b[0] + b[1], c1 = b[0] - b[1], c2 = b[2] + b[3], c3 = b[2] - b[3];
    const cpx<T> c4 = b[4] + b[5], c5 = b[4] - b[5], c6 = b[6] + b[7], c7 = b[6] - b[7];
    v[0] = c0;
    v[1] = c4;
    v[2] = c2;
    v[3] = c6;
    v[4] = c1;
    v[5] = c5;
    v[6] = c3;
    v[7] = c7;
}

// max over aligned groups of W lanes (W = 8, 16, 32 or 64), the result in every lane of the
// group: DPP lane permutations inside 16-lane rows (quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror, row_mirror; after the earlier steps each is a max with the partner half),
// then a shuffle across rows (W = 32) or lane reads of the four rows (W = 64).  All lanes active.
// The values are non-negative (row maxima of |x|), whose IEEE bit patterns order as the values:
// the lanes compare them as unsigned integers, which needs no NaN canonicalisation and lets the
// compiler fold each DPP move into the max.
__device__ __forceinline__ unsigned dpp_max_step(unsigned x, int ctrl_id) {
    int y;
    switch (ctrl_id) {
        case 0: y = __builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, true); break;
        case 1: y = __builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, true); break;
        case 2: y = __builtin_amdgcn_mov_dpp((int)x, 0x141, 0xF, 0xF, true); break;
        default: y = __builtin_amdgcn_mov_dpp((int)x, 0x140, 0xF, 0xF, true); break;
    }
    return max(x, (unsigned)y);
}
template <int W>
__device__ __forceinline__ float group_max(float v) {
    unsigned x = __float_as_uint(v);
    x = dpp_max_step(x, 0);
    x = dpp_max_step(x, 1);
    x = dpp_max_step(x, 2);
    if (W >= 16) x = dpp_max_step(x, 3);
    if (W == 32) x = max(x, (unsigned)__shfl_xor((int)x, 16));
    if (W == 64) {
        const int xi = (int)x;
        x = max(max((unsigned)__builtin_amdgcn_readlane(xi, 0), (unsigned)__builtin_amdgcn_readlane(xi, 16)),
                max((unsigned)__builtin_amdgcn_readlane(xi, 32), (unsigned)__builtin_amdgcn_readlane(xi, 48)));
    }
    return __uint_as_float(x);
}
__device__ __forceinline__ float wave_max(float v) { return group_max<64>(v); }
// power of two that brings the row maximum mb of the second field of a pair to that of the
// first (ma), so that one complex transform carries both with the rounding of each set by its
// own size; scaling by a power of two is exact
__device__ __forceinline__ float pair_scale(float ma, float mb) {
    if (!(mb > 0.f) || !(ma > 0.f)) return 1.f;
    int ea, eb;
    frexpf(ma, &ea);
    frexpf(mb, &eb);
    return ldexpf(1.f, max(-100, min(100, ea - eb)));
}
// 1 / s of a power of two s (exact; two bit-field instructions instead of a division)
__device__ __forceinline__ float inv_pow2(float s) {
    int e;
    frexpf(s, &e);  // s = 0.5 * 2^e
    return ldexpf(1.f, 1 - e);
}

// radix-4 butterfly in place (natural order outputs): X_p = sum_q a_q (-+i)^{pq}
template <bool INV>
__device__ __forceinline__ void dft4(cpx<float>& a0, cpx<float>& a1, cpx<float>& a2, cpx<float>& a3) {
    const cpx<float> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = a1 - a3;
    const cpx<float> it3 = INV ? mk(-t3.y, t3.x) : mk(t3.y, -t3.x);  // -+ i t3
    a0 = t0 + t2;
    a1 = t1 + it3;
    a2 = t0 - t2;
    a3 = t1 - it3;
}

// n-point FFT (n = 2^LOGN, 64 <= n <= 1024) of NF sequences on a group of W = n / 8 lanes in
// natural layout: element i of a sequence sits in register i / W of lane i % W of the group,
// on entry and on exit.  Stockham passes: radix 8 while three or more radix-2 stages remain,
// then one radix-2^(LOGN mod 3) pass.  Lane j runs the butterflies j + W b (b < 8 / R) of a
// radix-R pass, whose inputs x[j + W b + q n / R] are registers b + q 8 / R on the first pass
// and whose outputs y[j + W b + p n / R] are registers b + p 8 / R on the last; passes between
// exchange through the group's LDS scratch (element i at sb[pad8(i) S], so a column of an LDS
// tile of row stride S can serve).  The NF sequences go through the scratch one after the
// other.  tw[t] = exp(-2 pi i t / n); the inverse uses conjugate twiddles; unnormalised.
template <int LOGN, bool INV, int NF, int S = 1>
__device__ __forceinline__ void fft_group(cpx<float> (&v)[NF][8], cpx<float>* sb, const cpx<float>* tw, int j) {
    constexpr int n = 1 << LOGN, W = n / 8;
    constexpr int NP8 = LOGN / 3, REM = LOGN % 3, NPASS = NP8 + (REM ? 1 : 0);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        int logLs = 0;
#pragma unroll
        for (int ps = 0; ps < NPASS; ++ps) {
            const int LR = ps < NP8 ? 3 : REM, R = 1 << LR, NB = 8 / R, Ls = 1 << logLs;
            if (ps > 0) {
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int q = 0; q < R; ++q) v[f][b + q * NB] = sb[pad8(j + W * b + q * (n / R)) * S];
                wave_sync();  // every lane has read the pass inputs before the scratch is rewritten
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int k = (j + W * b) & (Ls - 1);
                if (ps > 0)
#pragma unroll
                    for (int q = 1; q < R; ++q) {
                        cpx<float> w = tw[(q * k) << (LOGN - LR - logLs)];
                        if (INV) w.y = -w.y;
                        v[f][b + q * NB] = cmul(v[f][b + q * NB], w);
                    }
                if (R == 8) {
                    dft8<float, INV>(v[f]);
                } else if (R == 4) {
                    dft4<INV>(v[f][b], v[f][b + 2], v[f][b + 4], v[f][b + 6]);
                } else {
                    const cpx<float> a0 = v[f][b], a1 = v[f][b + 4];
                    v[f][b] = a0 + a1;
                    v[f][b + 4] = a0 - a1;
                }
            }
            if (ps + 1 < NPASS) {
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int jb = j + W * b, k = jb & (Ls - 1);
#pragma unroll
                    for (int p = 0; p < R; ++p) sb[pad8((jb - k) * R + k + p * Ls) * S] = v[f][b + p * NB];
                }
                wave_sync();
            }
            logLs += LR;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Stockham FFT on nb sequences of length n = 2^logn held in LDS (a), with b as the ping-pong
// buffer: radix-4 passes while two or more radix-2 stages remain, then one radix-2 pass if
// logn is odd.  Pass with sub-length Ls and radix R: y[R j - (R-1) k + p Ls] =
// sum_q x[j + q n/R] w^{q k} e^{-+2 pi i p q / R}, k = j mod Ls, w = exp(-+2 pi i / (R Ls)).
// tw[k] = exp(-2 pi i k / n), k < n.  Unnormalised; inverse uses conjugate twiddles.  All
// threads of the block take part; returns the buffer holding the result.  Ends with a barrier.
// ---------------------------------------------------------------------------------------
template <typename T>
__device__ cpx<T>* lds_fft(cpx<T>* a, cpx<T>* b, int logn, int nb, const cpx<T>* __restrict__ tw,
                           bool inverse) {
    const int n = 1 << logn;
    int Ls = 1;
    for (int ls = 0; ls < logn;) {
        if (logn - ls >= 2) {
            const int quarter = n >> 2, total = nb * quarter;
            const int tsh = logn - 2 - ls;  // twiddle stride n / (4 Ls) = 2^tsh
            for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
                const int s = idx >> (logn - 2);
                const int j = idx & (quarter - 1);
                const int k = j & (Ls - 1);
                const cpx<T>* x = a + (s << logn);
                cpx<T>* y = b + (s << logn);
                cpx<T> a0 = x[j], a1 = x[j + quarter], a2 = x[j + 2 * quarter], a3 = x[j + 3 * quarter];
                if (k) {
                    cpx<T> w1 = tw[k << tsh], w2 = tw[(2 * k) << tsh], w3 = tw[(3 * k) << tsh];
                    if (inverse) {
                        w1.y = -w1.y;
                        w2.y = -w2.y;
                        w3.y = -w3.y;
                    }
                    a1 = cmul(a1, w1);
                    a2 = cmul(a2, w2);
                    a3 = cmul(a3, w3);
                }
                const cpx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = a1 - a3;
                const cpx<T> it3 = inverse ? mk(-t3.y, t3.x) : mk(t3.y, -t3.x);  // -+ i t3
                const int base = 4 * j - 3 * k;
                y[base] = t0 + t2;
                y[base + Ls] = t1 + it3;
                y[base + 2 * Ls] = t0 - t2;
                y[base + 3 * Ls] = t1 - it3;
            }
            ls += 2;
            Ls <<= 2;
        } else {
            const int half = n >> 1, total = nb * half;
            const int tsh = logn - 1 - ls;
            for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
                const int s = idx >> (logn - 1);
                const int j = idx & (half - 1);
                const int k = j & (Ls - 1);
                const cpx<T>* x = a + (s << logn);
                cpx<T>* y = b + (s << logn);
                cpx<T> u = x[j], v = x[j + half];
                cpx<T> w = tw[k << tsh];
                if (inverse) w.y = -w.y;
                const cpx<T> t = cmul(v, w);
                const int i0 = 2 * j - k;
                y[i0] = u + t;
                y[i0 + Ls] = u - t;
            }
            ls += 1;
            Ls <<= 1;
        }
        __syncthreads();
        cpx<T>* tmp = a;
        a = b;
        b = tmp;
    }
    return a;
}

// ---------------------------------------------------------------------------------------
// Stage 2: inverse y-FFT.  Block handles TX kx-columns of one (field, level): gathers the
// kept rows into zero-padded columns of length N, transforms, and writes the hybrid rows
// [field][l][y][kx] (kx fastest) through LDS so that each y-row segment is contiguous.
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_fft_y_inv(DevGeom g, int TX, const cpx<T>* __restrict__ lev,
                                                   const int* __restrict__ rows, const cpx<T>* __restrict__ twN,
                                                   cpx<T>* __restrict__ hin) {
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* a = reinterpret_cast<cpx<T>*>(smem_raw);
    const int N = g.N, NYc = g.NYc, P = g.P, nq = g.nq;
    cpx<T>* b = a + (size_t)TX * N;
    const int ftile = blockIdx.x;
    const int l = blockIdx.y, field = blockIdx.z;
    const int kx0 = ftile * TX;
    const int ncol = min(TX, g.KX - kx0);
    for (int e = threadIdx.x; e < TX * N; e += blockDim.x) a[e] = mk<T>(0, 0);
    __syncthreads();
    const cpx<T>* src = lev + ((size_t)field * nq + l) * P;
    for (int e = threadIdx.x; e < ncol * NYc; e += blockDim.x) {
        const int c = e / NYc, kyc = e - c * NYc;
        a[c * N + rows[kyc]] = src[(size_t)(kx0 + c) * NYc + kyc];
    }
    __syncthreads();
    cpx<T>* r = lds_fft(a, b, g.logN, TX, twN, true);
    cpx<T>* dst = hin + ((size_t)field * nq + l) * (size_t)N * g.KXp;
    for (int e = threadIdx.x; e < N * TX; e += blockDim.x) {
        const int y = e / TX, c = e - y * TX;
        if (c < ncol) dst[(size_t)y * g.KXp + kx0 + c] = r[c * N + y];
    }
}

// ---------------------------------------------------------------------------------------
// Stage 4: forward y-FFT of the seven product fields, keeping the kept rows only.
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_fft_y_fwd(DevGeom g, int TX, const cpx<T>* __restrict__ hout,
                                                   const int* __restrict__ rows, const cpx<T>* __restrict__ twN,
                                                   cpx<T>* __restrict__ prod) {
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* a = reinterpret_cast<cpx<T>*>(smem_raw);
    const int N = g.N, NYc = g.NYc, P = g.P, nq = g.nq;
    cpx<T>* b = a + (size_t)TX * N;
    const int kx0 = blockIdx.x * TX;
    const int l = blockIdx.y, q = blockIdx.z;
    const int ncol = min(TX, g.KX - kx0);
    const cpx<T>* src = hout + ((size_t)q * nq + l) * (size_t)N * g.KXp;
    for (int e = threadIdx.x; e < N * TX; e += blockDim.x) {
        const int y = e / TX, c = e - y * TX;
        a[c * N + y] = (c < ncol) ? src[(size_t)y * g.KXp + kx0 + c] : mk<T>(0, 0);
    }
    __syncthreads();
    cpx<T>* r = lds_fft(a, b, g.logN, TX, twN, false);
    cpx<T>* dst = prod + ((size_t)q * nq + l) * P;
    for (int e = threadIdx.x; e < ncol * NYc; e += blockDim.x) {
        const int c = e / NYc, kyc = e - c * NYc;
        dst[(size_t)(kx0 + c) * NYc + kyc] = r[c * N + rows[kyc]];
    }
}

// y-FFTs at N = 64, 128, 256 (FP32): a block of WPB waves handles CB = WPB (64 / W) kx columns
// of one (field, level), one group of W = N / 8 lanes per column (fft_group in natural layout);
// the LDS tile [pad8(y)][CB + 1] is each column's FFT scratch and the transpose to the
// row-major hybrid layout (rows of CB contiguous values).
template <int LOGN, int WPB>
__device__ __forceinline__ void k_fft_y_inv_w_body(int bz, DevGeom g, const cpx<float>* __restrict__ lev,
                                                   const cpx<float>* __restrict__ twN, cpx<float>* __restrict__ hin) {
    constexpr int N = 1 << LOGN, W = N / 8, G = 64 / W, CB = WPB * G, S = CB + 1;
    __shared__ cpx<float> tw[N];
    __shared__ cpx<float> tile[(N + N / 8) * S];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, grp = lane / W, j = lane - grp * W;
    constexpr int NYc = kept_rows(N), a = (NYc - 1) / 2, shift = N - NYc;  // kept rows [0, a] U [N - a, N)
    const int P = g.P, nq = g.nq, KX = g.KX;
    const int kx0 = blockIdx.x * CB, l = blockIdx.y, field = bz;
    const int c = wave * G + grp, kx = kx0 + c;
    for (int e = threadIdx.x; e < N; e += blockDim.x) tw[e] = twN[e];
    __syncthreads();
    const cpx<float>* src = lev + ((size_t)field * nq + l) * P + (size_t)kx * NYc;
    cpx<float> v[1][8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int n = j + W * r;
        v[0][r] = mk(0.f, 0.f);
        if (kx < KX) {
            if (n <= a) v[0][r] = src[n];
            else if (n >= N - a) v[0][r] = src[n - shift];
        }
    }
    fft_group<LOGN, true, 1, S>(v, tile + c, tw, j);
#pragma unroll
    for (int r = 0; r < 8; ++r) tile[pad8(j + W * r) * S + c] = v[0][r];
    __syncthreads();
    const int ncol = min(CB, KX - kx0);
    cpx<float>* dst = hin + ((size_t)field * nq + l) * (size_t)N * g.KXp + kx0;
    for (int e = threadIdx.x; e < N * CB; e += blockDim.x) {
        const int yy = e / CB, cc = e - yy * CB;
        if (cc < ncol) dst[(size_t)yy * g.KXp + cc] = tile[pad8(yy) * S + cc];
    }
}

template <int LOGN, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_fft_y_inv_w(DevGeom g, const cpx<float>* __restrict__ lev,
                                                     const cpx<float>* __restrict__ twN, cpx<float>* __restrict__ hin) {
    k_fft_y_inv_w_body<LOGN, WPB>(blockIdx.z, g, lev, twN, hin);
}
// one member's arguments of k_fft_y_inv_w_batch, which runs the k_fft_y_inv_w body for member
// blockIdx.z / zdim of a batch (the rest of blockIdx.z is the body's own z index)
struct k_fft_y_inv_w_args {
    DevGeom g;
    const cpx<float>* lev;
    const cpx<float>* twN;
    cpx<float>* hin;
};
template <int LOGN, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_fft_y_inv_w_batch(const k_fft_y_inv_w_args* __restrict__ tab, int zdim) {
    const int mb = blockIdx.z / zdim;
    const k_fft_y_inv_w_args& a = tab[mb];
    k_fft_y_inv_w_body<LOGN, WPB>(blockIdx.z - mb * zdim, a.g, a.lev, a.twN, a.hin);
}


template <int LOGN, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_fft_y_fwd_w(DevGeom g, const cpx<float>* __restrict__ hout,
                                                     const cpx<float>* __restrict__ twN, cpx<float>* __restrict__ prod) {
    constexpr int N = 1 << LOGN, W = N / 8, G = 64 / W, CB = WPB * G, S = CB + 1;
    __shared__ cpx<float> tw[N];
    __shared__ cpx<float> tile[(N + N / 8) * S];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, grp = lane / W, j = lane - grp * W;
    constexpr int NYc = kept_rows(N), a = (NYc - 1) / 2, shift = N - NYc;
    const int P = g.P, nq = g.nq, KX = g.KX;
    const int kx0 = blockIdx.x * CB, l = blockIdx.y, q = blockIdx.z;
    const int c = wave * G + grp, kx = kx0 + c;
    const int ncol = min(CB, KX - kx0);
    for (int e = threadIdx.x; e < N; e += blockDim.x) tw[e] = twN[e];
    const cpx<float>* src = hout + ((size_t)q * nq + l) * (size_t)N * g.KXp + kx0;
    for (int e = threadIdx.x; e < N * CB; e += blockDim.x) {
        const int yy = e / CB, cc = e - yy * CB;
        tile[pad8(yy) * S + cc] = (cc < ncol) ? src[(size_t)yy * g.KXp + cc] : mk(0.f, 0.f);
    }
    __syncthreads();
    cpx<float> v[1][8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[0][r] = tile[pad8(j + W * r) * S + c];
    wave_sync();  // the column is read before it becomes the scratch
    fft_group<LOGN, false, 1, S>(v, tile + c, tw, j);
    if (kx < KX) {
        cpx<float>* dst = prod + ((size_t)q * nq + l) * P + (size_t)kx * NYc;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int n = j + W * r;
            if (n <= a) dst[n] = v[0][r];
            else if (n >= N - a) dst[n - shift] = v[0][r];
        }
    }
}

// Launchers of the group y-FFTs for nf fields of g.nq levels each ([field][l][p] spectra,
// [field][l][y][kx] hybrid rows); cb = kx columns per block (WPB = cb / (512 / N) waves of 2, 4
// or 8; 0: 4 waves).  false (nothing launched) unless N is 64, 128 or 256.
template <int LOGN, int WPB>
inline void launch_y_group_t(bool inv, const DevGeom& g, int nf, const cpx<float>* in, const cpx<float>* twN,
                             cpx<float>* out, hipStream_t s) {
    constexpr int CB = WPB * 512 / (1 << LOGN);
    const dim3 gr((g.KX + CB - 1) / CB, g.nq, nf);
    if (inv) hipLaunchKernelGGL((k_fft_y_inv_w<LOGN, WPB>), gr, dim3(64 * WPB), 0, s, g, in, twN, out);
    else hipLaunchKernelGGL((k_fft_y_fwd_w<LOGN, WPB>), gr, dim3(64 * WPB), 0, s, g, in, twN, out);
}
template <int LOGN>
inline void launch_y_group_w(bool inv, int wpb, const DevGeom& g, int nf, const cpx<float>* in,
                             const cpx<float>* twN, cpx<float>* out, hipStream_t s) {
    if (wpb == 2) launch_y_group_t<LOGN, 2>(inv, g, nf, in, twN, out, s);
    else if (wpb == 8) launch_y_group_t<LOGN, 8>(inv, g, nf, in, twN, out, s);
    else launch_y_group_t<LOGN, 4>(inv, g, nf, in, twN, out, s);
}
inline bool launch_y_group(bool inv, const DevGeom& g, int nf, const cpx<float>* in, const cpx<float>* twN,
                           cpx<float>* out, hipStream_t s, int cb = 0) {
    if (g.N < 64 || g.N > 256 || g.NYc != kept_rows(g.N)) return false;
    const int wpb = cb > 0 ? cb / (512 / g.N) : 4;
    if (g.N == 64) launch_y_group_w<6>(inv, wpb, g, nf, in, twN, out, s);
    else if (g.N == 128) launch_y_group_w<7>(inv, wpb, g, nf, in, twN, out, s);
    else launch_y_group_w<8>(inv, wpb, g, nf, in, twN, out, s);
    return true;
}
inline bool launch_y_inv_group(const DevGeom& g, int nf, const cpx<float>* lev, const cpx<float>* twN,
                               cpx<float>* hin, hipStream_t s, int cb = 0) {
    return launch_y_group(true, g, nf, lev, twN, hin, s, cb);
}
inline bool launch_y_fwd_group(const DevGeom& g, int nf, const cpx<float>* hout, const cpx<float>* twN,
                               cpx<float>* prod, hipStream_t s, int cb = 0) {
    return launch_y_group(false, g, nf, hout, twN, prod, s, cb);
}

}  // namespace rbc
