// rbc_device.hpp -- device-side helpers shared by the kernel translation units of
// libsdc_rbc.so: complex arithmetic, XCD-aware block ids, wavefront LDS ordering, the
// radix-2 LDS FFT and the generic inverse y-FFT of level fields.
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

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2).  This maps the
// hardware block id to a logical id so that logical neighbours (which share cache lines) run
// on the same XCD: XCD x takes the contiguous logical range [x*per, (x+1)*per).  Grids using
// it are launched with 8*per blocks; blocks with logical id >= the real count return at once.
inline __device__ __forceinline__ int xcd_logical_block(int b, int per) { return (b & 7) * per + (b >> 3); }

// Orders LDS traffic between the lanes of one wavefront (no block-wide barrier).
inline __device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

}  // namespace rbc
