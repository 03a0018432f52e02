// rbc_kernels.hip -- gfx950 kernels of the Rayleigh-Benard time step.
//
// One IMEX step is a pipeline over three representations of the fields:
//   mode space   [p][m]        compact kept horizontal modes p = kx*NYc + kyc, vertical mode m
//   level space  [q][l][p]     the same horizontal modes at Gauss-Legendre level l
//   hybrid rows  [q][l][y][kx] kx spectral, y physical (one x-row per (l, y))
// Stages: eval (mode -> level), inverse y-FFT, per-row x-transforms with the pointwise
// products, forward y-FFT, projection (level -> mode), group update (inverse B, AB2,
// implicit Euler), make_real.  Everything is templated on the real type T (float/double).
#include <hip/hip_runtime.h>

#include <cstdlib>

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
__device__ __forceinline__ int xcd_logical_block(int b, int per) { return (b & 7) * per + (b >> 3); }

// Orders LDS traffic between the lanes of one wavefront (no block-wide barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------------------
// Radix-2 Stockham FFT on nb sequences of length n = 2^logn held in LDS (a), with b as the
// ping-pong buffer.  tw[k] = exp(-2 pi i k / n) (the first n/2 entries are used).  Unnormalised; inverse uses
// conjugate twiddles.  All threads of the block take part; returns the buffer holding the
// result.  Ends with a barrier.
// ---------------------------------------------------------------------------------------
template <typename T>
__device__ cpx<T>* lds_fft(cpx<T>* a, cpx<T>* b, int logn, int nb, const cpx<T>* __restrict__ tw,
                           bool inverse) {
    const int n = 1 << logn, half = n >> 1;
    const int total = nb * half;
    for (int ls = 0; ls < logn; ++ls) {
        const int Ls = 1 << ls;
        const int tws = logn - 1 - ls;  // twiddle stride = n / (2 Ls)
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int s = idx >> (logn - 1);
            const int j = idx & (half - 1);
            const int k = j & (Ls - 1);
            const cpx<T>* x = a + (s << logn);
            cpx<T>* y = b + (s << logn);
            cpx<T> u = x[j], v = x[j + half];
            cpx<T> w = tw[k << tws];
            if (inverse) w.y = -w.y;
            cpx<T> t = cmul(v, w);
            const int i0 = 2 * j - k;
            y[i0] = u + t;
            y[i0 + Ls] = u - t;
        }
        __syncthreads();
        cpx<T>* tmp = a;
        a = b;
        b = tmp;
    }
    return a;
}

// ---------------------------------------------------------------------------------------
// Stage 1: vertical evaluation.  Block = 64 modes x 4 waves; the mode data of the block is
// staged in LDS transposed to [m][64]; each wave walks levels, lane = mode.  Level fields:
//   0 u = i kx F' + i ky G_s (+ sum S F at k = 0)     1 v = i ky F' - i kx G_s (+ sum S G)
//   2 w = k^2 F_c      3 theta_x = i kx T_s      4 theta_y = i ky T_s      5 theta_z = T_s'
// with F' = sum C'_m f_m, F_c = sum C_m f_m, G_s = sum S_m g_m, T_s = sum S_m theta_m,
// T_s' = sum S'_m theta_m (calculate_velocity_operator.cu:426-455 for the velocity form).
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_eval(DevGeom g, int LG, int nblk, int per, const cpx<T>* __restrict__ Xt,
                                              const cpx<T>* __restrict__ Xf, const cpx<T>* __restrict__ Xg,
                                              const cpx<T>* __restrict__ XF, const cpx<T>* __restrict__ XG,
                                              const T* __restrict__ evtab, const T* __restrict__ kxa,
                                              const T* __restrict__ kya, cpx<T>* __restrict__ lev) {
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* sm = reinterpret_cast<cpx<T>*>(smem_raw);  // 3 * L * 64
    // logical block = (64-mode block, level group); the LG level groups of one mode block are
    // neighbours in logical order, so they run on one XCD and share its L2 copy of the modes
    const int lb = xcd_logical_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int L = g.L, P = g.P, nq = g.nq;
    const int p0 = (lb / LG) * 64;
    const int lgq = (nq + LG - 1) / LG, lbeg = (lb % LG) * lgq, lend = min(nq, lbeg + lgq);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // stage: element e of the block's [64][L] slab -> sm[field][m][mode]
    for (int e = threadIdx.x; e < 64 * L; e += blockDim.x) {
        const int mode = e / L, m = e - mode * L;
        const int p = p0 + mode;
        cpx<T> zt = mk<T>(0, 0), zf = zt, zg = zt;
        if (p < P) {
            const size_t gi = (size_t)p * L + m;
            zt = Xt[gi];
            zf = Xf[gi];
            zg = Xg[gi];
        }
        sm[(0 * L + m) * 64 + mode] = zt;
        sm[(1 * L + m) * 64 + mode] = zf;
        sm[(2 * L + m) * 64 + mode] = zg;
    }
    __syncthreads();
    const int p = p0 + lane;
    if (p >= P) return;
    const T kx = kxa[p], ky = kya[p];
    const T k2 = kx * kx + ky * ky;
    const size_t plane = (size_t)P;
    constexpr int LB = 4;  // levels per pass over the mode data
    for (int l0 = lbeg + wave * LB; l0 < lend; l0 += 4 * LB) {
        cpx<T> Fp[LB], Fc[LB], Gs[LB], Ts[LB], Tz[LB];
#pragma unroll
        for (int b = 0; b < LB; ++b) Fp[b] = Fc[b] = Gs[b] = Ts[b] = Tz[b] = mk<T>(0, 0);
        for (int m = 0; m < L; ++m) {
            const cpx<T> th = sm[(0 * L + m) * 64 + lane];
            const cpx<T> f = sm[(1 * L + m) * 64 + lane];
            const cpx<T> gg = sm[(2 * L + m) * 64 + lane];
#pragma unroll
            for (int b = 0; b < LB; ++b) {
                const int l = min(l0 + b, lend - 1);
                const T* tb = evtab + (size_t)l * 4 * L;  // C', C, S, S'
                const T cp = tb[m], cv = tb[L + m], sv = tb[2 * L + m], sp = tb[3 * L + m];
                Fp[b] = Fp[b] + cscale(f, cp);
                Fc[b] = Fc[b] + cscale(f, cv);
                Gs[b] = Gs[b] + cscale(gg, sv);
                Ts[b] = Ts[b] + cscale(th, sv);
                Tz[b] = Tz[b] + cscale(th, sp);
            }
        }
#pragma unroll
        for (int b = 0; b < LB; ++b) {
            const int l = l0 + b;
            if (l >= lend) break;
            const T* tb = evtab + (size_t)l * 4 * L;
            cpx<T> u = imul(Fp[b], kx) + imul(Gs[b], ky);
            cpx<T> v = imul(Fp[b], ky) - imul(Gs[b], kx);
            if (p == 0) {  // mean flow enters at k = 0 (add_dc_part, nonlinear...cu:1596-1608)
                cpx<T> sF = mk<T>(0, 0), sG = sF;
                for (int m = 0; m < L; ++m) {
                    sF = sF + cscale(XF[m], tb[2 * L + m]);
                    sG = sG + cscale(XG[m], tb[2 * L + m]);
                }
                u = u + sF;
                v = v + sG;
            }
            cpx<T>* o = lev + (size_t)l * plane + p;
            const size_t fs = (size_t)nq * plane;
            o[0 * fs] = u;
            o[1 * fs] = v;
            o[2 * fs] = cscale(Fc[b], k2);
            o[3 * fs] = imul(Ts[b], kx);
            o[4 * fs] = imul(Ts[b], ky);
            o[5 * fs] = Tz[b];
        }
    }
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
// Stage 3: one (level, y) row per block: three complex inverse FFTs of length M carry the
// six real fields in pairs (u+iv, w+i theta_x, theta_y+i theta_z); the seven products
// uu uv vv uw vw ww and u.grad(theta), scaled by 1/(M N) as the reference scales before its
// R2C (nonlinear...cu:252), go back through four forward FFTs (pairs) and are unpacked to
// kx = 0..KX-1.  The imaginary parts of the kx = 0 inputs are dropped, as a C2R does.
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_rows(DevGeom g, const cpx<T>* __restrict__ hin,
                                              const cpx<T>* __restrict__ twM, cpx<T>* __restrict__ hout) {
    extern __shared__ unsigned char smem_raw[];
    const int M = g.M, N = g.N, KX = g.KX, KXp = g.KXp, nq = g.nq;
    cpx<T>* a = reinterpret_cast<cpx<T>*>(smem_raw);  // 4 * M
    cpx<T>* b = a + 4 * M;                            // 4 * M
    const int y = blockIdx.x, l = blockIdx.y;
    const size_t fstride = (size_t)nq * N * KXp;
    const size_t roff = ((size_t)l * N + y) * KXp;
    // build the three Hermitian-extended pair spectra Z = X1 + i X2
    for (int e = threadIdx.x; e < 3 * M; e += blockDim.x) a[e] = mk<T>(0, 0);
    __syncthreads();
    for (int e = threadIdx.x; e < 3 * KX; e += blockDim.x) {
        const int pr = e / KX, k = e - pr * KX;
        cpx<T> x1 = hin[(size_t)(2 * pr) * fstride + roff + k];
        cpx<T> x2 = hin[(size_t)(2 * pr + 1) * fstride + roff + k];
        if (k == 0) {
            x1.y = 0;
            x2.y = 0;
        }
        cpx<T>* z = a + pr * M;
        z[k] = mk(x1.x - x2.y, x1.y + x2.x);  // X1 + i X2
        if (k > 0) z[M - k] = mk(x1.x + x2.y, -x1.y + x2.x);  // conj(X1) + i conj(X2)
    }
    __syncthreads();
    cpx<T>* r = lds_fft(a, b, g.logM, 3, twM, true);
    cpx<T>* o = (r == a) ? b : a;
    const T s = T(1) / (T(M) * T(N));
    for (int x = threadIdx.x; x < M; x += blockDim.x) {
        const cpx<T> z0 = r[x], z1 = r[M + x], z2 = r[2 * M + x];
        const T u = z0.x, v = z0.y, w = z1.x, tx = z1.y, ty = z2.x, tz = z2.y;
        o[x] = mk(s * u * u, s * u * v);
        o[M + x] = mk(s * v * v, s * u * w);
        o[2 * M + x] = mk(s * v * w, s * w * w);
        o[3 * M + x] = mk(s * (u * tx + v * ty + w * tz), T(0));
    }
    __syncthreads();
    cpx<T>* spare = (o == a) ? b : a;
    cpx<T>* R = lds_fft(o, spare, g.logM, 4, twM, false);
    // unpack pair spectra: X1_k = (Z_k + conj Z_{M-k}) / 2, X2_k = -i (Z_k - conj Z_{M-k}) / 2
    for (int e = threadIdx.x; e < 4 * KX; e += blockDim.x) {
        const int pr = e / KX, k = e - pr * KX;
        const cpx<T>* Z = R + pr * M;
        const cpx<T> zk = Z[k], zm = Z[(M - k) & (M - 1)];
        const cpx<T> x1 = mk(T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y));
        const cpx<T> x2 = mk(T(0.5) * (zk.y + zm.y), T(-0.5) * (zk.x - zm.x));
        hout[(size_t)(2 * pr) * fstride + roff + k] = x1;
        if (pr < 3) hout[(size_t)(2 * pr + 1) * fstride + roff + k] = x2;
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

// ---------------------------------------------------------------------------------------
// Stage 5: projection.  Block = 64 modes; wave w accumulates vertical modes
// m = w, w+4, ... in chunks.  Per level the seven product spectra give
//   A  = kx^2 Puu + 2 kx ky Puv + ky^2 Pvv - k^2 Pww         (tested with C')
//   Bf = i kx Puw + i ky Pvw                                  (with C'' + k^2 C)
//   Ag = -kx ky (Puu - Pvv) + (kx^2 - ky^2) Puv               (with S)
//   Bg = i kx Pvw - i ky Puw                                  (with S')
//   At = u.grad(theta)                                        (with S)
// which is the Galerkin sum of nonlinear...cu:479-1067 with each triple integral written
// as a quadrature over the levels.  Outputs are scaled by -1 (theta) and -1/Pr (:1366-1370).
// ---------------------------------------------------------------------------------------
// Variant for L <= 4 * MCH: per round of 4 levels, wave w forms the five level combinations
// of level l0 + w for the block's 64 modes in LDS once; then every wave accumulates its own
// MCH vertical modes from LDS (each product is read from memory once).
template <typename T, int MCH>
__global__ void __launch_bounds__(256) k_project_lds(DevGeom g, const cpx<T>* __restrict__ prod,
                                                     const T* __restrict__ prtab, const T* __restrict__ kxa,
                                                     const T* __restrict__ kya, T Pr, cpx<T>* __restrict__ Nt,
                                                     cpx<T>* __restrict__ Nf, cpx<T>* __restrict__ Ng,
                                                     cpx<T>* __restrict__ NF, cpx<T>* __restrict__ NG) {
    constexpr int LV = 4;
    __shared__ cpx<T> sD[LV][5][64];
    const int L = g.L, P = g.P, nq = g.nq;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    const bool valid = p < P;
    const int pc = valid ? p : 0;
    const T kx = kxa[pc], ky = kya[pc];
    const T k2 = kx * kx + ky * ky;
    const size_t fs = (size_t)nq * P;
    const T mt = T(-1), mp = T(-1) / Pr;
    const int mb = wave * MCH;
    cpx<T> af[MCH], ag[MCH], at[MCH];
#pragma unroll
    for (int i = 0; i < MCH; ++i) af[i] = ag[i] = at[i] = mk<T>(0, 0);
    for (int l0 = 0; l0 < nq; l0 += LV) {
        const int ls = l0 + wave;
        if (ls < nq) {
            const cpx<T>* src = prod + (size_t)ls * P + pc;
            const cpx<T> Puu = src[0], Puv = src[fs], Pvv = src[2 * fs], Puw = src[3 * fs];
            const cpx<T> Pvw = src[4 * fs], Pww = src[5 * fs], Pad = src[6 * fs];
            sD[wave][0][lane] = cscale(Puu, kx * kx) + cscale(Puv, T(2) * kx * ky) + cscale(Pvv, ky * ky) -
                                cscale(Pww, k2);
            sD[wave][1][lane] = imul(Puw, kx) + imul(Pvw, ky);
            sD[wave][2][lane] = cscale(Puu - Pvv, -kx * ky) + cscale(Puv, kx * kx - ky * ky);
            sD[wave][3][lane] = imul(Pvw, kx) - imul(Puw, ky);
            sD[wave][4][lane] = Pad;
        }
        __syncthreads();
        for (int lv = 0; lv < LV && l0 + lv < nq; ++lv) {
            const T* tb = prtab + (size_t)(l0 + lv) * 5 * L;  // wC', wC'', wC, wS, wS'
            const cpx<T> A = sD[lv][0][lane], Bf = sD[lv][1][lane], Ag = sD[lv][2][lane];
            const cpx<T> Bg = sD[lv][3][lane], At = sD[lv][4][lane];
#pragma unroll
            for (int i = 0; i < MCH; ++i) {
                const int m = mb + i;
                if (m < L) {
                    const T cp = tb[m], cpp = tb[L + m], cv = tb[2 * L + m], sv = tb[3 * L + m], sp = tb[4 * L + m];
                    af[i] = af[i] + cscale(A, cp) + cscale(Bf, cpp + k2 * cv);
                    ag[i] = ag[i] + cscale(Ag, sv) + cscale(Bg, sp);
                    at[i] = at[i] + cscale(At, sv);
                }
            }
        }
        __syncthreads();
    }
    if (valid) {
#pragma unroll
        for (int i = 0; i < MCH; ++i) {
            const int m = mb + i;
            if (m < L) {
                Nt[(size_t)p * L + m] = cscale(at[i], mt);
                Nf[(size_t)p * L + m] = cscale(af[i], mp);
                Ng[(size_t)p * L + m] = cscale(ag[i], mp);
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < L) {
        const int m = threadIdx.x;
        cpx<T> sF = mk<T>(0, 0), sG = sF;
        for (int l = 0; l < nq; ++l) {
            const T sp = prtab[(size_t)l * 5 * L + 4 * L + m];
            sF = sF + cscale(prod[3 * fs + (size_t)l * P], sp);
            sG = sG + cscale(prod[4 * fs + (size_t)l * P], sp);
        }
        NF[m] = cscale(sF, -mp);
        NG[m] = cscale(sG, -mp);
    }
}

template <typename T, int MCH>
__global__ void __launch_bounds__(256) k_project(DevGeom g, const cpx<T>* __restrict__ prod,
                                                 const T* __restrict__ prtab, const T* __restrict__ kxa,
                                                 const T* __restrict__ kya, T Pr, cpx<T>* __restrict__ Nt,
                                                 cpx<T>* __restrict__ Nf, cpx<T>* __restrict__ Ng,
                                                 cpx<T>* __restrict__ NF, cpx<T>* __restrict__ NG) {
    const int L = g.L, P = g.P, nq = g.nq;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    const bool valid = p < P;
    const int pc = valid ? p : 0;
    const T kx = kxa[pc], ky = kya[pc];
    const T k2 = kx * kx + ky * ky;
    const size_t fs = (size_t)nq * P;
    const T mt = T(-1), mp = T(-1) / Pr;
    for (int mb = wave * MCH; mb < L; mb += 4 * MCH) {
        cpx<T> af[MCH], ag[MCH], at[MCH];
#pragma unroll
        for (int i = 0; i < MCH; ++i) af[i] = ag[i] = at[i] = mk<T>(0, 0);
        for (int l = 0; l < nq; ++l) {
            const cpx<T>* src = prod + (size_t)l * P + pc;
            const cpx<T> Puu = src[0], Puv = src[fs], Pvv = src[2 * fs], Puw = src[3 * fs];
            const cpx<T> Pvw = src[4 * fs], Pww = src[5 * fs], Pad = src[6 * fs];
            const cpx<T> A = cscale(Puu, kx * kx) + cscale(Puv, T(2) * kx * ky) + cscale(Pvv, ky * ky) -
                             cscale(Pww, k2);
            const cpx<T> Bf = imul(Puw, kx) + imul(Pvw, ky);
            const cpx<T> Ag = cscale(Puu - Pvv, -kx * ky) + cscale(Puv, kx * kx - ky * ky);
            const cpx<T> Bg = imul(Pvw, kx) - imul(Puw, ky);
            const T* tb = prtab + (size_t)l * 5 * L;  // wC', wC'', wC, wS, wS'
#pragma unroll
            for (int i = 0; i < MCH; ++i) {
                const int m = mb + i;
                if (m < L) {
                    const T cp = tb[m], cpp = tb[L + m], cv = tb[2 * L + m], sv = tb[3 * L + m],
                            sp = tb[4 * L + m];
                    af[i] = af[i] + cscale(A, cp) + cscale(Bf, cpp + k2 * cv);
                    ag[i] = ag[i] + cscale(Ag, sv) + cscale(Bg, sp);
                    at[i] = at[i] + cscale(Pad, sv);
                }
            }
        }
        if (valid) {
#pragma unroll
            for (int i = 0; i < MCH; ++i) {
                const int m = mb + i;
                if (m < L) {
                    Nt[(size_t)p * L + m] = cscale(at[i], mt);
                    Nf[(size_t)p * L + m] = cscale(af[i], mp);
                    Ng[(size_t)p * L + m] = cscale(ag[i], mp);
                }
            }
        }
    }
    // mean flow (nonlinear...cu:920-1003): N_F = -<uw, S'>_{k=0}, N_G = -<vw, S'>_{k=0}
    if (blockIdx.x == 0 && threadIdx.x < L) {
        const int m = threadIdx.x;
        cpx<T> sF = mk<T>(0, 0), sG = sF;
        for (int l = 0; l < nq; ++l) {
            const T sp = prtab[(size_t)l * 5 * L + 4 * L + m];
            sF = sF + cscale(prod[3 * fs + (size_t)l * P], sp);
            sG = sG + cscale(prod[4 * fs + (size_t)l * P], sp);
        }
        NF[m] = cscale(sF, -mp);
        NG[m] = cscale(sG, -mp);
    }
}

// ---------------------------------------------------------------------------------------
// Stage 6: group update.  One block (64 lanes) per |k|^2 group; lanes 0..31 serve parity
// class 0 and 32..63 class 1 (rows of the class matrices, looped when a class is larger).
// Per member mode:
//   mode UPD_STEP:     R = B^-1 N;  X += a_cur R;  X += a_prev R_prev;  X = C^-1 X;  R_prev = R
//   mode UPD_INVB:     out = B^-1 N              (inverse_B_matrix::calculate_inverse)
//   mode UPD_IMPLICIT: out = C^-1 X              (linear_implicit_operator::calculate_operator)
// g uses its diagonal factors (inverse_B_matrix.cu:276-300, linear_implicit_operator.cu:419-458).
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(64) k_update(DevGeom g, int L0, int L1, int stride, T Pr, T dt, T a_cur,
                                               T a_prev, int use_prev, int mode,
                                               const int* __restrict__ gstart, const int* __restrict__ gmem,
                                               const T* __restrict__ gmat, const T* __restrict__ k2a,
                                               cpx<T>* __restrict__ Xt, cpx<T>* __restrict__ Xf,
                                               cpx<T>* __restrict__ Xg, const cpx<T>* __restrict__ Nt,
                                               const cpx<T>* __restrict__ Nf, const cpx<T>* __restrict__ Ng,
                                               cpx<T>* __restrict__ Rt, cpx<T>* __restrict__ Rf,
                                               cpx<T>* __restrict__ Rg) {
    extern __shared__ unsigned char smem_raw[];
    const int L = g.L;
    cpx<T>* rhs = reinterpret_cast<cpx<T>*>(smem_raw);  // 2L: [class0 theta, f][class1 theta, f]
    cpx<T>* nfv = rhs + 2 * L;                          // L: N_f reordered by class
    const int grp = blockIdx.x;
    const int b0 = gstart[grp], b1 = gstart[grp + 1];
    const int lane = threadIdx.x;
    const int cls = lane >> 5, cl = lane & 31;
    const int Lc = cls ? L1 : L0, nc = 2 * Lc;
    const int n0 = 2 * L0;
    const T* Cinv = gmat + (size_t)grp * stride + (cls ? (size_t)n0 * n0 : 0);
    const T* Binv = gmat + (size_t)grp * stride + (size_t)n0 * n0 + (size_t)(2 * L1) * (2 * L1) +
                    (cls ? (size_t)L0 * L0 : 0);
    const int rbase = cls ? 2 * L0 : 0;  // offset of this class in rhs
    for (int mi = b0; mi < b1; ++mi) {
        const int p = gmem[mi];
        const size_t base = (size_t)p * L;
        const T k2 = k2a[p];
        // ---- gather: class c holds modes m = c, c+2, ... ; rhs = [theta_c | f_c]
        for (int e = lane; e < L; e += 64) {
            const int c = e & 1, idx = e >> 1;
            const int off = (c ? 2 * L0 : 0);
            const int Lcc = c ? L1 : L0;
            if (mode == UPD_STEP) {
                rhs[off + idx] = Xt[base + e];
                rhs[off + Lcc + idx] = Xf[base + e];
                nfv[off / 2 + idx] = Nf[base + e];
            } else if (mode == UPD_INVB) {
                nfv[off / 2 + idx] = Nf[base + e];
            } else {
                rhs[off + idx] = Xt[base + e];
                rhs[off + Lcc + idx] = Xf[base + e];
            }
        }
        __syncthreads();
        // ---- inverse B on f (class matvec), theta x2, and the AB2 combine
        if (mode != UPD_IMPLICIT) {
            for (int r = cl; r < Lc; r += 32) {
                const int m = cls + 2 * r;
                cpx<T> rf = mk<T>(0, 0);
                for (int j = 0; j < Lc; ++j) rf = rf + cscale(nfv[rbase / 2 + j], Binv[(size_t)j * Lc + r]);
                const cpx<T> rt = cscale(Nt[base + m], T(2));
                if (mode == UPD_INVB) {
                    Rt[base + m] = rt;
                    Rf[base + m] = rf;
                } else {
                    cpx<T> xt = rhs[rbase + r], xf = rhs[rbase + Lc + r];
                    xt = xt + cscale(rt, a_cur);
                    xf = xf + cscale(rf, a_cur);
                    if (use_prev) {
                        xt = xt + cscale(Rt[base + m], a_prev);
                        xf = xf + cscale(Rf[base + m], a_prev);
                    }
                    Rt[base + m] = rt;
                    Rf[base + m] = rf;
                    rhs[rbase + r] = xt;
                    rhs[rbase + Lc + r] = xf;
                }
            }
            __syncthreads();
        }
        // ---- implicit Euler: x = C^-1 rhs per class
        if (mode != UPD_INVB) {
            cpx<T> outv[4];
            int cnt = 0;
            for (int r = cl; r < nc && cnt < 4; r += 32, ++cnt) {
                cpx<T> acc = mk<T>(0, 0);
                for (int j = 0; j < nc; ++j) acc = acc + cscale(rhs[rbase + j], Cinv[(size_t)j * nc + r]);
                outv[cnt] = acc;
            }
            cnt = 0;
            for (int r = cl; r < nc && cnt < 4; r += 32, ++cnt) {
                const int m = cls + 2 * (r < Lc ? r : r - Lc);
                if (r < Lc) Xt[base + m] = outv[cnt];
                else Xf[base + m] = outv[cnt];
            }
        }
        // ---- g: diagonal factors
        for (int m = lane; m < L; m += 64) {
            const T mpi = T(m + 1) * T(M_PI);
            if (mode == UPD_INVB) {
                Rg[base + m] = (k2 != T(0)) ? cscale(Ng[base + m], T(-2) * Pr / k2) : mk<T>(0, 0);
            } else if (mode == UPD_IMPLICIT) {
                Xg[base + m] = (k2 != T(0)) ? cscale(Xg[base + m], T(1) / (T(1) + dt * Pr * (k2 + mpi * mpi)))
                                            : mk<T>(0, 0);
            } else {
                const cpx<T> rg = (k2 != T(0)) ? cscale(Ng[base + m], T(-2) * Pr / k2) : mk<T>(0, 0);
                cpx<T> xg = Xg[base + m] + cscale(rg, a_cur);
                if (use_prev) xg = xg + cscale(Rg[base + m], a_prev);
                Rg[base + m] = rg;
                Xg[base + m] = (k2 != T(0)) ? cscale(xg, T(1) / (T(1) + dt * Pr * (k2 + mpi * mpi))) : mk<T>(0, 0);
            }
        }
        __syncthreads();
    }
}

// Group update for L <= 32 with the class matrices held in registers: one wave per group
// (4 groups per block), lane (class c, row r) keeps row r of C^-1 (n_c <= 32 values) and of
// B^-1_f (L_c <= 16 values) for every member of the group.  Same arithmetic and rounding
// order as k_update; mode UPD_STEP only.
template <typename T>
__global__ void __launch_bounds__(256) k_update32(DevGeom g, int ngroups, int L0, int L1, int stride, T Pr, T dt,
                                                  T a_cur, T a_prev, int use_prev, const int* __restrict__ gstart,
                                                  const int* __restrict__ gmem, const T* __restrict__ gmat,
                                                  const T* __restrict__ k2a, cpx<T>* __restrict__ Xt,
                                                  cpx<T>* __restrict__ Xf, cpx<T>* __restrict__ Xg,
                                                  const cpx<T>* __restrict__ Nt, const cpx<T>* __restrict__ Nf,
                                                  const cpx<T>* __restrict__ Ng, cpx<T>* __restrict__ Rt,
                                                  cpx<T>* __restrict__ Rf, cpx<T>* __restrict__ Rg) {
    __shared__ cpx<T> shm[4][2][2][32];  // [wave][class][0: N_f, 1: rhs][j]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = blockIdx.x * 4 + wave;
    if (grp >= ngroups) return;  // whole wave leaves; no block-wide barrier below
    const int L = g.L;
    const int cls = lane >> 5, r = lane & 31;
    const int Lc = cls ? L1 : L0, nc = 2 * Lc, n0 = 2 * L0;
    const T* rec = gmat + (size_t)grp * stride;
    const T* Cinv = rec + (cls ? (size_t)n0 * n0 : 0);
    const T* Binv = rec + (size_t)n0 * n0 + (size_t)(2 * L1) * (2 * L1) + (cls ? (size_t)L0 * L0 : 0);
    T crow[32], brow[16];
#pragma unroll
    for (int j = 0; j < 32; ++j) crow[j] = (j < nc && r < nc) ? Cinv[(size_t)j * nc + r] : T(0);
#pragma unroll
    for (int j = 0; j < 16; ++j) brow[j] = (j < Lc && r < Lc) ? Binv[(size_t)j * Lc + r] : T(0);
    cpx<T>* snf = shm[wave][cls][0];
    cpx<T>* srhs = shm[wave][cls][1];
    const int b0 = gstart[grp], b1 = gstart[grp + 1];
    for (int mi = b0; mi < b1; ++mi) {
        const int p = gmem[mi];
        const size_t base = (size_t)p * L;
        const T k2 = k2a[p];
        const int m = cls + 2 * r;  // theta/f mode of this lane (valid for r < Lc)
        cpx<T> xt = mk<T>(0, 0), xf = xt;
        if (r < Lc) {
            snf[r] = Nf[base + m];
            xt = Xt[base + m];
            xf = Xf[base + m];
        }
        wave_sync();
        if (r < Lc) {
            cpx<T> rf = mk<T>(0, 0);
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j < Lc) rf = rf + cscale(snf[j], brow[j]);
            const cpx<T> rt = cscale(Nt[base + m], T(2));
            xt = xt + cscale(rt, a_cur);
            xf = xf + cscale(rf, a_cur);
            if (use_prev) {
                xt = xt + cscale(Rt[base + m], a_prev);
                xf = xf + cscale(Rf[base + m], a_prev);
            }
            Rt[base + m] = rt;
            Rf[base + m] = rf;
            srhs[r] = xt;
            srhs[Lc + r] = xf;
        }
        wave_sync();
        if (r < nc) {
            cpx<T> acc = mk<T>(0, 0);
#pragma unroll
            for (int j = 0; j < 32; ++j)
                if (j < nc) acc = acc + cscale(srhs[j], crow[j]);
            const int mm = cls + 2 * (r < Lc ? r : r - Lc);
            if (r < Lc) Xt[base + mm] = acc;
            else Xf[base + mm] = acc;
        }
        // g: diagonal factors, one vertical mode per lane
        if (lane < L) {
            const int mg = lane;
            const T mpi = T(mg + 1) * T(M_PI);
            const cpx<T> rg = (k2 != T(0)) ? cscale(Ng[base + mg], T(-2) * Pr / k2) : mk<T>(0, 0);
            cpx<T> xg = Xg[base + mg] + cscale(rg, a_cur);
            if (use_prev) xg = xg + cscale(Rg[base + mg], a_prev);
            Rg[base + mg] = rg;
            Xg[base + mg] = (k2 != T(0)) ? cscale(xg, T(1) / (T(1) + dt * Pr * (k2 + mpi * mpi))) : mk<T>(0, 0);
        }
        wave_sync();
    }
}

// F, G: inverse B = 2 Pr (imag 0), implicit factor 1/(1 + dt Pr (m pi)^2) (imag 0).
template <typename T>
__global__ void k_update_FG(int L, T Pr, T dt, T a_cur, T a_prev, int use_prev, int mode, cpx<T>* XF,
                            cpx<T>* XG, const cpx<T>* NF, const cpx<T>* NG, cpx<T>* RF, cpx<T>* RG) {
    const int m = threadIdx.x;
    if (m >= L) return;
    const T mpi = T(m + 1) * T(M_PI);
    const T fac = T(1) / (T(1) + dt * Pr * mpi * mpi);
    cpx<T>* X[2] = {XF, XG};
    const cpx<T>* Nn[2] = {NF, NG};
    cpx<T>* R[2] = {RF, RG};
    for (int i = 0; i < 2; ++i) {
        if (mode == UPD_INVB) {
            R[i][m] = mk(T(2) * Pr * Nn[i][m].x, T(0));
        } else if (mode == UPD_IMPLICIT) {
            X[i][m] = mk(fac * X[i][m].x, T(0));
        } else {
            const cpx<T> r = mk(T(2) * Pr * Nn[i][m].x, T(0));
            cpx<T> x = X[i][m] + cscale(r, a_cur);
            if (use_prev) x = x + cscale(R[i][m], a_prev);
            R[i][m] = r;
            X[i][m] = mk(fac * x.x, T(0));
        }
    }
}

// make_real (util.cu): on the kx = 0 column set row r (0 < r < N/2) to conj(row N - r), and
// clear the imaginary part of (0, 0).  Kept rows are symmetric, so source and target are
// both kept or both removed.
template <typename T>
__global__ void k_make_real(DevGeom g, const int* __restrict__ rows, cpx<T>* X0, cpx<T>* X1, cpx<T>* X2) {
    const int L = g.L, NYc = g.NYc, N = g.N;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 3 * NYc * L) return;
    const int fld = e / (NYc * L), rem = e - fld * NYc * L;
    const int kyc = rem / L, m = rem - kyc * L;
    cpx<T>* X = fld == 0 ? X0 : (fld == 1 ? X1 : X2);
    const int row = rows[kyc];
    if (row == 0) {
        X[m].y = T(0);
    } else if (row < N / 2) {
        // source row N - row is at compact index NYc - kyc (rows symmetric, 0 first)
        const int src = NYc - kyc;
        const cpx<T> s = X[(size_t)src * L + m];
        X[(size_t)kyc * L + m] = conjg(s);
    }
}

// Nusselt reductions (FP64): blocks write partial sums of sum_p w_p sum_{m,n} Ia1(m,n)
// Re(k^2 f_m conj theta_n) to red[blockIdx]; red[nblk] = plate sums from mode p = 0.
template <typename T>
__global__ void __launch_bounds__(256) k_nusselt(DevGeom g, const cpx<T>* __restrict__ Xt,
                                                 const cpx<T>* __restrict__ Xf, const T* __restrict__ k2a,
                                                 const double* __restrict__ ia1, double* red) {
    __shared__ double part[256];
    const int L = g.L, NYc = g.NYc;
    const int p = blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (p < g.P) {
        const int kx = p / NYc;
        const double wgt = (kx == 0) ? 1.0 : 2.0;
        const double k2 = (double)k2a[p];
        for (int m = 0; m < L; ++m) {
            const cpx<T> f = Xf[(size_t)p * L + m];
            const double wr = k2 * (double)f.x, wi = k2 * (double)f.y;
            for (int n = (m & 1); n < L; n += 2) {
                const cpx<T> t = Xt[(size_t)p * L + n];
                s += wgt * ia1[m + n * L] * (wr * (double)t.x + wi * (double)t.y);
            }
        }
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) red[blockIdx.x] = part[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double sb = 0.0, st = 0.0;
        for (int m = 0; m < L; ++m) {
            const double a = (double)Xt[m].x, mp = (m + 1) * M_PI;
            sb += mp * a;
            st += mp * (((m + 1) % 2) ? -1.0 : 1.0) * a;
        }
        red[gridDim.x] = sb;
        red[gridDim.x + 1] = st;
    }
}

// =======================================================================================
// 512-point fast paths (FP32): one wavefront per transform, radix-8 Stockham in registers.
// Lane j holds x[j + 64 r], r = 0..7, on entry and X[j + 64 r] on exit.  Two exchanges
// through a per-wave LDS scratch of 576 complex (one pad slot every 8 to spread banks).
// =======================================================================================
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

// 512-point FFT of one wavefront.  tw = table exp(-2 pi i k / 512), k < 512 (LDS or global).
// The scratch element of internal index i is sb[pad8(i) * S] (S = 1 for a private buffer,
// S = tile row stride when a column of an LDS tile doubles as the scratch).
template <typename T, bool INV, int S = 1>
__device__ __forceinline__ void fft512_wave(cpx<T> v[8], cpx<T>* sb, const cpx<T>* tw, int lane) {
    dft8<T, INV>(v);
#pragma unroll
    for (int r = 0; r < 8; ++r) sb[pad8(8 * lane + r) * S] = v[r];
    wave_sync();
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = sb[pad8(lane + 64 * r) * S];
#pragma unroll
    for (int r = 1; r < 8; ++r) {
        cpx<T> w = tw[(lane & 7) * r * 8];
        if (INV) w.y = -w.y;
        v[r] = cmul(v[r], w);
    }
    dft8<T, INV>(v);
    wave_sync();
#pragma unroll
    for (int r = 0; r < 8; ++r) sb[pad8((lane >> 3) * 64 + (lane & 7) + 8 * r) * S] = v[r];
    wave_sync();
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = sb[pad8(lane + 64 * r) * S];
#pragma unroll
    for (int r = 1; r < 8; ++r) {
        cpx<T> w = tw[lane * r];
        if (INV) w.y = -w.y;
        v[r] = cmul(v[r], w);
    }
    dft8<T, INV>(v);
    wave_sync();  // the scratch may be reused right after
}

// x-rows at M = 512: one wave per (level, y) row.  Same arithmetic as k_rows.
__global__ void __launch_bounds__(64) k_rows512(DevGeom g, const cpx<float>* __restrict__ hin,
                                                const cpx<float>* __restrict__ twM, cpx<float>* __restrict__ hout) {
    __shared__ cpx<float> tw[512];
    __shared__ cpx<float> sb[576];
    const int lane = threadIdx.x;
    const int KX = g.KX, KXp = g.KXp, nq = g.nq, N = g.N;
    const int y = blockIdx.x, l = blockIdx.y;
    const size_t fstride = (size_t)nq * N * KXp;
    const size_t roff = ((size_t)l * N + y) * KXp;
#pragma unroll
    for (int r = 0; r < 8; ++r) tw[lane + 64 * r] = twM[lane + 64 * r];
    __syncthreads();
    cpx<float> z[3][8];
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {
        const cpx<float>* x1 = hin + (size_t)(2 * pr) * fstride + roff;
        const cpx<float>* x2 = hin + (size_t)(2 * pr + 1) * fstride + roff;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int n = lane + 64 * r;
            cpx<float> v = mk(0.f, 0.f);
            if (n < KX) {
                cpx<float> a = x1[n], b = x2[n];
                if (n == 0) {
                    a.y = 0.f;
                    b.y = 0.f;
                }
                v = mk(a.x - b.y, a.y + b.x);  // X1 + i X2
            } else if (n > 512 - KX) {
                const cpx<float> a = x1[512 - n], b = x2[512 - n];
                v = mk(a.x + b.y, -a.y + b.x);  // conj(X1) + i conj(X2)
            }
            z[pr][r] = v;
        }
        fft512_wave<float, true>(z[pr], sb, tw, lane);
    }
    const float s = 1.f / (512.f * (float)N);
#pragma unroll
    for (int pr = 0; pr < 4; ++pr) {
        // products of this pair only, so that at most one product pair is live at a time
        cpx<float> p[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const float u = z[0][r].x, v = z[0][r].y, w = z[1][r].x, tx = z[1][r].y, ty = z[2][r].x,
                        tz = z[2][r].y;
            if (pr == 0) p[r] = mk(s * u * u, s * u * v);
            else if (pr == 1) p[r] = mk(s * v * v, s * u * w);
            else if (pr == 2) p[r] = mk(s * v * w, s * w * w);
            else p[r] = mk(s * (u * tx + v * ty + w * tz), 0.f);
        }
        fft512_wave<float, false>(p, sb, tw, lane);
#pragma unroll
        for (int r = 0; r < 8; ++r) sb[pad8(lane + 64 * r)] = p[r];
        wave_sync();
        cpx<float>* o1 = hout + (size_t)(2 * pr) * fstride + roff;
        cpx<float>* o2 = hout + (size_t)(2 * pr + 1) * fstride + roff;
        for (int k = lane; k < KX; k += 64) {
            const cpx<float> zk = sb[pad8(k)], zm = sb[pad8((512 - k) & 511)];
            o1[k] = mk(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
            if (pr < 3) o2[k] = mk(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
        }
        wave_sync();
    }
}

// x-rows at M = 512 with four waves per (level, y) row: waves 0-2 run the three inverse
// pair transforms into LDS slots, then wave w forms product pair w from the slots and runs its
// forward transform, reusing slot w as scratch.  Same arithmetic as k_rows512.
__global__ void __launch_bounds__(256) k_rows512q(DevGeom g, const cpx<float>* __restrict__ hin,
                                                  const cpx<float>* __restrict__ twM, cpx<float>* __restrict__ hout) {
    __shared__ cpx<float> tw[512];
    __shared__ cpx<float> zs[4][576];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int KX = g.KX, KXp = g.KXp, nq = g.nq, N = g.N;
    const int y = blockIdx.x, l = blockIdx.y;
    const size_t fstride = (size_t)nq * N * KXp;
    const size_t roff = ((size_t)l * N + y) * KXp;
    for (int e = threadIdx.x; e < 512; e += 256) tw[e] = twM[e];
    __syncthreads();
    if (wave < 3) {
        const int pr = wave;
        const cpx<float>* x1 = hin + (size_t)(2 * pr) * fstride + roff;
        const cpx<float>* x2 = hin + (size_t)(2 * pr + 1) * fstride + roff;
        cpx<float> v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int n = lane + 64 * r;
            cpx<float> z = mk(0.f, 0.f);
            if (n < KX) {
                cpx<float> a = x1[n], b = x2[n];
                if (n == 0) {
                    a.y = 0.f;
                    b.y = 0.f;
                }
                z = mk(a.x - b.y, a.y + b.x);  // X1 + i X2
            } else if (n > 512 - KX) {
                const cpx<float> a = x1[512 - n], b = x2[512 - n];
                z = mk(a.x + b.y, -a.y + b.x);  // conj(X1) + i conj(X2)
            }
            v[r] = z;
        }
        fft512_wave<float, true>(v, zs[pr], tw, lane);
#pragma unroll
        for (int r = 0; r < 8; ++r) zs[pr][pad8(lane + 64 * r)] = v[r];
    }
    __syncthreads();
    const int pr = wave;
    const float s = 1.f / (512.f * (float)N);
    cpx<float> p[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int i = pad8(lane + 64 * r);
        const cpx<float> z0 = zs[0][i], z1 = zs[1][i], z2 = zs[2][i];
        const float u = z0.x, v = z0.y, w = z1.x, tx = z1.y, ty = z2.x, tz = z2.y;
        if (pr == 0) p[r] = mk(s * u * u, s * u * v);
        else if (pr == 1) p[r] = mk(s * v * v, s * u * w);
        else if (pr == 2) p[r] = mk(s * v * w, s * w * w);
        else p[r] = mk(s * (u * tx + v * ty + w * tz), 0.f);
    }
    __syncthreads();  // every wave has read the slots before they become scratch
    fft512_wave<float, false>(p, zs[pr], tw, lane);
#pragma unroll
    for (int r = 0; r < 8; ++r) zs[pr][pad8(lane + 64 * r)] = p[r];
    wave_sync();
    cpx<float>* o1 = hout + (size_t)(2 * pr) * fstride + roff;
    cpx<float>* o2 = hout + (size_t)(2 * pr + 1) * fstride + roff;
    for (int k = lane; k < KX; k += 64) {
        const cpx<float> zk = zs[pr][pad8(k)], zm = zs[pr][pad8((512 - k) & 511)];
        o1[k] = mk(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        if (pr < 3) o2[k] = mk(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
    }
}

// y inverse at N = 512: a block of 8 waves handles YT = 8 kx columns of one (field, level);
// each wave transforms one column.  The LDS tile [pad8(y)][YT + 1] holds the results and
// doubles as the FFT scratch of each column; it is written out row-wise (64-byte segments).
// Launched on a 1-D grid of 8 * per blocks (XCD-aware order, see xcd_logical_block).
constexpr int YT = 8;
constexpr int YS = YT + 1;
__global__ void __launch_bounds__(512) k_fft_y_inv512(DevGeom g, int ntile, int nblk, int per,
                                                      const cpx<float>* __restrict__ lev,
                                                      const cpx<float>* __restrict__ tw,
                                                      cpx<float>* __restrict__ hin) {
    __shared__ cpx<float> tile[576 * YS];
    const int lb = xcd_logical_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int NYc = g.NYc, P = g.P, nq = g.nq, KX = g.KX;
    const int tix = lb % ntile, rest = lb / ntile;
    const int l = rest % nq, field = rest / nq;
    const int kx0 = tix * YT;
    const int a = (NYc - 1) / 2, shift = 512 - NYc;  // kept rows [0, a] U [512 - a, 511]
    const cpx<float>* src = lev + ((size_t)field * nq + l) * P;
    {
        const int c = wave, kx = kx0 + c;  // one column per wave
        cpx<float> v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int n = lane + 64 * r;
            v[r] = mk(0.f, 0.f);
            if (kx < KX) {
                if (n <= a) v[r] = src[(size_t)kx * NYc + n];
                else if (n >= 512 - a) v[r] = src[(size_t)kx * NYc + n - shift];
            }
        }
        fft512_wave<float, true, YS>(v, tile + c, tw, lane);
#pragma unroll
        for (int r = 0; r < 8; ++r) tile[pad8(lane + 64 * r) * YS + c] = v[r];
    }
    __syncthreads();
    const int ncol = min(YT, KX - kx0);
    cpx<float>* dst = hin + ((size_t)field * nq + l) * (size_t)512 * g.KXp + kx0;
    for (int e = threadIdx.x; e < 512 * YT; e += 512) {
        const int y = e / YT, c = e - y * YT;
        if (c < ncol) dst[(size_t)y * g.KXp + c] = tile[pad8(y) * YS + c];
    }
}

// y forward at N = 512: the transpose of k_fft_y_inv512 (tile in, kept rows out).
__global__ void __launch_bounds__(512) k_fft_y_fwd512(DevGeom g, int ntile, int nblk, int per,
                                                      const cpx<float>* __restrict__ hout,
                                                      const cpx<float>* __restrict__ tw,
                                                      cpx<float>* __restrict__ prod) {
    __shared__ cpx<float> tile[576 * YS];
    const int lb = xcd_logical_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int NYc = g.NYc, P = g.P, nq = g.nq, KX = g.KX;
    const int tix = lb % ntile, rest = lb / ntile;
    const int l = rest % nq, q = rest / nq;
    const int kx0 = tix * YT;
    const int a = (NYc - 1) / 2, shift = 512 - NYc;
    const int ncol = min(YT, KX - kx0);
    const cpx<float>* src = hout + ((size_t)q * nq + l) * (size_t)512 * g.KXp + kx0;
    for (int e = threadIdx.x; e < 512 * YT; e += 512) {
        const int y = e / YT, c = e - y * YT;
        tile[pad8(y) * YS + c] = (c < ncol) ? src[(size_t)y * g.KXp + c] : mk(0.f, 0.f);
    }
    __syncthreads();
    cpx<float>* dst = prod + ((size_t)q * nq + l) * P;
    {
        const int c = wave, kx = kx0 + c;  // one column per wave
        cpx<float> v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = tile[pad8(lane + 64 * r) * YS + c];
        fft512_wave<float, false, YS>(v, tile + c, tw, lane);
        if (kx < KX) {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int n = lane + 64 * r;
                if (n <= a) dst[(size_t)kx * NYc + n] = v[r];
                else if (n >= 512 - a) dst[(size_t)kx * NYc + n - shift] = v[r];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------
// RBC_GENERIC_FFT=1 in the environment selects the generic radix-2 kernels everywhere (used
// by the tests to pin the 512-point kernels against them).
static bool generic_fft_forced() {
    const char* e = getenv("RBC_GENERIC_FFT");
    return e && e[0] == '1';
}
template <typename T>
static bool fast512_rows(const DevGeom& g) {
    return sizeof(T) == 4 && g.M == 512 && g.KX <= 192 && !generic_fft_forced();
}
template <typename T>
static bool fast512_y(const DevGeom& g) {
    return sizeof(T) == 4 && g.N == 512 && !generic_fft_forced();
}
static inline int fft_tile(int N, size_t csz) {
    // two LDS buffers of TX columns; stay within 64 KiB
    int tx = (int)(65536 / (2 * (size_t)N * csz));
    if (tx > 16) tx = 16;
    if (tx < 1) tx = 1;
    return tx;
}

template <typename T>
void Launch<T>::eval(const DevGeom& g, const Buffers& b, void* const X[3], const void* XF, const void* XG,
                     hipStream_t s) {
    const size_t sm = (size_t)3 * g.L * 64 * sizeof(cpx<T>);
    const int LG = g.nq >= 64 ? g.nq / 32 : 1;  // level groups per 64-mode block
    const int nblk = ((g.P + 63) / 64) * LG, per = (nblk + 7) / 8;
    hipLaunchKernelGGL(k_eval<T>, dim3(8 * per), dim3(256), sm, s, g, LG, nblk, per, (const cpx<T>*)X[0],
                       (const cpx<T>*)X[1],
                       (const cpx<T>*)X[2], (const cpx<T>*)XF, (const cpx<T>*)XG, (const T*)b.evtab,
                       (const T*)b.kx, (const T*)b.ky, (cpx<T>*)b.lev);
}

template <typename T>
void Launch<T>::fft_y_inv(const DevGeom& g, const Buffers& b, hipStream_t s) {
    if (fast512_y<T>(g)) {
        const int ntile = (g.KX + YT - 1) / YT, nblk = ntile * g.nq * 6, per = (nblk + 7) / 8;
        hipLaunchKernelGGL(k_fft_y_inv512, dim3(8 * per), dim3(512), 0, s, g, ntile, nblk, per,
                           (const cpx<float>*)b.lev, (const cpx<float>*)b.twN, (cpx<float>*)b.hin);
        return;
    }
    const int TX = fft_tile(g.N, sizeof(cpx<T>));
    dim3 grid((g.KX + TX - 1) / TX, g.nq, 6);
    const size_t sm = (size_t)2 * TX * g.N * sizeof(cpx<T>);
    hipLaunchKernelGGL(k_fft_y_inv<T>, grid, dim3(256), sm, s, g, TX, (const cpx<T>*)b.lev, b.rows,
                       (const cpx<T>*)b.twN, (cpx<T>*)b.hin);
}

template <typename T>
void Launch<T>::rows(const DevGeom& g, const Buffers& b, hipStream_t s) {
    dim3 grid(g.N, g.nq);
    if (fast512_rows<T>(g)) {
        hipLaunchKernelGGL(k_rows512q, grid, dim3(256), 0, s, g, (const cpx<float>*)b.hin, (const cpx<float>*)b.twM,
                           (cpx<float>*)b.hout);
        return;
    }
    const size_t sm = (size_t)8 * g.M * sizeof(cpx<T>);
    hipLaunchKernelGGL(k_rows<T>, grid, dim3(256), sm, s, g, (const cpx<T>*)b.hin, (const cpx<T>*)b.twM,
                       (cpx<T>*)b.hout);
}

template <typename T>
void Launch<T>::fft_y_fwd(const DevGeom& g, const Buffers& b, hipStream_t s) {
    if (fast512_y<T>(g)) {
        const int ntile = (g.KX + YT - 1) / YT, nblk = ntile * g.nq * 7, per = (nblk + 7) / 8;
        hipLaunchKernelGGL(k_fft_y_fwd512, dim3(8 * per), dim3(512), 0, s, g, ntile, nblk, per,
                           (const cpx<float>*)b.hout, (const cpx<float>*)b.twN, (cpx<float>*)b.prod);
        return;
    }
    const int TX = fft_tile(g.N, sizeof(cpx<T>));
    dim3 grid((g.KX + TX - 1) / TX, g.nq, 7);
    const size_t sm = (size_t)2 * TX * g.N * sizeof(cpx<T>);
    hipLaunchKernelGGL(k_fft_y_fwd<T>, grid, dim3(256), sm, s, g, TX, (const cpx<T>*)b.hout, b.rows,
                       (const cpx<T>*)b.twN, (cpx<T>*)b.prod);
}

template <typename T>
void Launch<T>::project(const DevGeom& g, const Buffers& b, double Pr, void* const Nout[3], void* NFo, void* NGo,
                        hipStream_t s) {
    constexpr int MCH = sizeof(T) == 8 ? 4 : 8;
    dim3 grid((g.P + 63) / 64);
    if (g.L <= 4 * MCH) {
        hipLaunchKernelGGL((k_project_lds<T, MCH>), grid, dim3(256), 0, s, g, (const cpx<T>*)b.prod,
                           (const T*)b.prtab, (const T*)b.kx, (const T*)b.ky, (T)Pr, (cpx<T>*)Nout[0],
                           (cpx<T>*)Nout[1], (cpx<T>*)Nout[2], (cpx<T>*)NFo, (cpx<T>*)NGo);
        return;
    }
    hipLaunchKernelGGL((k_project<T, MCH>), grid, dim3(256), 0, s, g, (const cpx<T>*)b.prod, (const T*)b.prtab,
                       (const T*)b.kx, (const T*)b.ky, (T)Pr, (cpx<T>*)Nout[0], (cpx<T>*)Nout[1],
                       (cpx<T>*)Nout[2], (cpx<T>*)NFo, (cpx<T>*)NGo);
}

template <typename T>
void Launch<T>::update(const DevGeom& g, const Buffers& b, int ngroups, int L0, int L1, int stride, double Pr,
                       double dt, StepCoef sc, int mode, void* const Xio[3], void* const Nin[3], void* const Rp[3],
                       hipStream_t s) {
    if (mode == UPD_STEP && g.L <= 32) {
        hipLaunchKernelGGL(k_update32<T>, dim3((ngroups + 3) / 4), dim3(256), 0, s, g, ngroups, L0, L1, stride, (T)Pr,
                           (T)dt, (T)sc.a_cur, (T)sc.a_prev, sc.use_prev, b.gstart, b.gmem, (const T*)b.gmat,
                           (const T*)b.k2, (cpx<T>*)Xio[0], (cpx<T>*)Xio[1], (cpx<T>*)Xio[2], (const cpx<T>*)Nin[0],
                           (const cpx<T>*)Nin[1], (const cpx<T>*)Nin[2], (cpx<T>*)Rp[0], (cpx<T>*)Rp[1],
                           (cpx<T>*)Rp[2]);
        return;
    }
    const size_t sm = (size_t)3 * g.L * sizeof(cpx<T>);
    hipLaunchKernelGGL(k_update<T>, dim3(ngroups), dim3(64), sm, s, g, L0, L1, stride, (T)Pr, (T)dt,
                       (T)sc.a_cur, (T)sc.a_prev, sc.use_prev, mode, b.gstart, b.gmem, (const T*)b.gmat,
                       (const T*)b.k2, (cpx<T>*)Xio[0], (cpx<T>*)Xio[1], (cpx<T>*)Xio[2], (const cpx<T>*)Nin[0],
                       (const cpx<T>*)Nin[1], (const cpx<T>*)Nin[2], (cpx<T>*)Rp[0], (cpx<T>*)Rp[1],
                       (cpx<T>*)Rp[2]);
}

template <typename T>
void Launch<T>::update_FG(const DevGeom& g, double Pr, double dt, StepCoef sc, int mode, void* XF, void* XG,
                          const void* NF, const void* NG, void* RpF, void* RpG, hipStream_t s) {
    hipLaunchKernelGGL(k_update_FG<T>, dim3(1), dim3(64 * ((g.L + 63) / 64)), 0, s, g.L, (T)Pr, (T)dt,
                       (T)sc.a_cur, (T)sc.a_prev, sc.use_prev, mode, (cpx<T>*)XF, (cpx<T>*)XG,
                       (const cpx<T>*)NF, (const cpx<T>*)NG, (cpx<T>*)RpF, (cpx<T>*)RpG);
}

template <typename T>
void Launch<T>::make_real(const DevGeom& g, const Buffers& b, void* const X[3], hipStream_t s) {
    const int n = 3 * g.NYc * g.L;
    hipLaunchKernelGGL(k_make_real<T>, dim3((n + 255) / 256), dim3(256), 0, s, g, b.rows, (cpx<T>*)X[0],
                       (cpx<T>*)X[1], (cpx<T>*)X[2]);
}

template <typename T>
void Launch<T>::nusselt(const DevGeom& g, const Buffers& b, const void* ia1, hipStream_t s) {
    const int nblk = (g.P + 255) / 256;
    hipLaunchKernelGGL(k_nusselt<T>, dim3(nblk), dim3(256), 0, s, g, (const cpx<T>*)b.X[0],
                       (const cpx<T>*)b.X[1], (const T*)b.k2, (const double*)ia1, b.red);
}

template struct Launch<float>;
template struct Launch<double>;

}  // namespace rbc
