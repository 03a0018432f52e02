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

#include "rbc_ctx.hpp"
#include "rbc_device.hpp"

namespace rbc {

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
// Stage 3: one (level, y) row per block: three complex inverse FFTs of length M carry the
// six real fields in pairs (u+iv, w+i theta_x, theta_y+i theta_z); the six product fields
// h = uu + vv - 2 ww, d = uu - vv, uv, uw, vw and u.grad(theta), scaled by 1/(M N) as the
// reference scales before its R2C (nonlinear...cu:252), go back through three forward FFTs
// (pairs) and are unpacked to kx = 0..KX-1.  They carry everything the level combinations need
// (kx^2 Puu + ky^2 Pvv - k^2 Pww = (k^2 / 2) h + ((kx^2 - ky^2) / 2) d, Puu - Pvv = d).  The
// imaginary parts of the kx = 0 inputs are dropped, as a C2R does.  In every product pair the
// second field is scaled by a power of two to the size of the first (block-wide maxima), so each
// keeps the rounding of its own size.
// ---------------------------------------------------------------------------------------
// power of two that brings a second pair member of row maximum mb to the first's ma (exact)
__device__ __forceinline__ double pair_scale_d(double ma, double mb) {
    if (!(mb > 0.0) || !(ma > 0.0)) return 1.0;
    int ea, eb;
    frexp(ma, &ea);
    frexp(mb, &eb);
    return ldexp(1.0, max(-200, min(200, ea - eb)));
}

// A block handles RB consecutive y-rows of one level (RB M <= 512 complex per field pair, so that
// the passes of the LDS transform carry enough butterflies to keep every thread busy); the
// blockDim / RB threads of a row take its maxima together.  TWL: the twiddles are staged in LDS.
template <typename T, bool TWL>
__global__ void __launch_bounds__(256) k_rows(DevGeom g, int RB, const cpx<T>* __restrict__ hin,
                                              const cpx<T>* __restrict__ twM, cpx<T>* __restrict__ hout) {
    extern __shared__ unsigned char smem_raw[];
    __shared__ int smaxb[8][6];
    __shared__ T sisg[8][3];  // inverse pair scales per (row, pair)
    const int M = g.M, N = g.N, KX = g.KX, KXp = g.KXp, nq = g.nq;
    const int FM = 3 * RB * M;                        // complex values per buffer
    cpx<T>* a = reinterpret_cast<cpx<T>*>(smem_raw);  // [RB][3][M]
    cpx<T>* b = a + FM;                               // [RB][3][M]
    cpx<T>* tws = b + FM;                             // [M] when TWL
    const int y0 = blockIdx.x * RB, l = blockIdx.y;
    const size_t fstride = (size_t)nq * N * KXp;
    const size_t roff0 = ((size_t)l * N + y0) * KXp;
    // build the three Hermitian-extended pair spectra Z = X1 + i X2 of every row
    for (int e = threadIdx.x; e < FM; e += blockDim.x) a[e] = mk<T>(0, 0);
    if (TWL)
        for (int e = threadIdx.x; e < M; e += blockDim.x) tws[e] = twM[e];
    if (threadIdx.x < 6 * RB) smaxb[threadIdx.x / 6][threadIdx.x % 6] = 0;
    __syncthreads();
    const cpx<T>* tw = TWL ? tws : twM;
    for (int e = threadIdx.x; e < 3 * RB * KX; e += blockDim.x) {
        const int rp = e / KX, k = e - rp * KX;  // rp = row * 3 + pair
        const int rr = rp / 3, pr = rp - 3 * rr;
        const size_t off = roff0 + (size_t)rr * KXp + k;
        cpx<T> x1 = hin[(size_t)(2 * pr) * fstride + off];
        cpx<T> x2 = hin[(size_t)(2 * pr + 1) * fstride + off];
        if (k == 0) {
            x1.y = 0;
            x2.y = 0;
        }
        cpx<T>* z = a + rp * M;
        z[k] = mk(x1.x - x2.y, x1.y + x2.x);  // X1 + i X2
        if (k > 0) z[M - k] = mk(x1.x + x2.y, -x1.y + x2.x);  // conj(X1) + i conj(X2)
    }
    __syncthreads();
    cpx<T>* r = lds_fft(a, b, g.logM, 3 * RB, tw, true);
    cpx<T>* o = (r == a) ? b : a;
    const T s = T(1) / (T(M) * T(N));
    // pass 1: row maxima of the six products (non-negative floats order like their bits)
    const int tpr = blockDim.x / RB, rr = threadIdx.x / tpr, tx0 = threadIdx.x - rr * tpr;
    const cpx<T>* rz = r + (size_t)rr * 3 * M;
    {
        float pm[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int x = tx0; x < M; x += tpr) {
            const cpx<T> z0 = rz[x], z1 = rz[M + x], z2 = rz[2 * M + x];
            const T u = z0.x, v = z0.y, w = z1.x, tx = z1.y, ty = z2.x, tz = z2.y;
            pm[0] = fmaxf(pm[0], fabsf((float)(u * u + v * v - T(2) * w * w)));
            pm[1] = fmaxf(pm[1], fabsf((float)(u * u - v * v)));
            pm[2] = fmaxf(pm[2], fabsf((float)(u * v)));
            pm[3] = fmaxf(pm[3], fabsf((float)(u * w)));
            pm[4] = fmaxf(pm[4], fabsf((float)(v * w)));
            pm[5] = fmaxf(pm[5], fabsf((float)(u * tx + v * ty + w * tz)));
        }
        // the tpr threads of a row are whole waves or aligned groups of 32, 16, ... lanes
        const int span = tpr < 64 ? tpr : 64;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            float x = pm[i];
            for (int o2 = span >> 1; o2 > 0; o2 >>= 1) x = fmaxf(x, __shfl_xor(x, o2));
            if ((threadIdx.x & (span - 1)) == 0) atomicMax(&smaxb[rr][i], __float_as_int(x));
        }
    }
    __syncthreads();
    T sg[3];
#pragma unroll
    for (int pr = 0; pr < 3; ++pr)
        sg[pr] = (T)pair_scale_d(__int_as_float(smaxb[rr][2 * pr]), __int_as_float(smaxb[rr][2 * pr + 1]));
    if (threadIdx.x < 3 * RB) {
        const int q = threadIdx.x / 3, pr = threadIdx.x - 3 * q;
        sisg[q][pr] = T(1) / (T)pair_scale_d(__int_as_float(smaxb[q][2 * pr]), __int_as_float(smaxb[q][2 * pr + 1]));
    }
    // pass 2: the scaled product pairs
    {
        cpx<T>* oz = o + (size_t)rr * 3 * M;
        for (int x = tx0; x < M; x += tpr) {
            const cpx<T> z0 = rz[x], z1 = rz[M + x], z2 = rz[2 * M + x];
            const T u = z0.x, v = z0.y, w = z1.x, tx = z1.y, ty = z2.x, tz = z2.y;
            oz[x] = mk(s * (u * u + v * v - T(2) * w * w), sg[0] * s * (u * u - v * v));
            oz[M + x] = mk(s * u * v, sg[1] * s * u * w);
            oz[2 * M + x] = mk(s * v * w, sg[2] * s * (u * tx + v * ty + w * tz));
        }
    }
    __syncthreads();
    cpx<T>* spare = (o == a) ? b : a;
    cpx<T>* R = lds_fft(o, spare, g.logM, 3 * RB, tw, false);
    // unpack pair spectra: X1_k = (Z_k + conj Z_{M-k}) / 2, X2_k = -i (Z_k - conj Z_{M-k}) / (2 s)
    for (int e = threadIdx.x; e < 3 * RB * KX; e += blockDim.x) {
        const int rp = e / KX, k = e - rp * KX;
        const int rq = rp / 3, pr = rp - 3 * rq;
        const T is = sisg[rq][pr];
        const cpx<T>* Z = R + rp * M;
        const cpx<T> zk = Z[k], zm = Z[(M - k) & (M - 1)];
        const cpx<T> x1 = mk(T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y));
        const cpx<T> x2 = mk(T(0.5) * is * (zk.y + zm.y), T(-0.5) * is * (zk.x - zm.x));
        const size_t off = roff0 + (size_t)rq * KXp + k;
        hout[(size_t)(2 * pr) * fstride + off] = x1;
        hout[(size_t)(2 * pr + 1) * fstride + off] = x2;
    }
}

// ---------------------------------------------------------------------------------------
// Stage 5: projection.  The six product spectra of a level (h, d, uv, uw, vw, u.grad(theta))
// combine into five level combinations (formed where (kx, ky) of the output are known: in the
// forward y-FFT, or by k_combine on the generic path):
//   D0 = kx^2 Puu + 2 kx ky Puv + ky^2 Pvv - k^2 Pww         (tested with C')
//   D1 = i kx Puw + i ky Pvw                                  (with C'' + k^2 C)
//   D2 = -kx ky (Puu - Pvv) + (kx^2 - ky^2) Puv               (with S)
//   D3 = i kx Pvw - i ky Puw                                  (with S')
//   D4 = u.grad(theta)                                        (with S)
// which is the Galerkin sum of nonlinear...cu:479-1067 with each triple integral written
// as a quadrature over the levels.  Outputs are scaled by -1 (theta) and -1/Pr (:1366-1370).
// The mean flow needs P_uw, P_vw at k = 0 per level (fgw), which D1, D3 lose (times k).
// ---------------------------------------------------------------------------------------
// product fields: 0 h = uu + vv - 2 ww, 1 d = uu - vv, 2 uv, 3 uw, 4 vw, 5 u.grad(theta)
template <typename T>
__device__ __forceinline__ void combine_add(int q, cpx<T> V, T kx, T ky, cpx<T> D[5]) {
    switch (q) {
        case 0: D[0] = D[0] + cscale(V, T(0.5) * (kx * kx + ky * ky)); break;                             // h
        case 1: D[0] = D[0] + cscale(V, T(0.5) * (kx * kx - ky * ky)); D[2] = D[2] + cscale(V, -kx * ky); break;  // d
        case 2: D[0] = D[0] + cscale(V, T(2) * kx * ky); D[2] = D[2] + cscale(V, kx * kx - ky * ky); break;  // uv
        case 3: D[1] = D[1] + imul(V, kx); D[3] = D[3] - imul(V, ky); break;                               // uw
        case 4: D[1] = D[1] + imul(V, ky); D[3] = D[3] + imul(V, kx); break;                               // vw
        default: D[4] = D[4] + V; break;                                                                   // u.grad
    }
}

// generic path: six product spectra [q][l][p] -> five combinations [t][l][p] and fgw
template <typename T>
__global__ void __launch_bounds__(256) k_combine(DevGeom g, const cpx<T>* __restrict__ prod,
                                                 const T* __restrict__ kxa, const T* __restrict__ kya,
                                                 cpx<T>* __restrict__ D, cpx<T>* __restrict__ fgw) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t P = g.P, fs = (size_t)g.nq * P;
    if (e >= fs) return;
    const size_t p = e % P;
    const T kx = kxa[p], ky = kya[p];
    cpx<T> d[5];
#pragma unroll
    for (int t = 0; t < 5; ++t) d[t] = mk<T>(0, 0);
#pragma unroll
    for (int q = 0; q < 6; ++q) combine_add(q, prod[q * fs + e], kx, ky, d);
#pragma unroll
    for (int t = 0; t < 5; ++t) D[t * fs + e] = d[t];
    if (p == 0) {
        const size_t l = e / P;
        fgw[2 * l] = prod[3 * fs + e];
        fgw[2 * l + 1] = prod[4 * fs + e];
    }
}

// mean flow (nonlinear...cu:920-1003): N_F = -<uw, S'>_{k=0}, N_G = -<vw, S'>_{k=0}, times -1/Pr
template <typename T>
__device__ __forceinline__ void mean_flow_projection(const DevGeom& g, const cpx<T>* __restrict__ fgw,
                                                     const T* __restrict__ prtab, T Pr, cpx<T>* NF, cpx<T>* NG) {
    // first wave of block 0: lane = (level parity, m) for L <= 32 (the two halves of the wave
    // take alternate levels and are added at the end), lane = m for L <= 64
    const int L = g.L, nq = g.nq;
    if (blockIdx.x != 0 || threadIdx.x >= 64) return;
    const int G = L <= 32 ? 2 : 1, m = threadIdx.x % (64 / G), par = threadIdx.x / (64 / G);
    cpx<T> sF = mk<T>(0, 0), sG = sF;
    if (m < L) {
#pragma unroll 4
        for (int l = par; l < nq; l += G) {
            const T sp = prtab[(size_t)l * 5 * L + 4 * L + m];
            sF = sF + cscale(fgw[2 * l], sp);
            sG = sG + cscale(fgw[2 * l + 1], sp);
        }
    }
    if (G == 2) {
        sF.x += __shfl_xor(sF.x, 32);
        sF.y += __shfl_xor(sF.y, 32);
        sG.x += __shfl_xor(sG.x, 32);
        sG.y += __shfl_xor(sG.y, 32);
    }
    if (par == 0 && m < L) {
        NF[m] = cscale(sF, T(1) / Pr);
        NG[m] = cscale(sG, T(1) / Pr);
    }
}

// Variant for L <= 4 * MCH: per round of 4 levels the block stages the combinations of 4 levels
// of its 64 modes in LDS; every wave then accumulates its own MCH vertical modes.
template <typename T, int MCH>
__global__ void __launch_bounds__(256) k_project_lds(DevGeom g, const cpx<T>* __restrict__ D,
                                                     const cpx<T>* __restrict__ fgw, const T* __restrict__ prtab,
                                                     const T* __restrict__ kxa, const T* __restrict__ kya, T Pr,
                                                     cpx<T>* __restrict__ Nt, cpx<T>* __restrict__ Nf,
                                                     cpx<T>* __restrict__ Ng, cpx<T>* __restrict__ NF,
                                                     cpx<T>* __restrict__ NG) {
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
#pragma unroll
            for (int t = 0; t < 5; ++t) sD[wave][t][lane] = D[t * fs + (size_t)ls * P + pc];
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
    mean_flow_projection(g, fgw, prtab, Pr, NF, NG);
}

// General L: block = 64 modes; wave w accumulates vertical modes m = w*MCH + 4*MCH*i.
template <typename T, int MCH>
__global__ void __launch_bounds__(256) k_project(DevGeom g, const cpx<T>* __restrict__ D,
                                                 const cpx<T>* __restrict__ fgw, const T* __restrict__ prtab,
                                                 const T* __restrict__ kxa, const T* __restrict__ kya, T Pr,
                                                 cpx<T>* __restrict__ Nt, cpx<T>* __restrict__ Nf,
                                                 cpx<T>* __restrict__ Ng, cpx<T>* __restrict__ NF,
                                                 cpx<T>* __restrict__ NG) {
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
            const cpx<T>* src = D + (size_t)l * P + pc;
            const cpx<T> A = src[0], Bf = src[fs], Ag = src[2 * fs], Bg = src[3 * fs], At = src[4 * fs];
            const T* tb = prtab + (size_t)l * 5 * L;
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
    mean_flow_projection(g, fgw, prtab, Pr, NF, NG);
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
// (4 groups per block), lane (class c, row r) keeps row r of C^-1 (n_c <= NC values) and of
// B^-1_f (L_c <= NC / 2 values) for every member of the group; NC = 32, or 16 for L <= 16.
// Same arithmetic and rounding order as k_update; mode UPD_STEP only.
template <typename T, int NC>
__device__ __forceinline__ void k_update32_body(DevGeom g, int nchunk, int L0, int L1, int stride, T Pr, T dt,
                                                T a_cur, T a_prev, int use_prev, const int* __restrict__ cgrp,
                                                const int* __restrict__ cbeg,
                                                const int* __restrict__ gmem, const T* __restrict__ gmat,
                                                const T* __restrict__ k2a, cpx<T>* __restrict__ Xt,
                                                cpx<T>* __restrict__ Xf, cpx<T>* __restrict__ Xg,
                                                const cpx<T>* __restrict__ Nt, const cpx<T>* __restrict__ Nf,
                                                const cpx<T>* __restrict__ Ng, cpx<T>* __restrict__ Rt,
                                                cpx<T>* __restrict__ Rf, cpx<T>* __restrict__ Rg) {
    __shared__ cpx<T> shm[4][2][2][32];  // [wave][class][0: N_f, 1: rhs][j]
    // the wave index as a wave-uniform value: chunk, group and member indices then live in
    // scalar registers and come from scalar loads, which the vector loads do not wait for
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ch = blockIdx.x * 4 + wave;  // one chunk (<= kChunk members of one group) per wave
    if (ch >= nchunk) return;               // whole wave leaves; no block-wide barrier below
    const int grp = cgrp[ch];
    const int L = g.L;
    const int cls = lane >> 5, r = lane & 31;
    const int Lc = cls ? L1 : L0, nc = 2 * Lc, n0 = 2 * L0;
    const T* rec = gmat + (size_t)grp * stride;
    const T* Cinv = rec + (cls ? (size_t)n0 * n0 : 0);
    const T* Binv = rec + (size_t)n0 * n0 + (size_t)(2 * L1) * (2 * L1) + (cls ? (size_t)L0 * L0 : 0);
    // Every load below is issued by every lane (lanes outside the class read element 0 of a
    // valid array and discard it), and the matrix-vector sums run over the full 16 / 32 terms
    // with zero matrix entries and zero LDS slots beyond the class size: no branch depends on
    // the class sizes, so the compiler can keep the next member's loads in flight across the
    // current member's arithmetic.  The extra terms add exact zeros (bit-identical sums).
    T crow[NC], brow[NC / 2];
    {
        const T* csrc = nc > 0 ? Cinv : rec;
        const T* bsrc = Lc > 0 ? Binv : rec;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            const bool in = j < nc && r < nc;
            const T v = csrc[in ? (size_t)j * nc + r : 0];
            crow[j] = in ? v : T(0);
        }
#pragma unroll
        for (int j = 0; j < NC / 2; ++j) {
            const bool in = j < Lc && r < Lc;
            const T v = bsrc[in ? (size_t)j * Lc + r : 0];
            brow[j] = in ? v : T(0);
        }
    }
    cpx<T>* snf = shm[wave][cls][0];
    cpx<T>* srhs = shm[wave][cls][1];
    for (int e = r; e < 64; e += 32) shm[wave][cls][e >> 5][e & 31] = mk<T>(0, 0);  // slots beyond the class stay 0
    const int b0 = cbeg[ch], b1 = cbeg[ch + 1];
    const bool tf = r < Lc, gl = lane < L;
    const int m = tf ? cls + 2 * r : 0;  // theta/f mode of this lane (element 0 outside the class)
    const int mg = gl ? lane : 0;        // g mode of this lane
    // one member's loads (issued together)
    struct Member {
        cpx<T> nf, xt, xf, nt, rt, rf, ng, xg, rg;
        T k2;
        size_t base;
    };
    auto fetch = [&](int mi, Member& d) {
        const int p = gmem[mi];
        d.base = (size_t)p * L;
        d.k2 = k2a[p];
        d.nf = Nf[d.base + m];
        d.xt = Xt[d.base + m];
        d.xf = Xf[d.base + m];
        d.nt = Nt[d.base + m];
        d.rt = Rt[d.base + m];  // read also without history (then unused)
        d.rf = Rf[d.base + m];
        d.ng = Ng[d.base + mg];
        d.xg = Xg[d.base + mg];
        d.rg = Rg[d.base + mg];
    };
    auto update = [&](const Member& d) {
        const size_t base = d.base;
        const T k2 = d.k2;
        if (tf) snf[r] = d.nf;
        wave_sync();
        if (tf) {
            cpx<T> rf = mk<T>(0, 0);
#pragma unroll
            for (int j = 0; j < NC / 2; ++j) rf = rf + cscale(snf[j], brow[j]);
            const cpx<T> rt = cscale(d.nt, T(2));
            cpx<T> x1 = d.xt + cscale(rt, a_cur), x2 = d.xf + cscale(rf, a_cur);
            if (use_prev) {
                x1 = x1 + cscale(d.rt, a_prev);
                x2 = x2 + cscale(d.rf, a_prev);
            }
            Rt[base + m] = rt;
            Rf[base + m] = rf;
            srhs[r] = x1;
            srhs[Lc + r] = x2;
        }
        wave_sync();
        if (r < nc) {
            cpx<T> acc = mk<T>(0, 0);
#pragma unroll
            for (int j = 0; j < NC; ++j) acc = acc + cscale(srhs[j], crow[j]);
            const int mm = cls + 2 * (r < Lc ? r : r - Lc);
            if (r < Lc) Xt[base + mm] = acc;
            else Xf[base + mm] = acc;
        }
        // g: diagonal factors, one vertical mode per lane
        if (gl) {
            const T mpi = T(lane + 1) * T(M_PI);
            const cpx<T> rg = (k2 != T(0)) ? cscale(d.ng, T(-2) * Pr / k2) : mk<T>(0, 0);
            cpx<T> x3 = d.xg + cscale(rg, a_cur);
            if (use_prev) x3 = x3 + cscale(d.rg, a_prev);
            Rg[base + lane] = rg;
            Xg[base + lane] = (k2 != T(0)) ? cscale(x3, T(1) / (T(1) + dt * Pr * (k2 + mpi * mpi))) : mk<T>(0, 0);
        }
        wave_sync();
    };
    // Members of a chunk (at most kMaxChunk) two ahead: member mi + 2's loads are in flight while
    // member mi is computed.  Three register sets taken in turn over a fully unrolled loop, so
    // that no set is copied (a copy would wait for its loads); loads return in issue order, so
    // member mi waits for its own only.
    constexpr int kMaxChunk = 4;
    Member d[3];
    if (b0 < b1) fetch(b0, d[0]);
    if (b0 + 1 < b1) fetch(b0 + 1, d[1]);
#pragma unroll
    for (int u = 0; u < kMaxChunk; ++u) {
        if (b0 + u >= b1) break;
        if (b0 + u + 2 < b1 && u + 2 < kMaxChunk) fetch(b0 + u + 2, d[(u + 2) % 3]);
        update(d[u % 3]);
    }
    for (int mi = b0 + kMaxChunk; mi < b1; ++mi) {  // longer chunks (the host makes none): in turn
        Member e;
        fetch(mi, e);
        update(e);
    }
}

template <typename T, int NC>
__global__ void __launch_bounds__(256) k_update32(DevGeom g, int nchunk, int L0, int L1, int stride, T Pr, T dt,
                                                  T a_cur, T a_prev, int use_prev, const int* __restrict__ cgrp,
                                                  const int* __restrict__ cbeg,
                                                  const int* __restrict__ gmem, const T* __restrict__ gmat,
                                                  const T* __restrict__ k2a, cpx<T>* __restrict__ Xt,
                                                  cpx<T>* __restrict__ Xf, cpx<T>* __restrict__ Xg,
                                                  const cpx<T>* __restrict__ Nt, const cpx<T>* __restrict__ Nf,
                                                  const cpx<T>* __restrict__ Ng, cpx<T>* __restrict__ Rt,
                                                  cpx<T>* __restrict__ Rf, cpx<T>* __restrict__ Rg) {
    k_update32_body<T, NC>(g, nchunk, L0, L1, stride, Pr, dt, a_cur, a_prev, use_prev, cgrp, cbeg, gmem, gmat, k2a, Xt,
                       Xf, Xg, Nt, Nf, Ng, Rt, Rf, Rg);
}
// one member's arguments of k_update32_batch, which runs the k_update32 body for member
// blockIdx.z / zdim of a batch (the rest of blockIdx.z is the body's own z index)
template <typename T>
struct k_update32_args {
    DevGeom g;
    int nchunk;
    int L0;
    int L1;
    int stride;
    T Pr;
    T dt;
    T a_cur;
    T a_prev;
    int use_prev;
    const int* cgrp;
    const int* cbeg;
    const int* gmem;
    const T* gmat;
    const T* k2a;
    cpx<T>* Xt;
    cpx<T>* Xf;
    cpx<T>* Xg;
    const cpx<T>* Nt;
    const cpx<T>* Nf;
    const cpx<T>* Ng;
    cpx<T>* Rt;
    cpx<T>* Rf;
    cpx<T>* Rg;
};
template <typename T, int NC>
__global__ void __launch_bounds__(256) k_update32_batch(const k_update32_args<T>* __restrict__ tab, int zdim) {
    const int mb = blockIdx.z / zdim;
    const k_update32_args<T>& a = tab[mb];
    k_update32_body<T, NC>(a.g, a.nchunk, a.L0, a.L1, a.stride, a.Pr, a.dt, a.a_cur, a.a_prev, a.use_prev, a.cgrp, a.cbeg,
                       a.gmem, a.gmat, a.k2a, a.Xt, a.Xf, a.Xg, a.Nt, a.Nf, a.Ng, a.Rt, a.Rf, a.Rg);
}


// F, G: inverse B = 2 Pr (imag 0), implicit factor 1/(1 + dt Pr (m pi)^2) (imag 0); thread m.
template <typename T>
__device__ __forceinline__ void update_FG_mode(int m, int L, T Pr, T dt, T a_cur, T a_prev, int use_prev, int mode,
                                               cpx<T>* XF, cpx<T>* XG, const cpx<T>* NF, const cpx<T>* NG,
                                               cpx<T>* RF, cpx<T>* RG) {
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
template <typename T>
__global__ void k_update_FG(int L, T Pr, T dt, T a_cur, T a_prev, int use_prev, int mode, cpx<T>* XF,
                            cpx<T>* XG, const cpx<T>* NF, const cpx<T>* NG, cpx<T>* RF, cpx<T>* RG) {
    update_FG_mode<T>(threadIdx.x, L, Pr, dt, a_cur, a_prev, use_prev, mode, XF, XG, NF, NG, RF, RG);
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

// the end of a time step in one launch: make_real on the blocks before the last, the F, G
// update (mode UPD_STEP) and the health check on the last block.  A non-finite coefficient
// anywhere in the state reaches every mode within one nonlinear evaluation (each Fourier mode
// of a product sums every grid point), so the mean modes theta(0,0), F and G, which the last
// block reads anyway, show it by the end of the step that consumed it.  health[1] counts the
// steps; health[0] keeps 1 + the count of the first step that ended non-finite.
template <typename T>
__device__ __forceinline__ void k_finish_body(DevGeom g, const int* __restrict__ rows, cpx<T>* X0, cpx<T>* X1,
                                              cpx<T>* X2, T Pr, T dt, T a_cur, T a_prev, int use_prev, cpx<T>* XF,
                                              cpx<T>* XG, const cpx<T>* NF, const cpx<T>* NG, cpx<T>* RF,
                                              cpx<T>* RG, int* __restrict__ health) {
    if (blockIdx.x == gridDim.x - 1) {
        update_FG_mode<T>(threadIdx.x, g.L, Pr, dt, a_cur, a_prev, use_prev, UPD_STEP, XF, XG, NF, NG, RF, RG);
        int bad = 0;
        if ((int)threadIdx.x < g.L) {
            const int m = threadIdx.x;
            // theta(0,0) real parts only: the make_real blocks clear its imaginary parts meanwhile
            bad = !(isfinite(XF[m].x) && isfinite(XG[m].x) && isfinite(X0[m].x));
        }
        bad = __syncthreads_or(bad);
        if (threadIdx.x == 0) {
            const int n = health[1];
            if (bad && health[0] == 0) health[0] = n + 1;
            health[1] = n + 1;
        }
        return;
    }
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
        const int src = NYc - kyc;
        const cpx<T> s = X[(size_t)src * L + m];
        X[(size_t)kyc * L + m] = conjg(s);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_finish(DevGeom g, const int* __restrict__ rows, cpx<T>* X0, cpx<T>* X1,
                                                cpx<T>* X2, T Pr, T dt, T a_cur, T a_prev, int use_prev, cpx<T>* XF,
                                                cpx<T>* XG, const cpx<T>* NF, const cpx<T>* NG, cpx<T>* RF,
                                                cpx<T>* RG, int* __restrict__ health) {
    k_finish_body<T>(g, rows, X0, X1, X2, Pr, dt, a_cur, a_prev, use_prev, XF, XG, NF, NG, RF, RG, health);
}
// one member's arguments of k_finish_batch, which runs the k_finish body for member
// blockIdx.z / zdim of a batch (the rest of blockIdx.z is the body's own z index)
template <typename T>
struct k_finish_args {
    DevGeom g;
    const int* rows;
    cpx<T>* X0;
    cpx<T>* X1;
    cpx<T>* X2;
    T Pr;
    T dt;
    T a_cur;
    T a_prev;
    int use_prev;
    cpx<T>* XF;
    cpx<T>* XG;
    const cpx<T>* NF;
    const cpx<T>* NG;
    cpx<T>* RF;
    cpx<T>* RG;
    int* health;
};
template <typename T>
__global__ void __launch_bounds__(256) k_finish_batch(const k_finish_args<T>* __restrict__ tab, int zdim) {
    const int mb = blockIdx.z / zdim;
    const k_finish_args<T>& a = tab[mb];
    k_finish_body<T>(a.g, a.rows, a.X0, a.X1, a.X2, a.Pr, a.dt, a.a_cur, a.a_prev, a.use_prev, a.XF, a.XG, a.NF, a.NG,
                     a.RF, a.RG, a.health);
}


// ---------------------------------------------------------------------------------------
// Device-resident exchange with the reference layout ((M/2+1) x N x L complex, kx fastest,
// matrix/matrix_device.h:13-63) in the context's element type.  The kept rows are
// [0, a] U [N - a, N - 1] (a = (NYc - 1) / 2, checked at create), so a row's compact index
// follows from the row itself.
// ---------------------------------------------------------------------------------------

// Reference layout [m][row][kx] (kx fastest, M/2 + 1 columns) <-> compact [p = kx NYc + kyc][m],
// three fields, as 32 x 32 tiles of (kx, m) of one y row staged in LDS: the reference side is read
// or written in runs of 32 kx and the compact side in runs of 32 m, so both sides are coalesced.
// blockIdx = (kx tile, kept row kyc or reference row, field * m tiles + m tile).
__device__ __forceinline__ int kept_row_of(const DevGeom& g, int kyc) {
    const int a = (g.NYc - 1) / 2;
    return kyc <= a ? kyc : g.N - (g.NYc - kyc);
}
__device__ __forceinline__ int kept_index_of(const DevGeom& g, int row) {  // -1: removed row
    const int a = (g.NYc - 1) / 2;
    return row <= a ? row : (row >= g.N - a ? g.NYc - (g.N - row) : -1);
}

template <typename T, typename TD = T>
__global__ void __launch_bounds__(256) k_ref_to_compact(DevGeom g, const cpx<T>* __restrict__ s0,
                                                        const cpx<T>* __restrict__ s1, const cpx<T>* __restrict__ s2,
                                                        cpx<TD>* __restrict__ d0, cpx<TD>* __restrict__ d1,
                                                        cpx<TD>* __restrict__ d2) {
    __shared__ cpx<T> tile[32][33];  // [m][kx]
    const int mtn = (g.L + 31) / 32, fld = blockIdx.z / mtn, m0 = (blockIdx.z - fld * mtn) * 32;
    const cpx<T>* src = fld == 0 ? s0 : (fld == 1 ? s1 : s2);
    cpx<TD>* dst = fld == 0 ? d0 : (fld == 1 ? d1 : d2);
    if (!src) return;  // the whole block
    const int kyc = blockIdx.y, row = kept_row_of(g, kyc), kx0 = blockIdx.x * 32, mox = g.M / 2 + 1;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + ty + 8 * i, kx = kx0 + tx;
        if (m < g.L && kx < g.KX) tile[ty + 8 * i][tx] = src[((size_t)m * g.N + row) * mox + kx];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int kx = kx0 + ty + 8 * i, m = m0 + tx;
        if (m < g.L && kx < g.KX) {
            const cpx<T> v = tile[tx][ty + 8 * i];
            dst[((size_t)kx * g.NYc + kyc) * g.L + m] = mk((TD)v.x, (TD)v.y);
        }
    }
}

// compact [p][m] -> reference layout, zeros in the removed band (columns kx >= KX and removed rows)
template <typename T, typename TD = T>
__global__ void __launch_bounds__(256) k_compact_to_ref(DevGeom g, const cpx<T>* __restrict__ s0,
                                                        const cpx<T>* __restrict__ s1, const cpx<T>* __restrict__ s2,
                                                        cpx<TD>* __restrict__ d0, cpx<TD>* __restrict__ d1,
                                                        cpx<TD>* __restrict__ d2) {
    __shared__ cpx<TD> tile[32][33];  // [kx][m]
    const int mtn = (g.L + 31) / 32, fld = blockIdx.z / mtn, m0 = (blockIdx.z - fld * mtn) * 32;
    const cpx<T>* src = fld == 0 ? s0 : (fld == 1 ? s1 : s2);
    cpx<TD>* dst = fld == 0 ? d0 : (fld == 1 ? d1 : d2);
    if (!dst) return;  // the whole block
    const int row = blockIdx.y, kyc = kept_index_of(g, row), kx0 = blockIdx.x * 32, mox = g.M / 2 + 1;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int kx = kx0 + ty + 8 * i, m = m0 + tx;
        cpx<TD> v = mk(TD(0), TD(0));
        if (kyc >= 0 && kx < g.KX && m < g.L) {
            const cpx<T> a = src[((size_t)kx * g.NYc + kyc) * g.L + m];
            v = mk((TD)a.x, (TD)a.y);
        }
        tile[ty + 8 * i][tx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + ty + 8 * i, kx = kx0 + tx;
        if (m < g.L && kx < mox) dst[((size_t)m * g.N + row) * mox + kx] = tile[tx][ty + 8 * i];
    }
}

// ---------------------------------------------------------------------------------------
// Initial states on the kept modes (the removed band is never stored: the stepper dealiases on
// entry, implicit_explicit_stepping.cu:40), one thread per compact (p, m), blockIdx.y = field.
//   kind 1 (random): theta = amp (2u - 1) + i amp (2u' - 1), u, u' uniform from SplitMix64 on
//          (seed, reference-layout index): the distribution of init_random (init.cu:9-20,
//          348-365), whose cuRAND XORWOW stream cannot be reproduced; f, g, F, G zero.
//   kind 2 (test): init_test_device (init.cu:263-301): theta and f at ky = 0, kx != 0 hold
//          1e-6 sin(kx) in the real part, g = 0, F_m = 1e-6 cos(m), G = 0, in the real type T.
//   kind 0: zeros.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t init_mix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename T>
__global__ void __launch_bounds__(256) k_init_state(DevGeom g, int kind, uint64_t seed, double amp,
                                                    cpx<T>* __restrict__ X0, cpx<T>* __restrict__ X1,
                                                    cpx<T>* __restrict__ X2, cpx<T>* __restrict__ XF,
                                                    cpx<T>* __restrict__ XG) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int fld = blockIdx.y;
    if (fld == 0 && e < (size_t)g.L) {  // mean flow
        const int m = (int)e;
        XF[m] = mk(kind == 2 ? T(1e-6) * cos(T(m)) : T(0), T(0));
        XG[m] = mk(T(0), T(0));
    }
    if (e >= (size_t)g.P * g.L) return;
    const int m = (int)(e % g.L), p = (int)(e / g.L);
    const int kx = p / g.NYc, kyc = p - kx * g.NYc;
    const int a = (g.NYc - 1) / 2;
    const int row = kyc <= a ? kyc : g.N - (g.NYc - kyc);
    cpx<T>* X = fld == 0 ? X0 : (fld == 1 ? X1 : X2);
    cpx<T> v = mk(T(0), T(0));
    if (kind == 1 && fld == 0) {
        const uint64_t idx = ((uint64_t)m * g.N + row) * (uint64_t)(g.M / 2 + 1) + kx;
        const uint64_t h0 = init_mix64(seed * 0x100000001B3ull + 2 * idx);
        const uint64_t h1 = init_mix64(seed * 0x100000001B3ull + 2 * idx + 1);
        const double u0 = (double)(h0 >> 11) * (1.0 / 9007199254740992.0);
        const double u1 = (double)(h1 >> 11) * (1.0 / 9007199254740992.0);
        v = mk(T(amp * (2.0 * u0 - 1.0)), T(amp * (2.0 * u1 - 1.0)));
    } else if (kind == 2 && fld != 2 && row == 0 && kx != 0) {
        v = mk(T(1e-6) * sin(T(kx)), T(0));
    }
    X[e] = v;
}

// force_mean (util/util.cu:74-81): F_0 = mean_F + 0 i
template <typename T>
__global__ void k_force_mean(cpx<T>* XF, double v) {
    if (threadIdx.x == 0) XF[0] = mk(T(v), T(0));
}

// ---------------------------------------------------------------------------------------
// Penalization (PENALIZED builds, nonlinear_operator_rayleigh_noslip.cu:1097-1343).  The mask
// chi(x, y) does not depend on z, so every sum over the source mode j of a summand commutes with
// the mask product and the transforms: the summands need, per output mode k, the transforms
// H = R2C(chi u / (M N)) of ten mode-space fields u (the reference forms six per j and sums
// after the transforms):
//   0: i kx a_k, 1: i ky a_k                 a_k = sum_j Iac(j, k) f_j   (f summands 1, 3)
//   2: i ky b_k + bF_k d_k0,                 b_k = sum_j Ia2(k, j) g_j, bF_k, bG_k from F, G
//   3: -i kx b_k + bG_k d_k0                                             (f summands 2, 4)
//   4: k^2 f_k                                                           (f summand 5)
//   5: i kx c_k, 6: i ky c_k                 c_k = sum_j Ia2(j, k) f_j   (g 1, 3; F, G 1)
//   7: i ky g_k + F_k d_k0, 8: -i kx g_k + G_k d_k0   (u12_k, u22_k: g 2, 4; F, G 2)
//   9: theta_k                                                           (theta)
// (d_k0: at kx = ky = 0 only, add_dc_part).  After the mask products and the transforms:
//   N_theta += -(1/(2 eta)) H9
//   N_f     += -(1/eta) (i kx (H0 + H2) + i ky (H1 + H3) + k^2 H4)
//   N_g     += -(1/eta) (-i ky H5 + i kx H6 + i ky (H7 - H8) / 2)   (summand 4 is d_y, as there)
//   N_F     += -(1/eta) (-H5 + H7 / 2) at k = 0,  N_G += -(1/eta) (-H6 + H8 / 2) at k = 0
// which are the reference's summands scaled by -1 (theta) and -1/Pr (:1366-1370).
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_pen_fields(DevGeom g, const T* __restrict__ iac, const T* __restrict__ ia2,
                                                    const cpx<T>* __restrict__ Xt, const cpx<T>* __restrict__ Xf,
                                                    const cpx<T>* __restrict__ Xg, const cpx<T>* __restrict__ XF,
                                                    const cpx<T>* __restrict__ XG, const T* __restrict__ kxa,
                                                    const T* __restrict__ kya, cpx<T>* __restrict__ out) {
    const int L = g.L, P = g.P;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)L * P) return;
    const int k = (int)(e / P), p = (int)(e % P);
    const cpx<T>* f = Xf + (size_t)p * L;
    const cpx<T>* gg = Xg + (size_t)p * L;
    cpx<T> a = mk(T(0), T(0)), b = a, c = a, bF = a, bG = a;
    for (int j = 0; j < L; ++j) {
        const T cac = iac[j + (size_t)k * L], ckj = ia2[k + (size_t)j * L], cjk = ia2[j + (size_t)k * L];
        const cpx<T> fj = f[j], gj = gg[j];
        a = a + cscale(fj, cac);
        b = b + cscale(gj, ckj);
        c = c + cscale(fj, cjk);
        if (p == 0) {
            bF = bF + cscale(XF[j], ckj);
            bG = bG + cscale(XG[j], ckj);
        }
    }
    const T kx = kxa[p], ky = kya[p];
    const cpx<T> fk = f[k], gk = gg[k];
    cpx<T> v[10];
    v[0] = imul(a, kx);
    v[1] = imul(a, ky);
    v[2] = imul(b, ky);
    v[3] = imul(b, -kx);
    v[4] = cscale(fk, kx * kx + ky * ky);
    v[5] = imul(c, kx);
    v[6] = imul(c, ky);
    v[7] = imul(gk, ky);
    v[8] = imul(gk, -kx);
    v[9] = Xt[(size_t)p * L + k];
    if (p == 0) {
        v[2] = v[2] + bF;
        v[3] = v[3] + bG;
        v[7] = v[7] + XF[k];
        v[8] = v[8] + XG[k];
    }
    const size_t fs = (size_t)L * P;
#pragma unroll
    for (int q = 0; q < 10; ++q) out[q * fs + e] = v[q];
}

// one (y row, mode k, field group) per block: the group's one or two hybrid rows as one complex
// M-point transform (X1 + i X2, the kx = 0 imaginary parts dropped as a C2R does), the product
// with chi(x, y) / (M N), the forward transform and the unpacking to kx < KX.  Groups pair
// fields of like size: (0, 1), (2, 3), (5, 6), (7, 8), (4), (9).
template <typename T>
__global__ void __launch_bounds__(256) k_rows_mask(DevGeom g, const T* __restrict__ mask,
                                                   const cpx<T>* __restrict__ hin, const cpx<T>* __restrict__ twM,
                                                   cpx<T>* __restrict__ hout) {
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* a = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* bb = a + g.M;
    const int M = g.M, N = g.N, KX = g.KX, KXp = g.KXp, L = g.L;
    const int y = blockIdx.x, k = blockIdx.y, grp = blockIdx.z;
    const int f1 = grp < 4 ? (grp == 0 ? 0 : grp == 1 ? 2 : grp == 2 ? 5 : 7) : (grp == 4 ? 4 : 9);
    const int f2 = grp < 4 ? f1 + 1 : -1;
    const size_t fstride = (size_t)L * N * KXp;
    const size_t roff = ((size_t)k * N + y) * KXp;
    for (int e = threadIdx.x; e < M; e += blockDim.x) a[e] = mk(T(0), T(0));
    __syncthreads();
    for (int kk = threadIdx.x; kk < KX; kk += blockDim.x) {
        cpx<T> x1 = hin[(size_t)f1 * fstride + roff + kk];
        cpx<T> x2 = f2 >= 0 ? hin[(size_t)f2 * fstride + roff + kk] : mk(T(0), T(0));
        if (kk == 0) {
            x1.y = T(0);
            x2.y = T(0);
        }
        a[kk] = mk(x1.x - x2.y, x1.y + x2.x);                  // X1 + i X2
        if (kk > 0) a[M - kk] = mk(x1.x + x2.y, -x1.y + x2.x);  // conj(X1) + i conj(X2)
    }
    __syncthreads();
    cpx<T>* r = lds_fft(a, bb, g.logM, 1, twM, true);
    const T s = T(1) / (T(M) * T(N));
    for (int x = threadIdx.x; x < M; x += blockDim.x) r[x] = cscale(r[x], s * mask[(size_t)y * M + x]);
    __syncthreads();
    cpx<T>* R = lds_fft(r, r == a ? bb : a, g.logM, 1, twM, false);
    for (int kk = threadIdx.x; kk < KX; kk += blockDim.x) {
        const cpx<T> zk = R[kk], zm = R[(M - kk) & (M - 1)];
        hout[(size_t)f1 * fstride + roff + kk] = mk(T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y));
        if (f2 >= 0) hout[(size_t)f2 * fstride + roff + kk] = mk(T(0.5) * (zk.y + zm.y), T(-0.5) * (zk.x - zm.x));
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_pen_combine(DevGeom g, T ie, const cpx<T>* __restrict__ H,
                                                     const T* __restrict__ kxa, const T* __restrict__ kya,
                                                     cpx<T>* __restrict__ Nt, cpx<T>* __restrict__ Nf,
                                                     cpx<T>* __restrict__ Ng, cpx<T>* __restrict__ NF,
                                                     cpx<T>* __restrict__ NG) {
    const int L = g.L, P = g.P;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)L * P) return;
    const int p = (int)(e / L), k = (int)(e % L);
    const size_t fs = (size_t)L * P, at = (size_t)k * P + p;
    cpx<T> h[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) h[q] = H[q * fs + at];
    const T kx = kxa[p], ky = kya[p];
    const cpx<T> ft = imul(h[0] + h[2], kx) + imul(h[1] + h[3], ky) + cscale(h[4], kx * kx + ky * ky);
    const cpx<T> gt = imul(h[5], -ky) + imul(h[6], kx) + imul(h[7] - h[8], T(0.5) * ky);
    Nt[e] = Nt[e] + cscale(h[9], T(-0.5) * ie);
    Nf[e] = Nf[e] + cscale(ft, -ie);
    Ng[e] = Ng[e] + cscale(gt, -ie);
    if (p == 0) {
        NF[k] = NF[k] + cscale(cscale(h[7], T(0.5)) - h[5], -ie);
        NG[k] = NG[k] + cscale(cscale(h[8], T(0.5)) - h[6], -ie);
    }
}

// ---------------------------------------------------------------------------------------
// Plane operators on the state's mode planes (operator/freeze_operator.cu, masktemp.cu,
// masking.cu): each mode plane m of theta (and f, g for the freeze) goes to physical space
// (C2R), is combined point by point under an integer mask, scaled by 1/(M N) and transformed
// back (R2C):
//   PLANE_CAPTURE  store the physical plane (freeze_mask, the frozen copy)
//   PLANE_FREEZE   mask ? u : frozen                  (freeze_operator apply_mask)
//   PLANE_HEAT     mask ? u : u + 1.5/100             (masktemp apply_mask)
//   PLANE_SCALE    mask ? u : u * factor              (masking julie_do_the_thing)
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_modes_to_planes(DevGeom g, int nf, const cpx<T>* __restrict__ X0,
                                                         const cpx<T>* __restrict__ X1,
                                                         const cpx<T>* __restrict__ X2, cpx<T>* __restrict__ out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (e >= (size_t)g.L * g.P || f >= nf) return;
    const int m = (int)(e / g.P), p = (int)(e % g.P);
    const cpx<T>* X = f == 0 ? X0 : (f == 1 ? X1 : X2);
    out[(size_t)f * g.L * g.P + e] = X[(size_t)p * g.L + m];
}

template <typename T>
__global__ void __launch_bounds__(256) k_planes_to_modes(DevGeom g, int nf, const cpx<T>* __restrict__ in,
                                                         cpx<T>* __restrict__ X0, cpx<T>* __restrict__ X1,
                                                         cpx<T>* __restrict__ X2) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (e >= (size_t)g.L * g.P || f >= nf) return;
    const int p = (int)(e / g.L), m = (int)(e % g.L);
    cpx<T>* X = f == 0 ? X0 : (f == 1 ? X1 : X2);
    X[e] = in[((size_t)f * g.L + m) * g.P + p];
}

// one (y row, mode m, field) per block; g.nq = L (the mode planes stand where the levels do)
template <typename T>
__global__ void __launch_bounds__(256) k_rows_planeop(DevGeom g, int op, const int* __restrict__ imask, T factor,
                                                      T* __restrict__ frozen, const cpx<T>* __restrict__ hin,
                                                      const cpx<T>* __restrict__ twM, cpx<T>* __restrict__ hout) {
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* a = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* bb = a + g.M;
    const int M = g.M, N = g.N, KX = g.KX, KXp = g.KXp, L = g.L;
    const int y = blockIdx.x, m = blockIdx.y, f = blockIdx.z;
    const size_t roff = (((size_t)f * L + m) * N + y) * KXp;
    for (int e = threadIdx.x; e < M; e += blockDim.x) a[e] = mk(T(0), T(0));
    __syncthreads();
    for (int kk = threadIdx.x; kk < KX; kk += blockDim.x) {
        cpx<T> x = hin[roff + kk];
        if (kk == 0) x.y = T(0);
        a[kk] = x;
        if (kk > 0) a[M - kk] = conjg(x);
    }
    __syncthreads();
    cpx<T>* r = lds_fft(a, bb, g.logM, 1, twM, true);  // the plane's row: real parts
    T* fr = frozen + (((size_t)f * L + m) * N + y) * M;
    if (op == PLANE_CAPTURE) {
        for (int x = threadIdx.x; x < M; x += blockDim.x) fr[x] = r[x].x;
        return;
    }
    const T s = T(1) / (T(M) * T(N));
    for (int x = threadIdx.x; x < M; x += blockDim.x) {
        const T u = r[x].x;
        T v = u;
        if (op == PLANE_SET) {
            // mode plane 0 of field 0 becomes the given real plane (frozen = [row][col]) under the
            // unnormalised R2C of util/init.cu fft(); the other planes pass through
            if (m == 0 && f == 0) v = frozen[(size_t)y * M + x] * (T(M) * T(N));
        } else if (imask[(size_t)y * M + x] != 1) {
            if (op == PLANE_FREEZE) v = fr[x];
            else if (op == PLANE_HEAT) v = u + T(1.5) / T(100);
            else v = u * factor;
        }
        r[x] = mk(s * v, T(0));
    }
    __syncthreads();
    cpx<T>* R = lds_fft(r, r == a ? bb : a, g.logM, 1, twM, false);
    for (int kk = threadIdx.x; kk < KX; kk += blockDim.x) hout[roff + kk] = R[kk];
}

// ---------------------------------------------------------------------------------------
// Initial patterns of theta (util/init.cu, main.cu:326-352): one real plane, amplitude amp, on
// the num_x = M by num_y = N grid, point (i, j) at x = i / M, y = j / N (fractions of the box),
// evaluated in FP64.  kind: 1 rectangle (init_rectangle: x_mode = mode, y_mode 0; the columns
// i < M / 8 hold uniform noise in [-amp, amp]), 2 dislocation (init_dislocation, x_mode = mode),
// 3 isr (init_isr, x_mode = mode, one random phase), 4 sv (init_sv, y_mode = mode, x_frac 0.2),
// 5 cr (init_cr, y_mode = mode, x_frac 0.8), 6 giantspiral (init_giantspiral, mode).  The
// reference draws the noise and the phase from the C library's rand(); here they come from
// SplitMix64 on (seed, point) as the random start does.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double init_uniform(uint64_t seed, uint64_t idx) {
    const uint64_t h = init_mix64(seed * 0x100000001B3ull + idx);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

template <typename T>
__global__ void __launch_bounds__(256) k_init_pattern(int M, int N, int kind, int mode, double amp, uint64_t seed,
                                                      T* __restrict__ plane) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)M * N) return;
    const int i = (int)(e % M), j = (int)(e / M);
    const double twopi = 2.0 * M_PI, dx = 1.0 / M, dy = 1.0 / N, x = i * dx, y = j * dy;
    double v = 0.0;
    if (kind == 1) {  // init.cu:70-100 (x_mode = mode, y_mode = 0)
        if (i < M / 8) v = (2.0 * init_uniform(seed, 2 * e) - 1.0) * amp;
        else v = amp * cos(twopi * (i * ((double)mode / M)));
    } else if (kind == 2) {  // init.cu:22-68: two opposite dislocations on the line x = 1/2 + dx/2
        const double x1 = (M / 2.0) * dx + dx / 2.0, y1 = (N / 4.0) * dy + dy / 2.0;
        const double y2 = (3.0 * N / 4.0) * dy + dy / 2.0;
        v = amp * cos(mode * twopi * x - atan2(y - y1, x - x1) + atan2(y - y2, x - x1));
    } else if (kind == 3) {  // init.cu:102-129 (x_mode = mode, y_mode = 0)
        const double phase = twopi * init_uniform(seed, (uint64_t)-1);
        v = amp * cos(twopi * (i * ((double)mode / M)) + phase);
    } else if (kind == 4) {  // init.cu:131-165 (x_frac = 0.2)
        const double xf = 0.2, xl = 0.5 * (1.0 - xf), xr = 1.0 - xl;
        double ph = -0.5 * sin(twopi * y) - x * x;
        if (x > xr) ph += 1.0;
        else if (x > xl) ph += 0.5 * (1.0 - cos(M_PI * (x - xl) / xf));
        v = amp * cos(twopi * (j * ((double)mode / N) + ph));
    } else if (kind == 5) {  // init.cu:167-213 (x_frac = 0.8)
        const double xf = 0.8, x1 = 0.5 * ((1.0 - xf) * M + 1) * dx, x2 = 1.0 - x1, y1 = 0.5 * (N + 1) * dy;
        v = amp * cos(mode * twopi * y - atan2(y - y1, x - x1) + atan2(y - y1, x - x2));
    } else if (kind == 6) {  // init.cu:215-242: centre (0.5, 0.5), second core (0.8, 0.5)
        const double r = sqrt((x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5));
        v = amp * cos(twopi * mode * r - atan2(y - 0.5, x - 0.5) + atan2(y - 0.5, x - 0.8));
    }
    plane[e] = (T)v;
}

template <typename T>
void Launch<T>::init_pattern(const DevGeom& g, int kind, int mode, double amp, uint64_t seed, void* plane,
                             hipStream_t s) {
    const size_t n = (size_t)g.M * g.N;
    hipLaunchKernelGGL(k_init_pattern<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g.M, g.N, kind, mode,
                       amp, seed, (T*)plane);
}

// temperature at one height from the physical mode planes of theta: sum_i S_i theta_i(x, y),
// accumulated in mode order (calculate_temperature_operator.cu:56-68)
template <typename T>
__global__ void __launch_bounds__(256) k_temp_sum(int L, size_t nxy, const T* __restrict__ planes,
                                                  const float* __restrict__ S, float* __restrict__ out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nxy) return;
    T acc = T(0);
    for (int i = 0; i < L; ++i) acc += T(S[i]) * planes[(size_t)i * nxy + e];
    out[e] = (float)acc;
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
// through a per-wave LDS scratch whose slots are laid out to spread the banks (below).
// =======================================================================================
// 512-point FFT of one wavefront.  tw = table exp(-2 pi i k / 512), k < 512 (LDS or global).
// The scratch element of internal index i is sb[Map::at(i) * S] (S = 1 for a private buffer,
// S = tile row stride when a column of an LDS tile doubles as the scratch).
//   Pad8Map: one pad slot every 8 (576 slots).
//   XorMap: bits 0-2 ^= bits 3-5 and bits 3-5 ^= bits 6-8, a permutation of the 512 slots.  With
//   8-byte slots it puts the 16 (32) lanes served per LDS cycle on distinct banks for all three
//   exchange patterns (8 j + r, j + 64 r, 64 (j / 8) + j % 8 + 8 r); the pad-8 layout leaves
//   the contiguous j + 64 r reads 2-way conflicted.
// Each map also gives the slots of the three exchange patterns of lane j (j < 64) for r < 8:
// col(j, r) = at(j + 64 r), row(j, r) = at(8 j + r), mid(j, r) = at(64 (j / 8) + j % 8 + 8 r).
struct Pad8Map {
    static constexpr int slots = 576;
    __device__ static __forceinline__ int at(int i) { return pad8(i); }
    __device__ static __forceinline__ int col(int j, int r) { return pad8(j + 64 * r); }
    __device__ static __forceinline__ int row(int j, int r) { return pad8(8 * j + r); }
    __device__ static __forceinline__ int mid(int j, int r) { return pad8((j >> 3) * 64 + (j & 7) + 8 * r); }
};
// XorMap's pattern slots as one lane constant XOR an immediate (the fields that r occupies are
// disjoint from the ones it is XORed into), so each costs one VALU instruction:
//   col = (8 (j / 8) + (j % 8 ^ j / 8)) ^ 8 r, plus 64 r;
//   row = (64 (j / 8) + 8 (j % 8 ^ j / 8) + j % 8) ^ r;
//   mid = (72 (j / 8) + j % 8) ^ 9 r.
struct XorMap {
    static constexpr int slots = 512;
    __device__ static __forceinline__ int at(int i) { return i ^ ((((i >> 6) & 7) << 3) | ((i >> 3) & 7)); }
    __device__ static __forceinline__ int col(int j, int r) {
        return ((8 * (j >> 3) + ((j & 7) ^ (j >> 3))) ^ (8 * r)) + 64 * r;
    }
    __device__ static __forceinline__ int row(int j, int r) {
        return ((64 * (j >> 3) + 8 * ((j & 7) ^ (j >> 3)) + (j & 7)) ^ r);
    }
    __device__ static __forceinline__ int mid(int j, int r) { return (72 * (j >> 3) + (j & 7)) ^ (9 * r); }
};
// Twiddle tables.  Tw512: tw[k] = exp(-2 pi i k / 512), k < 512.  TwByPass: the two passes' reads
// laid out contiguously, tw[64 r + j] = w^(j r) (pass 3, lane j) and tw[512 + 8 r + j] = w^(8 j r)
// (pass 2, lane group j = lane % 8), 576 entries: each read touches consecutive slots, where
// the strided reads of Tw512 share 2 to 4 banks.
struct Tw512 {
    static constexpr int size = 512;
    __device__ static __forceinline__ int pass2(int lane, int r) { return (lane & 7) * r * 8; }
    __device__ static __forceinline__ int pass3(int lane, int r) { return lane * r; }
};
struct TwByPass {
    static constexpr int size = 576;
    __device__ static __forceinline__ int pass2(int lane, int r) { return 512 + 8 * r + (lane & 7); }
    __device__ static __forceinline__ int pass3(int lane, int r) { return 64 * r + lane; }
    // the Tw512 exponent k of entry e
    __device__ static __forceinline__ int source(int e) {
        return e < 512 ? (e & 63) * (e >> 6) : 8 * (e & 7) * ((e - 512) >> 3);
    }
};
template <typename T, bool INV, int S = 1, typename Map = Pad8Map, typename Tw = Tw512>
__device__ __forceinline__ void fft512_wave(cpx<T> v[8], cpx<T>* sb, const cpx<T>* tw, int lane) {
    dft8<T, INV>(v);
#pragma unroll
    for (int r = 0; r < 8; ++r) sb[Map::row(lane, r) * S] = v[r];
    wave_sync();
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = sb[Map::col(lane, r) * S];
#pragma unroll
    for (int r = 1; r < 8; ++r) {
        cpx<T> w = tw[Tw::pass2(lane, r)];
        if (INV) w.y = -w.y;
        v[r] = cmul(v[r], w);
    }
    dft8<T, INV>(v);
    wave_sync();
#pragma unroll
    for (int r = 0; r < 8; ++r) sb[Map::mid(lane, r) * S] = v[r];
    wave_sync();
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = sb[Map::col(lane, r) * S];
#pragma unroll
    for (int r = 1; r < 8; ++r) {
        cpx<T> w = tw[Tw::pass3(lane, r)];
        if (INV) w.y = -w.y;
        v[r] = cmul(v[r], w);
    }
    dft8<T, INV>(v);
    wave_sync();  // the scratch may be reused right after
}

// Kept columns at M = 512 and kept rows at N = 512 (the dealias rule of make_geometry: kx <= 171,
// |ky| <= 170); the 512-point kernels take them as constants so that the zero band and the
// mirrored half resolve at compile time (the launchers check them).
constexpr int KX512 = kept_cols(512), NYC512 = kept_rows(512);
static_assert(KX512 == 172 && NYC512 == 341, "dealias rule at 512 points");

// Hybrid rows of the 512 x 512 kernels, tile-major: plane (field, level) = [KXT512 / 8 tiles][512
// y][8 kx], so that the y-FFT blocks (8 kx columns each) read and write one contiguous 32 KiB
// tile, and the x-rows read 64-byte pieces of a row (the other half of each 128-byte line is
// the neighbouring row, run on the same XCD).  Element (y, kx) of a plane:
constexpr int KXT512 = (KX512 + 7) & ~7;  // 176 columns, 22 tiles
__device__ __forceinline__ int hyb512(int y, int kx) { return (kx >> 3) * (512 * 8) + y * 8 + (kx & 7); }

// x-rows at M = N = 512 for the fused forward y-FFT: three waves per (level, y) row.  Phase 1:
// wave pr inverse-transforms the pair (2 pr, 2 pr + 1) of level fields.  Phase 2: the six
// real product fields h = uu + vv - 2 ww, d = uu - vv, uv, uw, vw, u.grad(theta) go back in
// pairs, one pair per wave, scaled by 1/(M N).  They are all the level combinations need:
// kx^2 Puu + ky^2 Pvv - k^2 Pww = (k^2 / 2) h + ((kx^2 - ky^2) / 2) d and Puu - Pvv = d, with
// d formed before any rounding of the larger ww (which would swamp uu - vv).  The second
// field of every pair is scaled to the size of the first for the transform (pair_scale).
__global__ void __launch_bounds__(192) k_rows512t(DevGeom g, int nblk, int per, const cpx<float>* __restrict__ hin,
                                                  const cpx<float>* __restrict__ twp, cpx<float>* __restrict__ hout) {
    __shared__ cpx<float> tw[TwByPass::size];
    __shared__ cpx<float> zs[3][XorMap::slots];
    // one (level, y) row per block; consecutive rows (which share the 128-byte lines of the
    // tile-major layout) are consecutive logical blocks and so run on one XCD
    const int lb = xcd_logical_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int lane = threadIdx.x & 63, pr = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int KX = KX512;
    const int nq = g.nq;
    const int y = lb % 512, l = lb / 512;
    const size_t fstride = (size_t)nq * 512 * KXT512;
    const size_t roff = (size_t)l * 512 * KXT512 + (size_t)y * 8;
    // twp: the twiddles already in TwByPass order (built at create), so the copy is plain: the
    // index arithmetic of TwByPass::source on every block cost 1.7 % of the kernel
    for (int e = threadIdx.x; e < TwByPass::size; e += 192) tw[e] = twp[e];
    __syncthreads();
    {
        const cpx<float>* x1 = hin + (size_t)(2 * pr) * fstride + roff;
        const cpx<float>* x2 = hin + (size_t)(2 * pr + 1) * fstride + roff;
        cpx<float> av[4], bv[4];  // n = lane + 64 r < KX <= 192 for r < 3
        float ma = 0.f, mb = 0.f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int n = lane + 64 * r;
            av[r] = bv[r] = mk(0.f, 0.f);
            if (n < KX) {
                av[r] = x1[hyb512(0, n)];
                bv[r] = x2[hyb512(0, n)];
                if (n == 0) {
                    av[r].y = 0.f;
                    bv[r].y = 0.f;
                }
            }
            ma = fmaxf(ma, fmaxf(fabsf(av[r].x), fabsf(av[r].y)));
            mb = fmaxf(mb, fmaxf(fabsf(bv[r].x), fabsf(bv[r].y)));
        }
        const float sg = pair_scale(wave_max(ma), wave_max(mb)), isg = inv_pow2(sg);
        // spectrum n < KX from (av, bv) at r = n / 64; n > 512 - KX mirrors 512 - n < KX,
        // which sits in lane (512 - n) % 64 of slot (512 - n) / 64 of another lane: read it
        // from LDS after a staging pass
        cpx<float>* st = zs[pr];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int n = lane + 64 * r;
            if (n < KX) {
                st[2 * n] = av[r];
                st[2 * n + 1] = mk(sg * bv[r].x, sg * bv[r].y);
            }
        }
        wave_sync();
        cpx<float> v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int n = lane + 64 * r;
            cpx<float> z = mk(0.f, 0.f);
            if (n < KX) {
                const cpx<float> a = st[2 * n], b = st[2 * n + 1];
                z = mk(a.x - b.y, a.y + b.x);  // X1 + i s X2
            } else if (n > 512 - KX) {
                const cpx<float> a = st[2 * (512 - n)], b = st[2 * (512 - n) + 1];
                z = mk(a.x + b.y, -a.y + b.x);  // conj(X1) + i s conj(X2)
            }
            v[r] = z;
        }
        wave_sync();  // the staging area becomes the transform's scratch
        fft512_wave<float, true, 1, XorMap, TwByPass>(v, zs[pr], tw, lane);
#pragma unroll
        for (int r = 0; r < 8; ++r) zs[pr][XorMap::col(lane, r)] = mk(v[r].x, isg * v[r].y);
    }
    __syncthreads();
    const float s = 1.f / (512.f * 512.f);
    cpx<float> p[8];
    float ma = 0.f, mb = 0.f;
    // pr is uniform over the wave: one branch around the whole slot loop, not one per slot
    if (pr == 0) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int i = XorMap::col(lane, r);
            const cpx<float> z0 = zs[0][i], z1 = zs[1][i];
            const float u = z0.x, v = z0.y, w = z1.x;
            p[r] = mk(s * (u * u + v * v - 2.f * w * w), s * (u * u - v * v));
        }
    } else if (pr == 1) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int i = XorMap::col(lane, r);
            const cpx<float> z0 = zs[0][i], z1 = zs[1][i];
            const float u = z0.x, v = z0.y, w = z1.x;
            p[r] = mk(s * u * v, s * u * w);
        }
    } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int i = XorMap::col(lane, r);
            const cpx<float> z0 = zs[0][i], z1 = zs[1][i], z2 = zs[2][i];
            const float u = z0.x, v = z0.y, w = z1.x, tx = z1.y, ty = z2.x, tz = z2.y;
            p[r] = mk(s * v * w, s * (u * tx + v * ty + w * tz));
        }
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        ma = fmaxf(ma, fabsf(p[r].x));
        mb = fmaxf(mb, fabsf(p[r].y));
    }
    const float sg = pair_scale(wave_max(ma), wave_max(mb)), isg = inv_pow2(sg);
#pragma unroll
    for (int r = 0; r < 8; ++r) p[r].y *= sg;
    __syncthreads();  // every wave has read the slots before they become scratch
    fft512_wave<float, false, 1, XorMap, TwByPass>(p, zs[pr], tw, lane);
#pragma unroll
    for (int r = 0; r < 8; ++r) zs[pr][XorMap::col(lane, r)] = p[r];
    wave_sync();
    cpx<float>* o1 = hout + (size_t)(2 * pr) * fstride + roff;
    cpx<float>* o2 = hout + (size_t)(2 * pr + 1) * fstride + roff;
    // bins k = lane + 64 r < KX (r < 3) and their mirrors (512 - k) % 512
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int k = lane + 64 * r;
        if (r < 2 || k < KX) {
            const cpx<float> zk = zs[pr][XorMap::col(lane, r)], zm = zs[pr][XorMap::at((512 - k) & 511)];
            o1[hyb512(0, k)] = mk(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
            o2[hyb512(0, k)] = mk(0.5f * isg * (zk.y + zm.y), -0.5f * isg * (zk.x - zm.x));
        }
    }
}

// x-rows at M = 64, 128, 256 (FP32): one group of W = M / 8 lanes per (level, y) row holding
// the row's three pair transforms in natural layout (fft_group), so the products form in
// registers between the inverse and the forward transforms; 64 / W rows per wave, WPB waves per
// block.  Same arithmetic as k_rows512t: both members of every input and product pair are
// balanced by a power of two (pair_scale over the row, a shuffle reduction over the group).
template <int LOGM, int WPB>
__device__ __forceinline__ void k_rows_w_body(DevGeom g, const cpx<float>* __restrict__ hin,
                                              const cpx<float>* __restrict__ twM, cpx<float>* __restrict__ hout) {
    constexpr int M = 1 << LOGM, W = M / 8, G = 64 / W, RPB = WPB * G, SP = M + M / 8, KX = kept_cols(M);
    __shared__ cpx<float> tw[M];
    __shared__ cpx<float> scr[RPB][SP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, grp = lane / W, j = lane - grp * W;
    const int KXp = g.KXp, nq = g.nq, N = g.N;
    const int y = blockIdx.x * RPB + wave * G + grp, l = blockIdx.y;
    const size_t fstride = (size_t)nq * N * KXp;
    const size_t roff = ((size_t)l * N + y) * KXp;
    for (int e = threadIdx.x; e < M; e += blockDim.x) tw[e] = twM[e];
    __syncthreads();
    cpx<float>* sb = scr[wave * G + grp];
    auto gmax = [](float x) { return group_max<W>(x); };
    cpx<float> v[3][8];
    float isg[3];
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {
        const cpx<float>* x1 = hin + (size_t)(2 * pr) * fstride + roff;
        const cpx<float>* x2 = hin + (size_t)(2 * pr + 1) * fstride + roff;
        cpx<float> a[8], b[8];
        float ma = 0.f, mb = 0.f;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int k = j + W * r;
            const int src = k < KX ? k : (k > M - KX ? M - k : -1);
            a[r] = b[r] = mk(0.f, 0.f);
            if (src >= 0) {
                a[r] = x1[src];
                b[r] = x2[src];
                if (src == 0) {  // a C2R ignores the imaginary part at kx = 0
                    a[r].y = 0.f;
                    b[r].y = 0.f;
                }
            }
            ma = fmaxf(ma, fmaxf(fabsf(a[r].x), fabsf(a[r].y)));
            mb = fmaxf(mb, fmaxf(fabsf(b[r].x), fabsf(b[r].y)));
        }
        const float sg = pair_scale(gmax(ma), gmax(mb));
        isg[pr] = inv_pow2(sg);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int k = j + W * r;
            const cpx<float> bs = mk(sg * b[r].x, sg * b[r].y);
            if (k < KX) v[pr][r] = mk(a[r].x - bs.y, a[r].y + bs.x);               // X1 + i s X2
            else if (k > M - KX) v[pr][r] = mk(a[r].x + bs.y, -a[r].y + bs.x);     // conj(X1) + i s conj(X2)
            else v[pr][r] = mk(0.f, 0.f);
        }
    }
    fft_group<LOGM, true, 3>(v, sb, tw, j);
    // products at x = j + W r, in place, scaled by 1 / (M N)
    const float s = 1.f / ((float)M * (float)N);
    float pa[3] = {0.f, 0.f, 0.f}, pb[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const float u = v[0][r].x, vv = isg[0] * v[0][r].y, w = v[1][r].x, tx = isg[1] * v[1][r].y;
        const float ty = v[2][r].x, tz = isg[2] * v[2][r].y;
        v[0][r] = mk(s * (u * u + vv * vv - 2.f * w * w), s * (u * u - vv * vv));
        v[1][r] = mk(s * u * vv, s * u * w);
        v[2][r] = mk(s * vv * w, s * (u * tx + vv * ty + w * tz));
#pragma unroll
        for (int pr = 0; pr < 3; ++pr) {
            pa[pr] = fmaxf(pa[pr], fabsf(v[pr][r].x));
            pb[pr] = fmaxf(pb[pr], fabsf(v[pr][r].y));
        }
    }
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {
        const float sg = pair_scale(gmax(pa[pr]), gmax(pb[pr]));
        isg[pr] = inv_pow2(sg);
#pragma unroll
        for (int r = 0; r < 8; ++r) v[pr][r].y *= sg;
    }
    fft_group<LOGM, false, 3>(v, sb, tw, j);
    // unpack X1_k = (Z_k + conj Z_{M-k}) / 2, X2_k = -i (Z_k - conj Z_{M-k}) / (2 s); Z_{M-k} of
    // another lane comes through the scratch
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {
#pragma unroll
        for (int r = 0; r < 8; ++r) sb[pad8(j + W * r)] = v[pr][r];
        wave_sync();
        cpx<float>* o1 = hout + (size_t)(2 * pr) * fstride + roff;
        cpx<float>* o2 = hout + (size_t)(2 * pr + 1) * fstride + roff;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int k = j + W * r;
            if (k < KX) {
                const cpx<float> zk = v[pr][r], zm = sb[pad8((M - k) & (M - 1))];
                o1[k] = mk(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
                o2[k] = mk(0.5f * isg[pr] * (zk.y + zm.y), -0.5f * isg[pr] * (zk.x - zm.x));
            }
        }
        wave_sync();  // the scratch is rewritten by the next pair
    }
}

template <int LOGM, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_rows_w(DevGeom g, const cpx<float>* __restrict__ hin,
                                                const cpx<float>* __restrict__ twM, cpx<float>* __restrict__ hout) {
    k_rows_w_body<LOGM, WPB>(g, hin, twM, hout);
}
// one member's arguments of k_rows_w_batch, which runs the k_rows_w body for member
// blockIdx.z / zdim of a batch (the rest of blockIdx.z is the body's own z index)
struct k_rows_w_args {
    DevGeom g;
    const cpx<float>* hin;
    const cpx<float>* twM;
    cpx<float>* hout;
};
template <int LOGM, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_rows_w_batch(const k_rows_w_args* __restrict__ tab, int zdim) {
    const int mb = blockIdx.z / zdim;
    const k_rows_w_args& a = tab[mb];
    k_rows_w_body<LOGM, WPB>(a.g, a.hin, a.twM, a.hout);
}


// y inverse at N = 512: a block of 8 waves handles YT = 8 kx columns of one (field, level) at a
// time (a tile), one column per wave.  The LDS tile [XorMap::at(y)][YT + 1] holds the results and
// doubles as the FFT scratch of each column; it is written out row-wise (64-byte segments).
// The blocks stay resident: the tiles are split into 8 contiguous ranges, one per XCD (block b
// runs on XCD b % 8), and the bpx blocks of an XCD walk their range together, so that the tiles
// in flight on an XCD are neighbours (their output rows share cache lines).  Each wave loads
// its column of the next tile before it transforms the current one.
constexpr int YT = 8;
constexpr int YS = YT + 1;
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6))) k_fft_y_inv512(DevGeom g, int ntile, int nblk, int bpx,
                                                      const cpx<float>* __restrict__ lev,
                                                      const cpx<float>* __restrict__ tw,
                                                      cpx<float>* __restrict__ hin) {
    __shared__ cpx<float> tile[XorMap::slots * YS];
    __shared__ cpx<float> twl[TwByPass::size];  // in LDS, so that no table entry stays in registers
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int NYc = NYC512;
    constexpr int a = (NYc - 1) / 2, shift = 512 - NYc;  // kept rows [0, a] U [512 - a, 511]
    const int P = g.P, KX = g.KX;
    for (int e = threadIdx.x; e < TwByPass::size; e += 512) twl[e] = tw[e];  // tw: in TwByPass order
    lds_barrier();
    const int range = (nblk + 7) / 8, xcd = blockIdx.x & 7;
    const int end = min(nblk, (xcd + 1) * range);
    auto load = [&](int lb, cpx<float> (&v)[8]) {
        const int tix = lb % ntile, rest = lb / ntile;
        const int kx = tix * YT + wave;
        const cpx<float>* src = lev + (size_t)rest * P + (size_t)kx * NYc;  // rest = field * nq + l
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int n = lane + 64 * r;
            v[r] = mk(0.f, 0.f);
            if (kx < KX) {
                if (n <= a) v[r] = src[n];
                else if (n >= 512 - a) v[r] = src[n - shift];
            }
        }
    };
    int lb = xcd * range + (blockIdx.x >> 3);
    cpx<float> nxt[8];
    if (lb < end) load(lb, nxt);
    for (; lb < end; lb += bpx) {
        cpx<float> v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = nxt[r];
        if (lb + bpx < end) load(lb + bpx, nxt);
        fft512_wave<float, true, YS, XorMap, TwByPass>(v, tile + wave, twl, lane);
#pragma unroll
        for (int r = 0; r < 8; ++r) tile[XorMap::col(lane, r) * YS + wave] = v[r];
        lds_barrier();
        const int tix = lb % ntile, rest = lb / ntile;
        // the tile is one contiguous 32 KiB piece of the tile-major plane; every lane stores on
        // every pass (the columns past KX are the transforms of zero columns), so that the number
        // of stores after the prefetch is fixed and the prefetch wait leaves them in flight
        cpx<float>* dst = hin + (size_t)rest * 512 * KXT512 + (size_t)tix * 512 * YT;
#pragma unroll
        for (int i = 0; i < YT; ++i) {
            const int e = threadIdx.x + 512 * i, y = e / YT;
            dst[e] = tile[XorMap::at(y) * YS + e % YT];
        }
        lds_barrier();  // the tile columns are the next tile's FFT scratch
    }
}

// =======================================================================================
// Vertical transforms on the matrix cores (FP32 storage).  MM = 0: v_mfma_f32_32x32x2f32
// (K = 2 per instruction); MM = 1: BF16 hi/lo split, v_mfma_f32_32x32x16_bf16 (K = 16),
// a*b ~ ah*bh + ah*bl + al*bh with FP32 accumulation.
// Operand layouts (wave64): A (32 x K) lane l holds row l%32, k = KPL*(l/32) + e;
// B (K x 32) lane l holds column l%32, k = KPL*(l/32) + e; D (32 x 32) lane l, register r
// holds column l%32, row 8*(r/4) + 4*(l/32) + r%4.  KPL = K/2 elements per lane.
// The A operands (basis tables) come pre-arranged per (table, tile, k-step, lane) by the host.
// =======================================================================================
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// FP32 -> BF16, round to nearest even (one v_cvt_pk_bf16_f32 per pair on gfx950)
__device__ __forceinline__ __bf16 bf16_rne(float x) { return (__bf16)x; }
__device__ __forceinline__ float bf16_to_f32(__bf16 h) {
    return __uint_as_float((unsigned)__builtin_bit_cast(unsigned short, h) << 16);
}

template <int MM>
struct Mma;
template <>
struct Mma<0> {
    static constexpr int K = 2, KPL = 1;
    __device__ static __forceinline__ void run(f32x16& c, const float* a, const float* b) {
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(a[0], b[0], c, 0, 0, 0);
    }
};
template <>
struct Mma<1> {
    static constexpr int K = 16, KPL = 8;
    __device__ static __forceinline__ void run(f32x16& c, const float* a, const float* b) {
        bf16x8 ah, al, bh, bl;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            ah[e] = bf16_rne(a[e]);
            al[e] = bf16_rne(a[e] - bf16_to_f32(ah[e]));
            bh[e] = bf16_rne(b[e]);
            bl[e] = bf16_rne(b[e] - bf16_to_f32(bh[e]));
        }
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, c, 0, 0, 0);
    }
};

// Evaluation on the matrix cores: D[level][column] = sum_m A_t[level][m] B_f[m][column], with
// columns = (mode, re/im) of 16 modes per wave and 4 tables C', C, S, S' applied to f, f, g,
// theta, theta (5 chains).  Block = 64 modes (4 waves); the block's mode data is staged in
// LDS as [field][m][128 columns].  evA: [NLT][NKS][64][4 tables] floats for FP32 (KPL = 1, one
// float4 per lane and k-step), [4][NLT][NKS][64][KPL] for BF16X3; ALDS: the block's A operands are
// copied to LDS with the modes (when both fit in 64 KiB).
template <int MM, bool ALDS>
__device__ __forceinline__ void k_eval_mma_body(DevGeom g, int nlt, int nks, int per,
                                                const cpx<float>* __restrict__ Xt,
                                                const cpx<float>* __restrict__ Xf, const cpx<float>* __restrict__ Xg,
                                                const cpx<float>* __restrict__ XF, const cpx<float>* __restrict__ XG,
                                                const float* __restrict__ evA, const float* __restrict__ evtab,
                                                const float* __restrict__ kxa, const float* __restrict__ kya,
                                                cpx<float>* __restrict__ lev) {
    using MT = Mma<MM>;
    extern __shared__ float sX[];  // 3 * 128 * (Lp + 1) modes, then sA: the block's A operands
    const int L = g.L, P = g.P, nq = g.nq;
    const int Lp = nks * MT::K;
    // logical block = (64-mode block, level tile); the level tiles of a mode block are logical
    // neighbours and so share one XCD's L2 copy of the staged modes
    const int lb = xcd_logical_block(blockIdx.x, per);
    if (lb >= ((P + 63) / 64) * nlt) return;
    const int p0 = (lb / nlt) * 64, lt_block = lb % nlt;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // LDS layout [field][column][LP] with column = 2 mode + re/im and an odd row length LP, so
    // that both the staging writes (m fastest) and the operand reads (column fastest) spread
    // over the banks
    const int LP = Lp + 1;
    // A operands of the block's level tile, [4][nks][64][KPL], copied next to the modes
    float* sA = sX + 3 * 128 * LP;
    if (ALDS) {
        if constexpr (MM == 0) {  // the tile's interleaved tables are one contiguous piece
            const float4* src = reinterpret_cast<const float4*>(evA) + (size_t)lt_block * nks * 64;
            for (int e = threadIdx.x; e < nks * 64; e += 256) reinterpret_cast<float4*>(sA)[e] = src[e];
        } else {
            const int na = nks * 64 * MT::KPL / 4;  // float4 per table
            for (int e = threadIdx.x; e < 4 * na; e += 256) {
                const int t = e / na, i = e - t * na;
                reinterpret_cast<float4*>(sA)[e] =
                    reinterpret_cast<const float4*>(evA + ((size_t)t * nlt + lt_block) * nks * 64 * MT::KPL)[i];
            }
        }
    }
    // mean flow u, v (sum_m C_m F_m, sum_m C_m G_m) at the block's 32 levels, for mode p = 0
    __shared__ cpx<float> smf[2][32];
    if (p0 == 0 && threadIdx.x < 64) {
        const int ll = threadIdx.x & 31, which = threadIdx.x >> 5, l = lt_block * 32 + ll;
        const cpx<float>* Xm = which ? XG : XF;
        cpx<float> sum = mk(0.f, 0.f);
        if (l < nq) {
            const float* tb = evtab + (size_t)l * 4 * L + 2 * L;
#pragma unroll 4
            for (int m = 0; m < L; ++m) sum = sum + cscale(Xm[m], tb[m]);
        }
        smf[which][ll] = sum;
    }
    // four elements per thread in flight: the loads of a round are all issued before its stores
    constexpr int SU = 4;
    for (int e0 = threadIdx.x; e0 < 64 * Lp; e0 += 256 * SU) {
        cpx<float> zt[SU], zf[SU], zg[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int e = e0 + 256 * u, mode = e / Lp, m = e - mode * Lp;
            const int p = p0 + mode;
            zt[u] = zf[u] = zg[u] = mk(0.f, 0.f);
            if (e < 64 * Lp && p < P && m < L) {
                const size_t gi = (size_t)p * L + m;
                zt[u] = Xt[gi];
                zf[u] = Xf[gi];
                zg[u] = Xg[gi];
            }
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int e = e0 + 256 * u, mode = e / Lp, m = e - mode * Lp;
            if (e < 64 * Lp) {
                sX[(0 * 128 + 2 * mode) * LP + m] = zt[u].x;
                sX[(0 * 128 + 2 * mode + 1) * LP + m] = zt[u].y;
                sX[(1 * 128 + 2 * mode) * LP + m] = zf[u].x;
                sX[(1 * 128 + 2 * mode + 1) * LP + m] = zf[u].y;
                sX[(2 * 128 + 2 * mode) * LP + m] = zg[u].x;
                sX[(2 * 128 + 2 * mode + 1) * LP + m] = zg[u].y;
            }
        }
    }
    __syncthreads();
    const int j = lane & 31, half = lane >> 5;
    const int col = wave * 32 + j;
    const int p = p0 + (col >> 1), comp = col & 1;
    const bool pvalid = p < P;
    const int pc = pvalid ? p : 0;
    const float kx = kxa[pc], ky = kya[pc], k2 = kx * kx + ky * ky;
    const size_t fs = (size_t)nq * P;
    {
        const int lt = lt_block;
        f32x16 cFp = {}, cFc = {}, cGs = {}, cTs = {}, cTz = {};
        for (int s = 0; s < nks; ++s) {
            float a[4][MT::KPL], bt[MT::KPL], bf[MT::KPL], bg[MT::KPL];
            if constexpr (MM == 0) {  // FP32: one float4 holds the lane's value of the four tables
                const float4 av = ALDS ? reinterpret_cast<const float4*>(sA)[s * 64 + lane]
                                       : reinterpret_cast<const float4*>(evA)[((size_t)lt * nks + s) * 64 + lane];
                a[0][0] = av.x;
                a[1][0] = av.y;
                a[2][0] = av.z;
                a[3][0] = av.w;
            } else {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float* src = ALDS ? sA + (((size_t)t * nks + s) * 64 + lane) * MT::KPL
                                            : evA + ((((size_t)t * nlt + lt) * nks + s) * 64 + lane) * MT::KPL;
#pragma unroll
                    for (int e = 0; e < MT::KPL; ++e) a[t][e] = src[e];
                }
            }
#pragma unroll
            for (int e = 0; e < MT::KPL; ++e) {
                const int m = s * MT::K + half * MT::KPL + e;
                bt[e] = sX[(0 * 128 + col) * LP + m];
                bf[e] = sX[(1 * 128 + col) * LP + m];
                bg[e] = sX[(2 * 128 + col) * LP + m];
            }
            MT::run(cFp, a[0], bf);  // C' . f
            MT::run(cFc, a[1], bf);  // C  . f
            MT::run(cGs, a[2], bg);  // S  . g
            MT::run(cTs, a[2], bt);  // S  . theta
            MT::run(cTz, a[3], bt);  // S' . theta
        }
        // epilogue: the partner component of the same mode sits in lane ^ 1 (a DPP swap of
        // adjacent lanes, no LDS); i k multiplies the real parts by -1 times the partner's
        // imaginary part (sg = -1) and the imaginary parts by the partner's real part (sg = 1)
        const float sg = comp ? 1.f : -1.f;
        auto partner = [](float x) {
            return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, true));
        };
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int l = lt * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
            const float oFp = partner(cFp[r]), oGs = partner(cGs[r]), oTs = partner(cTs[r]);
            float u = sg * (kx * oFp + ky * oGs);
            float v = sg * (ky * oFp - kx * oGs);
            const float w = k2 * cFc[r], tx = sg * (kx * oTs), ty = sg * (ky * oTs), tz = cTz[r];
            if (p == 0 && l < nq) {  // mean flow at k = 0 (add_dc_part)
                const cpx<float> mF = smf[0][l - lt * 32], mG = smf[1][l - lt * 32];
                u += comp ? mF.y : mF.x;
                v += comp ? mG.y : mG.x;
            }
            if (pvalid && l < nq) {
                float* o = reinterpret_cast<float*>(lev) + ((size_t)l * P + p) * 2 + comp;
                o[0 * 2 * fs] = u;
                o[1 * 2 * fs] = v;
                o[2 * 2 * fs] = w;
                o[3 * 2 * fs] = tx;
                o[4 * 2 * fs] = ty;
                o[5 * 2 * fs] = tz;
            }
        }
    }
}

template <int MM, bool ALDS>
__global__ void __launch_bounds__(256) k_eval_mma(DevGeom g, int nlt, int nks, int per,
                                                  const cpx<float>* __restrict__ Xt,
                                                  const cpx<float>* __restrict__ Xf, const cpx<float>* __restrict__ Xg,
                                                  const cpx<float>* __restrict__ XF, const cpx<float>* __restrict__ XG,
                                                  const float* __restrict__ evA, const float* __restrict__ evtab,
                                                  const float* __restrict__ kxa, const float* __restrict__ kya,
                                                  cpx<float>* __restrict__ lev) {
    k_eval_mma_body<MM, ALDS>(g, nlt, nks, per, Xt, Xf, Xg, XF, XG, evA, evtab, kxa, kya, lev);
}
// one member's arguments of k_eval_mma_batch, which runs the k_eval_mma body for member
// blockIdx.z / zdim of a batch (the rest of blockIdx.z is the body's own z index)
struct k_eval_mma_args {
    DevGeom g;
    int nlt;
    int nks;
    int per;
    const cpx<float>* Xt;
    const cpx<float>* Xf;
    const cpx<float>* Xg;
    const cpx<float>* XF;
    const cpx<float>* XG;
    const float* evA;
    const float* evtab;
    const float* kxa;
    const float* kya;
    cpx<float>* lev;
};
template <int MM, bool ALDS>
__global__ void __launch_bounds__(256) k_eval_mma_batch(const k_eval_mma_args* __restrict__ tab, int zdim) {
    const int mb = blockIdx.z / zdim;
    const k_eval_mma_args& a = tab[mb];
    k_eval_mma_body<MM, ALDS>(a.g, a.nlt, a.nks, a.per, a.Xt, a.Xf, a.Xg, a.XF, a.XG, a.evA, a.evtab, a.kxa, a.kya,
                              a.lev);
}


// Projection on the matrix cores: Out[m][column] = sum_l T_t[m][l] D_t[l][column]; the B
// operands are float components of the five level combinations (column = (mode, re/im), so a
// lane reads one float and the lanes of a wave read 128 contiguous bytes per level).
// prA: [5][NKS][64][KPL].  Block = WPB x KS waves: WPB groups of 16 modes (32 columns, rows
// m < L <= 32); the KS waves of a group split the level sum into contiguous k-step ranges and
// add their partial sums through LDS in fixed order (KS > 1 gives small grids more waves in
// flight).
template <int MM, int WPB, int KS>
__device__ __forceinline__ void k_project_mma_body(DevGeom g, int nks, const float* __restrict__ Dfl,
                                                   const cpx<float>* __restrict__ fgw, const float* __restrict__ prA,
                                                   const float* __restrict__ prtab, const float* __restrict__ kxa,
                                                   const float* __restrict__ kya, float Pr,
                                                   cpx<float>* __restrict__ Nt, cpx<float>* __restrict__ Nf,
                                                   cpx<float>* __restrict__ Ng, cpx<float>* __restrict__ NF,
                                                   cpx<float>* __restrict__ NG) {
    using MT = Mma<MM>;
    constexpr int OS = 66;               // staging row (2L <= 64 floats plus 2 to spread banks)
    // the FP32 form of the large grids stages one field at a time: 17 KB of LDS per block and
    // (with the hint on the kernel) 125 registers, so that 4 blocks fit a CU and a 512^2 x 32
    // projection (917 blocks) has all its blocks resident at once
    constexpr int NST = (MM == 0 && WPB == 4) ? 1 : 3;
    __shared__ float sO[WPB][NST][16 * OS];  // per wave: [field][mode][OS] staging for the write-out
    __shared__ f32x16 sPart[KS > 1 ? (KS - 1) * WPB : 1][KS > 1 ? 4 : 1][64];  // partials of waves wk > 0
    const int L = g.L, P = g.P, nq = g.nq;
    const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) % WPB, wk = (threadIdx.x >> 6) / WPB;
    const int j = lane & 31, half = lane >> 5;
    const int col = wave * 32 + j;
    const int p = blockIdx.x * 16 * WPB + (col >> 1), comp = col & 1;
    const bool pvalid = p < P;
    const int pc = pvalid ? p : 0;
    const float kx = kxa[pc], ky = kya[pc], k2 = kx * kx + ky * ky;
    const size_t ffs = (size_t)2 * nq * P;  // floats per combination
    f32x16 cF1 = {}, cF2 = {}, cG = {}, cT = {};
    const int spw = (nks + KS - 1) / KS, sbeg = wk * spw, send = min(nks, sbeg + spw);
    // Operands through buffer loads: a descriptor per level combination (its 2 nq P floats; a
    // level past nq, whose table entries are zero, lies outside the range and reads as 0) and
    // one for the tables, so that each load is a 32-bit offset that grows by a constant per
    // k-step, with no 64-bit address arithmetic.  Byte offsets stay below 2^31 (rbc_create
    // keeps the matrix-core forms only for grids where 2 nq P floats fit).
    __amdgpu_buffer_rsrc_t rD[5];
#pragma unroll
    for (int t = 0; t < 5; ++t)
        rD[t] = __builtin_amdgcn_make_buffer_rsrc((void*)(Dfl + t * ffs), (short)0, (int)(ffs * 4), 0x00020000);
    const __amdgpu_buffer_rsrc_t rA =
        __builtin_amdgcn_make_buffer_rsrc((void*)prA, (short)0, 5 * nks * 64 * MT::KPL * 4, 0x00020000);
    const int dbase = (2 * pc + comp) * 4, dstep = 2 * P * 4;  // bytes: column, one level
    auto fetch = [&](int s, float (&a)[5][MT::KPL], float (&b)[5][MT::KPL]) {
#pragma unroll
        for (int e = 0; e < MT::KPL; ++e) {
            const int off = dbase + (s * MT::K + half * MT::KPL + e) * dstep;
#pragma unroll
            for (int t = 0; t < 5; ++t) b[t][e] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rD[t], off, 0, 0));
        }
#pragma unroll
        for (int t = 0; t < 5; ++t) {
            const int off = ((t * nks + s) * 64 + lane) * MT::KPL * 4;
#pragma unroll
            for (int e = 0; e < MT::KPL; ++e) a[t][e] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rA, off + 4 * e, 0, 0));
        }
    };
    auto step = [&](const float (&a)[5][MT::KPL], const float (&b)[5][MT::KPL]) {
        MT::run(cF1, a[0], b[0]);  // wC'  . D0
        MT::run(cF1, a[1], b[1]);  // wC'' . D1
        MT::run(cF2, a[2], b[1]);  // wC   . D1   (times k^2 below)
        MT::run(cG, a[3], b[2]);   // wS   . D2
        MT::run(cG, a[4], b[3]);   // wS'  . D3
        MT::run(cT, a[3], b[4]);   // wS   . D4
    };
    // FP32: batches of U k-steps, the next batch's operands (tables and data, issued together)
    // loading while the current batch runs on the matrix cores (loads return in issue order,
    // so a batch waits for its own only).  Two buffers taken in turn over a loop unrolled twice,
    // so that no buffer is copied (a copy would wait for its loads).  A load past the wave's
    // range is harmless and unused: a step of another wave's range, or past the end of a
    // descriptor, which reads as 0.
    constexpr int U = 2;
    int s = sbeg;
    if constexpr (MM == 0) {
        float a0[U][5][MT::KPL], b0[U][5][MT::KPL], a1[U][5][MT::KPL], b1[U][5][MT::KPL];
        auto load = [&](int s0, float (&a)[U][5][MT::KPL], float (&b)[U][5][MT::KPL]) {
#pragma unroll
            for (int u = 0; u < U; ++u) fetch(s0 + u, a[u], b[u]);
        };
        auto run = [&](const float (&a)[U][5][MT::KPL], const float (&b)[U][5][MT::KPL]) {
#pragma unroll
            for (int u = 0; u < U; ++u) step(a[u], b[u]);
        };
        load(s, a0, b0);
        for (; s + 2 * U <= send; s += 2 * U) {
            __builtin_amdgcn_sched_barrier(0);
            load(s + U, a1, b1);
            __builtin_amdgcn_sched_barrier(0);
            run(a0, b0);
            __builtin_amdgcn_sched_barrier(0);
            load(s + 2 * U, a0, b0);
            __builtin_amdgcn_sched_barrier(0);
            run(a1, b1);
        }
        if (s + U <= send) {  // one more whole batch, already loaded
            run(a0, b0);
            s += U;
        }
    }
    for (; s < send; ++s) {
        float a[5][MT::KPL], b[5][MT::KPL];
        fetch(s, a, b);
        step(a, b);
    }
    if (KS > 1) {
        if (wk > 0) {
            f32x16* dst = sPart[(wk - 1) * WPB + wave][0];
            dst[0 * 64 + lane] = cF1;
            dst[1 * 64 + lane] = cF2;
            dst[2 * 64 + lane] = cG;
            dst[3 * 64 + lane] = cT;
        }
        __syncthreads();
        if (wk > 0) return;
#pragma unroll
        for (int k = 0; k < KS - 1; ++k) {
            const f32x16* src = sPart[k * WPB + wave][0];
            cF1 += src[0 * 64 + lane];
            cF2 += src[1 * 64 + lane];
            cG += src[2 * 64 + lane];
            cT += src[3 * 64 + lane];
        }
    }
    const float mt = -1.f, mp = -1.f / Pr;
    const int lm = j >> 1;  // local mode 0..15 within the wave
    const int pw0 = blockIdx.x * 16 * WPB + wave * 16;
    // 16 modes x L complex per field written contiguously ([p][m] layout) through the wave's
    // staging rows: all three fields at once, or (NST = 1) one after another, not unrolled, so
    // that one field's values are live at a time
    if constexpr (NST == 3) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int m = 8 * (r >> 2) + 4 * half + (r & 3);
            if (m < L) {
                sO[wave][0][lm * OS + 2 * m + comp] = mt * cT[r];
                sO[wave][1][lm * OS + 2 * m + comp] = mp * (cF1[r] + k2 * cF2[r]);
                sO[wave][2][lm * OS + 2 * m + comp] = mp * cG[r];
            }
        }
        wave_sync();
        for (int e = lane; e < 16 * 2 * L; e += 64) {
            const int lmode = e / (2 * L), q = e - lmode * 2 * L;
            const int pp = pw0 + lmode;
            if (pp < P) {
                reinterpret_cast<float*>(Nt)[(size_t)pp * 2 * L + q] = sO[wave][0][lmode * OS + q];
                reinterpret_cast<float*>(Nf)[(size_t)pp * 2 * L + q] = sO[wave][1][lmode * OS + q];
                reinterpret_cast<float*>(Ng)[(size_t)pp * 2 * L + q] = sO[wave][2][lmode * OS + q];
            }
        }
    } else {
        float* so = sO[wave][0];
#pragma unroll 1
        for (int fld = 0; fld < 3; ++fld) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = 8 * (r >> 2) + 4 * half + (r & 3);
                if (m < L)
                    so[lm * OS + 2 * m + comp] =
                        fld == 0 ? mt * cT[r] : (fld == 1 ? mp * (cF1[r] + k2 * cF2[r]) : mp * cG[r]);
            }
            wave_sync();
            float* out = reinterpret_cast<float*>(fld == 0 ? Nt : (fld == 1 ? Nf : Ng));
            for (int e = lane; e < 16 * 2 * L; e += 64) {
                const int lmode = e / (2 * L), q = e - lmode * 2 * L;
                const int pp = pw0 + lmode;
                if (pp < P) out[(size_t)pp * 2 * L + q] = so[lmode * OS + q];
            }
            wave_sync();  // the rows are read before the next field is staged
        }
    }
    mean_flow_projection(g, fgw, prtab, Pr, NF, NG);
}

template <int MM, int WPB, int KS>
__global__ void __launch_bounds__(64 * WPB * KS) __attribute__((amdgpu_waves_per_eu(MM == 0 && WPB == 4 ? 4 : 1)))
k_project_mma(DevGeom g, int nks, const float* __restrict__ Dfl,
                                                     const cpx<float>* __restrict__ fgw, const float* __restrict__ prA,
                                                     const float* __restrict__ prtab, const float* __restrict__ kxa,
                                                     const float* __restrict__ kya, float Pr,
                                                     cpx<float>* __restrict__ Nt, cpx<float>* __restrict__ Nf,
                                                     cpx<float>* __restrict__ Ng, cpx<float>* __restrict__ NF,
                                                     cpx<float>* __restrict__ NG) {
    k_project_mma_body<MM, WPB, KS>(g, nks, Dfl, fgw, prA, prtab, kxa, kya, Pr, Nt, Nf, Ng, NF, NG);
}
// one member's arguments of k_project_mma_batch, which runs the k_project_mma body for member
// blockIdx.z / zdim of a batch (the rest of blockIdx.z is the body's own z index)
struct k_project_mma_args {
    DevGeom g;
    int nks;
    const float* Dfl;
    const cpx<float>* fgw;
    const float* prA;
    const float* prtab;
    const float* kxa;
    const float* kya;
    float Pr;
    cpx<float>* Nt;
    cpx<float>* Nf;
    cpx<float>* Ng;
    cpx<float>* NF;
    cpx<float>* NG;
};
template <int MM, int WPB, int KS>
__global__ void __launch_bounds__(64 * WPB * KS) __attribute__((amdgpu_waves_per_eu(MM == 0 && WPB == 4 ? 4 : 1)))
k_project_mma_batch(const k_project_mma_args* __restrict__ tab, int zdim) {
    const int mb = blockIdx.z / zdim;
    const k_project_mma_args& a = tab[mb];
    k_project_mma_body<MM, WPB, KS>(a.g, a.nks, a.Dfl, a.fgw, a.prA, a.prtab, a.kxa, a.kya, a.Pr, a.Nt, a.Nf, a.Ng,
                                    a.NF, a.NG);
}


// Forward y-FFT fused with the level combinations, by product group:
//   group 0: h, d, uv -> D0, D2     group 1: uw, vw -> D1, D3     group 2: u.grad -> D4
//   group 3: groups 1 and 2 in one pass (uw, vw, u.grad -> D1, D3, D4)
// At N = 512 a block of 8 waves handles the 8 kx columns of one tile of one level, for group 0
// or group 3; the products of the group are transformed in turn (tile in LDS, one column per
// wave) and the group's combinations of the kept rows accumulate in registers; the next
// product's tile is read into registers while the current one is transformed.  The groups with
// uw and vw also write the k = 0 values of P_uw and P_vw (fgw) for the mean flow.
template <int GRP>
struct YGroup;
template <>
// product fields of k_rows512t: 0 h = uu + vv - 2 ww, 1 d = uu - vv, 2 uv, 3 uw, 4 vw, 5 u.grad(theta)
// h, d, uv -> D0 = (k^2 / 2) h + ((kx^2 - ky^2) / 2) d + 2 kx ky uv, D2 = -kx ky d + (kx^2 - ky^2) uv
struct YGroup<0> {
    static constexpr int n = 3, t0 = 0, t1 = 2, t2 = -1;
    __device__ static int q(int i) { return i; }
    __device__ static void add(int i, cpx<float> V, float kx, float ky, cpx<float>& d0, cpx<float>& d1, cpx<float>&) {
        if (i == 0) { d0 = d0 + cscale(V, 0.5f * (kx * kx + ky * ky)); }
        else if (i == 1) { d0 = d0 + cscale(V, 0.5f * (kx * kx - ky * ky)); d1 = d1 + cscale(V, -kx * ky); }
        else { d0 = d0 + cscale(V, 2.f * kx * ky); d1 = d1 + cscale(V, kx * kx - ky * ky); }
    }
};
template <>
struct YGroup<1> {  // uw, vw -> D1, D3
    static constexpr int n = 2, t0 = 1, t1 = 3, t2 = -1;
    __device__ static int q(int i) { return 3 + i; }
    __device__ static void add(int i, cpx<float> V, float kx, float ky, cpx<float>& d0, cpx<float>& d1, cpx<float>&) {
        if (i == 0) { d0 = d0 + imul(V, kx); d1 = d1 - imul(V, ky); }
        else { d0 = d0 + imul(V, ky); d1 = d1 + imul(V, kx); }
    }
};
template <>
struct YGroup<2> {  // u.grad(theta) -> D4
    static constexpr int n = 1, t0 = 4, t1 = -1, t2 = -1;
    __device__ static int q(int) { return 5; }
    __device__ static void add(int, cpx<float> V, float, float, cpx<float>& d0, cpx<float>&, cpx<float>&) {
        d0 = d0 + V;
    }
};
template <>
struct YGroup<3> {  // groups 1 and 2 in one pass: uw, vw, u.grad(theta) -> D1, D3, D4
    static constexpr int n = 3, t0 = 1, t1 = 3, t2 = 4;
    __device__ static int q(int i) { return 3 + i; }
    __device__ static void add(int i, cpx<float> V, float kx, float ky, cpx<float>& d0, cpx<float>& d1,
                               cpx<float>& d2) {
        if (i == 0) { d0 = d0 + imul(V, kx); d1 = d1 - imul(V, ky); }
        else if (i == 1) { d0 = d0 + imul(V, ky); d1 = d1 + imul(V, kx); }
        else { d2 = d2 + V; }
    }
};

// One block of the fused forward y-FFT: logical block lb = (level, 8-column tile) of group GRP.
// tile: the block's LDS tile (XorMap::slots * (CT + 1) complex); twl: LDS room for the
// per-pass twiddle table (used by groups of several products).
template <int GRP, int CT>
__device__ __forceinline__ void y_fwd512c_body(int lb, const DevGeom& g, int ntile,
                                               const cpx<float>* __restrict__ hout,
                                               const cpx<float>* __restrict__ tw,
                                               const float* __restrict__ kxa, const float* __restrict__ kya,
                                               cpx<float>* __restrict__ D, cpx<float>* __restrict__ fgw,
                                               cpx<float>* tile, cpx<float>* twl) {
    using G = YGroup<GRP>;
    constexpr int CS = CT + 1;  // tile row stride
    // groups of several products read the twiddles from LDS: the barrier between products keeps
    // the compiler from holding the table entries in registers across the transforms
    constexpr bool TWL = G::n > 1;
    if (TWL)
        for (int e = threadIdx.x; e < TwByPass::size; e += 64 * CT) twl[e] = tw[e];  // tw: in TwByPass order
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int NYc = NYC512;
    const int P = g.P, nq = g.nq, KX = g.KX;
    const int tix = lb % ntile, l = lb / ntile;
    const int kx0 = tix * CT;
    constexpr int a = (NYc - 1) / 2, shift = 512 - NYc;
    const int ncol = min(CT, KX - kx0);
    const int c = wave, kxc = kx0 + c;
    const bool cvalid = kxc < KX;
    int kyc[8];
    float kyv[8];
    const float kxv = cvalid ? kxa[(size_t)kxc * NYc] : 0.f;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int n = lane + 64 * r;
        kyc[r] = (n <= a) ? n : ((n >= 512 - a) ? n - shift : -1);
        kyv[r] = (cvalid && kyc[r] >= 0) ? kya[(size_t)kxc * NYc + kyc[r]] : 0.f;
    }
    cpx<float> d0[8], d1[8], d2[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) d0[r] = d1[r] = d2[r] = mk(0.f, 0.f);
    // tile element e = threadIdx.x + 64 CT i (i < 8): y = e / CT, column e % CT
    cpx<float> pre[8];
    // the block's 8 columns are one contiguous tile of the tile-major plane
    static_assert(CT == 8, "one 8-column tile of the hybrid layout per block");
    auto fetch = [&](int q) {
        const cpx<float>* src = hout + ((size_t)q * nq + l) * (size_t)512 * KXT512 + (size_t)tix * 512 * CT;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = threadIdx.x + 64 * CT * i, cc = e % CT;
            pre[i] = (cc < ncol) ? src[e] : mk(0.f, 0.f);
        }
    };
    fetch(G::q(0));
#pragma unroll
    for (int iq = 0; iq < G::n; ++iq) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = threadIdx.x + 64 * CT * i, y = e / CT, cc = e - y * CT;
            tile[XorMap::at(y) * CS + cc] = pre[i];
        }
        __syncthreads();
        if (iq + 1 < G::n) fetch(G::q(iq + 1));
        cpx<float> v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = tile[XorMap::col(lane, r) * CS + c];
        if constexpr (TWL) fft512_wave<float, false, CS, XorMap, TwByPass>(v, tile + c, twl, lane);
        else fft512_wave<float, false, CS, XorMap, TwByPass>(v, tile + c, tw, lane);
#pragma unroll
        for (int r = 0; r < 8; ++r) G::add(iq, v[r], kxv, kyv[r], d0[r], d1[r], d2[r]);
        if ((GRP == 1 || GRP == 3) && iq < 2 && kxc == 0 && lane == 0) fgw[2 * l + iq] = v[0];  // P_uw, P_vw, row 0
        __syncthreads();  // the tile is rewritten for the next product
    }
    if (cvalid) {
        const size_t fs = (size_t)nq * P;
#pragma unroll
        for (int r = 0; r < 8; ++r)
            if (kyc[r] >= 0) {
                cpx<float>* o = D + (size_t)l * P + (size_t)kxc * NYc + kyc[r];
                o[G::t0 * fs] = d0[r];
                if (G::t1 >= 0) o[G::t1 * fs] = d1[r];
                if (G::t2 >= 0) o[G::t2 * fs] = d2[r];
            }
    }
}

// Both groups of three products in one launch (group 0: h, d, uv -> D0, D2; group 3: uw, vw,
// u.grad(theta) -> D1, D3, D4): even logical blocks are group 0, odd ones group 3, so that every
// XCD (a contiguous range of logical blocks) carries both groups, whose traffic differs (306
// and 353 MB at 512^2 x 32).
// At 116-120 registers two 8-wave blocks fit a CU, so one launch of a group is 4.1 rounds of
// blocks; one grid over both leaves a single partial round instead of two and no drain between.
template <int CT>
__global__ void __launch_bounds__(64 * CT) k_fft_y_fwd512c2(DevGeom g, int ntile, int nblk, int per,
                                                        const cpx<float>* __restrict__ hout,
                                                        const cpx<float>* __restrict__ tw,
                                                        const float* __restrict__ kxa,
                                                        const float* __restrict__ kya, cpx<float>* __restrict__ D,
                                                        cpx<float>* __restrict__ fgw) {
    static_assert(YGroup<0>::n > 1 && YGroup<3>::n > 1, "both groups keep the twiddles in LDS");
    __shared__ cpx<float> tile[XorMap::slots * (CT + 1)];
    __shared__ cpx<float> twl[TwByPass::size];
    const int lb = xcd_logical_block(blockIdx.x, per);
    if (lb >= 2 * nblk) return;
    if (lb & 1) y_fwd512c_body<3, CT>(lb >> 1, g, ntile, hout, tw, kxa, kya, D, fgw, tile, twl);
    else y_fwd512c_body<0, CT>(lb >> 1, g, ntile, hout, tw, kxa, kya, D, fgw, tile, twl);
}

// Forward y-FFT at N = 64, 128, 256 fused with the level combinations (FP32): blockIdx.z = the
// combination group of YGroup (0: h, d, uv -> D0, D2; 1: uw, vw -> D1, D3 and the k = 0 values
// of P_uw, P_vw; 2: u.grad(theta) -> D4).  A block of WPB waves handles CB kx columns of one
// level, one group of W = N / 8 lanes per column (fft_group, natural layout); the group's
// products are transformed in turn through the LDS tile and their combinations accumulate in
// registers.  Writes D[t][l][p] into the level buffer (free once the inverse y-FFT has read it).
template <int LOGN, int WPB, int GRP>
__device__ __forceinline__ void y_fwd_group_comb(const DevGeom& g, cpx<float>* tile, const cpx<float>* tw,
                                                 const cpx<float>* __restrict__ hout, const float* __restrict__ kxa,
                                                 const float* __restrict__ kya, cpx<float>* __restrict__ D,
                                                 cpx<float>* __restrict__ fgw) {
    using Gr = YGroup<GRP>;
    constexpr int N = 1 << LOGN, W = N / 8, G = 64 / W, CB = WPB * G, S = CB + 1;
    constexpr int NYc = kept_rows(N), a = (NYc - 1) / 2, shift = N - NYc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, grp = lane / W, j = lane - grp * W;
    const int P = g.P, nq = g.nq, KX = g.KX;
    const int kx0 = blockIdx.x * CB, l = blockIdx.y;
    const int c = wave * G + grp, kx = kx0 + c;
    const bool cvalid = kx < KX;
    const int ncol = min(CB, KX - kx0);
    auto kyc_of = [](int n) { return n <= a ? n : (n >= N - a ? n - shift : -1); };
    const float kxv = cvalid ? kxa[(size_t)kx * NYc] : 0.f;
    float kyv[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int kc = kyc_of(j + W * r);
        kyv[r] = (cvalid && kc >= 0) ? kya[(size_t)kx * NYc + kc] : 0.f;
    }
    cpx<float> d0[8], d1[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) d0[r] = d1[r] = mk(0.f, 0.f);
#pragma unroll
    for (int iq = 0; iq < Gr::n; ++iq) {
        const cpx<float>* src = hout + ((size_t)Gr::q(iq) * nq + l) * (size_t)N * g.KXp + kx0;
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
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            cpx<float> unused;
            Gr::add(iq, v[0][r], kxv, kyv[r], d0[r], d1[r], unused);
        }
        if (GRP == 1 && kx == 0 && j == 0) fgw[2 * l + iq] = v[0][0];  // P_uw, P_vw at (0, 0)
        __syncthreads();  // the tile is refilled by the next product
    }
    if (cvalid) {
        const size_t fs = (size_t)nq * P;
        cpx<float>* o = D + (size_t)l * P + (size_t)kx * NYc;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int kc = kyc_of(j + W * r);
            if (kc >= 0) {
                o[Gr::t0 * fs + kc] = d0[r];
                if (Gr::t1 >= 0) o[Gr::t1 * fs + kc] = d1[r];
            }
        }
    }
}

template <int LOGN, int WPB>
__device__ __forceinline__ void k_fft_y_fwd_wc_body(int bz, DevGeom g, const cpx<float>* __restrict__ hout,
                                                    const cpx<float>* __restrict__ twN,
                                                    const float* __restrict__ kxa, const float* __restrict__ kya,
                                                    cpx<float>* __restrict__ D, cpx<float>* __restrict__ fgw) {
    constexpr int N = 1 << LOGN, W = N / 8, G = 64 / W, CB = WPB * G, S = CB + 1;
    __shared__ cpx<float> tw[N];
    __shared__ cpx<float> tile[(N + N / 8) * S];
    for (int e = threadIdx.x; e < N; e += blockDim.x) tw[e] = twN[e];
    __syncthreads();
    if (bz == 0) y_fwd_group_comb<LOGN, WPB, 0>(g, tile, tw, hout, kxa, kya, D, fgw);
    else if (bz == 1) y_fwd_group_comb<LOGN, WPB, 1>(g, tile, tw, hout, kxa, kya, D, fgw);
    else y_fwd_group_comb<LOGN, WPB, 2>(g, tile, tw, hout, kxa, kya, D, fgw);
}

template <int LOGN, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_fft_y_fwd_wc(DevGeom g, const cpx<float>* __restrict__ hout,
                                                          const cpx<float>* __restrict__ twN,
                                                          const float* __restrict__ kxa, const float* __restrict__ kya,
                                                          cpx<float>* __restrict__ D, cpx<float>* __restrict__ fgw) {
    k_fft_y_fwd_wc_body<LOGN, WPB>(blockIdx.z, g, hout, twN, kxa, kya, D, fgw);
}
// one member's arguments of k_fft_y_fwd_wc_batch, which runs the k_fft_y_fwd_wc body for member
// blockIdx.z / zdim of a batch (the rest of blockIdx.z is the body's own z index)
struct k_fft_y_fwd_wc_args {
    DevGeom g;
    const cpx<float>* hout;
    const cpx<float>* twN;
    const float* kxa;
    const float* kya;
    cpx<float>* D;
    cpx<float>* fgw;
};
template <int LOGN, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_fft_y_fwd_wc_batch(const k_fft_y_fwd_wc_args* __restrict__ tab, int zdim) {
    const int mb = blockIdx.z / zdim;
    const k_fft_y_fwd_wc_args& a = tab[mb];
    k_fft_y_fwd_wc_body<LOGN, WPB>(blockIdx.z - mb * zdim, a.g, a.hout, a.twN, a.kxa, a.kya, a.D, a.fgw);
}


// ---------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------
// RBC_GENERIC_FFT=1 in the environment selects the generic LDS Stockham kernels everywhere (used
// by the tests to pin the 512-point and group-transform kernels against them).
// Blocks of `kernel` (blocks of `threads`, static LDS only) that one XCD of the current device
// holds at once: resident blocks per CU x CUs / 8.
static int resident_blocks_per_xcd(const void* kernel, int threads) {
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess)
        return 32;
    return std::max(1, per_cu * ncu / 8);
}

static bool generic_fft_forced() {
    const char* e = getenv("RBC_GENERIC_FFT");
    return e && e[0] == '1';
}
template <typename T>
static bool fast512_rows(const DevGeom& g) {
    return sizeof(T) == 4 && g.M == 512 && g.N == 512 && g.KX == KX512 && g.NYc == NYC512 && g.KXp == KXT512 &&
           !generic_fft_forced();
}
// the 512-point y and x kernels share the tile-major hybrid layout: both run, or neither
template <typename T>
static bool fast512_y(const DevGeom& g) {
    return fast512_rows<T>(g);
}
// group (natural-layout) transforms for FP32 grids of 64, 128 and 256 points: log2 of the size,
// or 0 when the generic kernels run
template <typename T>
static int wave_rows_log(const DevGeom& g) {
    if (sizeof(T) != 4 || generic_fft_forced() || g.M < 64 || g.M > 256 || g.KX != kept_cols(g.M)) return 0;
    return g.logM;
}
template <typename T>
static int wave_y_log(const DevGeom& g) {
    if (sizeof(T) != 4 || generic_fft_forced() || g.N < 64 || g.N > 256) return 0;
    return g.logN;
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
    if (sizeof(T) == 4 && b.mma >= 0) {
        const int K = b.mma == 0 ? 2 : 16;
        const size_t sx = (size_t)3 * 128 * (b.ev_nks * K + 1) * sizeof(float);
        const size_t sa = (size_t)4 * b.ev_nks * 64 * (K / 2) * sizeof(float);
        // the A operands go to LDS with the modes only while the block stays within 64 KiB: at
        // 66.7 KiB (512^2 x 32; 2 blocks per CU instead of 3) the evaluation took 103 us against 97
        const bool alds = sx + sa <= 65536;
        const size_t smm = alds ? sx + sa : sx;
        const int nblk = ((g.P + 63) / 64) * b.ev_nlt, per = (nblk + 7) / 8;
        dim3 grid(8 * per);
#define RBC_EVAL_MMA(MMV, ALDSV)                                                                                 \
    hipLaunchKernelGGL((k_eval_mma<MMV, ALDSV>), grid, dim3(256), smm, s, g, b.ev_nlt, b.ev_nks, per,             \
                       (const cpx<float>*)X[0], (const cpx<float>*)X[1], (const cpx<float>*)X[2],                \
                       (const cpx<float>*)XF, (const cpx<float>*)XG, (const float*)b.evA, (const float*)b.evtab, \
                       (const float*)b.kx, (const float*)b.ky, (cpx<float>*)b.lev)
        if (b.mma == 0) {
            if (alds) RBC_EVAL_MMA(0, true);
            else RBC_EVAL_MMA(0, false);
        } else {
            if (alds) RBC_EVAL_MMA(1, true);
            else RBC_EVAL_MMA(1, false);
        }
#undef RBC_EVAL_MMA
        return;
    }
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
        const int ntile = (g.KX + YT - 1) / YT, nblk = ntile * g.nq * 6;
        static const int resident = resident_blocks_per_xcd((const void*)k_fft_y_inv512, 512);
        const int bpx = std::max(1, std::min(resident, (nblk + 7) / 8));
        hipLaunchKernelGGL(k_fft_y_inv512, dim3(8 * bpx), dim3(512), 0, s, g, ntile, nblk, bpx,
                           (const cpx<float>*)b.lev, (const cpx<float>*)b.twP, (cpx<float>*)b.hin);
        return;
    }
    if (wave_y_log<T>(g) && launch_y_inv_group(g, 6, (const cpx<float>*)b.lev, (const cpx<float>*)b.twN,
                                               (cpx<float>*)b.hin, s, b.tune_ytx))
        return;
    const int TX = b.tune_ytx > 0 ? b.tune_ytx : fft_tile(g.N, sizeof(cpx<T>));
    dim3 grid((g.KX + TX - 1) / TX, g.nq, 6);
    const size_t sm = (size_t)2 * TX * g.N * sizeof(cpx<T>);
    hipLaunchKernelGGL(k_fft_y_inv<T>, grid, dim3(256), sm, s, g, TX, (const cpx<T>*)b.lev, b.rows,
                       (const cpx<T>*)b.twN, (cpx<T>*)b.hin);
}

template <typename T>
void Launch<T>::rows(const DevGeom& g, const Buffers& b, hipStream_t s) {
    if (fast512_rows<T>(g)) {
        const int nblk = g.N * g.nq, per = (nblk + 7) / 8;
        hipLaunchKernelGGL(k_rows512t, dim3(8 * per), dim3(192), 0, s, g, nblk, per, (const cpx<float>*)b.hin,
                           (const cpx<float>*)b.twP, (cpx<float>*)b.hout);
        return;
    }
    if (const int lg = wave_rows_log<T>(g)) {
        const int wpb = b.tune_rows_threads > 0 ? b.tune_rows_threads / 64 : 4;
        if (g.N % (wpb * 512 / g.M) == 0) {
            const dim3 gr(g.N / (wpb * 512 / g.M), g.nq);
#define RBC_ROWSW(LG, WPBV)                                                                                        \
    hipLaunchKernelGGL((k_rows_w<LG, WPBV>), gr, dim3(64 * WPBV), 0, s, g, (const cpx<float>*)b.hin,                \
                       (const cpx<float>*)b.twM, (cpx<float>*)b.hout)
#define RBC_ROWSW_L(LG)              \
    if (wpb == 2) RBC_ROWSW(LG, 2);  \
    else if (wpb == 8) RBC_ROWSW(LG, 8); \
    else RBC_ROWSW(LG, 4)
            if (lg == 6) { RBC_ROWSW_L(6); }
            else if (lg == 7) { RBC_ROWSW_L(7); }
            else { RBC_ROWSW_L(8); }
#undef RBC_ROWSW_L
#undef RBC_ROWSW
            return;
        }
    }
    const int nth = b.tune_rows_threads > 0 ? b.tune_rows_threads : 256;
    // rows per block: RB M = 512 for M >= 64 (at most 8 rows, at most nth / 32 so that a row keeps
    // 32 threads or more, and dividing N)
    int RB = 1;
    if (g.M >= 64)
        while (RB * 2 * g.M <= 512 && RB * 2 <= 8 && RB * 2 <= nth / 32 && g.N % (RB * 2) == 0) RB *= 2;
    const size_t sm = (size_t)6 * RB * g.M * sizeof(cpx<T>);
    const bool twl = sm + (size_t)g.M * sizeof(cpx<T>) <= 65536;
    const dim3 gr(g.N / RB, g.nq);
    if (twl)
        hipLaunchKernelGGL((k_rows<T, true>), gr, dim3(nth), sm + (size_t)g.M * sizeof(cpx<T>), s, g, RB,
                           (const cpx<T>*)b.hin, (const cpx<T>*)b.twM, (cpx<T>*)b.hout);
    else
        hipLaunchKernelGGL((k_rows<T, false>), gr, dim3(nth), sm, s, g, RB, (const cpx<T>*)b.hin,
                           (const cpx<T>*)b.twM, (cpx<T>*)b.hout);
}

template <typename T>
void Launch<T>::fft_y_fwd(const DevGeom& g, const Buffers& b, hipStream_t s) {
    if (fast512_y<T>(g)) {
        // 8 kx columns (8 waves) per block, one tile of the tile-major layout (measured before
        // that layout: 4 columns 194 us, 8 178 us and 16 195 us for the launches together); two
        // launches of three products each: group 0 (h, d, uv -> D0, D2) and group 3 (uw, vw,
        // u.grad(theta) -> D1, D3, D4)
        constexpr int CT = 8;
        const int ntile = (g.KX + CT - 1) / CT, nblk = ntile * g.nq, per = (2 * nblk + 7) / 8;
        hipLaunchKernelGGL((k_fft_y_fwd512c2<CT>), dim3(8 * per), dim3(64 * CT), 0, s, g, ntile, nblk, per,
                           (const cpx<float>*)b.hout, (const cpx<float>*)b.twP, (const float*)b.kx,
                           (const float*)b.ky, (cpx<float>*)b.lev, (cpx<float>*)b.fgw);
        return;
    }
    if (wave_y_log<T>(g) && g.NYc == kept_rows(g.N)) {
        // default: 2 waves per block at N = 64 (more blocks on the small grids), else 4
        const int G = 512 / g.N, wpb = b.tune_ytx > 0 && b.tune_ytx / G >= 2 ? b.tune_ytx / G : (g.N <= 64 ? 2 : 4);
        const dim3 gr((g.KX + wpb * G - 1) / (wpb * G), g.nq, 3);
#define RBC_YFWDC(LG, WPBV)                                                                                        \
    hipLaunchKernelGGL((k_fft_y_fwd_wc<LG, WPBV>), gr, dim3(64 * WPBV), 0, s, g, (const cpx<float>*)b.hout,        \
                       (const cpx<float>*)b.twN, (const float*)b.kx, (const float*)b.ky, (cpx<float>*)b.lev,   \
                       (cpx<float>*)b.fgw)
#define RBC_YFWDC_L(LG)                  \
    if (wpb == 2) RBC_YFWDC(LG, 2);      \
    else if (wpb == 8) RBC_YFWDC(LG, 8); \
    else RBC_YFWDC(LG, 4)
        if (g.N == 64) { RBC_YFWDC_L(6); }
        else if (g.N == 128) { RBC_YFWDC_L(7); }
        else { RBC_YFWDC_L(8); }
#undef RBC_YFWDC_L
#undef RBC_YFWDC
        return;
    }
    {
        const int TX = b.tune_ytx > 0 ? b.tune_ytx : fft_tile(g.N, sizeof(cpx<T>));
        dim3 grid((g.KX + TX - 1) / TX, g.nq, 6);
        const size_t sm = (size_t)2 * TX * g.N * sizeof(cpx<T>);
        hipLaunchKernelGGL(k_fft_y_fwd<T>, grid, dim3(256), sm, s, g, TX, (const cpx<T>*)b.hout, b.rows,
                           (const cpx<T>*)b.twN, (cpx<T>*)b.prod);
    }
    const size_t n = (size_t)g.nq * g.P;
    hipLaunchKernelGGL(k_combine<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, (const cpx<T>*)b.prod,
                       (const T*)b.kx, (const T*)b.ky, (cpx<T>*)b.lev, (cpx<T>*)b.fgw);
}

template <typename T>
void Launch<T>::project(const DevGeom& g, const Buffers& b, double Pr, void* const Nout[3], void* NFo, void* NGo,
                        hipStream_t s) {
    constexpr int MCH = sizeof(T) == 8 ? 4 : 8;
    dim3 grid((g.P + 63) / 64);
    if (sizeof(T) == 4 && b.mma >= 0) {
        // when 64-mode blocks would leave most of the 256 CUs idle: 16-mode blocks whose four
        // waves split the level sum
        const bool small = (g.P + 63) / 64 < 512;
        const dim3 gr = small ? dim3((g.P + 15) / 16) : grid, bl = dim3(256);
#define RBC_PROJECT_MMA(MMV, WPBV)                                                                              \
    hipLaunchKernelGGL((k_project_mma<MMV, WPBV, 4 / WPBV>), gr, bl, 0, s, g, b.pr_nks, (const float*)b.lev,      \
                       (const cpx<float>*)b.fgw, (const float*)b.prA, (const float*)b.prtab, (const float*)b.kx,   \
                       (const float*)b.ky, (float)Pr, (cpx<float>*)Nout[0], (cpx<float>*)Nout[1],                \
                       (cpx<float>*)Nout[2], (cpx<float>*)NFo, (cpx<float>*)NGo)
        if (b.mma == 0) {
            if (small) RBC_PROJECT_MMA(0, 1);
            else RBC_PROJECT_MMA(0, 4);
        } else {
            if (small) RBC_PROJECT_MMA(1, 1);
            else RBC_PROJECT_MMA(1, 4);
        }
#undef RBC_PROJECT_MMA
        return;
    }
    if (g.L <= 4 * MCH) {
        hipLaunchKernelGGL((k_project_lds<T, MCH>), grid, dim3(256), 0, s, g, (const cpx<T>*)b.lev,
                           (const cpx<T>*)b.fgw, (const T*)b.prtab, (const T*)b.kx, (const T*)b.ky, (T)Pr,
                           (cpx<T>*)Nout[0], (cpx<T>*)Nout[1], (cpx<T>*)Nout[2], (cpx<T>*)NFo, (cpx<T>*)NGo);
        return;
    }
    hipLaunchKernelGGL((k_project<T, MCH>), grid, dim3(256), 0, s, g, (const cpx<T>*)b.lev, (const cpx<T>*)b.fgw,
                       (const T*)b.prtab, (const T*)b.kx, (const T*)b.ky, (T)Pr, (cpx<T>*)Nout[0],
                       (cpx<T>*)Nout[1], (cpx<T>*)Nout[2], (cpx<T>*)NFo, (cpx<T>*)NGo);
}

template <typename T>
void Launch<T>::update(const DevGeom& g, const Buffers& b, int ngroups, int L0, int L1, int stride, double Pr,
                       double dt, StepCoef sc, int mode, void* const Xio[3], void* const Nin[3], void* const Rp[3],
                       hipStream_t s) {
    if (mode == UPD_STEP && g.L <= 32) {
#define RBC_UPD32(NCV)                                                                                            \
    hipLaunchKernelGGL((k_update32<T, NCV>), dim3((b.nchunk + 3) / 4), dim3(256), 0, s, g, b.nchunk, L0, L1,       \
                       stride, (T)Pr, (T)dt, (T)sc.a_cur, (T)sc.a_prev, sc.use_prev, b.cgrp, b.cbeg, b.gmem,           \
                       (const T*)b.gmat,                                                                               \
                       (const T*)b.k2, (cpx<T>*)Xio[0], (cpx<T>*)Xio[1], (cpx<T>*)Xio[2], (const cpx<T>*)Nin[0],       \
                       (const cpx<T>*)Nin[1], (const cpx<T>*)Nin[2], (cpx<T>*)Rp[0], (cpx<T>*)Rp[1], (cpx<T>*)Rp[2])
        if (g.L <= 16) RBC_UPD32(16);
        else RBC_UPD32(32);
#undef RBC_UPD32
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
void Launch<T>::finish(const DevGeom& g, const Buffers& b, double Pr, double dt, StepCoef sc, hipStream_t s) {
    const int n = 3 * g.NYc * g.L;
    hipLaunchKernelGGL(k_finish<T>, dim3((n + 255) / 256 + 1), dim3(256), 0, s, g, b.rows, (cpx<T>*)b.X[0],
                       (cpx<T>*)b.X[1], (cpx<T>*)b.X[2], (T)Pr, (T)dt, (T)sc.a_cur, (T)sc.a_prev, sc.use_prev,
                       (cpx<T>*)b.XF, (cpx<T>*)b.XG, (const cpx<T>*)b.NF, (const cpx<T>*)b.NG, (cpx<T>*)b.RpF,
                       (cpx<T>*)b.RpG, b.health);
}

template <typename T>
void Launch<T>::nusselt(const DevGeom& g, const Buffers& b, const void* ia1, hipStream_t s) {
    const int nblk = (g.P + 255) / 256;
    hipLaunchKernelGGL(k_nusselt<T>, dim3(nblk), dim3(256), 0, s, g, (const cpx<T>*)b.X[0],
                       (const cpx<T>*)b.X[1], (const T*)b.k2, (const double*)ia1, b.red);
}

template <typename T>
void Launch<T>::from_ref(const DevGeom& g, const void* const src[3], void* const dst[3], hipStream_t s) {
    const int mtn = (g.L + 31) / 32;
    hipLaunchKernelGGL(k_ref_to_compact<T>, dim3((g.KX + 31) / 32, g.NYc, 3 * mtn), dim3(256), 0, s, g,
                       (const cpx<T>*)src[0], (const cpx<T>*)src[1], (const cpx<T>*)src[2], (cpx<T>*)dst[0],
                       (cpx<T>*)dst[1], (cpx<T>*)dst[2]);
}

template <typename T>
void Launch<T>::to_ref(const DevGeom& g, const void* const src[3], void* const dst[3], hipStream_t s) {
    const int mtn = (g.L + 31) / 32;
    hipLaunchKernelGGL(k_compact_to_ref<T>, dim3((g.M / 2 + 1 + 31) / 32, g.N, 3 * mtn), dim3(256), 0, s, g,
                       (const cpx<T>*)src[0], (const cpx<T>*)src[1], (const cpx<T>*)src[2], (cpx<T>*)dst[0],
                       (cpx<T>*)dst[1], (cpx<T>*)dst[2]);
}

// one field between an FP64 reference-layout buffer and the compact field of the context's real
// type (the host-buffer boundary: the host arrays are copied as they are and converted here)
void field_from_ref64(const DevGeom& g, int rsz, const void* ref, void* dst, hipStream_t s) {
    const int mtn = (g.L + 31) / 32;
    const dim3 grid((g.KX + 31) / 32, g.NYc, mtn);
    if (rsz == 8)
        hipLaunchKernelGGL((k_ref_to_compact<double, double>), grid, dim3(256), 0, s, g, (const cpx<double>*)ref,
                           nullptr, nullptr, (cpx<double>*)dst, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_ref_to_compact<double, float>), grid, dim3(256), 0, s, g, (const cpx<double>*)ref,
                           nullptr, nullptr, (cpx<float>*)dst, nullptr, nullptr);
}
void field_to_ref64(const DevGeom& g, int rsz, const void* src, void* ref, hipStream_t s) {
    const int mtn = (g.L + 31) / 32;
    const dim3 grid((g.M / 2 + 1 + 31) / 32, g.N, mtn);
    if (rsz == 8)
        hipLaunchKernelGGL((k_compact_to_ref<double, double>), grid, dim3(256), 0, s, g, (const cpx<double>*)src,
                           nullptr, nullptr, (cpx<double>*)ref, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_compact_to_ref<float, double>), grid, dim3(256), 0, s, g, (const cpx<float>*)src,
                           nullptr, nullptr, (cpx<double>*)ref, nullptr, nullptr);
}

template <typename T>
void Launch<T>::init_state(const DevGeom& g, const Buffers& b, int kind, uint64_t seed, double amp, hipStream_t s) {
    const size_t n = std::max((size_t)g.P * g.L, (size_t)g.L);
    hipLaunchKernelGGL(k_init_state<T>, dim3((unsigned)((n + 255) / 256), 3), dim3(256), 0, s, g, kind, seed, amp,
                       (cpx<T>*)b.X[0], (cpx<T>*)b.X[1], (cpx<T>*)b.X[2], (cpx<T>*)b.XF, (cpx<T>*)b.XG);
}

template <typename T>
void Launch<T>::force_mean(const Buffers& b, double v, hipStream_t s) {
    hipLaunchKernelGGL(k_force_mean<T>, dim3(1), dim3(64), 0, s, (cpx<T>*)b.XF, v);
}

template <typename T>
void Launch<T>::penalty(const DevGeom& g, const Buffers& b, const Penalty& pen, double Pr, void* const X[3],
                        const void* XF, const void* XG, void* const Nout[3], void* NF, void* NG, hipStream_t s) {
    (void)Pr;
    const size_t n = (size_t)g.L * g.P;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_pen_fields<T>, dim3(nb), dim3(256), 0, s, g, (const T*)pen.iac, (const T*)pen.ia2,
                       (const cpx<T>*)X[0], (const cpx<T>*)X[1], (const cpx<T>*)X[2], (const cpx<T>*)XF,
                       (const cpx<T>*)XG, (const T*)b.kx, (const T*)b.ky, (cpx<T>*)pen.spec);
    // the generic y-transforms with the L modes in place of the levels, ten fields
    DevGeom gp = g;
    gp.nq = g.L;
    const int TX = fft_tile(g.N, sizeof(cpx<T>));
    const size_t sm = (size_t)2 * TX * g.N * sizeof(cpx<T>);
    const dim3 gy((g.KX + TX - 1) / TX, g.L, 10);
    hipLaunchKernelGGL(k_fft_y_inv<T>, gy, dim3(256), sm, s, gp, TX, (const cpx<T>*)pen.spec, b.rows,
                       (const cpx<T>*)b.twN, (cpx<T>*)pen.hin);
    hipLaunchKernelGGL(k_rows_mask<T>, dim3(g.N, g.L, 6), dim3(256), (size_t)2 * g.M * sizeof(cpx<T>), s, gp,
                       (const T*)pen.mask, (const cpx<T>*)pen.hin, (const cpx<T>*)b.twM, (cpx<T>*)pen.hout);
    hipLaunchKernelGGL(k_fft_y_fwd<T>, gy, dim3(256), sm, s, gp, TX, (const cpx<T>*)pen.hout, b.rows,
                       (const cpx<T>*)b.twN, (cpx<T>*)pen.spec_out);
    hipLaunchKernelGGL(k_pen_combine<T>, dim3(nb), dim3(256), 0, s, g, (T)(1.0 / pen.eta),
                       (const cpx<T>*)pen.spec_out, (const T*)b.kx, (const T*)b.ky, (cpx<T>*)Nout[0],
                       (cpx<T>*)Nout[1], (cpx<T>*)Nout[2], (cpx<T>*)NF, (cpx<T>*)NG);
}

template <typename T>
void Launch<T>::plane_op(const DevGeom& g, const Buffers& b, int op, int nf, const int* imask, double factor,
                         void* frozen, void* const X[3], void* spec, void* hin, void* hout, void* spec_out,
                         hipStream_t s) {
    const size_t n = (size_t)g.L * g.P;
    const dim3 gm((unsigned)((n + 255) / 256), nf);
    hipLaunchKernelGGL(k_modes_to_planes<T>, gm, dim3(256), 0, s, g, nf, (const cpx<T>*)X[0], (const cpx<T>*)X[1],
                       (const cpx<T>*)X[2], (cpx<T>*)spec);
    DevGeom gp = g;
    gp.nq = g.L;
    const int TX = fft_tile(g.N, sizeof(cpx<T>));
    const size_t sm = (size_t)2 * TX * g.N * sizeof(cpx<T>);
    const dim3 gy((g.KX + TX - 1) / TX, g.L, nf);
    hipLaunchKernelGGL(k_fft_y_inv<T>, gy, dim3(256), sm, s, gp, TX, (const cpx<T>*)spec, b.rows,
                       (const cpx<T>*)b.twN, (cpx<T>*)hin);
    hipLaunchKernelGGL(k_rows_planeop<T>, dim3(g.N, g.L, nf), dim3(256), (size_t)2 * g.M * sizeof(cpx<T>), s, gp,
                       op, imask, (T)factor, (T*)frozen, (const cpx<T>*)hin, (const cpx<T>*)b.twM, (cpx<T>*)hout);
    if (op == PLANE_CAPTURE) return;
    hipLaunchKernelGGL(k_fft_y_fwd<T>, gy, dim3(256), sm, s, gp, TX, (const cpx<T>*)hout, b.rows,
                       (const cpx<T>*)b.twN, (cpx<T>*)spec_out);
    hipLaunchKernelGGL(k_planes_to_modes<T>, gm, dim3(256), 0, s, g, nf, (const cpx<T>*)spec_out, (cpx<T>*)X[0],
                       (cpx<T>*)X[1], (cpx<T>*)X[2]);
}

template <typename T>
void Launch<T>::temperature_sum(const DevGeom& g, const void* planes, const float* S, float* out, hipStream_t s) {
    const size_t n = (size_t)g.M * g.N;
    hipLaunchKernelGGL(k_temp_sum<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g.L, n, (const T*)planes,
                       S, out);
}

template struct Launch<float>;
template struct Launch<double>;

// ---------------------------------------------------------------------------------------
// Batched steps: the members of one ensemble size class advance together, one launch per stage
// for all of them (blockIdx.z = member, or member x fields for the kernels whose own grid has a
// z extent), each block reading its member's arguments from a table in device memory.  The
// kernels are the single-run kernels' bodies with the single-run launch shapes, so a member's
// arithmetic is the same as in a run of its own.
// ---------------------------------------------------------------------------------------
struct StepBatch {
    int B = 0;
    DevGeom g;            // the class geometry (launch shapes; Lx, Ly come from each member's table)
    int mma = 0;          // 0: FP32 MFMA, 1: BF16 x3
    bool alds = false;    // evaluation operand tables in LDS
    size_t ev_smem = 0;   // evaluation dynamic LDS
    int ev_blocks = 0;    // evaluation blocks per member (XCD-ordered 1-D grid)
    int yi_wpb = 0;       // waves per block of the inverse group y-FFT
    int yf_wpb = 0;       // waves per block of the forward group y-FFT
    int rows_wpb = 0;     // waves per row block
    bool small_proj = false;
    int nchunk_max = 0;
    k_eval_mma_args* ev = nullptr;
    k_fft_y_inv_w_args* yi = nullptr;
    k_rows_w_args* rw = nullptr;
    k_fft_y_fwd_wc_args* yf = nullptr;
    k_project_mma_args* pj = nullptr;
    k_update32_args<float>* up = nullptr;
    k_finish_args<float>* fi = nullptr;
};

template <typename A>
static A* upload_table(const std::vector<A>& v) {
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, v.size() * sizeof(A)));
    HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(A), hipMemcpyHostToDevice));
    return static_cast<A*>(p);
}

void batch_destroy(StepBatch* sb) {
    if (!sb) return;
    for (void* p : {(void*)sb->ev, (void*)sb->yi, (void*)sb->rw, (void*)sb->yf, (void*)sb->pj, (void*)sb->up,
                    (void*)sb->fi})
        if (p) hipFree(p);
    delete sb;
}

StepBatch* batch_create(const std::vector<BatchMember>& m, std::string& why) {
    if (m.empty()) {
        why = "no members";
        return nullptr;
    }
    const DevGeom& g = *m[0].g;
    const Buffers& b0 = *m[0].b;
    // the single-run path these kernels stand for: FP32 storage on the matrix cores, group
    // transforms in both directions, the register group update
    if (b0.mma < 0 || g.L > 32 || !wave_rows_log<float>(g) || !wave_y_log<float>(g) || g.NYc != kept_rows(g.N) ||
        fast512_rows<float>(g) || fast512_y<float>(g)) {
        why = "not on the batched path (FP32/BF16X3 storage, 64..256-point group transforms, L <= 32)";
        return nullptr;
    }
    // launch shapes as the single-run launchers choose them from the tuned values
    const int G = 512 / g.N;
    int yi_wpb = b0.tune_ytx > 0 ? b0.tune_ytx / G : 4;
    int yf_wpb = b0.tune_ytx > 0 && b0.tune_ytx / G >= 2 ? b0.tune_ytx / G : (g.N <= 64 ? 2 : 4);
    const int rows_wpb = b0.tune_rows_threads > 0 ? b0.tune_rows_threads / 64 : 4;
    if (yi_wpb != 2 && yi_wpb != 8) yi_wpb = 4;
    if (yf_wpb != 2 && yf_wpb != 8) yf_wpb = 4;
    if ((rows_wpb != 2 && rows_wpb != 4 && rows_wpb != 8) || g.N % (rows_wpb * 512 / g.M) != 0) {
        why = "row launch shape outside the group kernels";
        return nullptr;
    }
    for (const BatchMember& x : m) {
        const DevGeom& h = *x.g;
        const Buffers& b = *x.b;
        if (h.M != g.M || h.N != g.N || h.L != g.L || h.nq != g.nq || h.KX != g.KX || h.KXp != g.KXp ||
            h.NYc != g.NYc || h.P != g.P || b.mma != b0.mma || b.ev_nlt != b0.ev_nlt || b.ev_nks != b0.ev_nks ||
            b.pr_nks != b0.pr_nks || b.tune_ytx != b0.tune_ytx || b.tune_rows_threads != b0.tune_rows_threads) {
            why = "members differ in launch shape";
            return nullptr;
        }
    }
    StepBatch* sb = new StepBatch;
    sb->B = (int)m.size();
    sb->g = g;
    sb->mma = b0.mma;
    sb->yi_wpb = yi_wpb;
    sb->yf_wpb = yf_wpb;
    sb->rows_wpb = rows_wpb;
    {
        const int K = b0.mma == 0 ? 2 : 16;
        const size_t sx = (size_t)3 * 128 * (b0.ev_nks * K + 1) * sizeof(float);
        const size_t sa = (size_t)4 * b0.ev_nks * 64 * (K / 2) * sizeof(float);
        sb->alds = sx + sa <= 65536;
        sb->ev_smem = sb->alds ? sx + sa : sx;
    }
    const int nblk = ((g.P + 63) / 64) * b0.ev_nlt, per = (nblk + 7) / 8;
    sb->ev_blocks = 8 * per;
    sb->small_proj = (g.P + 63) / 64 < 512;
    std::vector<k_eval_mma_args> ev;
    std::vector<k_fft_y_inv_w_args> yi;
    std::vector<k_rows_w_args> rw;
    std::vector<k_fft_y_fwd_wc_args> yf;
    std::vector<k_project_mma_args> pj;
    std::vector<k_update32_args<float>> up;
    std::vector<k_finish_args<float>> fi;
    for (const BatchMember& x : m) {
        const DevGeom& h = *x.g;
        const Buffers& b = *x.b;
        using C = cpx<float>;
        ev.push_back({h, b.ev_nlt, b.ev_nks, per, (const C*)b.X[0], (const C*)b.X[1], (const C*)b.X[2],
                      (const C*)b.XF, (const C*)b.XG, (const float*)b.evA, (const float*)b.evtab,
                      (const float*)b.kx, (const float*)b.ky, (C*)b.lev});
        yi.push_back({h, (const C*)b.lev, (const C*)b.twN, (C*)b.hin});
        rw.push_back({h, (const C*)b.hin, (const C*)b.twM, (C*)b.hout});
        yf.push_back({h, (const C*)b.hout, (const C*)b.twN, (const float*)b.kx, (const float*)b.ky, (C*)b.lev,
                      (C*)b.fgw});
        pj.push_back({h, b.pr_nks, (const float*)b.lev, (const C*)b.fgw, (const float*)b.prA, (const float*)b.prtab,
                      (const float*)b.kx, (const float*)b.ky, (float)x.Pr, (C*)b.Nn[0], (C*)b.Nn[1], (C*)b.Nn[2],
                      (C*)b.NF, (C*)b.NG});
        // AB2 coefficients: a batch only runs members that have their history
        const float a_cur = (float)(1.5 * x.dt), a_prev = (float)(-0.5 * x.dt);
        up.push_back({h, b.nchunk, x.L0, x.L1, x.stride, (float)x.Pr, (float)x.dt, a_cur, a_prev, 1, b.cgrp, b.cbeg,
                      b.gmem, (const float*)b.gmat, (const float*)b.k2, (C*)b.X[0], (C*)b.X[1], (C*)b.X[2],
                      (const C*)b.Nn[0], (const C*)b.Nn[1], (const C*)b.Nn[2], (C*)b.Rp[0], (C*)b.Rp[1],
                      (C*)b.Rp[2]});
        fi.push_back({h, b.rows, (C*)b.X[0], (C*)b.X[1], (C*)b.X[2], (float)x.Pr, (float)x.dt, a_cur, a_prev, 1,
                      (C*)b.XF, (C*)b.XG, (const C*)b.NF, (const C*)b.NG, (C*)b.RpF, (C*)b.RpG, b.health});
        sb->nchunk_max = std::max(sb->nchunk_max, b.nchunk);
    }
    try {
        sb->ev = upload_table(ev);
        sb->yi = upload_table(yi);
        sb->rw = upload_table(rw);
        sb->yf = upload_table(yf);
        sb->pj = upload_table(pj);
        sb->up = upload_table(up);
        sb->fi = upload_table(fi);
    } catch (const HipError&) {
        batch_destroy(sb);
        throw;
    }
    return sb;
}

// One AB2 step of every member, in the stage order of a single step (evaluation, inverse
// y-FFT, rows, forward y-FFT with the combinations, projection, group update, finish).
void batch_step(const StepBatch* sb, hipStream_t s) {
    const DevGeom& g = sb->g;
    const int B = sb->B;
    {
        const dim3 gr(sb->ev_blocks, 1, B);
#define RBC_EVB(MMV, ALDSV) \
    hipLaunchKernelGGL((k_eval_mma_batch<MMV, ALDSV>), gr, dim3(256), sb->ev_smem, s, sb->ev, 1)
        if (sb->mma == 0) {
            if (sb->alds) RBC_EVB(0, true);
            else RBC_EVB(0, false);
        } else {
            if (sb->alds) RBC_EVB(1, true);
            else RBC_EVB(1, false);
        }
#undef RBC_EVB
    }
    const int G = 512 / g.N;
    {
        const int ywpb = sb->yi_wpb, cb = ywpb * G;
        const dim3 gr((g.KX + cb - 1) / cb, g.nq, 6 * B);
#define RBC_YIB(LG, WPBV) hipLaunchKernelGGL((k_fft_y_inv_w_batch<LG, WPBV>), gr, dim3(64 * WPBV), 0, s, sb->yi, 6)
#define RBC_YIB_L(LG)                 \
    if (ywpb == 2) RBC_YIB(LG, 2);      \
    else if (ywpb == 8) RBC_YIB(LG, 8); \
    else RBC_YIB(LG, 4)
        if (g.N == 64) { RBC_YIB_L(6); }
        else if (g.N == 128) { RBC_YIB_L(7); }
        else { RBC_YIB_L(8); }
#undef RBC_YIB_L
#undef RBC_YIB
    }
    {
        const int wpb = sb->rows_wpb;
        const dim3 gr(g.N / (wpb * 512 / g.M), g.nq, B);
#define RBC_RWB(LG, WPBV) hipLaunchKernelGGL((k_rows_w_batch<LG, WPBV>), gr, dim3(64 * WPBV), 0, s, sb->rw, 1)
#define RBC_RWB_L(LG)                \
    if (wpb == 2) RBC_RWB(LG, 2);      \
    else if (wpb == 8) RBC_RWB(LG, 8); \
    else RBC_RWB(LG, 4)
        if (g.M == 64) { RBC_RWB_L(6); }
        else if (g.M == 128) { RBC_RWB_L(7); }
        else { RBC_RWB_L(8); }
#undef RBC_RWB_L
#undef RBC_RWB
    }
    {
        const int wpb = sb->yf_wpb;
        const dim3 gr((g.KX + wpb * G - 1) / (wpb * G), g.nq, 3 * B);
#define RBC_YFB(LG, WPBV) hipLaunchKernelGGL((k_fft_y_fwd_wc_batch<LG, WPBV>), gr, dim3(64 * WPBV), 0, s, sb->yf, 3)
#define RBC_YFB_L(LG)                \
    if (wpb == 2) RBC_YFB(LG, 2);      \
    else if (wpb == 8) RBC_YFB(LG, 8); \
    else RBC_YFB(LG, 4)
        if (g.N == 64) { RBC_YFB_L(6); }
        else if (g.N == 128) { RBC_YFB_L(7); }
        else { RBC_YFB_L(8); }
#undef RBC_YFB_L
#undef RBC_YFB
    }
    {
        const dim3 gr(sb->small_proj ? (g.P + 15) / 16 : (g.P + 63) / 64, 1, B);
#define RBC_PJB(MMV, WPBV) \
    hipLaunchKernelGGL((k_project_mma_batch<MMV, WPBV, 4 / WPBV>), gr, dim3(256), 0, s, sb->pj, 1)
        if (sb->mma == 0) {
            if (sb->small_proj) RBC_PJB(0, 1);
            else RBC_PJB(0, 4);
        } else {
            if (sb->small_proj) RBC_PJB(1, 1);
            else RBC_PJB(1, 4);
        }
#undef RBC_PJB
    }
    {
        const dim3 gr((sb->nchunk_max + 3) / 4, 1, B);
        if (g.L <= 16) hipLaunchKernelGGL((k_update32_batch<float, 16>), gr, dim3(256), 0, s, sb->up, 1);
        else hipLaunchKernelGGL((k_update32_batch<float, 32>), gr, dim3(256), 0, s, sb->up, 1);
    }
    {
        const int n = 3 * g.NYc * g.L;
        hipLaunchKernelGGL(k_finish_batch<float>, dim3((n + 255) / 256 + 1, 1, B), dim3(256), 0, s, sb->fi, 1);
    }
}

}  // namespace rbc
