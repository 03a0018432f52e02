// rbc_abi.cpp -- C-ABI of libsdc_rbc.so: context lifetime, state exchange in the
// reference layout, the time step pipeline, operator-level calls and diagnostics.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "rbc_ctx.hpp"

using namespace rbc;

namespace {
thread_local std::string g_create_error;
}  // namespace

namespace {

template <typename T>
void upload_real(rbc_ctx* c, void* dst, const std::vector<double>& src) {
    std::vector<T> tmp(src.size());
    for (size_t i = 0; i < src.size(); ++i) tmp[i] = (T)src[i];
    HIPCHK(hipMemcpy(dst, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
}

void upload_reals(rbc_ctx* c, void* dst, const std::vector<double>& src) {
    if (c->rsz == 8) upload_real<double>(c, dst, src);
    else upload_real<float>(c, dst, src);
}

}  // namespace

namespace rbc {

// host fields in the reference layout: copied as they are into the context's FP64 staging
// buffer and converted to (from) the compact layout on the GPU
void put_field(rbc_ctx* c, const double* host, void* dev) {
    const size_t n = (size_t)c->geo.mox * c->geo.N * c->geo.L * 16;
    HIPCHK(hipMemcpyAsync(c->b.refstage, host, n, hipMemcpyHostToDevice, c->stream));
    field_from_ref64(c->dg, c->rsz, c->b.refstage, dev, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
}

void get_field(rbc_ctx* c, const void* dev, double* host) {
    const size_t n = (size_t)c->geo.mox * c->geo.N * c->geo.L * 16;
    field_to_ref64(c->dg, c->rsz, dev, c->b.refstage, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, c->b.refstage, n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}

void put_vec(rbc_ctx* c, const double* host, void* dev) {
    const int L = c->geo.L;
    std::vector<unsigned char> tmp((size_t)L * c->csz());
    for (int i = 0; i < 2 * L; ++i) {
        if (c->rsz == 8) ((double*)tmp.data())[i] = host[i];
        else ((float*)tmp.data())[i] = (float)host[i];
    }
    HIPCHK(hipMemcpyAsync(dev, tmp.data(), tmp.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}

void get_vec(rbc_ctx* c, const void* dev, double* host) {
    const int L = c->geo.L;
    std::vector<unsigned char> tmp((size_t)L * c->csz());
    HIPCHK(hipMemcpyAsync(tmp.data(), dev, tmp.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 2 * L; ++i) host[i] = (c->rsz == 8) ? ((double*)tmp.data())[i] : ((float*)tmp.data())[i];
}

}  // namespace rbc

namespace {

// ---- pipeline, dispatched on precision -------------------------------------------------
template <typename T>
void nonlinear_into(rbc_ctx* c, void* const X[3], const void* XF, const void* XG, void* const Nout[3], void* NF,
                    void* NG, int stage_base) {
    auto mark = [&](int i) {
        if (stage_base >= 0) HIPCHK(hipEventRecord(c->ev[stage_base + i], c->stream));
    };
    Launch<T>::eval(c->dg, c->b, X, XF, XG, c->stream);
    mark(1);
    Launch<T>::fft_y_inv(c->dg, c->b, c->stream);
    mark(2);
    Launch<T>::rows(c->dg, c->b, c->stream);
    mark(3);
    Launch<T>::fft_y_fwd(c->dg, c->b, c->stream);
    mark(4);
    Launch<T>::project(c->dg, c->b, c->p.Pr, Nout, NF, NG, c->stream);
    if (c->pen.eta > 0.0) Launch<T>::penalty(c->dg, c->b, c->pen, c->p.Pr, X, XF, XG, Nout, NF, NG, c->stream);
    mark(5);
}

template <typename T>
void one_step(rbc_ctx* c, bool timed) {
    StepCoef sc;
    if (c->have_hist) {
        sc.a_cur = 1.5 * c->p.dt;
        sc.a_prev = -0.5 * c->p.dt;
        sc.use_prev = 1;
    } else {
        sc.a_cur = c->p.dt;
        sc.a_prev = 0.0;
        sc.use_prev = 0;
    }
    if (timed) HIPCHK(hipEventRecord(c->ev[0], c->stream));
    nonlinear_into<T>(c, c->b.X, c->b.XF, c->b.XG, c->b.Nn, c->b.NF, c->b.NG, timed ? 0 : -1);
    if (c->post_nonlinear) c->post_nonlinear(c, c->post_user);
    Launch<T>::update(c->dg, c->b, c->fac.ngroups(), c->fac.L0, c->fac.L1, c->fac.stride, c->p.Pr, c->p.dt, sc,
                      UPD_STEP, c->b.X, c->b.Nn, c->b.Rp, c->stream);
    if (timed) HIPCHK(hipEventRecord(c->ev[6], c->stream));
    Launch<T>::finish(c->dg, c->b, c->p.Pr, c->p.dt, sc, c->stream);  // make_real + F, G update
    if (timed) HIPCHK(hipEventRecord(c->ev[7], c->stream));
    HIPCHK(hipGetLastError());
    c->have_hist = true;
}

}  // namespace

namespace rbc {

void nonlinear_op(rbc_ctx* c, void* const X[3], const void* XF, const void* XG, void* const Nout[3], void* NF,
                  void* NG) {
    if (c->rsz == 8) nonlinear_into<double>(c, X, XF, XG, Nout, NF, NG, -1);
    else nonlinear_into<float>(c, X, XF, XG, Nout, NF, NG, -1);
}

void run_steps(rbc_ctx* c, int n) {
    for (int i = 0; i < n; ++i) {
        if (c->rsz == 8) one_step<double>(c, false);
        else one_step<float>(c, false);
    }
}

int health_status(rbc_ctx* c) {
    int h[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(h, c->b.health, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (h[0] > 0 && c->bad_step == 0) c->bad_step = h[0];
    if (c->bad_step > 0)
        return fail(c, RBC_ERR_NUMERIC,
                    "non-finite state: step " + std::to_string(c->bad_step) +
                        " (counted from context creation) ended with a non-finite value (find_NaN)");
    return RBC_OK;
}

void clear_health(rbc_ctx* c) {
    c->bad_step = 0;
    HIPCHK(hipMemsetAsync(c->b.health, 0, sizeof(int), c->stream));
}

int fail(rbc_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

int hip_fail(rbc_ctx* c, const HipError& e) {
    return fail(c, RBC_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
}

}  // namespace rbc

namespace {

const char* kStageNames[RBC_NSTAGES] = {"eval", "fft_y_inv", "rows_products", "fft_y_fwd",
                                        "project", "update", "finish"};

}  // namespace

extern "C" {

int rbc_create(const rbc_params* p, rbc_ctx** out) {
    if (!p || !out) return fail(nullptr, RBC_ERR_ARG, "null argument");
    *out = nullptr;
    auto pow2 = [](int n) { return n >= 8 && n <= 4096 && (n & (n - 1)) == 0; };
    if (!pow2(p->M) || !pow2(p->N)) return fail(nullptr, RBC_ERR_ARG, "M and N must be powers of two in [8, 4096]");
    if (p->L < 1 || p->L > 64) return fail(nullptr, RBC_ERR_ARG, "L must be in [1, 64]");
    if (!(p->Lx > 0) || !(p->Ly > 0) || !(p->Pr > 0) || !(p->dt > 0) || !std::isfinite(p->Ra))
        return fail(nullptr, RBC_ERR_ARG, "Lx, Ly, Pr, dt must be positive and Ra finite");
    if (p->precision < RBC_F64 || p->precision > RBC_BF16X3) return fail(nullptr, RBC_ERR_ARG, "bad precision");
    if (p->precision == RBC_BF16X3 && p->L > 32)
        return fail(nullptr, RBC_ERR_ARG, "RBC_BF16X3 supports L <= 32 (one 32-row matrix-core tile)");
    if (p->nq < 0 || p->nq > 1024) return fail(nullptr, RBC_ERR_ARG, "nq out of range");
    {
        // LDS budgets of the row kernel (8 M complex) and the y-FFT tiles (2 N complex)
        const size_t cs = (p->precision == RBC_F64) ? 16 : 8;
        if (8 * (size_t)p->M * cs > 65536 || 2 * (size_t)p->N * cs > 65536)
            return fail(nullptr, RBC_ERR_ARG, "grid too large for the LDS budget at this precision");
    }
    std::unique_ptr<rbc_ctx> c(new rbc_ctx());
    c->p = *p;
    c->prec = p->precision;
    c->rsz = (p->precision == RBC_F64) ? 8 : 4;
    try {
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (p->device < 0 || p->device >= ndev) return fail(nullptr, RBC_ERR_ARG, "no such HIP device");
        HIPCHK(hipSetDevice(p->device));
        c->geo = make_geometry(p->M, p->N, p->L);
        const Geometry& g = c->geo;
        // kept rows must be [0, a] U [N-a, N-1] (true for the reference mask at every N)
        const int NYc = g.NYc(), a = (NYc - 1) / 2;
        for (int i = 0; i < NYc; ++i) {
            const int want = (i <= a) ? i : p->N - (NYc - i);
            if (NYc % 2 == 0 || g.rows[i] != want) return fail(nullptr, RBC_ERR_ARG, "unexpected dealias row set");
        }
        double qerr = 0.0;
        int nq = p->nq;
        if (nq == 0) nq = exact_level_count(p->L, p->precision == RBC_F64 ? 1e-13 : 1e-8, &qerr);
        c->p.nq = nq;
        DevGeom& d = c->dg;
        d.M = p->M;
        d.N = p->N;
        d.L = p->L;
        d.nq = nq;
        d.KX = g.KX;
        // hybrid row stride; the 512 x 512 kernels store the hybrid rows tile-major in tiles of 8
        // kx columns ([tile][y][8]), so the stride there is a whole number of tiles
        d.KXp = (p->M == 512 && p->N == 512) ? (g.KX + 7) & ~7 : (g.KX + 3) & ~3;
        d.NYc = NYc;
        d.P = g.P();
        d.logM = 0;
        while ((1 << d.logM) < p->M) ++d.logM;
        d.logN = 0;
        while ((1 << d.logN) < p->N) ++d.logN;
        d.Lx = p->Lx;
        d.Ly = p->Ly;

        std::string err;
        if (build_factors(g, p->Lx, p->Ly, p->Ra, p->Pr, p->dt, c->fac, err)) return fail(nullptr, RBC_ERR_NUMERIC, err);

        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto& row : c->ev_ring)
            for (auto& e : row) HIPCHK(hipEventCreate(&e));
        c->events = true;

        const size_t cs = c->csz(), P = d.P, L = d.L;
        Buffers& b = c->b;
        for (int i = 0; i < 3; ++i) {
            b.X[i] = c->dalloc(P * L * cs);
            b.Nn[i] = c->dalloc(P * L * cs);
            b.Rp[i] = c->dalloc(P * L * cs);
            b.Xo[i] = c->dalloc(P * L * cs);
            b.Oo[i] = c->dalloc(P * L * cs);
        }
        b.XF = c->dalloc(L * cs);
        b.XG = c->dalloc(L * cs);
        b.NF = c->dalloc(L * cs);
        b.NG = c->dalloc(L * cs);
        b.RpF = c->dalloc(L * cs);
        b.RpG = c->dalloc(L * cs);
        b.XoF = c->dalloc(L * cs);
        b.XoG = c->dalloc(L * cs);
        b.OoF = c->dalloc(L * cs);
        b.OoG = c->dalloc(L * cs);
        b.lev = c->dalloc((size_t)6 * nq * P * cs);
        b.hin = c->dalloc((size_t)6 * nq * d.N * d.KXp * cs);
        b.hout = c->dalloc((size_t)7 * nq * d.N * d.KXp * cs);
        b.prod = c->dalloc((size_t)7 * nq * P * cs);
        b.fgw = c->dalloc((size_t)2 * nq * cs);
        b.refstage = c->dalloc((size_t)(d.M / 2 + 1) * d.N * L * 16);
        // zero-initialise state and padding once (not in any step: nothing is memset later)
        for (int i = 0; i < 3; ++i) {
            HIPCHK(hipMemset(b.X[i], 0, P * L * cs));
            HIPCHK(hipMemset(b.Rp[i], 0, P * L * cs));
        }
        HIPCHK(hipMemset(b.XF, 0, L * cs));
        HIPCHK(hipMemset(b.XG, 0, L * cs));
        HIPCHK(hipMemset(b.RpF, 0, L * cs));
        HIPCHK(hipMemset(b.RpG, 0, L * cs));
        HIPCHK(hipMemset(b.hin, 0, (size_t)6 * nq * d.N * d.KXp * cs));
        HIPCHK(hipMemset(b.hout, 0, (size_t)7 * nq * d.N * d.KXp * cs));

        // vertical tables at the Gauss-Legendre levels
        std::vector<double> z, w, bv;
        gauss_legendre(nq, z, w);
        basis_values(p->L, z, bv);
        auto B = [&](int t, int m, int l) { return bv[((size_t)t * L + m) * nq + l]; };
        std::vector<double> ev((size_t)nq * 4 * L), pr((size_t)nq * 5 * L);
        for (int l = 0; l < nq; ++l)
            for (int m = 0; m < (int)L; ++m) {
                ev[((size_t)l * 4 + 0) * L + m] = B(1, m, l);  // C'
                ev[((size_t)l * 4 + 1) * L + m] = B(0, m, l);  // C
                ev[((size_t)l * 4 + 2) * L + m] = B(4, m, l);  // S
                ev[((size_t)l * 4 + 3) * L + m] = B(5, m, l);  // S'
                pr[((size_t)l * 5 + 0) * L + m] = w[l] * B(1, m, l);
                pr[((size_t)l * 5 + 1) * L + m] = w[l] * B(2, m, l);
                pr[((size_t)l * 5 + 2) * L + m] = w[l] * B(0, m, l);
                pr[((size_t)l * 5 + 3) * L + m] = w[l] * B(4, m, l);
                pr[((size_t)l * 5 + 4) * L + m] = w[l] * B(5, m, l);
            }
        b.evtab = c->dalloc(ev.size() * c->rsz);
        b.prtab = c->dalloc(pr.size() * c->rsz);
        upload_reals(c.get(), b.evtab, ev);
        upload_reals(c.get(), b.prtab, pr);
        // matrix-core operand tables in fragment order (see k_eval_mma / k_project_mma)
        {
            const char* no = getenv("RBC_NO_MFMA");
            const bool off = no && no[0] == '1';
            if (p->precision == RBC_BF16X3) b.mma = 1;
            else if (p->precision == RBC_F32 && p->L <= 32 && !off) b.mma = 0;
            // the matrix-core projection addresses one level combination (2 nq P floats) with
            // 32-bit buffer offsets
            if (b.mma >= 0 && (double)nq * P * 8.0 >= 2147483648.0) {
                if (p->precision == RBC_BF16X3)
                    return fail(nullptr, RBC_ERR_ARG, "BF16X3: grid too large for the 32-bit projection offsets");
                b.mma = -1;
            }
            if (b.mma >= 0) {
                const int K = b.mma == 0 ? 2 : 16, KPL = K / 2;
                b.ev_nlt = (nq + 31) / 32;
                b.ev_nks = ((int)L + K - 1) / K;
                b.pr_nks = (nq + K - 1) / K;
                std::vector<double> eA((size_t)4 * b.ev_nlt * b.ev_nks * 64 * KPL, 0.0);
                static const int evt[4] = {1, 0, 4, 5};  // C', C, S, S'
                for (int t = 0; t < 4; ++t)
                    for (int lt = 0; lt < b.ev_nlt; ++lt)
                        for (int s = 0; s < b.ev_nks; ++s)
                            for (int ln = 0; ln < 64; ++ln)
                                for (int e = 0; e < KPL; ++e) {
                                    const int l = lt * 32 + (ln & 31), m = s * K + KPL * (ln >> 5) + e;
                                    // FP32 (KPL = 1): the four tables interleaved, one float4 per (tile,
                                    // k-step, lane); BF16X3: [table][tile][k-step][lane][KPL]
                                    const size_t at =
                                        b.mma == 0 ? ((((size_t)lt * b.ev_nks + s) * 64 + ln) * 4 + t)
                                                   : ((((size_t)t * b.ev_nlt + lt) * b.ev_nks + s) * 64 + ln) * KPL + e;
                                    if (l < nq && m < (int)L) eA[at] = B(evt[t], m, l);
                                }
                std::vector<double> pA((size_t)5 * b.pr_nks * 64 * KPL, 0.0);
                static const int prt[5] = {1, 2, 0, 4, 5};  // w C', w C'', w C, w S, w S'
                for (int t = 0; t < 5; ++t)
                    for (int s = 0; s < b.pr_nks; ++s)
                        for (int ln = 0; ln < 64; ++ln)
                            for (int e = 0; e < KPL; ++e) {
                                const int m = ln & 31, l = s * K + KPL * (ln >> 5) + e;
                                if (l < nq && m < (int)L)
                                    pA[(((size_t)t * b.pr_nks + s) * 64 + ln) * KPL + e] = w[l] * B(prt[t], m, l);
                            }
                b.evA = c->dalloc(eA.size() * sizeof(float));
                b.prA = c->dalloc(pA.size() * sizeof(float));
                upload_real<float>(c.get(), b.evA, eA);
                upload_real<float>(c.get(), b.prA, pA);
            }
        }
        // twiddles exp(-2 pi i k / n), k < n, built in FP64
        auto tw = [&](int n) {
            std::vector<double> t((size_t)2 * n);
            for (int k = 0; k < n; ++k) {
                t[2 * k] = std::cos(2.0 * M_PI * k / n);
                t[2 * k + 1] = -std::sin(2.0 * M_PI * k / n);
            }
            return t;
        };
        auto tM = tw(p->M), tN = tw(p->N);
        b.twM = c->dalloc(tM.size() * c->rsz);
        b.twN = c->dalloc(tN.size() * c->rsz);
        upload_reals(c.get(), b.twM, tM);
        upload_reals(c.get(), b.twN, tN);
        if (c->rsz == 4 && p->M == 512 && p->N == 512) {
            // the per-pass table of the 512-point kernels: w^(j r) at 64 r + j (r < 8, j < 64),
            // w^(8 j r) at 512 + 8 r + j (r, j < 8), w = exp(-2 pi i / 512)
            std::vector<double> tP((size_t)2 * 576);
            for (int e = 0; e < 576; ++e) {
                const int k = e < 512 ? (e & 63) * (e >> 6) : 8 * (e & 7) * ((e - 512) >> 3);
                tP[2 * e] = tM[2 * k];
                tP[2 * e + 1] = tM[2 * k + 1];
            }
            b.twP = c->dalloc(tP.size() * c->rsz);
            upload_reals(c.get(), b.twP, tP);
        }
        b.kx = c->dalloc(P * c->rsz);
        b.ky = c->dalloc(P * c->rsz);
        b.k2 = c->dalloc(P * c->rsz);
        upload_reals(c.get(), b.kx, c->fac.kx);
        upload_reals(c.get(), b.ky, c->fac.ky);
        upload_reals(c.get(), b.k2, c->fac.k2);
        b.rows = (int*)c->dalloc(g.rows.size() * sizeof(int));
        HIPCHK(hipMemcpy(b.rows, g.rows.data(), g.rows.size() * sizeof(int), hipMemcpyHostToDevice));
        b.gstart = (int*)c->dalloc(c->fac.group_start.size() * sizeof(int));
        b.gmem = (int*)c->dalloc(c->fac.members.size() * sizeof(int));
        HIPCHK(hipMemcpy(b.gstart, c->fac.group_start.data(), c->fac.group_start.size() * sizeof(int),
                         hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.gmem, c->fac.members.data(), c->fac.members.size() * sizeof(int), hipMemcpyHostToDevice));
        {  // chunks of at most kChunk members per wave of the group update: a wave runs its members
           // one after another, so small grids (few modes, a few thousand waves at most) take one
           // member per wave; larger grids take 4, which matches the usual group size (the
           // (+-kx, +-ky) images of a mode) and shares the group's matrix loads (measured at
           // 512^2 x 32: 98 us with 4, 103 us with 2, 116 us with 8 members per chunk)
            const int kChunk = g.P() >= 8192 ? 4 : 1;
            std::vector<int> cg, cb;
            const auto& gs = c->fac.group_start;
            for (int gi = 0; gi + 1 < (int)gs.size(); ++gi)
                for (int m0 = gs[gi]; m0 < gs[gi + 1]; m0 += kChunk) {
                    cg.push_back(gi);
                    cb.push_back(m0);
                }
            cb.push_back(gs.back());
            b.nchunk = (int)cg.size();
            b.cgrp = (int*)c->dalloc(cg.size() * sizeof(int));
            b.cbeg = (int*)c->dalloc(cb.size() * sizeof(int));
            HIPCHK(hipMemcpy(b.cgrp, cg.data(), cg.size() * sizeof(int), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(b.cbeg, cb.data(), cb.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        b.gmat = c->dalloc(c->fac.mats.size() * c->rsz);
        upload_reals(c.get(), b.gmat, c->fac.mats);
        b.nred = (int)((P + 255) / 256) + 2;
        b.red = (double*)c->dalloc((size_t)b.nred * sizeof(double));
        b.health = (int*)c->dalloc(2 * sizeof(int));
        HIPCHK(hipMemset(b.health, 0, 2 * sizeof(int)));
        std::vector<double> iac, ia1;
        closed_form_iac_ia1(p->L, iac, ia1);
        c->ia1 = c->dalloc(ia1.size() * sizeof(double));
        HIPCHK(hipMemcpy(c->ia1, ia1.data(), ia1.size() * sizeof(double), hipMemcpyHostToDevice));
        velocity_setup(c.get());
        HIPCHK(hipDeviceSynchronize());
    } catch (const HipError& e) {
        return fail(nullptr, RBC_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
    } catch (const std::bad_alloc&) {
        return fail(nullptr, RBC_ERR_NOMEM, "host allocation failed");
    }
    *out = c.release();
    return RBC_OK;
}

void rbc_destroy(rbc_ctx* c) {
    if (!c) return;
    hipSetDevice(c->p.device);
    delete c;
}

const char* rbc_last_error(const rbc_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int rbc_get_info(const rbc_ctx* c, int* nq, int* kx_kept, int* ky_kept, int* n_unique_k) {
    if (!c) return RBC_ERR_ARG;
    if (nq) *nq = c->dg.nq;
    if (kx_kept) *kx_kept = c->dg.KX;
    if (ky_kept) *ky_kept = c->dg.NYc;
    if (n_unique_k) *n_unique_k = c->fac.ngroups();
    return RBC_OK;
}

int rbc_set_state(rbc_ctx* c, const double* th, const double* f, const double* g, const double* F, const double* G) {
    if (!c) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        if (th) put_field(c, th, c->b.X[0]);
        if (f) put_field(c, f, c->b.X[1]);
        if (g) put_field(c, g, c->b.X[2]);
        if (F) put_vec(c, F, c->b.XF);
        if (G) put_vec(c, G, c->b.XG);
        clear_health(c);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_get_state(rbc_ctx* c, double* th, double* f, double* g, double* F, double* G) {
    if (!c) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        if (th) get_field(c, c->b.X[0], th);
        if (f) get_field(c, c->b.X[1], f);
        if (g) get_field(c, c->b.X[2], g);
        if (F) get_vec(c, c->b.XF, F);
        if (G) get_vec(c, c->b.XG, G);
        return health_status(c);  // the state is copied out either way
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
}

int rbc_step(rbc_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return RBC_ERR_ARG;
    if (c->bad_step > 0)
        return fail(c, RBC_ERR_NUMERIC,
                    "refusing to step a non-finite state (step " + std::to_string(c->bad_step) +
                        " ended non-finite); replace the state (rbc_set_state) and reset the AB2 history first");
    try {
        hipSetDevice(c->p.device);
        run_steps(c, nsteps);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_reset_history(rbc_ctx* c) {
    if (!c) return RBC_ERR_ARG;
    c->have_hist = false;
    return RBC_OK;
}

int rbc_synchronize(rbc_ctx* c) {
    if (!c) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        return health_status(c);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
}

}  // extern "C"

namespace {

// The operator-level calls on the staging buffers: input in Xo / XoF / XoG, output in the
// buffers returned through outs / oF / oG (Oo for the nonlinear term and inverse B, Xo for the
// implicit solve, which works in place).
template <typename T>
void op_core_t(rbc_ctx* c, int which, void** outs, void** oF, void** oG) {
    Buffers& b = c->b;
    StepCoef sc{0.0, 0.0, 0};
    for (int i = 0; i < 3; ++i) outs[i] = b.Oo[i];
    *oF = b.OoF;
    *oG = b.OoG;
    if (which == 0) {
        nonlinear_into<T>(c, b.Xo, b.XoF, b.XoG, b.Oo, b.OoF, b.OoG, -1);
    } else if (which == 1) {  // inverse B: N in Xo -> R in Oo
        Launch<T>::update(c->dg, b, c->fac.ngroups(), c->fac.L0, c->fac.L1, c->fac.stride, c->p.Pr, c->p.dt, sc,
                          UPD_INVB, b.Xo, b.Xo, b.Oo, c->stream);
        Launch<T>::update_FG(c->dg, c->p.Pr, c->p.dt, sc, UPD_INVB, b.XoF, b.XoG, b.XoF, b.XoG, b.OoF, b.OoG,
                             c->stream);
    } else {  // implicit: in place on Xo
        Launch<T>::update(c->dg, b, c->fac.ngroups(), c->fac.L0, c->fac.L1, c->fac.stride, c->p.Pr, c->p.dt, sc,
                          UPD_IMPLICIT, b.Xo, b.Xo, b.Oo, c->stream);
        Launch<T>::update_FG(c->dg, c->p.Pr, c->p.dt, sc, UPD_IMPLICIT, b.XoF, b.XoG, b.XoF, b.XoG, b.OoF, b.OoG,
                             c->stream);
        for (int i = 0; i < 3; ++i) outs[i] = b.Xo[i];
        *oF = b.XoF;
        *oG = b.XoG;
    }
    HIPCHK(hipGetLastError());
}

void op_core(rbc_ctx* c, int which, void** outs, void** oF, void** oG) {
    if (c->rsz == 8) op_core_t<double>(c, which, outs, oF, oG);
    else op_core_t<float>(c, which, outs, oF, oG);
}

int op_call(rbc_ctx* c, int which, const double* th, const double* f, const double* g, const double* F,
                   const double* G, double* oth, double* of, double* og, double* oF, double* oG) {
    if (!c || !th || !f || !g || !F || !G || !oth || !of || !og || !oF || !oG) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        Buffers& b = c->b;
        put_field(c, th, b.Xo[0]);
        put_field(c, f, b.Xo[1]);
        put_field(c, g, b.Xo[2]);
        put_vec(c, F, b.XoF);
        put_vec(c, G, b.XoG);
        void* outs[3];
        void *oFp, *oGp;
        op_core(c, which, outs, &oFp, &oGp);
        get_field(c, outs[0], oth);
        get_field(c, outs[1], of);
        get_field(c, outs[2], og);
        get_vec(c, oFp, oF);
        get_vec(c, oGp, oG);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

// device buffers in the reference layout <-> the context's compact fields, on its stream
void dev_in(rbc_ctx* c, const void* const f3[3], const void* F, const void* G, void* const X[3], void* XF,
                   void* XG) {
    if (c->rsz == 8) Launch<double>::from_ref(c->dg, f3, X, c->stream);
    else Launch<float>::from_ref(c->dg, f3, X, c->stream);
    HIPCHK(hipGetLastError());
    const size_t vb = (size_t)c->dg.L * c->csz();
    if (F) HIPCHK(hipMemcpyAsync(XF, F, vb, hipMemcpyDeviceToDevice, c->stream));
    if (G) HIPCHK(hipMemcpyAsync(XG, G, vb, hipMemcpyDeviceToDevice, c->stream));
}

void dev_out(rbc_ctx* c, void* const X[3], const void* XF, const void* XG, void* const f3[3], void* F,
                    void* G) {
    if (c->rsz == 8) Launch<double>::to_ref(c->dg, X, f3, c->stream);
    else Launch<float>::to_ref(c->dg, X, f3, c->stream);
    HIPCHK(hipGetLastError());
    const size_t vb = (size_t)c->dg.L * c->csz();
    if (F) HIPCHK(hipMemcpyAsync(F, XF, vb, hipMemcpyDeviceToDevice, c->stream));
    if (G) HIPCHK(hipMemcpyAsync(G, XG, vb, hipMemcpyDeviceToDevice, c->stream));
}

int op_call_device(rbc_ctx* c, int which, const void* th, const void* f, const void* g, const void* F,
                          const void* G, void* oth, void* of, void* og, void* oF, void* oG) {
    if (!c || !th || !f || !g || !F || !G || !oth || !of || !og || !oF || !oG) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        Buffers& b = c->b;
        const void* in3[3] = {th, f, g};
        dev_in(c, in3, F, G, b.Xo, b.XoF, b.XoG);
        void* outs[3];
        void *oFp, *oGp;
        op_core(c, which, outs, &oFp, &oGp);
        void* out3[3] = {oth, of, og};
        dev_out(c, outs, oFp, oGp, out3, oF, oG);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

}  // namespace

extern "C" {

int rbc_nonlinear(rbc_ctx* c, const double* th, const double* f, const double* g, const double* F, const double* G,
                  double* oth, double* of, double* og, double* oF, double* oG) {
    return op_call(c, 0, th, f, g, F, G, oth, of, og, oF, oG);
}
int rbc_inverse_B(rbc_ctx* c, const double* th, const double* f, const double* g, const double* F, const double* G,
                  double* oth, double* of, double* og, double* oF, double* oG) {
    return op_call(c, 1, th, f, g, F, G, oth, of, og, oF, oG);
}
int rbc_implicit(rbc_ctx* c, const double* th, const double* f, const double* g, const double* F, const double* G,
                 double* oth, double* of, double* og, double* oF, double* oG) {
    return op_call(c, 2, th, f, g, F, G, oth, of, og, oF, oG);
}

int rbc_set_state_device(rbc_ctx* c, const void* th, const void* f, const void* g, const void* F, const void* G) {
    if (!c) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        const void* in3[3] = {th, f, g};
        dev_in(c, in3, F, G, c->b.X, c->b.XF, c->b.XG);
        clear_health(c);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_get_state_device(rbc_ctx* c, void* th, void* f, void* g, void* F, void* G) {
    if (!c) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        void* out3[3] = {th, f, g};
        dev_out(c, c->b.X, c->b.XF, c->b.XG, out3, F, G);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_nonlinear_device(rbc_ctx* c, const void* th, const void* f, const void* g, const void* F, const void* G,
                         void* oth, void* of, void* og, void* oF, void* oG) {
    return op_call_device(c, 0, th, f, g, F, G, oth, of, og, oF, oG);
}
int rbc_inverse_B_device(rbc_ctx* c, const void* th, const void* f, const void* g, const void* F, const void* G,
                         void* oth, void* of, void* og, void* oF, void* oG) {
    return op_call_device(c, 1, th, f, g, F, G, oth, of, og, oF, oG);
}
int rbc_implicit_device(rbc_ctx* c, const void* th, const void* f, const void* g, const void* F, const void* G,
                        void* oth, void* of, void* og, void* oF, void* oG) {
    return op_call_device(c, 2, th, f, g, F, G, oth, of, og, oF, oG);
}

void* rbc_stream(rbc_ctx* c) { return c ? (void*)c->stream : nullptr; }

int rbc_init_state(rbc_ctx* c, int kind, unsigned long long seed, double amplitude) {
    if (!c || kind < RBC_INIT_ZERO || kind > RBC_INIT_TEST || !std::isfinite(amplitude)) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        if (c->rsz == 8) {
            Launch<double>::init_state(c->dg, c->b, kind, seed, amplitude, c->stream);
            Launch<double>::make_real(c->dg, c->b, c->b.X, c->stream);  // main.cu:396
        } else {
            Launch<float>::init_state(c->dg, c->b, kind, seed, amplitude, c->stream);
            Launch<float>::make_real(c->dg, c->b, c->b.X, c->stream);
        }
        HIPCHK(hipGetLastError());
        c->have_hist = false;
        clear_health(c);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_set_penalization(rbc_ctx* c, double eta, int shape, const double* mask) {
    if (!c || !std::isfinite(eta) || eta < 0.0) return RBC_ERR_ARG;
    if (eta > 0.0 && (shape < RBC_MASK_RECTANGLE || shape > RBC_MASK_CUSTOM || (shape == RBC_MASK_CUSTOM && !mask)))
        return fail(c, RBC_ERR_ARG, "unknown mask shape (or no mask array for RBC_MASK_CUSTOM)");
    try {
        hipSetDevice(c->p.device);
        HIPCHK(hipStreamSynchronize(c->stream));  // steps in flight may read the old arrays
        for (void* a : c->pen_allocs) HIPCHK(hipFree(a));
        c->pen_allocs.clear();
        c->pen = Penalty();
        if (eta == 0.0) return RBC_OK;
        const DevGeom& g = c->dg;
        auto al = [&](size_t bytes) {
            void* ptr = nullptr;
            HIPCHK(hipMalloc(&ptr, bytes > 0 ? bytes : 16));
            c->pen_allocs.push_back(ptr);
            return ptr;
        };
        std::vector<double> m;
        if (shape == RBC_MASK_CUSTOM) m.assign(mask, mask + (size_t)g.M * g.N);
        else penalty_mask(shape, g.M, g.N, m);
        std::vector<double> iac, ia1, ia2;
        closed_form_iac_ia1(g.L, iac, ia1);
        quadrature_ia2(g.L, ia2);
        Penalty& pen = c->pen;
        pen.mask = al(m.size() * c->rsz);
        pen.iac = al(iac.size() * c->rsz);
        pen.ia2 = al(ia2.size() * c->rsz);
        upload_reals(c, pen.mask, m);
        upload_reals(c, pen.iac, iac);
        upload_reals(c, pen.ia2, ia2);
        // work arrays: the level-space buffers when they are large enough (the penalty runs after
        // the projection has consumed them), else arrays of their own
        const size_t cs = c->csz(), spec = (size_t)10 * g.L * g.P * cs, hyb = (size_t)10 * g.L * g.N * g.KXp * cs;
        const bool reuse = 10 * g.L <= 6 * g.nq;
        pen.spec = reuse ? c->b.lev : al(spec);
        pen.spec_out = reuse ? c->b.prod : al(spec);
        pen.hin = reuse ? c->b.hin : al(hyb);
        pen.hout = reuse ? c->b.hout : al(hyb);
        HIPCHK(hipDeviceSynchronize());
        pen.eta = eta;
    } catch (const HipError& e) {
        c->pen = Penalty();
        return hip_fail(c, e);
    }
    return RBC_OK;
}

}  // extern "C"

namespace {

// integer masks of the plane operators (1 = the operator leaves the point alone), index
// col + row M; real arithmetic in the context's real type as the reference's kernels use
// CUDA_FLOAT_REAL (freeze_operator.cu, masktemp.cu, masking.cu init_mask_physical_space_*)
template <typename R>
std::vector<int> plane_mask(int shape, int M, int N, double param, double cxf, double cyf) {
    std::vector<int> m((size_t)M * N, 0);
    for (int row = 0; row < N; ++row)
        for (int col = 0; col < M; ++col) {
            int v = 0;
            if (shape == RBC_PLANE_RECTANGLE) {
                const R w = (R)param;
                if (col < w * M || col > (1 - w) * M || row < w * N || row > (1 - w) * N) v = 1;
            } else {  // circle around (cxf M, cyf N), radius param * min(M, N) / 2
                const R dx = (R)col - (R)cxf * M, dy = (R)row - (R)cyf * N;
                const R d2 = dx * dx + dy * dy;
                const int mn = N < M ? N : M;
                const R r = (R)param * (R)0.5 * mn;
                if (d2 > r * r) v = 1;
            }
            m[(size_t)row * M + col] = v;
        }
    return m;
}

int* upload_mask(rbc_ctx* c, int*& dst, const std::vector<int>& m) {
    if (!dst) dst = (int*)c->dalloc(m.size() * sizeof(int));
    HIPCHK(hipMemcpy(dst, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice));
    return dst;
}

void run_plane_op(rbc_ctx* c, int op, int nf, const int* imask, double factor, void* planes = nullptr) {
    Buffers& b = c->b;
    void* fr = planes ? planes : c->frozen;
    if (c->rsz == 8)
        Launch<double>::plane_op(c->dg, b, op, nf, imask, factor, fr, b.X, b.lev, b.hin, b.hout, b.prod, c->stream);
    else
        Launch<float>::plane_op(c->dg, b, op, nf, imask, factor, fr, b.X, b.lev, b.hin, b.hout, b.prod, c->stream);
    HIPCHK(hipGetLastError());
}

bool plane_buffers_fit(const rbc_ctx* c, int nf) { return nf * c->dg.L <= 6 * c->dg.nq; }

}  // namespace

extern "C" {

int rbc_init_pattern(rbc_ctx* c, int kind, double init_mode, unsigned long long seed) {
    if (!c || kind < RBC_PATTERN_RECTANGLE || kind > RBC_PATTERN_GIANTSPIRAL || !std::isfinite(init_mode))
        return RBC_ERR_ARG;
    if (!plane_buffers_fit(c, 1)) return fail(c, RBC_ERR_ARG, "initial patterns need nq >= L / 6");
    int rc = rbc_init_state(c, RBC_INIT_RANDOM, seed, 1e-6);  // the random start of every level (main.cu:300)
    if (rc != RBC_OK) return rc;
    void* plane = nullptr;
    try {
        hipSetDevice(c->p.device);
        const DevGeom& g = c->dg;
        // the mode arguments are int parameters of the init_* functions (main.cu:329-345)
        const int mode = kind == RBC_PATTERN_RECTANGLE ? (int)((c->p.Lx / (g.M - 1.0)) * (g.M - 5.0) / 2.0)
                                                       : (int)init_mode;
        HIPCHK(hipMalloc(&plane, (size_t)g.M * g.N * c->rsz));
        if (c->rsz == 8) Launch<double>::init_pattern(g, kind, mode, 1e-6, seed, plane, c->stream);
        else Launch<float>::init_pattern(g, kind, mode, 1e-6, seed, plane, c->stream);
        run_plane_op(c, PLANE_SET, 1, nullptr, 1.0, plane);
        if (c->rsz == 8) Launch<double>::make_real(g, c->b, c->b.X, c->stream);
        else Launch<float>::make_real(g, c->b, c->b.X, c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(plane));
    } catch (const HipError& e) {
        if (plane) (void)hipFree(plane);
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_freeze(rbc_ctx* c, int shape, double param) {
    if (!c || (shape != RBC_PLANE_CIRCLE && shape != RBC_PLANE_RECTANGLE) || !std::isfinite(param))
        return RBC_ERR_ARG;
    if (!plane_buffers_fit(c, 3)) return fail(c, RBC_ERR_ARG, "plane operators need nq >= L / 2");
    try {
        hipSetDevice(c->p.device);
        const DevGeom& g = c->dg;
        const auto m = c->rsz == 8 ? plane_mask<double>(shape, g.M, g.N, param, 0.5, 0.5)
                                   : plane_mask<float>(shape, g.M, g.N, param, 0.5, 0.5);
        HIPCHK(hipStreamSynchronize(c->stream));
        upload_mask(c, c->freeze_mask, m);
        if (!c->frozen) c->frozen = c->dalloc((size_t)3 * g.L * g.M * g.N * c->rsz);
        run_plane_op(c, PLANE_CAPTURE, 3, c->freeze_mask, 1.0);
        c->frozen_set = true;
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_apply_freeze(rbc_ctx* c) {
    if (!c) return RBC_ERR_ARG;
    if (!c->frozen_set) return RBC_OK;  // freeze_operator::calculate_operator before freeze_*: nothing
    try {
        hipSetDevice(c->p.device);
        run_plane_op(c, PLANE_FREEZE, 3, c->freeze_mask, 1.0);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_temp_mask(rbc_ctx* c, int shape, double param) {
    if (!c || (shape != RBC_PLANE_CIRCLE && shape != RBC_PLANE_RECTANGLE) || !std::isfinite(param))
        return RBC_ERR_ARG;
    if (!plane_buffers_fit(c, 1)) return fail(c, RBC_ERR_ARG, "plane operators need nq >= L / 6");
    try {
        hipSetDevice(c->p.device);
        const DevGeom& g = c->dg;
        // masktemp's circle is centred at (0.2 M, 0.3 N) (masktemp.cu init_mask_physical_space_circle)
        const auto m = c->rsz == 8 ? plane_mask<double>(shape, g.M, g.N, param, 0.2, 0.3)
                                   : plane_mask<float>(shape, g.M, g.N, param, 0.2, 0.3);
        HIPCHK(hipStreamSynchronize(c->stream));
        upload_mask(c, c->temp_mask, m);
        c->temp_set = true;
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_apply_temp_mask(rbc_ctx* c) {
    if (!c) return RBC_ERR_ARG;
    if (!c->temp_set) return RBC_OK;  // masktemp::calculate_operator before temp_*: nothing
    try {
        hipSetDevice(c->p.device);
        run_plane_op(c, PLANE_HEAT, 1, c->temp_mask, 1.0);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_mask_theta(rbc_ctx* c, double cx, double cy, double radius, double factor) {
    if (!c || !std::isfinite(cx) || !std::isfinite(cy) || !std::isfinite(radius) || !std::isfinite(factor))
        return RBC_ERR_ARG;
    if (!plane_buffers_fit(c, 1)) return fail(c, RBC_ERR_ARG, "plane operators need nq >= L / 6");
    try {
        hipSetDevice(c->p.device);
        const DevGeom& g = c->dg;
        // masking.cu: the circle centre is (c_y M, c_x N), the coordinates swapped as there
        const auto m = c->rsz == 8 ? plane_mask<double>(RBC_PLANE_CIRCLE, g.M, g.N, radius, cy, cx)
                                   : plane_mask<float>(RBC_PLANE_CIRCLE, g.M, g.N, radius, cy, cx);
        HIPCHK(hipStreamSynchronize(c->stream));
        upload_mask(c, c->theta_mask, m);
        run_plane_op(c, PLANE_SCALE, 1, c->theta_mask, factor);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_temperature_grid(rbc_ctx* c, int nz, const double* z, float* out) {
    if (!c || nz < 1 || !z || !out) return RBC_ERR_ARG;
    if (!plane_buffers_fit(c, 1)) return fail(c, RBC_ERR_ARG, "the temperature grid needs nq >= L / 6");
    try {
        hipSetDevice(c->p.device);
        const DevGeom& g = c->dg;
        const size_t nxy = (size_t)g.M * g.N;
        // the physical mode planes of theta (the reference's C2R per mode) into the context's
        // product buffer, then per height the S-weighted sum into a float plane of the hybrid buffer
        Buffers& b = c->b;
        if (c->rsz == 8)
            Launch<double>::plane_op(g, b, PLANE_CAPTURE, 1, nullptr, 1.0, b.prod, b.X, b.lev, b.hin, b.hout, nullptr,
                                     c->stream);
        else
            Launch<float>::plane_op(g, b, PLANE_CAPTURE, 1, nullptr, 1.0, b.prod, b.X, b.lev, b.hin, b.hout, nullptr,
                                    c->stream);
        float* dS = (float*)b.fgw;  // L floats; fgw holds 2 nq complex >= L floats
        float* dout = (float*)b.hout;
        std::vector<float> S(g.L), plane(nxy);
        for (int k = 0; k < nz; ++k) {
            // S_i = sin(pi (i + 1) (z + 1/2)) with z in the build's real type (:59-65)
            const float zk = (float)z[k];
            for (int i = 0; i < g.L; ++i) S[i] = (float)std::sin(M_PI * (i + 1) * (zk + 0.5));
            HIPCHK(hipStreamSynchronize(c->stream));
            HIPCHK(hipMemcpy(dS, S.data(), S.size() * sizeof(float), hipMemcpyHostToDevice));
            if (c->rsz == 8) Launch<double>::temperature_sum(g, b.prod, dS, dout, c->stream);
            else Launch<float>::temperature_sum(g, b.prod, dS, dout, c->stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(out + (size_t)k * nxy, dout, nxy * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_force_mean(rbc_ctx* c, double mean_F) {
    if (!c || !std::isfinite(mean_F)) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        if (c->rsz == 8) Launch<double>::force_mean(c->b, mean_F, c->stream);
        else Launch<float>::force_mean(c->b, mean_F, c->stream);
        HIPCHK(hipGetLastError());
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_nusselt(rbc_ctx* c, double* nu3) {
    if (!c || !nu3) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        if (c->rsz == 8) Launch<double>::nusselt(c->dg, c->b, c->ia1, c->stream);
        else Launch<float>::nusselt(c->dg, c->b, c->ia1, c->stream);
        HIPCHK(hipGetLastError());
        std::vector<double> red(c->b.nred);
        HIPCHK(hipMemcpyAsync(red.data(), c->b.red, red.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        const int nblk = c->b.nred - 2;
        double s = 0.0;
        for (int i = 0; i < nblk; ++i) s += red[i];
        nu3[0] = 1.0 - red[nblk];
        nu3[1] = 1.0 - red[nblk + 1];
        nu3[2] = 1.0 + s;
        return health_status(c);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
}

int rbc_profile_steps(rbc_ctx* c, int nsteps, double* ms) {
    if (!c || !ms || nsteps < 1) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        for (int i = 0; i < RBC_NSTAGES; ++i) ms[i] = 0.0;
        // record up to kRing steps of stage events, then read them back; the host waits only
        // once per kRing steps, so the queue stays full as in rbc_step.
        for (int s0 = 0; s0 < nsteps; s0 += rbc_ctx::kRing) {
            const int n = std::min(rbc_ctx::kRing, nsteps - s0);
            for (int s = 0; s < n; ++s) {
                c->ev = c->ev_ring[s];
                if (c->rsz == 8) one_step<double>(c, true);
                else one_step<float>(c, true);
            }
            HIPCHK(hipEventSynchronize(c->ev_ring[n - 1][RBC_NSTAGES]));
            for (int s = 0; s < n; ++s)
                for (int i = 0; i < RBC_NSTAGES; ++i) {
                    float t = 0.f;
                    HIPCHK(hipEventElapsedTime(&t, c->ev_ring[s][i], c->ev_ring[s][i + 1]));
                    ms[i] += t;
                }
        }
        c->ev = c->ev_ring[0];
        for (int i = 0; i < RBC_NSTAGES; ++i) ms[i] /= nsteps;
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

const char* rbc_stage_name(int i) { return (i >= 0 && i < RBC_NSTAGES) ? kStageNames[i] : ""; }

}  // extern "C"
