// rbc_setup.cpp -- FP64 host setup of libsdc_rbc: vertical basis at the collocation levels,
// the closed-form integrals the linear operators need, and the per-wavenumber implicit and
// inverse-B factors (built once, as the reference builds them in its operator
// constructors).
#include <algorithm>
#include <cmath>
#include <map>
#include <thread>

#include "rbc_internal.hpp"

namespace rbc {

double lambda_k(int k) {
    // Roots of the Chandrasekhar boundary condition as tabulated by the reference
    // (util/coeff.cu:1232-1257): lambda_k / pi to 12 digits for k <= 8, pi (k + 1/2) beyond.
    static const double roots_over_pi[9] = {0.0,            1.505618731142, 2.499752670074,
                                            3.500010679436, 4.499999538484, 5.500000019944,
                                            6.499999999138, 7.500000000037, 8.499999999998};
    if ((unsigned)k > 8u) return M_PI * (k + 0.5);
    return M_PI * roots_over_pi[k];
}

void gauss_legendre(int n, std::vector<double>& z, std::vector<double>& w) {
    z.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double x = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double p0 = 1.0, p1 = x, dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            p0 = 1.0;
            p1 = x;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) p0 = 1.0, p1 = x;
            dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        p0 = 1.0;
        p1 = x;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        if (n == 1) p0 = 1.0, p1 = x;
        dp = n * (x * p1 - p0) / (x * x - 1.0);
        z[n - 1 - i] = 0.5 * x;
        w[n - 1 - i] = 1.0 / ((1.0 - x * x) * dp * dp);
    }
}

void basis_values(int L, const std::vector<double>& z, std::vector<double>& out) {
    const int nz = (int)z.size();
    out.assign((size_t)6 * L * nz, 0.0);
    auto at = [&](int t, int m, int q) -> double& { return out[((size_t)t * L + m) * nz + q]; };
    for (int m = 0; m < L; ++m) {
        const int k = m + 1;
        const double l = lambda_k(k);
        for (int q = 0; q < nz; ++q) {
            const double x = z[q];
            if (k % 2) {  // cosh/cos pair (coeff.h:31-36)
                double ch = std::cosh(0.5 * l), c = std::cos(0.5 * l);
                double a = std::cosh(l * x) / ch, b = std::cos(l * x) / c;
                double as = std::sinh(l * x) / ch, bs = std::sin(l * x) / c;
                at(0, m, q) = a - b;
                at(1, m, q) = l * (as + bs);
                at(2, m, q) = l * l * (a + b);
                at(3, m, q) = l * l * l * (as - bs);
            } else {  // sinh/sin pair
                double sh = std::sinh(0.5 * l), s = std::sin(0.5 * l);
                double a = std::sinh(l * x) / sh, b = std::sin(l * x) / s;
                double ac = std::cosh(l * x) / sh, bc = std::cos(l * x) / s;
                at(0, m, q) = a - b;
                at(1, m, q) = l * (ac - bc);
                at(2, m, q) = l * l * (a + b);
                at(3, m, q) = l * l * l * (ac + bc);
            }
            const double kp = k * M_PI, arg = kp * (x + 0.5);
            at(4, m, q) = std::sin(arg);  // S_k = sin(k pi (z + 1/2)), coeff.h:27
            at(5, m, q) = kp * std::cos(arg);
        }
    }
}

void closed_form_iac_ia1(int L, std::vector<double>& Iac, std::vector<double>& Ia1) {
    Iac.assign((size_t)L * L, 0.0);
    Ia1.assign((size_t)L * L, 0.0);
    // Iac = <C_i C''_j>: odd (1-based) pairs from lambda tan(lambda/2) (coeff.cu:77-95); the
    // even diagonal from lambda cot(lambda/2) (:97-106).  The even-even off-diagonal entries
    // are (lambdacotj - lambdacotj) = 0 in the reference (:112) and stay 0 here.
    for (int i = 1; i <= L; i += 2) {
        const double li = lambda_k(i), lti = li * std::tan(0.5 * li), li2 = li * li;
        Iac[(i - 1) + (size_t)(i - 1) * L] = -2.0 * lti - lti * lti;
        for (int j = i + 2; j <= L; j += 2) {
            const double lj = lambda_k(j), ltj = lj * std::tan(0.5 * lj), lj2 = lj * lj;
            const double rij = 8.0 * li2 * lj2 * (ltj - lti) / (li2 * li2 - lj2 * lj2);
            Iac[(i - 1) + (size_t)(j - 1) * L] = rij;
            Iac[(j - 1) + (size_t)(i - 1) * L] = rij;
        }
    }
    for (int i = 2; i <= L; i += 2) {
        const double li = lambda_k(i), lci = li / std::tan(0.5 * li);
        Iac[(i - 1) + (size_t)(i - 1) * L] = 2.0 * lci - lci * lci;
    }
    // Ia1 = <C_i S_j> (coeff.cu:146-178): nonzero for equal parity only.
    for (int i = 1; i <= L; ++i) {
        const double zi = lambda_k(i) / M_PI, zi2 = zi * zi;
        const double sgn = (i % 2) ? -1.0 : 1.0;
        for (int j = (i % 2) ? 1 : 2; j <= L; j += 2) {
            const int j2 = j * j;
            Ia1[(i - 1) + (size_t)(j - 1) * L] = sgn * 4.0 * j * zi2 / (M_PI * (j2 * j2 - zi2 * zi2));
        }
    }
}

void quadrature_ia2(int L, std::vector<double>& Ia2) {
    const int n = 4 * L + 64;
    std::vector<double> z, w, b;
    gauss_legendre(n, z, w);
    basis_values(L, z, b);
    Ia2.assign((size_t)L * L, 0.0);
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
            double s = 0.0;
            for (int q = 0; q < n; ++q) s += w[q] * b[((size_t)0 * L + i) * n + q] * b[((size_t)5 * L + j) * n + q];
            Ia2[i + (size_t)j * L] = s;
        }
}

void penalty_mask(int shape, int M, int N, std::vector<double>& mask) {
    // columns = M (x), rows = N (y); index col + row * M (get_current_matrix_indices, :1416-1424)
    mask.assign((size_t)M * N, 0.0);
    for (int row = 0; row < N; ++row)
        for (int col = 0; col < M; ++col) {
            double v = 0.0;
            if (shape == RBC_MASK_RECTANGLE) {  // three grid points along every side (:1654-1671)
                if (col < 3 || col > M - 4 || row < 3 || row > N - 4) v = 1.0;
            } else if (shape == RBC_MASK_CIRCLE) {  // outside the inscribed circle less 3 points (:1673-1693)
                const int dx = col - M / 2, dy = row - N / 2, r = std::min(M, N) / 2 - 3;
                if (dx * dx + dy * dy > r * r) v = 1.0;
            } else if (shape == RBC_MASK_DISKARRAY) {  // 3 x 2 disks (:1695-1727)
                const int nx = 3, ny = 2;
                const int r = std::min(M / (nx * 2), N / (ny * 2));
                for (int i = 0; i < nx; ++i)
                    for (int j = 0; j < ny; ++j) {
                        const int cx = i * M / nx + M / (nx * 2), cy = j * N / ny + N / (ny * 2);
                        // the squared distance is compared with the radius itself, as there
                        if ((cx - col) * (cx - col) + (cy - row) * (cy - row) < r) v = 1.0;
                    }
            } else {  // RBC_MASK_EVERYWHERE (:1729-1738)
                v = 1.0;
            }
            mask[(size_t)row * M + col] = v;
        }
}

int exact_level_count(int L, double tol, double* achieved) {
    // Compare the products the nonlinear term projects (C', C, C'', S, S' combinations)
    // integrated with n points against a 4L+64 point rule.
    const int nref = 4 * L + 64;
    std::vector<double> zr, wr, br;
    gauss_legendre(nref, zr, wr);
    basis_values(L, zr, br);
    // triples (t_i, t_j, t_k) used by the Galerkin nonlinear term: Ibb'(C'C'C'), Ibb1,
    // Ib11, Iabc, Iac1, Iaab, Iaa1, I111, Ia12, Iab2
    static const int trip[10][3] = {{1, 1, 1}, {1, 1, 4}, {1, 4, 4}, {0, 1, 2}, {0, 2, 4},
                                    {0, 0, 1}, {0, 0, 4}, {4, 4, 4}, {0, 4, 5}, {0, 1, 5}};
    auto tensors = [&](const std::vector<double>& w, const std::vector<double>& b, int nz,
                       std::vector<double>& out) {
        out.assign((size_t)10 * L * L * L, 0.0);
        for (int t = 0; t < 10; ++t)
            for (int k = 0; k < L; ++k)
                for (int j = 0; j < L; ++j)
                    for (int i = 0; i < L; ++i) {
                        const double* x = &b[((size_t)trip[t][0] * L + i) * nz];
                        const double* y = &b[((size_t)trip[t][1] * L + j) * nz];
                        const double* u = &b[((size_t)trip[t][2] * L + k) * nz];
                        double s = 0.0;
                        for (int q = 0; q < nz; ++q) s += w[q] * x[q] * y[q] * u[q];
                        out[(((size_t)t * L + k) * L + j) * L + i] = s;
                    }
    };
    std::vector<double> Tref;
    tensors(wr, br, nref, Tref);
    std::vector<double> scale(10, 0.0);
    for (int t = 0; t < 10; ++t)
        for (size_t e = 0; e < (size_t)L * L * L; ++e)
            scale[t] = std::max(scale[t], std::fabs(Tref[(size_t)t * L * L * L + e]));
    for (int n = 8; n <= nref; n += 8) {
        std::vector<double> z, w, b, T;
        gauss_legendre(n, z, w);
        basis_values(L, z, b);
        tensors(w, b, n, T);
        double err = 0.0;
        for (int t = 0; t < 10; ++t)
            for (size_t e = 0; e < (size_t)L * L * L; ++e)
                err = std::max(err, std::fabs(T[(size_t)t * L * L * L + e] - Tref[(size_t)t * L * L * L + e]) / scale[t]);
        if (err <= tol) {
            if (achieved) *achieved = err;
            return n;
        }
    }
    if (achieved) *achieved = -1.0;
    return nref;
}

Geometry make_geometry(int M, int N, int L) {
    Geometry g;
    g.M = M;
    g.N = N;
    g.L = L;
    g.mox = M / 2 + 1;
    g.KX = 0;
    for (int col = 0; col < g.mox; ++col)
        if (!(g.mox - col <= (1.0 / 3.0) * g.mox)) g.KX = col + 1;
    for (int r = 0; r < N; ++r)
        if (!(std::fabs(0.5 * N - r) <= (1.0 / 3.0) * 0.5 * N)) g.rows.push_back(r);
    return g;
}

// Partial-pivot LU inverse of an n x n row-major matrix (GSL LU_decomp + LU_invert restated).
static bool lu_invert(int n, std::vector<double>& a) {
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int j = 0; j < n - 1; ++j) {
        int p = j;
        double amax = std::fabs(a[(size_t)j * n + j]);
        for (int i = j + 1; i < n; ++i)
            if (std::fabs(a[(size_t)i * n + j]) > amax) amax = std::fabs(a[(size_t)i * n + j]), p = i;
        if (p != j) {
            for (int c = 0; c < n; ++c) std::swap(a[(size_t)j * n + c], a[(size_t)p * n + c]);
            std::swap(perm[j], perm[p]);
        }
        const double piv = a[(size_t)j * n + j];
        if (piv != 0.0)
            for (int i = j + 1; i < n; ++i) {
                const double l = a[(size_t)i * n + j] / piv;
                a[(size_t)i * n + j] = l;
                for (int c = j + 1; c < n; ++c) a[(size_t)i * n + c] -= l * a[(size_t)j * n + c];
            }
    }
    for (int i = 0; i < n; ++i)
        if (a[(size_t)i * n + i] == 0.0) return false;
    std::vector<double> inv((size_t)n * n);
    std::vector<double> x(n);
    for (int col = 0; col < n; ++col) {
        for (int i = 0; i < n; ++i) x[i] = (perm[i] == col) ? 1.0 : 0.0;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < i; ++k) x[i] -= a[(size_t)i * n + k] * x[k];
        for (int i = n - 1; i >= 0; --i) {
            for (int k = i + 1; k < n; ++k) x[i] -= a[(size_t)i * n + k] * x[k];
            x[i] /= a[(size_t)i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[(size_t)i * n + col] = x[i];
    }
    a.swap(inv);
    return true;
}

int build_factors(const Geometry& g, double Lx, double Ly, double Ra, double Pr, double dt,
                  Factors& F, std::string& err) {
    const int L = g.L;
    std::vector<double> Iac, Ia1;
    closed_form_iac_ia1(L, Iac, Ia1);
    std::vector<int> cls[2];
    for (int m = 0; m < L; ++m) cls[m % 2].push_back(m);
    F.L0 = (int)cls[0].size();
    F.L1 = (int)cls[1].size();
    const int n0 = 2 * F.L0, n1 = 2 * F.L1;
    F.stride = n0 * n0 + n1 * n1 + F.L0 * F.L0 + F.L1 * F.L1;

    const int NYc = g.NYc(), P = g.P();
    F.k2.assign(P, 0.0);
    F.kx.assign(P, 0.0);
    F.ky.assign(P, 0.0);
    // Group compact modes by the integer key the factor depends on: |k|^2 is a function of
    // (kx^2, ky^2), and of kx^2 + ky^2 alone when Lx == Ly.
    std::map<long long, std::vector<int>> groups;
    for (int kx = 0; kx < g.KX; ++kx)
        for (int c = 0; c < NYc; ++c) {
            const int p = kx * NYc + c;
            const int row = g.rows[c];
            // wave numbers as the reference's kernels form them (inverse_B_matrix.cu:187-194)
            const int r = (row >= g.N / 2 + 1) ? row - g.N : row;
            const double wx = (M_PI + M_PI) * (kx / Lx), wy = (M_PI + M_PI) * (r / Ly);
            F.kx[p] = wx;
            F.ky[p] = wy;
            F.k2[p] = wx * wx + wy * wy;
            long long key = (Lx == Ly) ? (long long)kx * kx + (long long)r * r
                                       : (long long)kx * kx * 4194304LL + (long long)r * r;
            groups[key].push_back(p);
        }
    F.group_start.clear();
    F.members.clear();
    F.group_start.push_back(0);
    std::vector<std::pair<int, int>> reps;  // (kx, row) of each group's first member
    for (auto& kv : groups) {
        for (int p : kv.second) F.members.push_back(p);
        F.group_start.push_back((int)F.members.size());
        const int p0 = kv.second.front();
        reps.push_back({p0 / NYc, g.rows[p0 % NYc]});
    }
    const int ng = F.ngroups();
    F.mats.assign((size_t)ng * F.stride, 0.0);

    int status = 0;
    auto work = [&](int g0, int g1) {
        const int n2 = 2 * L;
        std::vector<double> invB, tmpB, Lm, C, Bf;
        for (int gi = g0; gi < g1; ++gi) {
            const int k = reps[gi].first, l = reps[gi].second;
            // mode numbers as linear_implicit_operator.cu:49-60 forms them
            const double mode_x = k * 2.0 * M_PI / Lx;
            double mode_y = 2.0 * M_PI / Ly;
            if (l < (g.N / 2) + 1) mode_y *= l;
            else mode_y *= (g.N - l);
            const double ms = mode_x * mode_x + mode_y * mode_y;
            const bool zero = (k == 0 && l == 0);
            // inv_B (2L x 2L): theta block 2I; f block (k^2 (k^2 I - Iac)/Pr)^-1 (:70-104)
            invB.assign((size_t)n2 * n2, 0.0);
            std::vector<double> ident((size_t)n2 * n2, 0.0);
            for (int i = 0; i < n2; ++i) ident[(size_t)i * n2 + i] = 1.0;
            Bf.assign((size_t)L * L, 0.0);
            if (zero) {
                for (int m = L; m < n2; ++m) ident[(size_t)m * n2 + m] = 0.0;
                for (int m = 0; m < L; ++m) invB[(size_t)m * n2 + m] = 2.0;
            } else {
                tmpB.assign((size_t)n2 * n2, 0.0);
                for (int m = 0; m < L; ++m) {
                    tmpB[(size_t)m * n2 + m] = 0.5;
                    for (int m1 = 0; m1 < L; ++m1)
                        tmpB[(size_t)(m + L) * n2 + m1 + L] = ms * (ms * (m == m1) - Iac[m + (size_t)m1 * L]) / Pr;
                }
                invB = tmpB;
                if (!lu_invert(n2, invB)) { status = RBC_ERR_NUMERIC; return; }
                // inverse_B_matrix's own f-block inverse (inverse_B_matrix.cu:131-145)
                for (int m = 0; m < L; ++m)
                    for (int m1 = 0; m1 < L; ++m1)
                        Bf[(size_t)m * L + m1] = ms * (ms * (m == m1) - Iac[m + (size_t)m1 * L]) / Pr;
                if (!lu_invert(L, Bf)) { status = RBC_ERR_NUMERIC; return; }
            }
            // L (:109-128)
            Lm.assign((size_t)n2 * n2, 0.0);
            for (int m = 0; m < L; ++m) {
                const double pm = (m + 1) * M_PI, lam = lambda_k(m + 1), l2 = lam * lam;
                for (int m1 = 0; m1 < L; ++m1) {
                    Lm[(size_t)m * n2 + m1] = (-0.5) * (ms + pm * pm) * (m == m1);
                    Lm[(size_t)m * n2 + m1 + L] = ms * Ia1[m1 + (size_t)m * L];
                    Lm[(size_t)(m + L) * n2 + m1] = Ra * ms * Ia1[m + (size_t)m1 * L];
                    Lm[(size_t)(m + L) * n2 + m1 + L] =
                        -ms * ((l2 * l2 + ms * ms) * (m == m1) - 2.0 * ms * Iac[m + (size_t)m1 * L]);
                }
            }
            // C = ident - dt * invB * L (:132-137)
            C.assign((size_t)n2 * n2, 0.0);
            for (int i = 0; i < n2; ++i)
                for (int j = 0; j < n2; ++j) {
                    double s = 0.0;
                    for (int q = 0; q < n2; ++q) s += invB[(size_t)i * n2 + q] * Lm[(size_t)q * n2 + j];
                    C[(size_t)i * n2 + j] = -dt * s + ident[(size_t)i * n2 + j];
                }
            std::vector<double> Ci((size_t)n2 * n2, 0.0);
            if (zero) {  // invert the theta-theta block only (:141-168)
                std::vector<double> up((size_t)L * L);
                for (int m = 0; m < L; ++m)
                    for (int m1 = 0; m1 < L; ++m1) up[(size_t)m * L + m1] = C[(size_t)m * n2 + m1];
                if (!lu_invert(L, up)) { status = RBC_ERR_NUMERIC; return; }
                for (int m = 0; m < L; ++m)
                    for (int m1 = 0; m1 < L; ++m1) Ci[(size_t)m * n2 + m1] = up[(size_t)m * L + m1];
            } else {
                Ci = C;
                if (!lu_invert(n2, Ci)) { status = RBC_ERR_NUMERIC; return; }
            }
            // scatter into the parity-split, column-major group record
            double* rec = &F.mats[(size_t)gi * F.stride];
            size_t off = 0;
            for (int c = 0; c < 2; ++c) {
                const int Lc = (int)cls[c].size(), nc = 2 * Lc;
                for (int col = 0; col < nc; ++col)
                    for (int row = 0; row < nc; ++row) {
                        const int gr = row < Lc ? cls[c][row] : L + cls[c][row - Lc];
                        const int gc = col < Lc ? cls[c][col] : L + cls[c][col - Lc];
                        rec[off + (size_t)col * nc + row] = Ci[(size_t)gr * n2 + gc];
                    }
                off += (size_t)nc * nc;
            }
            for (int c = 0; c < 2; ++c) {
                const int Lc = (int)cls[c].size();
                for (int col = 0; col < Lc; ++col)
                    for (int row = 0; row < Lc; ++row)
                        rec[off + (size_t)col * Lc + row] = Bf[(size_t)cls[c][row] * L + cls[c][col]];
                off += (size_t)Lc * Lc;
            }
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (ng < 64) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        const int g0 = (int)((long long)ng * t / nt), g1 = (int)((long long)ng * (t + 1) / nt);
        th.emplace_back(work, g0, g1);
    }
    for (auto& t : th) t.join();
    if (status) err = "singular per-wavenumber matrix in factor build";
    return status;
}

}  // namespace rbc
