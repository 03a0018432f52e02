// Dumps the Galerkin integral tables computed by the reference's own Coeff<double>
// class (gpu_flow/code/util/coeff.cu, compiled unmodified from where it lies) into a
// flat binary file, so that the oracle's tables can be checked entry by entry.
//
// This file is test infrastructure: it is built only by oracle/Makefile into
// oracle/_ref/, and only tests read its output.
//
// Output layout (little endian):
//   int32 L
//   double lambda[L]                     lambda(m+1), m = 0..L-1   (coeff.cu:1232-1257)
//   for each 2-factor table in TWO[]:   double t[L*L]      t[i + j*L] = I(i,j)
//   for each 3-factor table in THREE[]: double t[L*L*L]    t[i + j*L + k*L*L] = I(i,j,k)
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "coeff.h"

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s L out.bin\n", argv[0]);
        return 2;
    }
    const int L = std::atoi(argv[1]);
    if (L < 1 || L > 64) {
        std::fprintf(stderr, "L out of range\n");
        return 2;
    }
    Coeff<double> c(L);
    std::FILE* fp = std::fopen(argv[2], "wb");
    if (!fp) {
        std::perror("fopen");
        return 1;
    }
    std::fwrite(&L, sizeof(int), 1, fp);
    std::vector<double> buf;
    buf.resize(L);
    for (int m = 0; m < L; ++m) buf[m] = c.lambda(m + 1);
    std::fwrite(buf.data(), sizeof(double), L, fp);

    // 2-factor tables: Iac, Iab, Ia1, Ia2 (coeff.h:75-78)
    buf.assign((size_t)L * L, 0.0);
    for (int t = 0; t < 4; ++t) {
        for (int j = 0; j < L; ++j)
            for (int i = 0; i < L; ++i) {
                double v = 0.0;
                switch (t) {
                    case 0: v = c.Iac(i, j); break;
                    case 1: v = c.Iab(i, j); break;
                    case 2: v = c.Ia1(i, j); break;
                    default: v = c.Ia2(i, j); break;
                }
                buf[i + (size_t)j * L] = v;
            }
        std::fwrite(buf.data(), sizeof(double), buf.size(), fp);
    }

    // 3-factor tables (coeff.h:80-92), in this order:
    // Iac1 Ibb1 Iaab Iaad Iabc Ia11 Ia12 I111 Ibbb Ib11 Iaa1 Iab2
    buf.assign((size_t)L * L * L, 0.0);
    for (int t = 0; t < 12; ++t) {
        for (int k = 0; k < L; ++k)
            for (int j = 0; j < L; ++j)
                for (int i = 0; i < L; ++i) {
                    double v = 0.0;
                    switch (t) {
                        case 0: v = c.Iac1(i, j, k); break;
                        case 1: v = c.Ibb1(i, j, k); break;
                        case 2: v = c.Iaab(i, j, k); break;
                        case 3: v = c.Iaad(i, j, k); break;
                        case 4: v = c.Iabc(i, j, k); break;
                        case 5: v = c.Ia11(i, j, k); break;
                        case 6: v = c.Ia12(i, j, k); break;
                        case 7: v = c.I111(i, j, k); break;
                        case 8: v = c.Ibbb(i, j, k); break;
                        case 9: v = c.Ib11(i, j, k); break;
                        case 10: v = c.Iaa1(i, j, k); break;
                        default: v = c.Iab2(i, j, k); break;
                    }
                    buf[i + (size_t)j * L + (size_t)k * L * L] = v;
                }
        std::fwrite(buf.data(), sizeof(double), buf.size(), fp);
    }
    std::fclose(fp);
    return 0;
}
