// rbc_io.cpp -- state files in the reference's binary format and the AB2 history.
//
// File format (output/matrix_folder_writer.cu:3-103): int32 number of matrices (1 per
// field file), then per matrix int32 ndims, int32 dims[ndims], and prod(dims) complex
// values of the writing build's CUDA_FLOAT (float2 or double2).  There is no type tag;
// the element size follows from the file size.  theta/f/g have dims {M/2+1, N, L}, F/G
// {1, 1, L} (main.cu:275-276, 359-381, util/util.cu:32-43).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rbc_ctx.hpp"

using namespace rbc;

namespace {

struct FileArray {
    std::vector<int> dims;
    std::vector<double> data;  // interleaved complex, converted to double
};

// Reads a one-matrix file; the element size (8 or 16 bytes per complex) is inferred from the
// file size.  Returns "" on success, else the reason.
std::string read_one(const char* path, FileArray& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return std::string("cannot open ") + path;
    std::fseek(f, 0, SEEK_END);
    const long fsize = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    auto rd = [&](void* p, size_t n) { return std::fread(p, 1, n, f) == n; };
    int32_t nmat = 0, nd = 0;
    std::string why;
    if (!rd(&nmat, 4) || nmat != 1) why = "expected exactly one matrix in the file";
    else if (!rd(&nd, 4) || nd < 1 || nd > 8) why = "bad dimension count";
    if (why.empty()) {
        out.dims.resize(nd);
        size_t n = 1;
        for (int i = 0; i < nd && why.empty(); ++i) {
            int32_t d = 0;
            if (!rd(&d, 4) || d < 1) why = "bad dimension";
            out.dims[i] = d;
            n *= (size_t)d;
        }
        const long head = 4 + 4 + 4L * nd;
        const long payload = fsize - head;
        if (why.empty()) {
            size_t esz = 0;
            if (payload == (long)(n * 8)) esz = 8;
            else if (payload == (long)(n * 16)) esz = 16;
            else why = "file size matches neither complex64 nor complex128 data";
            if (esz) {
                std::vector<unsigned char> raw(n * esz);
                if (!rd(raw.data(), raw.size())) why = "short read";
                out.data.resize(2 * n);
                if (esz == 8) {
                    const float* s = (const float*)raw.data();
                    for (size_t i = 0; i < 2 * n; ++i) out.data[i] = s[i];
                } else {
                    std::memcpy(out.data.data(), raw.data(), raw.size());
                }
            }
        }
    }
    std::fclose(f);
    return why.empty() ? "" : why + " (" + path + ")";
}

std::string write_one(const char* path, const std::vector<int>& dims, const double* data, int esz) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return std::string("cannot create ") + path;
    size_t n = 1;
    for (int d : dims) n *= (size_t)d;
    bool ok = true;
    const int32_t nmat = 1, nd = (int32_t)dims.size();
    ok = ok && std::fwrite(&nmat, 4, 1, f) == 1;
    ok = ok && std::fwrite(&nd, 4, 1, f) == 1;
    for (int d : dims) {
        const int32_t d32 = d;
        ok = ok && std::fwrite(&d32, 4, 1, f) == 1;
    }
    if (esz == 8) {
        std::vector<float> tmp(2 * n);
        for (size_t i = 0; i < 2 * n; ++i) tmp[i] = (float)data[i];
        ok = ok && std::fwrite(tmp.data(), 4, tmp.size(), f) == tmp.size();
    } else {
        ok = ok && std::fwrite(data, 8, 2 * n, f) == 2 * n;
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? "" : std::string("write failed: ") + path;
}

template <typename T>
void make_real_now(rbc_ctx* c) {
    Launch<T>::make_real(c->dg, c->b, c->b.X, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
}

// which = 0: state X; 1: AB2 history (previous B^-1 N)
int exchange(rbc_ctx* c, int which, bool put, double* const f5[5]) {
    void* const* fields = which == 0 ? c->b.X : c->b.Rp;
    void* F = which == 0 ? c->b.XF : c->b.RpF;
    void* G = which == 0 ? c->b.XG : c->b.RpG;
    for (int i = 0; i < 3; ++i)
        if (f5[i]) {
            if (put) put_field(c, f5[i], fields[i]);
            else get_field(c, fields[i], f5[i]);
        }
    if (f5[3]) {
        if (put) put_vec(c, f5[3], F);
        else get_vec(c, F, f5[3]);
    }
    if (f5[4]) {
        if (put) put_vec(c, f5[4], G);
        else get_vec(c, G, f5[4]);
    }
    return RBC_OK;
}

}  // namespace

extern "C" {

int rbc_read_state_files(rbc_ctx* c, const char* const paths[5], int which) {
    if (!c || !paths || (which != 0 && which != 1)) return RBC_ERR_ARG;
    const Geometry& g = c->geo;
    std::vector<FileArray> arr(5);
    for (int i = 0; i < 5; ++i) {
        if (!paths[i]) return fail(c, RBC_ERR_ARG, "all five paths are required (theta, f, g, F, G)");
        const std::string why = read_one(paths[i], arr[i]);
        if (!why.empty()) return fail(c, RBC_ERR_IO, why);
        const std::vector<int> want = i < 3 ? std::vector<int>{g.mox, g.N, g.L} : std::vector<int>{1, 1, g.L};
        if (arr[i].dims != want)
            return fail(c, RBC_ERR_ARG, std::string("matrix in ") + paths[i] + " does not have the right dimension");
    }
    try {
        hipSetDevice(c->p.device);
        double* f5[5];
        for (int i = 0; i < 5; ++i) f5[i] = arr[i].data.data();
        exchange(c, which, true, f5);
        if (which == 0) {  // main.cu:396: the initial state is made real after reading
            if (c->rsz == 8) make_real_now<double>(c);
            else make_real_now<float>(c);
            clear_health(c);
        } else {
            c->have_hist = true;
        }
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

int rbc_write_state_files(rbc_ctx* c, const char* const paths[5], int which, int elem_bytes) {
    if (!c || !paths || (which != 0 && which != 1)) return RBC_ERR_ARG;
    if (elem_bytes == 0) elem_bytes = c->rsz == 8 ? 16 : 8;
    if (elem_bytes != 8 && elem_bytes != 16) return fail(c, RBC_ERR_ARG, "elem_bytes must be 0, 8 or 16");
    if (which == 1 && !c->have_hist) return fail(c, RBC_ERR_STATE, "no AB2 history yet (no step since create/reset)");
    for (int i = 0; i < 5; ++i)
        if (!paths[i]) return fail(c, RBC_ERR_ARG, "all five paths are required (theta, f, g, F, G)");
    const Geometry& g = c->geo;
    const size_t nfield = (size_t)g.mox * g.N * g.L;
    std::vector<std::vector<double>> host(5);
    for (int i = 0; i < 5; ++i) host[i].resize(i < 3 ? 2 * nfield : 2 * (size_t)g.L);
    try {
        hipSetDevice(c->p.device);
        double* f5[5];
        for (int i = 0; i < 5; ++i) f5[i] = host[i].data();
        exchange(c, which, false, f5);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    for (int i = 0; i < 5; ++i) {
        const std::vector<int> dims = i < 3 ? std::vector<int>{g.mox, g.N, g.L} : std::vector<int>{1, 1, g.L};
        const std::string why = write_one(paths[i], dims, host[i].data(), elem_bytes);
        if (!why.empty()) return fail(c, RBC_ERR_IO, why);
    }
    return RBC_OK;
}

int rbc_get_history(rbc_ctx* c, double* theta, double* f, double* g, double* F, double* G) {
    if (!c) return RBC_ERR_ARG;
    if (!c->have_hist) return fail(c, RBC_ERR_STATE, "no AB2 history yet (no step since create/reset)");
    try {
        hipSetDevice(c->p.device);
        double* f5[5] = {theta, f, g, F, G};
        return exchange(c, 1, false, f5);
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
}

int rbc_set_history(rbc_ctx* c, const double* theta, const double* f, const double* g, const double* F,
                    const double* G) {
    if (!c || !theta || !f || !g || !F || !G) return RBC_ERR_ARG;
    try {
        hipSetDevice(c->p.device);
        double* f5[5] = {(double*)theta, (double*)f, (double*)g, (double*)F, (double*)G};
        exchange(c, 1, true, f5);
        c->have_hist = true;
    } catch (const HipError& e) {
        return hip_fail(c, e);
    }
    return RBC_OK;
}

}  // extern "C"
