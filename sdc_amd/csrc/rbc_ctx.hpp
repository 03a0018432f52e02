// rbc_ctx.hpp -- the context object behind the C-ABI and the host helpers shared by the
// translation units of libsdc_rbc.so (stepper, state files, diagnostics, tracers,
// ensembles, learned closure).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "rbc_abi.h"
#include "rbc_internal.hpp"

namespace rbc {

struct HipError {
    hipError_t e;
    const char* what;
};

#define HIPCHK(x)                                        \
    do {                                                 \
        hipError_t _e = (x);                             \
        if (_e != hipSuccess) throw rbc::HipError{_e, #x}; \
    } while (0)

}  // namespace rbc

struct rbc_ctx {
    rbc_params p{};
    rbc::Geometry geo;
    rbc::DevGeom dg{};
    rbc::Factors fac;
    int prec = RBC_F64;
    size_t rsz = 8;  // bytes per real on the device
    hipStream_t stream = nullptr;
    bool own_stream = true;
    rbc::Buffers b{};
    void* ia1 = nullptr;  // L*L doubles (Nusselt)
    std::vector<void*> allocs;
    bool have_hist = false;
    int bad_step = 0;  // host copy of the first non-finite step (health[0]) once read; 0 = none
    std::string err;
    static constexpr int kRing = 64;                 // steps of stage events in flight
    hipEvent_t ev_ring[kRing][RBC_NSTAGES + 1]{};
    hipEvent_t* ev = ev_ring[0];                     // events of the step being recorded
    bool events = false;
    // called inside every time step right after the nonlinear term (b.Nn, b.NF, b.NG) is
    // formed; the learned closure adds its correction there.  Null = plain method.
    void (*post_nonlinear)(rbc_ctx*, void*) = nullptr;
    void* post_user = nullptr;
    // velocity and spectrum workspace of rbc_velocity.hip: constants and the spectrum scratch
    // are made at create; the grid and point scratch of the host-buffer diagnostics grow on
    // demand (a call allocates only when it needs more than any earlier call)
    void* vel_ws = nullptr;
    void (*vel_ws_free)(void*) = nullptr;
    // penalization (rbc_set_penalization); its own device arrays, freed when it is set again
    rbc::Penalty pen;
    std::vector<void*> pen_allocs;
    // plane operators: frozen planes and masks (rbc_freeze, rbc_temp_mask, rbc_mask_theta)
    void* frozen = nullptr;     // 3 L M N reals
    int* freeze_mask = nullptr; // M N
    int* temp_mask = nullptr;   // M N
    int* theta_mask = nullptr;  // M N
    bool frozen_set = false, temp_set = false;

    size_t csz() const { return 2 * rsz; }
    void* dalloc(size_t bytes) {
        void* ptr = nullptr;
        HIPCHK(hipMalloc(&ptr, bytes > 0 ? bytes : 16));
        allocs.push_back(ptr);
        return ptr;
    }
    ~rbc_ctx() {
        if (stream) hipStreamSynchronize(stream);
        if (vel_ws && vel_ws_free) vel_ws_free(vel_ws);
        for (void* a : pen_allocs) hipFree(a);
        for (void* a : allocs) hipFree(a);
        if (events)
            for (auto& row : ev_ring)
                for (auto& e : row) hipEventDestroy(e);
        if (stream && own_stream) hipStreamDestroy(stream);
    }
};

namespace rbc {

// error text into the context (or the thread's create error when c is null); returns code
int fail(rbc_ctx* c, int code, const std::string& msg);
int hip_fail(rbc_ctx* c, const HipError& e);

// host <-> device exchange of one field in the reference layout (complex doubles,
// (M/2+1) x N x L, x fastest) and of an L-vector
void put_field(rbc_ctx* c, const double* host, void* dev);
void get_field(rbc_ctx* c, const void* dev, double* host);
void put_vec(rbc_ctx* c, const double* host, void* dev);
void get_vec(rbc_ctx* c, const void* dev, double* host);

// n IMEX steps on the context's stream (no host synchronisation)
void run_steps(rbc_ctx* c, int n);

// waits for the context's stream, then reads the end-of-step health words: RBC_OK, or
// RBC_ERR_NUMERIC (with the step in the error text) once a step has ended non-finite
int health_status(rbc_ctx* c);
// forgets a recorded non-finite step (the state was replaced); stream-ordered, no wait
void clear_health(rbc_ctx* c);

// velocity and spectrum workspace of a new context (rbc_velocity.hip), called by rbc_create
void velocity_setup(rbc_ctx* c);

// the nonlinear term of a state held in the context's device layout, on its stream
void nonlinear_op(rbc_ctx* c, void* const X[3], const void* XF, const void* XG, void* const Nout[3], void* NF,
                  void* NG);

}  // namespace rbc
