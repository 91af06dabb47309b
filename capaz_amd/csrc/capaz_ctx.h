// The single-domain context behind the C ABI (private to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/capaz.h"
#include "capaz_derive.h"
#include "capaz_launch.h"

struct capaz_ctx {
    capaz_params user;
    capaz::DevParams p;
    hipStream_t stream = nullptr;
    double* cur[capaz::F_COUNT] = {};
    double* alt[capaz::F_COUNT] = {};
    bool primary[capaz::F_COUNT] = {};
    bool stale = false;     // intermediates belong to an unrun materialisation
    bool m_checked = false; // m has been scanned for zeros since it last changed
    bool m_has_zero = false;
    int* d_flag = nullptr;
    unsigned long long* d_first = nullptr;  // non-finite sentinel: first bad cell
    double* d_dtm = nullptr;  // dt / m per cell, refreshed with the porosity scan
    double* d_lam = nullptr;  // energy model: lambda_eff per cell within a step
    unsigned char* d_redo = nullptr;  // energy model: cells the Newton pass redoes with pivoting
    capaz::Tuning tuning;
    // Event timing of the fused step's main kernel (capaz_tune "timing"):
    // a pool of event pairs, drained into the totals when it fills or is read.
    bool timing = false;
    std::vector<capaz::KernelEvents> ev;
    size_t ev_used = 0;
    double timed_ms = 0.;
    long timed_launches = 0;
    // adjoint workspace (checkpoints, segment states, adjoint fields)
    double* adj_mem = nullptr;
    size_t adj_bytes = 0;
    long adj_segment = 0;  // segment length C of the last capaz_adjoint
    // second stream of the reverse sweep (segment recompute beside the
    // adjoint passes) and its events: [0..1] segment buffer b recomputed,
    // [2..3] segment buffer b free again
    hipStream_t adj_stream = nullptr;
    hipEvent_t adj_ev[4] = {};
    // caller's device buffer for the z exchange planes each step writes
    // (capaz_set_halo_pack), or NULL
    double* halo_pack = nullptr;
    std::string err;
};

namespace capaz::detail {

// Records msg as the error of c (or of the failed create when c is NULL).
int fail(capaz_ctx* c, int code, const std::string& msg);
int hip_fail(capaz_ctx* c, hipError_t e, const char* what);
// Recomputes the intermediate fields after fused steps (no-op when current).
int materialize(capaz_ctx* c);
// Scans m for zeros and refreshes dt/m once per porosity field.
int scan_m(capaz_ctx* c);
// Makes the context's device current.
int check_ctx(capaz_ctx* c);

}  // namespace capaz::detail
