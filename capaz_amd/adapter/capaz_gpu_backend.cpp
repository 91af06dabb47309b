// Drop-in compute backend for capaz's driver: defines the device half of
// defines.h (the functions gpu.cu defines, gpu.cu:243-1128) with their exact
// signatures and forwards each to the C ABI of include/capaz.h.  A driver
// with main.cpp's structure links this file and libcapaz_amd.so where it
// linked gpu.o:
//
//   g++ -DTHREE_PHASE -I<capaz> -I<capaz_amd>/include main.cpp no_communication.cpp
//       shared_test.cpp gauss.cpp Three-phase/three-phase.cpp
//       capaz_gpu_backend.cpp -L<capaz_amd>/capaz_amd -lcapaz_amd -o capaz_gpu
//
// (add energy.cpp and -DENERGY for the energy model).  defines.h comes from
// the capaz tree; its struct layouts depend on THREE_PHASE / ENERGY, which is
// why this file is compiled together with the driver rather than shipped
// inside the library.
//
// Scope: one domain (no_communication.cpp).  The device arrays in
// ptr_Arrays are the backend's own field arrays; the entry points run the
// reference CPU path's arithmetic (cpu.cpp + model .cpp + energy.cpp) on the
// device, so results match the CPU build, not gpu.cu's divergent
// three-phase closures.
//
// Fused steps.  time_step_function (main.cpp:93-108) calls the step's entry
// points in a fixed order.  The adapter records them instead of running
// them; when a call completes the order (boundary_conditions), it runs the
// whole step as one capaz_step -- the fused kernels, the same bits as the
// eight entry points.  Any other call (data transfers, face exchanges) or a
// call out of order first runs what was recorded, one entry point at a time.
// A fused step ping-pongs the primary state between two device arrays, so
// the adapter maps the pointers it handed to the driver onto field ids and
// moves data with capaz_get_field / capaz_set_field, which read and write the
// current arrays (recomputing the intermediate fields where needed).
// CAPAZ_DROPIN_FUSED=0 in the environment runs every entry point as called.
#include "defines.h"

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

#include "capaz.h"

namespace {

capaz_ctx* g_ctx = nullptr;
consts g_def;
// The device pointers handed to the driver at allocation, by field id.
double* g_orig[CAPAZ_NUM_FIELDS] = {};

// The step's entry points in time_step_function's order (main.cpp:93-108).
enum Hot { H_S, H_PXR, H_U, H_ROS, H_HE, H_E, H_NEWTON, H_BOUNDARY };
#ifdef ENERGY
const Hot kOrder[] = {H_S, H_PXR, H_U, H_ROS, H_HE, H_E, H_NEWTON, H_BOUNDARY};
#else
const Hot kOrder[] = {H_S, H_PXR, H_U, H_ROS, H_NEWTON, H_BOUNDARY};
#endif
constexpr int kSteps = sizeof(kOrder) / sizeof(kOrder[0]);
bool g_fuse = true;
int g_pending = 0;   // entries of kOrder recorded and not yet run
double g_t = 0.;     // t of the recorded roS_calculation
long g_fused = 0;    // steps run as capaz_step
// The driver's host arrays, as last passed to an entry point.  save_data_plots
// downloads neither S_g nor T (main.cpp:478-491) and initialization uploads
// no T (main.cpp:329-339) -- gpu.cu had no energy step and its plots showed
// the initial S_g.  The adapter uploads the host T before the first step and
// refreshes S_g and T on the host whenever the driver pulls P_w for plots, so
// the plots match the CPU program's.
ptr_Arrays g_host = {};
bool g_have_host = false;
#ifdef ENERGY
bool g_T_on_device = false;
#endif

void remember_host(const ptr_Arrays& h) {
    g_host = h;
    g_have_host = true;
}

void check(int rc, const char* what) {
    if (rc != CAPAZ_OK) {
        char msg[512];
        snprintf(msg, sizeof msg, "%s: %s", what, capaz_last_error(g_ctx));
        print_error(msg, (char*)__FILE__, __LINE__);
    }
}

size_t cells(const consts& def) { return (size_t)def.locNx * def.locNy * def.locNz; }

double* dev_field(int id) {
    double* p = nullptr;
    check(capaz_field_device_ptr(g_ctx, id, &p), "capaz_field_device_ptr");
    return p;
}

// The field id of a device pointer the driver holds, -1 if it is not one.
int field_of(const double* dev) {
    for (int f = 0; f < CAPAZ_NUM_FIELDS; ++f)
        if (g_orig[f] && g_orig[f] == dev) return f;
    return -1;
}

void run_one(Hot h, double t) {
    switch (h) {
        case H_S: check(capaz_S(g_ctx), "capaz_S"); break;
        case H_PXR: check(capaz_P_Xi_ro(g_ctx), "capaz_P_Xi_ro"); break;
        case H_U: check(capaz_u(g_ctx), "capaz_u"); break;
        case H_ROS: check(capaz_roS(g_ctx, t), "capaz_roS"); break;
        case H_HE: check(capaz_energy_HE(g_ctx), "capaz_energy_HE"); break;
        case H_E: check(capaz_energy_E(g_ctx), "capaz_energy_E"); break;
        case H_NEWTON: check(capaz_newton(g_ctx), "capaz_newton"); break;
        case H_BOUNDARY: check(capaz_boundary(g_ctx), "capaz_boundary"); break;
    }
}

// Runs the recorded entry points one at a time.
void flush() {
    for (int q = 0; q < g_pending; ++q) run_one(kOrder[q], g_t);
    g_pending = 0;
}

// An entry point of the step: recorded while the calls follow
// time_step_function's order, the whole step run as one fused capaz_step when
// the order completes.
void hot(Hot h, double t = 0.) {
    if (g_fuse && kOrder[g_pending] == h) {
        if (h == H_ROS) g_t = t;
        if (++g_pending == kSteps) {
            g_pending = 0;
            // capaz_step's t = step * dt is unused: assign_roS_nr ignores t
            check(capaz_step(g_ctx, ++g_fused, 1), "capaz_step");
        }
        return;
    }
    flush();
    run_one(h, t);
}

capaz_params params_of(const consts& def) {
    capaz_params p;
    capaz_params_init(&p);
#ifdef THREE_PHASE
    p.model = CAPAZ_THREE_PHASE;
    p.mu_g = def.mu_g;
    p.ro0_g = def.ro0_g;
    p.S_nr = def.S_nr[0];
    p.S_gr = def.S_gr[0];
    p.S_w_gr = def.S_w_gr;
#else
    p.model = CAPAZ_TWO_PHASE;
    p.P_d = def.P_d[0];
#endif
#ifdef ENERGY
    p.energy = 1;
#endif
    p.closure = CAPAZ_CLOSURE_REFERENCE;
    p.Nx = def.locNx;
    p.Ny = def.locNy;
    p.Nz = def.locNz;
    p.source = def.source;
    p.newton_iterations = def.newton_iterations;
    p.device = 0;
    p.hx = def.hx;
    p.hy = def.hy;
    p.hz = def.hz;
    p.dt = def.dt;
    p.P_atm = def.P_atm;
    p.g_const = def.g_const;
    p.mu_w = def.mu_w;
    p.mu_n = def.mu_n;
    p.ro0_w = def.ro0_w;
    p.ro0_n = def.ro0_n;
    p.beta_w = def.beta_w;
    p.beta_n = def.beta_n;
    p.lambda = def.lambda[0];
    p.S_wr = def.S_wr[0];
    p.S_n_gr = def.S_n_gr;
    return p;
}

// Face copies go through a device buffer; the driver's HostBuffer is the
// staging area, as in gpu.cu:926-1128.
void face(double* DevArrayPtr, double* HostBuffer, double* DevBuffer, int axis, int side, bool save,
          const consts& def) {
    size_t plane = axis == 0 ? (size_t)def.locNy * def.locNz
                             : (axis == 1 ? (size_t)def.locNx * def.locNz : (size_t)def.locNx * def.locNy);
    flush();
    const int field = field_of(DevArrayPtr);
    if (field < 0) print_error((char*)"face exchange on an array the backend does not own", (char*)__FILE__, __LINE__);
    if (save) {
        check(hipMemcpy(DevBuffer, HostBuffer, plane * sizeof(double), hipMemcpyHostToDevice) == hipSuccess
                  ? CAPAZ_OK : CAPAZ_ERR_DEVICE, "hipMemcpy");
        check(capaz_save_face(g_ctx, field, axis, side, DevBuffer), "capaz_save_face");
        check(capaz_sync(g_ctx), "capaz_sync");
    } else {
        check(capaz_load_face(g_ctx, field, axis, side, DevBuffer), "capaz_load_face");
        check(capaz_sync(g_ctx), "capaz_sync");
        check(hipMemcpy(HostBuffer, DevBuffer, plane * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess
                  ? CAPAZ_OK : CAPAZ_ERR_DEVICE, "hipMemcpy");
    }
}

}  // namespace

// set_boundary_basic_coordinate (cpu.cpp:107-141): the model .cpp files'
// host-side boundary helpers call it; gpu.cu has only a device copy.
int set_boundary_basic_coordinate(int i, int j, int k, consts def) {
    int i1 = i, j1 = j, k1 = k;
    if (i == 0 && def.locNx > 2) i1++;
    if (i == def.locNx - 1 && def.locNx > 2) i1--;
    if (j == 0) j1++;
    if (j == def.locNy - 1) j1--;
    if (k == 0 && def.locNz > 2) k1++;
    if (k == def.locNz - 1 && def.locNz > 2) k1--;
    return i1 + j1 * def.locNx + k1 * def.locNx * def.locNy;
}

void device_initialization(consts* def) {
    if (def->size > 1)
        print_error((char*)"capaz_amd backend: one domain per process only (no_communication.cpp)", (char*)__FILE__,
                    __LINE__);
    def->blocksX = def->blocksY = def->blocksZ = 0;  // the backend chooses its own launch shapes
    g_def = *def;
    const char* fuse = std::getenv("CAPAZ_DROPIN_FUSED");
    g_fuse = !(fuse && std::strcmp(fuse, "0") == 0);
}

void device_finalization(void) {}

void device_memory_allocation(ptr_Arrays* ArraysPtr, double** DevBuffer, consts def) {
    capaz_params p = params_of(def);
    check(capaz_create(&p, &g_ctx), "capaz_create");
    double** slots[] = {
        &ArraysPtr->S_n, &ArraysPtr->S_w, &ArraysPtr->P_w, &ArraysPtr->P_n, &ArraysPtr->ro_w, &ArraysPtr->ro_n,
        &ArraysPtr->ux_w, &ArraysPtr->uy_w, &ArraysPtr->uz_w, &ArraysPtr->ux_n, &ArraysPtr->uy_n, &ArraysPtr->uz_n,
        &ArraysPtr->Xi_w, &ArraysPtr->Xi_n, &ArraysPtr->roS_w, &ArraysPtr->roS_w_old, &ArraysPtr->roS_n,
        &ArraysPtr->roS_n_old, &ArraysPtr->m, &ArraysPtr->K,
#ifdef THREE_PHASE
        &ArraysPtr->S_g, &ArraysPtr->P_g, &ArraysPtr->ro_g, &ArraysPtr->ux_g, &ArraysPtr->uy_g, &ArraysPtr->uz_g,
        &ArraysPtr->Xi_g, &ArraysPtr->roS_g, &ArraysPtr->roS_g_old,
#endif
#ifdef ENERGY
        &ArraysPtr->T, &ArraysPtr->H_w, &ArraysPtr->H_n, &ArraysPtr->H_g, &ArraysPtr->H_r, &ArraysPtr->E,
        &ArraysPtr->E_new,
#endif
    };
    const int ids[] = {
        CAPAZ_S_n, CAPAZ_S_w, CAPAZ_P_w, CAPAZ_P_n, CAPAZ_ro_w, CAPAZ_ro_n, CAPAZ_ux_w, CAPAZ_uy_w, CAPAZ_uz_w,
        CAPAZ_ux_n, CAPAZ_uy_n, CAPAZ_uz_n, CAPAZ_Xi_w, CAPAZ_Xi_n, CAPAZ_roS_w, CAPAZ_roS_w_old, CAPAZ_roS_n,
        CAPAZ_roS_n_old, CAPAZ_m, CAPAZ_K,
#ifdef THREE_PHASE
        CAPAZ_S_g, CAPAZ_P_g, CAPAZ_ro_g, CAPAZ_ux_g, CAPAZ_uy_g, CAPAZ_uz_g, CAPAZ_Xi_g, CAPAZ_roS_g, CAPAZ_roS_g_old,
#endif
#ifdef ENERGY
        CAPAZ_T, CAPAZ_H_w, CAPAZ_H_n, CAPAZ_H_g, CAPAZ_H_r, CAPAZ_E, CAPAZ_E_new,
#endif
    };
    for (size_t f = 0; f < sizeof(ids) / sizeof(ids[0]); ++f) *slots[f] = g_orig[ids[f]] = dev_field(ids[f]);
    *DevBuffer = nullptr;  // one domain: no exchange buffer (gpu.cu:676-685 sizes it from sizex/y/z)
}

void device_memory_free(ptr_Arrays DevArraysPtr, double* DevBuffer) {
    (void)DevArraysPtr;
    (void)DevBuffer;
    flush();
    capaz_destroy(g_ctx);
    g_ctx = nullptr;
}

void load_data_to_host(double* HostArrayPtr, double* DevArrayPtr, consts def) {
    flush();
    const int id = field_of(DevArrayPtr);
    if (id < 0) print_error((char*)"load_data_to_host: not a backend array", (char*)__FILE__, __LINE__);
    check(capaz_get_field(g_ctx, id, HostArrayPtr, cells(def)), "capaz_get_field");
    if (g_have_host && id == CAPAZ_P_w) {
#ifdef THREE_PHASE
        check(capaz_get_field(g_ctx, CAPAZ_S_g, g_host.S_g, cells(def)), "capaz_get_field(S_g)");
#endif
#ifdef ENERGY
        check(capaz_get_field(g_ctx, CAPAZ_T, g_host.T, cells(def)), "capaz_get_field(T)");
#endif
    }
}

void load_data_to_device(double* HostArrayPtr, double* DevArrayPtr, consts def) {
    flush();
    const int id = field_of(DevArrayPtr);
    if (id < 0) print_error((char*)"load_data_to_device: not a backend array", (char*)__FILE__, __LINE__);
    check(capaz_set_field(g_ctx, id, HostArrayPtr, cells(def)), "capaz_set_field");
}

void load_data_to_device_int(int* HostArrayPtr, int* DevArrayPtr, consts def) {
    flush();
    check(capaz_sync(g_ctx), "capaz_sync");
    if (hipMemcpy(DevArrayPtr, HostArrayPtr, cells(def) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        print_error((char*)"copy int data to device", (char*)__FILE__, __LINE__);
}

void S_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) {
    remember_host(HostArraysPtr);
#ifdef ENERGY
    if (!g_T_on_device) {
        flush();
        check(capaz_set_field(g_ctx, CAPAZ_T, HostArraysPtr.T, cells(def)), "capaz_set_field(T)");
        g_T_on_device = true;
    }
#endif
    hot(H_S);
}
void ro_P_Xi_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) { hot(H_PXR); }
void u_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) { hot(H_U); }
void roS_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, double t, consts def) { hot(H_ROS, t); }
void P_S_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) { hot(H_NEWTON); }
void boundary_conditions(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) { hot(H_BOUNDARY); }
#ifdef ENERGY
void H_E_current_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) { hot(H_HE); }
void E_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) { hot(H_E); }
#endif

#define CAPAZ_FACE(name, axis, side, save)                                                             \
    void name(double* HostArrayPtr, double* DevArrayPtr, double* HostBuffer, double* DevBuffer, consts def) { \
        face(DevArrayPtr, HostBuffer, DevBuffer, axis, side, save, def);                                    \
    }
CAPAZ_FACE(load_exchange_data_part_xl, 0, 0, false)
CAPAZ_FACE(load_exchange_data_part_xr, 0, 1, false)
CAPAZ_FACE(load_exchange_data_part_yl, 1, 0, false)
CAPAZ_FACE(load_exchange_data_part_yr, 1, 1, false)
CAPAZ_FACE(load_exchange_data_part_zl, 2, 0, false)
CAPAZ_FACE(load_exchange_data_part_zr, 2, 1, false)
CAPAZ_FACE(save_exchange_data_part_xl, 0, 0, true)
CAPAZ_FACE(save_exchange_data_part_xr, 0, 1, true)
CAPAZ_FACE(save_exchange_data_part_yl, 1, 0, true)
CAPAZ_FACE(save_exchange_data_part_yr, 1, 1, true)
CAPAZ_FACE(save_exchange_data_part_zl, 2, 0, true)
CAPAZ_FACE(save_exchange_data_part_zr, 2, 1, true)
