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
#include "defines.h"

#include <hip/hip_runtime.h>

#include "capaz.h"

namespace {

capaz_ctx* g_ctx = nullptr;
consts g_def;
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
    int field = -1;
    for (int f = 0; f < CAPAZ_NUM_FIELDS && field < 0; ++f)
        if (capaz_has_field(g_ctx, f) && dev_field(f) == DevArrayPtr) field = f;
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
    for (size_t f = 0; f < sizeof(ids) / sizeof(ids[0]); ++f) *slots[f] = dev_field(ids[f]);
    *DevBuffer = nullptr;  // one domain: no exchange buffer (gpu.cu:676-685 sizes it from sizex/y/z)
}

void device_memory_free(ptr_Arrays DevArraysPtr, double* DevBuffer) {
    (void)DevArraysPtr;
    (void)DevBuffer;
    capaz_destroy(g_ctx);
    g_ctx = nullptr;
}

void load_data_to_host(double* HostArrayPtr, double* DevArrayPtr, consts def) {
    check(capaz_sync(g_ctx), "capaz_sync");
    if (hipMemcpy(HostArrayPtr, DevArrayPtr, cells(def) * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        print_error((char*)"copy data to host", (char*)__FILE__, __LINE__);
    if (g_have_host && DevArrayPtr == dev_field(CAPAZ_P_w)) {
#ifdef THREE_PHASE
        check(capaz_get_field(g_ctx, CAPAZ_S_g, g_host.S_g, cells(def)), "capaz_get_field(S_g)");
#endif
#ifdef ENERGY
        check(capaz_get_field(g_ctx, CAPAZ_T, g_host.T, cells(def)), "capaz_get_field(T)");
#endif
    }
}

void load_data_to_device(double* HostArrayPtr, double* DevArrayPtr, consts def) {
    check(capaz_sync(g_ctx), "capaz_sync");
    if (hipMemcpy(DevArrayPtr, HostArrayPtr, cells(def) * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        print_error((char*)"copy double data to device", (char*)__FILE__, __LINE__);
}

void load_data_to_device_int(int* HostArrayPtr, int* DevArrayPtr, consts def) {
    check(capaz_sync(g_ctx), "capaz_sync");
    if (hipMemcpy(DevArrayPtr, HostArrayPtr, cells(def) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        print_error((char*)"copy int data to device", (char*)__FILE__, __LINE__);
}

void S_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) {
    remember_host(HostArraysPtr);
#ifdef ENERGY
    if (!g_T_on_device) {
        check(capaz_set_field(g_ctx, CAPAZ_T, HostArraysPtr.T, cells(def)), "capaz_set_field(T)");
        g_T_on_device = true;
    }
#endif
    check(capaz_S(g_ctx), "capaz_S");
}
void ro_P_Xi_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) {
    check(capaz_P_Xi_ro(g_ctx), "capaz_P_Xi_ro");
}
void u_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) { check(capaz_u(g_ctx), "capaz_u"); }
void roS_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, double t, consts def) {
    check(capaz_roS(g_ctx, t), "capaz_roS");
}
void P_S_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) {
    check(capaz_newton(g_ctx), "capaz_newton");
}
void boundary_conditions(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) {
    check(capaz_boundary(g_ctx), "capaz_boundary");
}
#ifdef ENERGY
void H_E_current_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) {
    check(capaz_energy_HE(g_ctx), "capaz_energy_HE");
}
void E_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def) {
    check(capaz_energy_E(g_ctx), "capaz_energy_E");
}
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
