// Times the drop-in path the way the reference's driver runs it: main.cpp's
// own read_defines, initialization and time_step_function (main.cpp:93-108,
// 296-340, 1280-1639; main.cpp is linked with its main renamed) over the
// capaz_amd adapter, with HIP device synchronisations around the timed
// steps.  Run in a directory holding defines.ini:
//
//   capaz_dropin_bench_<m> WARMUP STEPS
//
// prints one JSON line {"warmup": W, "steps": K, "seconds": S, "fused": F},
// F = 0 when CAPAZ_DROPIN_FUSED=0 made the adapter run every entry point as
// called.  Built by `make -C oracle dropin` next to the drop-in drivers.
// (the standard headers first: defines.h defines min and max macros)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "defines.h"

#include <hip/hip_runtime.h>

extern double* HostBuffer;
extern double* DevBuffer;

int main(int argc, char* argv[]) {
    const int warm = argc > 1 ? std::atoi(argv[1]) : 5;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 20;
    char* args[] = {argv[0], nullptr};
    consts def;
    read_defines(1, args, &def);
    ptr_Arrays host, dev;
    long int counter = 0;
    initialization(&host, &dev, &counter, 1, args, &def);
    long s = counter;
    for (int q = 0; q < warm; ++q, ++s) time_step_function(host, dev, DevBuffer, def, (s + 1) * def.dt);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    const auto t0 = std::chrono::steady_clock::now();
    for (int q = 0; q < steps; ++q, ++s) time_step_function(host, dev, DevBuffer, def, (s + 1) * def.dt);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    const auto t1 = std::chrono::steady_clock::now();
    const char* fuse = std::getenv("CAPAZ_DROPIN_FUSED");
    std::printf("{\"warmup\": %d, \"steps\": %d, \"seconds\": %.6f, \"fused\": %d}\n", warm, steps,
                std::chrono::duration<double>(t1 - t0).count(), (fuse && std::strcmp(fuse, "0") == 0) ? 0 : 1);
    std::fflush(stdout);
    finalization(host, dev, DevBuffer);
    return 0;
}
