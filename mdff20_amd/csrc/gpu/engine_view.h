// engine_view.h -- what another unit of the library (the RDF histogram)
// reads of an engine context: device pointers, the cell and the device
// trajectory.  Internal to libmdff_gpu.so; not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

struct mdff_ctx;

struct EngineView {
    int n, ntype, device;
    hipStream_t st;
    const double *r;        // positions, SoA (x block, y block, z block)
    const int *type;        // type index per ion
    double A[3][3], B[3][3];  // cell rows and the kardir matrix (cell.c:28-52)
    const double *frames;   // trajectory frames: r, v, f (9 n doubles each)
    int nframes;
    size_t frame_doubles;
};

extern "C" __attribute__((visibility("hidden"))) int mdff_engine_view(mdff_ctx *c, EngineView *v);
