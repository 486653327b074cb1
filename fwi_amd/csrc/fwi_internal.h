// Internal types shared by the HIP kernels and the host-side library code.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fwi {

constexpr int ACC = 3;        // stencil radius (solver_acoustic_SS.f90:91, acc=3)
constexpr int MAXDN = 96;     // largest CPML frame dn = dpml + 2*acc supported

// Snapshot storage precisions.
enum SnapKind : int { SNAP_F32 = 0, SNAP_BF16 = 1, SNAP_F16 = 2 };

// Per-shot device descriptor.  Every wavefield-shaped array is row-major with
// x (horizontal) contiguous: element (y, x) of a padded field lives at
// base[y * P + x], 0 <= y < NY, 0 <= x < NX, and the allocation extends 4
// floats before row 0 and past the last row so the vector loads of the edge
// lanes stay inside it.  The model window (rows dn..dn+ny-1, columns
// dn..dn+W-1) is stored separately for snapshots and the imaging kernel as
// [ny][Ws] with model column m at offset m.
struct ShotDev {
    int NX, P, W, Ws;             // padded width, field pitch, model width, model-window pitch
    int ntile;                    // number of x tiles (per launch configuration)
    int ntr_src, nrec, pad0;      // source traces, receivers
    float* p[2];                  // two time levels; p[(k&1)] holds p at time k-1 when stepping k
    const float* c2;              // c^2 dt^2 / d^2 on the padded grid
    float* phx[2];                // x-band CPML phi, ping-pong, [2 bands][NY][dn]
    float* chx;                   // x-band CPML chi, [2][NY][dn]
    float* phy[2];                // y-band CPML phi, ping-pong, [2 bands][dn][P]
    float* chy;                   // y-band CPML chi, [2][dn][P]
    const int* inj_ptr[2];        // CSR over (row, tile) of injection cells: [0] forward, [1] adjoint
    const int2* inj_ent[2];       // {padded column, trace index}
    const float* inj_tr[2];       // injected traces, time-major [t][ntr]
    int ntr[2];
    const int* rec_ptr;           // CSR over (row, tile) of receiver cells
    const int2* rec_ent;          // {padded column, receiver index}
    float* rec_tr;                // recorded traces, time-major [t][nrec]
    void* snap;                   // snapshot slots [slot][ny][Ws]
    const float* snap_scale;      // per-slot scale factors (compressed kinds), or null
    float* grad;                  // per-shot imaging kernel [ny][Ws] (FP32, reference accumulation)
    size_t snap_stride;           // elements per slot
};

// A unit of work: one x tile of one shot over rows [ya, yb).
struct Task {
    int shot, xs, xe, ya, yb, tile, pad0, pad1;
};

// Per-launch parameters of the time-step kernel.
struct StepParams {
    const ShotDev* shots;
    const Task* tasks;
    int ntasks;
    int NY, dn, ny;
    int k;              // Fortran time index of the step being taken (2..nt)
    int mode;           // 0 forward (inject sources, record receivers), 1 adjoint (inject at receivers)
    int snap_op;        // 0 none, 1 store snapshot slot, 2 correlate with slot
    int slot;
    int snap_kind;      // SnapKind
    int record;         // 1: sample receivers after the update
    float dt;
    float ssf;          // store_snap as float
    int free_surface;
    int pad0;
    float fA[MAXDN + 1];   // CPML factors indexed by max(dx, dy), 1-based distance into the frame
    float fB[MAXDN + 1];
};

// Launch configurations of the step kernel (cells per lane).
constexpr int step_halo_lanes(int vx) { return (ACC + vx - 1) / vx; }
constexpr int step_tile_cols(int vx) { return (64 - 2 * step_halo_lanes(vx)) * vx; }

hipError_t launch_step(int vx, const StepParams& prm, hipStream_t stream);

}  // namespace fwi
