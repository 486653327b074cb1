// Exact discrete adjoint (validation mode): device descriptors and launchers.
#pragma once
#include "fwi_internal.h"

#include <algorithm>

namespace fwi {

struct ExactShot {
    int NX, P;
    const float* c2;              // padded c2 [NY][P]
    float* a[2];                  // adjoint of p; a[k & 1] holds a_{k+2} on entry to step k, a_k on exit
    float* t[4];                  // t1x, t1y, t3x, t3y of level k+1 [NY][P]
    float* Xx[2];                 // adjoint of chi_x, [2 bands][NY][dn], ping-pong
    float* Px[2];                 // adjoint of phi_x
    float* Xy[2];                 // adjoint of chi_y, [2 bands][dn][P]
    float* Py[2];                 // adjoint of phi_y
    float* rte;                   // receiver term of the step [NY][P] (kept zero between steps)
    double* gc2;                  // dJ/dc2 on the padded grid [NY][P]
    const float* hist;            // p at levels 0..nt, [k][NY][P], or null (no gradient)
    float* hist_w;                // the same storage, written by the forward
    size_t hist_stride;           // NY * P
    const int* rec_cell;          // receiver -> y * P + x
    int nrec, nsrc;
    const float* e;               // dJ/d(recorded sample), time-major [t][nrec]
    const int* src_cell;          // source -> y * P + x, -1 when not injected (frame, or overridden)
    float* srcadj;                // dJ/d(source sample), time-major [t][nsrc]
};

struct ExactParams {
    const ExactShot* shots;
    int NY, dn, fs;
    int k, nt, cur;               // step, steps, index of the level-(k+1) memory buffers
    float fA[MAXDN + 1];
    float fB[MAXDN + 1];
};

struct ResidShot {                // e = dt (rec - obs), time-major, and the FP64 sum of (rec - obs)^2
    const float* rec;
    const float* obs;
    float* e;
    size_t n;
};

// One backward step k of every shot (intermediates, receiver term, adjoint
// update, source adjoint).
hipError_t launch_exact_step(const ExactParams& q, int nshots, size_t max_cells, hipStream_t st);
// History slot k <- p at time level k (after step k).
hipError_t launch_hist_store(const ShotDev* shots, const ExactShot* xs, int nshots, size_t max_stride, int k,
                             hipStream_t st);
// Residuals and per-block FP64 partial sums partial[shot * nblk + b].
hipError_t launch_residual_l2(const ResidShot* shots, int nshots, float dt, double* partial, int nblk,
                              hipStream_t st);

}  // namespace fwi
