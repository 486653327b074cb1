// Fused time-step kernel for the 2-D acoustic CPML propagator (gfx950).
//
// One launch advances every shot of a batch by one time step.  The work unit
// (Task) is one x tile of one shot over a range of rows.  A 64-lane wavefront
// owns the tile: lane l holds VX consecutive columns, the first and last H
// lanes only load halo columns.  The wavefront marches down its rows keeping
// the 7 rows p(yc-3..yc+3) of the current time level in registers (y
// stencil), stages the current row in LDS (x stencil), and in the same pass
//   * updates the CPML memory variables of the cells in the absorbing frame,
//   * injects source (or adjoint source) samples,
//   * writes p at the new time level over p at the previous-previous level,
//   * samples receivers,
//   * stores the forward snapshot or applies the imaging condition.
//
// Arithmetic follows the reference statement by statement so results match it
// bit for bit (the file is compiled with -ffp-contract=off):
//   derivatives    solver_space_derivs.f90:1-131 (pair sums, coefficient order,
//                  outer `acc` lines clamped to the nearest computed line)
//   CPML           solver_PML_matrices.f90:1-113 with the caller's updates at
//                  solver_acoustic_SS.f90:347-411; the factor of a cell is
//                  f(max(dx, dy)) where dx, dy are its 1-based distances into the
//                  frame (0 outside), which reproduces the strip profiles and the
//                  corner matrices of solver_PML_coeffs.f90:36-70
//   interior       solver_acoustic_SS.f90:433-434
//   snapshot       solver_acoustic_SS.f90:445-449
//   imaging        solver_acoustic_SS.f90:463-468
//   receivers      solver_acoustic_SS.f90:551-555
// The per-region local derivative of phi (clamped at the region edge) only
// runs along the region's own direction, so the x-strip and the two corners
// sharing its columns form one contiguous band per side for phi_x, and the
// same holds for phi_y along rows.
#include "fwi_internal.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

namespace fwi {
namespace {

// solver_FD_coeffs.f90:14-19 (first derivative) and :43-50,68 (second
// derivative, centre weight halved), evaluated in float.
constexpr float C1_1 = 3.0f / 4.0f;
constexpr float C1_2 = -3.0f / 20.0f;
constexpr float C1_3 = 1.0f / 60.0f;
constexpr float C2_0 = (-49.0f / 18.0f) / 2.0f;
constexpr float C2_1 = 3.0f / 2.0f;
constexpr float C2_2 = -3.0f / 20.0f;
constexpr float C2_3 = 1.0f / 90.0f;

// Second difference in grid units; the reference starts its sum from 0, which
// only affects the sign of an exactly-zero result.
__device__ __forceinline__ float dd2(float m3, float m2, float m1, float c0, float p1, float p2, float p3) {
    float s = C2_0 * (c0 + c0);
    s = s + C2_1 * (p1 + m1);
    s = s + C2_2 * (p2 + m2);
    s = s + C2_3 * (p3 + m3);
    return s;
}

// First difference in grid units (the i=1 term of the reference is (p-p)*0).
__device__ __forceinline__ float dd1(float m3, float m2, float m1, float p1, float p2, float p3) {
    float s = (p1 - m1) * C1_1;
    s = s + (p2 - m2) * C1_2;
    s = s + (p3 - m3) * C1_3;
    return s;
}

template <int VX> __device__ __forceinline__ void vload(float (&d)[VX], const float* src) {
    if constexpr (VX == 4) {
        const float4 t = *reinterpret_cast<const float4*>(src);
        d[0] = t.x; d[1] = t.y; d[2] = t.z; d[3] = t.w;
    } else if constexpr (VX == 2) {
        const float2 t = *reinterpret_cast<const float2*>(src);
        d[0] = t.x; d[1] = t.y;
    } else {
        d[0] = *src;
    }
}

template <int VX> __device__ __forceinline__ void vzero(float (&d)[VX]) {
#pragma unroll
    for (int i = 0; i < VX; ++i) d[i] = 0.0f;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ void snap_store(void* base, size_t idx, float v, int kind, float inv_scale) {
    if (kind == SNAP_F32) {
        static_cast<float*>(base)[idx] = v;
    } else if (kind == SNAP_BF16) {
        static_cast<__hip_bfloat16*>(base)[idx] = __float2bfloat16(v * inv_scale);
    } else {
        static_cast<__half*>(base)[idx] = __float2half(v * inv_scale);
    }
}

__device__ __forceinline__ float snap_load(const void* base, size_t idx, int kind, float scale) {
    if (kind == SNAP_F32) return static_cast<const float*>(base)[idx];
    if (kind == SNAP_BF16) return __bfloat162float(static_cast<const __hip_bfloat16*>(base)[idx]) * scale;
    return __half2float(static_cast<const __half*>(base)[idx]) * scale;
}

template <int VX>
__global__ __launch_bounds__(64) void step_kernel(const StepParams prm) {
    constexpr int H = step_halo_lanes(VX);
    constexpr int TW = 64 * VX;
    __shared__ float lds_p[2][TW];
    __shared__ float lds_src[TW];
    __shared__ float lds_out[TW];

    // Hardware hands block b to XCD b % 8; give each XCD a contiguous range of
    // tasks so tiles that share halos share an L2.  gridDim.x is a multiple of 8.
    const int per = gridDim.x >> 3;
    const int task = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (task >= prm.ntasks) return;
    const Task t = prm.tasks[task];
    const ShotDev& s = prm.shots[t.shot];

    const int lane = threadIdx.x;
    const int NY = prm.NY, dn = prm.dn, NX = s.NX, P = s.P;
    const int x0 = t.xs - VX * H;           // first loaded column of the tile
    const int xl = x0 + VX * lane;          // first column held by this lane
    const bool ld_ok = (xl + VX) <= P;      // xl >= -4 always holds (front pad)
    const bool compute_lane = lane >= H && lane < 64 - H;
    const int cur = prm.k & 1;
    const float* __restrict__ pc = s.p[cur];     // time k-1
    float* __restrict__ pn = s.p[cur ^ 1];       // time k-2, overwritten with time k
    const float* __restrict__ c2f = s.c2;
    const float* __restrict__ phx_o = s.phx[cur];
    float* __restrict__ phx_n = s.phx[cur ^ 1];
    const float* __restrict__ phy_o = s.phy[cur];
    float* __restrict__ phy_n = s.phy[cur ^ 1];
    float* __restrict__ chx = s.chx;
    float* __restrict__ chy = s.chy;

    const int inj_dir = prm.mode;
    const int* __restrict__ iptr = s.inj_ptr[inj_dir];
    const int2* __restrict__ ient = s.inj_ent[inj_dir];
    const float* __restrict__ itr_t = s.inj_tr[inj_dir] + (size_t)(prm.k - 2) * s.ntr[inj_dir];
    float* __restrict__ rec_t = prm.record ? s.rec_tr + (size_t)(prm.k - 2) * s.nrec : nullptr;

    // column classification
    int cx[VX], dxv[VX];
    bool in[VX];
#pragma unroll
    for (int i = 0; i < VX; ++i) {
        cx[i] = xl + i;
        in[i] = compute_lane && cx[i] >= t.xs && cx[i] < t.xe;
        dxv[i] = cx[i] < dn ? dn - cx[i] : (cx[i] >= NX - dn ? cx[i] - (NX - dn) + 1 : 0);
    }

    float snap_scale = 1.0f, snap_inv = 1.0f;
    void* snap_slot = nullptr;
    if (prm.snap_op != 0) {
        snap_slot = static_cast<char*>(s.snap) +
                    (size_t)prm.slot * s.snap_stride * (prm.snap_kind == SNAP_F32 ? 4 : 2);
        if (prm.snap_kind != SNAP_F32 && s.snap_scale) {
            snap_scale = s.snap_scale[prm.slot];
            snap_inv = snap_scale > 0.0f ? 1.0f / snap_scale : 0.0f;
        }
    }

#pragma unroll
    for (int i = 0; i < VX; ++i) lds_src[VX * lane + i] = 0.0f;

    // register window: rows yc-3 .. yc+3 of p at time k-1
    float w[7][VX];
    int yc = clampi(t.ya, ACC, NY - 1 - ACC);
#pragma unroll
    for (int r = 0; r < 7; ++r) {
        if (ld_ok) vload<VX>(w[r], pc + (size_t)(yc - 3 + r) * P + xl);
        else vzero<VX>(w[r]);
    }
    float pmr[VX], c2r[VX];
    if (ld_ok) {
        vload<VX>(pmr, pn + (size_t)t.ya * P + xl);
        vload<VX>(c2r, c2f + (size_t)t.ya * P + xl);
    } else {
        vzero<VX>(pmr);
        vzero<VX>(c2r);
    }

    int buf = 0;
    for (int y = t.ya; y < t.yb; ++y) {
        // prefetch the operands of the next row
        const int y1 = y + 1;
        const bool more = y1 < t.yb;
        const int yc1 = clampi(y1, ACC, NY - 1 - ACC);
        const bool shift = more && (yc1 != yc);
        float npm[VX], nc2[VX], npw[VX];
        vzero<VX>(npm);
        vzero<VX>(nc2);
        vzero<VX>(npw);
        if (more && ld_ok) {
            vload<VX>(npm, pn + (size_t)y1 * P + xl);
            vload<VX>(nc2, c2f + (size_t)y1 * P + xl);
            if (shift) vload<VX>(npw, pc + (size_t)(yc1 + 3) * P + xl);
        }

        // p on row y (the window is centred on the clamped row yc)
        const int wr = y - yc + 3;
        float prow[VX];
#pragma unroll
        for (int i = 0; i < VX; ++i) {
            float v = w[3][i];
            if (wr == 0) v = w[0][i];
            else if (wr == 1) v = w[1][i];
            else if (wr == 2) v = w[2][i];
            else if (wr == 4) v = w[4][i];
            else if (wr == 5) v = w[5][i];
            else if (wr == 6) v = w[6][i];
            prow[i] = v;
        }
#pragma unroll
        for (int i = 0; i < VX; ++i) lds_p[buf][VX * lane + i] = prow[i];

        // injection cells of this row and tile (model rows only)
        int ib = 0, ie = 0;
        if (y >= dn && y < dn + prm.ny) {
            const int ci = y * s.ntile + t.tile;
            ib = iptr[ci];
            ie = iptr[ci + 1];
        }
        const bool has_inj = ie > ib;
        if (has_inj)
            for (int e = ib + lane; e < ie; e += 64) {
                const int2 en = ient[e];
                lds_src[en.x - x0] = itr_t[en.y];
            }
        __syncthreads();

        const bool ytop = y < dn;
        const bool ybot = y >= NY - dn;
        const int dyv = ytop ? dn - y : (ybot ? y - (NY - dn) + 1 : 0);
        const int yband = ytop ? 0 : 1;
        const int rb = ytop ? y : y - (NY - dn);
        const int rbc = clampi(rb, ACC, dn - 1 - ACC);
        const bool model_row = y >= dn && y < dn + prm.ny;

        float pp[VX];
#pragma unroll
        for (int i = 0; i < VX; ++i) {
            pp[i] = 0.0f;
            if (!in[i]) continue;
            const int c = cx[i];
            const int li = clampi(c, ACC, NX - 1 - ACC) - x0;
            const float* L = &lds_p[buf][li];
            const float d2x = dd2(L[-3], L[-2], L[-1], L[0], L[1], L[2], L[3]);
            const float d2y = dd2(w[0][i], w[1][i], w[2][i], w[3][i], w[4][i], w[5][i], w[6][i]);
            const float base = (-pmr[i]) + 2.0f * prow[i];
            const float lap = d2x + d2y;
            const int dx = dxv[i];
            if (dx == 0 && dyv == 0) {
                float q = lap;
                if (has_inj) q = q + lds_src[c - x0];
                pp[i] = base + q * c2r[i];
            } else {
                const int m = dx > dyv ? dx : dyv;
                const float A = prm.fA[m], B = prm.fB[m];
                float propx = 0.0f, propy = 0.0f;
                if (dx > 0) {
                    const float d1x = dd1(L[-3], L[-2], L[-1], L[1], L[2], L[3]);
                    const int band = c < dn ? 0 : 1;
                    const int b = band == 0 ? c : c - (NX - dn);
                    const int bc = clampi(b, ACC, dn - 1 - ACC);
                    const size_t row = (size_t)(band * NY + y) * dn;
                    const float* ph = phx_o + row;
                    const float dph = dd1(ph[bc - 3], ph[bc - 2], ph[bc - 1], ph[bc + 1], ph[bc + 2], ph[bc + 3]);
                    const float phn = A * ph[b] + B * d1x;
                    const float chn = A * chx[row + b] + B * (d2x + dph);
                    phx_n[row + b] = phn;
                    chx[row + b] = chn;
                    propx = dph + chn;
                }
                if (dyv > 0) {
                    const float d1y = dd1(w[0][i], w[1][i], w[2][i], w[4][i], w[5][i], w[6][i]);
                    const float* ph = phy_o + (size_t)(yband * dn) * P + c;
                    const float dph = dd1(ph[(size_t)(rbc - 3) * P], ph[(size_t)(rbc - 2) * P], ph[(size_t)(rbc - 1) * P],
                                          ph[(size_t)(rbc + 1) * P], ph[(size_t)(rbc + 2) * P], ph[(size_t)(rbc + 3) * P]);
                    const size_t idx = (size_t)(yband * dn + rb) * P + c;
                    const float phn = A * phy_o[idx] + B * d1y;
                    const float chn = A * chy[idx] + B * (d2y + dph);
                    phy_n[idx] = phn;
                    chy[idx] = chn;
                    propy = dph + chn;
                }
                if (dx > 0 && dyv > 0) pp[i] = (base + c2r[i] * (lap + propx)) + c2r[i] * propy;
                else if (dx > 0) pp[i] = base + c2r[i] * (lap + propx);
                else pp[i] = base + c2r[i] * (lap + propy);
            }
            pn[(size_t)y * P + c] = pp[i];

            if (prm.snap_op != 0 && model_row && c >= dn && c < dn + s.W) {
                const size_t mi = (size_t)(y - dn) * s.Ws + (c - dn);
                if (prm.snap_op == 1) {
                    const float f = ((pp[i] - pmr[i]) * 0.5f) / prm.dt;
                    snap_store(snap_slot, mi, f, prm.snap_kind, snap_inv);
                } else {
                    const float f = snap_load(snap_slot, mi, prm.snap_kind, snap_scale);
                    const float k2 = -(((0.5f * (pp[i] - pmr[i])) * f) * prm.ssf);
                    s.grad[mi] = s.grad[mi] + k2;
                }
            }
        }

        // receivers on this row and tile
        if (rec_t) {
            const int ci = y * s.ntile + t.tile;
            const int rbeg = s.rec_ptr[ci], rend = s.rec_ptr[ci + 1];
            if (rend > rbeg) {
#pragma unroll
                for (int i = 0; i < VX; ++i) lds_out[VX * lane + i] = pp[i];
                __syncthreads();
                for (int e = rbeg + lane; e < rend; e += 64) {
                    const int2 en = s.rec_ent[e];
                    rec_t[en.y] = lds_out[en.x - x0];
                }
            }
        }
        if (has_inj) {
            __syncthreads();
            for (int e = ib + lane; e < ie; e += 64) lds_src[ient[e].x - x0] = 0.0f;
            __syncthreads();
        }

        // advance
#pragma unroll
        for (int i = 0; i < VX; ++i) {
            pmr[i] = npm[i];
            c2r[i] = nc2[i];
        }
        if (shift) {
#pragma unroll
            for (int r = 0; r < 6; ++r)
#pragma unroll
                for (int i = 0; i < VX; ++i) w[r][i] = w[r + 1][i];
#pragma unroll
            for (int i = 0; i < VX; ++i) w[6][i] = npw[i];
            yc = yc1;
        }
        buf ^= 1;
    }
}

}  // namespace

hipError_t launch_step(int vx, const StepParams& prm, hipStream_t stream) {
    const int nb = ((prm.ntasks + 7) / 8) * 8;
    if (nb == 0) return hipSuccess;
    switch (vx) {
        case 1: hipLaunchKernelGGL(step_kernel<1>, dim3(nb), dim3(64), 0, stream, prm); break;
        case 2: hipLaunchKernelGGL(step_kernel<2>, dim3(nb), dim3(64), 0, stream, prm); break;
        case 4: hipLaunchKernelGGL(step_kernel<4>, dim3(nb), dim3(64), 0, stream, prm); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fwi
