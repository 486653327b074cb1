// Auxiliary kernels of the gradient path: padded c2 construction, trace
// misfit / adjoint source, FP64 shot-summed gradient, snapshot quantisation,
// fills and copies for checkpoints.
#include "fwi_aux.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

namespace fwi {
namespace {

// c2 = c^2 dt^2 / d^2 on the padded grid of every shot, with the velocity
// extended by nearest-edge clamping into the CPML frame
// (solver_acoustic_SS.f90:271-292).  vp is the full model, column-major (ny, nx).
__global__ void build_c2_kernel(const float* __restrict__ vp, int ny, int nxm, const C2Shot* __restrict__ shots,
                                int dn, int NY, float dt, float d) {
    const C2Shot sh = shots[blockIdx.y];
    const size_t n = (size_t)NY * sh.P;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(e / sh.P), x = (int)(e % sh.P);
        float v = 0.0f;
        if (x < sh.NX) {
            int jm = y - dn; jm = jm < 0 ? 0 : (jm > ny - 1 ? ny - 1 : jm);
            int im = x - dn; im = im < 0 ? 0 : (im > sh.W - 1 ? sh.W - 1 : im);
            const float c = vp[(size_t)(sh.x0 + im) * ny + jm];
            v = ((c * c) * (dt * dt)) / (d * d);
        }
        sh.c2[e] = v;
    }
}

// Shared interior c2: out + q*copy_stride + 4 holds c2(row, j - q) at [row*Pm + j].
__global__ void build_c2_shared_kernel(const float* __restrict__ vp, int ny, int nxm, float dt, float d,
                                       float* __restrict__ out, int Pm, size_t copy_stride) {
    const int q = blockIdx.y;
    float* base = out + q * copy_stride + 4;
    const size_t n = (size_t)ny * Pm;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / Pm), j = (int)(e % Pm) - q;
        float v = 0.0f;
        if (j >= 0 && j < nxm) {
            const float c = vp[(size_t)j * ny + r];
            v = ((c * c) * (dt * dt)) / (d * d);
        }
        base[e] = v;
    }
}

// Per receiver: max |synth|, max |obs| and whether obs has a non-zero sample
// (Misfit.f90:31-51, is_not_cero in auxiliar.f90:282-299).  Traces are
// time-major [t][nrec_s] per shot; g indexes receivers over all shots.
__global__ void trace_norm_kernel(const float* __restrict__ synth, const float* __restrict__ obs,
                                  const RecMap* __restrict__ map, int nrec_total, int nt, float* __restrict__ invs) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nrec_total) return;
    const RecMap m = map[g];
    const float* s = synth + m.toff + m.r;
    const float* o = obs + m.toff + m.r;
    float maxS = 0.0f, maxT = 0.0f;
    bool nz = false;
    for (int t = 0; t < nt; ++t) {
        const float a = s[(size_t)t * m.nrec], b = o[(size_t)t * m.nrec];
        maxS = fmaxf(maxS, fabsf(a));
        maxT = fmaxf(maxT, fabsf(b));
        nz = nz || (b != 0.0f);
    }
    const bool ok = nz && maxS != 0.0f && maxT != 0.0f;
    invs[2 * g] = ok ? maxS : 0.0f;      // 0 marks a trace that stays zero
    invs[2 * g + 1] = ok ? maxT : 0.0f;
}

// Normalised residual, adjoint source and per-block misfit partials.
// Misfit_L2.f90:8-14: term = (0.5*dt)*(s-o)^2 in float, summed here in double;
// AdjSource(k) = -(s(nt-k+1) - o(nt-k+1)).
__global__ void residual_kernel(const float* __restrict__ synth, const float* __restrict__ obs,
                                const ResShot* __restrict__ shots, const float* __restrict__ invs, int nt, float dt,
                                float* __restrict__ adj, double* __restrict__ partial, int nblk_x) {
    const ResShot sh = shots[blockIdx.y];
    const size_t n = (size_t)nt * sh.nrec;
    double acc = 0.0;
    const float hdt = 0.5f * dt;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(e / sh.nrec), r = (int)(e % sh.nrec);
        const float mS = invs[2 * (sh.roff + r)], mT = invs[2 * (sh.roff + r) + 1];
        float df = 0.0f;
        if (mS != 0.0f) {
            const float a = synth[sh.toff + e] / mS;
            const float b = obs[sh.toff + e] / mT;
            df = a - b;
        }
        acc += (double)(hdt * (df * df));
        adj[sh.toff + (size_t)(nt - 1 - t) * sh.nrec + r] = -df;
    }
    // fixed-order block reduction
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)blockIdx.y * nblk_x + blockIdx.x] = red[0];
}

__global__ void misfit_sum_kernel(const double* __restrict__ partial, int nshots, int nblk_x, double* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    double m = 0.0;
    for (int b = 0; b < nblk_x; ++b) m += partial[(size_t)s * nblk_x + b];
    out[s] = m;
}

// grad(y, x) = scale * sum_s kernel_s(y, x - x0_s), shots in order, then the
// water rows zeroed.  Output column-major (ny, nx) to match the host layout.
__global__ void gradient_reduce_kernel(const GradShot* __restrict__ shots, int nshots, int ny, int nxm, double scale,
                                       const int* __restrict__ water_rows, double* __restrict__ grad) {
    const size_t n = (size_t)ny * nxm;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(e / ny), y = (int)(e % ny);
        double acc = 0.0;
        for (int s = 0; s < nshots; ++s) {
            const GradShot g = shots[s];
            const int xm = x - g.x0;
            if (xm >= 0 && xm < g.W) acc += (double)g.k[(size_t)y * g.Ws + xm + g.moff];
        }
        double v = acc * scale;
        if (water_rows && y < water_rows[x]) v = 0.0;
        grad[e] = v;
    }
}

// Power-of-two scale that maps the slot's largest |value| into [2^14, 2^15):
// exact to apply and to undo, so the 16-bit rounding is the only error, and
// the largest value stays below the half-precision limit 65504.
__device__ __forceinline__ float snap_scale_of(unsigned int bits) {
    const float mx = __uint_as_float(bits);
    if (!(mx > 0.0f) || !(mx <= 3.0e38f)) return 1.0f;
    int e;
    (void)frexpf(mx, &e);            // mx = f * 2^e, f in [0.5, 1)
    return ldexpf(1.0f, e - 15);
}

__global__ void quantize_kernel(const QuantShot* __restrict__ shots, int slot, int bf16) {
    const QuantShot q = shots[blockIdx.y];
    const float sc = snap_scale_of(q.maxbits[slot]);
    const float inv = 1.0f / sc;
    unsigned short* __restrict__ out = q.slots + (size_t)slot * q.stride;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < q.stride; e += (size_t)gridDim.x * blockDim.x) {
        const float v = q.stage[e] * inv;
        out[e] = bf16 ? __bfloat16_as_ushort(__float2bfloat16(v)) : __half_as_ushort(__float2half(v));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) q.scale[slot] = sc;
}

__global__ void fill_u32_kernel(unsigned int* __restrict__ p, size_t n, unsigned int v) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) p[e] = v;
}

__global__ void copy16_kernel(float4* __restrict__ dst, const float4* __restrict__ src, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        dst[e] = src[e];
}

}  // namespace

hipError_t launch_quantize(const QuantShot* shots, int nshots, size_t max_stride, int slot, int bf16, hipStream_t st) {
    if (nshots <= 0) return hipSuccess;
    const unsigned bx = (unsigned)std::min<size_t>(1024, std::max<size_t>(1, (max_stride + 1023) / 1024));
    hipLaunchKernelGGL(quantize_kernel, dim3(bx, nshots), dim3(256), 0, st, shots, slot, bf16);
    return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned int* p, size_t n, unsigned int v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const unsigned b = (unsigned)std::min<size_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(b), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_copy16(void* dst, const void* src, size_t n16, hipStream_t st) {
    if (n16 == 0) return hipSuccess;
    const unsigned b = (unsigned)std::min<size_t>(8192, (n16 + 255) / 256);
    hipLaunchKernelGGL(copy16_kernel, dim3(b), dim3(256), 0, st, static_cast<float4*>(dst),
                       static_cast<const float4*>(src), n16);
    return hipGetLastError();
}

hipError_t launch_build_c2(const float* vp, int ny, int nxm, const C2Shot* shots, int nshots, int dn, int NY,
                           float dt, float d, hipStream_t st) {
    if (nshots == 0) return hipSuccess;
    hipLaunchKernelGGL(build_c2_kernel, dim3(256, nshots), dim3(256), 0, st, vp, ny, nxm, shots, dn, NY, dt, d);
    return hipGetLastError();
}

hipError_t launch_build_c2_shared(const float* vp, int ny, int nxm, float dt, float d, float* out, int Pm,
                                  size_t copy_stride, hipStream_t st) {
    hipLaunchKernelGGL(build_c2_shared_kernel, dim3(256, 4), dim3(256), 0, st, vp, ny, nxm, dt, d, out, Pm,
                       copy_stride);
    return hipGetLastError();
}

hipError_t launch_misfit(const float* synth, const float* obs, const RecMap* map, int nrec_total, const ResShot* shots,
                         int nshots, int nt, float dt, float* invs, float* adj, double* partial, int nblk_x,
                         double* shot_misfit, hipStream_t st) {
    if (nrec_total > 0)
        hipLaunchKernelGGL(trace_norm_kernel, dim3((nrec_total + 255) / 256), dim3(256), 0, st, synth, obs, map,
                           nrec_total, nt, invs);
    if (nshots > 0) {
        hipLaunchKernelGGL(residual_kernel, dim3(nblk_x, nshots), dim3(256), 0, st, synth, obs, shots, invs, nt, dt,
                           adj, partial, nblk_x);
        hipLaunchKernelGGL(misfit_sum_kernel, dim3((nshots + 63) / 64), dim3(64), 0, st, partial, nshots, nblk_x,
                           shot_misfit);
    }
    return hipGetLastError();
}

hipError_t launch_gradient_reduce(const GradShot* shots, int nshots, int ny, int nxm, double scale,
                                  const int* water_rows, double* grad, hipStream_t st) {
    const size_t n = (size_t)ny * nxm;
    const int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(gradient_reduce_kernel, dim3(nb), dim3(256), 0, st, shots, nshots, ny, nxm, scale, water_rows,
                       grad);
    return hipGetLastError();
}

}  // namespace fwi
