// Exact discrete adjoint of the time step (validation mode; no counterpart
// in the reference, whose adjoint re-runs the forward propagator on
// time-reversed residuals).
//
// The forward step, per updated cell (all rows, or rows >= dn with the free
// surface on):
//   p_k = 2 p_{k-1} - p_{k-2} + c2 (D2x p_{k-1} + D2y p_{k-1} + X + Y + s_k)
//   X = Dbx phi_{k-1} + chi_k,  phi_k = A phi_{k-1} + B D1x p_{k-1},
//   chi_k = A chi_{k-1} + B (D2x p_{k-1} + Dbx phi_{k-1})           (y alike)
// with D2x, D1x clamped (the outer 3 columns take column 3 / NX-4) and Dbx
// the band-local first derivative clamped to band columns 3..dn-4.  A clamped
// operator is S C (stencil on the computed lines, C copying the clamp line),
// so its transpose folds the outer lines onto the clamp line and applies the
// mirrored stencil.  Backwards from k = nt to 2:
//   t1 = g + [band] B Xt,  t3 = [band] B Pt       (level k+1, g = c2 a)
//   a_k  = e_k + 2 a_{k+1} - a_{k+2} + sum_m w2(m) fold(t1)(.-m) + w1(m) fold(t3)(.-m)
//   Pt_k = A Pt_{k+1} + sum_m w1(m) foldband(t1)(.-m),  Xt_k = c2 a_k + A Xt_{k+1}
//   dJ/dc2 += a_k (p_k - 2 p_{k-1} + p_{k-2}) / c2     (FP64)
// One thread per padded cell; operation order identical to the CPU oracle's
// exact_adjoint, so both agree bit for bit.
#include "fwi_exact.h"

namespace fwi {
namespace {

constexpr float E_C1_1 = 3.0f / 4.0f;
constexpr float E_C1_2 = -3.0f / 20.0f;
constexpr float E_C1_3 = 1.0f / 60.0f;
constexpr float E_C2_0 = (-49.0f / 18.0f) / 2.0f;
constexpr float E_C2_1 = 3.0f / 2.0f;
constexpr float E_C2_2 = -3.0f / 20.0f;
constexpr float E_C2_3 = 1.0f / 90.0f;

__device__ __forceinline__ float w2(int m) {
    const int a = m < 0 ? -m : m;
    return a == 0 ? E_C2_0 * 2.0f : (a == 1 ? E_C2_1 : (a == 2 ? E_C2_2 : E_C2_3));
}
__device__ __forceinline__ float w1(int m) {
    const int a = m < 0 ? -m : m;
    const float c = a == 1 ? E_C1_1 : (a == 2 ? E_C1_2 : E_C1_3);
    return m == 0 ? 0.0f : (m > 0 ? c : -c);
}

struct Geo {
    int NX, P, NY, dn, fs;
    __device__ int dxv(int x) const { return x < dn ? dn - x : (x >= NX - dn ? x - (NX - dn) + 1 : 0); }
    __device__ int dyv(int y) const {
        if (y < dn) return fs ? 0 : dn - y;
        return y >= NY - dn ? y - (NY - dn) + 1 : 0;
    }
    __device__ bool upd(int y) const { return !fs || y >= dn; }
};

__device__ __forceinline__ Geo geo(const ExactShot& s, const ExactParams& q) { return Geo{s.NX, s.P, q.NY, q.dn, q.fs}; }

// level k+1 intermediates t1x, t1y, t3x, t3y
__global__ void exact_t_kernel(const ExactParams q) {
    const ExactShot& s = q.shots[blockIdx.y];
    const Geo g = geo(s, q);
    const size_t n = (size_t)g.NY * g.NX;
    const float* a1 = s.a[(q.k + 1) & 1];
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(e / g.NX), x = (int)(e % g.NX);
        const size_t i = (size_t)y * g.P + x;
        const float gk = s.c2[i] * a1[i];
        float t1x = gk, t1y = gk, t3x = 0.0f, t3y = 0.0f;
        if (g.upd(y)) {
            const int dxv = g.dxv(x), dyv = g.dyv(y), m = dxv > dyv ? dxv : dyv;
            const float B = q.fB[m];
            if (dxv > 0) {
                const int band = x < g.dn ? 0 : 1, b = x < g.dn ? x : x - (g.NX - g.dn);
                const size_t r = ((size_t)band * g.NY + y) * g.dn + b;
                t1x = gk + B * s.Xx[q.cur][r];
                t3x = B * s.Px[q.cur][r];
            }
            if (dyv > 0) {
                const int band = y < g.dn ? 0 : 1, rb = y < g.dn ? y : y - (g.NY - g.dn);
                const size_t r = (size_t)band * g.dn * g.P + (size_t)rb * g.P + x;
                t1y = gk + B * s.Xy[q.cur][r];
                t3y = B * s.Py[q.cur][r];
            }
        }
        s.t[0][i] = t1x;
        s.t[1][i] = t1y;
        s.t[2][i] = t3x;
        s.t[3][i] = t3y;
    }
}

// receiver term of step k, summed in receiver order (one thread per shot)
__global__ void exact_rec_kernel(const ExactParams q) {
    const ExactShot& s = q.shots[blockIdx.x];
    if (threadIdx.x != 0) return;
    for (int r = 0; r < s.nrec; ++r) {
        float v = s.e[(size_t)(q.k - 2) * s.nrec + r];
        if (q.k == q.nt) v = v + s.e[(size_t)(q.nt - 1) * s.nrec + r];
        s.rte[s.rec_cell[r]] += v;
    }
}

__device__ __forceinline__ float foldx(const float* t, const Geo& g, int y, int j) {
    if (j < 3 || j > g.NX - 4) return 0.0f;
    const float* row = t + (size_t)y * g.P;
    float v = row[j];
    if (j == 3) v = v + row[0] + row[1] + row[2];
    if (j == g.NX - 4) v = v + row[g.NX - 3] + row[g.NX - 2] + row[g.NX - 1];
    return v;
}
__device__ __forceinline__ float foldy(const float* t, const Geo& g, int j, int x) {
    if (j < 3 || j > g.NY - 4) return 0.0f;
    float v = t[(size_t)j * g.P + x];
    if (j == 3) v = v + t[x] + t[(size_t)g.P + x] + t[(size_t)2 * g.P + x];
    if (j == g.NY - 4)
        v = v + t[(size_t)(g.NY - 3) * g.P + x] + t[(size_t)(g.NY - 2) * g.P + x] + t[(size_t)(g.NY - 1) * g.P + x];
    return v;
}

// a_k, the memory-variable adjoints at level k, the gradient and the source
// adjoint
__global__ void exact_a_kernel(const ExactParams q) {
    const ExactShot& s = q.shots[blockIdx.y];
    const Geo g = geo(s, q);
    const size_t n = (size_t)g.NY * g.NX;
    const float* a1 = s.a[(q.k + 1) & 1];
    float* an = s.a[q.k & 1];                 // holds a_{k+2}, overwritten with a_k
    const int nx_ = q.cur ^ 1;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(e / g.NX), x = (int)(e % g.NX);
        const size_t i = (size_t)y * g.P + x;
        float v = 0.0f;
        if (g.upd(y)) {
            v = s.rte[i] + 2.0f * a1[i] - an[i];
            for (int m = -3; m <= 3; ++m) {
                v = v + w2(m) * foldx(s.t[0], g, y, x - m) + w1(m) * foldx(s.t[2], g, y, x - m);
                v = v + w2(m) * foldy(s.t[1], g, y - m, x) + w1(m) * foldy(s.t[3], g, y - m, x);
            }
            const int dxv = g.dxv(x), dyv = g.dyv(y), mm = dxv > dyv ? dxv : dyv;
            const float A = q.fA[mm];
            if (dxv > 0) {
                const int band = x < g.dn ? 0 : 1, b = x < g.dn ? x : x - (g.NX - g.dn);
                const int off = band == 0 ? 0 : g.NX - g.dn;
                const float* row = s.t[0] + (size_t)y * g.P + off;
                float dtu = 0.0f;
                for (int m = -3; m <= 3; ++m) {
                    const int j = b - m;
                    float u = 0.0f;
                    if (j >= 3 && j <= g.dn - 4) {
                        u = row[j];
                        if (j == 3) u = u + row[0] + row[1] + row[2];
                        if (j == g.dn - 4) u = u + row[g.dn - 3] + row[g.dn - 2] + row[g.dn - 1];
                    }
                    dtu = dtu + w1(m) * u;
                }
                const size_t r = ((size_t)band * g.NY + y) * g.dn + b;
                s.Px[nx_][r] = A * s.Px[q.cur][r] + dtu;
                s.Xx[nx_][r] = s.c2[i] * v + A * s.Xx[q.cur][r];
            }
            if (dyv > 0) {
                const int band = y < g.dn ? 0 : 1, rb = y < g.dn ? y : y - (g.NY - g.dn);
                const int off = band == 0 ? 0 : g.NY - g.dn;
                const float* col = s.t[1] + (size_t)off * g.P + x;
                float dtu = 0.0f;
                for (int m = -3; m <= 3; ++m) {
                    const int j = rb - m;
                    float u = 0.0f;
                    if (j >= 3 && j <= g.dn - 4) {
                        u = col[(size_t)j * g.P];
                        if (j == 3) u = u + col[0] + col[(size_t)g.P] + col[(size_t)2 * g.P];
                        if (j == g.dn - 4)
                            u = u + col[(size_t)(g.dn - 3) * g.P] + col[(size_t)(g.dn - 2) * g.P] +
                                col[(size_t)(g.dn - 1) * g.P];
                    }
                    dtu = dtu + w1(m) * u;
                }
                const size_t r = (size_t)band * g.dn * g.P + (size_t)rb * g.P + x;
                s.Py[nx_][r] = A * s.Py[q.cur][r] + dtu;
                s.Xy[nx_][r] = s.c2[i] * v + A * s.Xy[q.cur][r];
            }
        }
        an[i] = v;
        s.rte[i] = 0.0f;
        if (s.hist) {
            const float* hk = s.hist + (size_t)q.k * s.hist_stride;
            const float* h1 = hk - s.hist_stride;
            const float* h2 = h1 - s.hist_stride;
            s.gc2[i] += (double)v * (((double)hk[i] - 2.0 * (double)h1[i] + (double)h2[i]) / (double)s.c2[i]);
        }
    }
}

__global__ void exact_src_kernel(const ExactParams q) {
    const ExactShot& s = q.shots[blockIdx.y];
    const float* an = s.a[q.k & 1];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < s.nsrc; j += gridDim.x * blockDim.x) {
        const int c = s.src_cell[j];
        if (c >= 0) s.srcadj[(size_t)(q.k - 2) * s.nsrc + j] = s.c2[c] * an[c];
    }
}

// p at time level k of every shot into its history slot k
__global__ void hist_store_kernel(const ShotDev* __restrict__ shots, const ExactShot* __restrict__ xs, int k) {
    const ShotDev& sd = shots[blockIdx.y];
    const ExactShot& s = xs[blockIdx.y];
    const float* p = sd.p[(k & 1) ^ 1];
    float* h = s.hist_w + (size_t)k * s.hist_stride;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < s.hist_stride; e += (size_t)gridDim.x * blockDim.x)
        h[e] = p[e];
}

__global__ void residual_l2_kernel(const ResidShot* __restrict__ shots, float dt, double* __restrict__ partial) {
    const ResidShot& s = shots[blockIdx.y];
    __shared__ double sh[256];
    double acc = 0.0;
    const size_t n = s.n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const float r = s.rec[e] - s.obs[e];
        s.e[e] = dt * r;
        acc += (double)r * (double)r;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

}  // namespace

hipError_t launch_exact_step(const ExactParams& q, int nshots, size_t max_cells, hipStream_t st) {
    const unsigned bx = (unsigned)std::min<size_t>(2048, (max_cells + 255) / 256);
    hipLaunchKernelGGL(exact_t_kernel, dim3(bx, nshots), dim3(256), 0, st, q);
    hipLaunchKernelGGL(exact_rec_kernel, dim3(nshots), dim3(64), 0, st, q);
    hipLaunchKernelGGL(exact_a_kernel, dim3(bx, nshots), dim3(256), 0, st, q);
    hipLaunchKernelGGL(exact_src_kernel, dim3(1, nshots), dim3(256), 0, st, q);
    return hipGetLastError();
}

hipError_t launch_hist_store(const ShotDev* shots, const ExactShot* xs, int nshots, size_t max_stride, int k,
                             hipStream_t st) {
    const unsigned bx = (unsigned)std::min<size_t>(2048, (max_stride + 255) / 256);
    hipLaunchKernelGGL(hist_store_kernel, dim3(bx, nshots), dim3(256), 0, st, shots, xs, k);
    return hipGetLastError();
}

hipError_t launch_residual_l2(const ResidShot* shots, int nshots, float dt, double* partial, int nblk,
                              hipStream_t st) {
    hipLaunchKernelGGL(residual_l2_kernel, dim3(nblk, nshots), dim3(256), 0, st, shots, dt, partial);
    return hipGetLastError();
}

}  // namespace fwi
