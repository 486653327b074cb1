// Trace conditioning on the GPU: the Butterworth filter of time_filter.f90 /
// filters.f90 and the Akima resampling of interpolation_akima.f90 /
// akima433.f90, batched over traces.
//
// Filter: each trace is placed at the end of a 24x zero-padded record
// (time_filter.f90:17-19), transformed with a batched real-to-complex FFT
// (hipFFT), multiplied by the filter response on bins 0..N/2 and transformed
// back; the last nt samples, divided by N (FFTPACK's forward transform carries
// the 1/N), replace the trace.  The reference transforms the real record as a
// complex sequence and keeps the real part; with a real record and a response
// that is conjugate-symmetric except at the Nyquist bin, R2C/C2R gives the
// same real part.  The response follows filters.f90:81-184: order-6 analog
// Butterworth, H = 1/B(if) with the bin frequency accumulated in REAL(4)
// (f = f + fe/ts), the DC bin forced to 1 (low-pass) or 0.
//
// Akima: ACM Algorithm 433 as modified in akima433.f90 (INTRPL), one thread
// per output sample, the Fortran's statement order (its per-interval
// coefficient reuse and EQUIVALENCE'd temporaries do not change values).
#include "fwi_trace.h"

#include <hipfft/hipfft.h>

#include <cmath>
#include <complex>
#include <vector>

namespace fwi {

void butterworth_response(int ts, float dt, int typef, float f1, float fc, std::vector<float2>& H) {
    const int n = 6;                           // filters.f90:22
    const float pi = 3.14159265f;
    const float fe = 1.0f / dt;
    const float fs = 40.0f;                    // filters.f90:30 (high-pass start frequency)
    const float f2 = fc;
    const float f0 = std::sqrt(f1 * f2);
    const float Q = f0 / (f2 - f1);
    const int half = ts / 2 + 1;
    H.assign(half, float2{0.0f, 0.0f});
    const float step = fe / (float)ts;
    float f = 0.0f;
    for (int j = 0; j < half; ++j) {
        const double w = f;
        std::complex<double> z;
        if (typef == 1) z = std::complex<double>(0.0, w / (double)fc);
        else if (typef == 2) z = (double)fs / std::complex<double>(0.0, w);
        else z = (double)Q * (std::complex<double>(0.0, w / (double)f0) + (double)f0 / std::complex<double>(0.0, w));
        std::complex<double> B(1.0, 0.0);
        for (int k = 1; k <= n / 2; ++k) {
            const double c = std::cos((double)(((float)(2 * k + n - 1) * pi) / (float)(2 * n)));
            B = B * (z * z - 2.0 * z * c + 1.0);
        }
        const std::complex<double> h = 1.0 / B;
        H[j] = float2{(float)h.real(), (float)h.imag()};
        f = f + step;
    }
    H[0] = typef == 1 ? float2{1.0f, 0.0f} : float2{0.0f, 0.0f};
}

namespace {

__global__ void pad_kernel(const float* __restrict__ in, int nt, int N, int ntr, float* __restrict__ rec) {
    const size_t total = (size_t)ntr * N;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(e / N), i = (int)(e % N);
        rec[e] = i >= N - nt ? in[(size_t)t * nt + (i - (N - nt))] : 0.0f;
    }
}

__global__ void apply_response_kernel(float2* __restrict__ spec, const float2* __restrict__ H, int half, int ntr) {
    const size_t total = (size_t)ntr * half;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const float2 a = spec[e], h = H[e % half];
        spec[e] = make_float2(a.x * h.x - a.y * h.y, a.x * h.y + a.y * h.x);
    }
}

__global__ void extract_kernel(const float* __restrict__ rec, int nt, int N, int ntr, float inv_n,
                               float* __restrict__ out) {
    const size_t total = (size_t)ntr * nt;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(e / nt), i = (int)(e % nt);
        out[e] = rec[(size_t)t * N + (N - nt) + i] * inv_n;
    }
}

__global__ void akima_kernel(const float* __restrict__ in, float dt_in, int nt_in, float* __restrict__ out,
                             float dt_out, int nt_out, int ntr) {
    const size_t total = (size_t)ntr * nt_out;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(e / nt_out), k = (int)(e % nt_out);
        out[e] = akima_eval(nullptr, dt_in, in + (size_t)t * nt_in, nt_in, (float)k * dt_out);
    }
}

__global__ void time_major_kernel(const float* __restrict__ in, int nt, int ntr, float* __restrict__ out) {
    const size_t total = (size_t)ntr * nt;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(e / ntr), r = (int)(e % ntr);
        out[e] = in[(size_t)r * nt + t];
    }
}

// ---- Wiener source estimation (source_estimation.f90) ----
__global__ void window_maxabs_kernel(float* __restrict__ G, int nt, int nr, const int* __restrict__ it2,
                                     unsigned int* __restrict__ maxbits) {
    const size_t total = (size_t)nr * nt;
    float m = 0.0f;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / nt), i = (int)(e % nt);
        float v = G[e];
        if (i >= it2[r]) { v = 0.0f; G[e] = 0.0f; }
        m = fmaxf(m, fabsf(v));
    }
    atomicMax(maxbits, __float_as_uint(m));
}

// complex records [2 nr][npad]: Gs traces, then Go traces, each divided by its max |value|, zero padded
__global__ void pack_kernel(const float* __restrict__ Gs, const float* __restrict__ Go, int nt, int nr, int npad,
                            const unsigned int* __restrict__ maxbits, float2* __restrict__ rec) {
    const size_t total = (size_t)2 * nr * npad;
    const float ms = __uint_as_float(maxbits[0]), mo = __uint_as_float(maxbits[1]);
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(e / npad), i = (int)(e % npad);
        float v = 0.0f;
        if (i < nt) v = t < nr ? Gs[(size_t)t * nt + i] / ms : Go[(size_t)(t - nr) * nt + i] / mo;
        rec[e] = make_float2(v, 0.0f);
    }
}

// kk = sum_r |Ws|^2, bb = sum_r conj(Ws) Wo over bins [b0, b1), spectra scaled by 1/npad, FP64
__global__ void band_sums_kernel(const float2* __restrict__ W, int nr, int npad, int b0, int b1,
                                 double* __restrict__ kk, double2* __restrict__ bb) {
    const int b = b0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= b1) return;
    const double inv = 1.0 / (double)npad;
    double k = 0.0, br = 0.0, bi = 0.0;
    for (int r = 0; r < nr; ++r) {
        const float2 s = W[(size_t)r * npad + b], o = W[(size_t)(nr + r) * npad + b];
        const double sr = s.x * inv, si = s.y * inv, orr = o.x * inv, oi = o.y * inv;
        k += sr * sr + si * si;
        br += sr * orr + si * oi;
        bi += sr * oi - si * orr;
    }
    kk[b - b0] = k;
    bb[b - b0] = make_double2(br, bi);
}

unsigned blocks_for(size_t n) { return (unsigned)std::min<size_t>(16384, std::max<size_t>(1, (n + 255) / 256)); }

}  // namespace

int wiener_source(float* Gs, float* Go, int nt, int nr, float dt, float f1, float f2, const int* it2_host,
                  float* sst, hipStream_t st, std::string* err) {
    const int npad = 2 * nt;
    const float df = 1.0f / ((float)nt * dt);
    const int if1 = 1 + (int)std::floor(f1 / df), if2 = 1 + (int)std::floor(f2 / df);
    const int b0 = std::max(if1 - 1, 0), b1 = std::min(if2, npad);
    int status = 0;
    auto fail = [&](const std::string& m) { if (err) *err = m; status = -1; };
    int* it2 = nullptr;
    unsigned int* maxbits = nullptr;
    float2* rec = nullptr;
    double* kk = nullptr;
    double2* bb = nullptr;
    hipfftHandle plan = 0, plan1 = 0;
    bool have = false, have1 = false;
    do {
        if (b1 <= b0) { fail("source estimation: empty frequency band"); break; }
        if (hipMalloc(&it2, nr * sizeof(int)) != hipSuccess || hipMalloc(&maxbits, 2 * sizeof(unsigned)) != hipSuccess ||
            hipMalloc(&rec, (size_t)2 * nr * npad * sizeof(float2)) != hipSuccess ||
            hipMalloc(&kk, (size_t)(b1 - b0) * sizeof(double)) != hipSuccess ||
            hipMalloc(&bb, (size_t)(b1 - b0) * sizeof(double2)) != hipSuccess) {
            fail("source estimation: device allocation failed");
            break;
        }
        (void)hipMemcpyAsync(it2, it2_host, nr * sizeof(int), hipMemcpyHostToDevice, st);
        (void)hipMemsetAsync(maxbits, 0, 2 * sizeof(unsigned), st);
        hipLaunchKernelGGL(window_maxabs_kernel, dim3(blocks_for((size_t)nr * nt)), dim3(256), 0, st, Gs, nt, nr, it2,
                           maxbits);
        hipLaunchKernelGGL(window_maxabs_kernel, dim3(blocks_for((size_t)nr * nt)), dim3(256), 0, st, Go, nt, nr, it2,
                           maxbits + 1);
        hipLaunchKernelGGL(pack_kernel, dim3(blocks_for((size_t)2 * nr * npad)), dim3(256), 0, st, Gs, Go, nt, nr,
                           npad, maxbits, rec);
        int n = npad;
        if (hipfftPlanMany(&plan, 1, &n, nullptr, 1, npad, nullptr, 1, npad, HIPFFT_C2C, 2 * nr) != HIPFFT_SUCCESS) {
            fail("source estimation: hipfftPlanMany failed");
            break;
        }
        have = true;
        hipfftSetStream(plan, st);
        if (hipfftExecC2C(plan, reinterpret_cast<hipfftComplex*>(rec), reinterpret_cast<hipfftComplex*>(rec),
                          HIPFFT_FORWARD) != HIPFFT_SUCCESS) {
            fail("source estimation: forward FFT failed");
            break;
        }
        hipLaunchKernelGGL(band_sums_kernel, dim3((b1 - b0 + 255) / 256), dim3(256), 0, st, rec, nr, npad, b0, b1,
                           kk, bb);
        std::vector<double> hk(b1 - b0);
        std::vector<double2> hb(b1 - b0);
        if (hipMemcpyAsync(hk.data(), kk, hk.size() * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hb.data(), bb, hb.size() * sizeof(double2), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            fail("source estimation: download failed");
            break;
        }
        double mx = hk[0];
        for (double v : hk) mx = std::max(mx, v);
        const double eps = 0.001 * mx;
        std::vector<float2> ss(npad, float2{0.0f, 0.0f});
        for (int b = b0; b < b1; ++b) {
            const std::complex<double> q =
                std::complex<double>(hb[b - b0].x, hb[b - b0].y) / std::complex<double>(hk[b - b0] + eps, 0.0);
            ss[b] = float2{(float)q.real(), (float)q.imag()};
        }
        if (hipfftPlan1d(&plan1, npad, HIPFFT_C2C, 1) != HIPFFT_SUCCESS) {
            fail("source estimation: hipfftPlan1d failed");
            break;
        }
        have1 = true;
        hipfftSetStream(plan1, st);
        if (hipMemcpyAsync(rec, ss.data(), npad * sizeof(float2), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipfftExecC2C(plan1, reinterpret_cast<hipfftComplex*>(rec), reinterpret_cast<hipfftComplex*>(rec),
                          HIPFFT_BACKWARD) != HIPFFT_SUCCESS) {
            fail("source estimation: inverse FFT failed");
            break;
        }
        std::vector<float2> out(nt);
        if (hipMemcpyAsync(out.data(), rec, nt * sizeof(float2), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            fail("source estimation: download failed");
            break;
        }
        for (int i = 0; i < nt; ++i) sst[i] = out[i].x / (float)npad;
    } while (false);
    if (have) hipfftDestroy(plan);
    if (have1) hipfftDestroy(plan1);
    for (void* p : {(void*)it2, (void*)maxbits, (void*)rec, (void*)kk, (void*)bb})
        if (p) (void)hipFree(p);
    return status;
}

hipError_t traces_to_time_major(const float* in, int nt, int ntr, float* out, hipStream_t st) {
    if (ntr <= 0 || nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(time_major_kernel, dim3(blocks_for((size_t)ntr * nt)), dim3(256), 0, st, in, nt, ntr, out);
    return hipGetLastError();
}

int trace_filter(float* traces, int nt, int ntr, float dt, int typef, float f1, float fc, hipStream_t st,
                 std::string* err) {
    if (ntr <= 0 || nt <= 0) return 0;
    const int N = 24 * nt;                     // time_filter.f90:14-15
    const int half = N / 2 + 1;
    std::vector<float2> H;
    butterworth_response(N, dt, typef, f1, fc, H);
    // chunk of traces sized to ~1 GiB of FFT buffers
    const size_t per_trace = (size_t)N * sizeof(float) + (size_t)half * sizeof(float2);
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)ntr, ((size_t)1 << 30) / per_trace));
    float* rec = nullptr;
    float2* spec = nullptr;
    float2* dH = nullptr;
    hipfftHandle fwd = 0, bwd = 0;
    bool have_fwd = false, have_bwd = false;
    int status = 0;
    auto fail = [&](const std::string& m) { if (err) *err = m; status = -1; };
    do {
        if (hipMalloc(&rec, (size_t)chunk * N * sizeof(float)) != hipSuccess ||
            hipMalloc(&spec, (size_t)chunk * half * sizeof(float2)) != hipSuccess ||
            hipMalloc(&dH, (size_t)half * sizeof(float2)) != hipSuccess) {
            fail("trace filter: device allocation failed");
            break;
        }
        if (hipMemcpyAsync(dH, H.data(), half * sizeof(float2), hipMemcpyHostToDevice, st) != hipSuccess) {
            fail("trace filter: response upload failed");
            break;
        }
        int n = N;
        if (hipfftPlanMany(&fwd, 1, &n, nullptr, 1, N, nullptr, 1, half, HIPFFT_R2C, chunk) != HIPFFT_SUCCESS) {
            fail("trace filter: hipfftPlanMany (R2C) failed");
            break;
        }
        have_fwd = true;
        if (hipfftPlanMany(&bwd, 1, &n, nullptr, 1, half, nullptr, 1, N, HIPFFT_C2R, chunk) != HIPFFT_SUCCESS) {
            fail("trace filter: hipfftPlanMany (C2R) failed");
            break;
        }
        have_bwd = true;
        hipfftSetStream(fwd, st);
        hipfftSetStream(bwd, st);
        for (int t0 = 0; t0 < ntr && status == 0; t0 += chunk) {
            const int nb = std::min(chunk, ntr - t0);
            float* tr = traces + (size_t)t0 * nt;
            if (nb < chunk) (void)hipMemsetAsync(rec, 0, (size_t)chunk * N * sizeof(float), st);
            hipLaunchKernelGGL(pad_kernel, dim3(blocks_for((size_t)nb * N)), dim3(256), 0, st, tr, nt, N, nb, rec);
            if (hipfftExecR2C(fwd, rec, reinterpret_cast<hipfftComplex*>(spec)) != HIPFFT_SUCCESS) {
                fail("trace filter: R2C failed");
                break;
            }
            hipLaunchKernelGGL(apply_response_kernel, dim3(blocks_for((size_t)nb * half)), dim3(256), 0, st, spec, dH,
                               half, nb);
            if (hipfftExecC2R(bwd, reinterpret_cast<hipfftComplex*>(spec), rec) != HIPFFT_SUCCESS) {
                fail("trace filter: C2R failed");
                break;
            }
            hipLaunchKernelGGL(extract_kernel, dim3(blocks_for((size_t)nb * nt)), dim3(256), 0, st, rec, nt, N, nb,
                               1.0f / (float)N, tr);
            if (hipGetLastError() != hipSuccess) fail("trace filter: kernel launch failed");
        }
        if (status == 0 && hipStreamSynchronize(st) != hipSuccess) fail("trace filter: stream error");
    } while (false);
    if (have_fwd) hipfftDestroy(fwd);
    if (have_bwd) hipfftDestroy(bwd);
    if (rec) (void)hipFree(rec);
    if (spec) (void)hipFree(spec);
    if (dH) (void)hipFree(dH);
    return status;
}

hipError_t akima_resample(const float* in, float dt_in, int nt_in, float* out, float dt_out, int nt_out, int ntr,
                          hipStream_t st) {
    if (ntr <= 0 || nt_out <= 0) return hipSuccess;
    hipLaunchKernelGGL(akima_kernel, dim3(blocks_for((size_t)ntr * nt_out)), dim3(256), 0, st, in, dt_in, nt_in, out,
                       dt_out, nt_out, ntr);
    return hipGetLastError();
}

}  // namespace fwi
