// Trace conditioning on the GPU (Butterworth filter, Akima resampling).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include <string>
#include <vector>

namespace fwi {

// One output value of ACM Algorithm 433 as modified in akima433.f90 (INTRPL)
// at abscissa UK, from L >= 2 points with increasing abscissae X (or, when X
// is null, X(i) = (i-1) dx) and ordinates Y.  Statement order of the Fortran;
// its per-interval coefficient reuse and EQUIVALENCE'd temporaries do not
// change values, so every point is evaluated on its own.
__host__ __device__ inline float akima_eval(const float* X, float dx, const float* Y, int L, float UK) {
    const int L0 = L, LM1 = L0 - 1, LM2 = LM1 - 1, LP1 = L0 + 1;
    auto x = [&](int i) { return X ? X[i - 1] : (float)(i - 1) * dx; };
    auto y = [&](int i) { return Y[i - 1]; };
    int I;
    if (LM2 == 0) I = 2;
    else if (UK >= x(L0)) I = LP1;
    else if (UK < x(1)) I = 1;
    else {
        int IMN = 2, IMX = L0;
        do {
            I = (IMN + IMX) / 2;
            if (UK >= x(I)) IMN = I + 1;
            else IMX = I;
        } while (IMX > IMN);
        I = IMX;
    }
    int J = I;
    if (J == 1) J = 2;
    if (J == LP1) J = L0;
    float X3 = x(J - 1), Y3 = y(J - 1), X4 = x(J), Y4 = y(J);
    float A3 = X4 - X3;
    float M3 = (Y4 - Y3) / A3;
    float M1 = 0, M2 = 0, M4 = 0, M5 = 0, A2 = 0, A4 = 0, X2 = 0, Y2 = 0, X5 = 0, Y5 = 0;
    if (LM2 == 0) {
        M2 = M3;
        M4 = M3;
    } else {
        bool to41 = J == 2;
        if (!to41) {
            X2 = x(J - 2); Y2 = y(J - 2);
            A2 = X3 - X2;
            M2 = (Y3 - Y2) / A2;
            if (J == L0) M4 = M3 + M3 - M2;
            else to41 = true;
        }
        if (to41) {
            X5 = x(J + 1); Y5 = y(J + 1);
            A4 = X5 - X4;
            M4 = (Y5 - Y4) / A4;
            if (J == 2) M2 = M3 + M3 - M4;
        }
    }
    if (J <= 3) M1 = M2 + M2 - M3;
    else {
        const float A1 = X2 - x(J - 3);
        M1 = (Y2 - y(J - 3)) / A1;
    }
    if (J >= LM1) M5 = M4 + M4 - M3;
    else {
        const float A5 = x(J + 2) - X5;
        M5 = (y(J + 2) - Y5) / A5;
    }
    float T3 = 0, T4 = 0;
    const bool lp1 = I == LP1;
    if (!lp1) {
        float W2 = fabsf(M4 - M3), W3 = fabsf(M2 - M1);
        float SW = W2 + W3;
        if (!(SW != 0.0f)) { W2 = 0.5f; W3 = 0.5f; SW = 1.0f; }
        T3 = (W2 * M2 + W3 * M3) / SW;
    }
    if (I == 1) {
        T4 = T3;
        const float SA = A3 + A4;
        T3 = 0.5f * (M1 + M2 - A4 * (A3 - A4) * (M3 - M4) / (SA * SA));
        X3 = X3 - A4;
        Y3 = Y3 - M2 * A4;
        A3 = A4;
        M3 = M2;
    } else {
        float W3 = fabsf(M5 - M4), W4 = fabsf(M3 - M2);
        float SW = W3 + W4;
        if (!(SW != 0.0f)) { W3 = 0.5f; W4 = 0.5f; SW = 1.0f; }
        T4 = (W3 * M3 + W4 * M4) / SW;
        if (lp1) {
            T3 = T4;
            const float SA = A2 + A3;
            T4 = 0.5f * (M4 + M5 - A2 * (A2 - A3) * (M2 - M3) / (SA * SA));
            X3 = X4;
            Y3 = Y4;
            A3 = A2;
            M3 = M4;
        }
    }
    const float Q2 = (2.0f * (M3 - T3) + M3 - T4) / A3;
    const float Q3 = (-M3 - M3 + T3 + T4) / (A3 * A3);
    const float DX = UK - X3;
    return Y3 + DX * (T3 + DX * (Q2 + DX * Q3));
}

// Response of the order-6 Butterworth filter of filters.f90 on bins
// 0..ts/2 (typef 1 low-pass at fc, 2 high-pass at 40 Hz, 3 band-pass f1..fc).
void butterworth_response(int ts, float dt, int typef, float f1, float fc, std::vector<float2>& H);
// Filters ntr device traces (nt, ntr) column-major in place, as time_filter.f90
// does; returns 0 or -1 with a message.
int trace_filter(float* traces, int nt, int ntr, float dt, int typef, float f1, float fc, hipStream_t st,
                 std::string* err);
// Akima resampling of device traces (nt_in, ntr) on (i-1) dt_in onto (i-1) dt_out, (nt_out, ntr).
hipError_t akima_resample(const float* in, float dt_in, int nt_in, float* out, float dt_out, int nt_out, int ntr,
                          hipStream_t st);
// Wiener estimate of the source (source_estimation.f90:79-167): Gs, Go device
// (nt, nr) column-major, windowed in place after it2_host[r] samples and
// normalised; sst (host, nt) receives the wavelet.  Returns 0 or -1.
int wiener_source(float* Gs, float* Go, int nt, int nr, float dt, float f1, float f2, const int* it2_host,
                  float* sst, hipStream_t st, std::string* err);
// (nt, ntr) column-major -> time-major [t][ntr]
hipError_t traces_to_time_major(const float* in, int nt, int ntr, float* out, hipStream_t st);

}  // namespace fwi
