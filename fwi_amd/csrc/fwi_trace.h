// Trace conditioning on the GPU (Butterworth filter, Akima resampling).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace fwi {

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
