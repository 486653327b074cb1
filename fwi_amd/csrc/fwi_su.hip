// SU files and the host-side Akima interpolation of the geometry set-up.
//
// An SU file is a sequence of traces, each a 240-byte header (60 32-bit
// words, SEG-Y trace-header layout: ns at bytes 115-116 and dt in
// microseconds at bytes 117-118, both 16-bit) followed by ns float32
// samples, in little- or big-endian byte order (auxiliar.f90:9-50,
// get_mcs_data.f90:207-240, 291-326; the reference opens the files as
// Fortran streams with CONVERT='BIG_ENDIAN' when asked).  Written files carry
// tracl = 1..ntr, ns and dt in their headers, as the reference's
// `suaddhead ns=... | sushw key=dt a=...` pipeline produces
// (auxiliar.f90:512-519).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fwi.h"
#include "fwi_trace.h"

namespace {

thread_local std::string su_err;

int su_fail(const std::string& m) {
    su_err = m;
    return -1;
}

constexpr long kHeaderBytes = 240;

bool host_big_endian() {
    const uint16_t v = 1;
    uint8_t b;
    std::memcpy(&b, &v, 1);
    return b == 0;
}

void swap_bytes(void* p, int n) {
    uint8_t* b = static_cast<uint8_t*>(p);
    for (int i = 0; i < n / 2; ++i) std::swap(b[i], b[n - 1 - i]);
}

// file byte order -> host byte order for a value of n bytes
void to_host(void* p, int n, bool file_big) {
    if (file_big != host_big_endian()) swap_bytes(p, n);
}

struct File {
    std::FILE* f = nullptr;
    explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File() { if (f) std::fclose(f); }
};

long file_size(std::FILE* f) {
    if (std::fseek(f, 0, SEEK_END) != 0) return -1;
    const long n = std::ftell(f);
    std::rewind(f);
    return n;
}

}  // namespace

extern "C" {

const char* fwi_su_last_error(void) { return su_err.c_str(); }

int fwi_su_info(const char* path, int big_endian, int* ns, int* ntr, int* dt_us) {
    File F(path, "rb");
    if (!F.f) return su_fail(std::string("cannot open ") + path);
    const long size = file_size(F.f);
    uint8_t h[kHeaderBytes];
    if (size < kHeaderBytes || std::fread(h, 1, kHeaderBytes, F.f) != (size_t)kHeaderBytes)
        return su_fail(std::string(path) + ": shorter than one trace header");
    int16_t n16, dt16;
    std::memcpy(&n16, h + 114, 2);
    std::memcpy(&dt16, h + 116, 2);
    to_host(&n16, 2, big_endian != 0);
    to_host(&dt16, 2, big_endian != 0);
    const int n = (uint16_t)n16;
    if (n <= 0) return su_fail(std::string(path) + ": ns = 0 in the first header");
    const long trace_bytes = kHeaderBytes + 4L * n;
    if (size % trace_bytes != 0)
        return su_fail(std::string(path) + ": size is not a whole number of traces of ns = " + std::to_string(n));
    if (ns) *ns = n;
    if (ntr) *ntr = (int)(size / trace_bytes);
    if (dt_us) *dt_us = (uint16_t)dt16;
    return 0;
}

int fwi_su_read(const char* path, int big_endian, int ns, int first, int ntr, float* data) {
    if (ns <= 0 || first < 0 || ntr < 0) return su_fail("fwi_su_read: bad arguments");
    File F(path, "rb");
    if (!F.f) return su_fail(std::string("cannot open ") + path);
    const long trace_bytes = kHeaderBytes + 4L * ns;
    for (int j = 0; j < ntr; ++j) {
        const long pos = (long)(first + j) * trace_bytes + kHeaderBytes;
        float* out = data + (size_t)j * ns;
        if (std::fseek(F.f, pos, SEEK_SET) != 0 || std::fread(out, 4, ns, F.f) != (size_t)ns)
            return su_fail(std::string(path) + ": trace " + std::to_string(first + j) + " is incomplete");
        for (int i = 0; i < ns; ++i) to_host(out + i, 4, big_endian != 0);
    }
    return 0;
}

int fwi_su_header_word(const char* path, int big_endian, int ns, int trace, int byte_pos, int nbytes,
                       long long* value) {
    if (ns <= 0 || trace < 0 || byte_pos < 1 || byte_pos + nbytes - 1 > kHeaderBytes || (nbytes != 2 && nbytes != 4))
        return su_fail("fwi_su_header_word: bad arguments");
    File F(path, "rb");
    if (!F.f) return su_fail(std::string("cannot open ") + path);
    const long pos = (long)trace * (kHeaderBytes + 4L * ns) + (byte_pos - 1);
    uint8_t b[4];
    if (std::fseek(F.f, pos, SEEK_SET) != 0 || std::fread(b, 1, nbytes, F.f) != (size_t)nbytes)
        return su_fail(std::string(path) + ": header of trace " + std::to_string(trace) + " is incomplete");
    to_host(b, nbytes, big_endian != 0);
    if (nbytes == 2) {
        int16_t v;
        std::memcpy(&v, b, 2);
        *value = v;
    } else {
        int32_t v;
        std::memcpy(&v, b, 4);
        *value = v;
    }
    return 0;
}

int fwi_su_write(const char* path, int big_endian, const float* data, int ns, int ntr, float dt) {
    if (ns <= 0 || ns > 65535 || ntr < 0) return su_fail("fwi_su_write: bad arguments");
    File F(path, "wb");
    if (!F.f) return su_fail(std::string("cannot create ") + path);
    const int dt_head = (int)(dt * 1e6f);          // auxiliar.f90:503 (integer from REAL, truncated)
    std::vector<uint8_t> h(kHeaderBytes);
    std::vector<float> tr(ns);
    for (int j = 0; j < ntr; ++j) {
        std::fill(h.begin(), h.end(), 0);
        int32_t tracl = j + 1;
        int16_t n16 = (int16_t)(uint16_t)ns, dt16 = (int16_t)(uint16_t)dt_head;
        to_host(&tracl, 4, big_endian != 0);
        to_host(&n16, 2, big_endian != 0);
        to_host(&dt16, 2, big_endian != 0);
        std::memcpy(h.data(), &tracl, 4);
        std::memcpy(h.data() + 114, &n16, 2);
        std::memcpy(h.data() + 116, &dt16, 2);
        for (int i = 0; i < ns; ++i) {
            tr[i] = data[(size_t)j * ns + i];
            to_host(&tr[i], 4, big_endian != 0);
        }
        if (std::fwrite(h.data(), 1, kHeaderBytes, F.f) != (size_t)kHeaderBytes ||
            std::fwrite(tr.data(), 4, ns, F.f) != (size_t)ns)
            return su_fail(std::string(path) + ": write failed");
    }
    return 0;
}

int fwi_akima(int L, const float* X, const float* Y, int N, const float* U, float* V) {
    if (L < 2) return 1;                            // akima433.f90 ERR codes
    if (N <= 0) return 2;
    for (int i = 1; i < L; ++i) {
        if (X[i - 1] - X[i] < 0.0f) continue;
        return X[i - 1] - X[i] == 0.0f ? 6 : 7;
    }
    for (int k = 0; k < N; ++k) V[k] = fwi::akima_eval(X, 0.0f, Y, L, U[k]);
    return 0;
}

}  // extern "C"
