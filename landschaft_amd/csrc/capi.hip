// Library-wide entry points: error messages, device information, the MFMA
// layout self-test.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>

#include <cstring>
#include <string>

#include "../../include/landschaft_circuit.h"
#include "common.hpp"
#include "front.hpp"

namespace lsc {

void progress(const char* fmt, ...) {
    static const bool on = std::getenv("LSC_PROGRESS") != nullptr;
    static const auto t0 = std::chrono::steady_clock::now();
    if (!on) return;
    std::va_list ap;
    va_start(ap, fmt);
    std::fprintf(stderr, "[lsc %.3f s] ", std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    std::vfprintf(stderr, fmt, ap);
    std::fputc('\n', stderr);
    std::fflush(stderr);
    va_end(ap);
}

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }

// One wavefront: D = A B with A[i][k] = i + 16 k and B[k][j] = k + 4 j placed
// in the operand lanes the front kernels use, D read back through the same
// row map (Mfma<T>::row) and written row-major.
template <typename T>
__global__ void k_probe_mfma(double* out) {
    const int lane = threadIdx.x;
    typedef typename Mfma<T>::acc Acc;
    const int i = lane & 15, k = lane >> 4;
    const T a = static_cast<T>(i + 16 * k);
    const T b = static_cast<T>(k + 4 * i);  // B[k][j] with j = lane & 15
    Acc acc;
    acc[0] = acc[1] = acc[2] = acc[3] = T(0);
    acc = Mfma<T>::run(a, b, acc);
    for (int r = 0; r < 4; ++r) {
        const int row = Mfma<T>::row(lane, r), col = lane & 15;
        out[row * 16 + col] = static_cast<double>(acc[r]);
    }
}

}  // namespace lsc

extern "C" {

const char* lsc_last_error(void) { return lsc::g_error.c_str(); }

int lsc_device_info(char* name, int64_t name_len, int32_t* compute_units, int64_t* hbm_bytes) {
    return lsc::guarded([&] {
        int dev = 0;
        LSC_HIP(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        LSC_HIP(hipGetDeviceProperties(&prop, dev));
        if (name && name_len > 0) {
            std::strncpy(name, prop.gcnArchName, static_cast<std::size_t>(name_len) - 1);
            name[name_len - 1] = 0;
        }
        if (compute_units) *compute_units = prop.multiProcessorCount;
        if (hbm_bytes) *hbm_bytes = static_cast<int64_t>(prop.totalGlobalMem);
    });
}

int lsc_probe_mfma(int precision, double* out) {
    return lsc::guarded([&] {
        lsc::DevBuf<double> d(256);
        if (precision == LSC_PRECISION_FP64)
            hipLaunchKernelGGL(lsc::k_probe_mfma<double>, dim3(1), dim3(64), 0, nullptr, d.p);
        else
            hipLaunchKernelGGL(lsc::k_probe_mfma<float>, dim3(1), dim3(64), 0, nullptr, d.p);
        LSC_HIP(hipGetLastError());
        LSC_HIP(hipMemcpy(out, d.p, 256 * sizeof(double), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
