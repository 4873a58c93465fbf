// tpsc_noalloc.hip -- the sum-rule evaluation of tpsc_sum_rule (chi.cpp:738-843) for callers
// that cannot allocate device memory (code under HIP graph capture, tight serving loops): the
// caller passes chi0(q,iqn) and its tail coefficients as device arrays and receives the value
// in device memory; the wedge weights and the partial sums live in a fixed device array of this
// library, so the call makes no allocation of any kind.
//
// The scratch array below exists once per GPU and is shared by every call on that GPU: calls
// that overlap on one GPU (two host threads, or two streams without an order between them)
// race on it.  Every call rewrites what it reads (weights, partials) before reading it.
#include <mutex>
#include <string>

#include "../../include/tpsc_capi.h"
#include "tpsc_launch.hpp"

namespace tpsc {

// scratch layout (doubles): [0, kNaParts) sum-rule partial sums, [kNaParts, kNaDoubles) wedge
// weights.  kNaDoubles = 16384 (128 KiB) holds N <= 256 (nk8 = 8385 weights, 1056 partials).
constexpr int kNaParts = 2048;
constexpr int kNaDoubles = 16384;
constexpr long kNaMaxWedge = kNaDoubles - kNaParts;

static __device__ __align__(256) double scratch[kNaDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each GPU
thread_local std::string g_err;

int na_fail(int code, const std::string& msg)
{
    g_err = "tpsc_sum_rule_noalloc: " + msg;
    return code;
}

// the device address of `scratch` on the current GPU, looked up once per device
int scratch_ptr(double** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return na_fail(TPSC_E_DEVICE, "hipGetDevice failed");
    if (dev < 0 || dev >= kMaxDevices) return na_fail(TPSC_E_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return na_fail(TPSC_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return TPSC_OK;
}

__global__ void k_na_weights(int nk, double* wq)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= (long)nk * nk) return;
    const int l = (int)(i / nk), m = (int)(i % nk);
    if (m <= l) wq[widx(l, m)] = wedge_weight(l, m, nk);
}

// the partials in a fixed order, then the host-known terms of chi::sum_rule_sp / _ch
// (chi.cpp:833-843) in the order the context's evaluation uses
__global__ void k_na_value(const double* __restrict__ part, int np, int channel, double n, double nsr, double U,
                           double Ux, double Usp, double* out)
{
#pragma clang fp contract(off)
    __shared__ double sc[16];
    double a = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) a += part[i];
    a = block_sum(a, sc);
    if (threadIdx.x == 0) {
        if (channel == 0)
            out[0] = a + n - n * n / 2.0 - nsr + Ux * nsr * nsr / (2.0 * U);
        else
            out[0] = a + n - n * n / 2.0 - (nsr + Usp * nsr * nsr / (2.0 * U) - nsr * nsr);
    }
}

}  // namespace
}  // namespace tpsc

using namespace tpsc;

const char* tpsc_noalloc_error() { return tpsc::g_err.c_str(); }

extern "C" int tpsc_sum_rule_noalloc(const double* chi0_qn, const double* chi0_inf, const double* chi0_inf2, int N,
                                     int Nw, double T, double n, double U, double Usp, int channel, double Ux,
                                     double* value, void* stream)
{
    if (!chi0_qn || !chi0_inf || !chi0_inf2 || !value) return na_fail(TPSC_E_INVALID, "null array");
    if (!supported_N(N) || !supported_Nw(Nw))
        return na_fail(TPSC_E_INVALID, "unsupported grid N=" + std::to_string(N) + " Nw=" + std::to_string(Nw));
    const int nk = N / 2 + 1;
    const long nk8 = (long)nk * (nk + 1) / 2;
    if (nk8 > kNaMaxWedge || sum_rule_parts(N) > kNaParts)
        return na_fail(TPSC_E_INVALID, "N=" + std::to_string(N) + " exceeds the scratch capacity (" +
                                           std::to_string(kNaMaxWedge) + " wedge points, N <= 256)");
    if (channel != 0 && channel != 1) return na_fail(TPSC_E_INVALID, "channel must be 0 (spin) or 1 (charge)");
    if (!(T > 0.0) || !(U != 0.0)) return na_fail(TPSC_E_INVALID, "T must be > 0 and U != 0");
    double* scr = nullptr;
    if (int rc = scratch_ptr(&scr)) return rc;
    double* part = scr;
    double* wq = scr + kNaParts;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_na_weights, dim3((unsigned)(((long)nk * nk + 255) / 256)), dim3(256), 0, s, nk, wq);
    int np = 0;
    hipError_t e = launch_sum_rule_part(N, Nw, T, chi0_qn, chi0_inf, chi0_inf2, wq, channel == 0 ? -Ux : Ux, part,
                                        &np, s);
    if (e == hipSuccess) {
        const double nsr = (n > 1.0) ? 2.0 - n : n;  // chi.cpp:18-19
        hipLaunchKernelGGL(k_na_value, dim3(1), dim3(256), 0, s, part, np, channel, n, nsr, U, Ux, Usp, value);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return na_fail(TPSC_E_DEVICE, std::string("launch: ") + hipGetErrorString(e));
    return TPSC_OK;
}
