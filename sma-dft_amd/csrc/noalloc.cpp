// smad_r2r_3d_noalloc: the separable 3D r2r transform of smad_r2r_3d on
// caller-owned device buffers, for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops that forbid
// hipMalloc).  It has no counterpart in the reference.
//
// Device memory beyond the caller's input and output lives in one
// file-scope array, `scratch`, that every GPU holds until the process ends:
//   [0, TW64_BYTES)            quarter-wave cosine tables, FP64, every line length
//   [TW64_BYTES, +TW32_BYTES)  the same tables in FP32
//   [STAGE_OFF, +STAGE_BYTES)  FP32 staging of the field (precision 32 only),
//                              at most SMAD_NOALLOC_MAX_F32_POINTS points
// The tables are written once per device at set-up (the first call on that
// device) and only read afterwards; the staging area is overwritten by every
// FP32 call, so FP32 calls that overlap on one GPU race.
#include "../../include/smadft/smadft.h"
#include "kernels.hpp"

#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

using namespace smad;

namespace {

constexpr int kLengths[] = {16, 32, 64, 128, 256, 384, 512, 1024};
constexpr int kNumLengths = sizeof(kLengths) / sizeof(kLengths[0]);

constexpr size_t table_values()
{
    size_t v = 0;
    for (int i = 0; i < kNumLengths; ++i) v += (size_t)kLengths[i] + 1;
    return v;
}
constexpr size_t kTableValues = table_values();
constexpr size_t round256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t TW64_BYTES = round256(kTableValues * sizeof(double));
constexpr size_t TW32_BYTES = round256(kTableValues * sizeof(float));
constexpr size_t STAGE_OFF = TW64_BYTES + TW32_BYTES;
constexpr size_t STAGE_BYTES = (size_t)SMAD_NOALLOC_MAX_F32_POINTS * sizeof(float);
constexpr size_t SCRATCH_BYTES = STAGE_OFF + STAGE_BYTES;

// Device-side scratch.  In host code `scratch` is only a handle for
// hipGetSymbolAddress; kernels receive the per-device address from it.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
struct DeviceSlot {
    std::once_flag once;
    int rc = SMAD_OK;
    std::string err;
    unsigned char* ptr = nullptr;  // device address of `scratch` on this device
};
DeviceSlot g_slots[kMaxDevices];
thread_local std::string g_noalloc_error;

int table_offset(int n)
{
    size_t off = 0;
    for (int i = 0; i < kNumLengths; ++i) {
        if (kLengths[i] == n) return (int)off;
        off += (size_t)kLengths[i] + 1;
    }
    return -1;
}

// Host copies of the tables, in static storage: c[j] = cos(pi j / 2N),
// j = 0..N, c[N] = 0, computed in long double and rounded once (the values
// the context-owned tables of smad_r2r_3d hold).
double g_host64[kTableValues];
float g_host32[kTableValues];

void fill_host_tables()
{
    const long double pi = 3.141592653589793238462643383279502884L;
    size_t off = 0;
    for (int i = 0; i < kNumLengths; ++i) {
        const int n = kLengths[i];
        for (int j = 0; j <= n; ++j) {
            long double c = j == n ? 0.0L : cosl(pi * j / (2.0L * n));
            g_host64[off + j] = (double)c;
            g_host32[off + j] = (float)c;
        }
        off += (size_t)n + 1;
    }
}

// Once per device, with that device current: resolve the device address of
// `scratch` and write the twiddle tables into it.
void setup_device(int dev, DeviceSlot& slot)
{
    static std::once_flag host_once;
    std::call_once(host_once, fill_host_tables);
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        slot.rc = SMAD_EHIP;
        slot.err = std::string("hipGetSymbolAddress(scratch) on device ") + std::to_string(dev) + ": " +
                   hipGetErrorString(e);
        return;
    }
    unsigned char* base = static_cast<unsigned char*>(p);
    e = hipMemcpy(base, g_host64, sizeof(g_host64), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(base + TW64_BYTES, g_host32, sizeof(g_host32), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        slot.rc = SMAD_EHIP;
        slot.err = std::string("writing the twiddle tables into scratch: ") + hipGetErrorString(e);
        return;
    }
    slot.ptr = base;
}

__global__ __launch_bounds__(WG) void k_to_f32(const double* __restrict__ a, float* __restrict__ b, long n)
{
    for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) b[i] = (float)a[i];
}
__global__ __launch_bounds__(WG) void k_to_f64(const float* __restrict__ a, double* __restrict__ b, long n)
{
    for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) b[i] = (double)a[i];
}

dim3 conv_grid(long n)
{
    long b = (n + WG - 1) / WG;
    return dim3((unsigned)(b < 8192 ? b : 8192));
}

bool length_supported(int n) { return table_offset(n) >= 0; }

template <typename T>
void transform(hipStream_t s, const T* in, T* out, Dims d, const int kinds[3], const unsigned char* tables)
{
    const T* tab = reinterpret_cast<const T*>(tables);
    Tw<T> tx{tab + table_offset(d.nx)}, ty{tab + table_offset(d.ny)}, tz{tab + table_offset(d.nz)};
    Launch<T>::axis(s, in, out, d, 2, kinds[2], T(1), tz);
    Launch<T>::axis(s, out, out, d, 1, kinds[1], T(1), ty);
    Launch<T>::axis(s, out, out, d, 0, kinds[0], T(1), tx);
}

}  // namespace

extern "C" {

int smad_r2r_3d_noalloc(int nx, int ny, int nz, const int kinds[3], int precision, const double* in_dev,
                        double* out_dev, void* stream)
{
    g_noalloc_error.clear();
    if (!kinds || !in_dev || !out_dev) {
        g_noalloc_error = "smad_r2r_3d_noalloc: null argument";
        return SMAD_EINVAL;
    }
    if (!length_supported(nx) || !length_supported(ny) || !length_supported(nz)) {
        g_noalloc_error = "smad_r2r_3d_noalloc: each grid size must be one of 16, 32, 64, 128, 256, 384, 512, 1024";
        return SMAD_EINVAL;
    }
    for (int a = 0; a < 3; ++a)
        if (kinds[a] < SMAD_REDFT10 || kinds[a] > SMAD_RODFT01) {
            g_noalloc_error = "smad_r2r_3d_noalloc: bad kind";
            return SMAD_EINVAL;
        }
    if (precision != SMAD_FP64 && precision != SMAD_FP32_MIXED) {
        g_noalloc_error = "smad_r2r_3d_noalloc: precision must be 64 or 32";
        return SMAD_EINVAL;
    }
    const long n = (long)nx * ny * nz;
    if (precision == SMAD_FP32_MIXED && n > (long)SMAD_NOALLOC_MAX_F32_POINTS) {
        g_noalloc_error = "smad_r2r_3d_noalloc: FP32 arithmetic stages the field in a fixed scratch area of " +
                          std::to_string((long)SMAD_NOALLOC_MAX_F32_POINTS) + " points; this grid has " +
                          std::to_string(n);
        return SMAD_EINVAL;
    }
    if (((uintptr_t)in_dev | (uintptr_t)out_dev) & 15) {
        g_noalloc_error = "smad_r2r_3d_noalloc: device buffers must be 16-byte aligned";
        return SMAD_EINVAL;
    }
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        g_noalloc_error = std::string("smad_r2r_3d_noalloc: no usable current device: ") + hipGetErrorString(e);
        return SMAD_EHIP;
    }
    DeviceSlot& slot = g_slots[dev];
    std::call_once(slot.once, setup_device, dev, std::ref(slot));
    if (slot.rc != SMAD_OK) {
        g_noalloc_error = slot.err;
        return slot.rc;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Dims d{nx, ny, nz};
    try {
        if (precision == SMAD_FP64) {
            transform<double>(s, in_dev, out_dev, d, kinds, slot.ptr);
        } else {
            float* stage = reinterpret_cast<float*>(slot.ptr + STAGE_OFF);
            hipLaunchKernelGGL(k_to_f32, conv_grid(n), dim3(WG), 0, s, in_dev, stage, n);
            transform<float>(s, stage, stage, d, kinds, slot.ptr + TW64_BYTES);
            hipLaunchKernelGGL(k_to_f64, conv_grid(n), dim3(WG), 0, s, stage, out_dev, n);
        }
    } catch (const std::exception& e) {  // no C++ exception crosses the C ABI
        g_noalloc_error = std::string("smad_r2r_3d_noalloc: ") + e.what();
        return SMAD_EINVAL;
    }
    e = hipGetLastError();
    if (e != hipSuccess) {
        g_noalloc_error = std::string("smad_r2r_3d_noalloc: kernel launch: ") + hipGetErrorString(e);
        return SMAD_EHIP;
    }
    return SMAD_OK;
}

const char* smad_noalloc_last_error(void) { return g_noalloc_error.c_str(); }

}  // extern "C"
