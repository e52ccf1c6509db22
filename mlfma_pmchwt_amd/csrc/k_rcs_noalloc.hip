// Allocation-free far field and RCS (mlfma_rcs_device_noalloc).
//
// mlfma_rcs stages the observation directions and the per-block partial sums of
// the far-field integral in device buffers it allocates on every call.  This
// variant keeps both in one file-scope device array, `scratch`, of a fixed size
// (kRcsScratchBytes per GPU, held until the process ends), and its inputs
// (currents, angles) and output (dBsm values) are device memory the caller owns,
// so a call allocates nothing, on the device or the host.
//
// The array's name is a host-side stand-in in host code; its device address is
// looked up once per device at set-up (rcs_scratch_init, called when a handle
// is built, with that device current) and only that address is used afterwards.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (two handles on different host threads) race.  Nothing in the array needs to
// start at zero: every call writes each direction and partial sum it reads.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "cplx.hpp"
#include "engine.hpp"
#include "setup.hpp"

namespace mlfma {

constexpr size_t kRcsScratchBytes = size_t(8) << 20;   // 8 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kRcsScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Looks up this device's copy of `scratch` (the current device must be `device`).
void rcs_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(MLFMA_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    int cur = -1;
    MLFMA_HIP(hipGetDevice(&cur));
    if (cur != device) throw Error(MLFMA_ERR_STATE, "scratch lookup with another device current");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(MLFMA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
}

size_t rcs_scratch_capacity() { return kRcsScratchBytes; }

// Observation directions (sin t cos p, sin t sin p, cos t), t and p in radians from degrees.
__global__ __launch_bounds__(256) void k_rcs_dirs(int nang, const double* __restrict__ th,
                                                  const double* __restrict__ ph, double* __restrict__ dirs) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nang) return;
    const double t = th[a] / 180.0 * M_PI, p = ph[a] / 180.0 * M_PI;
    dirs[3 * a] = sin(t) * cos(p);
    dirs[3 * a + 1] = sin(t) * sin(p);
    dirs[3 * a + 2] = cos(t);
}

__device__ __forceinline__ double snap6(double v) { return fabs(v) < 1e-6 ? 0.0 : v; }

// One wavefront per direction: the block partial sums of k_farfield summed,
// scaled by the field constant, projected on the theta and phi unit vectors of
// the direction, and turned into dBsm (Code/RCS.h:54-192, :247-259).
__global__ __launch_bounds__(64) void k_rcs_finish(int nblk, const double2* __restrict__ partial, double2 coeff,
                                                   const double* __restrict__ th, const double* __restrict__ ph,
                                                   double* __restrict__ out) {
    const int a = blockIdx.x, lane = threadIdx.x;
    double2 F[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int b = lane; b < nblk; b += 64)
        for (int c = 0; c < 3; ++c) F[c] = cadd(F[c], partial[((size_t)a * nblk + b) * 3 + c]);
    for (int c = 0; c < 3; ++c) F[c] = cmul(wave_sum(F[c]), coeff);
    if (lane != 0) return;
    auto rad = [](double d) {
        const double b = d * M_PI / 180.0;
        return fabs(b) < 1e-6 ? 0.0 : b;
    };
    const double t = rad(th[a]), p = rad(ph[a]);
    const double vth[3] = {snap6(-cos(t) * cos(p)), snap6(-cos(t) * sin(p)), snap6(sin(t))};
    const double vph[3] = {snap6(-sin(p)), snap6(cos(p)), 0.0};
    double2 Ev = {0, 0}, Eh = {0, 0};
    for (int c = 0; c < 3; ++c) {
        Ev = cadd(Ev, cscale(F[c], vth[c]));
        Eh = cadd(Eh, cscale(F[c], vph[c]));
    }
    const double nv = Ev.x * Ev.x + Ev.y * Ev.y, nh = Eh.x * Eh.x + Eh.y * Eh.y;
    const double fpi = 4.0 * M_PI;
    out[3 * a] = 10.0 * log10((nv + nh) / fpi);
    out[3 * a + 1] = 10.0 * log10(nv / fpi);
    out[3 * a + 2] = 10.0 * log10(nh / fpi);
}

// Scratch layout: [directions: 3 * nang doubles, padded to 256 B][partials: nang * nblk * 3 double2].
void launch_rcs_noalloc(hipStream_t st, int device, int N, int pmchwt, const RwgGeo* geo, const int* perm,
                        const GaussRule& g, double2 k1, double2 coeff, int nang, const double* th_dev,
                        const double* ph_dev, const double2* I_dev, double* out_dev) {
    if (device < 0 || device >= kMaxDevices || !g_scratch_ptr[device])
        throw Error(MLFMA_ERR_STATE, "allocation-free RCS: scratch not set up on this device");
    const int nblk = (N + 255) / 256;
    const size_t dir_bytes = (3 * (size_t)nang * sizeof(double) + 255) / 256 * 256;
    const size_t need = dir_bytes + (size_t)nang * nblk * 3 * sizeof(double2);
    if (need > kRcsScratchBytes) {
        const size_t per = 3 * sizeof(double) + (size_t)nblk * 3 * sizeof(double2);
        throw Error(MLFMA_ERR_ARG, "mlfma_rcs_device_noalloc: " + std::to_string(nang) + " angles need " +
                                       std::to_string(need) + " B of scratch, more than its " +
                                       std::to_string(kRcsScratchBytes) + " B (at most about " +
                                       std::to_string((kRcsScratchBytes - 256) / per) + " angles per call for " +
                                       std::to_string(N) + " RWG functions)");
    }
    unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[device]);
    double* dirs = reinterpret_cast<double*>(base);
    double2* part = reinterpret_cast<double2*>(base + dir_bytes);
    k_rcs_dirs<<<(nang + 255) / 256, 256, 0, st>>>(nang, th_dev, ph_dev, dirs);
    launch_farfield(st, N, pmchwt, geo, perm, g, k1, nang, dirs, I_dev, 0, part, nblk);
    k_rcs_finish<<<nang, 64, 0, st>>>(nblk, part, coeff, th_dev, ph_dev, out_dev);
}

}  // namespace mlfma
