// Krylov-solver vector kernels (FP64).  Reductions use a fixed block partition
// and a fixed combination order, so repeated runs give identical results.
#include "cplx.hpp"
#include "setup.hpp"

namespace mlfma {

namespace {
__device__ __forceinline__ double2 block_sum(double2 v, double2* red) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    v = wave_sum(v);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double2 s = make_double2(0, 0);
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s = cadd(s, red[w]);
    return s;
}
}  // namespace

__global__ __launch_bounds__(256) void k_multidot(int n, const double2* __restrict__ V, size_t ldv,
                                                  const double2* __restrict__ w, double2* __restrict__ partial) {
    __shared__ double2 red[4];
    const int j = blockIdx.y;
    const double2* v = V + (size_t)j * ldv;
    double2 s = make_double2(0, 0);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s = cadd(s, cmulc(v[i], w[i]));
    s = block_sum(s, red);
    if (threadIdx.x == 0) partial[(size_t)j * gridDim.x + blockIdx.x] = s;
}

void launch_multidot(hipStream_t st, int n, int m, const double2* V, size_t ldv, const double2* w, double2* partial,
                     int nblk) {
    k_multidot<<<dim3(nblk, m), 256, 0, st>>>(n, V, ldv, w, partial);
}

__global__ __launch_bounds__(256) void k_reduce_partials(int nblk, const double2* __restrict__ partial,
                                                         double2* __restrict__ out) {
    __shared__ double2 red[4];
    const int j = blockIdx.x;
    double2 s = make_double2(0, 0);
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s = cadd(s, partial[(size_t)j * nblk + b]);
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[j] = s;
}

void launch_reduce_partials(hipStream_t st, int m, int nblk, const double2* partial, double2* out) {
    k_reduce_partials<<<m, 256, 0, st>>>(nblk, partial, out);
}

__global__ __launch_bounds__(256) void k_sub_combination(int n, int m, const double2* __restrict__ V, size_t ldv,
                                                         const double2* __restrict__ h, double2* __restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2 a = w[i];
    for (int j = 0; j < m; ++j) cfms(a, h[j], V[(size_t)j * ldv + i]);
    w[i] = a;
}

void launch_sub_combination(hipStream_t st, int n, int m, const double2* V, size_t ldv, const double2* h, double2* w) {
    k_sub_combination<<<(n + 255) / 256, 256, 0, st>>>(n, m, V, ldv, h, w);
}

__global__ __launch_bounds__(256) void k_add_combination(int n, int m, const double2* __restrict__ V, size_t ldv,
                                                         const double2* __restrict__ y, double2* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2 a = x[i];
    for (int j = 0; j < m; ++j) cfma(a, y[j], V[(size_t)j * ldv + i]);
    x[i] = a;
}

void launch_add_combination(hipStream_t st, int n, int m, const double2* V, size_t ldv, const double2* y, double2* x) {
    k_add_combination<<<(n + 255) / 256, 256, 0, st>>>(n, m, V, ldv, y, x);
}

__global__ void k_scale_copy(int n, double a, const double2* __restrict__ x, double2* __restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = cscale(x[i], a);
}
void launch_scale_copy(hipStream_t st, int n, double a, const double2* x, double2* y) {
    k_scale_copy<<<(n + 255) / 256, 256, 0, st>>>(n, a, x, y);
}

__global__ void k_residual(int n, const double2* __restrict__ b, double2* __restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = csub(b[i], r[i]);
}
void launch_residual(hipStream_t st, int n, const double2* b, double2* r) {
    k_residual<<<(n + 255) / 256, 256, 0, st>>>(n, b, r);
}

__global__ __launch_bounds__(256) void k_norm2_partial(int n, const double2* __restrict__ x,
                                                       double* __restrict__ partial) {
    __shared__ double2 red[4];
    double2 s = make_double2(0, 0);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        s.x += x[i].x * x[i].x + x[i].y * x[i].y;
    s = block_sum(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s.x;
}
void launch_norm2_partial(hipStream_t st, int n, const double2* x, double* partial, int nblk) {
    k_norm2_partial<<<nblk, 256, 0, st>>>(n, x, partial);
}

__global__ void k_axpy(int n, double2 a, const double2* __restrict__ z, const double2* __restrict__ x,
                       double2* __restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double2 t = x[i];
        cfma(t, a, z[i]);
        y[i] = t;
    }
}
void launch_axpy(hipStream_t st, int n, double2 a, const double2* z, const double2* x, double2* y) {
    k_axpy<<<(n + 255) / 256, 256, 0, st>>>(n, a, z, x, y);
}

__global__ __launch_bounds__(256) void k_bidot_partial(int n, const double2* __restrict__ a,
                                                       const double2* __restrict__ b, double2* __restrict__ partial) {
    __shared__ double2 red[4];
    double2 s = make_double2(0, 0);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) cfma(s, a[i], b[i]);
    s = block_sum(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
void launch_bidot_partial(hipStream_t st, int n, const double2* a, const double2* b, double2* partial, int nblk) {
    k_bidot_partial<<<nblk, 256, 0, st>>>(n, a, b, partial);
}

// y_t = D_t^{-1} x_t on every leaf; D_t^{-1} dense column-major of order
// nt (PEC) or 2 nt (PMCHWT, unknowns ordered J then M).
__global__ __launch_bounds__(256) void k_block_diag(int planes, int N, int max_leaf, const int* __restrict__ leaf_ptr,
                                                    const int* __restrict__ perm, const long long* __restrict__ off,
                                                    const double2* __restrict__ Dinv, const double2* __restrict__ x,
                                                    double2* __restrict__ y) {
    extern __shared__ double2 xs[];
    const int t = blockIdx.x;
    const int i0 = leaf_ptr[t], nt = leaf_ptr[t + 1] - i0;
    const int m = planes == 1 ? nt : 2 * nt;
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
        const int r = perm[i0 + (j % nt)];
        xs[j] = x[(j < nt) ? r : N + r];
    }
    __syncthreads();
    const double2* D = Dinv + off[t];
    for (int row = threadIdx.x; row < m; row += blockDim.x) {
        double2 a = make_double2(0, 0);
        for (int j = 0; j < m; ++j) cfma(a, D[(size_t)j * m + row], xs[j]);
        const int r = perm[i0 + (row % nt)];
        y[(row < nt) ? r : N + r] = a;
    }
}

void launch_block_diag(hipStream_t st, int planes, int nleaf, int N, const int* leaf_ptr, const int* perm,
                       const long long* inv_off, const double2* Dinv, int max_leaf, const double2* x, double2* y) {
    const int m = planes == 1 ? max_leaf : 2 * max_leaf;
    k_block_diag<<<nleaf, 256, (size_t)m * sizeof(double2), st>>>(planes, N, max_leaf, leaf_ptr, perm, inv_off, Dinv,
                                                                   x, y);
}

}  // namespace mlfma
