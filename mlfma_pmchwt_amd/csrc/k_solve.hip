// Krylov-solver vector kernels (FP64).  Reductions use a fixed block partition
// and a fixed combination order, so repeated runs give identical results.
#include <type_traits>

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

// ---------------------------------------------------------------- batched (R right-hand sides)
// Block vectors are RHS-fastest: element (u, r) at u * R + r, R <= 256.  A
// block of 256 threads covers rows of floor(256 / R) unknowns; thread t owns
// right-hand side t % R, and the per-r block sums are folded in a fixed order.
namespace {
__device__ __forceinline__ void fold_rows(double2 v, int R, double2* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    const int rows = blockDim.x / R;
    if ((int)threadIdx.x < R) {
        double2 s = red[threadIdx.x];
        for (int q = 1; q < rows; ++q) s = cadd(s, red[q * R + threadIdx.x]);
        red[threadIdx.x] = s;
    }
    __syncthreads();
}
}  // namespace

// partial[(j * nblk + b) * R + r] = sum over block b of conj(V_j[u][r]) w[u][r]
// (V in FP64 or, for FP32-mode solves, FP32: widened exactly before the FP64
// sums).  A block takes kDotJ basis vectors at once, so w is read once per
// kDotJ vectors; each vector's sum runs over the same rows in the same order
// whichever vectors share its block.
constexpr int kDotJ = 4;
template <typename TV>
__global__ __launch_bounds__(256) void k_multidot_b(int n, int R, int m, const TV* __restrict__ V, size_t ldv,
                                                    const double2* __restrict__ w, double2* __restrict__ partial) {
    __shared__ double2 red[256];
    const int j0 = blockIdx.y * kDotJ;
    const int rows = blockDim.x / R;
    const int r = threadIdx.x % R, q = threadIdx.x / R;
    const TV* v[kDotJ];
#pragma unroll
    for (int jj = 0; jj < kDotJ; ++jj) v[jj] = V + (size_t)min(j0 + jj, m - 1) * ldv;   // past m: read, not used
    double2 s[kDotJ];
#pragma unroll
    for (int jj = 0; jj < kDotJ; ++jj) s[jj] = make_double2(0, 0);
    if (q < rows)
        for (int u = blockIdx.x * rows + q; u < n; u += gridDim.x * rows) {
            const size_t e = (size_t)u * R + r;
            const double2 we = w[e];
#pragma unroll
            for (int jj = 0; jj < kDotJ; ++jj) s[jj] = cadd(s[jj], cmulc(widen(v[jj][e]), we));
        }
#pragma unroll
    for (int jj = 0; jj < kDotJ; ++jj) {
        if (j0 + jj >= m) break;   // uniform over the block
        fold_rows(s[jj], R, red);
        if ((int)threadIdx.x < R)
            partial[((size_t)(j0 + jj) * gridDim.x + blockIdx.x) * R + threadIdx.x] = red[threadIdx.x];
    }
}

void launch_multidot_b(hipStream_t st, int n, int R, int m, const double2* V, size_t ldv, const double2* w,
                       double2* partial, int nblk) {
    k_multidot_b<double2><<<dim3(nblk, (m + kDotJ - 1) / kDotJ), 256, 0, st>>>(n, R, m, V, ldv, w, partial);
}

void launch_multidot_b(hipStream_t st, int n, int R, int m, const float2* V, size_t ldv, const double2* w,
                       double2* partial, int nblk) {
    k_multidot_b<float2><<<dim3(nblk, (m + kDotJ - 1) / kDotJ), 256, 0, st>>>(n, R, m, V, ldv, w, partial);
}

// out[j * R + r] = sum_b partial[(j * nblk + b) * R + r], b ascending
__global__ void k_reduce_partials_b(int nblk, int R, const double2* __restrict__ partial, double2* __restrict__ out) {
    const int j = blockIdx.x;
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        double2 s = make_double2(0, 0);
        for (int b = 0; b < nblk; ++b) s = cadd(s, partial[((size_t)j * nblk + b) * R + r]);
        out[(size_t)j * R + r] = s;
    }
}

void launch_reduce_partials_b(hipStream_t st, int m, int R, int nblk, const double2* partial, double2* out) {
    k_reduce_partials_b<<<m, 256, 0, st>>>(nblk, R, partial, out);
}

// partial[b * R + r] = sum over block b of |x[u][r]|^2
__global__ __launch_bounds__(256) void k_norm2_partial_b(int n, int R, const double2* __restrict__ x,
                                                         double2* __restrict__ partial) {
    __shared__ double2 red[256];
    const int rows = blockDim.x / R;
    const int r = threadIdx.x % R, q = threadIdx.x / R;
    double2 s = make_double2(0, 0);
    if (q < rows)
        for (int u = blockIdx.x * rows + q; u < n; u += gridDim.x * rows) {
            const double2 a = x[(size_t)u * R + r];
            s.x += a.x * a.x + a.y * a.y;
        }
    fold_rows(s, R, red);
    if ((int)threadIdx.x < R) partial[(size_t)blockIdx.x * R + threadIdx.x] = red[threadIdx.x];
}

void launch_norm2_partial_b(hipStream_t st, int n, int R, const double2* x, double2* partial, int nblk) {
    k_norm2_partial_b<<<nblk, 256, 0, st>>>(n, R, x, partial);
}

// w[u][r] -= sum_j h[j * R + r] V_j[u][r]
template <typename TV>
__global__ __launch_bounds__(256) void k_sub_combination_b(size_t nR, int R, int m, const TV* __restrict__ V,
                                                           size_t ldv, const double2* __restrict__ h,
                                                           double2* __restrict__ w) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nR) return;
    const int r = (int)(e % R);
    double2 a = w[e];
    for (int j = 0; j < m; ++j) cfms(a, h[(size_t)j * R + r], widen(V[(size_t)j * ldv + e]));
    w[e] = a;
}

template <typename TV>
static void sub_combination_b(hipStream_t st, int n, int R, int m, const TV* V, size_t ldv, const double2* h,
                              double2* w) {
    const size_t nR = (size_t)n * R;
    k_sub_combination_b<TV><<<(unsigned)((nR + 255) / 256), 256, 0, st>>>(nR, R, m, V, ldv, h, w);
}

void launch_sub_combination_b(hipStream_t st, int n, int R, int m, const double2* V, size_t ldv, const double2* h,
                              double2* w) {
    sub_combination_b(st, n, R, m, V, ldv, h, w);
}

void launch_sub_combination_b(hipStream_t st, int n, int R, int m, const float2* V, size_t ldv, const double2* h,
                              double2* w) {
    sub_combination_b(st, n, R, m, V, ldv, h, w);
}

// x[u][r] += sum_j y[j * R + r] V_j[u][r]
template <typename TV>
__global__ __launch_bounds__(256) void k_add_combination_b(size_t nR, int R, int m, const TV* __restrict__ V,
                                                           size_t ldv, const double2* __restrict__ y,
                                                           double2* __restrict__ x) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nR) return;
    const int r = (int)(e % R);
    double2 a = x[e];
    for (int j = 0; j < m; ++j) cfma(a, y[(size_t)j * R + r], widen(V[(size_t)j * ldv + e]));
    x[e] = a;
}

template <typename TV>
static void add_combination_b(hipStream_t st, int n, int R, int m, const TV* V, size_t ldv, const double2* y,
                              double2* x) {
    const size_t nR = (size_t)n * R;
    k_add_combination_b<TV><<<(unsigned)((nR + 255) / 256), 256, 0, st>>>(nR, R, m, V, ldv, y, x);
}

void launch_add_combination_b(hipStream_t st, int n, int R, int m, const double2* V, size_t ldv, const double2* y,
                              double2* x) {
    add_combination_b(st, n, R, m, V, ldv, y, x);
}

void launch_add_combination_b(hipStream_t st, int n, int R, int m, const float2* V, size_t ldv, const double2* y,
                              double2* x) {
    add_combination_b(st, n, R, m, V, ldv, y, x);
}

// y[u][r] = a[r] * x[u][r]  (a real, one per right-hand side), y in FP64 or FP32
template <typename TY>
__global__ void k_scale_copy_b(size_t nR, int R, const double* __restrict__ a, const double2* __restrict__ x,
                               TY* __restrict__ y) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nR) narrow(cscale(x[e], a[e % R]), y[e]);
}

void launch_scale_copy_b(hipStream_t st, int n, int R, const double* a, const double2* x, double2* y) {
    const size_t nR = (size_t)n * R;
    k_scale_copy_b<double2><<<(unsigned)((nR + 255) / 256), 256, 0, st>>>(nR, R, a, x, y);
}

void launch_scale_copy_b(hipStream_t st, int n, int R, const double* a, const double2* x, float2* y) {
    const size_t nR = (size_t)n * R;
    k_scale_copy_b<float2><<<(unsigned)((nR + 255) / 256), 256, 0, st>>>(nR, R, a, x, y);
}

// FP32 -> FP64 (exact)
__global__ void k_widen(size_t n, const float2* __restrict__ in, double2* __restrict__ out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = widen(in[e]);
}

void launch_widen(hipStream_t st, size_t n, const float2* in, double2* out) {
    if (n == 0) return;
    k_widen<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, in, out);
}

// column-major (vector r at in + r * n) <-> RHS-fastest (element (u, r) at u * R + r)
__global__ void k_transpose(int rows, int cols, const double2* __restrict__ in, double2* __restrict__ out) {
    __shared__ double2 tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        const int r = r0 + k, c = c0 + threadIdx.x;
        if (r < rows && c < cols) tile[k][threadIdx.x] = in[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        const int c = c0 + k, r = r0 + threadIdx.x;
        if (r < rows && c < cols) out[(size_t)c * rows + r] = tile[threadIdx.x][k];
    }
}

// out (cols x rows, row-major) = transpose of in (rows x cols, row-major)
void launch_transpose(hipStream_t st, int rows, int cols, const double2* in, double2* out) {
    dim3 grid((cols + 31) / 32, (rows + 31) / 32);
    k_transpose<<<grid, dim3(32, 8), 0, st>>>(rows, cols, in, out);
}

// Column selection of RHS-fastest blocks (n rows): gather out[u][j] = in[u][idx[j]]
// (out has Rc columns, in has R), scatter out[u][idx[j]] = in[u][j].
__global__ void k_cols(size_t total, int R, int Rc, int scatter, const int* __restrict__ idx,
                       const double2* __restrict__ in, double2* __restrict__ out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const size_t u = e / Rc;
    const int j = (int)(e - u * Rc);
    if (scatter) out[u * R + idx[j]] = in[e];
    else out[e] = in[u * R + idx[j]];
}

void launch_gather_cols(hipStream_t st, int n, int R, int Rc, const int* idx, const double2* in, double2* out) {
    const size_t total = (size_t)n * Rc;
    if (total) k_cols<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(total, R, Rc, 0, idx, in, out);
}

void launch_scatter_cols(hipStream_t st, int n, int R, int Rc, const int* idx, const double2* in, double2* out) {
    const size_t total = (size_t)n * Rc;
    if (total) k_cols<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(total, R, Rc, 1, idx, in, out);
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

// Batched form on RHS-fastest blocks (element (u, r) at u * R + r): workgroup
// (leaf t, group g) applies leaf t's inverse to right-hand sides RB g .. RB g + RB - 1,
// so each inverse entry is read once per RB right-hand sides (RB as large as the
// leaf's block of RB right-hand sides fits 64 KiB of LDS).
template <int RB>
__global__ __launch_bounds__(256) void k_block_diag_b(int planes, int N, int R, const int* __restrict__ leaf_ptr,
                                                      const int* __restrict__ perm, const long long* __restrict__ off,
                                                      const double2* __restrict__ Dinv, const double2* __restrict__ x,
                                                      double2* __restrict__ y) {
    extern __shared__ double2 xs[];   // [m][RB]
    const int t = blockIdx.x, r0 = blockIdx.y * RB;
    const int nr = min(RB, R - r0);
    const int i0 = leaf_ptr[t], nt = leaf_ptr[t + 1] - i0;
    const int m = planes == 1 ? nt : 2 * nt;
    for (int e = threadIdx.x; e < m * RB; e += blockDim.x) {
        const int j = e / RB, q = e % RB;
        const int u = (j < nt) ? perm[i0 + j] : N + perm[i0 + j - nt];
        xs[e] = q < nr ? x[(size_t)u * R + r0 + q] : make_double2(0, 0);
    }
    __syncthreads();
    const double2* D = Dinv + off[t];
    for (int row = threadIdx.x; row < m; row += blockDim.x) {
        double2 a[RB];
        for (int q = 0; q < RB; ++q) a[q] = make_double2(0, 0);
        for (int j = 0; j < m; ++j) {
            const double2 d = D[(size_t)j * m + row];
            for (int q = 0; q < RB; ++q) cfma(a[q], d, xs[j * RB + q]);
        }
        const int u = (row < nt) ? perm[i0 + row] : N + perm[i0 + row - nt];
        for (int q = 0; q < nr; ++q) y[(size_t)u * R + r0 + q] = a[q];
    }
}

void launch_block_diag_b(hipStream_t st, int planes, int nleaf, int N, int R, const int* leaf_ptr, const int* perm,
                         const long long* inv_off, const double2* Dinv, int max_leaf, const double2* x, double2* y) {
    const size_t m = planes == 1 ? max_leaf : 2 * (size_t)max_leaf;
    const size_t lds = 65536;
    auto go = [&](auto rb) {
        constexpr int RB = decltype(rb)::value;
        dim3 grid(nleaf, (R + RB - 1) / RB);
        k_block_diag_b<RB><<<grid, 256, m * RB * sizeof(double2), st>>>(planes, N, R, leaf_ptr, perm, inv_off, Dinv,
                                                                         x, y);
    };
    if (m * 8 * sizeof(double2) <= lds) go(std::integral_constant<int, 8>{});
    else if (m * 4 * sizeof(double2) <= lds) go(std::integral_constant<int, 4>{});
    else if (m * 2 * sizeof(double2) <= lds) go(std::integral_constant<int, 2>{});
    else go(std::integral_constant<int, 1>{});
}

void launch_block_diag(hipStream_t st, int planes, int nleaf, int N, const int* leaf_ptr, const int* perm,
                       const long long* inv_off, const double2* Dinv, int max_leaf, const double2* x, double2* y) {
    const int m = planes == 1 ? max_leaf : 2 * max_leaf;
    k_block_diag<<<nleaf, 256, (size_t)m * sizeof(double2), st>>>(planes, N, max_leaf, leaf_ptr, perm, inv_off, Dinv,
                                                                   x, y);
}

}  // namespace mlfma
