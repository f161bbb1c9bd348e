// GetDt on caller-owned primitive arrays (src/solver_GetDt/GlobalDt_block.hpp:6-114):
// the CFL maxima max(|u_d| + sqrt(1.4 p / rho)) over all padded cells of the
// active directions, then dt = CFLnumber / sum_d max_d / Delta_d.  Two
// kernels: a grid-stride pass that leaves one maximum triple per block in a
// workspace, and a one-block fold that writes uvw_c_max and dt.
#include <hip/hip_runtime.h>
#include <climits>
#include <cstdio>
#include <mutex>
#include "../../include/xf_api.h"

namespace {
constexpr int DT_BLOCKS = 4096; // blocks of the reduction pass, at most
constexpr int DT_THREADS = 256;
} // namespace

// Workspace of xf_get_dt_flow_noalloc: one maximum triple per reduction
// block, DT_BLOCKS * 3 doubles (96 KiB), held by every GPU the library runs
// on until the process ends.  Host code never uses this name as an address:
// the device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[DT_BLOCKS * 3];

namespace {

constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// this device's address of `scratch`, looked up on its first use
int scratch_ptr(double **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return XF_EHIP;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            fprintf(stderr, "xfluids_amd: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return XF_EHIP;
        }
        g_scratch[dev] = (double *)ptr;
    }
    *out = g_scratch[dev];
    return XF_OK;
}

// maxima of non-negative finite values, as the reference's sycl maximum
// reduction: a NaN never replaces a value (v > m is false)
__device__ __forceinline__ double dmax_keep(double m, double v) { return v > m ? v : m; }

__global__ void __launch_bounds__(DT_THREADS) k_dt_flow_part(long long N, int dx, int dy, int dz,
                                                             const double *__restrict__ rho,
                                                             const double *__restrict__ p,
                                                             const double *__restrict__ u,
                                                             const double *__restrict__ v,
                                                             const double *__restrict__ w, double *part)
{
    double m[3] = {0.0, 0.0, 0.0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < N; id += stride) {
        const double c = sqrt(1.4 * p[id] / rho[id]);
        if (dx) m[0] = dmax_keep(m[0], fabs(u[id]) + c);
        if (dy) m[1] = dmax_keep(m[1], fabs(v[id]) + c);
        if (dz) m[2] = dmax_keep(m[2], fabs(w[id]) + c);
    }
    __shared__ double s[3][DT_THREADS];
    for (int q = 0; q < 3; q++) s[q][threadIdx.x] = m[q];
    __syncthreads();
    for (int h = DT_THREADS / 2; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < 3; q++) s[q][threadIdx.x] = dmax_keep(s[q][threadIdx.x], s[q][threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(1024) k_dt_flow_fold(const double *__restrict__ part, int nb, int dx, int dy,
                                                       int dz, double _dx, double _dy, double _dz, double cfl,
                                                       double *uvw_c_max, double *dt)
{
    __shared__ double s[3][1024];
    double m[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += 1024)
        for (int q = 0; q < 3; q++) m[q] = dmax_keep(m[q], part[3 * b + q]);
    for (int q = 0; q < 3; q++) s[q][threadIdx.x] = m[q];
    __syncthreads();
    for (int h = 512; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < 3; q++) s[q][threadIdx.x] = dmax_keep(s[q][threadIdx.x], s[q][threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double mx = dx ? s[0][0] : 0.0, my = dy ? s[1][0] : 0.0, mz = dz ? s[2][0] : 0.0;
        if (uvw_c_max) uvw_c_max[0] = mx, uvw_c_max[1] = my, uvw_c_max[2] = mz;
        const double dtref = mx * _dx + my * _dy + mz * _dz;
        *dt = cfl / dtref;
    }
}

int dt_flow_blocks(long long N)
{
    const long long nb = (N + DT_THREADS - 1) / DT_THREADS;
    return (int)(nb < DT_BLOCKS ? nb : DT_BLOCKS);
}

int dt_flow_check(const xf_case_desc *d, const double *rho, const double *p, const double *u, const double *v,
                  const double *w, const double *dt_out, long long &N)
{
    if (!d || !rho || !p || !dt_out) return XF_EINVAL;
    if ((d->DimX && !u) || (d->DimY && !v) || (d->DimZ && !w)) return XF_EINVAL;
    if (d->Xmax <= 0 || d->Ymax <= 0 || d->Zmax <= 0) return XF_EINVAL;
    N = (long long)d->Xmax * d->Ymax * d->Zmax;
    return XF_OK;
}

// both kernels on the stream; part holds dt_flow_blocks(N) triples
int dt_flow_enqueue(const xf_case_desc *d, long long N, const double *rho, const double *p, const double *u,
                    const double *v, const double *w, double *part, double *uvw_c_max, double *dt_out,
                    hipStream_t st)
{
    const int nb = dt_flow_blocks(N);
    hipLaunchKernelGGL(k_dt_flow_part, dim3(nb), dim3(DT_THREADS), 0, st, N, d->DimX, d->DimY, d->DimZ, rho, p, u, v,
                       w, part);
    hipLaunchKernelGGL(k_dt_flow_fold, dim3(1), dim3(1024), 0, st, part, nb, d->DimX, d->DimY, d->DimZ, d->_dx,
                       d->_dy, d->_dz, d->CFLnumber, uvw_c_max, dt_out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "xfluids_amd: dt launch error %s\n", hipGetErrorString(e));
        return XF_EHIP;
    }
    return XF_OK;
}

} // namespace

extern "C" {

int xf_get_dt_flow(const xf_case_desc *d, const double *rho, const double *p, const double *u, const double *v,
                   const double *w, double *uvw_c_max, double *dt_out, void *stream)
{
    long long N = 0;
    if (int e = dt_flow_check(d, rho, p, u, v, w, dt_out, N)) return e;
    const hipStream_t st = (hipStream_t)stream;
    double *part = nullptr;
    if (hipMalloc(&part, (size_t)dt_flow_blocks(N) * 3 * sizeof(double)) != hipSuccess) return XF_EHIP;
    int rc = dt_flow_enqueue(d, N, rho, p, u, v, w, part, uvw_c_max, dt_out, st);
    // the workspace is freed once the fold has read it
    if (hipStreamSynchronize(st) != hipSuccess) rc = XF_EHIP;
    (void)hipFree(part);
    return rc;
}

int xf_get_dt_flow_noalloc(const xf_case_desc *d, const double *rho, const double *p, const double *u,
                           const double *v, const double *w, double *uvw_c_max, double *dt_out, void *stream)
{
    long long N = 0;
    if (int e = dt_flow_check(d, rho, p, u, v, w, dt_out, N)) return e;
    if (N > (long long)INT_MAX) {
        fprintf(stderr, "xfluids_amd: xf_get_dt_flow_noalloc: %lld cells exceed the limit of %d\n", N, INT_MAX);
        return XF_EINVAL;
    }
    double *part = nullptr;
    if (int e = scratch_ptr(&part)) return e;
    // every block writes its triple before the fold reads it: nothing to clear
    return dt_flow_enqueue(d, N, rho, p, u, v, w, part, uvw_c_max, dt_out, (hipStream_t)stream);
}

} // extern "C"
