// hu.hip -- HU -> relative linear stopping power / mass density on gfx950:
// the machine's HU table interpolated per voxel like np.interp (FP64), then
// rounded to float32, as the plan layer's conversions do (plan_impt.py:112-122,
// plan_imrt.py:114-124).
#include "dh_runtime.h"

namespace {

constexpr int kBlock = 256;

// np.interp for one value (NumPy's arr_interp): NaN stays NaN; below the
// first node the first value, above the last node the last value; at a node
// (the last node at or below x, found by binary search) its value; between
// nodes slope * (x - x_j) + y_j with slope = (y_j+1 - y_j) / (x_j+1 - x_j),
// multiply and add rounded separately, and from the right node when that is
// NaN.  A one-node table gives its value everywhere.
__device__ static inline double interp(const dh_hu_table &t, double x)
{
    const int m = t.m;
    if (m == 1) return t.y[0];
    if (isnan(x)) return x;
    if (x < t.x[0]) return t.y[0];
    if (x > t.x[m - 1]) return t.y[m - 1];
    int lo = 0, hi = m - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t.x[mid] <= x) lo = mid; else hi = mid - 1;
    }
    const int j = lo;
    if (j == m - 1 || t.x[j] == x) return t.y[j];
    const double slope = (t.y[j + 1] - t.y[j]) / (t.x[j + 1] - t.x[j]);
    double r = __dadd_rn(__dmul_rn(slope, x - t.x[j]), t.y[j]);
    if (isnan(r)) {
        r = __dadd_rn(__dmul_rn(slope, x - t.x[j + 1]), t.y[j + 1]);
        if (isnan(r) && t.y[j] == t.y[j + 1]) r = t.y[j];
    }
    return r;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_hu_table(dh_hu_table t, const T *__restrict__ hu, int64_t n, float *__restrict__ out)
{
    for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
        out[v] = (float)interp(t, (double)hu[v]);
}

}  // namespace

void dh_launch_hu_table(const dh_hu_table &t, const void *hu, bool hu_double, int64_t n, float *out, hipStream_t s)
{
    if (n <= 0) return;
    const unsigned grid = (unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 8192);
    dh_timed tm("hu_table", s);
    if (hu_double)
        hipLaunchKernelGGL(k_hu_table<double>, dim3(grid), dim3(kBlock), 0, s, t, (const double *)hu, n, out);
    else
        hipLaunchKernelGGL(k_hu_table<float>, dim3(grid), dim3(kBlock), 0, s, t, (const float *)hu, n, out);
    DH_HIP(hipGetLastError());
}
