// metrics.hip -- beam and ray-frame figures of merit as device reductions, so
// that a PO result or a traced frame can be scored where it lies in HBM:
//
//  * aperture sums for the spillover and taper efficiencies
//    (src/PyPO/Efficiencies.py:90-145, mask of _generateMask 11-41):
//      s0 = sum |f|^2 dA            s1 = sum_mask |f|^2 dA
//      s2 = sum_mask Re(f) dA       s3 = sum_mask Im(f) dA      s4 = sum_mask dA
//    spillover = |s1|^2 / (s0 s1) (the reference's |sum conj(f_ap) f dA|^2 /
//    (sum |f|^2 dA * sum |f_ap|^2 dA_ap), whose numerator and second factor
//    are both s1), taper = (s2^2 + s3^2) / s1 / s4;
//  * power sums for the cross-polar efficiency (Efficiencies.py:147-159):
//      p0 = sum |co|^2, p1 = sum |cross|^2;
//  * ray-frame statistics (Efficiencies.py:43-88): over the rays with
//    |d|^2 > 0.8, the mean position, the normalised mean direction and the
//    RMS distance to the mean position (a second pass, so no cancellation).
//
// Every reduction is two-stage and deterministic: a grid-stride pass writes
// one partial per workgroup (wave shuffles, then LDS across the four waves),
// and a single workgroup folds the partials in a fixed order.  FP64
// throughout; each pass reads its inputs once (HBM-bound).

#include "common.hpp"

namespace pypo {
namespace {

constexpr double kPi = 3.14159265358979323846;   // numpy.pi as a double
constexpr int kRedThreads = 256;
constexpr int kRedMaxBlocks = 1024;

// Sum of v over the workgroup, valid in thread 0.  All threads must call it.
template <int K>
__device__ inline void block_sum(double (&v)[K], double (*sh)[kRedThreads / 64]) {
#pragma unroll
    for (int q = 0; q < K; q++)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < K; q++) sh[q][wave] = v[q];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < K; q++) {
            double s = 0.0;
            for (int w = 0; w < kRedThreads / 64; w++) s += sh[q][w];
            v[q] = s;
        }
}

// Partials [nblocks][K] -> out[K], folded by one workgroup in a fixed order.
template <int K>
__global__ __launch_bounds__(kRedThreads) void k_fold(const double *__restrict__ partial, int nblocks,
                                                      double *__restrict__ out) {
    __shared__ double sh[K][kRedThreads / 64];
    double v[K];
#pragma unroll
    for (int q = 0; q < K; q++) v[q] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kRedThreads)
#pragma unroll
        for (int q = 0; q < K; q++) v[q] += partial[(size_t)b * K + q];
    block_sum<K>(v, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < K; q++) out[q] = v[q];
}

// The aperture test of _generateMask (Efficiencies.py:11-41).  Ellipse: the
// squared distance from the centre is compared with (a cos t)^2 + (b sin t)^2
// where t = atan2(y, x) + pi is the angle of the point about the ORIGIN (not
// the centre), as in the reference.  Rectangle: open intervals of x - cx and
// y - cy.  Only the outer boundary selects (the reference computes the inner
// one and discards it).  Contraction is off so the squares and sums round as
// numpy's do.
__device__ inline bool in_aperture(const AperMask &m, double x, double y) {
#pragma clang fp contract(off)
    if (m.shape == PYPO_APER_NONE) return true;
    const double u = x - m.center[0], v = y - m.center[1];
    if (m.shape == PYPO_APER_ELLIPSE) {
        const double t = atan2(y, x) + kPi;
        const double ac = m.outer[0] * cos(t), bs = m.outer[1] * sin(t);
        return u * u + v * v < ac * ac + bs * bs;
    }
    return (u > m.outer_x[0]) && (u < m.outer_x[1]) && (v > m.outer_y[0]) && (v < m.outer_y[1]);
}

__global__ __launch_bounds__(kRedThreads) void k_aperture_partial(const double *__restrict__ x,
                                                                  const double *__restrict__ y,
                                                                  const double *__restrict__ area,
                                                                  const double *__restrict__ fr,
                                                                  const double *__restrict__ fi, int n,
                                                                  AperMask m, double *__restrict__ partial) {
    __shared__ double sh[5][kRedThreads / 64];
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads) {
        const double a = area[i], re = fr[i], im = fi[i];
        const double pa = (re * re + im * im) * a;
        const double w = in_aperture(m, x[i], y[i]) ? 1.0 : 0.0;
        v[0] += pa;
        v[1] += w * pa;
        v[2] += w * re * a;
        v[3] += w * im * a;
        v[4] += w * a;
    }
    block_sum<5>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; q++) partial[(size_t)blockIdx.x * 5 + q] = v[q];
}

__global__ __launch_bounds__(kRedThreads) void k_power_partial(const double *__restrict__ ar,
                                                               const double *__restrict__ ai,
                                                               const double *__restrict__ br,
                                                               const double *__restrict__ bi, int n,
                                                               double *__restrict__ partial) {
    __shared__ double sh[2][kRedThreads / 64];
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads) {
        v[0] += ar[i] * ar[i] + ai[i] * ai[i];
        v[1] += br[i] * br[i] + bi[i] * bi[i];
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0) {
        partial[(size_t)blockIdx.x * 2] = v[0];
        partial[(size_t)blockIdx.x * 2 + 1] = v[1];
    }
}

// A ray counts when |d|^2 > 0.8 (Efficiencies.py:53); misses carry d = 0.
__device__ inline bool ray_good(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    return dx * dx + dy * dy + dz * dz > 0.8;
}

// Pass 1 over a frame: {count, sum x, y, z, dx, dy, dz} of the good rays.
__global__ __launch_bounds__(kRedThreads) void k_frame_partial(FrameView f, int n, double *__restrict__ partial) {
    __shared__ double sh[7][kRedThreads / 64];
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads) {
        const double dx = f.dx[i], dy = f.dy[i], dz = f.dz[i];
        if (ray_good(dx, dy, dz)) {
            v[0] += 1.0;
            v[1] += f.x[i];
            v[2] += f.y[i];
            v[3] += f.z[i];
            v[4] += dx;
            v[5] += dy;
            v[6] += dz;
        }
    }
    block_sum<7>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 7; q++) partial[(size_t)blockIdx.x * 7 + q] = v[q];
}

// Pass 2: sum of squared distances of the good rays from the mean position
// sums[1..3] / sums[0] of pass 1.
__global__ __launch_bounds__(kRedThreads) void k_frame_rms_partial(FrameView f, int n,
                                                                   const double *__restrict__ sums,
                                                                   double *__restrict__ partial) {
    __shared__ double sh[1][kRedThreads / 64];
    const double cnt = sums[0];
    const double cx = sums[1] / cnt, cy = sums[2] / cnt, cz = sums[3] / cnt;
    double v[1] = {0.0};
    for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads) {
        if (ray_good(f.dx[i], f.dy[i], f.dz[i])) {
            const double ex = f.x[i] - cx, ey = f.y[i] - cy, ez = f.z[i] - cz;
            v[0] += ex * ex + ey * ey + ez * ez;
        }
    }
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

int red_blocks(int n) {
    const int b = (n + kRedThreads - 1) / kRedThreads;
    return b < 1 ? 1 : (b > kRedMaxBlocks ? kRedMaxBlocks : b);
}

}  // namespace

// Partials (at most 7 per workgroup) followed by 8 folded results.
size_t metrics_workspace_bytes() { return ((size_t)kRedMaxBlocks * 7 + 8) * sizeof(double); }
size_t metrics_result_offset() { return (size_t)kRedMaxBlocks * 7; }

int launch_aperture_sums(const double *x, const double *y, const double *area, const double *fr,
                         const double *fi, int n, const AperMask &m, double *ws, double *out5, hipStream_t st) {
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(k_aperture_partial, dim3(nb), dim3(kRedThreads), 0, st, x, y, area, fr, fi, n, m, ws);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fold<5>, dim3(1), dim3(kRedThreads), 0, st, (const double *)ws, nb, out5);
    PYPO_CHECK_LAUNCH();
    return 0;
}

int launch_power_sums(const double *ar, const double *ai, const double *br, const double *bi, int n, double *ws,
                      double *out2, hipStream_t st) {
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(k_power_partial, dim3(nb), dim3(kRedThreads), 0, st, ar, ai, br, bi, n, ws);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fold<2>, dim3(1), dim3(kRedThreads), 0, st, (const double *)ws, nb, out2);
    PYPO_CHECK_LAUNCH();
    return 0;
}

int launch_frame_stats(const FrameView &f, int n, double *ws, double *out8, hipStream_t st) {
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(k_frame_partial, dim3(nb), dim3(kRedThreads), 0, st, f, n, ws);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fold<7>, dim3(1), dim3(kRedThreads), 0, st, (const double *)ws, nb, out8);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_frame_rms_partial, dim3(nb), dim3(kRedThreads), 0, st, f, n, (const double *)out8, ws);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fold<1>, dim3(1), dim3(kRedThreads), 0, st, (const double *)ws, nb, out8 + 7);
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace pypo
