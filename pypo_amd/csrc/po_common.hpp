// po_common.hpp -- device helpers shared by the PO sum kernels.
#pragma once

#include "common.hpp"

namespace pypo {

constexpr int TILE = 256;   // sources per LDS tile == threads per workgroup

// Packed record slot of grid point i (i = xu ncy + yv): serpentine order, odd
// grid rows reversed, when the grid is ncx x ncy, else i.  The sums do not
// depend on the order; consecutive records stay grid neighbours across row
// ends, which the matrix far field's follower phase relies on.
__device__ __forceinline__ int packed_slot(int i, int n, int ncx, int ncy) {
    if (ncy < 2 || (long long)ncx * ncy != n) return i;
    const int xu = i / ncy, yv = i - xu * ncy;
    return (xu & 1) ? xu * ncy + (ncy - 1 - yv) : i;
}

// LDS tile staging: a tile of TILE records of REC_U4 16-byte words, loaded
// one record per lane.  Past the end of the split the lane writes a null
// record instead (all weights zero, position 1e7 mm away from anything) so
// that the inner loop always runs TILE iterations and a null record can
// never produce a zero distance.
template <int REC_U4>
struct TileLoader {
    uint4 buf[REC_U4];
    __device__ __forceinline__ void load(const uint4 *__restrict__ g, int s, int s_end) {
        if (s < s_end) {
#pragma unroll
            for (int k = 0; k < REC_U4; k++) buf[k] = g[(size_t)s * REC_U4 + k];
        } else {
#pragma unroll
            for (int k = 0; k < REC_U4; k++) buf[k] = make_uint4(0u, 0u, 0u, 0u);
            // every record starts with double x, y, z
            const unsigned long long far = (unsigned long long)__double_as_longlong(1.0e7);
            const unsigned lo = (unsigned)(far & 0xffffffffull), hi = (unsigned)(far >> 32);
            buf[0] = make_uint4(lo, hi, lo, hi);
            buf[1].x = lo;
            buf[1].y = hi;
        }
    }
    __device__ __forceinline__ void store(uint4 *lds, int tid) {
#pragma unroll
        for (int k = 0; k < REC_U4; k++) lds[tid * REC_U4 + k] = buf[k];
    }
};

// Wave-level reductions (64 lanes) through cross-lane shuffles.
// Work item (target block tb, source split sp) of a PO sum workgroup, for a
// grid of target blocks x splits.  The dispatcher deals workgroups out in
// linear-id order, round-robin over the XCDs (XCD = id % kXcds), so taken
// as they come every XCD works on the same split at once and its own L2
// fetches that split's records, again after each eviction.  This bijection
// hands XCD x a contiguous run of the split-major item list instead: the
// XCDs work on different splits and the workgroups sharing an XCD's L2 at
// one time read the same records.  Only with at least kXcds splits: a
// split's items then run on one XCD or two, and a short last split only
// shortens one XCD's run.  With fewer, the contiguous runs would put whole
// splits of different lengths on different XCDs (a 3-tile and a 2-tile
// split on four XCDs each ran 1.5x slower), and the records of so few
// tiles stay in L2 anyway.  Which workgroup computes which item does not
// change any result.
constexpr int kXcds = 8;
__device__ __forceinline__ void xcd_item(int &tb, int &sp) {
    if ((int)gridDim.y < kXcds) {
        tb = blockIdx.x;
        sp = blockIdx.y;
        return;
    }
    const int nx = gridDim.x, n = nx * gridDim.y;
    const int lin = blockIdx.y * nx + blockIdx.x;
    const int x = lin % kXcds, q = n / kXcds, r = n % kXcds;
    const int item = x * q + min(x, r) + lin / kXcds;   // XCD x: items [x q + min(x, r), ...)
    tb = item % nx;
    sp = item / nx;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

// Workgroup (256 lanes = 4 waves) reductions; every lane gets the result.
// `scratch` holds at least 4 floats of LDS; callers separate uses with the
// barriers inside.
__device__ __forceinline__ float block_sum(float v, float *scratch) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) scratch[w] = v;
    __syncthreads();
    return scratch[0] + scratch[1] + scratch[2] + scratch[3];
}

__device__ __forceinline__ float block_min(float v, float *scratch) {
    v = wave_min(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) scratch[w] = v;
    __syncthreads();
    return fminf(fminf(scratch[0], scratch[1]), fminf(scratch[2], scratch[3]));
}

// Cone of directions: unit axis and cos of the half-angle (-1 if none).
struct Cone {
    float ax, ay, az, cos_half;
};

// Scales (vx, vy, vz) to unit length; a zero or non-finite vector is left
// as it is.  Cones bound n.r sign tests, so they are built from directions:
// grid normals are unit only under an orthonormal transform (a transf whose
// rotation part scales gives normals of other lengths), and a cone built or
// widened with longer vectors would come out too narrow.
__device__ __forceinline__ void to_unit(float &vx, float &vy, float &vz) {
    const float big = fmaxf(fabsf(vx), fmaxf(fabsf(vy), fabsf(vz)));
    if (big > 0.f && big < INFINITY) {
        vx /= big; vy /= big; vz /= big;   // largest component 1: no overflow or underflow below
        const float s = 1.f / sqrtf(vx * vx + vy * vy + vz * vz);
        vx *= s; vy *= s; vz *= s;
    }
}

// The cone holding the directions of the vectors (vx, vy, vz) of the lanes
// with valid == true, over the workgroup (each vector made unit first; zero
// and non-finite ones widen the cone to the whole sphere, so every per-pair
// test stays on).
__device__ __forceinline__ Cone block_cone(float vx, float vy, float vz, bool valid, float *scratch) {
    if (valid) to_unit(vx, vy, vz);
    float sx = block_sum(valid ? vx : 0.f, scratch);
    float sy = block_sum(valid ? vy : 0.f, scratch);
    float sz = block_sum(valid ? vz : 0.f, scratch);
    float cnt = block_sum(valid ? 1.f : 0.f, scratch);
    float len = sqrtf(sx * sx + sy * sy + sz * sz);
    Cone c;
    if (!(len > 1e-4f * cnt) || cnt == 0.f) {
        c.ax = 0.f; c.ay = 0.f; c.az = 1.f; c.cos_half = -1.f;
        return c;
    }
    c.ax = sx / len; c.ay = sy / len; c.az = sz / len;
    float d = valid ? (vx * c.ax + vy * c.ay + vz * c.az) : 1.f;
    c.cos_half = block_min(d, scratch);
    return c;
}

// True when some pair (n in the normal cone, r in the direction cone) can
// have n.r < 0, i.e. the per-pair shadow test (Propagation.h:1719) can
// change the result.  Conservative: a 1e-3 rad margin covers the FP32
// rounding of both cones.
__device__ __forceinline__ bool cones_may_shadow(float nax, float nay, float naz, float ncos,
                                                 const Cone &t) {
    if (ncos <= 0.f || t.cos_half <= 0.f) return true;
    float ca = fminf(1.f, fmaxf(-1.f, nax * t.ax + nay * t.ay + naz * t.az));
    float ang = acosf(ca) + acosf(fminf(1.f, ncos)) + acosf(fminf(1.f, t.cos_half));
    return !(ang < 1.5697963f);   // pi/2 - 1e-3
}

// Output bundles of the finish kernels (pointer lists in struct order).
struct OutC2 { double *p[12]; };
struct OutC4 { double *p[24]; };

// Kernel timing with HIP events on the launch stream.
struct Timer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool on;
    hipStream_t st;
    Timer(bool enable, hipStream_t s) : on(enable), st(s) {
        if (on) {
            if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) on = false;
            else (void)hipEventRecord(e0, st);
        }
    }
    void stop() {
        if (!on) return;
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) == hipSuccess) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            record_kernel_ms(ms);
        }
    }
    ~Timer() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// Byte offset of the TileMeta array behind n_src packed records.
__host__ __device__ __forceinline__ size_t meta_offset(int rec_bytes, int n_src) {
    return ((size_t)rec_bytes * n_src + 255) / 256 * 256;
}

__host__ __device__ __forceinline__ int n_tiles(int n_src) { return (n_src + TILE - 1) / TILE; }

// Launchers of the individual sum kernels (po_ff.hip, po_near.hip, po_scalar.hip).
int launch_ff_sum(int precision, const SumPlan &p, const void *packed, int n_src, const GridPtrs &ct,
                  int n_tgt, double rev_scale, double *partial, hipStream_t st);
int launch_ff_mfma(const SumPlan &p, const void *packed, int n_src, const GridPtrs &ct, int n_tgt,
                   double rev_scale, double *partial, hipStream_t st);
// Workgroups of 256 lanes one CU holds at once for the far-field sum kernel
// that launch_ff_sum / launch_ff_mfma pick for (precision, targets per lane).
int ff_wgs_per_cu(int precision, int tpl);
int ff_mfma_wgs_per_cu(int tpl);

// Occupancy of a 256-lane kernel in workgroups per CU (1 if the query fails).
inline int blocks_per_cu(const void *kernel) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, TILE, 0) != hipSuccess || n < 1) return 1;
    return n;
}
int launch_ff_finish(const SumPlan &p, const GridPtrs &ct, int n_tgt, double tdir, double pref,
                     const c2Bundle &out, const double *partial, hipStream_t st);
int launch_near_sum(int mode, int precision, const SumPlan &p, const void *packed, int n_src,
                    const GridPtrs &ct, int n_tgt, double k, double tdir, double *partial,
                    hipStream_t st);
int launch_near_finish(int mode, const SumPlan &p, const GridPtrs &ct, int n_tgt, double pref,
                       double zeta_inv, const OutC4 &out, const double *partial, hipStream_t st);
int launch_scalar_sum(int precision, const SumPlan &p, const void *packed, int n_src,
                      const GridPtrs &ct, int n_tgt, double rev_scale, double *partial,
                      hipStream_t st);
int launch_scalar_finish(const SumPlan &p, int n_tgt, double scale, double *re, double *im,
                         const double *partial, hipStream_t st);

}  // namespace pypo
