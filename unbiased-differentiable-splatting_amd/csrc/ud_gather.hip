// ud_gather.hip -- the pixelgather and nearestsamples op surfaces (gfx950):
//   pixelgather.scatter  per-slot copy of the sampled splats' attribute rows
//   pixelgather.gather   its adjoint: per-splat sums of per-slot rows
//   (splatting/extensions/pixelGather.cpp:10-23, pixelGatherCuda.cu:7-152)
//   nearestsamples.nearest  per splat, its nearest co-sampled splats
//   (FMMKdTreeSampling/nearestSamples.cpp:30-75, nearestSamplesCuda.cu:290-564)
//
// Slot tensors are x-major (bn, W, H, k) with pixel (i, j) at j + H (i + W b);
// a slot's row of C channels is contiguous, so the scatter output
// (bn, W, H, k, C) is written by consecutive threads at consecutive addresses
// and the kernels are bound by HBM streaming plus the row gathers.
#include "ud_device.h"
#include "../../include/udsplat.h"

#include <hip/hip_fp16.h>

namespace {

inline unsigned grid1(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

inline int launch_status()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
}

// One thread per output element (slot, channel): out = values[b, id, c] or 0.
template <class T>
__global__ __launch_bounds__(256) void k_pixel_scatter(long nslots, int WHk, int N, int C,
                                                       const int32_t* __restrict__ idx,
                                                       const T* __restrict__ values, T* __restrict__ out)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nslots * C) return;
    const long slot = e / C;
    const int c = (int)(e - slot * C);
    const int id = idx[slot];
    const long b = slot / WHk;
    out[e] = id >= 0 ? values[((long)b * N + id) * C + c] : T(0);
}

// One thread per (slot, channel) input element, added into the splat's row.
// Zero entries add nothing and are skipped (the sums are the same).
template <class T>
__global__ __launch_bounds__(256) void k_pixel_gather(long nslots, int WHk, int N, int C,
                                                      const int32_t* __restrict__ idx,
                                                      const T* __restrict__ values, T* __restrict__ out)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nslots * C) return;
    const long slot = e / C;
    const int c = (int)(e - slot * C);
    const int id = idx[slot];
    if (id < 0) return;
    const T v = values[e];
    if (v == T(0)) return;
    const long b = slot / WHk;
    atomicAdd(out + ((long)b * N + id) * C + c, v);
}

template <class T>
int scatter_t(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const T* values, T* out, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || N < 0 || C <= 0) return UDE_ERR_ARG;
    const long nslots = (long)bn * W * H * k;
    k_pixel_scatter<T><<<grid1(nslots * C, 256), 256, 0, st>>>(nslots, W * H * k, N, C, idx, values, out);
    return launch_status();
}

template <class T>
int gather_t(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const T* values, T* out, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || N < 0 || C <= 0) return UDE_ERR_ARG;
    if (N == 0) return UDE_OK;
    const long nslots = (long)bn * W * H * k;
    k_pixel_gather<T><<<grid1(nslots * C, 256), 256, 0, st>>>(nslots, W * H * k, N, C, idx, values, out);
    return launch_status();
}

// ---------------------------------------------------------------------------
// nearestsamples.nearest.  A pixel's surface splat is the first slot of largest
// weight (nearestSamplesCuda.cu:306-321); every other splat sampled at the pixel
// is a candidate neighbour of the surface splat at squared distance d2 (float32,
// ((dx dx + dy dy) + dz dz)).  Each splat keeps the K smallest (d2, index) words
// (distance bits high, index low: ties by index) in a sorted row of K 64-bit
// slots.  A candidate enters by a cascade of atomicMin over the row: each slot
// keeps the smaller word and the larger one moves on to the next slot, so every
// word that leaves slot j reaches slot j + 1; a slot returning the candidate's
// own word means another pixel inserted the same neighbour, and the copy is
// dropped.  The final rows are the K nearest distinct candidates in order,
// whatever order the pixels ran in (the reference merges under a spin lock per
// splat, nearestSamplesCuda.cu:370-404).
// ---------------------------------------------------------------------------
constexpr unsigned long long kNearestEmpty = ~0ull;

__global__ void k_fill_u64(unsigned long long* __restrict__ p, long n, unsigned long long v)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

template <class T>
__global__ __launch_bounds__(256) void k_nearest_insert(long npix, int k, int K, int N,
                                                        const int32_t* __restrict__ idx, const T* __restrict__ wts,
                                                        const T* __restrict__ pts, unsigned long long* rows)
{
    const long pix = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    const long base = pix * k;
    int s = idx[base];
    T sw = wts[base];
    for (int q = 1; q < k; q++) {
        const T w = wts[base + q];
        if (w > sw) { sw = w; s = idx[base + q]; }
    }
    if (s < 0 || s >= N) return;
    const T sx = pts[3L * s], sy = pts[3L * s + 1], sz = pts[3L * s + 2];
    unsigned long long* row = rows + (long)s * K;
    for (int q = 0; q < k; q++) {
        const int c = idx[base + q];
        if (c < 0 || c == s || c >= N) continue;
        const float dx = (float)(sx - pts[3L * c]), dy = (float)(sy - pts[3L * c + 1]), dz = (float)(sz - pts[3L * c + 2]);
        float d2 = dx * dx;
        d2 = d2 + dy * dy;
        d2 = d2 + dz * dz;
        unsigned long long key = ((unsigned long long)__float_as_uint(d2) << 32) | (uint32_t)c;
        if (key >= row[K - 1]) continue;            // cannot enter (a stale read only costs work)
        for (int j = 0; j < K; j++) {
            const unsigned long long old = atomicMin(row + j, key);
            if (old == key || old == kNearestEmpty) break;   // a copy already there / an empty slot filled
            if (old > key) key = old;                        // carry the displaced word on
        }
    }
}

__global__ void k_nearest_finish(long n, const unsigned long long* __restrict__ rows, int32_t* __restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const unsigned long long v = rows[t];
    out[t] = v == kNearestEmpty ? -1 : (int32_t)(uint32_t)(v & 0xffffffffull);
}

template <class T>
int nearest_t(int bn, int W, int H, int k, int N, int K, const int32_t* idx, const T* wts, const T* pts,
              int32_t* out, void* ws, size_t wsBytes, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || N < 0 || K <= 0) return UDE_ERR_ARG;
    if (N == 0) return UDE_OK;
    if (wsBytes < (size_t)N * K * 8) return UDE_ERR_WORKSPACE;
    unsigned long long* rows = (unsigned long long*)ws;
    const long n = (long)N * K, npix = (long)bn * W * H;
    k_fill_u64<<<grid1(n, 256), 256, 0, st>>>(rows, n, kNearestEmpty);
    k_nearest_insert<T><<<grid1(npix, 256), 256, 0, st>>>(npix, k, K, N, idx, wts, pts, rows);
    k_nearest_finish<<<grid1(n, 256), 256, 0, st>>>(n, rows, out);
    return launch_status();
}

} // namespace

extern "C" {

size_t ude_nearest_workspace_bytes(int N, int K) { return (size_t)(N > 0 ? N : 1) * (K > 0 ? K : 1) * 8 + 256; }

int ude_nearest_samples(int bn, int W, int H, int k, int N, int K, const int32_t* idx, const float* wts,
                        const float* pts, int32_t* out, void* ws, size_t wsBytes, hipStream_t st)
{
    return nearest_t<float>(bn, W, H, k, N, K, idx, wts, pts, out, ws, wsBytes, st);
}

int ude_nearest_samples_f64(int bn, int W, int H, int k, int N, int K, const int32_t* idx, const double* wts,
                            const double* pts, int32_t* out, void* ws, size_t wsBytes, hipStream_t st)
{
    return nearest_t<double>(bn, W, H, k, N, K, idx, wts, pts, out, ws, wsBytes, st);
}


int ude_pixel_scatter(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const float* values,
                      float* out, hipStream_t st)
{
    return scatter_t<float>(bn, W, H, k, N, C, idx, values, out, st);
}

int ude_pixel_scatter_f64(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const double* values,
                          double* out, hipStream_t st)
{
    return scatter_t<double>(bn, W, H, k, N, C, idx, values, out, st);
}

int ude_pixel_scatter_f16(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const void* values,
                          void* out, hipStream_t st)
{
    return scatter_t<__half>(bn, W, H, k, N, C, idx, (const __half*)values, (__half*)out, st);
}

int ude_pixel_gather(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const float* values, float* out,
                     hipStream_t st)
{
    return gather_t<float>(bn, W, H, k, N, C, idx, values, out, st);
}

int ude_pixel_gather_f64(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const double* values,
                         double* out, hipStream_t st)
{
    return gather_t<double>(bn, W, H, k, N, C, idx, values, out, st);
}

} // extern "C"
