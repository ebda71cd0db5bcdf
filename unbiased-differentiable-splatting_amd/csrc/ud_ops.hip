// ud_ops.hip -- per-slot operators of the splatting op surface (gfx950):
//   splatting.forward / backward   (EWA weights and their adjoint)
//   sampling.covariance
//   filtering.forward / backward   (z-sort + alpha compositing)
//   transform.{affine,linear}{Forward,Backward}
//   interpolation.forward / backward
//   shading.{forward,backward,forwardSh,backwardSh}
//
// Layouts follow the reference op surface: slot tensors are x-major
// (bn, W, H, k) with pixel (i = x, j = y) at j + H (i + W b); point tensors
// are (bn, N, 3) fp32, std-devs (1 | bn, N, 2).
#include "ud_device.h"
#include "../../include/udsplat.h"

#include <hip/hip_fp16.h>

#include <mutex>

using namespace ud;

// Scratch of the allocation-free splatting.backward (ude_splat_backward_static):
// per (view, splat) the 32 B screen-Gaussian record and the 32 B moment record,
// UDE_STATIC_SCRATCH_BYTES in all, held by every device the library runs on until
// the process ends.  Host code never passes this name to a kernel (on the host it
// is a stand-in, not a device address): it uses the per-device address below.
static __device__ __align__(256) unsigned char scratch[UDE_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices];           // device address of `scratch` on each device, once known
std::mutex g_scratch_mu;

// The device address of `scratch` on the current device, taken once per device
// (hipGetSymbolAddress with that device current); nullptr on failure.
void* scratch_on_current_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch[dev] = ptr;
    }
    return g_scratch[dev];
}
} // namespace

namespace {

// ---------------------------------------------------------------------------
// EWA geometry in FP64.  A = inverse Jacobian of the local plane frame
// (ewaSplatting.cuh:8-44), evaluated for T = double or T = D6 (forward-mode
// duals over point (3) and normal (3)).
// ---------------------------------------------------------------------------
struct D6 {
    double v, d[6];
};
__device__ inline D6 mk(double v) { D6 r; r.v = v; for (int i = 0; i < 6; i++) r.d[i] = 0.0; return r; }
__device__ inline D6 operator+(D6 a, D6 b) { D6 r; r.v = a.v + b.v; for (int i = 0; i < 6; i++) r.d[i] = a.d[i] + b.d[i]; return r; }
__device__ inline D6 operator-(D6 a, D6 b) { D6 r; r.v = a.v - b.v; for (int i = 0; i < 6; i++) r.d[i] = a.d[i] - b.d[i]; return r; }
__device__ inline D6 operator-(D6 a) { D6 r; r.v = -a.v; for (int i = 0; i < 6; i++) r.d[i] = -a.d[i]; return r; }
__device__ inline D6 operator*(D6 a, D6 b) { D6 r; r.v = a.v * b.v; for (int i = 0; i < 6; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
__device__ inline D6 operator*(D6 a, double s) { D6 r; r.v = a.v * s; for (int i = 0; i < 6; i++) r.d[i] = a.d[i] * s; return r; }
__device__ inline D6 operator+(D6 a, double s) { a.v += s; return a; }
__device__ inline D6 operator/(D6 a, D6 b)
{
    D6 r; r.v = a.v / b.v;
    double ib = 1.0 / b.v;
    for (int i = 0; i < 6; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
__device__ inline D6 dsqrt(D6 a) { D6 r; r.v = sqrt(a.v); double h = 0.5 / r.v; for (int i = 0; i < 6; i++) r.d[i] = a.d[i] * h; return r; }
__device__ inline double dsqrt(double a) { return sqrt(a); }
__device__ inline double val(double a) { return a; }
__device__ inline double val(const D6& a) { return a.v; }
__device__ inline D6 operator/(D6 a, double s) { return a * (1.0 / s); }
template <class T> __device__ inline T mkT(double v);
template <> __device__ inline double mkT<double>(double v) { return v; }
template <> __device__ inline D6 mkT<D6>(double v) { return mk(v); }

template <class T>
__device__ inline T dot3T(const T* a, const T* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// screen point s and A = [[A00, A01], [0, A11]] for camera-space point p, normal n
template <class T>
__device__ inline void ewa_geom(const T* p, const T* n, double W, double H, double f,
                                T& sx, T& sy, T& A00, T& A01, T& A11)
{
    sx = (-(p[0] * f)) / p[2] + W * 0.5;
    sy = (p[1] * f) / p[2] + H * 0.5;
    const double i2f = 1.0 / (2.0 * f);
    T x0[3], x1[3];
    // backproject(s + e_x), backproject(s + e_y): (W - 2 y_x) / 2f, -(H - 2 y_y) / 2f, 1
    T y0x = sx + 1.0, y1y = sy + 1.0;
    x0[0] = (mkT<T>(W) - y0x * 2.0) * i2f;
    x0[1] = -((mkT<T>(H) - sy * 2.0) * i2f);
    x0[2] = mkT<T>(1.0);
    x1[0] = (mkT<T>(W) - sx * 2.0) * i2f;
    x1[1] = -((mkT<T>(H) - y1y * 2.0) * i2f);
    x1[2] = mkT<T>(1.0);
    T a = dot3T(p, n);
    T xt0 = a / dot3T(x0, n), xt1 = a / dot3T(x1, n);
    T t0[3], t1[3];
    for (int c = 0; c < 3; c++) { t0[c] = x0[c] * xt0 - p[c]; t1[c] = x1[c] * xt1 - p[c]; }
    T nt0 = dsqrt(dot3T(t0, t0));
    T u0[3], u1[3];
    for (int c = 0; c < 3; c++) u0[c] = t0[c] / nt0;
    u1[0] = u0[1] * n[2] - u0[2] * n[1];
    u1[1] = u0[2] * n[0] - u0[0] * n[2];
    u1[2] = u0[0] * n[1] - u0[1] * n[0];
    A00 = nt0;
    A01 = dot3T(t1, u0);
    A11 = dot3T(t1, u1);
}


template <class T>
struct SplatRecT {
    T sxi, syi, sxf, syf;       // screen point split into integer part and fraction
    T t00, t01, t11;            // upper-triangular T with T^T T = V^-1 (computed in FP64)
    T norm;                     // 1 / (2 pi sqrt(det V) |det A|)
};
using SplatRec = SplatRecT<float>;
// With u = t00 dx + t01 dy, v = t11 dy the exponent is u^2 + v^2: rounding T to
// FP32 perturbs it by O(eps sqrt(e lambda_max) |d|), while rounding V^-1 itself
// would cost O(eps cond(V) e) for nearly edge-on splats.

// Per (view, splat) FP64 precompute of the screen Gaussian (ewaSplatting.cuh:24-67),
// rounded to the op's real type T (float, or double for the FP64 op surface).
template <class T>
__global__ void k_splat_prepare(int bn, int N, int W, int H, double f, double smoothing,
                                const T* __restrict__ pts, const T* __restrict__ nrm,
                                const T* __restrict__ sd, int sdBatch, SplatRecT<T>* __restrict__ rec)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)bn * N) return;
    long b = t / N, id = t % N;
    double p[3] = { pts[3 * t], pts[3 * t + 1], pts[3 * t + 2] };
    double n[3] = { nrm[3 * t], nrm[3 * t + 1], nrm[3 * t + 2] };
    long so = 2 * ((sdBatch > 1 ? b * N : 0) + id);
    double s0 = sd[so], s1 = sd[so + 1];
    double sx, sy, A00, A01, A11;
    ewa_geom<double>(p, n, W, H, f, sx, sy, A00, A01, A11);
    double detA = A00 * A11;
    double J00 = A11 / detA, J01 = -A01 / detA, J11 = A00 / detA;
    double V00 = J00 * J00 * s0 * s0 + J01 * J01 * s1 * s1 + smoothing;
    double V01 = J01 * J11 * s1 * s1;
    double V11 = J11 * J11 * s1 * s1 + smoothing;
    double detV = V00 * V11 - V01 * V01;
    SplatRecT<T> r;
    double fx = floor(sx), fy = floor(sy);
    r.sxi = (T)fx; r.syi = (T)fy;
    r.sxf = (T)(sx - fx); r.syf = (T)(sy - fy);
    r.t00 = (T)sqrt(V11 / detV);
    r.t01 = (T)(-V01 / sqrt(V11 * detV));
    r.t11 = (T)(1.0 / sqrt(V11));
    r.norm = (T)(1.0 / (2.0 * UD_PI * sqrt(detV) * fabs(detA)));
    rec[t] = r;
}

__device__ __forceinline__ float ewa_exp(float x) { return __expf(x); }
__device__ __forceinline__ double ewa_exp(double x) { return exp(x); }

template <class T>
__device__ __forceinline__ T ewa_eval(const SplatRecT<T>& r, T px, T py, T& u, T& v)
{
    T dx = (px - r.sxi) - r.sxf;
    T dy = (py - r.syi) - r.syf;
    u = r.t00 * dx + r.t01 * dy;
    v = r.t11 * dy;
    T e = u * u + v * v;
    return r.norm * ewa_exp(T(-0.5) * e);
}

// splatting.forward: one thread per slot (splattingCuda.cu:18-66)
template <class T>
__global__ void k_splat_forward(long nslots, int N, int W, int H, int k,
                                const int32_t* __restrict__ idx, const SplatRecT<T>* __restrict__ rec,
                                T* __restrict__ out)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nslots) return;
    int id = idx[t];
    if (id < 0) { out[t] = T(0); return; }
    long pix = t / k;
    long b = pix / ((long)W * H);
    int r = (int)(pix - b * (long)W * H);
    int i = r / H, j = r - (r / H) * H;
    SplatRecT<T> s = rec[b * N + id];
    T dx, dy;
    T w = ewa_eval(s, (T)i, (T)j, dx, dy);
    out[t] = fmin(w, T(1));
}

// floats per (view, splat) record of the backward passes: splatting.backward's
// six weight moments, and the fused estimator's moments + attribute gradient
constexpr int kMomRec = 8;
constexpr int kHtRec = 16;

// splatting.backward, slot pass: accumulate per-splat moments in the whitened
// coordinates z = T d = (u, v):
//   m0 = sum g w, (m1u, m1v) = sum g w z, (m2uu, m2uv, m2vv) = sum g w z z^T
// for slots whose unclamped weight is <= 1 (splattingCuda.cu:173).
template <class T>
__global__ __launch_bounds__(256) void k_splat_moments(long nslots, int N, int W, int H, int k,
                                                       const int32_t* __restrict__ idx,
                                                       const T* __restrict__ gin,
                                                       const SplatRecT<T>* __restrict__ rec, T* __restrict__ mom)
{
    // each lane computes its slot's six sums; the wavefront then adds them with six
    // consecutive lanes per record, so one atomic instruction covers ten records'
    // lines instead of 64 (L2 serves the dwords of one line in one instruction
    // together)
    __shared__ T sv[4][64][6];
    __shared__ long sr[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long si = -1;
    T v[6] = { T(0), T(0), T(0), T(0), T(0), T(0) };
    const int id = t < nslots ? idx[t] : -1;
    if (id >= 0) {
        long pix = t / k;
        long b = pix / ((long)W * H);
        int r = (int)(pix - b * (long)W * H);
        int i = r / H, j = r - (r / H) * H;
        const long s_ = b * N + id;
        SplatRecT<T> s = rec[s_];
        T dx, dy;
        T w = ewa_eval(s, (T)i, (T)j, dx, dy);
        if (w <= T(1)) {
            T gw = gin[t] * w;
            if (gw != T(0)) {
                si = s_;
                v[0] = gw;
                v[1] = gw * dx;
                v[2] = gw * dy;
                v[3] = (gw * dx) * dx;
                v[4] = (gw * dx) * dy;
                v[5] = (gw * dy) * dy;
            }
        }
    }
    for (int c = 0; c < 6; c++) sv[wv][lane][c] = v[c];
    sr[wv][lane] = si;
    __syncthreads();
    const int sub = lane / 6, c = lane - 6 * sub;
    for (int q = 0; q < 64; q += 10) {
        const int slot = q + sub;
        if (lane < 60 && slot < 64) {
            const long s_ = sr[wv][slot];
            const T x = sv[wv][slot][c];
            if (s_ >= 0 && x != T(0)) atomicAdd(mom + kMomRec * s_ + c, x);
        }
    }
}

// splatting.backward, splat pass: chain the moments through V(A, sigma),
// A(p, n, s(p)) and s(p) in FP64 with forward-mode derivatives of A.  A record
// holds the six moments at [0, 6); records of kHtRec floats (the fused
// estimator's) also hold the attribute gradient of the (view, splat) at
// [8, 11), written to gAttr (ab == bn) or added into it (ab == 1 < bn).
// The reference's EWASplatting.backward post-processing of a gradient row
// (splatting/extensions/__init__.py:219-225 with utils.py:147-150's
// conditionalNormalization): g / max(|g| + 1e-4, 1), then NaN -> 0 (infinities
// kept), in the element type.
template <class T, int C>
__device__ __forceinline__ void cond_normalize(T* g)
{
    T s = g[0] * g[0];
    for (int c = 1; c < C; c++) s = s + g[c] * g[c];
    T den = sqrt(s) + T(1e-4);
    den = den < T(1) ? T(1) : den;
    for (int c = 0; c < C; c++) {
        const T v = g[c] / den;
        g[c] = v != v ? T(0) : v;
    }
}

template <class T, bool POST = false>
__global__ __launch_bounds__(128) void k_splat_finalize(int bn, int N, int W, int H, double f, double smoothing,
                                                        const T* __restrict__ pts, const T* __restrict__ nrm,
                                                        const T* __restrict__ sd, int sdBatch,
                                                        const SplatRecT<T>* __restrict__ rec,
                                                        const T* __restrict__ mom, int recLen, int ab,
                                                        T* __restrict__ gAttr, T* __restrict__ gPts,
                                                        T* __restrict__ gNrm, T* __restrict__ gSd)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)bn * N) return;
    long b = t / N, id = t % N;
    long so = 2 * ((sdBatch > 1 ? b * N : 0) + id);
    const T* m = mom + (long)recLen * t;
    if (gAttr) {
        T* ga = gAttr + 3 * ((ab > 1 ? b * N : 0) + id);
        for (int c = 0; c < 3; c++) {
            if (ab == bn) ga[c] = m[8 + c];
            else if (m[8 + c] != T(0)) atomicAdd(ga + c, m[8 + c]);
        }
    }
    double m0 = m[0], m1x = m[1], m1y = m[2], m2xx = m[3], m2xy = m[4], m2yy = m[5];
    if (m0 == 0.0 && m1x == 0.0 && m1y == 0.0 && m2xx == 0.0 && m2xy == 0.0 && m2yy == 0.0) {
        for (int c = 0; c < 3; c++) { gPts[3 * t + c] = T(0); gNrm[3 * t + c] = T(0); }
        if (sdBatch == bn) { gSd[so] = T(0); gSd[so + 1] = T(0); }
        return;
    }
    D6 p[3], n[3];
    for (int c = 0; c < 3; c++) {
        p[c] = mk(pts[3 * t + c]); p[c].d[c] = 1.0;
        n[c] = mk(nrm[3 * t + c]); n[c].d[3 + c] = 1.0;
    }
    double s0 = sd[so], s1 = sd[so + 1];
    D6 sx, sy, A00, A01, A11;
    ewa_geom<D6>(p, n, W, H, f, sx, sy, A00, A01, A11);
    double a00 = A00.v, a01 = A01.v, a11 = A11.v;
    double detA = a00 * a11;
    // J = A^-1 (upper triangular)
    double J00 = a11 / detA, J01 = -a01 / detA, J10 = 0.0, J11 = a00 / detA;
    double D0 = s0 * s0, D1 = s1 * s1;
    double V00 = J00 * J00 * D0 + J01 * J01 * D1 + smoothing;
    double V01 = J00 * J10 * D0 + J01 * J11 * D1;
    double V11 = J10 * J10 * D0 + J11 * J11 * D1 + smoothing;
    double detV = V00 * V11 - V01 * V01;
    (void)detV;
    // T (as rounded in the slot pass) maps d to the whitened z = T d
    const SplatRecT<T>& R = rec[t];
    double T00 = R.t00, T01 = R.t01, T11 = R.t11;
    // G_s = V^-1 sum g w d = T^T (sum g w z)
    double Gsx = T00 * m1x;
    double Gsy = T01 * m1x + T11 * m1y;
    // G_V = 1/2 (V^-1 S2 V^-1 - m0 V^-1) = 1/2 T^T (S2' - m0 I) T
    double Q00 = 0.5 * (m2xx - m0), Q01 = 0.5 * m2xy, Q11 = 0.5 * (m2yy - m0);
    // R = Q T
    double R00 = Q00 * T00, R01 = Q00 * T01 + Q01 * T11;
    double R10 = Q01 * T00, R11 = Q01 * T01 + Q11 * T11;
    // G = T^T R
    double G00 = T00 * R00;
    double G01 = T00 * R01;
    double G11 = T01 * R01 + T11 * R11;
    (void)R10;
    double G10 = G01;
    // dL/dsigma_k = 2 sigma_k J[:,k]^T G_V J[:,k]
    double gs0 = 2.0 * s0 * (J00 * (G00 * J00 + G01 * J10) + J10 * (G10 * J00 + G11 * J10));
    double gs1 = 2.0 * s1 * (J01 * (G00 * J01 + G01 * J11) + J11 * (G10 * J01 + G11 * J11));
    // G_A = -2 J^T G_V J D J^T - m0 J^T
    double JD00 = J00 * D0, JD01 = J01 * D1, JD10 = J10 * D0, JD11 = J11 * D1;
    double C00 = JD00 * J00 + JD01 * J01, C01 = JD00 * J10 + JD01 * J11;   // C = J D J^T
    double C10 = JD10 * J00 + JD11 * J01, C11 = JD10 * J10 + JD11 * J11;
    double E00 = G00 * C00 + G01 * C10, E01 = G00 * C01 + G01 * C11;       // E = G_V C
    double E10 = G10 * C00 + G11 * C10, E11 = G10 * C01 + G11 * C11;
    // J^T E : (J^T)_{ij} = J_{ji}
    double F00 = J00 * E00 + J10 * E10, F01 = J00 * E01 + J10 * E11;
    double F10 = J01 * E00 + J11 * E10, F11 = J01 * E01 + J11 * E11;
    double GA00 = -2.0 * F00 - m0 * J00;
    double GA01 = -2.0 * F01 - m0 * J10;
    double GA11 = -2.0 * F11 - m0 * J11;
    (void)F10;
    // dL/dp_c = sum GA : dA/dp_c + G_s . ds/dp_c ; dL/dn_c = sum GA : dA/dn_c
    T gpt[3];
    for (int c = 0; c < 3; c++) {
        double gp = GA00 * A00.d[c] + GA01 * A01.d[c] + GA11 * A11.d[c] + Gsx * sx.d[c] + Gsy * sy.d[c];
        double gn = GA00 * A00.d[3 + c] + GA01 * A01.d[3 + c] + GA11 * A11.d[3 + c];
        gpt[c] = (T)gp;
        gNrm[3 * t + c] = (T)gn;
    }
    T gsd[2] = { (T)gs0, (T)gs1 };
    if (POST) {                      // sdBatch == bn: this thread owns the std-dev row
        cond_normalize<T, 3>(gpt);
        cond_normalize<T, 2>(gsd);
    }
    for (int c = 0; c < 3; c++) gPts[3 * t + c] = gpt[c];
    if (sdBatch == bn) { gSd[so] = gsd[0]; gSd[so + 1] = gsd[1]; }     // one writer per std-dev
    else { atomicAdd(gSd + so, gsd[0]); atomicAdd(gSd + so + 1, gsd[1]); }
}

// ---------------------------------------------------------------------------
// Horvitz-Thompson image estimate (no reference counterpart; it composes
// splatting.forward, the 1 / pi weighting and interpolation.forward of a
// three-channel attribute, and their adjoints).  A pixel's samples come first
// in its slots, then -1 padding -- the layout sampling.samplePoisson writes --
// so a pixel's slots are read up to its first -1.  Each slot is computed with
// the FP32 operations of the unfused ops (w = min(ewa, 1), wt = w * ipi,
// gW = sum_c g_c a_c, gw = gW * ipi, moments of gw * ewa), so the image equals
// the unfused composition bit for bit and the gradients differ from it only in
// the order of their atomic sums.  Neither pass stores a per-slot tensor.
// ---------------------------------------------------------------------------

// Four consecutive slots of a pixel (k % 4 == 0, 16-byte aligned rows): the
// records and attributes of all four are requested before any is used, so a
// lane has eight gathers in flight instead of two; ids past the first -1 are
// read as splat 0 (in bounds) and never used.
struct SlotQuad {
    int id[4];
    float q[4];
    SplatRec r[4];
    float a[4][3];
};

__device__ __forceinline__ int slot_quad(const int4* ip, const float4* qp, int c, const SplatRec* rv,
                                         const float* av, SlotQuad& Q)
{
    const int4 i4 = ip[c];
    const float4 q4 = qp[c];
    Q.id[0] = i4.x; Q.id[1] = i4.y; Q.id[2] = i4.z; Q.id[3] = i4.w;
    Q.q[0] = q4.x; Q.q[1] = q4.y; Q.q[2] = q4.z; Q.q[3] = q4.w;
    int n = 0;
    while (n < 4 && Q.id[n] >= 0) n++;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int id = u < n ? Q.id[u] : 0;
        Q.r[u] = rv[id];
        Q.a[u][0] = av[3 * id]; Q.a[u][1] = av[3 * id + 1]; Q.a[u][2] = av[3 * id + 2];
    }
    return n;                                    // samples in this quad, in slot order
}

// one thread per pixel; the channel sums run over the slots in order
__global__ __launch_bounds__(256) void k_ht_forward(long npix, int N, int W, int H, int k, int ab, int vec4,
                                                    const int32_t* __restrict__ idx, const float* __restrict__ ipi,
                                                    const SplatRec* __restrict__ rec,
                                                    const float* __restrict__ attr, float* __restrict__ out)
{
    // XCD-grouped block order: the columns a block shades share their splats with
    // the neighbouring blocks', which then run on the same XCD
    const long pix = (long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    const long b = pix / ((long)W * H);
    const int r = (int)(pix - b * (long)W * H);
    const int i = r / H, j = r - (r / H) * H;
    const SplatRec* rv = rec + b * N;
    const float* av = attr + 3 * (ab > 1 ? b * N : 0);
    const float px = (float)i, py = (float)j;
    float o0 = 0.0f, o1 = 0.0f, o2 = 0.0f;
    auto add = [&](const SplatRec& sr, const float* a, float q) {
        float u, v;
        const float wt = fminf(ewa_eval(sr, px, py, u, v), 1.0f) * q;
        o0 += wt * a[0]; o1 += wt * a[1]; o2 += wt * a[2];
    };
    if (vec4) {
        const int4* ip = (const int4*)(idx + pix * k);
        const float4* qp = (const float4*)(ipi + pix * k);
        for (int c = 0; c < (k >> 2); c++) {
            SlotQuad Q;
            const int n = slot_quad(ip, qp, c, rv, av, Q);
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (u < n) add(Q.r[u], Q.a[u], Q.q[u]);
            if (n < 4) break;
        }
    } else {
        for (int s = 0; s < k; s++) {
            const int id = idx[pix * k + s];
            if (id < 0) break;
            add(rv[id], av + 3 * id, ipi[pix * k + s]);
        }
    }
    out[3 * pix] = o0; out[3 * pix + 1] = o1; out[3 * pix + 2] = o2;
}

// The adjoint of one slot: interpolation.backward's attribute gradient g wt and
// weight gradient gW, times 1 / pi, into splatting.backward's moments.  v[0, 6)
// are the weight moments (zero for a slot whose unclamped weight exceeds 1),
// v[6, 9) the attribute gradient.  With compositing factors cf (the composited
// estimate), the attribute gradient is g wt T and the gradient of the slot's
// alpha = wt is cf.y = T (g . a - S) instead of g . a (k_htc's weight pass).
__device__ __forceinline__ void ht_slot_adjoint(const SplatRec& s, float px, float py, float g0, float g1, float g2,
                                                const float* a, float q, const float2* cf, float v[9])
{
    float u, w;
    const float we = ewa_eval(s, px, py, u, w);
    float wt = fminf(we, 1.0f) * q;
    float acc;
    if (cf) {
        const float2 c = *cf;
        wt = wt * c.x;
        acc = c.y;
    } else {
        acc = 0.0f;
        acc += g0 * a[0];
        acc += g1 * a[1];
        acc += g2 * a[2];
    }
    v[6] = g0 * wt;
    v[7] = g1 * wt;
    v[8] = g2 * wt;
    const float gw = (we <= 1.0f) ? (acc * q) * we : 0.0f;
    v[0] = gw;
    v[1] = gw * u;
    v[2] = gw * w;
    v[3] = (gw * u) * u;
    v[4] = (gw * u) * w;
    v[5] = (gw * w) * w;
}

__device__ __forceinline__ void ht_add_record(float* m, const float v[9])
{
    for (int c = 0; c < 6; c++)
        if (v[c] != 0.0f) atomicAdd(m + c, v[c]);
    for (int c = 0; c < 3; c++)
        if (v[6 + c] != 0.0f) atomicAdd(m + 8 + c, v[6 + c]);
}

// One workgroup per 16 x 16 pixel tile (tiles on blockIdx.x, views on
// blockIdx.y), one thread per pixel over its samples in slot order.  A splat is
// sampled by about 2.5 pixels of a tile on average, so the tile first sums its
// samples per splat in an LDS hash table (open addressing on the splat id) and
// then adds each splat's sums to its record with one set of device atomics; a
// sample that finds no free entry within kHtProbes adds its own values directly.
constexpr int kHtTile = 16;
constexpr int kHtSlots = 2048;
constexpr int kHtProbes = 32;

__global__ __launch_bounds__(kHtTile * kHtTile) void k_ht_backward(int N, int W, int H, int k, int ab, int vec4,
                                                                   const float* __restrict__ g,
                                                                   const int32_t* __restrict__ idx,
                                                                   const float* __restrict__ ipi,
                                                                   const SplatRec* __restrict__ rec,
                                                                   const float* __restrict__ attr,
                                                                   const float2* __restrict__ coef,
                                                                   float* __restrict__ mom)
{
    __shared__ int key[kHtSlots];
    __shared__ float sum[9][kHtSlots];
    for (int e = threadIdx.x; e < kHtSlots; e += blockDim.x) {
        key[e] = -1;
        for (int c = 0; c < 9; c++) sum[c][e] = 0.0f;
    }
    __syncthreads();
    const long b = blockIdx.y;
    const int tilesY = (H + kHtTile - 1) / kHtTile;
    const int ti = (int)blockIdx.x / tilesY, tj = (int)blockIdx.x - ti * tilesY;
    const int i = ti * kHtTile + (int)(threadIdx.x / kHtTile), j = tj * kHtTile + (int)(threadIdx.x % kHtTile);
    float* mv = mom + (long)kHtRec * b * N;
    if (i < W && j < H) {
        const long pix = b * W * H + j + (long)H * i;
        const float px = (float)i, py = (float)j;
        const float g0 = g[3 * pix], g1 = g[3 * pix + 1], g2 = g[3 * pix + 2];
        const float* av = attr + 3 * (ab > 1 ? b * N : 0);
        const SplatRec* rv = rec + b * N;
        const float2* cp = coef ? coef + pix * k : nullptr;
        auto sample = [&](int id, int slot_, const SplatRec& sr, const float* a, float q) {
            float v[9];
            ht_slot_adjoint(sr, px, py, g0, g1, g2, a, q, cp ? cp + slot_ : nullptr, v);
            int e = (int)(((uint32_t)id * 2654435761u) >> 21);            // 11 bits: kHtSlots entries
            int slot = -1;
            for (int t = 0; t < kHtProbes; t++) {
                const int old = atomicCAS(&key[e], -1, id);
                if (old == -1 || old == id) { slot = e; break; }
                e = (e + 1) & (kHtSlots - 1);
            }
            if (slot >= 0) {
                for (int c = 0; c < 9; c++)
                    if (v[c] != 0.0f) atomicAdd(&sum[c][slot], v[c]);
            } else {
                ht_add_record(mv + (long)kHtRec * id, v);
            }
        };
        if (vec4) {
            const int4* ip = (const int4*)(idx + pix * k);
            const float4* qp = (const float4*)(ipi + pix * k);
            for (int c = 0; c < (k >> 2); c++) {
                SlotQuad Q;
                const int n = slot_quad(ip, qp, c, rv, av, Q);
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (u < n) sample(Q.id[u], 4 * c + u, Q.r[u], Q.a[u], Q.q[u]);
                if (n < 4) break;
            }
        } else {
            for (int s = 0; s < k; s++) {
                const int id = idx[pix * k + s];
                if (id < 0) break;
                sample(id, s, rv[id], av + 3 * id, ipi[pix * k + s]);
            }
        }
    }
    __syncthreads();
    // Flush: nine consecutive lanes add the nine sums of one entry, so one atomic
    // instruction covers seven records (one 64 B line each) instead of 64 lines;
    // L2 serves a line's dwords of one instruction together (about 9x faster).
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int per = kHtSlots / (kHtTile * kHtTile / 64);
    const int sub = lane / 9, c = lane - 9 * sub;
    const int field = c < 6 ? c : c + 2;          // moments at [0, 6), attribute sums at [8, 11)
    for (int e = wv * per + sub; e < (wv + 1) * per + sub; e += 7) {
        if (lane < 63 && e < (wv + 1) * per) {
            const int id = key[e];
            if (id >= 0) {
                const float x = sum[c][e];
                if (x != 0.0f) atomicAdd(mv + (long)kHtRec * id + field, x);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// L2 image loss of a frame (no reference counterpart; the reference's scripts
// compute it with torch ops): grad = 2 (image - target) and loss =
// sum (image - target)^2 in one pass, the per-workgroup FP64 sums added by one
// workgroup in a fixed order (a reproducible loss), and no memset, so a captured
// frame can include it.
// ---------------------------------------------------------------------------
constexpr int kL2Blocks = 1024;

__global__ __launch_bounds__(256) void k_l2_grad(long n, const float* __restrict__ a, const float* __restrict__ b,
                                                 float* __restrict__ grad, double* __restrict__ part)
{
    __shared__ double red[4];
    double acc = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const float d = a[i] - b[i];
        grad[i] = 2.0f * d;
        acc += (double)d * (double)d;
    }
    acc = wave_sum_d(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ __launch_bounds__(256) void k_sum_parts(int nb, const double* __restrict__ part, double* __restrict__ out)
{
    __shared__ double red[4];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) acc += part[i];
    acc = wave_sum_d(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = ((red[0] + red[1]) + red[2]) + red[3];
}

// sampling.covariance (samplingCuda.cu:323-361), FP64 internally
__global__ void k_covariance(int bn, int N, int W, int H, double f, double smoothing,
                             const float* __restrict__ pts, const float* __restrict__ nrm,
                             const float* __restrict__ sd, int sdBatch, float* __restrict__ out)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)bn * N) return;
    long b = t / N, id = t % N;
    double p[3] = { pts[3 * t], pts[3 * t + 1], pts[3 * t + 2] };
    double n[3] = { nrm[3 * t], nrm[3 * t + 1], nrm[3 * t + 2] };
    long so = 2 * ((sdBatch > 1 ? b * N : 0) + id);
    double s0 = sd[so], s1 = sd[so + 1];
    double sx, sy, A00, A01, A11;
    ewa_geom<double>(p, n, W, H, f, sx, sy, A00, A01, A11);
    double detA = A00 * A11;
    double J00 = A11 / detA, J01 = -A01 / detA, J11 = A00 / detA;
    double V00 = J00 * J00 * s0 * s0 + J01 * J01 * s1 * s1 + smoothing;
    double V01 = J01 * J11 * s1 * s1;
    double V11 = J11 * J11 * s1 * s1 + smoothing;
    out[4 * t] = (float)V00; out[4 * t + 1] = (float)V01;
    out[4 * t + 2] = (float)V01; out[4 * t + 3] = (float)V11;
}

// ---------------------------------------------------------------------------
// Compositing.  Depth of slot s along the pixel ray (alphacompositing.cuh:6-36),
// in a fixed operation order.  The storage type S of the weights and points is
// half, float or double (the reference dispatches all three, zSortingCuda.cu:100,
// 220); arithmetic runs in C = float for half and float, double for double.
// ---------------------------------------------------------------------------
template <class S> struct CompOf { using type = float; };
template <> struct CompOf<double> { using type = double; };
__device__ __forceinline__ float ld_c(const float* p, long i) { return p[i]; }
__device__ __forceinline__ double ld_c(const double* p, long i) { return p[i]; }
__device__ __forceinline__ float ld_c(const __half* p, long i) { return __half2float(p[i]); }
__device__ __forceinline__ void st_c(float* p, long i, float v) { p[i] = v; }
__device__ __forceinline__ void st_c(double* p, long i, double v) { p[i] = v; }
__device__ __forceinline__ void st_c(__half* p, long i, float v) { p[i] = __float2half(v); }

template <class C>
__device__ __forceinline__ C dot3c(const C* a, const C* b)
{
    C x = a[0] * b[0];
    x = x + a[1] * b[1];
    x = x + a[2] * b[2];
    return x;
}

template <class C>
__device__ __forceinline__ C ray_depth(int i, int j, C Wf, C Hf, C ff, const C* p, const C* n)
{
    C bx = (Wf - C(2) * (C)i) / (C(2) * ff);
    C by = -((Hf - C(2) * (C)j) / (C(2) * ff));
    C bz = C(1);
    C nn = bx * bx;
    nn = nn + by * by;
    nn = nn + bz * bz;
    C inv = C(1) / sqrt(nn);
    C dir[3] = { -(bx * inv), -(by * inv), -(bz * inv) };
    C num = dot3c(p, n);
    C den = dot3c(dir, n);
    C d = num / den;
    return (d > C(0)) ? d : C(10000);
}

// Cross-lane moves of the 64-lane wavefront through DPP (gfx9 controls:
// row_shr:n = 0x110 + n inside 16-lane rows, row_bcast:15 = 0x142 and
// row_bcast:31 = 0x143 across rows).  Lanes whose source is outside the row, or
// whose row the row mask leaves out, return `ident`.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_from(float ident, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(ident), __float_as_int(v), CTRL, ROWMASK, 0xf,
                                                      false));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_from(double ident, double v)
{
    const long long a = __double_as_longlong(ident), b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)a, (int)b, CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(a >> 32), (int)(b >> 32), CTRL, ROWMASK, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Inclusive product over lanes 0..s (Hillis-Steele: shifts 1, 2, 4, 8 inside
// rows, then the row broadcasts).
template <int CTRL, int RM, class C>
__device__ __forceinline__ void prod_step(C& v) { v = v * dpp_from<CTRL, RM>(C(1), v); }
template <class C>
__device__ __forceinline__ C wave_prod_incl(C v)
{
    prod_step<0x111, 0xf>(v);
    prod_step<0x112, 0xf>(v);
    prod_step<0x114, 0xf>(v);
    prod_step<0x118, 0xf>(v);
    prod_step<0x142, 0xa>(v);
    prod_step<0x143, 0xc>(v);
    return v;
}

// Inclusive composition of affine maps x -> A + B x over lanes 0..s, the map of
// the lower lane applied first: (A, B) o (tA, tB) = (A + B tA, B tB).
template <int CTRL, int RM, class C>
__device__ __forceinline__ void affine_step(C& A, C& B)
{
    const C tA = dpp_from<CTRL, RM>(C(0), A);
    const C tB = dpp_from<CTRL, RM>(C(1), B);
    A = A + B * tA;
    B = B * tB;
}
template <class C>
__device__ __forceinline__ void wave_affine_incl(C& A, C& B)
{
    affine_step<0x111, 0xf>(A, B);
    affine_step<0x112, 0xf>(A, B);
    affine_step<0x114, 0xf>(A, B);
    affine_step<0x118, 0xf>(A, B);
    affine_step<0x142, 0xa>(A, B);
    affine_step<0x143, 0xc>(A, B);
}

__device__ __forceinline__ float quiet_nan(float) { return __int_as_float(0x7fc00000); }
__device__ __forceinline__ double quiet_nan(double) { return __longlong_as_double(0x7ff8000000000000ll); }

// One wavefront per pixel (k <= 64): lane s owns slot s.  The stable
// insertion argsort of the reference becomes a rank per lane:
//   rank_s = #{d_j < d_s} + #{j < s : d_j == d_s}.
// The first term compares the lane's depth with every slot's, read from LDS by
// broadcast (lanes past k hold NaN, which compares false); lanes of equal depth
// get the same first term, so each ORs its bit into the mask at that index and
// the second term is the count of mask bits below the lane.  Depths are never
// NaN (ray_depth maps a non-positive or NaN quotient to 10000).
template <bool BACKWARD, class S>
__global__ __launch_bounds__(256) void k_filter_wave(long npix, int N, int W, int H, int k,
                                                     typename CompOf<S>::type ff, typename CompOf<S>::type gate,
                                                     const int32_t* __restrict__ idx, const S* __restrict__ wts,
                                                     const S* __restrict__ gin,
                                                     const S* __restrict__ pts, const S* __restrict__ nrm, int pb,
                                                     S* __restrict__ out)
{
    using C = typename CompOf<S>::type;
    __shared__ C sbuf[4][2][64];
    __shared__ C skey[4][64];
    __shared__ unsigned long long sgrp[4][64];
    // the wavefront index is uniform: read it into a scalar so the pixel
    // arithmetic below runs on the scalar unit
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), s = threadIdx.x & 63;
    long pix = (long)blockIdx.x * 4 + wv;
    if (pix >= npix) return;   // whole wave exits together
    long b = pix / ((long)W * H);
    int r = (int)(pix - b * (long)W * H);
    int i = r / H, j = r - (r / H) * H;
    long base = pix * k;
    const bool act = s < k;
    int id = act ? idx[base + s] : -1;
    C w = act ? ld_c(wts, base + s) : C(0);
    C gi = (BACKWARD && act) ? ld_c(gin, base + s) : C(0);
    C d = C(10000);
    if (id >= 0) {
        long po = 3 * ((pb > 1 ? b * N : 0) + id);
        C p[3] = { ld_c(pts, po), ld_c(pts, po + 1), ld_c(pts, po + 2) };
        C n[3] = { ld_c(nrm, po), ld_c(nrm, po + 1), ld_c(nrm, po + 2) };
        d = ray_depth<C>(i, j, (C)W, (C)H, ff, p, n);
    }
    C* key = skey[wv];
    key[s] = act ? d : quiet_nan(C(0));
    sgrp[wv][s] = 0ull;
    __builtin_amdgcn_wave_barrier();
    int lt = 0;
    const int k16 = (k + 15) & ~15;   // <= 64: the reads stay inside the row
    for (int q0 = 0; q0 < k16; q0 += 16) {
        C kq[16];
#pragma unroll
        for (int q = 0; q < 16; q++) kq[q] = key[q0 + q];
#pragma unroll
        for (int q = 0; q < 16; q++) lt += kq[q] < d;
    }
    if (act) atomicOr(&sgrp[wv][lt], 1ull << s);
    __builtin_amdgcn_wave_barrier();
    const unsigned long long grp = act ? sgrp[wv][lt] : 0ull;
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(grp >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((unsigned)grp, lt));
    // exclusive prefix product of (1 - w) in rank order
    C* sa = sbuf[wv][0];
    C* sb = sbuf[wv][1];
    sb[s] = C(1);
    if (BACKWARD) sa[s] = C(0);
    __builtin_amdgcn_wave_barrier();
    if (act) {
        sb[rank] = C(1) - w;
        if (BACKWARD) sa[rank] = gi * w;
    }
    __builtin_amdgcn_wave_barrier();
    const C incl = wave_prod_incl(sb[s]);
    // transmittance in front of this slot: the inclusive product of the rank before it
    C rem = __shfl(incl, rank > 0 ? rank - 1 : 0, 64);
    if (rank == 0) rem = C(1);
    if (!BACKWARD) {
        if (act) st_c(out, base + s, rem * w);
        return;
    }
    // backward (alphacompositingGrad.cuh:7-26).  In depth order, with T_r the
    // transmittance in front of rank r, the reference's sum
    //   g_r = [|w_r| > eps] gin_r out_r / w_r - [|1 - w_r| > eps] sum_{i>r} gin_i out_i / (1 - w_r)
    // is [|w_r| > eps] gin_r T_r - [|1 - w_r| > eps] T_r S_r with
    //   S_r = sum_{i>r} gin_i w_i prod_{r<t<i} (1 - w_t),
    // the suffix recurrence S_r = a_{r+1} + b_{r+1} S_{r+1} (a = gin w, b = 1 - w).
    // Nothing is divided, so a small w or 1 - w amplifies no rounding.
    // Lane l holds the affine map x -> A + B x of rank 63 - l (identity past the
    // last slot); the inclusive scan composes ranks 63 .. 63 - l, outer (smaller
    // rank) applied last.
    C A = sa[63 - s], B = sb[63 - s];
    wave_affine_incl(A, B);
    // S_r = U_{r+1}, held by lane 63 - (r + 1)
    int src = 62 - rank;
    C after = __shfl(A, src < 0 ? 0 : src, 64);
    if (rank == 63) after = C(0);
    // the reference's gates (|w| > 1e-4, |1 - w| > 1e-4); gate < 0 keeps every term
    // (the exact derivative: nothing here is divided, so no term needs dropping)
    C g = C(0);
    if (fabs(w) > gate) g += gi * rem;
    if (fabs(C(1) - w) > gate) g -= rem * after;
    if (act) st_c(out, base + s, g);
}

// General k (64 < k <= 1024): one workgroup per pixel, ranks from LDS, serial scans.
template <bool BACKWARD, class S>
__global__ __launch_bounds__(1024) void k_filter_block(long npix, int N, int W, int H, int k,
                                                       typename CompOf<S>::type ff, typename CompOf<S>::type gate,
                                                       const int32_t* __restrict__ idx, const S* __restrict__ wts,
                                                       const S* __restrict__ gin,
                                                       const S* __restrict__ pts, const S* __restrict__ nrm, int pb,
                                                       S* __restrict__ out)
{
    using C = typename CompOf<S>::type;
    __shared__ C sd_[1024], sw[1024], sx[1024];
    __shared__ int srank[1024];
    long pix = blockIdx.x;
    long b = pix / ((long)W * H);
    int r = (int)(pix - b * (long)W * H);
    int i = r / H, j = r - (r / H) * H;
    long base = pix * k;
    int s = threadIdx.x;
    bool act = s < k;
    int id = act ? idx[base + s] : -1;
    C w = act ? ld_c(wts, base + s) : C(0);
    C d = C(10000);
    if (id >= 0) {
        long po = 3 * ((pb > 1 ? b * N : 0) + id);
        C p[3] = { ld_c(pts, po), ld_c(pts, po + 1), ld_c(pts, po + 2) };
        C n[3] = { ld_c(nrm, po), ld_c(nrm, po + 1), ld_c(nrm, po + 2) };
        d = ray_depth<C>(i, j, (C)W, (C)H, ff, p, n);
    }
    if (act) { sd_[s] = d; sw[s] = w; }
    __syncthreads();
    int rank = 0;
    if (act) {
        for (int q = 0; q < k; q++) { C dq = sd_[q]; rank += (dq < d) || (dq == d && q < s); }
        srank[rank] = s;
    }
    __syncthreads();
    if (s == 0) {   // sx[rank] = exclusive product of (1 - w) in rank order
        C rem = C(1);
        for (int q = 0; q < k; q++) { sx[q] = rem; rem *= (C(1) - sw[srank[q]]); }
    }
    __syncthreads();
    C rem = act ? sx[rank] : C(0);
    if (!BACKWARD) { if (act) st_c(out, base + s, rem * w); return; }
    // backward: the division-free form of the wave kernel, S in a serial suffix pass
    C gi = act ? ld_c(gin, base + s) : C(0);
    __syncthreads();
    if (act) sd_[rank] = gi * w;
    __syncthreads();
    if (s == 0) {   // sd_[rank] = S_rank = sum over later ranks, recurrence from the back
        C acc = C(0);
        for (int q = k - 1; q >= 0; q--) {
            C a = sd_[q];
            sd_[q] = acc;
            acc = a + (C(1) - sw[srank[q]]) * acc;
        }
    }
    __syncthreads();
    if (!act) return;
    C g = C(0);
    if (fabs(w) > gate) g += gi * rem;
    if (fabs(C(1) - w) > gate) g -= rem * sd_[rank];
    st_c(out, base + s, g);
}

// ---------------------------------------------------------------------------
// Composited Horvitz-Thompson estimate (no reference counterpart: the
// Poisson-rate counterpart of splatting.forward -> filtering.forward ->
// interpolation.forward).  With independent inclusions I_s ~ Bernoulli(pi_s)
// and alpha_s = I_s min(w_s, 1) / pi_s, and a pixel's splats ordered by ray
// depth as filtering orders them (alphacompositing.cuh:6-54),
//     image = sum_s alpha_s a_s prod_{t in front of s} (1 - alpha_t)
// has expectation sum_s w_s a_s prod_{t in front of s} (1 - w_t): the factors of
// different splats are independent, each of mean (1 - w_t), so the estimate of
// the image composited over every splat of the proposal's support is unbiased
// (the reference's sampled set composited by filtering is not).  Its adjoint in
// depth order, with T_r the product in front of rank r and
// S_r = sum_{i > r} alpha_i (g . a_i) prod_{r < t < i} (1 - alpha_t):
//     d/d alpha_r = T_r (g . a_r - S_r),   d/d a_r = alpha_r T_r g.
// One lane per pixel over its samples (read up to the first -1); a sample's
// depth, alpha and depth rank live in LDS arrays [entry][lane].
// ---------------------------------------------------------------------------
struct PlaneRec {
    float nx, ny, nz, c;      // camera-space normal and c = p . n (ray depth c / (dir . n))
};

__global__ void k_plane_prepare(long n, const float* __restrict__ pts, const float* __restrict__ nrm,
                                PlaneRec* __restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float p[3] = { pts[3 * t], pts[3 * t + 1], pts[3 * t + 2] };
    const float q[3] = { nrm[3 * t], nrm[3 * t + 1], nrm[3 * t + 2] };
    PlaneRec r;
    r.nx = q[0]; r.ny = q[1]; r.nz = q[2];
    r.c = dot3(p, q);                          // ray_depth's numerator, same operation order
    out[t] = r;
}

// the unit direction towards the camera through pixel (i, j), ray_depth's operation order
__device__ __forceinline__ void pixel_ray(int i, int j, float Wf, float Hf, float ff, float dir[3])
{
    float bx = (Wf - 2.0f * (float)i) / (2.0f * ff);
    float by = -((Hf - 2.0f * (float)j) / (2.0f * ff));
    float bz = 1.0f;
    float nn = bx * bx;
    nn = nn + by * by;
    nn = nn + bz * bz;
    float inv = 1.0f / sqrtf(nn);
    dir[0] = -(bx * inv); dir[1] = -(by * inv); dir[2] = -(bz * inv);
}

__device__ __forceinline__ float plane_depth(const PlaneRec& P, const float dir[3])
{
    const float n[3] = { P.nx, P.ny, P.nz };
    const float d = P.c / dot3(dir, n);
    return (d > 0.0f) ? d : 10000.0f;
}

constexpr int kHtcMaxK = 112;          // per-lane LDS entries (9 B each): at most 63 KB per wavefront

// BWD = false: the image.  BWD = true: the compositing factors (T_s, T_s (g . a_s - S_s))
// of every sample slot, read by k_ht_backward.
template <bool BWD>
__global__ __launch_bounds__(64) void k_htc(long npix, int N, int W, int H, int k, int ab, float ff,
                                            const int32_t* __restrict__ idx, const float* __restrict__ ipi,
                                            const SplatRec* __restrict__ rec, const PlaneRec* __restrict__ pl,
                                            const float* __restrict__ attr, const float* __restrict__ g,
                                            float* __restrict__ out, float2* __restrict__ coef)
{
    extern __shared__ float lds[];
    float* sdep = lds;                                  // [k][64] depth by slot; later T by rank
    float* salp = lds + (long)k * 64;                   // [k][64] alpha by slot
    uint8_t* sord = (uint8_t*)(lds + 2L * k * 64);      // [k][64] slot of each depth rank
    const int lane = threadIdx.x;
    const long pix = (long)xcd_remap(blockIdx.x, gridDim.x) * 64 + lane;
    if (pix >= npix) return;
    const long b = pix / ((long)W * H);
    const int r = (int)(pix - b * (long)W * H);
    const int i = r / H, j = r - (r / H) * H;
    const SplatRec* rv = rec + b * N;
    const PlaneRec* pv = pl + b * N;
    const float* av = attr + 3 * (ab > 1 ? b * N : 0);
    const float px = (float)i, py = (float)j;
    float dir[3];
    pixel_ray(i, j, (float)W, (float)H, ff, dir);
    const int32_t* ip = idx + pix * k;
    const float* qp = ipi + pix * k;
    int n = 0;
    for (; n < k; n++) {
        const int id = ip[n];
        if (id < 0) break;
        float u, v;
        const float alpha = fminf(ewa_eval(rv[id], px, py, u, v), 1.0f) * qp[n];
        const float d = plane_depth(pv[id], dir);
        sdep[n * 64 + lane] = d;
        salp[n * 64 + lane] = alpha;
        // stable insertion: behind every earlier slot of equal depth
        int q = n;
        while (q > 0) {
            const int o = sord[(q - 1) * 64 + lane];
            if (!(sdep[o * 64 + lane] > d)) break;
            sord[q * 64 + lane] = (uint8_t)o;
            q--;
        }
        sord[q * 64 + lane] = (uint8_t)n;
    }
    if (!BWD) {
        float T = 1.0f, o0 = 0.0f, o1 = 0.0f, o2 = 0.0f;
        for (int q = 0; q < n; q++) {
            const int s = sord[q * 64 + lane];
            const float al = salp[s * 64 + lane];
            const float* a = av + 3 * ip[s];
            const float c = al * T;
            o0 += c * a[0]; o1 += c * a[1]; o2 += c * a[2];
            T = T * (1.0f - al);
        }
        out[3 * pix] = o0; out[3 * pix + 1] = o1; out[3 * pix + 2] = o2;
        return;
    }
    const float g0 = g[3 * pix], g1 = g[3 * pix + 1], g2 = g[3 * pix + 2];
    float T = 1.0f;
    for (int q = 0; q < n; q++) {             // T in front of each rank, over the depths no longer needed
        const int s = sord[q * 64 + lane];
        sdep[q * 64 + lane] = T;
        T = T * (1.0f - salp[s * 64 + lane]);
    }
    float S = 0.0f;
    float2* cp = coef + pix * k;
    for (int q = n - 1; q >= 0; q--) {
        const int s = sord[q * 64 + lane];
        const float al = salp[s * 64 + lane];
        const float Tq = sdep[q * 64 + lane];
        const float* a = av + 3 * ip[s];
        float ga = 0.0f;
        ga += g0 * a[0];
        ga += g1 * a[1];
        ga += g2 * a[2];
        cp[s] = make_float2(Tq, Tq * (ga - S));
        S = al * ga + (1.0f - al) * S;
    }
}

// ---------------------------------------------------------------------------
// Transform (transformCuda.cu:11-117).  affine = 1 adds the translation column.
// ---------------------------------------------------------------------------
// facing (optional, linear transforms of normals): the transformed vector
// negated where its z component is positive, so it faces the camera
// (render.py:200-203)
template <class T>
__global__ __launch_bounds__(256) void k_transform_forward(int affine, int N, int vb, const T* __restrict__ M,
                                                           const T* __restrict__ v, T* __restrict__ out,
                                                           T* __restrict__ facing)
{
    // view b = blockIdx.y: its matrix is uniform (scalar loads), no division per point
    const int b = blockIdx.y;
    const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= N) return;
    const T* m = M + 16 * b;
    const long t = (long)b * N + x;
    const long vo = 3 * ((vb > 1 ? (long)b * N : 0) + x);
    T a0 = v[vo], a1 = v[vo + 1], a2 = v[vo + 2];
    T o[3];
    for (int r = 0; r < 3; r++) {
        T s = m[4 * r] * a0 + m[4 * r + 1] * a1 + m[4 * r + 2] * a2;
        if (affine) s += m[4 * r + 3];
        out[3 * t + r] = s;
        o[r] = s;
    }
    if (facing) {
        const bool flip = o[2] > T(0);
        for (int r = 0; r < 3; r++) facing[3 * t + r] = flip ? -o[r] : o[r];
    }
}

__device__ __forceinline__ float wave_sum_t(float v) { return wave_sum(v); }
__device__ __forceinline__ double wave_sum_t(double v) { return wave_sum_d(v); }

// gradient: dM[r][c] = sum_n g_r v_c (+ g_r in column 3), dv = M33^T g.
// The 12 matrix partial sums are reduced per workgroup, then one atomic each.
// The 12 words share one cache line, so their atomics serialise: 256 workgroups
// of 1024 threads per view (16 waves per CU) stream the points with 3072 such
// atomics (1M points: 11.7 us, where 512 x 256 threads took 15.5 us and
// 2048 x 256 33.5 us).
template <class T>
__global__ __launch_bounds__(1024) void k_transform_backward(int affine, int bn, int vb, int N, const T* __restrict__ g,
                                                            const T* __restrict__ M, const T* __restrict__ v,
                                                            T* __restrict__ gM, T* __restrict__ gV)
{
    __shared__ T red[16][12];
    int b = blockIdx.y;
    T acc[12];
    for (int q = 0; q < 12; q++) acc[q] = T(0);
    const T* m = M + 16 * b;
    // grid-stride over the points: few workgroups per view, so the 12 matrix
    // sums see few atomics (they all land on the same 12 words)
    for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < N; x += (long)gridDim.x * blockDim.x) {
        long vo = 3 * ((vb > 1 ? (long)b * N : 0) + x);
        long go = 3 * ((long)b * N + x);
        T g0 = g[go], g1 = g[go + 1], g2 = g[go + 2];
        T a[3] = { v[vo], v[vo + 1], v[vo + 2] };
        T gg[3] = { g0, g1, g2 };
        for (int r = 0; r < 3; r++) {
            for (int c = 0; c < 3; c++) acc[4 * r + c] += gg[r] * a[c];
            if (affine) acc[4 * r + 3] += gg[r];
        }
        T gv[3];
        for (int c = 0; c < 3; c++) gv[c] = m[c] * g0 + m[4 + c] * g1 + m[8 + c] * g2;
        if (vb > 1 || bn == 1) { for (int c = 0; c < 3; c++) gV[vo + c] = gv[c]; }
        else { for (int c = 0; c < 3; c++) atomicAdd(gV + vo + c, gv[c]); }
    }
    for (int q = 0; q < 12; q++) acc[q] = wave_sum_t(acc[q]);
    int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) for (int q = 0; q < 12; q++) red[wv][q] = acc[q];
    __syncthreads();
    if (threadIdx.x < 12) {
        T s = red[0][threadIdx.x];
        for (int w = 1; w < 16; w++) s += red[w][threadIdx.x];
        atomicAdd(gM + 16 * b + threadIdx.x, s);
    }
}

// ---------------------------------------------------------------------------
// Interpolation (interpolationCuda.cu:5-129).
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_interp_forward(long npix, int W, int H, int k, int ab, int N, int C, int vec4,
                                 const int32_t* __restrict__ idx, const T* __restrict__ wts,
                                 const T* __restrict__ attr, T* __restrict__ out)
{
    long pix = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    long b = pix / ((long)W * H);
    const long abase = ab > 1 ? b * N : 0;
    if (C == 3) {
        T o0 = T(0), o1 = T(0), o2 = T(0);
        auto add = [&](int id, T w) {
            if (id < 0) return;
            const T* a = attr + 3 * (abase + id);
            o0 += w * a[0]; o1 += w * a[1]; o2 += w * a[2];
        };
        if constexpr (sizeof(T) == 4) {
            if (vec4) {
                // a pixel's k slots are contiguous and 16-byte aligned: read them four at a time
                const int4* ip = (const int4*)(idx + pix * k);
                const float4* wp = (const float4*)(wts + pix * k);
                for (int s = 0; s < (k >> 2); s++) {
                    const int4 i4 = ip[s];
                    const float4 w4 = wp[s];
                    add(i4.x, w4.x); add(i4.y, w4.y); add(i4.z, w4.z); add(i4.w, w4.w);
                }
                out[3 * pix] = o0; out[3 * pix + 1] = o1; out[3 * pix + 2] = o2;
                return;
            }
        }
        for (int s = 0; s < k; s++) add(idx[pix * k + s], wts[pix * k + s]);
        out[3 * pix] = o0; out[3 * pix + 1] = o1; out[3 * pix + 2] = o2;
        return;
    }
    for (int c = 0; c < C; c++) out[(long)C * pix + c] = T(0);
    for (int s = 0; s < k; s++) {
        int id = idx[pix * k + s];
        if (id < 0) continue;
        T w = wts[pix * k + s];
        const T* a = attr + (long)C * (abase + id);
        for (int c = 0; c < C; c++) out[(long)C * pix + c] += w * a[c];
    }
}

// interpolation.forward for float, C = 3, k <= 64: one wavefront per 64
// consecutive pixels.  Their slot rows are one contiguous block of 64 k indices
// and weights, read with coalesced 16-byte loads into LDS (rows padded to k + 1
// words, so the lanes' row reads fall in distinct banks); each lane then sums its
// pixel's slots in slot order, as the per-pixel kernel does (same result).
constexpr int kInterpTileMaxK = 64;
__global__ __launch_bounds__(64) void k_interp_forward_tile(long npix, int W, int H, int k, int ab, int N,
                                                            const int32_t* __restrict__ idx,
                                                            const float* __restrict__ wts,
                                                            const float* __restrict__ attr, float* __restrict__ out)
{
    extern __shared__ int32_t sh[];
    int32_t* sI = sh;                                  // [64][k + 1]
    float* sW = (float*)(sh + 64 * (k + 1));
    const int lane = threadIdx.x;
    const long pix0 = (long)blockIdx.x * 64;
    const int np = npix - pix0 < 64 ? (int)(npix - pix0) : 64;
    const int kp = k + 1;
    const int total = np * k;                          // contiguous words of this block's rows
    const int32_t* gi = idx + pix0 * k;
    const float* gw = wts + pix0 * k;
    if ((k & 3) == 0) {
        // 16-byte loads (k % 4 == 0 keeps every row, and the block, 16-byte aligned)
        const int4* gi4 = (const int4*)gi;
        const float4* gw4 = (const float4*)gw;
        for (int e = lane; e < (total >> 2); e += 64) {
            const int4 a = gi4[e];
            const float4 b = gw4[e];
            const int r = (4 * e) / k, c = 4 * e - r * k;
            int32_t* di = sI + r * kp + c;
            float* dw = sW + r * kp + c;
            di[0] = a.x; di[1] = a.y; di[2] = a.z; di[3] = a.w;
            dw[0] = b.x; dw[1] = b.y; dw[2] = b.z; dw[3] = b.w;
        }
    } else {
        for (int e = lane; e < total; e += 64) {
            const int r = e / k, c = e - r * k;
            sI[r * kp + c] = gi[e];
            sW[r * kp + c] = gw[e];
        }
    }
    __syncthreads();
    if (lane >= np) return;
    const long pix = pix0 + lane;
    const long b = pix / ((long)W * H);
    const long abase = ab > 1 ? b * N : 0;
    float o0 = 0.0f, o1 = 0.0f, o2 = 0.0f;
    const int32_t* ri = sI + lane * kp;
    const float* rw = sW + lane * kp;
    for (int s = 0; s < k; s++) {
        const int id = ri[s];
        if (id < 0) continue;
        const float w = rw[s];
        const float* a = attr + 3 * (abase + id);
        o0 += w * a[0]; o1 += w * a[1]; o2 += w * a[2];
    }
    out[3 * pix] = o0; out[3 * pix + 1] = o1; out[3 * pix + 2] = o2;
}

template <class T>
__global__ __launch_bounds__(256) void k_interp_backward(long nslots, int W, int H, int k, int ab, int N, int C,
                                                         const T* __restrict__ g,
                                                         const int32_t* __restrict__ idx,
                                                         const T* __restrict__ wts,
                                                         const T* __restrict__ attr, T* __restrict__ gW,
                                                         T* __restrict__ gA)
{
    // three-channel attributes: the wavefront adds the slots' g w with three
    // consecutive lanes per attribute row (21 rows per atomic instruction)
    __shared__ T sv[4][64][3];
    __shared__ long sa[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long ao = -1;
    T v3[3] = { T(0), T(0), T(0) };
    if (t < nslots) {
        const int id = idx[t];
        if (id < 0) gW[t] = T(0);
        else {
            long pix = t / k;
            long b = pix / ((long)W * H);
            ao = (long)C * ((ab > 1 ? b * N : 0) + id);
            T w = wts[t];
            T acc = T(0);
            for (int c = 0; c < C; c++) {
                T gc = g[(long)C * pix + c];
                acc += gc * attr[ao + c];
                const T v = gc * w;
                if (C == 3) v3[c] = v;
                else if (v != T(0)) atomicAdd(gA + ao + c, v);      // zero terms (occluded slots) add nothing
            }
            gW[t] = acc;
        }
    }
    if (C != 3) return;                                // block-uniform
    for (int c = 0; c < 3; c++) sv[wv][lane][c] = v3[c];
    sa[wv][lane] = ao;
    __syncthreads();
    const int sub = lane / 3, c = lane - 3 * sub;
    for (int q = 0; q < 64; q += 21) {
        const int slot = q + sub;
        if (lane < 63 && slot < 64) {
            const long a = sa[wv][slot];
            const T x = sv[wv][slot][c];
            if (a >= 0 && x != T(0)) atomicAdd(gA + a + c, x);
        }
    }
}

// ---------------------------------------------------------------------------
// Shading (shadingCuda.cu, diffuseShading*.cuh).
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void sh_matrix(const T* c, T M[16])
{
    const T c1 = T(0.429043), c2 = T(0.511664), c3 = T(0.743125), c4 = T(0.886227), c5 = T(0.247708);
    M[0] = c1 * c[8];  M[1] = c1 * c[4];  M[2] = c1 * c[7];  M[3] = c2 * c[3];
    M[4] = c1 * c[4];  M[5] = -c1 * c[8]; M[6] = c1 * c[5];  M[7] = c2 * c[1];
    M[8] = c1 * c[7];  M[9] = c1 * c[5];  M[10] = c3 * c[6]; M[11] = c2 * c[2];
    M[12] = c2 * c[3]; M[13] = c2 * c[1]; M[14] = c2 * c[2]; M[15] = c4 * c[0] - c5 * c[6];
}

template <class T>
__device__ __forceinline__ T sh_n4(const T* n, T n4[4])
{
    T nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (nn > T(1e-5)) { T inv = T(1) / nn; n4[0] = n[0] * inv; n4[1] = n[1] * inv; n4[2] = n[2] * inv; n4[3] = T(1); }
    else { n4[0] = n4[1] = n4[2] = T(0); n4[3] = T(1); }
    return nn;
}

template <class T>
__global__ void k_sh_forward(long npix, int W, int H, int cb, const T* __restrict__ coef,
                             const T* __restrict__ nrm, const T* __restrict__ diff, T* __restrict__ out)
{
    long pix = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    long b = pix / ((long)W * H);
    T n4[4];
    sh_n4(nrm + 3 * pix, n4);
    for (int ch = 0; ch < 3; ch++) {
        T M[16];
        sh_matrix(coef + 9 * (cb > 1 ? ch + 3 * b : ch), M);
        T res = T(0);
        for (int rr = 0; rr < 4; rr++) {
            T y = n4[0] * M[rr] + n4[1] * M[4 + rr] + n4[2] * M[8 + rr] + n4[3] * M[12 + rr];
            res += y * n4[rr];
        }
        out[3 * pix + ch] = diff[3 * pix + ch] * res / T(UD_PI);
    }
}

// backwardSh keeps the reference's per-component normal gradient
// (shadingCuda.cu:283 pairs colour channel i with normal component i).
template <class T>
__global__ __launch_bounds__(256) void k_sh_backward(long npix, int W, int H, int cb, const T* __restrict__ g,
                                                     const T* __restrict__ coef, const T* __restrict__ nrm,
                                                     const T* __restrict__ diff, T* __restrict__ gCoef,
                                                     T* __restrict__ gN, T* __restrict__ gD)
{
    __shared__ T red[4][27];
    const T c1 = T(0.429043), c2 = T(0.511664), c3 = T(0.743125), c4 = T(0.886227), c5 = T(0.247708);
    long pix = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // one workgroup never spans two views when W*H % 256 == 0; otherwise fall back to per-lane atomics
    const long WH = (long)W * H;
    long b0 = ((long)blockIdx.x * blockDim.x) / WH;
    long bl = ((long)blockIdx.x * blockDim.x + blockDim.x - 1) / WH;
    bool uniformView = (b0 == bl) || cb == 1;
    T acc[27];
    for (int q = 0; q < 27; q++) acc[q] = T(0);
    long b = pix < npix ? pix / WH : b0;
    if (pix < npix) {
        const T* n = nrm + 3 * pix;
        T n4[4];
        T nn = sh_n4(n, n4);
        T Gn[9] = { T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0) };
        if (nn > T(1e-5)) {
            T inv = T(1) / nn, inv3 = inv * inv * inv;
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) Gn[3 * r + c] = (r == c ? inv : T(0)) - n[r] * n[c] * inv3;
        }
        for (int ch = 0; ch < 3; ch++) {
            T M[16];
            sh_matrix(coef + 9 * (cb > 1 ? ch + 3 * b : ch), M);
            T res = T(0), gN4[4];
            for (int rr = 0; rr < 4; rr++) {
                T y = n4[0] * M[rr] + n4[1] * M[4 + rr] + n4[2] * M[8 + rr] + n4[3] * M[12 + rr];
                res += y * n4[rr];
            }
            for (int rr = 0; rr < 4; rr++) {
                gN4[rr] = T(0);
                for (int c = 0; c < 4; c++) gN4[rr] += (M[4 * rr + c] + M[4 * c + rr]) * n4[c];
            }
            T gnv = Gn[3 * ch] * gN4[0] + Gn[3 * ch + 1] * gN4[1] + Gn[3 * ch + 2] * gN4[2];
            T gi = g[3 * pix + ch], di = diff[3 * pix + ch];
            gD[3 * pix + ch] = gi * res / T(UD_PI);
            gN[3 * pix + ch] = gi * di * gnv / T(UD_PI);
            T sc = gi * di / T(UD_PI);
            T gM[16];
            for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) gM[4 * r + c] = n4[r] * n4[c];
            T gc[9];
            gc[0] = c4 * gM[15];
            gc[1] = c2 * (gM[7] + gM[13]);
            gc[2] = c2 * (gM[14] + gM[11]);
            gc[3] = c2 * (gM[12] + gM[3]);
            gc[4] = c1 * (gM[4] + gM[1]);
            gc[5] = c1 * (gM[9] + gM[6]);
            gc[6] = c3 * gM[10] - c5 * gM[15];
            gc[7] = c1 * (gM[8] + gM[2]);
            gc[8] = c1 * (gM[0] - gM[5]);
            for (int c = 0; c < 9; c++) acc[9 * ch + c] = sc * gc[c];
        }
    }
    if (!uniformView) {
        if (pix < npix)
            for (int ch = 0; ch < 3; ch++)
                for (int c = 0; c < 9; c++) atomicAdd(gCoef + 9 * (cb > 1 ? ch + 3 * b : ch) + c, acc[9 * ch + c]);
        return;
    }
    for (int q = 0; q < 27; q++) acc[q] = wave_sum_t(acc[q]);
    int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) for (int q = 0; q < 27; q++) red[wv][q] = acc[q];
    __syncthreads();
    if (threadIdx.x < 27) {
        int q = threadIdx.x, ch = q / 9, c = q % 9;
        T s = red[0][q] + red[1][q] + red[2][q] + red[3][q];
        atomicAdd(gCoef + 9 * (cb > 1 ? ch + 3 * b0 : ch) + c, s);
    }
}

template <class T>
__global__ void k_lambert_forward(long npix, int W, int H, int L, int ib, const T* __restrict__ dirs,
                                  const T* __restrict__ ints, const T* __restrict__ nrm,
                                  const T* __restrict__ diff, T* __restrict__ out)
{
    long pix = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    long b = pix / ((long)W * H);
    const T* n = nrm + 3 * pix;
    const T* d = diff + 3 * pix;
    T acc[3] = { T(0), T(0), T(0) };
    for (int l = 0; l < L; l++) {
        const T* dir = dirs + 3 * (b * L + l);
        const T* I = ints + 3 * ((ib > 1 ? b * L : 0) + l);
        T c = dir[0] * n[0] + dir[1] * n[1] + dir[2] * n[2];
        c = fmax(T(0), c);
        for (int ch = 0; ch < 3; ch++) acc[ch] += d[ch] / T(UD_PI) * I[ch] * c;
    }
    for (int ch = 0; ch < 3; ch++) out[3 * pix + ch] = acc[ch];
}

template <class T>
__global__ void k_lambert_backward(long npix, int W, int H, int L, int ib, const T* __restrict__ g,
                                   const T* __restrict__ dirs, const T* __restrict__ ints,
                                   const T* __restrict__ nrm, const T* __restrict__ diff,
                                   T* __restrict__ gDir, T* __restrict__ gInt,
                                   T* __restrict__ gN, T* __restrict__ gD)
{
    long pix = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    long b = pix / ((long)W * H);
    const T* n = nrm + 3 * pix;
    const T* d = diff + 3 * pix;
    const T* gi = g + 3 * pix;
    T gn[3] = { T(0), T(0), T(0) }, gd[3] = { T(0), T(0), T(0) };
    for (int l = 0; l < L; l++) {
        const T* dir = dirs + 3 * (b * L + l);
        long io = 3 * ((ib > 1 ? b * L : 0) + l);
        const T* I = ints + io;
        T c = dir[0] * n[0] + dir[1] * n[1] + dir[2] * n[2];
        if (!(c > T(0))) continue;
        T dco[3] = { d[0] * I[0], d[1] * I[1], d[2] * I[2] };
        T gdc = gi[0] * dco[0] + gi[1] * dco[1] + gi[2] * dco[2];
        T sc = c / T(UD_PI);
        for (int ch = 0; ch < 3; ch++) {
            atomicAdd(gDir + 3 * (b * L + l) + ch, gdc * n[ch] / T(UD_PI));
            atomicAdd(gInt + io + ch, gi[ch] * d[ch] * sc);
            gn[ch] += gdc * dir[ch] / T(UD_PI);
            gd[ch] += gi[ch] * I[ch] * sc;
        }
    }
    for (int ch = 0; ch < 3; ch++) { gN[3 * pix + ch] = gn[ch]; gD[3 * pix + ch] = gd[ch]; }
}

__global__ void k_fill_f32(float* __restrict__ p, long n, float v)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

template <class T>
__global__ void k_fill(T* __restrict__ p, long n, T v)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

inline unsigned grid1(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

inline int launch_status()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
}

// ---------------------------------------------------------------------------
// Host bodies of the per-slot ops, shared by the float entry points and the
// float64 (and, for filtering, half) ones the reference's dispatch also accepts.
// ---------------------------------------------------------------------------
template <class T>
size_t splat_ws_bytes(int bn, int N)
{
    return (size_t)bn * N * (sizeof(SplatRecT<T>) + kMomRec * sizeof(T)) + 256;
}

template <class T>
int splat_forward_t(int bn, int N, int W, int H, int k, double f, double smoothing, const int32_t* idx,
                    const T* pts, const T* nrm, const T* sd, int sdBatch, T* out, void* ws, size_t wsBytes,
                    hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || N < 0) return UDE_ERR_ARG;
    long nslots = (long)bn * W * H * k;
    if (N == 0) {
        k_fill<T><<<grid1(nslots, 256), 256, 0, st>>>(out, nslots, T(0));
        return launch_status();
    }
    if (wsBytes < splat_ws_bytes<T>(bn, N)) return UDE_ERR_WORKSPACE;
    SplatRecT<T>* rec = (SplatRecT<T>*)ws;
    k_splat_prepare<T><<<grid1((long)bn * N, 256), 256, 0, st>>>(bn, N, W, H, f, smoothing, pts, nrm, sd, sdBatch,
                                                                 rec);
    k_splat_forward<T><<<grid1(nslots, 256), 256, 0, st>>>(nslots, N, W, H, k, idx, rec, out);
    return launch_status();
}

template <class T>
int splat_backward_t(int bn, int N, int W, int H, int k, double f, double smoothing, const T* gradIn,
                     const int32_t* idx, const T* pts, const T* nrm, const T* sd, int sdBatch, T* gPts, T* gNrm,
                     T* gSd, void* ws, size_t wsBytes, hipStream_t st, bool post = false)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || N < 0) return UDE_ERR_ARG;
    if (post && sdBatch != bn) return UDE_ERR_ARG;   // a shared std-dev row is a sum over views
    if (N == 0) return UDE_OK;
    if (wsBytes < splat_ws_bytes<T>(bn, N)) return UDE_ERR_WORKSPACE;
    long nslots = (long)bn * W * H * k;
    SplatRecT<T>* rec = (SplatRecT<T>*)ws;
    T* mom = (T*)(rec + (long)bn * N);
    k_splat_prepare<T><<<grid1((long)bn * N, 256), 256, 0, st>>>(bn, N, W, H, f, smoothing, pts, nrm, sd, sdBatch,
                                                                 rec);
    k_fill<T><<<grid1((long)kMomRec * bn * N, 256), 256, 0, st>>>(mom, (long)kMomRec * bn * N, T(0));
    k_splat_moments<T><<<grid1(nslots, 256), 256, 0, st>>>(nslots, N, W, H, k, idx, gradIn, rec, mom);
    if (post)
        k_splat_finalize<T, true><<<grid1((long)bn * N, 128), 128, 0, st>>>(bn, N, W, H, f, smoothing, pts, nrm, sd,
                                                                            sdBatch, rec, mom, kMomRec, 1, nullptr,
                                                                            gPts, gNrm, gSd);
    else
        k_splat_finalize<T><<<grid1((long)bn * N, 128), 128, 0, st>>>(bn, N, W, H, f, smoothing, pts, nrm, sd,
                                                                      sdBatch, rec, mom, kMomRec, 1, nullptr, gPts,
                                                                      gNrm, gSd);
    return launch_status();
}

template <bool BWD, class S>
int filter_t(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const S* wts, const S* gin,
             const S* pts, const S* nrm, int pb, S* out, hipStream_t st, bool exact = false)
{
    using C = typename CompOf<S>::type;
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || k > 1024 || (pb != 1 && pb != bn)) return UDE_ERR_ARG;
    long npix = (long)bn * W * H;
    const C gate = exact ? C(-1) : C(1e-4);
    if (k <= 64)
        k_filter_wave<BWD, S><<<grid1(npix, 4), 256, 0, st>>>(npix, N, W, H, k, (C)f, gate, idx, wts, gin, pts, nrm,
                                                               pb, out);
    else
        k_filter_block<BWD, S><<<(unsigned)npix, ((k + 63) / 64) * 64, 0, st>>>(npix, N, W, H, k, (C)f, gate, idx,
                                                                                 wts, gin, pts, nrm, pb, out);
    return launch_status();
}

template <class T>
int transform_forward_t(int affine, int bn, int vb, int N, const T* M, const T* v, T* out, T* facing,
                        hipStream_t st)
{
    if (bn <= 0 || bn > 65535 || N < 0 || (vb != 1 && vb != bn)) return UDE_ERR_ARG;
    if (N == 0) return UDE_OK;
    k_transform_forward<T><<<dim3(grid1(N, 256), bn), 256, 0, st>>>(affine, N, vb, M, v, out, facing);
    return launch_status();
}

template <class T>
int transform_backward_t(int affine, int bn, int vb, int N, const T* g, const T* M, const T* v, T* gM, T* gV,
                         hipStream_t st)
{
    if (bn <= 0 || bn > 65535 || N < 0 || (vb != 1 && vb != bn)) return UDE_ERR_ARG;
    if (N == 0) return UDE_OK;
    const unsigned blocks = grid1(N, 1024) < 256u ? grid1(N, 1024) : 256u;
    dim3 grid(blocks, bn);
    k_transform_backward<T><<<grid, 1024, 0, st>>>(affine, bn, vb, N, g, M, v, gM, gV);
    return launch_status();
}

template <class T>
int interp_forward_t(int bn, int W, int H, int k, int ab, int N, int C, const int32_t* idx, const T* wts,
                     const T* attr, T* out, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || C <= 0 || (ab != 1 && ab != bn)) return UDE_ERR_ARG;
    long npix = (long)bn * W * H;
    const int vec4 = (k % 4 == 0) && ((uintptr_t)idx % 16 == 0) && ((uintptr_t)wts % 16 == 0);
    if constexpr (sizeof(T) == 4) {
        if (C == 3 && k <= kInterpTileMaxK && ((k & 3) != 0 || vec4)) {
            const size_t lds = sizeof(float) * 2 * 64 * (size_t)(k + 1);
            k_interp_forward_tile<<<(unsigned)((npix + 63) / 64), 64, lds, st>>>(
                npix, W, H, k, ab, N, idx, (const float*)wts, (const float*)attr, (float*)out);
            return launch_status();
        }
    }
    k_interp_forward<T><<<grid1(npix, 256), 256, 0, st>>>(npix, W, H, k, ab, N, C, vec4, idx, wts, attr, out);
    return launch_status();
}

template <class T>
int interp_backward_t(int bn, int W, int H, int k, int ab, int N, int C, const T* g, const int32_t* idx,
                      const T* wts, const T* attr, T* gW, T* gA, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || C <= 0 || (ab != 1 && ab != bn)) return UDE_ERR_ARG;
    long nslots = (long)bn * W * H * k;
    k_interp_backward<T><<<grid1(nslots, 256), 256, 0, st>>>(nslots, W, H, k, ab, N, C, g, idx, wts, attr, gW, gA);
    return launch_status();
}

template <class T>
int sh_forward_t(int bn, int W, int H, int cb, const T* coef, const T* nrm, const T* diff, T* out, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || (cb != 1 && cb != bn)) return UDE_ERR_ARG;
    long npix = (long)bn * W * H;
    k_sh_forward<T><<<grid1(npix, 256), 256, 0, st>>>(npix, W, H, cb, coef, nrm, diff, out);
    return launch_status();
}

template <class T>
int sh_backward_t(int bn, int W, int H, int cb, const T* g, const T* coef, const T* nrm, const T* diff, T* gCoef,
                  T* gN, T* gD, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || (cb != 1 && cb != bn)) return UDE_ERR_ARG;
    long npix = (long)bn * W * H;
    k_sh_backward<T><<<grid1(npix, 256), 256, 0, st>>>(npix, W, H, cb, g, coef, nrm, diff, gCoef, gN, gD);
    return launch_status();
}

template <class T>
int lambert_forward_t(int bn, int W, int H, int L, int ib, const T* dirs, const T* ints, const T* nrm,
                      const T* diff, T* out, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || L < 0 || (ib != 1 && ib != bn)) return UDE_ERR_ARG;
    long npix = (long)bn * W * H;
    k_lambert_forward<T><<<grid1(npix, 256), 256, 0, st>>>(npix, W, H, L, ib, dirs, ints, nrm, diff, out);
    return launch_status();
}

template <class T>
int lambert_backward_t(int bn, int W, int H, int L, int ib, const T* g, const T* dirs, const T* ints, const T* nrm,
                       const T* diff, T* gDir, T* gInt, T* gN, T* gD, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || L < 0 || (ib != 1 && ib != bn)) return UDE_ERR_ARG;
    long npix = (long)bn * W * H;
    k_lambert_backward<T><<<grid1(npix, 256), 256, 0, st>>>(npix, W, H, L, ib, g, dirs, ints, nrm, diff, gDir, gInt,
                                                            gN, gD);
    return launch_status();
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

size_t ude_splat_workspace_bytes(int bn, int N) { return splat_ws_bytes<float>(bn, N); }
size_t ude_splat_workspace_bytes_f64(int bn, int N) { return splat_ws_bytes<double>(bn, N); }

int ude_splat_forward(int bn, int N, int W, int H, int k, double f, double smoothing,
                      const int32_t* idx, const float* pts, const float* nrm, const float* sd, int sdBatch,
                      float* out, void* ws, size_t wsBytes, hipStream_t st)
{
    return splat_forward_t<float>(bn, N, W, H, k, f, smoothing, idx, pts, nrm, sd, sdBatch, out, ws, wsBytes, st);
}

int ude_splat_forward_f64(int bn, int N, int W, int H, int k, double f, double smoothing,
                          const int32_t* idx, const double* pts, const double* nrm, const double* sd, int sdBatch,
                          double* out, void* ws, size_t wsBytes, hipStream_t st)
{
    return splat_forward_t<double>(bn, N, W, H, k, f, smoothing, idx, pts, nrm, sd, sdBatch, out, ws, wsBytes, st);
}

int ude_splat_backward(int bn, int N, int W, int H, int k, double f, double smoothing,
                       const float* gradIn, const int32_t* idx, const float* pts, const float* nrm,
                       const float* sd, int sdBatch, float* gPts, float* gNrm, float* gSd,
                       void* ws, size_t wsBytes, hipStream_t st)
{
    return splat_backward_t<float>(bn, N, W, H, k, f, smoothing, gradIn, idx, pts, nrm, sd, sdBatch, gPts, gNrm,
                                   gSd, ws, wsBytes, st);
}

int ude_splat_backward_f64(int bn, int N, int W, int H, int k, double f, double smoothing,
                           const double* gradIn, const int32_t* idx, const double* pts, const double* nrm,
                           const double* sd, int sdBatch, double* gPts, double* gNrm, double* gSd,
                           void* ws, size_t wsBytes, hipStream_t st)
{
    return splat_backward_t<double>(bn, N, W, H, k, f, smoothing, gradIn, idx, pts, nrm, sd, sdBatch, gPts, gNrm,
                                    gSd, ws, wsBytes, st);
}

int ude_splat_backward_normalized(int bn, int N, int W, int H, int k, double f, double smoothing,
                                  const float* gradIn, const int32_t* idx, const float* pts, const float* nrm,
                                  const float* sd, float* gPts, float* gNrm, float* gSd, void* ws, size_t wsBytes,
                                  hipStream_t st)
{
    return splat_backward_t<float>(bn, N, W, H, k, f, smoothing, gradIn, idx, pts, nrm, sd, bn, gPts, gNrm, gSd,
                                   ws, wsBytes, st, true);
}

int ude_splat_backward_normalized_f64(int bn, int N, int W, int H, int k, double f, double smoothing,
                                      const double* gradIn, const int32_t* idx, const double* pts,
                                      const double* nrm, const double* sd, double* gPts, double* gNrm,
                                      double* gSd, void* ws, size_t wsBytes, hipStream_t st)
{
    return splat_backward_t<double>(bn, N, W, H, k, f, smoothing, gradIn, idx, pts, nrm, sd, bn, gPts, gNrm, gSd,
                                    ws, wsBytes, st, true);
}

int ude_splat_backward_static(int bn, int N, int W, int H, int k, double f, double smoothing,
                              const float* gradIn, const int32_t* idx, const float* pts, const float* nrm,
                              const float* sd, int sdBatch, float* gPts, float* gNrm, float* gSd, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || N < 0) return UDE_ERR_ARG;
    if (N == 0) return UDE_OK;
    const size_t need = ude_splat_workspace_bytes(bn, N);
    if (need > (size_t)UDE_STATIC_SCRATCH_BYTES) return UDE_ERR_SCRATCH;
    void* ws = scratch_on_current_device();
    if (!ws) return UDE_ERR_LAUNCH;
    // the kernels of ude_splat_backward on the static scratch; a fill kernel clears the
    // moments, so the previous call's contents never leak into this one
    return splat_backward_t<float>(bn, N, W, H, k, f, smoothing, gradIn, idx, pts, nrm, sd, sdBatch, gPts, gNrm,
                                   gSd, ws, (size_t)UDE_STATIC_SCRATCH_BYTES, st);
}

int ude_covariance(int bn, int N, int W, int H, double f, double smoothing, const float* pts,
                   const float* nrm, const float* sd, int sdBatch, float* out, hipStream_t st)
{
    if (bn <= 0 || N < 0) return UDE_ERR_ARG;
    if (N == 0) return UDE_OK;
    k_covariance<<<grid1((long)bn * N, 256), 256, 0, st>>>(bn, N, W, H, f, smoothing, pts, nrm, sd, sdBatch, out);
    return launch_status();
}

int ude_filter_forward(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const float* wts,
                       const float* pts, const float* nrm, int pb, float* out, hipStream_t st)
{
    return filter_t<false, float>(bn, N, W, H, k, f, idx, wts, nullptr, pts, nrm, pb, out, st);
}

int ude_filter_forward_f64(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const double* wts,
                           const double* pts, const double* nrm, int pb, double* out, hipStream_t st)
{
    return filter_t<false, double>(bn, N, W, H, k, f, idx, wts, nullptr, pts, nrm, pb, out, st);
}

int ude_filter_forward_f16(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const void* wts,
                           const void* pts, const void* nrm, int pb, void* out, hipStream_t st)
{
    return filter_t<false, __half>(bn, N, W, H, k, f, idx, (const __half*)wts, nullptr, (const __half*)pts,
                                   (const __half*)nrm, pb, (__half*)out, st);
}

int ude_filter_backward(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const float* wts,
                        const float* gin, const float* pts, const float* nrm, int pb, float* out, hipStream_t st)
{
    return filter_t<true, float>(bn, N, W, H, k, f, idx, wts, gin, pts, nrm, pb, out, st);
}

int ude_filter_backward_f64(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const double* wts,
                            const double* gin, const double* pts, const double* nrm, int pb, double* out,
                            hipStream_t st)
{
    return filter_t<true, double>(bn, N, W, H, k, f, idx, wts, gin, pts, nrm, pb, out, st);
}

int ude_filter_backward_exact(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const float* wts,
                              const float* gin, const float* pts, const float* nrm, int pb, float* out,
                              hipStream_t st)
{
    return filter_t<true, float>(bn, N, W, H, k, f, idx, wts, gin, pts, nrm, pb, out, st, true);
}

int ude_filter_backward_exact_f64(int bn, int N, int W, int H, int k, double f, const int32_t* idx,
                                  const double* wts, const double* gin, const double* pts, const double* nrm, int pb,
                                  double* out, hipStream_t st)
{
    return filter_t<true, double>(bn, N, W, H, k, f, idx, wts, gin, pts, nrm, pb, out, st, true);
}

int ude_filter_backward_f16(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const void* wts,
                            const void* gin, const void* pts, const void* nrm, int pb, void* out, hipStream_t st)
{
    return filter_t<true, __half>(bn, N, W, H, k, f, idx, (const __half*)wts, (const __half*)gin,
                                  (const __half*)pts, (const __half*)nrm, pb, (__half*)out, st);
}

int ude_transform_forward(int affine, int bn, int vb, int N, const float* M, const float* v, float* out, hipStream_t st)
{
    return transform_forward_t<float>(affine, bn, vb, N, M, v, out, nullptr, st);
}

int ude_transform_forward_f64(int affine, int bn, int vb, int N, const double* M, const double* v, double* out,
                              hipStream_t st)
{
    return transform_forward_t<double>(affine, bn, vb, N, M, v, out, nullptr, st);
}

int ude_transform_forward_facing(int bn, int vb, int N, const float* M, const float* v, float* out, float* facing,
                                 hipStream_t st)
{
    if (!facing) return UDE_ERR_ARG;
    return transform_forward_t<float>(0, bn, vb, N, M, v, out, facing, st);
}

int ude_transform_backward(int affine, int bn, int vb, int N, const float* g, const float* M, const float* v,
                           float* gM, float* gV, hipStream_t st)
{
    return transform_backward_t<float>(affine, bn, vb, N, g, M, v, gM, gV, st);
}

int ude_transform_backward_f64(int affine, int bn, int vb, int N, const double* g, const double* M,
                               const double* v, double* gM, double* gV, hipStream_t st)
{
    return transform_backward_t<double>(affine, bn, vb, N, g, M, v, gM, gV, st);
}

int ude_interp_forward(int bn, int W, int H, int k, int ab, int N, int C, const int32_t* idx, const float* wts,
                       const float* attr, float* out, hipStream_t st)
{
    return interp_forward_t<float>(bn, W, H, k, ab, N, C, idx, wts, attr, out, st);
}

int ude_interp_forward_f64(int bn, int W, int H, int k, int ab, int N, int C, const int32_t* idx, const double* wts,
                           const double* attr, double* out, hipStream_t st)
{
    return interp_forward_t<double>(bn, W, H, k, ab, N, C, idx, wts, attr, out, st);
}

int ude_interp_backward(int bn, int W, int H, int k, int ab, int N, int C, const float* g, const int32_t* idx,
                        const float* wts, const float* attr, float* gW, float* gA, hipStream_t st)
{
    return interp_backward_t<float>(bn, W, H, k, ab, N, C, g, idx, wts, attr, gW, gA, st);
}

int ude_interp_backward_f64(int bn, int W, int H, int k, int ab, int N, int C, const double* g,
                            const int32_t* idx, const double* wts, const double* attr, double* gW, double* gA,
                            hipStream_t st)
{
    return interp_backward_t<double>(bn, W, H, k, ab, N, C, g, idx, wts, attr, gW, gA, st);
}

size_t ude_ht_workspace_bytes(int bn, int N)
{
    return (size_t)bn * N * (sizeof(SplatRec) + kHtRec * sizeof(float)) + 512;
}

// backward records after the splat records, on a 128 B boundary (one record per 64 B half-line)
static float* ht_records(void* ws, int bn, int N)
{
    const uintptr_t a = (uintptr_t)((SplatRec*)ws + (long)bn * N);
    return (float*)((a + 127) & ~(uintptr_t)127);
}

int ude_ht_forward(int bn, int N, int W, int H, int k, double f, double smoothing, const int32_t* idx,
                   const float* ipi, const float* pts, const float* nrm, const float* sd, int sdBatch, int ab,
                   const float* attr, float* out, void* ws, size_t wsBytes, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || N < 0 || (ab != 1 && ab != bn)) return UDE_ERR_ARG;
    const long npix = (long)bn * W * H;
    if (N == 0) {
        k_fill_f32<<<grid1(3 * npix, 256), 256, 0, st>>>(out, 3 * npix, 0.0f);
        return launch_status();
    }
    if (wsBytes < ude_ht_workspace_bytes(bn, N)) return UDE_ERR_WORKSPACE;
    SplatRec* rec = (SplatRec*)ws;
    k_splat_prepare<float><<<grid1((long)bn * N, 256), 256, 0, st>>>(bn, N, W, H, f, smoothing, pts, nrm, sd, sdBatch, rec);
    const int vec4 = (k % 4 == 0) && ((uintptr_t)idx % 16 == 0) && ((uintptr_t)ipi % 16 == 0);
    k_ht_forward<<<grid1(npix, 256), 256, 0, st>>>(npix, N, W, H, k, ab, vec4, idx, ipi, rec, attr, out);
    return launch_status();
}

int ude_ht_backward(int bn, int N, int W, int H, int k, double f, double smoothing, const float* gradImage,
                    const int32_t* idx, const float* ipi, const float* pts, const float* nrm, const float* sd,
                    int sdBatch, int ab, const float* attr, float* gPts, float* gNrm, float* gSd, float* gAttr,
                    int prepared, void* ws, size_t wsBytes, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || N < 0 || (ab != 1 && ab != bn)) return UDE_ERR_ARG;
    if (N == 0) return UDE_OK;
    if (wsBytes < ude_ht_workspace_bytes(bn, N)) return UDE_ERR_WORKSPACE;
    SplatRec* rec = (SplatRec*)ws;
    float* mom = ht_records(ws, bn, N);
    if (!prepared)
        k_splat_prepare<float><<<grid1((long)bn * N, 256), 256, 0, st>>>(bn, N, W, H, f, smoothing, pts, nrm, sd, sdBatch, rec);
    k_fill_f32<<<grid1((long)kHtRec * bn * N, 256), 256, 0, st>>>(mom, (long)kHtRec * bn * N, 0.0f);
    const dim3 tiles((unsigned)(((W + kHtTile - 1) / kHtTile) * ((H + kHtTile - 1) / kHtTile)), (unsigned)bn);
    const int vec4 = (k % 4 == 0) && ((uintptr_t)idx % 16 == 0) && ((uintptr_t)ipi % 16 == 0);
    k_ht_backward<<<tiles, kHtTile * kHtTile, 0, st>>>(N, W, H, k, ab, vec4, gradImage, idx, ipi, rec, attr, nullptr,
                                                        mom);
    k_splat_finalize<float><<<grid1((long)bn * N, 128), 128, 0, st>>>(bn, N, W, H, f, smoothing, pts, nrm, sd, sdBatch,
                                                              rec, mom, kHtRec, ab, gAttr, gPts, gNrm, gSd);
    return launch_status();
}

// Composited estimate: the fused estimator's workspace, then the plane records of
// the depth order and the per-slot compositing factors of the backward
size_t ude_htc_workspace_bytes(int bn, int N, int W, int H, int k)
{
    return ude_ht_workspace_bytes(bn, N) + (size_t)bn * N * sizeof(PlaneRec) +
           (size_t)bn * W * H * k * sizeof(float2) + 256;
}

static PlaneRec* htc_planes(void* ws, int bn, int N)
{
    return (PlaneRec*)(ht_records(ws, bn, N) + (long)kHtRec * bn * N);
}

static int htc_check(int bn, int N, int W, int H, int k, int ab)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || k > kHtcMaxK || N < 0 || (ab != 1 && ab != bn)) return UDE_ERR_ARG;
    return UDE_OK;
}

int ude_htc_forward(int bn, int N, int W, int H, int k, double f, double smoothing, const int32_t* idx,
                    const float* ipi, const float* pts, const float* nrm, const float* sd, int sdBatch, int ab,
                    const float* attr, float* out, void* ws, size_t wsBytes, hipStream_t st)
{
    if (htc_check(bn, N, W, H, k, ab) != UDE_OK) return UDE_ERR_ARG;
    const long npix = (long)bn * W * H;
    if (N == 0) {
        k_fill_f32<<<grid1(3 * npix, 256), 256, 0, st>>>(out, 3 * npix, 0.0f);
        return launch_status();
    }
    if (wsBytes < ude_htc_workspace_bytes(bn, N, W, H, k)) return UDE_ERR_WORKSPACE;
    SplatRec* rec = (SplatRec*)ws;
    PlaneRec* pl = htc_planes(ws, bn, N);
    k_splat_prepare<float><<<grid1((long)bn * N, 256), 256, 0, st>>>(bn, N, W, H, f, smoothing, pts, nrm, sd, sdBatch, rec);
    k_plane_prepare<<<grid1((long)bn * N, 256), 256, 0, st>>>((long)bn * N, pts, nrm, pl);
    k_htc<false><<<grid1(npix, 64), 64, (size_t)k * 64 * 9, st>>>(npix, N, W, H, k, ab, (float)f, idx, ipi, rec, pl,
                                                                  attr, nullptr, out, nullptr);
    return launch_status();
}

int ude_htc_backward(int bn, int N, int W, int H, int k, double f, double smoothing, const float* gradImage,
                     const int32_t* idx, const float* ipi, const float* pts, const float* nrm, const float* sd,
                     int sdBatch, int ab, const float* attr, float* gPts, float* gNrm, float* gSd, float* gAttr,
                     int prepared, void* ws, size_t wsBytes, hipStream_t st)
{
    if (htc_check(bn, N, W, H, k, ab) != UDE_OK) return UDE_ERR_ARG;
    if (N == 0) return UDE_OK;
    if (wsBytes < ude_htc_workspace_bytes(bn, N, W, H, k)) return UDE_ERR_WORKSPACE;
    const long npix = (long)bn * W * H;
    SplatRec* rec = (SplatRec*)ws;
    float* mom = ht_records(ws, bn, N);
    PlaneRec* pl = htc_planes(ws, bn, N);
    float2* coef = (float2*)(pl + (long)bn * N);
    if (!prepared) {
        k_splat_prepare<float><<<grid1((long)bn * N, 256), 256, 0, st>>>(bn, N, W, H, f, smoothing, pts, nrm, sd, sdBatch, rec);
        k_plane_prepare<<<grid1((long)bn * N, 256), 256, 0, st>>>((long)bn * N, pts, nrm, pl);
    }
    k_htc<true><<<grid1(npix, 64), 64, (size_t)k * 64 * 9, st>>>(npix, N, W, H, k, ab, (float)f, idx, ipi, rec, pl,
                                                                 attr, gradImage, nullptr, coef);
    k_fill_f32<<<grid1((long)kHtRec * bn * N, 256), 256, 0, st>>>(mom, (long)kHtRec * bn * N, 0.0f);
    const dim3 tiles((unsigned)(((W + kHtTile - 1) / kHtTile) * ((H + kHtTile - 1) / kHtTile)), (unsigned)bn);
    const int vec4 = (k % 4 == 0) && ((uintptr_t)idx % 16 == 0) && ((uintptr_t)ipi % 16 == 0);
    k_ht_backward<<<tiles, kHtTile * kHtTile, 0, st>>>(N, W, H, k, ab, vec4, gradImage, idx, ipi, rec, attr, coef,
                                                        mom);
    k_splat_finalize<float><<<grid1((long)bn * N, 128), 128, 0, st>>>(bn, N, W, H, f, smoothing, pts, nrm, sd, sdBatch,
                                                              rec, mom, kHtRec, ab, gAttr, gPts, gNrm, gSd);
    return launch_status();
}

int ude_l2_grad(long n, const float* image, const float* target, float* grad, double* sums, hipStream_t st)
{
    if (n < 0) return UDE_ERR_ARG;
    if (n == 0) {
        k_sum_parts<<<1, 256, 0, st>>>(0, sums + 1, sums);
        return launch_status();
    }
    const unsigned blocks = grid1(n, 256) < (unsigned)kL2Blocks ? grid1(n, 256) : (unsigned)kL2Blocks;
    k_l2_grad<<<blocks, 256, 0, st>>>(n, image, target, grad, sums + 1);
    k_sum_parts<<<1, 256, 0, st>>>((int)blocks, sums + 1, sums);
    return launch_status();
}

int ude_sh_forward(int bn, int W, int H, int cb, const float* coef, const float* nrm, const float* diff,
                   float* out, hipStream_t st)
{
    return sh_forward_t<float>(bn, W, H, cb, coef, nrm, diff, out, st);
}

int ude_sh_forward_f64(int bn, int W, int H, int cb, const double* coef, const double* nrm, const double* diff,
                       double* out, hipStream_t st)
{
    return sh_forward_t<double>(bn, W, H, cb, coef, nrm, diff, out, st);
}

int ude_sh_backward(int bn, int W, int H, int cb, const float* g, const float* coef, const float* nrm,
                    const float* diff, float* gCoef, float* gN, float* gD, hipStream_t st)
{
    return sh_backward_t<float>(bn, W, H, cb, g, coef, nrm, diff, gCoef, gN, gD, st);
}

int ude_sh_backward_f64(int bn, int W, int H, int cb, const double* g, const double* coef, const double* nrm,
                        const double* diff, double* gCoef, double* gN, double* gD, hipStream_t st)
{
    return sh_backward_t<double>(bn, W, H, cb, g, coef, nrm, diff, gCoef, gN, gD, st);
}

int ude_lambert_forward(int bn, int W, int H, int L, int ib, const float* dirs, const float* ints,
                        const float* nrm, const float* diff, float* out, hipStream_t st)
{
    return lambert_forward_t<float>(bn, W, H, L, ib, dirs, ints, nrm, diff, out, st);
}

int ude_lambert_forward_f64(int bn, int W, int H, int L, int ib, const double* dirs, const double* ints,
                            const double* nrm, const double* diff, double* out, hipStream_t st)
{
    return lambert_forward_t<double>(bn, W, H, L, ib, dirs, ints, nrm, diff, out, st);
}

int ude_lambert_backward(int bn, int W, int H, int L, int ib, const float* g, const float* dirs, const float* ints,
                         const float* nrm, const float* diff, float* gDir, float* gInt, float* gN, float* gD,
                         hipStream_t st)
{
    return lambert_backward_t<float>(bn, W, H, L, ib, g, dirs, ints, nrm, diff, gDir, gInt, gN, gD, st);
}

int ude_lambert_backward_f64(int bn, int W, int H, int L, int ib, const double* g, const double* dirs,
                             const double* ints, const double* nrm, const double* diff, double* gDir, double* gInt,
                             double* gN, double* gD, hipStream_t st)
{
    return lambert_backward_t<double>(bn, W, H, L, ib, g, dirs, ints, nrm, diff, gDir, gInt, gN, gD, st);
}

} // extern "C"
