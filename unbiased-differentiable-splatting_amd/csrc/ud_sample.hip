// ud_sample.hip -- the sampling module on gfx950: LBVH build, Hermite FGT,
// the per-pixel sequential-Poisson sampler (sampling.sample; reference:
// samplingCuda.cu:197-290, buildDynamicRadixTree.cuh,
// evaluateDynamicRadixTree.cuh, sequentialPoissonDynamicTree.cuh), the
// fixed-order densities (approximateFMMWeights; transformRadixTree.cuh) and
// Poisson-rate sampling with inclusion probabilities (no reference kernel).
//
// Every stage runs on the device with sizes kept in device memory, so a view
// costs no host synchronisation:
//   1 k_bandwidth     per-splat isotropic h_k -> fixed-point sum (2^-40 units)
//   2 k_view_params   h, query radius, screen bounds
//   3 k_cull_*        order-preserving compaction of visible splats + Morton
//   4 k_sort_*        3-pass LSD radix sort of 20-bit Morton keys (stable,
//                     wavefront ballot ranking)
//   5 k_hierarchy     Karras 2012 split on 64-bit keys (morton << 32 | pos)
//   6 k_bottom_up     atomic fan-in: boxes, counts, centres, Hermite degree
//   7 k_coef_*        coefficient slots (deg+1)^2 of the reachable expanded
//                     nodes, device-wide exclusive scan
//   8 k_moments       Hermite moments per node over its contiguous leaf range,
//                     accumulated in FP64 in leaf order
//   9 k_sample_walk   per pixel tile: union FGT walk, capacity descents
//  10 k_sample_draw   per pixel: sequential Poisson keys u/p from
//                     Philox4x32-10, top-k in LDS
// The fixed-order and Poisson-rate paths share stages 1-6 and add expansion
// selection and moments per band (k_fixed_*), walk records (k_fmm_records,
// k_walk_records), the density walk (k_fmm_eval) and the thinning pass
// (k_poisson).  Batches of views run each stage as one launch (views on
// blockIdx.y) where the workspace holds them all.
// All FP32 arithmetic on the tree and in the samplers follows a fixed operation
// order (no contraction) so that index sets are reproducible bit for bit.
#include "ud_device.h"
#include "../../include/udsplat.h"

#include <utility>
#include <vector>

using namespace ud;

namespace {

constexpr int kMaxDegree = 16;
constexpr int kMaxDepth = 64;        // descent trace bound (reference STACKSIZE)
// runs of consecutive 8x8 tiles dealt to each XCD in turn by the sampler's
// passes (k_sample_walk / draw / replay): their cost follows the splat density,
// so each XCD takes runs from the whole image (4 tiles: walk + draw 2.73 -> 2.59 ms
// at 1M splats, 512x512, against one contiguous block of tiles per XCD)
constexpr uint32_t kSampTileRun = 4;
constexpr int kTrieFields = 5;       // 4-byte fields per descent-trie entry (TrieRef)
constexpr int kCullTile = 1024;      // elements per workgroup in the cull pass
constexpr int kSortTile = 2048;      // elements per workgroup in a sort pass
constexpr int kFixedDegree = 6;      // default expansion order of the fixed-order densities (hermitefmm.cuh:11)
constexpr int kFixedPMin = 4, kFixedPMax = 8;              // orders ude_set_fixed_degree accepts
constexpr int kFixedCoefsMax = (kFixedPMax + 1) * (kFixedPMax + 1);
constexpr int kScanTile = 1024;      // elements per workgroup in a scan pass

// Gaussian bandwidth of a fixed-order density: h, 1 / sqrt(2 h^2), the leaf
// coefficient 1 / (2 pi sqrt(h)) and the query radius
struct Band {
    float h, is2h, c, radius;
};

struct ViewState {
    unsigned long long hsum;
    unsigned int hcnt;
    int nc;
    float h, radius, lo, hix, hiy, s2h, prec, is2h;
    int err;
    int ncoef;
    int nseeds;       // block-spanning fan-in starts (bottom-up, second level)
    int nmom;         // internal nodes with a Hermite expansion
    int pad;
    Band band[2];     // fixed-order densities: tree bandwidth, wide proposal bandwidth
    int nfix[2];      // accepted internal nodes per band
    int hhist[256];   // visible splats' proposal bandwidths, 1/16 px bins over [0, 16)
};

struct NodeGeo {      // traversal record
    float lox, loy, hix, hiy;
    float cx, cy, r2;
    int deg;
};
struct NodeLnk {
    int left, right, parent, count;
};
struct NodeAux {
    int off, first, last, pad;   // pad: a leaf's splat id
};
// per internal node: its two children as the sampler's walk reads them (one
// record per descent; a node's own fields arrive with its parent's record)
struct SampChild {
    float lox, loy, hix, hiy;    // box
    float cx, cy, r2;            // expansion centre, squared truncation radius
    int deg;                     // Hermite degree, -1 when the walk goes below
    int off, count, id, pad;     // coefficient offset, leaves under it, node index
};
struct SampRec {
    SampChild l, r;
};
// per node: what the per-pixel Poisson walk reads of it in one load
struct PNode {
    float lox, loy, hix, hiy;    // box
    int left;                    // left child, -1 for leaves
    int esc;                     // next node in preorder after the subtree, -1 after the last
    int first, last;             // sorted leaf range
};
// per sorted leaf: the footprint proposal's record (foot_f32)
struct FootLeaf {
    float x, y;                  // screen position (the tree's)
    float t00, t01, t11, amp;    // whitening and peak of the EWA weight
    float pad[2];
};
// the footprint term is capped by one envelope for every splat, E(d) =
// exp(-max(0, d - kFootRcap)^2 a1) with a1 = 1 / 2h_1^2, the wide band's decay:
// min(1, w) equals w where w is below E, which holds beyond kFootRcap pixels
// for every splat whose largest axis is within h_1 and whose core min(1, w) = 1
// reaches at most kFootRcap; for the others (edge-on needles, the widest 0.1 %)
// the cap trims the footprint term and the wide band's half still covers them.
// A node's bound is then E at its box distance, with no per-node data.
constexpr float kFootRcap = 2.0f;

// per internal node: what the fixed-order density walk needs of its two
// children, so that one wave-uniform load serves a descent
struct FmmRec {
    float lbox[4], rbox[4];      // children's boxes (lo.x, lo.y, hi.x, hi.y)
    float lc[2], rc[2];          // children's centres
    int left, right;
    int lfix[2], rfix[2];        // children's expansion offsets per band, -1 for none / leaves
    int pad[2];
};

struct TreeWs {
    ViewState* vs;
    uint32_t *keysA, *keysB;
    int32_t *idsA, *idsB;
    int *cullCounts, *sortHist, *scanTmp, *scanBlocks;
    NodeGeo* geo;
    NodeLnk* lnk;
    NodeAux* aux;
    int* flags;
    int* seeds;
    int* momList;
    float* trie;
    SampRec* srec;        // per internal node: children records of the sampler's walk
    int2* reach;          // per launch and pixel lane: leaf range the capacity descent reached
    int* tieQueue;                    // [0]: count, then (sampler lane, final key bits) per lane that met equal keys
    float2* leafXY;       // sorted position -> screen position of the leaf
    float* coef;
    long coefCap;
    // fixed-order densities (approximateFMMWeights, Poisson-rate sampling)
    int* fixOff[2];       // per internal node and band: coefficient offset, -1 = no expansion
    int* fixList[2];      // per band: accepted internal nodes
    float* fixCoef[2];    // per band: 49 coefficients per accepted node
    float* hraw;          // per splat id: proposal bandwidth sqrt(lambda_max(V))
    PNode* pnode;         // per node: Poisson walk record
    FootLeaf* footLeaf;   // per sorted leaf: footprint proposal record
    double* footSum;      // sum of the visible splats' footprint masses
    FmmRec* frec;         // per internal node: children for the density walk
    float* dens[2];       // per pixel and band: density D_b(q)
};

enum WsMode { kModeSample = 0, kModeFixed = 1, kModePoisson = 2 };

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int trie_entries(int k);

// carve the workspace; returns total bytes (ws may be null to only measure)
// mode kModeSample: dynamic degrees + descent trie; kModeFixed: degree-6 moments of
// the accepted internal nodes only; kModePoisson: fixed tree + proposal bandwidths
// + density image
size_t carve(void* ws, int N, int W, int H, int k, int mode, TreeWs* t)
{
    long n = N > 0 ? N : 1;
    long nn = 2 * n;
    long nCull = (n + kCullTile - 1) / kCullTile;
    long nSort = (n + kSortTile - 1) / kSortTile;
    long nScan = (nn + kScanTile - 1) / kScanTile;
    const bool fixedOrder = mode != kModeSample;
    long coefCap = (long)(kMaxDegree + 1) * (kMaxDegree + 1) * (n - 1 > 0 ? n - 1 : 1) + n;
    size_t off = 0;
    char* base = (char*)ws;
    auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return base ? (void*)(base + o) : nullptr; };
    TreeWs w;
    w.vs = (ViewState*)take(sizeof(ViewState));
    w.keysA = (uint32_t*)take(4 * n);
    w.keysB = (uint32_t*)take(4 * n);
    w.idsA = (int32_t*)take(4 * n);
    w.idsB = (int32_t*)take(4 * n);
    w.cullCounts = (int*)take(4 * (nCull + 1));
    w.sortHist = (int*)take(4 * 256 * (nSort + 1));
    w.scanTmp = (int*)take(4 * nn);
    w.scanBlocks = (int*)take(4 * (nScan + 1) > 4 * (256 * (nSort + 1)) ? 4 * (nScan + 1) : 4 * 256 * (nSort + 1));
    w.geo = (NodeGeo*)take(sizeof(NodeGeo) * nn);
    w.lnk = (NodeLnk*)take(sizeof(NodeLnk) * nn);
    w.aux = (NodeAux*)take(sizeof(NodeAux) * nn);
    w.flags = (int*)take(4 * n);
    w.seeds = (int*)take(4 * n);
    w.momList = (int*)take(4 * n);
    long lanes = (long)((W + 7) / 8) * ((H + 7) / 8) * 64;
    w.trie = fixedOrder ? nullptr : (float*)take(4L * kTrieFields * trie_entries(k) * lanes);
    w.srec = fixedOrder ? nullptr : (SampRec*)take(sizeof(SampRec) * n);
    w.reach = fixedOrder ? nullptr : (int2*)take(8L * 16 * lanes);
    w.tieQueue = fixedOrder ? nullptr : (int*)take(4 * (2 * lanes + 1));
    w.leafXY = (float2*)take(8 * n);
    w.coef = fixedOrder ? nullptr : (float*)take(4 * coefCap);
    w.coefCap = fixedOrder ? 0 : coefCap;
    w.hraw = nullptr;
    w.pnode = nullptr;
    w.footLeaf = nullptr;
    w.footSum = nullptr;
    w.frec = nullptr;
    for (int b = 0; b < 2; b++) { w.fixOff[b] = w.fixList[b] = nullptr; w.fixCoef[b] = w.dens[b] = nullptr; }
    const int nb = mode == kModePoisson ? 2 : mode == kModeFixed ? 1 : 0;
    for (int b = 0; b < nb; b++) {
        w.fixOff[b] = (int*)take(4 * n);
        w.fixList[b] = (int*)take(4 * n);
        w.fixCoef[b] = (float*)take(4L * kFixedCoefsMax * (n - 1 > 0 ? n - 1 : 1));
    }
    if (nb > 0) w.frec = (FmmRec*)take(sizeof(FmmRec) * n);
    if (mode == kModePoisson) {
        w.hraw = (float*)take(4 * n);
        w.pnode = (PNode*)take(sizeof(PNode) * nn);
        w.footLeaf = (FootLeaf*)take(sizeof(FootLeaf) * n);
        w.footSum = (double*)take(8);
        w.dens[0] = (float*)take(4L * W * H);
        w.dens[1] = (float*)take(4L * W * H);
    }
    if (t) *t = w;
    return off;
}

// ---------------------------------------------------------------------------
// Screen-space covariance V (row-major) of one splat in FP32, the operation
// order of samplingCuda.cu:106-150.
__device__ void cov_f32(float Wf, float Hf, float ff, const float* p, const float* n, const float* sd,
                        float smoothing, float V[4], float* detAOut = nullptr)
{
    float sx, sy;
    screen_f32(Wf, Hf, ff, p, sx, sy);
    float y0x = sx + 1.0f, y0y = sy, y1x = sx, y1y = sy + 1.0f;
    float tf = 2.0f * ff;
    float x0[3] = { (Wf - 2.0f * y0x) / tf, -((Hf - 2.0f * y0y) / tf), 1.0f };
    float x1[3] = { (Wf - 2.0f * y1x) / tf, -((Hf - 2.0f * y1y) / tf), 1.0f };
    float a = dot3(p, n);
    float xt0 = a / dot3(x0, n), xt1 = a / dot3(x1, n);
    float t0[3], t1[3];
    for (int c = 0; c < 3; c++) { t0[c] = x0[c] * xt0 - p[c]; t1[c] = x1[c] * xt1 - p[c]; }
    float nt0 = sqrtf(dot3(t0, t0));
    float inv0 = 1.0f / nt0;
    float u0[3] = { t0[0] * inv0, t0[1] * inv0, t0[2] * inv0 };
    float u1[3] = { u0[1] * n[2] - u0[2] * n[1], u0[2] * n[0] - u0[0] * n[2], u0[0] * n[1] - u0[1] * n[0] };
    float A00 = nt0, A01 = dot3(t1, u0), A10 = 0.0f, A11 = dot3(t1, u1);
    float detA = A00 * A11 - A01 * A10;
    if (detAOut) *detAOut = detA;
    float idA = 1.0f / detA;
    float J00 = A11 * idA, J01 = -A01 * idA, J10 = -A10 * idA, J11 = A00 * idA;
    float d0 = sd[0] * sd[0], d1 = sd[1] * sd[1];
    float M00 = J00 * d0 + J01 * 0.0f, M01 = J00 * 0.0f + J01 * d1;
    float M10 = J10 * d0 + J11 * 0.0f, M11 = J10 * 0.0f + J11 * d1;
    float V00 = M00 * J00 + M01 * J01, V01 = M00 * J10 + M01 * J11;
    float V10 = M10 * J00 + M11 * J01, V11 = M10 * J10 + M11 * J11;
    V[0] = V00 + smoothing; V[1] = V01 + 0.0f; V[2] = V10 + 0.0f; V[3] = V11 + smoothing;
}

// Per-splat bandwidth before the one-pixel cap, 1 / sqrt(lambda_min(V^-1))
// (samplingCuda.cu:106-163) in FP32, fixed operation order.
__device__ float iso_hraw(float Wf, float Hf, float ff, const float* p, const float* n, const float* sd, float smoothing)
{
    float V[4];
    cov_f32(Wf, Hf, ff, p, n, sd, smoothing, V);
    float V00 = V[0], V01 = V[1], V10 = V[2], V11 = V[3];
    float dV = V00 * V11 - V01 * V10;
    float idV = 1.0f / dV;
    float I00 = V11 * idV, I01 = -V01 * idV, I10 = -V10 * idV, I11 = V00 * idV;
    float tr = I00 + I11;
    float det = I00 * I11 - I01 * I10;
    float disc = 0.25f * (tr * tr) - det;
    float sq = sqrtf(disc);
    float l1 = 0.5f * tr + sq, l2 = 0.5f * tr - sq;
    float lk = fminf(l1, l2);
    return 1.0f / sqrtf(lk);
}

// Splat bandwidth for the Poisson-rate proposal: sqrt(lambda_max(V)) from V
// itself (well conditioned, unlike the smallest eigenvalue of V^-1 for splats
// seen at grazing angles); NaN for degenerate geometry.
__device__ float prop_h(float Wf, float Hf, float ff, const float* p, const float* n, const float* sd, float smoothing)
{
    float V[4];
    cov_f32(Wf, Hf, ff, p, n, sd, smoothing, V);
    float tr = V[0] + V[3];
    float det = V[0] * V[3] - V[1] * V[2];
    float disc = 0.25f * (tr * tr) - det;
    float lmax = 0.5f * tr + sqrtf(fmaxf(disc, 0.0f));
    return sqrtf(lmax);
}

// Footprint proposal of Poisson-rate sampling (no reference counterpart): the
// splat's own EWA weight w(q) = amp exp(-(u^2 + v^2) / 2), (u, v) = T (q - s),
// T^T T = V^-1 (upper triangular), amp = 1 / (2 pi sqrt(det V) |det A|), with V
// and A the FP32 screen covariance and Jacobian of cov_f32, and its isotropic
// envelope amp exp(-|q - s|^2 m), m = 1 / (2 lambda_max(V)) >= the envelope's
// exponent along any direction.  Degenerate geometry gives amp = 0 (the splat is
// then drawn through the wide-band half of the mixture alone).
__device__ void foot_f32(float Wf, float Hf, float ff, const float* p, const float* n, const float* sd,
                         float smoothing, float& t00, float& t01, float& t11, float& amp, float& m, float& mass)
{
    float V[4], detA;
    cov_f32(Wf, Hf, ff, p, n, sd, smoothing, V, &detA);
    float tr = V[0] + V[3];
    float det = V[0] * V[3] - V[1] * V[2];
    float disc = 0.25f * (tr * tr) - det;
    float lmax = 0.5f * tr + sqrtf(fmaxf(disc, 0.0f));
    t00 = sqrtf(V[3] / det);
    t01 = -V[1] / sqrtf(V[3] * det);
    t11 = 1.0f / sqrtf(V[3]);
    const float area = (2.0f * 3.14159265358979f) * sqrtf(det);     // 2 pi sqrt(det V)
    amp = 1.0f / (area * fabsf(detA));
    m = 0.5f / lmax;
    // integral of min(w, 1) over the plane
    mass = area * (amp <= 1.0f ? amp : 1.0f + plogf(amp));
    const bool ok = det > 0.0f && lmax > 0.0f && V[3] > 0.0f && isfinite(t00) && isfinite(t01) && isfinite(t11) &&
                    isfinite(amp) && isfinite(m) && isfinite(mass);
    if (!ok) { t00 = t01 = t11 = 0.0f; amp = 0.0f; m = 0.0f; mass = 0.0f; }
}

// the footprint cap E(d) at distance d (see kFootRcap)
__device__ __forceinline__ float foot_env(float a1, float d)
{
    const float e = fmaxf(0.0f, d - kFootRcap);
    return pexpf(-((e * e) * a1));
}

// the footprint term of one leaf at q: min(1, w(q), E(|q - s|))
__device__ __forceinline__ float foot_leaf(const FootLeaf& L, float a1, float qx, float qy, float d2)
{
    const float dx = qx - L.x, dy = qy - L.y;
    float u = L.t00 * dx;
    u = u + L.t01 * dy;
    const float v = L.t11 * dy;
    float e = u * u;
    e = e + v * v;
    const float w = L.amp * pexpf(-(0.5f * e));
    return fminf(1.0f, fminf(w, foot_env(a1, sqrtf(d2))));
}

// node bound of the footprint term at squared box distance d2: E, inflated by
// 2^-20 so that a one-ulp step of the exponential cannot put a leaf above it
__device__ __forceinline__ float foot_bound(float a1, float d2)
{
    return fminf(1.0f, foot_env(a1, sqrtf(d2)) * 1.00000095367431640625f);
}

// isotropic bandwidth capped at one pixel (samplingCuda.cu:160)
__device__ float iso_h(float Wf, float Hf, float ff, const float* p, const float* n, const float* sd, float smoothing)
{
    return fminf(1.0f, iso_hraw(Wf, Hf, ff, p, n, sd, smoothing));
}

// Tree-build kernels run over a batch of views at once, view = blockIdx.y: the
// views' workspaces lie ws bytes apart, their points / normals pts bytes apart
// and their std-devs sd bytes apart (0 when shared).  A single view passes {}.
struct VStride {
    size_t ws, pts, sd;
};

template <class T>
__device__ __forceinline__ T* vw(T* p, size_t stride)
{
    return p ? (T*)((char*)p + (size_t)blockIdx.y * stride) : p;
}

__global__ void k_zero_state(ViewState* vs, VStride vsd)
{
    vs = vw(vs, vsd.ws);
    if (threadIdx.x == 0) {
        vs->hsum = 0; vs->hcnt = 0; vs->nc = 0; vs->err = 0; vs->ncoef = 0; vs->nseeds = 0; vs->nmom = 0;
        vs->nfix[0] = 0; vs->nfix[1] = 0;
    }
    for (int b = threadIdx.x; b < 256; b += blockDim.x) vs->hhist[b] = 0;
}

__global__ __launch_bounds__(1024) void k_bandwidth(int N, float Wf, float Hf, float ff, float smoothing,
                                                   const float* __restrict__ pts, const float* __restrict__ nrm,
                                                   const float* __restrict__ sd, ViewState* vs, float* hraw, VStride vsd)
{
    pts = vw(pts, vsd.pts);
    nrm = vw(nrm, vsd.pts);
    sd = vw(sd, vsd.sd);
    vs = vw(vs, vsd.ws);
    hraw = vw(hraw, vsd.ws);
    __shared__ unsigned long long ssum[16];
    __shared__ unsigned int scnt[16];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long q = 0;
    unsigned int c = 0;
    // grid-stride over the splats (a capped grid keeps the two device-wide atomics
    // per workgroup few; the fixed-point sum does not depend on the grouping)
    for (; t < N; t += gridDim.x * blockDim.x) {
        float hk = iso_h(Wf, Hf, ff, pts + 3L * t, nrm + 3L * t, sd + 2L * t, smoothing);
        if (hraw) hraw[t] = prop_h(Wf, Hf, ff, pts + 3L * t, nrm + 3L * t, sd + 2L * t, smoothing);
        if (isfinite(hk)) { q += (unsigned long long)((double)hk * 1099511627776.0); c += 1; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        q += __shfl_xor(q, o, 64);
        c += __shfl_xor(c, o, 64);
    }
    int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { ssum[wv] = q; scnt[wv] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        unsigned int cc = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) { s += ssum[q]; cc += scnt[q]; }
        if (cc) { atomicAdd(&vs->hsum, s); atomicAdd(&vs->hcnt, cc); }
    }
}

// query radius r with exp(-r^2 / 2h^2) / (2 pi sqrt(h)) = precision (samplingCuda.cu:251-252)
__device__ float band_radius(float h, double prec)
{
    float normalization = (float)(2.0 * UD_PI * (double)sqrtf(h));
    return (float)(sqrt(2.0) * (double)h * sqrt(-log(prec) - log((double)normalization)));
}

__device__ Band make_band(float h, double prec)
{
    Band b;
    b.h = h;
    b.is2h = 1.0f / sqrtf((2.0f * h) * h);
    b.c = (float)(1.0 / (2.0 * UD_PI * (double)sqrtf(h)));    // buildRadixTree.cuh:580
    b.radius = band_radius(h, prec);
    return b;
}

__global__ void k_view_params(ViewState* vs, float Wf, float Hf, double prec, VStride vsd)
{
    vs = vw(vs, vsd.ws);
    if (threadIdx.x != 0) return;
    double hd = ((double)vs->hsum * (1.0 / 1099511627776.0)) / (double)vs->hcnt;
    float h = (float)hd;
    float radius = band_radius(h, prec);
    vs->h = h;
    vs->radius = radius;
    vs->lo = -radius;
    vs->hix = Wf + radius;
    vs->hiy = Hf + radius;
    vs->s2h = sqrtf((2.0f * h) * h);
    vs->is2h = 1.0f / vs->s2h;
    vs->prec = (float)prec;
    vs->nc = 0;
    vs->band[0] = make_band(h, prec);
}

__device__ __forceinline__ bool visible(const ViewState& v, float Wf, float Hf, float ff, const float* p, float& sx, float& sy)
{
    screen_f32(Wf, Hf, ff, p, sx, sy);
    bool vis = (p[2] <= 0.0f);
    vis &= (v.lo <= sx); vis &= (v.hix >= sx);
    vis &= (v.lo <= sy); vis &= (v.hiy >= sy);
    return vis;
}

// 10-bit cell coordinates interleaved into a 20-bit key (x bits above y bits).
// The reference spreads the bits three apart (a 3D code with z = 0,
// buildRadixTree.cuh:18-54); the compact code orders the splats identically and
// its highest differing bit is a monotone function of the spread one, so the
// Karras hierarchy on it is the same tree, and the sort needs three 8-bit
// passes instead of four.
__device__ __forceinline__ uint32_t expand_bits2(uint32_t v)
{
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ __forceinline__ uint32_t morton2(float ux, float uy)
{
    float cx = fminf(fmaxf(ux * 1024.0f, 0.0f), 1023.0f);
    float cy = fminf(fmaxf(uy * 1024.0f, 0.0f), 1023.0f);
    return 2u * expand_bits2((uint32_t)cx) + expand_bits2((uint32_t)cy);
}

// exclusive scan of one int per thread over a 256-thread workgroup
__device__ __forceinline__ int block_excl_scan(int v, int* sh /*[4]*/, int& total)
{
    int incl = wave_incl_scan(v);
    int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) sh[wv] = incl;
    __syncthreads();
    int off = 0;
    for (int q = 0; q < wv; q++) off += sh[q];
    total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return off + incl - v;
}

__global__ __launch_bounds__(256) void k_cull_count(int N, float Wf, float Hf, float ff,
                                                    const float* __restrict__ pts, const ViewState* __restrict__ vs,
                                                    int* __restrict__ counts, VStride vsd)
{
    pts = vw(pts, vsd.pts);
    vs = vw(vs, vsd.ws);
    counts = vw(counts, vsd.ws);
    __shared__ int sh[4];
    ViewState v = *vs;
    int base = blockIdx.x * kCullTile;
    int c = 0;
    for (int r = 0; r < kCullTile / 256; r++) {
        int e = base + r * 256 + threadIdx.x;
        float sx, sy;
        if (e < N && visible(v, Wf, Hf, ff, pts + 3L * e, sx, sy)) c++;
    }
    int total;
    block_excl_scan(c, sh, total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// single-workgroup exclusive scan of n ints in place; total to *totalOut
__global__ __launch_bounds__(1024) void k_scan_small(int* data, int n, int* totalOut, VStride vsd)
{
    data = vw(data, vsd.ws);
    totalOut = vw(totalOut, vsd.ws);
    __shared__ int sh[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        int e = base + threadIdx.x;
        int v = e < n ? data[e] : 0;
        int incl = wave_incl_scan(v);
        int wv = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 63) sh[wv] = incl;
        __syncthreads();
        int off = 0;
        for (int q = 0; q < wv; q++) off += sh[q];
        int tot = 0;
        for (int q = 0; q < 16; q++) tot += sh[q];
        int c = carry;
        if (e < n) data[e] = c + off + incl - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *totalOut = carry;
}

__global__ __launch_bounds__(256) void k_cull_scatter(int N, float Wf, float Hf, float ff,
                                                      const float* __restrict__ pts, const ViewState* __restrict__ vs,
                                                      const int* __restrict__ offs, uint32_t* __restrict__ keys,
                                                      int32_t* __restrict__ ids, VStride vsd)
{
    pts = vw(pts, vsd.pts);
    vs = vw(vs, vsd.ws);
    offs = vw(offs, vsd.ws);
    keys = vw(keys, vsd.ws);
    ids = vw(ids, vsd.ws);
    __shared__ int sh[4];
    ViewState v = *vs;
    int base = blockIdx.x * kCullTile;
    int run = offs[blockIdx.x];
    for (int r = 0; r < kCullTile / 256; r++) {
        int e = base + r * 256 + threadIdx.x;
        float sx = 0.0f, sy = 0.0f;
        bool vis = e < N && visible(v, Wf, Hf, ff, pts + 3L * e, sx, sy);
        int total;
        int pos = run + block_excl_scan(vis ? 1 : 0, sh, total);
        if (vis) {
            float ux = (sx - v.lo) / (v.hix - v.lo);
            float uy = (sy - v.lo) / (v.hiy - v.lo);
            keys[pos] = morton2(ux, uy);
            ids[pos] = e;
        }
        run += total;
    }
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits.  Histogram -> digit-major scan -> stable scatter.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sort_hist(const ViewState* __restrict__ vs, const uint32_t* __restrict__ keys,
                                                   int shift, int nblk, int* __restrict__ hist, VStride vsd)
{
    vs = vw(vs, vsd.ws);
    keys = vw(keys, vsd.ws);
    hist = vw(hist, vsd.ws);
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int n = vs->nc;
    int base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortTile / 256; r++) {
        int e = base + r * 256 + threadIdx.x;
        if (e < n) atomicAdd(&h[(keys[e] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of hist[d][0..nblk) for each digit d (one workgroup per digit);
// digit totals to tot[d]
__global__ __launch_bounds__(256) void k_sort_scan_digits(int* hist, int nblk, int* tot, VStride vsd)
{
    hist = vw(hist, vsd.ws);
    tot = vw(tot, vsd.ws);
    __shared__ int sh[4];
    __shared__ int carry;
    int* h = hist + (long)blockIdx.x * nblk;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nblk; base += 256) {
        int e = base + threadIdx.x;
        int v = e < nblk ? h[e] : 0;
        int total;
        int ex = block_excl_scan(v, sh, total);
        int c = carry;
        if (e < nblk) h[e] = c + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// digit bases: exclusive scan of the 256 digit totals (in place)
__global__ __launch_bounds__(256) void k_sort_scan_top(int* tot, VStride vsd)
{
    tot = vw(tot, vsd.ws);
    __shared__ int sh[4];
    int total;
    int v = tot[threadIdx.x];
    int ex = block_excl_scan(v, sh, total);
    tot[threadIdx.x] = ex;
}

__global__ __launch_bounds__(256) void k_sort_scatter(const ViewState* __restrict__ vs, const uint32_t* __restrict__ kin,
                                                      const int32_t* __restrict__ iin, int shift, int nblk,
                                                      const int* __restrict__ hist, const int* __restrict__ digitBase,
                                                      uint32_t* __restrict__ kout, int32_t* __restrict__ iout, VStride vsd)
{
    vs = vw(vs, vsd.ws);
    kin = vw(kin, vsd.ws);
    iin = vw(iin, vsd.ws);
    hist = vw(hist, vsd.ws);
    digitBase = vw(digitBase, vsd.ws);
    kout = vw(kout, vsd.ws);
    iout = vw(iout, vsd.ws);
    __shared__ int run[256];
    __shared__ int wc[4][256];
    int n = vs->nc;
    int base = blockIdx.x * kSortTile;
    run[threadIdx.x] = hist[threadIdx.x * nblk + blockIdx.x] + digitBase[threadIdx.x];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kSortTile / 256; r++) {
        for (int q = 0; q < 4; q++) wc[q][threadIdx.x] = 0;
        __syncthreads();
        int e = base + r * 256 + threadIdx.x;
        bool act = e < n;
        uint32_t key = act ? kin[e] : 0u;
        int32_t id = act ? iin[e] : 0;
        uint32_t dg = (key >> shift) & 255u;
        // lanes of this wave with the same digit
        unsigned long long m = __ballot(act);
        for (int b = 0; b < 8; b++) {
            unsigned long long bb = __ballot((dg >> b) & 1u);
            m &= ((dg >> b) & 1u) ? bb : ~bb;
        }
        int rk = __popcll(m & lt);
        if (act && rk == 0) wc[wv][dg] = __popcll(m);
        __syncthreads();
        if (act) {
            int off = run[dg];
            for (int q = 0; q < wv; q++) off += wc[q][dg];
            int pos = off + rk;
            kout[pos] = key;
            iout[pos] = id;
        }
        __syncthreads();
        run[threadIdx.x] += wc[0][threadIdx.x] + wc[1][threadIdx.x] + wc[2][threadIdx.x] + wc[3][threadIdx.x];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Karras hierarchy on 64-bit unique keys (buildRadixTree.cuh:76-167).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t key64(const uint32_t* k, int i) { return ((uint64_t)k[i] << 32) | (uint32_t)i; }

__device__ __forceinline__ int delta(const uint32_t* k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    return __clzll(key64(k, i) ^ key64(k, j));
}

__device__ float trunc_r2(float h, float prec, int cnt)
{
    float v = -2.0f * (h * h);
    v = v * plogf(prec / (float)cnt);
    float R = sqrtf(v);
    return R * R;
}

__global__ __launch_bounds__(256) void k_hierarchy(float Wf, float Hf, float ff, const float* __restrict__ pts,
                                                   const ViewState* __restrict__ vs, const uint32_t* __restrict__ keys,
                                                   const int32_t* __restrict__ ids, NodeGeo* __restrict__ geo,
                                                   NodeLnk* __restrict__ lnk, NodeAux* __restrict__ aux,
                                                   float2* __restrict__ leafXY, int* __restrict__ flags, VStride vsd)
{
    pts = vw(pts, vsd.pts);
    vs = vw(vs, vsd.ws);
    keys = vw(keys, vsd.ws);
    ids = vw(ids, vsd.ws);
    geo = vw(geo, vsd.ws);
    lnk = vw(lnk, vsd.ws);
    aux = vw(aux, vsd.ws);
    leafXY = vw(leafXY, vsd.ws);
    flags = vw(flags, vsd.ws);
    int n = vs->nc;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int L0 = n - 1;
    float h = vs->h, prec = vs->prec;
    // leaf t
    {
        int nd = L0 + t;
        float sx, sy;
        screen_f32(Wf, Hf, ff, pts + 3L * ids[t], sx, sy);
        NodeGeo g;
        g.lox = sx; g.loy = sy; g.hix = sx; g.hiy = sy; g.cx = sx; g.cy = sy;
        g.r2 = trunc_r2(h, prec, 1);
        g.deg = 0;
        geo[nd] = g;
        leafXY[t] = make_float2(sx, sy);
        lnk[nd].left = -1; lnk[nd].right = -1; lnk[nd].count = 1;
        NodeAux a; a.off = 0; a.first = t; a.last = t; a.pad = ids[t];   // pad: the leaf's splat id
        aux[nd] = a;
    }
    if (t == 0) lnk[0].parent = -1;
    if (t >= n - 1) return;
    flags[t] = 0;                                  // fan-in arrival count of internal node t
    // internal node t
    int i0, j0;
    if (t == 0) { i0 = 0; j0 = n - 1; }
    else {
        int dl = delta(keys, n, t, t - 1), dr = delta(keys, n, t, t + 1);
        int d = dr > dl ? 1 : -1;
        int dmin = dl < dr ? dl : dr;
        int lmax = 2;
        while (delta(keys, n, t, t + d * lmax) > dmin) lmax <<= 1;
        int l = 0;
        for (int s = lmax >> 1; s > 0; s >>= 1)
            if (delta(keys, n, t, t + (l + s) * d) > dmin) l += s;
        int jdx = t + l * d;
        i0 = d < 0 ? jdx : t;
        j0 = d < 0 ? t : jdx;
    }
    uint64_t fc = key64(keys, i0), lc = key64(keys, j0);
    int sp;
    if (fc == lc) sp = (i0 + j0) >> 1;
    else {
        int dn = __clzll(fc ^ lc);
        sp = i0;
        int stride = j0 - i0;
        do {
            stride = (stride + 1) >> 1;
            int mid = sp + stride;
            if (mid < j0 && __clzll(fc ^ key64(keys, mid)) > dn) sp = mid;
        } while (stride > 1);
    }
    int lc_ = sp, rc_ = sp + 1;
    if (i0 == sp) lc_ += L0;
    if (j0 == sp + 1) rc_ += L0;
    lnk[t].left = lc_; lnk[t].right = rc_;
    lnk[lc_].parent = t; lnk[rc_].parent = t;
    aux[t].first = i0; aux[t].last = j0;
}

// hermiteError (hermitefmm.cuh:184-195), integer powers by repeated products
__device__ float hermite_error(float r, float nr, int p)
{
    if (r >= 1.0f) return r;
    float om = 1.0f - r;
    float a = nr / (om * om);
    float rp = 1.0f;
    for (int i = 0; i < p; i++) rp = rp * r;
    float fac = 1.0f;
    for (int i = 1; i <= p; i++) fac = fac * (float)i;
    float b = rp / sqrtf(fac);
    float c = b * b;
    float d = 2.0f * (1.0f - rp) * b;
    return a * (c + d);
}

// The record of an internal node from its two complete children, left then
// right (buildDynamicRadixTree.cuh:126-223, with the children's own centres).
__device__ __forceinline__ NodeGeo merge_geo(const NodeGeo& gl, int cl, const NodeGeo& gr, int cr, float h,
                                             float prec)
{
    NodeGeo g;
    g.lox = fminf(gl.lox, gr.lox); g.loy = fminf(gl.loy, gr.loy);
    g.hix = fmaxf(gl.hix, gr.hix); g.hiy = fmaxf(gl.hiy, gr.hiy);
    int cn = cl + cr;
    float num = (float)cl * gl.cx;
    num = num + (float)cr * gr.cx;
    g.cx = num / (float)cn;
    num = (float)cl * gl.cy;
    num = num + (float)cr * gr.cy;
    g.cy = num / (float)cn;
    float mn = 0.0f;
    mn = fmaxf(mn, fmaxf(g.cx - g.lox, g.hix - g.cx));
    mn = fmaxf(mn, fmaxf(g.cy - g.loy, g.hiy - g.cy));
    float rr = mn / h;
    int deg;
    if (gl.deg < 0 && gr.deg < 0) deg = -1;
    else if (!(rr < 1.0f)) deg = -1;
    else {
        // the lowest degree from the children's up whose hermiteError meets the
        // precision; the error's power r^d and factorial d! are carried from one
        // degree to the next, with the same products as hermite_error evaluates
        int d = gl.deg > gr.deg ? gl.deg : gr.deg;
        const float om = 1.0f - rr;
        const float a = (float)cn / (om * om);
        float rp = 1.0f, fac = 1.0f;
        for (int i = 0; i < d; i++) rp = rp * rr;
        for (int i = 1; i <= d; i++) fac = fac * (float)i;
        while (d <= kMaxDegree) {
            const float b = rp / sqrtf(fac);
            const float c = b * b;
            const float e = 2.0f * (1.0f - rp) * b;
            if (!(a * (c + e) > prec)) break;
            d++;
            rp = rp * rr;
            fac = fac * (float)d;
        }
        deg = d <= kMaxDegree ? d : -1;
    }
    g.deg = deg;
    g.r2 = trunc_r2(h, prec, cn);
    return g;
}

// One fan-in step at parent p of node nd, whose record g (count cnt) the thread
// holds: the second arrival loads the sibling's record (the sibling's count is
// p's leaf count minus cnt), merges, stores p and continues from it.
__device__ __forceinline__ void merge_up(int p, const NodeLnk& L, const NodeAux& a, int& nd, NodeGeo& g, int& cnt,
                                         float h, float prec, NodeGeo* geo, NodeLnk* lnk)
{
    const bool isLeft = L.left == nd;
    const NodeGeo gs = geo[isLeft ? L.right : L.left];
    const int cs = (a.last - a.first + 1) - cnt;
    const NodeGeo gl = isLeft ? g : gs, gr = isLeft ? gs : g;
    const int cl = isLeft ? cnt : cs, cr = isLeft ? cs : cnt;
    const NodeGeo m = merge_geo(gl, cl, gr, cr, h, prec);
    geo[p] = m;
    lnk[p].count = cl + cr;
    g = m;
    cnt = cl + cr;
    nd = p;
}

// Bottom-up pass, level 1: one workgroup per block of 1024 sorted leaves.  The
// second thread to reach a node whose leaf range lies inside the block merges
// it.  Such a node's index lies inside the block too (a Karras node's index is
// one end of its range), so the block's internal nodes are staged in LDS first
// (links and leaf ranges in, merged records out, one arrival counter each) and
// the fan-in never waits on device memory: both arrivals come from the same
// workgroup, so workgroup-scope fences order the exchange.  Merged records are
// also stored to device memory for the later passes.  A thread that reaches a
// node spanning the block's boundary hands its (complete) node to level 2.
constexpr int kFanBlock = 1024;
__global__ __launch_bounds__(kFanBlock) void k_bottom_up_local(ViewState* vs, NodeGeo* geo, NodeLnk* lnk,
                                                               const NodeAux* __restrict__ aux, int* seeds, VStride vsd)
{
    __shared__ int arrivals[kFanBlock];
    __shared__ float2 sLeaf[kFanBlock];          // the block's leaf positions
    __shared__ NodeGeo sNode[kFanBlock];         // merged records of internal nodes b0 + i
    __shared__ int4 sLink[kFanBlock];            // left, right, parent, first leaf of node b0 + i
    __shared__ int sLast[kFanBlock];             // last leaf of node b0 + i
    __shared__ int sSeeds[kFanBlock];            // nodes handed to level 2
    __shared__ int nSeeds, seedBase;
    vs = vw(vs, vsd.ws);
    geo = vw(geo, vsd.ws);
    lnk = vw(lnk, vsd.ws);
    aux = vw(aux, vsd.ws);
    seeds = vw(seeds, vsd.ws);
    const int n = vs->nc;
    const int L0 = n - 1;
    const int b0 = blockIdx.x * kFanBlock;
    const int tid = threadIdx.x;
    const int pos = b0 + tid;
    arrivals[tid] = 0;
    if (tid == 0) nSeeds = 0;
    NodeGeo g;
    int p = -1;
    if (pos < n && n >= 2) {
        g = geo[L0 + pos];
        sLeaf[tid] = make_float2(g.cx, g.cy);
        p = lnk[L0 + pos].parent;
    }
    if (pos < L0) {
        const NodeLnk L = lnk[pos];
        const NodeAux a = aux[pos];
        sLink[tid] = make_int4(L.left, L.right, L.parent, a.first);
        sLast[tid] = a.last;
    }
    __syncthreads();
    const float h = vs->h, prec = vs->prec;
    int nd = L0 + pos;
    int cnt = 1;
    if (pos >= n || n < 2) p = -1;
    while (p >= 0) {
        const int pi = p - b0;
        const int4 Lk = (pi >= 0 && pi < kFanBlock) ? sLink[pi] : make_int4(0, 0, 0, -1);
        const int last = (pi >= 0 && pi < kFanBlock) ? sLast[pi] : b0 + kFanBlock;
        if (Lk.w < b0 || last >= b0 + kFanBlock) {   // p spans the block's boundary
            sSeeds[atomicAdd(&nSeeds, 1)] = nd;       // at most one per thread
            break;
        }
        __threadfence_block();
        int old = atomicAdd(&arrivals[pi], 1);
        if (old == 0) break;
        __threadfence_block();
        const bool isLeft = Lk.x == nd;
        const int sib = isLeft ? Lk.y : Lk.x;
        NodeGeo gs;
        if (sib >= L0) {                             // a leaf: its box and centre are its position
            const float2 q = sLeaf[sib - L0 - b0];
            gs.lox = q.x; gs.loy = q.y; gs.hix = q.x; gs.hiy = q.y; gs.cx = q.x; gs.cy = q.y;
            gs.r2 = 0.0f; gs.deg = 0;
        } else {
            gs = sNode[sib - b0];
        }
        const int cs = (last - Lk.w + 1) - cnt;
        const NodeGeo gl = isLeft ? g : gs, gr = isLeft ? gs : g;
        const int cl = isLeft ? cnt : cs, cr = isLeft ? cs : cnt;
        const NodeGeo m = merge_geo(gl, cl, gr, cr, h, prec);
        sNode[pi] = m;
        geo[p] = m;
        lnk[p].count = cl + cr;
        g = m;
        cnt = cl + cr;
        nd = p;
        p = Lk.z;
    }
    // one device-wide slot reservation per workgroup (thousands of single-word
    // atomics on one counter would serialize)
    __syncthreads();
    if (tid == 0) seedBase = nSeeds ? atomicAdd(&vs->nseeds, nSeeds) : 0;
    __syncthreads();
    if (tid < nSeeds) seeds[seedBase + tid] = sSeeds[tid];
}

// Bottom-up pass, level 2: the few block-spanning nodes, device-scope fan-in.
__global__ __launch_bounds__(256) void k_bottom_up_global(const ViewState* __restrict__ vs, NodeGeo* geo, NodeLnk* lnk,
                                                          const NodeAux* __restrict__ aux, int* flags,
                                                          const int* __restrict__ seeds, VStride vsd)
{
    vs = vw(vs, vsd.ws);
    geo = vw(geo, vsd.ws);
    lnk = vw(lnk, vsd.ws);
    aux = vw(aux, vsd.ws);
    flags = vw(flags, vsd.ws);
    seeds = vw(seeds, vsd.ws);
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= vs->nseeds) return;
    float h = vs->h, prec = vs->prec;
    int nd = seeds[t];
    NodeGeo g = geo[nd];
    int cnt = lnk[nd].count;
    int p = lnk[nd].parent;
    while (p >= 0) {
        const NodeAux a = aux[p];
        const NodeLnk L = lnk[p];
        // release this thread's node record before the arrival count, acquire the
        // sibling's after it (the second arrival reads the first's increment)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        int old = atomicAdd(&flags[p], 1);
        if (old == 0) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        merge_up(p, L, a, nd, g, cnt, h, prec, geo, lnk);
        p = L.parent;
    }
}

// Coefficient block sizes.  FGT walks start at the root and at the children of
// nodes with more than k sources (the capacity descent asks for their weights);
// a walk stops at the first node with an expansion.  An internal node below an
// expanded node, with no start in between, is therefore never evaluated and
// gets no block.
__global__ __launch_bounds__(256) void k_coef_sizes(ViewState* vs, const NodeGeo* __restrict__ geo,
                                                    const NodeLnk* __restrict__ lnk, int k, int* sizes, int* momList)
{
    int n = vs->nc;
    int nn = 2 * n - 1;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int d = t < nn ? geo[t].deg : -1;
    if (d >= 0 && t < n - 1) {
        int x = t;
        while (true) {
            const int p = lnk[x].parent;
            if (p < 0 || lnk[p].count > k) break;           // x starts a walk
            if (geo[p].deg >= 0) { d = -1; break; }          // every walk stops at p
            x = p;
        }
    }
    if (t < nn) sizes[t] = d >= 0 ? (d + 1) * (d + 1) : 0;
    // list the internal nodes that need moments: one slot reservation per workgroup
    __shared__ int wcnt[4];
    __shared__ int bbase;
    const bool want = d >= 0 && t < n - 1;
    const unsigned long long m = __ballot(want);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wcnt[wid] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        bbase = tot ? atomicAdd(&vs->nmom, tot) : 0;
    }
    __syncthreads();
    int base = bbase;
    for (int q = 0; q < wid; q++) base += wcnt[q];
    if (want) momList[base + __popcll(m & ((1ull << lane) - 1ull))] = t;
}

// device-wide exclusive scan, phase 1: per-tile scan + tile totals
__global__ __launch_bounds__(256) void k_scan_tiles(const ViewState* __restrict__ vs, int* data, int* tileSums)
{
    __shared__ int sh[4];
    int nn = 2 * vs->nc - 1;
    int base = blockIdx.x * kScanTile;
    if (base >= nn) { if (threadIdx.x == 0) tileSums[blockIdx.x] = 0; return; }
    int v[4], s = 0;
    for (int q = 0; q < 4; q++) {
        int e = base + 4 * threadIdx.x + q;
        v[q] = e < nn ? data[e] : 0;
        s += v[q];
    }
    int total;
    int ex = block_excl_scan(s, sh, total);
    for (int q = 0; q < 4; q++) {
        int e = base + 4 * threadIdx.x + q;
        if (e < nn) data[e] = ex;
        ex += v[q];
    }
    if (threadIdx.x == 0) tileSums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_apply(const ViewState* __restrict__ vs, const int* data,
                                                    const int* tileOffs, NodeAux* aux)
{
    int nn = 2 * vs->nc - 1;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn) return;
    aux[t].off = data[t] + tileOffs[t / kScanTile];
}

// Hermite moments A[a][b] = sum_leaves tx^a ty^b / (a! b!), t = (x - c) / sqrt(2 h^2)
// (hermitefmm.cuh:119-138).  Leaves carry the single coefficient 1.  Internal
// nodes with an expansion are listed by k_coef_sizes; thread (node, row a)
// computes row a of one node, accumulating in FP64 in leaf order.  A workgroup
// of 255 threads covers 15 nodes x 17 rows.
constexpr int kMomNodes = 15;

// 1 / a! as correctly rounded doubles (IEEE division, evaluated at compile time)
constexpr double ifact_c(int a) { double f = 1.0; for (int q = 2; q <= a; q++) f = f * q; return 1.0 / f; }
__constant__ double kInvFact[kMaxDegree + 1] = {
    ifact_c(0), ifact_c(1), ifact_c(2), ifact_c(3), ifact_c(4), ifact_c(5), ifact_c(6), ifact_c(7), ifact_c(8),
    ifact_c(9), ifact_c(10), ifact_c(11), ifact_c(12), ifact_c(13), ifact_c(14), ifact_c(15), ifact_c(16) };
__global__ __launch_bounds__(256) void k_moments(const ViewState* __restrict__ vs, const NodeGeo* __restrict__ geo,
                                                 const NodeAux* __restrict__ aux, const int* __restrict__ momList,
                                                 float* __restrict__ coef, long coefCap)
{
    const int n = vs->nc;
    if (vs->ncoef > coefCap) return;
    const int L0 = n - 1;
    const int t = threadIdx.x;
    // leaves: 256 per workgroup
    {
        long lf = (long)blockIdx.x * 256 + t;
        if (lf < n) coef[aux[L0 + lf].off] = 1.0f;
    }
    if (t >= kMomNodes * (kMaxDegree + 1)) return;
    const long slot = (long)blockIdx.x * kMomNodes + t / (kMaxDegree + 1);
    const int ax = t % (kMaxDegree + 1);
    if (slot >= vs->nmom) return;
    const int nd = momList[slot];
    NodeGeo g = geo[nd];
    const int p = g.deg;
    if (ax > p) return;
    const float is2h = vs->is2h;
    NodeAux a = aux[nd];
    const int P1 = p + 1;
    const double ifx = kInvFact[ax];
    double acc[kMaxDegree + 1];
#pragma unroll
    for (int bb = 0; bb <= kMaxDegree; bb++) acc[bb] = 0.0;
    for (int pos = a.first; pos <= a.last; pos++) {
        NodeGeo lf = geo[L0 + pos];
        float tx = (lf.cx - g.cx) * is2h;
        float ty = (lf.cy - g.cy) * is2h;
        double px = 1.0;
        for (int q = 0; q < ax; q++) px = px * (double)tx;
        double mx = px * ifx;
        double py = 1.0;
#pragma unroll
        for (int bb = 0; bb <= kMaxDegree; bb++) {
            if (bb <= p) {
                acc[bb] += mx * (py * ifact_c(bb));
                py = py * (double)ty;
            }
        }
    }
#pragma unroll
    for (int bb = 0; bb <= kMaxDegree; bb++)
        if (bb <= p) coef[a.off + ax * P1 + bb] = (float)acc[bb];
}

__global__ void k_check_coef(ViewState* vs, long coefCap)
{
    if (threadIdx.x == 0 && (long)vs->ncoef > coefCap) vs->err |= 2;
}

// ---------------------------------------------------------------------------
// FGT evaluation (evaluateDynamicRadixTree.cuh:128-242, hermitefmm.cuh:35-47, 213-237)
// ---------------------------------------------------------------------------
typedef float f4v __attribute__((ext_vector_type(4)));
typedef int i4v __attribute__((ext_vector_type(4)));

struct TreeView {
    const NodeGeo UD_GLOBAL* geo;
    const NodeLnk UD_GLOBAL* lnk;
    const NodeAux UD_GLOBAL* aux;
    const float UD_GLOBAL* coef;
    float s2h, is2h;
};

__device__ __forceinline__ NodeGeo ld_geo(const TreeView& T, int i)
{
    const f4v UD_GLOBAL* q = (const f4v UD_GLOBAL*)(T.geo + i);
    f4v a = q[0], b = q[1];
    NodeGeo g;
    g.lox = a.x; g.loy = a.y; g.hix = a.z; g.hiy = a.w;
    g.cx = b.x; g.cy = b.y; g.r2 = b.z; g.deg = __float_as_int(b.w);
    return g;
}

__device__ __forceinline__ NodeLnk ld_lnk(const TreeView& T, int i)
{
    i4v a = *(const i4v UD_GLOBAL*)(T.lnk + i);
    NodeLnk l;
    l.left = a.x; l.right = a.y; l.parent = a.z; l.count = a.w;
    return l;
}

__device__ __forceinline__ float eval_node(const TreeView& T, const NodeGeo& g, int nd, float qx, float qy)
{
    int p = g.deg;
    float tx = (qx - g.cx) * T.is2h;
    float ty = (qy - g.cy) * T.is2h;
    float ex = pexpf(-(tx * tx)), ey = pexpf(-(ty * ty));
    const float UD_GLOBAL* c = T.coef + T.aux[nd].off;
    if (p == 0) {
        float w = 0.0f;
        w = w + (c[0] * ex) * ey;
        return w;
    }
    // Hermite functions h_n(t) = 2t h_{n-1} - 2(n-1) h_{n-2}, h_0 = e^{-t^2}
    // (loops unrolled to the maximum degree and predicated, so hy stays in registers).
    // Each term of the series is added with one fused multiply-add,
    // w = fma(c[a][b] h_a, h_b, w): the reference's `weight += coeffs[index] *
    // hermiteX[a] * hermiteY[b]` (hermitefmm.cuh:233) as nvcc's default
    // contraction (-fmad=true) compiles it; each recurrence step is
    // h_n = fma(2t, h_{n-1}, -(2(n-1) h_{n-2})) (hermitefmm.cuh:45), the form
    // of the fixed-order densities below.
    float hy[kMaxDegree + 1];
    hy[0] = ey;
    hy[1] = (2.0f * ty) * ey;
#pragma unroll
    for (int n = 2; n <= kMaxDegree; n++)
        hy[n] = (n <= p) ? __builtin_fmaf(2.0f * ty, hy[n - 1], -((2.0f * (float)(n - 1)) * hy[n - 2])) : 0.0f;
    float w = 0.0f;
    float hprev = 0.0f, ha = ex;
    for (int a = 0; a <= p; a++) {
        if (a == 1) { hprev = ex; ha = (2.0f * tx) * ex; }
        else if (a >= 2) {
            float hn = __builtin_fmaf(2.0f * tx, ha, -((2.0f * (float)(a - 1)) * hprev));
            hprev = ha; ha = hn;
        }
        const float UD_GLOBAL* row = c + a * (p + 1);
#pragma unroll
        for (int b = 0; b <= kMaxDegree; b++)
            if (b <= p) w = __builtin_fmaf(row[b] * ha, hy[b], w);
    }
    return w;
}

// Hermite evaluation with the coefficient block staged in LDS: called by all 64
// lanes of a wavefront for the same node (uniform p and block), so the block
// is fetched with one coalesced load per 64 coefficients and then read with
// LDS broadcasts.  Lanes with act == false only help with the load.
// the degree-P series sum_a sum_b c[a][b] h_a(tx) h_b(ty) with the loops of
// eval_node unrolled exactly (same operation order)
template <int P>
__device__ __forceinline__ float hermite_series(const float* sc, float tx, float ty, float ex, float ey)
{
    float hy[P + 1];
    hy[0] = ey;
    if (P >= 1) hy[P >= 1 ? 1 : 0] = (2.0f * ty) * ey;
#pragma unroll
    for (int n = 2; n <= P; n++) hy[n] = __builtin_fmaf(2.0f * ty, hy[n - 1], -((2.0f * (float)(n - 1)) * hy[n - 2]));
    float w = 0.0f, hprev = 0.0f, ha = ex;
#pragma unroll
    for (int a = 0; a <= P; a++) {
        if (a == 1) { hprev = ex; ha = (2.0f * tx) * ex; }
        else if (a >= 2) {
            float hn = __builtin_fmaf(2.0f * tx, ha, -((2.0f * (float)(a - 1)) * hprev));
            hprev = ha; ha = hn;
        }
#pragma unroll
        for (int b = 0; b <= P; b++) w = __builtin_fmaf(sc[a * (P + 1) + b] * ha, hy[b], w);
    }
    return w;
}

// eval_node for a wave-uniform node: the coefficient block is staged in LDS
// (read by broadcast) and the series is unrolled for the node's degree
__device__ __forceinline__ float eval_node_staged(const TreeView& T, const NodeGeo& g, int off, float qx, float qy,
                                                  bool act, float* sc, int lane)
{
    const int p = g.deg;
    const int np = (p + 1) * (p + 1);
    const float UD_GLOBAL* c = T.coef + off;
    for (int q = lane; q < np; q += 64) sc[q] = c[q];
    __builtin_amdgcn_wave_barrier();
    float w = 0.0f;
    if (act) {
        float tx = (qx - g.cx) * T.is2h;
        float ty = (qy - g.cy) * T.is2h;
        float ex = pexpf(-(tx * tx)), ey = pexpf(-(ty * ty));
        switch (p) {
#define UD_SERIES(P) case P: w = hermite_series<P>(sc, tx, ty, ex, ey); break;
        UD_SERIES(0) UD_SERIES(1) UD_SERIES(2) UD_SERIES(3) UD_SERIES(4) UD_SERIES(5) UD_SERIES(6)
        UD_SERIES(7) UD_SERIES(8)
#undef UD_SERIES
        default: {
            // degrees 9 .. 16 (few nodes): loops to the maximum degree, predicated
            float hy[kMaxDegree + 1];
            hy[0] = ey;
            hy[1] = (2.0f * ty) * ey;
#pragma unroll
            for (int n = 2; n <= kMaxDegree; n++)
                hy[n] = (n <= p) ? __builtin_fmaf(2.0f * ty, hy[n - 1], -((2.0f * (float)(n - 1)) * hy[n - 2])) : 0.0f;
            float hprev = 0.0f, ha = ex;
            for (int a = 0; a <= p; a++) {
                if (a == 1) { hprev = ex; ha = (2.0f * tx) * ex; }
                else if (a >= 2) {
                    float hn = __builtin_fmaf(2.0f * tx, ha, -((2.0f * (float)(a - 1)) * hprev));
                    hprev = ha; ha = hn;
                }
                const float* row = sc + a * (p + 1);
#pragma unroll
                for (int b = 0; b <= kMaxDegree; b++)
                    if (b <= p) w = __builtin_fmaf(row[b] * ha, hy[b], w);
            }
        }
        }
    }
    __builtin_amdgcn_wave_barrier();
    return w;
}

__device__ __forceinline__ bool truncated(const NodeGeo& g, float qx, float qy)
{
    float cx = fminf(fmaxf(qx, g.lox), g.hix);
    float cy = fminf(fmaxf(qy, g.loy), g.hiy);
    float dx = qx - cx, dy = qy - cy;
    float d2 = dx * dx;
    d2 = d2 + dy * dy;
    return !(d2 < g.r2);
}

__device__ __noinline__ float tree_weight(const TreeView& T, int entry, float qx, float qy)
{
    NodeGeo g = ld_geo(T, entry);
    if (g.deg >= 0) return eval_node(T, g, entry, qx, qy);
    int rootparent = T.lnk[entry].parent;
    int nd = entry;
    float w = 0.0f;
    while (true) {
        while (g.deg < 0 && !truncated(g, qx, qy)) { nd = T.lnk[nd].left; g = ld_geo(T, nd); }
        if (g.deg >= 0) w = w + eval_node(T, g, nd, qx, qy);
        int pr = T.lnk[nd].parent;
        while (pr != rootparent && nd == T.lnk[pr].right) { nd = pr; pr = T.lnk[pr].parent; }
        if (pr != rootparent) { nd = T.lnk[pr].right; g = ld_geo(T, nd); }
        else break;   // climbed back to the entry node
    }
    return w;
}


__device__ __forceinline__ float key_of(unsigned long long v) { return __uint_as_float((uint32_t)(v >> 32)); }

// Children records of the sampler's walk, one per internal node (after the tree's
// degrees and coefficient offsets are final).
__device__ __forceinline__ SampChild samp_child(int c, const NodeGeo* __restrict__ geo,
                                                const NodeLnk* __restrict__ lnk, const NodeAux* __restrict__ aux)
{
    const NodeGeo g = geo[c];
    SampChild r;
    r.lox = g.lox; r.loy = g.loy; r.hix = g.hix; r.hiy = g.hiy;
    r.cx = g.cx; r.cy = g.cy; r.r2 = g.r2; r.deg = g.deg;
    r.off = aux[c].off; r.count = lnk[c].count; r.id = c; r.pad = 0;
    return r;
}

__global__ __launch_bounds__(256) void k_samp_records(const ViewState* __restrict__ vs,
                                                      const NodeGeo* __restrict__ geo,
                                                      const NodeLnk* __restrict__ lnk,
                                                      const NodeAux* __restrict__ aux, SampRec* __restrict__ rec)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = vs->nc;
    if (nc <= 1 || p >= nc - 1) return;
    const NodeLnk lk = lnk[p];
    SampRec r;
    r.l = samp_child(lk.left, geo, lnk, aux);
    r.r = samp_child(lk.right, geo, lnk, aux);
    rec[p] = r;
}

// Sampler: one lane per pixel, 8x8 pixel tile per wavefront.
//
// Every FGT weight the descents need is a sequential sum over a subset of the
// frontier of the root's walk (nodes with degree >= 0, or truncated ones that
// contribute 0): the walk from any node n stops exactly where the root's walk
// stops inside n's subtree, because the stop rule depends only on the node and
// the pixel.  So one depth-first pass over the root's walk, adding each
// frontier value to the accumulators of its ancestors in traversal order,
// produces W(n) for every node on it bit for bit as the reference's
// per-call walks (evaluateDynamicRadixTree.cuh:128-174) would.  The pass
// records, for each walked node with more than k sources, the weights of its
// two children in a trie of descent states; the ceil(sqrt(k)) descents then
// read weights and capacities from the trie and evaluate the few remaining
// weights (below clusters that were not walked) on demand.  The reference
// keeps the k smallest keys in a sorted insertion list (ties keep insertion
// order, sequentialPoissonRadixTree.cuh:15-41); the same set and order are the
// k smallest (key, draw number) words, kept here in a per-lane max-heap in LDS
// (slot-major) and heap-sorted once at the end: O(log k) LDS steps per
// accepted key instead of a lane-divergent O(k) shift.
// Accumulators of the current DFS path, one per depth and lane, kept only for
// the depths a descent compares (the leading nodes with more than k sources and
// the first one below them): the first kDepthLds depths in LDS ([depth][lane],
// indices are wave-uniform), deeper ones (rare: about 20 levels cover k = 40 at
// 1M splats) in private memory.
constexpr int kDepthLds = 20;
struct DepthF {
    float* lds;                     // this wavefront's [kDepthLds][64] slice, offset by the lane
    float a[kMaxDepth - kDepthLds];
    __device__ float get(int i) const { return i < kDepthLds ? lds[i * 64] : a[i - kDepthLds]; }
    __device__ void set(int i, float v)
    {
        if (i < kDepthLds) lds[i * 64] = v;
        else a[i - kDepthLds] = v;
    }
};

// An entry holds the weights of its node's two children and their capacities, so
// a descent step reads everything it compares in one round of loads:
// field 0 the left child's capacity, 1 / 2 the left / right weight, 3 the child
// entries, 4 the right child's capacity.  Each wavefront owns one contiguous
// block of the workspace ([tile][entry][field][lane]): lanes that walk the same
// nodes create their entries in step, so their accesses coalesce, and a
// wavefront's whole trie (under 1 MB) sits in one or two pages, so its loads do
// not miss the address-translation caches (with the fields of all tiles
// interleaved, consecutive fields of one entry lay 1 MB apart).
struct TrieRef {
    float UD_GLOBAL* base;          // this lane's slot in its tile's block
    __device__ float get(int e, int f) const { return base[(kTrieFields * e + f) * 64]; }
    __device__ void set(int e, int f, float v) const { base[(kTrieFields * e + f) * 64] = v; }
    __device__ int kids(int e) const { return __float_as_int(get(e, 3)); }
    __device__ void setKids(int e, int v) const { set(e, 3, __int_as_float(v)); }
    // one half of field 3 (left child entry + 1 in the low half, right in the high
    // half), stored alone so that no load of the other half is needed
    __device__ void setKidHalf(int e, bool right, int v) const
    {
        unsigned short UD_GLOBAL* h = (unsigned short UD_GLOBAL*)(base + (kTrieFields * e + 3) * 64);
        h[right ? 1 : 0] = (unsigned short)v;
    }
    __device__ void init(int e, float wUnset) const
    {
        set(e, 0, 1.0f); set(e, 1, wUnset); set(e, 2, wUnset); setKids(e, 0); set(e, 4, 1.0f);
    }
    // an entry of the union walk: both children's weights are stored for every
    // lane when the walk leaves them, so only capacities and child links start here
    __device__ void initCaps(int e) const { set(e, 0, 1.0f); setKids(e, 0); set(e, 4, 1.0f); }
};

// Replace the top of a per-lane max-heap of n words (LDS, slot-major
// [slot][lane]) by x and restore the heap order.
__device__ __forceinline__ void heap_replace_top(unsigned long long* list, int lane, int n, unsigned long long x)
{
    int i = 0;
    while (true) {
        int c = 2 * i + 1;
        if (c >= n) break;
        unsigned long long vc = list[c * 64 + lane];
        if (c + 1 < n) {
            unsigned long long vr = list[(c + 1) * 64 + lane];
            if (vr > vc) { vc = vr; c = c + 1; }
        }
        if (vc <= x) break;
        list[i * 64 + lane] = vc;
        i = c;
    }
    list[i * 64 + lane] = x;
}

__global__ __launch_bounds__(64) void k_sample_walk(int W, int H, int k, int L, const ViewState* __restrict__ vs,
                                                    const NodeGeo* geo, const NodeLnk* lnk, const NodeAux* aux,
                                                    const float* coef, const SampRec* __restrict__ srec,
                                                    float* trieWs, int tcap,
                                                    int2* __restrict__ reach, int* __restrict__ errOut)
{
    TreeView T;
    T.geo = (const NodeGeo UD_GLOBAL*)geo;
    T.lnk = (const NodeLnk UD_GLOBAL*)lnk;
    T.aux = (const NodeAux UD_GLOBAL*)aux;
    T.coef = (const float UD_GLOBAL*)coef;
    T.s2h = vs->s2h;
    T.is2h = vs->is2h;
    const int TMAX = tcap;
    const int DFSMAX = tcap - L * 32;                 // entries the walk may create
    const int lane = threadIdx.x;
    TrieRef tr;
    tr.base = (float UD_GLOBAL*)trieWs + (long)blockIdx.x * kTrieFields * tcap * 64 + lane;
    const int tilesY = (H + 7) >> 3;
    const uint32_t tile = xcd_chunk_remap<kSampTileRun>(blockIdx.x, gridDim.x);
    const int ti = tile / tilesY, tj = tile - (tile / tilesY) * tilesY;
    const int i = ti * 8 + (lane >> 3), j = tj * 8 + (lane & 7);
    const bool valid = (i < W) && (j < H);
    int2* rch = reach + (long)blockIdx.x * L * 64 + lane;   // [launch][lane] of this tile
    int drawn = 0;                                      // launches whose leaf range is recorded
    const int n = vs->nc;
    const float prec = vs->prec;
    const bool runTree = n > 0 && vs->err == 0;          // wave-uniform
    const int L0 = n - 1;
    const float qx = (float)i, qy = (float)j;
    const float unset = __builtin_nanf("");              // weight not computed yet
    // trie fields: cap, wl, wr, kids = (left entry + 1) | (right entry + 1) << 16
    int nt = 1;
    bool bad = false;
    float total = 0.0f;
    if (runTree) {
        tr.init(0, unset);                                // entry 0: the root
        // ---- one depth-first pass over the union of the 64 lanes' FGT walks ----
        // The wavefront visits nodes in the reference's left-to-right order; a lane
        // is active inside a node exactly when its own walk enters that node, so
        // each lane sees its own frontier in its own order.  A node's fields come
        // with its parent's children record (SampRec): a descent loads one record,
        // a truncated node or a right sibling popped from the stack loads nothing,
        // an expanded node only its coefficients.  Trie entries are numbered for
        // the wavefront (a big node any lane walks gets one; a lane that does not
        // walk it finds unset weights there, as in an entry it would create), so
        // the path's entries are wave-uniform and every trie store coalesces.
        __shared__ SampChild s_rinfo[kMaxDepth];           // right siblings of the path
        __shared__ unsigned long long s_walk[kMaxDepth];
        __shared__ int s_pent[kMaxDepth];                  // trie entry of each path node, -1 none
        __shared__ float s_coef[(kMaxDegree + 1) * (kMaxDegree + 1)];
        __shared__ float s_pacc[kDepthLds * 64];
        DepthF pacc;
        pacc.lds = s_pacc + lane;
        int big = 0;                 // leading path nodes with more than k sources
        uint64_t rightBits = 0;      // bit d: node at depth d is its parent's right child
        int d = 0;
        bool act = valid;
        bool overflow = false;
        SampChild cur;               // the node at depth d (wave-uniform)
        {
            const NodeGeo g0 = ld_geo(T, 0);
            cur.lox = g0.lox; cur.loy = g0.loy; cur.hix = g0.hix; cur.hiy = g0.hiy;
            cur.cx = g0.cx; cur.cy = g0.cy; cur.r2 = g0.r2; cur.deg = g0.deg;
            cur.off = T.aux[0].off; cur.count = T.lnk[0].count; cur.id = 0; cur.pad = 0;
        }
        // a node's weight into its parent's entry (the root's: the total)
        // a completed node's weight into its parent's entry, for every lane: the
        // lanes that reached it store theirs, the others the unset marker (their
        // descents evaluate it on demand); the root's weight is the total
        auto report = [&](int dd, bool has, float w) {
            if (dd == 0) { if (has) total = w; return; }
            const int pe = s_pent[dd - 1];
            if (pe >= 0) tr.set(pe, ((rightBits >> dd) & 1ull) ? 2 : 1, has ? w : unset);
        };
        auto uni = [](float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); };
        while (true) {
            // cur is wave-uniform (the root's record, a children record, or the LDS
            // stack); say so, so that it lives in scalar registers
            cur.id = __builtin_amdgcn_readfirstlane(cur.id);
            cur.deg = __builtin_amdgcn_readfirstlane(cur.deg);
            cur.off = __builtin_amdgcn_readfirstlane(cur.off);
            cur.count = __builtin_amdgcn_readfirstlane(cur.count);
            NodeGeo g;
            g.lox = uni(cur.lox); g.loy = uni(cur.loy); g.hix = uni(cur.hix); g.hiy = uni(cur.hiy);
            g.cx = uni(cur.cx); g.cy = uni(cur.cy); g.r2 = uni(cur.r2); g.deg = cur.deg;
            float wv = 0.0f;
            bool walk = false;
            if (g.deg >= 0) {
                wv = eval_node_staged(T, g, cur.off, qx, qy, act, s_coef, lane);
                if (act) {
                    // ancestors whose weight a descent can compare: depth 0 and every
                    // node whose parent has more than k sources (a prefix of the path,
                    // since source counts shrink with depth)
                    const int amax = d < big + 1 ? d : big + 1;
                    for (int a = 0; a < amax; a++) pacc.set(a, pacc.get(a) + wv);
                }
            } else {
                walk = act && !truncated(g, qx, qy);
            }
            unsigned long long wm = __ballot(walk);
            if (wm != 0ull) {
                if (d >= kMaxDepth - 1) { overflow = true; break; }
                const SampRec R = srec[cur.id];           // the one load of a descent
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) { s_rinfo[d] = R.r; s_walk[d] = wm; }
                int e = -1;
                if (d == 0) e = 0;
                else if (cur.count > k && s_pent[d - 1] >= 0 && nt < DFSMAX) {
                    e = nt++;
                    tr.initCaps(e);
                    tr.setKidHalf(s_pent[d - 1], (rightBits >> d) & 1ull, e + 1);
                }
                if (lane == 0) s_pent[d] = e;
                __builtin_amdgcn_wave_barrier();
                if (walk && d <= big) pacc.set(d, 0.0f);
                if (big == d && cur.count > k) big = d + 1;
                d++;
                cur = R.l;
                rightBits &= ~(1ull << d);
                act = walk;
                continue;
            }
            // the node is complete for every active lane (wv: its expansion, or 0
            // where it is truncated); report and advance
            report(d, act, wv);
            bool done = false;
            while (true) {
                if (d == 0) { done = true; break; }
                if (!((rightBits >> d) & 1ull)) {
                    rightBits |= (1ull << d);
                    __builtin_amdgcn_wave_barrier();
                    cur = s_rinfo[d - 1];
                    act = (s_walk[d - 1] >> lane) & 1ull;
                    break;
                }
                d--;
                if (big > d) big = d;
                act = (s_walk[d] >> lane) & 1ull;
                // its accumulator is kept exactly when its parent has an entry; the
                // lanes that reached it without walking it (truncated) report 0
                if (d == 0 || s_pent[d - 1] >= 0) {
                    const bool reached = d == 0 ? act : ((s_walk[d - 1] >> lane) & 1ull) != 0;
                    report(d, reached, act ? pacc.get(d) : 0.0f);
                }
            }
            if (done) break;
        }
        if (overflow) bad = true;
    }
    if (valid && runTree) {
        total = fmaxf(0.0f, total);
        if (!bad && total >= prec) {
            int pe[kMaxDepth];
            float pp[kMaxDepth], pc[kMaxDepth];
            for (int la = 0; la < L && !bad; la++) {
                int e = 0, nd = 0, depth = 0;
                uint64_t sideBits = 0;         // bit d: the step to depth d went right
                pe[0] = 0;
                NodeLnk lk = ld_lnk(T, nd);
                while (lk.count > k) {
                    // one round of loads: the node's link (issued at the end of the last
                    // step) and the entry's weights, children and capacities
                    float wl = tr.get(e, 1), wr = tr.get(e, 2);
                    const int kids = tr.kids(e);
                    const float cl = tr.get(e, 0), cr = tr.get(e, 4);
                    if (wl != wl) {
                        wl = tree_weight(T, lk.left, qx, qy);
                        wr = tree_weight(T, lk.right, qx, qy);
                        tr.set(e, 1, wl);
                        tr.set(e, 2, wr);
                    }
                    wl = fmaxf(0.0f, wl); wr = fmaxf(0.0f, wr);
                    float sum = wl + wr;
                    wl = wl / sum; wr = wr / sum;
                    int el = (kids & 0xffff) - 1, er = (kids >> 16) - 1;
                    float sl = wl * cl, sr = wr * cr;
                    bool goRight = sl < sr;
                    int ce = goRight ? er : el;
                    int cnode = goRight ? lk.right : lk.left;
                    if (ce < 0) {
                        if (nt >= TMAX) { bad = true; break; }
                        ce = nt++;
                        tr.init(ce, unset);
                        tr.setKidHalf(e, goRight, ce + 1);
                    }
                    depth++;
                    if (depth >= kMaxDepth) { bad = true; break; }
                    pe[depth] = ce;
                    pp[depth] = goRight ? wr : wl;
                    pc[depth] = goRight ? cr : cl;     // its capacity, unchanged until the update below
                    if (goRight) sideBits |= 1ull << depth;
                    e = ce;
                    nd = cnode;
                    lk = ld_lnk(T, nd);
                }
                if (bad) break;
                // capacities: an entry's lives in its parent's entry (field 4 if it is
                // the right child, else 0); the path's were read on the way down
                float prob = 1.0f;
                for (int dd = depth; dd >= 1; dd--) {
                    const int par = pe[dd - 1];
                    const int fld = ((sideBits >> dd) & 1ull) ? 4 : 0;
                    tr.set(par, fld, fmaxf(0.0f, pc[dd] - prob));
                    prob = prob * pp[dd];
                }
                // the reached subtree's leaves are the contiguous sorted range
                // [first, first + count); k_sample_draw keys them
                const int cnt = lk.count;
                const int first = (nd >= L0) ? nd - L0 : T.aux[nd].first;
                rch[(long)la * 64] = make_int2(first, cnt);
                drawn = la + 1;
            }
        }
        if (bad) atomicOr(errOut, 1);
    }
    for (int la = drawn; la < L; la++) rch[(long)la * 64] = make_int2(0, 0);
}

// The key word of one candidate: (u / E(leaf) bits, draw number (launch << 16 |
// ordinal)); false where the key is not finite (E = 0: never inserted).  Keys are
// >= 0, so their bit patterns order like the floats.
__device__ __forceinline__ bool draw_word(float qx, float qy, float2 xy, float is2h, uint32_t x, int la, int ord,
                                          unsigned long long& ent)
{
    float tx = (qx - xy.x) * is2h;
    float ty = (qy - xy.y) * is2h;
    float ex = pexpf(-(tx * tx)), ey = pexpf(-(ty * ty));
    float wgt = 0.0f;
    wgt = wgt + (1.0f * ex) * ey;
    float p = fmaxf(0.0f, wgt);
    float key = u01(x) / p;
    if (!(key < __builtin_huge_valf())) return false;
    ent = ((unsigned long long)__float_as_uint(key) << 32) | (uint32_t)((la << 16) | ord);
    return true;
}

// Sequential Poisson keys over the leaves each launch reached (the walk's
// capacity descents do not depend on the keys, so they run first, in
// k_sample_walk, at the occupancy their long dependent loads need): in a Karras
// tree the leaves of a subtree are the contiguous sorted range [first, first +
// count), in the left-to-right order nextLeaf visits
// (sequentialPoissonDynamicTree.cuh:16-53), and every leaf carries the single
// Hermite coefficient 1, so E(leaf) = e^{-tx^2} e^{-ty^2}.  Key u / E with u the
// Philox word of (leaf ordinal, launch, pixel, view).  The reference keeps the
// keys in a sorted list of k slots, inserting with a strict < and carrying each
// displaced entry down (insertSample, sequentialPoissonRadixTree.cuh:15-41).
// With distinct keys that list is the k smallest (key, draw number) words in
// order, and they are kept here in a per-lane bounded max-heap in LDS,
// heap-sorted at the end.  Equal keys are where the two differ: a carried entry
// passes over entries equal to it, so an insertion ahead of a run of equal keys
// moves the run's first entry to the run's end, and a full list drops the front
// of its last run.  That matters exactly when two equal keys are in the heap at
// once: the sorted heap holds two equal neighbours, or an entry it dropped ties
// with the final maximum -- dropped keys never increase, so it is enough that
// the last one dropped equals the final maximum (a run dropped whole leaves the
// sample as the heap has it).  Such a lane (rare: none in 16,384 pixels of the
// 1M-splat frame) queues itself, and k_sample_replay redraws it through the
// reference's insertion.  (A replay path inside this kernel, never
// taken, cost 45 % of its time in code generation.)
__global__ __launch_bounds__(64) void k_sample_draw(int W, int H, int k, int L, uint32_t seedLo, uint32_t seedHi,
                                                    uint32_t view, const ViewState* __restrict__ vs,
                                                    const float2* __restrict__ leafXY,
                                                    const int32_t* __restrict__ leafId,
                                                    const int2* __restrict__ reach, int32_t* __restrict__ out,
                                                    int* __restrict__ tieQueue)
{
    extern __shared__ unsigned long long list[];
    const int lane = threadIdx.x;
    const int tilesY = (H + 7) >> 3;
    const uint32_t tile = xcd_chunk_remap<kSampTileRun>(blockIdx.x, gridDim.x);
    const int ti = tile / tilesY, tj = tile - (tile / tilesY) * tilesY;
    const int i = ti * 8 + (lane >> 3), j = tj * 8 + (lane & 7);
    const bool valid = (i < W) && (j < H);
    if (!valid) return;
    // the padding word (key +inf, draw 0xffffffff) is larger than any real entry
    const unsigned long long empty = ((unsigned long long)0x7f800000u << 32) | 0xffffffffull;
    for (int s = 0; s < k; s++) list[s * 64 + lane] = empty;
    const float is2h = vs->is2h;
    const float qx = (float)i, qy = (float)j;
    const uint32_t pixel = (uint32_t)(j + H * i);
    const int2* rch = reach + (long)blockIdx.x * L * 64 + lane;
    int firstPos[16];                                   // per launch: first leaf position drawn
    unsigned long long lastDrop = empty;               // the last word the heap dropped
    // The lane's groups of 4 leaves, launch after launch, in one loop: the
    // wavefront runs as many iterations as its busiest lane has groups (a loop
    // per launch would run the busiest lane of each launch).  Each lane keys
    // the same leaves in the same order as a loop per launch.
    int la = 0, kk0 = 0;
    int2 fc = rch[0];
    firstPos[0] = fc.x;
    while (la < L && fc.y == 0) {
        if (++la < L) { fc = rch[(long)la * 64]; firstPos[la] = fc.x; }
    }
    while (la < L) {
        const int first = fc.x, cnt = fc.y;
        u32x4 c4 = { (uint32_t)(kk0 >> 2), (uint32_t)la, pixel, view };
        u32x4 rnd = philox4x32_10(c4, seedLo, seedHi);
        float2 xy[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int pos = first + kk0 + q;
            bool ok = kk0 + q < cnt;
            xy[q] = ok ? leafXY[pos] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (kk0 + q >= cnt) break;
            unsigned long long ent;
            if (!draw_word(qx, qy, xy[q], is2h, (q == 0) ? rnd.x : (q == 1) ? rnd.y : (q == 2) ? rnd.z : rnd.w,
                           la, kk0 + q, ent))
                continue;
            const unsigned long long top = list[lane];
            if (ent < top) {
                lastDrop = top;
                heap_replace_top(list, lane, k, ent);
            }
        }
        kk0 += 4;
        if (kk0 >= cnt) {
            kk0 = 0;
            do {
                if (++la < L) { fc = rch[(long)la * 64]; firstPos[la] = fc.x; }
            } while (la < L && fc.y == 0);
        }
    }
    // heap -> ascending order (heapsort)
    for (int end = k - 1; end > 0; end--) {
        unsigned long long x = list[end * 64 + lane];
        list[end * 64 + lane] = list[lane];
        heap_replace_top(list, lane, end, x);
    }
    bool tie = lastDrop != empty && key_of(lastDrop) == key_of(list[(k - 1) * 64 + lane]);
    for (int s = 1; s < k && !tie; s++) {
        const unsigned long long b = list[s * 64 + lane];
        tie = b != empty && key_of(b) == key_of(list[(s - 1) * 64 + lane]);
    }
    if (tie) {
        const int e = atomicAdd(tieQueue, 1);
        tieQueue[1 + 2 * e] = (int)blockIdx.x * 64 + lane;
        tieQueue[2 + 2 * e] = (int)(uint32_t)(list[(k - 1) * 64 + lane] >> 32);   // the final maximum key
    }
    // draw numbers -> splat ids
    int32_t* o = out + ((long)j + (long)H * i) * k;
    for (int s = 0; s < k; s++) {
        uint32_t dn = (uint32_t)(list[s * 64 + lane] & 0xffffffffull);
        o[s] = dn == 0xffffffffu ? -1 : leafId[firstPos[dn >> 16] + (int)(dn & 0xffffu)];
    }
}

// The pixels k_sample_draw queued, drawn again through the reference's insertion,
// one wavefront per pixel: list position s lives in row s / 64 of lane s % 64
// (k <= 256: at most 4 rows).  Inserting word c with the strict < of the reference
// puts c at s0, the count of keys <= c's; behind it a displaced entry is carried
// on and passes over entries of its own key, so each run of equal keys behind s0
// keeps its entries in place except its first, which moves to the position after
// the run (and the first entry of the last run leaves a full list; a carried
// empty slot ends the insertion, and the empty slots form the last run).  So
// position s > s0 keeps its entry inside a run and otherwise takes the first
// entry of the run ending at s - 1; s0 takes c.  Candidates are keyed 64 at a
// time, one per lane, and inserted in the reference's order; one whose key is
// above the final list's largest (queued by k_sample_draw: the heap ends with the
// same largest key) is skipped, as it only ever joins runs the list drops whole.
// A persistent grid drains the queue.
constexpr int kReplayRows = 4;
__global__ __launch_bounds__(64) void k_sample_replay(int W, int H, int k, int L, unsigned drawBlocks, uint32_t seedLo,
                                                      uint32_t seedHi, uint32_t view, const ViewState* __restrict__ vs,
                                                      const float2* __restrict__ leafXY,
                                                      const int32_t* __restrict__ leafId,
                                                      const int2* __restrict__ reach, const int* __restrict__ tieQueue,
                                                      int32_t* __restrict__ out)
{
    const int lane = threadIdx.x;
    const int count = tieQueue[0];
    const unsigned long long empty = ((unsigned long long)0x7f800000u << 32) | 0xffffffffull;
    const int R = (k + 63) >> 6;
    for (int e = blockIdx.x; e < count; e += gridDim.x) {
        const int sl = tieQueue[1 + 2 * e];             // sampler lane: draw block * 64 + lane
        const float kmax = __uint_as_float((uint32_t)tieQueue[2 + 2 * e]);
        const int blk = sl >> 6, pl = sl & 63;
        const int tilesY = (H + 7) >> 3;
        const uint32_t tile = xcd_chunk_remap<kSampTileRun>((uint32_t)blk, drawBlocks);
        const int ti = tile / tilesY, tj = tile - (tile / tilesY) * tilesY;
        const int i = ti * 8 + (pl >> 3), j = tj * 8 + (pl & 7);
        const float is2h = vs->is2h;
        const float qx = (float)i, qy = (float)j;
        const uint32_t pixel = (uint32_t)(j + H * i);
        const int2* rch = reach + (long)blk * L * 64 + pl;
        unsigned long long v[kReplayRows];
#pragma unroll
        for (int r = 0; r < kReplayRows; r++) v[r] = empty;
        for (int la = 0; la < L; la++) {
            const int2 fc = rch[(long)la * 64];
            const int first = fc.x, cnt = fc.y;
            for (int c0 = 0; c0 < cnt; c0 += 64) {
                // this lane's candidate of the chunk
                const int kk = c0 + lane;
                unsigned long long mine = empty;
                bool okc = false;
                if (kk < cnt) {
                    const u32x4 rnd = philox4x32_10(u32x4{ (uint32_t)(kk >> 2), (uint32_t)la, pixel, view }, seedLo,
                                                    seedHi);
                    const int q = kk & 3;
                    okc = draw_word(qx, qy, leafXY[first + kk], is2h,
                                    (q == 0) ? rnd.x : (q == 1) ? rnd.y : (q == 2) ? rnd.z : rnd.w, la, kk, mine) &&
                          !(key_of(mine) > kmax);
                }
                const unsigned long long okm = __ballot(okc);
                const int nc = cnt - c0 < 64 ? cnt - c0 : 64;
                for (int t = 0; t < nc; t++) {
                    if (!((okm >> t) & 1ull)) continue;       // key not finite (never inserted) or above kmax
                    const unsigned long long c = __shfl(mine, t, 64);
                    const float kc = key_of(c);
                    int s0 = 0;
#pragma unroll
                    for (int r = 0; r < kReplayRows; r++)
                        if (r < R) s0 += __popcll(__ballot(r * 64 + lane < k && !(kc < key_of(v[r]))));
                    if (s0 >= k) continue;                                // not below any key: not inserted
                    // run starts at or after s0 (s0 itself is one), prefix maximum over positions
                    // every shuffle runs on all 64 lanes; lane 0 then takes row r - 1's
                    // lane 63 (or nothing in row 0) as its predecessor
                    int start[kReplayRows];
                    int carry = -1;
#pragma unroll
                    for (int r = 0; r < kReplayRows; r++) {
                        if (r >= R) break;                                // wave-uniform
                        const int pos = r * 64 + lane;
                        const unsigned long long up = __shfl_up(v[r], 1, 64);
                        const unsigned long long wrap = __shfl(v[r > 0 ? r - 1 : 0], 63, 64);
                        const unsigned long long prev = lane > 0 ? up : r > 0 ? wrap : empty;
                        const bool st = pos >= s0 && pos < k && (pos == s0 || key_of(v[r]) != key_of(prev));
                        int m = st ? pos : -1;
#pragma unroll
                        for (int o = 1; o < 64; o <<= 1) {
                            const int u = __shfl_up(m, o, 64);
                            if (lane >= o && u > m) m = u;
                        }
                        if (m < carry) m = carry;
                        start[r] = m;
                        carry = __shfl(m, 63, 64);
                    }
                    // the first entry of the run ending at pos - 1: start of pos - 1, then its entry
                    unsigned long long nv[kReplayRows];
#pragma unroll
                    for (int r = 0; r < kReplayRows; r++) {
                        if (r >= R) break;
                        const int pos = r * 64 + lane;
                        const int spUp = __shfl_up(start[r], 1, 64);
                        const int spWrap = __shfl(start[r > 0 ? r - 1 : 0], 63, 64);
                        const int sp = lane > 0 ? spUp : r > 0 ? spWrap : -1;
                        unsigned long long from = empty;
#pragma unroll
                        for (int rr = 0; rr < kReplayRows; rr++) {
                            if (rr >= R) break;
                            const unsigned long long g = __shfl(v[rr], sp & 63, 64);
                            if (sp >= 0 && (sp >> 6) == rr) from = g;
                        }
                        const unsigned long long up = __shfl_up(v[r], 1, 64);
                        const unsigned long long wrap = __shfl(v[r > 0 ? r - 1 : 0], 63, 64);
                        const unsigned long long prev = lane > 0 ? up : r > 0 ? wrap : empty;
                        // inside a run (the empty slots, which end a list that is not full,
                        // are one run: a carried empty slot ends the insertion there)
                        const bool inside = pos > s0 && key_of(v[r]) == key_of(prev);
                        nv[r] = pos >= k ? empty : pos < s0 ? v[r] : pos == s0 ? c : inside ? v[r] : from;
                    }
#pragma unroll
                    for (int r = 0; r < kReplayRows; r++)
                        if (r < R) v[r] = nv[r];
                }
            }
        }
        int32_t* o = out + ((long)j + (long)H * i) * k;
#pragma unroll
        for (int r = 0; r < kReplayRows; r++) {
            const int pos = r * 64 + lane;
            if (pos < k) {
                const uint32_t dn = (uint32_t)(v[r] & 0xffffffffull);
                o[pos] = dn == 0xffffffffu ? -1 : leafId[rch[(long)(dn >> 16) * 64].x + (int)(dn & 0xffffu)];
            }
        }
    }
}

__global__ void k_fill_i32(int* p, long n, int v)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void k_fill_f32(float* p, long n, float v)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void k_fill_idx(int32_t* p, long n)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = -1;
}

__global__ void k_copy_err(const ViewState* vs, int* err)
{
    if (threadIdx.x == 0 && vs->err) atomicOr(err, vs->err);
}

// ---------------------------------------------------------------------------
// Fixed-order (degree 6) Hermite densities D_b(q) = c_b sum_i exp(-|q - s_i|^2 /
// 2 h_b^2) on the shared tree: approximateFMMWeights (band 0, the tree's h;
// sampling.cpp:136-175, samplingCuda.cu:398-493, buildRadixTree.cuh:449-606,
// transformRadixTree.cuh:15-69, 152-222) and the normalisers of Poisson-rate
// sampling (bands 0 and 1).
// ---------------------------------------------------------------------------

// Internal nodes whose degree-6 truncation bound at band b meets the
// precision, and whose parent's does not, carry an expansion: a 49-float
// block at fixOff[node].
__global__ __launch_bounds__(256) void k_fixed_select(ViewState* vs, int nbands, int p, const NodeGeo* __restrict__ geo,
                                                      const NodeLnk* __restrict__ lnk, int* fixOff0, int* fixOff1,
                                                      int* list0, int* list1, VStride vsd)
{
    vs = vw(vs, vsd.ws);
    geo = vw(geo, vsd.ws);
    lnk = vw(lnk, vsd.ws);
    fixOff0 = vw(fixOff0, vsd.ws);
    fixOff1 = vw(fixOff1, vsd.ws);
    list0 = vw(list0, vsd.ws);
    list1 = vw(list1, vsd.ws);
    __shared__ int wcnt[2][4];
    __shared__ int bbase[2];
    const int n = vs->nc;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const float prec = vs->prec;
    bool acc[2] = { false, false };
    unsigned long long m[2] = { 0ull, 0ull };
    for (int b = 0; b < nbands; b++) {
        const float hb = vs->band[b].h;
        auto meets = [&](int nd) {
            NodeGeo g = geo[nd];
            float mn = 0.0f;
            mn = fmaxf(mn, fmaxf(g.cx - g.lox, g.hix - g.cx));
            mn = fmaxf(mn, fmaxf(g.cy - g.loy, g.hiy - g.cy));
            float rr = mn / hb;
            return hermite_error(rr, (float)lnk[nd].count, p) <= prec;
        };
        // the walk from the root evaluates the first node on its path that meets the
        // bound and skips its subtree, so only those nodes need an expansion
        if (t < n - 1) acc[b] = meets(t) && (t == 0 || !meets(lnk[t].parent));
        m[b] = __ballot(acc[b]);
        if (lane == 0) wcnt[b][wid] = __popcll(m[b]);
    }
    __syncthreads();
    // one slot reservation per band and workgroup
    if (threadIdx.x < nbands) {
        const int b = threadIdx.x;
        const int tot = wcnt[b][0] + wcnt[b][1] + wcnt[b][2] + wcnt[b][3];
        bbase[b] = tot ? atomicAdd(&vs->nfix[b], tot) : 0;
    }
    __syncthreads();
    if (t >= n - 1) return;
    for (int b = 0; b < nbands; b++) {
        int base = bbase[b];
        for (int q = 0; q < wid; q++) base += wcnt[b][q];
        const int slot = base + __popcll(m[b] & ((1ull << lane) - 1ull));
        int* fixOff = b ? fixOff1 : fixOff0;
        int* list = b ? list1 : list0;
        fixOff[t] = acc[b] ? slot * (p + 1) * (p + 1) : -1;
        if (acc[b]) list[slot] = t;
    }
}

// Degree-P moments at band b, scaled by the band's leaf coefficient: thread
// (node, row a) sums row a over the node's leaf range in FP64 (leaf order).
// 256 / (P + 1) nodes x (P + 1) rows per 256-thread workgroup.
template <int P>
__global__ __launch_bounds__(256) void k_fixed_moments(const ViewState* __restrict__ vs, int b,
                                                       const NodeGeo* __restrict__ geo, const NodeAux* __restrict__ aux,
                                                       const int* __restrict__ list, const int* __restrict__ fixOff,
                                                       float* __restrict__ coef, VStride vsd)
{
    vs = vw(vs, vsd.ws);
    geo = vw(geo, vsd.ws);
    aux = vw(aux, vsd.ws);
    list = vw(list, vsd.ws);
    fixOff = vw(fixOff, vsd.ws);
    coef = vw(coef, vsd.ws);
    const int t = threadIdx.x;
    if (t >= (256 / (P + 1)) * (P + 1)) return;
    const long slot = (long)blockIdx.x * (256 / (P + 1)) + t / (P + 1);
    const int ax = t % (P + 1);
    if (slot >= vs->nfix[b]) return;
    const int nd = list[slot];
    const int L0 = vs->nc - 1;
    const float is2h = vs->band[b].is2h;
    const double c = (double)vs->band[b].c;
    NodeGeo g = geo[nd];
    NodeAux a = aux[nd];
    const int off = fixOff[nd];
    const double ifx = kInvFact[ax];
    double acc[P + 1];
#pragma unroll
    for (int bb = 0; bb <= P; bb++) acc[bb] = 0.0;
    for (int pos = a.first; pos <= a.last; pos++) {
        NodeGeo lf = geo[L0 + pos];
        float tx = (lf.cx - g.cx) * is2h;
        float ty = (lf.cy - g.cy) * is2h;
        double px = 1.0;
        for (int q = 0; q < ax; q++) px = px * (double)tx;
        double mx = px * ifx;
        double py = 1.0;
#pragma unroll
        for (int bb = 0; bb <= P; bb++) {
            acc[bb] += mx * (py * ifact_c(bb));
            py = py * (double)ty;
        }
    }
#pragma unroll
    for (int bb = 0; bb <= P; bb++)
        coef[off + ax * (P + 1) + bb] = (float)(c * acc[bb]);
}

__device__ __forceinline__ bool box_meets_disc(const NodeGeo& g, float qx, float qy, float rad)
{
    float cx = fminf(fmaxf(qx, g.lox), g.hix);
    float cy = fminf(fmaxf(qy, g.loy), g.hiy);
    float dx = qx - cx, dy = qy - cy;
    float d2 = 0.0f;
    d2 = d2 + dx * dx;
    d2 = d2 + dy * dy;
    return d2 < rad * rad;
}

// c exp(-tx^2) exp(-ty^2), t = (q - x) / sqrt(2 h^2), in leaf-evaluation order
__device__ __forceinline__ float band_point(const Band& B, float qx, float qy, float x, float y)
{
    float tx = (qx - x) * B.is2h;
    float ty = (qy - y) * B.is2h;
    float ex = pexpf(-(tx * tx)), ey = pexpf(-(ty * ty));
    float v = 0.0f;
    v = v + (B.c * ex) * ey;
    return v;
}

// ---------------------------------------------------------------------------
// Launch shape of the per-pixel passes (k_fmm_eval, k_poisson): one wavefront
// per pixel tile of tw x 2^lth pixels (tw 2^lth = 64), wpb wavefronts per
// workgroup, and optionally an XCD-grouped block order: workgroups are handed
// to the 8 XCDs round-robin, so block b runs on XCD b mod 8; with xcd set,
// block b takes the tile range of position b / 8 within XCD b mod 8's
// contiguous share, so each XCD's L2 serves one band of the image (and of the
// leaves under it).  A pixel's result does not depend on the shape.
// The views of a batch are blockIdx.y; their workspaces lie vstride bytes apart.
// ---------------------------------------------------------------------------
struct TileCfg {
    int tw, lth, wpb, xcd;
};
constexpr int kMaxWavesPerBlock = 8;


// pixel (i, j) of this lane; false for a wavefront past the last tile (wave-uniform)
__device__ __forceinline__ bool tile_lane(const TileCfg& c, int W, int H, int lane, int& i, int& j)
{
    const int blk = c.xcd ? (int)xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int tile = blk * c.wpb + (int)(threadIdx.x >> 6);
    const int th = 1 << c.lth;
    const int tilesY = (H + th - 1) >> c.lth;
    const int ntiles = ((W + c.tw - 1) / c.tw) * tilesY;
    if (tile >= ntiles) return false;
    const int ti = tile / tilesY, tj = tile - ti * tilesY;
    i = ti * c.tw + (lane >> c.lth);
    j = (tj << c.lth) + (lane & (th - 1));
    return true;
}


constexpr int kFmmStack = 136;

// children record of internal node p for the density walk (after the expansions
// of every band are selected)
__device__ __forceinline__ FmmRec fmm_record(int p, int L0, const NodeGeo* __restrict__ geo,
                                             const NodeLnk* __restrict__ lnk, const int* __restrict__ fixOff0,
                                             const int* __restrict__ fixOff1)
{
    const NodeLnk lk = lnk[p];
    const NodeGeo gl = geo[lk.left], gr = geo[lk.right];
    FmmRec r;
    r.lbox[0] = gl.lox; r.lbox[1] = gl.loy; r.lbox[2] = gl.hix; r.lbox[3] = gl.hiy;
    r.rbox[0] = gr.lox; r.rbox[1] = gr.loy; r.rbox[2] = gr.hix; r.rbox[3] = gr.hiy;
    r.lc[0] = gl.cx; r.lc[1] = gl.cy; r.rc[0] = gr.cx; r.rc[1] = gr.cy;
    r.left = lk.left; r.right = lk.right;
    r.lfix[0] = lk.left < L0 ? fixOff0[lk.left] : -1;
    r.rfix[0] = lk.right < L0 ? fixOff0[lk.right] : -1;
    r.lfix[1] = (fixOff1 && lk.left < L0) ? fixOff1[lk.left] : -1;
    r.rfix[1] = (fixOff1 && lk.right < L0) ? fixOff1[lk.right] : -1;
    r.pad[0] = r.pad[1] = 0;
    return r;
}

// children records of the internal nodes for the density walk
__global__ __launch_bounds__(256) void k_fmm_records(const ViewState* __restrict__ vs, const NodeGeo* __restrict__ geo,
                                                     const NodeLnk* __restrict__ lnk,
                                                     const int* __restrict__ fixOff0,
                                                     const int* __restrict__ fixOff1, FmmRec* __restrict__ rec, VStride vsd)
{
    vs = vw(vs, vsd.ws);
    geo = vw(geo, vsd.ws);
    lnk = vw(lnk, vsd.ws);
    fixOff0 = vw(fixOff0, vsd.ws);
    fixOff1 = vw(fixOff1, vsd.ws);
    rec = vw(rec, vsd.ws);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int L0 = vs->nc - 1;
    if (p >= L0) return;
    rec[p] = fmm_record(p, L0, geo, lnk, fixOff0, fixOff1);
}

__device__ __forceinline__ bool box4_meets_disc(const float* bx, float qx, float qy, float rad)
{
    float cx = fminf(fmaxf(qx, bx[0]), bx[2]);
    float cy = fminf(fmaxf(qy, bx[1]), bx[3]);
    float dx = qx - cx, dy = qy - cy;
    float d2 = 0.0f;
    d2 = d2 + dx * dx;
    d2 = d2 + dy * dy;
    return d2 < rad * rad;
}

// One lane per pixel, one pixel tile per wavefront (TileCfg), bands on
// blockIdx.y (nbands per view).
// The reference walks each pixel with an explicit stack: pop a node; evaluate
// it if it is a leaf or carries an expansion, else push its left child and then
// its right child, each only if its box meets the query disc
// (transformRadixTree.cuh:15-69).  Here the wavefront runs that walk once for
// the union of its 64 pixels: stack entries carry the mask of lanes for which
// the node was pushed, so each lane sees its own nodes in its own order (the
// shared stack is its stack with other lanes' entries interleaved, in the same
// last-in-first-out order) and sums bit for bit like the reference.  A pushed
// entry also carries the child's centre and expansion offset from its parent's
// FmmRec, so leaves and expansions need no node load and a descent needs one
// wave-uniform record load.  An internal entry just below the top is handled
// in the same step as the top entry (two record loads in flight, or a record
// load under a leaf's or expansion's term), its children pushed in the order
// that popping one after the other leaves, so the accumulation order is
// unchanged.  Karras trees on 64-bit keys are at most 64 levels deep; a level
// keeps at most two entries pending, so the stack stays within 2 * 64 + 4.
// A fixed-order expansion at scaled offset t = (q - c) / sqrt(2 h^2):
// sum_a h_a(tx) sum_b C_ab h_b(ty), Hermite functions by their three-term
// recurrence, every product-sum one fused multiply-add (the oracle uses fmaf in
// the same order).
template <int P>
__device__ __forceinline__ float fixed_series(const float* co, float tx, float ty)
{
    const float ex = pexpf(-(tx * tx)), ey = pexpf(-(ty * ty));
    float hx[P + 1], hy[P + 1];
    hx[0] = ex; hy[0] = ey;
    hx[1] = (2.0f * tx) * ex; hy[1] = (2.0f * ty) * ey;
#pragma unroll
    for (int q = 2; q <= P; q++) {
        hx[q] = __builtin_fmaf(2.0f * tx, hx[q - 1], -((2.0f * (float)(q - 1)) * hx[q - 2]));
        hy[q] = __builtin_fmaf(2.0f * ty, hy[q - 1], -((2.0f * (float)(q - 1)) * hy[q - 2]));
    }
    float val = 0.0f;
#pragma unroll
    for (int a = 0; a <= P; a++) {
        float sa = 0.0f;
#pragma unroll
        for (int b = 0; b <= P; b++) sa = __builtin_fmaf(co[a * (P + 1) + b], hy[b], sa);
        val = __builtin_fmaf(hx[a], sa, val);
    }
    return val;
}

template <int P>
__global__ __launch_bounds__(64 * kMaxWavesPerBlock) void k_fmm_eval(TileCfg cfg, int W, int H,
                                                                    const ViewState* __restrict__ vs, size_t vstride,
                                                                    int nbands, const NodeGeo* geo,
                                                                    const FmmRec* __restrict__ frec,
                                                                    const int* __restrict__ fixOff0,
                                                                    const int* __restrict__ fixOff1,
                                                                    const float* coef0, const float* coef1,
                                                                    float* __restrict__ out, size_t outStride,
                                                                    size_t bandStride)
{
    __shared__ int stkN[kMaxWavesPerBlock][kFmmStack];
    __shared__ int stkF[kMaxWavesPerBlock][kFmmStack];
    __shared__ float2 stkC[kMaxWavesPerBlock][kFmmStack];
    __shared__ unsigned long long stkM[kMaxWavesPerBlock][kFmmStack];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    int i, j;
    if (!tile_lane(cfg, W, H, lane, i, j)) return;           // wave-uniform
    const int b = (int)blockIdx.y % nbands;                  // band
    const size_t v = blockIdx.y / nbands;                    // view
    vs = (const ViewState*)((const char*)vs + v * vstride);
    geo = (const NodeGeo*)((const char*)geo + v * vstride);
    frec = (const FmmRec*)((const char*)frec + v * vstride);
    const int* fixOff = (const int*)((const char*)(b ? fixOff1 : fixOff0) + v * vstride);
    const float* coef = (const float*)((const char*)(b ? coef1 : coef0) + v * vstride);
    out = (float*)((char*)out + v * outStride + (size_t)b * bandStride);
    const bool valid = i < W && j < H;
    const int n = vs->nc;
    float w = 0.0f;
    if (n > 0) {
        const float UD_GLOBAL* cf = (const float UD_GLOBAL*)coef;
        const Band B = vs->band[b];
        const int L0 = n - 1;
        const float qx = (float)i, qy = (float)j;
        int* sN = stkN[wid];
        int* sF = stkF[wid];
        float2* sC = stkC[wid];
        unsigned long long* sM = stkM[wid];
        const unsigned long long all = __ballot(valid);
        if (lane == 0) {
            sN[0] = 0;
            sF[0] = L0 > 0 ? fixOff[0] : -1;
            sC[0] = make_float2(geo[0].cx, geo[0].cy);
            sM[0] = all;
        }
        int sp = 1;
        while (sp > 0) {
            sp--;
            __builtin_amdgcn_wave_barrier();
            const int nd = __builtin_amdgcn_readfirstlane(sN[sp]);
            const int off = __builtin_amdgcn_readfirstlane(sF[sp]);
            const int sp2 = sp > 0 ? sp - 1 : 0;
            const int nd2 = __builtin_amdgcn_readfirstlane(sN[sp2]);
            const int off2 = __builtin_amdgcn_readfirstlane(sF[sp2]);
            const float2 c = sC[sp];
            const unsigned long long mask = sM[sp];
            const unsigned long long mask2 = sM[sp2];
            __builtin_amdgcn_wave_barrier();
            const bool act = (mask >> lane) & 1ull;
            // the term of a leaf or an expansion on top of the stack
            auto add_top = [&]() {
                if (nd >= L0) {
                    if (act) w = w + band_point(B, qx, qy, c.x, c.y);
                } else {
                    // wave-uniform block, fetched by broadcast vector loads issued
                    // together (faster here than scalar loads or staging in LDS)
                    float co[(P + 1) * (P + 1)];
                    const float UD_GLOBAL* cc = cf + off;
#pragma unroll
                    for (int q = 0; q < (P + 1) * (P + 1); q++) co[q] = cc[q];
                    if (act) w = w + fixed_series<P>(co, (qx - c.x) * B.is2h, (qy - c.y) * B.is2h);
                }
            };
            // An internal entry without an expansion just below the top is taken
            // along: its record loads with the top's record (two internal entries)
            // or while the top's term is evaluated (a leaf or expansion on top,
            // whose term is added first).  The lower entry's children are pushed
            // below the upper entry's: the stack a pop of each in turn leaves.
            const bool topInt = nd < L0 && off < 0;
            const bool lowInt = sp > 0 && nd2 < L0 && off2 < 0;
            if (!topInt && !lowInt) {
                add_top();
                continue;
            }
            if (lowInt) sp--;
            const bool two = topInt && lowInt;
            const int ndA = topInt ? nd : nd2;               // the upper internal entry
            const FmmRec r = frec[ndA];
            const FmmRec r2 = frec[two ? nd2 : ndA];
            if (!topInt) add_top();
            const bool actA = topInt ? act : (bool)((mask2 >> lane) & 1ull);
            const bool act2 = two && ((mask2 >> lane) & 1ull);
            const unsigned long long ml = __ballot(actA && box4_meets_disc(r.lbox, qx, qy, B.radius));
            const unsigned long long mr = __ballot(actA && box4_meets_disc(r.rbox, qx, qy, B.radius));
            const unsigned long long ml2 = __ballot(act2 && box4_meets_disc(r2.lbox, qx, qy, B.radius));
            const unsigned long long mr2 = __ballot(act2 && box4_meets_disc(r2.rbox, qx, qy, B.radius));
            if (sp + 4 > kFmmStack) break;                   // unreachable for trees of depth <= 64
            if (ml2) {
                if (lane == 0) {
                    sN[sp] = r2.left; sF[sp] = b ? r2.lfix[1] : r2.lfix[0];
                    sC[sp] = make_float2(r2.lc[0], r2.lc[1]); sM[sp] = ml2;
                }
                sp++;
            }
            if (mr2) {
                if (lane == 0) {
                    sN[sp] = r2.right; sF[sp] = b ? r2.rfix[1] : r2.rfix[0];
                    sC[sp] = make_float2(r2.rc[0], r2.rc[1]); sM[sp] = mr2;
                }
                sp++;
            }
            if (ml) {
                if (lane == 0) {
                    sN[sp] = r.left; sF[sp] = b ? r.lfix[1] : r.lfix[0];
                    sC[sp] = make_float2(r.lc[0], r.lc[1]); sM[sp] = ml;
                }
                sp++;
            }
            if (mr) {
                if (lane == 0) {
                    sN[sp] = r.right; sF[sp] = b ? r.rfix[1] : r.rfix[0];
                    sC[sp] = make_float2(r.rc[0], r.rc[1]); sM[sp] = mr;
                }
                sp++;
            }
        }
    }
    if (valid) out[(long)j + (long)H * i] = w;
}

// ---------------------------------------------------------------------------
// Poisson-rate sampling with inclusion probabilities (no reference kernel; its
// dead plain-Poisson kernel, poissonRadixTree.cuh:70-200, draws Bernoulli(k p)
// only inside the reached subtree, which leaves the inclusion probabilities
// unknown).  Every culled splat i is included at pixel q independently with the
// defensive-mixture probability (Hesterberg 1995)
//   pi_i = min(1, lambda p_i(q)),  p_i = (g0_i(q) / N0(q) + g1_i(q) / N1(q)) / 2,
// g_b = c_b exp(-|q - s_i|^2 / 2 h_b^2) for the tree bandwidth h_0 = h and the
// wide bandwidth h_1 = max(h, 99.9th percentile of the visible splats' own
// bandwidths sqrt(lambda_max(V_i))), and N_b = max(D_b, c_b) with D_b the
// fixed-order density of band b: the expected count stays near lambda, and EWA
// weight / pi stays bounded for splats not wider than h_1.  Pixels with
// D_0 < precision draw nothing.  The support is the set of pairs with
// p >= 1e-4 (independent of the rate), so 1 / pi <= 1e4 / lambda: far tails of
// splats wider than h_1, whose weights there would be multiplied by huge
// 1 / pi, are outside it.
// ---------------------------------------------------------------------------
constexpr float kProposalHCap = 16.0f;
constexpr float kProposalPMin = 1e-4f;   // support: pairs whose rate-free proposal p >= 1e-4

// histogram of the visible splats' proposal bandwidths (1/16 px bins, the
// last bin holds everything >= 16; degenerate splats are skipped)
__global__ __launch_bounds__(256) void k_wide_hist(ViewState* vs, const int32_t* __restrict__ ids,
                                                   const float* __restrict__ hraw, VStride vsd)
{
    vs = vw(vs, vsd.ws);
    ids = vw(ids, vsd.ws);
    hraw = vw(hraw, vsd.ws);
    __shared__ int sh[256];
    sh[threadIdx.x] = 0;
    __syncthreads();
    const int nc = vs->nc;
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < nc; pos += gridDim.x * blockDim.x) {
        float hi = hraw[ids[pos]];
        if (hi >= 0.0f) atomicAdd(&sh[hi < kProposalHCap ? (int)(hi * 16.0f) : 255], 1);
    }
    __syncthreads();
    if (sh[threadIdx.x]) atomicAdd(&vs->hhist[threadIdx.x], sh[threadIdx.x]);
}

// wide band: h_1 = max(h, 99.9th percentile of the histogram, rounded up to 1/16 px)
// (one workgroup of 256 threads, a thread per bin: the chosen bin is the first
// whose inclusive count reaches the target, 0 when the histogram is empty)
__global__ __launch_bounds__(256) void k_wide_band(ViewState* vs, VStride vsd)
{
    vs = vw(vs, vsd.ws);
    __shared__ int wsum[4];
    __shared__ int qbs;
    const int t = threadIdx.x, wid = t >> 6;
    const int c = vs->hhist[t];
    const int inc = wave_incl_scan(c);
    if ((t & 63) == 63) wsum[wid] = inc;
    if (t == 0) qbs = 0;
    __syncthreads();
    int base = 0;
    for (int q = 0; q < wid; q++) base += wsum[q];
    const long cum = (long)base + inc;
    const long total = (long)wsum[0] + wsum[1] + wsum[2] + wsum[3];
    const long target = total - total / 1000;
    if (cum >= target && cum - c < target) qbs = t;
    __syncthreads();
    if (t != 0) return;
    const float hw = fmaxf(vs->h, (float)(qbs + 1) * 0.0625f);
    vs->band[1] = make_band(hw, (double)vs->prec);
}

// rate-free mixture proposal t0 g0 + t1 g1 at (x, y)
// The proposal mixture at squared distance d2: k0 e^{-a0 d2} + k1 e^{-a1 d2},
// with k_b = t_b c_b and a_b = is2h_b^2 = 1 / 2h_b^2 (one exponential per band).
__device__ __forceinline__ float mix_d2(float k0, float a0, float k1, float a1, float d2)
{
    const float e0 = pexpf(-(d2 * a0));
    const float e1 = pexpf(-(d2 * a1));
    float m = k0 * e0;
    m = m + k1 * e1;
    return m;
}

// wide-band half plus footprint half: k1 e^{-a1 d2} + kf f (same order as mix_d2)
__device__ __forceinline__ float foot_mix(float k0, float a0, float kf, float d2, float f)
{
    const float e0 = pexpf(-(d2 * a0));
    float m = k0 * e0;
    m = m + kf * f;
    return m;
}

__device__ __forceinline__ float dist2(float qx, float qy, float x, float y)
{
    const float dx = qx - x, dy = qy - y;
    float d2 = dx * dx;
    d2 = d2 + dy * dy;
    return d2;
}

// Walk records of the Poisson pass, and (for internal nodes) the density walk's
// children records, in one launch after the expansions are selected.  The
// escape pointer of a node is the right sibling of its first ancestor-or-self
// that is a left child (-1 on the rightmost path).  That sibling's leaf range
// starts at m = last + 1, and in a Karras tree a right child's index is the
// first leaf of its range, so it is internal node m when node m's range starts
// at m and leaf m otherwise (no climb over parent links).
__global__ __launch_bounds__(256) void k_walk_records(const ViewState* __restrict__ vs,
                                                      const NodeGeo* __restrict__ geo,
                                                      const NodeLnk* __restrict__ lnk,
                                                      const NodeAux* __restrict__ aux,
                                                      const int* __restrict__ fixOff0,
                                                      const int* __restrict__ fixOff1, PNode* __restrict__ pnode,
                                                      FmmRec* __restrict__ frec, VStride vsd)
{
    vs = vw(vs, vsd.ws);
    geo = vw(geo, vsd.ws);
    lnk = vw(lnk, vsd.ws);
    aux = vw(aux, vsd.ws);
    fixOff0 = vw(fixOff0, vsd.ws);
    fixOff1 = vw(fixOff1, vsd.ws);
    pnode = vw(pnode, vsd.ws);
    frec = vw(frec, vsd.ws);
    const int nd = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = vs->nc;
    if (nc <= 0 || nd >= 2 * nc - 1) return;
    const int L0 = nc - 1;
    const NodeGeo g = geo[nd];
    const NodeAux a = aux[nd];
    const int m = a.last + 1;
    int esc = -1;
    if (m < nc) esc = (m < L0 && aux[m].first == m) ? m : L0 + m;
    PNode q;
    q.lox = g.lox; q.loy = g.loy; q.hix = g.hix; q.hiy = g.hiy;
    q.left = nd < L0 ? lnk[nd].left : -1;
    q.esc = esc;
    q.first = a.first; q.last = a.last;
    pnode[nd] = q;
    if (nd < L0) frec[nd] = fmm_record(nd, L0, geo, lnk, fixOff0, fixOff1);
}

// Footprint records of the sorted leaves, and the sum of the visible splats'
// footprint masses (integrals of min(w, 1)) for the normaliser of the footprint
// half (footSum starts at zero).
__global__ __launch_bounds__(256) void k_foot_leaves(float Wf, float Hf, float ff, float smoothing,
                                                     const float* __restrict__ pts, const float* __restrict__ nrm,
                                                     const float* __restrict__ sd, const ViewState* __restrict__ vs,
                                                     const int32_t* __restrict__ ids,
                                                     const float2* __restrict__ leafXY, FootLeaf* __restrict__ footLeaf,
                                                     double* footSum, VStride vsd)
{
    pts = vw(pts, vsd.pts);
    nrm = vw(nrm, vsd.pts);
    sd = vw(sd, vsd.sd);
    vs = vw(vs, vsd.ws);
    ids = vw(ids, vsd.ws);
    leafXY = vw(leafXY, vsd.ws);
    footLeaf = vw(footLeaf, vsd.ws);
    footSum = vw(footSum, vsd.ws);
    __shared__ double sm[4];
    const int n = vs->nc;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double mass = 0.0;
    if (t < n) {
        const int id = ids[t];
        FootLeaf L;
        const float2 xy = leafXY[t];
        L.x = xy.x; L.y = xy.y;
        float m, ms;
        foot_f32(Wf, Hf, ff, pts + 3L * id, nrm + 3L * id, sd + 2L * id, smoothing, L.t00, L.t01, L.t11, L.amp, m,
                 ms);
        L.pad[0] = L.pad[1] = 0.0f;
        footLeaf[t] = L;
        mass = (double)ms;
    }
    for (int o = 32; o > 0; o >>= 1) mass += __shfl_xor(mass, o, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mass;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = (sm[0] + sm[1]) + (sm[2] + sm[3]);
        if (b != 0.0) atomicAdd(footSum, b);
    }
}

__global__ void k_zero_f64(double* p, size_t stride, int views)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < views) *(double*)((char*)p + stride * v) = 0.0;
}

// One lane per pixel (one pixel tile per wavefront, TileCfg).  The walk
// visits nodes left to right; a node's bound U (the mixture at the point of its
// box closest to q) is >= pi_i of each of its leaves.  U = 0: skip the
// subtree.  A leaf, or a node with count U <= 1, is thinned: candidate leaves at
// rate U by geometric skips over its contiguous leaf range (skip = floor(log u /
// log(1 - U))), each kept with probability pi_i / U, i.e. included with
// probability pi_i, independently.  Uniforms are Philox4x32-10 words at counter
// (step, 0x80000000, pixel, view), one block per geometric-skip step of the
// pixel (word x the skip, word y the acceptance test).  A skipped or thinned
// subtree is left through its escape pointer.
// The first kPoissonStage samples of each lane are staged in LDS (slot-major,
// [slot][lane], so the walk's stores hit 64 distinct banks) and written out at
// the end together with the empty slots, one tile row (2^lth contiguous pixels
// x kcap slots) per coalesced pass of the wavefront; only samples past the
// staged ones are stored directly.  A lane's scattered 4-byte stores would
// otherwise leave partly written lines in L2 for the length of the walk.
constexpr int kPoissonStage = 16;
constexpr size_t kPoissonStageBytes = (size_t)kPoissonStage * 64 * 8;    // per wavefront

__global__ __launch_bounds__(64 * kMaxWavesPerBlock) void k_poisson(TileCfg cfg, int W, int H, int kcap, float rate,
                                                                   const float* __restrict__ tscale, int tsH,
                                                                   uint32_t seedLo, uint32_t seedHi,
                                                                   const uint64_t* __restrict__ seedDev,
                                                                   uint32_t view0, const ViewState* __restrict__ vs,
                                                                   size_t vstride, const PNode* __restrict__ pnode,
                                                                   const float2* __restrict__ leafXY,
                                                                   const int32_t* __restrict__ leafId,
                                                                   const float* __restrict__ dens0,
                                                                   const float* __restrict__ dens1, size_t densStride,
                                                                   int32_t* __restrict__ outIds,
                                                                   float* __restrict__ outIpi, int* err, int fmode,
                                                                   float fscale, const FootLeaf* __restrict__ footLeaf,
                                                                   const double* __restrict__ footSum)
{
    extern __shared__ int32_t stage[];
    const int lane = threadIdx.x & 63;
    int i, j;
    if (!tile_lane(cfg, W, H, lane, i, j)) return;           // wave-uniform
    int32_t* sid = stage + (threadIdx.x >> 6) * (2 * kPoissonStage * 64);
    float* sip = (float*)(sid + kPoissonStage * 64);
    const bool valid = i < W && j < H;
    const uint32_t view = view0 + blockIdx.y;
    if (seedDev) {
        const uint64_t sd = *seedDev;
        seedLo = (uint32_t)sd;
        seedHi = (uint32_t)(sd >> 32);
    }
    vs = vw(vs, vstride);
    pnode = vw(pnode, vstride);
    leafXY = vw(leafXY, vstride);
    leafId = vw(leafId, vstride);
    dens0 = vw(dens0, densStride);
    dens1 = vw(dens1, densStride);
    if (fmode) {
        footLeaf = vw(footLeaf, vstride);
        footSum = vw(footSum, vstride);
    }
    const long perView = (long)W * H;
    outIds += (long)blockIdx.y * perView * kcap;
    outIpi += (long)blockIdx.y * perView * kcap;
    if (tscale) tscale += (long)blockIdx.y * ((W + 15) / 16) * tsH;
    const long pix = valid ? (long)j + (long)H * i : 0;
    int32_t* oid = outIds + pix * kcap;
    float* oip = outIpi + pix * kcap;
    int ns = 0;
    bool over = false;
    const int n = vs->nc;
    const float D0 = (n > 0 && valid) ? dens0[pix] : 0.0f;
    if (valid && n > 0 && D0 >= vs->prec) {
        const int L0 = n - 1;
        const Band B0 = vs->band[0], B1 = vs->band[1];
        float lam = rate;
        if (tscale) lam = lam * tscale[(long)(i >> 4) * tsH + (j >> 4)];
        // a negative or non-finite multiplier would make every candidate's pi 1
        // (fminf drops a NaN): draw nothing there and say so
        if (!(lam >= 0.0f && lam <= 3.0e38f)) { atomicOr(err, 16); lam = 0.0f; }
        const float t0 = 0.5f / fmaxf(D0, B0.c);
        const float t1 = 0.5f / fmaxf(dens1[pix], B1.c);
        const float k0 = t0 * B0.c, k1 = t1 * B1.c;
        const float a0 = B0.is2h * B0.is2h, a1 = B1.is2h * B1.is2h;
        // footprint mode: the wide band's half as above, and in place of band 0
        // f_i(q) / (kw max(D1, c1)), kw = the visible splats' mean footprint mass
        // over h_1^1.5 (so that both halves expect about lam / 2 samples), or the
        // caller's scale
        float kf = 0.0f;
        if (fmode) {
            float kw = fscale;
            if (!(kw > 0.0f)) {
                const double hh = (double)B1.h;
                kw = (float)((*footSum / (double)n) / (hh * sqrt(hh)));
            }
            if (kw > 0.0f && kw < 3.0e38f) kf = t1 / kw;
        }
        const float qx = (float)i, qy = (float)j;
        const uint32_t pixel = (uint32_t)pix;
        uint32_t step = 0;                            // geometric-skip steps of this pixel
        int nd = 0;
        int visits = 0;                               // a preorder walk visits each node at most once
        while (true) {
            if (++visits > 2 * n) { atomicOr(err, 8); break; }
            const PNode pn = pnode[nd];
            float cx = fminf(fmaxf(qx, pn.lox), pn.hix);
            float cy = fminf(fmaxf(qy, pn.loy), pn.hiy);
            const float d2b = dist2(qx, qy, cx, cy);
            // proposal bound over the node
            float pB = fmode ? foot_mix(k1, a1, kf, d2b, foot_bound(a1, d2b)) : mix_d2(k0, a0, k1, a1, d2b);
            float U = fminf(1.0f, lam * pB);
            const int last = pn.last;
            int next = pn.esc;                        // past the subtree
            if (pB >= kProposalPMin && U > 0.0f) {
                const bool leaf = nd >= L0;
                if (!leaf && (float)(last - pn.first + 1) * U > 1.0f) next = pn.left;
                else {
                    const float il1 = 1.0f / plog1mf(U);
                    int pos = pn.first - 1;
                    while (true) {
                        // one Philox block per step, computed by every lane of the
                        // wavefront in the same iteration: word x draws the skip,
                        // word y the acceptance test
                        const u32x4 blk = philox4x32_10(u32x4{ step, 0x80000000u, pixel, view }, seedLo, seedHi);
                        step++;
                        float u = u01(blk.x);
                        float r = plogf(u) * il1;         // failures before the next candidate
                        if (!(r < (float)(last - pos))) break;
                        pos += (int)r + 1;
                        float pl;
                        if (fmode) {
                            const FootLeaf L = footLeaf[pos];
                            const float d2 = dist2(qx, qy, L.x, L.y);
                            pl = foot_mix(k1, a1, kf, d2, foot_leaf(L, a1, qx, qy, d2));
                        } else {
                            const float2 lp = leafXY[pos];
                            pl = mix_d2(k0, a0, k1, a1, dist2(qx, qy, lp.x, lp.y));
                        }
                        if (!(pl >= kProposalPMin)) continue;     // outside the support
                        float pi = fminf(1.0f, lam * pl);
                        float a = u01(blk.y);
                        if (a * U <= pi) {            // P = pi / U, exactly 1 when pi = U
                            if (ns < kcap) {
                                const int32_t id = leafId[pos];
                                const float ip = 1.0f / pi;
                                if (ns < kPoissonStage) { sid[ns * 64 + lane] = id; sip[ns * 64 + lane] = ip; }
                                else { oid[ns] = id; oip[ns] = ip; }
                                ns++;
                            } else over = true;
                        }
                    }
                }
            }
            nd = next;
            if (nd < 0) break;
        }
    }
    // staged samples and empty slots: a row of the tile (2^lth pixels) is contiguous
    // in memory and is written by the whole wavefront, so the stores coalesce
    __builtin_amdgcn_wave_barrier();
    const int th = 1 << cfg.lth;
    for (int r = 0; r < (64 >> cfg.lth); r++) {
        const int l0 = r << cfg.lth;
        const int ir = __shfl(i, l0, 64), jr = __shfl(j, l0, 64);
        const int cnt = ir < W ? (H - jr < th ? H - jr : th) : 0;          // wave-uniform
        if (cnt <= 0) continue;
        int32_t* rid = outIds + ((long)jr + (long)H * ir) * kcap;
        float* rip = outIpi + ((long)jr + (long)H * ir) * kcap;
        const int total = cnt * kcap;
        for (int e0 = 0; e0 < total; e0 += 64) {
            const int e = e0 + lane;
            const int c = e < total ? e / kcap : 0;
            const int nsc = __shfl(ns, l0 + c, 64);
            const int q = e - c * kcap;                   // slot of pixel l0 + c
            if (e < total) {
                if (q >= nsc) { rid[e] = -1; rip[e] = 0.0f; }
                else if (q < kPoissonStage) { rid[e] = sid[q * 64 + l0 + c]; rip[e] = sip[q * 64 + l0 + c]; }
            }
        }
    }
    if (over) atomicOr(err, 4);
}

// ---------------------------------------------------------------------------
// Per-tile features for the rate network (no reference counterpart): 16x16
// pixel tiles, 32 features each (layout in ude_tile_features).  Splat
// statistics are accumulated in fixed point with integer atomics, pixel
// statistics by fixed-order reductions, so the features are reproducible.
// ---------------------------------------------------------------------------
constexpr int kTileFeat = 32;
constexpr double kFix = 1048576.0;          // 2^20 fixed-point unit

struct TileAcc {
    unsigned long long cnt, sz, szz, sh;
    unsigned int hmax, pad;
};

__global__ __launch_bounds__(256) void k_tile_splats(int N, int W, int H, float ff, float smoothing,
                                                     const float* __restrict__ pts, const float* __restrict__ nrm,
                                                     const float* __restrict__ sd, TileAcc* acc)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const float* p = pts + 3L * t;
    if (!(p[2] < 0.0f)) return;
    float sx, sy;
    screen_f32((float)W, (float)H, ff, p, sx, sy);
    if (!(sx >= 0.0f && sx < (float)W && sy >= 0.0f && sy < (float)H)) return;
    const int tH = (H + 15) / 16;
    TileAcc* a = acc + (long)((int)sx >> 4) * tH + ((int)sy >> 4);
    float z = -p[2];
    float h = prop_h((float)W, (float)H, ff, p, nrm + 3L * t, sd + 2L * t, smoothing);
    if (!(h >= 0.0f)) h = 0.0f;
    h = fminf(h, kProposalHCap);
    atomicAdd(&a->cnt, 1ull);
    atomicAdd(&a->sz, (unsigned long long)((double)z * kFix));
    atomicAdd(&a->szz, (unsigned long long)((double)z * (double)z * kFix));
    atomicAdd(&a->sh, (unsigned long long)((double)h * kFix));
    atomicMax(&a->hmax, __float_as_uint(h));
}

// block-wide sum / max / min over 256 threads, fixed order
__device__ __forceinline__ float blk_sum(float v, float* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
        __syncthreads();
    }
    float r = sh[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ float blk_max(float v, float* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    float r = sh[0];
    __syncthreads();
    return r;
}

// one workgroup per tile, one thread per pixel
__global__ __launch_bounds__(256) void k_tile_pixels(int W, int H, float prec, const float* __restrict__ dens,
                                                     const float* __restrict__ gimg, const TileAcc* __restrict__ acc,
                                                     float* __restrict__ out)
{
    __shared__ float sh[256];
    const int tH = (H + 15) / 16;
    const int tile = blockIdx.x;
    const int ti = tile / tH, tj = tile - (tile / tH) * tH;
    const int i = ti * 16 + (threadIdx.x >> 4), j = tj * 16 + (threadIdx.x & 15);
    const bool in = i < W && j < H;
    const long pix = (long)j + (long)H * i;
    const float d0 = in ? log1pf(fmaxf(dens[pix], 0.0f)) : 0.0f;
    const float d1 = in ? log1pf(fmaxf(dens[(long)W * H + pix], 0.0f)) : 0.0f;
    float gm = 0.0f;
    if (in && gimg) {
        const float* g = gimg + 3 * pix;
        float s2 = g[0] * g[0];
        s2 = s2 + g[1] * g[1];
        s2 = s2 + g[2] * g[2];
        gm = sqrtf(s2);
    }
    const float np = blk_sum(in ? 1.0f : 0.0f, sh);
    const float inv = 1.0f / fmaxf(np, 1.0f);
    float f[kTileFeat];
    // splat statistics
    const TileAcc a = acc[tile];
    const float cnt = (float)a.cnt;
    const float ic = a.cnt ? 1.0f / cnt : 0.0f;
    const float mz = (float)((double)a.sz / kFix) * ic;
    const float mzz = (float)((double)a.szz / kFix) * ic;
    f[0] = log1pf(cnt);
    f[1] = mz;
    f[2] = sqrtf(fmaxf(mzz - mz * mz, 0.0f));
    f[3] = (float)((double)a.sh / kFix) * ic;
    f[4] = __uint_as_float(a.hmax);
    // densities (log1p), coverage, gradient magnitudes
    const float m0 = blk_sum(d0, sh) * inv, m1 = blk_sum(d1, sh) * inv;
    f[5] = m0;
    f[6] = blk_max(in ? d0 : 0.0f, sh);
    f[7] = -blk_max(in ? -d0 : -1e30f, sh);
    f[8] = sqrtf(fmaxf(blk_sum(in ? (d0 - m0) * (d0 - m0) : 0.0f, sh) * inv, 0.0f));
    f[9] = m1;
    f[10] = blk_max(in ? d1 : 0.0f, sh);
    f[11] = -blk_max(in ? -d1 : -1e30f, sh);
    f[12] = sqrtf(fmaxf(blk_sum(in ? (d1 - m1) * (d1 - m1) : 0.0f, sh) * inv, 0.0f));
    f[13] = blk_sum((in && dens[pix] >= prec) ? 1.0f : 0.0f, sh) * inv;
    const float mg = blk_sum(gm, sh) * inv;
    f[14] = mg;
    f[15] = blk_max(gm, sh);
    f[16] = sqrtf(fmaxf(blk_sum(in ? (gm - mg) * (gm - mg) : 0.0f, sh) * inv, 0.0f));
    f[17] = log1pf(blk_sum(gm * gm, sh));
    // quadrants: mean log density and mean gradient magnitude
    const int quad = ((threadIdx.x >> 7) << 1) | ((threadIdx.x >> 3) & 1);
    for (int q = 0; q < 4; q++) {
        const bool mine = quad == q && in;
        const float nq = fmaxf(blk_sum(mine ? 1.0f : 0.0f, sh), 1.0f);
        f[18 + 2 * q] = blk_sum(mine ? d0 : 0.0f, sh) / nq;
        f[19 + 2 * q] = blk_sum(mine ? gm : 0.0f, sh) / nq;
    }
    f[26] = ((float)ti * 16.0f + 8.0f) / (float)W;
    f[27] = ((float)tj * 16.0f + 8.0f) / (float)H;
    f[28] = log1pf(blk_sum(in ? fmaxf(dens[pix], 0.0f) : 0.0f, sh));
    f[29] = log1pf(blk_sum(in ? gm * fmaxf(dens[pix], 0.0f) : 0.0f, sh));
    f[30] = np * (1.0f / 256.0f);
    f[31] = log1pf(cnt * np * (1.0f / 256.0f));
    if (threadIdx.x < kTileFeat) out[(long)tile * kTileFeat + threadIdx.x] = f[threadIdx.x];
}

__global__ void k_zero_u64(unsigned long long* p, long n)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0ull;
}

inline unsigned grid1(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Tree structure of one view: bandwidth, cull, sort, hierarchy, bottom-up
// boxes / centres / counts / dynamic degrees.
int build_structure(int N, int W, int H, double f, double smoothing, double prec, const float* pts, const float* nrm,
                    const float* sd, const TreeWs& w, hipStream_t st, int views = 1, VStride vsd = VStride{})
{
    float Wf = (float)W, Hf = (float)H, ff = (float)f;
    long nCull = (N + kCullTile - 1) / kCullTile;
    long nSort = (N + kSortTile - 1) / kSortTile;
    k_zero_state<<<dim3(1, views), 64, 0, st>>>(w.vs, vsd);
    // 256 workgroups of 1024 threads per view: 16 waves per CU stream the splats,
    // and the two sums take 512 device atomics (1M splats: 23.7 -> 17.2 us)
    const unsigned bwBlocks = grid1(N, 1024) < 256u ? grid1(N, 1024) : 256u;
    k_bandwidth<<<dim3(bwBlocks, views), 1024, 0, st>>>(N, Wf, Hf, ff, (float)smoothing, pts, nrm, sd, w.vs, w.hraw,
                                                       vsd);
    k_view_params<<<dim3(1, views), 64, 0, st>>>(w.vs, Wf, Hf, prec, vsd);
    k_cull_count<<<dim3((unsigned)nCull, views), 256, 0, st>>>(N, Wf, Hf, ff, pts, w.vs, w.cullCounts, vsd);
    k_scan_small<<<dim3(1, views), 1024, 0, st>>>(w.cullCounts, (int)nCull, &w.vs->nc, vsd);
    k_cull_scatter<<<dim3((unsigned)nCull, views), 256, 0, st>>>(N, Wf, Hf, ff, pts, w.vs, w.cullCounts, w.keysB, w.idsB, vsd);
    uint32_t* kin = w.keysB; uint32_t* kout = w.keysA;
    int32_t* iin = w.idsB; int32_t* iout = w.idsA;
    for (int shift = 0; shift < 24; shift += 8) {                  // 20-bit keys
        k_sort_hist<<<dim3((unsigned)nSort, views), 256, 0, st>>>(w.vs, kin, shift, (int)nSort, w.sortHist, vsd);
        k_sort_scan_digits<<<dim3(256, views), 256, 0, st>>>(w.sortHist, (int)nSort, w.scanBlocks, vsd);
        k_sort_scan_top<<<dim3(1, views), 256, 0, st>>>(w.scanBlocks, vsd);
        k_sort_scatter<<<dim3((unsigned)nSort, views), 256, 0, st>>>(w.vs, kin, iin, shift, (int)nSort, w.sortHist, w.scanBlocks,
                                                        kout, iout, vsd);
        uint32_t* tk = kin; kin = kout; kout = tk;
        int32_t* ti = iin; iin = iout; iout = ti;
    }
    // after 3 passes the sorted data is in keysA / idsA
    k_hierarchy<<<dim3(grid1(N, 256), views), 256, 0, st>>>(Wf, Hf, ff, pts, w.vs, kin, iin, w.geo, w.lnk, w.aux, w.leafXY,
                                                            w.flags, vsd);
    k_bottom_up_local<<<dim3(grid1(N, kFanBlock), views), kFanBlock, 0, st>>>(w.vs, w.geo, w.lnk, w.aux, w.seeds, vsd);
    k_bottom_up_global<<<dim3(grid1(N, 256), views), 256, 0, st>>>(w.vs, w.geo, w.lnk, w.aux, w.flags, w.seeds, vsd);
    hipError_t e0 = hipGetLastError();
    return e0 == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
}

// Dynamic-degree Hermite moments of the sampler's tree.
int build_moments(int N, int k, const TreeWs& w, hipStream_t st)
{
    long nScan = (2L * N + kScanTile - 1) / kScanTile;
    k_coef_sizes<<<grid1(2L * N, 256), 256, 0, st>>>(w.vs, w.geo, w.lnk, k, w.scanTmp, w.momList);
    k_scan_tiles<<<(unsigned)nScan, 256, 0, st>>>(w.vs, w.scanTmp, w.scanBlocks);
    k_scan_small<<<1, 1024, 0, st>>>(w.scanBlocks, (int)nScan, &w.vs->ncoef, VStride{});
    k_scan_apply<<<grid1(2L * N, 256), 256, 0, st>>>(w.vs, w.scanTmp, w.scanBlocks, w.aux);
    k_check_coef<<<1, 64, 0, st>>>(w.vs, w.coefCap);
    // enough workgroups for max(N leaves / 256, N nodes / 15)
    k_moments<<<grid1(N, kMomNodes), 256, 0, st>>>(w.vs, w.geo, w.aux, w.momList, w.coef, w.coefCap);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
}

int num_launches(int k)
{
    int L = 0;
    while ((double)L < sqrt((double)k)) L++;
    return L;
}

// trie entries per lane: room for the walk plus 32 new entries per descent
int trie_entries(int k) { return 384 + 32 * num_launches(k); }

// Optional timing of one sampling kernel (0 the walk + draw pair of
// sampling.sample, 1 k_fmm_eval, 2 k_poisson): while enabled, its launches are
// bracketed by pairs of HIP events recorded on their stream (single host
// thread; events are pooled and reused across profiling windows).
struct SamplerTiming {
    bool on = false;
    int kernel = 0;
    int used = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
};
SamplerTiming g_timing;

bool timing_pair(int kernel, hipEvent_t* a, hipEvent_t* b)
{
    if (!g_timing.on || g_timing.kernel != kernel) return false;
    if (g_timing.used == (int)g_timing.pool.size()) {
        hipEvent_t x, y;
        if (hipEventCreate(&x) != hipSuccess) return false;
        if (hipEventCreate(&y) != hipSuccess) { (void)hipEventDestroy(x); return false; }
        g_timing.pool.push_back({x, y});
    }
    *a = g_timing.pool[g_timing.used].first;
    *b = g_timing.pool[g_timing.used].second;
    g_timing.used++;
    return true;
}

void launch_sample(int N, int W, int H, int k, uint32_t seedLo, uint32_t seedHi, uint32_t view, const TreeWs& w,
                   int32_t* out, int* err, hipStream_t st)
{
    k_samp_records<<<grid1(N, 256), 256, 0, st>>>(w.vs, w.geo, w.lnk, w.aux, w.srec);
    unsigned tiles = (unsigned)(((W + 7) / 8) * ((H + 7) / 8));
    size_t lds = (size_t)k * 64 * sizeof(unsigned long long);
    const int L = num_launches(k);
    hipEvent_t e0, e1;
    bool timed = timing_pair(0, &e0, &e1);
    if (timed) (void)hipEventRecord(e0, st);
    k_sample_walk<<<tiles, 64, 0, st>>>(W, H, k, L, w.vs, w.geo, w.lnk, w.aux, w.coef, w.srec, w.trie,
                                        trie_entries(k), w.reach, err);
    k_fill_i32<<<1, 64, 0, st>>>(w.tieQueue, 1, 0);
    k_sample_draw<<<tiles, 64, lds, st>>>(W, H, k, L, seedLo, seedHi, view, w.vs, w.leafXY, w.idsA, w.reach, out,
                                          w.tieQueue);
    k_sample_replay<<<64, 64, 0, st>>>(W, H, k, L, tiles, seedLo, seedHi, view, w.vs, w.leafXY, w.idsA, w.reach,
                                       w.tieQueue, out);
    if (timed) (void)hipEventRecord(e1, st);
}

// proposal of ude_sample_poisson: 0 the two density bands, 1 the wide band plus the
// footprint half; scale 0 = normaliser from the view, > 0 the caller's
int g_pois_mode = 0;
double g_pois_scale = 0.0;

// launch shapes of k_fmm_eval (0) and k_poisson (1); set by ude_set_tile_config
TileCfg g_tile[2] = {{8, 3, 4, 0}, {8, 3, 4, 0}};

dim3 tile_grid(const TileCfg& c, int W, int H, int views)
{
    const int th = 1 << c.lth;
    const long ntiles = (long)((W + c.tw - 1) / c.tw) * ((H + th - 1) / th);
    return dim3((unsigned)((ntiles + c.wpb - 1) / c.wpb), (unsigned)views);
}

// degree-6 expansions of bands 0 .. nbands-1 on one view's tree (or a batch)
void fixed_expansions(int N, int nbands, int p, const TreeWs& w, hipStream_t st, int views = 1,
                      VStride vsd = VStride{})
{
    k_fixed_select<<<dim3(grid1(N, 256), views), 256, 0, st>>>(w.vs, nbands, p, w.geo, w.lnk, w.fixOff[0],
                                                                nbands > 1 ? w.fixOff[1] : nullptr, w.fixList[0],
                                                                nbands > 1 ? w.fixList[1] : nullptr, vsd);
    for (int b = 0; b < nbands; b++) {
        const dim3 grid(grid1(N, 256 / (p + 1)), views);
        switch (p) {
#define UD_MOMENTS(P) case P: k_fixed_moments<P><<<grid, 256, 0, st>>>(w.vs, b, w.geo, w.aux, w.fixList[b], \
                                                                        w.fixOff[b], w.fixCoef[b], vsd); break;
        UD_MOMENTS(4) UD_MOMENTS(5) UD_MOMENTS(6) UD_MOMENTS(7) UD_MOMENTS(8)
#undef UD_MOMENTS
        default: break;
        }
    }
}

// densities of bands 0 .. nbands-1 of `views` views in one launch: w is the
// first view's workspace, the others follow vstride bytes apart; view v writes
// band b to out + v outStride + b bandStride bytes
void fmm_eval(int W, int H, int views, int nbands, int p, const TreeWs& w, size_t vstride, float* out,
              size_t outStride, size_t bandStride, hipStream_t st)
{
    const TileCfg c = g_tile[0];
    hipEvent_t e0, e1;
    bool timed = timing_pair(1, &e0, &e1);
    if (timed) (void)hipEventRecord(e0, st);
    const dim3 grid = tile_grid(c, W, H, views * nbands);
    switch (p) {
#define UD_EVAL(P) case P: k_fmm_eval<P><<<grid, 64 * c.wpb, 0, st>>>(                                             \
        c, W, H, w.vs, vstride, nbands, w.geo, w.frec, w.fixOff[0], nbands > 1 ? w.fixOff[1] : nullptr,            \
        w.fixCoef[0], nbands > 1 ? w.fixCoef[1] : nullptr, out, outStride, bandStride); break;
    UD_EVAL(4) UD_EVAL(5) UD_EVAL(6) UD_EVAL(7) UD_EVAL(8)
#undef UD_EVAL
    default: break;
    }
    if (timed) (void)hipEventRecord(e1, st);
}

// degree of the fixed-order expansions; set by ude_set_fixed_degree
int g_fixed_p = kFixedDegree;

} // namespace

extern "C" {

size_t ude_sample_workspace_bytes(int N, int W, int H, int k)
{
    return carve(nullptr, N, W, H, k, kModeSample, nullptr) + 256;
}

int ude_sample(int bn, int N, int W, int H, int k, double f, double smoothing, double prec,
               const float* pts, const float* nrm, const float* sd, int sdBatch,
               uint64_t seed, int32_t* out, int* errFlag, void* ws, size_t wsBytes, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || k <= 0 || N < 0) return UDE_ERR_ARG;
    if (k > 256) return UDE_ERR_ARG;
    long perView = (long)W * H * k;
    if (N == 0) {
        k_fill_idx<<<grid1(perView * bn, 256), 256, 0, st>>>(out, perView * bn);
        return UDE_OK;
    }
    if (wsBytes < ude_sample_workspace_bytes(N, W, H, k)) return UDE_ERR_WORKSPACE;
    TreeWs w;
    carve(ws, N, W, H, k, kModeSample, &w);
    uint32_t sLo = (uint32_t)seed, sHi = (uint32_t)(seed >> 32);
    for (int b = 0; b < bn; b++) {
        const float* pb = pts + 3L * N * b;
        const float* nb = nrm + 3L * N * b;
        const float* sb = sd + (sdBatch > 1 ? 2L * N * b : 0);
        int rc = build_structure(N, W, H, f, smoothing, prec, pb, nb, sb, w, st);
        if (rc == UDE_OK) rc = build_moments(N, k, w, st);
        if (rc != UDE_OK) return rc;
        int32_t* ob = out + perView * b;
        launch_sample(N, W, H, k, sLo, sHi, (uint32_t)b, w, ob, errFlag, st);
        k_copy_err<<<1, 64, 0, st>>>(w.vs, errFlag);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
}

size_t ude_fmm_weights_workspace_bytes(int N, int W, int H)
{
    return carve(nullptr, N, W, H, 0, kModeFixed, nullptr) + 256;
}

int ude_fmm_weights(int bn, int N, int W, int H, double f, double smoothing, double prec,
                    const float* pts, const float* nrm, const float* sd, int sdBatch,
                    float* out, void* ws, size_t wsBytes, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || N < 0) return UDE_ERR_ARG;
    long perView = (long)W * H;
    if (N == 0) {
        k_fill_f32<<<grid1(perView * bn, 256), 256, 0, st>>>(out, perView * bn, 0.0f);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
    }
    if (wsBytes < ude_fmm_weights_workspace_bytes(N, W, H)) return UDE_ERR_WORKSPACE;
    const int p = g_fixed_p;
    TreeWs w;
    carve(ws, N, W, H, 0, kModeFixed, &w);
    for (int b = 0; b < bn; b++) {
        const float* pb = pts + 3L * N * b;
        const float* nb = nrm + 3L * N * b;
        const float* sb = sd + (sdBatch > 1 ? 2L * N * b : 0);
        int rc = build_structure(N, W, H, f, smoothing, prec, pb, nb, sb, w, st);
        if (rc != UDE_OK) return rc;
        fixed_expansions(N, 1, p, w, st);
        k_fmm_records<<<grid1(N, 256), 256, 0, st>>>(w.vs, w.geo, w.lnk, w.fixOff[0], nullptr, w.frec, VStride{});
        fmm_eval(W, H, 1, 1, p, w, 0, out + perView * b, 0, 0, st);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
}

size_t ude_poisson_workspace_bytes(int bn, int N, int W, int H)
{
    return carve(nullptr, N, W, H, 0, kModePoisson, nullptr) * (size_t)(bn > 1 ? bn : 1) + 256;
}

int ude_sample_poisson(int bn, int N, int W, int H, int kcap, double f, double smoothing, double prec,
                       double rate, const float* tileScale, const float* pts, const float* nrm, const float* sd,
                       int sdBatch, uint64_t seed, const uint64_t* seedDev, int32_t* outIds, float* outIpi,
                       float* densOut, int* errFlag, void* ws, size_t wsBytes, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || kcap <= 0 || N < 0 || !(rate >= 0.0)) return UDE_ERR_ARG;
    long perView = (long)W * H;
    if (N == 0) {
        k_fill_idx<<<grid1(perView * kcap * bn, 256), 256, 0, st>>>(outIds, perView * kcap * bn);
        k_fill_f32<<<grid1(perView * kcap * bn, 256), 256, 0, st>>>(outIpi, perView * kcap * bn, 0.0f);
        if (densOut) k_fill_f32<<<grid1(perView * 2 * bn, 256), 256, 0, st>>>(densOut, perView * 2 * bn, 0.0f);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
    }
    // all views' trees resident at once when the workspace holds them: every
    // per-pixel pass is then one launch over the whole batch; otherwise view by view
    const size_t per = carve(nullptr, N, W, H, 0, kModePoisson, nullptr);
    const bool batched = bn > 1 && wsBytes >= per * (size_t)bn + 256;
    if (!batched && wsBytes < per + 256) return UDE_ERR_WORKSPACE;
    const int span = batched ? bn : 1;
    const int tsW = (W + 15) / 16, tsH = (H + 15) / 16;
    const uint32_t sLo = (uint32_t)seed, sHi = (uint32_t)(seed >> 32);
    const TileCfg pc = g_tile[1];
    const int p = g_fixed_p;
    const int fmode = g_pois_mode;
    const float fscale = (float)g_pois_scale;
    for (int g = 0; g < bn; g += span) {
        // every view of the group in one launch per build stage (view = blockIdx.y)
        TreeWs w;
        carve(ws, N, W, H, 0, kModePoisson, &w);
        const VStride vsd{per, sizeof(float) * 3 * (size_t)N, sdBatch > 1 ? sizeof(float) * 2 * (size_t)N : 0};
        const float* pb = pts + 3L * N * g;
        const float* nb = nrm + 3L * N * g;
        const float* sb = sd + (sdBatch > 1 ? 2L * N * g : 0);
        int rc = build_structure(N, W, H, f, smoothing, prec, pb, nb, sb, w, st, span, vsd);
        if (rc != UDE_OK) return rc;
        k_wide_hist<<<dim3(grid1(N, 256) < 256u ? grid1(N, 256) : 256u, span), 256, 0, st>>>(w.vs, w.idsA, w.hraw,
                                                                                          vsd);
        k_wide_band<<<dim3(1, span), 256, 0, st>>>(w.vs, vsd);
        fixed_expansions(N, 2, p, w, st, span, vsd);
        k_walk_records<<<dim3(grid1(2L * N, 256), span), 256, 0, st>>>(w.vs, w.geo, w.lnk, w.aux, w.fixOff[0],
                                                                       w.fixOff[1], w.pnode, w.frec, vsd);
        // footprint proposal: the leaves' footprint records and their node bounds
        if (fmode) {
            k_zero_f64<<<(span + 63) / 64, 64, 0, st>>>(w.footSum, per, span);
            k_foot_leaves<<<dim3(grid1(N, 256), span), 256, 0, st>>>((float)W, (float)H, (float)f, (float)smoothing,
                                                                     pb, nb, sb, w.vs, w.idsA, w.leafXY, w.footLeaf,
                                                                     w.footSum, vsd);
        }
        float* d0 = densOut ? densOut + perView * 2 * g : w.dens[0];
        float* d1 = densOut ? d0 + perView : w.dens[1];
        const size_t dstride = densOut ? sizeof(float) * 2 * perView : per;
        fmm_eval(W, H, span, 2, p, w, per, d0, dstride, (size_t)((char*)d1 - (char*)d0), st);
        const float* ts = tileScale ? tileScale + (long)tsW * tsH * g : nullptr;
        hipEvent_t e0, e1;
        bool timed = timing_pair(2, &e0, &e1);
        if (timed) (void)hipEventRecord(e0, st);
        k_poisson<<<tile_grid(pc, W, H, span), 64 * pc.wpb, kPoissonStageBytes * pc.wpb, st>>>(
            pc, W, H, kcap, (float)rate, ts, tsH, sLo, sHi, seedDev, (uint32_t)g, w.vs, per, w.pnode, w.leafXY,
            w.idsA, d0, d1, dstride, outIds + perView * kcap * g, outIpi + perView * kcap * g, errFlag, fmode,
            fscale, w.footLeaf, w.footSum);
        if (timed) (void)hipEventRecord(e1, st);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
}

int ude_set_poisson_proposal(int mode, double footprintScale)
{
    if (mode != 0 && mode != 1) return UDE_ERR_ARG;
    if (!(footprintScale <= 3.0e38)) return UDE_ERR_ARG;          // NaN or too large for float
    g_pois_mode = mode;
    g_pois_scale = footprintScale > 0.0 ? footprintScale : 0.0;
    return UDE_OK;
}

int ude_get_poisson_proposal(double* footprintScale)
{
    if (footprintScale) *footprintScale = g_pois_scale;
    return g_pois_mode;
}

int ude_set_fixed_degree(int p)
{
    if (p < kFixedPMin || p > kFixedPMax) return UDE_ERR_ARG;
    g_fixed_p = p;
    return UDE_OK;
}

int ude_get_fixed_degree(void)
{
    return g_fixed_p;
}

int ude_set_tile_config(int kernel, int tileW, int tileH, int wavesPerBlock, int xcdOrder)
{
    if (kernel < 0 || kernel > 1) return UDE_ERR_ARG;
    if (tileW <= 0 || tileH <= 0 || tileW * tileH != 64 || (tileW & (tileW - 1))) return UDE_ERR_ARG;
    if (wavesPerBlock != 1 && wavesPerBlock != 2 && wavesPerBlock != 4 && wavesPerBlock != 8) return UDE_ERR_ARG;
    if (xcdOrder != 0 && xcdOrder != 1) return UDE_ERR_ARG;
    int lth = 0;
    while ((1 << lth) < tileH) lth++;
    g_tile[kernel] = TileCfg{tileW, lth, wavesPerBlock, xcdOrder};
    return UDE_OK;
}

int ude_get_tile_config(int kernel, int* cfg)
{
    if (kernel < 0 || kernel > 1 || !cfg) return UDE_ERR_ARG;
    const TileCfg c = g_tile[kernel];
    cfg[0] = c.tw; cfg[1] = 1 << c.lth; cfg[2] = c.wpb; cfg[3] = c.xcd;
    return UDE_OK;
}

size_t ude_tile_features_workspace_bytes(int W, int H)
{
    return sizeof(TileAcc) * (size_t)((W + 15) / 16) * ((H + 15) / 16) + 256;
}

int ude_tile_features(int bn, int N, int W, int H, double f, double smoothing, double prec, const float* pts,
                      const float* nrm, const float* sd, int sdBatch, const float* dens, const float* gimg,
                      float* out, void* ws, size_t wsBytes, hipStream_t st)
{
    if (bn <= 0 || W <= 0 || H <= 0 || N < 0 || !dens) return UDE_ERR_ARG;
    if (wsBytes < ude_tile_features_workspace_bytes(W, H)) return UDE_ERR_WORKSPACE;
    const int tiles = ((W + 15) / 16) * ((H + 15) / 16);
    TileAcc* acc = (TileAcc*)ws;
    const long perView = (long)W * H;
    for (int b = 0; b < bn; b++) {
        k_zero_u64<<<grid1((long)tiles * 5, 256), 256, 0, st>>>((unsigned long long*)acc, (long)tiles * 5);
        if (N > 0)
            k_tile_splats<<<grid1(N, 256), 256, 0, st>>>(N, W, H, (float)f, (float)smoothing, pts + 3L * N * b,
                                                         nrm + 3L * N * b, sd + (sdBatch > 1 ? 2L * N * b : 0), acc);
        k_tile_pixels<<<tiles, 256, 0, st>>>(W, H, (float)prec, dens + 2 * perView * b,
                                             gimg ? gimg + 3 * perView * b : nullptr, acc,
                                             out + (long)tiles * kTileFeat * b);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
}

int ude_profile_sampler_begin(int kernel)
{
    if (kernel < 0 || kernel > 2) return UDE_ERR_ARG;
    g_timing.kernel = kernel;
    g_timing.on = true;
    g_timing.used = 0;
    return UDE_OK;
}

int ude_profile_sampler_end(double* totalMs, int* launches)
{
    double tot = 0.0;
    int n = g_timing.used;
    g_timing.on = false;
    for (int q = 0; q < n; q++) {
        float ms = 0.0f;
        if (hipEventSynchronize(g_timing.pool[q].second) != hipSuccess) return UDE_ERR_LAUNCH;
        if (hipEventElapsedTime(&ms, g_timing.pool[q].first, g_timing.pool[q].second) != hipSuccess)
            return UDE_ERR_LAUNCH;
        tot += ms;
    }
    g_timing.used = 0;
    if (totalMs) *totalMs = tot;
    if (launches) *launches = n;
    return UDE_OK;
}

} // extern "C"
