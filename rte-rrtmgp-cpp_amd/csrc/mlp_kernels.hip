// MLP gas-optics surrogate: training and inference kernels (BF16 MFMA, FP32 accumulation).
// No counterpart in the reference; the network and its targets come from the plain gas optics.
//
// Network, per (column, layer): x = [T, ln p, ln h2o, ln o3, ln co2, ln ch4, ln n2o, f], with f = 1
// in the lower atmosphere (p > p_trop, where the k-distribution switches tables) and 0 above,
// normalised with the training set's mean and standard deviation, then
//     h1 = softplus(W1 x + b1),  h2 = softplus(W2 h1 + b2),  y = W3 h2 + b3      (hidden 128)
// and outputs denormalised with per-output mean/std.  Outputs are interleaved per g-point:
//     LW row 2g = ln(tau_g / col_dry), row 2g+1 = Planck fraction_g           (no = 2 ngpt)
//     SW row 2g = ln(tau_abs_g / col_dry), row 2g+1 = ln(tau_ray_g / col_dry)
// Weights are row-major (out, in) in BF16; FP32 master copies live with the optimiser.
//
// Matrix cores: v_mfma_f32_32x32x16_bf16.  For D(32x32) += A(32x16) B(16x32), lane l holds
// A[l%32][8(l/32) .. 8(l/32)+7] and B[8(l/32) .. +7][l%32] (8 BF16 each) and D rows
// 8(r/4) + 4(l/32) + r%4 (register r = 0..15) of column l%32.
#include <algorithm>
#include <cmath>

#include "common.h"

namespace rrtmgp_hip
{
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int MLP_H = 128;     // hidden width
constexpr int MLP_IN = 8;      // inputs: T, ln p, 5 ln vmr, lower-atmosphere flag
constexpr int MLP_INP = 16;    // inputs padded to one MFMA K step
constexpr float MLP_LN_FLOOR = -100.f;   // ln(tau / col_dry) target where tau = 0

__device__ __forceinline__ float softplus(float z) { return z > 20.f ? z : log1pf(expf(z)); }
// max(z, 0) + ln(1 + e^-|z|) on the hardware exp2/log2 (absolute error ~1e-7, far below the BF16
// rounding of the activations it feeds).
__device__ __forceinline__ float softplus_fast(float z)
{
    return fmaxf(z, 0.f) + __logf(1.f + __expf(-fabsf(z)));
}

// Store at a wave-uniform base plus a 32-bit lane offset (scalar base, one VGPR of offset),
// streamed (non-temporal): the g-point outputs are read back by the solvers long after any cache
// could hold them.
__device__ __forceinline__ void store_at(float* base, uint32_t i, float v)
{
    __builtin_nontemporal_store(v, reinterpret_cast<float*>(reinterpret_cast<char*>(base) + i * uint32_t(sizeof(float))));
}
__device__ __forceinline__ float sigmoid(float z) { return 1.f / (1.f + expf(-z)); }

__device__ __forceinline__ int mfma_row(int r, int lane) { return 8 * (r / 4) + 4 * (lane / 32) + (r % 4); }

// ---------------------------------------------------------------------------------------------
// Training GEMM: C[m][n] = sum_k A(m, k) B(k, n) (+ C if accumulate), with strided BF16 operands
// (A(m, k) = A[m sAm + k sAk], B(k, n) = B[k sBk + n sBn]) and an FP32 result C[m sCm + n sCn].
// Work-group: 64x64 tile of C, 4 wavefronts of 32x32; K in steps of 16 through LDS.  Bounds are
// checked (out-of-range operands read as zero).
constexpr int GT = 64, GK = 16;

// With split-K (gridDim.z > 1) work-group z covers K range [z kchunk, (z + 1) kchunk) and writes its
// partial tile to C + z M N (row-major, no accumulate); sum_split_kernel then adds the partials in
// order of z, so the result does not depend on scheduling.
__global__ __launch_bounds__(256)
void gemm_bf16_kernel(int64_t M, int64_t N, int64_t K, const __bf16* __restrict__ A, int64_t sAm, int64_t sAk,
                      const __bf16* __restrict__ B, int64_t sBk, int64_t sBn, float* __restrict__ C,
                      int64_t sCm, int64_t sCn, int accumulate, int64_t kchunk)
{
    __shared__ __bf16 sA[GT][GK + 8];     // [m][k]
    __shared__ __bf16 sB[GT][GK + 8];     // [n][k]
    const int tid = threadIdx.x, w = tid / 64, lane = tid % 64;
    const int64_t m0 = int64_t(blockIdx.y) * GT, n0 = int64_t(blockIdx.x) * GT;
    const int wm = (w / 2) * 32, wn = (w % 2) * 32;
    f32x16 acc = {};
    const int64_t kb = int64_t(blockIdx.z) * kchunk, ke = std::min<int64_t>(K, kb + kchunk);
    if (gridDim.z > 1)
    {
        C += int64_t(blockIdx.z) * M * N;
        sCm = N;
        sCn = 1;
        accumulate = 0;
    }
    for (int64_t k0 = kb; k0 < ke; k0 += GK)
    {
        // 64 x 16 elements of each operand, 4 per thread.
        for (int e = tid; e < GT * GK; e += 256)
        {
            const int r = e / GK, c = e % GK;
            const int64_t gm = m0 + r, gn = n0 + r, gk = k0 + c;
            sA[r][c] = (gm < M && gk < ke) ? A[gm * sAm + gk * sAk] : __bf16(0.f);
            sB[r][c] = (gn < N && gk < ke) ? B[gk * sBk + gn * sBn] : __bf16(0.f);
        }
        __syncthreads();
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e)
        {
            a[e] = sA[wm + lane % 32][8 * (lane / 32) + e];
            b[e] = sB[wn + lane % 32][8 * (lane / 32) + e];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 16; ++r)
    {
        const int64_t gm = m0 + wm + mfma_row(r, lane), gn = n0 + wn + lane % 32;
        if (gm < M && gn < N)
        {
            float* c = C + gm * sCm + gn * sCn;
            *c = accumulate ? *c + acc[r] : acc[r];
        }
    }
}

// C[m sCm + n sCn] (+)= sum over z of P[z][m][n], z in order.
__global__ void sum_split_kernel(int64_t M, int64_t N, int nsplit, const float* __restrict__ P, float* __restrict__ C,
                                 int64_t sCm, int64_t sCn, int accumulate)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= M * N) return;
    float s = 0.f;
    for (int z = 0; z < nsplit; ++z) s += P[int64_t(z) * M * N + i];
    float* c = C + (i / N) * sCm + (i % N) * sCn;
    *c = accumulate ? *c + s : s;
}

// Z (rows x n, FP32) += bias[n]; H = softplus(Z) as BF16 (both row-major, ld = n).
__global__ void bias_softplus_kernel(int64_t rows, int n, float* __restrict__ Z, const float* __restrict__ bias,
                                     __bf16* __restrict__ H)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows * n) return;
    const float z = Z[i] + bias[i % n];
    Z[i] = z;
    H[i] = __bf16(softplus(z));
}

// Y += bias; dY = 2 (Y - T) / (rows n) (mean squared error) as BF16; loss partial sums per block.
__global__ void mse_grad_kernel(int64_t rows, int n, float* __restrict__ Y, const float* __restrict__ bias,
                                const float* __restrict__ T, __bf16* __restrict__ dY, float* __restrict__ dYf,
                                float* __restrict__ loss_parts)
{
    __shared__ float red[256];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < rows * n)
    {
        const float y = Y[i] + bias[i % n];
        Y[i] = y;
        const float d = y - T[i];
        l = d * d;
        const float g = 2.f * d / float(rows * n);
        dY[i] = __bf16(g);
        dYf[i] = g;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_parts[blockIdx.x] = red[0];
}

// dZ = dH * softplus'(Z) = dH * sigmoid(Z); FP32 dH in, BF16 and FP32 dZ out.
__global__ void softplus_back_kernel(int64_t n, const float* __restrict__ dH, const float* __restrict__ Z,
                                     __bf16* __restrict__ dZ, float* __restrict__ dZf)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = dH[i] * sigmoid(Z[i]);
    dZ[i] = __bf16(g);
    dZf[i] = g;
}

// Partial column sums of dZ (rows x n, row-major): work-group (x, z) sums rows [z chunk, (z + 1)
// chunk) of columns [64 x, 64 x + 64), four row lanes per column combined in LDS in fixed order;
// P[z][j] receives the partial (summed over z by sum_split_kernel).
__global__ __launch_bounds__(256)
void bias_partial_kernel(int64_t rows, int n, int64_t chunk, const float* __restrict__ dZ, float* __restrict__ P)
{
    __shared__ float red[4][64];
    const int c = threadIdx.x % 64, q = threadIdx.x / 64;
    const int j = blockIdx.x * 64 + c;
    const int64_t r0 = int64_t(blockIdx.z) * chunk, r1 = std::min<int64_t>(rows, r0 + chunk);
    float s = 0.f;
    if (j < n)
        for (int64_t r = r0 + q; r < r1; r += 4) s += dZ[r * n + j];
    red[q][c] = s;
    __syncthreads();
    if (q == 0 && j < n) P[int64_t(blockIdx.z) * n + j] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
}

// Adam on FP32 master weights; writes the BF16 copy (Kingma & Ba 2015, bias-corrected).
__global__ void adam_kernel(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, __bf16* __restrict__ p16, float lr, float b1, float b2, float eps,
                            float c1, float c2)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (p16) p16[i] = __bf16(pi);
}

// FP32 -> BF16.
__global__ void to_bf16_kernel(int64_t n, const float* __restrict__ x, __bf16* __restrict__ y)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = __bf16(x[i]);
}

// ---------------------------------------------------------------------------------------------
// Fused inference, weight-stationary.  A persistent work-group of 8 wavefronts loops over units
// of 64 consecutive columns of one layer.  Each wavefront keeps its share of the weights in
// registers for the whole launch, as matrix-core A operands: row tile w % 4 of W1 (one K step)
// and row tiles w and w + 8 of W3 (8 K steps each); W2 is staged in LDS.  Per unit:
//   0. wavefront 0 turns the unit's inputs (loaded during the previous unit) into normalised BF16
//      inputs in LDS;
//   1. h1 = softplus(W1 x + b1): D (128 x 64) = W1 (128 x 16) x^T, wavefront w taking row tile
//      w % 4 and column tile w / 4; meanwhile (LW) the band Planck values of the layer, its levels
//      and the surface, bands spread over the wavefronts;
//   2. h2 = softplus(W2 h1 + b2) with the same split, K = 128;
//   3. y = W3 h2 (+ b3, denormalised) and the epilogue, which turns each (g-point, column) pair of
//      rows into the gas-optics outputs, adds the band cloud properties as the plain kernels do,
//      and stores 32 consecutive columns per half-wavefront; wavefront 0 first issues the next
//      unit's input loads.
// Only inputs and outputs touch HBM; the weights are read once per work-group.
// A pair outside the training envelope (any raw input beyond its range) puts its (64-column
// tile, layer) unit on the fallback list: fallback[0] counts the listed units (zeroed by the
// launcher) and fallback[1 ..] holds them as ilay * ntile + tile, in no particular order, for
// the plain gas optics to recompute.
struct MlpNet
{
    const __bf16* W1;        // (128, 16): 7 inputs, zero padding
    const __bf16* W2;        // (128, 128)
    const __bf16* W3;        // (no, 128)
    const float* b1;
    const float* b2;
    const float* b3;
    const float* x_mean;     // (7)
    const float* x_std;
    const float* x_lo;       // training envelope (raw inputs)
    const float* x_hi;
    const float* y_mean;     // (no)
    const float* y_std;
    int no;
    float ln_p_trop;         // ln of the k-distribution's tropopause pressure
};

// The inputs' gases in network order: h2o, o3, co2, ch4, n2o (vmr with broadcast strides).
struct MlpGases
{
    const float* p[5];
    int64_t scol[5], slay[5];
};

constexpr int MLP_P = 64;             // columns per unit
constexpr int MLP_MAXPLANCK = 4096;   // totplnk entries (nplanck x nbnd) staged in LDS
constexpr int MLP_MAXNO = 512;        // outputs the register-resident W3 can hold
constexpr int MLP_WAVES = 8;
constexpr int MLP_RT = MLP_MAXNO / 32 / MLP_WAVES;   // W3 row tiles per wavefront

__device__ __forceinline__ bf16x8 load_bf16x8(const __bf16* p)
{
    return *reinterpret_cast<const bf16x8*>(p);
}

// Raw inputs of one column of a unit (wavefront 0, lane = column).
struct MlpIn
{
    float t, p, v[5], cd, th, td, ts;
};

template<bool LW>
__device__ __forceinline__ void mlp_load_inputs(MlpIn& in, int64_t u, int64_t ntile, int64_t ncol, int lane,
                                                const float* __restrict__ play, const float* __restrict__ tlay,
                                                const float* __restrict__ tlev, const float* __restrict__ tsfc,
                                                const float* __restrict__ col_dry, const MlpGases& gases)
{
    const int64_t ilay = u / ntile, icol = (u % ntile) * MLP_P + lane;
    if (icol >= ncol) return;
    const int64_t idx = icol + ncol * ilay;
    in.t = tlay[idx];
    in.p = play[idx];
#pragma unroll
    for (int g = 0; g < 5; ++g) in.v[g] = gases.p[g][icol * gases.scol[g] + ilay * gases.slay[g]];
    in.cd = col_dry[idx];
    if constexpr (LW)
    {
        in.th = tlev[icol + ncol * (ilay + 1)];
        in.td = tlev[idx];
        in.ts = tsfc[icol];
    }
}

template<bool LW>
__global__ __launch_bounds__(64 * MLP_WAVES) __attribute__((amdgpu_waves_per_eu(2)))
void mlp_gas_optics_kernel(MlpNet net, int64_t ncol, int64_t nlay, int ngpt, int nbnd, const int* __restrict__ gpt2band,
                           const float* __restrict__ play, const float* __restrict__ tlay,
                           const float* __restrict__ tlev, const float* __restrict__ tsfc,
                           const float* __restrict__ col_dry, MlpGases gases,
                           const float* __restrict__ totplnk, int nplanck, float temp_min, float dplanck,
                           int sfc_lay, const float* __restrict__ ctau, const float* __restrict__ cssa,
                           const float* __restrict__ cg, float* __restrict__ out0, float* __restrict__ out1,
                           float* __restrict__ out2, float* __restrict__ out3, float* __restrict__ sfc_src,
                           float* __restrict__ sfc_jac, int* __restrict__ fallback,
                           float* __restrict__ bnd_lay, float* __restrict__ bnd_lev)
{
    constexpr int P = MLP_P;
    __shared__ __align__(16) __bf16 xs[P][MLP_INP + 8];
    __shared__ __align__(16) __bf16 h1[P][MLP_H + 8];
    __shared__ __align__(16) __bf16 h2[P][MLP_H + 8];
    __shared__ __align__(16) __bf16 sW2[MLP_H][MLP_H + 8];
    __shared__ float sB[LW ? 4 : 1][16][P];         // LW: Planck values per (band, column)
    __shared__ float sT[LW ? 4 : 1][P];             // LW: T of the layer, upper and lower level, surface
    __shared__ float sC[LW ? 1 : 3][16][P];         // band cloud tau (LW); tau, ssa, g (SW)
    __shared__ float sPl[LW ? MLP_MAXPLANCK : 1];   // LW: totplnk (nplanck, nbnd)
    __shared__ float s_cd[P];
    __shared__ float s_b1[MLP_H], s_b2[MLP_H];
    __shared__ float4 s_so[MLP_MAXNO / 2];        // per g-point: (scale, offset) of rows 2g, 2g+1
    __shared__ int s_g2b[MLP_MAXNO / 2];
    __shared__ int s_out;
    const int tid = threadIdx.x, lane = tid % 64;
    const int w = __builtin_amdgcn_readfirstlane(tid / 64);
    const int no = net.no, nrt = no / 32;
    const int64_t ntile = (ncol + P - 1) / P;
    const int64_t nunits = ntile * nlay;
    const int64_t ncl = ncol * nlay;

    // Register-resident weights (A operands: lane l holds row l % 32, K half l / 32).
    const int ar = lane % 32, ak = 8 * (lane / 32);
    const int rt12 = w % 4, pt12 = w / 4;           // layers 1 and 2: row tile, column tile
    const bf16x8 A1 = load_bf16x8(net.W1 + (rt12 * 32 + ar) * MLP_INP + ak);
    for (int e = tid; e < MLP_H * MLP_H / 8; e += 64 * MLP_WAVES)
    {
        const int r = e / (MLP_H / 8), c = 8 * (e % (MLP_H / 8));
        *reinterpret_cast<bf16x8*>(&sW2[r][c]) = load_bf16x8(net.W2 + r * MLP_H + c);
    }
    bf16x8 A3[MLP_RT][MLP_H / 16];
#pragma unroll
    for (int i = 0; i < MLP_RT; ++i)
    {
        const int rt = w + MLP_WAVES * i;
#pragma unroll
        for (int ks = 0; ks < MLP_H / 16; ++ks)
            A3[i][ks] = rt < nrt ? load_bf16x8(net.W3 + int64_t(rt * 32 + ar) * MLP_H + ks * 16 + ak) : bf16x8{};
    }
    for (int e = tid; e < MLP_H; e += 64 * MLP_WAVES) { s_b1[e] = net.b1[e]; s_b2[e] = net.b2[e]; }
    for (int e = tid; e < no / 2; e += 64 * MLP_WAVES)
    {
        const int a = 2 * e, b = 2 * e + 1;
        s_so[e] = make_float4(net.y_std[a], net.b3[a] * net.y_std[a] + net.y_mean[a],
                              net.y_std[b], net.b3[b] * net.y_std[b] + net.y_mean[b]);
    }
    for (int e = tid; e < no / 2; e += 64 * MLP_WAVES) s_g2b[e] = gpt2band[e];
    if constexpr (LW)
        for (int e = tid; e < nplanck * nbnd; e += 64 * MLP_WAVES) sPl[e] = totplnk[e];

    MlpIn in{};
    if (w == 0 && int64_t(blockIdx.x) < nunits)
        mlp_load_inputs<LW>(in, blockIdx.x, ntile, ncol, lane, play, tlay, tlev, tsfc, col_dry, gases);
    // Cloud properties of bands w and w + 8 for the next unit's columns, one unit ahead.
    float cl[LW ? 1 : 3][2] = {};
    auto load_cloud = [&](int64_t v) {
        if (!ctau) return;
        const int64_t vt = v % ntile, vl = v / ntile;
        const int64_t icol = vt * P + lane;
        if (icol >= ncol) return;
        const int64_t idx = icol + ncol * vl;
#pragma unroll
        for (int k = 0; k < 2; ++k)
        {
            const int b = w + MLP_WAVES * k;
            if (b < nbnd)
            {
                cl[0][k] = ctau[idx + ncl * b];
                if constexpr (!LW) { cl[1][k] = cssa[idx + ncl * b]; cl[2][k] = cg[idx + ncl * b]; }
            }
        }
    };
    if (int64_t(blockIdx.x) < nunits) load_cloud(blockIdx.x);

    for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x)
    {
        const int64_t ilay = u / ntile, col0 = (u % ntile) * P;
        const bool at_sfc = LW && (ilay == sfc_lay - 1);
        if (tid == 0) s_out = 0;
        __syncthreads();     // the previous unit is done with the LDS buffers; constants staged

        // 0. Normalised inputs (wavefront 0, lane = column).
        if (w == 0)
        {
            const bool valid = col0 + lane < ncol;
            float x[MLP_IN] = {};
            bool out = false;
            if (valid)
            {
                x[0] = in.t;
                x[1] = __logf(in.p);
#pragma unroll
                for (int g = 0; g < 5; ++g) x[2 + g] = __logf(in.v[g]);
                x[7] = x[1] > net.ln_p_trop ? 1.f : 0.f;
#pragma unroll
                for (int i = 0; i < MLP_IN; ++i) out |= !(x[i] >= net.x_lo[i] && x[i] <= net.x_hi[i]);
            }
            s_cd[lane] = valid ? in.cd : 0.f;
#pragma unroll
            for (int i = 0; i < MLP_INP; ++i)
                xs[lane][i] = i < MLP_IN ? __bf16((x[i] - net.x_mean[i]) / net.x_std[i]) : __bf16(0.f);
            if constexpr (LW)
            {
                sT[0][lane] = valid ? in.t : temp_min;
                sT[1][lane] = valid ? in.th : temp_min;
                sT[2][lane] = valid ? in.td : temp_min;
                sT[3][lane] = valid ? in.ts : temp_min;
            }
            if (out) atomicOr(&s_out, 1);
        }
        __syncthreads();

        // 1. Layer 1 on the matrix cores (K = 16: the 7 inputs and padding); LW Planck values.
        {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&xs[pt12 * 32 + lane % 32][ak]);
            f32x16 acc = {};
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, b, acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r)
            {
                const int j = rt12 * 32 + mfma_row(r, lane);
                h1[pt12 * 32 + lane % 32][j] = __bf16(softplus_fast(acc[r] + s_b1[j]));
            }
        }
        if (ctau)
        {
            // Band cloud properties of the unit's columns (loaded during the previous unit),
            // bands w and w + 8.
#pragma unroll
            for (int k = 0; k < 2; ++k)
            {
                const int b = w + MLP_WAVES * k;
                if (b < nbnd)
                {
                    sC[0][b][lane] = cl[0][k];
                    if constexpr (!LW) { sC[1][b][lane] = cl[1][k]; sC[2][b][lane] = cl[2][k]; }
                }
            }
        }
        if constexpr (LW)
        {
            // interpolate1D over totplnk for the band's Planck values of this column; with factored
            // sources also the band Planck functions at the layer and its lower level (and the top
            // level from the top layer).
            const int64_t icol = col0 + lane, nlev = nlay + 1;
            for (int b = w; b < nbnd; b += MLP_WAVES)
            {
                const float* tab = sPl + b * nplanck;
                float bq[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    const float v0 = (sT[q][lane] - temp_min) / dplanck;
                    const float fr = v0 - float(int(v0));
                    const int id = min(nplanck - 1, max(1, int(v0) + 1));
                    bq[q] = tab[id - 1] + fr * (tab[id] - tab[id - 1]);
                    sB[q][b][lane] = bq[q];
                }
                if (bnd_lay && icol < ncol)
                {
                    bnd_lay[icol + ncol * ilay + ncl * b] = bq[0];
                    bnd_lev[icol + ncol * ilay + ncol * nlev * b] = bq[2];
                    if (ilay == nlay - 1) bnd_lev[icol + ncol * nlay + ncol * nlev * b] = bq[1];
                }
            }
        }
        __syncthreads();

        // 2. Layer 2.
        {
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < MLP_H / 16; ++ks)
            {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sW2[rt12 * 32 + ar][ks * 16 + ak]);
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(&h1[pt12 * 32 + lane % 32][ks * 16 + ak]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 16; ++r)
            {
                const int j = rt12 * 32 + mfma_row(r, lane);
                h2[pt12 * 32 + lane % 32][j] = __bf16(softplus_fast(acc[r] + s_b2[j]));
            }
        }
        __syncthreads();

        // Next unit's inputs and cloud properties, in flight during layer 3.
        if (w == 0 && u + gridDim.x < nunits)
            mlp_load_inputs<LW>(in, u + gridDim.x, ntile, ncol, lane, play, tlay, tlev, tsfc, col_dry, gases);
        if (u + gridDim.x < nunits) load_cloud(u + gridDim.x);

        // 3. Layer 3 and the epilogue.  Rows 2g, 2g+1 of D are g-point g; for register r the
        // g-point is gbase + c(r), gbase = 16 rt + 2 (lane / 32), c(r) = 4 (r / 4) + (r % 4) / 2.
#pragma unroll 1
        for (int pt = 0; pt < 2; ++pt)
        {
            bf16x8 bop[MLP_H / 16];
#pragma unroll
            for (int ks = 0; ks < MLP_H / 16; ++ks)
                bop[ks] = *reinterpret_cast<const bf16x8*>(&h2[pt * 32 + lane % 32][ks * 16 + ak]);
            const int pl = pt * 32 + lane % 32;
            const int64_t icol = col0 + pl;
            const int64_t idx = icol + ncol * ilay;
            const float cd = s_cd[pl];
#pragma unroll
            for (int i = 0; i < MLP_RT; ++i)
            {
                const int rt = w + MLP_WAVES * i;
                if (rt >= nrt) break;                           // uniform over the wavefront
                f32x16 acc = {};
#pragma unroll
                for (int ks = 0; ks < MLP_H / 16; ++ks)
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A3[i][ks], bop[ks], acc, 0, 0, 0);
                if (icol >= ncol) continue;
                const int gbase = 16 * rt + 2 * (lane / 32);
                // Output (column icol, g-point gbase + c): uniform part col0 + ncol ilay + ncl 16 rt
                // (+ ncl c), lane part pl + 2 ncl (lane / 32), below 2^32 bytes by the host check.
                const int64_t ubase = col0 + ncol * ilay + ncl * (16 * rt);
                const uint32_t vofs = uint32_t(pl) + uint32_t(2 * ncl) * uint32_t(lane / 32);
#pragma unroll
                for (int r = 0; r < 16; r += 2)
                {
                    const int c = 4 * (r / 4) + (r % 4) / 2;
                    const int g = gbase + c;
                    const int64_t uo = ubase + ncl * c;                     // uniform
                    const float4 so = s_so[g];
                    const float y0 = acc[r] * so.x + so.y;
                    const float y1 = acc[r + 1] * so.z + so.w;
                    const int b = s_g2b[g];
                    if constexpr (LW)
                    {
                        const float tau = __expf(y0) * cd;
                        const float pf = y1;
                        store_at(out0 + uo, vofs, ctau ? tau + sC[0][b][pl] : tau);
                        if (bnd_lay)
                            store_at(out1 + uo, vofs, pf);               // factored: the fraction only
                        else
                        {
                            store_at(out1 + uo, vofs, pf * sB[0][b][pl]);
                            store_at(out2 + uo, vofs, pf * sB[1][b][pl]);
                            store_at(out3 + uo, vofs, pf * sB[2][b][pl]);
                        }
                        if (at_sfc)
                        {
                            // sfc_jac: pf (B(T_sfc + 1) - B(T_sfc)), as the plain gas optics.
                            const float v0 = (sT[3][pl] + 1.f - temp_min) / dplanck;
                            const float fr = v0 - float(int(v0));
                            const int id = min(nplanck - 1, max(1, int(v0) + 1));
                            const float* tab = sPl + b * nplanck;
                            const float b1s = tab[id - 1] + fr * (tab[id] - tab[id - 1]);
                            const int64_t os = icol + ncol * g;
                            sfc_src[os] = pf * sB[3][b][pl];
                            sfc_jac[os] = pf * (b1s - sB[3][b][pl]);
                        }
                    }
                    else
                    {
                        const float tabs = __expf(y0) * cd, tray = __expf(y1) * cd;
                        const float tt = tabs + tray;
                        const float eps2 = float(2 * prop_eps), op_eps = 3.f * FLT_MIN;
                        float ss = tt > eps2 ? __fdividef(tray, tt) : 0.f, tt_out = tt, gg = 0.f;
                        if (ctau)
                        {
                            const float t2 = sC[0][b][pl], s2 = sC[1][b][pl], g2 = sC[2][b][pl];
                            const float tau12 = tt + t2, tauscat = tt * ss + t2 * s2;
                            gg = __fdividef(t2 * s2 * g2, fmaxf(tauscat, op_eps));
                            ss = __fdividef(tauscat, fmaxf(op_eps, tau12));
                            tt_out = tau12;
                        }
                        store_at(out0 + uo, vofs, tt_out);
                        store_at(out1 + uo, vofs, ss);
                        store_at(out2 + uo, vofs, gg);
                    }
                }
            }
        }
        if (tid == 0 && s_out)
        {
            const int pos = atomicAdd(fallback, 1);
            fallback[1 + pos] = int(u);
        }
    }
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

// ---------------------------------------------------------------------------------------------
// Training data from the plain gas optics (clear sky).  Row r = (icol, ilay) pair, icol fastest.
// X (rows, 16): raw inputs, zero padded.  Y (rows, no): interleaved targets.
__global__ void mlp_inputs_kernel(int64_t ncol, int64_t nlay, const float* __restrict__ play,
                                  const float* __restrict__ tlay, MlpGases gases, float ln_p_trop,
                                  float* __restrict__ X)
{
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= ncol * nlay) return;
    const int64_t icol = r % ncol, ilay = r / ncol;
    float* x = X + r * MLP_INP;
    x[0] = tlay[r];
    x[1] = logf(play[r]);
#pragma unroll
    for (int g = 0; g < 5; ++g) x[2 + g] = logf(gases.p[g][icol * gases.scol[g] + ilay * gases.slay[g]]);
    x[7] = x[1] > ln_p_trop ? 1.f : 0.f;
#pragma unroll
    for (int i = MLP_IN; i < MLP_INP; ++i) x[i] = 0.f;
}

template<bool LW>
__global__ void mlp_targets_kernel(int64_t ncol, int64_t nlay, int ngpt, const int* __restrict__ gpt2band,
                                   const float* __restrict__ tlay, const float* __restrict__ col_dry,
                                   const float* __restrict__ a, const float* __restrict__ b,
                                   const float* __restrict__ totplnk, int nplanck, float temp_min, float dplanck,
                                   float* __restrict__ Y)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t ncl = ncol * nlay;
    if (i >= ncl * ngpt) return;
    const int64_t r = i % ncl;
    const int g = int(i / ncl);
    // ln(tau / col_dry) as ln tau - ln col_dry (tau / col_dry can be below the FP32 range), floored
    // at MLP_LN_FLOOR where tau = 0: a constant the network can learn, independent of col_dry.
    const float lcd = logf(col_dry[r]);
    auto lnr = [&](float t) { return t > 0.f ? fmaxf(logf(t) - lcd, MLP_LN_FLOOR) : MLP_LN_FLOOR; };
    float y0, y1;
    if (LW)
    {
        // a = tau, b = lay_src: pf = lay_src / B(T_lay) of the g-point's band.
        const float v0 = (tlay[r] - temp_min) / dplanck;
        const float fr = v0 - float(int(v0));
        const int id = min(nplanck - 1, max(1, int(v0) + 1));
        const float* tab = totplnk + int64_t(gpt2band[g]) * nplanck;
        const float B = tab[id - 1] + fr * (tab[id] - tab[id - 1]);
        y0 = lnr(a[i]);
        y1 = b[i] / B;
    }
    else
    {
        // a = tau, b = ssa: tau_abs = tau (1 - ssa), tau_ray = tau ssa.
        y0 = lnr(a[i] * (1.f - b[i]));
        y1 = lnr(a[i] * b[i]);
    }
    Y[r * (2 * ngpt) + 2 * g] = y0;
    Y[r * (2 * ngpt) + 2 * g + 1] = y1;
}

// dst[i] = src[idx[i]] for rows of n FP32 values.
__global__ void gather_rows_kernel(int64_t rows, int n, const int64_t* __restrict__ idx, const float* __restrict__ src,
                                   float* __restrict__ dst)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows * n) return;
    const int64_t r = i / n, c = i % n;
    dst[i] = src[idx[r] * n + c];
}

// Row-wise (x - mean) / std over n columns (the first n_norm of each row), FP32 -> BF16.
__global__ void normalize_rows_kernel(int64_t rows, int n, int n_norm, const float* __restrict__ x,
                                      const float* __restrict__ mean, const float* __restrict__ std_,
                                      __bf16* __restrict__ y16, float* __restrict__ y32)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows * n) return;
    const int c = int(i % n);
    const float v = c < n_norm ? (x[i] - mean[c]) / std_[c] : 0.f;
    if (y16) y16[i] = __bf16(v);
    if (y32) y32[i] = v;
}

// ---------------------------------------------------------------------------------------------
// Launchers.
int gemm_bf16(int64_t M, int64_t N, int64_t K, const void* A, int64_t sAm, int64_t sAk, const void* B, int64_t sBk,
              int64_t sBn, float* C, int64_t sCm, int64_t sCn, int accumulate, void* stream)
{
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return arg_error("mlp_gemm: bad arguments");
    const int64_t tiles = grid1(N, GT) * grid1(M, GT);
    // Split K when there are few output tiles and a long K (the weight gradients): at least 256 of
    // K per split, about 256 work-groups in all.
    int64_t nsplit = 1;
    if (tiles < 128 && K >= 512) nsplit = std::max<int64_t>(1, std::min<int64_t>(K / 256, 256 / tiles));
    const int64_t kchunk = ((K + nsplit - 1) / nsplit + GK - 1) / GK * GK;
    nsplit = (K + kchunk - 1) / kchunk;
    const hipStream_t s = as_stream(stream);
    float* part = nullptr;
    if (nsplit > 1) RRTMGP_TRY(scratch_alloc(&part, size_t(nsplit * M * N), s));
    const dim3 grid(grid1(N, GT), grid1(M, GT), unsigned(nsplit));
    RRTMGP_TIMED("gemm_bf16_kernel", s);
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, s, M, N, K, static_cast<const __bf16*>(A), sAm, sAk,
                       static_cast<const __bf16*>(B), sBk, sBn, nsplit > 1 ? part : C, sCm, sCn, accumulate, kchunk);
    RRTMGP_LAUNCHED("gemm_bf16_kernel");
    if (nsplit > 1)
    {
        hipLaunchKernelGGL(sum_split_kernel, dim3(grid1(M * N, 256)), dim3(256), 0, s, M, N, int(nsplit), part, C,
                           sCm, sCn, accumulate);
        scratch_free(part, s);
    }
    return RRTMGP_OK;
}
}  // namespace rrtmgp_hip

using namespace rrtmgp_hip;

extern "C" {

int rrtmgp_hip_mlp_gemm_bf16(int64_t M, int64_t N, int64_t K, const void* A, int64_t sAm, int64_t sAk, const void* B,
                             int64_t sBk, int64_t sBn, float* C, int64_t sCm, int64_t sCn, int accumulate, void* stream)
{
    return gemm_bf16(M, N, K, A, sAm, sAk, B, sBk, sBn, C, sCm, sCn, accumulate, stream);
}

int rrtmgp_hip_mlp_bias_softplus(int64_t rows, int n, float* Z, const float* bias, void* H, void* stream)
{
    if (rows <= 0 || n <= 0 || !Z || !bias || !H) return arg_error("mlp_bias_softplus: bad arguments");
    RRTMGP_TIMED("bias_softplus_kernel", as_stream(stream));
    hipLaunchKernelGGL(bias_softplus_kernel, dim3(grid1(rows * n, 256)), dim3(256), 0, as_stream(stream), rows, n, Z,
                       bias, static_cast<__bf16*>(H));
    RRTMGP_LAUNCHED("bias_softplus_kernel");
    return RRTMGP_OK;
}

int rrtmgp_hip_mlp_mse_grad(int64_t rows, int n, float* Y, const float* bias, const float* T, void* dY, float* dYf,
                            float* loss_parts, void* stream)
{
    if (rows <= 0 || n <= 0 || !Y || !bias || !T || !dY || !dYf || !loss_parts)
        return arg_error("mlp_mse_grad: bad arguments");
    RRTMGP_TIMED("mse_grad_kernel", as_stream(stream));
    hipLaunchKernelGGL(mse_grad_kernel, dim3(grid1(rows * n, 256)), dim3(256), 0, as_stream(stream), rows, n, Y, bias,
                       T, static_cast<__bf16*>(dY), dYf, loss_parts);
    RRTMGP_LAUNCHED("mse_grad_kernel");
    return RRTMGP_OK;
}

int rrtmgp_hip_mlp_softplus_back(int64_t n, const float* dH, const float* Z, void* dZ, float* dZf, void* stream)
{
    if (n <= 0 || !dH || !Z || !dZ || !dZf) return arg_error("mlp_softplus_back: bad arguments");
    RRTMGP_TIMED("softplus_back_kernel", as_stream(stream));
    hipLaunchKernelGGL(softplus_back_kernel, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream), n, dH, Z,
                       static_cast<__bf16*>(dZ), dZf);
    RRTMGP_LAUNCHED("softplus_back_kernel");
    return RRTMGP_OK;
}

int rrtmgp_hip_mlp_bias_grad(int64_t rows, int n, const float* dZ, float* db, void* stream)
{
    if (rows <= 0 || n <= 0 || !dZ || !db) return arg_error("mlp_bias_grad: bad arguments");
    const hipStream_t s = as_stream(stream);
    const int64_t nz = std::max<int64_t>(1, std::min<int64_t>(64, rows / 64));
    const int64_t chunk = (rows + nz - 1) / nz;
    float* part = nullptr;
    RRTMGP_TRY(scratch_alloc(&part, size_t(nz * n), s));
    RRTMGP_TIMED("bias_partial_kernel", s);
    hipLaunchKernelGGL(bias_partial_kernel, dim3(grid1(n, 64), 1, unsigned(nz)), dim3(256), 0, s, rows, n, chunk, dZ,
                       part);
    RRTMGP_LAUNCHED("bias_partial_kernel");
    hipLaunchKernelGGL(sum_split_kernel, dim3(grid1(n, 256)), dim3(256), 0, s, int64_t(1), int64_t(n), int(nz), part,
                       db, int64_t(n), int64_t(1), 0);
    scratch_free(part, s);
    return RRTMGP_OK;
}

int rrtmgp_hip_mlp_adam(int64_t n, float* p, const float* g, float* m, float* v, void* p_bf16, float lr, float beta1,
                        float beta2, float eps, int64_t step, void* stream)
{
    if (n <= 0 || !p || !g || !m || !v || step < 1) return arg_error("mlp_adam: bad arguments");
    const float c1 = 1.f - std::pow(beta1, float(step)), c2 = 1.f - std::pow(beta2, float(step));
    RRTMGP_TIMED("adam_kernel", as_stream(stream));
    hipLaunchKernelGGL(adam_kernel, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream), n, p, g, m, v,
                       static_cast<__bf16*>(p_bf16), lr, beta1, beta2, eps, c1, c2);
    RRTMGP_LAUNCHED("adam_kernel");
    return RRTMGP_OK;
}

int rrtmgp_hip_mlp_to_bf16(int64_t n, const float* x, void* y, void* stream)
{
    if (n <= 0 || !x || !y) return arg_error("mlp_to_bf16: bad arguments");
    RRTMGP_TIMED("to_bf16_kernel", as_stream(stream));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream), n, x,
                       static_cast<__bf16*>(y));
    RRTMGP_LAUNCHED("to_bf16_kernel");
    return RRTMGP_OK;
}


// Persistent grid of the inference kernel: one 8-wavefront work-group per compute unit (two
// wavefronts per SIMD: its register budget), no more than there are units.
static unsigned mlp_grid(int64_t nunits)
{
    return unsigned(std::min<int64_t>(nunits, int64_t(device_cu_count())));
}

// The register-resident weights are read as 16-byte vectors.
static int check_mlp_weights(const MlpNet& net)
{
    auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    if (!al(net.W1) || !al(net.W2) || !al(net.W3)) return arg_error("mlp: weights must be 16-byte aligned");
    if (net.no > MLP_MAXNO) return arg_error("mlp: at most 512 outputs (256 g-points)");
    return RRTMGP_OK;
}

// The epilogue's lane offsets (2 ncol nlay + 63 elements) must fit 32 bits of bytes.
static int check_mlp_offsets(int64_t ncol, int64_t nlay)
{
    if ((2 * ncol * nlay + 64) * int64_t(sizeof(float)) >= (int64_t(1) << 32))
        return arg_error("mlp gas optics: ncol * nlay too large for one call; split the columns into blocks");
    return RRTMGP_OK;
}

static int make_mlp_net(const rrtmgp_mlp_net* n, MlpNet& net)
{
    if (!n || !n->W1 || !n->W2 || !n->W3 || !n->b1 || !n->b2 || !n->b3 || !n->x_mean || !n->x_std || !n->x_lo ||
        !n->x_hi || !n->y_mean || !n->y_std || n->n_out <= 0 || n->n_out % 32 != 0)
        return arg_error("mlp: incomplete network (the output count must be a multiple of 32)");
    net.W1 = static_cast<const __bf16*>(n->W1); net.W2 = static_cast<const __bf16*>(n->W2);
    net.W3 = static_cast<const __bf16*>(n->W3);
    net.b1 = n->b1; net.b2 = n->b2; net.b3 = n->b3; net.x_mean = n->x_mean; net.x_std = n->x_std;
    net.x_lo = n->x_lo; net.x_hi = n->x_hi; net.y_mean = n->y_mean; net.y_std = n->y_std; net.no = int(n->n_out);
    if (!(n->p_trop > 0.0)) return arg_error("mlp: the network needs the tropopause pressure of its training");
    net.ln_p_trop = float(std::log(n->p_trop));
    return RRTMGP_OK;
}

static int make_mlp_gases(const rrtmgp_gas_vmr* g, int64_t ncol, int64_t nlay, MlpGases& out)
{
    if (!g || g->ngas != 5) return arg_error("mlp: expects the 5 input gases h2o, o3, co2, ch4, n2o");
    for (int i = 0; i < 5; ++i)
    {
        const int64_t d1 = g->dim1[i], d2 = g->dim2[i];
        if (!g->vmr[i] || !((d1 == 1 || d1 == ncol) && (d2 == 1 || d2 == nlay))) return arg_error("mlp: bad gas vmr");
        out.p[i] = static_cast<const float*>(g->vmr[i]);
        out.scol[i] = d1 == 1 ? 0 : 1;
        out.slay[i] = d2 == 1 ? 0 : d1;
    }
    return RRTMGP_OK;
}

namespace
{
int mlp_lw(const rrtmgp_mlp_net* n, int64_t ncol, int64_t nlay, int64_t ngpt, int64_t nbnd, const int* gpt2band,
           const float* play, const float* tlay, const float* tlev, const float* tsfc, const float* col_dry,
           const rrtmgp_gas_vmr* gases, const float* totplnk, int64_t nplanck, double temp_min, double dplanck,
           int sfc_lay, const float* cloud_tau, float* tau, float* lay_src, float* lev_src_inc, float* lev_src_dec,
           float* sfc_src, float* sfc_src_jac, int* fallback, float* bnd_lay, float* bnd_lev, void* stream)
{
    MlpNet net;
    RRTMGP_TRY(make_mlp_net(n, net));
    MlpGases g;
    RRTMGP_TRY(make_mlp_gases(gases, ncol, nlay, g));
    const bool factored = bnd_lay != nullptr;
    if (ncol <= 0 || nlay <= 0 || net.no != 2 * ngpt || nbnd < 1 || nbnd > 16 || nplanck < 2 ||
        nplanck * nbnd > MLP_MAXPLANCK || !gpt2band || !play ||
        !tlay || !tlev || !tsfc || !col_dry || !totplnk || !tau || !lay_src ||
        (factored ? !bnd_lev : (!lev_src_inc || !lev_src_dec)) ||
        !sfc_src || !sfc_src_jac || !fallback || (sfc_lay != 1 && sfc_lay != nlay))
        return arg_error("mlp_gas_optics_lw: bad arguments");
    const int64_t nunits = ((ncol + 63) / 64) * nlay;
    if (nunits >= (int64_t(1) << 31)) return arg_error("mlp_gas_optics_lw: too many column tiles");
    RRTMGP_TRY(check_mlp_weights(net));
    RRTMGP_TRY(check_mlp_offsets(ncol, nlay));
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, as_stream(stream), fallback);
    RRTMGP_TIMED("mlp_gas_optics_lw_kernel", as_stream(stream));
    hipLaunchKernelGGL(mlp_gas_optics_kernel<true>, dim3(mlp_grid(nunits)), dim3(64 * MLP_WAVES), 0, as_stream(stream), net,
                       ncol, nlay, int(ngpt), int(nbnd), gpt2band, play, tlay, tlev, tsfc, col_dry, g, totplnk,
                       int(nplanck), float(temp_min), float(dplanck), sfc_lay, cloud_tau, nullptr, nullptr, tau,
                       lay_src, lev_src_inc, lev_src_dec, sfc_src, sfc_src_jac, fallback, bnd_lay, bnd_lev);
    RRTMGP_LAUNCHED("mlp_gas_optics_lw_kernel");
    return RRTMGP_OK;
}
}  // namespace

int rrtmgp_hip_mlp_gas_optics_lw(const rrtmgp_mlp_net* n, int64_t ncol, int64_t nlay, int64_t ngpt, int64_t nbnd,
                                 const int* gpt2band, const float* play, const float* tlay, const float* tlev,
                                 const float* tsfc, const float* col_dry, const rrtmgp_gas_vmr* gases,
                                 const float* totplnk, int64_t nplanck, double temp_min, double dplanck, int sfc_lay,
                                 const float* cloud_tau, float* tau, float* lay_src, float* lev_src_inc,
                                 float* lev_src_dec, float* sfc_src, float* sfc_src_jac, int* fallback, void* stream)
{
    return mlp_lw(n, ncol, nlay, ngpt, nbnd, gpt2band, play, tlay, tlev, tsfc, col_dry, gases, totplnk, nplanck,
                  temp_min, dplanck, sfc_lay, cloud_tau, tau, lay_src, lev_src_inc, lev_src_dec, sfc_src, sfc_src_jac,
                  fallback, nullptr, nullptr, stream);
}

int rrtmgp_hip_mlp_gas_optics_lw_factored(const rrtmgp_mlp_net* n, int64_t ncol, int64_t nlay, int64_t ngpt,
                                          int64_t nbnd, const int* gpt2band, const float* play, const float* tlay,
                                          const float* tlev, const float* tsfc, const float* col_dry,
                                          const rrtmgp_gas_vmr* gases, const float* totplnk, int64_t nplanck,
                                          double temp_min, double dplanck, int sfc_lay, const float* cloud_tau,
                                          float* tau, float* planck_frac, float* band_lay_src, float* band_lev_src,
                                          float* sfc_src, float* sfc_src_jac, int* fallback, void* stream)
{
    if (!band_lay_src || !band_lev_src) return arg_error("mlp_gas_optics_lw_factored: NULL band array");
    return mlp_lw(n, ncol, nlay, ngpt, nbnd, gpt2band, play, tlay, tlev, tsfc, col_dry, gases, totplnk, nplanck,
                  temp_min, dplanck, sfc_lay, cloud_tau, tau, planck_frac, nullptr, nullptr, sfc_src, sfc_src_jac,
                  fallback, band_lay_src, band_lev_src, stream);
}

int rrtmgp_hip_mlp_gas_optics_sw(const rrtmgp_mlp_net* n, int64_t ncol, int64_t nlay, int64_t ngpt, int64_t nbnd,
                                 const int* gpt2band, const float* play, const float* tlay, const float* col_dry,
                                 const rrtmgp_gas_vmr* gases, const float* cloud_tau, const float* cloud_ssa,
                                 const float* cloud_g, float* tau, float* ssa, float* g, int* fallback, void* stream)
{
    MlpNet net;
    RRTMGP_TRY(make_mlp_net(n, net));
    MlpGases gs;
    RRTMGP_TRY(make_mlp_gases(gases, ncol, nlay, gs));
    if (ncol <= 0 || nlay <= 0 || net.no != 2 * ngpt || nbnd < 1 || nbnd > 16 || !gpt2band || !play || !tlay || !col_dry ||
        !tau || !ssa || !g || !fallback || (cloud_tau && (!cloud_ssa || !cloud_g)))
        return arg_error("mlp_gas_optics_sw: bad arguments");
    const int64_t nunits = ((ncol + 63) / 64) * nlay;
    if (nunits >= (int64_t(1) << 31)) return arg_error("mlp_gas_optics_sw: too many column tiles");
    RRTMGP_TRY(check_mlp_weights(net));
    RRTMGP_TRY(check_mlp_offsets(ncol, nlay));
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, as_stream(stream), fallback);
    RRTMGP_TIMED("mlp_gas_optics_sw_kernel", as_stream(stream));
    hipLaunchKernelGGL(mlp_gas_optics_kernel<false>, dim3(mlp_grid(nunits)), dim3(64 * MLP_WAVES), 0, as_stream(stream), net,
                       ncol, nlay, int(ngpt), int(nbnd), gpt2band, play, tlay, nullptr, nullptr, col_dry, gs,
                       nullptr, 2, 0.f, 1.f, 1, cloud_tau, cloud_ssa, cloud_g, tau, ssa, g, nullptr, nullptr,
                       nullptr, fallback, nullptr, nullptr);
    RRTMGP_LAUNCHED("mlp_gas_optics_sw_kernel");
    return RRTMGP_OK;
}

int rrtmgp_hip_mlp_inputs(int64_t ncol, int64_t nlay, const float* play, const float* tlay,
                          const rrtmgp_gas_vmr* gases, double p_trop, float* X, void* stream)
{
    MlpGases g;
    RRTMGP_TRY(make_mlp_gases(gases, ncol, nlay, g));
    if (!play || !tlay || !X || !(p_trop > 0.0)) return arg_error("mlp_inputs: bad arguments");
    RRTMGP_TIMED("mlp_inputs_kernel", as_stream(stream));
    hipLaunchKernelGGL(mlp_inputs_kernel, dim3(grid1(ncol * nlay, 256)), dim3(256), 0, as_stream(stream), ncol,
                       nlay, play, tlay, g, float(std::log(p_trop)), X);
    RRTMGP_LAUNCHED("mlp_inputs_kernel");
    return RRTMGP_OK;
}

int rrtmgp_hip_mlp_targets(int lw, int64_t ncol, int64_t nlay, int64_t ngpt, const int* gpt2band, const float* tlay,
                           const float* col_dry, const float* a, const float* b, const float* totplnk,
                           int64_t nplanck, double temp_min, double dplanck, float* Y, void* stream)
{
    if (ncol <= 0 || nlay <= 0 || ngpt <= 0 || !gpt2band || !tlay || !col_dry || !a || !b || !Y ||
        (lw && (!totplnk || nplanck < 2)))
        return arg_error("mlp_targets: bad arguments");
    const int64_t n = ncol * nlay * ngpt;
    RRTMGP_TIMED("mlp_targets_kernel", as_stream(stream));
    if (lw)
        hipLaunchKernelGGL(mlp_targets_kernel<true>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream), ncol, nlay,
                           int(ngpt), gpt2band, tlay, col_dry, a, b, totplnk, int(nplanck), float(temp_min),
                           float(dplanck), Y);
    else
        hipLaunchKernelGGL(mlp_targets_kernel<false>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream), ncol,
                           nlay, int(ngpt), gpt2band, tlay, col_dry, a, b, nullptr, 2, 0.f, 1.f, Y);
    RRTMGP_LAUNCHED("mlp_targets_kernel");
    return RRTMGP_OK;
}

int rrtmgp_hip_mlp_gather_rows(int64_t rows, int n, const int64_t* idx, const float* src, float* dst, void* stream)
{
    if (rows <= 0 || n <= 0 || !idx || !src || !dst) return arg_error("mlp_gather_rows: bad arguments");
    RRTMGP_TIMED("gather_rows_kernel", as_stream(stream));
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid1(rows * n, 256)), dim3(256), 0, as_stream(stream), rows, n, idx,
                       src, dst);
    RRTMGP_LAUNCHED("gather_rows_kernel");
    return RRTMGP_OK;
}

int rrtmgp_hip_mlp_normalize_rows(int64_t rows, int n, int n_norm, const float* x, const float* mean,
                                  const float* std_, void* y_bf16, float* y_f32, void* stream)
{
    if (rows <= 0 || n <= 0 || n_norm < 0 || n_norm > n || !x || !mean || !std_ || (!y_bf16 && !y_f32))
        return arg_error("mlp_normalize_rows: bad arguments");
    RRTMGP_TIMED("normalize_rows_kernel", as_stream(stream));
    hipLaunchKernelGGL(normalize_rows_kernel, dim3(grid1(rows * n, 256)), dim3(256), 0, as_stream(stream), rows, n,
                       n_norm, x, mean, std_, static_cast<__bf16*>(y_bf16), y_f32);
    RRTMGP_LAUNCHED("normalize_rows_kernel");
    return RRTMGP_OK;
}

}  // extern "C"
