// The warm-start network (include/espice_mlp.h) on gfx950: every layer is
// one BF16 MFMA GEMM (v_mfma_f32_32x32x16_bf16, FP32 accumulation), one
// 32x32 output tile per wavefront, with the bias and tanh fused into the
// store. Training keeps FP32 master weights and Adam moments on the GPU and
// refreshes the BF16 copies (row-major W for the forward pass, its
// transpose for the input gradient) after every step.
//
// Buffers are padded to multiples of 32 in every dimension and the padding
// is kept at exactly zero (zero weights, zero inputs, zero output
// gradients), so no kernel needs a bounds test inside its tile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "mlp.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

namespace {

constexpr uint32_t TILE = 32;

// The hidden layers' activation everywhere the network evaluates it
// (forward, its derivative, the warm start): tanh z = 1 - 2 / (e^{2z} + 1)
// from the hardware exponential and reciprocal. Its error (a few FP32
// ulps, at most ~1e-7 absolute) is far below the BF16 rounding every
// activation then takes; saturates to +-1 for large |z|.
__device__ __forceinline__ float act_tanh(float z)
{
    const float e = __expf(2.0f * z);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}
constexpr uint32_t MAX_GRAD_BATCH = 4096;

uint32_t pad32(uint32_t v) { return (v + TILE - 1) / TILE * TILE; }

template <class T>
struct Dev {
    T *p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count)
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count ? count : 1;
        return hipMalloc(reinterpret_cast<void **>(&p), n * sizeof(T));
    }
    ~Dev()
    {
        if (p) (void)hipFree(p);
    }
};

espice_status fail(espice_status s, const char *name)
{
    esp::set_error(name);
    return s;
}

#define HIP_TRY(expr)                                                                \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) return fail(ESPICE_FAILED, hipGetErrorName(e_)); \
    } while (0)

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

double unit(uint64_t &s) { return ((splitmix64(s) >> 11) + 0.5) * 0x1.0p-53; }

// ------------------------------------------------------------- kernels

enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_TANH = 2 };

// C(m, n) = sum_k A(m, k) B(n, k) over k < K (K a multiple of 16), with
// A(m, k) = A[m * sam + k * sak] and B(n, k) = B[n * sbn + k * sbk]; one
// wavefront per 32x32 tile of C (grid: N/32 x M/32). CONTIG: sak == sbk == 1,
// so each lane's 8 consecutive k are one 16-byte load.
//
// Operand layout of v_mfma_f32_32x32x16_bf16: lane l supplies row l % 32
// of A (column l % 32 of B) at k = 8 (l / 32) + 0..7; accumulator v of lane
// l is C(8 (v / 4) + 4 (l / 32) + v % 4, l % 32).
//
// EPI_BIAS adds bias[n]; EPI_BIAS_TANH also stores bf16(tanh) to H (same
// indexing as C).
template <int EPI, bool CONTIG>
__global__ __launch_bounds__(64) void mlp_gemm_kernel(const bf16 *__restrict__ A, size_t sam, size_t sak,
                                                      const bf16 *__restrict__ B, size_t sbn, size_t sbk, uint32_t K,
                                                      const float *__restrict__ bias, float *__restrict__ C,
                                                      size_t ldc, bf16 *__restrict__ H)
{
    const uint32_t l = threadIdx.x;
    const uint32_t m0 = blockIdx.y * TILE, n0 = blockIdx.x * TILE;
    const uint32_t i = l & 31, kb = (l >> 5) * 8;
    const bf16 *pa = A + (size_t)(m0 + i) * sam;
    const bf16 *pb = B + (size_t)(n0 + i) * sbn;
    floatx16 acc = {};
    for (uint32_t k0 = 0; k0 < K; k0 += 16) {
        bf16x8 a, b;
        if constexpr (CONTIG) {
            a = *reinterpret_cast<const bf16x8 *>(pa + k0 + kb);
            b = *reinterpret_cast<const bf16x8 *>(pb + k0 + kb);
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                a[j] = pa[(size_t)(k0 + kb + j) * sak];
                b[j] = pb[(size_t)(k0 + kb + j) * sbk];
            }
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const uint32_t col = n0 + i;
    const float bv = EPI == EPI_STORE ? 0.0f : bias[col];
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const uint32_t row = m0 + 8 * (v / 4) + 4 * (l >> 5) + (v & 3);
        float z = acc[v];
        if constexpr (EPI != EPI_STORE) z += bv;
        C[(size_t)row * ldc + col] = z;
        if constexpr (EPI == EPI_BIAS_TANH) H[(size_t)row * ldc + col] = (bf16)act_tanh(z);
    }
}

// H0[m][k] = bf16(float((x[m][k] - mean[k]) / std[k])) for m < count and
// k < n_in with std[k] != 0, else 0 (rows to rows_pad, columns to ld).
__global__ void mlp_input_kernel(const double *__restrict__ x, uint32_t count, uint32_t n_in,
                                 const double *__restrict__ mean, const double *__restrict__ sd, bf16 *__restrict__ H0,
                                 uint32_t ld, uint32_t rows_pad)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)rows_pad * ld) return;
    const uint32_t m = (uint32_t)(e / ld), k = (uint32_t)(e % ld);
    float v = 0.0f;
    if (m < count && k < n_in && sd[k] != 0.0) v = (float)((x[(size_t)m * n_in + k] - mean[k]) / sd[k]);
    H0[e] = (bf16)v;
}

// Standardized targets, FP32: T[m][k] = (y - mean) / std, 0 off the data.
__global__ void mlp_target_kernel(const double *__restrict__ y, uint32_t count, uint32_t n_out,
                                  const double *__restrict__ mean, const double *__restrict__ sd, float *__restrict__ T,
                                  uint32_t ld, uint32_t rows_pad)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)rows_pad * ld) return;
    const uint32_t m = (uint32_t)(e / ld), k = (uint32_t)(e % ld);
    float v = 0.0f;
    if (m < count && k < n_out && sd[k] != 0.0) v = (float)((y[(size_t)m * n_out + k] - mean[k]) / sd[k]);
    T[e] = v;
}

// A training batch: rows idx[0..b) of the standardized data set.
__global__ void mlp_gather_kernel(const bf16 *__restrict__ Xs, const float *__restrict__ Ts,
                                  const uint32_t *__restrict__ idx, uint32_t b, bf16 *__restrict__ H0, uint32_t ld_in,
                                  float *__restrict__ T, uint32_t ld_out, uint32_t rows_pad)
{
    const uint32_t m = blockIdx.x;
    if (m >= rows_pad) return;
    const bool live = m < b;
    const size_t src = live ? idx[m] : 0;
    for (uint32_t k = threadIdx.x; k < ld_in; k += blockDim.x)
        H0[(size_t)m * ld_in + k] = live ? Xs[src * ld_in + k] : (bf16)0.0f;
    for (uint32_t k = threadIdx.x; k < ld_out; k += blockDim.x)
        T[(size_t)m * ld_out + k] = live ? Ts[src * ld_out + k] : 0.0f;
}

// Loss and its gradient on the output layer (no activation): over m < b,
// k < n_out, d = Y - T; dZ = 2 d / (b n_out) (FP32 and BF16); loss_sum +=
// sum d^2 (one atomic add per block of 256 threads).
__global__ __launch_bounds__(256) void mlp_loss_kernel(const float *__restrict__ Y, const float *__restrict__ T,
                                                       uint32_t b, uint32_t n_out, uint32_t ld, uint32_t rows_pad,
                                                       float *__restrict__ dZ, bf16 *__restrict__ dZb,
                                                       float *__restrict__ loss_sum)
{
    __shared__ float part[256];
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float sq = 0.0f;
    if (e < (size_t)rows_pad * ld) {
        const uint32_t m = (uint32_t)(e / ld), k = (uint32_t)(e % ld);
        float g = 0.0f;
        if (m < b && k < n_out) {
            const float d = Y[e] - T[e];
            sq = d * d;
            g = 2.0f * d / (float)((size_t)b * n_out);
        }
        dZ[e] = g;
        dZb[e] = (bf16)g;
    }
    part[threadIdx.x] = sq;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss_sum, part[0]);
}

// dZ = dH (1 - tanh(Z)^2), FP32 and BF16.
__global__ void mlp_tanh_back_kernel(const float *__restrict__ dH, const float *__restrict__ Z, size_t count,
                                     float *__restrict__ dZ, bf16 *__restrict__ dZb)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const float t = act_tanh(Z[e]);
    const float g = dH[e] * (1.0f - t * t);
    dZ[e] = g;
    dZb[e] = (bf16)g;
}

// db[n] = sum over the rows of dZ[:, n]: 32 columns per block of 256
// threads, row group g = 0..7 sums rows g, g + 8, ... and the eight partial
// sums are added in group order (grid: ld / 32).
__global__ __launch_bounds__(256) void mlp_colsum_kernel(const float *__restrict__ dZ, uint32_t rows, uint32_t ld,
                                                         float *__restrict__ db)
{
    __shared__ float part[8][32];
    const uint32_t c = threadIdx.x & 31, g = threadIdx.x >> 5;
    const uint32_t n = blockIdx.x * 32 + c;
    float s = 0.0f;
#pragma unroll 4
    for (uint32_t m = g; m < rows; m += 8) s += dZ[(size_t)m * ld + n];
    part[g][c] = s;
    __syncthreads();
    if (g == 0) {
        float t = 0.0f;
        for (int k = 0; k < 8; k++) t += part[k][c];
        db[n] = t;
    }
}

// Adam on count FP32 parameters (bias-corrected step sizes c1, c2 folded
// by the host).
__global__ void mlp_adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ mo,
                                float *__restrict__ ve, size_t count, float lr, float b1, float b2, float eps,
                                float c1, float c2)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const float gi = g[e];
    const float m = b1 * mo[e] + (1.0f - b1) * gi;
    const float v = b2 * ve[e] + (1.0f - b2) * gi * gi;
    mo[e] = m;
    ve[e] = v;
    w[e] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// BF16 copies of a rows x cols FP32 weight matrix: W and its transpose.
__global__ void mlp_cast_kernel(const float *__restrict__ w, uint32_t rows, uint32_t cols, bf16 *__restrict__ wb,
                                bf16 *__restrict__ wt)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)rows * cols) return;
    const uint32_t r = (uint32_t)(e / cols), c = (uint32_t)(e % cols);
    const bf16 v = (bf16)w[e];
    wb[e] = v;
    wt[(size_t)c * rows + r] = v;
}

// The lanes' start point: x[i * N + l] = mean[i] + std[i] * Y[l][i] (std 0:
// exactly mean[i]) for l < count, i < n.
__global__ void mlp_write_x_kernel(const float *__restrict__ Y, uint32_t ld, uint32_t count, uint32_t n,
                                   const double *__restrict__ mean, const double *__restrict__ sd, double *__restrict__ x,
                                   uint32_t N)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)count * n) return;
    const uint32_t l = (uint32_t)(e / n), i = (uint32_t)(e % n);
    const double s = sd[i];
    x[(size_t)i * N + l] = s != 0.0 ? mean[i] + s * (double)Y[(size_t)l * ld + i] : mean[i];
}

// The whole network for the warm start in one launch: workgroup b owns rows
// 32b..32b+31, standardizes them into LDS, runs every layer with its
// activations in LDS (two buffers of 32 rows, ld bf16 apart) and each of
// its four wavefronts taking every fourth 32-column tile, and writes the
// de-standardized outputs straight into the lanes' x. Each tile is the
// same MFMA sequence as mlp_gemm_kernel's (k ascending) with the same
// bias, tanh and BF16 rounding, so the result is bit for bit the
// layer-by-layer path's.
constexpr uint32_t INFER_MAX_LAYERS = 17;
struct MlpNet {
    const bf16 *w[INFER_MAX_LAYERS];
    const float *b[INFER_MAX_LAYERS];
    uint32_t pin[INFER_MAX_LAYERS], pout[INFER_MAX_LAYERS];
    uint32_t layers, n_in, n_out;
    const double *in_mean, *in_sd, *out_mean, *out_sd;
};

constexpr uint32_t INFER_THREADS = 512;
// inputs of up to this many values travel in the kernel's arguments
constexpr uint32_t INFER_INLINE = 128;
struct InlineInputs {
    double v[INFER_INLINE];
};

__global__ __launch_bounds__(INFER_THREADS) void mlp_infer_kernel(MlpNet net, const double *__restrict__ xin,
                                                                  InlineInputs inl, uint32_t count,
                                                                  double *__restrict__ x, uint32_t N, uint32_t ld)
{
    extern __shared__ bf16 act[];
    const uint32_t t = threadIdx.x, wave = t >> 6, l = t & 63, i = l & 31, kb = (l >> 5) * 8;
    const uint32_t m0 = blockIdx.x * TILE;
    const uint32_t p0 = net.pin[0];
    const bool inline_in = xin == nullptr;
    // the output scaling of this lane's column of the last layer's first
    // tile, requested before anything waits on memory
    const uint32_t Lz = net.layers - 1, zcol = wave * TILE + i;
    const bool zlive = zcol < net.n_out;
    const double z_sd = zlive ? net.out_sd[zcol] : 0.0, z_mean = zlive ? net.out_mean[zcol] : 0.0;
    // each wave's first tile of the next layer: its weight fragments and
    // bias are requested while the current layer runs (weights do not
    // depend on the activations; the first layer's while the inputs are
    // standardized), for layers of K <= 256
    constexpr uint32_t PF = 16;
    const uint32_t nw = blockDim.x / 64;
    auto fetch = [&](uint32_t L, bf16x8 *b, float &bias) {
        const uint32_t K = net.pin[L];
        const bool live = wave * TILE < net.pout[L] && K <= PF * 16;
        const bf16 *pb = net.w[L] + (size_t)(wave * TILE + i) * K + kb;
        if (live) bias = net.b[L][wave * TILE + i];
#pragma unroll
        for (uint32_t j = 0; j < PF; j++)
            if (live && 16 * j < K) b[j] = *reinterpret_cast<const bf16x8 *>(pb + 16 * j);
    };
    bf16x8 bn[PF];
    float bias_n = 0.0f;
    fetch(0, bn, bias_n);
    for (uint32_t e = t; e < TILE * p0; e += blockDim.x) {
        const uint32_t r = e / p0, k = e % p0, m = m0 + r;
        float v = 0.0f;
        if (m < count && k < net.n_in) {
            const size_t at = (size_t)m * net.n_in + k;
            const double xv = inline_in ? inl.v[at] : xin[at];
            const double mean = net.in_mean[k], sd = net.in_sd[k];
            if (sd != 0.0) v = (float)((xv - mean) / sd);
        }
        act[r * ld + k] = (bf16)v;
    }
    __syncthreads();
    for (uint32_t L = 0; L < net.layers; L++) {
        const bf16 *in = act + (L & 1) * TILE * ld;
        bf16 *out = act + ((L + 1) & 1) * TILE * ld;
        const uint32_t K = net.pin[L], tiles = net.pout[L] / TILE;
        const bool last = L + 1 == net.layers;
        bf16x8 bc[PF];
#pragma unroll
        for (uint32_t j = 0; j < PF; j++) bc[j] = bn[j];
        const float bias_c = bias_n;
        if (!last) fetch(L + 1, bn, bias_n);
        for (uint32_t tile = wave; tile < tiles; tile += nw) {
            const uint32_t n0 = tile * TILE;
            const bf16 *pa = in + i * ld + kb;
            const bf16 *pb = net.w[L] + (size_t)(n0 + i) * K + kb;
            floatx16 acc = {};
            if (tile == wave && K <= PF * 16) {
                // activations from LDS four k-steps at a time
#pragma unroll
                for (uint32_t j0 = 0; j0 < PF; j0 += 4) {
                    bf16x8 a[4];
#pragma unroll
                    for (uint32_t j = 0; j < 4; j++)
                        if (16 * (j0 + j) < K) a[j] = *reinterpret_cast<const bf16x8 *>(pa + 16 * (j0 + j));
#pragma unroll
                    for (uint32_t j = 0; j < 4; j++)
                        if (16 * (j0 + j) < K)
                            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j], bc[j0 + j], acc, 0, 0, 0);
                }
            } else {
                // four k-steps' fragments in flight before their MFMAs,
                // which then run in k order
                for (uint32_t k0 = 0; k0 < K; k0 += 4 * 16) {
                    bf16x8 a[4], b[4];
#pragma unroll
                    for (uint32_t j = 0; j < 4; j++)
                        if (k0 + 16 * j < K) {
                            a[j] = *reinterpret_cast<const bf16x8 *>(pa + k0 + 16 * j);
                            b[j] = *reinterpret_cast<const bf16x8 *>(pb + k0 + 16 * j);
                        }
#pragma unroll
                    for (uint32_t j = 0; j < 4; j++)
                        if (k0 + 16 * j < K) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j], b[j], acc, 0, 0, 0);
                }
            }
            const uint32_t col = n0 + i;
            const bool first = tile == wave && K <= PF * 16;
            const float bv = first ? bias_c : net.b[L][col];
            const bool zc = first && L == Lz;
            const double sd = zc ? z_sd : (last && col < net.n_out ? net.out_sd[col] : 0.0);
            const double mean = zc ? z_mean : (last && col < net.n_out ? net.out_mean[col] : 0.0);
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const uint32_t row = 8 * (v / 4) + 4 * (l >> 5) + (v & 3);
                const float z = acc[v] + bv;
                if (!last) {
                    out[row * ld + col] = (bf16)act_tanh(z);
                } else {
                    const uint32_t m = m0 + row;
                    if (m < count && col < net.n_out) x[(size_t)col * N + m] = sd != 0.0 ? mean + sd * (double)z : mean;
                }
            }
        }
        __syncthreads();
    }
}

// Y rows (de-standardized) to a row-major host-bound buffer.
__global__ void mlp_unscale_kernel(const float *__restrict__ Y, uint32_t ld, uint32_t count, uint32_t n,
                                   const double *__restrict__ mean, const double *__restrict__ sd, double *__restrict__ y)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)count * n) return;
    const uint32_t l = (uint32_t)(e / n), i = (uint32_t)(e % n);
    const double s = sd[i];
    y[e] = s != 0.0 ? mean[i] + s * (double)Y[(size_t)l * ld + i] : mean[i];
}

uint32_t blocks(size_t count, uint32_t per) { return (uint32_t)((count + per - 1) / per); }

}  // namespace

// ----------------------------------------------------------------- host

// Mapped, page-locked host memory that kernels read in place.
template <typename T>
struct Pinned {
    T *p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count)
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        const hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&p), count * sizeof(T), hipHostMallocMapped);
        if (e == hipSuccess) n = count;
        return e;
    }
    ~Pinned()
    {
        if (p) (void)hipHostFree(p);
    }
};

struct espice_mlp {
    espice_mlp_shape shape{};
    uint32_t layers = 0;
    std::vector<uint32_t> fin, fout;  // true fan-in / fan-out per layer
    std::vector<uint32_t> pin, pout;  // padded
    std::vector<size_t> woff, boff;   // offsets into the packed FP32 parameter vector (padded)
    size_t n_params = 0;              // padded FP32 parameter count
    size_t n_packed = 0;              // unpadded (API) count
    hipStream_t stream = nullptr;
    Dev<float> w, g, mo, ve;          // padded master params, grads, Adam moments
    std::vector<Dev<bf16>> wb, wt;    // per layer BF16 copies
    Dev<double> in_mean, in_sd, out_mean, out_sd;
    // activations for up to `rows` rows
    uint32_t rows = 0;
    Dev<bf16> h0;
    std::vector<Dev<float>> z;     // per layer pre-activation
    std::vector<Dev<bf16>> h;      // per hidden layer post-activation
    Dev<float> dzf, dh, tgt;       // scratch: output-layer sized and widest layer sized
    Dev<bf16> dzb;
    Dev<float> loss;
    Dev<double> xin;               // uploaded inputs (rows x n_in)
    Pinned<double> h_in;           // the warm start's inputs, mapped for the GPU to read
    Dev<double> yout;
    uint64_t steps = 0;
};

namespace {

size_t widest(const espice_mlp &m)
{
    uint32_t w = m.pin[0];
    for (uint32_t l = 0; l < m.layers; l++) w = std::max(w, m.pout[l]);
    return w;
}

espice_status ensure_rows(espice_mlp &m, uint32_t rows)
{
    rows = pad32(std::max(rows, 1u));
    if (rows <= m.rows) return ESPICE_OK;
    const size_t W = widest(m);
    HIP_TRY(m.h0.alloc((size_t)rows * m.pin[0]));
    for (uint32_t l = 0; l < m.layers; l++) {
        HIP_TRY(m.z[l].alloc((size_t)rows * m.pout[l]));
        if (l + 1 < m.layers) HIP_TRY(m.h[l].alloc((size_t)rows * m.pout[l]));
    }
    HIP_TRY(m.dzf.alloc((size_t)rows * W));
    HIP_TRY(m.dzb.alloc((size_t)rows * W));
    HIP_TRY(m.dh.alloc((size_t)rows * W));
    HIP_TRY(m.tgt.alloc((size_t)rows * m.pout[m.layers - 1]));
    HIP_TRY(m.xin.alloc((size_t)rows * std::max(m.shape.n_in, m.shape.n_out)));
    HIP_TRY(m.yout.alloc((size_t)rows * m.shape.n_out));
    m.rows = rows;
    return ESPICE_OK;
}

// BF16 copies of every layer from the master weights.
espice_status refresh_bf16(espice_mlp &m, hipStream_t s)
{
    for (uint32_t l = 0; l < m.layers; l++) {
        const size_t cnt = (size_t)m.pout[l] * m.pin[l];
        mlp_cast_kernel<<<blocks(cnt, 256), 256, 0, s>>>(m.w.p + m.woff[l], m.pout[l], m.pin[l], m.wb[l].p, m.wt[l].p);
    }
    HIP_TRY(hipGetLastError());
    return ESPICE_OK;
}

// Forward over rows_pad rows of h0; the output layer's Z is the prediction.
espice_status forward(espice_mlp &m, uint32_t rows_pad, hipStream_t s)
{
    for (uint32_t l = 0; l < m.layers; l++) {
        const bf16 *in = l == 0 ? m.h0.p : m.h[l - 1].p;
        const dim3 grid(m.pout[l] / TILE, rows_pad / TILE);
        const float *bias = m.w.p + m.boff[l];
        if (l + 1 < m.layers)
            mlp_gemm_kernel<EPI_BIAS_TANH, true><<<grid, 64, 0, s>>>(in, m.pin[l], 1, m.wb[l].p, m.pin[l], 1, m.pin[l],
                                                                     bias, m.z[l].p, m.pout[l], m.h[l].p);
        else
            mlp_gemm_kernel<EPI_BIAS, true><<<grid, 64, 0, s>>>(in, m.pin[l], 1, m.wb[l].p, m.pin[l], 1, m.pin[l], bias,
                                                                m.z[l].p, m.pout[l], nullptr);
    }
    HIP_TRY(hipGetLastError());
    return ESPICE_OK;
}

// Loss of the standardized targets in tgt over b rows, then the gradient of
// every parameter into g (padded layout).
espice_status backward(espice_mlp &m, uint32_t b, uint32_t rows_pad, hipStream_t s)
{
    const uint32_t L = m.layers;
    HIP_TRY(hipMemsetAsync(m.loss.p, 0, sizeof(float), s));
    {
        const size_t cnt = (size_t)rows_pad * m.pout[L - 1];
        mlp_loss_kernel<<<blocks(cnt, 256), 256, 0, s>>>(m.z[L - 1].p, m.tgt.p, b, m.shape.n_out, m.pout[L - 1],
                                                         rows_pad, m.dzf.p, m.dzb.p, m.loss.p);
    }
    for (uint32_t l = L; l-- > 0;) {
        const uint32_t po = m.pout[l], pi = m.pin[l];
        const bf16 *in = l == 0 ? m.h0.p : m.h[l - 1].p;
        // dW(n, k) = sum over rows r of dZ(r, n) in(r, k)
        mlp_gemm_kernel<EPI_STORE, false><<<dim3(pi / TILE, po / TILE), 64, 0, s>>>(
            m.dzb.p, 1, po, in, 1, pi, rows_pad, nullptr, m.g.p + m.woff[l], pi, nullptr);
        mlp_colsum_kernel<<<po / TILE, 256, 0, s>>>(m.dzf.p, rows_pad, po, m.g.p + m.boff[l]);
        if (l == 0) break;
        // dH(r, k) = sum_n dZ(r, n) W(n, k) = sum_n dZ(r, n) Wt(k, n)
        mlp_gemm_kernel<EPI_STORE, true><<<dim3(pi / TILE, rows_pad / TILE), 64, 0, s>>>(
            m.dzb.p, po, 1, m.wt[l].p, po, 1, po, nullptr, m.dh.p, pi, nullptr);
        const size_t cnt = (size_t)rows_pad * pi;
        mlp_tanh_back_kernel<<<blocks(cnt, 256), 256, 0, s>>>(m.dh.p, m.z[l - 1].p, cnt, m.dzf.p, m.dzb.p);
    }
    HIP_TRY(hipGetLastError());
    return ESPICE_OK;
}

espice_status upload_inputs(espice_mlp &m, const double *x, uint32_t count, hipStream_t s)
{
    HIP_TRY(hipMemcpyAsync(m.xin.p, x, sizeof(double) * count * m.shape.n_in, hipMemcpyHostToDevice, s));
    const uint32_t rp = pad32(count);
    const size_t cnt = (size_t)rp * m.pin[0];
    mlp_input_kernel<<<blocks(cnt, 256), 256, 0, s>>>(m.xin.p, count, m.shape.n_in, m.in_mean.p, m.in_sd.p, m.h0.p,
                                                      m.pin[0], rp);
    HIP_TRY(hipGetLastError());
    return ESPICE_OK;
}

// Packed (unpadded, API order) <-> padded layout.
void pack_params(const espice_mlp &m, const std::vector<float> &padded, float *out)
{
    size_t at = 0;
    for (uint32_t l = 0; l < m.layers; l++) {
        for (uint32_t r = 0; r < m.fout[l]; r++)
            for (uint32_t c = 0; c < m.fin[l]; c++) out[at++] = padded[m.woff[l] + (size_t)r * m.pin[l] + c];
        for (uint32_t r = 0; r < m.fout[l]; r++) out[at++] = padded[m.boff[l] + r];
    }
}

void unpack_params(const espice_mlp &m, const float *in, std::vector<float> &padded)
{
    padded.assign(m.n_params, 0.0f);
    size_t at = 0;
    for (uint32_t l = 0; l < m.layers; l++) {
        for (uint32_t r = 0; r < m.fout[l]; r++)
            for (uint32_t c = 0; c < m.fin[l]; c++) padded[m.woff[l] + (size_t)r * m.pin[l] + c] = in[at++];
        for (uint32_t r = 0; r < m.fout[l]; r++) padded[m.boff[l] + r] = in[at++];
    }
}

}  // namespace

extern "C" {

espice_status espice_mlp_create(const espice_mlp_shape *shape, espice_mlp **out)
{
    if (!shape || !out) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    *out = nullptr;
    const espice_mlp_shape &S = *shape;
    if (S.n_in == 0 || S.n_out == 0 || S.hidden == 0 || S.n_in > 4096 || S.n_out > 4096 || S.hidden > 4096 ||
        S.depth > 16)
        return fail(ESPICE_INVALID_ARGUMENT, "InvalidShape");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) return fail(ESPICE_FAILED, "NoGpu");
    std::unique_ptr<espice_mlp> m(new (std::nothrow) espice_mlp());
    if (!m) return fail(ESPICE_OUT_OF_MEMORY, "OutOfMemory");
    m->shape = S;
    m->layers = S.depth + 1;
    for (uint32_t l = 0; l < m->layers; l++) {
        const uint32_t fi = l == 0 ? S.n_in : S.hidden, fo = l + 1 == m->layers ? S.n_out : S.hidden;
        m->fin.push_back(fi);
        m->fout.push_back(fo);
        m->pin.push_back(pad32(fi));
        m->pout.push_back(pad32(fo));
    }
    size_t at = 0;
    for (uint32_t l = 0; l < m->layers; l++) {
        m->woff.push_back(at);
        at += (size_t)m->pout[l] * m->pin[l];
        m->boff.push_back(at);
        at += m->pout[l];
        m->n_packed += (size_t)m->fout[l] * m->fin[l] + m->fout[l];
    }
    m->n_params = at;
    HIP_TRY(hipStreamCreate(&m->stream));
    HIP_TRY(m->w.alloc(at));
    HIP_TRY(m->g.alloc(at));
    HIP_TRY(m->mo.alloc(at));
    HIP_TRY(m->ve.alloc(at));
    HIP_TRY(hipMemset(m->g.p, 0, at * sizeof(float)));
    HIP_TRY(hipMemset(m->mo.p, 0, at * sizeof(float)));
    HIP_TRY(hipMemset(m->ve.p, 0, at * sizeof(float)));
    m->wb.resize(m->layers);
    m->wt.resize(m->layers);
    m->z.resize(m->layers);
    m->h.resize(m->layers);
    for (uint32_t l = 0; l < m->layers; l++) {
        HIP_TRY(m->wb[l].alloc((size_t)m->pout[l] * m->pin[l]));
        HIP_TRY(m->wt[l].alloc((size_t)m->pout[l] * m->pin[l]));
    }
    HIP_TRY(m->in_mean.alloc(S.n_in));
    HIP_TRY(m->in_sd.alloc(S.n_in));
    HIP_TRY(m->out_mean.alloc(S.n_out));
    HIP_TRY(m->out_sd.alloc(S.n_out));
    HIP_TRY(m->loss.alloc(1));
    // identity scaling until the caller sets one
    std::vector<double> zero(std::max(S.n_in, S.n_out), 0.0), one(std::max(S.n_in, S.n_out), 1.0);
    HIP_TRY(hipMemcpy(m->in_mean.p, zero.data(), 8 * S.n_in, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->in_sd.p, one.data(), 8 * S.n_in, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->out_mean.p, zero.data(), 8 * S.n_out, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->out_sd.p, one.data(), 8 * S.n_out, hipMemcpyHostToDevice));
    // Xavier-uniform weights, zero biases
    std::vector<float> host(at, 0.0f);
    uint64_t st = S.seed;
    for (uint32_t l = 0; l < m->layers; l++) {
        const double a = std::sqrt(6.0 / (double)(m->fin[l] + m->fout[l]));
        for (uint32_t r = 0; r < m->fout[l]; r++)
            for (uint32_t c = 0; c < m->fin[l]; c++)
                host[m->woff[l] + (size_t)r * m->pin[l] + c] = (float)((2.0 * unit(st) - 1.0) * a);
    }
    HIP_TRY(hipMemcpy(m->w.p, host.data(), at * sizeof(float), hipMemcpyHostToDevice));
    if (espice_status e = refresh_bf16(*m, m->stream)) return e;
    HIP_TRY(hipStreamSynchronize(m->stream));
    *out = m.release();
    return ESPICE_OK;
}

espice_status espice_mlp_destroy(espice_mlp *m)
{
    if (m) {
        if (m->stream) (void)hipStreamSynchronize(m->stream);
        hipStream_t s = m->stream;
        delete m;
        if (s) (void)hipStreamDestroy(s);
    }
    return ESPICE_OK;
}

espice_status espice_mlp_param_count(const espice_mlp *m, size_t *count)
{
    if (!m || !count) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    *count = m->n_packed;
    return ESPICE_OK;
}

espice_status espice_mlp_set_scaling(espice_mlp *m, const double *in_mean, const double *in_std, const double *out_mean,
                                     const double *out_std)
{
    if (!m || !in_mean || !in_std || !out_mean || !out_std) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    for (uint32_t k = 0; k < m->shape.n_in; k++)
        if (!std::isfinite(in_mean[k]) || !std::isfinite(in_std[k]) || in_std[k] < 0)
            return fail(ESPICE_INVALID_ARGUMENT, "InvalidScaling");
    for (uint32_t k = 0; k < m->shape.n_out; k++)
        if (!std::isfinite(out_mean[k]) || !std::isfinite(out_std[k]) || out_std[k] < 0)
            return fail(ESPICE_INVALID_ARGUMENT, "InvalidScaling");
    HIP_TRY(hipMemcpy(m->in_mean.p, in_mean, 8 * m->shape.n_in, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->in_sd.p, in_std, 8 * m->shape.n_in, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->out_mean.p, out_mean, 8 * m->shape.n_out, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->out_sd.p, out_std, 8 * m->shape.n_out, hipMemcpyHostToDevice));
    return ESPICE_OK;
}

espice_status espice_mlp_get_params(espice_mlp *m, float *params)
{
    if (!m || !params) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    std::vector<float> host(m->n_params);
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(host.data(), m->w.p, m->n_params * sizeof(float), hipMemcpyDeviceToHost));
    pack_params(*m, host, params);
    return ESPICE_OK;
}

espice_status espice_mlp_set_params(espice_mlp *m, const float *params)
{
    if (!m || !params) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    std::vector<float> host;
    unpack_params(*m, params, host);
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(m->w.p, host.data(), m->n_params * sizeof(float), hipMemcpyHostToDevice));
    if (espice_status e = refresh_bf16(*m, m->stream)) return e;
    HIP_TRY(hipStreamSynchronize(m->stream));
    return ESPICE_OK;
}

espice_status espice_mlp_train(espice_mlp *m, const double *x, const double *y, uint32_t count,
                               const espice_mlp_train_opts *opts, float *loss)
{
    if (!m || !x || !y || !opts) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    const espice_mlp_train_opts &o = *opts;
    if (count == 0 || o.batch == 0 || o.batch > MAX_GRAD_BATCH || o.epochs == 0 || !(o.lr > 0))
        return fail(ESPICE_INVALID_ARGUMENT, "InvalidTraining");
    const uint32_t b = std::min(o.batch, count);
    const uint32_t rp = pad32(b);
    if (espice_status e = ensure_rows(*m, rp)) return e;
    hipStream_t s = m->stream;
    // the standardized data set, resident for the whole run
    const uint32_t cp = pad32(count), li = m->pin[0], lo = m->pout[m->layers - 1];
    Dev<double> xd, yd;
    Dev<bf16> xs;
    Dev<float> ts;
    Dev<uint32_t> idx;
    HIP_TRY(xd.alloc((size_t)count * m->shape.n_in));
    HIP_TRY(yd.alloc((size_t)count * m->shape.n_out));
    HIP_TRY(xs.alloc((size_t)cp * li));
    HIP_TRY(ts.alloc((size_t)cp * lo));
    HIP_TRY(idx.alloc(count));
    HIP_TRY(hipMemcpyAsync(xd.p, x, 8 * (size_t)count * m->shape.n_in, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(yd.p, y, 8 * (size_t)count * m->shape.n_out, hipMemcpyHostToDevice, s));
    mlp_input_kernel<<<blocks((size_t)cp * li, 256), 256, 0, s>>>(xd.p, count, m->shape.n_in, m->in_mean.p,
                                                                  m->in_sd.p, xs.p, li, cp);
    mlp_target_kernel<<<blocks((size_t)cp * lo, 256), 256, 0, s>>>(yd.p, count, m->shape.n_out, m->out_mean.p,
                                                                   m->out_sd.p, ts.p, lo, cp);
    HIP_TRY(hipGetLastError());
    std::vector<uint32_t> perm(count);
    uint64_t st = o.seed;
    const uint32_t per_epoch = (count + b - 1) / b;
    std::vector<float> losses((size_t)o.epochs * per_epoch);
    for (uint32_t ep = 0; ep < o.epochs; ep++) {
        for (uint32_t i = 0; i < count; i++) perm[i] = i;
        for (uint32_t i = count; i > 1; i--) std::swap(perm[i - 1], perm[splitmix64(st) % i]);
        HIP_TRY(hipMemcpyAsync(idx.p, perm.data(), 4 * (size_t)count, hipMemcpyHostToDevice, s));
        for (uint32_t bi = 0; bi < per_epoch; bi++) {
            const uint32_t first = bi * b, nb = std::min(b, count - first);
            const uint32_t nrp = pad32(nb);
            mlp_gather_kernel<<<nrp, 64, 0, s>>>(xs.p, ts.p, idx.p + first, nb, m->h0.p, li, m->tgt.p, lo, nrp);
            HIP_TRY(hipGetLastError());
            if (espice_status e = forward(*m, nrp, s)) return e;
            if (espice_status e = backward(*m, nb, nrp, s)) return e;
            m->steps++;
            const float c1 = 1.0f - std::pow(o.beta1, (float)m->steps), c2 = 1.0f - std::pow(o.beta2, (float)m->steps);
            mlp_adam_kernel<<<blocks(m->n_params, 256), 256, 0, s>>>(m->w.p, m->g.p, m->mo.p, m->ve.p, m->n_params,
                                                                     o.lr, o.beta1, o.beta2, o.eps, c1, c2);
            HIP_TRY(hipGetLastError());
            if (espice_status e = refresh_bf16(*m, s)) return e;
            HIP_TRY(hipMemcpyAsync(&losses[(size_t)ep * per_epoch + bi], m->loss.p, sizeof(float),
                                   hipMemcpyDeviceToHost, s));
        }
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (loss)
        for (uint32_t ep = 0; ep < o.epochs; ep++) {
            double acc = 0;
            for (uint32_t bi = 0; bi < per_epoch; bi++) {
                const uint32_t nb = std::min(b, count - bi * b);
                acc += losses[(size_t)ep * per_epoch + bi] / ((double)nb * m->shape.n_out);
            }
            loss[ep] = (float)(acc / per_epoch);
        }
    return ESPICE_OK;
}

espice_status espice_mlp_gradients(espice_mlp *m, const double *x, const double *y, uint32_t count, float *grads,
                                   float *loss)
{
    if (!m || !x || !y || !grads) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (count == 0 || count > MAX_GRAD_BATCH) return fail(ESPICE_INVALID_ARGUMENT, "InvalidCount");
    if (espice_status e = ensure_rows(*m, count)) return e;
    hipStream_t s = m->stream;
    const uint32_t rp = pad32(count), lo = m->pout[m->layers - 1];
    if (espice_status e = upload_inputs(*m, x, count, s)) return e;
    HIP_TRY(hipMemcpyAsync(m->yout.p, y, 8 * (size_t)count * m->shape.n_out, hipMemcpyHostToDevice, s));
    mlp_target_kernel<<<blocks((size_t)rp * lo, 256), 256, 0, s>>>(m->yout.p, count, m->shape.n_out, m->out_mean.p,
                                                                   m->out_sd.p, m->tgt.p, lo, rp);
    HIP_TRY(hipGetLastError());
    if (espice_status e = forward(*m, rp, s)) return e;
    if (espice_status e = backward(*m, count, rp, s)) return e;
    std::vector<float> host(m->n_params);
    float lsum = 0;
    HIP_TRY(hipMemcpyAsync(host.data(), m->g.p, m->n_params * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&lsum, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    pack_params(*m, host, grads);
    if (loss) *loss = lsum / ((float)count * m->shape.n_out);
    return ESPICE_OK;
}

espice_status espice_mlp_predict(espice_mlp *m, const double *x, uint32_t count, double *y)
{
    if (!m || !x || !y) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (count == 0) return ESPICE_OK;
    if (espice_status e = ensure_rows(*m, count)) return e;
    hipStream_t s = m->stream;
    const uint32_t rp = pad32(count), n = m->shape.n_out;
    if (espice_status e = upload_inputs(*m, x, count, s)) return e;
    if (espice_status e = forward(*m, rp, s)) return e;
    mlp_unscale_kernel<<<blocks((size_t)count * n, 256), 256, 0, s>>>(m->z[m->layers - 1].p, m->pout[m->layers - 1],
                                                                      count, n, m->out_mean.p, m->out_sd.p, m->yout.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(y, m->yout.p, 8 * (size_t)count * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ESPICE_OK;
}

}  // extern "C"

espice_status esp::mlp_warm_start(espice_mlp *m, const double *params, uint32_t count, double *x, uint32_t n,
                                  uint32_t N, hipStream_t s)
{
    if (m->shape.n_out != n) return fail(ESPICE_INVALID_ARGUMENT, "ShapeMismatch");
    if (count > N) return fail(ESPICE_INVALID_ARGUMENT, "InvalidCount");
    if (espice_status e = ensure_rows(*m, count)) return e;
    // the network's own stream may still be training or predicting
    HIP_TRY(hipStreamSynchronize(m->stream));
    const uint32_t rp = pad32(count);
    uint32_t widest_in = 0;
    for (uint32_t l = 0; l < m->layers; l++) widest_in = std::max(widest_in, std::max(m->pin[l], m->pout[l]));
    const uint32_t ld = widest_in + 8;  // 16 bytes of padding between rows spreads them over the LDS banks
    const size_t lds = 2 * (size_t)TILE * ld * sizeof(bf16);
    if (m->layers <= INFER_MAX_LAYERS && lds <= 64 * 1024) {
        MlpNet net{};
        for (uint32_t l = 0; l < m->layers; l++) {
            net.w[l] = m->wb[l].p;
            net.b[l] = m->w.p + m->boff[l];
            net.pin[l] = m->pin[l];
            net.pout[l] = m->pout[l];
        }
        net.layers = m->layers;
        net.n_in = m->shape.n_in;
        net.n_out = m->shape.n_out;
        net.in_mean = m->in_mean.p;
        net.in_sd = m->in_sd.p;
        net.out_mean = m->out_mean.p;
        net.out_sd = m->out_sd.p;
        // the inputs in the kernel's arguments when they are few, else
        // staged in mapped pinned memory, which the kernel reads in place
        // (no copy on the stream); every caller synchronizes before
        // returning, so the next call may reuse the staging
        const size_t in_n = (size_t)count * m->shape.n_in;
        InlineInputs inl{};
        double *d_in = nullptr;
        if (in_n <= INFER_INLINE) {
            std::memcpy(inl.v, params, sizeof(double) * in_n);
        } else {
            if (m->h_in.n < in_n) HIP_TRY(m->h_in.alloc(in_n));
            std::memcpy(m->h_in.p, params, sizeof(double) * in_n);
            HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void **>(&d_in), m->h_in.p, 0));
        }
        mlp_infer_kernel<<<rp / TILE, INFER_THREADS, lds, s>>>(net, d_in, inl, count, x, N, ld);
        HIP_TRY(hipGetLastError());
        return ESPICE_OK;
    }
    if (espice_status e = upload_inputs(*m, params, count, s)) return e;
    if (espice_status e = forward(*m, rp, s)) return e;
    mlp_write_x_kernel<<<blocks((size_t)count * n, 256), 256, 0, s>>>(m->z[m->layers - 1].p, m->pout[m->layers - 1],
                                                                      count, n, m->out_mean.p, m->out_sd.p, x, N);
    HIP_TRY(hipGetLastError());
    return ESPICE_OK;
}
