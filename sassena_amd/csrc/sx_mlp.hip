// sx_mlp.hip -- the small network of the learned control variate, trained and
// evaluated by hand-written gfx950 kernels.
//
// y = W3 . tanh(W2 tanh(W1 x + b1) + b2) + b3 with 16 inputs (zero padded),
// two hidden layers of 128 and one output.  W1 and W2 are BF16 operands of
// v_mfma_f32_16x16x16_bf16 (FP32 accumulation); the output layer, biases,
// activations and optimizer state are FP32.  One wavefront owns 16 rows:
//   forward   x[16x16] W1 -> h1[16x128]; h1 W2 -> h2[16x128]; h2 . W3 -> y
//   backward  dW2 = h1^T dz2 and dW1 = x^T dz1 (MFMA, K = the 16 rows),
//             dh1 = dz2 W2^T (MFMA), reductions over rows in LDS.
// Operand layouts of the 16x16x16 MFMA (lane l, register j < 4):
//   A[m][k]: m = l % 16, k = 4 (l / 16) + j;  B[k][n]: k = 4 (l / 16) + j,
//   n = l % 16;  D[m][n]: m = 4 (l / 16) + j, n = l % 16.
// A wavefront's D tile of h1 is therefore the A operand of h1^T and the B
// operand of dz2 is its D tile, so the weight gradients need no shuffles.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/sassena_amd.h"

namespace sxm {

constexpr int IN = 16, H = 128;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + IN * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H, NPARAM = OFF_B3 + 1;
constexpr int ROWS_PER_BLOCK = 64;  // 4 wavefronts x 16 rows

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f) {
    // round to nearest even (finite inputs)
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float from_bf16(short s) { return __uint_as_float(((unsigned)(unsigned short)s) << 16); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Row index of batch row r of step `step`: a hashed draw from [0, n).
__device__ __forceinline__ unsigned pick_row(unsigned long long seed, unsigned step, unsigned r, unsigned n) {
    unsigned long long z = seed ^ ((unsigned long long)step << 32) ^ r;
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (unsigned)(z % n);
}

struct StepArgs {
    const float *x;        // [n][IN] inputs
    const float *t;        // [n] targets (nullptr: forward only)
    unsigned n;            // rows available
    unsigned rows;         // rows in this batch
    int sample;            // 1: draw rows with pick_row, 0: rows 0..rows-1
    unsigned long long seed;
    unsigned step;
    const short *w1b;      // BF16 copies [IN][H], [H][H]
    const short *w2b;
    const float *p;        // FP32 parameters (b1, b2, W3, b3 used from here)
    float *y;              // [rows] outputs (optional)
    float *partial;        // [blocks][NPARAM + 1] gradient partials and loss (training)
};

// Forward (and backward when a.t != nullptr) of ROWS_PER_BLOCK rows.
__global__ __launch_bounds__(256) void mlp_step_kernel(StepArgs a) {
    __shared__ short w2s[H * H];            // 32 KB, W2 as BF16 [k][n]
    __shared__ short w1s[IN * H];           // W1 as BF16 [k][n]
    __shared__ float gW2[H * H];            // 64 KB block accumulator of dW2
    __shared__ float gW1[IN * H];
    __shared__ float gsmall[3 * H + 2];     // db1, db2, dW3, db3, loss
    __shared__ short scratch[4][16 * H];    // per-wave re-layout tile (BF16)
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool train = a.t != nullptr;
    for (int i = tid; i < H * H; i += 256) w2s[i] = a.w2b[i];
    for (int i = tid; i < IN * H; i += 256) w1s[i] = a.w1b[i];
    if (train) {
        for (int i = tid; i < H * H; i += 256) gW2[i] = 0.f;
        for (int i = tid; i < IN * H; i += 256) gW1[i] = 0.f;
        for (int i = tid; i < 3 * H + 2; i += 256) gsmall[i] = 0.f;
    }
    __syncthreads();

    const int lr = lane & 15, lg = lane >> 4;  // MFMA lane row/col and k group
    const unsigned row0 = blockIdx.x * ROWS_PER_BLOCK + wave * 16;
    // the input row this lane loads for the A operand (row lr) and the D-layout rows 4*lg + j
    auto src_row = [&](unsigned r) -> int {
        const unsigned br = row0 + r;
        if (br >= a.rows) return -1;
        return (int)(a.sample ? pick_row(a.seed, a.step, br, a.n) : br);
    };
    const int ra = src_row(lr);
    bf16x4 xa;
    for (int j = 0; j < 4; ++j) xa[j] = ra >= 0 ? to_bf16(a.x[(size_t)ra * IN + 4 * lg + j]) : (short)0;

    // ---- layer 1: z1 = x W1 + b1, h1 = tanh(z1); D layout: rows 4*lg + j, col 16*nt + lr
    float h1[8][4];
    for (int nt = 0; nt < 8; ++nt) {
        bf16x4 bw;
        for (int j = 0; j < 4; ++j) bw[j] = w1s[(4 * lg + j) * H + 16 * nt + lr];
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(xa, bw, acc);
        const float bias = a.p[OFF_B1 + 16 * nt + lr];
        for (int j = 0; j < 4; ++j) h1[nt][j] = tanhf(acc[j] + bias);
    }
    // re-layout h1 (BF16) through LDS into the A operand of layer 2
    short *sc = scratch[wave];
    for (int nt = 0; nt < 8; ++nt)
        for (int j = 0; j < 4; ++j) sc[(4 * lg + j) * H + 16 * nt + lr] = to_bf16(h1[nt][j]);
    __syncthreads();
    // ---- layer 2
    float h2[8][4];
    for (int nt = 0; nt < 8; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < 8; ++kt) {
            bf16x4 av, bw;
            for (int j = 0; j < 4; ++j) {
                av[j] = sc[lr * H + 16 * kt + 4 * lg + j];
                bw[j] = w2s[(16 * kt + 4 * lg + j) * H + 16 * nt + lr];
            }
            acc = mfma(av, bw, acc);
        }
        const float bias = a.p[OFF_B2 + 16 * nt + lr];
        for (int j = 0; j < 4; ++j) h2[nt][j] = tanhf(acc[j] + bias);
    }
    // ---- layer 3: y[r] = h2[r] . W3 + b3; reduce over the 16 lanes of a k group
    float yr[4];
    for (int j = 0; j < 4; ++j) {
        float s = 0.f;
        for (int nt = 0; nt < 8; ++nt) s += h2[nt][j] * a.p[OFF_W3 + 16 * nt + lr];
        for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        yr[j] = s + a.p[OFF_B3];
    }
    int rd[4];
    for (int j = 0; j < 4; ++j) rd[j] = src_row(4 * lg + j);
    if (a.y && lr == 0)
        for (int j = 0; j < 4; ++j)
            if (row0 + 4 * lg + j < a.rows) a.y[row0 + 4 * lg + j] = yr[j];
    if (!train) return;

    // ---- loss and dy = 2 (y - t) / rows (rows beyond the batch contribute 0)
    float dy[4];
    float lsum = 0.f;
    for (int j = 0; j < 4; ++j) {
        if (rd[j] >= 0) {
            const float e = yr[j] - a.t[rd[j]];
            dy[j] = 2.f * e / (float)a.rows;
            lsum += e * e;
        } else {
            dy[j] = 0.f;
        }
    }
    // ---- dz2 = dy W3 (1 - h2^2); dW3 += dy h2; db2 += dz2
    float dz2[8][4];
    for (int nt = 0; nt < 8; ++nt) {
        const float w3 = a.p[OFF_W3 + 16 * nt + lr];
        float gw3 = 0.f, gb2 = 0.f;
        for (int j = 0; j < 4; ++j) {
            dz2[nt][j] = dy[j] * w3 * (1.f - h2[nt][j] * h2[nt][j]);
            gw3 += dy[j] * h2[nt][j];
            gb2 += dz2[nt][j];
        }
        atomicAdd(&gsmall[2 * H + 16 * nt + lr], gw3);
        atomicAdd(&gsmall[H + 16 * nt + lr], gb2);
    }
    if (lr == 0) {
        atomicAdd(&gsmall[3 * H], dy[0] + dy[1] + dy[2] + dy[3]);
        atomicAdd(&gsmall[3 * H + 1], lsum);
    }
    // ---- dW2[k][c] += sum_r h1[r][k] dz2[r][c]: A = h1^T (D tile of h1), B = dz2 (D tile)
    bf16x4 h1b[8], dzb[8];
    for (int t8 = 0; t8 < 8; ++t8)
        for (int j = 0; j < 4; ++j) {
            h1b[t8][j] = to_bf16(h1[t8][j]);
            dzb[t8][j] = to_bf16(dz2[t8][j]);
        }
    for (int mt = 0; mt < 8; ++mt)
        for (int ct = 0; ct < 8; ++ct) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(h1b[mt], dzb[ct], acc);
            for (int j = 0; j < 4; ++j) atomicAdd(&gW2[(16 * mt + 4 * lg + j) * H + 16 * ct + lr], acc[j]);
        }
    // ---- dh1 = dz2 W2^T: A = dz2 (re-laid through LDS), B[c][k] = W2[k][c]
    __syncthreads();  // every wave is done reading its h1 scratch tile
    for (int nt = 0; nt < 8; ++nt)
        for (int j = 0; j < 4; ++j) sc[(4 * lg + j) * H + 16 * nt + lr] = dzb[nt][j];
    __syncthreads();
    float dz1[8][4];
    for (int nt = 0; nt < 8; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < 8; ++kt) {
            bf16x4 av, bw;
            for (int j = 0; j < 4; ++j) {
                av[j] = sc[lr * H + 16 * kt + 4 * lg + j];
                bw[j] = w2s[(16 * nt + lr) * H + 16 * kt + 4 * lg + j];
            }
            acc = mfma(av, bw, acc);
        }
        float gb1 = 0.f;
        for (int j = 0; j < 4; ++j) {
            dz1[nt][j] = acc[j] * (1.f - h1[nt][j] * h1[nt][j]);
            gb1 += dz1[nt][j];
        }
        atomicAdd(&gsmall[16 * nt + lr], gb1);
    }
    // ---- dW1[k][c] += sum_r x[r][k] dz1[r][c]: A = x^T (lane rows 4 lg + j, k = lr)
    bf16x4 xt;
    for (int j = 0; j < 4; ++j) xt[j] = rd[j] >= 0 ? to_bf16(a.x[(size_t)rd[j] * IN + lr]) : (short)0;
    for (int ct = 0; ct < 8; ++ct) {
        bf16x4 bz;
        for (int j = 0; j < 4; ++j) bz[j] = to_bf16(dz1[ct][j]);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(xt, bz, acc);
        for (int j = 0; j < 4; ++j) atomicAdd(&gW1[(4 * lg + j) * H + 16 * ct + lr], acc[j]);
    }
    __syncthreads();
    float *dst = a.partial + (size_t)blockIdx.x * (NPARAM + 1);
    for (int i = tid; i < IN * H; i += 256) dst[OFF_W1 + i] = gW1[i];
    for (int i = tid; i < H * H; i += 256) dst[OFF_W2 + i] = gW2[i];
    for (int i = tid; i < H; i += 256) {
        dst[OFF_B1 + i] = gsmall[i];
        dst[OFF_B2 + i] = gsmall[H + i];
        dst[OFF_W3 + i] = gsmall[2 * H + i];
    }
    if (tid == 0) {
        dst[OFF_B3] = gsmall[3 * H];
        dst[NPARAM] = gsmall[3 * H + 1];
    }
}

// Sums the block partials (fixed order) into grad, and optionally applies
// one Adam step to the FP32 parameters and refreshes the BF16 copies.
__global__ __launch_bounds__(256) void mlp_reduce_adam_kernel(const float *__restrict__ partial, int blocks,
                                                              float *__restrict__ grad, float *__restrict__ loss,
                                                              float *__restrict__ p, float *__restrict__ m,
                                                              float *__restrict__ v, short *__restrict__ w1b,
                                                              short *__restrict__ w2b, int apply, float lr,
                                                              float b1c, float b2c) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i > NPARAM) return;
    float g = 0.f;
    for (int b = 0; b < blocks; ++b) g += partial[(size_t)b * (NPARAM + 1) + i];
    if (i == NPARAM) {
        if (loss) *loss = g;
        return;
    }
    grad[i] = g;
    if (!apply) return;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float mi = beta1 * m[i] + (1.f - beta1) * g;
    const float vi = beta2 * v[i] + (1.f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / b1c) / (sqrtf(vi / b2c) + eps);
    p[i] = pi;
    if (i >= OFF_W1 && i < OFF_B1) w1b[i - OFF_W1] = to_bf16(pi);
    if (i >= OFF_W2 && i < OFF_B2) w2b[i - OFF_W2] = to_bf16(pi);
}

__global__ void mlp_refresh_bf16_kernel(const float *__restrict__ p, short *__restrict__ w1b, short *__restrict__ w2b) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < IN * H) w1b[i] = to_bf16(p[OFF_W1 + i]);
    if (i < H * H) w2b[i] = to_bf16(p[OFF_W2 + i]);
}

}  // namespace sxm

struct sx_mlp {
    int device = 0;
    std::string err;
    hipStream_t stream = nullptr;
    float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *partial = nullptr, *loss = nullptr;
    short *w1b = nullptr, *w2b = nullptr;
    float *x = nullptr, *t = nullptr, *y = nullptr;
    size_t xcap = 0, pcap_blocks = 0;
    unsigned long long adam_t = 0;
};

namespace {

int mfail(sx_mlp *n, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (n) n->err = buf;
    return code;
}

#define SXM_HIP(n, expr)                                                                                    \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            return mfail((n), e_ == hipErrorOutOfMemory ? SX_ERR_NOMEM : SX_ERR_HIP, "%s failed: %s (%s:%d)", \
                         #expr, hipGetErrorString(e_), __FILE__, __LINE__);                                 \
    } while (0)

int ensure_rows(sx_mlp *n, size_t rows) {
    if (rows <= n->xcap) return SX_OK;
    for (float **q : {&n->x, &n->t, &n->y})
        if (*q) {
            (void)hipFree(*q);
            *q = nullptr;
        }
    SXM_HIP(n, hipMalloc(&n->x, sizeof(float) * sxm::IN * rows));
    SXM_HIP(n, hipMalloc(&n->t, sizeof(float) * rows));
    SXM_HIP(n, hipMalloc(&n->y, sizeof(float) * rows));
    n->xcap = rows;
    return SX_OK;
}

int ensure_partials(sx_mlp *n, size_t blocks) {
    if (blocks <= n->pcap_blocks) return SX_OK;
    if (n->partial) (void)hipFree(n->partial);
    n->partial = nullptr;
    SXM_HIP(n, hipMalloc(&n->partial, sizeof(float) * (sxm::NPARAM + 1) * blocks));
    n->pcap_blocks = blocks;
    return SX_OK;
}

int check_rows(sx_mlp *n, const float *x, size_t rows) {
    if (!x || rows == 0 || rows > (size_t)1 << 28) return mfail(n, SX_ERR_ARG, "bad row count %zu", rows);
    for (size_t i = 0; i < rows * sxm::IN; ++i)
        if (!std::isfinite(x[i])) return mfail(n, SX_ERR_ARG, "non-finite input at %zu", i);
    return SX_OK;
}

// One pass over `rows` rows (sampled from n when sample != 0): forward,
// and backward + reduction (+ Adam when apply) when t is given.
int step(sx_mlp *n, unsigned nrows_avail, unsigned rows, int sample, unsigned long long seed, unsigned stepno,
         bool train, int apply, float lr) {
    const unsigned blocks = (rows + sxm::ROWS_PER_BLOCK - 1) / sxm::ROWS_PER_BLOCK;
    if (train) {
        int rc = ensure_partials(n, blocks);
        if (rc != SX_OK) return rc;
    }
    sxm::StepArgs a;
    a.x = n->x;
    a.t = train ? n->t : nullptr;
    a.n = nrows_avail;
    a.rows = rows;
    a.sample = sample;
    a.seed = seed;
    a.step = stepno;
    a.w1b = n->w1b;
    a.w2b = n->w2b;
    a.p = n->p;
    a.y = train ? nullptr : n->y;
    a.partial = n->partial;
    hipLaunchKernelGGL(sxm::mlp_step_kernel, dim3(blocks), dim3(256), 0, n->stream, a);
    SXM_HIP(n, hipGetLastError());
    if (train) {
        float b1c = 1.f, b2c = 1.f;
        if (apply) {
            n->adam_t++;
            b1c = 1.f - powf(0.9f, (float)n->adam_t);
            b2c = 1.f - powf(0.999f, (float)n->adam_t);
        }
        hipLaunchKernelGGL(sxm::mlp_reduce_adam_kernel, dim3((sxm::NPARAM + 1 + 255) / 256), dim3(256), 0, n->stream,
                           n->partial, (int)blocks, n->grad, n->loss, n->p, n->m, n->v, n->w1b, n->w2b, apply, lr,
                           b1c, b2c);
        SXM_HIP(n, hipGetLastError());
    }
    return SX_OK;
}

}  // namespace

extern "C" {

size_t sx_mlp_param_count(void) { return (size_t)sxm::NPARAM; }

const char *sx_mlp_last_error(const sx_mlp *n) { return n ? n->err.c_str() : "null network"; }

int sx_mlp_create(int device, unsigned long long seed, sx_mlp **out) {
    if (!out) return SX_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SX_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return SX_ERR_HIP;
    sx_mlp *n = new (std::nothrow) sx_mlp();
    if (!n) return SX_ERR_NOMEM;
    n->device = device;
    int rc = SX_OK;
    auto fail = [&](hipError_t e) {
        if (e != hipSuccess && rc == SX_OK) rc = SX_ERR_HIP;
    };
    fail(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    fail(hipMalloc(&n->p, sizeof(float) * sxm::NPARAM));
    fail(hipMalloc(&n->m, sizeof(float) * sxm::NPARAM));
    fail(hipMalloc(&n->v, sizeof(float) * sxm::NPARAM));
    fail(hipMalloc(&n->grad, sizeof(float) * sxm::NPARAM));
    fail(hipMalloc(&n->loss, sizeof(float)));
    fail(hipMalloc(&n->w1b, sizeof(short) * sxm::IN * sxm::H));
    fail(hipMalloc(&n->w2b, sizeof(short) * sxm::H * sxm::H));
    if (rc != SX_OK) {
        sx_mlp_destroy(n);
        return rc;
    }
    // Glorot-uniform weights from a seeded generator, zero biases
    std::vector<float> p(sxm::NPARAM, 0.f);
    unsigned long long s = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(s >> 11) / 9007199254740992.0;
    };
    const double l1 = std::sqrt(6.0 / (sxm::IN + sxm::H)), l2 = std::sqrt(6.0 / (2 * sxm::H)),
                 l3 = std::sqrt(6.0 / (sxm::H + 1));
    for (int i = 0; i < sxm::IN * sxm::H; ++i) p[sxm::OFF_W1 + i] = (float)((2 * uni() - 1) * l1);
    for (int i = 0; i < sxm::H * sxm::H; ++i) p[sxm::OFF_W2 + i] = (float)((2 * uni() - 1) * l2);
    for (int i = 0; i < sxm::H; ++i) p[sxm::OFF_W3 + i] = (float)((2 * uni() - 1) * l3);
    rc = sx_mlp_set_params(n, p.data(), p.size());
    if (rc != SX_OK) {
        sx_mlp_destroy(n);
        return rc;
    }
    *out = n;
    return SX_OK;
}

void sx_mlp_destroy(sx_mlp *n) {
    if (!n) return;
    (void)hipSetDevice(n->device);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (float *q : {n->p, n->m, n->v, n->grad, n->partial, n->loss, n->x, n->t, n->y})
        if (q) (void)hipFree(q);
    if (n->w1b) (void)hipFree(n->w1b);
    if (n->w2b) (void)hipFree(n->w2b);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

int sx_mlp_set_params(sx_mlp *n, const float *params, size_t count) {
    if (!n) return SX_ERR_ARG;
    if (!params || count != (size_t)sxm::NPARAM) return mfail(n, SX_ERR_ARG, "expected %d parameters", sxm::NPARAM);
    SXM_HIP(n, hipSetDevice(n->device));
    SXM_HIP(n, hipMemcpyAsync(n->p, params, sizeof(float) * sxm::NPARAM, hipMemcpyHostToDevice, n->stream));
    SXM_HIP(n, hipMemsetAsync(n->m, 0, sizeof(float) * sxm::NPARAM, n->stream));
    SXM_HIP(n, hipMemsetAsync(n->v, 0, sizeof(float) * sxm::NPARAM, n->stream));
    hipLaunchKernelGGL(sxm::mlp_refresh_bf16_kernel, dim3((sxm::H * sxm::H + 255) / 256), dim3(256), 0, n->stream,
                       n->p, n->w1b, n->w2b);
    SXM_HIP(n, hipGetLastError());
    SXM_HIP(n, hipStreamSynchronize(n->stream));
    n->adam_t = 0;
    return SX_OK;
}

int sx_mlp_get_params(sx_mlp *n, float *params, size_t count) {
    if (!n) return SX_ERR_ARG;
    if (!params || count != (size_t)sxm::NPARAM) return mfail(n, SX_ERR_ARG, "expected %d parameters", sxm::NPARAM);
    SXM_HIP(n, hipSetDevice(n->device));
    SXM_HIP(n, hipMemcpyAsync(params, n->p, sizeof(float) * sxm::NPARAM, hipMemcpyDeviceToHost, n->stream));
    SXM_HIP(n, hipStreamSynchronize(n->stream));
    return SX_OK;
}

int sx_mlp_gradients(sx_mlp *n, const float *x, const float *t, size_t rows, float *grads, float *loss) {
    if (!n) return SX_ERR_ARG;
    int rc = check_rows(n, x, rows);
    if (rc != SX_OK) return rc;
    if (!t || !grads) return mfail(n, SX_ERR_ARG, "null targets or gradient buffer");
    SXM_HIP(n, hipSetDevice(n->device));
    rc = ensure_rows(n, rows);
    if (rc != SX_OK) return rc;
    SXM_HIP(n, hipMemcpyAsync(n->x, x, sizeof(float) * sxm::IN * rows, hipMemcpyHostToDevice, n->stream));
    SXM_HIP(n, hipMemcpyAsync(n->t, t, sizeof(float) * rows, hipMemcpyHostToDevice, n->stream));
    rc = step(n, (unsigned)rows, (unsigned)rows, 0, 0, 0, true, 0, 0.f);
    if (rc != SX_OK) return rc;
    SXM_HIP(n, hipMemcpyAsync(grads, n->grad, sizeof(float) * sxm::NPARAM, hipMemcpyDeviceToHost, n->stream));
    float l = 0.f;
    SXM_HIP(n, hipMemcpyAsync(&l, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    SXM_HIP(n, hipStreamSynchronize(n->stream));
    if (loss) *loss = l / (float)rows;
    return SX_OK;
}

int sx_mlp_train(sx_mlp *n, const float *x, const float *t, size_t rows, int steps, int batch, float lr,
                 unsigned long long seed, float *final_loss) {
    if (!n) return SX_ERR_ARG;
    int rc = check_rows(n, x, rows);
    if (rc != SX_OK) return rc;
    if (!t || steps <= 0 || batch <= 0 || batch > (1 << 20) || !(lr > 0.f))
        return mfail(n, SX_ERR_ARG, "bad training arguments");
    SXM_HIP(n, hipSetDevice(n->device));
    rc = ensure_rows(n, rows);
    if (rc != SX_OK) return rc;
    SXM_HIP(n, hipMemcpyAsync(n->x, x, sizeof(float) * sxm::IN * rows, hipMemcpyHostToDevice, n->stream));
    SXM_HIP(n, hipMemcpyAsync(n->t, t, sizeof(float) * rows, hipMemcpyHostToDevice, n->stream));
    for (int s = 0; s < steps; ++s) {
        // cosine decay to 5 % of the initial rate
        const float lrs = lr * (0.05f + 0.95f * 0.5f * (1.f + cosf(3.14159265f * (float)s / (float)steps)));
        rc = step(n, (unsigned)rows, (unsigned)batch, 1, seed, (unsigned)s, true, 1, lrs);
        if (rc != SX_OK) return rc;
    }
    float l = 0.f;
    SXM_HIP(n, hipMemcpyAsync(&l, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    SXM_HIP(n, hipStreamSynchronize(n->stream));
    if (final_loss) *final_loss = l / (float)batch;
    return SX_OK;
}

int sx_mlp_predict(sx_mlp *n, const float *x, size_t rows, float *y) {
    if (!n) return SX_ERR_ARG;
    int rc = check_rows(n, x, rows);
    if (rc != SX_OK) return rc;
    if (!y) return mfail(n, SX_ERR_ARG, "null output");
    SXM_HIP(n, hipSetDevice(n->device));
    rc = ensure_rows(n, rows);
    if (rc != SX_OK) return rc;
    SXM_HIP(n, hipMemcpyAsync(n->x, x, sizeof(float) * sxm::IN * rows, hipMemcpyHostToDevice, n->stream));
    rc = step(n, (unsigned)rows, (unsigned)rows, 0, 0, 0, false, 0, 0.f);
    if (rc != SX_OK) return rc;
    SXM_HIP(n, hipMemcpyAsync(y, n->y, sizeof(float) * rows, hipMemcpyDeviceToHost, n->stream));
    SXM_HIP(n, hipStreamSynchronize(n->stream));
    return SX_OK;
}

}  // extern "C"
