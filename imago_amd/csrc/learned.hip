// learned.hip - the learned initial SCF potential: per-site environment
// descriptors and the per-site MLP (forward, backward, Adam), FP32 except
// the descriptor sums.
//
// The reference's SCF starts from scfV.dat (initPotCoeffs,
// src/imago/potential.f90:433-564), whose coefficients makeinput copies
// from isolated-atom database entries (makeinput.py:5805-5941).  This
// module predicts, per potential site, the correction of those atomic
// coefficients from the site's environment; the prediction is written to
// the same coefficient vector the reference reads.
//
// Descriptor of site i (96 values): for 32 Gaussian radial channels
// g_k(r) = exp(-eta (r - mu_k)^2), mu_k = r_min + k (r_c - r_min) / 31, and
// the cosine cut-off f(r) = (cos(pi r / r_c) + 1) / 2, with a_j = w_j g_k f
// over every neighbour image j (w_j a species weight):
//   l = 0:  sum_j a_j
//   l = 1:  |sum_j a_j u_j|^2                      (u_j the unit vector)
//   l = 2:  ||sum_j a_j (u_j u_j^T - I/3)||_F^2
// each invariant under rotation, translation and neighbour order.
//
// MLP: layers of widths w_0 -> w_1 -> ... -> w_L, tanh on hidden layers,
// linear output; row-major activations [batch][width]; parameters packed
// per layer as W_l [w_{l+1}][w_l] row-major, then b_l [w_{l+1}].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "common.hpp"

int imago_ctx_device(imago_ctx *c);
hipStream_t imago_ctx_stream(imago_ctx *c);

namespace {

constexpr int kRad = 32;
constexpr int kDesc = 3 * kRad;

// one thread per (site, radial channel); double sums
__global__ __launch_bounds__(256) void k_site_descriptor(int ns, const double *__restrict__ pos,
                                                         const double *__restrict__ w, int nc,
                                                         const double *__restrict__ cells,
                                                         double rc, double rmin, double eta,
                                                         double *__restrict__ out)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= ns * kRad) return;
    const int i = g / kRad, k = g % kRad;
    const double mu = rmin + k * (rc - rmin) / (kRad - 1);
    const double pi = 3.141592653589793;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    double s0 = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    double qxx = 0.0, qyy = 0.0, qzz = 0.0, qxy = 0.0, qxz = 0.0, qyz = 0.0;
    for (int j = 0; j < ns; ++j) {
        const double bx = pos[3 * j] - xi, by = pos[3 * j + 1] - yi, bz = pos[3 * j + 2] - zi;
        for (int c = 0; c < nc; ++c) {
            const double dx = bx + cells[3 * c], dy = by + cells[3 * c + 1],
                         dz = bz + cells[3 * c + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 >= rc * rc || r2 < 1e-20) continue;
            const double r = sqrt(r2);
            const double t = r - mu;
            const double a = w[j] * exp(-eta * t * t) * 0.5 * (cos(pi * r / rc) + 1.0);
            const double ux = dx / r, uy = dy / r, uz = dz / r;
            s0 += a;
            vx += a * ux;
            vy += a * uy;
            vz += a * uz;
            qxx += a * (ux * ux - 1.0 / 3.0);
            qyy += a * (uy * uy - 1.0 / 3.0);
            qzz += a * (uz * uz - 1.0 / 3.0);
            qxy += a * ux * uy;
            qxz += a * ux * uz;
            qyz += a * uy * uz;
        }
    }
    double *o = out + (size_t)i * kDesc;
    o[k] = s0;
    o[kRad + k] = vx * vx + vy * vy + vz * vz;
    o[2 * kRad + k] = qxx * qxx + qyy * qyy + qzz * qzz + 2.0 * (qxy * qxy + qxz * qxz + qyz * qyz);
}

// C[M x N] = op(A) op(B) (+ C when accumulate), row-major FP32;
// op(A) is A [M][K] or, with TA, A stored [K][M]; op(B) is B [K][N] or,
// with TB, B stored [N][K].  16 x 16 output tiles, K in steps of 16
// through LDS.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_sgemm(int M, int N, int K, const float *__restrict__ A,
                                               const float *__restrict__ B, float *__restrict__ C,
                                               int accumulate)
{
    __shared__ float As[16][17], Bs[16][17];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
    float acc = 0.0f;
    for (int k0 = 0; k0 < K; k0 += 16) {
        {   // As[ty][tx] = op(A)(blockIdx.y*16 + ty, k0 + tx)
            const int r = blockIdx.y * 16 + ty, kk = k0 + tx;
            float v = 0.0f;
            if (r < M && kk < K) v = TA ? A[(size_t)kk * M + r] : A[(size_t)r * K + kk];
            As[ty][tx] = v;
        }
        {   // Bs[ty][tx] = op(B)(k0 + ty, blockIdx.x*16 + tx)
            const int kk = k0 + ty, cc = blockIdx.x * 16 + tx;
            float v = 0.0f;
            if (kk < K && cc < N) v = TB ? B[(size_t)cc * K + kk] : B[(size_t)kk * N + cc];
            Bs[ty][tx] = v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q) acc = fmaf(As[ty][q], Bs[q][tx], acc);
        __syncthreads();
    }
    if (row < M && col < N) {
        float *c = C + (size_t)row * N + col;
        *c = accumulate ? *c + acc : acc;
    }
}

// H = act(Z + b) in place; act = tanh or identity
__global__ __launch_bounds__(256) void k_bias_act(float *__restrict__ Z, const float *__restrict__ b,
                                                  int M, int N, int use_tanh)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)M * N) return;
    const float z = Z[g] + b[g % N];
    Z[g] = use_tanh ? tanhf(z) : z;
}

// dZ = dH * (1 - H^2) in place (tanh)
__global__ __launch_bounds__(256) void k_tanh_bwd(float *__restrict__ dH, const float *__restrict__ H,
                                                  size_t n)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g < n) dH[g] = dH[g] * (1.0f - H[g] * H[g]);
}

// db[n] = sum_m dZ[m][n]; one work-group per column, fixed-order tree
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ dZ, int M, int N,
                                                float *__restrict__ db)
{
    __shared__ float s[256];
    const int n = blockIdx.x;
    float a = 0.0f;
    for (int m = threadIdx.x; m < M; m += 256) a += dZ[(size_t)m * N + n];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[n] = s[0];
}

// mean squared error over M x N: dY = 2 (Y - T) / (M N); per-block partial
// sums of (Y - T)^2 in double
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ Y, const float *__restrict__ T,
                                             size_t n, float scale, float *__restrict__ dY,
                                             double *__restrict__ part)
{
    __shared__ double s[256];
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    double e2 = 0.0;
    if (g < n) {
        const float d = Y[g] - T[g];
        dY[g] = 2.0f * d * scale;
        e2 = (double)d * d;
    }
    s[threadIdx.x] = e2;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Adam (Kingma and Ba 2015) with bias correction at step t (1-based)
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, size_t n,
                                              float lr, float b1, float b2, float eps, float c1,
                                              float c2)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.0f - b1) * gq;
    const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

// rows idx[0..B) of X [*, D] into Xb [B][D]
__global__ __launch_bounds__(256) void k_gather_rows(const float *__restrict__ X,
                                                     const int *__restrict__ idx, int B, int D,
                                                     float *__restrict__ Xb)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)B * D) return;
    const int b = (int)(g / D), d = (int)(g % D);
    Xb[g] = X[(size_t)idx[b] * D + d];
}

unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct imago_mlp {
    imago_ctx *ctx = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<int> w;           // widths w_0 .. w_L
    std::vector<size_t> off;      // offset of W_l in the packed parameters; b_l follows
    size_t np = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // device
    long step = 0;                // Adam steps taken
    int cap = 0;                  // batch capacity of the buffers below
    std::vector<float *> h;       // activations per layer [cap][w_l]
    std::vector<float *> dh;      // their gradients
    float *target = nullptr;
    double *part = nullptr;
    // training data resident on the device (imago_mlp_train)
    float *X = nullptr, *T = nullptr;
    int *idx = nullptr;
    int nx = 0;
};

namespace {

void mlp_free_batch(imago_mlp *m)
{
    for (float *q : m->h) (void)hipFree(q);
    for (float *q : m->dh) (void)hipFree(q);
    m->h.clear();
    m->dh.clear();
    if (m->target) (void)hipFree(m->target);
    if (m->part) (void)hipFree(m->part);
    if (m->idx) (void)hipFree(m->idx);
    m->target = nullptr;
    m->part = nullptr;
    m->idx = nullptr;
    m->cap = 0;
}

int mlp_reserve(imago_mlp *m, int batch)
{
    if (batch <= m->cap) return IMAGO_OK;
    mlp_free_batch(m);
    const int L = (int)m->w.size() - 1;
    for (int l = 0; l <= L; ++l) {
        float *a = nullptr, *b = nullptr;
        HIPCHK(hipMalloc(&a, sizeof(float) * (size_t)batch * m->w[l]));
        HIPCHK(hipMalloc(&b, sizeof(float) * (size_t)batch * m->w[l]));
        m->h.push_back(a);
        m->dh.push_back(b);
    }
    HIPCHK(hipMalloc(&m->target, sizeof(float) * (size_t)batch * m->w[L]));
    HIPCHK(hipMalloc(&m->part, sizeof(double) * blocks((size_t)batch * m->w[L])));
    HIPCHK(hipMalloc(&m->idx, sizeof(int) * (size_t)batch));
    m->cap = batch;
    return IMAGO_OK;
}

void gemm(hipStream_t s, bool ta, bool tb, int M, int N, int K, const float *A, const float *B,
          float *C, int acc)
{
    const dim3 grid((N + 15) / 16, (M + 15) / 16);
    if (!ta && !tb)
        hipLaunchKernelGGL((k_sgemm<false, false>), grid, dim3(256), 0, s, M, N, K, A, B, C, acc);
    else if (!ta && tb)
        hipLaunchKernelGGL((k_sgemm<false, true>), grid, dim3(256), 0, s, M, N, K, A, B, C, acc);
    else if (ta && !tb)
        hipLaunchKernelGGL((k_sgemm<true, false>), grid, dim3(256), 0, s, M, N, K, A, B, C, acc);
    else
        hipLaunchKernelGGL((k_sgemm<true, true>), grid, dim3(256), 0, s, M, N, K, A, B, C, acc);
}

// forward over the batch held in h[0]; output in h[L]
int mlp_forward_dev(imago_mlp *m, int B)
{
    const int L = (int)m->w.size() - 1;
    for (int l = 0; l < L; ++l) {
        const int in = m->w[l], out = m->w[l + 1];
        const float *W = m->p + m->off[l], *b = W + (size_t)out * in;
        gemm(m->stream, false, true, B, out, in, m->h[l], W, m->h[l + 1], 0);   // X W^T
        hipLaunchKernelGGL(k_bias_act, dim3(blocks((size_t)B * out)), dim3(256), 0, m->stream,
                           m->h[l + 1], b, B, out, l + 1 < L ? 1 : 0);
    }
    HIPCHK(hipGetLastError());
    return IMAGO_OK;
}

// loss and gradient for the batch in h[0] against m->target; returns the
// sum of squared errors through *sse (mean over batch x outputs is the loss)
int mlp_backward_dev(imago_mlp *m, int B, double *sse)
{
    const int L = (int)m->w.size() - 1;
    const size_t ny = (size_t)B * m->w[L];
    const unsigned nb = blocks(ny);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, m->stream, m->h[L], m->target, ny,
                       1.0f / (float)ny, m->dh[L], m->part);
    for (int l = L - 1; l >= 0; --l) {
        const int in = m->w[l], out = m->w[l + 1];
        if (l + 1 < L)
            hipLaunchKernelGGL(k_tanh_bwd, dim3(blocks((size_t)B * out)), dim3(256), 0,
                               m->stream, m->dh[l + 1], m->h[l + 1], (size_t)B * out);
        float *gW = m->g + m->off[l], *gb = gW + (size_t)out * in;
        gemm(m->stream, true, false, out, in, B, m->dh[l + 1], m->h[l], gW, 0);   // dZ^T X
        hipLaunchKernelGGL(k_colsum, dim3(out), dim3(256), 0, m->stream, m->dh[l + 1], B, out, gb);
        if (l > 0) {
            const float *W = m->p + m->off[l];
            gemm(m->stream, false, false, B, in, out, m->dh[l + 1], W, m->dh[l], 0);   // dZ W
        }
    }
    HIPCHK(hipGetLastError());
    std::vector<double> part(nb);
    HIPCHK(hipMemcpyAsync(part.data(), m->part, sizeof(double) * nb, hipMemcpyDeviceToHost,
                          m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    double s = 0.0;
    for (double x : part) s += x;
    *sse = s;
    return IMAGO_OK;
}

int mlp_adam_dev(imago_mlp *m, float lr, float b1, float b2, float eps)
{
    m->step += 1;
    const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
    hipLaunchKernelGGL(k_adam, dim3(blocks(m->np)), dim3(256), 0, m->stream, m->p, m->g, m->m, m->v,
                       m->np, lr, b1, b2, eps, c1, c2);
    HIPCHK(hipGetLastError());
    return IMAGO_OK;
}

}  // namespace

extern "C" int imago_site_descriptors(imago_ctx *c, int num_sites, const double *pos,
                                      const double *weight, int num_cells, const double *cells,
                                      double cutoff, double r_min, double eta, double *desc)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (num_sites < 1 || num_cells < 1 || !pos || !weight || !cells || !desc)
        return fail(IMAGO_E_ARG, "bad descriptor arguments");
    if (!(cutoff > r_min) || !(eta > 0.0)) return fail(IMAGO_E_ARG, "bad radial channels");
    HIPCHK(hipSetDevice(imago_ctx_device(c)));
    hipStream_t s = imago_ctx_stream(c);
    double *dpos = nullptr, *dw = nullptr, *dc = nullptr, *dout = nullptr;
    const size_t nout = (size_t)num_sites * kDesc;
    HIPCHK(hipMalloc(&dpos, sizeof(double) * 3 * num_sites));
    HIPCHK(hipMalloc(&dw, sizeof(double) * num_sites));
    HIPCHK(hipMalloc(&dc, sizeof(double) * 3 * num_cells));
    HIPCHK(hipMalloc(&dout, sizeof(double) * nout));
    HIPCHK(hipMemcpyAsync(dpos, pos, sizeof(double) * 3 * num_sites, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dw, weight, sizeof(double) * num_sites, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dc, cells, sizeof(double) * 3 * num_cells, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_site_descriptor, dim3(blocks((size_t)num_sites * kRad)), dim3(256), 0, s,
                       num_sites, dpos, dw, num_cells, dc, cutoff, r_min, eta, dout);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(desc, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    (void)hipFree(dpos);
    (void)hipFree(dw);
    (void)hipFree(dc);
    (void)hipFree(dout);
    return IMAGO_OK;
}

extern "C" int imago_mlp_create(imago_ctx *c, int num_layers, const int *widths, unsigned seed,
                                imago_mlp **out)
{
    if (!c || !out || !widths || num_layers < 1) return fail(IMAGO_E_ARG, "bad MLP arguments");
    for (int l = 0; l <= num_layers; ++l)
        if (widths[l] < 1) return fail(IMAGO_E_ARG, "layer %d has width %d", l, widths[l]);
    HIPCHK(hipSetDevice(imago_ctx_device(c)));
    imago_mlp *m = new imago_mlp;
    m->ctx = c;
    m->device = imago_ctx_device(c);
    m->stream = imago_ctx_stream(c);
    m->w.assign(widths, widths + num_layers + 1);
    size_t o = 0;
    for (int l = 0; l < num_layers; ++l) {
        m->off.push_back(o);
        o += (size_t)m->w[l] * m->w[l + 1] + m->w[l + 1];
    }
    m->np = o;
    // Glorot-uniform weights, zero biases
    std::vector<float> p(m->np, 0.0f);
    std::mt19937 rng(seed);
    for (int l = 0; l < num_layers; ++l) {
        const float lim = std::sqrt(6.0f / (float)(m->w[l] + m->w[l + 1]));
        std::uniform_real_distribution<float> U(-lim, lim);
        for (size_t q = 0; q < (size_t)m->w[l] * m->w[l + 1]; ++q) p[m->off[l] + q] = U(rng);
    }
    hipError_t e = hipSuccess;
    for (float **q : {&m->p, &m->g, &m->m, &m->v})
        if (e == hipSuccess) e = hipMalloc(q, sizeof(float) * m->np);
    // on the context's stream: the legacy default stream would wait on the
    // streams of other contexts in the process
    if (e == hipSuccess)
        e = hipMemcpyAsync(m->p, p.data(), sizeof(float) * m->np, hipMemcpyHostToDevice,
                           m->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->m, 0, sizeof(float) * m->np, m->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->v, 0, sizeof(float) * m->np, m->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->g, 0, sizeof(float) * m->np, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) {
        for (float *q : {m->p, m->g, m->m, m->v})
            if (q) (void)hipFree(q);
        delete m;
        return fail(IMAGO_E_NOMEM, "MLP parameters: %s", hipGetErrorString(e));
    }
    *out = m;
    return IMAGO_OK;
}

extern "C" int imago_mlp_destroy(imago_mlp *m)
{
    if (!m) return IMAGO_OK;
    if (hipSetDevice(m->device) == hipSuccess) {
        (void)hipStreamSynchronize(m->stream);
        mlp_free_batch(m);
        for (float *q : {m->p, m->g, m->m, m->v, m->X, m->T})
            if (q) (void)hipFree(q);
    }
    delete m;
    return IMAGO_OK;
}

extern "C" int imago_mlp_num_params(imago_mlp *m, long *n)
{
    if (!m || !n) return fail(IMAGO_E_ARG, "null argument");
    *n = (long)m->np;
    return IMAGO_OK;
}

extern "C" int imago_mlp_params(imago_mlp *m, float *get, const float *set)
{
    if (!m) return fail(IMAGO_E_ARG, "null MLP");
    HIPCHK(hipSetDevice(m->device));
    if (set) {
        HIPCHK(hipMemcpyAsync(m->p, set, sizeof(float) * m->np, hipMemcpyHostToDevice, m->stream));
        HIPCHK(hipMemsetAsync(m->m, 0, sizeof(float) * m->np, m->stream));
        HIPCHK(hipMemsetAsync(m->v, 0, sizeof(float) * m->np, m->stream));
        m->step = 0;
    }
    if (get)
        HIPCHK(hipMemcpyAsync(get, m->p, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return IMAGO_OK;
}

extern "C" int imago_mlp_forward(imago_mlp *m, int batch, const float *x, float *y)
{
    if (!m || batch < 1 || !x || !y) return fail(IMAGO_E_ARG, "bad forward arguments");
    HIPCHK(hipSetDevice(m->device));
    RCHK(mlp_reserve(m, batch));
    const int L = (int)m->w.size() - 1;
    HIPCHK(hipMemcpyAsync(m->h[0], x, sizeof(float) * (size_t)batch * m->w[0],
                          hipMemcpyHostToDevice, m->stream));
    RCHK(mlp_forward_dev(m, batch));
    HIPCHK(hipMemcpyAsync(y, m->h[L], sizeof(float) * (size_t)batch * m->w[L],
                          hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return IMAGO_OK;
}

extern "C" int imago_mlp_gradient(imago_mlp *m, int batch, const float *x, const float *target,
                                  double *loss, float *grad)
{
    if (!m || batch < 1 || !x || !target) return fail(IMAGO_E_ARG, "bad gradient arguments");
    HIPCHK(hipSetDevice(m->device));
    RCHK(mlp_reserve(m, batch));
    const int L = (int)m->w.size() - 1;
    HIPCHK(hipMemcpyAsync(m->h[0], x, sizeof(float) * (size_t)batch * m->w[0],
                          hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->target, target, sizeof(float) * (size_t)batch * m->w[L],
                          hipMemcpyHostToDevice, m->stream));
    RCHK(mlp_forward_dev(m, batch));
    double sse = 0.0;
    RCHK(mlp_backward_dev(m, batch, &sse));
    if (loss) *loss = sse / ((double)batch * m->w[L]);
    if (grad) {
        HIPCHK(hipMemcpyAsync(grad, m->g, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
    }
    return IMAGO_OK;
}

extern "C" int imago_mlp_adam_step(imago_mlp *m, float lr, float beta1, float beta2, float eps)
{
    if (!m) return fail(IMAGO_E_ARG, "null MLP");
    HIPCHK(hipSetDevice(m->device));
    RCHK(mlp_adam_dev(m, lr, beta1, beta2, eps));
    HIPCHK(hipStreamSynchronize(m->stream));
    return IMAGO_OK;
}

extern "C" int imago_mlp_train(imago_mlp *m, int num_samples, const float *x, const float *target,
                               int epochs, int minibatch, float lr, float beta1, float beta2,
                               float eps, unsigned seed, double *epoch_loss)
{
    if (!m || num_samples < 1 || !x || !target || epochs < 1 || minibatch < 1)
        return fail(IMAGO_E_ARG, "bad training arguments");
    HIPCHK(hipSetDevice(m->device));
    const int L = (int)m->w.size() - 1, din = m->w[0], dout = m->w[L];
    const int B = std::min(minibatch, num_samples);
    RCHK(mlp_reserve(m, B));
    if (m->X) (void)hipFree(m->X);
    if (m->T) (void)hipFree(m->T);
    m->X = m->T = nullptr;
    HIPCHK(hipMalloc(&m->X, sizeof(float) * (size_t)num_samples * din));
    HIPCHK(hipMalloc(&m->T, sizeof(float) * (size_t)num_samples * dout));
    HIPCHK(hipMemcpyAsync(m->X, x, sizeof(float) * (size_t)num_samples * din,
                          hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->T, target, sizeof(float) * (size_t)num_samples * dout,
                          hipMemcpyHostToDevice, m->stream));
    std::vector<int> order(num_samples);
    for (int i = 0; i < num_samples; ++i) order[i] = i;
    std::mt19937 rng(seed);
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        double sse = 0.0;
        long cnt = 0;
        for (int b0 = 0; b0 + B <= num_samples; b0 += B) {
            HIPCHK(hipMemcpyAsync(m->idx, order.data() + b0, sizeof(int) * B,
                                  hipMemcpyHostToDevice, m->stream));
            hipLaunchKernelGGL(k_gather_rows, dim3(blocks((size_t)B * din)), dim3(256), 0,
                               m->stream, m->X, m->idx, B, din, m->h[0]);
            hipLaunchKernelGGL(k_gather_rows, dim3(blocks((size_t)B * dout)), dim3(256), 0,
                               m->stream, m->T, m->idx, B, dout, m->target);
            RCHK(mlp_forward_dev(m, B));
            double s = 0.0;
            RCHK(mlp_backward_dev(m, B, &s));
            RCHK(mlp_adam_dev(m, lr, beta1, beta2, eps));
            sse += s;
            cnt += (long)B * dout;
        }
        if (epoch_loss) epoch_loss[ep] = cnt ? sse / cnt : 0.0;
    }
    HIPCHK(hipStreamSynchronize(m->stream));
    return IMAGO_OK;
}

// ---- forward pass that allocates nothing -----------------------------------
//
// For callers that cannot allocate (code under HIP graph capture, serving
// loops): the hidden activations live in a file-scope device array of
// kNoAllocFloats floats (4 MiB on every device the library runs on, held
// until the process ends).  Its address on each device comes from
// hipGetSymbolAddress on that device's first call; host code never uses
// the array's name as a device address.  Every call on a device shares
// the device's array: calls that overlap on one device race.

namespace {

constexpr size_t kNoAllocFloats = (size_t)1 << 20;
__device__ __align__(256) float scratch[kNoAllocFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};

int scratch_on_device(int dev, float **out)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(IMAGO_E_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(IMAGO_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_ptr[dev];
    return IMAGO_OK;
}

}  // namespace

extern "C" int imago_mlp_forward_noalloc(imago_mlp *m, int batch, const float *x_dev,
                                         float *y_dev)
{
    if (!m || batch < 1 || !x_dev || !y_dev) return fail(IMAGO_E_ARG, "bad forward arguments");
    const int L = (int)m->w.size() - 1;
    // two hidden buffers used alternately: the widest hidden layer, twice
    size_t widest = 0;
    for (int l = 1; l < L; ++l) widest = std::max(widest, (size_t)m->w[l]);
    const size_t need = 2 * widest * (size_t)batch;
    if (need > kNoAllocFloats)
        return fail(IMAGO_E_ARG, "batch %d needs %zu scratch floats, the limit is %zu (batch <= "
                                 "%zu for this network)", batch, need, kNoAllocFloats,
                    widest ? kNoAllocFloats / (2 * widest) : (size_t)0);
    HIPCHK(hipSetDevice(m->device));
    float *base = nullptr;
    RCHK(scratch_on_device(m->device, &base));
    float *buf[2] = {base, base + widest * (size_t)batch};
    const float *in = x_dev;
    for (int l = 0; l < L; ++l) {
        const int din = m->w[l], dout = m->w[l + 1];
        float *out = l + 1 < L ? buf[l % 2] : y_dev;
        const float *W = m->p + m->off[l], *b = W + (size_t)dout * din;
        gemm(m->stream, false, true, batch, dout, din, in, W, out, 0);
        hipLaunchKernelGGL(k_bias_act, dim3(blocks((size_t)batch * dout)), dim3(256), 0,
                           m->stream, out, b, batch, dout, l + 1 < L ? 1 : 0);
        in = out;
    }
    HIPCHK(hipGetLastError());
    return IMAGO_OK;
}
