// hetrd.hip - launcher of the cooperative tridiagonal reduction
// (hetrd_coop.hpp).
#include "hetrd_coop.hpp"
#include "hetrd_launch.hpp"

namespace imago {

namespace {

template <int S>
hipError_t trd_one(const TrdLaunch &L)
{
    const void *fn = reinterpret_cast<const void *>(&k_hetrd_coop<S>);
    int dev = 0, ncu = 0, per_cu = 0;
    hipError_t st = hipGetDevice(&dev);
    if (st != hipSuccess) return st;
    st = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (st != hipSuccess) return st;
    st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kTrdThreads, 0);
    if (st != hipSuccess) return st;
    const int P = (L.n + kTrdRows - 1) / kTrdRows;
    const int fit = per_cu * ncu / P;   // matrices whose work-groups are co-resident
    if (fit < 1) return hipErrorCooperativeLaunchTooLarge;
    const size_t nn = (size_t)L.n * L.n;
    for (int m0 = 0; m0 < L.nb; m0 += fit) {
        const int nb = std::min(fit, L.nb - m0);
        st = hipMemsetAsync(L.bar, 0, sizeof(unsigned) * ((size_t)nb * P * kTrdFlagStride + 1),
                            L.stream);
        if (st != hipSuccess) return st;
        TrdArgs T;
        T.n = L.n;
        T.P = P;
        T.nb_launch = nb;
        T.A = L.A + m0 * nn;
        T.V = L.V + m0 * nn;
        T.d = L.d + (size_t)m0 * L.n;
        T.e = L.e + (size_t)m0 * L.n;
        T.tau = L.tau + (size_t)m0 * L.n;
        T.vbuf = L.vbuf;
        T.pbuf = L.pbuf;
        T.dbuf = L.dbuf;
        T.bar = L.bar;
        T.err = L.err;
        T.spin_limit = 2000000;
        void *args[] = {&T};
        st = hipLaunchCooperativeKernel(fn, dim3((unsigned)(nb * P)), dim3(kTrdThreads), args, 0,
                                        L.stream);
        if (st != hipSuccess) return st;
    }
    return hipSuccess;
}

}  // namespace

// one work-group per eigenvector q: the block of its largest entry
__global__ __launch_bounds__(256) void k_trd_owner(const double *__restrict__ Z, int N, int n,
                                                   int *__restrict__ owner)
{
    __shared__ double bv[256];
    __shared__ int bi[256];
    const int q = blockIdx.x;
    const double *z = Z + (size_t)q * N;
    double best = -1.0;
    int at = 0;
    for (int r = threadIdx.x; r < N; r += 256) {
        const double a = fabs(z[r]);
        if (a > best) {
            best = a;
            at = r;
        }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = at;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = bv[threadIdx.x + s];
            const int oi = bi[threadIdx.x + s];
            if (o > bv[threadIdx.x] || (o == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
                bv[threadIdx.x] = o;
                bi[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) owner[q] = bi[0] / n;
}

// rank[q] = number of q' < q with owner[q'] == owner[q]; one work-group of
// 1024 threads, an exclusive scan of the flags of each block in turn
__global__ __launch_bounds__(1024) void k_trd_rank(const int *__restrict__ owner, int N, int nb,
                                                   int *__restrict__ rank)
{
    __shared__ int sc[1024];
    const int t = threadIdx.x;
    const int per = (N + 1023) / 1024;
    const int q0 = t * per, q1 = min(N, q0 + per);
    for (int b = 0; b < nb; ++b) {
        int cnt = 0;
        for (int q = q0; q < q1; ++q) cnt += owner[q] == b;
        sc[t] = cnt;
        __syncthreads();
        // inclusive Hillis-Steele scan
        for (int d = 1; d < 1024; d <<= 1) {
            const int x = t >= d ? sc[t - d] : 0;
            __syncthreads();
            sc[t] += x;
            __syncthreads();
        }
        int base = sc[t] - cnt;
        for (int q = q0; q < q1; ++q)
            if (owner[q] == b) rank[q] = base++;
        __syncthreads();
    }
}

// one work-group per eigenvector q: its block's part as complex column
// rank[q] of X[owner[q]], its eigenvalue to w[owner[q]][rank[q]]
__global__ __launch_bounds__(256) void k_trd_gather(const double *__restrict__ Z,
                                                    const double *__restrict__ wall, int N, int n,
                                                    const int *__restrict__ owner,
                                                    const int *__restrict__ rank,
                                                    double *__restrict__ w, double2 *__restrict__ X)
{
    const int q = blockIdx.x;
    const int b = owner[q], p = rank[q];
    if (p >= n) return;
    const double *z = Z + (size_t)q * N + (size_t)b * n;
    double2 *x = X + (size_t)b * n * n + (size_t)p * n;
    for (int r = threadIdx.x; r < n; r += 256) x[r] = make_double2(z[r], 0.0);
    if (threadIdx.x == 0) w[(size_t)b * n + p] = wall[q];
}

hipError_t launch_trd_unmix(const double *Z, const double *wall, int n, int nb, int b0,
                            int *owner, int *rank, double *w, double2 *X, hipStream_t stream)
{
    const int N = n * nb;
    hipLaunchKernelGGL(k_trd_owner, dim3(N), dim3(256), 0, stream, Z, N, n, owner);
    hipLaunchKernelGGL(k_trd_rank, dim3(1), dim3(1024), 0, stream, owner, N, nb, rank);
    hipLaunchKernelGGL(k_trd_gather, dim3(N), dim3(256), 0, stream, Z, wall, N, n, owner, rank,
                       w + (size_t)b0 * n, X + (size_t)b0 * n * n);
    return hipGetLastError();
}

// zlarft (direct = 'F', storev = 'C') for block b of matrix m, one
// work-group each: G = V_b^H V_b over the rows below the block's first
// reflector (64-row tiles of V_b staged in LDS, four entries of G per
// thread), then T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i), T(i, i) = tau_i.
// Rows of a reflector above its unit entry are zero in V.
__global__ __launch_bounds__(256) void k_trd_larft(const double2 *__restrict__ V,
                                                   const double2 *__restrict__ tau, int n,
                                                   int nblocks, double2 *__restrict__ T)
{
    constexpr int K = kTrdBlock, R = 64;
    __shared__ double2 tile[R][K + 1];
    __shared__ double2 G[K][K + 1];
    __shared__ double2 Ts[K][K + 1];
    const int m = blockIdx.x / nblocks, b = blockIdx.x % nblocks;
    const int j0 = b * K, nbk = min(K, n - 1 - j0);
    const size_t nn = (size_t)n * n;
    const double2 *Vm = V + m * nn;
    const double2 *tm = tau + (size_t)m * n;
    double2 acc[K * K / 256];
#pragma unroll
    for (int q = 0; q < K * K / 256; ++q) acc[q] = trd_zero();
    for (int r0 = j0 + 1; r0 < n; r0 += R) {
        for (int e = threadIdx.x; e < R * K; e += 256) {
            const int rr = e % R, cc = e / R, r = r0 + rr;
            tile[rr][cc] = (r < n && cc < nbk) ? Vm[(size_t)(j0 + cc) * n + r] : trd_zero();
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < K * K / 256; ++q) {
            const int e = threadIdx.x + 256 * q, kr = e % K, ic = e / K;
            if (kr < ic) {
                double2 s = acc[q];
                for (int rr = 0; rr < R; ++rr) {
                    const double2 x = trd_mulc(tile[rr][ic], tile[rr][kr]);   // conj(vk) vi
                    s.x += x.x;
                    s.y += x.y;
                }
                acc[q] = s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < K * K / 256; ++q) {
        const int e = threadIdx.x + 256 * q, kr = e % K, ic = e / K;
        G[kr][ic] = acc[q];
        Ts[kr][ic] = trd_zero();
    }
    __syncthreads();
    for (int i = 0; i < nbk; ++i) {
        const double2 ti = tm[j0 + i];
        if ((int)threadIdx.x < i) {
            // -tau_i * sum_k T(row, k) G(k, i), k in [row, i)
            const int row = threadIdx.x;
            double2 s = trd_zero();
            for (int k = row; k < i; ++k) {
                const double2 q = trd_mul(Ts[row][k], G[k][i]);
                s.x += q.x;
                s.y += q.y;
            }
            const double2 v = trd_mul(ti, s);
            Ts[row][i] = make_double2(-v.x, -v.y);
        }
        if (threadIdx.x == 0) Ts[i][i] = ti;
        __syncthreads();
    }
    double2 *To = T + ((size_t)m * nblocks + b) * K * K;
    for (int e = threadIdx.x; e < K * K; e += 256) To[e] = Ts[e % K][e / K];
}

hipError_t launch_trd_larft(const double2 *V, const double2 *tau, int n, int nb, double2 *T,
                            hipStream_t stream)
{
    const int nblocks = (n - 1 + kTrdBlock - 1) / kTrdBlock;
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trd_larft, dim3((unsigned)(nb * nblocks)), dim3(256), 0, stream, V, tau,
                       n, nblocks, T);
    return hipGetLastError();
}

bool hetrd_coop_supported(int n) { return n >= 2 && n <= kTrdMaxOrder; }

hipError_t launch_hetrd_coop(const TrdLaunch &L)
{
    if (!hetrd_coop_supported(L.n) || L.nb < 1) return hipErrorInvalidValue;
    if (L.n <= 128) return trd_one<4>(L);
    if (L.n <= 256) return trd_one<8>(L);
    if (L.n <= 512) return trd_one<16>(L);
    return trd_one<32>(L);
}

}  // namespace imago
