// hetrd.hip - launcher of the cooperative tridiagonal reduction
// (hetrd_coop.hpp).
#include "hetrd_coop.hpp"
#include "hetrd_launch.hpp"

namespace imago {

namespace {

template <class T, int S>
hipError_t trd_one(const TrdLaunch &L)
{
    const void *fn = reinterpret_cast<const void *>(&k_hetrd_coop<T, S>);
    int dev = 0, ncu = 0, per_cu = 0;
    hipError_t st = hipGetDevice(&dev);
    if (st != hipSuccess) return st;
    st = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (st != hipSuccess) return st;
    st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kTrdThreads, 0);
    if (st != hipSuccess) return st;
    const int P = (L.n + kTrdRows - 1) / kTrdRows;
    if (P > kTrdMaxGroups) return hipErrorInvalidValue;
    const int fit = per_cu * ncu / P;   // matrices whose work-groups are co-resident
    if (fit < 1) return hipErrorCooperativeLaunchTooLarge;
    const size_t nn = (size_t)L.n * L.n;
    for (int m0 = 0; m0 < L.nb; m0 += fit) {
        const int nb = std::min(fit, L.nb - m0);
        // every exchange slot empty, the abort word clear
        const size_t words = sizeof(T) / 4;
        st = hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(L.rbuf), (int)kTrdEmptyWord,
                               nb * nn * words, L.stream);
        if (st == hipSuccess)
            st = hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(L.pbuf), (int)kTrdEmptyWord,
                                   nb * nn * words, L.stream);
        if (st == hipSuccess)
            st = hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(L.dbuf), (int)kTrdEmptyWord,
                                   (size_t)nb * L.n * kTrdMaxGroups * words, L.stream);
        if (st == hipSuccess) st = hipMemsetAsync(L.abortw, 0, sizeof(unsigned), L.stream);
        if (st != hipSuccess) return st;
        TrdArgs<T> A;
        A.n = L.n;
        A.P = P;
        A.nb_launch = nb;
        A.A = static_cast<const T *>(L.A) + m0 * nn;
        A.V = static_cast<T *>(L.V) + m0 * nn;
        A.d = L.d + (size_t)m0 * L.n;
        A.e = L.e + (size_t)m0 * L.n;
        A.tau = static_cast<T *>(L.tau) + (size_t)m0 * L.n;
        A.rbuf = static_cast<T *>(L.rbuf);
        A.pbuf = static_cast<T *>(L.pbuf);
        A.dbuf = static_cast<T *>(L.dbuf);
        A.abortw = L.abortw;
        A.err = L.err;
        A.spin_limit = 2000000;
        void *args[] = {&A};
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
// 1024 threads, an exclusive scan of the flags of each block in turn.  A
// block that does not own exactly n vectors (N = nb * n) sets bit 2 of err:
// its matrix would keep stale columns.
__global__ __launch_bounds__(1024) void k_trd_rank(const int *__restrict__ owner, int N, int nb,
                                                   int *__restrict__ rank, int *__restrict__ err)
{
    const int n = N / nb;
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
        if (t == 1023 && sc[1023] != n)
            __hip_atomic_fetch_or(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
    }
}

// one work-group per eigenvector q: its block's part as column rank[q] of
// X[owner[q]] (complex with zero imaginary part, or real), its eigenvalue
// to w[owner[q]][rank[q]]
template <class T>
__global__ __launch_bounds__(256) void k_trd_gather(const double *__restrict__ Z,
                                                    const double *__restrict__ wall, int N, int n,
                                                    const int *__restrict__ owner,
                                                    const int *__restrict__ rank,
                                                    double *__restrict__ w, T *__restrict__ X)
{
    const int q = blockIdx.x;
    const int b = owner[q], p = rank[q];
    if (p >= n) return;
    const double *z = Z + (size_t)q * N + (size_t)b * n;
    T *x = X + (size_t)b * n * n + (size_t)p * n;
    for (int r = threadIdx.x; r < n; r += 256) x[r] = sc_make<T>(z[r], 0.0);
    if (threadIdx.x == 0) w[(size_t)b * n + p] = wall[q];
}

hipError_t launch_trd_unmix(const double *Z, const double *wall, int n, int nb, int b0,
                            int *owner, int *rank, double *w, void *X, bool real, int *err,
                            hipStream_t stream)
{
    const int N = n * nb;
    hipLaunchKernelGGL(k_trd_owner, dim3(N), dim3(256), 0, stream, Z, N, n, owner);
    hipLaunchKernelGGL(k_trd_rank, dim3(1), dim3(1024), 0, stream, owner, N, nb, rank, err);
    if (real)
        hipLaunchKernelGGL(k_trd_gather<double>, dim3(N), dim3(256), 0, stream, Z, wall, N, n,
                           owner, rank, w + (size_t)b0 * n,
                           static_cast<double *>(X) + (size_t)b0 * n * n);
    else
        hipLaunchKernelGGL(k_trd_gather<double2>, dim3(N), dim3(256), 0, stream, Z, wall, N, n,
                           owner, rank, w + (size_t)b0 * n,
                           static_cast<double2 *>(X) + (size_t)b0 * n * n);
    return hipGetLastError();
}

// zlarft / dlarft (direct = 'F', storev = 'C') for block b of matrix m,
// one work-group each.  On entry T[m][b] holds G = V_b^H V_b (the caller's
// batched GEMM over all n rows: V is zero above each reflector's unit
// entry); the strict upper triangle of G is read into LDS and T built in
// place of it: T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i), T(i, i) = tau_i.
template <class T>
__global__ __launch_bounds__(256) void k_trd_larft(const T *__restrict__ tau, int n, int nblocks,
                                                   T *__restrict__ TG)
{
    constexpr int K = kTrdBlock;
    __shared__ T G[K][K + 1];
    __shared__ T Ts[K][K + 1];
    const int m = blockIdx.x / nblocks, b = blockIdx.x % nblocks;
    const int j0 = b * K, nbk = min(K, n - 1 - j0);
    const T *tm = tau + (size_t)m * n;
    T *Tg = TG + ((size_t)m * nblocks + b) * K * K;
    for (int e = threadIdx.x; e < K * K; e += 256) {
        const int kr = e % K, ic = e / K;
        G[kr][ic] = kr < ic ? Tg[e] : sc_zero<T>();
        Ts[kr][ic] = sc_zero<T>();
    }
    __syncthreads();
    for (int i = 0; i < nbk; ++i) {
        const T ti = tm[j0 + i];
        if ((int)threadIdx.x < i) {
            // -tau_i * sum_k T(row, k) G(k, i), k in [row, i)
            const int row = threadIdx.x;
            T s = sc_zero<T>();
            for (int k = row; k < i; ++k) s = sc_add(s, sc_mul(Ts[row][k], G[k][i]));
            Ts[row][i] = sc_scale(-1.0, sc_mul(ti, s));
        }
        if (threadIdx.x == 0) Ts[i][i] = ti;
        __syncthreads();
    }
    for (int e = threadIdx.x; e < K * K; e += 256) Tg[e] = Ts[e % K][e / K];
}

// The UT form of Q = H(0) ... H(n-2) = I - V T V^H for all n - 1
// reflectors at once: T^-1 = diag(1 / tau) + the strict upper triangle of
// G = V^H V (the zlarft recurrence T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i),
// T(i, i) = tau_i, inverted column by column).  A reflector with tau = 0 is
// the identity (zhetd2 leaves v = e_{j+1} then): its column of V is zeroed
// first and its diagonal entry of T^-1 set to 1, so it adds nothing.
template <class T>
__global__ __launch_bounds__(256) void k_ut_null(const T *__restrict__ tau, int n,
                                                 T *__restrict__ V)
{
    const int m = blockIdx.y, j = blockIdx.x;
    const T t = tau[(size_t)m * n + j];
    if (sc_re(t) != 0.0 || sc_im(t) != 0.0) return;
    T *col = V + (size_t)m * n * n + (size_t)j * n;
    for (int r = threadIdx.x; r < n; r += 256) col[r] = sc_zero<T>();
}

// in place of G (order n - 1, leading dimension n): the strict upper
// triangle kept, 1 / tau on the diagonal, zero below
template <class T>
__global__ __launch_bounds__(256) void k_ut_tinv(const T *__restrict__ tau, int n,
                                                 T *__restrict__ G)
{
    const int m = blockIdx.y, n1 = n - 1;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (size_t)n1 * n1) return;
    const int i = (int)(e % n1), j = (int)(e / n1);
    if (j > i) return;
    T *g = G + (size_t)m * n * n + (size_t)j * n + i;
    if (j < i) {
        *g = sc_zero<T>();
        return;
    }
    const T t = tau[(size_t)m * n + i];
    const double tr = sc_re(t), ti = sc_im(t), d = tr * tr + ti * ti;
    *g = d == 0.0 ? sc_one<T>() : sc_make<T>(tr / d, -ti / d);
}

hipError_t launch_ut_null(const void *tau, int n, int nb, bool real, void *V, hipStream_t stream)
{
    if (n < 2 || nb <= 0) return hipSuccess;
    const dim3 grid((unsigned)(n - 1), (unsigned)nb);
    if (real)
        hipLaunchKernelGGL(k_ut_null<double>, grid, dim3(256), 0, stream,
                           static_cast<const double *>(tau), n, static_cast<double *>(V));
    else
        hipLaunchKernelGGL(k_ut_null<double2>, grid, dim3(256), 0, stream,
                           static_cast<const double2 *>(tau), n, static_cast<double2 *>(V));
    return hipGetLastError();
}

hipError_t launch_ut_tinv(const void *tau, int n, int nb, bool real, void *G, hipStream_t stream)
{
    if (n < 2 || nb <= 0) return hipSuccess;
    const size_t n1 = (size_t)(n - 1);
    const dim3 grid((unsigned)((n1 * n1 + 255) / 256), (unsigned)nb);
    if (real)
        hipLaunchKernelGGL(k_ut_tinv<double>, grid, dim3(256), 0, stream,
                           static_cast<const double *>(tau), n, static_cast<double *>(G));
    else
        hipLaunchKernelGGL(k_ut_tinv<double2>, grid, dim3(256), 0, stream,
                           static_cast<const double2 *>(tau), n, static_cast<double2 *>(G));
    return hipGetLastError();
}

hipError_t launch_trd_larft(const void *tau, int n, int nb, bool real, void *T, hipStream_t stream)
{
    const int nblocks = (n - 1 + kTrdBlock - 1) / kTrdBlock;
    if (nblocks <= 0) return hipSuccess;
    if (real)
        hipLaunchKernelGGL(k_trd_larft<double>, dim3((unsigned)(nb * nblocks)), dim3(256), 0,
                           stream, static_cast<const double *>(tau), n, nblocks,
                           static_cast<double *>(T));
    else
        hipLaunchKernelGGL(k_trd_larft<double2>, dim3((unsigned)(nb * nblocks)), dim3(256), 0,
                           stream, static_cast<const double2 *>(tau), n, nblocks,
                           static_cast<double2 *>(T));
    return hipGetLastError();
}

// zpotf2 / dpotf2 (uplo = 'U') of the diagonal block A(j:j+jb, j:j+jb) of
// every matrix, one work-group each, the block in LDS: for each column k
// the pivot d = sqrt(a(k,k)), row k of U scaled by 1/d, then the trailing
// block updated, a(r,c) -= conj(u(k,r)) u(k,c) for k < r <= c.  The first
// pivot that is not positive (or NaN) sets info[m] = j + k + 1 unless an
// earlier block already set it; the launch with j = 0 clears info first.
template <class T>
__global__ __launch_bounds__(256) void k_potf2_diag(T *__restrict__ A, int n, size_t stride, int j,
                                                    int jb, int *__restrict__ info)
{
    constexpr int B = kPotrfBlock;
    constexpr int NP = 256 / B;   // row phases: a thread updates rows q + NP i of its column
    static_assert(256 % B == 0 && B <= 64, "columns of the block over the work-group");
    __shared__ T a[B][B + 1];   // a[c][r] = A(j + r, j + c)
    __shared__ int first;
    const int t = threadIdx.x, m = blockIdx.x;
    T *Am = A + m * stride;
    for (int e = t; e < jb * jb; e += 256) {
        const int r = e % jb, c = e / jb;
        a[c][r] = r <= c ? Am[(size_t)(j + c) * n + j + r] : sc_zero<T>();
    }
    if (t == 0) first = j == 0 ? 0 : info[m];
    __syncthreads();
    // thread t updates column c = t % B at the rows q + NP i, q = t / B
    const int c = t % B, q = t / B;
    for (int k = 0; k < jb; ++k) {
        const double akk = sc_re(a[k][k]);
        const double d = sqrt(akk);
        const double rd = 1.0 / d;
        __syncthreads();
        if (t == 0) {
            if (!(akk > 0.0) && first == 0) first = j + k + 1;
            a[k][k] = sc_make<T>(d, 0.0);
        }
        if (t > k && t < jb) a[t][k] = sc_scale(rd, a[t][k]);
        __syncthreads();
        if (c > k && c < jb) {
            const T ukc = a[c][k];
#pragma unroll
            for (int i = 0; i < B / NP; ++i) {
                const int r = q + NP * i;
                if (r > k && r <= c) a[c][r] = sc_sub(a[c][r], sc_mul(sc_conj(a[r][k]), ukc));
            }
        }
        __syncthreads();
    }
    for (int e = t; e < jb * jb; e += 256) {
        const int r = e % jb, c = e / jb;
        if (r <= c) Am[(size_t)(j + c) * n + j + r] = a[c][r];
    }
    if (t == 0) info[m] = first;
}

hipError_t launch_potf2_diag(void *A, int n, int nb, int j, int jb, bool real, int *info,
                             hipStream_t stream)
{
    if (jb < 1 || jb > kPotrfBlock || j < 0 || j + jb > n) return hipErrorInvalidValue;
    const size_t stride = (size_t)n * n;
    if (real)
        hipLaunchKernelGGL(k_potf2_diag<double>, dim3((unsigned)nb), dim3(256), 0, stream,
                           static_cast<double *>(A), n, stride, j, jb, info);
    else
        hipLaunchKernelGGL(k_potf2_diag<double2>, dim3((unsigned)nb), dim3(256), 0, stream,
                           static_cast<double2 *>(A), n, stride, j, jb, info);
    return hipGetLastError();
}

bool hetrd_coop_supported(int n, bool real)
{
    return n >= 2 && n <= (real ? kTrdMaxOrderReal : kTrdMaxOrderComplex);
}

hipError_t launch_hetrd_coop(const TrdLaunch &L)
{
    if (!hetrd_coop_supported(L.n, L.real) || L.nb < 1) return hipErrorInvalidValue;
    if (L.real) {
        if (L.n <= 256) return trd_one<double, 8>(L);
        if (L.n <= 512) return trd_one<double, 16>(L);
        if (L.n <= 1024) return trd_one<double, 32>(L);
        return trd_one<double, 64>(L);
    }
    if (L.n <= 128) return trd_one<double2, 4>(L);
    if (L.n <= 256) return trd_one<double2, 8>(L);
    if (L.n <= 512) return trd_one<double2, 16>(L);
    return trd_one<double2, 32>(L);
}

}  // namespace imago
