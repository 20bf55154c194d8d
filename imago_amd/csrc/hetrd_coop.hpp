// hetrd_coop.hpp - Householder reduction of a batch of Hermitian (complex)
// or symmetric (real) matrices to real symmetric tridiagonal form (LAPACK
// zhetd2 / dsytd2, uplo = 'L': A = Q T Q^H, Q = H(0) H(1) ... H(n-2),
// H(j) = I - tau_j v_j v_j^H) in one persistent cooperative kernel.
//
// Each matrix lives in the registers of P = ceil(n / 16) work-groups for the
// whole reduction: work-group w holds rows 16w .. 16w+15, thread t of the
// work-group rows 16w + (t % 8) and 16w + (t % 8) + 8 at the columns
// (t / 8) + 64 s, s < S / 2.  One
// column step is one exchange between the work-groups of a matrix:
//   every work-group holds the pivot row j of the current matrix in LDS and
//   builds the reflector from conj(row j) (zlarfg / dlarfg) itself,
//   bit-identically;
//   forms p = tau A v on its own rows and the partial p^H v;
//   publishes p, the partial and (the owner of row j+1) row j+1;
//   waits for p, the partials and row j+1 of the others, forms
//   w = p - tau/2 (p^H v) v, applies A -= v w^H + w v^H to its rows and
//   the same update to row j+1, the next pivot row.
// Only p, the dot partials and one row per step go through memory.  The
// tridiagonal (d, e) and tau are written by the last work-group, the
// reflectors (LAPACK's lower layout) by the work-group owning their rows; a
// work-group leaves once all its rows are reduced.
#pragma once
#include <hip/hip_runtime.h>

#include "lane_xor.hpp"

namespace imago {

constexpr int kTrdThreads = 512;
constexpr int kTrdRows = 16;
constexpr int kTrdGroups = 32;
constexpr int kTrdMaxGroups = 128;   // work-groups per matrix (order <= 2048)

// scalar helpers for double (real symmetric) and double2 (complex)
__device__ __forceinline__ double2 trd_zero() { return make_double2(0.0, 0.0); }
template <class T> __device__ __forceinline__ T sc_zero();
template <> __device__ __forceinline__ double sc_zero<double>() { return 0.0; }
template <> __device__ __forceinline__ double2 sc_zero<double2>() { return trd_zero(); }
template <class T> __device__ __forceinline__ T sc_one();
template <> __device__ __forceinline__ double sc_one<double>() { return 1.0; }
template <> __device__ __forceinline__ double2 sc_one<double2>() { return make_double2(1.0, 0.0); }
template <class T> __device__ __forceinline__ T sc_make(double re, double im);
template <> __device__ __forceinline__ double sc_make<double>(double re, double) { return re; }
template <> __device__ __forceinline__ double2 sc_make<double2>(double re, double im)
{
    return make_double2(re, im);
}
__device__ __forceinline__ double sc_re(double a) { return a; }
__device__ __forceinline__ double sc_re(double2 a) { return a.x; }
__device__ __forceinline__ double sc_im(double) { return 0.0; }
__device__ __forceinline__ double sc_im(double2 a) { return a.y; }
__device__ __forceinline__ double sc_conj(double a) { return a; }
__device__ __forceinline__ double2 sc_conj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double sc_abs2(double a) { return a * a; }
__device__ __forceinline__ double sc_abs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double sc_add(double a, double b) { return a + b; }
__device__ __forceinline__ double2 sc_add(double2 a, double2 b)
{
    return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double sc_sub(double a, double b) { return a - b; }
__device__ __forceinline__ double2 sc_sub(double2 a, double2 b)
{
    return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double sc_mul(double a, double b) { return a * b; }
__device__ __forceinline__ double2 sc_mul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double sc_mulc(double a, double b) { return a * b; }
__device__ __forceinline__ double2 sc_mulc(double2 a, double2 b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double sc_scale(double s, double a) { return s * a; }
__device__ __forceinline__ double2 sc_scale(double s, double2 a)
{
    return make_double2(s * a.x, s * a.y);
}
// acc += a * b as fused multiply-adds
__device__ __forceinline__ void sc_fma_acc(double &acc, double a, double b) { acc = fma(a, b, acc); }
__device__ __forceinline__ void sc_fma_acc(double2 &acc, double2 a, double2 b)
{
    acc.x = fma(-a.y, b.y, fma(a.x, b.x, acc.x));
    acc.y = fma(a.y, b.x, fma(a.x, b.y, acc.y));
}
// x - u conj(w) - y conj(v) as fused multiply-adds (the rank-2 update)
__device__ __forceinline__ double sc_rank2(double x, double u, double w, double y, double v)
{
    return fma(-y, v, fma(-u, w, x));
}
__device__ __forceinline__ double2 sc_rank2(double2 x, double2 u, double2 w, double2 y, double2 v)
{
    double re = fma(-u.x, w.x, x.x);
    re = fma(-u.y, w.y, re);
    re = fma(-y.x, v.x, re);
    re = fma(-y.y, v.y, re);
    double im = fma(-u.y, w.x, x.y);
    im = fma(u.x, w.y, im);
    im = fma(-y.y, v.x, im);
    im = fma(y.x, v.y, im);
    return make_double2(re, im);
}
// c ? x : y, component by component (keeps register arrays in registers)
__device__ __forceinline__ double sc_sel(bool c, double x, double y) { return c ? x : y; }
__device__ __forceinline__ double2 sc_sel(bool c, double2 x, double2 y)
{
    return make_double2(c ? x.x : y.x, c ? x.y : y.y);
}
// An empty asm that takes x as read and written: the load producing x is
// issued where it stands and not moved into a later conditional block.
__device__ __forceinline__ void sc_pin(double &x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void sc_pin(double2 &x) { asm volatile("" : "+v"(x.x), "+v"(x.y)); }
// v of lane (this lane xor O), by the VALU exchanges of lane_xor.hpp
template <int O>
__device__ __forceinline__ double sc_xor_lane(double v)
{
    return xor_lane<O>(v);
}
template <int O>
__device__ __forceinline__ double2 sc_xor_lane(double2 v)
{
    return make_double2(sc_xor_lane<O>(v.x), sc_xor_lane<O>(v.y));
}
// butterfly sums (every lane of the group ends with the same total): over
// the lanes xor 8, 16, 32 (the column groups of one row phase), over 16
// lanes, over the wave
template <class T>
__device__ __forceinline__ T sc_sum_xor_8_16_32(T v)
{
    v = sc_add(v, sc_xor_lane<8>(v));
    v = sc_add(v, sc_xor_lane<16>(v));
    return sc_add(v, sc_xor_lane<32>(v));
}
template <class T>
__device__ __forceinline__ T sc_sum16(T v)
{
    v = sc_add(v, sc_xor_lane<8>(v));
    v = sc_add(v, sc_xor_lane<4>(v));
    v = sc_add(v, sc_xor_lane<2>(v));
    return sc_add(v, sc_xor_lane<1>(v));
}
template <class T>
__device__ __forceinline__ T sc_sum64(T v)
{
    v = sc_add(v, sc_xor_lane<32>(v));
    v = sc_add(v, sc_xor_lane<16>(v));
    return sc_sum16(v);
}

// The vectors exchanged between work-groups (p, pivot rows, dot partials)
// are written and read with agent-scope atomic accesses, which are coherent
// across the XCDs' L2 caches without a cache-wide write-back or invalidate.
__device__ __forceinline__ void pub_word(double *p, double v)
{
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(p),
                       (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double pub_word_load(const double *p)
{
    const unsigned long long x = __hip_atomic_load(
        reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)x);
}
__device__ __forceinline__ void pub_store(double *p, double v) { pub_word(p, v); }
__device__ __forceinline__ void pub_store(double2 *p, double2 v)
{
    pub_word(reinterpret_cast<double *>(p), v.x);
    pub_word(reinterpret_cast<double *>(p) + 1, v.y);
}
__device__ __forceinline__ double pub_load(const double *p) { return pub_word_load(p); }
__device__ __forceinline__ double2 pub_load(const double2 *p)
{
    const double *q = reinterpret_cast<const double *>(p);
    return make_double2(pub_word_load(q), pub_word_load(q + 1));
}

// Every exchanged value has a slot of its own per column step, filled with
// kTrdEmpty before the launch: a reader waits on the words it needs until
// none holds the fill, so a value and its arrival travel together and no
// flag or barrier is needed.  kTrdEmpty is a signalling NaN; arithmetic
// only produces quiet NaNs, so no computed value equals it (an input entry
// with exactly that bit pattern would make the wait time out).
constexpr unsigned kTrdEmptyWord = 0xFFF5A5A5u;
constexpr unsigned long long kTrdEmpty = 0xFFF5A5A5FFF5A5A5ull;
__device__ __forceinline__ bool trd_word_ready(double v)
{
    return (unsigned long long)__double_as_longlong(v) != kTrdEmpty;
}
__device__ __forceinline__ bool trd_ready(double v) { return trd_word_ready(v); }
__device__ __forceinline__ bool trd_ready(double2 v)
{
    return trd_word_ready(v.x) && trd_word_ready(v.y);
}

template <class T>
struct TrdArgs {
    int n, P;             // order; work-groups per matrix (ceil(n / 16))
    int nb_launch;        // matrices of this launch
    const T *A;           // [nb][n * n] column-major, full Hermitian / symmetric
    T *V;                 // [nb][n * n] reflectors below the diagonal, zero above
    double *d, *e;        // [nb][n] diagonal, subdiagonal (e[n-1] = 0)
    T *tau;               // [nb][n] (tau[n-1] = 0)
    T *rbuf;              // [nb][n][n] step j: pivot row j+1, columns j+1 ..
    T *pbuf;              // [nb][n][n] step j: p, rows j+1 ..
    T *dbuf;              // [nb][n][kTrdMaxGroups] step j: dot partials
    unsigned *abortw;     // zero at launch; set when a wait gives up
    int *err;             // set to 1 when a wait times out
    int spin_limit;       // polls before a wait gives up
};

// wave sum of a double; every lane gets the total
__device__ __forceinline__ double trd_wave_sum(double v) { return sc_sum64(v); }

// grid = nb * P work-groups of kTrdThreads, all co-resident (cooperative
// launch); n <= 32 * S and P <= kTrdMaxGroups.  One column step j:
//   the owner of row j+1 publishes it (staged in LDS at the end of step
//   j-1, so the stores leave before the reflector is formed);
//   every work-group forms the reflector from the pivot row j in LDS and the
//   norm partials of step j-1, p = tau A v on its rows and its part of
//   p^H v, and publishes both;
//   every thread waits for its columns of p and of row j+1, wave 0 for the
//   dot partials (summed in a fixed order, the same in every work-group);
//   the rank-2 update of the work-group's rows and of row j+1, whose norm
//   partials and (its owner) row j+2 are staged for the next step.
template <class T, int S>
__global__ __launch_bounds__(kTrdThreads) void k_hetrd_coop(TrdArgs<T> A_)
{
    constexpr int NW = kTrdThreads / 64;
    constexpr int MC = (kTrdGroups * S + kTrdThreads - 1) / kTrdThreads;   // columns a thread waits on
    // a lane holds RPL rows (r, r + LR, ...) of the columns g + GC s,
    // s < S2: each v (and w) read from LDS serves RPL matrix entries.  Two
    // rows measured best: four take 6.6 ms instead of 6.2 at 4 x 896
    // complex (the update's registers), the same at order 1944 real.
    constexpr int RPL = 2, LR = kTrdRows / RPL, GC = kTrdThreads / LR;
    constexpr int S2 = S / RPL, QB = S2 < 4 ? S2 : 4;
    static_assert(S % RPL == 0 && S2 % QB == 0 && GC * S2 == kTrdGroups * S, "column batches");
    static_assert(2 * MC <= 30, "wait mask");
    __shared__ T vl[kTrdGroups * S];   // v of this step
    __shared__ T pl[kTrdGroups * S];   // p of this step
    __shared__ T rl[kTrdGroups * S];   // the pivot row
    __shared__ T stg[kTrdGroups * S];  // the next row to publish (its owner)
    __shared__ T red[NW][kTrdRows];
    __shared__ double nrm[NW];         // partial |x(2:)|^2 per wave
    __shared__ T dot;                  // p^H v
    const int n = A_.n, P = A_.P;
    const int m = blockIdx.x / P, w = blockIdx.x % P;
    const int t = threadIdx.x, r = t % LR, g = t / LR;
    const int lane = t & 63, wave = t >> 6;
    const int i0 = kTrdRows * w + r;                  // this lane's rows: i0 + LR h
    const int it = kTrdRows * w + (t & (kTrdRows - 1)); // row of thread t < 16
    // the largest column group of this wave's lanes (wave-uniform)
    const int gtop = (64 / LR) * __builtin_amdgcn_readfirstlane(wave) + 64 / LR - 1;
    const size_t nn = (size_t)n * n;
    const T *A = A_.A + m * nn;
    T *V = A_.V + m * nn;
    T *rbuf = A_.rbuf + m * nn, *pbuf = A_.pbuf + m * nn;
    T *dbuf = A_.dbuf + (size_t)m * n * kTrdMaxGroups;
    // the last work-group, which takes part in every step, writes d, e
    // and tau
    const bool out = w == P - 1;

    T a[S2][RPL];
#pragma unroll
    for (int s = 0; s < S2; ++s) {
        const int c = g + GC * s;
#pragma unroll
        for (int h = 0; h < RPL; ++h) {
            const int ih = i0 + LR * h;
            a[s][h] = (ih < n && c < n) ? A[(size_t)c * n + ih] : sc_zero<T>();
        }
    }
    // pivot row 0: A(0, c) for every column; its norm partials; row 1 staged
    double x2 = 0.0;
    for (int c = t; c < n; c += kTrdThreads) {
        const T x = A[(size_t)c * n];
        rl[c] = x;
        if (c >= 2) x2 += sc_abs2(x);
    }
    x2 = trd_wave_sum(x2);
    if (lane == 0) nrm[wave] = x2;
    if (w == 0 && r == 1 % LR) {
#pragma unroll
        for (int s = 0; s < S2; ++s) stg[g + GC * s] = a[s][1 / LR];
    }
    // The loads of a[] complete here.  Left pending on some paths, they make
    // the compiler wait for every outstanding memory access at each a[s]'s
    // first use inside the column loop, on every step; vmcnt also counts
    // stores, so the owner of the pivot row would stall there until its
    // device-scope row stores were acknowledged.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int j = 0; j + 1 < n; ++j) {
        // A work-group whose rows all lie at or before j has nothing left
        // to do: it leaves, and the others no longer wait for its (zero)
        // dot partial.
        if (kTrdRows * w + kTrdRows - 1 <= j) break;
        T *ps = pbuf + (size_t)j * n, *rs = rbuf + (size_t)j * n;
        T *ds = dbuf + (size_t)j * kTrdMaxGroups;
        if (w == ((j + 1) >> 4))
            for (int c = j + 1 + t; c < n; c += kTrdThreads) pub_store(rs + c, stg[c]);
        // ---- the reflector; every thread forms it from the same partials,
        // so all work-groups and threads hold bit-identical beta, tau, scale
        T tau, scale;
        double beta;
        {
            double xn2 = 0.0;
            for (int q = 0; q < NW; ++q) xn2 += nrm[q];
            const T alpha = sc_conj(rl[j + 1]);
            const double ar = sc_re(alpha), ai = sc_im(alpha);
            if (xn2 == 0.0 && ai == 0.0) {
                beta = ar;
                tau = sc_zero<T>();
                scale = sc_zero<T>();
            } else {
                beta = -copysign(sqrt(ar * ar + ai * ai + xn2), ar);
                // one reciprocal each for tau and 1 / (alpha - beta)
                const double rb = 1.0 / beta;
                tau = sc_make<T>((beta - ar) * rb, -ai * rb);
                const double zr = ar - beta, zi = ai;
                const double rden = 1.0 / (zr * zr + zi * zi);
                scale = sc_make<T>(zr * rden, -zi * rden);
            }
        }
        // v into LDS for the hemv, the dot and the update
        for (int c = j + 1 + t; c < n; c += kTrdThreads)
            vl[c] = c == j + 1 ? sc_one<T>() : sc_mul(scale, sc_conj(rl[c]));
        if (out && t == 0) {
            A_.tau[(size_t)m * n + j] = tau;
            A_.e[(size_t)m * n + j] = beta;
            A_.d[(size_t)m * n + j] = sc_re(rl[j]);
        }
        __syncthreads();

        // ---- p = tau A v on this work-group's rows, partial p^H v ----
        // Columns go in batches of four without branches (masked by
        // selects, so their LDS loads are in flight together); a wave skips
        // a batch whose columns all lie at or before j, a work-group whose
        // rows all do skips the product.
        bool on[RPL], any_on = false;
#pragma unroll
        for (int h = 0; h < RPL; ++h) {
            on[h] = i0 + LR * h > j && i0 + LR * h < n;
            any_on = any_on || on[h];
        }
        const bool row_on = it > j && it < n;   // thread t < 16's row
        const bool wg_on = kTrdRows * w + kTrdRows - 1 > j;
        T sp[RPL];
#pragma unroll
        for (int h = 0; h < RPL; ++h) sp[h] = sc_zero<T>();
        if (wg_on) {
#pragma unroll
            for (int s0 = 0; s0 < S2; s0 += QB) {
                if (gtop + GC * (s0 + QB - 1) > j) {
                    T vc[QB];
#pragma unroll
                    for (int q = 0; q < QB; ++q) vc[q] = vl[g + GC * (s0 + q)];
#pragma unroll
                    for (int q = 0; q < QB; ++q) sc_pin(vc[q]);
#pragma unroll
                    for (int q = 0; q < QB; ++q) {
                        const int c = g + GC * (s0 + q);
                        vc[q] = sc_sel(c > j && c < n, vc[q], sc_zero<T>());
#pragma unroll
                        for (int h = 0; h < RPL; ++h) sc_fma_acc(sp[h], a[s0 + q][h], vc[q]);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // sums over the column groups of the wave (lanes xor 8, 16, 32),
        // each of the lane's rows
        static_assert(LR == 8, "column groups of a row phase");
#pragma unroll
        for (int h = 0; h < RPL; ++h) {
            sp[h] = on[h] ? sp[h] : sc_zero<T>();
            sp[h] = sc_sum_xor_8_16_32(sp[h]);
        }
        if (lane < LR) {
#pragma unroll
            for (int h = 0; h < RPL; ++h) red[wave][lane + LR * h] = sp[h];
        }
        __syncthreads();
        if (t < kTrdRows) {
            T p = sc_zero<T>();
            for (int q = 0; q < NW; ++q) p = sc_add(p, red[q][t]);
            p = row_on ? sc_mul(tau, p) : sc_zero<T>();
            if (row_on) pub_store(ps + it, p);
            // conj(p_i) v_i over the 16 rows, the same tree in every work-group
            T dp = row_on ? sc_mulc(vl[it], p) : sc_zero<T>();
            dp = sc_sum16(dp);
            if (t == 0) pub_store(ds + w, dp);
        }

        // ---- wait for p and row j+1 (every thread its columns) and the
        // dot partials (wave 0): all loads of a round in flight together
        {
            unsigned pend = 0;
#pragma unroll
            for (int k = 0; k < MC; ++k)
                if (j + 1 + t + kTrdThreads * k < n) pend |= 3u << (2 * k);
            if (wave == 0) {
                const int first = (j + 1) / kTrdRows;   // the first work-group still active
                if (lane < P && lane >= first) pend |= 1u << 30;
                if (lane + 64 < P && lane + 64 >= first) pend |= 1u << 31;
            }
            T d0 = sc_zero<T>(), d1 = sc_zero<T>();
            int polls = 0;
            while (pend) {
                T gp[MC], gr[MC], g0 = sc_zero<T>(), g1 = sc_zero<T>();
#pragma unroll
                for (int k = 0; k < MC; ++k) {
                    const int c = j + 1 + t + kTrdThreads * k;
                    if (pend & (1u << (2 * k))) gp[k] = pub_load(ps + c);
                    if (pend & (2u << (2 * k))) gr[k] = pub_load(rs + c);
                }
                if (pend & (1u << 30)) g0 = pub_load(ds + lane);
                if (pend & (1u << 31)) g1 = pub_load(ds + lane + 64);
#pragma unroll
                for (int k = 0; k < MC; ++k) {
                    const int c = j + 1 + t + kTrdThreads * k;
                    if ((pend & (1u << (2 * k))) && trd_ready(gp[k])) {
                        pl[c] = gp[k];
                        pend &= ~(1u << (2 * k));
                    }
                    if ((pend & (2u << (2 * k))) && trd_ready(gr[k])) {
                        rl[c] = gr[k];
                        pend &= ~(2u << (2 * k));
                    }
                }
                if ((pend & (1u << 30)) && trd_ready(g0)) {
                    d0 = g0;
                    pend &= ~(1u << 30);
                }
                if ((pend & (1u << 31)) && trd_ready(g1)) {
                    d1 = g1;
                    pend &= ~(1u << 31);
                }
                if (!pend) break;
                if ((++polls & 63) == 0) {
                    if (__hip_atomic_load(A_.abortw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
                        break;
                    if (polls > A_.spin_limit) {
                        __hip_atomic_store(A_.abortw, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_fetch_or(A_.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
                __builtin_amdgcn_s_sleep(1);
            }
            // Every path out of the loop has waited for its loads; saying so
            // keeps the compiler from waiting again (on this step's stores
            // too) before each reuse of their registers in the next step.
            __builtin_amdgcn_s_waitcnt(0);
            if (wave == 0) {
                // partials of work-groups lane and lane + 64, then a fixed tree
                T x = sc_add(d0, d1);
                x = sc_sum64(x);
                if (lane == 0) dot = x;
            }
        }
        __syncthreads();

        // ---- w = p - tau/2 (p^H v) v in place of p; A -= v w^H + w v^H;
        // next pivot row ----
        const T a2 = sc_mul(sc_scale(-0.5, tau), dot);
        for (int c = j + 1 + t; c < n; c += kTrdThreads) pl[c] = sc_add(pl[c], sc_mul(a2, vl[c]));
        // Each work-group writes the part of V on its own rows here, a step
        // before the next wait on its memory accesses.
        if (t < kTrdRows && row_on) V[(size_t)j * n + it] = vl[it];
        __syncthreads();
        if (wg_on && any_on) {
            // rows at or before j keep their values (the selects); their v
            // and w entries may be stale or unset
            T vr[RPL], wr[RPL];
#pragma unroll
            for (int h = 0; h < RPL; ++h) {
                const int ih = i0 + LR * h < n ? i0 + LR * h : 0;
                vr[h] = vl[ih];
                wr[h] = pl[ih];
            }
#pragma unroll
            for (int s0 = 0; s0 < S2; s0 += QB) {
                if (gtop + GC * (s0 + QB - 1) > j) {
                    T vc[QB], wc[QB];
#pragma unroll
                    for (int q = 0; q < QB; ++q) {
                        const int c = g + GC * (s0 + q);
                        vc[q] = vl[c];
                        wc[q] = pl[c];
                    }
#pragma unroll
                    for (int q = 0; q < QB; ++q) {
                        sc_pin(vc[q]);
                        sc_pin(wc[q]);
                    }
#pragma unroll
                    for (int q = 0; q < QB; ++q) {
                        const int c = g + GC * (s0 + q);
                        const bool col = c > j && c < n;
#pragma unroll
                        for (int h = 0; h < RPL; ++h) {
                            const T x = sc_rank2(a[s0 + q][h], vr[h], wc[q], wr[h], vc[q]);
                            a[s0 + q][h] = sc_sel(col && on[h], x, a[s0 + q][h]);
                        }
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // the owner of row j+2 stages it for the next step's publication
        // (columns up to j+1 are staged too and never published)
        {
            const int rj = (j + 2) & (kTrdRows - 1);
            if (w == ((j + 2) >> 4) && r == rj % LR && j + 2 < n) {
                const int hr = rj / LR;
#pragma unroll
                for (int s = 0; s < S2; ++s) {
                    T x = a[s][0];
#pragma unroll
                    for (int h = 1; h < RPL; ++h) x = sc_sel(hr == h, a[s][h], x);
                    stg[g + GC * s] = x;
                }
            }
        }
        // row j+1 of the updated matrix (v_{j+1} = 1): the next pivot row,
        // and the partials of its norm beyond the next subdiagonal
        {
            const T wj1 = pl[j + 1];
            double y2 = 0.0;
            for (int c = j + 1 + t; c < n; c += kTrdThreads) {
                const T x = sc_rank2(rl[c], sc_one<T>(), pl[c], wj1, vl[c]);
                rl[c] = x;
                if (c >= j + 3) y2 += sc_abs2(x);
            }
            y2 = trd_wave_sum(y2);
            if (lane == 0) nrm[wave] = y2;
        }
        __syncthreads();
    }
    if (out && t == 0) {
        A_.d[(size_t)m * n + n - 1] = sc_re(rl[n - 1]);
        A_.e[(size_t)m * n + n - 1] = 0.0;
        A_.tau[(size_t)m * n + n - 1] = sc_zero<T>();
    }
}

}  // namespace imago
