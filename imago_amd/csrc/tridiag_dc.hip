// tridiag_dc.hip - batched divide and conquer for the real symmetric
// tridiagonal eigenproblems left by the reduction of the transformed
// Hamiltonians (the tridiagonal stage of the zhegv / dsygv behind
// solveZHEGV / solveDSYGV, src/imago/interfaces.F90:73-200).
//
// The numerics are in tridiag_dc.hpp.  Each height of the merge tree is
// five launches over every merge of every matrix of the batch:
//   k_dc_deflate   one work-group per merge: z from the children's
//                  eigenvectors, the merged sort, the deflation scan;
//   k_dc_secular   one wave per eigenpair: its secular root, the sums
//                  strided over the 64 lanes;
//   k_dc_zhat      one wave per eigenpair: Loewner's z and the rank of the
//                  eigenvalue among the merge's;
//   k_dc_vectors   one wave per eigenpair: its column of U in LDS, the
//                  deflating rotations, the column scattered unsorted;
//   k_dc_gemm      64 x 64 tiles of blockdiag(Q_L, Q_R) U on FP64 MFMA
//                  (v_mfma_f64_16x16x4f64), tiles of all merges in one grid.
// The children of a merge of height h have heights h-1 and h-2 at most, so
// three eigenvector buffers indexed by height mod 3 never overlap a merge's
// inputs with its output.
#include "lane_xor.hpp"
#include "tridiag_dc.hpp"
#include "tridiag_dc_launch.hpp"

#include <algorithm>

namespace imago {

using namespace dc;

namespace {

// The device lane policy: sums strided over a wave of 64, reduced by
// butterfly and broadcast from lane 0 so every lane holds the same bits.
struct Wave {
    __device__ int lane() const { return threadIdx.x & 63; }
    __device__ int width() const { return 64; }
    __device__ double sum(double x) const { return bcast0(wave_sum_xor(x)); }
    __device__ double prod(double x) const { return bcast0(wave_prod_xor(x)); }
    static __device__ double bcast0(double x)
    {
        const int lo = __builtin_amdgcn_readlane(__double2loint(x), 0);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 0);
        return __hiloint2double(hi, lo);
    }
};

struct MergeRec {
    int r0, n1, n2, hL, hR, h;
};

__device__ __forceinline__ MergeRec load_merge(const int *p)
{
    return MergeRec{p[0], p[1], p[2], p[3], p[4], p[5]};
}

// per problem: the scale max(|d|, |e|), the torn leaves d'_i and their
// eigenvectors [1] on the diagonal of buffer 0
__global__ __launch_bounds__(256) void k_dc_init(const double *__restrict__ d,
                                                 const double *__restrict__ e, int n,
                                                 double *__restrict__ scale,
                                                 double *__restrict__ w0,
                                                 double *__restrict__ Z0)
{
    __shared__ double red[256];
    const int b = blockIdx.x;
    const double *db = d + (size_t)b * n, *eb = e + (size_t)b * n;
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        mx = fmax(mx, fabs(db[i]));
        if (i < n - 1) mx = fmax(mx, fabs(eb[i]));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double sc = red[0] > 0.0 ? red[0] : 1.0;
    if (threadIdx.x == 0) scale[b] = sc;
    for (int i = threadIdx.x; i < n; i += 256) {
        double v = db[i] / sc;
        if (i > 0) v -= fabs(eb[i - 1] / sc);
        if (i < n - 1) v -= fabs(eb[i] / sc);
        w0[(size_t)b * n + i] = v;
        Z0[(size_t)b * n * n + (size_t)i * n + i] = 1.0;
    }
}

// deflate_scan on one wave with the same decisions and the same output
// order, 64 entries a round.  Each lane takes one entry: the |z| test, its
// predecessor (the last earlier entry that survived it, or the pending one
// carried from the last round) and the neighbour test against it.  All
// entries up to the first lane whose neighbour test deflates are final; that
// rotation changes the entry the next test reads, so the round ends there
// and the next one starts behind it.  Output slots come from prefix counts of
// each kind of event in lane order, as the sequential scan appends them.
__device__ int deflate_scan_wave(int m, double rho, double tol, double *Ds, double *zs,
                                 double *dl, double *zz, int *kpos, int *dpos, double *dval,
                                 int *rp, int *rq, double *rc, double *rs, int *nrot_out)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    int K = 0, nd = 0, nrot = 0, pj = -1;
    for (int t0 = 0; t0 < m;) {
        const int t = t0 + lane;
        const bool in = t < m;
        const double z = in ? zs[t] : 0.0, D = in ? Ds[t] : 0.0;
        const bool f1 = in && rho * fabs(z) <= tol;
        const bool cand = in && !f1;
        const unsigned long long cm = __ballot(cand);
        const unsigned long long prev = cm & below;
        const int pred = prev ? t0 + 63 - __clzll((long long)prev) : pj;
        bool rot = false;
        double c = 0.0, sn = 0.0, tau = 0.0, Dp = 0.0, zp = 0.0;
        if (cand && pred >= 0) {
            zp = zs[pred];
            Dp = Ds[pred];
            rot = pair_deflates(zp, Dp, z, D, tol, &c, &sn, &tau);
        }
        const unsigned long long rm = __ballot(rot);
        const int f = rm ? __ffsll((long long)rm) - 1 : 64;
        const bool act = in && lane <= f;
        const bool ev_def = act && (f1 || lane == f);
        const bool ev_keep = act && cand && pred >= 0 && lane != f;
        const unsigned long long dm = __ballot(ev_def), km = __ballot(ev_keep);
        if (ev_keep) {
            const int k = K + __popcll(km & below);
            dl[k] = Dp;
            zz[k] = zp;
            kpos[k] = pred;
        }
        if (ev_def) {
            const int d = nd + __popcll(dm & below);
            if (f1) {
                dpos[d] = t;
                dval[d] = D;
            } else {
                const double dp = Dp * c * c + D * sn * sn;
                zs[t] = tau;
                zs[pred] = 0.0;
                Ds[t] = Dp * sn * sn + D * c * c;
                Ds[pred] = dp;
                rp[nrot] = pred;
                rq[nrot] = t;
                rc[nrot] = c;
                rs[nrot] = sn;
                dpos[d] = pred;
                dval[d] = dp;
            }
        }
        K += __popcll(km);
        nd += __popcll(dm);
        if (rm) {
            ++nrot;
            pj = t0 + f;
            t0 += f + 1;
        } else {
            if (cm) pj = t0 + 63 - __clzll((long long)cm);
            t0 += 64;
        }
        // the rotated entries are read by other lanes in the next round
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (pj >= 0 && lane == 0) {
        dl[K] = Ds[pj];
        zz[K] = zs[pj];
        kpos[K] = pj;
    }
    if (pj >= 0) ++K;
    *nrot_out = nrot;
    return K;
}

// one work-group per merge: z, the sort of the children's eigenvalues, the
// deflation scan (wave 0)
__global__ __launch_bounds__(256) void k_dc_deflate(const int *__restrict__ merges, int nmerge,
                                                    int n, const double *__restrict__ e,
                                                    DcWork w)
{
    extern __shared__ double lds[];
    __shared__ double red[2][256];
    const MergeRec mg = load_merge(merges + 6 * (blockIdx.x % nmerge));
    const int b = blockIdx.x / nmerge;
    const int m = mg.n1 + mg.n2, s = mg.r0 + mg.n1;
    double *Ds = lds, *zs = lds + m;
    const size_t vb = (size_t)b * n, mb = (size_t)b * n * n, base = vb + mg.r0;
    const double *wL = w.wb[mg.hL % 3] + vb, *wR = w.wb[mg.hR % 3] + vb;
    const double *ZL = w.Zb[mg.hL % 3] + mb, *ZR = w.Zb[mg.hR % 3] + mb;
    const double beta = e[vb + s - 1] / w.scale[b];
    const double sigma = beta >= 0.0 ? 1.0 : -1.0;
    const double rsq = 0.70710678118654752440;
    double dmax = 0.0, zmax = 0.0;
    for (int x = threadIdx.x; x < m; x += 256) {
        const bool left = x < mg.n1;
        const double D = left ? wL[mg.r0 + x] : wR[mg.r0 + x];
        const double z = (left ? ZL[(size_t)(mg.r0 + x) * n + s - 1]
                               : sigma * ZR[(size_t)(mg.r0 + x) * n + s]) * rsq;
        // position in the merged order: left entries before equal right ones
        int lo = 0, hi;
        if (left) {
            const double *o = wR + s;
            hi = mg.n2;
            while (lo < hi) {
                const int md = (lo + hi) >> 1;
                if (o[md] < D) lo = md + 1; else hi = md;
            }
        } else {
            const double *o = wL + mg.r0;
            hi = mg.n1;
            while (lo < hi) {
                const int md = (lo + hi) >> 1;
                if (o[md] <= D) lo = md + 1; else hi = md;
            }
        }
        const int pos = (left ? x : x - mg.n1) + lo;
        Ds[pos] = D;
        zs[pos] = z;
        w.perm[base + pos] = x;
        dmax = fmax(dmax, fabs(D));
        zmax = fmax(zmax, fabs(z));
    }
    red[0][threadIdx.x] = dmax;
    red[1][threadIdx.x] = zmax;
    __syncthreads();
    for (int q = 128; q > 0; q >>= 1) {
        if ((int)threadIdx.x < q) {
            red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + q]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + q]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 64) {
        const double tol = 8.0 * kEps * fmax(red[0][0], red[1][0]);
        const double rho = 2.0 * fabs(beta);
        int nrot = 0;
        const int K = deflate_scan_wave(m, rho, tol, Ds, zs, w.dl + base, w.zz + base,
                                        w.kpos + base, w.dpos + base, w.dv + base, w.rp + base,
                                        w.rq + base, w.rc + base, w.rs + base, &nrot);
        if (threadIdx.x == 0) {
            w.K[base] = K;
            w.nrot[base] = nrot;
            w.rho[base] = rho;
        }
    }
}

__device__ __forceinline__ bool dc_item(const int *items, int nitem, const int *merges,
                                        int wave, int n, MergeRec *mg, int *x, size_t *base,
                                        int *b)
{
    const int it = wave % nitem;
    *b = wave / nitem;
    *mg = load_merge(merges + 6 * items[2 * it]);
    *x = items[2 * it + 1];
    *base = (size_t)(*b) * n + mg->r0;
    return true;
}

// one wave per eigenpair: the secular root of item x (x < K)
__global__ __launch_bounds__(256) void k_dc_secular(const int *__restrict__ merges,
                                                    const int *__restrict__ items, int nitem,
                                                    int nwaves, int n, DcWork w)
{
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= nwaves) return;
    MergeRec mg;
    int x, b;
    size_t base;
    dc_item(items, nitem, merges, wave, n, &mg, &x, &base, &b);
    const int K = w.K[base];
    if (x >= K) return;
    int o;
    double tau;
    Wave ln;
    secular_root(ln, x, K, w.dl + base, w.zz + base, w.rho[base], &o, &tau);
    if (ln.lane() == 0) {
        w.org[base + x] = o;
        w.tau[base + x] = tau;
        w.val[base + x] = w.dl[base + o] + tau;
    }
}

// one wave per eigenpair: Loewner's z_x (x < K) and the rank of the
// eigenvalue of item x among the merge's m (ties by item order)
__global__ __launch_bounds__(256) void k_dc_zhat(const int *__restrict__ merges,
                                                 const int *__restrict__ items, int nitem,
                                                 int nwaves, int n, DcWork w)
{
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= nwaves) return;
    MergeRec mg;
    int x, b;
    size_t base;
    dc_item(items, nitem, merges, wave, n, &mg, &x, &base, &b);
    const int K = w.K[base], m = mg.n1 + mg.n2;
    Wave ln;
    if (x < K) {
        const double z = loewner_z(ln, x, K, w.dl + base, w.zz + base, w.org + base,
                                   w.tau + base, w.rho[base]);
        if (ln.lane() == 0) w.zh[base + x] = z;
    }
    const double vx = x < K ? w.val[base + x] : w.dv[base + x - K];
    double c = 0.0;
    for (int y = ln.lane(); y < m; y += 64) {
        const double vy = y < K ? w.val[base + y] : w.dv[base + y - K];
        c += (vy < vx || (vy == vx && y < x)) ? 1.0 : 0.0;
    }
    c = ln.sum(c);
    if (ln.lane() == 0) w.rank[base + x] = (int)c;
}

// one wave (the whole work-group) per eigenpair: its column of U in the
// sorted basis, the deflating rotations, the column written unsorted to
// column r0 + rank of the merge's block of Vu; the merged eigenvalue to the
// merge's height buffer
__global__ __launch_bounds__(64) void k_dc_vectors(const int *__restrict__ merges,
                                                   const int *__restrict__ items, int nitem,
                                                   int n, DcWork w)
{
    extern __shared__ double col[];
    MergeRec mg;
    int x, b;
    size_t base;
    dc_item(items, nitem, merges, blockIdx.x, n, &mg, &x, &base, &b);
    const int K = w.K[base], m = mg.n1 + mg.n2, lane = threadIdx.x;
    for (int t = lane; t < m; t += 64) col[t] = 0.0;
    __syncthreads();
    Wave ln;
    if (x < K) {
        const int o = w.org[base + x];
        const double tau = w.tau[base + x], dlo = w.dl[base + o];
        double nrm = 0.0;
        for (int k = lane; k < K; k += 64) {
            const double u = w.zh[base + k] / ((w.dl[base + k] - dlo) - tau);
            col[w.kpos[base + k]] = u;
            nrm += u * u;
        }
        nrm = 1.0 / sqrt(ln.sum(nrm));
        for (int k = lane; k < K; k += 64) col[w.kpos[base + k]] *= nrm;
    } else if (lane == 0) {
        col[w.dpos[base + x - K]] = 1.0;
    }
    __syncthreads();
    const int nrot = w.nrot[base];
    if (nrot > 0 && lane == 0)
        apply_rotations(col, nrot, w.rp + base, w.rq + base, w.rc + base, w.rs + base);
    __syncthreads();
    const int r = w.rank[base + x];
    double *dst = w.Vu + (size_t)b * n * n + (size_t)(mg.r0 + r) * n + mg.r0;
    for (int t = lane; t < m; t += 64) dst[w.perm[base + t]] = col[t];
    if (lane == 0)
        w.wb[mg.h % 3][base + r] = x < K ? w.val[base + x] : w.dv[base + x - K];
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// C = A B for the jobs of one height (column-major, leading dimension n):
// one 64 x 64 tile of C per work-group, 4 waves of 32 x 32, each 2 x 2
// blocks of v_mfma_f64_16x16x4f64.  The MFMA computes the transposed
// product (B^T as its A operand, A^T as its B operand) so the 16 lanes of
// a group hold 16 consecutive rows of one column of C and write them as
// one 128-byte segment.
constexpr int kDcLd = 80;   // LDS row of 64 doubles + 16: rows k, k+1 on other banks

__global__ __launch_bounds__(256) void k_dc_gemm(const int *__restrict__ tiles, int ntile,
                                                 const int *__restrict__ jobs, int n,
                                                 const double *__restrict__ Z0,
                                                 const double *__restrict__ Z1,
                                                 const double *__restrict__ Z2,
                                                 const double *__restrict__ Vu,
                                                 double *__restrict__ C)
{
    __shared__ double As[16][kDcLd];
    __shared__ double Bs[16][kDcLd];
    const int t = blockIdx.x % ntile, b = blockIdx.x / ntile;
    const int *tl = tiles + 4 * t;
    const int *jb = jobs + 8 * tl[0];
    const int M = jb[3], N = jb[4], Kd = jb[5];
    const size_t mb = (size_t)b * n * n;
    const double *A = (jb[6] == 0 ? Z0 : jb[6] == 1 ? Z1 : Z2) + mb + jb[0];
    const double *B = Vu + mb + jb[1];
    double *Cp = C + mb + jb[2];
    const int i0 = tl[1] * 64, j0 = tl[2] * 64;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wi = wv & 1, wj = wv >> 1;
    dbl4 acc[2][2];
    for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 2; ++q) acc[p][q] = dbl4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < Kd; k0 += 16) {
        {
            const int i = tid & 63, kk = tid >> 6;
            for (int q = 0; q < 4; ++q) {
                const int k = kk + 4 * q;
                As[k][i] = (i0 + i < M && k0 + k < Kd) ? A[(size_t)(k0 + k) * n + i0 + i] : 0.0;
            }
            const int k = tid & 15, cc = tid >> 4;
            for (int q = 0; q < 4; ++q) {
                const int c = cc + 16 * q;
                Bs[k][c] = (j0 + c < N && k0 + k < Kd) ? B[(size_t)(j0 + c) * n + k0 + k] : 0.0;
            }
        }
        __syncthreads();
        for (int s4 = 0; s4 < 4; ++s4) {
            const int kk = 4 * s4 + (lane >> 4), l16 = lane & 15;
            const double a0 = As[kk][wi * 32 + l16], a1 = As[kk][wi * 32 + 16 + l16];
            const double b0 = Bs[kk][wj * 32 + l16], b1 = Bs[kk][wj * 32 + 16 + l16];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a1, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    // acc[p][q] holds D[c][i] = C[i][c] for i in row block p, c in column
    // block q.  The 16 x 16 result of v_mfma_f64_16x16x4f64 sits in lane l,
    // register r at row l / 16 + 4 r, column l % 16 (rows interleaved over
    // the registers), so here c = l / 16 + 4 r and i = l % 16.
    for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 2; ++q) {
            const int i = i0 + wi * 32 + 16 * p + (lane & 15);
            if (i >= M) continue;
            for (int r = 0; r < 4; ++r) {
                const int c = j0 + wj * 32 + 16 * q + (lane >> 4) + 4 * r;
                if (c < N) Cp[(size_t)c * n + i] = acc[p][q][r];
            }
        }
}

// the eigenvalues unscaled, the eigenvectors into the caller's matrices
template <class T>
__global__ __launch_bounds__(256) void k_dc_finish(const double *__restrict__ wH,
                                                   const double *__restrict__ ZH,
                                                   const double *__restrict__ scale, int n,
                                                   int nb, double *__restrict__ evals,
                                                   T *__restrict__ X)
{
    const size_t nn = (size_t)n * n, g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= nn * nb) return;
    const int b = (int)(g / nn);
    const size_t r = g % nn;
    if (r < (size_t)n) evals[(size_t)b * n + r] = wH[(size_t)b * n + r] * scale[b];
    if constexpr (sizeof(T) == sizeof(double))
        X[g] = ZH[g];
    else
        X[g] = T{ZH[g], 0.0};
}

}   // namespace

void dc_make_plan(int n, DcPlan &plan)
{
    plan.n = n;
    plan.levels.clear();
    plan.host.clear();
    std::vector<Merge> merges;
    plan.H = build_tree(n, merges);
    std::vector<int> &h = plan.host;
    size_t i = 0;
    while (i < merges.size()) {
        size_t j = i;
        while (j < merges.size() && merges[j].h == merges[i].h) ++j;
        DcLevel L{};
        L.h = merges[i].h;
        L.nmerge = (int)(j - i);
        L.merge_off = h.size();
        for (size_t q = i; q < j; ++q) {
            const Merge &mg = merges[q];
            h.insert(h.end(), {mg.r0, mg.n1, mg.n2, mg.hL, mg.hR, mg.h});
        }
        L.item_off = h.size();
        for (size_t q = i; q < j; ++q)
            for (int x = 0; x < merges[q].n1 + merges[q].n2; ++x) {
                h.push_back((int)(q - i));
                h.push_back(x);
                ++L.nitem;
            }
        L.job_off = h.size();
        for (size_t q = i; q < j; ++q) {
            const Merge &mg = merges[q];
            const int m = mg.n1 + mg.n2, s = mg.r0 + mg.n1;
            // left: rows r0.. of columns r0.., A the left child's block
            h.insert(h.end(), {mg.r0 * n + mg.r0, mg.r0 * n + mg.r0, mg.r0 * n + mg.r0, mg.n1, m,
                               mg.n1, mg.hL % 3, 0});
            h.insert(h.end(), {s * n + s, mg.r0 * n + s, mg.r0 * n + s, mg.n2, m, mg.n2,
                               mg.hR % 3, 0});
            L.njob += 2;
        }
        L.tile_off = h.size();
        for (int jb = 0; jb < L.njob; ++jb) {
            const int M = h[L.job_off + 8 * jb + 3], N = h[L.job_off + 8 * jb + 4];
            for (int ti = 0; ti < (M + 63) / 64; ++ti)
                for (int tj = 0; tj < (N + 63) / 64; ++tj) {
                    h.insert(h.end(), {jb, ti, tj, 0});
                    ++L.ntile;
                }
        }
        plan.levels.push_back(L);
        i = j;
    }
}

size_t dc_work_bytes(int n, int nb)
{
    const size_t nn = (size_t)n * n * nb, v = (size_t)n * nb;
    auto al = [](size_t x) { return (x + 31) / 32 * 32; };   // 256-byte alignment
    return 8 * (4 * al(nn) + 4 * al(v) + al(nb) + 9 * al(v)) + 4 * 9 * al(v) * 2;
}

void dc_work_carve(void *base, int n, int nb, DcWork &w)
{
    const size_t nn = (size_t)n * n * nb, v = (size_t)n * nb;
    auto al = [](size_t x) { return (x + 31) / 32 * 32; };
    double *p = static_cast<double *>(base);
    auto take = [&](size_t count) {
        double *q = p;
        p += al(count);
        return q;
    };
    for (int q = 0; q < 3; ++q) w.Zb[q] = take(nn);
    w.Vu = take(nn);
    for (int q = 0; q < 3; ++q) w.wb[q] = take(v);
    w.scale = take(nb);
    w.dl = take(v);
    w.zz = take(v);
    w.dv = take(v);
    w.val = take(v);
    w.rc = take(v);
    w.rs = take(v);
    w.tau = take(v);
    w.zh = take(v);
    w.rho = take(v);
    int *ip = reinterpret_cast<int *>(p);
    auto takei = [&](size_t count) {
        int *q = ip;
        ip += al(count) * 2;
        return q;
    };
    w.perm = takei(v);
    w.kpos = takei(v);
    w.dpos = takei(v);
    w.rp = takei(v);
    w.rq = takei(v);
    w.org = takei(v);
    w.rank = takei(v);
    w.K = takei(v);
    w.nrot = takei(v);
}

hipError_t launch_tridiag_dc(const DcPlan &plan, int nb, const double *d, const double *e,
                             const DcWork &w, double *evals, void *X, bool real,
                             hipStream_t stream)
{
    const int n = plan.n;
    hipLaunchKernelGGL(k_dc_init, dim3(nb), dim3(256), 0, stream, d, e, n, w.scale, w.wb[0],
                       w.Zb[0]);
    const size_t lds_sort = (size_t)2 * n * sizeof(double), lds_col = (size_t)n * sizeof(double);
    for (const DcLevel &L : plan.levels) {
        const int *mg = plan.dev + L.merge_off, *it = plan.dev + L.item_off;
        hipLaunchKernelGGL(k_dc_deflate, dim3(L.nmerge * nb), dim3(256), lds_sort, stream, mg,
                           L.nmerge, n, e, w);
        const int nw = L.nitem * nb;
        hipLaunchKernelGGL(k_dc_secular, dim3((nw + 3) / 4), dim3(256), 0, stream, mg, it,
                           L.nitem, nw, n, w);
        hipLaunchKernelGGL(k_dc_zhat, dim3((nw + 3) / 4), dim3(256), 0, stream, mg, it, L.nitem,
                           nw, n, w);
        hipLaunchKernelGGL(k_dc_vectors, dim3(nw), dim3(64), lds_col, stream, mg, it, L.nitem, n,
                           w);
        hipLaunchKernelGGL(k_dc_gemm, dim3(L.ntile * nb), dim3(256), 0, stream,
                           plan.dev + L.tile_off, L.ntile, plan.dev + L.job_off, n, w.Zb[0],
                           w.Zb[1], w.Zb[2], w.Vu, w.Zb[L.h % 3]);
    }
    const size_t total = (size_t)n * n * nb;
    const unsigned g = (unsigned)((total + 255) / 256);
    if (real)
        hipLaunchKernelGGL(k_dc_finish<double>, dim3(g), dim3(256), 0, stream, w.wb[plan.H % 3],
                           w.Zb[plan.H % 3], w.scale, n, nb, evals, static_cast<double *>(X));
    else
        hipLaunchKernelGGL(k_dc_finish<double2>, dim3(g), dim3(256), 0, stream,
                           w.wb[plan.H % 3], w.Zb[plan.H % 3], w.scale, n, nb, evals,
                           static_cast<double2 *>(X));
    return hipGetLastError();
}

}   // namespace imago
