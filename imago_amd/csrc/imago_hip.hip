// imago_hip.hip - C ABI of the MI355X integral + k-point eigensolve library.
//
// Declarations and conventions: include/imago_hip.h.  Host-side set-up in
// this file restates the parts of the reference's setup that decide which
// primitive tuples are evaluated (src/imago/integrals.F90:323-430 pair and
// image screening, lattice.f90:784-890 minimum image, kpoints.f90:578-609
// phase factors, intgSaving.F90:241-292 origin shift, gaussRelations.f90
// screening tables); the kernels are in assemble.hpp / prim.hpp and below.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <map>
#include <mutex>
#include <atomic>
#include <chrono>
#include <thread>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/imago_hip.h"
#include "common.hpp"
#include "assemble.hpp"
#include "assemble_launch.hpp"
#include "geom_kernel.hpp"
#include "hetrd_launch.hpp"
#include "gemm_mfma_launch.hpp"
#include "lane_xor.hpp"
#include "tridiag_dc_launch.hpp"
#include "dos_launch.hpp"
#include "core_ortho_launch.hpp"
#include "elecstat_launch.hpp"
#include "xcmesh_launch.hpp"
#include "populate.hpp"

using namespace imago;

// ---------------------------------------------------------------------------
// error reporting

static thread_local std::string g_err;

int imago_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char *imago_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// host-side system copy and derived tables

namespace {

struct HostSys {
    int nt = 0, maxa = 0, maxs = 0, nel = 0, maxpa = 1;
    int natoms = 0, npt = 0, nsites = 0, ncells = 0, nk = 0, gamma = 0;
    double logth = 0;
    double rv[9];
    double inv[9];
    std::vector<int> type_el, noa, nstates, lm, lma, atom_type, cumul;
    std::vector<double> alphas, basis, atom_pos, cell_dims, cell_sizes, kpts, kw;
    std::vector<int> pot_el, pot_nalpha, pot_first, pot_mult, site_type;
    std::vector<double> pot_alphas, pot_z, pot_nuc, site_pos;
    // element-keyed tables (first type of each element)
    std::vector<int> el_nalpha, el_npot;
    std::vector<double> el_alpha, el_pot_alpha, el_nuc_alpha;
    std::vector<double> adist, acenter, andist, apdist;
    std::vector<int> term_type, term_alpha, type_term0;
    int n = 0, pot_dim = 0;
    // core states per type (stored after the valence states in basis / lm),
    // the first core function of each atom, and their total (coreDim)
    std::vector<int> ncore, cumul_core;
    int core_dim = 0;
};

// Device buffers of the current system.  Released buffers are kept and
// handed out again to later requests of a similar size (set_system is
// called once per SCF run and repeatedly in benchmarks; returning GBs to
// the driver and mapping them again costs milliseconds).  The cache is
// emptied when an allocation fails and when the context is destroyed.
struct DevBuf {
    struct Blk {
        void *p;
        size_t bytes;
    };
    std::vector<Blk> live, cache;
    void *take(size_t bytes)
    {
        int best = -1;
        for (int i = 0; i < (int)cache.size(); ++i)
            if (cache[i].bytes >= bytes && cache[i].bytes <= 2 * bytes + (1u << 20) &&
                (best < 0 || cache[i].bytes < cache[best].bytes))
                best = i;
        if (best < 0) return nullptr;
        const Blk b = cache[best];
        cache.erase(cache.begin() + best);
        live.push_back(b);
        return b.p;
    }
    void release()
    {
        cache.insert(cache.end(), live.begin(), live.end());
        live.clear();
    }
    void release_one(void *p)
    {
        for (size_t i = 0; i < live.size(); ++i)
            if (live[i].p == p) {
                cache.push_back(live[i]);
                live.erase(live.begin() + i);
                return;
            }
    }
    void purge_cache()
    {
        for (const Blk &b : cache) (void)hipFree(b.p);
        cache.clear();
    }
};

}  // namespace

// A device buffer with guard bands (IMAGO_DEBUG_GUARD)
struct GuardedBuf {
    void *user, *raw;
    size_t bytes;
    int line;
    long seq;
};

struct imago_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    bool has_system = false;
    HostSys h;
    DevSys d{};
    DevBuf sysbuf;
    std::vector<GuardedBuf> guarded;
    long guard_seq = 0;
    // geometry tables
    PairRec *pairs = nullptr;
    ImgRec *imgs = nullptr;
    double2 *phase = nullptr;
    CandRec *cands = nullptr;
    int *cand_off = nullptr;
    int npairs = 0;
    long nimgs = 0, ncands = 0;
    const int *type_term0 = nullptr;
    const int *pair_order = nullptr;
    // block layout of the panels: each pair's block (row states x column
    // states, column-major) back to back; nbe entries per k-point
    long nbe = 0;
    int max_img = 0;   // largest lattice-image count of a pair
    std::vector<PairRec> hpairs;
    const int *atom_of_state = nullptr;   // [n]
    const int *pair_index = nullptr;      // [natoms][natoms]
    int lmax = 0, ne = 0;
    // panels: 0 OL, 1 KE, 2 NP, 3 EP
    double *panel[4] = {nullptr, nullptr, nullptr, nullptr};
    bool assembled[4] = {false, false, false, false};
    // core orthogonalisation (coreDim > 0): the extended view of the
    // system, each type's valence then core states, in its own block
    // layout; the dense extended overlap of every k-point is kept after the
    // overlap's assembly for the other kinds (their Scv)
    DevSys d_ext{};
    PairRec *pairs_ext = nullptr;
    long nbe_ext = 0;
    const int *fn_atom = nullptr, *fn_state = nullptr;   // [valeDim + coreDim]
    const int *atom_nstates_ext = nullptr;               // [atoms]
    double *core_S = nullptr;
    // (term, k-point) slices of each panel written by imago_set_packed;
    // a panel loaded from the host counts as assembled once all are
    std::vector<unsigned char> filled[4];
    bool chol_valid = false;   // Sfull/Sreal hold the Cholesky factor of S(k)
    // secular buffers
    double2 *Hfull = nullptr, *Sfull = nullptr;
    double *Hreal = nullptr, *Sreal = nullptr;
    // U^-1 of every k-point's Cholesky factor (upper triangle; made with the
    // factor) and an n x n scratch per k-point for the products around the
    // eigensolve, typed like Sfull / Sreal
    void *Sinv = nullptr, *Gtmp = nullptr;
    double *evals = nullptr, *ework = nullptr;
    int *info = nullptr;
    double *coef = nullptr;
    double *blockH = nullptr;   // H(k) in the block layout
    int num_states = 0;
    bool have_vectors = false;
    // density / traces
    double *rho = nullptr;      // full n x n per k (complex or real)
    double *rhow = nullptr;     // weighted packed
    double *occ = nullptr;
    double *partial = nullptr;
    // EP screening: tuples of every (potential type, pair) and the primitive
    // block of every (tuple, exponent) item, built once per system
    EpTuple *ep_tuples = nullptr;
    long *ep_tseg = nullptr;     // [types * pairs + 1] offsets into ep_tuples
    long *ep_iseg = nullptr;     // [types * pairs + 1] first item of each task
    int *ep_item_tuple = nullptr;
    double *ep_blocks = nullptr;
    long ep_ntuples = 0, ep_nitems = 0;
    bool ep_screened = false;
    // direct EP path (IMAGO_EP_DIRECT): H(k) and the traces from the
    // primitive blocks each iteration, no potDim panels
    bool ep_direct = false;
    bool ep_runs = false;         // run tables below built
    double *coef_eff = nullptr;   // [pot_dim]
    int *pitem = nullptr;         // [items] items sorted by (pair, lattice image)
    unsigned char *item_jexp = nullptr;   // [items] potential exponent of each item
    EpdSub *hruns = nullptr;      // H sum: runs in pair order
    long nhruns = 0;
    long *hrun_lo = nullptr;      // [pairs + 1] first run of each pair
    long *hrun_woff = nullptr;    // [pairs + 1] first partial double of each pair
    double *hpart = nullptr;      // run partial blocks
    EpdSub *truns = nullptr;      // traces: runs in pair order
    long ntruns = 0;
    double *ival = nullptr;       // [items] trace value of each item
    double *foldX = nullptr;      // [images][64 * ne] density folded per image
    double *tchunk = nullptr;     // [types][kEpdTraceChunks][max_pot_alphas] partial sums
    double timing[23] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    // second stream and handle: the two halves of the batched tridiagonal
    // divide and conquer run side by side (created on first use)
    hipStream_t stream2 = nullptr;
    rocblas_handle blas2 = nullptr;
    hipEvent_t evf = nullptr, evj = nullptr;
    hipEvent_t evh0 = nullptr, evh1 = nullptr;   // around the tridiagonal reduction
    // stages of the eigensolve: [0, 1] the Cholesky transform, [2] the end
    // of the tridiagonal eigensolve, [3] of the compact-WY factors, [4, 5,
    // 6] around the back-transform and the solve of finished columns
    hipEvent_t evs[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    bool evs_pending = false;   // evs[4..6] recorded, not yet read
    // exchange-correlation mesh of the system (imago_xc_mesh)
    bool xc_ready = false;
    imago_xc_mesh_params xcp{};
    double xc_th = 0.0;
    std::vector<int> xc_npts, xc_off;     // per site: ray points, first point
    int xc_maxpts = 0;
    long xc_total = 0;
    double *xc_pts = nullptr, *xc_w = nullptr;   // [points][3], [points]
    double *xc_pls = nullptr, *xc_vec = nullptr, *xc_sitecov = nullptr;
    double *xc_sizes = nullptr;                  // real-space cell sizes
    double *xc_dmax = nullptr;                   // [site] longest ray
    double xc_gemm_flops = 0.0;                  // of the last imago_xc_overlap
    // unscaled operator blocks of every batch of sites, built by the first
    // imago_xc_potential after imago_xc_mesh and kept for the next ones
    struct XcBatch {
        int s0, nb, p0, npb, mm, nn;
        double *B;
        int *rows, *cnt, *rank, *slot, *first;
    };
    std::vector<XcBatch> xc_cache;
    bool xc_cached = false;
    // potential-side overlaps kept for imago_pot_solve: [0] potAlphaOverlap
    // (imago_elecstat), [1] exchCorrOverlap (imago_xc_overlap); the
    // equilibrated upper triangle, its Cholesky factor and the scales,
    // made on the first solve
    struct SpdMat {
        double *A = nullptr;      // the matrix as computed (upper triangle used)
        double *As = nullptr;     // equilibrated copy
        double *U = nullptr;      // its Cholesky factor (upper)
        double *Af = nullptr;     // the equilibrated matrix, both triangles
        double *Ainv = nullptr;   // its inverse from the factor, both triangles
        double *s = nullptr;      // scale factors (1 when not equilibrated)
        bool factored = false;
        int info = 0;
    };
    SpdMat spd[2];
    // run-time choices of the batched eigensolve (imago_tune): columns per
    // diagonal block of the blocked Cholesky (<= kPotrfBlock); the
    // eigensolver (0: this library's, 1: rocSOLVER's heevd / syevd); the
    // Cholesky (0: this library's blocked one, 1: rocSOLVER's potrf).
    // Initialised from IMAGO_HEEVD / IMAGO_POTRF = rocsolver.
    int tune_potrf_block = kPotrfBlock;
    int tune_heevd = 0;
    int tune_potrf = 0;
    // the tridiagonal stage of this library's eigensolver (0: its batched
    // divide and conquer, 1: rocSOLVER's dstedc; IMAGO_STEDC = rocsolver)
    // and the divide and conquer's merge plan for the last order, with its
    // device copy
    int tune_tridiag = 0;
    DcPlan dcplan;
    // the divide and conquer's launches captured once in a graph and
    // replayed (IMAGO_TUNE_GRAPH, default on): its work arrays are kept
    // for the graph's lifetime, which is keyed by the order, the batch and
    // the buffers it reads and writes
    int tune_graph = 1;
    // the back-transform of the reduction's reflectors: 1 (the default) in
    // compact-WY blocks of kTrdBlock, the Gram blocks on the matrix cores and
    // the block products with rocBLAS; 0 all n - 1 reflectors at once in the
    // UT form (one Gram GEMM and T = (T^-1)^-1 per eigensolve, three GEMMs
    // per finished column range on the matrix cores; IMAGO_BACKTRANSFORM =
    // ut), slower at the workloads' orders because of the triangular inverse
    int tune_wy = 1;
    struct DcGraph {
        hipGraphExec_t exec = nullptr;
        int n = 0, nb = 0;
        bool real = false;
        const void *d = nullptr, *e = nullptr, *X = nullptr;
        const int *plan = nullptr;   // the plan's device copy the graph reads
        char *work = nullptr;
        size_t work_bytes = 0;
    } dcg;
    // eigenvector columns of the last imago_secular still to finish: the
    // tridiagonal eigenvectors sit in the result matrices, the reflectors of
    // the reduction (V) and their compact-WY factors (Tf) are kept, and
    // columns [done, n) wait for the back-transform Y = Q Z and the solve
    // X = U^-1 Y.  The density needs only the occupied columns; the rest
    // are finished when they are asked for.
    struct PendingVectors {
        bool active = false;
        bool reflectors = false;   // false: the back-transform is done, only U^-1 left
        bool ut = false;           // Tf holds the UT factor of all reflectors (n x n each)
        bool cx = true;
        int n = 0, nb = 0, done = 0;
        char *V = nullptr, *Tf = nullptr;
        void *A = nullptr;
        const void *B = nullptr;
        const void *Binv = nullptr;   // U^-1 when known (the solve is then a GEMM)
    } pend;
    // makeSCFPot state (imago_scf_setup / imago_scf_update): the
    // electrostatic operators on the device, the per-iteration vectors on
    // the host, the mixing history the reference keeps in module variables
    // (xl1, xl2, yl1, yl2; potentialUpdate.F90:38-45)
    struct ScfState {
        bool ready = false;
        imago_scf_params p{};
        std::vector<double> intg, nuc_nl, nuc_l;
        double *es_buf = nullptr;   // device: the elecstat arrays (one allocation)
        double *neut_nl = nullptr, *neut_l = nullptr, *resq = nullptr;   // device
        // device scratch of imago_scf_update (its vectors, the XC potential's
        // point values and partial sums, the convergence measure's rows and
        // point values), sized by imago_scf_setup
        double *work = nullptr, *xc_v = nullptr, *xc_part = nullptr, *td_u = nullptr,
               *td_d = nullptr;
        int iteration = 1;
        bool have_hist = false;
        std::vector<double> xl1, xl2, yl1, yl2;
    } scf;
    // partial sums of the row-form matrix-vector product (dev_gemv),
    // grown on demand, held with the system's buffers
    double *gv_part = nullptr;
    size_t gv_cap = 0;
};

static void guard_release_all(imago_ctx *c);

static void ctx_free_system(imago_ctx *c)
{
    c->sysbuf.release();
    guard_release_all(c);
    for (int k = 0; k < 4; ++k) {
        c->panel[k] = nullptr;
        c->assembled[k] = false;
        c->filled[k].clear();
    }
    c->chol_valid = false;
    c->d_ext = DevSys{};
    c->pairs_ext = nullptr;
    c->nbe_ext = 0;
    c->fn_atom = c->fn_state = c->atom_nstates_ext = nullptr;
    c->core_S = nullptr;
    c->ep_tuples = nullptr;
    c->ep_tseg = nullptr;
    c->ep_iseg = nullptr;
    c->ep_item_tuple = nullptr;
    c->ep_blocks = nullptr;
    c->ep_ntuples = c->ep_nitems = 0;
    c->ep_screened = false;
    c->ep_direct = false;
    c->ep_runs = false;
    c->coef_eff = c->hpart = c->ival = c->foldX = c->tchunk = nullptr;
    c->pitem = nullptr;
    c->item_jexp = nullptr;
    c->hruns = c->truns = nullptr;
    c->hrun_lo = c->hrun_woff = nullptr;
    c->nhruns = c->ntruns = 0;
    c->pairs = nullptr;
    c->type_term0 = nullptr;
    c->pair_order = nullptr;
    c->nbe = 0;
    c->hpairs.clear();
    c->atom_of_state = nullptr;
    c->pair_index = nullptr;
    c->imgs = nullptr;
    c->phase = nullptr;
    c->cands = nullptr;
    c->cand_off = nullptr;
    c->Hfull = c->Sfull = nullptr;
    c->Hreal = c->Sreal = nullptr;
    c->Sinv = c->Gtmp = nullptr;
    c->evals = c->ework = nullptr;
    c->info = nullptr;
    c->coef = nullptr;
    c->blockH = nullptr;
    c->rho = c->rhow = c->occ = c->partial = nullptr;
    c->have_vectors = false;
    c->xc_ready = false;
    c->xc_npts.clear();
    c->xc_off.clear();
    c->xc_maxpts = 0;
    c->xc_total = 0;
    c->xc_pts = c->xc_w = c->xc_pls = c->xc_vec = c->xc_sitecov = c->xc_sizes = nullptr;
    c->xc_dmax = nullptr;
    c->xc_cache.clear();
    c->xc_cached = false;
    for (auto &m : c->spd) m = imago_ctx::SpdMat{};
    c->scf = imago_ctx::ScfState{};
    c->pend = imago_ctx::PendingVectors{};
    c->gv_part = nullptr;
    c->gv_cap = 0;
    c->has_system = false;
}

// IMAGO_DEBUG_GUARD (a check for writes outside a buffer's bounds; slow):
// every device buffer is a fresh allocation between two guard bands of a
// fixed byte pattern, and the bands are compared with the pattern when the
// buffer is returned or the system's buffers are; a changed band is
// reported on stderr with the buffer's size and the source line that
// allocated it.
constexpr size_t kGuardBytes = 256 << 10;
constexpr unsigned char kGuardByte = 0xcb;

static bool guard_mode()
{
    static const bool on = [] {
        const bool v = getenv("IMAGO_DEBUG_GUARD") != nullptr;
        if (v) fprintf(stderr, "[imago guard] guard bands of %zu bytes on\n", kGuardBytes);
        return v;
    }();
    return on;
}

static void guard_check(imago_ctx *c, const GuardedBuf &g)
{
    std::vector<unsigned char> band(kGuardBytes);
    for (int side = 0; side < 2; ++side) {
        const unsigned char *src = static_cast<const unsigned char *>(g.raw) +
                                   (side ? kGuardBytes + g.bytes : 0);
        if (hipMemcpyAsync(band.data(), src, kGuardBytes, hipMemcpyDeviceToHost, c->stream) !=
                hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            fprintf(stderr, "[imago guard] could not read the guard of a buffer from line %d\n",
                    g.line);
            return;
        }
        size_t changed = 0, first = 0, last = 0;
        for (size_t b = 0; b < kGuardBytes; ++b)
            if (band[b] != kGuardByte) {
                if (!changed) first = b;
                last = b;
                ++changed;
            }
        if (changed)
            fprintf(stderr,
                    "[imago guard] %zu bytes %s the buffer of %zu bytes allocated at line %d "
                    "(#%ld) changed, offsets %ld .. %ld from its %s\n",
                    changed, side ? "after" : "before", g.bytes, g.line, g.seq,
                    side ? (long)first : (long)first - (long)kGuardBytes,
                    side ? (long)last : (long)last - (long)kGuardBytes, side ? "end" : "start");
    }
}

static void guard_release(imago_ctx *c, void *p)
{
    for (size_t i = 0; i < c->guarded.size(); ++i)
        if (c->guarded[i].user == p) {
            guard_check(c, c->guarded[i]);
            (void)hipFree(c->guarded[i].raw);
            c->guarded.erase(c->guarded.begin() + i);
            return;
        }
}

static void guard_release_all(imago_ctx *c)
{
    if (c->guarded.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (const GuardedBuf &g : c->guarded) {
        guard_check(c, g);
        (void)hipFree(g.raw);
    }
    c->guarded.clear();
}

template <class T>
static int dev_alloc_at(imago_ctx *c, T **p, size_t count, int line)
{
    const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    if (guard_mode()) {
        void *raw = nullptr;
        HIPCHK(hipMalloc(&raw, bytes + 2 * kGuardBytes));
        unsigned char *r = static_cast<unsigned char *>(raw);
        HIPCHK(hipMemsetAsync(r, kGuardByte, kGuardBytes, c->stream));
        HIPCHK(hipMemsetAsync(r + kGuardBytes + bytes, kGuardByte, kGuardBytes, c->stream));
        c->guarded.push_back({r + kGuardBytes, raw, bytes, line, c->guard_seq++});
        *p = (T *)(r + kGuardBytes);
        return IMAGO_OK;
    }
    void *q = c->sysbuf.take(bytes);
    if (!q) {
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            c->sysbuf.purge_cache();
            e = hipMalloc(&q, bytes);
        }
        if (e != hipSuccess)
            return fail(IMAGO_E_NOMEM, "hipMalloc of %zu bytes failed: %s", bytes,
                        hipGetErrorString(e));
        c->sysbuf.live.push_back({q, bytes});
    }
    *p = (T *)q;
    return IMAGO_OK;
}

#define dev_alloc(c, p, count) dev_alloc_at((c), (p), (count), __LINE__)

static void dev_release(imago_ctx *c, void *p)
{
    if (guard_mode()) {
        (void)hipStreamSynchronize(c->stream);
        guard_release(c, p);
        return;
    }
    c->sysbuf.release_one(p);
}

// Returns an SPD matrix's buffers (the matrix, its factor and inverse) and
// leaves it empty.
static void spd_release(imago_ctx *c, imago_ctx::SpdMat &m)
{
    for (void *q : {(void *)m.A, (void *)m.As, (void *)m.U, (void *)m.Af, (void *)m.Ainv,
                    (void *)m.s})
        if (q) dev_release(c, q);
    m = imago_ctx::SpdMat{};
}

// A blocking copy on the context's own stream.  The library never uses the
// legacy default stream: a copy there would wait on (and, under a graph
// capture, invalidate) the streams of other contexts in the process.
static hipError_t ctx_memcpy(imago_ctx *c, void *dst, const void *src, size_t bytes,
                             hipMemcpyKind kind)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

template <class T>
static int dev_upload(imago_ctx *c, const T **dst, const std::vector<T> &v)
{
    T *p = nullptr;
    RCHK(dev_alloc(c, &p, v.size()));
    if (!v.empty()) HIPCHK(ctx_memcpy(c, p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *dst = p;
    return IMAGO_OK;
}

static void find_lattice_vector(const HostSys &h, const double *v, double *L)
{
    imago::find_lattice_vector(h.rv, h.inv, h.cell_dims.data(), h.cell_sizes.data(), h.ncells, v, L);
}

static size_t adx(const HostSys &h, int e1, int e2, int a1, int a2)
{
    return (((size_t)e1 * h.nel + e2) * h.maxa + a1) * h.maxa + a2;
}

static int build_host(HostSys &h, const imago_system_desc *s)
{
    if (!s) return fail(IMAGO_E_ARG, "null system description");
    h = HostSys();
    h.nt = s->num_atom_types;
    h.maxa = s->max_alphas;
    h.maxs = s->max_states;
    h.natoms = s->num_atoms;
    h.npt = s->num_pot_types;
    h.maxpa = s->max_pot_alphas > 0 ? s->max_pot_alphas : 1;
    h.nsites = s->num_pot_sites;
    h.ncells = s->num_cells;
    h.gamma = s->gamma ? 1 : 0;
    h.nk = h.gamma ? 1 : s->num_kpoints;
    h.logth = s->log_basis_fn_thresh;
    if (h.nt <= 0 || h.natoms <= 0 || h.ncells <= 0 || h.nk <= 0 || h.npt <= 0)
        return fail(IMAGO_E_ARG, "empty system (types %d atoms %d cells %d k %d pot types %d)",
                    h.nt, h.natoms, h.ncells, h.nk, h.npt);
    if (h.maxa > 255 || h.maxs > 32)
        return fail(IMAGO_E_ARG, "max_alphas %d (<=255) / max_states %d (<=32) out of range",
                    h.maxa, h.maxs);
    if (!(h.logth > 0.0)) return fail(IMAGO_E_ARG, "log_basis_fn_thresh must be > 0");
    std::memcpy(h.rv, s->real_vectors, sizeof h.rv);
    h.type_el.assign(s->type_element, s->type_element + h.nt);
    h.noa.assign(s->type_num_orb_alphas, s->type_num_orb_alphas + 4 * h.nt);
    h.nstates.assign(s->type_num_states, s->type_num_states + h.nt);
    h.ncore.assign(h.nt, 0);
    if (s->type_num_core_states)
        h.ncore.assign(s->type_num_core_states, s->type_num_core_states + h.nt);
    h.alphas.assign(s->type_alphas, s->type_alphas + (size_t)h.nt * h.maxa);
    h.basis.assign(s->type_basis, s->type_basis + (size_t)h.nt * h.maxs * h.maxa);
    h.lm.assign(s->type_lm_index, s->type_lm_index + (size_t)h.nt * h.maxs);
    h.atom_type.assign(s->atom_type, s->atom_type + h.natoms);
    h.atom_pos.assign(s->atom_pos, s->atom_pos + 3 * (size_t)h.natoms);
    h.cell_dims.assign(s->cell_dims, s->cell_dims + 3 * (size_t)h.ncells);
    h.cell_sizes.assign(s->cell_sizes, s->cell_sizes + h.ncells);
    if (!h.gamma) {
        h.kpts.assign(s->kpoints, s->kpoints + 3 * (size_t)h.nk);
        h.kw.assign(s->kpoint_weights, s->kpoint_weights + h.nk);
    } else {
        h.kpts.assign(3, 0.0);
        h.kw.assign(1, 2.0);
    }
    h.pot_el.assign(s->pot_type_element, s->pot_type_element + h.npt);
    h.pot_nalpha.assign(s->pot_type_num_alphas, s->pot_type_num_alphas + h.npt);
    h.pot_first.assign(s->pot_type_first_site, s->pot_type_first_site + h.npt);
    h.pot_mult.assign(s->pot_type_multiplicity, s->pot_type_multiplicity + h.npt);
    h.pot_alphas.assign(h.npt * (size_t)h.maxpa, 0.0);
    for (int t = 0; t < h.npt; ++t)
        for (int a = 0; a < h.pot_nalpha[t]; ++a)
            h.pot_alphas[t * h.maxpa + a] = s->pot_type_alphas[t * s->max_pot_alphas + a];
    h.pot_z.assign(s->pot_type_nuc_charge, s->pot_type_nuc_charge + h.npt);
    h.pot_nuc.assign(s->pot_type_nuc_alpha, s->pot_type_nuc_alpha + h.npt);
    h.site_type.assign(s->pot_site_type, s->pot_site_type + h.nsites);
    h.site_pos.assign(s->pot_site_pos, s->pot_site_pos + 3 * (size_t)h.nsites);

    // validation
    for (int t = 0; t < h.nt; ++t) {
        const int *noa = &h.noa[4 * t];
        if (noa[0] <= 0 || noa[0] > h.maxa)
            return fail(IMAGO_E_ARG, "type %d: numOrbAlphas(1)=%d out of range", t, noa[0]);
        for (int l = 0; l < 3; ++l)
            if (noa[l + 1] > noa[l])
                return fail(IMAGO_E_ARG, "type %d: numOrbAlphas not non-increasing in l", t);
        if (h.nstates[t] <= 0 || h.nstates[t] > h.maxs)
            return fail(IMAGO_E_ARG, "type %d: num_states %d out of range", t, h.nstates[t]);
        if (h.ncore[t] < 0 || h.nstates[t] + h.ncore[t] > h.maxs)
            return fail(IMAGO_E_ARG, "type %d: %d valence + %d core states exceed max_states %d", t,
                        h.nstates[t], h.ncore[t], h.maxs);
        for (int m = 0; m < h.nstates[t] + h.ncore[t]; ++m) {
            int lmv = h.lm[t * h.maxs + m];
            if (lmv < 1 || lmv > 16) return fail(IMAGO_E_ARG, "type %d: lm index %d", t, lmv);
        }
        for (int a = 1; a < noa[0]; ++a)
            if (!(h.alphas[t * h.maxa + a] > h.alphas[t * h.maxa + a - 1]))
                return fail(IMAGO_E_ARG, "type %d: exponents not ascending", t);
    }
    for (int i = 0; i < h.natoms; ++i)
        if (h.atom_type[i] < 0 || h.atom_type[i] >= h.nt)
            return fail(IMAGO_E_ARG, "atom %d: bad type %d", i, h.atom_type[i]);
    for (int i = 0; i < h.nsites; ++i)
        if (h.site_type[i] < 0 || h.site_type[i] >= h.npt)
            return fail(IMAGO_E_ARG, "potential site %d: bad type", i);
    for (int t = 0; t < h.npt; ++t) {
        if (h.pot_nalpha[t] <= 0 || h.pot_nalpha[t] > kEpMaxPotAlphas || h.pot_nalpha[t] > h.maxpa)
            return fail(IMAGO_E_ARG, "potential type %d: %d exponents (1..%d supported)", t,
                        h.pot_nalpha[t], kEpMaxPotAlphas);
        if (h.pot_first[t] < 0 || h.pot_mult[t] <= 0 || h.pot_first[t] + h.pot_mult[t] > h.nsites)
            return fail(IMAGO_E_ARG, "potential type %d: bad site range", t);
        for (int s2 = h.pot_first[t]; s2 < h.pot_first[t] + h.pot_mult[t]; ++s2)
            if (h.site_type[s2] != t)
                return fail(IMAGO_E_ARG, "potential type %d: sites not contiguous", t);
    }
    // potential coefficients are read in type order (initPotCoeffs,
    // potential.f90:456-470) while the terms are built in the order of each
    // type's first site (integrals.F90:2411-2476): the two agree only when
    // types are numbered by first site, as the reference's input writer does
    for (int t = 1; t < h.npt; ++t)
        if (h.pot_first[t] <= h.pot_first[t - 1])
            return fail(IMAGO_E_ARG, "potential types must be numbered in the order of their "
                                     "first site (type %d starts at site %d)", t + 1,
                        h.pot_first[t] + 1);
    for (int k = 1; k < h.ncells; ++k)
        if (h.cell_sizes[k] < h.cell_sizes[k - 1])
            return fail(IMAGO_E_ARG, "cell list not sorted by |R|^2");

    // lmAlphaIndex (basis.f90:124-148)
    h.lma.assign((size_t)h.nt * h.maxa, 1);
    for (int t = 0; t < h.nt; ++t) {
        const int *noa = &h.noa[4 * t];
        for (int a = 0; a < noa[0]; ++a)
            h.lma[t * h.maxa + a] = a < noa[3] ? 16 : a < noa[2] ? 9 : a < noa[1] ? 4 : 1;
    }
    // cumulValeStates, cumulCoreStates
    h.cumul.resize(h.natoms);
    h.cumul_core.resize(h.natoms);
    h.n = 0;
    h.core_dim = 0;
    for (int i = 0; i < h.natoms; ++i) {
        h.cumul[i] = h.n;
        h.n += h.nstates[h.atom_type[i]];
        h.cumul_core[i] = h.core_dim;
        h.core_dim += h.ncore[h.atom_type[i]];
    }
    // inverse of the real vectors (columns)
    {
        double m[3][3];
        for (int i = 0; i < 3; ++i)
            for (int x = 0; x < 3; ++x) m[x][i] = h.rv[3 * i + x];
        double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                     m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        if (det == 0.0) return fail(IMAGO_E_ARG, "singular real-space vectors");
        double id = 1.0 / det;
        h.inv[0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
        h.inv[1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
        h.inv[2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
        h.inv[3] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
        h.inv[4] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
        h.inv[5] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
        h.inv[6] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
        h.inv[7] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
        h.inv[8] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    }
    // element-keyed screening tables (gaussRelations.f90:64-278)
    h.nel = 0;
    for (int t = 0; t < h.nt; ++t) h.nel = std::max(h.nel, h.type_el[t] + 1);
    for (int t = 0; t < h.npt; ++t) h.nel = std::max(h.nel, h.pot_el[t] + 1);
    h.el_nalpha.assign(h.nel, 0);
    h.el_alpha.assign((size_t)h.nel * h.maxa, 0.0);
    h.el_npot.assign(h.nel, 0);
    h.el_pot_alpha.assign((size_t)h.nel * h.maxpa, 0.0);
    h.el_nuc_alpha.assign(h.nel, 0.0);
    for (int t = 0; t < h.nt; ++t) {
        int e = h.type_el[t];
        if (h.el_nalpha[e] == 0) {
            h.el_nalpha[e] = h.noa[4 * t];
            for (int a = 0; a < h.el_nalpha[e]; ++a)
                h.el_alpha[e * h.maxa + a] = h.alphas[t * h.maxa + a];
        }
    }
    for (int t = 0; t < h.npt; ++t) {
        int e = h.pot_el[t];
        if (h.el_npot[e] == 0) {
            h.el_npot[e] = h.pot_nalpha[t];
            for (int a = 0; a < h.el_npot[e]; ++a)
                h.el_pot_alpha[e * h.maxpa + a] = h.pot_alphas[t * h.maxpa + a];
        }
        if (h.el_nuc_alpha[e] == 0.0) h.el_nuc_alpha[e] = h.pot_nuc[t];
    }
    for (int t = 0; t < h.nt; ++t)
        if (h.noa[4 * t] > h.el_nalpha[h.type_el[t]])
            return fail(IMAGO_E_ARG, "type %d has more exponents than the first type of its element",
                        t);
    const int M = h.maxa, E = h.nel, PA = h.maxpa;
    h.adist.assign((size_t)E * E * M * M, 0.0);
    h.acenter.assign((size_t)E * E * M * M, 0.0);
    h.andist.assign((size_t)E * E * E * M * M, 0.0);
    h.apdist.assign((size_t)E * PA * E * E * M * M, 0.0);
    for (int e1 = 0; e1 < E; ++e1)
        for (int e2 = 0; e2 < E; ++e2)
            for (int a1 = 0; a1 < h.el_nalpha[e1]; ++a1)
                for (int a2 = 0; a2 < h.el_nalpha[e2]; ++a2) {
                    double x = h.el_alpha[e1 * M + a1], y = h.el_alpha[e2 * M + a2];
                    double c = x / (x + y);
                    h.acenter[adx(h, e1, e2, a1, a2)] = c;
                    h.adist[adx(h, e1, e2, a1, a2)] = h.logth / y / c;
                    for (int pe = 0; pe < E; ++pe) {
                        double an = h.el_nuc_alpha[pe];
                        if (an > 0.0)
                            h.andist[(size_t)pe * E * E * M * M + adx(h, e1, e2, a1, a2)] =
                                h.logth * (x + y + an) / (x + y) / an;
                        for (int j = 0; j < h.el_npot[pe]; ++j) {
                            double ap = h.el_pot_alpha[pe * PA + j];
                            h.apdist[((size_t)pe * PA + j) * E * E * M * M + adx(h, e1, e2, a1, a2)] =
                                h.logth * (x + y + ap) / (x + y) / ap;
                        }
                    }
                }
    // the walk of integrals.F90:451-568 equals the set alphaDist >= sep only
    // when alphaDist is non-increasing along both alpha indices
    for (int e1 = 0; e1 < E; ++e1)
        for (int e2 = 0; e2 < E; ++e2)
            for (int a1 = 0; a1 < h.el_nalpha[e1]; ++a1)
                for (int a2 = 0; a2 < h.el_nalpha[e2]; ++a2) {
                    double v = h.adist[adx(h, e1, e2, a1, a2)];
                    if ((a1 > 0 && v > h.adist[adx(h, e1, e2, a1 - 1, a2)]) ||
                        (a2 > 0 && v > h.adist[adx(h, e1, e2, a1, a2 - 1)]))
                        return fail(IMAGO_E_SCREEN,
                                    "alphaDist not monotone at elements (%d,%d) alphas (%d,%d)",
                                    e1, e2, a1, a2);
                }
    // potential terms in the reference's order (integrals.F90:2411-2476)
    std::vector<int> order(h.npt);
    for (int t = 0; t < h.npt; ++t) order[t] = t;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return h.pot_first[a] < h.pot_first[b]; });
    h.type_term0.assign(h.npt, 0);
    for (int t : order) {
        h.type_term0[t] = (int)h.term_type.size();
        for (int a = 0; a < h.pot_nalpha[t]; ++a) {
            h.term_type.push_back(t);
            h.term_alpha.push_back(a);
        }
    }
    h.pot_dim = (int)h.term_type.size();
    return IMAGO_OK;
}

// Pair/image table (integrals.F90:323-430) with the Bloch phases of every
// image, built for the pairs (i, j >= i) of one atom i; atoms run in
// parallel and their parts are concatenated in atom order.  The 3-centre
// candidate images are listed on the device (k_geom_cands).
struct GeomPart {
    std::vector<PairRec> pairs;
    std::vector<ImgRec> imgs;
    std::vector<double2> phase;
    std::vector<int> img_pair;   // pair of each image, relative to this part's pairs
    int rc = IMAGO_OK;
    std::string err;
};

static void build_geometry_atom(const HostSys &h, int i, GeomPart &g)
{
    const int nk = h.nk;
    auto &pairs = g.pairs;
    auto &imgs = g.imgs;
    auto &phase = g.phase;
    auto &img_pair = g.img_pair;
    const double lastsz = h.cell_sizes[h.ncells - 1];
    {
        const int t1 = h.atom_type[i], e1 = h.type_el[t1];
        const double *A = &h.atom_pos[3 * i];
        for (int j = i; j < h.natoms; ++j) {
            const int t2 = h.atom_type[j], e2 = h.type_el[t2];
            const double *B = &h.atom_pos[3 * j];
            PairRec pr{i, j, (int)imgs.size(), 0};
            double dv[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]}, L[3];
            find_lattice_vector(h, dv, L);
            double sepsq = 0;
            for (int x = 0; x < 3; ++x) sepsq += (A[x] - B[x] - L[x]) * (A[x] - B[x] - L[x]);
            const double negl = h.adist[adx(h, e1, e2, 0, 0)];
            if (sepsq > negl) {
                pairs.push_back(pr);
                continue;
            }
            const double maxLR = sepsq + negl + 2.0 * std::sqrt(sepsq * negl);
            if (maxLR > lastsz) {
                char buf[256];
                snprintf(buf, sizeof buf,
                         "More lattice points needed for atom pair (%d,%d): requested %g, "
                         "available %g", i + 1, j + 1, maxLR, lastsz);
                g.rc = IMAGO_E_LATTICE;
                g.err = buf;
                return;
            }
            // origin shift e^{ik.L}, skipped when |k.L| < 1e-8
            std::vector<double2> shift(nk, make_double2(1.0, 0.0));
            if (!h.gamma)
                for (int q = 0; q < nk; ++q) {
                    const double *kp = &h.kpts[3 * q];
                    double dot = kp[0] * L[0] + kp[1] * L[1] + kp[2] * L[2];
                    if (std::fabs(dot) >= 1.0e-8) shift[q] = make_double2(std::cos(dot), std::sin(dot));
                }
            for (int k = 0; k < h.ncells; ++k) {
                if (h.cell_sizes[k] > maxLR) break;
                const double *R = &h.cell_dims[3 * k];
                ImgRec ir;
                ir.bx = B[0] + L[0] + R[0];
                ir.by = B[1] + L[1] + R[1];
                ir.bz = B[2] + L[2] + R[2];
                double sep = (A[0] - ir.bx) * (A[0] - ir.bx) + (A[1] - ir.by) * (A[1] - ir.by) +
                             (A[2] - ir.bz) * (A[2] - ir.bz);
                if (sep > negl) continue;
                ir.sep = sep;
                imgs.push_back(ir);
                for (int q = 0; q < nk; ++q) {
                    double2 pf = make_double2(1.0, 0.0);
                    if (!h.gamma) {
                        const double *kp = &h.kpts[3 * q];
                        double dot = kp[0] * R[0] + kp[1] * R[1] + kp[2] * R[2];
                        pf = make_double2(std::cos(dot), std::sin(dot));
                    }
                    double2 s = shift[q];
                    phase.push_back(make_double2(pf.x * s.x - pf.y * s.y, pf.x * s.y + pf.y * s.x));
                }
                img_pair.push_back((int)pairs.size());
                pr.img_cnt++;
                if (pr.img_cnt >= 65536) {
                    g.rc = IMAGO_E_LATTICE;
                    g.err = "more than 65535 lattice images for one atom pair";
                    return;
                }
            }
            pairs.push_back(pr);
        }
    }
}

// 3-centre candidate images of every (image, potential site) on the
// device (k_geom_cands): count, scan into cand_off[image * sites + site],
// fill.  The largest threshold of each (potential element, element pair)
// bounds the distance of a passing image from the segment [B', A].
static int build_candidates(imago_ctx *c, const std::vector<int> &img_pair)
{
    const HostSys &h = c->h;
    const long nimg = c->nimgs, ntask = nimg * h.nsites;
    if (ntask >= (1L << 31)) return fail(IMAGO_E_ARG, "%ld (image, site) pairs", ntask);
    std::vector<double> bound0((size_t)h.nel * h.nel * h.nel, 0.0);
    for (int pel = 0; pel < h.nel; ++pel)
        for (int e1 = 0; e1 < h.nel; ++e1)
            for (int e2 = 0; e2 < h.nel; ++e2) {
                const size_t base = adx(h, e1, e2, 0, 0);
                double b = 0.0;
                if (h.el_nuc_alpha[pel] > 0.0)
                    b = std::max(b, h.andist[(size_t)pel * h.nel * h.nel * h.maxa * h.maxa + base]);
                if (h.el_npot[pel] > 0)
                    b = std::max(b, h.apdist[((size_t)pel * h.maxpa) * h.nel * h.nel * h.maxa * h.maxa + base]);
                bound0[((size_t)pel * h.nel + e1) * h.nel + e2] = b;
            }
    const int *dimg_pair = nullptr;
    const double *dbound0 = nullptr, *dsizes = nullptr;
    RCHK(dev_upload(c, &dimg_pair, img_pair));
    RCHK(dev_upload(c, &dbound0, bound0));
    RCHK(dev_upload(c, &dsizes, h.cell_sizes));
    GeomArgs G{};
    G.pairs = c->pairs;
    G.imgs = c->imgs;
    G.img_pair = dimg_pair;
    G.atom_pos = c->d.atom_pos;
    G.atom_type = c->d.atom_type;
    G.type_element = c->d.type_element;
    G.site_pos = c->d.pot_pos;
    G.site_type = c->d.pot_site_type;
    G.pot_el = c->d.pot_el;
    G.adist = c->d.adist;
    G.bound0 = dbound0;
    G.cell_dims = c->d.cell_dims;
    G.cell_sizes = dsizes;
    std::memcpy(G.rv, h.rv, sizeof G.rv);
    std::memcpy(G.inv, h.inv, sizeof G.inv);
    G.ncells = h.ncells;
    G.nsites = h.nsites;
    G.nel = h.nel;
    G.maxa = h.maxa;
    G.nimgs = nimg;
    int *cnt = nullptr, *err = nullptr;
    RCHK(dev_alloc(c, &cnt, (size_t)ntask + 1));
    RCHK(dev_alloc(c, &c->cand_off, (size_t)ntask + 1));
    RCHK(dev_alloc(c, &err, 1));
    HIPCHK(hipMemsetAsync(err, 0, sizeof(int), c->stream));
    HIPCHK(hipMemsetAsync(cnt + ntask, 0, sizeof(int), c->stream));
    const unsigned g = (unsigned)((ntask + 255) / 256);
    if (ntask > 0)
        hipLaunchKernelGGL(k_geom_cands<false>, dim3(g), dim3(256), 0, c->stream, G, cnt,
                           (const int *)nullptr, (CandRec *)nullptr, err);
    HIPCHK(hipGetLastError());
    size_t tmp_bytes = 0;
    HIPCHK(int_exclusive_scan(cnt, c->cand_off, ntask + 1, nullptr, &tmp_bytes, c->stream));
    char *tmp = nullptr;
    RCHK(dev_alloc(c, &tmp, tmp_bytes));
    HIPCHK(int_exclusive_scan(cnt, c->cand_off, ntask + 1, tmp, &tmp_bytes, c->stream));
    int herr = 0, total = 0;
    HIPCHK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&total, c->cand_off + ntask, sizeof(int), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dev_release(c, cnt);
    dev_release(c, tmp);
    dev_release(c, err);
    if (herr)
        return fail(IMAGO_E_LATTICE, "More lattice points needed for the potential sites near "
                                     "some atom pair: extend the cell list");
    if (total < 0) return fail(IMAGO_E_ARG, "more than 2^31 candidate images");
    c->ncands = total;
    RCHK(dev_alloc(c, &c->cands, (size_t)std::max(total, 1)));
    if (ntask > 0)
        hipLaunchKernelGGL(k_geom_cands<true>, dim3(g), dim3(256), 0, c->stream, G, (int *)nullptr,
                           c->cand_off, c->cands, (int *)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return IMAGO_OK;
}

static int build_geometry(imago_ctx *c)
{
    HostSys &h = c->h;
    std::vector<GeomPart> parts(h.natoms);
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<int> next{0};
    auto worker = [&]() {
        for (int i = next++; i < h.natoms; i = next++) build_geometry_atom(h, i, parts[i]);
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nth; ++t) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    std::vector<PairRec> pairs;
    std::vector<ImgRec> imgs;
    std::vector<double2> phase;
    std::vector<int> img_pair;
    for (auto &g : parts) {
        if (g.rc != IMAGO_OK) return fail(g.rc, "%s", g.err.c_str());
        const int ib = (int)imgs.size(), pb = (int)pairs.size();
        for (PairRec pr : g.pairs) {
            pr.img_off += ib;
            pairs.push_back(pr);
        }
        imgs.insert(imgs.end(), g.imgs.begin(), g.imgs.end());
        phase.insert(phase.end(), g.phase.begin(), g.phase.end());
        for (int q : g.img_pair) img_pair.push_back(q + pb);
    }
    c->npairs = (int)pairs.size();
    // pairs ordered by column atom j, then row atom i (the packed layout's
    // order) for the EP launch
    {
        std::vector<int> order(pairs.size());
        for (size_t q = 0; q < pairs.size(); ++q) order[q] = (int)q;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return pairs[a].j != pairs[b].j ? pairs[a].j < pairs[b].j : pairs[a].i < pairs[b].i;
        });
        const int *po = nullptr;
        RCHK(dev_upload(c, &po, order));
        c->pair_order = po;
        // block layout: the pairs' blocks back to back in the same order, so
        // consecutive work-groups of the EP launch write consecutive memory
        long off = 0;
        for (int q : order) {
            PairRec &pr = pairs[q];
            const long sz = (long)h.nstates[h.atom_type[pr.i]] * h.nstates[h.atom_type[pr.j]];
            if (off + sz >= (1L << 31)) return fail(IMAGO_E_ARG, "block layout exceeds 2^31 entries");
            pr.blk = (int)off;
            off += sz;
        }
        c->nbe = off;
        c->d.nbe = off;
    }
    // unpack tables: the atom of each basis function and the pair of each
    // (row atom, column atom), -1 below the diagonal
    {
        std::vector<int> atom_of(h.n), pidx((size_t)h.natoms * h.natoms, -1);
        for (int a = 0; a < h.natoms; ++a)
            for (int m = 0; m < h.nstates[h.atom_type[a]]; ++m) atom_of[h.cumul[a] + m] = a;
        for (size_t q = 0; q < pairs.size(); ++q)
            pidx[(size_t)pairs[q].i * h.natoms + pairs[q].j] = (int)q;
        const int *ao = nullptr, *pi = nullptr;
        RCHK(dev_upload(c, &ao, atom_of));
        RCHK(dev_upload(c, &pi, pidx));
        c->atom_of_state = ao;
        c->pair_index = pi;
    }
    c->hpairs = pairs;
    c->max_img = 0;
    for (const PairRec &pr : pairs) c->max_img = std::max(c->max_img, pr.img_cnt);
    c->nimgs = (long)imgs.size();
    RCHK(dev_alloc(c, &c->pairs, pairs.size()));
    RCHK(dev_alloc(c, &c->imgs, imgs.size()));
    RCHK(dev_alloc(c, &c->phase, phase.size()));
    HIPCHK(ctx_memcpy(c, c->pairs, pairs.data(), pairs.size() * sizeof(PairRec), hipMemcpyHostToDevice));
    if (!imgs.empty()) {
        HIPCHK(ctx_memcpy(c, c->imgs, imgs.data(), imgs.size() * sizeof(ImgRec), hipMemcpyHostToDevice));
        HIPCHK(ctx_memcpy(c, c->phase, phase.data(), phase.size() * sizeof(double2), hipMemcpyHostToDevice));
    }
    RCHK(build_candidates(c, img_pair));
    return IMAGO_OK;
}

// The extended view for the core orthogonalisation: the same pairs with
// block offsets for each type's valence + core states (in the valence
// layout's block order), the atom and extended state of every function
// (valence functions first, then core functions), and a DevSys whose state
// counts are the extended ones.
static int build_core_tables(imago_ctx *c)
{
    const HostSys &h = c->h;
    const int N = h.n + h.core_dim;
    std::vector<int> next(h.nt);
    for (int t = 0; t < h.nt; ++t) next[t] = h.nstates[t] + h.ncore[t];
    std::vector<PairRec> pe = c->hpairs;
    std::vector<int> order(pe.size());
    for (size_t q = 0; q < pe.size(); ++q) order[q] = (int)q;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return pe[a].blk < pe[b].blk; });
    long off = 0;
    for (int q : order) {
        const long sz = (long)next[h.atom_type[pe[q].i]] * next[h.atom_type[pe[q].j]];
        if (off + sz >= (1L << 31)) return fail(IMAGO_E_ARG, "extended block layout exceeds 2^31 entries");
        pe[q].blk = (int)off;
        off += sz;
    }
    c->nbe_ext = off;
    std::vector<int> fa(N), fs(N), ans(h.natoms);
    for (int a = 0; a < h.natoms; ++a) {
        const int t = h.atom_type[a];
        ans[a] = next[t];
        for (int m = 0; m < h.nstates[t]; ++m) {
            fa[h.cumul[a] + m] = a;
            fs[h.cumul[a] + m] = m;
        }
        for (int m = 0; m < h.ncore[t]; ++m) {
            fa[h.n + h.cumul_core[a] + m] = a;
            fs[h.n + h.cumul_core[a] + m] = h.nstates[t] + m;
        }
    }
    RCHK(dev_alloc(c, &c->pairs_ext, pe.size()));
    HIPCHK(ctx_memcpy(c, c->pairs_ext, pe.data(), pe.size() * sizeof(PairRec), hipMemcpyHostToDevice));
    RCHK(dev_upload(c, &c->fn_atom, fa));
    RCHK(dev_upload(c, &c->fn_state, fs));
    RCHK(dev_upload(c, &c->atom_nstates_ext, ans));
    c->d_ext = c->d;
    RCHK(dev_upload(c, &c->d_ext.type_nstates, next));
    c->d_ext.nbe = c->nbe_ext;
    c->d_ext.vale_dim = N;
    return IMAGO_OK;
}

static int lma_to_l_host(int lma) { return lma == 1 ? 0 : lma == 4 ? 1 : lma == 9 ? 2 : 3; }

// ---------------------------------------------------------------------------
// context

extern "C" int imago_ctx_create(int device, imago_ctx **out)
{
    if (!out) return fail(IMAGO_E_ARG, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(IMAGO_E_ARG, "device %d not present (%d devices)", device, ndev);
    imago_ctx *c = new imago_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipEventCreate(&c->ev2);
    if (e == hipSuccess) e = hipEventCreate(&c->ev3);
    if (e != hipSuccess) {
        delete c;
        return fail(IMAGO_E_HIP, "context set-up failed: %s", hipGetErrorString(e));
    }
    // rocBLAS and rocSOLVER without atomic accumulation (split-K GEMMs
    // summing partial tiles in arrival order): results then do not depend
    // on how work groups are scheduled, which other contexts on the device
    // change from run to run
    if (rocblas_create_handle(&c->blas) != rocblas_status_success ||
        rocblas_set_stream(c->blas, c->stream) != rocblas_status_success ||
        rocblas_set_atomics_mode(c->blas, rocblas_atomics_not_allowed) != rocblas_status_success) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(IMAGO_E_HIP, "rocblas handle creation failed");
    }
    {
        const char *e1 = getenv("IMAGO_HEEVD"), *e2 = getenv("IMAGO_POTRF");
        c->tune_heevd = (e1 && strcmp(e1, "rocsolver") == 0) ? 1 : 0;
        c->tune_potrf = (e2 && strcmp(e2, "rocsolver") == 0) ? 1 : 0;
        const char *e3 = getenv("IMAGO_STEDC");
        c->tune_tridiag = (e3 && strcmp(e3, "rocsolver") == 0) ? 1 : 0;
        const char *e4 = getenv("IMAGO_GRAPH");
        c->tune_graph = (e4 && strcmp(e4, "0") == 0) ? 0 : 1;
        const char *e5 = getenv("IMAGO_BACKTRANSFORM");
        c->tune_wy = (e5 && strcmp(e5, "ut") == 0) ? 0 : 1;
    }
    *out = c;
    return IMAGO_OK;
}

extern "C" int imago_tune(imago_ctx *c, int param, int value)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    switch (param) {
    case IMAGO_TUNE_POTRF_BLOCK:
        if (value < 1 || value > kPotrfBlock)
            return fail(IMAGO_E_ARG, "Cholesky block %d outside 1..%d", value, kPotrfBlock);
        c->tune_potrf_block = value;
        break;
    case IMAGO_TUNE_EIGENSOLVER:
        if (value != 0 && value != 1) return fail(IMAGO_E_ARG, "eigensolver %d", value);
        c->tune_heevd = value;
        break;
    case IMAGO_TUNE_CHOLESKY:
        if (value != 0 && value != 1) return fail(IMAGO_E_ARG, "Cholesky %d", value);
        c->tune_potrf = value;
        break;
    case IMAGO_TUNE_TRIDIAG:
        if (value != 0 && value != 1) return fail(IMAGO_E_ARG, "tridiagonal solver %d", value);
        c->tune_tridiag = value;
        break;
    case IMAGO_TUNE_GRAPH:
        if (value != 0 && value != 1) return fail(IMAGO_E_ARG, "graph replay %d", value);
        c->tune_graph = value;
        break;
    default:
        return fail(IMAGO_E_ARG, "unknown tuning parameter %d", param);
    }
    c->chol_valid = false;   // the next secular call factors the overlap again
    return IMAGO_OK;
}

static int dc_ensure_plan(imago_ctx *c, int n);

extern "C" int imago_tridiag_eig(imago_ctx *c, int n, int nb, const double *d, const double *e,
                                 double *w, double *Z)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (n < 1 || nb < 1 || !d || !w || !Z || (n > 1 && !e))
        return fail(IMAGO_E_ARG, "bad tridiagonal arguments");
    if (n > kDcMaxOrder) return fail(IMAGO_E_ARG, "order %d above %d", n, kDcMaxOrder);
    HIPCHK(hipSetDevice(c->device));
    RCHK(dc_ensure_plan(c, n));
    const size_t N = (size_t)n * nb;
    std::vector<double> eh(N, 0.0);
    for (int b = 0; b < nb; ++b)
        for (int i = 0; i + 1 < n; ++i) eh[(size_t)b * n + i] = e[(size_t)b * (n - 1) + i];
    double *dd = nullptr, *de = nullptr, *dZ = nullptr;
    char *wk = nullptr;
    RCHK(dev_alloc(c, &dd, N));
    RCHK(dev_alloc(c, &de, N));
    RCHK(dev_alloc(c, &dZ, N * n));
    RCHK(dev_alloc(c, &wk, dc_work_bytes(n, nb)));
    HIPCHK(hipMemcpyAsync(dd, d, N * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(de, eh.data(), N * 8, hipMemcpyHostToDevice, c->stream));
    DcWork dw;
    dc_work_carve(wk, n, nb, dw);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(launch_tridiag_dc(c->dcplan, nb, dd, de, dw, dd, dZ, true, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(w, dd, N * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(Z, dZ, N * n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->timing[17] = ms;
    for (void *p : {(void *)dd, (void *)de, (void *)dZ, (void *)wk}) dev_release(c, p);
    return IMAGO_OK;
}

extern "C" int imago_ctx_destroy(imago_ctx *c)
{
    if (!c) return IMAGO_OK;
    // a runtime that is already being torn down (a destroy from an exit
    // handler) refuses the device: the device memory goes with the process,
    // only the host side is freed
    if (hipSetDevice(c->device) != hipSuccess) {
        delete c;
        return IMAGO_OK;
    }
    (void)hipStreamSynchronize(c->stream);
    ctx_free_system(c);
    c->sysbuf.purge_cache();
    if (c->blas) rocblas_destroy_handle(c->blas);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->ev3) (void)hipEventDestroy(c->ev3);
    if (c->blas2) rocblas_destroy_handle(c->blas2);
    if (c->evf) (void)hipEventDestroy(c->evf);
    if (c->evj) (void)hipEventDestroy(c->evj);
    if (c->evh0) (void)hipEventDestroy(c->evh0);
    if (c->evh1) (void)hipEventDestroy(c->evh1);
    for (hipEvent_t ev : c->evs)
        if (ev) (void)hipEventDestroy(ev);
    if (c->dcplan.dev) (void)hipFree(c->dcplan.dev);
    if (c->dcg.exec) (void)hipGraphExecDestroy(c->dcg.exec);
    if (c->dcg.work) (void)hipFree(c->dcg.work);
    if (c->stream2) {
        (void)hipStreamSynchronize(c->stream2);
        (void)hipStreamDestroy(c->stream2);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return IMAGO_OK;
}

extern "C" int imago_set_system(imago_ctx *c, const imago_system_desc *s, int *vale_dim,
                                int *pot_dim)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    ctx_free_system(c);
    RCHK(build_host(c->h, s));
    HostSys &h = c->h;
    DevSys &d = c->d;
    d = DevSys{};
    d.num_types = h.nt;
    d.max_alphas = h.maxa;
    d.max_states = h.maxs;
    d.num_elements = h.nel;
    d.max_pot_alphas = h.maxpa;
    d.num_atoms = h.natoms;
    d.num_pot_sites = h.nsites;
    d.num_pot_types = h.npt;
    d.num_cells = h.ncells;
    d.nk = h.nk;
    d.vale_dim = h.n;
    d.pot_dim = h.pot_dim;
    d.gamma = h.gamma;
    std::vector<int> nal(h.nt);
    for (int t = 0; t < h.nt; ++t) nal[t] = h.noa[4 * t];
    RCHK(dev_upload(c, &d.type_alphas, h.alphas));
    RCHK(dev_upload(c, &d.type_basis, h.basis));
    RCHK(dev_upload(c, &d.type_lm, h.lm));
    RCHK(dev_upload(c, &d.type_lma, h.lma));
    RCHK(dev_upload(c, &d.type_nalpha, nal));
    RCHK(dev_upload(c, &d.type_nstates, h.nstates));
    RCHK(dev_upload(c, &d.type_element, h.type_el));
    RCHK(dev_upload(c, &d.adist, h.adist));
    RCHK(dev_upload(c, &d.acenter, h.acenter));
    RCHK(dev_upload(c, &d.andist, h.andist));
    RCHK(dev_upload(c, &d.apdist, h.apdist));
    RCHK(dev_upload(c, &d.atom_type, h.atom_type));
    RCHK(dev_upload(c, &d.cumul, h.cumul));
    RCHK(dev_upload(c, &d.atom_pos, h.atom_pos));
    RCHK(dev_upload(c, &d.pot_pos, h.site_pos));
    RCHK(dev_upload(c, &d.pot_site_type, h.site_type));
    RCHK(dev_upload(c, &d.pot_z, h.pot_z));
    RCHK(dev_upload(c, &d.pot_nuc_alpha, h.pot_nuc));
    RCHK(dev_upload(c, &d.pot_alphas, h.pot_alphas));
    RCHK(dev_upload(c, &d.pot_el, h.pot_el));
    RCHK(dev_upload(c, &d.pot_first, h.pot_first));
    RCHK(dev_upload(c, &d.pot_mult, h.pot_mult));
    RCHK(dev_upload(c, &d.term_type, h.term_type));
    RCHK(dev_upload(c, &d.term_alpha, h.term_alpha));
    RCHK(dev_upload(c, &d.cell_dims, h.cell_dims));
    RCHK(dev_upload(c, &d.pot_nalpha, h.pot_nalpha));
    {
        const int *tt0 = nullptr;
        RCHK(dev_upload(c, &tt0, h.type_term0));
        c->type_term0 = tt0;
    }
    const auto t1 = clk::now();
    RCHK(build_geometry(c));
    if (h.core_dim > 0) RCHK(build_core_tables(c));
    const auto t2 = clk::now();
    c->timing[7] = std::chrono::duration<double, std::milli>(t2 - t1).count();
    int lmax = 0;
    for (int t = 0; t < h.nt; ++t)
        for (int a = 0; a < h.noa[4 * t]; ++a) lmax = std::max(lmax, lma_to_l_host(h.lma[t * h.maxa + a]));
    c->lmax = lmax;
    c->ne = (h.maxs * h.maxs + 63) / 64;
    c->has_system = true;
    c->timing[6] = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    if (vale_dim) *vale_dim = h.n;
    if (pot_dim) *pot_dim = h.pot_dim;
    return IMAGO_OK;
}

// ---------------------------------------------------------------------------
// assembly

static AssembleLaunch make_launch(imago_ctx *c, double *out, long term_stride, int term0,
                                  int nterms, size_t lds, bool ext = false)
{
    const HostSys &h = c->h;
    return AssembleLaunch{ext ? c->d_ext : c->d, ext ? c->pairs_ext : c->pairs, c->imgs, c->phase,
                          c->cands, c->cand_off, term0, out,
                          term_stride, c->npairs, nterms, lds, c->stream, c->lmax,
                          c->ne <= 4 ? 4 : 16, h.gamma != 0, c->type_term0, h.npt, c->pair_order,
                          c->ep_tuples, c->ep_tseg, c->ep_item_tuple, c->ep_blocks,
                          c->ep_nitems, h.nk};
}


// Screening of the EP terms (k_ep_screen) and their primitive blocks
// (k_ep_prims): count the tuples and items of every (potential type, pair),
// scan the counts into offsets, fill the tuples, then evaluate every item.
// Runs once per system: none of it depends on the potential coefficients.
static int ep_prepare(imago_ctx *c)
{
    if (c->ep_screened) return IMAGO_OK;
    const HostSys &h = c->h;
    const size_t lds = ep_screen_lds_bytes(h.maxa);
    if (lds > 64 * 1024)
        return fail(IMAGO_E_ARG, "EP screening needs %zu bytes of LDS (max_alphas %d)", lds, h.maxa);
    const long ntask = (long)c->npairs * h.npt;
    // one 64-lane work-group per task; a dispatch counts its work-items in 32 bits
    if (ntask * 64 > 0xffffffffL)
        return fail(IMAGO_E_ARG, "%ld (pair, potential type) tasks exceed one dispatch", ntask);
    long *tcnt = nullptr, *icnt = nullptr, *toff = nullptr, *ioff = nullptr;
    RCHK(dev_alloc(c, &tcnt, (size_t)ntask + 1));
    RCHK(dev_alloc(c, &icnt, (size_t)ntask + 1));
    RCHK(dev_alloc(c, &toff, (size_t)ntask + 1));
    RCHK(dev_alloc(c, &ioff, (size_t)ntask + 1));
    HIPCHK(hipMemsetAsync(tcnt + ntask, 0, sizeof(long), c->stream));
    HIPCHK(hipMemsetAsync(icnt + ntask, 0, sizeof(long), c->stream));
    AssembleLaunch L = make_launch(c, nullptr, 0, 0, 0, 0);
    HIPCHK(launch_ep_screen(L, false, tcnt, icnt, lds));
    size_t tmp_bytes = 0;
    HIPCHK(ep_exclusive_scan(tcnt, toff, ntask + 1, nullptr, &tmp_bytes, c->stream));
    char *tmp = nullptr;
    RCHK(dev_alloc(c, &tmp, tmp_bytes));
    HIPCHK(ep_exclusive_scan(tcnt, toff, ntask + 1, tmp, &tmp_bytes, c->stream));
    HIPCHK(ep_exclusive_scan(icnt, ioff, ntask + 1, tmp, &tmp_bytes, c->stream));
    long tot[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&tot[0], toff + ntask, sizeof(long), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&tot[1], ioff + ntask, sizeof(long), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (void *p : {(void *)tcnt, (void *)icnt, (void *)tmp}) dev_release(c, p);
    if (tot[1] >= (1L << 31))
        return fail(IMAGO_E_ARG, "%ld EP primitive items exceed the 2^31 item index", tot[1]);
    const int sh = (c->lmax + 1) * (c->lmax + 1);
    RCHK(dev_alloc(c, &c->ep_tuples, (size_t)std::max(tot[0], 1L)));
    RCHK(dev_alloc(c, &c->ep_item_tuple, (size_t)std::max(tot[1], 1L)));
    RCHK(dev_alloc(c, &c->ep_blocks, (size_t)std::max(tot[1], 1L) * sh * sh));
    c->ep_tseg = toff;
    c->ep_ntuples = tot[0];
    c->ep_nitems = tot[1];
    L = make_launch(c, nullptr, 0, 0, 0, 0);
    HIPCHK(launch_ep_screen(L, true, toff, ioff, lds));
    HIPCHK(launch_ep_prims(L));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ep_iseg = ioff;
    c->ep_screened = true;
    return IMAGO_OK;
}

// Run tables of the direct EP path, once per system: every item's lattice
// image as a key, a stable radix sort by it (so each pair's items are
// contiguous and image-ordered, types in order within an image), and the
// runs of at most kEpdRunItems over each pair's items: H-sum runs with
// their partial offsets, trace runs, and every item's potential exponent.
static int ep_build_runs(imago_ctx *c)
{
    if (c->ep_runs) return IMAGO_OK;
    const HostSys &h = c->h;
    const int np = c->npairs;
    const long ni = c->ep_nitems, C = kEpdRunItems;
    if (c->nimgs >= (1L << 30)) return fail(IMAGO_E_ARG, "%ld lattice images", c->nimgs);
    int bits = 1;
    while ((1L << bits) <= c->nimgs) ++bits;
    int *key = nullptr, *skey = nullptr, *val = nullptr;
    long *hcnt = nullptr, *wcnt = nullptr, *tcnt = nullptr, *hoff = nullptr, *woff = nullptr,
         *toff = nullptr;
    const size_t nis = (size_t)std::max(ni, 1L);
    RCHK(dev_alloc(c, &key, nis));
    RCHK(dev_alloc(c, &skey, nis));
    RCHK(dev_alloc(c, &val, nis));
    RCHK(dev_alloc(c, &c->pitem, nis));
    RCHK(dev_alloc(c, &c->item_jexp, nis));
    RCHK(dev_alloc(c, &c->ival, nis));
    RCHK(dev_alloc(c, &hcnt, (size_t)np + 1));
    RCHK(dev_alloc(c, &wcnt, (size_t)np + 1));
    RCHK(dev_alloc(c, &tcnt, (size_t)np + 1));
    RCHK(dev_alloc(c, &hoff, (size_t)np + 1));
    RCHK(dev_alloc(c, &woff, (size_t)np + 1));
    RCHK(dev_alloc(c, &toff, (size_t)np + 1));
    HIPCHK(launch_epd_item_keys(c->ep_tuples, c->ep_item_tuple, c->pairs, ni, np, key, val,
                                c->item_jexp, c->stream));
    size_t sb = 0, t1 = 0, t2 = 0, t3 = 0;
    HIPCHK(int_sort_pairs(key, skey, val, c->pitem, ni, bits, nullptr, &sb, c->stream));
    HIPCHK(ep_exclusive_scan(hcnt, hoff, np + 1, nullptr, &t1, c->stream));
    HIPCHK(ep_exclusive_scan(wcnt, woff, np + 1, nullptr, &t2, c->stream));
    HIPCHK(ep_exclusive_scan(tcnt, toff, np + 1, nullptr, &t3, c->stream));
    const size_t tb = std::max(std::max(sb, t1), std::max(t2, t3));
    char *tmp = nullptr;
    RCHK(dev_alloc(c, &tmp, tb));
    size_t tbx = tb;
    if (ni > 0) HIPCHK(int_sort_pairs(key, skey, val, c->pitem, ni, bits, tmp, &tbx, c->stream));
    HIPCHK(hipMemsetAsync(hcnt + np, 0, sizeof(long), c->stream));
    HIPCHK(hipMemsetAsync(wcnt + np, 0, sizeof(long), c->stream));
    HIPCHK(hipMemsetAsync(tcnt + np, 0, sizeof(long), c->stream));
    HIPCHK(launch_epd_pair_runs(false, skey, ni, c->pairs, np, c->d, C, hcnt, wcnt, tcnt, nullptr,
                                nullptr, nullptr, nullptr, nullptr, c->stream));
    tbx = tb;
    HIPCHK(ep_exclusive_scan(hcnt, hoff, np + 1, tmp, &tbx, c->stream));
    tbx = tb;
    HIPCHK(ep_exclusive_scan(wcnt, woff, np + 1, tmp, &tbx, c->stream));
    tbx = tb;
    HIPCHK(ep_exclusive_scan(tcnt, toff, np + 1, tmp, &tbx, c->stream));
    long tot[3] = {0, 0, 0};
    HIPCHK(hipMemcpyAsync(&tot[0], hoff + np, sizeof(long), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&tot[1], woff + np, sizeof(long), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&tot[2], toff + np, sizeof(long), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (std::max(tot[0], tot[2]) * 64 > 0xffffffffL)
        return fail(IMAGO_E_ARG, "%ld / %ld EP runs exceed one dispatch", tot[0], tot[2]);
    RCHK(dev_alloc(c, &c->hruns, (size_t)std::max(tot[0], 1L)));
    RCHK(dev_alloc(c, &c->truns, (size_t)std::max(tot[2], 1L)));
    RCHK(dev_alloc(c, &c->hpart, (size_t)std::max(tot[1], 1L)));
    HIPCHK(launch_epd_pair_runs(true, skey, ni, c->pairs, np, c->d, C, nullptr, nullptr, nullptr,
                                hoff, woff, toff, c->hruns, c->truns, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (void *p : {(void *)key, (void *)skey, (void *)val, (void *)hcnt, (void *)wcnt,
                    (void *)tcnt, (void *)toff, (void *)tmp})
        dev_release(c, p);
    c->hrun_lo = hoff;
    c->hrun_woff = woff;
    c->nhruns = tot[0];
    c->ntruns = tot[2];
    c->ep_runs = true;
    return IMAGO_OK;
}

// ext: the extended (valence + core) view into an extended block layout
static int launch_assemble(imago_ctx *c, int kind, double *out, long term_stride, int term0,
                           int nterms, bool ext = false)
{
    const HostSys &h = c->h;
    const size_t lds =
        kind == IMAGO_EP ? ep_store_lds_bytes(h.maxs, h.nk, h.gamma || h.nk <= 4)
                         : assemble_lds_bytes(c->lmax, h.maxa, h.maxs, h.nk);
    if (lds > 64 * 1024)
        return fail(IMAGO_E_ARG, "LDS requirement %zu bytes exceeds 64 KiB (max_states %d, nk %d)",
                    lds, h.maxs, h.nk);
    AssembleLaunch L = make_launch(c, out, term_stride, term0, nterms, lds, ext);
    if (kind == IMAGO_NP && c->lmax == 3) {
        // the f-shell nuclear blocks accumulate in device memory
        RCHK(dev_alloc(c, &L.acc_work, (size_t)kAccSlots * kAccSlotDoubles));
    }
    hipError_t e;
    switch (kind) {
    case IMAGO_OL: e = launch_assemble_ol(L); break;
    case IMAGO_KE: e = launch_assemble_ke(L); break;
    case IMAGO_NP: e = launch_assemble_np(L); break;
    default: e = launch_assemble_ep(L); break;
    }
    if (e == hipErrorInvalidValue)
        return fail(IMAGO_E_ARG, "no assembly kernel for kind %d lmax %d", kind, c->lmax);
    HIPCHK(e);
    HIPCHK(hipGetLastError());
    if (L.acc_work) {
        HIPCHK(hipStreamSynchronize(c->stream));
        dev_release(c, L.acc_work);
    }
    return IMAGO_OK;
}

static long panel_doubles(const imago_ctx *c)
{
    return c->h.gamma ? c->nbe : 2 * c->nbe * c->h.nk;
}

static int core_assemble(imago_ctx *c, int kind);

extern "C" int imago_assemble(imago_ctx *c, int kind)
{
    if (!c || !c->has_system) return fail(IMAGO_E_STATE, "imago_assemble before imago_set_system");
    if (kind < IMAGO_OL || kind > IMAGO_EP_DIRECT)
        return fail(IMAGO_E_ARG, "unknown integral kind %d", kind);
    if (kind == IMAGO_EP_DIRECT && c->h.core_dim > 0)
        return fail(IMAGO_E_ARG, "the direct EP path has no core orthogonalisation: assemble "
                                 "IMAGO_EP for a system with core states");
    if (kind == IMAGO_EP_DIRECT) {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        RCHK(ep_prepare(c));
        RCHK(ep_build_runs(c));
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->timing[0] = c->timing[5] = ms;
        c->timing[4] = 0.0;
        c->ep_direct = true;
        return IMAGO_OK;
    }
    HIPCHK(hipSetDevice(c->device));
    const int slot = kind - 1;
    const long per = panel_doubles(c);
    const int nterms = kind == IMAGO_EP ? c->h.pot_dim : 1;
    if (!c->panel[slot]) RCHK(dev_alloc(c, &c->panel[slot], (size_t)per * nterms));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if (c->h.core_dim > 0) {
        // extended assembly and orthogonalisation against the core states
        RCHK(core_assemble(c, kind));
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->timing[0] = ms;
        c->assembled[slot] = true;
        c->filled[slot].assign((size_t)nterms * c->h.nk, 1);
        if (kind == IMAGO_OL) c->chol_valid = false;
        return IMAGO_OK;
    }
    if (kind == IMAGO_EP) {
        // stage 5: screening and primitive blocks (first EP assembly of a
        // system); stage 4: the k_ep_store launch alone; stage 0: both
        const bool screen = !c->ep_screened;
        RCHK(ep_prepare(c));
        HIPCHK(hipEventRecord(c->ev2, c->stream));
        RCHK(launch_assemble(c, kind, c->panel[slot], per, 0, nterms));
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms_s = 0, ms_e = 0;
        HIPCHK(hipEventElapsedTime(&ms_s, c->ev0, c->ev2));
        HIPCHK(hipEventElapsedTime(&ms_e, c->ev2, c->ev1));
        c->timing[5] = screen ? ms_s : 0.0;
        c->timing[4] = ms_e;
    } else {
        RCHK(launch_assemble(c, kind, c->panel[slot], per, 0, 1));
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->timing[0] = ms;
    c->assembled[slot] = true;
    c->filled[slot].assign((size_t)nterms * c->h.nk, 1);
    if (kind == IMAGO_OL) c->chol_valid = false;
    return IMAGO_OK;
}

extern "C" int imago_get_packed(imago_ctx *c, int kind, int term, int kpoint, double *out)
{
    if (!c || !c->has_system) return fail(IMAGO_E_STATE, "no system");
    if (kind < IMAGO_OL || kind > IMAGO_EP || !c->assembled[kind - 1])
        return fail(IMAGO_E_STATE, "kind %d not assembled", kind);
    const int nterms = kind == IMAGO_EP ? c->h.pot_dim : 1;
    if (term < 0 || term >= nterms || kpoint < 0 || kpoint >= c->h.nk || !out)
        return fail(IMAGO_E_ARG, "term %d / k-point %d out of range", term, kpoint);
    HIPCHK(hipSetDevice(c->device));
    const HostSys &h = c->h;
    const int dsz = h.gamma ? 1 : 2;
    const long per = panel_doubles(c);
    const long len = dsz * c->nbe;
    const double *src = c->panel[kind - 1] + (size_t)term * per + (size_t)kpoint * len;
    std::vector<double> blk(len);
    HIPCHK(ctx_memcpy(c, blk.data(), src, len * sizeof(double), hipMemcpyDeviceToHost));
    // block layout -> packed upper triangle, column by column (matrixSubs.F90:506)
    for (const PairRec &pr : c->hpairs) {
        const int ns1 = h.nstates[h.atom_type[pr.i]], ns2 = h.nstates[h.atom_type[pr.j]];
        const long r0 = h.cumul[pr.i], c0 = h.cumul[pr.j];
        for (int s2 = 0; s2 < ns2; ++s2)
            for (int s1 = 0; s1 < ns1; ++s1) {
                const long row = r0 + s1, col = c0 + s2;
                if (row > col) continue;
                const long pk = col * (col + 1) / 2 + row;
                const long be = (long)pr.blk + s1 + (long)ns1 * s2;
                for (int x = 0; x < dsz; ++x) out[dsz * pk + x] = blk[dsz * be + x];
            }
    }
    return IMAGO_OK;
}

extern "C" int imago_set_packed(imago_ctx *c, int kind, int term, int kpoint, const double *in)
{
    if (!c || !c->has_system) return fail(IMAGO_E_STATE, "imago_set_packed before imago_set_system");
    if (kind < IMAGO_OL || kind > IMAGO_EP) return fail(IMAGO_E_ARG, "unknown integral kind %d", kind);
    const HostSys &h = c->h;
    const int nterms = kind == IMAGO_EP ? h.pot_dim : 1;
    if (term < 0 || term >= nterms || kpoint < 0 || kpoint >= h.nk || !in)
        return fail(IMAGO_E_ARG, "term %d / k-point %d out of range", term, kpoint);
    HIPCHK(hipSetDevice(c->device));
    const int slot = kind - 1;
    const int dsz = h.gamma ? 1 : 2;
    const long per = panel_doubles(c);
    const long len = dsz * c->nbe;
    const long npk = (long)h.n * (h.n + 1) / 2;
    // packed upper triangle -> block layout (the inverse of
    // imago_get_packed); the lower triangles of diagonal blocks stay zero
    std::vector<double> blk(len, 0.0);
    std::vector<unsigned char> cover(npk, 0);
    for (const PairRec &pr : c->hpairs) {
        const int ns1 = h.nstates[h.atom_type[pr.i]], ns2 = h.nstates[h.atom_type[pr.j]];
        const long r0 = h.cumul[pr.i], c0 = h.cumul[pr.j];
        for (int s2 = 0; s2 < ns2; ++s2)
            for (int s1 = 0; s1 < ns1; ++s1) {
                const long row = r0 + s1, col = c0 + s2;
                if (row > col) continue;
                const long pk = col * (col + 1) / 2 + row;
                const long be = (long)pr.blk + s1 + (long)ns1 * s2;
                for (int x = 0; x < dsz; ++x) blk[dsz * be + x] = in[dsz * pk + x];
                cover[pk] = 1;
            }
    }
    // atom pairs beyond every lattice image's reach have no block: their
    // entries must be zero, as the assembly leaves them
    for (long pk = 0; pk < npk; ++pk)
        if (!cover[pk])
            for (int x = 0; x < dsz; ++x)
                if (in[dsz * pk + x] != 0.0)
                    return fail(IMAGO_E_ARG,
                                "packed entry %ld is nonzero but its atom pair has no block in "
                                "this system's pair list",
                                pk);
    if (!c->panel[slot]) {
        RCHK(dev_alloc(c, &c->panel[slot], (size_t)per * nterms));
        c->filled[slot].assign((size_t)nterms * h.nk, 0);
    }
    if (c->filled[slot].size() != (size_t)nterms * h.nk)
        c->filled[slot].assign((size_t)nterms * h.nk, 0);
    double *dst = c->panel[slot] + (size_t)term * per + (size_t)kpoint * len;
    HIPCHK(hipMemcpyAsync(dst, blk.data(), len * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->filled[slot][(size_t)term * h.nk + kpoint] = 1;
    bool all = true;
    for (unsigned char f : c->filled[slot]) all = all && f;
    c->assembled[slot] = all;
    if (kind == IMAGO_OL) c->chol_valid = false;
    return IMAGO_OK;
}

// ---------------------------------------------------------------------------
// secular equation: H(k) build (HBM-bound stream over the potential panel)

// packed H(k) = NP + KE + sum_j c_j V_j with readPackedMatrixAccum's rule
// that a factor of exactly 0.0 adds unscaled (matrixSubs.F90:369-373).
// len = doubles per term (all k); each thread owns 2 consecutive doubles.
__global__ __launch_bounds__(256) void k_hbuild(const double *__restrict__ np,
                                                const double *__restrict__ ke,
                                                const double *__restrict__ V,
                                                const double *__restrict__ coef, int pot_dim,
                                                long len, double *__restrict__ H)
{
#pragma clang fp contract(off)
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len) return;
    double a = np[e];
    a = a + ke[e];
    int j = 0;
    for (; j + 4 <= pot_dim; j += 4) {
        double v0 = V[(size_t)(j + 0) * len + e];
        double v1 = V[(size_t)(j + 1) * len + e];
        double v2 = V[(size_t)(j + 2) * len + e];
        double v3 = V[(size_t)(j + 3) * len + e];
        const double c0 = coef[j], c1 = coef[j + 1], c2 = coef[j + 2], c3 = coef[j + 3];
        if (c0 != 0.0) v0 = v0 * c0;
        a = a + v0;
        if (c1 != 0.0) v1 = v1 * c1;
        a = a + v1;
        if (c2 != 0.0) v2 = v2 * c2;
        a = a + v2;
        if (c3 != 0.0) v3 = v3 * c3;
        a = a + v3;
    }
    for (; j < pot_dim; ++j) {
        double v = V[(size_t)j * len + e];
        const double cj = coef[j];
        if (cj != 0.0) v = v * cj;
        a = a + v;
    }
    H[e] = a;
}

__device__ __forceinline__ int cumul_states(int a, const int *cumul, int natoms, int n)
{
    return (a + 1 < natoms ? cumul[a + 1] : n) - cumul[a];
}

// Block layout -> full n x n Hermitian (symmetric for Gamma) column-major
// matrix per k, one thread per (row, col): the entry comes from the block
// of the pair (atom of min(row, col), atom of max(row, col)), conjugated
// below the diagonal (unpackMatrix, matrixSubs.F90:405-460).
template <bool COMPLEX>
__global__ void k_unpack(const double *__restrict__ blk, long nbe, int n, int nk,
                         const PairRec *__restrict__ pairs, const int *__restrict__ atom_of,
                         const int *__restrict__ pair_index, const int *__restrict__ cumul,
                         int natoms, double *__restrict__ full)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nn = (long)n * n;
    if (t >= nn * nk) return;
    const int k = (int)(t / nn);
    const long rc = t % nn;
    const int r = (int)(rc % n), col = (int)(rc / n);
    const int lo = min(r, col), hi = max(r, col);
    const int a = atom_of[lo], b = atom_of[hi];
    const PairRec pr = pairs[pair_index[(size_t)a * natoms + b]];
    const int ns1 = cumul_states(a, cumul, natoms, n);
    const long e = (long)pr.blk + (lo - cumul[a]) + (long)ns1 * (hi - cumul[b]);
    if (COMPLEX) {
        const double2 v = reinterpret_cast<const double2 *>(blk)[(size_t)k * nbe + e];
        reinterpret_cast<double2 *>(full)[t] = make_double2(v.x, r <= col ? v.y : -v.y);
    } else {
        full[t] = blk[(size_t)k * nbe + e];
    }
}

static int ensure_secular_buffers(imago_ctx *c)
{
    const HostSys &h = c->h;
    const long nn = (long)h.n * h.n;
    if (!c->coef) RCHK(dev_alloc(c, &c->coef, std::max(h.pot_dim, 1)));
    if (!c->blockH) RCHK(dev_alloc(c, &c->blockH, (size_t)panel_doubles(c)));
    if (!c->evals) RCHK(dev_alloc(c, &c->evals, (size_t)h.nk * h.n));
    if (!c->ework) RCHK(dev_alloc(c, &c->ework, (size_t)h.nk * h.n));
    if (!c->info) RCHK(dev_alloc(c, &c->info, (size_t)h.nk));
    if (h.gamma) {
        if (!c->Hreal) RCHK(dev_alloc(c, &c->Hreal, (size_t)nn));
        if (!c->Sreal) RCHK(dev_alloc(c, &c->Sreal, (size_t)nn));
    } else {
        if (!c->Hfull) RCHK(dev_alloc(c, &c->Hfull, (size_t)nn * h.nk));
        if (!c->Sfull) RCHK(dev_alloc(c, &c->Sfull, (size_t)nn * h.nk));
    }
    const size_t mat_bytes = (size_t)nn * (h.gamma ? sizeof(double) : sizeof(double2) * h.nk);
    if (!c->Sinv) RCHK(dev_alloc(c, (char **)&c->Sinv, mat_bytes));
    if (!c->Gtmp) RCHK(dev_alloc(c, (char **)&c->Gtmp, mat_bytes));
    return IMAGO_OK;
}

// Generalized Hermitian/symmetric eigensolve H c = e S c for a batch of
// problems, the algorithm of LAPACK zhegv/dsygv itype=1 (interfaces.F90:
// 73-200) with every stage on BLAS3 kernels:
//   S = U^H U (potrf), C = U^-H H U^-1 (two trsm), C = Y E Y^H (heevd,
//   divide and conquer), X = U^-1 Y (trsm).
// factor_gv_batched runs the Cholesky step on B (S on entry, U on exit);
// solve_gv_factored runs the rest on A (H on entry, X on exit) given U.
// Matrices are n x n column-major with stride n*n; info[k] follows zhegv:
// n + i when the leading minor i of S(k) is not positive definite, i when
// heevd fails.
// C = alpha op(A) op(B) + beta C over a strided batch, complex or real
static rocblas_status gemm_sb(rocblas_handle hb, bool cx, rocblas_operation ta,
                              rocblas_operation tb, int m, int n, int k, double alpha,
                              const void *A, int lda, rocblas_stride sa, const void *B, int ldb,
                              rocblas_stride sb, double beta, void *C, int ldc, rocblas_stride sc,
                              int batch)
{
    if (cx) {
        const rocblas_double_complex al(alpha, 0.0), be(beta, 0.0);
        return rocblas_zgemm_strided_batched(hb, ta, tb, m, n, k, &al,
                                             (const rocblas_double_complex *)A, lda, sa,
                                             (const rocblas_double_complex *)B, ldb, sb, &be,
                                             (rocblas_double_complex *)C, ldc, sc, batch);
    }
    const rocblas_operation ra =
        ta == rocblas_operation_conjugate_transpose ? rocblas_operation_transpose : ta;
    const rocblas_operation rb =
        tb == rocblas_operation_conjugate_transpose ? rocblas_operation_transpose : tb;
    return rocblas_dgemm_strided_batched(hb, ra, rb, m, n, k, &alpha, (const double *)A, lda, sa,
                                         (const double *)B, ldb, sb, &beta, (double *)C, ldc, sc,
                                         batch);
}

// Orthogonalisation against the core states (orthoOL / ortho,
// integrals.F90:3608-3910): for each term the extended matrix is assembled
// into an extended block layout, spread into dense [valence | core]
// matrices X per k-point, and
//   X'vv = Xvv + Scv^H (Xcc Scv) - Scv^H Xcv - Xcv^H Scv
// is formed by four batched GEMMs over the k-points (only Xvv is written;
// Scv, Xcv and Xcc are read from disjoint blocks), then the upper valence
// triangle goes back to the valence block layout with the 1e-10 flush.
// For the overlap X is S itself, which gives Svv - 2 Scv^H Scv + Scv^H Scc
// Scv; its dense matrices are kept (core_S) for the other kinds' Scv.
static int core_assemble(imago_ctx *c, int kind)
{
    const HostSys &h = c->h;
    const int n = h.n, cd = h.core_dim, N = n + cd, nk = h.nk, na = h.natoms;
    const bool cx = !h.gamma;
    const int dsz = cx ? 2 : 1;
    if (kind != IMAGO_OL && !c->core_S)
        return fail(IMAGO_E_STATE, "a system with core states needs IMAGO_OL assembled before the "
                                   "other kinds (their core-valence overlap)");
    const long per_ext = (long)dsz * c->nbe_ext * nk;
    const long per = panel_doubles(c);
    const size_t dn = (size_t)dsz * N * N * nk;
    const int nterms = kind == IMAGO_EP ? h.pot_dim : 1;
    if (kind == IMAGO_EP) RCHK(ep_prepare(c));
    double *ext = nullptr, *T = nullptr, *X = nullptr;
    RCHK(dev_alloc(c, &ext, (size_t)per_ext));
    RCHK(dev_alloc(c, &T, (size_t)dsz * cd * n * nk));
    if (kind == IMAGO_OL) {
        if (!c->core_S) RCHK(dev_alloc(c, &c->core_S, dn));
        X = c->core_S;
    } else {
        RCHK(dev_alloc(c, &X, dn));
    }
    const double *S = c->core_S;
    const rocblas_stride sN = (rocblas_stride)N * N, sT = (rocblas_stride)cd * n;
    const rocblas_operation opN = rocblas_operation_none,
                            opC = rocblas_operation_conjugate_transpose;
    // offsets (elements) of the blocks in a dense N x N column-major matrix
    const size_t cv = (size_t)n, cc = (size_t)n + (size_t)n * N;
    auto at = [&](const double *base, size_t el) { return base + (size_t)dsz * el; };
    for (int term = 0; term < nterms; ++term) {
        // term_stride 0: the term's matrix lands at ext itself
        RCHK(launch_assemble(c, kind, ext, 0, term, 1, true));
        HIPCHK(launch_core_unpack(ext, c->nbe_ext, N, nk, cx, c->pairs_ext, c->fn_atom, c->fn_state,
                                  c->pair_index, c->atom_nstates_ext, na, X, c->stream));
        // T = Xcc Scv; Xvv += Scv^H T; Xvv -= Scv^H Xcv; Xvv -= Xcv^H Scv
        BLASCHK(gemm_sb(c->blas, cx, opN, opN, cd, n, cd, 1.0, at(X, cc), N, sN, at(S, cv), N, sN,
                        0.0, T, cd, sT, nk));
        BLASCHK(gemm_sb(c->blas, cx, opC, opN, n, n, cd, 1.0, at(S, cv), N, sN, T, cd, sT, 1.0, X,
                        N, sN, nk));
        BLASCHK(gemm_sb(c->blas, cx, opC, opN, n, n, cd, -1.0, at(S, cv), N, sN, at(X, cv), N, sN,
                        1.0, X, N, sN, nk));
        BLASCHK(gemm_sb(c->blas, cx, opC, opN, n, n, cd, -1.0, at(X, cv), N, sN, at(S, cv), N, sN,
                        1.0, X, N, sN, nk));
        HIPCHK(launch_core_pack(X, N, n, nk, cx, c->pairs, c->nbe, c->atom_of_state, c->pair_index,
                                c->d.cumul, na, c->panel[kind - 1] + (size_t)term * per,
                                c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    dev_release(c, ext);
    dev_release(c, T);
    if (kind != IMAGO_OL) dev_release(c, X);
    return IMAGO_OK;
}

// The Cholesky factor S = U^H U of a batch, blocked right-looking as
// LAPACK zpotrf: for each block of kPotrfBlock columns the diagonal block
// by k_potf2_diag, the block row beside it by a triangular solve
// (U_jj^-H A_j,rest) and the trailing matrix by a rank-k update, each one
// batched call.  The update is a GEMM over the whole trailing square
// (rocBLAS's herk runs at a fraction of its speed at these shapes), so the
// strictly lower triangle is overwritten as well; every reader of the
// factor takes the upper triangle only.  IMAGO_POTRF=rocsolver selects
// rocSOLVER's batched potrf instead where allow_rocsolver.
static int potrf_batched(imago_ctx *c, bool cx, int n, int nb, void *B, int *dinfo,
                         bool allow_rocsolver = true)
{
    const rocblas_stride nn = (rocblas_stride)n * n;
    if (c->tune_potrf == 1 && allow_rocsolver) {
        if (cx)
            BLASCHK(rocsolver_zpotrf_strided_batched(c->blas, rocblas_fill_upper, n,
                                                     (rocblas_double_complex *)B, n, nn, dinfo, nb));
        else
            BLASCHK(rocsolver_dpotrf_strided_batched(c->blas, rocblas_fill_upper, n, (double *)B, n,
                                                     nn, dinfo, nb));
        return IMAGO_OK;
    }
    const size_t es = cx ? sizeof(double2) : sizeof(double);
    char *base = static_cast<char *>(B);
    const int blk = std::max(1, std::min(c->tune_potrf_block, kPotrfBlock));
    for (int j = 0; j < n; j += blk) {
        const int jb = std::min(blk, n - j), rest = n - j - jb;
        HIPCHK(launch_potf2_diag(B, n, nb, j, jb, !cx, dinfo, c->stream));
        if (rest == 0) break;
        void *Ajj = base + ((size_t)j * n + j) * es;
        void *Ajr = base + ((size_t)(j + jb) * n + j) * es;
        void *Arr = base + ((size_t)(j + jb) * n + j + jb) * es;
        if (cx) {
            const rocblas_double_complex one(1.0, 0.0);
            const double m1 = -1.0, p1 = 1.0;
            BLASCHK(rocblas_ztrsm_strided_batched(
                c->blas, rocblas_side_left, rocblas_fill_upper,
                rocblas_operation_conjugate_transpose, rocblas_diagonal_non_unit, jb, rest, &one,
                (const rocblas_double_complex *)Ajj, n, nn, (rocblas_double_complex *)Ajr, n, nn,
                nb));
            BLASCHK(gemm_sb(c->blas, true, rocblas_operation_conjugate_transpose,
                            rocblas_operation_none, rest, rest, jb, m1, Ajr, n, nn, Ajr, n, nn, p1,
                            Arr, n, nn, nb));
        } else {
            const double one = 1.0, m1 = -1.0;
            BLASCHK(rocblas_dtrsm_strided_batched(c->blas, rocblas_side_left, rocblas_fill_upper,
                                                  rocblas_operation_transpose,
                                                  rocblas_diagonal_non_unit, jb, rest, &one,
                                                  (const double *)Ajj, n, nn, (double *)Ajr, n, nn,
                                                  nb));
            BLASCHK(gemm_sb(c->blas, false, rocblas_operation_transpose, rocblas_operation_none,
                            rest, rest, jb, m1, Ajr, n, nn, Ajr, n, nn, one, Arr, n, nn, nb));
        }
    }
    return IMAGO_OK;
}

static int factor_gv_batched(imago_ctx *c, bool cx, int n, int nb, void *B, int *dinfo, int *hinfo,
                             bool *bad)
{
    RCHK(potrf_batched(c, cx, n, nb, B, dinfo));
    HIPCHK(hipMemcpyAsync(hinfo, dinfo, sizeof(int) * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *bad = false;
    for (int k = 0; k < nb; ++k)
        if (hinfo[k] != 0) {
            hinfo[k] += n;
            *bad = true;
        }
    return IMAGO_OK;
}

// Hermitian (complex) or symmetric (real) eigendecomposition C = Y E Y^H of
// a batch, the three stages of LAPACK zheevd / dsyevd:
//   C = Q T Q^H, T real tridiagonal: k_hetrd_coop, one persistent
//     cooperative kernel for the whole batch;
//   T = Z E Z^T: this library's batched divide and conquer
//     (tridiag_dc.hip), or with IMAGO_TUNE_TRIDIAG = 1 rocSOLVER's dstedc
//     on each half of the batch's tridiagonals as one block-diagonal
//     problem (e[n-1] = 0 splits it), the halves on two streams, the
//     eigenvectors sorted back into their matrices (launch_trd_unmix);
//   Y = Q Z: left to finish_columns, which applies the reflectors kept in
//     c->pend in compact-WY blocks of kTrdBlock (factors by k_trd_larft) to
//     the columns its caller needs.
// A holds C on entry and Z on exit; evals the eigenvalues in ascending
// order; dinfo the per-matrix codes of rocSOLVER's dstedc (zero with the
// divide and conquer of this library).
// The default up to order 1024 (complex) / 2048 (real); IMAGO_HEEVD =
// rocsolver selects rocSOLVER's batched heevd / syevd instead.
__device__ __forceinline__ double part_add(double a, double b) { return a + b; }
__device__ __forceinline__ double2 part_add(double2 a, double2 b)
{
    return make_double2(a.x + b.x, a.y + b.y);
}

// W[m](r, j) = sum over c < nch in order of P[m][c](r, j) for r < kb
// (each partial K x N with leading dimension K, rows kb.. unused)
template <class T>
__global__ __launch_bounds__(256) void k_sum_partials(const T *__restrict__ P, int nch, int kb,
                                                      int K, int N, T *__restrict__ W)
{
    const int m = blockIdx.y;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x, per = (size_t)K * N;
    if (e >= per || (int)(e % K) >= kb) return;
    const T *p = P + (size_t)m * nch * per + e;
    T s = p[0];
    for (int q = 1; q < nch; ++q) s = part_add(s, p[(size_t)q * per]);
    W[(size_t)m * per + e] = s;
}

static hipError_t launch_sum_partials(const void *P, int nch, int kb, int K, int N, int nb,
                                      void *W, bool cx, hipStream_t st)
{
    const dim3 grid((unsigned)(((size_t)K * N + 255) / 256), (unsigned)nb);
    if (cx)
        hipLaunchKernelGGL(k_sum_partials<double2>, grid, dim3(256), 0, st,
                           static_cast<const double2 *>(P), nch, kb, K, N,
                           static_cast<double2 *>(W));
    else
        hipLaunchKernelGGL(k_sum_partials<double>, grid, dim3(256), 0, st,
                           static_cast<const double *>(P), nch, kb, K, N,
                           static_cast<double *>(W));
    return hipGetLastError();
}

static bool use_coop_heevd(const imago_ctx *c, int n, bool cx)
{
    return hetrd_coop_supported(n, !cx) && c->tune_heevd == 0;
}

// heevd_coop's return when the cooperative launch is refused (A untouched)
constexpr int kCoopRefused = 1000;

// One cooperative reduction in flight per process: contexts on other host
// threads (the batched stream runs several) wait here rather than launch a
// second grid that needs to be co-resident with the first.
static std::mutex g_coop_mu;

static int ensure_stage_events(imago_ctx *c)
{
    for (hipEvent_t &ev : c->evs)
        if (!ev) HIPCHK(hipEventCreate(&ev));
    return IMAGO_OK;
}

// times of the last finished columns (timing stages 20, 21), read once the
// stream has passed them
static int read_finish_events(imago_ctx *c)
{
    if (!c->evs_pending) return IMAGO_OK;
    float a = 0.0f, b = 0.0f;
    HIPCHK(hipEventElapsedTime(&a, c->evs[4], c->evs[5]));
    HIPCHK(hipEventElapsedTime(&b, c->evs[5], c->evs[6]));
    c->timing[20] = a;
    c->timing[21] = b;
    c->evs_pending = false;
    return IMAGO_OK;
}

// the divide and conquer's merge plan for order n, on the device
static int dc_ensure_plan(imago_ctx *c, int n)
{
    if (c->dcplan.n == n && c->dcplan.dev) return IMAGO_OK;
    if (c->dcplan.dev) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(c->dcplan.dev));
        c->dcplan.dev = nullptr;
        c->dcplan.n = 0;
    }
    DcPlan p;
    dc_make_plan(n, p);
    const size_t bytes = std::max<size_t>(p.host.size(), 1) * sizeof(int);
    HIPCHK(hipMalloc(&p.dev, bytes));
    if (!p.host.empty())
        HIPCHK(ctx_memcpy(c, p.dev, p.host.data(), p.host.size() * sizeof(int), hipMemcpyHostToDevice));
    c->dcplan = std::move(p);
    return IMAGO_OK;
}

// The batched divide and conquer on the context's stream as a replay of a
// graph captured on the first call with these buffers (the plan is on the
// device already; the capture records kernel launches only, no allocation,
// copy or memset)
static int dc_graph_launch(imago_ctx *c, int n, int nb, double *evals, double *ework, void *A,
                           bool real)
{
    imago_ctx::DcGraph &g = c->dcg;
    const size_t bytes = dc_work_bytes(n, nb);
    const bool same = g.exec && g.n == n && g.nb == nb && g.real == real && g.d == evals &&
                      g.e == ework && g.X == A && g.plan == c->dcplan.dev &&
                      g.work_bytes >= bytes;
    if (!same) {
        // one capture at a time in the process, with no allocation or free
        // of another context's graph work area during it
        static std::mutex capture_mu;
        std::lock_guard<std::mutex> capture_lock(capture_mu);
        if (g.exec || g.work) HIPCHK(hipStreamSynchronize(c->stream));
        if (g.exec) {
            HIPCHK(hipGraphExecDestroy(g.exec));
            g.exec = nullptr;
        }
        if (g.work_bytes < bytes) {
            if (g.work) HIPCHK(hipFree(g.work));
            g.work = nullptr;
            g.work_bytes = 0;
            HIPCHK(hipMalloc(&g.work, bytes));
            g.work_bytes = bytes;
        }
        DcWork w;
        dc_work_carve(g.work, n, nb, w);
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t le =
            launch_tridiag_dc(c->dcplan, nb, evals, ework, w, evals, A, real, c->stream);
        const hipError_t ee = hipStreamEndCapture(c->stream, &graph);
        if (le != hipSuccess || ee != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            return fail(IMAGO_E_HIP, "capture of the divide and conquer: %s",
                        hipGetErrorString(le != hipSuccess ? le : ee));
        }
        const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(ie);
        g.n = n;
        g.nb = nb;
        g.real = real;
        g.d = evals;
        g.e = ework;
        g.X = A;
        g.plan = c->dcplan.dev;
    }
    HIPCHK(hipGraphLaunch(g.exec, c->stream));
    return IMAGO_OK;
}

static int heevd_coop(imago_ctx *c, bool cx, int n, int nb, void *A, double *evals,
                      double *ework, int *dinfo)
{
    std::lock_guard<std::mutex> coop_lock(g_coop_mu);
    const size_t nn = (size_t)n * n, es = cx ? sizeof(double2) : sizeof(double);
    const int N = n * nb, K = kTrdBlock, nblk = (n - 1 + K - 1) / K;
    char *V = nullptr, *tau = nullptr, *vbuf = nullptr, *pbuf = nullptr, *dbuf = nullptr,
         *Tf = nullptr;
    double *Z = nullptr, *wall = nullptr;
    int *owner = nullptr, *rank = nullptr;
    unsigned *bar = nullptr;
    int *err = nullptr, *sinfo = nullptr;
    const int h0 = (nb + 1) / 2, h1 = nb - h0;
    const int N0 = h0 * n, N1 = h1 * n;
    const bool own_dc = c->tune_tridiag == 0;
    char *dcw = nullptr;
    // V carries K*n zero elements past its last matrix: the GEMM forming the
    // compact-WY Gram blocks reads whole K-column blocks
    RCHK(dev_alloc(c, &V, (nn * nb + (size_t)K * n) * es));
    if (own_dc) {
        RCHK(dc_ensure_plan(c, n));
        if (!c->tune_graph) RCHK(dev_alloc(c, &dcw, dc_work_bytes(n, nb)));
    } else {
        RCHK(dev_alloc(c, &Z, (size_t)N0 * N0 + (size_t)N1 * N1));
        RCHK(dev_alloc(c, &wall, (size_t)N));
        RCHK(dev_alloc(c, &owner, (size_t)N));
        RCHK(dev_alloc(c, &rank, (size_t)N));
    }
    RCHK(dev_alloc(c, &tau, (size_t)n * nb * es));
    const bool ut = c->tune_wy == 0;
    RCHK(dev_alloc(c, &Tf, ut ? nn * nb * es : (size_t)nb * std::max(nblk, 1) * K * K * es));
    RCHK(dev_alloc(c, &vbuf, nn * nb * es));
    RCHK(dev_alloc(c, &pbuf, nn * nb * es));
    RCHK(dev_alloc(c, &dbuf, (size_t)n * 128 * nb * es));
    RCHK(dev_alloc(c, &bar, 1));
    RCHK(dev_alloc(c, &err, 1));
    RCHK(dev_alloc(c, &sinfo, 2));
    HIPCHK(hipMemsetAsync(err, 0, sizeof(int), c->stream));
    HIPCHK(hipMemsetAsync(sinfo, 0, 2 * sizeof(int), c->stream));
    HIPCHK(hipMemsetAsync(dinfo, 0, sizeof(int) * nb, c->stream));
    HIPCHK(hipMemsetAsync(V, 0, es * (nn * nb + (size_t)K * n), c->stream));
    TrdLaunch L{n,     nb,   !cx,  A,    V,   evals, ework, tau,
                vbuf,  pbuf, dbuf, bar,  err, c->stream};
    if (!c->evh0) {
        HIPCHK(hipEventCreate(&c->evh0));
        HIPCHK(hipEventCreate(&c->evh1));
    }
    RCHK(ensure_stage_events(c));
    HIPCHK(hipEventRecord(c->evh0, c->stream));
    const hipError_t le = launch_hetrd_coop(L);
    if (le == hipErrorIllegalState || le == hipErrorCooperativeLaunchTooLarge) {
        // the runtime refuses the cooperative launch (another process holds
        // the device's cooperative queue, or its work-groups would not all
        // be resident); A is untouched, the caller takes rocSOLVER's heevd
        (void)hipGetLastError();
        for (void *p : {(void *)V, (void *)Z, (void *)wall, (void *)owner, (void *)rank,
                        (void *)tau, (void *)vbuf, (void *)pbuf, (void *)dbuf, (void *)Tf,
                        (void *)bar, (void *)err, (void *)sinfo, (void *)dcw})
            dev_release(c, p);
        return kCoopRefused;
    }
    HIPCHK(le);
    HIPCHK(hipEventRecord(c->evh1, c->stream));
    if (own_dc) {
        // the batched divide and conquer: eigenvalues ascending per matrix
        // into evals, eigenvectors into A
        if (c->tune_graph) {
            RCHK(dc_graph_launch(c, n, nb, evals, ework, A, !cx));
        } else {
            DcWork dw;
            dc_work_carve(dcw, n, nb, dw);
            HIPCHK(launch_tridiag_dc(c->dcplan, nb, evals, ework, dw, evals, A, !cx, c->stream));
        }
    } else {
    // d and e of the batch are back to back, e[n-1] = 0 between problems:
    // each half of the batch is one block-diagonal problem, the two halves
    // on two streams (one call on the whole batch merges across the blocks
    // at its top levels for nothing)
    HIPCHK(hipMemcpyAsync(wall, evals, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream));
    double *Z1 = Z + (size_t)N0 * N0;
    if (h1 > 0) {
        if (!c->stream2) {
            HIPCHK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
            if (rocblas_create_handle(&c->blas2) != rocblas_status_success ||
                rocblas_set_stream(c->blas2, c->stream2) != rocblas_status_success ||
                rocblas_set_atomics_mode(c->blas2, rocblas_atomics_not_allowed) !=
                    rocblas_status_success)
                return fail(IMAGO_E_HIP, "rocblas handle for the second stream");
            HIPCHK(hipEventCreateWithFlags(&c->evf, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&c->evj, hipEventDisableTiming));
        }
        HIPCHK(hipEventRecord(c->evf, c->stream));
        HIPCHK(hipStreamWaitEvent(c->stream2, c->evf, 0));
        BLASCHK(rocsolver_dstedc(c->blas2, rocblas_evect_tridiagonal, N1, wall + N0, ework + N0,
                                 Z1, N1, sinfo + 1));
        HIPCHK(hipEventRecord(c->evj, c->stream2));
    }
    BLASCHK(rocsolver_dstedc(c->blas, rocblas_evect_tridiagonal, N0, wall, ework, Z, N0, sinfo));
    if (h1 > 0) HIPCHK(hipStreamWaitEvent(c->stream, c->evj, 0));
    HIPCHK(launch_trd_unmix(Z, wall, n, h0, 0, owner, rank, evals, A, !cx, err, c->stream));
    if (h1 > 0)
        HIPCHK(launch_trd_unmix(Z1, wall + N0, n, h1, h0, owner + N0, rank + N0, evals, A, !cx,
                                err, c->stream));
    }
    HIPCHK(hipEventRecord(c->evs[2], c->stream));
    if (ut && n > 1) {
        // G = V^H V over the n - 1 reflectors (V^H upper, V lower: the K range
        // of tile (i, j) starts at max(i0, j0); only the lower tiles, mirrored),
        // T^-1 from it in place, then T = (T^-1)^-1
        char *G = nullptr;
        RCHK(dev_alloc(c, &G, nn * nb * es));
        HIPCHK(launch_ut_null(tau, n, nb, !cx, V, c->stream));
        GemmArgs g;
        g.M = g.N = n - 1;
        g.K = n;
        g.lda = g.ldb = g.ldc = n;
        g.batch = nb;
        g.A = g.B = V;
        g.sa1 = g.sb1 = g.sc1 = (long)nn;
        g.C = G;
        g.amask = kTriUpper;
        g.bmask = kTriLower;
        g.herm = true;
        HIPCHK(launch_gemm_mfma(g, cx, true, false, c->stream));
        HIPCHK(launch_ut_tinv(tau, n, nb, !cx, G, c->stream));
        const rocblas_stride snn = (rocblas_stride)nn;
        if (cx)
            BLASCHK(rocblas_ztrtri_strided_batched(
                c->blas, rocblas_fill_upper, rocblas_diagonal_non_unit, n - 1,
                (const rocblas_double_complex *)G, n, snn, (rocblas_double_complex *)Tf, n, snn,
                nb));
        else
            BLASCHK(rocblas_dtrtri_strided_batched(c->blas, rocblas_fill_upper,
                                                   rocblas_diagonal_non_unit, n - 1,
                                                   (const double *)G, n, snn, (double *)Tf, n, snn,
                                                   nb));
        HIPCHK(hipStreamSynchronize(c->stream));
        dev_release(c, G);
    } else if (!ut) {
        const rocblas_stride sw = (rocblas_stride)K * n,
                             st = (rocblas_stride)std::max(nblk, 1) * K * K;
        // G_b = V_b^H V_b for every block of every matrix in one launch on the
        // matrix cores (V_b is zero above row K b, so its K range starts
        // there), then the factors T_b in place of G_b; the columns a last
        // block reads beyond n only reach G entries the factor ignores
        if (nblk > 0) {
            // K split into chunks of kGramChunk rows (a launch entry each),
            // the partials summed in order into T's buffer
            constexpr int kGramChunk = 128;
            const int nch = (n + kGramChunk - 1) / kGramChunk;
            char *Gp = nullptr;
            RCHK(dev_alloc(c, &Gp, (size_t)nb * nblk * nch * K * K * es));
            GemmArgs g;
            g.M = g.N = K;
            g.K = n;
            g.lda = g.ldb = n;
            g.ldc = K;
            g.batch = nb * nblk * nch;
            g.nb2 = nblk;
            g.nb3 = nch;
            g.A = g.B = V;
            g.sa1 = g.sb1 = (long)nn;
            g.sa2 = g.sb2 = (long)sw;
            g.C = Gp;
            g.sc1 = (long)nblk * nch * K * K;
            g.sc2 = (long)nch * K * K;
            g.sc3 = (long)K * K;
            g.koff2 = K;
            g.klen2 = kGramChunk;
            HIPCHK(launch_gemm_mfma(g, cx, true, false, c->stream));
            HIPCHK(launch_sum_partials(Gp, nch, K, K, K, nb * nblk, Tf, cx, c->stream));
            dev_release(c, Gp);
        }
        HIPCHK(launch_trd_larft(tau, n, nb, !cx, Tf, c->stream));
    }
    HIPCHK(hipEventRecord(c->evs[3], c->stream));
    // the reflectors and their factors stay for the back-transform
    c->pend.V = V;
    c->pend.Tf = Tf;
    c->pend.reflectors = true;
    c->pend.ut = ut;
    V = nullptr;
    Tf = nullptr;
    int herr = 0, hs[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(hs, sinfo, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float tms = 0.0f, tdc = 0.0f, twy = 0.0f;
    HIPCHK(hipEventElapsedTime(&tms, c->evh0, c->evh1));
    HIPCHK(hipEventElapsedTime(&tdc, c->evh1, c->evs[2]));
    HIPCHK(hipEventElapsedTime(&twy, c->evs[2], c->evs[3]));
    c->timing[10] = tms;
    c->timing[19] = tdc;
    c->timing[22] = twy;
    for (void *p : {(void *)V, (void *)Z, (void *)wall, (void *)owner, (void *)rank, (void *)tau,
                    (void *)vbuf, (void *)pbuf, (void *)dbuf, (void *)Tf, (void *)bar, (void *)err,
                    (void *)sinfo, (void *)dcw})
        dev_release(c, p);
    if (herr & 1) return fail(IMAGO_E_HIP, "tridiagonal reduction: a barrier wait timed out");
    if (herr & 2)
        return fail(IMAGO_E_HIP, "eigenvector sort: a matrix did not receive exactly n vectors");
    if (hs[0] != 0 || hs[1] != 0) {
        // divide and conquer failed in a half: its matrices report zheevd's
        // "did not converge" (1..n) through dinfo, read by the caller
        std::vector<int> kinfo(nb, 0);
        map_stedc_info(n, h0, hs[0], kinfo.data());
        if (h1 > 0) map_stedc_info(n, h1, hs[1], kinfo.data() + h0);
        HIPCHK(ctx_memcpy(c, dinfo, kinfo.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
    }
    return IMAGO_OK;
}

static void pending_clear(imago_ctx *c)
{
    dev_release(c, c->pend.V);
    dev_release(c, c->pend.Tf);
    c->pend = imago_ctx::PendingVectors{};
}

// Eigenvector columns [done, upto) of the pending eigensolve: Y = Q Z by
// the compact-WY blocks from the last to the first (W = V_b^H Y,
// W = T_b W, Y -= V_b W as batched GEMMs over these columns only), then
// X = U^-1 Y.  Columns are independent, so later calls continue where this
// one stops.
static int finish_columns(imago_ctx *c, int upto)
{
    imago_ctx::PendingVectors &p = c->pend;
    if (!p.active || upto <= p.done) return IMAGO_OK;
    const int n = p.n, nb = p.nb, c0 = p.done, N = std::min(upto, n) - c0;
    const bool cx = p.cx;
    const size_t es = cx ? sizeof(double2) : sizeof(double), nn = (size_t)n * n;
    char *Ac = static_cast<char *>(p.A) + (size_t)c0 * n * es;
    RCHK(ensure_stage_events(c));
    HIPCHK(hipEventRecord(c->evs[4], c->stream));
    if (p.reflectors && p.ut) {
        // Y = (I - V T V^H) Y: W = V^H Y (V^H upper: row tile i from row i0),
        // W2 = T W (T upper), Y -= V W2 (V lower: row tile i up to its last row)
        if (n > 1) {
            const int n1 = n - 1;
            char *W = nullptr, *W2 = nullptr;
            RCHK(dev_alloc(c, &W, (size_t)nb * n1 * N * es));
            RCHK(dev_alloc(c, &W2, (size_t)nb * n1 * N * es));
            GemmArgs g;
            g.M = n1;
            g.N = N;
            g.K = n;
            g.lda = n;
            g.ldb = n;
            g.ldc = n1;
            g.batch = nb;
            g.A = p.V;
            g.sa1 = (long)nn;
            g.B = Ac;
            g.sb1 = (long)nn;
            g.C = W;
            g.sc1 = (long)n1 * N;
            g.amask = kTriUpper;
            HIPCHK(launch_gemm_mfma(g, cx, true, false, c->stream));
            GemmArgs g2;
            g2.M = n1;
            g2.N = N;
            g2.K = n1;
            g2.lda = n;
            g2.ldb = g2.ldc = n1;
            g2.batch = nb;
            g2.A = p.Tf;
            g2.sa1 = (long)nn;
            g2.B = W;
            g2.sb1 = (long)n1 * N;
            g2.C = W2;
            g2.sc1 = (long)n1 * N;
            g2.amask = kTriUpper;
            HIPCHK(launch_gemm_mfma(g2, cx, false, false, c->stream));
            GemmArgs g3;
            g3.M = n;
            g3.N = N;
            g3.K = n1;
            g3.lda = n;
            g3.ldb = n1;
            g3.ldc = n;
            g3.batch = nb;
            g3.A = p.V;
            g3.sa1 = (long)nn;
            g3.B = W2;
            g3.sb1 = (long)n1 * N;
            g3.C = Ac;
            g3.sc1 = (long)nn;
            g3.alpha_re = -1.0;
            g3.beta_re = 1.0;
            g3.amask = kTriLower;
            HIPCHK(launch_gemm_mfma(g3, cx, false, false, c->stream));
            dev_release(c, W);
            dev_release(c, W2);
        }
    } else if (p.reflectors) {
        const int K = kTrdBlock, nblk = (n - 1 + K - 1) / K;
        // W = V_b^H Y_b is K x N over a long K range: it is split into
        // chunks of kBtChunk rows, each product a separate matrix-core
        // launch entry, and the partials summed in order
        constexpr int kBtChunk = 64;
        const int nchm = (n - 1 + kBtChunk - 1) / kBtChunk;
        char *W = nullptr, *W2 = nullptr, *Wp = nullptr;
        RCHK(dev_alloc(c, &W, (size_t)nb * K * N * es));
        RCHK(dev_alloc(c, &W2, (size_t)nb * K * N * es));
        RCHK(dev_alloc(c, &Wp, (size_t)nb * nchm * K * N * es));
        const rocblas_stride snn = (rocblas_stride)nn, sw = (rocblas_stride)K * N,
                             st = (rocblas_stride)std::max(nblk, 1) * K * K;
        for (int b = nblk - 1; b >= 0; --b) {
            const int j0 = b * K, kb = std::min(K, n - 1 - j0), r0 = j0 + 1, mb = n - r0;
            const char *Vb = p.V + ((size_t)j0 * n + r0) * es;
            char *Yb = Ac + (size_t)r0 * es;
            const int nch = (mb + kBtChunk - 1) / kBtChunk;
            GemmArgs gw;
            gw.M = kb;
            gw.N = N;
            gw.K = mb;
            gw.lda = n;
            gw.ldb = n;
            gw.ldc = K;
            gw.batch = nb * nch;
            gw.nb2 = nch;
            gw.A = Vb;
            gw.sa1 = (long)nn;
            gw.B = Yb;
            gw.sb1 = (long)nn;
            gw.C = Wp;
            gw.sc1 = (long)nch * K * N;
            gw.sc2 = (long)K * N;
            gw.koff2 = kBtChunk;
            gw.klen2 = kBtChunk;
            HIPCHK(launch_gemm_mfma(gw, cx, true, false, c->stream));
            HIPCHK(launch_sum_partials(Wp, nch, kb, K, N, nb, W, cx, c->stream));
            BLASCHK(gemm_sb(c->blas, cx, rocblas_operation_none, rocblas_operation_none, kb, N, kb,
                            1.0, p.Tf + (size_t)b * K * K * es, K, st, W, K, sw, 0.0, W2, K, sw,
                            nb));
            BLASCHK(gemm_sb(c->blas, cx, rocblas_operation_none, rocblas_operation_none, mb, N, kb,
                            -1.0, Vb, n, snn, W2, K, sw, 1.0, Yb, n, snn, nb));
        }
        dev_release(c, W);
        dev_release(c, W2);
        dev_release(c, Wp);
    }
    HIPCHK(hipEventRecord(c->evs[5], c->stream));
    const rocblas_stride nns = (rocblas_stride)nn;
    if (p.Binv) {
        // X = U^-1 Y on the matrix cores (U^-1 upper: the K range of row
        // tile i starts at its first row) into the scratch, then back
        GemmArgs g;
        g.M = n;
        g.N = N;
        g.K = n;
        g.lda = g.ldb = g.ldc = n;
        g.batch = nb;
        g.A = p.Binv;
        g.sa1 = (long)nn;
        g.B = Ac;
        g.sb1 = (long)nn;
        g.C = c->Gtmp;
        g.sc1 = (long)n * N;
        g.amask = kTriUpper;
        HIPCHK(launch_gemm_mfma(g, cx, false, false, c->stream));
        HIPCHK(hipMemcpy2DAsync(Ac, nn * es, c->Gtmp, (size_t)n * N * es, (size_t)n * N * es, nb,
                                hipMemcpyDeviceToDevice, c->stream));
    } else if (cx) {
        const rocblas_double_complex one(1.0, 0.0);
        BLASCHK(rocblas_ztrsm_strided_batched(
            c->blas, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
            rocblas_diagonal_non_unit, n, N, &one, (const rocblas_double_complex *)p.B, n, nns,
            (rocblas_double_complex *)Ac, n, nns, nb));
    } else {
        const double one = 1.0;
        BLASCHK(rocblas_dtrsm_strided_batched(c->blas, rocblas_side_left, rocblas_fill_upper,
                                              rocblas_operation_none, rocblas_diagonal_non_unit, n,
                                              N, &one, (const double *)p.B, n, nns, (double *)Ac,
                                              n, nns, nb));
    }
    HIPCHK(hipEventRecord(c->evs[6], c->stream));
    c->evs_pending = true;
    p.done += N;
    if (p.done >= n) {
        HIPCHK(hipStreamSynchronize(c->stream));
        pending_clear(c);
    }
    return IMAGO_OK;
}

// C = U^-H H U^-1 from the explicit inverse on the matrix cores: G = H U^-1
// (U^-1 upper: the K range of column tile j ends at its last column), then
// the lower tiles of C = U^-H G (U^-H lower: the K range of row tile i ends
// at its last row), each written to both triangles with a real diagonal; H
// is overwritten with C.
static int gemm_transform(imago_ctx *c, bool cx, int n, int nb, void *H, const void *Uinv)
{
    const long nn = (long)n * n;
    GemmArgs g;
    g.M = g.N = g.K = n;
    g.lda = g.ldb = g.ldc = n;
    g.batch = nb;
    g.A = H;
    g.sa1 = nn;
    g.B = Uinv;
    g.sb1 = nn;
    g.C = c->Gtmp;
    g.sc1 = nn;
    g.bmask = kTriUpper;
    HIPCHK(launch_gemm_mfma(g, cx, false, false, c->stream));
    GemmArgs g2 = g;
    g2.A = Uinv;
    g2.B = c->Gtmp;
    g2.C = H;
    g2.amask = kTriLower;
    g2.bmask = kTriNone;
    g2.herm = true;
    HIPCHK(launch_gemm_mfma(g2, cx, true, false, c->stream));
    return IMAGO_OK;
}

// C = U^-H A U^-1, its eigenpairs, and the eigenvectors X = U^-1 Y of the
// generalized problem: all columns when defer is false; with defer the
// columns are left to finish_columns.
static int solve_gv_factored(imago_ctx *c, bool cx, int n, int nb, void *A, const void *B,
                             double *evals, double *ework, int *dinfo, int *hinfo,
                             bool defer = false, const void *Binv = nullptr)
{
    if (c->pend.active || c->pend.V || c->pend.Tf) {
        HIPCHK(hipStreamSynchronize(c->stream));
        pending_clear(c);
    }
    c->timing[10] = c->timing[19] = c->timing[22] = 0.0;
    RCHK(ensure_stage_events(c));
    const rocblas_stride nn = (rocblas_stride)n * n;
    HIPCHK(hipEventRecord(c->evs[0], c->stream));
    if (cx) {
        const rocblas_double_complex one(1.0, 0.0);
        auto *a = (rocblas_double_complex *)A;
        auto *b = (const rocblas_double_complex *)B;
        if (Binv) {
            RCHK(gemm_transform(c, true, n, nb, A, Binv));
        } else {
            BLASCHK(rocblas_ztrsm_strided_batched(c->blas, rocblas_side_left, rocblas_fill_upper,
                                                  rocblas_operation_conjugate_transpose,
                                                  rocblas_diagonal_non_unit, n, n, &one, b, n, nn,
                                                  a, n, nn, nb));
            BLASCHK(rocblas_ztrsm_strided_batched(c->blas, rocblas_side_right, rocblas_fill_upper,
                                                  rocblas_operation_none, rocblas_diagonal_non_unit,
                                                  n, n, &one, b, n, nn, a, n, nn, nb));
        }
        HIPCHK(hipEventRecord(c->evs[1], c->stream));
        const int rc = use_coop_heevd(c, n, true) ? heevd_coop(c, true, n, nb, A, evals, ework, dinfo)
                                               : kCoopRefused;
        if (rc != kCoopRefused) {
            RCHK(rc);
        } else {
            BLASCHK(rocsolver_zheevd_strided_batched(c->blas, rocblas_evect_original,
                                                     rocblas_fill_upper, n, a, n, nn, evals, n,
                                                     ework, n, dinfo, nb));
        }
    } else {
        const double one = 1.0;
        auto *a = (double *)A;
        auto *b = (const double *)B;
        if (Binv) {
            RCHK(gemm_transform(c, false, n, nb, A, Binv));
        } else {
            BLASCHK(rocblas_dtrsm_strided_batched(c->blas, rocblas_side_left, rocblas_fill_upper,
                                                  rocblas_operation_transpose,
                                                  rocblas_diagonal_non_unit, n, n, &one, b, n, nn,
                                                  a, n, nn, nb));
            BLASCHK(rocblas_dtrsm_strided_batched(c->blas, rocblas_side_right, rocblas_fill_upper,
                                                  rocblas_operation_none, rocblas_diagonal_non_unit,
                                                  n, n, &one, b, n, nn, a, n, nn, nb));
        }
        HIPCHK(hipEventRecord(c->evs[1], c->stream));
        const int rc = use_coop_heevd(c, n, false)
                           ? heevd_coop(c, false, n, nb, A, evals, ework, dinfo)
                           : kCoopRefused;
        if (rc != kCoopRefused) {
            RCHK(rc);
        } else {
            BLASCHK(rocsolver_dsyevd_strided_batched(c->blas, rocblas_evect_original,
                                                     rocblas_fill_upper, n, a, n, nn, evals, n,
                                                     ework, n, dinfo, nb));
        }
    }
    // heevd_coop leaves its reflectors in c->pend; rocSOLVER's heevd has
    // back-transformed already and leaves only X = U^-1 Y
    imago_ctx::PendingVectors &p = c->pend;
    p.active = true;
    p.cx = cx;
    p.n = n;
    p.nb = nb;
    p.done = 0;
    p.A = A;
    p.B = B;
    p.Binv = Binv;
    if (!defer) RCHK(finish_columns(c, n));
    HIPCHK(hipMemcpyAsync(hinfo, dinfo, sizeof(int) * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float tt = 0.0f;
    HIPCHK(hipEventElapsedTime(&tt, c->evs[0], c->evs[1]));
    c->timing[18] = tt;
    RCHK(read_finish_events(c));
    return IMAGO_OK;
}

static int solve_gv_batched(imago_ctx *c, bool cx, int n, int nb, void *A, void *B,
                            double *evals, double *ework, int *dinfo, int *hinfo)
{
    bool bad = false;
    RCHK(factor_gv_batched(c, cx, n, nb, B, dinfo, hinfo, &bad));
    if (bad) return IMAGO_OK;
    return solve_gv_factored(c, cx, n, nb, A, B, evals, ework, dinfo, hinfo);
}

extern "C" int imago_secular(imago_ctx *c, const double *pot_coeffs, int num_states,
                             double *eigenvalues, int *info_out)
{
    if (!c || !c->has_system) return fail(IMAGO_E_STATE, "imago_secular before imago_set_system");
    const HostSys &h = c->h;
    if (!c->assembled[0] || !c->assembled[1] || !c->assembled[2] ||
        !(c->assembled[3] || c->ep_direct))
        return fail(IMAGO_E_STATE,
                    "imago_secular needs OL, KE, NP and EP (or EP_DIRECT) assembled");
    // the stored EP panels when they exist, else the direct path
    const bool direct = !c->assembled[3];
    if (num_states <= 0 || num_states > h.n)
        return fail(IMAGO_E_ARG, "num_states %d out of range 1..%d", num_states, h.n);
    if (!pot_coeffs && h.pot_dim > 0) return fail(IMAGO_E_ARG, "null potential coefficients");
    HIPCHK(hipSetDevice(c->device));
    RCHK(ensure_secular_buffers(c));
    const long per = panel_doubles(c);
    HIPCHK(hipMemcpyAsync(c->coef, pot_coeffs, sizeof(double) * h.pot_dim, hipMemcpyHostToDevice,
                          c->stream));
    if (direct) {
        // readPackedMatrixAccum: a factor of exactly 0.0 adds the term unscaled
        std::vector<double> ce(h.pot_dim);
        for (int j = 0; j < h.pot_dim; ++j) ce[j] = pot_coeffs[j] == 0.0 ? 1.0 : pot_coeffs[j];
        if (!c->coef_eff) RCHK(dev_alloc(c, &c->coef_eff, std::max(h.pot_dim, 1)));
        HIPCHK(hipMemcpyAsync(c->coef_eff, ce.data(), sizeof(double) * h.pot_dim,
                              hipMemcpyHostToDevice, c->stream));
    }
    if (direct) RCHK(ep_build_runs(c));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if (direct) {
        EpdLaunch L{c->d,          c->pairs,     c->phase,       c->ep_tuples,
                    c->ep_item_tuple, c->pitem, c->ep_blocks,   c->coef_eff,
                    nullptr,       0,            c->type_term0,  c->hruns,
                    c->nhruns,     c->hpart,     c->npairs,      c->lmax,
                    c->ne <= 4 ? 4 : 16,         h.gamma != 0,   c->stream};
        HIPCHK(launch_epd_hsum(L));
        HIPCHK(hipEventRecord(c->ev2, c->stream));
        HIPCHK(launch_epd_hfinal(L, c->panel[2], c->panel[1], c->hrun_lo, c->hrun_woff, c->blockH));
    } else {
        hipLaunchKernelGGL(k_hbuild, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, c->stream,
                           c->panel[2], c->panel[1], c->panel[3], c->coef, h.pot_dim, per,
                           c->blockH);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    const long nn = (long)h.n * h.n;
    if (nn * h.nk > 0xffffffffL - 255)
        return fail(IMAGO_E_ARG, "%ld matrix entries exceed one dispatch", nn * h.nk);
    const unsigned ug = (unsigned)((nn * h.nk + 255) / 256);
    const bool cx = !h.gamma;
    void *Hm = cx ? (void *)c->Hfull : (void *)c->Hreal;
    void *Sm = cx ? (void *)c->Sfull : (void *)c->Sreal;
    const int nb = cx ? h.nk : 1;
    if (cx)
        hipLaunchKernelGGL(k_unpack<true>, dim3(ug), dim3(256), 0, c->stream, c->blockH, c->nbe,
                           h.n, h.nk, c->pairs, c->atom_of_state, c->pair_index, c->d.cumul,
                           h.natoms, (double *)Hm);
    else
        hipLaunchKernelGGL(k_unpack<false>, dim3(ug), dim3(256), 0, c->stream, c->blockH, c->nbe,
                           h.n, 1, c->pairs, c->atom_of_state, c->pair_index, c->d.cumul,
                           h.natoms, (double *)Hm);
    HIPCHK(hipGetLastError());
    hipEvent_t e2, e3;
    HIPCHK(hipEventCreate(&e2));
    HIPCHK(hipEventCreate(&e3));
    HIPCHK(hipEventRecord(e2, c->stream));
    std::vector<int> inf(h.nk, 0);
    // S(k) does not change between calls on one assembled overlap, so its
    // Cholesky factor is computed once and reused until OL is reassembled.
    bool bad = false;
    if (!c->chol_valid) {
        if (cx)
            hipLaunchKernelGGL(k_unpack<true>, dim3(ug), dim3(256), 0, c->stream, c->panel[0],
                               c->nbe, h.n, h.nk, c->pairs, c->atom_of_state, c->pair_index,
                               c->d.cumul, h.natoms, (double *)Sm);
        else
            hipLaunchKernelGGL(k_unpack<false>, dim3(ug), dim3(256), 0, c->stream, c->panel[0],
                               c->nbe, h.n, 1, c->pairs, c->atom_of_state, c->pair_index,
                               c->d.cumul, h.natoms, (double *)Sm);
        HIPCHK(hipGetLastError());
        RCHK(factor_gv_batched(c, cx, h.n, nb, Sm, c->info, inf.data(), &bad));
        c->chol_valid = !bad;
        // U^-1 once per factor: every later transform and solve is a GEMM
        if (!bad) {
            const rocblas_stride snn = (rocblas_stride)nn;
            if (cx)
                BLASCHK(rocblas_ztrtri_strided_batched(
                    c->blas, rocblas_fill_upper, rocblas_diagonal_non_unit, h.n,
                    (const rocblas_double_complex *)Sm, h.n, snn,
                    (rocblas_double_complex *)c->Sinv, h.n, snn, nb));
            else
                BLASCHK(rocblas_dtrtri_strided_batched(
                    c->blas, rocblas_fill_upper, rocblas_diagonal_non_unit, h.n,
                    (const double *)Sm, h.n, snn, (double *)c->Sinv, h.n, snn, nb));
        }
    }
    // the eigenvector columns are finished by the consumer (the density
    // needs the occupied ones only)
    if (!bad)
        RCHK(solve_gv_factored(c, cx, h.n, nb, Hm, Sm, c->evals, c->ework, c->info, inf.data(),
                               true, c->Sinv));
    HIPCHK(hipEventRecord(e3, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms1 = 0, ms2 = 0;
    HIPCHK(hipEventElapsedTime(&ms1, c->ev0, c->ev1));
    if (direct) {
        float ms8 = 0;
        HIPCHK(hipEventElapsedTime(&ms8, c->ev0, c->ev2));
        c->timing[8] = ms8;
    }
    HIPCHK(hipEventElapsedTime(&ms2, e2, e3));
    (void)hipEventDestroy(e2);
    (void)hipEventDestroy(e3);
    c->timing[1] = ms1;
    c->timing[2] = ms2;
    std::vector<double> ev((size_t)h.nk * h.n);
    HIPCHK(ctx_memcpy(c, ev.data(), c->evals, sizeof(double) * ev.size(), hipMemcpyDeviceToHost));
    if (info_out)
        for (int k = 0; k < h.nk; ++k) info_out[k] = inf[k];
    if (eigenvalues)
        for (int k = 0; k < h.nk; ++k)
            for (int s2 = 0; s2 < num_states; ++s2)
                eigenvalues[(size_t)k * num_states + s2] = ev[(size_t)k * h.n + s2];
    c->num_states = num_states;
    c->have_vectors = true;
    for (int k = 0; k < h.nk; ++k) {
        if (inf[k] > h.n) {
            c->have_vectors = false;
            return fail(IMAGO_E_NOTPD,
                        "ZHEGV error code=%d at k-point %d: overlap not positive definite "
                        "(leading minor %d)",
                        inf[k], k + 1, inf[k] - h.n);
        }
        if (inf[k] != 0) {
            c->have_vectors = false;
            return fail(IMAGO_E_SOLVE, "ZHEGV error code=%d at k-point %d", inf[k], k + 1);
        }
    }
    return IMAGO_OK;
}

extern "C" int imago_get_eigenvectors(imago_ctx *c, int kpoint, double *out)
{
    if (!c || !c->has_system || !c->have_vectors)
        return fail(IMAGO_E_STATE, "no eigenvectors: call imago_secular first");
    if (kpoint < 0 || kpoint >= c->h.nk || !out) return fail(IMAGO_E_ARG, "bad k-point %d", kpoint);
    HIPCHK(hipSetDevice(c->device));
    RCHK(finish_columns(c, c->num_states));
    HIPCHK(hipStreamSynchronize(c->stream));
    RCHK(read_finish_events(c));
    const long nn = (long)c->h.n * c->h.n;
    const long cols = (long)c->h.n * c->num_states;
    if (c->h.gamma)
        HIPCHK(ctx_memcpy(c, out, c->Hreal, sizeof(double) * cols, hipMemcpyDeviceToHost));
    else
        HIPCHK(ctx_memcpy(c, out, c->Hfull + (size_t)kpoint * nn, sizeof(double2) * cols,
                         hipMemcpyDeviceToHost));
    return IMAGO_OK;
}

extern "C" int imago_dos(imago_ctx *c, const double *energies, int spin, const int *channel,
                         int num_channels, double emin, double delta, int num_points,
                         double sigma, double *pdos, double *electron_number,
                         double *localization)
{
    if (!c || !c->has_system || !c->have_vectors)
        return fail(IMAGO_E_STATE, "imago_dos needs the eigenpairs of imago_secular");
    if (!c->assembled[0]) return fail(IMAGO_E_STATE, "imago_dos needs the overlap (IMAGO_OL)");
    if (!energies || !channel || !pdos || num_channels < 1 || num_points < 1 ||
        !(sigma > 0.0) || !(delta > 0.0) || (spin != 1 && spin != 2))
        return fail(IMAGO_E_ARG, "bad density-of-states arguments");
    const HostSys &h = c->h;
    const int n = h.n, ns = c->num_states, nk = h.nk;
    for (int v = 0; v < n; ++v)
        if (channel[v] < 0 || channel[v] >= num_channels)
            return fail(IMAGO_E_ARG, "basis function %d has channel %d of %d", v, channel[v],
                        num_channels);
    HIPCHK(hipSetDevice(c->device));
    RCHK(finish_columns(c, ns));
    const bool cx = !h.gamma;
    const size_t nn = (size_t)n * n, dz = cx ? 2 : 1, rows = (size_t)nk * ns;
    double *Sf = nullptr, *SC = nullptr, *Q = nullptr, *M = nullptr, *P = nullptr, *G = nullptr,
           *Ed = nullptr, *wd = nullptr, *out = nullptr, *en = nullptr, *loc = nullptr;
    RCHK(dev_alloc(c, &Sf, nn * nk * dz));
    RCHK(dev_alloc(c, &SC, (size_t)n * ns * nk * dz));
    RCHK(dev_alloc(c, &Q, (size_t)n * ns * nk));
    RCHK(dev_alloc(c, &M, (size_t)n * num_channels));
    RCHK(dev_alloc(c, &P, rows * num_channels));
    RCHK(dev_alloc(c, &G, rows * num_points));
    RCHK(dev_alloc(c, &Ed, rows));
    RCHK(dev_alloc(c, &wd, (size_t)nk));
    RCHK(dev_alloc(c, &out, (size_t)num_channels * num_points));
    RCHK(dev_alloc(c, &en, (size_t)n));
    RCHK(dev_alloc(c, &loc, (size_t)ns));
    std::vector<double> Mh((size_t)n * num_channels, 0.0);
    for (int v = 0; v < n; ++v) Mh[(size_t)channel[v] * n + v] = 1.0 / spin;
    HIPCHK(hipMemcpyAsync(M, Mh.data(), Mh.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(Ed, energies, rows * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(wd, h.kw.data(), (size_t)nk * 8, hipMemcpyHostToDevice, c->stream));
    // the overlap S(k) in full (both triangles), then S C for the states
    const unsigned ug = (unsigned)((nn * nk + 255) / 256);
    if (cx)
        hipLaunchKernelGGL(k_unpack<true>, dim3(ug), dim3(256), 0, c->stream, c->panel[0], c->nbe,
                           n, nk, c->pairs, c->atom_of_state, c->pair_index, c->d.cumul, h.natoms,
                           Sf);
    else
        hipLaunchKernelGGL(k_unpack<false>, dim3(ug), dim3(256), 0, c->stream, c->panel[0],
                           c->nbe, n, 1, c->pairs, c->atom_of_state, c->pair_index, c->d.cumul,
                           h.natoms, Sf);
    HIPCHK(hipGetLastError());
    const rocblas_stride snn = (rocblas_stride)nn, sns = (rocblas_stride)n * ns;
    if (cx) {
        const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
        BLASCHK(rocblas_zgemm_strided_batched(
            c->blas, rocblas_operation_none, rocblas_operation_none, n, ns, n, &one,
            (const rocblas_double_complex *)Sf, n, snn, (const rocblas_double_complex *)c->Hfull,
            n, snn, &zero, (rocblas_double_complex *)SC, n, sns, nk));
    } else {
        const double one = 1.0, zero = 0.0;
        BLASCHK(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, n, ns, n,
                              &one, Sf, n, c->Hreal, n, &zero, SC, n));
    }
    HIPCHK(launch_dos_mulliken(cx ? (const double *)c->Hfull : c->Hreal, (long)nn, SC, n, ns, nk,
                               cx, Q, c->stream));
    {
        // channel sums P[(k, j)][ch] = sum_v Q_kjv M_v,ch, stacked over k
        const double one = 1.0, zero = 0.0;
        BLASCHK(rocblas_dgemm_strided_batched(
            c->blas, rocblas_operation_transpose, rocblas_operation_none, ns, num_channels, n,
            &one, Q, n, sns, M, n, 0, &zero, P, (rocblas_int)rows, ns, nk));
        HIPCHK(launch_dos_gauss(Ed, wd, ns, nk, emin, delta, num_points, sigma, G, c->stream));
        // PDOS = P^T G
        BLASCHK(rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none,
                              num_channels, num_points, (rocblas_int)rows, &one, P,
                              (rocblas_int)rows, G, (rocblas_int)rows, &zero, out, num_channels));
    }
    HIPCHK(launch_dos_counts(Q, Ed, wd, n, ns, nk, spin, en, loc, c->stream));
    HIPCHK(hipMemcpyAsync(pdos, out, (size_t)num_channels * num_points * 8,
                          hipMemcpyDeviceToHost, c->stream));
    std::vector<double> enh(n), loch(ns);
    HIPCHK(hipMemcpyAsync(enh.data(), en, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(loch.data(), loc, (size_t)ns * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (electron_number) std::copy(enh.begin(), enh.end(), electron_number);
    if (localization) std::copy(loch.begin(), loch.end(), localization);
    for (void *p : {(void *)Sf, (void *)SC, (void *)Q, (void *)M, (void *)P, (void *)G,
                    (void *)Ed, (void *)wd, (void *)out, (void *)en, (void *)loc})
        dev_release(c, p);
    return IMAGO_OK;
}

// ---------------------------------------------------------------------------
// valence density and traces (valeCharge.F90:230-393)

// W(:,s) = sqrt(f_s) C(:,s): zher with alpha = f_s summed over states is
// rho = W W^H because every population is >= 0 (populate.F90:190-485).
template <bool COMPLEX>
__global__ void k_scale_vectors(const double *__restrict__ C, const double *__restrict__ occ, int n,
                                int nocc, int ns, int nk, long ldstride, double *__restrict__ W)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)n * nocc;
    if (t >= per * nk) return;
    const int k = (int)(t / per);
    const long q = t % per;
    const int s = (int)(q / n);
    const double f = occ[(size_t)k * ns + s];
    const double w = f > 0.0 ? sqrt(f) : 0.0;
    if (COMPLEX) {
        const double2 v = reinterpret_cast<const double2 *>(C)[(size_t)k * ldstride + q];
        reinterpret_cast<double2 *>(W)[t] = make_double2(w * v.x, w * v.y);
    } else {
        W[t] = w * C[(size_t)k * ldstride + q];
    }
}

// rho in the block layout with the matrixElementMult weights folded in
// (matrixSubs.F90:66-112): off-diagonal entries x2, diagonal entries real
// part only, the strict lower triangle of a diagonal block 0.  rho holds
// the upper triangle (zherk); one work-group per (pair, k).
template <bool COMPLEX>
__global__ void k_pack_weighted(const double *__restrict__ rho, int n, long nbe,
                                const PairRec *__restrict__ pairs, const int *__restrict__ cumul,
                                int natoms, double *__restrict__ rw)
{
    const PairRec pr = pairs[blockIdx.x];
    const int k = blockIdx.y;
    const int ns1 = cumul_states(pr.i, cumul, natoms, n), ns2 = cumul_states(pr.j, cumul, natoms, n);
    const long r0 = cumul[pr.i], c0 = cumul[pr.j];
    const long nn = (long)n * n;
    for (int e = threadIdx.x; e < ns1 * ns2; e += blockDim.x) {
        const long row = r0 + e % ns1, col = c0 + e / ns1;
        const size_t o = (size_t)k * nbe + pr.blk + e;
        if (COMPLEX) {
            double2 w = make_double2(0.0, 0.0);
            if (row <= col) {
                const double2 v = reinterpret_cast<const double2 *>(rho)[(size_t)k * nn + row + col * n];
                w = row == col ? make_double2(v.x, 0.0) : make_double2(2.0 * v.x, 2.0 * v.y);
            }
            reinterpret_cast<double2 *>(rw)[o] = w;
        } else {
            double w = 0.0;
            if (row <= col) {
                const double v = rho[(size_t)k * nn + row + col * n];
                w = row == col ? v : 2.0 * v;
            }
            rw[o] = w;
        }
    }
}

// partial[term][blk] = sum over a chunk of the term's panel of M . rw
// (matrixElementMult summed over k-points); TG terms per work-group share
// each chunk of the weighted density rw.
template <int TG>
__global__ __launch_bounds__(256) void k_trace_partial(const double *__restrict__ M, long len,
                                                       int nterms, const double *__restrict__ rw,
                                                       int nblk, double *__restrict__ partial)
{
    const int blk = blockIdx.x;
    const int t0 = blockIdx.y * TG;
    M += (size_t)t0 * len;
    const long chunk = ((len + nblk - 1) / nblk + 1) & ~1L;
    const long lo = blk * chunk, hi = min(len, lo + chunk);
    double s[TG];
#pragma unroll
    for (int t = 0; t < TG; ++t) s[t] = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const double w = rw[i];
#pragma unroll
        for (int t = 0; t < TG; ++t)
            if (t0 + t < nterms) s[t] += M[(size_t)t * len + i] * w;
    }
    __shared__ double red[TG][257];
#pragma unroll
    for (int t = 0; t < TG; ++t) red[t][threadIdx.x] = s[t];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int t = 0; t < TG; ++t) red[t][threadIdx.x] += red[t][threadIdx.x + w];
        __syncthreads();
    }
    if ((int)threadIdx.x < TG && t0 + (int)threadIdx.x < nterms)
        partial[(size_t)(t0 + threadIdx.x) * nblk + blk] = red[threadIdx.x][0];
}

__global__ void k_trace_final(const double *__restrict__ partial, int nblk, int nterms,
                              double *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nterms) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += partial[(size_t)t * nblk + b];
    out[t] = s;
}

// the same sum for few terms with many chunks: one work-group per term,
// 256 fixed strides then a tree
__global__ __launch_bounds__(256) void k_trace_final_wide(const double *__restrict__ partial,
                                                          int nblk, double *__restrict__ out)
{
    __shared__ double red[256];
    const int t = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += partial[(size_t)t * nblk + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[t] = red[0];
}

// rho = W W^H for every k-point on the matrix cores: the lower 64 x 64
// tiles, each written to both triangles (W is n x nocc per k-point)
static int density_gemm(imago_ctx *c, bool cx, int n, int nocc, int nk, const double *W)
{
    GemmArgs g;
    g.M = g.N = n;
    g.K = nocc;
    g.lda = g.ldb = g.ldc = n;
    g.batch = nk;
    g.A = g.B = W;
    g.sa1 = g.sb1 = (long)n * nocc;
    g.C = c->rho;
    g.sc1 = (long)n * n;
    g.herm = true;
    HIPCHK(launch_gemm_mfma(g, cx, false, true, c->stream));
    return IMAGO_OK;
}

extern "C" int imago_density_traces(imago_ctx *c, const double *occupations, double *traces3,
                                    double *pot_rho)
{
    if (!c || !c->has_system || !c->have_vectors)
        return fail(IMAGO_E_STATE, "imago_density_traces needs imago_secular first");
    if (!occupations) return fail(IMAGO_E_ARG, "null occupations");
    const HostSys &h = c->h;
    const int ns = c->num_states;
    // states past the last one with a positive occupation at any k-point add
    // nothing to rho, so the rank-k update stops there
    int nocc = 0;
    for (long q = 0; q < (long)h.nk * ns; ++q) {
        if (occupations[q] < 0.0) return fail(IMAGO_E_ARG, "negative occupation at %ld", q);
        if (occupations[q] > 0.0) nocc = std::max(nocc, (int)(q % ns) + 1);
    }
    HIPCHK(hipSetDevice(c->device));
    RCHK(finish_columns(c, nocc));
    const long nn = (long)h.n * h.n;
    const long per = panel_doubles(c);
    const bool cx = !h.gamma;
    const int dsz = cx ? 2 : 1;
    if (!c->rho) RCHK(dev_alloc(c, &c->rho, (size_t)nn * h.nk * dsz));
    if (!c->rhow) RCHK(dev_alloc(c, &c->rhow, (size_t)per));
    if (!c->occ) RCHK(dev_alloc(c, &c->occ, (size_t)h.nk * h.n));
    double *W = nullptr;
    RCHK(dev_alloc(c, &W, (size_t)h.n * std::max(nocc, 1) * h.nk * dsz));
    HIPCHK(hipMemcpyAsync(c->occ, occupations, sizeof(double) * h.nk * ns, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    const long tot = (long)h.n * nocc * h.nk;
    if (tot > 0) {
        if (cx)
            hipLaunchKernelGGL(k_scale_vectors<true>, dim3((unsigned)((tot + 255) / 256)), dim3(256),
                               0, c->stream, (const double *)c->Hfull, c->occ, h.n, nocc, ns, h.nk,
                               nn, W);
        else
            hipLaunchKernelGGL(k_scale_vectors<false>, dim3((unsigned)((tot + 255) / 256)),
                               dim3(256), 0, c->stream, c->Hreal, c->occ, h.n, nocc, ns, h.nk, nn,
                               W);
        HIPCHK(hipGetLastError());
    }
    if (nocc == 0) {
        HIPCHK(hipMemsetAsync(c->rho, 0, sizeof(double) * nn * h.nk * dsz, c->stream));
    } else if (cx) {
        // rho = W W^H as a full product: rocBLAS's zherk runs at ~1 TFLOP/s
        // here (n = 896, k = 130), its zgemm near the FP64 matrix rate
        RCHK(density_gemm(c, true, h.n, nocc, h.nk, W));
    } else {
        RCHK(density_gemm(c, false, h.n, nocc, 1, W));
    }
    if (cx)
        hipLaunchKernelGGL(k_pack_weighted<true>, dim3(c->npairs, h.nk), dim3(256), 0, c->stream,
                           c->rho, h.n, c->nbe, c->pairs, c->d.cumul, h.natoms, c->rhow);
    else
        hipLaunchKernelGGL(k_pack_weighted<false>, dim3(c->npairs, 1), dim3(256), 0, c->stream,
                           c->rho, h.n, c->nbe, c->pairs, c->d.cumul, h.natoms, c->rhow);
    HIPCHK(hipGetLastError());
    // chunks per term: the three single-panel traces get enough work-groups
    // to fill the device on their own; the potDim panel terms 64 each
    const int nblk3 = 1024, nblk = 64;
    const int nterms = h.pot_dim + 3;
    if (!c->partial) RCHK(dev_alloc(c, &c->partial, (size_t)3 * nblk3 + (size_t)h.pot_dim * nblk));
    double *part_ep = c->partial + (size_t)3 * nblk3;
    double *outd = nullptr;
    RCHK(dev_alloc(c, &outd, (size_t)nterms));
    // S, NP, KE (one term each), then the potential terms: from the stored
    // panel, or directly from the primitive blocks
    const bool direct = !c->assembled[3];
    const int order3[3] = {0, 2, 1};
    for (int q = 0; q < 3; ++q)
        hipLaunchKernelGGL(k_trace_partial<1>, dim3(nblk3, 1), dim3(256), 0, c->stream,
                           c->panel[order3[q]], per, 1, c->rhow, nblk3,
                           c->partial + (size_t)q * nblk3);
    if (h.pot_dim > 0 && !direct)
        hipLaunchKernelGGL(k_trace_partial<1>, dim3(nblk, h.pot_dim), dim3(256), 0, c->stream,
                           c->panel[3], per, h.pot_dim, c->rhow, nblk, part_ep);
    if (h.pot_dim > 0 && direct) {
        RCHK(ep_build_runs(c));
        const int ne = c->ne <= 4 ? 4 : 16, ent = 64 * ne;
        if (!h.gamma && !c->foldX) RCHK(dev_alloc(c, &c->foldX, (size_t)c->nimgs * ent));
        EpdLaunch L{c->d,          c->pairs,     c->phase,       c->ep_tuples,
                    c->ep_item_tuple, c->pitem, c->ep_blocks,   nullptr,
                    h.gamma ? c->rhow : c->foldX, ent, c->type_term0, c->truns,
                    c->ntruns,     c->ival,      c->npairs,      c->lmax,
                    ne,            h.gamma != 0, c->stream};
        HIPCHK(hipEventRecord(c->ev2, c->stream));
        HIPCHK(launch_epd_fold(L, c->rhow, c->foldX));
        HIPCHK(launch_epd_trace(L));
        HIPCHK(hipEventRecord(c->ev3, c->stream));
        if (!c->tchunk)
            RCHK(dev_alloc(c, &c->tchunk, (size_t)h.npt * kEpdTraceChunks * h.maxpa));
        HIPCHK(launch_epd_trace_final(c->ival, c->item_jexp, c->ep_iseg, c->npairs, h.npt, h.maxpa,
                                      c->d.pot_nalpha, c->d.term_type, c->d.term_alpha,
                                      h.pot_dim, c->tchunk, outd + 3, c->stream));
    }
    hipLaunchKernelGGL(k_trace_final_wide, dim3(3), dim3(256), 0, c->stream, c->partial, nblk3,
                       outd);
    if (h.pot_dim > 0 && !direct)
        hipLaunchKernelGGL(k_trace_final, dim3((h.pot_dim + 255) / 256), dim3(256), 0, c->stream,
                           part_ep, nblk, h.pot_dim, outd + 3);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->timing[3] = ms;
    RCHK(read_finish_events(c));
    if (h.pot_dim > 0 && direct) {
        float ms9 = 0;
        HIPCHK(hipEventElapsedTime(&ms9, c->ev2, c->ev3));
        c->timing[9] = ms9;
    }
    std::vector<double> res(nterms);
    HIPCHK(ctx_memcpy(c, res.data(), outd, sizeof(double) * nterms, hipMemcpyDeviceToHost));
    if (traces3)
        for (int q = 0; q < 3; ++q) traces3[q] = res[q];
    if (pot_rho)
        for (int j = 0; j < h.pot_dim; ++j) pot_rho[j] = res[3 + j];
    dev_release(c, W);
    dev_release(c, outd);
    return IMAGO_OK;
}

// ---------------------------------------------------------------------------
// single generalized eigenproblems (solveZHEGV / solveDSYGV)

// A(i, j) = conj(A(j, i)) for i > j and Im A(i, i) = 0: the matrix as zhegv
// 'U' / dsygv 'U' define it from their upper triangle.  The reference's
// caller fills only the upper triangle (secularEqn.F90:140 zeroes valeVale,
// unpackMatrix with fullFlag 0 writes rows <= column, matrixSubs.F90:
// 426-433), and the solve below reads both triangles.
template <bool COMPLEX>
__global__ __launch_bounds__(256) void k_herm_from_upper(double *__restrict__ A, int n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * n) return;
    const int r = (int)(t % n), col = (int)(t / n);
    if (COMPLEX) {
        double2 *a = reinterpret_cast<double2 *>(A);
        if (r == col) {
            a[t].y = 0.0;
        } else if (r > col) {
            const double2 u = a[(size_t)r * n + col];
            a[t] = make_double2(u.x, -u.y);
        }
    } else if (r > col) {
        A[t] = A[(size_t)r * n + col];
    }
}

static int single_gv(imago_ctx *c, bool cx, int n, double *A, double *B, double *w, int *info)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (n <= 0 || !A || !B || !w) return fail(IMAGO_E_ARG, "bad eigenproblem arguments (n=%d)", n);
    HIPCHK(hipSetDevice(c->device));
    const size_t nn = (size_t)n * n * (cx ? 2 : 1);
    double *dA, *dB, *dw, *de;
    int *di;
    HIPCHK(hipMalloc(&dA, nn * sizeof(double)));
    HIPCHK(hipMalloc(&dB, nn * sizeof(double)));
    HIPCHK(hipMalloc(&dw, n * sizeof(double)));
    HIPCHK(hipMalloc(&de, n * sizeof(double)));
    HIPCHK(hipMalloc(&di, sizeof(int)));
    HIPCHK(ctx_memcpy(c, dA, A, nn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(ctx_memcpy(c, dB, B, nn * sizeof(double), hipMemcpyHostToDevice));
    {
        const unsigned g = (unsigned)(((long)n * n + 255) / 256);
        if (cx)
            hipLaunchKernelGGL(k_herm_from_upper<true>, dim3(g), dim3(256), 0, c->stream, dA, n);
        else
            hipLaunchKernelGGL(k_herm_from_upper<false>, dim3(g), dim3(256), 0, c->stream, dA, n);
        HIPCHK(hipGetLastError());
    }
    int inf = 0;
    int rc = solve_gv_batched(c, cx, n, 1, dA, dB, dw, de, di, &inf);
    // B returns the factor in its upper triangle and, as in LAPACK, its
    // strictly lower triangle as it came in (the device copy's is overwritten)
    std::vector<double> hB(nn);
    if (rc == IMAGO_OK && ctx_memcpy(c, A, dA, nn * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
        ctx_memcpy(c, hB.data(), dB, nn * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
        ctx_memcpy(c, w, dw, n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess) {
        const int es = cx ? 2 : 1;
        for (int col = 0; col < n; ++col)
            for (int row = 0; row <= col; ++row)
                for (int x = 0; x < es; ++x) {
                    const size_t at = ((size_t)col * n + row) * es + x;
                    B[at] = hB[at];
                }
    } else if (rc == IMAGO_OK) {
        rc = fail(IMAGO_E_HIP, "copy of eigenproblem results failed");
    }
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dw);
    (void)hipFree(de);
    (void)hipFree(di);
    if (rc != IMAGO_OK) return rc;
    if (info) *info = inf;
    if (inf > n) return fail(IMAGO_E_NOTPD, "%s error code=%d: B not positive definite",
                             cx ? "ZHEGV" : "DSYGV", inf);
    if (inf != 0) return fail(IMAGO_E_SOLVE, "%s error code=%d", cx ? "ZHEGV" : "DSYGV", inf);
    return IMAGO_OK;
}

extern "C" int imago_zhegv(imago_ctx *c, int n, double *A, double *B, double *w, int *info)
{
    return single_gv(c, true, n, A, B, w, info);
}

extern "C" int imago_dsygv(imago_ctx *c, int n, double *A, double *B, double *w, int *info)
{
    return single_gv(c, false, n, A, B, w, info);
}

extern "C" int imago_ep_counts(imago_ctx *c, long *tuples, long *items, int *block_doubles,
                               long *partial_doubles)
{
    if (!c || !c->has_system) return fail(IMAGO_E_STATE, "no system");
    if (!c->ep_screened) return fail(IMAGO_E_STATE, "EP terms not assembled");
    const int sh = (c->lmax + 1) * (c->lmax + 1);
    if (tuples) *tuples = c->ep_ntuples;
    if (items) *items = c->ep_nitems;
    if (block_doubles) *block_doubles = sh * sh;
    if (partial_doubles) {
        long w = 0;
        if (c->ep_runs)
            HIPCHK(ctx_memcpy(c, &w, c->hrun_woff + c->npairs, sizeof(long), hipMemcpyDeviceToHost));
        *partial_doubles = w;
    }
    return IMAGO_OK;
}

// ---- electrostatics of the potential-side Gaussians (elecStat.f90) ------

static int es_args(imago_ctx *c, double th, imago::EsArgs &A, const double **dsizes)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (!c->has_system) return fail(IMAGO_E_STATE, "imago_set_system must come first");
    if (!(th > 0.0) || !std::isfinite(th))
        return fail(IMAGO_E_ARG, "log_elec_thresh must be positive and finite (got %g)", th);
    const HostSys &h = c->h;
    if (h.maxpa > 64) return fail(IMAGO_E_ARG, "more than 64 alphas in a potential type");
    RCHK(dev_upload(c, dsizes, h.cell_sizes));
    A = imago::EsArgs{};
    A.nsites = h.nsites;
    A.npt = h.npt;
    A.pot_dim = h.pot_dim;
    A.ncells = h.ncells;
    A.maxpa = h.maxpa;
    A.site_pos = c->d.pot_pos;
    A.site_type = c->d.pot_site_type;
    A.type_first = c->d.pot_first;
    A.type_mult = c->d.pot_mult;
    A.type_nalpha = c->d.pot_nalpha;
    A.type_alphas = c->d.pot_alphas;
    A.type_z = c->d.pot_z;
    A.type_nuc = c->d.pot_nuc_alpha;
    A.type_term0 = c->type_term0;
    A.term_type = c->d.term_type;
    A.term_alpha = c->d.term_alpha;
    A.cell_dims = c->d.cell_dims;
    A.cell_sizes = *dsizes;
    std::memcpy(A.rv, h.rv, sizeof A.rv);
    std::memcpy(A.inv, h.inv, sizeof A.inv);
    A.th = th;
    RCHK(dev_alloc(c, &A.msep, (size_t)h.nsites * h.nsites));
    return IMAGO_OK;
}

// neutralAndNuclearQPot into the device buffer *out: potAlphaOverlap,
// nonLocalNeutQPot, localNeutQPot (potDim x potDim each), nonLocalNucQPot,
// localNucQPot (potDim each); potAlphaOverlap is also kept for
// imago_pot_solve.  The caller releases *out.
static int elecstat_run(imago_ctx *c, double log_elec_thresh, double **out)
{
    imago::EsArgs A;
    const double *dsizes = nullptr;
    if (c) HIPCHK(hipSetDevice(c->device));
    RCHK(es_args(c, log_elec_thresh, A, &dsizes));
    const size_t P = (size_t)c->h.pot_dim, PP = P * P;
    double *buf = nullptr;
    RCHK(dev_alloc(c, &buf, 3 * PP + 2 * P));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(imago::launch_es_neutral_nuclear(A, buf, buf + PP, buf + 2 * PP, buf + 3 * PP,
                                            buf + 3 * PP + P, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    // potAlphaOverlap stays on the device for imago_pot_solve
    imago_ctx::SpdMat &m = c->spd[0];
    spd_release(c, m);
    RCHK(dev_alloc(c, &m.A, PP));
    HIPCHK(hipMemcpyAsync(m.A, buf, PP * 8, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->timing[11] = ms;
    dev_release(c, (void *)dsizes);
    dev_release(c, A.msep);
    *out = buf;
    return IMAGO_OK;
}

extern "C" int imago_elecstat(imago_ctx *c, double log_elec_thresh, double *pot_alpha_overlap,
                              double *non_local_neut_q_pot, double *local_neut_q_pot,
                              double *non_local_nuc_q_pot, double *local_nuc_q_pot)
{
    if (!pot_alpha_overlap || !non_local_neut_q_pot || !local_neut_q_pot ||
        !non_local_nuc_q_pot || !local_nuc_q_pot)
        return fail(IMAGO_E_ARG, "null output array");
    if (!c || !c->has_system) return fail(IMAGO_E_STATE, "imago_set_system must come first");
    // potAlphaOverlap is formed again: imago_scf_setup must follow before
    // the next imago_scf_update
    c->scf.ready = false;
    double *buf = nullptr;
    RCHK(elecstat_run(c, log_elec_thresh, &buf));
    const size_t P = (size_t)c->h.pot_dim, PP = P * P;
    HIPCHK(hipMemcpyAsync(pot_alpha_overlap, buf, PP * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(non_local_neut_q_pot, buf + PP, PP * 8, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(local_neut_q_pot, buf + 2 * PP, PP * 8, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(non_local_nuc_q_pot, buf + 3 * PP, P * 8, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(local_nuc_q_pot, buf + 3 * PP + P, P * 8, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dev_release(c, buf);
    return IMAGO_OK;
}

// residualQ into the device buffer *outp (numPotTypes x potDim, released by
// the caller)
static int residual_q_run(imago_ctx *c, double log_elec_thresh, int num_cells_recip,
                          const double *cell_dims_recip, const double *cell_sizes_recip,
                          double **outp)
{
    if (!cell_dims_recip || !cell_sizes_recip) return fail(IMAGO_E_ARG, "null array");
    if (num_cells_recip < 1) return fail(IMAGO_E_ARG, "no reciprocal cells");
    for (int q = 1; q < num_cells_recip; ++q)
        if (!(cell_sizes_recip[q] > 0.0))
            return fail(IMAGO_E_ARG, "reciprocal cell %d has size %g: cell 0 must be the only "
                                     "origin", q, cell_sizes_recip[q]);
    imago::EsArgs A;
    const double *dsizes = nullptr;
    if (c) HIPCHK(hipSetDevice(c->device));
    RCHK(es_args(c, log_elec_thresh, A, &dsizes));
    const HostSys &h = c->h;
    // element runs in site order: the reference recomputes its reciprocal
    // tables when the element of a type's first site changes
    // (elecStat.f90:814-858) and keeps them for the types that follow
    std::vector<int> site_src(h.nsites), src_type, ord_type;
    int cur_el = -1;
    for (int i = 0; i < h.nsites; ++i) {
        const int t = h.site_type[i];
        if (i == 0 || h.site_type[i - 1] != t) {
            ord_type.push_back(t);
            if (h.pot_el[t] != cur_el) {
                cur_el = h.pot_el[t];
                src_type.push_back(t);
            }
            if (h.pot_nalpha[t] > h.pot_nalpha[src_type.back()]) {
                dev_release(c, (void *)dsizes);
                return fail(IMAGO_E_ARG, "potential type %d has more alphas than type %d, whose "
                                         "tables its element run shares", t, src_type.back());
            }
        }
        site_src[i] = (int)src_type.size() - 1;
    }
    if ((int)ord_type.size() != h.npt) {
        dev_release(c, (void *)dsizes);
        return fail(IMAGO_E_ARG, "potential types without sites");
    }
    imago::EsRecip R{};
    R.nk = imago::es_recip_nk(h.maxpa);
    const int bi = imago::es_recip_bi(R.nk);
    std::vector<int> blk;
    for (int i = 0; i < h.nsites;) {
        int n = 1;
        while (n < bi && i + n < h.nsites && site_src[i + n] == site_src[i]) ++n;
        blk.push_back(i);
        blk.push_back(n);
        blk.push_back(site_src[i]);
        i += n;
    }
    R.nblk = (int)blk.size() / 3;
    R.nsrc = (int)src_type.size();

    double min_pot = 10000.0;   // potTypes.f90:213-250
    for (int t = 0; t < h.npt; ++t) min_pot = std::min(min_pot, h.pot_alphas[(size_t)t * h.maxpa]);
    R.min_pot = min_pot;
    {
        // realCellVolume as lattice.f90:195-216 leaves it: (a x b) . c
        const double *a = h.rv, *b = h.rv + 3, *cc = h.rv + 6;
        double v = 0.0;
        for (int j = 0; j < 3; ++j) {
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            v = v + (a[j1] * b[j2] - a[j2] * b[j1]) * cc[j];
        }
        R.volume = std::fabs(v);
    }
    // The reciprocal sum leaves out cell 0, the origin (elecStat.f90:986-990).
    // +G and -G add the same term (cos is even), so one cell of each exact
    // pair is summed with weight 2; a cell without a partner keeps weight 1.
    std::vector<double> cdr, csr, cw;
    {
        auto key = [](const double *v, double sg) {
            std::array<uint64_t, 3> k;
            for (int x = 0; x < 3; ++x) {
                double t = sg * v[x];
                if (t == 0.0) t = 0.0;
                std::memcpy(&k[x], &t, sizeof t);
            }
            return k;
        };
        std::map<std::array<uint64_t, 3>, int> where;
        for (int q = 1; q < num_cells_recip; ++q)
            where.emplace(key(cell_dims_recip + 3 * (size_t)q, 1.0), q);
        std::vector<char> used(num_cells_recip, 0);
        for (int q = 1; q < num_cells_recip; ++q) {
            if (used[q]) continue;
            const auto it = where.find(key(cell_dims_recip + 3 * (size_t)q, -1.0));
            double w = 1.0;
            if (it != where.end() && it->second > q && !used[it->second]) {
                used[it->second] = 1;
                w = 2.0;
            }
            cdr.insert(cdr.end(), cell_dims_recip + 3 * (size_t)q,
                       cell_dims_recip + 3 * (size_t)q + 3);
            csr.push_back(cell_sizes_recip[q]);
            cw.push_back(w);
        }
        if (csr.empty()) {   // no cell but the origin: an empty sum
            cdr.assign(3, 0.0);
            csr.push_back(1.0);
            cw.push_back(0.0);
        }
    }
    const double *dcdr = nullptr, *dcsr = nullptr, *dcw = nullptr;
    const int *dsrc = nullptr, *dsite_src = nullptr, *dord = nullptr, *dblk = nullptr;
    R.ncr = (int)csr.size();
    R.chunks = imago::es_recip_chunks(h.nsites, R.nblk, R.ncr);
    RCHK(dev_upload(c, &dcdr, cdr));
    RCHK(dev_upload(c, &dcsr, csr));
    RCHK(dev_upload(c, &dcw, cw));
    RCHK(dev_upload(c, &dsrc, src_type));
    RCHK(dev_upload(c, &dsite_src, site_src));
    RCHK(dev_upload(c, &dord, ord_type));
    RCHK(dev_upload(c, &dblk, blk));
    R.cdr = dcdr;
    R.csr = dcsr;
    R.cw = dcw;
    R.src_type = dsrc;
    R.site_src = dsite_src;
    R.ord_type = dord;
    R.blk = dblk;
    const size_t nout = (size_t)h.npt * h.pot_dim;
    double *E = nullptr, *S = nullptr, *out = nullptr;
    double2 *CS = nullptr;
    RCHK(dev_alloc(c, &E, imago::es_recip_e_doubles(R.nsrc, R.ncr, R.nk)));
    RCHK(dev_alloc(c, &CS, (size_t)R.ncr * h.nsites));
    RCHK(dev_alloc(c, &S, imago::es_recip_s_doubles(R.chunks, h.nsites, R.nk)));
    RCHK(dev_alloc(c, &out, nout));
    R.E = E;
    R.CS = CS;
    R.S = S;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(imago::launch_es_residual_q(A, R, out, c->stream, c->ev2, c->ev3));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f, ms_r = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    HIPCHK(hipEventElapsedTime(&ms_r, c->ev2, c->ev3));
    c->timing[12] = ms;
    c->timing[13] = ms_r;
    for (const void *p : {(const void *)dcdr, (const void *)dcsr, (const void *)dcw,
                          (const void *)dsrc,
                          (const void *)dsite_src, (const void *)dord, (const void *)dblk,
                          (const void *)E, (const void *)CS, (const void *)S,
                          (const void *)dsizes, (const void *)A.msep})
        dev_release(c, const_cast<void *>(p));
    *outp = out;
    return IMAGO_OK;
}

extern "C" int imago_residual_q(imago_ctx *c, double log_elec_thresh, int num_cells_recip,
                                const double *cell_dims_recip, const double *cell_sizes_recip,
                                double *non_local_residual_q)
{
    if (!non_local_residual_q) return fail(IMAGO_E_ARG, "null array");
    double *out = nullptr;
    RCHK(residual_q_run(c, log_elec_thresh, num_cells_recip, cell_dims_recip, cell_sizes_recip,
                        &out));
    const size_t nout = (size_t)c->h.npt * c->h.pot_dim;
    HIPCHK(ctx_memcpy(c, non_local_residual_q, out, nout * 8, hipMemcpyDeviceToHost));
    dev_release(c, out);
    return IMAGO_OK;
}

// ---- exchange-correlation mesh (exchCorr.f90) ---------------------------

// makeSampleVectors (exchCorr.f90:920-980): the generalised spiral of n
// points on the sphere (E. B. Saff and A. B. J. Kuijlaars, Math.
// Intelligencer 19 (1997) 5), with the reference's constant 3.6.
static void xc_sample_vectors(int n, std::vector<double> &v)
{
    std::vector<double> h(n), th(n), ph(n);
    h[0] = -1.0;
    h[n - 1] = 1.0;
    for (int k = 2; k <= n - 1; ++k) h[k - 1] = -1.0 + 2.0 * (k - 1.0) / (n - 1.0);
    for (int k = 0; k < n; ++k) th[k] = std::acos(h[k]);
    ph[0] = 0.0;
    ph[n - 1] = 0.0;
    const double two_pi = 2.0 * 3.141592653589793;
    for (int k = 2; k <= n - 1; ++k) {
        ph[k - 1] = ph[k - 2] + 3.6 / std::sqrt((double)n) * 1.0 / std::sqrt(1.0 - h[k - 1] * h[k - 1]);
        while (ph[k - 1] > two_pi) ph[k - 1] = ph[k - 1] - two_pi;
    }
    v.assign(3 * (size_t)n, 0.0);
    for (int k = 0; k < n; ++k) {
        v[3 * k] = std::sin(th[k]) * std::cos(ph[k]);
        v[3 * k + 1] = std::sin(th[k]) * std::sin(ph[k]);
        v[3 * k + 2] = std::cos(th[k]);
    }
}

static imago::XcArgs xc_args(imago_ctx *c)
{
    const HostSys &h = c->h;
    imago::XcArgs X{};
    X.nsites = h.nsites;
    X.npt = h.npt;
    X.pot_dim = h.pot_dim;
    X.ncells = h.ncells;
    X.maxpa = h.maxpa;
    X.nreps = std::min(64, h.ncells);
    X.nvec = c->xcp.num_sample_vectors;
    X.site_pos = c->d.pot_pos;
    X.site_type = c->d.pot_site_type;
    X.site_cov = c->xc_sitecov;
    X.type_first = c->d.pot_first;
    X.type_mult = c->d.pot_mult;
    X.type_nalpha = c->d.pot_nalpha;
    X.type_alphas = c->d.pot_alphas;
    X.type_term0 = c->type_term0;
    X.cell_dims = c->d.cell_dims;
    X.cell_sizes = c->xc_sizes;
    X.pls = c->xc_pls;
    X.vec = c->xc_vec;
    std::memcpy(X.rv, h.rv, sizeof X.rv);
    std::memcpy(X.inv, h.inv, sizeof X.inv);
    X.th = c->xc_th;
    double min_pot = 10000.0;   // potTypes.f90:213-250
    for (int t = 0; t < h.npt; ++t) min_pot = std::min(min_pot, h.pot_alphas[(size_t)t * h.maxpa]);
    X.max_radius = std::sqrt(c->xcp.r_sample_out / min_pot);
    return X;
}

extern "C" int imago_xc_mesh(imago_ctx *c, const imago_xc_mesh_params *params,
                             double log_elec_thresh, const double *covalent_radius,
                             int *max_num_ray_points)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (!c->has_system) return fail(IMAGO_E_STATE, "imago_set_system must come first");
    if (!params || !covalent_radius) return fail(IMAGO_E_ARG, "null argument");
    // a new mesh replaces the operators imago_scf_setup built
    c->scf.ready = false;
    const imago_xc_mesh_params &p = *params;
    if (p.num_sample_vectors < 1 || p.num_sample_vectors > 100000)
        return fail(IMAGO_E_ARG, "num_sample_vectors %d out of range", p.num_sample_vectors);
    if (!(p.r_sample_space > 0.0 && p.r_sample_space < 1.0) || !(p.r_sample_in > 0.0) ||
        !(p.r_sample_out > 0.0))
        return fail(IMAGO_E_ARG, "radial sampling needs rIn > 0, rOut > 0, 0 < spacing < 1");
    if (!(p.weight_in >= 0.0) || !(p.weight_out >= 0.0))
        return fail(IMAGO_E_ARG, "angular weights must be >= 0");
    if (!(log_elec_thresh > 0.0) || !std::isfinite(log_elec_thresh))
        return fail(IMAGO_E_ARG, "log_elec_thresh must be positive and finite");
    HIPCHK(hipSetDevice(c->device));
    HostSys &h = c->h;
    if (h.maxpa > 64) return fail(IMAGO_E_ARG, "more than 64 alphas in a potential type");
    for (void *q : {(void *)c->xc_pts, (void *)c->xc_w, (void *)c->xc_pls, (void *)c->xc_vec,
                    (void *)c->xc_sitecov, (void *)c->xc_sizes, (void *)c->xc_dmax})
        if (q) dev_release(c, q);
    c->xc_pts = c->xc_w = c->xc_pls = c->xc_vec = c->xc_sitecov = c->xc_sizes = nullptr;
    c->xc_dmax = nullptr;
    for (const auto &q : c->xc_cache)
        for (void *x : {(void *)q.B, (void *)q.rows, (void *)q.cnt, (void *)q.rank,
                        (void *)q.slot, (void *)q.first})
            if (x) dev_release(c, x);
    c->xc_cache.clear();
    c->xc_cached = false;
    {
        // an overlap of the previous mesh no longer applies
        spd_release(c, c->spd[1]);
    }
    c->xc_ready = false;
    c->xcp = p;
    c->xc_th = log_elec_thresh;
    const int ns = h.nsites, nv = p.num_sample_vectors;
    std::vector<double> vec, cov(ns), pls(3 * (size_t)ns, 0.0);
    xc_sample_vectors(nv, vec);
    const double small = 1.0e-8;
    for (int i = 0; i < ns; ++i) {
        cov[i] = covalent_radius[h.site_type[i]];
        if (cov[i] < small) continue;
        double L[3];
        find_lattice_vector(h, &h.site_pos[3 * (size_t)i], L);   // exchCorr.f90:137-148
        for (int x = 0; x < 3; ++x) pls[3 * i + x] = h.site_pos[3 * i + x] - L[x];
    }
    {
        const double *q = nullptr;
        RCHK(dev_upload(c, &q, vec));
        c->xc_vec = const_cast<double *>(q);
        RCHK(dev_upload(c, &q, cov));
        c->xc_sitecov = const_cast<double *>(q);
        RCHK(dev_upload(c, &q, pls));
        c->xc_pls = const_cast<double *>(q);
        RCHK(dev_upload(c, &q, h.cell_sizes));
        c->xc_sizes = const_cast<double *>(q);
    }
    const imago::XcArgs X = xc_args(c);
    double *dmax = nullptr;
    RCHK(dev_alloc(c, &dmax, (size_t)ns * nv));
    HIPCHK(imago::launch_xc_maxrad(X, dmax, c->stream));
    std::vector<double> maxrad((size_t)ns * nv);
    HIPCHK(hipMemcpyAsync(maxrad.data(), dmax, maxrad.size() * 8, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dev_release(c, dmax);
    // ray points and radial weights of every site (exchCorr.f90:492-540)
    const double angw[2] = {p.weight_in, p.weight_out};
    const double wf = (1.0 - p.r_sample_space) / std::sqrt(p.r_sample_space);
    const double sqsp = std::sqrt(p.r_sample_space);
    std::vector<double> pts, w;
    c->xc_npts.assign(ns, 0);
    c->xc_off.assign(ns + 1, 0);
    int most = 0;
    for (int i = 0; i < ns; ++i) {
        c->xc_off[i] = (int)w.size();
        const int t = h.site_type[i];
        if (cov[i] < small) continue;
        const double cov_half = cov[i] / 2.0;
        const double min_radius =
            std::sqrt(p.r_sample_in / h.pot_alphas[(size_t)t * h.maxpa + h.pot_nalpha[t] - 1]);
        const double *ls = &pls[3 * (size_t)i];
        int np = 0;
        for (int j = 0; j < nv; ++j) {
            const double *v = &vec[3 * (size_t)j];
            double r = maxrad[(size_t)i * nv + j];
            int wi = 1;
            double epf = 0.5;
            for (int guard = 0;; ++guard) {
                if (guard > 100000)
                    return fail(IMAGO_E_ARG, "ray of site %d does not reach rSampleIn", i);
                for (int x = 0; x < 3; ++x) pts.push_back(r * v[x] + ls[x]);
                w.push_back(angw[wi] * wf * epf * (r * r * r));
                ++np;
                epf = 1.0;
                if (r < cov_half || r < min_radius) wi = 0;
                if (angw[wi] < small) break;
                if (r < min_radius) {
                    r = r * sqsp;
                    for (int x = 0; x < 3; ++x) pts.push_back(0.75 * r * v[x] + ls[x]);
                    w.push_back(angw[0] * (r * r * r) / 3.0);
                    ++np;
                    break;
                }
                r = r * p.r_sample_space;
            }
        }
        c->xc_npts[i] = np;
        most = std::max(most, np);
    }
    c->xc_off[ns] = (int)w.size();
    c->xc_total = (long)w.size();
    c->xc_maxpts = most;
    {
        std::vector<double> dm(ns, 0.0);
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < nv; ++j) dm[i] = std::max(dm[i], maxrad[(size_t)i * nv + j]);
        const double *q = nullptr;
        RCHK(dev_upload(c, &q, dm));
        c->xc_dmax = const_cast<double *>(q);
    }
    if (w.empty()) {
        pts.assign(3, 0.0);
        w.assign(1, 0.0);
    }
    {
        const double *q = nullptr;
        RCHK(dev_upload(c, &q, pts));
        c->xc_pts = const_cast<double *>(q);
        RCHK(dev_upload(c, &q, w));
        c->xc_w = const_cast<double *>(q);
    }
    c->xc_ready = true;
    if (max_num_ray_points) *max_num_ray_points = most;
    return IMAGO_OK;
}

extern "C" int imago_xc_rho_op(imago_ctx *c, int site, int gga, int *num_ray_points,
                               double *radial_weight, double *exch_rho_op)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (!c->xc_ready) return fail(IMAGO_E_STATE, "imago_xc_mesh must come first");
    if (site < 0 || site >= c->h.nsites) return fail(IMAGO_E_ARG, "site %d out of range", site);
    if (!radial_weight || !exch_rho_op) return fail(IMAGO_E_ARG, "null output array");
    HIPCHK(hipSetDevice(c->device));
    const int P = c->h.pot_dim, mp = std::max(c->xc_maxpts, 1), np = c->xc_npts[site];
    const int nop = gga ? 10 : 1;
    const size_t nout = (size_t)nop * mp * P;
    double *op = nullptr;
    RCHK(dev_alloc(c, &op, nout));
    HIPCHK(hipMemsetAsync(op, 0, nout * 8, c->stream));
    const imago::XcArgs X = xc_args(c);
    HIPCHK(imago::launch_xc_op(X, gga != 0, c->xc_pts + 3 * (size_t)c->xc_off[site], np, mp, op,
                               c->stream));
    std::fill(radial_weight, radial_weight + mp, 0.0);
    if (np > 0)
        HIPCHK(hipMemcpyAsync(radial_weight, c->xc_w + c->xc_off[site], (size_t)np * 8,
                              hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(exch_rho_op, op, nout * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dev_release(c, op);
    if (num_ray_points) *num_ray_points = np;
    return IMAGO_OK;
}

// exchCorrOverlap, kept on the device for imago_pot_solve; copied to
// exch_corr_overlap unless it is null
static int xc_overlap_run(imago_ctx *c, double *exch_corr_overlap)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (!c->xc_ready) return fail(IMAGO_E_STATE, "imago_xc_mesh must come first");
    HIPCHK(hipSetDevice(c->device));
    const int P = c->h.pot_dim, ns = c->h.nsites;
    const int mp = std::max(c->xc_maxpts, 1);
    const imago::XcArgs X = xc_args(c);
    // Sites in batches.  For each site, the terms its points can touch
    // (a superset from the geometry) are its rows; the operator of its
    // points goes straight into a block B_b (points x rows, scaled by
    // sqrt(w)), one strided-batched GEMM forms B_b^T B_b for every site of
    // the batch (zero-padded to the batch's largest row and point counts),
    // and the upper triangles are added into C site by site in order.
    const int batch = std::max(1, std::min(ns, 64));
    double *C = nullptr;
    unsigned char *nz = nullptr;
    int *rows = nullptr, *cnt = nullptr, *slot = nullptr, *first = nullptr, *rank = nullptr;
    int *err = nullptr;
    RCHK(dev_alloc(c, &C, (size_t)P * P));
    RCHK(dev_alloc(c, &nz, (size_t)batch * P));
    RCHK(dev_alloc(c, &rows, (size_t)batch * P));
    RCHK(dev_alloc(c, &rank, (size_t)batch * P));
    RCHK(dev_alloc(c, &cnt, (size_t)batch));
    RCHK(dev_alloc(c, &slot, (size_t)batch * mp));
    RCHK(dev_alloc(c, &first, (size_t)batch));
    RCHK(dev_alloc(c, &err, 1));
    double *A = nullptr, *Cs = nullptr;
    size_t a_cap = 0, c_cap = 0;
    HIPCHK(hipMemsetAsync(C, 0, (size_t)P * P * 8, c->stream));
    HIPCHK(hipMemsetAsync(err, 0, sizeof(int), c->stream));
    std::vector<int> hslot((size_t)batch * mp), hcnt(batch), hfirst(batch);
    std::vector<hipEvent_t> gev;   // pairs around the batched GEMMs
    double gflop = 0.0;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int s0 = 0; s0 < ns; s0 += batch) {
        const int nb = std::min(batch, ns - s0);
        const int p0 = c->xc_off[s0], p1 = c->xc_off[s0 + nb];
        const int npb = p1 - p0;
        if (npb == 0) continue;
        for (int b = 0; b < nb; ++b) {
            hfirst[b] = c->xc_off[s0 + b] - p0;
            for (int q = c->xc_off[s0 + b]; q < c->xc_off[s0 + b + 1]; ++q) hslot[q - p0] = b;
        }
        HIPCHK(hipMemcpyAsync(slot, hslot.data(), (size_t)npb * 4, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipMemcpyAsync(first, hfirst.data(), (size_t)nb * 4, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(imago::launch_xc_rows(X, s0, nb, c->xc_dmax, nz, c->stream));
        HIPCHK(imago::launch_xc_compact(nz, nb, P, rows, cnt, c->stream));
        HIPCHK(hipMemcpyAsync(hcnt.data(), cnt, (size_t)nb * 4, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        int mm = 0, nn = 0;
        for (int b = 0; b < nb; ++b) {
            if (c->xc_npts[s0 + b] == 0) continue;
            mm = std::max(mm, hcnt[b]);
            nn = std::max(nn, c->xc_npts[s0 + b]);
        }
        if (mm == 0 || nn == 0) continue;
        const size_t a_need = (size_t)nb * mm * nn, c_need = (size_t)nb * mm * mm;
        if (a_need > a_cap) {
            if (A) dev_release(c, A);
            RCHK(dev_alloc(c, &A, a_need));
            a_cap = a_need;
        }
        if (c_need > c_cap) {
            if (Cs) dev_release(c, Cs);
            RCHK(dev_alloc(c, &Cs, c_need));
            c_cap = c_need;
        }
        HIPCHK(hipMemsetAsync(rank, 0xff, (size_t)nb * P * 4, c->stream));
        HIPCHK(imago::launch_xc_rank(rows, cnt, nb, P, rank, c->stream));
        HIPCHK(hipMemsetAsync(A, 0, a_need * 8, c->stream));
        HIPCHK(imago::launch_xc_op_rows(X, c->xc_pts + 3 * (size_t)p0, c->xc_w + p0, npb, slot,
                                        first, rank, mm, nn, A, err, c->stream));
        // A holds B_b = (sqrt(w) op)^T, points x rows, so C_b = B_b^T B_b
        const double one = 1.0, zero = 0.0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        gev.push_back(e0);
        gev.push_back(e1);
        HIPCHK(hipEventRecord(e0, c->stream));
        BLASCHK(rocblas_dgemm_strided_batched(
            c->blas, rocblas_operation_transpose, rocblas_operation_none, mm, mm, nn, &one, A, nn,
            (rocblas_stride)mm * nn, A, nn, (rocblas_stride)mm * nn, &zero, Cs, mm,
            (rocblas_stride)mm * mm, nb));
        HIPCHK(hipEventRecord(e1, c->stream));
        gflop += 2.0 * mm * (double)mm * nn * nb;
        for (int b = 0; b < nb; ++b)
            if (c->xc_npts[s0 + b] > 0)
                HIPCHK(imago::launch_xc_scatter(Cs + (size_t)b * mm * mm, mm, rows + (size_t)b * P,
                                                hcnt[b], P, C, c->stream));
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (exch_corr_overlap)
        HIPCHK(hipMemcpyAsync(exch_corr_overlap, C, (size_t)P * P * 8, hipMemcpyDeviceToHost,
                              c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->timing[14] = ms;
    double gms = 0.0;
    for (size_t q = 0; q + 1 < gev.size(); q += 2) {
        float x = 0.f;
        HIPCHK(hipEventElapsedTime(&x, gev[q], gev[q + 1]));
        gms += x;
    }
    for (hipEvent_t e : gev) (void)hipEventDestroy(e);
    c->timing[15] = gms;
    c->xc_gemm_flops = gflop;
    {
        // exchCorrOverlap stays on the device for imago_pot_solve
        imago_ctx::SpdMat &m = c->spd[1];
        spd_release(c, m);
        m.A = C;
    }
    for (void *q : {(void *)A, (void *)Cs, (void *)nz, (void *)rows, (void *)rank, (void *)cnt,
                    (void *)slot, (void *)first, (void *)err})
        if (q) dev_release(c, q);
    if (herr)
        return fail(IMAGO_E_HIP, "exchange-correlation overlap: a point touched a term outside "
                                 "its site's row bound (internal error)");
    return IMAGO_OK;
}

extern "C" int imago_xc_overlap(imago_ctx *c, double *exch_corr_overlap)
{
    if (!exch_corr_overlap) return fail(IMAGO_E_ARG, "null output array");
    if (!c) return fail(IMAGO_E_ARG, "null context");
    // exchCorrOverlap is formed again: imago_scf_setup must follow before
    // the next imago_scf_update
    c->scf.ready = false;
    return xc_overlap_run(c, exch_corr_overlap);
}

// The operator blocks of every batch of sites (rows bounded from the
// geometry, the unscaled operator of the batch's points in B_b, points x
// rows), built once per mesh and kept: about potDim/6 x points doubles
// (4.4 GB for a-SiO2).
static int xc_build_cache(imago_ctx *c)
{
    const int P = c->h.pot_dim, ns = c->h.nsites;
    const int mp = std::max(c->xc_maxpts, 1);
    const imago::XcArgs X = xc_args(c);
    const int batch = std::max(1, std::min(ns, 64));
    unsigned char *nz = nullptr;
    int *err = nullptr;
    RCHK(dev_alloc(c, &nz, (size_t)batch * P));
    RCHK(dev_alloc(c, &err, 1));
    HIPCHK(hipMemsetAsync(err, 0, sizeof(int), c->stream));
    std::vector<int> hslot((size_t)batch * mp), hcnt(batch), hfirst(batch);
    for (int s0 = 0; s0 < ns; s0 += batch) {
        imago_ctx::XcBatch q{};
        q.s0 = s0;
        q.nb = std::min(batch, ns - s0);
        q.p0 = c->xc_off[s0];
        q.npb = c->xc_off[s0 + q.nb] - q.p0;
        if (q.npb == 0) continue;
        for (int b = 0; b < q.nb; ++b) {
            hfirst[b] = c->xc_off[s0 + b] - q.p0;
            for (int k = c->xc_off[s0 + b]; k < c->xc_off[s0 + b + 1]; ++k) hslot[k - q.p0] = b;
        }
        RCHK(dev_alloc(c, &q.rows, (size_t)q.nb * P));
        RCHK(dev_alloc(c, &q.rank, (size_t)q.nb * P));
        RCHK(dev_alloc(c, &q.cnt, (size_t)q.nb));
        RCHK(dev_alloc(c, &q.slot, (size_t)q.npb));
        RCHK(dev_alloc(c, &q.first, (size_t)q.nb));
        c->xc_cache.push_back(q);   // owned by the cache from here on
        HIPCHK(hipMemcpyAsync(q.slot, hslot.data(), (size_t)q.npb * 4, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipMemcpyAsync(q.first, hfirst.data(), (size_t)q.nb * 4, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(imago::launch_xc_rows(X, s0, q.nb, c->xc_dmax, nz, c->stream));
        HIPCHK(imago::launch_xc_compact(nz, q.nb, P, q.rows, q.cnt, c->stream));
        HIPCHK(hipMemcpyAsync(hcnt.data(), q.cnt, (size_t)q.nb * 4, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        int mm = 0, nn = 0;
        for (int b = 0; b < q.nb; ++b) {
            mm = std::max(mm, hcnt[b]);
            nn = std::max(nn, c->xc_npts[s0 + b]);
        }
        imago_ctx::XcBatch &qq = c->xc_cache.back();
        qq.mm = mm;
        qq.nn = nn;
        if (mm == 0 || nn == 0) continue;
        RCHK(dev_alloc(c, &qq.B, (size_t)q.nb * mm * nn));
        HIPCHK(hipMemsetAsync(qq.rank, 0xff, (size_t)q.nb * P * 4, c->stream));
        HIPCHK(imago::launch_xc_rank(qq.rows, qq.cnt, q.nb, P, qq.rank, c->stream));
        HIPCHK(hipMemsetAsync(qq.B, 0, (size_t)q.nb * mm * nn * 8, c->stream));
        HIPCHK(imago::launch_xc_op_rows(X, c->xc_pts + 3 * (size_t)q.p0, nullptr, q.npb, qq.slot,
                                        qq.first, qq.rank, mm, nn, qq.B, err, c->stream));
    }
    int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dev_release(c, nz);
    dev_release(c, err);
    if (herr)
        return fail(IMAGO_E_HIP, "exchange-correlation potential: a point touched a term outside "
                                 "its site's row bound (internal error)");
    c->xc_cached = true;
    return IMAGO_OK;
}

// Point-value and partial-sum scratch the cached operator blocks need.
static void xc_scratch_sizes(const imago_ctx *c, size_t *v_need, size_t *p_need)
{
    *v_need = 1;
    *p_need = 1;
    for (const auto &q : c->xc_cache) {
        *v_need = std::max(*v_need, (size_t)q.nb * 3 * q.nn);
        *p_need = std::max(*p_need, (size_t)q.nb * 3 * q.mm);
    }
}

// The XC potential terms on the device: rho holds rho_total then rho_core
// (2 potDim), pot receives the 3 potDim terms; v and part scratch of the
// sizes above.  Enqueued on the context's stream.
static int xc_potential_dev(imago_ctx *c, const double *rho, double *pot, double *v,
                            double *part)
{
    const int P = c->h.pot_dim;
    HIPCHK(hipMemsetAsync(pot, 0, 3 * (size_t)P * 8, c->stream));
    for (const auto &q : c->xc_cache) {
        if (!q.B) continue;
        HIPCHK(hipMemsetAsync(v, 0, (size_t)q.nb * 3 * q.nn * 8, c->stream));
        HIPCHK(imago::launch_xc_point_pot(q.B, q.mm, q.nn, q.slot, q.first, q.rows, q.cnt, P,
                                          c->xc_w + q.p0, rho, rho + P, q.npb, v, c->stream));
        // part_b = B_b^T V_b (rows x 3)
        const double one = 1.0, zero = 0.0;
        BLASCHK(rocblas_dgemm_strided_batched(
            c->blas, rocblas_operation_transpose, rocblas_operation_none, q.mm, 3, q.nn, &one, q.B,
            q.nn, (rocblas_stride)q.mm * q.nn, v, q.nn, (rocblas_stride)3 * q.nn, &zero, part, q.mm,
            (rocblas_stride)3 * q.mm, q.nb));
        HIPCHK(imago::launch_xc_pot_add(part, q.mm, q.rank, q.nb, P, pot, c->stream));
    }
    return IMAGO_OK;
}

extern "C" int imago_xc_potential(imago_ctx *c, const double *rho_total, const double *rho_core,
                                  double *exch_corr_pot)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (!c->xc_ready) return fail(IMAGO_E_STATE, "imago_xc_mesh must come first");
    if (!rho_total || !exch_corr_pot) return fail(IMAGO_E_ARG, "null array");
    HIPCHK(hipSetDevice(c->device));
    const int P = c->h.pot_dim;
    if (!c->xc_cached) RCHK(xc_build_cache(c));
    std::vector<double> hr(2 * (size_t)P, 0.0);
    std::copy(rho_total, rho_total + P, hr.begin());
    if (rho_core) std::copy(rho_core, rho_core + P, hr.begin() + P);
    size_t v_need = 1, p_need = 1;
    xc_scratch_sizes(c, &v_need, &p_need);
    double *rho = nullptr, *pot = nullptr, *v = nullptr, *part = nullptr;
    RCHK(dev_alloc(c, &rho, 2 * (size_t)P));
    RCHK(dev_alloc(c, &pot, 3 * (size_t)P));
    RCHK(dev_alloc(c, &v, v_need));
    RCHK(dev_alloc(c, &part, p_need));
    HIPCHK(hipMemcpyAsync(rho, hr.data(), 2 * (size_t)P * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    RCHK(xc_potential_dev(c, rho, pot, v, part));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(exch_corr_pot, pot, 3 * (size_t)P * 8, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->timing[16] = ms;
    for (void *x : {(void *)rho, (void *)pot, (void *)v, (void *)part}) dev_release(c, x);
    return IMAGO_OK;
}

// ---- potential-side SPD solves (solveDPOSVX, interfaces.F90:236-293) ----

// d(i) = A(i, i) of an n x n column-major matrix
__global__ __launch_bounds__(256) void k_spd_diag(const double *__restrict__ A, int n,
                                                  double *__restrict__ d)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) d[i] = A[(size_t)i * (n + 1)];
}

// As(i, j) = s_i A(i, j) s_j on the upper triangle (j >= i), 0 below
__global__ __launch_bounds__(256) void k_spd_scale(const double *__restrict__ A,
                                                   const double *__restrict__ s, int n,
                                                   double *__restrict__ As)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)n * n) return;
    const int i = (int)(g % n), j = (int)(g / n);
    As[g] = i <= j ? s[i] * A[g] * s[j] : 0.0;
}

// F = the symmetric matrix whose upper triangle U holds (both triangles)
__global__ __launch_bounds__(256) void k_sym_full(const double *__restrict__ U, int n,
                                                  double *__restrict__ F)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)n * n) return;
    const int i = (int)(g % n), j = (int)(g / n);
    F[g] = i <= j ? U[g] : U[(size_t)i * n + j];
}

// B(i, r) *= s_i for an n x nrhs block
__global__ __launch_bounds__(256) void k_spd_rows(double *__restrict__ B,
                                                  const double *__restrict__ s, int n, int nrhs)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)n * nrhs) return;
    B[g] *= s[g % n];
}

// Matrix-vector products of the potential-side operators (potDim x potDim,
// a few right-hand sides), spread over the whole device: rocBLAS runs these
// shapes as GEMMs or row-block GEMVs on 16 work-groups (35-60 us for 8 MB).
// Every sum has a fixed order, so results do not depend on timing.
__device__ __forceinline__ double gv_wave_sum(double v) { return imago::wave_sum_xor(v); }

// y(j, r) = alpha * sum_i M(i, j) x(i, r) + beta * y(j, r) for an m x n
// column-major M (leading dimension ld) and R right-hand sides: one wave per
// column, lanes over rows
template <int R>
__global__ __launch_bounds__(256) void k_gemv_cols(const double *__restrict__ M, int m, int n,
                                                   int ld, const double *__restrict__ x, int ldx,
                                                   double alpha, double beta,
                                                   double *__restrict__ y, int ldy)
{
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (j >= n) return;
    const double *col = M + (size_t)j * ld;
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    int i = lane;
    for (; i + 192 < m; i += 256) {
        double a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = col[i + 64 * u];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = fma(a[u], x[i + 64 * u + (size_t)r * ldx], acc[r]);
    }
    for (; i < m; i += 64) {
        const double a = col[i];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = fma(a, x[i + (size_t)r * ldx], acc[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = gv_wave_sum(acc[r]);
    if (lane == 0)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            double *yr = y + j + (size_t)r * ldy;
            *yr = beta == 0.0 ? alpha * acc[r] : alpha * acc[r] + beta * *yr;
        }
}

// part(i, r, chunk) = sum over the chunk's columns j of M(i, j) x(j, r):
// one thread per row, column chunks of cw columns along blockIdx.y
template <int R>
__global__ __launch_bounds__(256) void k_gemv_rows_part(const double *__restrict__ M, int m,
                                                        int n, int ld,
                                                        const double *__restrict__ x, int ldx,
                                                        int cw, double *__restrict__ part)
{
    const int i = blockIdx.x * 256 + threadIdx.x, ch = blockIdx.y;
    if (i >= m) return;
    const int j0 = ch * cw, j1 = min(n, j0 + cw);
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    for (int j = j0; j < j1; ++j) {
        const double a = M[i + (size_t)j * ld];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = fma(a, x[j + (size_t)r * ldx], acc[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) part[((size_t)ch * R + r) * m + i] = acc[r];
}

// y(i, r) = alpha * sum over chunks (in order) of part(i, r, chunk) + beta * y(i, r)
template <int R>
__global__ __launch_bounds__(256) void k_gemv_rows_sum(const double *__restrict__ part, int m,
                                                       int nch, double alpha, double beta,
                                                       double *__restrict__ y, int ldy)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        double s = 0.0;
        for (int ch = 0; ch < nch; ++ch) s += part[((size_t)ch * R + r) * m + i];
        double *yr = y + i + (size_t)r * ldy;
        *yr = beta == 0.0 ? alpha * s : alpha * s + beta * *yr;
    }
}

template <int R>
static hipError_t gemv_launch(bool trans, int m, int n, const double *M, int ld, const double *x,
                              int ldx, double alpha, double beta, double *y, int ldy,
                              double *part, int cw, hipStream_t st)
{
    if (trans) {
        hipLaunchKernelGGL(k_gemv_cols<R>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, M, m,
                           n, ld, x, ldx, alpha, beta, y, ldy);
    } else {
        const int nch = (n + cw - 1) / cw;
        hipLaunchKernelGGL(k_gemv_rows_part<R>, dim3((unsigned)((m + 255) / 256), (unsigned)nch),
                           dim3(256), 0, st, M, m, n, ld, x, ldx, cw, part);
        hipLaunchKernelGGL(k_gemv_rows_sum<R>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st,
                           part, m, nch, alpha, beta, y, ldy);
    }
    return hipGetLastError();
}

// y = alpha op(M) x + beta y on the device (op(M) = M^T when trans) for
// nrhs <= 4 right-hand sides; x, y device arrays (leading dimensions ldx,
// ldy), M m x n column-major with leading dimension ld
static int dev_gemv(imago_ctx *c, bool trans, int m, int n, const double *M, int ld,
                    const double *x, int ldx, int nrhs, double alpha, double beta, double *y,
                    int ldy)
{
    if (nrhs < 1 || nrhs > 4) return fail(IMAGO_E_ARG, "dev_gemv: %d right-hand sides", nrhs);
    if (m <= 0 || n <= 0) return IMAGO_OK;
    constexpr int cw = 32;
    const int nch = (n + cw - 1) / cw;
    double *part = nullptr;
    if (!trans) {
        const size_t need = (size_t)nch * nrhs * m;
        if (need > c->gv_cap) {
            if (c->gv_part) {
                HIPCHK(hipStreamSynchronize(c->stream));
                dev_release(c, c->gv_part);
                c->gv_part = nullptr;
                c->gv_cap = 0;
            }
            RCHK(dev_alloc(c, &c->gv_part, need));
            c->gv_cap = need;
        }
        part = c->gv_part;
    }
    hipError_t e = hipSuccess;
    switch (nrhs) {
    case 1: e = gemv_launch<1>(trans, m, n, M, ld, x, ldx, alpha, beta, y, ldy, part, cw, c->stream); break;
    case 2: e = gemv_launch<2>(trans, m, n, M, ld, x, ldx, alpha, beta, y, ldy, part, cw, c->stream); break;
    case 3: e = gemv_launch<3>(trans, m, n, M, ld, x, ldx, alpha, beta, y, ldy, part, cw, c->stream); break;
    default: e = gemv_launch<4>(trans, m, n, M, ld, x, ldx, alpha, beta, y, ldy, part, cw, c->stream); break;
    }
    HIPCHK(e);
    return IMAGO_OK;
}

// The equilibration of dposvx('E') (dpoequ / dlaqsy): s_i = 1/sqrt(a_ii),
// applied when min(s)/max(s) < 0.1 or the largest diagonal is outside
// [tiny/eps, eps/tiny]; then the Cholesky factor of the (scaled) upper
// triangle.  info as dpotrf (i > 0: leading minor i not positive), or
// i for a_ii <= 0 as dpoequ.
static int spd_factor(imago_ctx *c, imago_ctx::SpdMat &m, int n)
{
    // the diagonal, gathered on the device
    std::vector<double> diag(n), sc(n, 1.0);
    {
        double *dd = nullptr;
        RCHK(dev_alloc(c, &dd, (size_t)n));
        hipLaunchKernelGGL(k_spd_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           m.A, n, dd);
        HIPCHK(hipGetLastError());
        HIPCHK(ctx_memcpy(c, diag.data(), dd, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        dev_release(c, dd);
    }
    double smin = diag[0], amax = diag[0];
    for (int i = 0; i < n; ++i) {
        smin = std::min(smin, diag[i]);
        amax = std::max(amax, diag[i]);
    }
    m.info = 0;
    for (int i = 0; i < n; ++i)
        if (!(diag[i] > 0.0)) {
            m.info = i + 1;
            m.factored = true;
            return IMAGO_OK;
        }
    const double eps = 1.1102230246251565e-16, tiny = 2.2250738585072014e-308;
    const double small = tiny / eps, large = 1.0 / small;
    const double scond = std::sqrt(smin) / std::sqrt(amax);
    const bool equed = scond < 0.1 || amax < small || amax > large;
    if (equed)
        for (int i = 0; i < n; ++i) sc[i] = 1.0 / std::sqrt(diag[i]);
    const size_t nn = (size_t)n * n;
    RCHK(dev_alloc(c, &m.As, nn));
    RCHK(dev_alloc(c, &m.U, nn));
    const double *ds = nullptr;
    RCHK(dev_upload(c, &ds, sc));
    m.s = const_cast<double *>(ds);
    hipLaunchKernelGGL(k_spd_scale, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream,
                       m.A, m.s, n, m.As);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(m.U, m.As, nn * 8, hipMemcpyDeviceToDevice, c->stream));
    int *dinfo = nullptr;
    RCHK(dev_alloc(c, &dinfo, 1));
    RCHK(potrf_batched(c, false, n, 1, m.U, dinfo, false));
    int hinfo = 0;
    HIPCHK(ctx_memcpy(c, &hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost));
    m.info = hinfo;
    m.factored = true;
    if (hinfo == 0) {
        // the inverse from the factor (dpotri: U^-1 U^-T), kept with the
        // equilibrated matrix in full storage: every later solve is
        // products with them
        const unsigned g = (unsigned)((nn + 255) / 256);
        double *Ui = nullptr;
        RCHK(dev_alloc(c, &m.Af, nn));
        RCHK(dev_alloc(c, &m.Ainv, nn));
        RCHK(dev_alloc(c, &Ui, nn));
        BLASCHK(rocblas_dtrtri(c->blas, rocblas_fill_upper, rocblas_diagonal_non_unit, n, m.U, n,
                               Ui, n));
        imago::GemmArgs ga;
        ga.M = ga.N = ga.K = n;
        ga.lda = ga.ldb = ga.ldc = n;
        ga.A = Ui;
        ga.B = Ui;
        ga.C = m.Ainv;
        ga.amask = imago::kTriUpper;
        ga.bmask = imago::kTriLower;
        ga.herm = true;
        HIPCHK(imago::launch_gemm_mfma(ga, false, false, true, c->stream));
        hipLaunchKernelGGL(k_sym_full, dim3(g), dim3(256), 0, c->stream, m.As, n, m.Af);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        dev_release(c, Ui);
    }
    dev_release(c, dinfo);
    return IMAGO_OK;
}

// x = S (As^-1 (S b)) for a factored overlap with the inverse from its
// factor, then two steps of refinement on the scaled system
// (r = S b - As x, x += As^-1 r); b and x device arrays of n x nrhs, scr
// 2 n nrhs doubles of device scratch.  Enqueued on the context's stream.
static int pot_solve_dev(imago_ctx *c, imago_ctx::SpdMat &m, int n, int nrhs, const double *b,
                         double *x, double *scr)
{
    if (!m.factored || m.info != 0 || !m.Ainv || !m.Af || !m.s)
        return fail(IMAGO_E_STATE, "solve with an overlap that has no factor");
    const size_t nb = (size_t)n * nrhs;
    double *bs = scr, *r = scr + nb;
    const unsigned g = (unsigned)((nb + 255) / 256);
    HIPCHK(hipMemcpyAsync(bs, b, nb * 8, hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_spd_rows, dim3(g), dim3(256), 0, c->stream, bs, m.s, n, nrhs);
    HIPCHK(hipGetLastError());
    // Ainv and Af hold both triangles of symmetric matrices, so up to four
    // right-hand sides go through the column-dot product M^T x = M x
    const bool gv = nrhs <= 4;
    const double one = 1.0, mone = -1.0, zero = 0.0;
    if (gv)
        RCHK(dev_gemv(c, true, n, n, m.Ainv, n, bs, n, nrhs, 1.0, 0.0, x, n));
    else
        BLASCHK(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, n, nrhs, n,
                              &one, m.Ainv, n, bs, n, &zero, x, n));
    for (int it = 0; it < 2; ++it) {
        HIPCHK(hipMemcpyAsync(r, bs, nb * 8, hipMemcpyDeviceToDevice, c->stream));
        if (gv) {
            RCHK(dev_gemv(c, true, n, n, m.Af, n, x, n, nrhs, -1.0, 1.0, r, n));
            RCHK(dev_gemv(c, true, n, n, m.Ainv, n, r, n, nrhs, 1.0, 1.0, x, n));
        } else {
            BLASCHK(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, n,
                                  nrhs, n, &mone, m.Af, n, x, n, &one, r, n));
            BLASCHK(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, n,
                                  nrhs, n, &one, m.Ainv, n, r, n, &one, x, n));
        }
    }
    hipLaunchKernelGGL(k_spd_rows, dim3(g), dim3(256), 0, c->stream, x, m.s, n, nrhs);
    HIPCHK(hipGetLastError());
    return IMAGO_OK;
}

extern "C" int imago_pot_solve(imago_ctx *c, int matrix, int nrhs, double *b, int *info)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (matrix != IMAGO_SOLVE_POT_OVERLAP && matrix != IMAGO_SOLVE_XC_OVERLAP)
        return fail(IMAGO_E_ARG, "unknown matrix %d", matrix);
    if (nrhs < 1 || !b) return fail(IMAGO_E_ARG, "no right-hand sides");
    imago_ctx::SpdMat &m = c->spd[matrix == IMAGO_SOLVE_POT_OVERLAP ? 0 : 1];
    if (!m.A)
        return fail(IMAGO_E_STATE, "%s first", matrix == IMAGO_SOLVE_POT_OVERLAP
                                                  ? "imago_elecstat" : "imago_xc_overlap");
    HIPCHK(hipSetDevice(c->device));
    const int n = c->h.pot_dim;
    if (!m.factored) RCHK(spd_factor(c, m, n));
    if (info) *info = m.info;
    if (m.info != 0)
        return fail(IMAGO_E_NOTPD, "the overlap is not positive definite (minor %d)", m.info);
    const size_t nb = (size_t)n * nrhs;
    double *db = nullptr, *x = nullptr, *scr = nullptr;
    RCHK(dev_alloc(c, &db, nb));
    RCHK(dev_alloc(c, &x, nb));
    RCHK(dev_alloc(c, &scr, 2 * nb));
    HIPCHK(hipMemcpyAsync(db, b, nb * 8, hipMemcpyHostToDevice, c->stream));
    RCHK(pot_solve_dev(c, m, n, nrhs, db, x, scr));
    HIPCHK(hipMemcpyAsync(b, x, nb * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dev_release(c, db);
    dev_release(c, x);
    dev_release(c, scr);
    return IMAGO_OK;
}

extern "C" int imago_xc_counts(imago_ctx *c, long *ray_points, double *gemm_flops)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (!c->xc_ready) return fail(IMAGO_E_STATE, "imago_xc_mesh must come first");
    if (ray_points) *ray_points = c->xc_total;
    if (gemm_flops) *gemm_flops = c->xc_gemm_flops;
    return IMAGO_OK;
}

// ---- the SCF iteration: populateStandard and makeSCFPot -----------------

extern "C" int imago_populate(int num_kpoints, int num_states, const double *eigenvalues,
                              const double *kpoint_weights, double num_electrons,
                              double *populations, double *occupied_energy, double *band_gap,
                              int *gap_kind)
{
    if (num_kpoints < 1 || num_states < 1 || !eigenvalues || !kpoint_weights || !populations)
        return fail(IMAGO_E_ARG, "bad populate arguments");
    imago::PopulateResult r;
    if (imago::populate_standard(num_kpoints, 1, num_states, eigenvalues, kpoint_weights,
                                 num_electrons, populations, &r) != 0)
        return fail(IMAGO_E_ARG, "%d states at %d k-points cannot hold %g electrons", num_states,
                    num_kpoints, num_electrons);
    if (occupied_energy) *occupied_energy = r.occupied_energy;
    if (band_gap) *band_gap = r.band_gap;
    if (gap_kind) *gap_kind = r.gap_kind;
    return IMAGO_OK;
}

// d_b = B_b u_b for the operator blocks of a batch of sites needs u_b, the
// potential change at the batch's compact rows (0 past a site's count)
__global__ __launch_bounds__(256) void k_scf_gather_rows(const double *__restrict__ v,
                                                         const int *__restrict__ rows,
                                                         const int *__restrict__ cnt, int P,
                                                         int mm, int nb, double *__restrict__ u)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= nb * mm) return;
    const int b = g / mm, r = g % mm;
    u[g] = r < cnt[b] ? v[rows[(size_t)b * P + r]] : 0.0;
}

// per site b of a batch: sum_k w_k d_k^2 and sum_k w_k over its points
// (d_b laid out [b][nn], the point order of the operator blocks)
__global__ __launch_bounds__(256) void k_scf_site_rms(const double *__restrict__ d, int nn,
                                                      const double *__restrict__ w,
                                                      const int *__restrict__ first, int nb,
                                                      int npb, double *__restrict__ out)
{
    __shared__ double s0[256], s1[256];
    const int b = blockIdx.x;
    const int k0 = first[b], k1 = b + 1 < nb ? first[b + 1] : npb;
    double a = 0.0, ww = 0.0;
    for (int k = k0 + (int)threadIdx.x; k < k1; k += 256) {
        const double x = d[(size_t)b * nn + (k - k0)];
        a += w[k] * x * x;
        ww += w[k];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = ww;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s0[threadIdx.x] += s0[threadIdx.x + h];
            s1[threadIdx.x] += s1[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * b] = s0[0];
        out[2 * b + 1] = s1[0];
    }
}

static void td_scratch_sizes(const imago_ctx *c, size_t *u_need, size_t *d_need);

// imago_scf_update's device vectors, carved out of ScfState::work in this
// order (P = potDim, NPT = max(P, potential types)); the groups marked
// "one copy" are contiguous so each phase of the update moves them in one
// transfer.
struct ScfWork {
    double *fit_b, *fit_x;        // 2P each: [potRho intgConsts], the fit
    double *scr;                  // 6P: solve scratch
    double *g4, *gv, *g6, *rho;   // P, P, NPT, 2P (g1 then zeros): one copy in
    double *e1, *e2, *e6, *xc;    // P, P, P, 3P: one copy out
    double *pes_b;                // 2P
    double *pes_x, *xc_x;         // 2P, 3P: one copy out
    double *outpot, *dd;          // P, 2P: one copy in
    double *sx, *uu, *sums;       // P, 2P, 2 nsites: one copy out
};

static ScfWork scf_work(const imago_ctx *c, double *base, size_t *total)
{
    const size_t P = (size_t)c->h.pot_dim, NPT = std::max<size_t>(P, (size_t)c->h.npt);
    ScfWork w;
    double *p = base;
    auto take = [&p](size_t n) {
        double *q = p;
        p += n;
        return q;
    };
    w.fit_b = take(2 * P);
    w.fit_x = take(2 * P);
    w.scr = take(6 * P);
    w.g4 = take(P);
    w.gv = take(P);
    w.g6 = take(NPT);
    w.rho = take(2 * P);
    w.e1 = take(P);
    w.e2 = take(P);
    w.e6 = take(P);
    w.xc = take(3 * P);
    w.pes_b = take(2 * P);
    w.pes_x = take(2 * P);
    w.xc_x = take(3 * P);
    w.outpot = take(P);
    w.dd = take(2 * P);
    w.sx = take(P);
    w.uu = take(2 * P);
    w.sums = take(2 * (size_t)c->h.nsites);
    if (total) *total = (size_t)(p - base);
    return w;
}

static size_t scf_work_doubles(const imago_ctx *c)
{
    size_t total = 0;
    (void)scf_work(c, nullptr, &total);
    return total;
}

extern "C" int imago_scf_setup(imago_ctx *c, const imago_scf_params *params,
                               const imago_xc_mesh_params *mesh, const double *covalent_radius,
                               int num_cells_recip, const double *cell_dims_recip,
                               const double *cell_sizes_recip)
{
    if (!c || !c->has_system) return fail(IMAGO_E_STATE, "imago_scf_setup before imago_set_system");
    if (!params || !mesh || !covalent_radius) return fail(IMAGO_E_ARG, "null argument");
    if (c->h.core_dim > 0)
        return fail(IMAGO_E_ARG, "the SCF update has no core charge (makeCoreRho): systems with "
                                 "core states run the integral and secular stages only");
    if (params->xc_code != 100)
        return fail(IMAGO_E_ARG, "xc_code %d: only 100 (Wigner, spin 1, LDA) is built",
                    params->xc_code);
    if (params->feedback_level < 0 || params->feedback_level > 2)
        return fail(IMAGO_E_ARG, "feedback level %d", params->feedback_level);
    HIPCHK(hipSetDevice(c->device));
    const HostSys &h = c->h;
    const int P = h.pot_dim;
    imago_ctx::ScfState &S = c->scf;
    for (void *q : {(void *)S.es_buf, (void *)S.resq, (void *)S.work,
                    (void *)S.xc_v, (void *)S.xc_part, (void *)S.td_u, (void *)S.td_d})
        if (q) dev_release(c, q);
    S = imago_ctx::ScfState{};
    S.p = *params;
    // intgConsts (potential.f90:150-160)
    const double pi32 = std::pow(3.141592653589793, 1.5);
    S.intg.assign(P, 0.0);
    for (int t = 0; t < h.npt; ++t)
        for (int a = 0; a < h.pot_nalpha[t]; ++a)
            S.intg[h.type_term0[t] + a] = (double)h.pot_mult[t] * pi32 /
                                          std::pow(h.pot_alphas[(size_t)t * h.maxpa + a], 1.5);
    // electrostatic operators: potAlphaOverlap (kept in spd[0]),
    // nonLocalNeutQPot and localNeutQPot on the device, the nuclear vectors
    // on the host
    double *es = nullptr;
    RCHK(elecstat_run(c, params->log_elec_thresh, &es));
    const size_t PP = (size_t)P * P;
    S.es_buf = es;
    S.neut_nl = es + PP;
    S.neut_l = es + 2 * PP;
    S.nuc_nl.resize(P);
    S.nuc_l.resize(P);
    HIPCHK(ctx_memcpy(c, S.nuc_nl.data(), es + 3 * PP, (size_t)P * 8, hipMemcpyDeviceToHost));
    HIPCHK(ctx_memcpy(c, S.nuc_l.data(), es + 3 * PP + P, (size_t)P * 8, hipMemcpyDeviceToHost));
    RCHK(residual_q_run(c, params->log_elec_thresh, num_cells_recip, cell_dims_recip,
                        cell_sizes_recip, &S.resq));
    // exchange-correlation mesh, overlap and operator blocks
    int maxpts = 0;
    RCHK(imago_xc_mesh(c, mesh, params->log_elec_thresh, covalent_radius, &maxpts));
    RCHK(xc_overlap_run(c, nullptr));
    if (!c->xc_cached) RCHK(xc_build_cache(c));
    // factor both overlaps now (the first solve would)
    for (int m = 0; m < 2; ++m)
        if (!c->spd[m].factored) RCHK(spd_factor(c, c->spd[m], P));
    // imago_scf_update's device scratch
    {
        size_t v_need, p_need, u_need, d_need;
        xc_scratch_sizes(c, &v_need, &p_need);
        td_scratch_sizes(c, &u_need, &d_need);
        RCHK(dev_alloc(c, &S.work, scf_work_doubles(c)));
        RCHK(dev_alloc(c, &S.xc_v, v_need));
        RCHK(dev_alloc(c, &S.xc_part, p_need));
        RCHK(dev_alloc(c, &S.td_u, u_need));
        RCHK(dev_alloc(c, &S.td_d, d_need));
    }
    S.iteration = 1;
    S.ready = true;
    return IMAGO_OK;
}

// The convergence measure testableDelta: the largest site RMS of the
// real-space potential change over the ray points
// (potentialUpdate.F90:1350-1431), per site sqrt(sum w d^2 / sum w) with
// d = exchRhoOp(:, k) . outputPot at each of its points.
// Row and point scratch of the convergence measure.
static void td_scratch_sizes(const imago_ctx *c, size_t *u_need, size_t *d_need)
{
    *u_need = 1;
    *d_need = 1;
    for (const auto &q : c->xc_cache) {
        *u_need = std::max(*u_need, (size_t)q.nb * q.mm);
        *d_need = std::max(*d_need, (size_t)q.nb * q.nn);
    }
}

// The per-site sums of the measure for the device vector v (potDim): sums
// receives (sum w d^2, sum w) per site; u, d scratch of the sizes above.
// Enqueued on the context's stream.
static int testable_delta_dev(imago_ctx *c, const double *v, double *sums, double *u, double *d)
{
    const int P = c->h.pot_dim;
    HIPCHK(hipMemsetAsync(sums, 0, 2 * (size_t)c->h.nsites * 8, c->stream));
    for (const auto &q : c->xc_cache) {
        if (!q.B) continue;
        const int ng = (q.nb * q.mm + 255) / 256;
        hipLaunchKernelGGL(k_scf_gather_rows, dim3(ng), dim3(256), 0, c->stream, v, q.rows, q.cnt,
                           P, q.mm, q.nb, u);
        HIPCHK(hipGetLastError());
        const double one = 1.0, zero = 0.0;
        BLASCHK(rocblas_dgemv_strided_batched(
            c->blas, rocblas_operation_none, q.nn, q.mm, &one, q.B, q.nn,
            (rocblas_stride)q.mm * q.nn, u, 1, (rocblas_stride)q.mm, &zero, d, 1,
            (rocblas_stride)q.nn, q.nb));
        hipLaunchKernelGGL(k_scf_site_rms, dim3(q.nb), dim3(256), 0, c->stream, d, q.nn,
                           c->xc_w + q.p0, q.first, q.nb, q.npb, sums + 2 * (size_t)q.s0);
        HIPCHK(hipGetLastError());
    }
    return IMAGO_OK;
}

// the largest site RMS from the sums (sites with ray points only)
static double testable_delta_of(const imago_ctx *c, const double *hs)
{
    double best = 0.0;
    for (int i = 0; i < c->h.nsites; ++i)
        if (c->xc_npts[i] > 0) best = std::max(std::sqrt(hs[2 * i] / hs[2 * i + 1]), best);
    return best;
}

extern "C" int imago_scf_update(imago_ctx *c, const double *traces3, const double *pot_rho,
                                double *pot_coeffs, double *rho_coeffs,
                                imago_scf_record *record)
{
    if (!c || !c->scf.ready)
        return fail(IMAGO_E_STATE, "imago_scf_update before imago_scf_setup (or after a call that "
                                   "formed its operators again)");
    if (!traces3 || !pot_rho || !pot_coeffs || !record) return fail(IMAGO_E_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    const HostSys &h = c->h;
    imago_ctx::ScfState &S = c->scf;
    const int P = h.pot_dim, NPT = std::max(P, h.npt);
    const std::vector<double> &ic = S.intg;
    const ScfWork W = scf_work(c, S.work, nullptr);
    imago_ctx::SpdMat &mp = c->spd[0], &mx = c->spd[1];
    auto dot = [P](const double *a, const double *b) {
        double t = 0.0;
        for (int i = 0; i < P; ++i) t += a[i] * b[i];
        return t;
    };
    // The update runs in four phases, each one transfer in, its kernels and
    // one transfer out with a single wait; the host does the vector
    // arithmetic between them in the reference's order.
    // ---- 1: charge fit, potAlphaOverlap [g1 g2] = [potRho intgConsts]
    // (:191-239)
    if (mp.info != 0) return fail(IMAGO_E_SOLVE, "Fitting dposvx failed. INFO= %d", mp.info);
    std::vector<double> g(2 * (size_t)P);
    std::copy(pot_rho, pot_rho + P, g.begin());
    std::copy(ic.begin(), ic.end(), g.begin() + P);
    HIPCHK(hipMemcpyAsync(W.fit_b, g.data(), 2 * (size_t)P * 8, hipMemcpyHostToDevice, c->stream));
    RCHK(pot_solve_dev(c, mp, P, 2, W.fit_b, W.fit_x, W.scr));
    HIPCHK(hipMemcpyAsync(g.data(), W.fit_x, 2 * (size_t)P * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<double> g1(g.begin(), g.begin() + P), g2(g.begin() + P, g.end());
    const double fitted = dot(ic.data(), g1.data());
    const double electron_diff = S.p.num_electrons - fitted;
    {
        const double scale = electron_diff / dot(ic.data(), g2.data());   // (:300-301)
        for (int i = 0; i < P; ++i) g2[i] = g1[i] + g2[i] * scale;
    }
    // no core charge: total = valence (:329-342)
    g1 = g2;
    std::vector<double> g4 = g1, g5(P, 0.0), g6(NPT, 0.0), g7(P, 0.0);
    for (int t = 0; t < h.npt; ++t) {   // neutral / residual / nuclear split (:416-484)
        const int i0 = h.type_term0[t], i1 = i0 + h.pot_nalpha[t] - 1;
        const double nuc = h.pot_z[t] * (double)h.pot_mult[t];
        double sum = 0.0;
        for (int j = i0; j <= i1; ++j) sum += ic[j] * g1[j];
        g5[i0] = (sum - nuc) / ic[i0];
        g4[i0] = g1[i0] - g5[i0];
        g6[t] = g5[i0];
        g7[i1] = -nuc / ic[i1];
    }
    // ---- 2: the electrostatic potential components (:536-606; the core
    // terms vanish) and the exchange-correlation potential on the mesh
    // (:720-1015) of rho_total = g1
    std::vector<double> in2(4 * (size_t)P + NPT, 0.0);
    std::copy(g4.begin(), g4.end(), in2.begin());
    for (int i = 0; i < P; ++i) in2[P + i] = g2[i] - g5[i];   // gv
    std::copy(g6.begin(), g6.end(), in2.begin() + 2 * P);
    std::copy(g1.begin(), g1.end(), in2.begin() + 2 * P + NPT);   // rho_total, then rho_core = 0
    HIPCHK(hipMemcpyAsync(W.g4, in2.data(), in2.size() * 8, hipMemcpyHostToDevice, c->stream));
    RCHK(dev_gemv(c, true, P, P, S.neut_nl, P, W.g4, P, 1, 1.0, 0.0, W.e1, P));
    RCHK(dev_gemv(c, true, P, P, S.neut_l, P, W.gv, P, 1, 1.0, 0.0, W.e2, P));
    RCHK(dev_gemv(c, true, h.npt, P, S.resq, h.npt, W.g6, h.npt, 1, 1.0, 0.0, W.e6, P));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    RCHK(xc_potential_dev(c, W.rho, W.xc, S.xc_v, S.xc_part));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    std::vector<double> out2(6 * (size_t)P);
    HIPCHK(hipMemcpyAsync(out2.data(), W.e1, out2.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->timing[16] = ms;
    }
    const double *e1 = out2.data(), *e2 = e1 + P, *e6 = e2 + P;
    const std::vector<double> &e4 = S.nuc_nl, &e5 = S.nuc_l;
    double es_energy = 0.0;
    for (int i = 0; i < P; ++i)
        es_energy += 0.5 * (g1[i] + g7[i]) * (e1[i] + e6[i]) + 0.5 * g2[i] * (e2[i] + e5[i]) +
                     0.5 * (g2[i] - g5[i]) * e5[i];
    es_energy = es_energy - dot(g2.data(), e4.data()) + traces3[1];
    std::vector<double> pes(2 * (size_t)P);
    for (int i = 0; i < P; ++i) {
        pes[i] = e1[i] + e2[i] - e4[i] + e5[i];
        pes[P + i] = e6[i];
    }
    // ---- 3: the electrostatic and exchange-correlation solves
    if (mx.info != 0)
        return fail(IMAGO_E_SOLVE, "Exchange-correlation dposvx failed. INFO= %d", mx.info);
    HIPCHK(hipMemcpyAsync(W.pes_b, pes.data(), pes.size() * 8, hipMemcpyHostToDevice, c->stream));
    RCHK(pot_solve_dev(c, mp, P, 2, W.pes_b, W.pes_x, W.scr));
    RCHK(pot_solve_dev(c, mx, P, 3, W.xc, W.xc_x, W.scr));
    std::vector<double> out3(5 * (size_t)P);
    HIPCHK(hipMemcpyAsync(out3.data(), W.pes_x, out3.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::copy(out3.begin(), out3.begin() + 2 * P, pes.begin());
    const double *xc = out3.data() + 2 * P;
    // ---- 4: exchCorrEnergy's product (:1033-1039; the core term is zero),
    // the convergence measure of outputPot = yl0 - xl0 and the feedback's
    // products with potAlphaOverlap
    std::vector<double> yl0(P), xl0(pot_coeffs, pot_coeffs + P), outpot(P);
    for (int i = 0; i < P; ++i) yl0[i] = pes[i] + pes[P + i] + xc[i];
    for (int i = 0; i < P; ++i) outpot[i] = yl0[i] - xl0[i];
    const bool mixing = S.p.feedback_level != 0 && S.iteration > 1;
    std::vector<double> in4(3 * (size_t)P, 0.0), rl0, d1, d2;
    std::copy(outpot.begin(), outpot.end(), in4.begin());
    if (mixing) {
        rl0.resize(P);
        d1.resize(P);
        d2.resize(P);
        for (int i = 0; i < P; ++i) {
            rl0[i] = yl0[i] - xl0[i];
            d1[i] = rl0[i] - (S.yl1[i] - S.xl1[i]);
            d2[i] = rl0[i] - (S.yl2[i] - S.xl2[i]);
        }
        std::copy(d1.begin(), d1.end(), in4.begin() + P);
        std::copy(d2.begin(), d2.end(), in4.begin() + 2 * P);
    }
    HIPCHK(hipMemcpyAsync(W.outpot, in4.data(), in4.size() * 8, hipMemcpyHostToDevice, c->stream));
    RCHK(dev_gemv(c, true, P, P, mp.A, P, W.xc_x + P, P, 1, 1.0, 0.0, W.sx, P));
    RCHK(testable_delta_dev(c, W.outpot, W.sums, S.td_u, S.td_d));
    if (mixing) RCHK(dev_gemv(c, false, P, P, mp.A, P, W.dd, P, 2, 1.0, 0.0, W.uu, P));
    std::vector<double> out4(3 * (size_t)P + 2 * (size_t)h.nsites);
    HIPCHK(hipMemcpyAsync(out4.data(), W.sx, out4.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double *sx = out4.data(), *u1 = sx + P, *u2 = u1 + P, *sums = u2 + P;
    const double xc_energy = dot(sx, g1.data());
    const double kinetic = traces3[2];
    const double total = kinetic + es_energy + xc_energy;
    record->iteration = S.iteration;
    record->electron_diff = electron_diff;
    record->kinetic_energy = kinetic;
    record->elec_stat_energy = es_energy;
    record->exch_corr_energy = xc_energy;
    record->total_energy = total;
    record->testable_delta = 0.0;
    record->converged = 0;
    if (rho_coeffs) {
        std::copy(g1.begin(), g1.end(), rho_coeffs);
        std::copy(g2.begin(), g2.end(), rho_coeffs + P);
    }
    if (total > 0.0)
        return fail(IMAGO_E_SOLVE, "Total energy is positive (%.8f Ha): the structure might have a "
                                   "serious problem", total);
    // mixing history (:1301-1340)
    if (!S.have_hist) {
        S.xl1 = S.xl2 = xl0;
        S.yl1 = S.yl2 = yl0;
        S.have_hist = true;
    }
    const double delta = testable_delta_of(c, sums);
    record->testable_delta = delta;
    // the two-step feedback (:1604-1666)
    double th1 = 0.0, th2 = 0.0;
    if (mixing) {
        const double s11 = dot(u1, d1.data()), t1 = dot(u1, rl0.data());
        const double s12 = dot(u2, d1.data()), s22 = dot(u2, d2.data());
        const double t2 = dot(u2, rl0.data());
        th1 = t1 / s11;
        if (S.p.feedback_level != 1 && S.iteration > 2) {
            const double det = s11 * s22 - s12 * s12;
            if (det / (s11 * s22) >= 0.01) {
                th1 = (s22 * t1 - s12 * t2) / det;
                th2 = (-s12 * t1 + s11 * t2) / det;
            }
        }
    }
    const double r = S.p.relax_factor;
    for (int i = 0; i < P; ++i)
        pot_coeffs[i] = (1.0 - r) * ((1.0 - th1 - th2) * xl0[i] + th1 * S.xl1[i] + th2 * S.xl2[i]) +
                        r * ((1.0 - th1 - th2) * yl0[i] + th1 * S.yl1[i] + th2 * S.yl2[i]);
    S.xl2 = S.xl1;
    S.xl1 = xl0;
    S.yl2 = S.yl1;
    S.yl1 = yl0;
    if (delta < S.p.convg_test)
        record->converged = 1;
    else
        S.iteration += 1;
    return IMAGO_OK;
}

extern "C" int imago_last_timing(imago_ctx *c, int stage, double *ms)
{
    if (!c || !ms || stage < 0 || stage > 22) return fail(IMAGO_E_ARG, "bad timing query");
    *ms = c->timing[stage];
    return IMAGO_OK;
}

int imago_ctx_device(imago_ctx *c) { return c->device; }
hipStream_t imago_ctx_stream(imago_ctx *c) { return c->stream; }
