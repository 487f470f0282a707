// Batched apply and factored solve over many independent plans, in FP64
// (precision 0), FP32 (1) or with BF16 factor matrices and FP32 arithmetic
// (2), both directions, and a BF16 solve refined with FP32 residuals.
//
// Every phase is one launch that covers all matrices: the blocks of matrix i
// are [pre[i], pre[i+1]) of the grid (a prefix table per phase, found by a
// binary search in the block), and each block runs the same device bodies
// as the single-plan kernels (plr_mm.h and plr_cap.h in FP64, plr_sp.h in
// single precision) on matrix i's table.
//   apply:    reduce (all levels) -> finish -> expand (leaves + all levels)
//   solve N:  leaf expand (Y = Dinv X), then for depth d = Lmax-1 .. 0 the
//             matrices deeper than d run reduce -> finish -> capacitance ->
//             expand at level d on Y in place
//   solve T:  Z = X, depth d = 0 .. Lmax-1 with Ut and V exchanged and the
//             transposed capacitance step, then Y = Dinv^T Z
//   refined:  Y = S(X); then iters - 1 times { r = X - A Y (FP32 apply),
//             Y += S(r) }, S the solve with BF16 factors
// (a matrix's own levels keep their order).  The whole sequence of a call
// is captured once into a HIP graph per (operation, precision, direction,
// RHS count) and replayed; the tables are rebuilt when a plan's device
// memory changes (its workspace generation).  The single-precision phases
// reuse each plan's FP64 workspaces (fewer bytes, same element counts).
//
// Tile shape: the rows per reduction task (tmax) trades parallelism (more,
// shorter tasks) against partial sums (more work for the finishing pass).
// dplrh_batch_tune times the apply + solve of the caller's batch for each
// candidate and keeps the fastest; without it the batch uses MM_TMAX.
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "plr_cap.h"
#include "plr_common.h"
#include "plr_mm.h"
#include "plr_sp.h"

// arithmetic (and right-hand side) type of a storage type
template <typename T>
struct BPrec
{
    typedef float C;
};
template <>
struct BPrec<double>
{
    typedef double C;
};

// per-matrix device view of one precision's factors and the workspaces
template <typename T>
struct BWs
{
    typedef typename BPrec<T>::C C;
    const T *D, *Dinv;
    const C *Kinv, *S;
    size_t koff[PLR_MAXL + 1], soff[PLR_MAXL + 1];
    int R[PLR_MAXL];
    int L, N;
    C *w; // apply coefficients, stride wstride per column
    size_t wstride;
    C *Tr, *W; // solve reduction / coefficients, stride tstride per column
    size_t tstride;
    C *part; // partial blocks of long segments
};

// per-matrix vectors of one call: X read, Y written, Z the solve's working
// block (Y itself for solve N)
struct BIo
{
    const void *X;
    void *Y, *Z;
    int ldx, ldy, ldz;
};

__device__ __forceinline__ int batch_find(const int *__restrict__ pre, int cnt, int bx)
{
    int lo = 0, hi = cnt; // pre[lo] <= bx < pre[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= bx)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// mode 0: apply (X -> w, alpha * t to the partner segment); mode 1: solve
// level (Z -> T, raw)
template <typename T>
__global__ __launch_bounds__(64) void kb_reduce(const MmTab *__restrict__ tab, int tsl, int slot,
                                                const int *__restrict__ pre, int cnt,
                                                const BIo *__restrict__ io,
                                                const BWs<T> *__restrict__ ws, int mode,
                                                typename BPrec<T>::C alpha)
{
    typedef typename BPrec<T>::C C;
    const int mi = batch_find(pre, cnt, blockIdx.x);
    const MmTab &a = tab[(size_t)mi * tsl + slot];
    const BWs<T> &b = ws[mi];
    const BIo &o = io[mi];
    const int task = blockIdx.x - pre[mi];
    const C *src = (const C *)(mode == 0 ? o.X : o.Z);
    const int lds = mode == 0 ? o.ldx : o.ldz;
    C *dst = mode == 0 ? b.w : b.Tr;
    const size_t dstride = mode == 0 ? b.wstride : b.tstride;
    if constexpr (std::is_same<T, double>::value)
        mm_reduce_body<1>(a, task, src, lds, alpha, 0, mode, b.part, dst, dstride);
    else
        sp_reduce_body<T>(a, task, src, lds, alpha, mode, b.part, dst, dstride);
}

template <typename T>
__global__ __launch_bounds__(256) void kb_finish(const MmTab *__restrict__ tab, int tsl, int slot,
                                                 const int *__restrict__ pre, int cnt,
                                                 const BWs<T> *__restrict__ ws, int mode,
                                                 typename BPrec<T>::C alpha)
{
    const int mi = batch_find(pre, cnt, blockIdx.x);
    const MmTab &a = tab[(size_t)mi * tsl + slot];
    const BWs<T> &b = ws[mi];
    const int bx = blockIdx.x - pre[mi];
    if constexpr (std::is_same<T, double>::value)
        mm_finish_body(a, bx, alpha, 0, mode, b.part, mode ? b.Tr : b.w,
                       mode ? b.tstride : b.wstride);
    else
        sp_finish_body(a, bx, alpha, mode, b.part, mode ? b.Tr : b.w, mode ? b.tstride : b.wstride);
}

// mode 0: apply Y = alpha*op(A) X + beta*Y; mode 1: leaves Y = op(Dinv) X;
// mode 2: level Z += Q W; mode 3: leaves Y = op(Dinv) Z.  NT: column tiles
// in registers (1 for at most 16 right-hand sides: more waves per SIMD)
template <typename T, int NT>
__global__ __launch_bounds__(256) void kb_expand(const MmTab *__restrict__ tab, int tsl, int slot,
                                                 const int *__restrict__ pre, int cnt,
                                                 const BIo *__restrict__ io,
                                                 const BWs<T> *__restrict__ ws, int mode, int trans,
                                                 typename BPrec<T>::C alpha,
                                                 typename BPrec<T>::C beta)
{
    typedef typename BPrec<T>::C C;
    const int mi = batch_find(pre, cnt, blockIdx.x);
    const MmTab &a = tab[(size_t)mi * tsl + slot];
    const BWs<T> &b = ws[mi];
    const BIo &o = io[mi];
    const int loc = blockIdx.x - pre[mi];
    const int rb = loc / a.ncg, cg = loc - rb * a.ncg;
    const T *D = mode == 0 ? b.D : mode == 2 ? nullptr : b.Dinv;
    const C *X = (const C *)(mode == 3 ? o.Z : mode == 2 ? nullptr : o.X);
    const int ldx = mode == 3 ? o.ldz : o.ldx;
    const C *w = mode == 0 ? b.w : mode == 2 ? b.W : nullptr;
    const size_t wstride = mode == 0 ? b.wstride : b.tstride;
    C *Y = (C *)(mode == 2 ? o.Z : o.Y);
    const int ldy = mode == 2 ? o.ldz : o.ldy;
    const C al = mode == 0 ? alpha : (C)1, be = mode == 0 ? beta : mode == 2 ? (C)1 : (C)0;
    if constexpr (std::is_same<T, double>::value)
        mm_expand_body<NT>(a, rb, cg, D, X, ldx, al, be, trans, w, wstride, Y, ldy);
    else
        sp_expand_body<T, NT>(a, rb, cg, D, X, ldx, al, be, trans, w, wstride, Y, ldy);
}

// capacitance step of level d, 16 columns per work-group (FP64 MFMA in every
// precision: the inputs and outputs have the batch's type)
template <typename T>
__global__ __launch_bounds__(256) void kb_cap(int d, const int *__restrict__ pre, int cnt,
                                              const BWs<T> *__restrict__ ws, int ncols, int trans)
{
    typedef typename BPrec<T>::C C;
    const int mi = batch_find(pre, cnt, blockIdx.x);
    const BWs<T> &b = ws[mi];
    const int loc = blockIdx.x - pre[mi], nodes = 1 << d;
    const int node = loc % nodes, cblk = loc / nodes;
    cap_body<C, C, false, 1>(node, cblk, b.Kinv + b.koff[d], b.R[d], b.S + b.soff[d], b.Tr,
                             b.tstride, b.W, b.tstride, trans, ncols, nullptr, 0);
}

// dst = alpha * src + beta * dst over each matrix's N x ncols block, src and
// dst one of X (0), Y (1), Z (2) of the call's vectors
template <typename C>
__global__ __launch_bounds__(256) void kb_axpby(const int *__restrict__ pre, int cnt,
                                                const BIo *__restrict__ io, const int *__restrict__ nrow,
                                                int ncols, int src, int dst, C alpha, C beta)
{
    const int mi = batch_find(pre, cnt, blockIdx.x);
    const BIo &o = io[mi];
    const int N = nrow[mi];
    const size_t e = (size_t)(blockIdx.x - pre[mi]) * 256 + threadIdx.x;
    const int c = (int)(e / N), r = (int)(e - (size_t)c * N);
    if (c >= ncols) return;
    const C *s = (const C *)(src == 0 ? o.X : src == 1 ? o.Y : o.Z);
    const int lds = src == 0 ? o.ldx : src == 1 ? o.ldy : o.ldz;
    C *t = (C *)(dst == 1 ? o.Y : o.Z);
    const int ldt = dst == 1 ? o.ldy : o.ldz;
    const C v = alpha * s[(size_t)c * lds + r];
    t[(size_t)c * ldt + r] = beta == (C)0 ? v : v + beta * t[(size_t)c * ldt + r];
}

int plr_mm_part(dplrh_plan *p, size_t blocks);

// phases: 0 apply reduce, 1 apply finish, 2 apply expand, 3 solve leaves,
// 4 vector update (N x nrhs), then per depth d four phases
// 5 + 4d + {0 reduce, 1 finish, 2 cap, 3 expand}
static constexpr int PH_LVL = 5;
// table slots per matrix: 0 apply, 1 leaves, 2 transposed apply, 3 + d solve
// level d, 3 + Lmax + d transposed solve level d
static constexpr int SL_LVL = 3;

struct batch_state
{
    int built_nrhs, built_fact, tmax;
    std::vector<unsigned> gen;
    MmTab *dtab;
    void *dws;
    int *dpre, *dnrow;
    std::vector<size_t> pre_off; // offset of each phase's prefix in dpre
    std::vector<int> pre_tot;    // blocks of each phase
    size_t cap_lds;
};

struct batch_graph
{
    int prec, op, trans, nrhs, iters;
    double alpha, beta;
    std::vector<BIo> hio; // the call's vectors, compared on every call
    BIo *dio;
    hipGraphExec_t exec;
};

struct dplrh_batch
{
    int count, Lmax, tsl;
    std::vector<dplrh_plan *> plans;
    batch_state st[3];
    std::vector<batch_graph> graphs;
    hipStream_t cap;
    // working blocks of the transposed and refined solves: per matrix two
    // N x nrhs blocks of doubles (ld N) at zoff[i]
    char *scratch;
    int scratch_nrhs;
    std::vector<size_t> zoff;
};

static void batch_drop_graphs(dplrh_batch *B)
{
    for (auto &g : B->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.dio) (void)hipFree(g.dio);
    }
    B->graphs.clear();
}

static void state_free(batch_state &s)
{
    void *bufs[] = {s.dtab, s.dws, s.dpre, s.dnrow};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    s.dtab = nullptr;
    s.dws = nullptr;
    s.dpre = s.dnrow = nullptr;
    s.built_nrhs = 0;
    s.built_fact = 0;
}

extern "C" int dplrh_batch_create(int count, dplrh_plan *const *plans, dplrh_batch **out)
{
    *out = nullptr;
    if (count <= 0) {
        plr_set_error("batch of %d plans", count);
        return -1;
    }
    for (int i = 0; i < count; ++i) {
        dplrh_plan *p = plans[i];
        if (!p || p->N == 0 || !plr_mm_ok(p)) {
            plr_set_error("batch member %d: the batched kernels need leaves of a multiple of 64 "
                          "rows and N >= 128",
                          i);
            return -1;
        }
    }
    dplrh_batch *B = new dplrh_batch();
    B->count = count;
    B->plans.assign(plans, plans + count);
    B->Lmax = 0;
    for (auto *p : B->plans) B->Lmax = p->L > B->Lmax ? p->L : B->Lmax;
    B->tsl = SL_LVL + 2 * B->Lmax;
    for (auto &s : B->st) {
        s.tmax = MM_TMAX;
        s.gen.assign(count, 0);
    }
    if (hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking) != hipSuccess) {
        delete B;
        plr_set_error("stream creation failed");
        return -10;
    }
    *out = B;
    return 0;
}

extern "C" void dplrh_batch_destroy(dplrh_batch *B)
{
    if (!B) return;
    batch_drop_graphs(B);
    for (auto &s : B->st) state_free(s);
    if (B->scratch) (void)hipFree(B->scratch);
    if (B->cap) (void)hipStreamDestroy(B->cap);
    delete B;
}

// factor operands of one precision
template <typename T>
struct BOps
{
    const T *U, *V, *Ut, *D, *Dinv;
    const typename BPrec<T>::C *S, *Kinv;
};

template <typename T>
static int batch_ops(dplrh_plan *p, int prec, int fact, BOps<T> *o)
{
    if (prec == 0) {
        const BOps<double> d = {p->U, p->V, p->f.Ut, p->D, p->f.Dinv, p->S, p->f.Kinv};
        memcpy(o, &d, sizeof(d));
        return fact && !p->f.done ? -4 : 0;
    }
    const plr_sp *s = p->sp[prec - 1];
    if (!s || (fact && !s->fact)) return -4;
    o->U = (const T *)s->U;
    o->V = (const T *)s->V;
    o->Ut = (const T *)s->Ut;
    o->D = (const T *)s->D;
    o->Dinv = (const T *)s->Dinv;
    o->S = (const typename BPrec<T>::C *)s->S;
    o->Kinv = (const typename BPrec<T>::C *)s->Kinv;
    return 0;
}

template <typename T>
static int batch_build(dplrh_batch *B, int prec, int nrhs, int fact)
{
    typedef typename BPrec<T>::C C;
    batch_state &S = B->st[prec];
    const int cnt = B->count, Lm = B->Lmax, tsl = B->tsl;
    // workspaces first: they may reallocate plan memory
    for (int i = 0; i < cnt; ++i) {
        dplrh_plan *p = B->plans[i];
        BOps<T> o;
        if (batch_ops<T>(p, prec, fact, &o)) {
            plr_set_error(prec == 0 ? "batched solve before a successful factorization of every plan"
                                    : "batch member %d has no single-precision factors of type %d%s",
                          i, prec, fact ? " with its factorization" : "");
            return -4;
        }
        int rc = plr_workspace(p, nrhs);
        if (!rc && fact) rc = plr_solve_reserve(p, nrhs);
        if (rc) return rc;
    }
    std::vector<MmTab> tabs((size_t)cnt * tsl);
    std::vector<BWs<T>> ws(cnt);
    std::vector<int> nrow(cnt);
    const int nph = PH_LVL + 4 * Lm;
    std::vector<std::vector<int>> blocks(nph, std::vector<int>(cnt, 0));
    size_t lds = 0;
    for (int i = 0; i < cnt; ++i) {
        dplrh_plan *p = B->plans[i];
        BOps<T> o;
        batch_ops<T>(p, prec, fact, &o);
        const double *P[PLR_MAXL], *Q[PLR_MAXL], *Sg[PLR_MAXL];
        for (int d = 0; d < p->L; ++d) {
            P[d] = (const double *)(o.V + p->uoff[d]);
            Q[d] = (const double *)(o.U + p->uoff[d]);
            Sg[d] = (const double *)(o.S + p->soff[d]);
        }
        MmTab *ti = &tabs[(size_t)i * tsl];
        mm_tab(p, 0, p->L, P, Q, Sg, nrhs, 0, &ti[0], S.tmax);
        mm_tab(p, 0, 0, P, Q, Sg, nrhs, 1, &ti[1], S.tmax);
        mm_tab(p, 0, p->L, Q, P, Sg, nrhs, 0, &ti[2], S.tmax);
        size_t maxpart = (size_t)ti[0].part_off[ti[0].L];
        for (int d = 0; d < p->L; ++d) {
            const double *Vd = P[d], *Sd = Sg[d];
            const double *Ud = (const double *)((fact ? o.Ut : o.U) + p->uoff[d]);
            mm_tab(p, d, d + 1, &Vd, &Ud, &Sd, nrhs, 1, &ti[SL_LVL + d], S.tmax);
            mm_tab(p, d, d + 1, &Ud, &Vd, &Sd, nrhs, 1, &ti[SL_LVL + Lm + d], S.tmax);
            if ((size_t)ti[SL_LVL + d].part_off[1] > maxpart) maxpart = ti[SL_LVL + d].part_off[1];
        }
        int rc = plr_mm_part(p, maxpart);
        if (rc) return rc;
        BWs<T> &b = ws[i];
        memset(&b, 0, sizeof(b));
        b.D = o.D;
        b.Dinv = o.Dinv;
        b.Kinv = o.Kinv;
        b.S = o.S;
        memcpy(b.koff, p->f.koff, sizeof(b.koff));
        memcpy(b.soff, p->soff, sizeof(b.soff));
        memcpy(b.R, p->R, sizeof(b.R));
        b.L = p->L;
        b.N = p->N;
        b.w = (C *)p->w;
        b.wstride = p->w_per_col;
        b.Tr = (C *)p->sw.T;
        b.W = (C *)p->sw.W;
        b.tstride = p->sw.t_per_col;
        b.part = (C *)p->tmp;
        nrow[i] = p->N;
        const int eb = (p->N / 64) * ti[0].ncg;
        blocks[0][i] = ti[0].task_off[ti[0].L];
        blocks[1][i] = ti[0].fin_off[ti[0].L] * MM_FCH;
        blocks[2][i] = eb;
        blocks[3][i] = eb;
        blocks[4][i] = (int)(((size_t)p->N * nrhs + 255) / 256);
        for (int d = 0; d < p->L && fact; ++d) {
            if (p->R[d] == 0) continue;
            const MmTab &t = ti[SL_LVL + d];
            blocks[PH_LVL + 4 * d + 0][i] = t.task_off[1];
            blocks[PH_LVL + 4 * d + 1][i] = t.fin_off[1] * MM_FCH;
            blocks[PH_LVL + 4 * d + 2][i] = (1 << d) * ((nrhs + 15) / 16);
            blocks[PH_LVL + 4 * d + 3][i] = eb;
            const size_t n2p = (size_t)((2 * p->R[d] + 63) & ~63);
            if (n2p * 17 * sizeof(double) > lds) lds = n2p * 17 * sizeof(double);
        }
    }
    if (lds > 65536) {
        plr_set_error("rank too large for the batched capacitance kernel");
        return -3;
    }
    std::vector<int> hpre;
    S.pre_off.assign(nph, 0);
    S.pre_tot.assign(nph, 0);
    for (int ph = 0; ph < nph; ++ph) {
        S.pre_off[ph] = hpre.size();
        int acc = 0;
        for (int i = 0; i < cnt; ++i) {
            hpre.push_back(acc);
            acc += blocks[ph][i];
        }
        hpre.push_back(acc);
        S.pre_tot[ph] = acc;
    }
    batch_drop_graphs(B); // they hold this precision's table pointers
    state_free(S);
    PLR_HIP(hipMalloc(&S.dtab, tabs.size() * sizeof(MmTab)));
    PLR_HIP(hipMalloc(&S.dws, ws.size() * sizeof(BWs<T>)));
    PLR_HIP(hipMalloc(&S.dpre, hpre.size() * sizeof(int)));
    PLR_HIP(hipMalloc(&S.dnrow, nrow.size() * sizeof(int)));
    PLR_HIP(hipMemcpy(S.dtab, tabs.data(), tabs.size() * sizeof(MmTab), hipMemcpyHostToDevice));
    PLR_HIP(hipMemcpy(S.dws, ws.data(), ws.size() * sizeof(BWs<T>), hipMemcpyHostToDevice));
    PLR_HIP(hipMemcpy(S.dpre, hpre.data(), hpre.size() * sizeof(int), hipMemcpyHostToDevice));
    PLR_HIP(hipMemcpy(S.dnrow, nrow.data(), nrow.size() * sizeof(int), hipMemcpyHostToDevice));
    S.cap_lds = lds;
    S.built_nrhs = nrhs;
    S.built_fact = fact;
    for (int i = 0; i < cnt; ++i) S.gen[i] = B->plans[i]->ws_gen;
    return 0;
}

static int build_prec(dplrh_batch *B, int prec, int nrhs, int fact)
{
    return prec == 0 ? batch_build<double>(B, 0, nrhs, fact)
           : prec == 1 ? batch_build<float>(B, 1, nrhs, fact)
                       : batch_build<uint16_t>(B, 2, nrhs, fact);
}

static bool state_stale(const dplrh_batch *B, int prec, int nrhs, int fact)
{
    const batch_state &S = B->st[prec];
    if (S.built_nrhs != nrhs || S.built_fact < fact) return true;
    for (int i = 0; i < B->count; ++i)
        if (S.gen[i] != B->plans[i]->ws_gen) return true;
    return false;
}

// the state is current for nrhs (and the solve, with fact); a rebuild can
// reallocate plan workspaces, so repeat until every state in `need` is current
static int batch_ready(dplrh_batch *B, const int *need, int nneed, int nrhs)
{
    for (int round = 0; round < 4; ++round) {
        bool any = false;
        for (int j = 0; j < nneed; j += 2)
            if (state_stale(B, need[j], nrhs, need[j + 1])) {
                any = true;
                const int fact = need[j + 1] || B->st[need[j]].built_fact;
                int rc = build_prec(B, need[j], nrhs, fact);
                if (rc) return rc;
            }
        if (!any) return 0;
    }
    plr_set_error("batch tables did not settle");
    return -10;
}

template <typename T>
static int launch_op(dplrh_batch *B, int prec, int op, int trans, int nrhs, const BIo *dio,
                     typename BPrec<T>::C alpha, typename BPrec<T>::C beta, hipStream_t st)
{
    typedef typename BPrec<T>::C C;
    const batch_state &S = B->st[prec];
    const int cnt = B->count, tsl = B->tsl, Lm = B->Lmax;
    const MmTab *tab = S.dtab;
    const BWs<T> *ws = (const BWs<T> *)S.dws;
    auto pre = [&](int ph) { return S.dpre + S.pre_off[ph]; };
    auto grid = [&](int ph) { return dim3((unsigned)S.pre_tot[ph]); };
    auto kexp = nrhs <= 16 ? kb_expand<T, 1> : kb_expand<T, 4>;
    if (op == 0) {
        const int slot = trans ? 2 : 0;
        if (S.pre_tot[0] > 0)
            hipLaunchKernelGGL(kb_reduce<T>, grid(0), dim3(64), 0, st, tab, tsl, slot, pre(0), cnt,
                               dio, ws, 0, alpha);
        if (S.pre_tot[1] > 0)
            hipLaunchKernelGGL(kb_finish<T>, grid(1), dim3(256), 0, st, tab, tsl, slot, pre(1), cnt,
                               ws, 0, alpha);
        hipLaunchKernelGGL(kexp, grid(2), dim3(256), 0, st, tab, tsl, slot, pre(2), cnt, dio,
                           ws, 0, trans, alpha, beta);
        PLR_HIP(hipGetLastError());
        return 0;
    }
    auto level = [&](int d) {
        const int ph = PH_LVL + 4 * d, slot = SL_LVL + (trans ? Lm : 0) + d;
        if (S.pre_tot[ph] > 0)
            hipLaunchKernelGGL(kb_reduce<T>, grid(ph), dim3(64), 0, st, tab, tsl, slot, pre(ph),
                               cnt, dio, ws, 1, (C)1);
        if (S.pre_tot[ph + 1] > 0)
            hipLaunchKernelGGL(kb_finish<T>, grid(ph + 1), dim3(256), 0, st, tab, tsl, slot,
                               pre(ph + 1), cnt, ws, 1, (C)1);
        if (S.pre_tot[ph + 2] > 0)
            hipLaunchKernelGGL(kb_cap<T>, grid(ph + 2), dim3(256), S.cap_lds, st, d, pre(ph + 2),
                               cnt, ws, nrhs, trans);
        if (S.pre_tot[ph + 3] > 0)
            hipLaunchKernelGGL(kexp, grid(ph + 3), dim3(256), 0, st, tab, tsl, slot,
                               pre(ph + 3), cnt, dio, ws, 2, 0, (C)1, (C)1);
    };
    if (!trans) {
        hipLaunchKernelGGL(kexp, grid(3), dim3(256), 0, st, tab, tsl, 1, pre(3), cnt, dio,
                           ws, 1, 0, (C)1, (C)0);
        for (int d = Lm - 1; d >= 0; --d) level(d);
    } else {
        hipLaunchKernelGGL(kb_axpby<C>, grid(4), dim3(256), 0, st, pre(4), cnt, dio, S.dnrow, nrhs,
                           0, 2, (C)1, (C)0);
        for (int d = 0; d < Lm; ++d) level(d);
        hipLaunchKernelGGL(kexp, grid(3), dim3(256), 0, st, tab, tsl, 1, pre(3), cnt, dio,
                           ws, 3, 1, (C)1, (C)0);
    }
    PLR_HIP(hipGetLastError());
    return 0;
}

static int launch_prec(dplrh_batch *B, int prec, int op, int trans, int nrhs, const BIo *dio,
                       double alpha, double beta, hipStream_t st)
{
    if (prec == 0) return launch_op<double>(B, 0, op, trans, nrhs, dio, alpha, beta, st);
    if (prec == 1)
        return launch_op<float>(B, 1, op, trans, nrhs, dio, (float)alpha, (float)beta, st);
    return launch_op<uint16_t>(B, 2, op, trans, nrhs, dio, (float)alpha, (float)beta, st);
}

// the refined solve; dio holds five vector sets of count entries:
// 0 {X: B, Y: X} first solve, 1 {X: B, Y: r} r = B, 2 {X: X, Y: r} r -= A X,
// 3 {X: r, Y: dx, Z: dx} correction, 4 {X: dx, Y: X} X += dx
static int launch_refined(dplrh_batch *B, int nrhs, int iters, const BIo *dio, hipStream_t st)
{
    const int cnt = B->count;
    const batch_state &S = B->st[1];
    int rc = launch_prec(B, 2, 1, 0, nrhs, dio, 1.0, 0.0, st);
    for (int k = 1; k < iters && !rc; ++k) {
        hipLaunchKernelGGL(kb_axpby<float>, dim3((unsigned)S.pre_tot[4]), dim3(256), 0, st,
                           S.dpre + S.pre_off[4], cnt, dio + cnt, S.dnrow, nrhs, 0, 1, 1.f, 0.f);
        if ((rc = launch_prec(B, 1, 0, 0, nrhs, dio + 2 * cnt, -1.0, 1.0, st))) break;
        if ((rc = launch_prec(B, 2, 1, 0, nrhs, dio + 3 * cnt, 1.0, 0.0, st))) break;
        hipLaunchKernelGGL(kb_axpby<float>, dim3((unsigned)S.pre_tot[4]), dim3(256), 0, st,
                           S.dpre + S.pre_off[4], cnt, dio + 4 * cnt, S.dnrow, nrhs, 0, 1, 1.f, 1.f);
        PLR_HIP(hipGetLastError());
    }
    return rc;
}

static int batch_scratch(dplrh_batch *B, int nrhs)
{
    if (B->scratch && B->scratch_nrhs >= nrhs) return 0;
    batch_drop_graphs(B);
    if (B->scratch) (void)hipFree(B->scratch);
    B->scratch = nullptr;
    B->zoff.assign(B->count, 0);
    size_t off = 0;
    for (int i = 0; i < B->count; ++i) {
        B->zoff[i] = off;
        off += ((size_t)B->plans[i]->N * nrhs * 2 * sizeof(double) + 255) & ~(size_t)255;
    }
    PLR_HIP(hipMalloc(&B->scratch, off ? off : 1));
    B->scratch_nrhs = nrhs;
    return 0;
}

static int check_io(const dplrh_batch *B, int prec, int nrhs, const void *const *X,
                    const int *ldx, void *const *Y, const int *ldy)
{
    for (int i = 0; i < B->count; ++i) {
        const int N = B->plans[i]->N;
        bool ok = ldx[i] >= N && ldy[i] >= N;
        if (prec == 0)
            ok = ok && !(nrhs > 1 && (ldx[i] % 2 || ldy[i] % 2));
        else // FP32 vectors are read 16 bytes at a time
            ok = ok && ldx[i] % 4 == 0 && ldy[i] % 4 == 0 && ((uintptr_t)X[i] & 15) == 0 &&
                 ((uintptr_t)Y[i] & 15) == 0;
        if (!ok) {
            plr_set_error(prec == 0 ? "batch member %d: leading dimensions must be >= N and even"
                                    : "batch member %d: leading dimensions must be >= N and "
                                      "multiples of 4, the blocks 16-byte aligned",
                          i);
            return -1;
        }
    }
    return 0;
}

// op 0: Y := alpha op(A) X + beta Y; op 1: Y := op(A)^{-1} X; op 2: the
// refined solve (BF16 solves, FP32 residuals, `iters` solves)
static int batch_run(dplrh_batch *B, int prec, int op, int trans, int nrhs, const void *const *X,
                     const int *ldx, void *const *Y, const int *ldy, double alpha, double beta,
                     int iters, hipStream_t st)
{
    if (nrhs <= 0) return 0;
    const int cnt = B->count;
    int rc = check_io(B, prec, nrhs, X, ldx, Y, ldy);
    if (rc) return rc;
    if (op == 2 && (iters < 1 || prec != 2)) {
        plr_set_error("refined batch solve: BF16 factors and at least one solve (%d)", iters);
        return -1;
    }
    int need[4] = {prec, op != 0, 0, 0}, nneed = 2;
    if (op == 2) {
        need[2] = 1; // FP32 apply for the residuals
        nneed = 4;
    }
    if ((rc = batch_ready(B, need, nneed, nrhs))) return rc;
    const bool zs = (op == 1 && trans) || op == 2;
    if (zs && (rc = batch_scratch(B, nrhs))) return rc;
    // the call's vector sets (five for the refined solve, see launch_refined)
    std::vector<BIo> hio;
    for (int i = 0; i < cnt; ++i) {
        const int N = B->plans[i]->N;
        if (op == 1 && trans)
            hio.push_back(BIo{X[i], Y[i], B->scratch + B->zoff[i], ldx[i], ldy[i], N});
        else
            hio.push_back(BIo{X[i], Y[i], Y[i], ldx[i], ldy[i], ldy[i]});
    }
    if (op == 2)
        for (int set = 1; set < 5; ++set)
            for (int i = 0; i < cnt; ++i) {
                const int N = B->plans[i]->N;
                char *r = B->scratch + B->zoff[i];
                char *dx = r + (size_t)N * nrhs * sizeof(float);
                hio.push_back(set == 1   ? BIo{X[i], r, r, ldx[i], N, N}
                              : set == 2 ? BIo{Y[i], r, r, ldy[i], N, N}
                              : set == 3 ? BIo{r, dx, dx, N, N, N}
                                         : BIo{dx, Y[i], Y[i], N, ldy[i], ldy[i]});
            }
    const int it = op == 2 ? iters : 0;
    const double al = op == 0 ? alpha : 0.0, be = op == 0 ? beta : 0.0;
    batch_graph *g = nullptr;
    for (auto &e : B->graphs)
        if (e.prec == prec && e.op == op && e.trans == trans && e.nrhs == nrhs && e.iters == it &&
            e.alpha == al && e.beta == be)
            g = &e;
    if (!g) {
        B->graphs.push_back(
            batch_graph{prec, op, trans, nrhs, it, al, be, {}, nullptr, nullptr});
        g = &B->graphs.back();
        PLR_HIP(hipMalloc(&g->dio, hio.size() * sizeof(BIo)));
    }
    if (g->hio.size() != hio.size() || memcmp(g->hio.data(), hio.data(), hio.size() * sizeof(BIo))) {
        PLR_HIP(hipMemcpyAsync(g->dio, hio.data(), hio.size() * sizeof(BIo), hipMemcpyHostToDevice,
                               st));
        PLR_HIP(hipStreamSynchronize(st));
        g->hio = hio;
    }
    if (!g->exec) {
        // capture on the private stream, ordered after the user's stream
        hipEvent_t ev;
        PLR_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        PLR_HIP(hipEventRecord(ev, st));
        PLR_HIP(hipStreamWaitEvent(B->cap, ev, 0));
        PLR_HIP(hipStreamSynchronize(B->cap));
        (void)hipEventDestroy(ev);
        hipGraph_t graph;
        PLR_HIP(hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal));
        rc = op == 2 ? launch_refined(B, nrhs, iters, g->dio, B->cap)
                     : launch_prec(B, prec, op, trans, nrhs, g->dio, alpha, beta, B->cap);
        hipError_t e2 = hipStreamEndCapture(B->cap, &graph);
        if (rc) return rc;
        PLR_HIP(e2);
        hipError_t e3 = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        PLR_HIP(e3);
    }
    PLR_HIP(hipGraphLaunch(g->exec, st));
    return 0;
}

extern "C" int dplrh_batch_gemm(dplrh_batch *B, int nrhs, const double *const *X, const int *ldx,
                                double *const *Y, const int *ldy, void *stream)
{
    return batch_run(B, 0, 0, 0, nrhs, (const void *const *)X, ldx, (void *const *)Y, ldy, 1.0,
                     0.0, 0, (hipStream_t)stream);
}

extern "C" int dplrh_batch_solve(dplrh_batch *B, int nrhs, const double *const *X, const int *ldx,
                                 double *const *Y, const int *ldy, void *stream)
{
    return batch_run(B, 0, 1, 0, nrhs, (const void *const *)X, ldx, (void *const *)Y, ldy, 0.0,
                     0.0, 0, (hipStream_t)stream);
}

static int check_prec(int prec)
{
    if (prec < 0 || prec > 2) {
        plr_set_error("batch precision %d (0: FP64, 1: FP32, 2: BF16 factors)", prec);
        return -1;
    }
    return 0;
}

// Any precision and direction: op 0 Y := alpha op(A) X + beta Y, op 1
// Y := op(A)^{-1} X.  Vectors are FP64 for precision 0, FP32 otherwise.
extern "C" int dplrh_batch_op(dplrh_batch *B, int prec, int op, int trans, int nrhs,
                              const void *const *X, const int *ldx, void *const *Y,
                              const int *ldy, double alpha, double beta, void *stream)
{
    if (check_prec(prec)) return -1;
    if (op != 0 && op != 1) {
        plr_set_error("batch operation %d (0: apply, 1: solve)", op);
        return -1;
    }
    return batch_run(B, prec, op, trans ? 1 : 0, nrhs, X, ldx, Y, ldy, alpha, beta, 0,
                     (hipStream_t)stream);
}

extern "C" int dplrh_batch_refined(dplrh_batch *B, int nrhs, const float *const *X, const int *ldx,
                                   float *const *Y, const int *ldy, int iters, void *stream)
{
    return batch_run(B, 2, 2, 0, nrhs, (const void *const *)X, ldx, (void *const *)Y, ldy, 0.0, 0.0,
                     iters, (hipStream_t)stream);
}

// Single-plan single-precision operations through a batch of the plan alone.
static int self_batch(dplrh_plan *p, dplrh_batch **B)
{
    if (!p->self_batch) {
        int rc = dplrh_batch_create(1, &p, &p->self_batch);
        if (rc) return rc;
    }
    *B = p->self_batch;
    return 0;
}

extern "C" int dplrh_sp_gemm(dplrh_plan *p, int prec, int trans, int n, float alpha, float beta,
                             const float *X, int ldx, float *Y, int ldy, void *stream)
{
    dplrh_batch *B;
    int rc = self_batch(p, &B);
    if (rc) return rc;
    const void *xs[1] = {X};
    void *ys[1] = {Y};
    return dplrh_batch_op(B, prec, 0, trans, n, xs, &ldx, ys, &ldy, alpha, beta, stream);
}

extern "C" int dplrh_sp_solve(dplrh_plan *p, int prec, int trans, int n, const float *X, int ldx,
                              float *Y, int ldy, void *stream)
{
    dplrh_batch *B;
    int rc = self_batch(p, &B);
    if (rc) return rc;
    const void *xs[1] = {X};
    void *ys[1] = {Y};
    return dplrh_batch_op(B, prec, 1, trans, n, xs, &ldx, ys, &ldy, 0.0, 0.0, stream);
}

// Autotuner: for each candidate tile (rows per reduction task) rebuild the
// tables, run the batch's apply and solve eagerly (one warm-up, then `reps`
// timed with events) and keep the fastest.  ms[i] receives the time of
// candidate i (apply + solve, per repetition) when not null.  The tables
// and graphs are rebuilt for the chosen tile on the next call.
static const int kTuneTiles[] = {128, 256, 512, 1024, 2048, 4096};

extern "C" int dplrh_batch_tune_prec(dplrh_batch *B, int prec, int nrhs, const void *const *X,
                                     const int *ldx, void *const *Y, const int *ldy, int reps,
                                     int *tmax_out, double *ms, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (check_prec(prec)) return -1;
    if (nrhs <= 0 || reps <= 0) {
        plr_set_error("tuning needs nrhs > 0 and reps > 0");
        return -1;
    }
    int rc = check_io(B, prec, nrhs, X, ldx, Y, ldy);
    if (rc) return rc;
    std::vector<BIo> hio(B->count);
    for (int i = 0; i < B->count; ++i) hio[i] = BIo{X[i], Y[i], Y[i], ldx[i], ldy[i], ldy[i]};
    BIo *dio = nullptr;
    PLR_HIP(hipMalloc(&dio, hio.size() * sizeof(BIo)));
    PLR_HIP(hipMemcpy(dio, hio.data(), hio.size() * sizeof(BIo), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    PLR_HIP(hipEventCreate(&e0));
    PLR_HIP(hipEventCreate(&e1));
    batch_state &S = B->st[prec];
    int best = S.tmax;
    float best_ms = 0.f;
    const int ncand = (int)(sizeof(kTuneTiles) / sizeof(kTuneTiles[0]));
    const int need[2] = {prec, 1};
    for (int c = 0; c < ncand && !rc; ++c) {
        S.tmax = kTuneTiles[c];
        S.built_nrhs = 0;
        if ((rc = batch_ready(B, need, 2, nrhs))) break;
        float t = 0.f;
        for (int r = 0; r <= reps && !rc; ++r) {
            if (r == 1) PLR_HIP(hipEventRecord(e0, st));
            if (!(rc = launch_prec(B, prec, 0, 0, nrhs, dio, 1.0, 0.0, st)))
                rc = launch_prec(B, prec, 1, 0, nrhs, dio, 0.0, 0.0, st);
        }
        if (rc) break;
        PLR_HIP(hipEventRecord(e1, st));
        PLR_HIP(hipEventSynchronize(e1));
        PLR_HIP(hipEventElapsedTime(&t, e0, e1));
        t /= reps;
        if (ms) ms[c] = t;
        if (c == 0 || t < best_ms) {
            best_ms = t;
            best = kTuneTiles[c];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(dio);
    S.tmax = best;
    batch_drop_graphs(B); // rebuild (and recapture) with the chosen tile
    state_free(S);
    if (tmax_out) *tmax_out = best;
    return rc;
}

extern "C" int dplrh_batch_tune(dplrh_batch *B, int nrhs, const double *const *X, const int *ldx,
                                double *const *Y, const int *ldy, int reps, int *tmax_out,
                                double *ms, void *stream)
{
    return dplrh_batch_tune_prec(B, 0, nrhs, (const void *const *)X, ldx, (void *const *)Y, ldy,
                                 reps, tmax_out, ms, stream);
}

// every precision's tile
extern "C" int dplrh_batch_set_tile(dplrh_batch *B, int tmax)
{
    bool ok = false;
    for (int t : kTuneTiles) ok = ok || t == tmax;
    if (!ok) {
        plr_set_error("tile %d is not one of 128, 256, 512, 1024, 2048, 4096", tmax);
        return -1;
    }
    batch_drop_graphs(B);
    for (auto &s : B->st)
        if (s.tmax != tmax) {
            s.tmax = tmax;
            state_free(s);
        }
    return 0;
}
