// Batched FP64 apply and factored solve over many independent plans.
//
// Every phase is one launch that covers all matrices: the blocks of matrix i
// are [pre[i], pre[i+1]) of the grid (a prefix table per phase, found by a
// binary search in the block), and each block runs the same device bodies
// as the single-plan kernels (plr_mm.h, plr_cap.h) on matrix i's table.
//   apply:  reduce (all levels) -> finish -> expand (leaves + all levels)
//   solve:  leaf expand (X = Dinv B), then for depth d = Lmax-1 .. 0 the
//           matrices deeper than d run reduce -> finish -> capacitance ->
//           expand at level d (a matrix's own levels keep their order).
// The whole sequence is captured once into a HIP graph per (operation,
// RHS count) and replayed; the tables are rebuilt when a plan's device
// memory changes (its workspace generation).
//
// Tile shape: the rows per reduction task (tmax) trades parallelism (more,
// shorter tasks) against partial sums (more work for the finishing pass).
// dplrh_batch_tune times the apply + solve of the caller's batch for each
// candidate and keeps the fastest; without it the batch uses MM_TMAX.
#include <cstring>
#include <vector>

#include "plr_cap.h"
#include "plr_common.h"
#include "plr_mm.h"

// per-matrix device view of the factors and workspaces
struct BWs
{
    const double *D, *Dinv, *Kinv, *S;
    size_t koff[PLR_MAXL + 1], soff[PLR_MAXL + 1];
    int R[PLR_MAXL];
    int L, N;
    double *w; // apply coefficients, stride wstride per column
    size_t wstride;
    double *T, *W; // solve reduction / coefficients, stride tstride per column
    size_t tstride;
    double *part; // partial blocks of long segments
};

// per-matrix right-hand sides of one call
struct BIo
{
    const double *X;
    double *Y;
    int ldx, ldy;
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

// mode 0: apply (X -> w, scaled), mode 1: solve level (Y in place -> T, raw)
__global__ __launch_bounds__(64) void kb_reduce(const MmTab *__restrict__ tab, int tsl, int slot,
                                                const int *__restrict__ pre, int cnt,
                                                const BIo *__restrict__ io,
                                                const BWs *__restrict__ ws, int mode, double alpha)
{
    const int mi = batch_find(pre, cnt, blockIdx.x);
    const MmTab &a = tab[(size_t)mi * tsl + slot];
    const BWs &b = ws[mi];
    const BIo &o = io[mi];
    const int task = blockIdx.x - pre[mi];
    if (mode == 0)
        mm_reduce_body<1>(a, task, o.X, o.ldx, alpha, 0, 0, b.part, b.w, b.wstride);
    else
        mm_reduce_body<1>(a, task, o.Y, o.ldy, 1.0, 0, 1, b.part, b.T, b.tstride);
}

__global__ __launch_bounds__(256) void kb_finish(const MmTab *__restrict__ tab, int tsl, int slot,
                                                 const int *__restrict__ pre, int cnt,
                                                 const BWs *__restrict__ ws, int mode, double alpha)
{
    const int mi = batch_find(pre, cnt, blockIdx.x);
    const MmTab &a = tab[(size_t)mi * tsl + slot];
    const BWs &b = ws[mi];
    const int bx = blockIdx.x - pre[mi];
    if (mode == 0)
        mm_finish_body(a, bx, alpha, 0, 0, b.part, b.w, b.wstride);
    else
        mm_finish_body(a, bx, 1.0, 0, 1, b.part, b.T, b.tstride);
}

// mode 0: apply Y = alpha*A X + beta*Y; mode 1: solve leaves Y = Dinv X;
// mode 2: solve level Y += Q W
__global__ __launch_bounds__(256) void kb_expand(const MmTab *__restrict__ tab, int tsl, int slot,
                                                 const int *__restrict__ pre, int cnt,
                                                 const BIo *__restrict__ io,
                                                 const BWs *__restrict__ ws, int mode, double alpha,
                                                 double beta)
{
    const int mi = batch_find(pre, cnt, blockIdx.x);
    const MmTab &a = tab[(size_t)mi * tsl + slot];
    const BWs &b = ws[mi];
    const BIo &o = io[mi];
    const int loc = blockIdx.x - pre[mi];
    const int rb = loc / a.ncg, cg = loc - rb * a.ncg;
    if (mode == 0)
        mm_expand_body(a, rb, cg, b.D, o.X, o.ldx, alpha, beta, 0, b.w, b.wstride, o.Y, o.ldy);
    else if (mode == 1)
        mm_expand_body(a, rb, cg, b.Dinv, o.X, o.ldx, 1.0, 0.0, 0, nullptr, 0, o.Y, o.ldy);
    else
        mm_expand_body(a, rb, cg, nullptr, nullptr, 0, 1.0, 1.0, 0, b.W, b.tstride, o.Y, o.ldy);
}

// capacitance step of level d, 16 columns per work-group
__global__ __launch_bounds__(256) void kb_cap(int d, const int *__restrict__ pre, int cnt,
                                              const BWs *__restrict__ ws, int ncols)
{
    const int mi = batch_find(pre, cnt, blockIdx.x);
    const BWs &b = ws[mi];
    const int loc = blockIdx.x - pre[mi], nodes = 1 << d;
    const int node = loc % nodes, cblk = loc / nodes;
    cap_body<double, double, false, 1>(node, cblk, b.Kinv + b.koff[d], b.R[d], b.S + b.soff[d],
                                       b.T, b.tstride, b.W, b.tstride, 0, ncols, nullptr, 0);
}

int plr_mm_part(dplrh_plan *p, size_t blocks);

struct batch_graph
{
    int kind, nrhs;
    std::vector<BIo> hio; // the call's pointers, compared on every call
    BIo *dio;
    hipGraphExec_t exec;
};

struct dplrh_batch
{
    int count, Lmax;
    std::vector<dplrh_plan *> plans;
    std::vector<unsigned> gen; // workspace generation of each plan when the tables were built
    int built_nrhs, built_kind;
    // device tables
    MmTab *dtab;
    int tsl; // tables per matrix: apply, leaves, one per level
    BWs *dws;
    int *dpre;
    std::vector<size_t> pre_off; // offset of each phase's prefix in dpre
    std::vector<int> pre_tot;    // blocks of each phase
    size_t cap_lds;
    hipStream_t cap;
    std::vector<batch_graph> graphs;
    int tmax; // rows per reduction task (tile shape)
};

static void batch_free_tables(dplrh_batch *B)
{
    if (B->dtab) (void)hipFree(B->dtab);
    if (B->dws) (void)hipFree(B->dws);
    if (B->dpre) (void)hipFree(B->dpre);
    B->dtab = nullptr;
    B->dws = nullptr;
    B->dpre = nullptr;
    for (auto &g : B->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.dio) (void)hipFree(g.dio);
    }
    B->graphs.clear();
    B->built_nrhs = 0;
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
    B->gen.assign(count, 0);
    B->Lmax = 0;
    for (auto *p : B->plans) B->Lmax = p->L > B->Lmax ? p->L : B->Lmax;
    B->tsl = B->Lmax + 2;
    B->tmax = MM_TMAX;
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
    batch_free_tables(B);
    if (B->cap) (void)hipStreamDestroy(B->cap);
    delete B;
}

// phases: 0 apply reduce, 1 apply finish, 2 apply expand, 3 solve leaves,
// then per depth d four phases 4 + 4d + {0 reduce, 1 finish, 2 cap, 3 expand}
static int batch_build(dplrh_batch *B, int kind, int nrhs)
{
    const int cnt = B->count, Lm = B->Lmax, tsl = B->tsl;
    // workspaces first: they may reallocate plan memory
    for (auto *p : B->plans) {
        if (kind == 1 && !p->f.done) {
            plr_set_error("batched solve before a successful factorization of every plan");
            return -4;
        }
        int rc = plr_workspace(p, nrhs);
        if (!rc && kind == 1) rc = plr_solve_reserve(p, nrhs);
        if (rc) return rc;
    }
    std::vector<MmTab> tabs((size_t)cnt * tsl);
    std::vector<BWs> ws(cnt);
    const int nph = 4 + 4 * Lm;
    std::vector<std::vector<int>> blocks(nph, std::vector<int>(cnt, 0));
    size_t lds = 0;
    for (int i = 0; i < cnt; ++i) {
        dplrh_plan *p = B->plans[i];
        const double *P[PLR_MAXL], *Q[PLR_MAXL], *S[PLR_MAXL];
        for (int d = 0; d < p->L; ++d) {
            P[d] = p->V + p->uoff[d];
            Q[d] = p->U + p->uoff[d];
            S[d] = p->S + p->soff[d];
        }
        MmTab *ti = &tabs[(size_t)i * tsl];
        mm_tab(p, 0, p->L, P, Q, S, nrhs, 0, &ti[0], B->tmax);
        mm_tab(p, 0, 0, P, Q, S, nrhs, 1, &ti[1], B->tmax);
        size_t maxpart = (size_t)ti[0].part_off[ti[0].L];
        for (int d = 0; d < p->L; ++d) {
            const double *Pd = p->V + p->uoff[d];
            const double *Qd = p->f.done ? p->f.Ut + p->uoff[d] : p->U + p->uoff[d];
            const double *Sd = p->S + p->soff[d];
            mm_tab(p, d, d + 1, &Pd, &Qd, &Sd, nrhs, 1, &ti[2 + d], B->tmax);
            if ((size_t)ti[2 + d].part_off[1] > maxpart) maxpart = ti[2 + d].part_off[1];
        }
        int rc = plr_mm_part(p, maxpart);
        if (rc) return rc;
        BWs &b = ws[i];
        memset(&b, 0, sizeof(b));
        b.D = p->D;
        b.Dinv = p->f.Dinv;
        b.Kinv = p->f.Kinv;
        b.S = p->S;
        memcpy(b.koff, p->f.koff, sizeof(b.koff));
        memcpy(b.soff, p->soff, sizeof(b.soff));
        memcpy(b.R, p->R, sizeof(b.R));
        b.L = p->L;
        b.N = p->N;
        b.w = p->w;
        b.wstride = p->w_per_col;
        b.T = p->sw.T;
        b.W = p->sw.W;
        b.tstride = p->sw.t_per_col;
        b.part = p->tmp;
        const int eb = (p->N / 64) * ti[0].ncg;
        blocks[0][i] = ti[0].task_off[ti[0].L];
        blocks[1][i] = ti[0].fin_off[ti[0].L] * MM_FCH;
        blocks[2][i] = eb;
        blocks[3][i] = eb;
        for (int d = 0; d < p->L; ++d) {
            if (p->R[d] == 0) continue;
            const MmTab &t = ti[2 + d];
            blocks[4 + 4 * d + 0][i] = t.task_off[1];
            blocks[4 + 4 * d + 1][i] = t.fin_off[1] * MM_FCH;
            blocks[4 + 4 * d + 2][i] = (1 << d) * ((nrhs + 15) / 16);
            blocks[4 + 4 * d + 3][i] = eb;
            const size_t n2p = (size_t)((2 * p->R[d] + 63) & ~63);
            if (n2p * 17 * sizeof(double) > lds) lds = n2p * 17 * sizeof(double);
        }
        B->gen[i] = p->ws_gen;
    }
    if (lds > 65536) {
        plr_set_error("rank too large for the batched capacitance kernel");
        return -3;
    }
    // prefix tables
    std::vector<int> hpre;
    B->pre_off.assign(nph, 0);
    B->pre_tot.assign(nph, 0);
    for (int ph = 0; ph < nph; ++ph) {
        B->pre_off[ph] = hpre.size();
        int acc = 0;
        for (int i = 0; i < cnt; ++i) {
            hpre.push_back(acc);
            acc += blocks[ph][i];
        }
        hpre.push_back(acc);
        B->pre_tot[ph] = acc;
    }
    batch_free_tables(B);
    PLR_HIP(hipMalloc(&B->dtab, tabs.size() * sizeof(MmTab)));
    PLR_HIP(hipMalloc(&B->dws, ws.size() * sizeof(BWs)));
    PLR_HIP(hipMalloc(&B->dpre, hpre.size() * sizeof(int)));
    PLR_HIP(hipMemcpy(B->dtab, tabs.data(), tabs.size() * sizeof(MmTab), hipMemcpyHostToDevice));
    PLR_HIP(hipMemcpy(B->dws, ws.data(), ws.size() * sizeof(BWs), hipMemcpyHostToDevice));
    PLR_HIP(hipMemcpy(B->dpre, hpre.data(), hpre.size() * sizeof(int), hipMemcpyHostToDevice));
    B->cap_lds = lds;
    B->built_nrhs = nrhs;
    B->built_kind = kind;
    return 0;
}

static int batch_launch(dplrh_batch *B, int kind, int nrhs, const BIo *dio, hipStream_t st)
{
    const int cnt = B->count, tsl = B->tsl;
    auto pre = [&](int ph) { return B->dpre + B->pre_off[ph]; };
    if (kind == 0) {
        if (B->pre_tot[0] > 0)
            hipLaunchKernelGGL(kb_reduce, dim3(B->pre_tot[0]), dim3(64), 0, st, B->dtab, tsl, 0,
                               pre(0), cnt, dio, B->dws, 0, 1.0);
        if (B->pre_tot[1] > 0)
            hipLaunchKernelGGL(kb_finish, dim3(B->pre_tot[1]), dim3(256), 0, st, B->dtab, tsl, 0,
                               pre(1), cnt, B->dws, 0, 1.0);
        hipLaunchKernelGGL(kb_expand, dim3(B->pre_tot[2]), dim3(256), 0, st, B->dtab, tsl, 0,
                           pre(2), cnt, dio, B->dws, 0, 1.0, 0.0);
        PLR_HIP(hipGetLastError());
        return 0;
    }
    hipLaunchKernelGGL(kb_expand, dim3(B->pre_tot[3]), dim3(256), 0, st, B->dtab, tsl, 1, pre(3),
                       cnt, dio, B->dws, 1, 1.0, 0.0);
    for (int d = B->Lmax - 1; d >= 0; --d) {
        const int ph = 4 + 4 * d;
        if (B->pre_tot[ph] > 0)
            hipLaunchKernelGGL(kb_reduce, dim3(B->pre_tot[ph]), dim3(64), 0, st, B->dtab, tsl,
                               2 + d, pre(ph), cnt, dio, B->dws, 1, 1.0);
        if (B->pre_tot[ph + 1] > 0)
            hipLaunchKernelGGL(kb_finish, dim3(B->pre_tot[ph + 1]), dim3(256), 0, st, B->dtab, tsl,
                               2 + d, pre(ph + 1), cnt, B->dws, 1, 1.0);
        if (B->pre_tot[ph + 2] > 0)
            hipLaunchKernelGGL(kb_cap, dim3(B->pre_tot[ph + 2]), dim3(256), B->cap_lds, st, d,
                               pre(ph + 2), cnt, B->dws, nrhs);
        if (B->pre_tot[ph + 3] > 0)
            hipLaunchKernelGGL(kb_expand, dim3(B->pre_tot[ph + 3]), dim3(256), 0, st, B->dtab, tsl,
                               2 + d, pre(ph + 3), cnt, dio, B->dws, 2, 1.0, 1.0);
    }
    PLR_HIP(hipGetLastError());
    return 0;
}

// kind 0: Y_i := A_i X_i; kind 1: Y_i := inv(A_i) X_i (factored plans)
static int batch_run(dplrh_batch *B, int kind, int nrhs, const double *const *X, const int *ldx,
                     double *const *Y, const int *ldy, hipStream_t st)
{
    if (nrhs <= 0) return 0;
    for (int i = 0; i < B->count; ++i)
        if ((nrhs > 1 && (ldx[i] % 2 || ldy[i] % 2)) || ldx[i] < B->plans[i]->N ||
            ldy[i] < B->plans[i]->N) {
            plr_set_error("batch member %d: leading dimensions must be >= N and even", i);
            return -1;
        }
    bool stale = B->built_nrhs != nrhs || B->built_kind < kind;
    for (int i = 0; i < B->count && !stale; ++i) stale = B->gen[i] != B->plans[i]->ws_gen;
    if (stale) {
        int rc = batch_build(B, kind > B->built_kind ? kind : B->built_kind, nrhs);
        if (rc) return rc;
        // a rebuilt table may follow a reallocation: make sure no plan changed again
        for (int i = 0; i < B->count; ++i) B->gen[i] = B->plans[i]->ws_gen;
    }
    std::vector<BIo> hio(B->count);
    for (int i = 0; i < B->count; ++i) hio[i] = BIo{X[i], Y[i], ldx[i], ldy[i]};
    batch_graph *g = nullptr;
    for (auto &e : B->graphs)
        if (e.kind == kind && e.nrhs == nrhs) g = &e;
    if (!g) {
        B->graphs.push_back(batch_graph{kind, nrhs, {}, nullptr, nullptr});
        g = &B->graphs.back();
        PLR_HIP(hipMalloc(&g->dio, (size_t)B->count * sizeof(BIo)));
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
        int rc = batch_launch(B, kind, nrhs, g->dio, B->cap);
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
    return batch_run(B, 0, nrhs, X, ldx, Y, ldy, (hipStream_t)stream);
}

extern "C" int dplrh_batch_solve(dplrh_batch *B, int nrhs, const double *const *X, const int *ldx,
                                 double *const *Y, const int *ldy, void *stream)
{
    return batch_run(B, 1, nrhs, X, ldx, Y, ldy, (hipStream_t)stream);
}

// Autotuner: for each candidate tile (rows per reduction task) rebuild the
// tables, run the batch's apply and solve eagerly (one warm-up, then `reps`
// timed with events) and keep the fastest.  ms[i] receives the time of
// candidate i (apply + solve, per repetition) when not null.  The tables
// and graphs are rebuilt for the chosen tile on the next call.
static const int kTuneTiles[] = {128, 256, 512, 1024};

extern "C" int dplrh_batch_tune(dplrh_batch *B, int nrhs, const double *const *X, const int *ldx,
                                double *const *Y, const int *ldy, int reps, int *tmax_out,
                                double *ms, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (nrhs <= 0 || reps <= 0) {
        plr_set_error("tuning needs nrhs > 0 and reps > 0");
        return -1;
    }
    for (int i = 0; i < B->count; ++i)
        if (!B->plans[i]->f.done) {
            plr_set_error("tuning runs the solve: factor every plan first");
            return -4;
        }
    std::vector<BIo> hio(B->count);
    for (int i = 0; i < B->count; ++i) hio[i] = BIo{X[i], Y[i], ldx[i], ldy[i]};
    BIo *dio = nullptr;
    PLR_HIP(hipMalloc(&dio, hio.size() * sizeof(BIo)));
    PLR_HIP(hipMemcpy(dio, hio.data(), hio.size() * sizeof(BIo), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    PLR_HIP(hipEventCreate(&e0));
    PLR_HIP(hipEventCreate(&e1));
    int best = B->tmax, rc = 0;
    float best_ms = 0.f;
    const int ncand = (int)(sizeof(kTuneTiles) / sizeof(kTuneTiles[0]));
    for (int c = 0; c < ncand && !rc; ++c) {
        B->tmax = kTuneTiles[c];
        if ((rc = batch_build(B, 1, nrhs))) break;
        for (int i = 0; i < B->count; ++i) B->gen[i] = B->plans[i]->ws_gen;
        float t = 0.f;
        for (int r = 0; r <= reps && !rc; ++r) {
            if (r == 1) PLR_HIP(hipEventRecord(e0, st));
            if (!(rc = batch_launch(B, 0, nrhs, dio, st))) rc = batch_launch(B, 1, nrhs, dio, st);
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
    B->tmax = best;
    B->built_nrhs = -1; // rebuild (and recapture) with the chosen tile
    batch_free_tables(B);
    if (tmax_out) *tmax_out = best;
    return rc;
}

extern "C" int dplrh_batch_set_tile(dplrh_batch *B, int tmax)
{
    bool ok = false;
    for (int t : kTuneTiles) ok = ok || t == tmax;
    if (!ok) {
        plr_set_error("tile %d is not one of 128, 256, 512, 1024", tmax);
        return -1;
    }
    if (B->tmax != tmax) {
        B->tmax = tmax;
        B->built_nrhs = -1;
        batch_free_tables(B);
    }
    return 0;
}
