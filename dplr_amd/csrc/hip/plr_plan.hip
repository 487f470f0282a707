// Device plan: upload of the level-wise FP64 arrays, kernel workspace, small
// utility kernels and the C entry points of the HIP layer.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "plr_common.h"

// one message per thread: concurrent calls on distinct plans do not share it
static thread_local char g_err[512];

void plr_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *dplrh_last_error(void) { return g_err; }

static constexpr int kMaxTile = 512; // rows per reduction tile (8 per lane)
static constexpr int kKG = 16;       // rank group of the reduction kernel

void plr_level_tab(const dplrh_plan *p, int trans, LevelTab *t)
{
    memset(t, 0, sizeof(*t));
    t->L = p->L;
    t->N = p->N;
    t->b = p->b;
    int task = 0, part = 0, cnt = 0;
    for (int d = 0; d < p->L; ++d) {
        const int seg = p->N >> (d + 1);
        const int nseg = 2 << d;
        const int R = p->R[d];
        const int tps = (seg + kMaxTile - 1) / kMaxTile;
        int T = (seg + tps - 1) / tps;
        T = (T + 63) & ~63;
        const int nkg = (R + kKG - 1) / kKG;
        t->R[d] = R;
        t->seglen[d] = seg;
        t->T[d] = T;
        t->tps[d] = tps;
        t->nkg[d] = nkg;
        t->task_off[d] = task;
        task += nkg * nseg * tps;
        t->part_off[d] = part;
        if (tps > 1) part += nseg * tps * R;
        t->cnt_off[d] = cnt;
        if (tps > 1) cnt += nseg * nkg;
        t->w_off[d] = (int)p->soff[d];
        const double *Ud = p->U + p->uoff[d], *Vd = p->V + p->uoff[d];
        t->P[d] = trans ? Ud : Vd;
        t->Q[d] = trans ? Vd : Ud;
        t->S[d] = p->S + p->soff[d];
    }
    t->task_off[p->L] = task;
    t->cnt_off[p->L] = cnt;
    t->w_off[p->L] = (int)p->soff[p->L];
}

void plr_level_tab1(const dplrh_plan *p, int d, const double *P, const double *Q,
                    const double *S, LevelTab *t)
{
    LevelTab all;
    plr_level_tab(p, 0, &all);
    memset(t, 0, sizeof(*t));
    t->L = 1;
    t->N = p->N;
    t->b = p->b;
    t->R[0] = all.R[d];
    t->seglen[0] = all.seglen[d];
    t->T[0] = all.T[d];
    t->tps[0] = all.tps[d];
    t->nkg[0] = all.nkg[d];
    t->task_off[0] = 0;
    t->task_off[1] = all.task_off[d + 1] - all.task_off[d];
    t->part_off[0] = 0;
    t->cnt_off[0] = 0;
    t->cnt_off[1] = all.cnt_off[d + 1] - all.cnt_off[d];
    t->w_off[0] = 0;
    t->w_off[1] = (2 << d) * all.R[d];
    t->P[0] = P;
    t->Q[0] = Q;
    t->S[0] = S;
}

void plr_level_tab_from(const dplrh_plan *p, int trans, int d0, LevelTab *t)
{
    LevelTab all;
    plr_level_tab(p, trans, &all);
    memset(t, 0, sizeof(*t));
    t->L = p->L - d0;
    t->N = p->N;
    t->b = p->b;
    // the workspace offsets stay those of the whole table (the workspace is
    // sized for it); only the task numbering starts again at zero
    for (int i = 0; i < t->L; ++i) {
        const int d = d0 + i;
        t->R[i] = all.R[d];
        t->seglen[i] = all.seglen[d];
        t->T[i] = all.T[d];
        t->tps[i] = all.tps[d];
        t->nkg[i] = all.nkg[d];
        t->task_off[i] = all.task_off[d] - all.task_off[d0];
        t->part_off[i] = all.part_off[d];
        t->cnt_off[i] = all.cnt_off[d];
        t->w_off[i] = all.w_off[d];
        t->P[i] = all.P[d];
        t->Q[i] = all.Q[d];
        t->S[i] = all.S[d];
    }
    t->task_off[t->L] = all.task_off[p->L] - all.task_off[d0];
    t->cnt_off[t->L] = all.cnt_off[p->L];
    t->w_off[t->L] = all.w_off[p->L];
}

void plr_level_tab0(const dplrh_plan *p, LevelTab *t)
{
    memset(t, 0, sizeof(*t));
    t->L = 0;
    t->N = p->N;
    t->b = p->b;
}

void plr_ws_per_col(const dplrh_plan *p, size_t *part_out, size_t *w_out, size_t *cnt_out)
{
    LevelTab t;
    plr_level_tab(p, 0, &t);
    size_t part = 0, part2 = 0;
    for (int d = 0; d < p->L; ++d) {
        if (t.tps[d] > 1) part += (size_t)(2 << d) * t.tps[d] * t.R[d];
        // layout of the two-launch matvec: one partial per 1024 rows of the
        // segments longer than 1024 rows
        const int seg = p->N >> (d + 1);
        if (seg > 1024) part2 += (size_t)(2 << d) * (seg / 1024) * t.R[d];
    }
    *part_out = part2 > part ? part2 : part;
    *w_out = p->soff[p->L];
    *cnt_out = (size_t)t.cnt_off[p->L];
}

int plr_workspace(dplrh_plan *p, int ncols)
{
    if (ncols <= p->wk_cols) return 0;
    size_t part, w, cnt;
    plr_ws_per_col(p, &part, &w, &cnt);
    ++p->ws_gen;
    if (p->part) (void)hipFree(p->part);
    if (p->w) (void)hipFree(p->w);
    if (p->cnt) (void)hipFree(p->cnt);
    p->part = nullptr;
    p->w = nullptr;
    p->cnt = nullptr;
    p->wk_cols = 0;
    p->part_per_col = part ? part : 1;
    p->w_per_col = w ? w : 1;
    p->cnt_per_col = cnt ? cnt : 1;
    PLR_HIP(hipMalloc(&p->part, p->part_per_col * ncols * sizeof(double)));
    PLR_HIP(hipMalloc(&p->w, p->w_per_col * ncols * sizeof(double)));
    PLR_HIP(hipMalloc(&p->cnt, p->cnt_per_col * ncols * sizeof(int)));
    // Arrival counters start at zero; every kernel that uses one returns it
    // to zero, so this is the only clear they need.
    PLR_HIP(hipMemset(p->cnt, 0, p->cnt_per_col * ncols * sizeof(int)));
    PLR_HIP(hipDeviceSynchronize());
    p->wk_cols = ncols;
    return 0;
}

// ---------------------------------------------------------------------------
// HIP graph cache

static int graphs_enabled()
{
    static int on = -1;
    if (on < 0) {
        const char *e = getenv("DPLR_GRAPHS");
        on = (e && e[0] == '0') ? 0 : 1;
    }
    return on;
}

void plr_graphs_clear(dplrh_plan *p)
{
    for (auto &g : p->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        memset(&g, 0, sizeof(g));
    }
}

int plr_run_cached_impl(dplrh_plan *p, const unsigned long long *key, hipStream_t user,
                        int (*op)(void *ctx, hipStream_t), void *ctx)
{
    if (!graphs_enabled()) return op(ctx, user);
    ++p->tick;
    int slot = -1, oldest = 0;
    for (int i = 0; i < 16; ++i) {
        if (p->graphs[i].seen && !memcmp(p->graphs[i].key, key, sizeof(p->graphs[i].key))) {
            slot = i;
            break;
        }
        if (p->graphs[i].tick < p->graphs[oldest].tick) oldest = i;
    }
    if (slot >= 0 && p->graphs[slot].gen != p->ws_gen) { // stale: buffers moved
        if (p->graphs[slot].exec) (void)hipGraphExecDestroy(p->graphs[slot].exec);
        p->graphs[slot].exec = nullptr;
        p->graphs[slot].seen = 0;
        slot = -1;
    }
    if (slot < 0) { // first time: run eagerly, remember the call
        const unsigned gen0 = p->ws_gen;
        int rc = op(ctx, user);
        if (rc) return rc;
        auto &g = p->graphs[oldest];
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        memset(&g, 0, sizeof(g));
        memcpy(g.key, key, sizeof(g.key));
        g.seen = 1;
        g.gen = p->ws_gen;
        g.tick = p->tick;
        (void)gen0;
        return 0;
    }
    auto &g = p->graphs[slot];
    g.tick = p->tick;
    if (!g.exec) {
        if (!p->cap) PLR_HIP(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
        PLR_HIP(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
        int rc = op(ctx, p->cap);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(p->cap, &graph);
        if (rc || e != hipSuccess || !graph || p->ws_gen != g.gen) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            g.seen = 0;
            return op(ctx, user); // could not capture: run eagerly
        }
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g.exec = nullptr;
            g.seen = 0;
            (void)hipGetLastError();
            return op(ctx, user);
        }
    }
    PLR_HIP(hipGraphLaunch(g.exec, user));
    return 0;
}

extern "C" int dplrh_plan_create(const plr_flat *f, dplrh_plan **out)
{
    *out = nullptr;
    dplrh_plan *p = new dplrh_plan;
    memset(p, 0, sizeof(*p));
    p->N = f->N;
    p->L = f->L;
    p->b = f->b;
    for (int d = 0; d < f->L; ++d) p->R[d] = f->R[d];
    for (int d = 0; d <= f->L; ++d) {
        p->uoff[d] = f->uoff[d];
        p->soff[d] = f->soff[d];
    }
    const size_t nuv = f->uoff[f->L], ns = f->soff[f->L];
    const size_t nd = (size_t)f->N * f->b;
    const size_t nr = (size_t)2 << f->L;
    int rc = 0;
    auto up = [&](void **dst, const void *src, size_t bytes) -> int {
        PLR_HIP(hipMalloc(dst, bytes ? bytes : 8));
        if (bytes) PLR_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return 0;
    };
    if ((rc = up((void **)&p->U, f->U, nuv * sizeof(double))) ||
        (rc = up((void **)&p->V, f->V, nuv * sizeof(double))) ||
        (rc = up((void **)&p->S, f->S, ns * sizeof(double))) ||
        (rc = up((void **)&p->D, f->D, nd * sizeof(double))) ||
        (rc = up((void **)&p->rank, f->rank, nr * sizeof(int))) ||
        (rc = plr_workspace(p, 1))) {
        dplrh_plan_destroy(p);
        return rc;
    }
    *out = p;
    return 0;
}

extern "C" void dplrh_plan_destroy(dplrh_plan *p)
{
    if (!p) return;
    plr_graphs_clear(p);
    if (p->cap) (void)hipStreamDestroy(p->cap);
    void *bufs[] = {p->U, p->V, p->S, p->D, p->rank, p->part, p->w, p->cnt, p->tmp};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    plr_free_fact(p);
    plr_lp_release(p);
    plr_up_release(p);
    plr_sp_release(p);
    if (p->self_batch) dplrh_batch_destroy(p->self_batch);
    delete p;
}

extern "C" int dplrh_plan_dims(const dplrh_plan *p, int *N, int *L, int *b, int *R)
{
    *N = p->N;
    *L = p->L;
    *b = p->b;
    for (int d = 0; d < p->L; ++d) R[d] = p->R[d];
    return 0;
}

extern "C" size_t dplrh_plan_bytes(const dplrh_plan *p)
{
    return (2 * p->uoff[p->L] + p->soff[p->L] + (size_t)p->N * p->b) * sizeof(double);
}

// ---------------------------------------------------------------------------
// utilities

__global__ void k_gather(const double *__restrict__ src, int inc, int n, double *__restrict__ dst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[(size_t)i * inc];
}

// dst[c*ldd + r] = src[r*lds + c]: dst is rows x cols, src is cols x rows.
__global__ void k_transpose(int rows, int cols, const double *__restrict__ src, int lds,
                            double *__restrict__ dst, int ldd)
{
    __shared__ double tile[32][33];
    const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 256 threads: 8 rows of 32
    for (int j = ty; j < 32; j += 8) {
        int c = c0 + j, r = r0 + tx;
        if (c < cols && r < rows) tile[j][tx] = src[(size_t)r * lds + c];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        int r = r0 + j, c = c0 + tx;
        if (c < cols && r < rows) dst[(size_t)c * ldd + r] = tile[tx][j];
    }
}

__global__ void k_scale(int n, int cols, double beta, int twice, double *__restrict__ y, int ldy)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i >= n || c >= cols) return;
    double *p = y + (size_t)c * ldy + i;
    if (beta == 0.0) {
        *p = 0.0;
    } else {
        double v = beta * *p;
        *p = twice ? beta * v : v;
    }
}

extern "C" int dplrh_gather(const double *src, int inc, int n, double *dst, void *stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       src, inc, n, dst);
    PLR_HIP(hipGetLastError());
    return 0;
}

extern "C" int dplrh_transpose(int rows, int cols, const double *src, int lds, double *dst,
                               int ldd, void *stream)
{
    if (rows <= 0 || cols <= 0) return 0;
    dim3 grid((rows + 31) / 32, (cols + 31) / 32);
    hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, (hipStream_t)stream, rows, cols, src,
                       lds, dst, ldd);
    PLR_HIP(hipGetLastError());
    return 0;
}

static int launch_scale(int n, int cols, double beta, int twice, double *dy, int ldy,
                        void *stream)
{
    if (n <= 0 || cols <= 0) return 0;
    hipLaunchKernelGGL(k_scale, dim3((n + 255) / 256, cols), dim3(256), 0,
                       (hipStream_t)stream, n, cols, beta, twice, dy, ldy);
    PLR_HIP(hipGetLastError());
    return 0;
}

extern "C" int dplrh_scale2(int n, int cols, double beta, double *dy, int ldy, void *stream)
{
    return launch_scale(n, cols, beta, 1, dy, ldy, stream);
}

extern "C" int dplrh_scale(int n, int cols, double beta, double *dy, int ldy, void *stream)
{
    return launch_scale(n, cols, beta, 0, dy, ldy, stream);
}

extern "C" void *dplrh_dev_alloc(size_t bytes)
{
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) {
        plr_set_error("hipMalloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}

extern "C" void dplrh_dev_free(void *p)
{
    if (p) (void)hipFree(p);
}

extern "C" int dplrh_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!bytes) return 0;
    PLR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return 0;
}

extern "C" int dplrh_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!bytes) return 0;
    PLR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    PLR_HIP(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

extern "C" int dplrh_sync(void *stream)
{
    PLR_HIP(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}
