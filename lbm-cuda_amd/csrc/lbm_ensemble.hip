// Batched ensembles of independent lattices of uneven sizes, advanced by one persistent kernel per AA
// step, and stream-capture graphs that replay any sequence of liblbm launches.
//
// No reference counterpart: the reference runs one lattice per process (lbm_smemaa_main.py:8-10).  The
// kernel walks a tile list that spans every lattice (tiles of 128 x 2 cells, as the single-step kernel):
// tile t belongs to the lattice whose first tile is the largest <= t.  Workgroups are dealt round-robin to
// the 8 XCDs, so XCD k takes the k-th contiguous eighth of the list (neighbouring rows of a lattice then
// share an L2), and each workgroup strides through its XCD's share.  Each lattice's update is exactly the
// single-step kernel's (aa_cell), so every member is bit-identical to its solo run.
#include "lbm_aa_cell.h"
#include "lbm_common.h"

#include <vector>

namespace lbm {

struct EnsDesc {
    KPar P;
    const uint8_t *mask;
    void *f;
    const float *inlet_u;
    int tiles_x;
    int tile0; // first tile of this lattice in the ensemble's list
};

} // namespace lbm

struct lbm_ensemble {
    int n = 0;
    int fmt = 0, model = 0, bc = 0;
    int default_blocks = 0; // workgroups that are resident at once on the whole device
    int total_tiles = 0;
    long long total_cells = 0;
    lbm::EnsDesc *d_desc = nullptr;
    std::vector<lbm::EnsDesc> h_desc;
};

struct lbm_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace lbm {

// Steps members [m0, m1) whose tiles are [t0, t1) of the list.
template <int MODEL, int FMT, int PARITY, bool PERIODIC>
__global__ __launch_bounds__(TILE_X *TILE_Y) void k_ensemble(const EnsDesc *__restrict__ d, int m0, int m1, int t0,
                                                              int t1)
{
    const int G = gridDim.x, b = blockIdx.x;
    const int gper = G / 8;
    const int xcd = b % 8, l = b / 8;
    const int per = (t1 - t0 + 7) / 8;
    const int start = t0 + xcd * per, end = min(start + per, t1);
    for (int t = start + l; t < end; t += gper) {
        int lo = m0, hi = m1 - 1; // uniform binary search over the first tiles
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (d[mid].tile0 <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const EnsDesc &e = d[lo];
        const int lt = t - e.tile0;
        const int ty = lt / e.tiles_x, tx = lt - ty * e.tiles_x;
        aa_cell<MODEL, FMT, PARITY, PERIODIC, false>(e.P, e.mask, e.f, e.inlet_u, nullptr, nullptr, nullptr,
                                                     tx * TILE_X + threadIdx.x, ty * TILE_Y + threadIdx.y);
    }
}

using EnsKernel = void (*)(const EnsDesc *, int, int, int, int);

template <int MODEL, int FMT>
static EnsKernel ens_kernel_m(int parity, bool periodic)
{
    if (parity)
        return periodic ? k_ensemble<MODEL, FMT, 1, true> : k_ensemble<MODEL, FMT, 1, false>;
    return periodic ? k_ensemble<MODEL, FMT, 0, true> : k_ensemble<MODEL, FMT, 0, false>;
}

template <int FMT>
static EnsKernel ens_kernel_f(int model, int parity, bool periodic)
{
    return model == LBM_BGK ? ens_kernel_m<MODEL_BGK, FMT>(parity, periodic)
                            : ens_kernel_m<MODEL_MRT, FMT>(parity, periodic);
}

static EnsKernel ens_kernel(const lbm_ensemble *e, int parity)
{
    const bool periodic = e->bc == LBM_BC_PERIODIC;
    if (e->fmt == LBM_FMT_F32) return ens_kernel_f<FMT_F32>(e->model, parity, periodic);
    if (e->fmt == LBM_FMT_F16) return ens_kernel_f<FMT_F16>(e->model, parity, periodic);
    return ens_kernel_f<FMT_BF16>(e->model, parity, periodic);
}

} // namespace lbm

using namespace lbm;

extern "C" {

int lbm_ensemble_create(const lbm_instance *inst, int n, int fmt, lbm_ensemble **out)
{
    if (!inst || !out || n < 1) return fail(LBM_ERR_ARG, "lbm_ensemble_create: need n >= 1 instances and an output");
    if (fmt != LBM_FMT_F32 && fmt != LBM_FMT_F16 && fmt != LBM_FMT_BF16)
        return fail(LBM_ERR_ARG, "lbm_ensemble_create: bad format %d", fmt);
    *out = nullptr;
    auto *e = new lbm_ensemble();
    e->n = n;
    e->fmt = fmt;
    e->model = inst[0].p.model;
    e->bc = inst[0].p.bc;
    e->h_desc.resize(n);
    long long tiles = 0;
    for (int i = 0; i < n; ++i) {
        EnsDesc &d = e->h_desc[i];
        int rc = make_kpar(&inst[i].p, &d.P);
        if (rc) {
            delete e;
            return rc;
        }
        if (inst[i].p.model != e->model || inst[i].p.bc != e->bc) {
            delete e;
            return fail(LBM_ERR_ARG, "lbm_ensemble_create: instance %d differs in model or bc", i);
        }
        if (!inst[i].mask || !inst[i].f || (e->bc == LBM_BC_CHANNEL && !inst[i].inlet_u)) {
            delete e;
            return fail(LBM_ERR_ARG, "lbm_ensemble_create: instance %d has a NULL buffer", i);
        }
        d.mask = inst[i].mask;
        d.f = inst[i].f;
        d.inlet_u = inst[i].inlet_u;
        d.tiles_x = (d.P.W + TILE_X - 1) / TILE_X;
        d.tile0 = (int)tiles;
        tiles += (long long)d.tiles_x * ((d.P.H + TILE_Y - 1) / TILE_Y);
        e->total_cells += (long long)d.P.plane;
    }
    if (tiles >= (1LL << 31)) {
        delete e;
        return fail(LBM_ERR_SIZE, "lbm_ensemble_create: too many tiles");
    }
    e->total_tiles = (int)tiles;
    // A persistent grid must be resident at once: a workgroup that starts only after others finish would
    // run its whole share of tiles at the end.  Size it by the occupancy of the slower parity's kernel.
    int dev = 0, cus = 0, occ = 1 << 30;
    hipError_t he = hipGetDevice(&dev);
    if (he == hipSuccess) he = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    for (int par = 0; par < 2 && he == hipSuccess; ++par) {
        int o = 0;
        he = hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, reinterpret_cast<const void *>(ens_kernel(e, par)),
                                                          TILE_X * TILE_Y, 0);
        occ = o < occ ? o : occ;
    }
    if (he == hipSuccess) {
        e->default_blocks = (occ > 0 ? occ : 1) * (cus > 0 ? cus : 1) / 8 * 8;
        if (e->default_blocks < 8) e->default_blocks = 8;
        he = hipMalloc(&e->d_desc, sizeof(EnsDesc) * n);
    }
    if (he == hipSuccess) he = hipMemcpy(e->d_desc, e->h_desc.data(), sizeof(EnsDesc) * n, hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        if (e->d_desc) (void)hipFree(e->d_desc);
        delete e;
        return fail((int)he, "lbm_ensemble_create: %s", hipGetErrorString(he));
    }
    *out = e;
    return LBM_OK;
}

int lbm_ensemble_info(const lbm_ensemble *e, int *total_tiles, long long *total_cells, int *default_blocks)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_info: NULL ensemble");
    if (total_tiles) *total_tiles = e->total_tiles;
    if (default_blocks) *default_blocks = e->default_blocks;
    if (total_cells) *total_cells = e->total_cells;
    return LBM_OK;
}

static int ens_launch(const lbm_ensemble *e, int parity, int blocks, int m0, int m1, hipStream_t s)
{
    const EnsDesc *d = e->d_desc;
    int t0 = e->h_desc[m0].tile0;
    int t1 = m1 < e->n ? e->h_desc[m1].tile0 : e->total_tiles;
    void *args[] = {&d, &m0, &m1, &t0, &t1};
    hipError_t he = hipLaunchKernel(reinterpret_cast<const void *>(ens_kernel(e, parity)), dim3(blocks),
                                    dim3(TILE_X, TILE_Y), args, 0, s);
    if (he != hipSuccess) return fail((int)he, "lbm_ensemble_step: %s", hipGetErrorString(he));
    return LBM_OK;
}

static int ens_blocks(const lbm_ensemble *e, int blocks)
{
    if (blocks <= 0) blocks = e->default_blocks;
    return (blocks + 7) / 8 * 8; // a multiple of the XCD count
}

int lbm_ensemble_step(lbm_ensemble *e, int parity, int blocks, lbm_stream_t stream)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_step: NULL ensemble");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_ensemble_step: parity must be 0 or 1");
    int rc = ens_launch(e, parity, ens_blocks(e, blocks), 0, e->n, to_hip(stream));
    return rc ? rc : check_launch("lbm_ensemble_step");
}

int lbm_ensemble_step_n(lbm_ensemble *e, int parity, int nsteps, long long group_bytes, int blocks,
                        lbm_stream_t stream)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_step_n: NULL ensemble");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_ensemble_step_n: parity must be 0 or 1");
    if (nsteps < 0) return fail(LBM_ERR_ARG, "lbm_ensemble_step_n: nsteps < 0");
    blocks = ens_blocks(e, blocks);
    hipStream_t s = to_hip(stream);
    const long long elem = e->fmt == LBM_FMT_F32 ? 4 : 2;
    int m0 = 0;
    while (m0 < e->n) {
        // the next group: consecutive members whose states fit in group_bytes (at least one member)
        int m1 = m0 + 1;
        long long bytes = 9LL * elem * e->h_desc[m0].P.plane;
        if (group_bytes > 0) {
            while (m1 < e->n) {
                const long long b = 9LL * elem * e->h_desc[m1].P.plane;
                if (bytes + b > group_bytes) break;
                bytes += b;
                ++m1;
            }
        } else {
            m1 = e->n;
        }
        for (int n = 0; n < nsteps; ++n) {
            int rc = ens_launch(e, (parity + n) & 1, blocks, m0, m1, s);
            if (rc) return rc;
        }
        m0 = m1;
    }
    return check_launch("lbm_ensemble_step_n");
}

int lbm_ensemble_destroy(lbm_ensemble *e)
{
    if (!e) return LBM_OK;
    hipError_t he = e->d_desc ? hipFree(e->d_desc) : hipSuccess;
    delete e;
    if (he != hipSuccess) return fail((int)he, "lbm_ensemble_destroy: %s", hipGetErrorString(he));
    return LBM_OK;
}

int lbm_graph_begin(lbm_stream_t stream)
{
    if (!stream) return fail(LBM_ERR_ARG, "lbm_graph_begin: capture needs a non-default stream");
    hipError_t he = hipStreamBeginCapture(to_hip(stream), hipStreamCaptureModeThreadLocal);
    if (he != hipSuccess) return fail((int)he, "lbm_graph_begin: %s", hipGetErrorString(he));
    return LBM_OK;
}

int lbm_graph_end(lbm_stream_t stream, lbm_graph **out)
{
    if (!out) return fail(LBM_ERR_ARG, "lbm_graph_end: NULL output");
    *out = nullptr;
    auto *g = new lbm_graph();
    hipError_t he = hipStreamEndCapture(to_hip(stream), &g->graph);
    if (he == hipSuccess) he = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (he != hipSuccess) {
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return fail((int)he, "lbm_graph_end: %s", hipGetErrorString(he));
    }
    *out = g;
    return LBM_OK;
}

int lbm_graph_launch(lbm_graph *g, lbm_stream_t stream)
{
    if (!g) return fail(LBM_ERR_ARG, "lbm_graph_launch: NULL graph");
    hipError_t he = hipGraphLaunch(g->exec, to_hip(stream));
    if (he != hipSuccess) return fail((int)he, "lbm_graph_launch: %s", hipGetErrorString(he));
    return LBM_OK;
}

int lbm_graph_destroy(lbm_graph *g)
{
    if (!g) return LBM_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return LBM_OK;
}

} // extern "C"
