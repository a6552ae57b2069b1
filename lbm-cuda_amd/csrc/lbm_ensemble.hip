// Batched ensembles of independent lattices of uneven sizes, advanced by persistent kernels that walk a work
// list spanning every member, and stream-capture graphs that replay any sequence of liblbm launches (the
// reference driver's CuPy graph, lbm_smemaa_main.py:136-151).
//
// No reference counterpart: the reference runs one lattice per process (lbm_smemaa_main.py:8-10).  Two work
// lists are kept per ensemble:
//   k = 1  tiles of 128 x 2 cells, each updated by aa_cell (the single-step kernel's cell update), in place;
//   k = 2  tiles of 128 x CH cells, each advanced two steps by pair_tile (the two-step tile kernel), from a
//          member's current buffer into its second one.
// Tile t belongs to the member whose first tile is the largest <= t.  Workgroups are dealt round-robin to the
// 8 XCDs, so XCD k takes the k-th contiguous eighth of the list (neighbouring rows of a member share an L2), and
// each workgroup strides through its XCD's share.  Every member ends bit-identical to its solo run.
#include "lbm_aa_cell.h"
#include "lbm_common.h"
#include "lbm_pair_tile.h"

#include <vector>

namespace lbm {

constexpr int ENS_NCH = 4;
constexpr int ENS_CH[ENS_NCH] = {8, 12, 16, 24}; // tile rows of the two-step list

struct EnsDesc {
    KPar P;
    const uint8_t *mask;
    void *f[2]; // the state lives in f[side]; f[1] is the second buffer of the two-step path (may be NULL)
    const float *inlet_u;
    int tiles_x;
    int tile0;              // first 128 x 2 tile of this member
    int ptiles_x;           // 128-column tiles of the two-step list
    int ptile0[ENS_NCH];    // first 128 x CH tile of this member, per CH
};

} // namespace lbm

struct lbm_ensemble {
    int n = 0;
    int fmt = 0, model = 0, bc = 0;
    int side = 0;           // which of each member's buffers holds the state
    bool has_alt = false;   // every member has a second buffer
    int default_blocks = 0; // single-step workgroups resident at once on the whole device
    int pair_blocks[lbm::ENS_NCH] = {0, 0, 0, 0}; // the same for the two-step kernels (0 = not yet computed)
    int total_tiles = 0;
    int total_ptiles[lbm::ENS_NCH] = {0, 0, 0, 0};
    long long total_cells = 0;
    lbm::EnsDesc *d_desc = nullptr;
    std::vector<lbm::EnsDesc> h_desc;
};

struct lbm_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace lbm {

// the member of list tile t among [m0, m1): uniform binary search over the members' first tiles
template <typename First>
__device__ __forceinline__ int ens_member(const EnsDesc *__restrict__ d, int m0, int m1, int t, First first)
{
    int lo = m0, hi = m1 - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first(d[mid]) <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// One single step of members [m0, m1) whose tiles are [t0, t1) of the k = 1 list, in their buffer `side`.
template <int MODEL, int FMT, int PARITY, bool PERIODIC>
__global__ __launch_bounds__(TILE_X *TILE_Y) void k_ensemble(const EnsDesc *__restrict__ d, int m0, int m1, int t0,
                                                              int t1, int side)
{
    const int G = gridDim.x, b = blockIdx.x;
    const int gper = G / 8;
    const int xcd = b % 8, l = b / 8;
    const int per = (t1 - t0 + 7) / 8;
    const int start = t0 + xcd * per, end = min(start + per, t1);
    for (int t = start + l; t < end; t += gper) {
        const EnsDesc &e = d[ens_member(d, m0, m1, t, [](const EnsDesc &x) { return x.tile0; })];
        const int lt = t - e.tile0;
        const int ty = lt / e.tiles_x, tx = lt - ty * e.tiles_x;
        aa_tile_cell<MODEL, FMT, PARITY, PERIODIC, false>(e.P, e.mask, e.f[side], e.inlet_u, nullptr, nullptr, nullptr,
                                                          tx, ty);
    }
}

// Two steps (even then odd) of members [m0, m1) whose tiles are [t0, t1) of the k = 2 list (tile rows
// ENS_CH[chi] = CH), from buffer `side` into buffer `side ^ 1`.
template <int MODEL, int FMT, bool PERIODIC, int CH>
__global__ __launch_bounds__(pairk::NT) void k_ensemble_pair(const EnsDesc *__restrict__ d, int m0, int m1, int t0,
                                                             int t1, int side, int chi)
{
    __shared__ pairk::TileLds<FMT, CH> L;
    const int G = gridDim.x, b = blockIdx.x;
    const int gper = G / 8;
    const int xcd = b % 8, l = b / 8;
    const int per = (t1 - t0 + 7) / 8;
    const int start = t0 + xcd * per, end = min(start + per, t1);
    for (int t = start + l; t < end; t += gper) {
        const EnsDesc &e = d[ens_member(d, m0, m1, t, [chi](const EnsDesc &x) { return x.ptile0[chi]; })];
        const int lt = t - e.ptile0[chi];
        const int ty = lt / e.ptiles_x, tx = lt - ty * e.ptiles_x;
        pairk::pair_tile_any<MODEL, PERIODIC, FMT, CH>(e.P, e.mask, e.f[side], e.f[side ^ 1], e.inlet_u, nullptr, nullptr,
                                                   nullptr, tx, ty, L);
        __syncthreads(); // the next tile's phase A overwrites the LDS this tile's phase B read
    }
}

using EnsKernel = void (*)(const EnsDesc *, int, int, int, int, int);
using PairKernel = void (*)(const EnsDesc *, int, int, int, int, int, int);

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

template <int MODEL, int FMT, bool PERIODIC>
static PairKernel pair_kernel_c(int chi)
{
    switch (chi) {
    case 0: return k_ensemble_pair<MODEL, FMT, PERIODIC, 8>;
    case 1: return k_ensemble_pair<MODEL, FMT, PERIODIC, 12>;
    case 2: return k_ensemble_pair<MODEL, FMT, PERIODIC, 16>;
    default: return k_ensemble_pair<MODEL, FMT, PERIODIC, 24>;
    }
}

template <int FMT>
static PairKernel pair_kernel_f(int model, bool periodic, int chi)
{
    if (model == LBM_BGK)
        return periodic ? pair_kernel_c<MODEL_BGK, FMT, true>(chi) : pair_kernel_c<MODEL_BGK, FMT, false>(chi);
    return periodic ? pair_kernel_c<MODEL_MRT, FMT, true>(chi) : pair_kernel_c<MODEL_MRT, FMT, false>(chi);
}

static PairKernel pair_kernel(const lbm_ensemble *e, int chi)
{
    const bool periodic = e->bc == LBM_BC_PERIODIC;
    if (e->fmt == LBM_FMT_F32) return pair_kernel_f<FMT_F32>(e->model, periodic, chi);
    if (e->fmt == LBM_FMT_F16) return pair_kernel_f<FMT_F16>(e->model, periodic, chi);
    return pair_kernel_f<FMT_BF16>(e->model, periodic, chi);
}

// workgroups of `kernel` (block size `threads`) resident at once on the current device, a multiple of 8
static hipError_t resident_blocks(const void *kernel, int threads, int &blocks)
{
    int dev = 0, cus = 0, occ = 0;
    hipError_t he = hipGetDevice(&dev);
    if (he == hipSuccess) he = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (he == hipSuccess) he = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0);
    if (he != hipSuccess) return he;
    blocks = (occ > 0 ? occ : 1) * (cus > 0 ? cus : 1) / 8 * 8;
    if (blocks < 8) blocks = 8;
    return hipSuccess;
}

static int ens_blocks(int blocks, int dflt)
{
    if (blocks <= 0) blocks = dflt;
    return (blocks + 7) / 8 * 8; // a multiple of the XCD count
}

// The next cache group after member m0: consecutive members whose states total at most group_bytes (at
// least one member); group_bytes <= 0 puts every member in one group.
static int next_group(const lbm_ensemble *e, int m0, long long group_bytes)
{
    if (group_bytes <= 0) return e->n;
    const long long elem = e->fmt == LBM_FMT_F32 ? 4 : 2;
    long long bytes = 9LL * elem * e->h_desc[m0].P.plane;
    int m1 = m0 + 1;
    while (m1 < e->n) {
        const long long b = 9LL * elem * e->h_desc[m1].P.plane;
        if (bytes + b > group_bytes) break;
        bytes += b;
        ++m1;
    }
    return m1;
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
    e->has_alt = true;
    e->h_desc.resize(n);
    long long tiles = 0, ptiles[ENS_NCH] = {0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        EnsDesc &d = e->h_desc[i];
        int rc = make_kpar(&inst[i].p, &d.P, fmt == LBM_FMT_F32 ? 4u : 2u);
        if (rc) {
            delete e;
            return rc;
        }
        if (d.P.sums) {
            delete e;
            return fail(LBM_ERR_UNSUPPORTED, "lbm_ensemble_create: compensated sums are single-step (AA) only");
        }
        if (inst[i].p.model != e->model || inst[i].p.bc != e->bc) {
            delete e;
            return fail(LBM_ERR_ARG, "lbm_ensemble_create: instance %d differs in model or bc", i);
        }
        if (!inst[i].mask || !inst[i].f || (e->bc == LBM_BC_CHANNEL && !inst[i].inlet_u)) {
            delete e;
            return fail(LBM_ERR_ARG, "lbm_ensemble_create: instance %d has a NULL buffer", i);
        }
        if (inst[i].f_alt == inst[i].f) {
            delete e;
            return fail(LBM_ERR_ARG, "lbm_ensemble_create: instance %d: f_alt must differ from f", i);
        }
        d.mask = inst[i].mask;
        d.f[0] = inst[i].f;
        d.f[1] = inst[i].f_alt;
        e->has_alt = e->has_alt && inst[i].f_alt != nullptr;
        d.inlet_u = inst[i].inlet_u;
        d.tiles_x = (d.P.W + TILE_X - 1) / TILE_X;
        d.tile0 = (int)tiles;
        tiles += (long long)d.tiles_x * ((d.P.H + TILE_Y - 1) / TILE_Y);
        d.ptiles_x = (d.P.W + pairk::CW - 1) / pairk::CW;
        for (int c = 0; c < ENS_NCH; ++c) {
            d.ptile0[c] = (int)ptiles[c];
            ptiles[c] += (long long)d.ptiles_x * ((d.P.H + ENS_CH[c] - 1) / ENS_CH[c]);
        }
        e->total_cells += (long long)d.P.plane;
    }
    if (tiles >= (1LL << 31)) {
        delete e;
        return fail(LBM_ERR_SIZE, "lbm_ensemble_create: too many tiles");
    }
    e->total_tiles = (int)tiles;
    for (int c = 0; c < ENS_NCH; ++c) e->total_ptiles[c] = (int)ptiles[c];
    // A persistent grid must be resident at once: a workgroup that starts only after others finish would
    // run its whole share of tiles at the end.  Size it by the occupancy of the slower parity's kernel.
    int occ_blocks = 1 << 30;
    hipError_t he = hipSuccess;
    for (int par = 0; par < 2 && he == hipSuccess; ++par) {
        int b = 0;
        he = resident_blocks(reinterpret_cast<const void *>(ens_kernel(e, par)), TILE_X * TILE_Y, b);
        occ_blocks = b < occ_blocks ? b : occ_blocks;
    }
    if (he == hipSuccess) {
        e->default_blocks = occ_blocks;
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

static int pair_chi(int chunk)
{
    for (int c = 0; c < ENS_NCH; ++c)
        if (ENS_CH[c] == chunk) return c;
    return -1;
}

// resident two-step workgroups for tile rows ENS_CH[chi] (computed on first use)
static int pair_default(lbm_ensemble *e, int chi, const char *who)
{
    if (!e->pair_blocks[chi]) {
        hipError_t he = resident_blocks(reinterpret_cast<const void *>(pair_kernel(e, chi)), pairk::NT,
                                        e->pair_blocks[chi]);
        if (he != hipSuccess) return fail((int)he, "%s: %s", who, hipGetErrorString(he));
    }
    return LBM_OK;
}

int lbm_ensemble_pair_info(lbm_ensemble *e, int chunk, int *total_tiles, int *default_blocks)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_pair_info: NULL ensemble");
    const int chi = pair_chi(chunk);
    if (chi < 0) return fail(LBM_ERR_ARG, "lbm_ensemble_pair_info: chunk must be 8, 12, 16 or 24");
    int rc = pair_default(e, chi, "lbm_ensemble_pair_info");
    if (rc) return rc;
    if (total_tiles) *total_tiles = e->total_ptiles[chi];
    if (default_blocks) *default_blocks = e->pair_blocks[chi];
    return LBM_OK;
}

int lbm_ensemble_side(const lbm_ensemble *e)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_side: NULL ensemble");
    return e->side;
}

static int ens_launch(const lbm_ensemble *e, int parity, int blocks, int m0, int m1, hipStream_t s)
{
    const EnsDesc *d = e->d_desc;
    int t0 = e->h_desc[m0].tile0;
    int t1 = m1 < e->n ? e->h_desc[m1].tile0 : e->total_tiles;
    int side = e->side;
    void *args[] = {&d, &m0, &m1, &t0, &t1, &side};
    hipError_t he = hipLaunchKernel(reinterpret_cast<const void *>(ens_kernel(e, parity)), dim3(blocks),
                                    dim3(TILE_X, TILE_Y), args, 0, s);
    if (he != hipSuccess) return fail((int)he, "lbm_ensemble_step: %s", hipGetErrorString(he));
    return LBM_OK;
}

int lbm_ensemble_step(lbm_ensemble *e, int parity, int blocks, lbm_stream_t stream)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_step: NULL ensemble");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_ensemble_step: parity must be 0 or 1");
    int rc = ens_launch(e, parity, ens_blocks(blocks, e->default_blocks), 0, e->n, to_hip(stream));
    return rc ? rc : check_launch("lbm_ensemble_step");
}

int lbm_ensemble_step_n(lbm_ensemble *e, int parity, int nsteps, long long group_bytes, int blocks,
                        lbm_stream_t stream)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_step_n: NULL ensemble");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_ensemble_step_n: parity must be 0 or 1");
    if (nsteps < 0) return fail(LBM_ERR_ARG, "lbm_ensemble_step_n: nsteps < 0");
    blocks = ens_blocks(blocks, e->default_blocks);
    hipStream_t s = to_hip(stream);
    for (int m0 = 0; m0 < e->n;) {
        const int m1 = next_group(e, m0, group_bytes);
        for (int n = 0; n < nsteps; ++n) {
            int rc = ens_launch(e, (parity + n) & 1, blocks, m0, m1, s);
            if (rc) return rc;
        }
        m0 = m1;
    }
    return check_launch("lbm_ensemble_step_n");
}

int lbm_ensemble_pair_n(lbm_ensemble *e, int npairs, long long group_bytes, int chunk, int blocks,
                        lbm_stream_t stream)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_pair_n: NULL ensemble");
    if (!e->has_alt) return fail(LBM_ERR_ARG, "lbm_ensemble_pair_n: every member needs a second buffer (f_alt)");
    if (npairs < 0) return fail(LBM_ERR_ARG, "lbm_ensemble_pair_n: npairs < 0");
    const int chi = pair_chi(chunk);
    if (chi < 0) return fail(LBM_ERR_ARG, "lbm_ensemble_pair_n: chunk must be 8, 12, 16 or 24");
    const void *kern = reinterpret_cast<const void *>(pair_kernel(e, chi));
    int rc = pair_default(e, chi, "lbm_ensemble_pair_n");
    if (rc) return rc;
    blocks = ens_blocks(blocks, e->pair_blocks[chi]);
    hipStream_t s = to_hip(stream);
    const EnsDesc *d = e->d_desc;
    for (int m0 = 0; m0 < e->n;) {
        int m1 = next_group(e, m0, group_bytes);
        int t0 = e->h_desc[m0].ptile0[chi];
        int t1 = m1 < e->n ? e->h_desc[m1].ptile0[chi] : e->total_ptiles[chi];
        for (int n = 0; n < npairs; ++n) {
            int side = e->side ^ (n & 1);
            int a0 = m0, c = chi;
            void *args[] = {&d, &a0, &m1, &t0, &t1, &side, &c};
            hipError_t he = hipLaunchKernel(kern, dim3(blocks), dim3(pairk::NT), args, 0, s);
            if (he != hipSuccess) return fail((int)he, "lbm_ensemble_pair_n: %s", hipGetErrorString(he));
        }
        m0 = m1;
    }
    e->side ^= npairs & 1;
    return check_launch("lbm_ensemble_pair_n");
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
