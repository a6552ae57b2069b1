// Batched ensembles of independent lattices of uneven sizes, advanced by kernels over a work list spanning every
// member (a persistent grid for single steps, one workgroup per tile for the two- and four-step tiles), and
// stream-capture graphs that replay any sequence of liblbm launches (the reference driver's CuPy graph,
// lbm_smemaa_main.py:136-151).
//
// No reference counterpart: the reference runs one lattice per process (lbm_smemaa_main.py:8-10).  Work lists
// kept per ensemble:
//   k = 1  tiles of 128 x 2 cells, each updated by aa_cell (the single-step kernel's cell update), in place;
//   k = 2  tiles of 128 x CH cells, each advanced two steps by pair_tile (the two-step tile kernel), from a
//          member's current buffer into its second one;
//   k = 4  (FP32 populations) tiles of the four-step kernel's shapes (128 x 8, 64 x 12, 64 x 20, 64 x 24), each advanced
//          four steps by quad_tile, from one buffer into the other.
// lbm_ensemble_run advances any contiguous range of members with any of these, so a tuner can give each size
// class of an ensemble its own k and tile shape; each member's buffer side is tracked on its own.
// Tile t belongs to the member whose first tile is the largest <= t.  Workgroups are dealt round-robin to the
// 8 XCDs, so XCD k takes the k-th contiguous eighth of the list (neighbouring rows of a member share an L2), and
// each workgroup strides through its XCD's share.  Every member ends bit-identical to its solo run.
#include "lbm_ensemble_impl.h"

#include <vector>

struct lbm_ensemble {
    int n = 0;
    int fmt = 0, model = 0, bc = 0, sums = 0, fast = 0;
    std::vector<int> side;  // per member: which of its buffers holds the state
    bool has_alt = false;   // every member has a second buffer
    int default_blocks = 0; // single-step workgroups resident at once on the whole device
    int total_tiles = 0;
    int total_ptiles[lbm::ENS_NCH] = {0, 0, 0, 0};
    int total_qtiles[lbm::ENS_NQ] = {0, 0, 0, 0};
    long long total_cells = 0;
    lbm::EnsDesc *d_desc = nullptr;
    std::vector<lbm::EnsDesc> h_desc;
};

struct lbm_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace lbm {

static PairKernel quad_kernel(const lbm_ensemble *e, int qi)
{
    const bool periodic = e->bc == LBM_BC_PERIODIC;
    if (e->sums) return quad_kernel_sums(periodic, qi);
    if (e->fast) return quad_kernel_fast(periodic, qi);
    if (e->model == LBM_BGK) return periodic ? quad_kernel_c<MODEL_BGK, true>(qi) : quad_kernel_c<MODEL_BGK, false>(qi);
    return periodic ? quad_kernel_c<MODEL_MRT, true>(qi) : quad_kernel_c<MODEL_MRT, false>(qi);
}

static int quad_threads(int qi) { return quadk::quad_nt(ENS_QCH[qi]); }

template <int FMT>
static EnsKernel ens_kernel_f(int model, int parity, bool periodic)
{
    return model == LBM_BGK ? ens_kernel_m<MODEL_BGK, FMT>(parity, periodic)
                            : ens_kernel_m<MODEL_MRT, FMT>(parity, periodic);
}

static EnsKernel ens_kernel(const lbm_ensemble *e, int parity)
{
    const bool periodic = e->bc == LBM_BC_PERIODIC;
    if (e->sums) return ens_kernel_sums(e->fmt, parity, periodic);
    if (e->fast) return ens_kernel_fast(e->fmt, parity, periodic);
    if (e->fmt == LBM_FMT_F32) return ens_kernel_f<FMT_F32>(e->model, parity, periodic);
    if (e->fmt == LBM_FMT_F16) return ens_kernel_f<FMT_F16>(e->model, parity, periodic);
    return ens_kernel_f<FMT_BF16>(e->model, parity, periodic);
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
    if (e->sums) return pair_kernel_sums(e->fmt, periodic, chi);
    if (e->fast) return pair_kernel_fast(e->fmt, periodic, chi);
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
static int next_group(const lbm_ensemble *e, int m0, long long group_bytes, int mend)
{
    if (group_bytes <= 0) return mend;
    const long long elem = e->fmt == LBM_FMT_F32 ? 4 : 2;
    long long bytes = 9LL * elem * e->h_desc[m0].P.plane;
    int m1 = m0 + 1;
    while (m1 < mend) {
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
    e->side.assign(n, 0);
    long long tiles = 0, ptiles[ENS_NCH] = {0, 0, 0, 0}, qtiles[ENS_NQ] = {0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        EnsDesc &d = e->h_desc[i];
        int rc = make_kpar(&inst[i].p, &d.P, fmt == LBM_FMT_F32 ? 4u : 2u);
        if (rc) {
            delete e;
            return rc;
        }
        if (inst[i].p.row_base) {
            delete e;
            return fail(LBM_ERR_UNSUPPORTED, "lbm_ensemble_create: member %d: row-based storage (row_base) is for "
                                             "single lattices", i);
        }
        if (i == 0) {
            e->sums = d.P.sums;
            e->fast = d.P.fast;
        }
        if (d.P.sums && e->model != LBM_MRT_LES) {
            delete e;
            return fail(LBM_ERR_UNSUPPORTED, "lbm_ensemble_create: compensated sums need the MRT model");
        }
        if (inst[i].p.model != e->model || inst[i].p.bc != e->bc || d.P.sums != e->sums || d.P.fast != e->fast) {
            delete e;
            return fail(LBM_ERR_ARG, "lbm_ensemble_create: instance %d differs in model, bc, sums or fast", i);
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
        for (int c = 0; c < ENS_NQ; ++c) {
            const int cw = quadk::quad_cw(ENS_QCH[c]);
            d.qtiles_x[c] = (d.P.W + cw - 1) / cw;
            d.qtile0[c] = (int)qtiles[c];
            qtiles[c] += (long long)d.qtiles_x[c] * ((d.P.H + ENS_QCH[c] - 1) / ENS_QCH[c]);
        }
        e->total_cells += (long long)d.P.plane;
    }
    bool too_many = tiles >= (1LL << 31);
    for (int c = 0; c < ENS_NCH; ++c) too_many = too_many || ptiles[c] >= (1LL << 31);
    for (int c = 0; c < ENS_NQ; ++c) too_many = too_many || qtiles[c] >= (1LL << 31);
    if (too_many) {
        delete e;
        return fail(LBM_ERR_SIZE, "lbm_ensemble_create: too many tiles");
    }
    e->total_tiles = (int)tiles;
    for (int c = 0; c < ENS_NCH; ++c) e->total_ptiles[c] = (int)ptiles[c];
    for (int c = 0; c < ENS_NQ; ++c) e->total_qtiles[c] = (int)qtiles[c];
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

static int quad_qi(int chunk)
{
    for (int c = 0; c < ENS_NQ; ++c)
        if (ENS_QCH[c] == chunk) return c;
    return -1;
}

int lbm_ensemble_pair_info(lbm_ensemble *e, int chunk, int *total_tiles, int *default_blocks)
{
    return lbm_ensemble_run_info(e, 2, chunk, total_tiles, default_blocks);
}

int lbm_ensemble_run_info(lbm_ensemble *e, int k, int chunk, int *total_tiles, int *default_blocks)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_run_info: NULL ensemble");
    int tiles = 0, blocks = 0;
    if (k == 1) {
        tiles = e->total_tiles;
        blocks = e->default_blocks;
    } else if (k == 2) {
        const int chi = pair_chi(chunk);
        if (chi < 0) return fail(LBM_ERR_ARG, "lbm_ensemble_run_info: k = 2 needs chunk 8, 12, 16 or 24");
        tiles = e->total_ptiles[chi];
        blocks = tiles; // one workgroup per tile
    } else if (k == 4) {
        const int qi = quad_qi(chunk);
        if (qi < 0) return fail(LBM_ERR_ARG, "lbm_ensemble_run_info: k = 4 needs chunk 8, 12, 20 or 24");
        if (e->fmt != LBM_FMT_F32)
            return fail(LBM_ERR_UNSUPPORTED, "lbm_ensemble_run_info: k = 4 takes FP32 populations only");
        tiles = e->total_qtiles[qi];
        blocks = tiles; // one workgroup per tile
    } else {
        return fail(LBM_ERR_ARG, "lbm_ensemble_run_info: k must be 1, 2 or 4");
    }
    if (total_tiles) *total_tiles = tiles;
    if (default_blocks) *default_blocks = blocks;
    return LBM_OK;
}

int lbm_ensemble_side(const lbm_ensemble *e)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_side: NULL ensemble");
    for (int i = 1; i < e->n; ++i)
        if (e->side[i] != e->side[0])
            return fail(LBM_ERR_ARG, "lbm_ensemble_side: members hold their states on different sides "
                                     "(lbm_ensemble_member_side)");
    return e->side[0];
}

int lbm_ensemble_member_side(const lbm_ensemble *e, int i)
{
    if (!e || i < 0 || i >= e->n) return fail(LBM_ERR_ARG, "lbm_ensemble_member_side: no member %d", i);
    return e->side[i];
}

// one launch of the k-step kernel (list `li`: chunk index) over members [m0, m1), tiles [t0, t1), from `side`
static int ens_launch_k(const lbm_ensemble *e, int k, int li, int parity, int blocks, int m0, int m1, int side,
                        hipStream_t s)
{
    const EnsDesc *d = e->d_desc;
    hipError_t he;
    if (k == 1) {
        int t0 = e->h_desc[m0].tile0;
        int t1 = m1 < e->n ? e->h_desc[m1].tile0 : e->total_tiles;
        void *args[] = {&d, &m0, &m1, &t0, &t1, &side};
        he = hipLaunchKernel(reinterpret_cast<const void *>(ens_kernel(e, parity)), dim3(blocks),
                             dim3(TILE_X, TILE_Y), args, 0, s);
    } else if (k == 2) {
        int t0 = e->h_desc[m0].ptile0[li];
        int t1 = m1 < e->n ? e->h_desc[m1].ptile0[li] : e->total_ptiles[li];
        if (t1 <= t0) return LBM_OK;
        void *args[] = {&d, &m0, &m1, &t0, &t1, &side, &li};
        he = hipLaunchKernel(reinterpret_cast<const void *>(pair_kernel(e, li)), dim3(t1 - t0), dim3(pairk::NT), args,
                             0, s);
    } else {
        int t0 = e->h_desc[m0].qtile0[li];
        int t1 = m1 < e->n ? e->h_desc[m1].qtile0[li] : e->total_qtiles[li];
        if (t1 <= t0) return LBM_OK;
        void *args[] = {&d, &m0, &m1, &t0, &t1, &side, &li};
        he = hipLaunchKernel(reinterpret_cast<const void *>(quad_kernel(e, li)), dim3(t1 - t0), dim3(quad_threads(li)),
                             args, 0, s);
    }
    if (he != hipSuccess) return fail((int)he, "lbm_ensemble_run: %s", hipGetErrorString(he));
    return LBM_OK;
}

int lbm_ensemble_run(lbm_ensemble *e, int m0, int m1, int k, int chunk, int nsteps, int parity,
                     long long group_bytes, int blocks, lbm_stream_t stream)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_run: NULL ensemble");
    if (m0 < 0 || m1 > e->n || m0 >= m1) return fail(LBM_ERR_ARG, "lbm_ensemble_run: members [%d, %d) of %d", m0, m1, e->n);
    if (nsteps < 0) return fail(LBM_ERR_ARG, "lbm_ensemble_run: nsteps < 0");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_ensemble_run: parity must be 0 or 1");
    for (int i = m0 + 1; i < m1; ++i)
        if (e->side[i] != e->side[m0])
            return fail(LBM_ERR_ARG, "lbm_ensemble_run: members %d and %d hold their states on different sides", m0, i);
    int li = 0, dflt = e->default_blocks;
    if (k == 2 || k == 4) {
        if (!e->has_alt) return fail(LBM_ERR_ARG, "lbm_ensemble_run: k = %d needs a second buffer (f_alt) per member", k);
        if (parity != 0) return fail(LBM_ERR_UNSUPPORTED, "lbm_ensemble_run: k = %d starts from an even state", k);
        if (nsteps % k) return fail(LBM_ERR_ARG, "lbm_ensemble_run: nsteps %d is not a multiple of k = %d", nsteps, k);
        int rc = lbm_ensemble_run_info(e, k, chunk, nullptr, &dflt);
        if (rc) return rc;
        li = k == 2 ? pair_chi(chunk) : quad_qi(chunk);
    } else if (k != 1) {
        return fail(LBM_ERR_ARG, "lbm_ensemble_run: k must be 1, 2 or 4");
    }
    blocks = ens_blocks(blocks, dflt);
    hipStream_t s = to_hip(stream);
    const int nl = k == 1 ? nsteps : nsteps / k;
    const int side0 = e->side[m0];
    for (int g0 = m0; g0 < m1;) {
        const int g1 = next_group(e, g0, group_bytes, m1);
        for (int n = 0; n < nl; ++n) {
            const int side = k == 1 ? side0 : side0 ^ (n & 1);
            int rc = ens_launch_k(e, k, li, (parity + n) & 1, blocks, g0, g1, side, s);
            if (rc) return rc;
        }
        g0 = g1;
    }
    if (k != 1)
        for (int i = m0; i < m1; ++i) e->side[i] ^= nl & 1;
    return check_launch("lbm_ensemble_run");
}

int lbm_ensemble_step(lbm_ensemble *e, int parity, int blocks, lbm_stream_t stream)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_step: NULL ensemble");
    return lbm_ensemble_run(e, 0, e->n, 1, 0, 1, parity, 0, blocks, stream);
}

int lbm_ensemble_step_n(lbm_ensemble *e, int parity, int nsteps, long long group_bytes, int blocks,
                        lbm_stream_t stream)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_step_n: NULL ensemble");
    return lbm_ensemble_run(e, 0, e->n, 1, 0, nsteps, parity, group_bytes, blocks, stream);
}

int lbm_ensemble_pair_n(lbm_ensemble *e, int npairs, long long group_bytes, int chunk, int blocks,
                        lbm_stream_t stream)
{
    if (!e) return fail(LBM_ERR_ARG, "lbm_ensemble_pair_n: NULL ensemble");
    if (npairs < 0) return fail(LBM_ERR_ARG, "lbm_ensemble_pair_n: npairs < 0");
    return lbm_ensemble_run(e, 0, e->n, 2, chunk, 2 * npairs, 0, group_bytes, blocks, stream);
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
    // the instantiated graph's device-side setup now, not in its first launch
    if (he == hipSuccess) he = hipGraphUpload(g->exec, to_hip(stream));
    if (he == hipSuccess) he = hipStreamSynchronize(to_hip(stream));
    if (he != hipSuccess) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
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
