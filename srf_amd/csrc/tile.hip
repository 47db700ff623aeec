// Tiled TOR ops: plan construction and the C ABI entry points
// (include/srfx.h, srfx_tor_plan_*).  Device code in tile_core.h.
#include <rocprim/device/device_scan.hpp>

#include <cstring>
#include <vector>

#include "tile_core.h"

using namespace srfx;



namespace {

template <class T> int dev_alloc(T **ptr, size_t count)
{
    *ptr = nullptr;
    if (count == 0)
        count = 1;
    if (hipMalloc((void **)ptr, sizeof(T) * count) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return SRFX_ENOMEM;
    }
    return SRFX_OK;
}

void plan_free(srfx_tor_plan_s *pl)
{
    if (!pl)
        return;
    if (pl->arena) {
        // records, projection buffer, list and chunks live in the set's blocks
        if (pl->arena->refs.fetch_sub(1) == 1) {
            (void)hipFree(pl->arena->mem[0]);
            (void)hipFree(pl->arena->mem[1]);
            delete pl->arena;
        }
    } else {
        (void)hipFree(pl->list);
        (void)hipFree(pl->rec);
        (void)hipFree(pl->chunks);
        (void)hipFree(pl->proj);
    }
    (void)hipFree(pl->ev_off);
    (void)hipFree(pl->ev_slot);
    (void)hipFree(pl->partial);
    delete pl;
}

// LDS of the gather (FP32 tile) and scatter (FP64 tile) kernels.
unsigned lds_bytes(const TileGeom &g)
{
    return (unsigned)(sizeof(float) * g.S * g.Wx * g.Wy);
}

unsigned lds_bytes_scatter(const TileGeom &g)
{
    return (unsigned)(sizeof(double) * g.S * g.Wx * g.Wy);
}

// Largest tile count whose per-workgroup counters (3 per tile) fit in 64 KB.
constexpr int kLdsTiles = 16384 / 3;
constexpr unsigned kLdsPerCU = 160u * 1024u; // LDS of one CU

template <bool FILL>
void launch_pass(const TorParams &p, const TileGeom &g, const PlanArgs &a, hipStream_t st)
{
    const dim3 grid(blocks_for(p.n, 256)), block(256);
    if (a.ntiles <= kLdsTiles)
        hipLaunchKernelGGL((plan_pass<FILL, true>), grid, block,
                           (FILL ? 3u : 1u) * sizeof(int) * (unsigned)a.ntiles, st, p, g, a);
    else
        hipLaunchKernelGGL((plan_pass<FILL, false>), grid, block, 0, st, p, g, a);
}

// A read image in contiguous op-frame layout of fewer than 2^31 bytes, with
// fewer than 2^24 voxels per slice, goes through the frame gathers (buffer
// loads with 32-bit offsets formed by 24-bit products).
bool frame_image(const TorParams &p, int &bytes)
{
    const int64_t b = (int64_t)4 * p.gx * p.gy * p.gz;
    bytes = (int)b;
    return p.sx == 1 && p.sy == p.gx && p.sz == p.gx * p.gy && b > 0 &&
           b < ((int64_t)1 << 31) && (int64_t)p.gx * p.gy < ((int64_t)1 << 24);
}

template <int NB, int LPE>
int launch_record_pg(const TorParams &p, float4 *rec, float *pout, int block, hipStream_t st)
{
    TorParams q = p;
    if (frame_image(q, q.img_bytes))
        hipLaunchKernelGGL((record_proj_kernel<NB, LPE, true>), dim3(blocks_for(q.n * LPE, block)),
                           dim3(block), 0, st, q, rec, pout);
    else
        hipLaunchKernelGGL((record_proj_kernel<NB, LPE, false>),
                           dim3(blocks_for(q.n * LPE, block)), dim3(block), 0, st, q, rec, pout);
    return launch_status();
}

template <int LPE>
int launch_record_lpe(const TorParams &p, float4 *rec, float *pout, int block, hipStream_t st)
{
    switch (p.full_patch ? 0 : p.box) {
    case 3:
        return launch_record_pg<3, LPE>(p, rec, pout, block, st);
    case 4:
        return launch_record_pg<4, LPE>(p, rec, pout, block, st);
    default:
        return launch_record_pg<0, LPE>(p, rec, pout, block, st);
    }
}

// The record projection with the schedule of the flags (lanes per event,
// block size), as launch<kProj> takes it; into the records' value slots
// (1/p), or with pout the projection itself into pout.
int launch_record_project(const TorParams &p, float4 *rec, uint32_t flags, hipStream_t st,
                          float *pout = nullptr)
{
    if (p.n == 0)
        return SRFX_OK;
    static const int kBlock[4] = {256, 64, 128, 256};
    const int block = kBlock[(flags >> SRFX_BLOCK_SHIFT) & 3u];
    // no lanes-per-event flag: four lanes per event, the fastest measured
    switch ((flags >> SRFX_LPE_SHIFT) & 3u) {
    case 1:
        return launch_record_lpe<1>(p, rec, pout, block, st);
    case 3:
        return launch_record_lpe<16>(p, rec, pout, block, st);
    default:
        return launch_record_lpe<4>(p, rec, pout, block, st);
    }
}

// Count, scan (host: per tile; device: per event when slots are kept), fill,
// and without slot tables the listing sort.
int plan_build(srfx_tor_plan_s *pl, hipStream_t st, int chunk, bool keep_slots)
{
    TorParams &p = pl->p;
    const TileGeom &g = pl->g;
    const int64_t n = p.n;
    const int ntiles = g.nslab * g.nwx * g.nwy;
    int *cnt_d = nullptr, *ev_cnt = nullptr, *tile_off_d = nullptr;
    void *tmp = nullptr;
    int rc = SRFX_OK;
    std::vector<int> cnt(ntiles), off(ntiles);
    std::vector<TileChunk> ch;
    int64_t total = 0;
    size_t tmp_bytes = 0;
    int last_cnt = 0, last_off = 0, mismatch = 0, total32 = 0, other = 0;
    PlanArgs a{};
    // cnt_d: tile counts [ntiles], tile fill [ntiles], error count [1], the
    // flag of an event that needs another walk than the NB x (NB - 1) box [1]
    if ((rc = dev_alloc(&cnt_d, 2 * (size_t)ntiles + 2)) ||
        (rc = dev_alloc(&pl->rec, (size_t)kWalkF4 * n)) ||
        (rc = dev_alloc(&tile_off_d, ntiles)) || (rc = dev_alloc(&pl->proj, n)))
        goto done;
    if (keep_slots && ((rc = dev_alloc(&ev_cnt, n)) || (rc = dev_alloc(&pl->ev_off, n + 1))))
        goto done;
    hipLaunchKernelGGL(fill_int, dim3(blocks_for(2 * (int64_t)ntiles + 2, 256)), dim3(256), 0,
                       st, cnt_d, 2 * (int64_t)ntiles + 2, 0);
    a.cols7 = pl->cols7;
    a.ntiles = ntiles;
    a.tile_cnt = cnt_d;
    a.ev_cnt = ev_cnt;
    a.rec = pl->rec;
    a.other = cnt_d + 2 * ntiles + 1;
    if (n > 0)
        launch_pass<false>(p, g, a, st);
    if ((rc = launch_status()))
        goto done;
    if (keep_slots && n > 0) {
        // ev_off = exclusive scan of the per-event listing counts
        if (rocprim::exclusive_scan(nullptr, tmp_bytes, ev_cnt, pl->ev_off, 0, (size_t)n,
                                    rocprim::plus<int>(), st) != hipSuccess) {
            rc = SRFX_ELAUNCH;
            goto done;
        }
        if ((rc = dev_alloc((char **)&tmp, tmp_bytes)))
            goto done;
        if (rocprim::exclusive_scan(tmp, tmp_bytes, ev_cnt, pl->ev_off, 0, (size_t)n,
                                    rocprim::plus<int>(), st) != hipSuccess ||
            hipMemcpyAsync(&last_cnt, ev_cnt + (n - 1), sizeof(int), hipMemcpyDeviceToHost,
                           st) != hipSuccess ||
            hipMemcpyAsync(&last_off, pl->ev_off + (n - 1), sizeof(int), hipMemcpyDeviceToHost,
                           st) != hipSuccess) {
            rc = SRFX_ELAUNCH;
            goto done;
        }
    }
    if (hipMemcpyAsync(cnt.data(), cnt_d, sizeof(int) * ntiles, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipMemcpyAsync(&other, cnt_d + 2 * ntiles + 1, sizeof(int), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        rc = SRFX_ELAUNCH;
        goto done;
    }
    pl->one = p.box > 0 && !p.full_patch && other == 0;
    for (int t = 0; t < ntiles; t++) {
        off[t] = (int)total;
        total += cnt[t];
        if (total > (int64_t)INT32_MAX - 1)
            break;
    }
    if (total > (int64_t)INT32_MAX - 1 ||
        (keep_slots && n > 0 && (int64_t)last_off + last_cnt != total)) {
        rc = SRFX_EUNSUPPORTED;
        goto done;
    }
    total32 = (int)total;
    // chunks window by window, each window's slabs in order: the listings of
    // one event in consecutive slabs mostly share a window, so its record is
    // read again while the L2 still holds it
    {
        const int nw = g.nwx * g.nwy;
        for (int w = 0; w < nw; w++)
            for (int t = w; t < ntiles; t += nw)
                for (int s = 0; s < cnt[t]; s += chunk)
                    ch.push_back(
                        TileChunk{t, off[t] + s, cnt[t] - s < chunk ? cnt[t] - s : chunk, 0});
    }
    pl->nlist = total;
    pl->nchunks = (int)ch.size();
    pl->hchunks = ch;
    if ((rc = dev_alloc(&pl->list, total)) || (rc = dev_alloc(&pl->chunks, ch.size())))
        goto done;
    if (keep_slots && ((rc = dev_alloc(&pl->ev_slot, total)) ||
                       (rc = dev_alloc(&pl->partial, total))))
        goto done;
    if (hipMemcpyAsync(tile_off_d, off.data(), sizeof(int) * ntiles, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        (!ch.empty() && hipMemcpyAsync(pl->chunks, ch.data(), sizeof(TileChunk) * ch.size(),
                                       hipMemcpyHostToDevice, st) != hipSuccess) ||
        (keep_slots && hipMemcpyAsync(pl->ev_off + n, &total32, sizeof(int),
                                      hipMemcpyHostToDevice, st) != hipSuccess)) {
        rc = SRFX_ELAUNCH;
        goto done;
    }
    a.ev_cnt = nullptr;
    a.ev_off = keep_slots ? pl->ev_off : nullptr;
    a.tile_off = tile_off_d;
    a.tile_fill = cnt_d + ntiles;
    a.list = pl->list;
    a.ev_slot = keep_slots ? pl->ev_slot : nullptr;
    a.err = cnt_d + 2 * ntiles;
    // without the tiled projection's slot tables each list entry also carries
    // its listing's slices where the event index and the slab fit its fields
    pl->g.spans = !keep_slots && g.S <= kSpanMaxSlab && n < ((int64_t)1 << kSpanEventBits);
    if (n > 0)
        launch_pass<true>(p, g, a, st);
    if ((rc = launch_status()))
        goto done;
    if (hipMemcpyAsync(&mismatch, cnt_d + 2 * ntiles, sizeof(int), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        rc = SRFX_ELAUNCH;
        goto done;
    }
    if (mismatch) {
        rc = SRFX_ELAUNCH; // count and fill passes disagreed: no valid plan
        goto done;
    }
    // without the tiled projection's slot tables the listings of a chunk may
    // be reordered: longest walks first
    if (!keep_slots && pl->nchunks > 0 && chunk <= 16 * 256) {
        hipLaunchKernelGGL(sort_chunks, dim3(pl->nchunks), dim3(256),
                           sizeof(int) * (33 + chunk), st, pl->p, pl->g, pl->chunks, pl->list,
                           pl->rec);
        if ((rc = launch_status()))
            goto done;
        pl->g.sorted = 1;
    }
done:
    // host buffers above are read by the copies: finish them before returning
    (void)hipStreamSynchronize(st);
    (void)hipFree(cnt_d);
    (void)hipFree(ev_cnt);
    (void)hipFree(tile_off_d);
    (void)hipFree(tmp);
    return rc;
}

size_t align256(size_t b)
{
    return (b + 255) & ~(size_t)255;
}

// Whether plans of this tile shape and event count can be built as a set
// (their list entries carry spans).
bool set_buildable(const TileGeom &g, int64_t n, bool keep_slots)
{
    return !keep_slots && g.S <= kSpanMaxSlab && n < ((int64_t)1 << kSpanEventBits);
}

// One set of plans built together: plan k of pls[0..K) holds events
// first[k] .. first[k + 1] - 1 of the array at lors (row stride ld), and
// pls[k]->p is already set up for its own events.
struct PlanSet {
    srfx_tor_plan_s *const *pls;
    int K;
    const int64_t *first; // K + 1
    const float *lors;
    int chunk;
};

// Build every set of `sets` in shared phases: the count passes of all sets,
// one host synchronisation, the fill passes of all sets, one host
// synchronisation, then the chunk sorts.  One count pass and one fill pass
// per set cover all its plans; all plans of all sets share two device blocks
// (records and projection buffers; chunks and lists), freed with the last
// plan.  Every plan's entries carry their spans; chunks are sorted.
int plan_build_sets(const std::vector<PlanSet> &sets, hipStream_t st)
{
    const int S = (int)sets.size();
    int nplans = 0;
    for (const PlanSet &ps : sets)
        nplans += ps.K;
    if (nplans == 0)
        return SRFX_OK;
    auto *arena = new srfx_tor_plan_s::Arena;
    arena->mem[0] = arena->mem[1] = nullptr;
    arena->refs.store(nplans);
    for (const PlanSet &ps : sets)
        for (int k = 0; k < ps.K; k++)
            ps.pls[k]->arena = arena;
    // per set: tile geometry, tile count, and where its tables go
    struct Layout {
        TileGeom g;
        int nt;
        size_t head;  // offset of first [K + 1], record pointers [K], list pointers [K]
        size_t ctr;   // offset (ints) of cnt [K nt], fill [K nt], off [K nt]
        std::vector<int64_t> totals;
        std::vector<int> nch;
        size_t ch0;   // first chunk of the set in the chunk table
    };
    std::vector<Layout> L(S);
    // block 0: records and projection buffers of every plan
    size_t b0 = 0, head_b = 0, ctr_n = 0;
    std::vector<size_t> rec_off(nplans), proj_off(nplans);
    for (int s = 0, j = 0; s < S; s++) {
        const PlanSet &ps = sets[s];
        Layout &l = L[s];
        l.g = ps.pls[0]->g;
        l.g.spans = 1;
        l.nt = l.g.nslab * l.g.nwx * l.g.nwy;
        for (int k = 0; k < ps.K; k++, j++) {
            const size_t n = (size_t)std::max<int64_t>(ps.first[k + 1] - ps.first[k], 1);
            rec_off[j] = b0;
            b0 += align256(sizeof(float4) * kWalkF4 * n);
            proj_off[j] = b0;
            b0 += align256(sizeof(float) * n);
        }
        l.head = head_b;
        head_b += align256(sizeof(int64_t) * (ps.K + 1) + 2 * sizeof(void *) * ps.K);
        l.ctr = ctr_n;
        ctr_n += (size_t)3 * ps.K * l.nt;
    }
    // temporaries: the sets' heads, their counters, one error count per set,
    // per plan the flag of an event that needs another walk than the
    // NB x (NB - 1) box, the chunk sort's jobs
    const size_t ctr_b = align256(sizeof(int) * std::max<size_t>(ctr_n, 1));
    const size_t sjobs_at = align256(head_b + ctr_b + sizeof(int) * (S + nplans));
    const size_t tmp_b = sjobs_at + sizeof(SortJob) * nplans;
    char *tmp = nullptr, *blk0 = nullptr, *blk1 = nullptr;
    std::vector<char> head(head_b);
    std::vector<int> ctr(ctr_n);
    std::vector<int> errs(S, 0), others(nplans, 0);
    std::vector<SortJob> sjobs;
    int64_t sblocks = 0;
    int max_chunk = 0; // largest chunk of the sorted plans (the sort's LDS)
    std::vector<TileChunk> ch;
    int rc = SRFX_OK;
    int *ctr_d = nullptr, *err_d = nullptr, *other_d = nullptr;
    std::vector<int> plan0(S + 1, 0); // first plan of each set
    for (int s = 0; s < S; s++)
        plan0[s + 1] = plan0[s] + sets[s].K;
    auto args = [&](int s) {
        const PlanSet &ps = sets[s];
        const Layout &l = L[s];
        PlanSetArgs a{};
        a.cols7 = ps.pls[0]->cols7;
        a.ntiles = l.nt;
        a.nplans = ps.K;
        a.lds = 3u * sizeof(int) * l.nt <= 65536u;
        // the longest listing queues that keep three workgroups of the fill
        // pass per CU (with fewer, at C5's 2816 tiles, queues of 959 entries
        // in 64 KB: 115 ms against 101 ms without queues)
        a.queue = 0;
        const unsigned qbudget = kLdsPerCU / 3u;
        if (l.nt < (1 << 24) &&
            fill_queue_lds(kFillQueueMin) + 3u * sizeof(int) * l.nt <= qbudget)
            a.queue = std::min<int>(kFillQueue, (int)((qbudget - 3u * sizeof(int) * l.nt -
                                                       fill_queue_lds(0)) /
                                                      (4u * sizeof(int))));
        a.first = (const int64_t *)(tmp + l.head);
        a.rec = (float4 *const *)(tmp + l.head + sizeof(int64_t) * (ps.K + 1));
        a.list = (int *const *)(tmp + l.head + sizeof(int64_t) * (ps.K + 1) +
                                sizeof(void *) * ps.K);
        a.cnt = ctr_d + l.ctr;
        a.fill = a.cnt + (size_t)ps.K * l.nt;
        a.off = a.fill + (size_t)ps.K * l.nt;
        a.err = err_d + s;
        a.other = other_d + plan0[s];
        return a;
    };
    auto launch = [&](bool fill) {
        for (int s = 0; s < S; s++) {
            const PlanSet &ps = sets[s];
            const int64_t n = ps.first[ps.K];
            if (n <= 0)
                continue;
            const PlanSetArgs a = args(s);
            TorParams p = ps.pls[0]->p;
            p.lors = ps.lors;
            const unsigned lds = (a.lds ? (fill ? 3u : 1u) * sizeof(int) * L[s].nt : 0u) +
                                 (fill && a.queue ? fill_queue_lds(a.queue) : 0u);
            if (fill)
                hipLaunchKernelGGL(plan_set_pass<true>, dim3(blocks_for(n, 256)), dim3(256), lds,
                                   st, p, L[s].g, a);
            else
                hipLaunchKernelGGL(plan_set_pass<false>, dim3(blocks_for(n, 256)), dim3(256), lds,
                                   st, p, L[s].g, a);
        }
        return launch_status();
    };
    if ((rc = dev_alloc(&blk0, b0)) || (rc = dev_alloc(&tmp, tmp_b)))
        goto done;
    arena->mem[0] = blk0;
    ctr_d = (int *)(tmp + head_b);
    err_d = (int *)(tmp + head_b + ctr_b);
    other_d = err_d + S;
    for (int s = 0, j = 0; s < S; s++) {
        const PlanSet &ps = sets[s];
        char *h = head.data() + L[s].head;
        std::memcpy(h, ps.first, sizeof(int64_t) * (ps.K + 1));
        auto **rp = reinterpret_cast<void **>(h + sizeof(int64_t) * (ps.K + 1));
        for (int k = 0; k < ps.K; k++, j++) {
            ps.pls[k]->rec = (float4 *)(blk0 + rec_off[j]);
            ps.pls[k]->proj = (float *)(blk0 + proj_off[j]);
            rp[k] = ps.pls[k]->rec;
        }
    }
    {
        // the counters, the error counts and the other-walk flags behind them
        const int64_t nz = (int64_t)(ctr_b / sizeof(int) + S + nplans);
        hipLaunchKernelGGL(fill_int, dim3(blocks_for(nz, 256)), dim3(256), 0, st, ctr_d, nz, 0);
    }
    if (hipMemcpyAsync(tmp, head.data(), head_b, hipMemcpyHostToDevice, st) != hipSuccess) {
        rc = SRFX_ELAUNCH;
        goto done;
    }
    if ((rc = launch(false)))
        goto done;
    if ((ctr_n > 0 && hipMemcpyAsync(ctr.data(), ctr_d, sizeof(int) * ctr_n,
                                     hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipMemcpyAsync(others.data(), other_d, sizeof(int) * nplans, hipMemcpyDeviceToHost,
                       st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        rc = SRFX_ELAUNCH;
        goto done;
    }
    for (int s = 0; s < S; s++)
        for (int k = 0; k < sets[s].K; k++) {
            srfx_tor_plan_s *pl = sets[s].pls[k];
            pl->one = pl->p.box > 0 && !pl->p.full_patch && others[plan0[s] + k] == 0;
        }
    // per plan: slot offsets per tile (into the counters' off part), then
    // chunks window by window, each window's slabs in order (an event's
    // listings in consecutive slabs mostly share a window, so its record is
    // read again while the L2 holds it)
    for (int s = 0; s < S; s++) {
        const PlanSet &ps = sets[s];
        Layout &l = L[s];
        const int nt = l.nt, nw = l.g.nwx * l.g.nwy;
        l.totals.assign(ps.K, 0);
        l.nch.assign(ps.K, 0);
        l.ch0 = ch.size();
        for (int k = 0; k < ps.K; k++) {
            const int *c = ctr.data() + l.ctr + (size_t)k * nt;
            int *o = ctr.data() + l.ctr + (size_t)(2 * ps.K + k) * nt;
            int64_t tot = 0;
            for (int t = 0; t < nt; t++) {
                o[t] = (int)tot;
                tot += c[t];
                if (tot > (int64_t)INT32_MAX - 1) {
                    rc = SRFX_EUNSUPPORTED;
                    goto done;
                }
            }
            l.totals[k] = tot;
            const size_t before = ch.size();
            for (int w = 0; w < nw; w++)
                for (int t = w; t < nt; t += nw)
                    for (int q = 0; q < c[t]; q += ps.chunk)
                        ch.push_back(TileChunk{t, o[t] + q,
                                               c[t] - q < ps.chunk ? c[t] - q : ps.chunk, 0});
            l.nch[k] = (int)(ch.size() - before);
        }
    }
    {
        // block 1: every plan's chunks, contiguous, then the lists
        const size_t chunk_b = align256(sizeof(TileChunk) * std::max<size_t>(ch.size(), 1));
        size_t b1 = chunk_b;
        std::vector<size_t> list_off(nplans);
        for (int s = 0, j = 0; s < S; s++)
            for (int k = 0; k < sets[s].K; k++, j++) {
                list_off[j] = b1;
                b1 += align256(sizeof(int) * (size_t)std::max<int64_t>(L[s].totals[k], 1));
            }
        if ((rc = dev_alloc(&blk1, b1)))
            goto done;
        arena->mem[1] = blk1;
        for (int s = 0, j = 0; s < S; s++) {
            const PlanSet &ps = sets[s];
            const Layout &l = L[s];
            char *h = head.data() + l.head;
            auto **lp = reinterpret_cast<void **>(h + sizeof(int64_t) * (ps.K + 1) +
                                                  sizeof(void *) * ps.K);
            size_t c0 = l.ch0;
            for (int k = 0; k < ps.K; k++, j++) {
                srfx_tor_plan_s *pl = ps.pls[k];
                pl->chunks = (TileChunk *)blk1 + c0;
                pl->nchunks = l.nch[k];
                pl->hchunks.assign(ch.begin() + c0, ch.begin() + c0 + l.nch[k]);
                c0 += l.nch[k];
                pl->list = (int *)(blk1 + list_off[j]);
                pl->nlist = l.totals[k];
                pl->g = l.g;
                lp[k] = pl->list;
            }
        }
    }
    if ((ctr_n > 0 && hipMemcpyAsync(ctr_d, ctr.data(), sizeof(int) * ctr_n,
                                     hipMemcpyHostToDevice, st) != hipSuccess) ||
        hipMemcpyAsync(tmp, head.data(), head_b, hipMemcpyHostToDevice, st) != hipSuccess) {
        rc = SRFX_ELAUNCH;
        goto done;
    }
    if ((rc = launch(true)))
        goto done;
    // the chunk table, read by the sort only: its copy from pageable memory
    // is staged while the fill pass runs
    if (!ch.empty() && hipMemcpyAsync(blk1, ch.data(), sizeof(TileChunk) * ch.size(),
                                      hipMemcpyHostToDevice, st) != hipSuccess) {
        rc = SRFX_ELAUNCH;
        goto done;
    }
    // listings of each chunk ordered by slice count, longest first: one
    // launch for the chunks of every plan.  It runs before the fill's error
    // count is read (one host synchronisation less): it only permutes the
    // entries inside each chunk's range, whatever they hold.
    for (const PlanSet &ps : sets) {
        if (ps.chunk > 16 * 256)
            continue;
        for (int k = 0; k < ps.K; k++) {
            srfx_tor_plan_s *pl = ps.pls[k];
            if (pl->nchunks > 0) {
                sjobs.push_back(SortJob{pl->chunks, pl->list, sblocks});
                sblocks += pl->nchunks;
                max_chunk = std::max(max_chunk, ps.chunk);
            }
            pl->g.sorted = 1;
        }
    }
    if (!sjobs.empty()) {
        if (hipMemcpyAsync(tmp + sjobs_at, sjobs.data(), sizeof(SortJob) * sjobs.size(),
                           hipMemcpyHostToDevice, st) != hipSuccess) {
            rc = SRFX_ELAUNCH;
            goto done;
        }
        hipLaunchKernelGGL(sort_chunks_many, dim3((unsigned)sblocks), dim3(256),
                           sizeof(int) * (33 + max_chunk), st, (const SortJob *)(tmp + sjobs_at),
                           (int)sjobs.size());
    }
    if ((rc = launch_status()))
        goto done;
    if (hipMemcpyAsync(errs.data(), err_d, sizeof(int) * S, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        rc = SRFX_ELAUNCH;
        goto done;
    }
    for (int s = 0; s < S; s++)
        if (errs[s]) {
            rc = SRFX_ELAUNCH; // count and fill passes disagreed: no valid plan
            goto done;
        }
done:
    // host buffers above are read by the copies: finish them before returning
    (void)hipStreamSynchronize(st);
    (void)hipFree(tmp);
    return rc;
}

int plan_build_set(srfx_tor_plan_s *const *pls, int K, const int64_t *first, const float *lors,
                   int chunk, hipStream_t st)
{
    return plan_build_sets({PlanSet{pls, K, first, lors, chunk}}, st);
}

template <int MODE, bool VREC, bool YF>
void launch_scatter_l(const TorParams &p, const srfx_tor_plan_s *pl, dim3 grid, dim3 block,
                      unsigned sh, hipStream_t st)
{
    switch (p.box) {
    case 3:
        // plans whose events all walk the 3 x 2 box, with tiles that let five
        // workgroups share a CU's LDS: the single-walk kernel
        if constexpr (YF) {
            if (pl->one && 5u * sh <= kLdsPerCU) {
                hipLaunchKernelGGL((tile_scatter<MODE, 3, VREC, true, true>), grid, block, sh, st,
                                   p, pl->g, pl->chunks, pl->list, pl->rec);
                break;
            }
        }
        hipLaunchKernelGGL((tile_scatter<MODE, 3, VREC, YF>), grid, block, sh, st, p, pl->g,
                           pl->chunks, pl->list, pl->rec);
        break;
    case 4:
        // the 4 x 3 box walk alone: four waves per SIMD instead of three
        if constexpr (YF) {
            if (pl->one && 4u * sh <= kLdsPerCU) {
                hipLaunchKernelGGL((tile_scatter<MODE, 4, VREC, true, true>), grid, block, sh, st,
                                   p, pl->g, pl->chunks, pl->list, pl->rec);
                break;
            }
        }
        hipLaunchKernelGGL((tile_scatter<MODE, 4, VREC, YF>), grid, block, sh, st, p, pl->g,
                           pl->chunks, pl->list, pl->rec);
        break;
    default:
        hipLaunchKernelGGL((tile_scatter<MODE, 0, VREC, YF>), grid, block, sh, st, p, pl->g,
                           pl->chunks, pl->list, pl->rec);
        break;
    }
}

// The LDS tile runs fastest along the output's smaller stride (op y for the
// world layout of the axis groups, op x for a contiguous op frame), so the
// flush's atomics are coalesced.
template <int MODE, bool VREC>
void launch_scatter(const TorParams &p, const srfx_tor_plan_s *pl, dim3 grid, dim3 block,
                    unsigned sh, hipStream_t st)
{
    if (flush_y_fastest(p))
        launch_scatter_l<MODE, VREC, true>(p, pl, grid, block, sh, st);
    else
        launch_scatter_l<MODE, VREC, false>(p, pl, grid, block, sh, st);
}

int plan_scatter(srfx_tor_plan_s *pl, int mode, const float *vals, float *image,
                 const srfx_grid *out_grid, uint32_t flags, hipStream_t st)
{
    TorParams p = pl->p;
    p.val = vals;
    p.out = image;
    if (out_grid) {
        Strides so = grid_strides(*out_grid);
        p.ox = (int)so.x;
        p.oy = (int)so.y;
        p.oz = (int)so.z;
    }
    if (pl->nchunks == 0)
        return SRFX_OK;
    const dim3 grid(pl->nchunks), block(pl->block);
    const unsigned sh = lds_bytes_scatter(pl->g);
    switch (mode) {
    case kBack:
        if (vals)
            launch_scatter<kBack, false>(p, pl, grid, block, sh, st);
        else
            launch_scatter<kBack, true>(p, pl, grid, block, sh, st);
        break;
    case kAdj:
        launch_scatter<kAdj, false>(p, pl, grid, block, sh, st);
        break;
    default:
        launch_scatter<kMap, false>(p, pl, grid, block, sh, st);
        break;
    }
    (void)flags;
    return launch_status();
}

int plan_project(srfx_tor_plan_s *pl, const float *image, float *proj, hipStream_t st)
{
    TorParams p = pl->p;
    p.img = image;
    if (!pl->ev_slot)
        return SRFX_EUNSUPPORTED; // plan built without the tiled projection tables
    if (p.n == 0)
        return SRFX_OK;
    if (pl->nchunks > 0) {
        const dim3 grid(pl->nchunks), block(pl->block);
        const unsigned sh = lds_bytes(pl->g);
        switch (p.box) {
        case 3:
            hipLaunchKernelGGL(tile_gather<3>, grid, block, sh, st, p, pl->g, pl->chunks,
                               pl->list, pl->rec, pl->partial);
            break;
        case 4:
            hipLaunchKernelGGL(tile_gather<4>, grid, block, sh, st, p, pl->g, pl->chunks,
                               pl->list, pl->rec, pl->partial);
            break;
        default:
            hipLaunchKernelGGL(tile_gather<0>, grid, block, sh, st, p, pl->g, pl->chunks,
                               pl->list, pl->rec, pl->partial);
            break;
        }
    }
    hipLaunchKernelGGL(tile_reduce, dim3(blocks_for(p.n, 256)), dim3(256), 0, st, pl->partial,
                       pl->ev_off, pl->ev_slot, p.n, proj);
    return launch_status();
}

} // namespace

namespace {

// A plan object for n events of one group frame, set up but not built.
int plan_new(const float *lors, int64_t n, int64_t ld, int32_t ncols, const srfx_grid &grid,
             const srfx_tor_model &m, const srfx_tile &tile, srfx_tor_plan_s **out)
{
    *out = nullptr;
    if (ncols != 7 && ncols != 10)
        return SRFX_EINVAL;
    if (n > (int64_t)INT32_MAX - 1)
        return SRFX_EUNSUPPORTED;
    if (tile.slices < 1 || tile.wx < 1 || tile.wy < 1 || tile.chunk < 1)
        return SRFX_EINVAL;
    if ((int64_t)tile.slices * tile.wx * tile.wy > 8192)
        return SRFX_EUNSUPPORTED; // FP64 LDS tile above 64 KB
    if (tile.block != 0 && tile.block != 64 && tile.block != 128 && tile.block != 256)
        return SRFX_EINVAL;
    srfx_tor_plan_s *pl = new srfx_tor_plan_s{};
    srfx_tor_model mm = ncols == 7 ? model_no_tof(m.kernel_width, m.geo_sigma2_flag) : m;
    int rc = setup(pl->p, lors, n, ld, grid, mm.kernel_width, mm.geo_sigma2_flag, mm.tof_bin,
                   mm.tof_sigma2);
    if (rc) {
        delete pl;
        return rc;
    }
    pl->grid = grid;
    pl->cols7 = ncols == 7;
    pl->block = tile.block ? tile.block : 256;
    TileGeom &g = pl->g;
    g.S = tile.slices;
    g.Wx = tile.wx;
    g.Wy = tile.wy;
    g.nwx = (pl->p.gx + g.Wx - 1) / g.Wx;
    g.nwy = (pl->p.gy + g.Wy - 1) / g.Wy;
    g.nslab = (pl->p.gz + g.S - 1) / g.S;
    if ((int64_t)g.nwx * g.nwy * g.nslab > (int64_t)INT32_MAX / 2) {
        delete pl;
        return SRFX_EUNSUPPORTED;
    }
    *out = pl;
    return SRFX_OK;
}

} // namespace

extern "C" int srfx_tor_plan_create(const float *lors, int64_t n, int64_t ld, int32_t ncols,
                                    srfx_grid grid, srfx_tor_model m, srfx_tile tile,
                                    srfx_tor_plan *out, void *stream)
{
    if (!out)
        return SRFX_EINVAL;
    srfx_tor_plan_s *pl = nullptr;
    int rc = plan_new(lors, n, ld, ncols, grid, m, tile, &pl);
    if (rc)
        return rc;
    const bool keep_slots = tile.projection != 0;
    if (set_buildable(pl->g, n, keep_slots)) {
        const int64_t first[2] = {0, n};
        rc = plan_build_set(&pl, 1, first, lors, tile.chunk, (hipStream_t)stream);
    } else {
        rc = plan_build(pl, (hipStream_t)stream, tile.chunk, keep_slots);
    }
    if (rc) {
        plan_free(pl);
        return rc;
    }
    *out = pl;
    return SRFX_OK;
}

namespace srfx {

int plan_create_subsets_many(const SubsetRequest *req, int nreq, int32_t ncols,
                             const srfx_tor_model &m, srfx_tor_plan *plans, hipStream_t st)
{
    if (!plans || (nreq > 0 && !req) || nreq < 0)
        return SRFX_EINVAL;
    int total = 0;
    for (int r = 0; r < nreq; r++) {
        const SubsetRequest &q = req[r];
        if (!q.bounds || q.nplans < 1 || q.n < 0)
            return SRFX_EINVAL;
        total += q.nplans;
    }
    for (int i = 0; i < total; i++)
        plans[i] = nullptr;
    for (int r = 0; r < nreq; r++) {
        const SubsetRequest &q = req[r];
        if (q.bounds[0] < 0 || q.bounds[q.nplans] > q.n)
            return SRFX_EINVAL;
        for (int k = 0; k < q.nplans; k++)
            if (q.bounds[k + 1] < q.bounds[k])
                return SRFX_EINVAL;
    }
    std::vector<srfx_tor_plan_s *> pls(total, nullptr);
    std::vector<std::vector<int64_t>> first(nreq);
    int rc = SRFX_OK;
    bool together = true;
    for (int r = 0, i = 0; r < nreq && !rc; r++) {
        const SubsetRequest &q = req[r];
        together = together && q.tile.projection == 0;
        first[r].resize(q.nplans + 1);
        for (int k = 0; k < q.nplans && !rc; k++, i++) {
            const int64_t nk = q.bounds[k + 1] - q.bounds[k];
            rc = plan_new(q.lors ? q.lors + q.bounds[k] : nullptr, nk, q.ld, ncols, q.grid, m,
                          q.tile, &pls[i]);
            if (!rc)
                together = together && set_buildable(pls[i]->g, nk, false);
            first[r][k] = q.bounds[k] - q.bounds[0];
        }
        first[r][q.nplans] = q.bounds[q.nplans] - q.bounds[0];
    }
    if (!rc) {
        if (together) {
            std::vector<PlanSet> sets;
            for (int r = 0, i = 0; r < nreq; i += req[r].nplans, r++)
                sets.push_back(PlanSet{pls.data() + i, req[r].nplans, first[r].data(),
                                       req[r].lors ? req[r].lors + req[r].bounds[0] : nullptr,
                                       req[r].tile.chunk});
            rc = plan_build_sets(sets, st);
        } else {
            for (int r = 0, i = 0; r < nreq && !rc; r++)
                for (int k = 0; k < req[r].nplans && !rc; k++, i++)
                    rc = plan_build(pls[i], st, req[r].tile.chunk, req[r].tile.projection != 0);
        }
    }
    if (rc) {
        for (srfx_tor_plan_s *pl : pls)
            plan_free(pl);
        return rc;
    }
    for (int i = 0; i < total; i++)
        plans[i] = pls[i];
    return SRFX_OK;
}

} // namespace srfx

extern "C" int srfx_tor_plan_create_subsets(const float *lors, int64_t n, int64_t ld,
                                            int32_t ncols, srfx_grid grid, srfx_tor_model m,
                                            srfx_tile tile, int32_t nplans, const int64_t *bounds,
                                            srfx_tor_plan *plans, void *stream)
{
    if (!plans || !bounds || nplans < 1 || n < 0)
        return SRFX_EINVAL;
    const SubsetRequest q{lors, n, ld, grid, tile, nplans, bounds};
    return plan_create_subsets_many(&q, 1, ncols, m, plans, (hipStream_t)stream);
}

extern "C" int srfx_tor_plan_create_subset_groups(const srfx_subset_group *groups,
                                                  int32_t ngroups, int32_t ncols,
                                                  srfx_tor_model m, srfx_tor_plan *plans,
                                                  void *stream)
{
    if (!plans || ngroups < 0 || (ngroups > 0 && !groups))
        return SRFX_EINVAL;
    std::vector<SubsetRequest> req;
    for (int r = 0; r < ngroups; r++) {
        const srfx_subset_group &q = groups[r];
        if (!q.bounds || q.nplans < 1 || q.n < 0 || q.tile.projection != 0)
            return SRFX_EINVAL;
        req.push_back(SubsetRequest{q.lors, q.n, q.ld, q.grid, q.tile, q.nplans, q.bounds});
    }
    return plan_create_subsets_many(req.data(), ngroups, ncols, m, plans, (hipStream_t)stream);
}

extern "C" int srfx_tor_plan_info(srfx_tor_plan plan, int64_t *nlistings, int32_t *nchunks)
{
    if (!plan)
        return SRFX_EINVAL;
    if (nlistings)
        *nlistings = plan->nlist;
    if (nchunks)
        *nchunks = plan->nchunks;
    return SRFX_OK;
}

extern "C" int srfx_tor_plan_walks(srfx_tor_plan plan, int32_t *box, int32_t *single_walk)
{
    if (!plan)
        return SRFX_EINVAL;
    if (box)
        *box = plan->p.full_patch ? 0 : plan->p.box;
    if (single_walk)
        *single_walk = plan->one;
    return SRFX_OK;
}

extern "C" int srfx_tor_plan_destroy(srfx_tor_plan plan)
{
    if (plan)
        (void)hipDeviceSynchronize();
    plan_free(plan);
    return SRFX_OK;
}

extern "C" int srfx_tor_plan_project(srfx_tor_plan plan, const float *image, float *proj,
                                     uint32_t flags, void *stream)
{
    if (!plan || plan->cols7 || (plan->p.n > 0 && (!image || !proj)))
        return SRFX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (plan->ev_slot && !(flags & SRFX_IMAGE_OP_FRAME))
        return plan_project(plan, image, proj, st);
    // a plan without the tiled gather's tables: the projection from the walk
    // records (the first kernel of the plan's OSEM step), with the flags'
    // schedule; SRFX_IMAGE_OP_FRAME: image is the op-frame copy
    TorParams q = plan->p;
    q.img = image;
    if (flags & SRFX_IMAGE_OP_FRAME) {
        q.sx = 1;
        q.sy = q.gx;
        q.sz = q.gx * q.gy;
    }
    return launch_record_project(q, plan->rec, flags, st, proj);
}

static int plan_scatter_entry(srfx_tor_plan plan, int mode, const float *vals, float *image,
                              uint32_t flags, void *stream)
{
    if (!plan || !image || (plan->p.n > 0 && !vals))
        return SRFX_EINVAL;
    if ((mode == kMap) != (plan->cols7 != 0))
        return SRFX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    int rc = zero_grid(image, plan->grid, flags, st);
    if (rc)
        return rc;
    return plan_scatter(plan, mode, vals, image, nullptr, flags, st);
}

extern "C" int srfx_tor_plan_backproject(srfx_tor_plan plan, const float *lors_value,
                                         float *image, uint32_t flags, void *stream)
{
    return plan_scatter_entry(plan, kBack, lors_value, image, flags, stream);
}

extern "C" int srfx_tor_plan_backproject_adjoint(srfx_tor_plan plan, const float *g,
                                                 float *image, uint32_t flags, void *stream)
{
    return plan_scatter_entry(plan, kAdj, g, image, flags, stream);
}

extern "C" int srfx_tor_plan_maplors(srfx_tor_plan plan, const float *lors_value, float *image,
                                     uint32_t flags, void *stream)
{
    return plan_scatter_entry(plan, kMap, lors_value, image, flags, stream);
}

namespace {
// out[e] = the value slot of event e's walk record.
__global__ __launch_bounds__(256) void record_values(const float4 *__restrict__ rec, int64_t n,
                                                     float *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n)
        out[e] = reinterpret_cast<const float *>(rec)[4 * kWalkHead * e + kWalkValue];
}
} // namespace

extern "C" int srfx_tor_plan_scales(srfx_tor_plan plan, float *out, void *stream)
{
    if (!plan || plan->cols7 || (plan->p.n > 0 && !out))
        return SRFX_EINVAL;
    if (plan->p.n == 0)
        return SRFX_OK;
    hipLaunchKernelGGL(record_values, dim3(blocks_for(plan->p.n, 256)), dim3(256), 0,
                       (hipStream_t)stream, plan->rec, plan->p.n, out);
    return launch_status();
}

extern "C" int srfx_tor_plan_osem_subset(srfx_tor_plan plan, const float *image,
                                         srfx_grid bp_grid, float *bp, uint32_t flags,
                                         void *stream)
{
    if (!plan || plan->cols7 || !bp || (plan->p.n > 0 && !image) || !grid_ok(bp_grid))
        return SRFX_EINVAL;
    const TorParams &p = plan->p;
    if (bp_grid.shape[0] != p.gx || bp_grid.shape[1] != p.gy || bp_grid.shape[2] != p.gz)
        return SRFX_EINVAL;
    const bool proj_only = flags & SRFX_OSEM_PROJECT_ONLY;
    const bool scat_only = flags & SRFX_OSEM_SCATTER_ONLY;
    if ((proj_only && scat_only) || ((proj_only || scat_only) && (flags & SRFX_TILED_PROJECTION)))
        return SRFX_EINVAL;
    if ((flags & SRFX_IMAGE_OP_FRAME) && (flags & SRFX_TILED_PROJECTION))
        return SRFX_EUNSUPPORTED; // the tiled gather stages tiles with the plan's strides
    hipStream_t st = (hipStream_t)stream;
    int rc = proj_only ? SRFX_OK : zero_grid(bp, bp_grid, flags, st);
    if (rc)
        return rc;
    if (p.n == 0)
        return SRFX_OK;
    // Projection with the chosen schedule (untiled by default), then the
    // tiled scatter of 1 / p over events with p > 1e-7.
    if (flags & SRFX_TILED_PROJECTION) {
        rc = plan_project(plan, image, plan->proj, st);
        if (rc)
            return rc;
        return plan_scatter(plan, kBack, plan->proj, bp, &bp_grid, flags, st);
    }
    // The projection goes straight into the value slot of each event's walk
    // record, which the scatter loads anyway (no separate value gather).
    if (!scat_only) {
        TorParams q = p;
        q.img = image;
        if (flags & SRFX_IMAGE_OP_FRAME) {
            q.sx = 1;
            q.sy = p.gx;
            q.sz = p.gx * p.gy;
        }
        rc = launch_record_project(q, plan->rec, flags, st);
        if (rc || proj_only)
            return rc;
    }
    return plan_scatter(plan, kBack, nullptr, bp, &bp_grid, flags, st);
}
