// The step engine of libeqd: advances one scenario (a handle) or an ensemble
// of scenarios.  One step is (src/driver.f90:9-34)
//   step start (clock, stations) + regular elements -> records -> PML elements
//     -> fault pairs | nodes (3-comp, PML)
// with the step start in the first blocks of the regular-element launch and
// the fault pairs on a second stream beside the node kernels.  Every
// launch covers all scenarios that share its kernel variant: the kernels find
// their scenario through a block-to-scenario table and read its arrays from a
// per-scenario descriptor (Member in eqd_device.h), so a 64-scenario ensemble
// costs the same number of launches per step as one scenario.  Step number and
// time live in device memory (Clock), so one step can be captured as a HIP
// graph and replayed for the whole run.  The launch shapes (block size, XCD
// remapping per kind) are chosen by a tuner that times candidates on the
// scenarios themselves and restores their state afterwards.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "eqd_internal.h"

using namespace eqd;

namespace {

const char *const kKindName[kKinds] = {"element", "element_pml", "node", "node_pml", "fault",
                                       "station", "clock", "pending", "pending_pml", "thermop", "record"};

template <typename R>
const RegElems<R> &elems(const eqd_handle *h);
template <>
const RegElems<double> &elems<double>(const eqd_handle *h) { return h->E64; }
template <>
const RegElems<float> &elems<float>(const eqd_handle *h) { return h->E32; }
template <typename R>
const PmlElems<R> &pmls(const eqd_handle *h);
template <>
const PmlElems<double> &pmls<double>(const eqd_handle *h) { return h->P64; }
template <>
const PmlElems<float> &pmls<float>(const eqd_handle *h) { return h->P32; }

template <typename R>
Member<R> make_member(const eqd_engine *e, const eqd_handle *h)
{
    Member<R> M{};
    M.E = elems<R>(h);
    M.P = pmls<R>(h);
    M.N = h->N;
    M.PN = h->PN;
    M.F = h->F;
    if (e->no_harvest) M.F.harvest = nullptr;
    M.S = h->S;
    M.C = h->C;
    M.node3_free = h->d_node3_free;
    M.n_node3_free = (int64_t)h->node3_free.size();
    M.ell3 = h->d_ell3;
    M.node3_all = h->d_node3_all;
    M.n_node3_all = (int64_t)h->node3_all.size();
    M.clk = h->d_clock;
    M.rec = e->no_harvest ? nullptr : h->d_rec;  // the tuner's trial steps record nothing
    return M;
}

int64_t count_of(const eqd_handle *h, int kind)
{
    switch (kind) {
    case kKElem: return h->n_reg;
    case kKPml: return h->n_pml;
    case kKNode3: return (int64_t)h->node3_free.size();
    case kKNodePml: return h->PN.n;
    case kKFault: return h->F.nf;
    case kKStation: return 0;  // the step start: its own launch or blocks of the first element launch
    case kKPend3: return (int64_t)h->node3_all.size();
    case kKPendPml: return h->PN.n;
    case kKTp: return h->F.tph ? h->F.nf * 64 : 0;  // one wavefront per pair
    case kKRecord: return h->d_rec ? h->rec.n_nodes : 0;
    default: return 0;
    }
}

// class tables of an element kind in LDS: both tables present and at most 16 KiB
template <typename X>
size_t class_lds(const X &E, size_t rsize)
{
    if (!E.optab || !E.phitab) return 0;
    const size_t b = rsize * ((size_t)E.n_opc * kOpW + (size_t)E.n_phic * 32);
    return b <= 16384 ? b : 0;
}

// where an element kind's kernel finds its operators (kOps*)
template <typename X>
int ops_of(const X &E, size_t lds)
{
    if (lds) return kOpsLds;
    if (E.opb && E.opb_geom) return kOpsGeom;
    return E.opb && E.opb_compact ? kOpsCompact : kOpsRows;
}

// kernel variant of a scenario in a launch kind, and its dynamic LDS bytes
template <typename R>
int variant(const eqd_handle *h, int kind, size_t *lds)
{
    const eqd_scalars &s = h->s;
    const int q = s.C_Q == 1, dp = s.C_elastic == 0, hg = s.C_hg & 3;
    const int mass = (s.rdampm != 0.0 || s.C_elastic == 0) ? 1 : 0;
    *lds = 0;
    if (kind == kKElem) {
        *lds = class_lds(elems<R>(h), sizeof(R));
        return q | dp << 1 | hg << 2 | mass << 4 | ops_of(elems<R>(h), *lds) << 5;
    }
    if (kind == kKPml) {
        *lds = class_lds(pmls<R>(h), sizeof(R));
        return hg | mass << 2 | ops_of(pmls<R>(h), *lds) << 3;
    }
    if (kind == kKFault) return h->learn ? 1 : 0;
    return 0;
}

int64_t version_sum(const eqd_engine *e)
{
    int64_t v = 0;
    for (const eqd_handle *h : e->hs) v += h->version;
    return v;
}

void drop_graph(eqd_engine *e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

void free_tables(eqd_engine *e)
{
    for (auto &gs : e->groups) {
        for (Group &g : gs)
            if (g.d_bmem) (void)hipFree(g.d_bmem);
        gs.clear();
    }
}

// launch groups and per-scenario descriptors for the current tune
template <typename R>
int build(eqd_engine *e)
{
    drop_graph(e);
    free_tables(e);
    const int n = (int)e->hs.size();
    std::vector<Member<R>> M(n);
    for (int m = 0; m < n; m++) M[m] = make_member<R>(e, e->hs[m]);
    for (int kind = 0; kind < kKinds; kind++) {
        if (kind == kKClock) continue;
        int block = e->tune.block[kind];
        if (kind == kKStation) block = 64;
        std::map<int, Group> by;
        for (int m = 0; m < n; m++) {
            if (count_of(e->hs[m], kind) == 0) continue;
            size_t lds = 0;
            const int sig = variant<R>(e->hs[m], kind, &lds);
            Group &g = by[sig];
            g.sig = sig;
            g.members.push_back(m);
            g.g.lds = std::max(g.g.lds, lds);
        }
        for (auto &kv : by) {
            Group g = kv.second;
            int64_t blocks = 0;
            for (int m : g.members) {
                M[m].blk0[kind] = (int32_t)blocks;
                blocks += (count_of(e->hs[m], kind) + block - 1) / block;
            }
            if (blocks > 0x7fffffff) return fail(e, EQD_EINVAL, "launch %s too large", kKindName[kind]);
            g.g.grid = (unsigned)blocks;
            g.g.block = block;
            g.g.remap = e->tune.remap[kind];
            // without a table every block reads scenario 0: only right when the
            // engine has no other scenario (a group of one scenario m > 0
            // still needs its table)
            if (n > 1) {
                std::vector<uint16_t> t((size_t)blocks);
                for (int m : g.members) {
                    const int64_t b0 = M[m].blk0[kind], nb = (count_of(e->hs[m], kind) + block - 1) / block;
                    for (int64_t b = 0; b < nb; b++) t[(size_t)(b0 + b)] = (uint16_t)m;
                }
                if (hipMalloc(&g.d_bmem, sizeof(uint16_t) * t.size()) != hipSuccess)
                    return fail(e, EQD_EOOM, "block table of %lld entries", (long long)blocks);
                HIPCHK(e, hipMemcpy(g.d_bmem, t.data(), sizeof(uint16_t) * t.size(), hipMemcpyHostToDevice));
            }
            g.g.bmem = g.d_bmem;
            e->groups[kind].push_back(g);
        }
    }
    HIPCHK(e, hipMemcpy(e->d_members, M.data(), sizeof(Member<R>) * n, hipMemcpyHostToDevice));
    e->built = true;
    e->built_version = version_sum(e);
    return EQD_OK;
}

int ensure_built(eqd_engine *e)
{
    if (e->built && e->built_version == version_sum(e)) return EQD_OK;
    return e->fp32 ? build<float>(e) : build<double>(e);
}

// ---------------- per-kind timers (eager runs only)
constexpr size_t kTimerPool = 256;

// The timer events of one launch of a kind in this step: the kind's first
// launch records the start, its last the end, both from the dispatch itself
// (no event packets of their own); null events when not timing.
struct Stamp {
    hipEvent_t a = nullptr, b = nullptr;
    size_t n = 0;  // launches of the kind this step
    hipEvent_t start(size_t i) const { return i == 0 ? a : nullptr; }
    hipEvent_t stop(size_t i) const { return i + 1 == n ? b : nullptr; }
};

Stamp take_timer(eqd_engine *e, int kind, size_t launches)
{
    Stamp p;
    p.n = launches;
    if (!e->sample || launches == 0) return p;
    KTimer &t = e->timers[kKindName[kind]];
    if (t.used == t.ev.size()) {
        hipEvent_t a, b;
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        t.ev.push_back({a, b});
    }
    p.a = t.ev[t.used].first;
    p.b = t.ev[t.used].second;
    t.used++;
    return p;
}

// add the recorded launches' durations to the timers and recycle their events
int drain_timers(eqd_engine *e)
{
    HIPCHK(e, hipStreamSynchronize(e->st));
    HIPCHK(e, hipStreamSynchronize(e->st2));
    for (auto &kv : e->timers) {
        KTimer &t = kv.second;
        for (size_t i = 0; i < t.used; i++) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, t.ev[i].first, t.ev[i].second);
            t.total_ms += ms;
        }
        t.launches += (int64_t)t.used;
        t.used = 0;
    }
    return EQD_OK;
}

// ---------------- one step
template <typename R>
int enqueue_step(eqd_engine *e)
{
    const Member<R> *dM = (const Member<R> *)e->d_members;
    hipStream_t st = e->st;
    // The step start (clock advance and off-fault stations, one block per
    // scenario) rides in the first blocks of the first regular-element
    // launch, which needs no clock itself; a step without regular elements
    // launches it on its own (and so does EQD_STEP_START_LAUNCH=1).
    const auto &ge = e->groups[kKElem];
    const int nstart = ge.empty() || e->step_start_launch ? 0 : (int)e->hs.size();
    if (nstart == 0) {
        const Stamp p = take_timer(e, kKStation, 1);
        HIPCHK(e, launch_step_start<R>(dM, (int)e->hs.size(), st, p.start(0), p.stop(0)));
    }
    {
        const Stamp p = take_timer(e, kKElem, ge.size());
        for (size_t i = 0; i < ge.size(); i++) {
            const Group &g = ge[i];
            HIPCHK(e, launch_elem<R>(dM, g.g, g.sig & 1, (g.sig >> 1) & 1, (g.sig >> 2) & 3, (g.sig >> 4) & 1,
                                     (g.sig >> 5) & 3, i == 0 ? nstart : 0, st, p.start(i), p.stop(i)));
        }
    }
    {
        // records read the advanced clock; the element kernels leave the node
        // state they record untouched
        const auto &gr = e->groups[kKRecord];
        const Stamp p = take_timer(e, kKRecord, gr.size());
        for (size_t i = 0; i < gr.size(); i++) HIPCHK(e, launch_record<R>(dM, gr[i].g, st, p.start(i), p.stop(i)));
    }
    {
        const auto &gs = e->groups[kKPml];
        const Stamp p = take_timer(e, kKPml, gs.size());
        for (size_t i = 0; i < gs.size(); i++) {
            const Group &g = gs[i];
            HIPCHK(e, launch_pml<R>(dM, g.g, g.sig & 3, (g.sig >> 2) & 1, (g.sig >> 3) & 3, st, p.start(i),
                                    p.stop(i)));
        }
    }
    // fork: the fault pairs on the second stream, beside the node kernels
    // (EQD_SINGLE_STREAM=1 keeps them on the main stream, in step order)
    hipStream_t fst = e->single_stream ? st : e->st2;
    if (!e->single_stream) {
        HIPCHK(e, hipEventRecord(e->ev_fork, st));
        HIPCHK(e, hipStreamWaitEvent(e->st2, e->ev_fork, 0));
    }
    // thermal pressurization (friclaw 5) before the fault pairs, src/driver.f90:27
    {
        const auto &gs = e->groups[kKTp];
        const Stamp p = take_timer(e, kKTp, gs.size());
        for (size_t i = 0; i < gs.size(); i++) HIPCHK(e, launch_thermop<R>(dM, gs[i].g, fst, p.start(i), p.stop(i)));
    }
    {
        const auto &gs = e->groups[kKFault];
        const Stamp p = take_timer(e, kKFault, gs.size());
        for (size_t i = 0; i < gs.size(); i++)
            HIPCHK(e, launch_fault<R>(dM, gs[i].g, gs[i].sig, fst, p.start(i), p.stop(i)));
    }
    if (!e->single_stream) HIPCHK(e, hipEventRecord(e->ev_join, e->st2));
    {
        const auto &gs = e->groups[kKNode3];
        const Stamp p = take_timer(e, kKNode3, gs.size());
        for (size_t i = 0; i < gs.size(); i++) HIPCHK(e, launch_node3<R>(dM, gs[i].g, st, p.start(i), p.stop(i)));
    }
    {
        const auto &gs = e->groups[kKNodePml];
        const Stamp p = take_timer(e, kKNodePml, gs.size());
        for (size_t i = 0; i < gs.size(); i++)
            HIPCHK(e, launch_node_pml<R>(dM, gs[i].g, st, p.start(i), p.stop(i)));
    }
    // join: the next step sees the fault pairs' update
    if (!e->single_stream) HIPCHK(e, hipStreamWaitEvent(st, e->ev_join, 0));
    return EQD_OK;
}

template <typename R>
int capture(eqd_engine *e)
{
    HIPCHK(e, hipStreamBeginCapture(e->st, hipStreamCaptureModeRelaxed));
    const int rc = enqueue_step<R>(e);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(e->st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return fail(e, EQD_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
    e->graph = g;
    HIPCHK(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    return EQD_OK;
}

// ---------------- records (eqd_record): the window a run may write, and the
// copies of its records into the pinned host ring behind its last step

bool recording(const eqd_engine *e, const eqd_handle *h) { return h->d_rec && !e->no_harvest; }

// Enqueue (on the handle's copy stream, behind everything st holds so far)
// the copy of the run's records r0 <= r < r1 that fit the window starting at
// lo, and queue the run's segment.
int rec_copy(eqd_engine *e, eqd_handle *h, int32_t r0, int32_t r1, int32_t lo)
{
    RecStream &q = h->rq;
    const Recorder &R = h->rec;
    RecSeg g{r0, std::max(r0, std::min(r1, lo + R.capacity)), r1, nullptr};
    if (g.end > g.begin) {
        if (q.spare.empty()) {
            hipEvent_t ev;
            HIPCHK(e, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            q.spare.push_back(ev);
        }
        g.ev = q.spare.back();
        q.spare.pop_back();
        HIPCHK(e, hipEventRecord(e->ev_rec, e->st));
        HIPCHK(e, hipStreamWaitEvent(q.cst, e->ev_rec, 0));
        const int64_t nv = 3 * (int64_t)R.n_nodes, nf = h->F.nf;
        for (int32_t r = g.begin; r < g.end;) {
            const int32_t a = r % R.capacity, n = std::min(g.end - r, R.capacity - a);  // up to the ring's end
            if (nv > 0)
                HIPCHK(e, hipMemcpyAsync(q.hvel + a * nv, R.vel + a * nv, sizeof(double) * nv * n,
                                         hipMemcpyDeviceToHost, q.cst));
            if (nf > 0 && R.slip_rate)
                HIPCHK(e, hipMemcpyAsync(q.hslip + a * nf, R.slip_rate + a * nf, sizeof(double) * nf * n,
                                         hipMemcpyDeviceToHost, q.cst));
            r += n;
        }
        HIPCHK(e, hipEventRecord(g.ev, q.cst));
    }
    q.segs.push_back(g);
    q.next = r1;
    return EQD_OK;
}

// nsteps steps of every scenario, enqueued (finish waits for them); with
// mirrors the records are copied out (the tuner's trial runs do neither)
template <typename R>
int launch(eqd_engine *e, int32_t nsteps, bool mirrors)
{
    const Member<R> *dM = (const Member<R> *)e->d_members;
    hipStream_t st = e->st;
    // clocks from the host mirrors; the run ends at nt + nsteps, and writes
    // the records from the first one not yet handed over
    std::vector<int32_t> lo(e->hs.size(), 0);
    for (size_t m = 0; m < e->hs.size(); m++) {
        eqd_handle *h = e->hs[m];
        if (mirrors && recording(e, h)) lo[m] = h->rq.drained;
        const Clock c = {h->time, h->nt, 0, h->nt + nsteps, lo[m]};
        HIPCHK(e, hipMemcpy(h->d_clock, &c, sizeof(Clock), hipMemcpyHostToDevice));
    }
    for (auto &kv : e->timers) {
        kv.second.used = 0;
        kv.second.total_ms = 0;
        kv.second.launches = 0;
    }
    // the update due from the previous run's last step (its stored F/m)
    const LaunchGrid none{};
    for (const Group &g : e->groups[kKPend3]) HIPCHK(e, launch_pending<R>(dM, g.g, none, st));
    for (const Group &g : e->groups[kKPendPml]) HIPCHK(e, launch_pending<R>(dM, none, g.g, st));
    const bool graph = e->graph_on && !e->timing;
    if (graph && !e->exec) {
        int rc = capture<R>(e);
        if (rc) return rc;
    }
    for (int32_t k = 0; k < nsteps; k++) {
        if (graph) {
            HIPCHK(e, hipGraphLaunch(e->exec, st));
        } else {
            e->sample = e->timing > 0 && k % e->timing == 0;
            int rc = enqueue_step<R>(e);
            e->sample = false;
            if (rc) return rc;
            if (e->timing) {
                size_t most = 0;
                for (auto &kv : e->timers) most = std::max(most, kv.second.used);
                if (most >= kTimerPool && (rc = drain_timers(e))) return rc;
            }
        }
    }
    if (mirrors)
        for (size_t m = 0; m < e->hs.size(); m++) {
            eqd_handle *h = e->hs[m];
            if (!recording(e, h)) continue;
            const int32_t every = h->rec.every;
            int rc = rec_copy(e, h, next_record(h->nt, every), next_record(h->nt + nsteps, every), lo[m]);
            if (rc) return rc;
        }
    return EQD_OK;
}

// Wait for the run that launch enqueued; with mirrors the handles' host
// clocks advance and NaN velocities are reported.
int finish(eqd_engine *e, int32_t nsteps, bool mirrors)
{
    HIPCHK(e, hipStreamSynchronize(e->st));
    HIPCHK(e, hipStreamSynchronize(e->st2));
    if (e->timing) {
        int rc = drain_timers(e);
        if (rc) return rc;
    }
    if (!mirrors) return EQD_OK;
    // The NaN flag is read once per run, so a run goes on to its end past a
    // NaN; the handle's clock then stops at the step of the first NaN (the
    // reference stops there, src/driver.f90:147-152) and the device state is
    // that of the run's end, which the caller should not continue from.
    int bad = -1;
    int32_t bad_node = 0, bad_step = 0;
    for (size_t m = 0; m < e->hs.size(); m++) {
        eqd_handle *h = e->hs[m];
        unsigned long long key;
        HIPCHK(e, hipMemcpy(&key, h->N.nanflag, sizeof(key), hipMemcpyDeviceToHost));
        int32_t steps = nsteps;
        if (key != ~0ull) {
            const int32_t node = (int32_t)(uint32_t)(key & 0xFFFFFFFFu), step = (int32_t)(uint32_t)(key >> 32);
            steps = std::max(0, std::min(nsteps, step - h->nt));
            if (bad < 0) {
                bad = (int)m;
                bad_node = node;
                bad_step = step;
            }
        }
        h->nt += steps;
        for (int32_t k = 0; k < steps; k++) h->time = h->time + h->s.dt;
        if (steps < nsteps && recording(e, h) && !h->rq.segs.empty()) {
            // the records after the NaN step are not the reference's: the
            // run's segment ends there and the next run writes from there on
            RecSeg &g = h->rq.segs.back();
            g.next = std::max(g.begin, next_record(h->nt, h->rec.every));
            g.end = std::min(g.end, g.next);
            h->rq.next = g.next;
        }
    }
    if (bad >= 0) {
        if (e->hs.size() > 1)
            return fail(e, EQD_ENAN, "scenario %d: Velocity NaN at node %d at time step %d", bad, bad_node, bad_step);
        return fail(e, EQD_ENAN, "Velocity NaN at node %d at time step %d", bad_node, bad_step);
    }
    return EQD_OK;
}

template <typename R>
int run(eqd_engine *e, int32_t nsteps, bool mirrors)
{
    const int rc = launch<R>(e, nsteps, mirrors);
    return rc ? rc : finish(e, nsteps, mirrors);
}

// device state a step changes, for the tuner's save and restore
struct Region {
    void *p;
    size_t bytes;
    void *backup;
};

std::vector<Region> mutable_regions(const eqd_engine *e)
{
    std::vector<Region> r;
    const size_t rs = e->fp32 ? sizeof(float) : sizeof(double);
    auto add = [&](const void *p, size_t b) {
        if (p && b) r.push_back({const_cast<void *>(p), b, nullptr});
    };
    for (const eqd_handle *h : e->hs) {
        // node, element and PML-node arrays are in blocks of 64 items (blk_at)
        auto blocked = [](int64_t n, int w) { return (size_t)((n + 63) / 64 * 64 * w); };
        const int64_t nn = h->N.nn;
        add(h->N.vel, blocked(nn, 3) * sizeof(double));
        add(h->N.disp, blocked(nn, 3) * sizeof(double));
        add(h->N.vel32, blocked(nn, 3) * sizeof(float));
        add(h->N.dhi, blocked(nn, 3) * sizeof(float));
        add(h->N.dlo, blocked(nn, 3) * sizeof(uint16_t));
        add(h->N.acc, blocked(nn, 3) * sizeof(double));
        add(h->N.nanflag, sizeof(unsigned long long));
        add(e->fp32 ? (const void *)h->E32.stress : (const void *)h->E64.stress, blocked(h->n_reg, h->nreg_stress) * rs);
        add(e->fp32 ? (const void *)h->E32.pstrain : (const void *)h->E64.pstrain, (size_t)h->n_reg * sizeof(double));
        add(e->fp32 ? (const void *)h->P32.stress : (const void *)h->P64.stress, blocked(h->n_pml, 21) * rs);
        add(h->PN.v12, blocked(h->PN.n, 12) * sizeof(double));
        add(h->PN.v12f, blocked(h->PN.n, 12) * sizeof(float));
        add(h->PN.acc, blocked(h->PN.n, 12) * sizeof(double));
        if (h->F.nf > 0) {
            add(h->F.fric, 100 * (size_t)h->F.ld * sizeof(double));
            add(h->F.sv, 6 * (size_t)h->F.ld * sizeof(double));
            add(h->F.su, 6 * (size_t)h->F.ld * sizeof(double));
            add(h->F.fnft, (size_t)h->F.nf * sizeof(double));
            add(h->F.nstat, 5 * (size_t)h->F.ld * sizeof(int32_t));
            add(h->F.hist, (size_t)(h->s.non > 0 ? h->s.non : 1) * h->s.nstep * 12 * sizeof(double));
            add(h->F.hyp_hist, (size_t)h->F.nhyp * h->s.nstep * kHypVals * sizeof(double));
        }
        if (h->S.nrow > 0) add(h->S.hist, (size_t)(h->S.nrow + 1) * h->s.nstep * sizeof(double));
        if (h->F.tph) add(h->F.tph, 2 * (size_t)h->F.nf * h->s.nstep * sizeof(double));
        add(h->d_clock, sizeof(Clock));
    }
    return r;
}

int copy_regions(eqd_engine *e, std::vector<Region> &r, bool save)
{
    for (Region &x : r) {
        if (save)
            HIPCHK(e, hipMemcpyAsync(x.backup, x.p, x.bytes, hipMemcpyDeviceToDevice, e->st));
        else
            HIPCHK(e, hipMemcpyAsync(x.p, x.backup, x.bytes, hipMemcpyDeviceToDevice, e->st));
    }
    HIPCHK(e, hipStreamSynchronize(e->st));
    return EQD_OK;
}

template <typename R>
int tune_impl(eqd_engine *e, int32_t steps, std::string &report)
{
    for (const eqd_handle *h : e->hs)
        if (h->nt + 1 + steps > h->s.nstep)
            return fail(e, EQD_EINVAL, "tuning runs %d steps past step %d but nstep is %d", 1 + steps, h->nt,
                        h->s.nstep);
    std::vector<Region> regs = mutable_regions(e);
    for (Region &x : regs)
        if (hipMalloc(&x.backup, x.bytes) != hipSuccess) {
            for (Region &y : regs)
                if (y.backup) (void)hipFree(y.backup);
            return fail(e, EQD_EOOM, "tuner backup of %zu bytes", x.bytes);
        }
    int rc = copy_regions(e, regs, true);
    const int32_t timing = e->timing;
    e->timing = 1;
    e->no_harvest = true;
    struct Cand {
        int block, remap;
        double ms;
    };
    report = "{";
    const int kinds[] = {kKElem, kKPml, kKNode3, kKNodePml, kKFault};
    bool first = true;
    for (int kind : kinds) {
        if (rc) break;
        bool any = false;
        for (const eqd_handle *h : e->hs) any = any || count_of(h, kind) > 0;
        if (!any) continue;
        std::vector<Cand> cands;
        if (kind == kKFault) {
            for (int b : {64, 128}) cands.push_back({b, 0, 0.0});
        } else {
            for (int b : {64, 128, 256})
                for (int r : {0, 1}) cands.push_back({b, r, 0.0});
        }
        for (Cand &c : cands) {
            e->tune.block[kind] = c.block;
            e->tune.remap[kind] = c.remap;
            if ((rc = build<R>(e)) || (rc = copy_regions(e, regs, false))) break;
            // one untimed step (first launch of a variant), then the timed ones
            if ((rc = run<R>(e, 1, false)) || (rc = run<R>(e, steps, false))) break;
            const KTimer &t = e->timers[kKindName[kind]];
            c.ms = t.launches ? t.total_ms / (double)t.launches : 0.0;
        }
        if (rc) break;
        const Cand *best = &cands[0];
        for (const Cand &c : cands)
            if (c.ms < best->ms) best = &c;
        e->tune.block[kind] = best->block;
        e->tune.remap[kind] = best->remap;
        char buf[160];
        snprintf(buf, sizeof(buf), "%s\"%s\": {\"block\": %d, \"remap\": %d, \"ms\": %.5f, \"candidates\": [", first ? "" : ", ",
                 kKindName[kind], best->block, best->remap, best->ms);
        report += buf;
        for (size_t i = 0; i < cands.size(); i++) {
            snprintf(buf, sizeof(buf), "%s[%d, %d, %.5f]", i ? ", " : "", cands[i].block, cands[i].remap, cands[i].ms);
            report += buf;
        }
        report += "]}";
        first = false;
    }
    report += "}";
    e->no_harvest = false;
    e->timing = timing;
    const int rc2 = copy_regions(e, regs, false);
    for (Region &x : regs) (void)hipFree(x.backup);
    if (rc) return rc;
    if (rc2) return rc2;
    for (auto &kv : e->timers) {
        kv.second.total_ms = 0;
        kv.second.launches = 0;
    }
    return build<R>(e);
}

}  // namespace

namespace eqd {

int32_t next_record(int32_t nt, int32_t every) { return nt <= 0 ? 0 : (nt - 1) / every + 1; }

int engine_init(eqd_engine *e, const std::vector<eqd_handle *> &hs)
{
    if (hs.empty()) return fail(e, EQD_EINVAL, "no scenarios");
    e->hs = hs;
    e->fp32 = hs[0]->s.precision == EQD_PREC_FP32;
    e->device = hs[0]->s.device;
    for (const eqd_handle *h : hs) {
        if ((h->s.precision == EQD_PREC_FP32) != e->fp32) return fail(e, EQD_EINVAL, "scenarios differ in precision");
        if (h->s.device != e->device) return fail(e, EQD_EINVAL, "scenarios are on different devices");
    }
    for (int k = 0; k < kKinds; k++) {
        e->tune.block[k] = 256;
        e->tune.remap[k] = 0;
    }
    e->tune.block[kKFault] = 128;
    e->tune.block[kKStation] = 64;
    {
        const char *s = std::getenv("EQD_XCD_REMAP");
        if (s && s[0] == '1') e->tune.remap[kKElem] = e->tune.remap[kKPml] = 1;
        s = std::getenv("EQD_SINGLE_STREAM");
        e->single_stream = s && s[0] == '1';
        s = std::getenv("EQD_STEP_START_LAUNCH");
        e->step_start_launch = s && s[0] == '1';
    }
    HIPCHK(e, hipSetDevice(e->device));
    // a second stream (default priority) for the fault pairs, which overlap the
    // node kernels within a step; events join the two every step
    HIPCHK(e, hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking));
    HIPCHK(e, hipStreamCreateWithFlags(&e->st2, hipStreamNonBlocking));
    // The fork and join events order two streams of this device and nothing
    // else, so they skip the system-scope fence (a cache writeback and
    // invalidation at every record: each kernel's own device-scope release
    // and acquire already order the data); EQD_EVENT_SYSTEM_FENCE=1 keeps it.
    // The record event orders a copy into host memory and keeps its fence.
    unsigned sync_flags = hipEventDisableTiming | hipEventDisableSystemFence;
    {
        const char *s = std::getenv("EQD_EVENT_SYSTEM_FENCE");
        if (s && s[0] == '1') sync_flags = hipEventDisableTiming;
    }
    HIPCHK(e, hipEventCreateWithFlags(&e->ev_fork, sync_flags));
    HIPCHK(e, hipEventCreateWithFlags(&e->ev_join, sync_flags));
    HIPCHK(e, hipEventCreateWithFlags(&e->ev_rec, hipEventDisableTiming));
    const size_t msize = e->fp32 ? sizeof(Member<float>) : sizeof(Member<double>);
    HIPCHK(e, hipMalloc(&e->d_members, msize * hs.size()));
    return EQD_OK;
}

int engine_launch(eqd_engine *e, int32_t nsteps)
{
    if (nsteps < 0) return fail(e, EQD_EINVAL, "nsteps < 0");
    if (nsteps == 0) return EQD_OK;
    HIPCHK(e, hipSetDevice(e->device));
    for (size_t m = 0; m < e->hs.size(); m++) {
        const eqd_handle *h = e->hs[m];
        if (h->nt + nsteps > h->s.nstep)
            return fail(e, EQD_EINVAL, "step %d exceeds nstep %d (station history length)%s", h->nt + nsteps,
                        h->s.nstep, e->hs.size() > 1 ? " in an ensemble scenario" : "");
    }
    int rc = ensure_built(e);
    if (rc) return rc;
    return e->fp32 ? launch<float>(e, nsteps, true) : launch<double>(e, nsteps, true);
}

int engine_finish(eqd_engine *e, int32_t nsteps)
{
    if (nsteps <= 0) return EQD_OK;
    HIPCHK(e, hipSetDevice(e->device));
    return finish(e, nsteps, true);
}

int engine_run(eqd_engine *e, int32_t nsteps)
{
    const int rc = engine_launch(e, nsteps);
    return rc ? rc : engine_finish(e, nsteps);
}

int engine_tune(eqd_engine *e, int32_t steps, std::string &report)
{
    if (steps < 1 || steps > 1000) return fail(e, EQD_EINVAL, "tuning steps %d outside 1..1000", steps);
    HIPCHK(e, hipSetDevice(e->device));
    int rc = ensure_built(e);
    if (rc) return rc;
    return e->fp32 ? tune_impl<float>(e, steps, report) : tune_impl<double>(e, steps, report);
}

int engine_set_launch(eqd_engine *e, int kind, int block, int remap)
{
    if (kind < 0 || kind > kKFault) return fail(e, EQD_EINVAL, "launch kind %d outside 0..4", kind);
    if (kind == kKFault ? (block != 64 && block != 128) : (block != 64 && block != 128 && block != 256))
        return fail(e, EQD_EINVAL, "block %d is not a launch shape of %s", block, kKindName[kind]);
    if (remap != 0 && remap != 1) return fail(e, EQD_EINVAL, "remap must be 0 or 1");
    e->tune.block[kind] = block;
    e->tune.remap[kind] = kind == kKFault ? 0 : remap;
    e->built = false;
    drop_graph(e);
    return EQD_OK;
}

void engine_free(eqd_engine *e)
{
    (void)hipSetDevice(e->device);
    if (e->st) (void)hipStreamSynchronize(e->st);
    if (e->st2) (void)hipStreamSynchronize(e->st2);
    drop_graph(e);
    free_tables(e);
    if (e->d_members) (void)hipFree(e->d_members);
    e->d_members = nullptr;
    for (auto &kv : e->timers)
        for (auto &p : kv.second.ev) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    e->timers.clear();
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    if (e->ev_join) (void)hipEventDestroy(e->ev_join);
    if (e->ev_rec) (void)hipEventDestroy(e->ev_rec);
    e->ev_rec = nullptr;
    if (e->st2) (void)hipStreamDestroy(e->st2);
    if (e->st) (void)hipStreamDestroy(e->st);
    e->ev_fork = e->ev_join = nullptr;
    e->st = e->st2 = nullptr;
}

}  // namespace eqd

// ---------------- C ABI: launch shapes and graphs of one handle, ensembles

struct eqd_ensemble {
    eqd_engine e;
};

static int copy_report(const std::string &r, char *out, int64_t cap)
{
    if (out && cap > 0) {
        const size_t n = std::min((size_t)(cap - 1), r.size());
        memcpy(out, r.data(), n);
        out[n] = 0;
    }
    return (int64_t)r.size() + 1 > cap && out ? EQD_EINVAL : EQD_OK;
}

extern "C" {

int eqd_tune(eqd_handle *h, int32_t steps, char *report, int64_t report_cap)
{
    if (!h) return EQD_EINVAL;
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_tune before eqd_commit");
    SETTLE(h);
    if (h->ensemble) return fail(h, EQD_ESTATE, "the handle belongs to an ensemble: tune the ensemble");
    std::string r;
    const int rc = engine_tune(h->eng, steps, r);
    if (rc) {
        h->err = h->eng->err;
        return rc;
    }
    if (copy_report(r, report, report_cap)) return fail(h, EQD_EINVAL, "report buffer too small (%zu bytes)", r.size() + 1);
    return EQD_OK;
}

int eqd_set_launch(eqd_handle *h, int32_t kind, int32_t block, int32_t remap)
{
    if (!h) return EQD_EINVAL;
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_set_launch before eqd_commit");
    SETTLE(h);
    const int rc = engine_set_launch(h->eng, kind, block, remap);
    if (rc) h->err = h->eng->err;
    return rc;
}

int eqd_get_launch(eqd_handle *h, int32_t kind, int32_t *block, int32_t *remap)
{
    if (!h || !block || !remap) return EQD_EINVAL;
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_get_launch before eqd_commit");
    if (kind < 0 || kind > kKFault) return fail(h, EQD_EINVAL, "launch kind %d outside 0..4", kind);
    *block = h->eng->tune.block[kind];
    *remap = h->eng->tune.remap[kind];
    return EQD_OK;
}

int eqd_set_graph(eqd_handle *h, int32_t on)
{
    if (!h) return EQD_EINVAL;
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_set_graph before eqd_commit");
    SETTLE(h);
    h->eng->graph_on = on != 0;
    return EQD_OK;
}

int eqd_ensemble_create(eqd_handle *const *members, int32_t n, eqd_ensemble **out)
{
    if (!members || !out || n < 1 || n > 65535) return EQD_EINVAL;
    *out = nullptr;
    std::vector<eqd_handle *> hs(members, members + n);
    for (int i = 0; i < n; i++) {
        if (!hs[i] || !hs[i]->committed || hs[i]->ensemble) return EQD_ESTATE;
        for (int j = 0; j < i; j++)
            if (hs[j] == hs[i]) return EQD_EINVAL;
    }
    for (eqd_handle *h : hs) SETTLE(h);
    eqd_ensemble *E = new eqd_ensemble();
    const int rc = engine_init(&E->e, hs);
    if (rc) {
        engine_free(&E->e);
        delete E;
        return rc;
    }
    E->e.graph_on = true;
    for (eqd_handle *h : hs) {
        (void)hipStreamSynchronize(h->stream);
        h->ensemble = &E->e;
    }
    *out = E;
    return EQD_OK;
}

int eqd_ensemble_run(eqd_ensemble *E, int32_t nsteps)
{
    if (!E) return EQD_EINVAL;
    return engine_run(&E->e, nsteps);
}

int eqd_ensemble_tune(eqd_ensemble *E, int32_t steps, char *report, int64_t report_cap)
{
    if (!E) return EQD_EINVAL;
    std::string r;
    const int rc = engine_tune(&E->e, steps, r);
    if (rc) return rc;
    if (copy_report(r, report, report_cap))
        return fail(&E->e, EQD_EINVAL, "report buffer too small (%zu bytes)", r.size() + 1);
    return EQD_OK;
}

int eqd_ensemble_set_launch(eqd_ensemble *E, int32_t kind, int32_t block, int32_t remap)
{
    if (!E) return EQD_EINVAL;
    return engine_set_launch(&E->e, kind, block, remap);
}

int eqd_ensemble_set_graph(eqd_ensemble *E, int32_t on)
{
    if (!E) return EQD_EINVAL;
    E->e.graph_on = on != 0;
    return EQD_OK;
}

int eqd_ensemble_set_timing(eqd_ensemble *E, int32_t on)
{
    if (!E) return EQD_EINVAL;
    E->e.timing = on < 0 ? 0 : on;
    return EQD_OK;
}

int eqd_ensemble_kernel_time(eqd_ensemble *E, const char *name, double *avg_ms, int64_t *launches)
{
    if (!E || !name) return EQD_EINVAL;
    auto it = E->e.timers.find(name);
    const bool none = it == E->e.timers.end() || it->second.launches == 0;
    if (avg_ms) *avg_ms = none ? 0.0 : it->second.total_ms / (double)it->second.launches;
    if (launches) *launches = none ? 0 : it->second.launches;
    return EQD_OK;
}

const char *eqd_ensemble_last_error(eqd_ensemble *E) { return E ? E->e.err.c_str() : "null ensemble"; }

int eqd_ensemble_destroy(eqd_ensemble *E)
{
    if (!E) return EQD_OK;
    engine_free(&E->e);
    for (eqd_handle *h : E->e.hs) h->ensemble = nullptr;
    delete E;
    return EQD_OK;
}

}  // extern "C"
