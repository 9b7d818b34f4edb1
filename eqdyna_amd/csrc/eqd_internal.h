// Host-side internals of libeqd shared by the C-ABI (eqd_api.cpp) and the
// step engine (eqd_engine.cpp): the handle, the engine that launches steps
// for one or many scenarios, and the kernel launchers (eqd_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/eqdyna_amd.h"
#include "eqd_device.h"

namespace eqd {
template <typename R>
hipError_t launch_elem(const Member<R> *, const LaunchGrid &, int q, int dp, int hg, int mass, int ops, int nstart,
                       hipStream_t, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
template <typename R>
hipError_t launch_pml(const Member<R> *, const LaunchGrid &, int hg, int mass, int ops, hipStream_t,
                      hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
template <typename R>
hipError_t launch_node3(const Member<R> *, const LaunchGrid &, hipStream_t, hipEvent_t t0 = nullptr,
                        hipEvent_t t1 = nullptr);
template <typename R>
hipError_t launch_node_pml(const Member<R> *, const LaunchGrid &, hipStream_t, hipEvent_t t0 = nullptr,
                           hipEvent_t t1 = nullptr);
template <typename R>
hipError_t launch_pending(const Member<R> *, const LaunchGrid &, const LaunchGrid &, hipStream_t);
template <typename R>
hipError_t launch_fault(const Member<R> *, const LaunchGrid &, int learn, hipStream_t, hipEvent_t t0 = nullptr,
                        hipEvent_t t1 = nullptr);
template <typename R>
hipError_t launch_step_start(const Member<R> *, int nmem, hipStream_t, hipEvent_t t0 = nullptr,
                             hipEvent_t t1 = nullptr);
template <typename R>
hipError_t launch_thermop(const Member<R> *, const LaunchGrid &, hipStream_t, hipEvent_t t0 = nullptr,
                          hipEvent_t t1 = nullptr);
template <typename R>
hipError_t launch_record(const Member<R> *, const LaunchGrid &, hipStream_t, hipEvent_t t0 = nullptr,
                         hipEvent_t t1 = nullptr);
hipError_t launch_mlp_prepare(const float *p, __bf16 *wb, float *w4r, hipStream_t s);

struct KTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t used = 0;
    double total_ms = 0;
    int64_t launches = 0;
};

// launch shape of each kind: block size and XCD remapping
struct Tune {
    int block[kKinds];
    int remap[kKinds];
};

// one launch of one kind for the scenarios that share a kernel variant
struct Group {
    int sig = 0;               // kernel variant (template flags) of the launch
    std::vector<int> members;  // indices into the engine's scenarios
    LaunchGrid g{};
    uint16_t *d_bmem = nullptr;  // block-to-scenario table, owned (null when the engine has one scenario)
};

}  // namespace eqd

struct eqd_handle;

// The step engine: one or many committed scenarios of one precision on one
// device, advanced together.  Each step is one launch per kind and kernel
// variant over all scenarios (block-to-scenario tables), optionally captured
// once as a HIP graph and replayed.
struct eqd_engine {
    std::vector<eqd_handle *> hs;
    bool fp32 = false;
    int device = 0;
    hipStream_t st = nullptr, st2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_rec = nullptr;  // a run's end, for the record copies
    bool single_stream = false;
    bool step_start_launch = false;  // the step start as a launch of its own
    void *d_members = nullptr;  // Member<R>[hs.size()]
    std::vector<eqd::Group> groups[eqd::kKinds];
    eqd::Tune tune{};
    bool built = false;
    int64_t built_version = -1;  // sum of the scenarios' versions at the last build
    bool no_harvest = false;     // set while tuning
    bool graph_on = false;  // replay one captured step instead of launching every step
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t timing = 0;   // per-kernel event timing: 0 off, k: every k-th step of a run
    bool sample = false;  // the step being enqueued carries timer events
    std::map<std::string, eqd::KTimer> timers;
    std::string err;
};

namespace eqd {

// The records of one run: [begin, end) were copied (or are being copied) into
// the pinned host ring on the handle's copy stream, `ev` marks the copy's end,
// and [end, next) were dropped for want of room (next is the following run's
// first record).
struct RecSeg {
    int32_t begin, end, next;
    hipEvent_t ev;
};

// Host side of the record ring: the device writes record r into slot
// r % capacity of its ring, each run's records are copied to the same slot of
// the pinned host ring on `cst` behind the run's last step, and eqd_record_get
// hands them over from there.  drained: records before it were handed over
// (or dropped); the window a run may write is [drained, drained + capacity).
struct RecStream {
    hipStream_t cst = nullptr;
    double *hvel = nullptr, *hslip = nullptr;  // pinned [capacity][n_nodes][3], [capacity][nf]
    int32_t drained = 0, next = 0;
    std::deque<RecSeg> segs;
    std::vector<hipEvent_t> spare;  // finished copy events for reuse
};

}  // namespace eqd

struct eqd_handle {
    eqd_scalars s{};
    std::string err;
    std::map<int, const void *> src;         // caller's arrays, read during eqd_commit
    std::map<int, std::vector<double>> fd;   // f64 fields kept after commit
    std::map<int, std::vector<int32_t>> fi;  // i32 fields kept after commit
    bool committed = false;
    int32_t nt = 0;
    double time = 0.0;
    hipStream_t stream = nullptr;  // the handle's own engine's stream
    std::vector<void *> allocs;
    int64_t bytes = 0;
    // host-side copies of the layout needed for downloads
    std::vector<int32_t> reg_ids, pml_ids;     // element ids (0-based) per class
    std::vector<int32_t> node3_all, node3_free, pml_node_ids;
    std::vector<int32_t> pml_node_index;       // node -> PML node index or -1
    int32_t n_reg = 0, n_pml = 0, nreg_stress = 6;
    std::vector<int32_t> hyp_pairs;  // pairs (0-based) at the hypocentre, in pair order
    // device objects
    eqd::NodeState N{};
    eqd::RegElems<double> E64{};
    eqd::RegElems<float> E32{};
    eqd::PmlElems<double> P64{};
    eqd::PmlElems<float> P32{};
    int64_t op_classes_reg = 0, phi_classes_reg = 0, op_classes_pml = 0, phi_classes_pml = 0;
    int32_t phi_compact_reg = 0, phi_compact_pml = 0;  // hourglass operators in compact form
    eqd::PmlNodes PN{};
    eqd::FaultState F{};
    eqd::OffStations S{};
    eqd::StepConsts C{};
    eqd::Clock *d_clock = nullptr;
    int32_t *d_node3_free = nullptr, *d_node3_all = nullptr;
    uint32_t *d_ell3 = nullptr;  // slot-major gather entries of the free 3-component nodes (or null)
    // learned Newton guess: float32 parameters and their bfloat16 operands
    bool learn = false;
    float *mlp_p = nullptr, *mlp_w4r = nullptr;
    __bf16 *mlp_wb = nullptr;
    // harvest buffer of training rows (its schedule is in F)
    float *harvest = nullptr;
    // ground-motion / source-evolution records: host copy and device objects,
    // and the copies of finished records into pinned host memory
    eqd::Recorder rec{};
    eqd::Recorder *d_rec = nullptr;
    void *rec_bufs[3] = {nullptr, nullptr, nullptr};
    eqd::RecStream rq{};
    int32_t inflight = 0;  // steps of a run launched by eqd_run_async and not yet waited for
    int32_t timing = 0;  // per-kernel event timing of the handle's own runs (every k-th step)
    // bumped by every change the engine's launches depend on
    int64_t version = 0;
    eqd_engine *eng = nullptr;       // the handle's own engine (one scenario)
    eqd_engine *ensemble = nullptr;  // the ensemble the handle belongs to, if any
};

namespace eqd {

inline int vfail(std::string &err, int code, const char *fmt, va_list ap)
{
    char buf[1024];
    vsnprintf(buf, sizeof(buf), fmt, ap);
    err = buf;
    return code;
}

inline int fail(eqd_handle *h, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::string tmp;
    const int r = vfail(h ? h->err : tmp, code, fmt, ap);
    va_end(ap);
    return r;
}

inline int fail(eqd_engine *e, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::string tmp;
    const int r = vfail(e ? e->err : tmp, code, fmt, ap);
    va_end(ap);
    return r;
}

}  // namespace eqd

#define SETTLE(h)                                                                                      \
    do {                                                                                               \
        const int s_ = eqd::settle(h);                                                                 \
        if (s_) return s_;                                                                             \
    } while (0)

#define HIPCHK(h, x)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return eqd::fail(h, EQD_EHIP, "%s: %s", #x, hipGetErrorString(e_));      \
    } while (0)

namespace eqd {

// engine entry points (eqd_engine.cpp)
int engine_init(eqd_engine *e, const std::vector<eqd_handle *> &hs);
int engine_run(eqd_engine *e, int32_t nsteps);
int engine_launch(eqd_engine *e, int32_t nsteps);  // enqueue a run (engine_finish waits for it)
int engine_finish(eqd_engine *e, int32_t nsteps);
int32_t next_record(int32_t nt, int32_t every);      // first record index at a step after nt
// Wait for the handle's run in flight (eqd_run_async), if any: its clock
// advances and its error (a NaN) is returned here.
int settle(eqd_handle *h);
int engine_tune(eqd_engine *e, int32_t steps, std::string &report);
int engine_set_launch(eqd_engine *e, int kind, int block, int remap);
void engine_free(eqd_engine *e);

}  // namespace eqd
