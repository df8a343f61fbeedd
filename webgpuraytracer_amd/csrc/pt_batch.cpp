// pt_batch.cpp — one graph launch per progressive frame of many jobs.
//
// Each job is a pt_ctx with its own scene, frame, options and stream.  A
// frame prepares every job (allocation, seed upload: pt::prepare_trace), then
// records the jobs' kernel chains (pt::enqueue_trace, kernels only) on their
// own streams forked from and joined back into the batch stream under stream
// capture.  The first frame instantiates the graph; later frames update the
// executable graph in place when the topology is unchanged (same active jobs
// and batch counts: only kernel arguments such as sample numbers differ) and
// re-instantiate otherwise.  Nothing in the graph is a memset or a copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pt_batch.h"
#include "pt_ctx.h"

struct pt_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> join;
    std::vector<pt_ctx*> jobs;
    hipGraphExec_t exec = nullptr;
    uint32_t frames = 0, instantiations = 0, updates = 0;
    std::string err;

    int fail(int code, const char* fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        return fail(e == hipErrorOutOfMemory ? PT_ERR_OOM : PT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    }
};

extern "C" {

int pt_batch_create(pt_batch** out, int device) {
    if (!out) return PT_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PT_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return PT_ERR_ARG;
    pt_batch* b = new (std::nothrow) pt_batch();
    if (!b) return PT_ERR_OOM;
    b->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        pt_batch_destroy(b);
        return PT_ERR_HIP;
    }
    *out = b;
    return PT_OK;
}

void pt_batch_destroy(pt_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* pt_batch_last_error(const pt_batch* b) { return b ? b->err.c_str() : "null batch"; }

int pt_batch_add(pt_batch* b, pt_ctx* job, uint32_t* index) {
    if (!b || !job) return PT_ERR_ARG;
    if (job->device != b->device) return b->fail(PT_ERR_ARG, "job on device %d, batch on device %d", job->device, b->device);
    if (std::find(b->jobs.begin(), b->jobs.end(), job) != b->jobs.end())
        return b->fail(PT_ERR_ARG, "context already in the batch (its buffers cannot serve two jobs at once)");
    (void)hipSetDevice(b->device);
    hipEvent_t ev = nullptr;
    const hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) return b->hip_fail(e, "batch event");
    b->join.push_back(ev);
    b->jobs.push_back(job);
    if (index) *index = uint32_t(b->jobs.size() - 1);
    return PT_OK;
}

int pt_batch_size(const pt_batch* b, uint32_t* n) {
    if (!b || !n) return PT_ERR_ARG;
    *n = uint32_t(b->jobs.size());
    return PT_OK;
}

int pt_batch_graph_info(const pt_batch* b, uint32_t* frames, uint32_t* instantiations, uint32_t* updates) {
    if (!b) return PT_ERR_ARG;
    if (frames) *frames = b->frames;
    if (instantiations) *instantiations = b->instantiations;
    if (updates) *updates = b->updates;
    return PT_OK;
}

int pt_batch_frame(pt_batch* b, const pt_scene_params* params, const uint32_t* n_passes,
                   const uint32_t* const* seeds, pt_stats* stats) {
    if (!b || !params || !n_passes) return PT_ERR_ARG;
    const size_t n = b->jobs.size();
    if (n == 0) return b->fail(PT_ERR_STATE, "no jobs in the batch");
    (void)hipSetDevice(b->device);
    std::vector<pt::TracePlan> plans(n);
    std::vector<char> active(n, 0);
    size_t n_active = 0;
    for (size_t i = 0; i < n; ++i) {
        if (n_passes[i] == 0) continue;
        pt_ctx* c = b->jobs[i];
        int rc = pt::check_trace_params(c, &params[i]);
        if (!rc) rc = pt::prepare_trace(c, &params[i], 0, 0, n_passes[i], seeds ? seeds[i] : nullptr, nullptr, 0, plans[i]);
        if (rc) return b->fail(rc, "job %zu: %s", i, c->err.c_str());
        active[i] = 1;
        ++n_active;
    }
    if (n_active == 0) {
        if (stats) std::memset(stats, 0, sizeof(pt_stats) * n);
        return PT_OK;
    }
    // seed uploads and allocations are done before anything is recorded
    for (size_t i = 0; i < n; ++i) {
        if (!active[i]) continue;
        const hipError_t e = hipStreamSynchronize(b->jobs[i]->stream);
        if (e != hipSuccess) return b->hip_fail(e, "job stream");
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return b->hip_fail(e, "begin capture");
    hipError_t rec = hipEventRecord(b->fork, b->stream);
    for (size_t i = 0; rec == hipSuccess && i < n; ++i) {
        if (!active[i]) continue;
        pt_ctx* c = b->jobs[i];
        rec = hipStreamWaitEvent(c->stream, b->fork, 0);
        if (rec == hipSuccess) rec = pt::enqueue_trace(c, plans[i], c->stream, false);
        if (rec == hipSuccess) rec = hipEventRecord(b->join[i], c->stream);
        if (rec == hipSuccess) rec = hipStreamWaitEvent(b->stream, b->join[i], 0);
    }
    e = hipStreamEndCapture(b->stream, &g);  // always ends the capture
    if (rec != hipSuccess || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return b->hip_fail(rec != hipSuccess ? rec : e, "record frame");
    }
    if (b->exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(b->exec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            ++b->updates;
        } else {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
        }
    }
    if (!b->exec) {
        e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            b->exec = nullptr;
            (void)hipGraphDestroy(g);
            return b->hip_fail(e, "instantiate frame graph");
        }
        ++b->instantiations;
    }
    (void)hipGraphDestroy(g);
    e = hipEventRecord(b->ev0, b->stream);
    if (e == hipSuccess) e = hipGraphLaunch(b->exec, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return b->hip_fail(e, "frame graph");
    ++b->frames;
    if (stats) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
        for (size_t i = 0; i < n; ++i) {
            if (!active[i]) {
                std::memset(&stats[i], 0, sizeof(pt_stats));
                continue;
            }
            const int rc = pt::read_stats(b->jobs[i], &stats[i], ms, 1);
            if (rc) return b->fail(rc, "job %zu: %s", i, b->jobs[i]->err.c_str());
        }
    }
    return PT_OK;
}

}  // extern "C"
