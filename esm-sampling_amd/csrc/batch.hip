// batch.hip — operation sweeps of several replica sets (uneven networks, horizons and replica
// counts) captured into one HIP graph and launched as one, and the launch-shape autotuner.
//
// A batch is a list of handles on one device. esm_batch_sweeps captures, on the first handle's
// stream, a fork event that every other handle's stream waits on, then each handle's sweep
// launches on its own stream (the statistics cleared by a kernel, not a memset node), and a
// join of every stream back into the first: the graph holds one branch per replica set, whose
// kernels are exactly those of esm_sweeps, in the same order, with the same arguments. The graph
// is instantiated once; a later sweep whose kernel nodes have the same functions, grids,
// workgroups and LDS sizes in the same order updates the executable graph in place
// (hipGraphExecUpdate: new kernel arguments only), any other sweep instantiates it again. The
// handles are captured in their graph-stable form: the storage sweeps' per-sweep draws (whether
// they run, the order of the block lengths, offsets and classes) are made by the kernels from the
// sweep's counter, so the launch sequence is the same every sweep and the graph is instantiated
// once (a change of the capacities that makes a unit join or leave the long-range sweep, or of
// the sweep options, is the exception).
//
// esm_autotune times the launch forms of the random families on the handle's own state: it saves
// v and the caches, runs each candidate's sweeps from that saved state with the same key and
// counter (so every candidate does identical work), restores the state afterwards and keeps the
// fastest form.
#include <vector>

#include "esm_internal.h"

using namespace esm;

namespace {

__global__ void k_clear_stats(unsigned long long *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

constexpr int NSTAT_ALL = ESM_NFAM * ESM_NSTAT * STAT_SLOTS;

int read_stats(esm_handle *h, int64_t *out) {
    std::vector<unsigned long long> st((size_t)NSTAT_ALL);
    ESM_HIP_CHECK(hipMemcpyAsync(st.data(), h->d_stats, st.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                 h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    for (int i = 0; i < ESM_NFAM * ESM_NSTAT; i++) {
        unsigned long long acc = 0;
        for (int q = 0; q < STAT_SLOTS; q++) acc += st[(size_t)i * STAT_SLOTS + q];
        out[i] = (int64_t)acc;
    }
    return 0;
}

int sweeps_on_stream(esm_handle *h, int nsweeps, const uint32_t key[2], uint32_t counter, double maxnats) {
    hipLaunchKernelGGL(k_clear_stats, dim3((NSTAT_ALL + 255) / 256), dim3(256), 0, h->stream, h->d_stats, NSTAT_ALL);
    ESM_HIP_CHECK(hipGetLastError());
    return h->prec == ESM_FP64 ? launch_sweeps<double>(h, nsweeps, key[0], key[1], counter, maxnats)
                               : launch_sweeps<float>(h, nsweeps, key[0], key[1], counter, maxnats);
}

}  // namespace

// The launch shape of every node of a captured graph, in node order: type, and for kernels the
// function, grid, workgroup and dynamic LDS size. Two graphs with equal signatures differ only in
// kernel arguments.
static bool graph_signature(hipGraph_t g, std::vector<uintptr_t> &sig) {
    sig.clear();
    size_t n = 0;
    if (hipGraphGetNodes(g, nullptr, &n) != hipSuccess) return false;
    std::vector<hipGraphNode_t> nodes(n);
    if (n && hipGraphGetNodes(g, nodes.data(), &n) != hipSuccess) return false;
    for (size_t i = 0; i < n; i++) {
        hipGraphNodeType t;
        if (hipGraphNodeGetType(nodes[i], &t) != hipSuccess) return false;
        sig.push_back((uintptr_t)t);
        if (t == hipGraphNodeTypeKernel) {
            hipKernelNodeParams kp{};
            if (hipGraphKernelNodeGetParams(nodes[i], &kp) != hipSuccess) return false;
            sig.insert(sig.end(), {(uintptr_t)kp.func, kp.gridDim.x, kp.gridDim.y, kp.gridDim.z, kp.blockDim.x,
                                   kp.blockDim.y, kp.blockDim.z, kp.sharedMemBytes});
        }
    }
    return true;
}

struct esm_batch {
    std::vector<esm_handle *> hs;
    int device = 0;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraphExec_t exec = nullptr;
    std::vector<uintptr_t> sig;   // the signature of the graph the executable was made from
    int64_t counts[3] = {0, 0, 0};   // graph launches, instantiations, in-place updates
};

static void free_batch(esm_batch *b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    delete b;
}

extern "C" {

int32_t esm_batch_create(esm_handle *const *hs, int32_t n, esm_batch **out) {
    if (!hs || !out) { set_error("esm_batch_create: null argument"); return ESM_EINVAL; }
    *out = nullptr;
    if (n < 1 || n > ESM_BATCH_MAX) { set_error("esm_batch_create: n must be in [1, ESM_BATCH_MAX]"); return ESM_EINVAL; }
    for (int i = 0; i < n; i++) {
        if (!hs[i]) { set_error("esm_batch_create: null handle"); return ESM_EINVAL; }
        if (hs[i]->device != hs[0]->device) { set_error("esm_batch_create: the handles are on different devices"); return ESM_EINVAL; }
        for (int q = 0; q < i; q++)
            if (hs[q] == hs[i]) { set_error("esm_batch_create: a handle appears twice"); return ESM_EINVAL; }
    }
    esm_batch *b = new esm_batch;
    b->hs.assign(hs, hs + n);
    b->device = hs[0]->device;
    b->join.assign((size_t)n, nullptr);
    bool ok = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < n; i++) ok = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        set_error("esm_batch_create: hipEventCreate failed");
        free_batch(b);
        return ESM_EHIP;
    }
    *out = b;
    return ESM_OK;
}

int32_t esm_batch_destroy(esm_batch *b) {
    free_batch(b);
    return ESM_OK;
}

int32_t esm_batch_sweeps(esm_batch *b, int32_t nsweeps, const uint32_t *keys, const uint32_t *counters, double maxnats,
                         int64_t *stats) {
    if (!b || !keys || !counters) { set_error("esm_batch_sweeps: null argument"); return ESM_EINVAL; }
    if (nsweeps < 0 || !(maxnats > 0)) { set_error("esm_batch_sweeps: nsweeps must be >= 0 and maxnats > 0"); return ESM_EINVAL; }
    const int n = (int)b->hs.size();
    for (esm_handle *h : b->hs)
        if (h->prof) { set_error("esm_batch_sweeps: kernel profiling is on for a handle of the batch (esm_profile)"); return ESM_EINVAL; }
    int cur = 0;
    ESM_HIP_CHECK(hipGetDevice(&cur));
    if (cur != b->device) { set_error("esm_batch_sweeps: the batch's device is not the current device"); return ESM_EINVAL; }
    hipStream_t s0 = b->hs[0]->stream;
    ESM_HIP_CHECK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    auto fail = [&](const char *what) {
        set_error(std::string("esm_batch_sweeps: ") + what + ": " + hipGetErrorString(hipGetLastError()));
        rc = ESM_EHIP;
    };
    if (hipEventRecord(b->fork, s0) != hipSuccess) fail("fork event");
    for (int i = 0; i < n && !rc; i++) {
        esm_handle *h = b->hs[i];
        if (i > 0 && hipStreamWaitEvent(h->stream, b->fork, 0) != hipSuccess) { fail("fork wait"); break; }
        // the graph-stable launch sequence: the storage sweeps' draws are made on the device, so
        // every capture has the same nodes and the executable graph is updated, not rebuilt
        h->graph_stable = true;
        rc = sweeps_on_stream(h, nsweeps, keys + 2 * i, counters[i], maxnats);
        h->graph_stable = false;
        if (rc) break;
        if (i > 0 && (hipEventRecord(b->join[i], h->stream) != hipSuccess || hipStreamWaitEvent(s0, b->join[i], 0) != hipSuccess))
            fail("join");
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s0, &g);
    if (rc || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (!rc) set_error(std::string("esm_batch_sweeps: hipStreamEndCapture: ") + hipGetErrorString(ec));
        return rc ? rc : ESM_EHIP;
    }
    std::vector<uintptr_t> sig;
    const bool have_sig = graph_signature(g, sig);
    bool updated = false;
    if (b->exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res;
        updated = have_sig && sig == b->sig && hipGraphExecUpdate(b->exec, g, &bad, &res) == hipSuccess &&
                  res == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
        }
    }
    if (!b->exec) {
        const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            b->exec = nullptr;
            (void)hipGraphDestroy(g);
            set_error(std::string("esm_batch_sweeps: hipGraphInstantiate: ") + hipGetErrorString(ei));
            return ESM_EHIP;
        }
        b->sig = have_sig ? sig : std::vector<uintptr_t>();
        b->counts[1]++;
    } else {
        b->counts[2]++;
    }
    const hipError_t el = hipGraphLaunch(b->exec, s0);
    (void)hipGraphDestroy(g);
    if (el != hipSuccess) { set_error(std::string("esm_batch_sweeps: hipGraphLaunch: ") + hipGetErrorString(el)); return ESM_EHIP; }
    b->counts[0]++;
    ESM_HIP_CHECK(hipStreamSynchronize(s0));
    for (int i = 0; i < n; i++) {
        esm_handle *h = b->hs[i];
        h->scores_valid = false;
        if (stats && (rc = read_stats(h, stats + (size_t)i * ESM_NFAM * ESM_NSTAT))) return rc;
    }
    return ESM_OK;
}

int32_t esm_batch_counts(const esm_batch *b, int64_t *counts) {
    if (!b || !counts) { set_error("esm_batch_counts: null argument"); return ESM_EINVAL; }
    for (int i = 0; i < 3; i++) counts[i] = b->counts[i];
    return ESM_OK;
}

int32_t esm_sweep_block(esm_handle *h, int32_t block) {
    if (!h) { set_error("esm_sweep_block: null handle"); return ESM_EINVAL; }
    if (block != 64 && block != 128 && block != 256) { set_error("esm_sweep_block: block must be 64, 128 or 256"); return ESM_EINVAL; }
    h->rand_block = block;
    return ESM_OK;
}

int32_t esm_sweep_pad(esm_handle *h, int32_t pad) {
    if (!h) { set_error("esm_sweep_pad: null handle"); return ESM_EINVAL; }
    if (pad != 0 && pad != 1) { set_error("esm_sweep_pad: pad must be 0 or 1"); return ESM_EINVAL; }
    h->rand_pad = pad != 0;
    return ESM_OK;
}

int32_t esm_autotune(esm_handle *h, int32_t nsweeps, const uint32_t key[2], uint32_t counter, double maxnats,
                     int32_t *best, double *ms, int32_t *ncand) {
    if (!h || !key || !best) { set_error("esm_autotune: null argument"); return ESM_EINVAL; }
    if (nsweeps < 1 || nsweeps > 64 || !(maxnats > 0)) { set_error("esm_autotune: nsweeps must be in [1, 64] and maxnats > 0"); return ESM_EINVAL; }
    if (h->prof) { set_error("esm_autotune: kernel profiling is on (esm_profile)"); return ESM_EINVAL; }
    // candidates c: with R a multiple of 64, c = 3·f + b over the forms 0 (a launch per family),
    // 1 (merged), 2 and 4 (merged, waves per step) and the one-wave launches' blocks 64, 128, 256;
    // otherwise c = 6·p + 3·f + b over padding p (0: R threads per step, 1: whole waves per
    // step), forms 0 and 1 and the blocks
    static const int forms[4] = {0, 1, 2, 4}, blocks[3] = {64, 128, 256};
    const bool r64 = h->R % 64 == 0;
    const int nc = 12;
    auto form_of = [&](int c) { return r64 ? forms[c / 3] : forms[(c % 6) / 3]; };
    auto pad_of = [&](int c) { return r64 ? 0 : c / 6; };
    const size_t TR = (size_t)h->T * h->R;
    const size_t nv = (size_t)ESM_NV * h->N * TR * h->es, nr = (size_t)h->N * TR * h->es, nF = (size_t)h->L * TR * h->es;
    void *save = nullptr;
    ESM_HIP_CHECK(hipMallocAsync(&save, nv + nr + nF, h->stream));
    auto snap = [&](bool to_save) -> hipError_t {
        char *p = (char *)save;
        void *dst[3] = {p, p + nv, p + nv + nr}, *src[3] = {h->d_v, h->d_r, h->d_F};
        const size_t len[3] = {nv, nr, nF};
        for (int i = 0; i < 3; i++) {
            const hipError_t e = to_save ? hipMemcpyAsync(dst[i], src[i], len[i], hipMemcpyDeviceToDevice, h->stream)
                                         : hipMemcpyAsync(src[i], dst[i], len[i], hipMemcpyDeviceToDevice, h->stream);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    const bool merged0 = h->sweep_merged, pad0 = h->rand_pad;
    const int waves0 = h->sweep_waves, block0 = h->rand_block;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    double tbest = 0.0;
    int ibest = -1;
    auto fail = [&](const char *what) {
        set_error(std::string("esm_autotune: ") + what + ": " + hipGetErrorString(hipGetLastError()));
        rc = ESM_EHIP;
    };
    if (snap(true) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        fail("saving the state");
    for (int c = 0; c < nc && !rc; c++) {
        const int f = form_of(c);
        h->sweep_merged = f != 0;
        h->sweep_waves = f > 1 ? f : 0;
        h->rand_block = blocks[c % 3];
        h->rand_pad = pad_of(c) != 0;
        float t = 0.f;
        // one untimed sweep (code objects, caches), then the timed ones from the same state
        if ((rc = sweeps_on_stream(h, 1, key, counter, maxnats))) break;
        if (snap(false) != hipSuccess || hipEventRecord(e0, h->stream) != hipSuccess) { fail("restoring the state"); break; }
        if ((rc = sweeps_on_stream(h, nsweeps, key, counter, maxnats))) break;
        if (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess || snap(false) != hipSuccess) {
            fail("timing a candidate");
            break;
        }
        const double tm = (double)t / nsweeps;
        if (ms) ms[c] = tm;
        if (ibest < 0 || tm < tbest) { tbest = tm; ibest = c; }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (save) {
        (void)snap(false);
        (void)hipFreeAsync(save, h->stream);
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess && !rc) fail("stream synchronization");
    h->scores_valid = false;
    if (rc) {
        h->sweep_merged = merged0;
        h->sweep_waves = waves0;
        h->rand_block = block0;
        h->rand_pad = pad0;
        return rc;
    }
    const int f = form_of(ibest);
    h->sweep_merged = f != 0;
    h->sweep_waves = f > 1 ? f : 0;
    h->rand_block = blocks[ibest % 3];
    h->rand_pad = pad_of(ibest) != 0;
    best[0] = f;
    best[1] = h->rand_block;
    best[2] = pad_of(ibest);
    if (ncand) *ncand = nc;
    return ESM_OK;
}

}  // extern "C"
