// Batched circuits: host planning, descriptor image, launch and hipGraph
// replay (see qsim_batch.h).
#include "qsim_batch.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/qgpusim.h"

using namespace qsim;

struct qsim_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int K_by_n[41];            // chunk bits of streamed circuits, by qubit count
    int gang_by_n[41];         // workgroups per streamed circuit, by qubit count
    int fuse_by_n[41];         // forward sweep fused into dense 16x16 steps (1) or slot steps (0)
    int blocks_per_cu = 0;     // 0: occupancy maximum
    int ncirc = 0;
    int ntasks = 0;
    int ngangs = 0;
    int64_t nprobe_total = 0;
    int64_t passes_total = 0;
    char *d_desc = nullptr;
    size_t desc_bytes = 0;
    cx<float> *d_scratch = nullptr;
    size_t scratch_elems = 0;
    unsigned *d_counter = nullptr;
    unsigned *d_gang_bar = nullptr;
    int32_t *d_gang_owner = nullptr;
    double *d_gang_part = nullptr;
    double *d_expect = nullptr;
    double *d_probes = nullptr;
    uint64_t *d_times = nullptr;
    BatchArgs args{};
    int blocks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double tick_hz = 0.0;
};

namespace {

#define HIP_OK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return set_error(std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

void free_buffers(qsim_batch *b)
{
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    void *bufs[] = {b->d_desc, b->d_scratch, b->d_counter, b->d_gang_bar, b->d_gang_owner, b->d_gang_part,
                    b->d_expect, b->d_probes, b->d_times};
    for (void *p : bufs)
        if (p) hipFree(p);
    b->d_desc = nullptr;
    b->d_scratch = nullptr;
    b->d_counter = nullptr;
    b->d_gang_bar = nullptr;
    b->d_gang_owner = nullptr;
    b->d_gang_part = nullptr;
    b->d_expect = nullptr;
    b->d_probes = nullptr;
    b->d_times = nullptr;
    b->desc_bytes = b->scratch_elems = 0;
}

std::vector<PlanOp> plan_ops(const qsim_op *ops, int64_t n)
{
    std::vector<PlanOp> v((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        v[i].kind = ops[i].kind;
        v[i].q0 = ops[i].q0;
        v[i].q1 = ops[i].kind == QSIM_OP_2Q ? ops[i].q1 : -1;
        v[i].mat_off = ops[i].mat_offset;
        v[i].probe = ops[i].kind == QSIM_OP_PROBE ? ops[i].q1 : -1;
    }
    return v;
}

// Append `p` to the combined image, rebasing its step, probe and matrix
// offsets.  Returns the index of its first pass.
int append_plan(const Plan &p, std::vector<PassDev> &passes, std::vector<StepDev> &steps,
                std::vector<int32_t> &ids, std::vector<cx<float>> &mats)
{
    if (!p.fdiag.empty()) return -1;  // free diagonals need specialised kernels
    const int pass0 = (int)passes.size();
    const int step0 = (int)steps.size(), id0 = (int)ids.size(), mat0 = (int)mats.size();
    for (PassDev d : p.passes) {
        d.step_off += step0;
        d.probe_off += id0;
        d.mat_lo += mat0;
        d.mat_hi += mat0;
        passes.push_back(d);
    }
    for (StepDev s : p.steps) {
        for (int k = 0; k < kSlots; ++k) s.mat[k] += mat0;
        if (s.dense >= 0) s.dense += mat0;
        if (s.scale >= 0) s.scale += mat0;
        steps.push_back(s);
    }
    ids.insert(ids.end(), p.probe_ids.begin(), p.probe_ids.end());
    for (size_t i = 0; i + 1 < p.mats.size(); i += 2) mats.push_back(cx<float>{(float)p.mats[i], (float)p.mats[i + 1]});
    return pass0;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

extern "C" {

int qsim_batch_create(qsim_batch **out)
{
    if (!out) return set_error("qsim_batch_create: null output");
    *out = nullptr;
    qsim_batch *b = new qsim_batch();
    for (int n = 0; n <= 40; ++n) {
        b->K_by_n[n] = kBatchK;
        b->gang_by_n[n] = n > 20 ? std::min(kBatchMaxGang, 1 << (n - 20)) : 1;
        b->fuse_by_n[n] = 1;
    }
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    int khz = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device);
    if (e != hipSuccess) {
        delete b;
        return set_error(std::string("HIP error: ") + hipGetErrorString(e));
    }
    b->tick_hz = 1e3 * khz;
    *out = b;
    return 0;
}

int qsim_batch_destroy(qsim_batch *b)
{
    if (!b) return 0;
    free_buffers(b);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int qsim_batch_set_tuning(qsim_batch *b, const int32_t *K_by_n, int nk, int blocks_per_cu)
{
    if (!b) return set_error("null batch");
    for (int n = 0; n < nk && n <= 40; ++n) {
        if (K_by_n[n] < 6 || K_by_n[n] > kBatchK) return set_error("batch chunk bits must be in [6, 12]");
        b->K_by_n[n] = K_by_n[n];
    }
    if (blocks_per_cu < 0) return set_error("blocks_per_cu must be >= 0");
    b->blocks_per_cu = blocks_per_cu;
    return 0;
}

int qsim_batch_set_gang(qsim_batch *b, const int32_t *gang_by_n, int nk)
{
    if (!b || !gang_by_n) return set_error("null argument");
    for (int n = 0; n < nk && n <= 40; ++n)
        if (gang_by_n[n] < 1 || gang_by_n[n] > kBatchMaxGang)
            return set_error("batch workgroups per circuit must be in [1, 64]");
    for (int n = 0; n < nk && n <= 40; ++n) b->gang_by_n[n] = gang_by_n[n];
    return 0;
}

int qsim_batch_set_fusion(qsim_batch *b, const int32_t *fuse_by_n, int nk)
{
    if (!b || !fuse_by_n) return set_error("null argument");
    for (int n = 0; n < nk && n <= 40; ++n)
        if (fuse_by_n[n] != 0 && fuse_by_n[n] != 1) return set_error("batch fusion mode must be 0 or 1");
    for (int n = 0; n < nk && n <= 40; ++n) b->fuse_by_n[n] = fuse_by_n[n];
    return 0;
}

int qsim_batch_load(qsim_batch *b, int ncirc, const int32_t *num_qubits, const uint64_t *zmask,
                    const qsim_op *fwd_ops, const int64_t *fwd_off, const qsim_op *rev_ops, const int64_t *rev_off,
                    const void *mats, uint64_t mat_count, const int32_t *nprobes)
{
    if (!b) return set_error("null batch");
    if (ncirc < 1) return set_error("batch: no circuits");
    if (!num_qubits || !zmask || !fwd_ops || !fwd_off || !rev_ops || !rev_off || !mats || !nprobes)
        return set_error("batch: null argument");
    free_buffers(b);
    std::vector<BatchCirc> circs((size_t)ncirc);
    std::vector<PassDev> passes;
    std::vector<StepDev> steps;
    std::vector<int32_t> ids;
    std::vector<cx<float>> cmats;
    std::vector<double> cost((size_t)ncirc);
    // Validate and plan every circuit on worker threads (the planner keeps no
    // shared state); the first error in circuit order is reported.
    std::vector<Plan> pf((size_t)ncirc), pr((size_t)ncirc);
    std::vector<std::string> errs((size_t)ncirc);
    auto plan_one = [&](int i) -> std::string {
        const int n = num_qubits[i];
        if (n < 6 || n > 34) return "batch: qubit counts must be in [6, 34]";
        if (nprobes[i] < 0) return "batch: negative probe count";
        if (fwd_off[i] < 0 || rev_off[i] < 0) return "batch: op offsets must be non-negative";
        const int64_t nf = fwd_off[i + 1] - fwd_off[i], nr = rev_off[i + 1] - rev_off[i];
        if (nf < 0 || nr < 0) return "batch: op offsets must be non-decreasing";
        const qsim_op *fo = fwd_ops + fwd_off[i], *ro = rev_ops + rev_off[i];
        for (int64_t k = 0; k < nf; ++k)
            if (fo[k].kind == QSIM_OP_PROBE) return "batch: probe in a forward list";
        for (int pass = 0; pass < 2; ++pass) {
            const qsim_op *o = pass ? ro : fo;
            const int64_t m = pass ? nr : nf;
            for (int64_t k = 0; k < m; ++k) {
                const bool two = o[k].kind == QSIM_OP_2Q;
                if (o[k].kind != QSIM_OP_1Q && o[k].kind != QSIM_OP_2Q && o[k].kind != QSIM_OP_PROBE)
                    return "batch: unknown op kind";
                if (o[k].q0 < 0 || o[k].q0 >= n || (two && (o[k].q1 < 0 || o[k].q1 >= n || o[k].q1 == o[k].q0)))
                    return "batch: invalid qubit index";
                if (o[k].kind == QSIM_OP_PROBE && (o[k].q1 < 0 || o[k].q1 >= nprobes[i]))
                    return "batch: probe slot out of range";
                if (o[k].mat_offset < 0 || (uint64_t)o[k].mat_offset + (two ? 16 : 4) > mat_count)
                    return "batch: matrix offset out of range";
            }
        }
        const int K = n <= kBatchK ? n : b->K_by_n[std::min(n, 40)];
        std::string err;
        std::vector<PlanOp> vf = plan_ops(fo, nf), vr = plan_ops(ro, nr);
        // the batch kernel runs the generic slot forms
        if (!plan_passes(n, vf.data(), (int)nf, mats, 8, K, kBatchCmin, 1, pf[i], err, b->fuse_by_n[std::min(n, 40)],
                         nullptr, nullptr, true) ||
            !plan_passes(n, vr.data(), (int)nr, mats, 8, K, kBatchCmin, kBatchProbes, pr[i], err, 0, nullptr, nullptr,
                         true))
            return "batch circuit " + std::to_string(i) + ": " + err;
        return std::string();
    };
    {
        const int nthreads = std::max(1, std::min({ncirc, 16, (int)std::thread::hardware_concurrency()}));
        std::atomic<int> next{0};
        auto worker = [&]() {
            for (int i = next++; i < ncirc; i = next++) errs[i] = plan_one(i);
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker);
        worker();
        for (auto &t : pool) t.join();
        for (int i = 0; i < ncirc; ++i)
            if (!errs[i].empty()) return set_error(errs[i]);
    }
    int nmax_streamed = 0;
    int64_t probe0 = 0;
    for (int i = 0; i < ncirc; ++i) {
        const int n = num_qubits[i];
        const bool resident = n <= kBatchK;
        BatchCirc &C = circs[i];
        C.n = n;
        C.K = resident ? n : b->K_by_n[std::min(n, 40)];
        C.fwd0 = append_plan(pf[i], passes, steps, ids, cmats);
        C.nfwd = (int)pf[i].passes.size();
        C.rev0 = append_plan(pr[i], passes, steps, ids, cmats);
        if (C.fwd0 < 0 || C.rev0 < 0) return set_error("internal: batch plan with free diagonals");
        C.nrev = (int)pr[i].passes.size();
        C.probe0 = (int32_t)probe0;
        C.nprobe = nprobes[i];
        C.zmask = zmask[i];
        probe0 += nprobes[i];
        if (!resident) nmax_streamed = std::max(nmax_streamed, n);
        // cost model for the queue order: amplitudes touched (two vectors backwards)
        cost[i] = (double)(1ull << n) * (C.nfwd + 2.0 * C.nrev + 2.0) * (resident ? 0.25 : 1.0);
        pf[i] = Plan();  // appended: release the circuit's own copy
        pr[i] = Plan();
    }
    std::vector<int32_t> order((size_t)ncirc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });

    int cus = 0;
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    const int per_cu = b->blocks_per_cu > 0 ? std::min(b->blocks_per_cu, batch_blocks_per_cu_max())
                                            : batch_blocks_per_cu_max();
    b->scratch_elems = nmax_streamed ? (size_t(1) << nmax_streamed) : 0;
    int blocks = std::max(1, cus * per_cu);
    // keep the HBM slots (psi and lam per workgroup) within a budget: fewer
    // workgroups for very large streamed states
    const size_t slot_bytes = 2 * b->scratch_elems * sizeof(cx<float>);
    const size_t budget = size_t(64) << 30;
    if (slot_bytes && slot_bytes * (size_t)blocks > budget) blocks = (int)std::max<size_t>(1, budget / slot_bytes);

    // tasks: one per circuit, or one per rank of a gang (at most a quarter of
    // the workgroups, so a gang never waits on more ranks than can run)
    std::vector<BatchTask> tasks;
    int64_t part_total = 0;
    int ngangs = 0;
    const int gang_cap = std::max(1, blocks / 4);
    for (int i : order) {
        BatchCirc &C = circs[i];
        const int G = C.K == C.n ? 1 : std::min(b->gang_by_n[std::min(C.n, 40)], gang_cap);
        C.gang = G;
        C.gslot = -1;
        C.part0 = 0;
        if (G > 1) {
            C.gslot = ngangs++;
            C.part0 = part_total;
            part_total += (int64_t)G * (1 + C.nprobe);
        }
        for (int r = 0; r < G; ++r) tasks.push_back(BatchTask{i, r});
    }
    b->blocks = std::max(1, std::min((int)tasks.size(), blocks));

    // one device image: circuits | tasks | passes | steps | probe ids | matrices
    const size_t o_circ = 0;
    const size_t o_task = align256(o_circ + circs.size() * sizeof(BatchCirc));
    const size_t o_pass = align256(o_task + tasks.size() * sizeof(BatchTask));
    const size_t o_step = align256(o_pass + passes.size() * sizeof(PassDev));
    const size_t o_ids = align256(o_step + steps.size() * sizeof(StepDev));
    const size_t o_mats = align256(o_ids + ids.size() * sizeof(int32_t));
    const size_t total = align256(o_mats + cmats.size() * sizeof(cx<float>)) + 256;
    std::vector<char> img(total, 0);
    std::memcpy(&img[o_circ], circs.data(), circs.size() * sizeof(BatchCirc));
    std::memcpy(&img[o_task], tasks.data(), tasks.size() * sizeof(BatchTask));
    if (!passes.empty()) std::memcpy(&img[o_pass], passes.data(), passes.size() * sizeof(PassDev));
    if (!steps.empty()) std::memcpy(&img[o_step], steps.data(), steps.size() * sizeof(StepDev));
    if (!ids.empty()) std::memcpy(&img[o_ids], ids.data(), ids.size() * sizeof(int32_t));
    if (!cmats.empty()) std::memcpy(&img[o_mats], cmats.data(), cmats.size() * sizeof(cx<float>));

    HIP_OK(hipMalloc(reinterpret_cast<void **>(&b->d_desc), total));
    b->desc_bytes = total;
    if (b->scratch_elems)
        HIP_OK(hipMalloc(reinterpret_cast<void **>(&b->d_scratch),
                         2 * b->scratch_elems * sizeof(cx<float>) * (size_t)b->blocks));
    HIP_OK(hipMalloc(reinterpret_cast<void **>(&b->d_counter), 256));
    HIP_OK(hipMalloc(reinterpret_cast<void **>(&b->d_gang_bar), sizeof(unsigned) * std::max(1, ngangs)));
    HIP_OK(hipMalloc(reinterpret_cast<void **>(&b->d_gang_owner), sizeof(int32_t) * std::max(1, ngangs)));
    HIP_OK(hipMalloc(reinterpret_cast<void **>(&b->d_gang_part), sizeof(double) * std::max<int64_t>(1, part_total)));
    HIP_OK(hipMalloc(reinterpret_cast<void **>(&b->d_expect), sizeof(double) * ncirc));
    HIP_OK(hipMalloc(reinterpret_cast<void **>(&b->d_probes), sizeof(double) * std::max<int64_t>(1, probe0)));
    HIP_OK(hipMalloc(reinterpret_cast<void **>(&b->d_times), sizeof(uint64_t) * 2 * ncirc));
    HIP_OK(hipMemcpyAsync(b->d_desc, img.data(), total, hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));

    BatchArgs &a = b->args;
    a.circs = reinterpret_cast<const BatchCirc *>(b->d_desc + o_circ);
    a.tasks = reinterpret_cast<const BatchTask *>(b->d_desc + o_task);
    a.ncirc = ncirc;
    a.ntasks = (int32_t)tasks.size();
    a.ngangs = ngangs;
    a.passes = reinterpret_cast<const PassDev *>(b->d_desc + o_pass);
    a.steps = reinterpret_cast<const StepDev *>(b->d_desc + o_step);
    a.probe_ids = reinterpret_cast<const int32_t *>(b->d_desc + o_ids);
    a.mats = reinterpret_cast<const cx<float> *>(b->d_desc + o_mats);
    a.scratch = b->d_scratch;
    a.slot_elems = b->scratch_elems;
    a.counter = b->d_counter;
    a.gang_bar = b->d_gang_bar;
    a.gang_owner = b->d_gang_owner;
    a.gang_part = b->d_gang_part;
    a.part_total = part_total;
    a.probe_total = probe0;
    a.expect = b->d_expect;
    a.probes = b->d_probes;
    a.times = b->d_times;
    b->ncirc = ncirc;
    b->ntasks = (int)tasks.size();
    b->ngangs = ngangs;
    b->nprobe_total = probe0;
    b->passes_total = (int64_t)passes.size();
    return 0;
}

int qsim_batch_run(qsim_batch *b, int use_graph, double *expect_out, double *probes_out, double *exec_ms_out,
                   double *total_ms_out)
{
    if (!b || !b->d_desc) return set_error("batch: nothing loaded");
    struct Events {  // destroyed on every return
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events()
        {
            if (e0) hipEventDestroy(e0);
            if (e1) hipEventDestroy(e1);
        }
    } ev;
    hipEvent_t &e0 = ev.e0, &e1 = ev.e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, b->stream));
    if (use_graph) {
        if (!b->exec) {
            HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t r1 = launch_batch_reset(b->args, b->stream);
            hipError_t r2 = launch_batch(b->args, b->blocks, b->stream);
            hipGraph_t g = nullptr;
            hipError_t r3 = hipStreamEndCapture(b->stream, &g);
            if (r1 != hipSuccess || r2 != hipSuccess || r3 != hipSuccess) {
                if (g) hipGraphDestroy(g);
                return set_error("batch: graph capture failed");
            }
            b->graph = g;
            HIP_OK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        }
        HIP_OK(hipGraphLaunch(b->exec, b->stream));
    } else {
        HIP_OK(launch_batch_reset(b->args, b->stream));
        HIP_OK(launch_batch(b->args, b->blocks, b->stream));
    }
    HIP_OK(hipEventRecord(e1, b->stream));
    std::vector<uint64_t> times(2 * (size_t)b->ncirc);
    if (expect_out)
        HIP_OK(hipMemcpyAsync(expect_out, b->d_expect, sizeof(double) * b->ncirc, hipMemcpyDeviceToHost, b->stream));
    if (probes_out && b->nprobe_total)
        HIP_OK(hipMemcpyAsync(probes_out, b->d_probes, sizeof(double) * b->nprobe_total, hipMemcpyDeviceToHost,
                              b->stream));
    if (exec_ms_out)
        HIP_OK(hipMemcpyAsync(times.data(), b->d_times, sizeof(uint64_t) * times.size(), hipMemcpyDeviceToHost,
                              b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    if (total_ms_out) *total_ms_out = ms;
    if (exec_ms_out)
        for (int i = 0; i < b->ncirc; ++i)
            exec_ms_out[i] = (double)(times[2 * i + 1] - times[2 * i]) / b->tick_hz * 1e3;
    return 0;
}

int qsim_batch_info(const qsim_batch *b, int64_t *out, int n)
{
    if (!b || !out) return set_error("null argument");
    const int64_t v[7] = {b->ncirc, b->nprobe_total, b->passes_total, b->blocks,
                          (int64_t)(2 * b->scratch_elems * sizeof(cx<float>) * (size_t)b->blocks), b->ntasks,
                          b->ngangs};
    const int m = std::min(n, 7);
    for (int i = 0; i < m; ++i) out[i] = v[i];
    return m;
}

}  // extern "C"
