// C ABI of the statevector backend: state objects, validation, dispatch of
// the tile modes, hipGraph cache, descriptor uploads, reductions, adjoint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <list>
#include <memory>
#include <string>
#include <vector>

#include "../../include/qgpusim.h"
#include "qsim_internal.h"
#include "qsim_jit.h"
#include "qsim_plan.h"
#include "qsim_sample.h"

using namespace qsim;

namespace {

thread_local std::string g_err;

int fail(const std::string &msg)
{
    g_err = msg;
    return -1;
}

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) return fail(std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

// Pinned staging + device ring for descriptor uploads on one stream.  Host
// regions are never reused before the stream has drained past their copy: a
// wrap synchronises the stream first.
struct Uploader {
    char *host = nullptr;
    char *dev = nullptr;
    size_t cap = 0, head = 0;

    void release()
    {
        if (host) hipHostFree(host);
        if (dev) hipFree(dev);
        host = dev = nullptr;
        cap = head = 0;
    }

    // Returns the device address holding a copy of src[0..bytes).
    int upload(const void *src, size_t bytes, hipStream_t s, void **out)
    {
        bytes = (bytes + 255) & ~size_t(255);
        if (bytes > cap) {
            HIP_TRY(hipStreamSynchronize(s));
            release();
            cap = std::max(bytes * 2, size_t(4) << 20);
            HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&host), cap, hipHostMallocDefault));
            HIP_TRY(hipMalloc(reinterpret_cast<void **>(&dev), cap));
            head = 0;
        }
        if (head + bytes > cap) {
            HIP_TRY(hipStreamSynchronize(s));
            head = 0;
        }
        std::memcpy(host + head, src, bytes);
        HIP_TRY(hipMemcpyAsync(dev + head, host + head, bytes, hipMemcpyHostToDevice, s));
        *out = dev + head;
        head += bytes;
        return 0;
    }
};

// Flat descriptor image: passes | steps | probe ids | matrices (state precision).
struct DescImage {
    std::vector<char> bytes;
    size_t off_pass = 0, off_step = 0, off_ids = 0, off_mats = 0;

    template <typename R>
    void build(const Plan &p)
    {
        auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
        const size_t nm = p.mats.size() / 2;
        off_pass = 0;
        off_step = al(off_pass + p.passes.size() * sizeof(PassDev));
        off_ids = al(off_step + p.steps.size() * sizeof(StepDev));
        off_mats = al(off_ids + p.probe_ids.size() * sizeof(int32_t));
        // the raw words follow the matrices at Plan::raw_offset from them
        const size_t off_raw = off_mats + p.raw_offset(sizeof(cx<R>));
        bytes.assign(al(off_raw + p.raw.size() * sizeof(uint32_t)) + 256, 0);
        if (!p.raw.empty()) std::memcpy(&bytes[off_raw], p.raw.data(), p.raw.size() * sizeof(uint32_t));
        if (!p.passes.empty()) std::memcpy(&bytes[off_pass], p.passes.data(), p.passes.size() * sizeof(PassDev));
        if (!p.steps.empty()) std::memcpy(&bytes[off_step], p.steps.data(), p.steps.size() * sizeof(StepDev));
        if (!p.probe_ids.empty())
            std::memcpy(&bytes[off_ids], p.probe_ids.data(), p.probe_ids.size() * sizeof(int32_t));
        cx<R> *m = reinterpret_cast<cx<R> *>(&bytes[off_mats]);
        for (size_t i = 0; i < nm; ++i) m[i] = cx<R>{(R)p.mats[2 * i], (R)p.mats[2 * i + 1]};
    }
};

struct GraphEntry {
    std::vector<char> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *desc = nullptr;
    int swaps = 0;  // passes writing the other buffer (folded layout changes): the buffers swap this often
    int folds = 0;  // layout changes those passes carry (a single pass can carry two)
    bool post_done = false;      // its last pass writes the layout change that follows the list
    bool after_pending = false;  // stored layout after the replay (perm_map)
    int32_t after_map[64];
};

size_t amp_bytes(int dtype) { return dtype == QSIM_C64 ? 8 : 16; }

}  // namespace

int qsim::set_error(const std::string &msg) { return fail(msg); }

struct qsim_state {
    int n = 0;
    int dtype = QSIM_C64;
    uint64_t dim = 0;
    int device = 0;
    void *psi = nullptr;
    void *tmp = nullptr;
    hipStream_t stream = nullptr;
    int K = 12, cmin = 4;
    Uploader up;
    std::list<GraphEntry> graphs;  // most recent first
    double *d_small = nullptr;     // reduction scratch (partials)
    size_t small_cap = 0;          // doubles
    double *d_probe = nullptr;     // probe partials
    size_t probe_cap = 0;          // doubles
    void *d_gatek = nullptr;       // matrix of a 3- or 4-qubit op in the sequential mode (16x16 complex128)
    // kernel launches, graph launches, graph captures, chunk passes,
    // permutations run, chunk passes run by specialised kernels, permutations
    // folded into a chunk pass
    int64_t counters[7] = {0, 0, 0, 0, 0, 0, 0};
    // The stored bit order may differ from the logical one (the layout the
    // caller's permutations define): perm_map[p] = logical bit position held
    // by stored bit p.  qsim_permute only composes it; op lists run on the
    // stored order (their qubits mapped through it), moving op qubits into
    // the low bits inside their first chunk pass; readers apply it first.
    bool perm_pending = false;  // perm_map is not the identity
    int32_t perm_map[64];
    // The state is |0...0> but not written yet (reset): the next chunked op
    // list's first pass generates it; any other use writes it first.
    bool zero_pending = false;
    bool jit = true;               // specialise forward chunk passes at run time
    int fuse = 0;                  // fusion mode (qsim_set_fusion)
    double fuse_bar = 0.0, fuse_bar5 = 0.0;  // its cost bars (qsim_set_fusion_bars; 0: defaults)
    std::string arch;              // device gcnArchName (for run-time compiles)
    // per-kernel-class timing with events on `stream` (off by default)
    bool prof = false, capturing = false;
    struct Rec {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
};

namespace {

constexpr size_t kMaxGraphs = 64;
constexpr int kProfClasses = 4;  // 0 chunk pass, 1 permutation, 2 gate stream, 3 Pauli chunk

hipEvent_t take_event(qsim_state *st)
{
    if (!st->pool.empty()) {
        hipEvent_t e = st->pool.back();
        st->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets the launches issued during its lifetime with two events.
struct ProfScope {
    qsim_state *st;
    int cls;
    hipEvent_t a = nullptr;
    ProfScope(qsim_state *s, int c) : st(s), cls(c)
    {
        if (st->prof && !st->capturing) {
            a = take_event(st);
            if (a) hipEventRecord(a, st->stream);
        }
    }
    ~ProfScope()
    {
        if (!a) return;
        hipEvent_t b = take_event(st);
        if (!b) return;
        hipEventRecord(b, st->stream);
        st->recs.push_back({cls, a, b});
    }
};

int ensure_tmp(qsim_state *st)
{
    if (!st->tmp) HIP_TRY(hipMalloc(&st->tmp, st->dim * amp_bytes(st->dtype)));
    return 0;
}

int build_perm_args(int n, const int32_t *map, int smax, PermArgs &P);

// Writes a pending |0...0>.
int materialize_zero(qsim_state *st)
{
    if (!st->zero_pending) return 0;
    if (st->dtype == QSIM_C64)
        HIP_TRY(launch_init_basis<float>(static_cast<cx<float> *>(st->psi), st->dim, 0, st->stream));
    else
        HIP_TRY(launch_init_basis<double>(static_cast<cx<double> *>(st->psi), st->dim, 0, st->stream));
    st->zero_pending = false;
    st->counters[0]++;
    return 0;
}

// Brings the buffer to the logical state: writes a pending |0...0>, applies a
// stored order that differs from the logical one with the permutation kernel.
int flush_perm(qsim_state *st)
{
    if (materialize_zero(st)) return -1;
    if (!st->perm_pending) return 0;
    if (ensure_tmp(st)) return -1;
    PermArgs P;
    build_perm_args(st->n, st->perm_map, st->dtype == QSIM_C64 ? 12 : 11, P);
    {
        ProfScope ps(st, 1);
        if (st->dtype == QSIM_C64)
            HIP_TRY(launch_permute<float>(static_cast<cx<float> *>(st->psi), static_cast<cx<float> *>(st->tmp), P,
                                          st->stream));
        else
            HIP_TRY(launch_permute<double>(static_cast<cx<double> *>(st->psi), static_cast<cx<double> *>(st->tmp), P,
                                           st->stream));
    }
    std::swap(st->psi, st->tmp);
    st->perm_pending = false;
    st->counters[0]++;
    st->counters[4]++;
    return 0;
}

int ensure_small(qsim_state *st, size_t doubles)
{
    if (doubles <= st->small_cap) return 0;
    HIP_TRY(hipStreamSynchronize(st->stream));
    if (st->d_small) hipFree(st->d_small);
    st->d_small = nullptr;
    HIP_TRY(hipMalloc(reinterpret_cast<void **>(&st->d_small), doubles * sizeof(double)));
    st->small_cap = doubles;
    return 0;
}

int ensure_probe(qsim_state *st, size_t doubles)
{
    if (doubles <= st->probe_cap) return 0;
    HIP_TRY(hipStreamSynchronize(st->stream));
    if (st->d_probe) hipFree(st->d_probe);
    if (st->d_gatek) hipFree(st->d_gatek);
    st->d_probe = nullptr;
    HIP_TRY(hipMalloc(reinterpret_cast<void **>(&st->d_probe), doubles * sizeof(double)));
    st->probe_cap = doubles;
    return 0;
}

void drop_graphs(qsim_state *st)
{
    if (!st->graphs.empty()) hipStreamSynchronize(st->stream);
    for (auto &g : st->graphs) {
        if (g.exec) hipGraphExecDestroy(g.exec);
        if (g.graph) hipGraphDestroy(g.graph);
        if (g.desc) hipFree(g.desc);
    }
    st->graphs.clear();
}

// Qubits of a QSIM_OP_3Q / QSIM_OP_4Q op: q0, then q1's 8-bit fields.
int kq_qubits(const qsim_op &o, int *q)
{
    const int k = o.kind == QSIM_OP_3Q ? 3 : 4;
    q[0] = o.q0;
    for (int j = 1; j < k; ++j) q[j] = (int)(((uint32_t)o.q1 >> (8 * (j - 1))) & 255u);
    return k;
}

int32_t kq_pack(const int *q, int k)
{
    uint32_t v = 0;
    for (int j = 1; j < k; ++j) v |= (uint32_t)q[j] << (8 * (j - 1));
    return (int32_t)v;
}

int check_ops(int n, const qsim_op *ops, int nops, uint64_t mat_count, bool allow_probe)
{
    if (nops < 0) return fail("apply_tile_dev: negative op count");
    if (nops > 0 && !ops) return fail("apply_tile_dev: null op array");
    for (int i = 0; i < nops; ++i) {
        const qsim_op &o = ops[i];
        if (o.kind == QSIM_OP_1Q || (allow_probe && o.kind == QSIM_OP_PROBE)) {
            if (o.q0 < 0 || o.q0 >= n) return fail("apply_tile_dev: invalid q0");
            if (o.mat_offset < 0 || (uint64_t)o.mat_offset + 4 > mat_count)
                return fail("apply_tile_dev: matrix offset out of range");
        } else if (o.kind == QSIM_OP_2Q) {
            if (o.q0 < 0 || o.q0 >= n || o.q1 < 0 || o.q1 >= n) return fail("apply_tile_dev: invalid q0/q1");
            if (o.q0 == o.q1) return fail("apply_tile_dev: q0 == q1");
            if (o.mat_offset < 0 || (uint64_t)o.mat_offset + 16 > mat_count)
                return fail("apply_tile_dev: matrix offset out of range");
        } else if (o.kind == QSIM_OP_3Q || o.kind == QSIM_OP_4Q) {
            int q[4];
            const int k = kq_qubits(o, q);
            if (o.q1 < 0 || (k == 3 && ((uint32_t)o.q1 >> 16) != 0u) || ((uint32_t)o.q1 >> 24) != 0u)
                return fail("apply_tile_dev: invalid qubit field");
            uint64_t seen = 0;
            for (int j = 0; j < k; ++j) {
                if (q[j] < 0 || q[j] >= n) return fail("apply_tile_dev: invalid qubit");
                if ((seen >> q[j]) & 1) return fail("apply_tile_dev: repeated qubit");
                seen |= 1ull << q[j];
            }
            if (o.mat_offset < 0 || (uint64_t)o.mat_offset + (1u << (2 * k)) > mat_count)
                return fail("apply_tile_dev: matrix offset out of range");
        } else {
            return fail("apply_tile_dev: unknown op kind: " + std::to_string(o.kind));
        }
    }
    return 0;
}

std::vector<PlanOp> to_plan_ops(const qsim_op *ops, int nops)
{
    std::vector<PlanOp> v(nops);
    for (int i = 0; i < nops; ++i) {
        v[i].kind = ops[i].kind;
        v[i].q0 = ops[i].q0;
        v[i].q1 = ops[i].kind == QSIM_OP_2Q ? ops[i].q1 : -1;
        v[i].mat_off = ops[i].mat_offset;
        v[i].probe = ops[i].kind == QSIM_OP_PROBE ? ops[i].q1 : -1;
        if (ops[i].kind == QSIM_OP_3Q || ops[i].kind == QSIM_OP_4Q) {
            int q[4];
            const int k = kq_qubits(ops[i], q);
            v[i].q1 = q[1];
            v[i].q2 = q[2];
            v[i].q3 = k == 4 ? q[3] : -1;
        }
    }
    return v;
}

template <typename R>
int run_sequential(qsim_state *st, const qsim_op *ops, int nops, const void *mats)
{
    const cx<R> *m = static_cast<const cx<R> *>(mats);
    cx<R> *psi = static_cast<cx<R> *>(st->psi);
    for (int i = 0; i < nops; ++i) {
        const qsim_op &o = ops[i];
        ProfScope ps(st, 2);
        if (o.kind == QSIM_OP_1Q) {
            Mat1<R> U;
            std::memcpy(U.u, m + o.mat_offset, sizeof(U));
            HIP_TRY(launch_1q<R>(psi, st->n, o.q0, U, st->stream));
        } else if (o.kind == QSIM_OP_3Q || o.kind == QSIM_OP_4Q) {
            // the matrix goes to a device buffer (stream-ordered copy: the
            // previous op's kernel has read it before it is overwritten)
            int q[4];
            const int k = kq_qubits(o, q);
            if (!st->d_gatek) HIP_TRY(hipMalloc(&st->d_gatek, 256 * sizeof(cx<double>)));
            HIP_TRY(hipMemcpyAsync(st->d_gatek, m + o.mat_offset, sizeof(cx<R>) << (2 * k), hipMemcpyHostToDevice,
                                   st->stream));
            HIP_TRY(launch_gate_k<R>(psi, st->n, q, k, static_cast<const cx<R> *>(st->d_gatek), st->stream));
        } else {
            Mat2<R> U;
            std::memcpy(U.u, m + o.mat_offset, sizeof(U));
            HIP_TRY(launch_2q<R>(psi, st->n, o.q0, o.q1, U, st->stream));
        }
        st->counters[0]++;
    }
    return 0;
}

// Enqueue the passes of `plan` whose descriptor image lives at `dev`.
template <typename R>
int enqueue_passes(qsim_state *st, const Plan &plan, const DescImage &img, char *dev, void *lam,
                   double *probe_out, const std::vector<hipFunction_t> *jit = nullptr)
{
    const PassDev *dp = reinterpret_cast<const PassDev *>(dev + img.off_pass);
    const StepDev *ds = reinterpret_cast<const StepDev *>(dev + img.off_step);
    const int32_t *di = reinterpret_cast<const int32_t *>(dev + img.off_ids);
    const cx<R> *dm = reinterpret_cast<const cx<R> *>(dev + img.off_mats);
    for (size_t p = 0; p < plan.passes.size(); ++p) {
        const PassDev &hp = plan.passes[p];
        const int nblocks = 1 << (st->n - hp.K);
        ProfScope ps(st, 0);
        const bool remap = (p == 0 && !plan.in_map.empty()) || (p + 1 == plan.passes.size() && !plan.out_map.empty());
        if (remap) {
            // folded layout change: read psi in the old layout (first pass)
            // and/or write the output order (last pass) into tmp; the other
            // passes work in place
            if (!jit || p >= jit->size() || !(*jit)[p]) return fail("internal: folded layout change without its kernel");
            void *in = st->psi, *out = st->tmp;
            const void *mats = dm;
            void *args[] = {&in, &out, &mats};
            HIP_TRY(hipModuleLaunchKernel((*jit)[p], (unsigned)nblocks, 1, 1, (unsigned)jit_block_threads(hp.K, 1),
                                          1, 1, 0, st->stream, args, nullptr));
            std::swap(st->psi, st->tmp);
            st->counters[5]++;
            st->counters[6] += (p == 0 && !plan.in_map.empty()) + (p + 1 == plan.passes.size() && !plan.out_map.empty());
        } else if (jit && p < jit->size() && (*jit)[p]) {
            void *psi = st->psi;
            void *lamp = lam;
            double *part = st->d_probe;
            const void *mats = dm;
            void *args1[] = {&psi, &mats};
            void *args2[] = {&psi, &lamp, &part, &mats};
            const int nv = lam ? 2 : 1;
            HIP_TRY(hipModuleLaunchKernel((*jit)[p], (unsigned)nblocks, 1, 1, (unsigned)jit_block_threads(hp.K, nv),
                                          1, 1, 0, st->stream, lam ? args2 : args1, nullptr));
            st->counters[5]++;
        } else {
            // 32x32 dense steps and free diagonals exist only in specialised kernels
            for (int si = hp.step_off; si < hp.step_off + hp.nsteps; ++si) {
                const FreeDiag *fd = nullptr;
                if (plan.steps[si].bit4 >= 0 || plan.step_fdiag(si, &fd) > 0)
                    return fail("internal: step form without its specialised kernel");
            }
            HIP_TRY(launch_tile_pass<R>(static_cast<cx<R> *>(st->psi), static_cast<cx<R> *>(lam), st->n, hp, dp + p,
                                        ds, dm, st->d_probe, plan.generic, st->stream));
        }
        st->counters[0]++;
        st->counters[3]++;
        if (lam && hp.nprobe > 0) {
            HIP_TRY(launch_probe_reduce(st->d_probe, nblocks, hp.nprobe, di + hp.probe_off, probe_out, st->stream));
            st->counters[0]++;
        }
    }
    return 0;
}

int plan_for(qsim_state *st, const qsim_op *ops, int nops, const void *mats, int max_probes, Plan &plan,
             const int32_t *in_map = nullptr, const int32_t *out_map = nullptr)
{
    std::vector<PlanOp> pops = to_plan_ops(ops, nops);
    std::string err;
    // two-vector (probe) passes hold two chunks in LDS: at most 2^12 complex64
    // amplitudes each
    const int K = max_probes > 1 ? std::min(st->K, st->dtype == QSIM_C64 ? 12 : 11) : st->K;
    // kernels compiled ahead of time take the generic slot forms
    if (!plan_passes(st->n, pops.data(), nops, mats, (int)amp_bytes(st->dtype), K, st->cmin, max_probes, plan,
                     err, max_probes == 1 ? st->fuse : 0, in_map, out_map, !st->jit, st->fuse_bar, st->fuse_bar5))
        return fail(err);
    return 0;
}

// Specialised kernels for the plan's passes (empty when switched off).
int jit_for(qsim_state *st, const Plan &plan, bool dbl, int nv, std::vector<hipFunction_t> &fns)
{
    fns.clear();
    if (!st->jit) return 0;
    if (st->arch.empty()) {
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, st->device));
        st->arch = prop.gcnArchName;
    }
    std::string err;
    if (!jit_functions(plan, dbl, nv, st->device, st->arch, fns, err))
        return fail("run-time kernel compile: " + err);
    return 0;
}

// A stored order that differs from the logical one is kept across op lists
// when their chunk passes run specialised kernels (a layout change then folds
// into a first pass; complex64 folds read amplitude pairs, so they need at
// least one row bit); otherwise it is applied first.
bool can_fold(const qsim_state *st)
{
    return st->perm_pending && st->jit && (st->dtype == QSIM_C128 || st->cmin >= 1);
}

// Stored bits below low_region() span one 2 MiB window of amplitudes (2^18
// complex64, 2^17 complex128): chunk passes over them gather and scatter
// within few address-translation pages, while chunks over higher bits touch
// up to one page per row.
int low_region(const qsim_state *st) { return st->dtype == QSIM_C64 ? 18 : 17; }

// Stored position tgt[q] for each logical bit q while running `ops` on a
// stored order that differs from the logical one.  The op qubits should sit
// below r = max(#op qubits, low_region()).  Each op qubit stored at or above r
// takes the slot of a non-op qubit below r (the row bits 0..cmin-1 first, so
// every pass's reserved low bits hold op qubits); that qubit moves to the
// vacated position, except that one leaving a row bit takes the lowest
// position at or above r (displacing a non-op qubit there to a vacated
// position): the first chunk pass then scatters to near rows only, while
// qubits outside its chunk move with the chunk base at no cost.  Every other
// qubit stays.
void choose_layout_impl(int n, int cmin, int region, const int32_t *at, const qsim_op *ops, int nops,
                        int32_t *tgt)
{
    // at[p]: logical bit held by stored bit p
    for (int p = 0; p < n; ++p) tgt[at[p]] = p;
    uint64_t T = 0;
    for (int i = 0; i < nops; ++i) {
        T |= 1ull << ops[i].q0;
        if (ops[i].kind == QSIM_OP_2Q) T |= 1ull << ops[i].q1;
        if (ops[i].kind == QSIM_OP_3Q || ops[i].kind == QSIM_OP_4Q) {
            int q[4];
            const int k = kq_qubits(ops[i], q);
            for (int j = 1; j < k; ++j) T |= 1ull << q[j];
        }
    }
    const int r = std::min(n, std::max(__builtin_popcountll(T), region));
    std::vector<int> slots, in;  // ascending stored positions
    for (int p = 0; p < r; ++p)
        if (!((T >> at[p]) & 1)) slots.push_back(p);
    for (int p = r; p < n; ++p)
        if ((T >> at[p]) & 1) in.push_back(p);
    // in.size() <= slots.size(): the op qubits number at most r
    bool vacated[64] = {}, taken[64] = {};
    for (int p : in) vacated[p] = true;
    std::vector<int> movers;  // qubits to place on the vacated positions left
    int scan = r;
    for (size_t i = 0; i < in.size(); ++i) {
        const int slot = slots[i], leaving = at[slot];
        tgt[at[in[i]]] = slot;
        if (slot < cmin) {
            while (taken[scan]) ++scan;  // stays below n: at most in.size() positions are taken
            taken[scan] = true;
            tgt[leaving] = scan;
            if (!vacated[scan]) movers.push_back(at[scan]);
        } else {
            movers.push_back(leaving);
        }
    }
    size_t k = 0;
    for (int p : in)
        if (!taken[p]) tgt[movers[k++]] = p;
}

void choose_layout(const qsim_state *st, const qsim_op *ops, int nops, int32_t *tgt)
{
    choose_layout_impl(st->n, st->cmin, low_region(st), st->perm_map, ops, nops, tgt);
}

// Stored order after an op list: `after` (perm_map form) and whether it
// differs from the logical order.
struct LayoutAfter {
    bool pending = false;
    bool post_done = false;  // a layout change after the list was written by its last pass
    int32_t map[64];
};

// out_map of a layout change `post` (pmap form: position p moves to
// post[p]) applied after an op list that leaves stored bit p holding logical
// bit at[p] (identity when null); false when it moves nothing.
bool post_out_map(int n, const int32_t *at, const int32_t *post, int32_t *out)
{
    bool move = false;
    for (int p = 0; p < n; ++p) {
        out[p] = post[at ? at[p] : p];
        move |= out[p] != p;
    }
    return move;
}

// With `post` (a layout change applied after the list, pmap form), the last
// pass writes the state in the order that change leaves (when it has room
// for the output's row bits and specialised kernels apply): after.post_done
// then says so, and after describes the stored order after both.
template <typename R>
int plan_and_compile(qsim_state *st, const qsim_op *ops, int nops, const void *mats, Plan &plan,
                     std::vector<hipFunction_t> &fns, LayoutAfter &after, const int32_t *post = nullptr)
{
    const bool keep = can_fold(st);
    if (st->perm_pending && !keep && flush_perm(st)) return -1;
    after.pending = false;
    after.post_done = false;
    int32_t omap[64];
    if (!keep) {
        const bool fold_out = post && st->jit && post_out_map(st->n, nullptr, post, omap);
        if (plan_for(st, ops, nops, mats, 1, plan, nullptr, fold_out ? omap : nullptr)) return -1;
        // a pending |0...0> is generated by a specialised first pass; forward
        // lists carry no probes, so with specialised kernels on, every pass has
        // one.  A graph captured from this plan is keyed on the pending |0...0>,
        // so it must never rely on an eager write made outside the capture.
        plan.zero_input = st->zero_pending && st->jit;
        if (!plan.out_map.empty() && !jit_mapped_last(plan, sizeof(R) == 8, 1)) plan.out_map.clear();
        if (!plan.out_map.empty()) {
            after.post_done = true;  // the state leaves in the logical order after `post`
            if (ensure_tmp(st)) return -1;
        }
        if (jit_for(st, plan, sizeof(R) == 8, 1, fns)) return -1;
        if (plan.zero_input && (fns.empty() || !fns[0]))
            return fail("internal: generated first pass without its kernel");
        if (st->zero_pending && !plan.zero_input && materialize_zero(st)) return -1;
        return 0;
    }
    const int n = st->n;
    int32_t tgt[64], in_map[64];
    choose_layout(st, ops, nops, tgt);
    std::vector<qsim_op> mapped(ops, ops + nops);
    for (qsim_op &o : mapped) {
        if (o.kind == QSIM_OP_3Q || o.kind == QSIM_OP_4Q) {
            int q[4];
            const int k = kq_qubits(o, q);
            for (int j = 0; j < k; ++j) q[j] = tgt[q[j]];
            o.q0 = q[0];
            o.q1 = kq_pack(q, k);
            continue;
        }
        o.q0 = tgt[o.q0];
        if (o.kind == QSIM_OP_2Q) o.q1 = tgt[o.q1];
    }
    bool move = false;
    for (int p = 0; p < n; ++p) {
        in_map[p] = tgt[st->perm_map[p]];
        move |= in_map[p] != p;
        after.map[tgt[p]] = p;
    }
    for (int p = 0; p < n; ++p) after.pending |= after.map[p] != p;
    const bool fold_out = post && post_out_map(n, after.map, post, omap);
    if (plan_for(st, mapped.data(), nops, mats, 1, plan, move ? in_map : nullptr, fold_out ? omap : nullptr))
        return -1;
    if (!plan.out_map.empty() && !jit_mapped_last(plan, sizeof(R) == 8, 1)) plan.out_map.clear();
    if (!plan.out_map.empty()) {
        after.post_done = true;  // the state leaves in the logical order after `post`
        after.pending = false;
        for (int p = 0; p < n; ++p) after.map[p] = p;
    }
    if (jit_for(st, plan, sizeof(R) == 8, 1, fns)) return -1;
    if (move) {
        if (!jit_mapped_first(plan, sizeof(R) == 8, 1) || fns.empty() || !fns[0])
            return fail("internal: folded layout change without its kernel");
    }
    if (move || !plan.out_map.empty())
        if (ensure_tmp(st)) return -1;
    return 0;
}

void set_layout(qsim_state *st, const LayoutAfter &after)
{
    st->perm_pending = after.pending;
    std::memcpy(st->perm_map, after.map, sizeof(int32_t) * st->n);
}

// Records a layout change (pmap form: position p moves to map[p]) in the
// stored bit order, as qsim_permute does: no data moves.
void compose_layout(qsim_state *st, const int32_t *map)
{
    if (st->zero_pending) return;  // a bit permutation leaves |0...0> as it is
    int32_t comp[64];
    for (int p = 0; p < st->n; ++p) comp[p] = st->perm_pending ? map[st->perm_map[p]] : map[p];
    bool ident = true;
    for (int p = 0; p < st->n; ++p) ident &= comp[p] == p;
    st->perm_pending = !ident;
    std::memcpy(st->perm_map, comp, sizeof(int32_t) * st->n);
}

// `post` (or null): a layout change applied after the list (pmap form),
// written by the last pass when it can be, else recorded as by qsim_permute.
template <typename R>
int run_chunked(qsim_state *st, const qsim_op *ops, int nops, const void *mats, uint64_t mat_count,
                const int32_t *post = nullptr)
{
    (void)mat_count;
    Plan plan;
    std::vector<hipFunction_t> fns;
    LayoutAfter after;
    if (plan_and_compile<R>(st, ops, nops, mats, plan, fns, after, post)) return -1;
    DescImage img;
    img.build<R>(plan);
    void *dev = nullptr;
    if (st->up.upload(img.bytes.data(), img.bytes.size(), st->stream, &dev)) return -1;
    if (enqueue_passes<R>(st, plan, img, static_cast<char *>(dev), nullptr, nullptr, &fns)) return -1;
    set_layout(st, after);
    st->zero_pending = false;
    if (post && !after.post_done) compose_layout(st, post);
    return 0;
}

template <typename R>
int run_graphed(qsim_state *st, const qsim_op *ops, int nops, const void *mats, uint64_t mat_count,
                const int32_t *post = nullptr)
{
    // a layout change that cannot fold is applied first, so the key sees
    // the buffer the passes will use; so is a pending |0...0> that no
    // specialised first pass will generate (the graph must not need it)
    if (st->perm_pending && !can_fold(st) && flush_perm(st)) return -1;
    if (st->zero_pending && !st->jit && materialize_zero(st)) return -1;
    // key: buffer address, geometry, ops, matrices
    std::vector<char> key;
    auto put = [&key](const void *p, size_t b) {
        const char *c = static_cast<const char *>(p);
        key.insert(key.end(), c, c + b);
    };
    put(&st->psi, sizeof(void *));
    put(&st->K, sizeof(int));
    put(&st->cmin, sizeof(int));
    put(&st->jit, sizeof(bool));
    put(&st->fuse, sizeof(int));
    put(&st->fuse_bar, sizeof(double));
    put(&st->fuse_bar5, sizeof(double));
    const bool keep = can_fold(st);
    put(&keep, sizeof(bool));
    if (keep) put(st->perm_map, sizeof(int32_t) * st->n);
    put(&st->zero_pending, sizeof(bool));
    put(ops, sizeof(qsim_op) * nops);
    put(mats, sizeof(cx<R>) * mat_count);
    const bool has_post = post != nullptr;
    put(&has_post, sizeof(bool));
    if (post) put(post, sizeof(int32_t) * st->n);
    for (auto it = st->graphs.begin(); it != st->graphs.end(); ++it) {
        if (it->key == key) {
            st->graphs.splice(st->graphs.begin(), st->graphs, it);
            HIP_TRY(hipGraphLaunch(st->graphs.front().exec, st->stream));
            const GraphEntry &g = st->graphs.front();
            if (g.swaps & 1) std::swap(st->psi, st->tmp);  // replayed folded layout changes
            st->counters[6] += g.folds;
            st->perm_pending = g.after_pending;
            std::memcpy(st->perm_map, g.after_map, sizeof(int32_t) * st->n);
            st->zero_pending = false;
            st->counters[1]++;
            if (post && !g.post_done) compose_layout(st, post);
            return 0;
        }
    }
    Plan plan;
    std::vector<hipFunction_t> fns;
    LayoutAfter after;
    if (plan_and_compile<R>(st, ops, nops, mats, plan, fns, after, post)) return -1;
    DescImage img;
    img.build<R>(plan);
    GraphEntry e;
    e.key.swap(key);
    e.folds = (int)!plan.in_map.empty() + (int)!plan.out_map.empty();
    e.swaps = (int)!plan.in_map.empty() + (int)!plan.out_map.empty() -
              (int)(plan.passes.size() == 1 && !plan.in_map.empty() && !plan.out_map.empty());
    e.post_done = after.post_done;
    e.after_pending = after.pending;
    std::memcpy(e.after_map, after.map, sizeof(int32_t) * st->n);
    HIP_TRY(hipMalloc(&e.desc, img.bytes.size()));
    HIP_TRY(hipMemcpyAsync(e.desc, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    // Capturing a folded first pass swaps psi and tmp (enqueue_passes), but no
    // pass runs until the launch below: every failure return before it puts
    // the buffers back, so the state keeps describing the data it holds.
    void *const psi0 = st->psi, *const tmp0 = st->tmp;
    auto undo = [&]() {
        st->psi = psi0;
        st->tmp = tmp0;
    };
    hipError_t eb = hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal);
    if (eb != hipSuccess) {
        hipFree(e.desc);
        return fail(std::string("HIP error: ") + hipGetErrorString(eb));
    }
    st->capturing = true;
    int rc = enqueue_passes<R>(st, plan, img, static_cast<char *>(e.desc), nullptr, nullptr, &fns);
    st->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t ec = hipStreamEndCapture(st->stream, &g);
    if (rc || ec != hipSuccess) {
        undo();
        if (g) hipGraphDestroy(g);
        hipFree(e.desc);
        if (rc) return rc;
        return fail(std::string("HIP error: ") + hipGetErrorString(ec));
    }
    e.graph = g;
    hipError_t ei = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        undo();
        hipGraphDestroy(g);
        hipFree(e.desc);
        return fail(std::string("HIP error: ") + hipGetErrorString(ei));
    }
    st->counters[2]++;
    st->graphs.push_front(std::move(e));
    hipError_t el = hipGraphLaunch(st->graphs.front().exec, st->stream);
    if (el != hipSuccess) {
        undo();
        GraphEntry &bad = st->graphs.front();
        hipGraphExecDestroy(bad.exec);
        hipGraphDestroy(bad.graph);
        hipFree(bad.desc);
        st->graphs.pop_front();
        return fail(std::string("HIP error: ") + hipGetErrorString(el));
    }
    st->counters[1]++;
    set_layout(st, after);
    st->zero_pending = false;
    if (post && !after.post_done) compose_layout(st, post);
    while (st->graphs.size() > kMaxGraphs) {
        HIP_TRY(hipStreamSynchronize(st->stream));
        GraphEntry &old = st->graphs.back();
        hipGraphExecDestroy(old.exec);
        hipGraphDestroy(old.graph);
        hipFree(old.desc);
        st->graphs.pop_back();
    }
    return 0;
}

// Chunk planning needs at least kG qubits; tiny states use the gate streams.
bool chunk_ok(const qsim_state *st) { return st->n >= 6; }

template <typename R>
int apply_ops_t(qsim_state *st, int mode, const qsim_op *ops, int nops, const void *mats, uint64_t mat_count,
                const int32_t *post = nullptr)
{
    if (nops > 0 && mode != QSIM_SEQUENTIAL && chunk_ok(st)) {
        if (mode == QSIM_COOPERATIVE) return run_chunked<R>(st, ops, nops, mats, mat_count, post);
        if (mode == QSIM_GRAPHED) return run_graphed<R>(st, ops, nops, mats, mat_count, post);
        return fail("apply_tile: unknown tile mode " + std::to_string(mode));
    }
    if (nops > 0) {
        if (mode != QSIM_SEQUENTIAL && mode != QSIM_COOPERATIVE && mode != QSIM_GRAPHED)
            return fail("apply_tile: unknown tile mode " + std::to_string(mode));
        if (flush_perm(st)) return -1;
        if (run_sequential<R>(st, ops, nops, mats)) return -1;
    }
    if (post) compose_layout(st, post);
    return 0;
}

int build_perm_args(int n, const int32_t *map, int smax, PermArgs &P)
{
    int inv[64];
    for (int p = 0; p < n; ++p) inv[map[p]] = p;
    int a = std::min(n, smax);
    uint64_t S = 0;
    for (; a >= 1; --a) {
        S = (a >= 64 ? ~0ull : ((1ull << a) - 1));
        for (int k = 0; k < a; ++k) S |= 1ull << inv[k];
        if (__builtin_popcountll(S) <= smax) break;
    }
    std::memset(&P, 0, sizeof(P));
    P.n = n;
    P.a = P.b = a;
    P.s = __builtin_popcountll(S);
    int sin[64], ns = 0, comp = 0;
    for (int p = 0; p < n; ++p) {
        if (S >> p & 1) sin[ns++] = p;
        else {
            P.comp_in[comp] = p;
            P.comp_out[comp] = map[p];
            ++comp;
        }
    }
    P.ncomp = comp;
    int sout[64];
    for (int k = 0; k < ns; ++k) sout[k] = map[sin[k]];
    std::sort(sout, sout + ns);
    for (int k = 0; k < ns - P.a; ++k) P.sin_hi[k] = sin[P.a + k];
    for (int k = 0; k < ns - P.b; ++k) P.sout_hi[k] = sout[P.b + k];
    for (int k = 0; k < ns; ++k) {
        const int in_pos = inv[sout[k]];
        int r = 0;
        while (sin[r] != in_pos) ++r;
        P.w2r[k] = r;
    }
    return 0;
}

int permute_map_impl(int n, const int32_t *lo, const int32_t *ln, int32_t *map)
{
    std::vector<int> new_pos(n, -1);
    for (int p = 0; p < n; ++p) {
        const int gq = ln[p];
        if (gq < 0 || gq >= n) return fail("layout_new contains invalid qubit id");
        if (new_pos[gq] != -1) return fail("layout_new is not a permutation");
        new_pos[gq] = p;
    }
    for (int p = 0; p < n; ++p) {
        const int gq = lo[p];
        if (gq < 0 || gq >= n) return fail("layout_old contains invalid qubit id");
        if (new_pos[gq] < 0) return fail("layout_old contains qubit not in layout_new");
        map[p] = new_pos[gq];
    }
    return 0;
}

template <typename RA, typename RB = RA>
int dot_t(qsim_state *st, const void *a, const void *b, double *re, double *im)
{
    if (ensure_small(st, 2 * (size_t)max_dot_blocks() + 8)) return -1;
    int nb = 0;
    HIP_TRY((launch_dot<RA, RB>(static_cast<const cx<RA> *>(a), static_cast<const cx<RB> *>(b), st->dim,
                                st->d_small, &nb, st->stream)));
    double *out = st->d_small + 2 * (size_t)max_dot_blocks();
    HIP_TRY(launch_sum_partials(st->d_small, nb, 2, out, st->stream));
    double h[2];
    HIP_TRY(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    st->counters[0] += 2;
    *re = h[0];
    if (im) *im = h[1];
    return 0;
}

// Pauli chunk passes.  lam == nullptr: expectation only.
template <typename R>
int pauli_t(qsim_state *st, const qsim_pauli *terms, int nterms, void *lam, double *re, double *im)
{
    std::vector<uint64_t> xm(nterms);
    for (int t = 0; t < nterms; ++t) xm[t] = terms[t].xmask;
    std::vector<PauliPass> passes;
    std::string err;
    if (!plan_pauli(st->n, xm.data(), nterms, st->K, st->cmin, passes, err)) return fail(err);
    // descriptor image: per pass one PassDev followed by its terms
    std::vector<char> img;
    std::vector<size_t> pass_off, term_off;
    for (auto &pp : passes) {
        pass_off.push_back(img.size());
        img.resize(img.size() + ((sizeof(PassDev) + 255) & ~size_t(255)));
        std::memcpy(&img[pass_off.back()], &pp.pd, sizeof(PassDev));
        term_off.push_back(img.size());
        const size_t tb = pp.terms.size() * sizeof(PauliDev);
        img.resize(img.size() + ((tb + 255) & ~size_t(255)) + 256);
        for (size_t k = 0; k < pp.terms.size(); ++k) {
            const qsim_pauli &T = terms[pp.terms[k]];
            PauliDev d{};
            d.xmask = T.xmask;
            d.zmask = T.zmask;
            d.xloc = pp.xloc[k];
            d.cre = T.coef_re;
            d.cim = T.coef_im;
            std::memcpy(&img[term_off.back() + k * sizeof(PauliDev)], &d, sizeof(d));
        }
    }
    void *dev = nullptr;
    if (!img.empty() && st->up.upload(img.data(), img.size(), st->stream, &dev)) return -1;
    size_t maxblocks = 1;
    for (auto &pp : passes) maxblocks = std::max(maxblocks, size_t(1) << (st->n - pp.pd.K));
    if (ensure_small(st, 2 * maxblocks + 2 * passes.size() + 8)) return -1;
    double *sums = st->d_small + 2 * maxblocks;
    for (size_t p = 0; p < passes.size(); ++p) {
        const PauliPass &pp = passes[p];
        int mode = PAULI_EXPECT;
        if (lam) mode |= (p == 0 ? PAULI_WRITE : PAULI_ACC);
        const int nblocks = 1 << (st->n - pp.pd.K);
        char *base = static_cast<char *>(dev);
        ProfScope ps(st, 3);
        HIP_TRY(launch_pauli_chunk<R>(static_cast<const cx<R> *>(st->psi), static_cast<cx<R> *>(lam), st->n, pp.pd,
                                      reinterpret_cast<const PassDev *>(base + pass_off[p]),
                                      reinterpret_cast<const PauliDev *>(base + term_off[p]), (int)pp.terms.size(),
                                      mode, st->d_small, st->stream));
        HIP_TRY(launch_sum_partials(st->d_small, nblocks, 2, sums + 2 * p, st->stream));
        st->counters[0] += 2;
        st->counters[3]++;
    }
    std::vector<double> h(2 * passes.size() + 2, 0.0);
    if (!passes.empty())
        HIP_TRY(hipMemcpyAsync(h.data(), sums, 2 * passes.size() * sizeof(double), hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    double r = 0, i = 0;
    for (size_t p = 0; p < passes.size(); ++p) {
        r += h[2 * p];
        i += h[2 * p + 1];
    }
    if (re) *re = r;
    if (im) *im = i;
    return 0;
}

template <typename R>
int adjoint_t(qsim_state *st, const qsim_pauli *terms, int nterms, const qsim_op *ops, int nops, const void *mats,
              uint64_t mat_count, int nprobes, double *probes_out, double *expect_out)
{
    if (ensure_tmp(st)) return -1;
    double er = 0, ei = 0;
    if (pauli_t<R>(st, terms, nterms, st->tmp, &er, &ei)) return -1;
    if (expect_out) *expect_out = er;
    (void)mat_count;
    Plan plan;
    if (plan_for(st, ops, nops, mats, kMaxProbes, plan)) return -1;
    size_t need = 1;
    for (auto &p : plan.passes) need = std::max(need, (size_t(1) << (st->n - p.K)) * (size_t)std::max(1, p.nprobe));
    if (ensure_probe(st, need)) return -1;
    double *d_out = nullptr;
    HIP_TRY(hipMalloc(reinterpret_cast<void **>(&d_out), sizeof(double) * std::max(1, nprobes)));
    std::unique_ptr<void, void (*)(void *)> guard(d_out, [](void *p) { hipFree(p); });
    HIP_TRY(hipMemsetAsync(d_out, 0, sizeof(double) * std::max(1, nprobes), st->stream));
    std::vector<hipFunction_t> fns;
    if (jit_for(st, plan, sizeof(R) == 8, 2, fns)) return -1;
    DescImage img;
    img.build<R>(plan);
    void *dev = nullptr;
    if (st->up.upload(img.bytes.data(), img.bytes.size(), st->stream, &dev)) return -1;
    if (enqueue_passes<R>(st, plan, img, static_cast<char *>(dev), st->tmp, d_out, &fns)) return -1;
    if (nprobes > 0)
        HIP_TRY(hipMemcpyAsync(probes_out, d_out, sizeof(double) * nprobes, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return 0;
}

}  // namespace

// ============================================================================
// exported C ABI
// ============================================================================
namespace {
size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

// Sampling (qsim_sample).  Device memory: the block sums, the shot groups,
// the residuals and the indices; allocated per call when `scratch` is null,
// else carved from [scratch, scratch + scratch_bytes), which is then the only
// device memory the call touches besides the state (refused beforehand when
// too small; nothing in it needs to start at zero).
int qsim::sample_impl(qsim_state *st, uint64_t shots, uint64_t seed, uint64_t *indices_out, void *scratch,
                      size_t scratch_bytes)
{
    if (!st) return fail("null state");
    if (shots == 0) return 0;
    if (!indices_out) return fail("sample: null output");
    if (shots > (1ull << 31)) return fail("sample: at most 2^31 shots per call");
    const uint64_t nb = (st->dim + (1ull << kSampleB) - 1) >> kSampleB;
    const size_t blk_bytes = align256(nb * sizeof(double));
    const size_t rbytes = shots * sizeof(double);
    if (scratch) {
        const size_t need = blk_bytes + align256(std::min<uint64_t>(shots, nb) * sizeof(SampleGroup)) +
                            align256(rbytes) + shots * sizeof(uint64_t);
        if (need > scratch_bytes)
            return fail("sample_noalloc: " + std::to_string(shots) + " shots of a " + std::to_string(st->n) +
                        "-qubit state need " + std::to_string(need) + " bytes of scratch; the library holds " +
                        std::to_string(scratch_bytes));
        // a pending layout change is applied through the second buffer
        if (st->perm_pending && !st->tmp)
            return fail("sample_noalloc: the state holds a pending layout change and has no second buffer to "
                        "apply it with (call qsim_sample once, or read the state, first)");
    }
    if (flush_perm(st)) return -1;
    double *d_blk = nullptr;
    std::unique_ptr<void, void (*)(void *)> g1(nullptr, [](void *p) { hipFree(p); });
    if (scratch) {
        d_blk = static_cast<double *>(scratch);
    } else {
        HIP_TRY(hipMalloc(reinterpret_cast<void **>(&d_blk), nb * sizeof(double)));
        g1.reset(d_blk);
    }
    if (st->dtype == QSIM_C64)
        HIP_TRY(launch_block_prob<float>(static_cast<cx<float> *>(st->psi), st->dim, d_blk, st->stream));
    else
        HIP_TRY(launch_block_prob<double>(static_cast<cx<double> *>(st->psi), st->dim, d_blk, st->stream));
    std::vector<double> cdf(nb);
    HIP_TRY(hipMemcpyAsync(cdf.data(), d_blk, nb * sizeof(double), hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    for (uint64_t b = 1; b < nb; ++b) cdf[b] += cdf[b - 1];
    const double total = cdf[nb - 1];
    if (!(total > 0.0)) return fail("sample: the state has zero norm");
    // shot k: u = U(k) * total with U from splitmix64 of (seed, k), 53 bits
    std::vector<std::pair<uint64_t, double>> draws(shots);  // (block, residual)
    for (uint64_t k = 0; k < shots; ++k) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * (k + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = (double)(z >> 11) * 0x1p-53 * total;
        uint64_t b = (uint64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        // u rounded up to the total: the last block with nonzero probability
        if (b >= nb) b = (uint64_t)(std::lower_bound(cdf.begin(), cdf.end(), total) - cdf.begin());
        draws[k] = {b, u - (b ? cdf[b - 1] : 0.0)};
    }
    std::sort(draws.begin(), draws.end());
    std::vector<SampleGroup> groups;
    std::vector<double> resid(shots);
    for (uint64_t k = 0; k < shots; ++k) {
        resid[k] = draws[k].second;
        if (groups.empty() || groups.back().block != draws[k].first)
            groups.push_back(SampleGroup{draws[k].first, (uint32_t)k, 0});
        groups.back().count++;
    }
    const size_t gbytes = align256(groups.size() * sizeof(SampleGroup));
    char *base = nullptr;
    std::unique_ptr<void, void (*)(void *)> g2(nullptr, [](void *p) { hipFree(p); });
    if (scratch) {
        base = static_cast<char *>(scratch) + blk_bytes;
    } else {
        void *d_buf = nullptr;
        HIP_TRY(hipMalloc(&d_buf, gbytes + align256(rbytes) + shots * sizeof(uint64_t)));
        g2.reset(d_buf);
        base = static_cast<char *>(d_buf);
    }
    SampleGroup *d_groups = reinterpret_cast<SampleGroup *>(base);
    double *d_resid = reinterpret_cast<double *>(base + gbytes);
    uint64_t *d_idx = reinterpret_cast<uint64_t *>(base + gbytes + align256(rbytes));
    HIP_TRY(hipMemcpyAsync(d_groups, groups.data(), groups.size() * sizeof(SampleGroup), hipMemcpyHostToDevice,
                           st->stream));
    HIP_TRY(hipMemcpyAsync(d_resid, resid.data(), rbytes, hipMemcpyHostToDevice, st->stream));
    if (st->dtype == QSIM_C64)
        HIP_TRY(launch_sample_blocks<float>(static_cast<cx<float> *>(st->psi), st->dim, d_groups, (int)groups.size(),
                                            d_resid, d_idx, st->stream));
    else
        HIP_TRY(launch_sample_blocks<double>(static_cast<cx<double> *>(st->psi), st->dim, d_groups,
                                             (int)groups.size(), d_resid, d_idx, st->stream));
    HIP_TRY(hipMemcpyAsync(indices_out, d_idx, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    st->counters[0] += 2;
    return 0;
}

int qsim::state_device(const qsim_state *st) { return st ? st->device : -1; }

extern "C" {

const char *qsim_last_error(void) { return g_err.c_str(); }

int qsim_abi_version(void) { return 1; }

int qsim_create(int num_qubits, int dtype, qsim_state **out)
{
    if (!out) return fail("qsim_create: null output");
    *out = nullptr;
    if (num_qubits < 1 || num_qubits > 40) return fail("num_qubits must be in [1, 40]");
    if (dtype != QSIM_C64 && dtype != QSIM_C128) return fail("unknown dtype");
    std::unique_ptr<qsim_state> st(new qsim_state());
    st->n = num_qubits;
    st->dtype = dtype;
    st->dim = 1ull << num_qubits;
    // chunk bits: 2^13 complex64 or 2^12 complex128 amplitudes (64 KB of LDS)
    st->K = dtype == QSIM_C64 ? 13 : 11;
    st->cmin = dtype == QSIM_C64 ? 4 : 3;
    if (const char *e = std::getenv("QSIM_JIT")) st->jit = std::atoi(e) != 0;
    HIP_TRY(hipGetDevice(&st->device));
    static bool prepared = false;
    if (!prepared) {
        HIP_TRY(prepare_kernels());
        prepared = true;
    }
    HIP_TRY(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
    hipError_t e = hipMalloc(&st->psi, st->dim * amp_bytes(dtype));
    if (e != hipSuccess) {
        hipStreamDestroy(st->stream);
        return fail(std::string("HIP error: ") + hipGetErrorString(e) + " (out of memory allocating statevector)");
    }
    if (qsim_reset(st.get())) {
        hipFree(st->psi);
        hipStreamDestroy(st->stream);
        return -1;
    }
    *out = st.release();
    return 0;
}

int qsim_destroy(qsim_state *st)
{
    if (!st) return 0;
    hipStreamSynchronize(st->stream);
    drop_graphs(st);
    st->up.release();
    if (st->psi) hipFree(st->psi);
    if (st->tmp) hipFree(st->tmp);
    if (st->d_small) hipFree(st->d_small);
    if (st->d_probe) hipFree(st->d_probe);
    if (st->d_gatek) hipFree(st->d_gatek);
    for (auto &r : st->recs) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    for (auto e : st->pool) hipEventDestroy(e);
    if (st->stream) hipStreamDestroy(st->stream);
    delete st;
    return 0;
}

int qsim_num_qubits(const qsim_state *st) { return st ? st->n : -1; }
uint64_t qsim_dim(const qsim_state *st) { return st ? st->dim : 0; }
int qsim_dtype(const qsim_state *st) { return st ? st->dtype : -1; }

int qsim_device_ptr(qsim_state *st, void **out)
{
    if (!st || !out) return fail("null argument");
    if (flush_perm(st)) return -1;
    *out = st->psi;
    return 0;
}

int qsim_reset(qsim_state *st)
{
    if (!st) return fail("null state");
    st->perm_pending = false;  // the state is overwritten
    st->zero_pending = true;   // written by the next first pass, or by the next reader
    return 0;
}

int qsim_from_host(qsim_state *st, const void *src, uint64_t count)
{
    if (!st) return fail("null state");
    if (count != st->dim) return fail("Input array must have size 2^num_qubits");
    st->perm_pending = false;  // the state is overwritten
    st->zero_pending = false;
    HIP_TRY(hipMemcpyAsync(st->psi, src, count * amp_bytes(st->dtype), hipMemcpyHostToDevice, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return 0;
}

int qsim_to_host(qsim_state *st, void *dst, uint64_t count)
{
    if (!st) return fail("null state");
    if (count != st->dim) return fail("Output array must have size 2^num_qubits");
    if (flush_perm(st)) return -1;
    HIP_TRY(hipMemcpyAsync(dst, st->psi, count * amp_bytes(st->dtype), hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return 0;
}

int qsim_get_amplitude(qsim_state *st, uint64_t index, double *re, double *im)
{
    if (!st || !re || !im) return fail("null argument");
    if (index >= st->dim) return fail("amplitude index out of range");
    if (flush_perm(st)) return -1;
    if (st->dtype == QSIM_C64) {
        float h[2];
        HIP_TRY(hipMemcpyAsync(h, static_cast<char *>(st->psi) + 8 * index, 8, hipMemcpyDeviceToHost, st->stream));
        HIP_TRY(hipStreamSynchronize(st->stream));
        *re = h[0];
        *im = h[1];
    } else {
        double h[2];
        HIP_TRY(hipMemcpyAsync(h, static_cast<char *>(st->psi) + 16 * index, 16, hipMemcpyDeviceToHost, st->stream));
        HIP_TRY(hipStreamSynchronize(st->stream));
        *re = h[0];
        *im = h[1];
    }
    return 0;
}

int qsim_synchronize(qsim_state *st)
{
    if (!st) return fail("null state");
    if (flush_perm(st)) return -1;  // all requested work, a pending layout change included
    HIP_TRY(hipStreamSynchronize(st->stream));
    return 0;
}

int qsim_permute_map(int num_qubits, const int32_t *layout_old, const int32_t *layout_new, int32_t *map_out)
{
    if (num_qubits < 1 || num_qubits > 63 || !layout_old || !layout_new || !map_out)
        return fail("layout_old/layout_new must have length num_qubits");
    return permute_map_impl(num_qubits, layout_old, layout_new, map_out);
}

int qsim_permute(qsim_state *st, const int32_t *layout_old, const int32_t *layout_new, int num_qubits)
{
    if (!st) return fail("null state");
    if (num_qubits != st->n) return fail("layout_old/layout_new must have length num_qubits");
    int32_t map[64];
    if (permute_map_impl(st->n, layout_old, layout_new, map)) return -1;
    if (st->zero_pending) return 0;  // a bit permutation leaves |0...0> as it is
    // Recorded, not run: the next op list folds it into its first chunk pass
    // (or it runs as one permutation pass when anything else reads the state).
    if (st->perm_pending) {
        int32_t comp[64];
        for (int p = 0; p < st->n; ++p) comp[p] = map[st->perm_map[p]];
        std::memcpy(map, comp, sizeof(int32_t) * st->n);
    }
    bool ident = true;
    for (int p = 0; p < st->n; ++p) ident &= map[p] == p;
    st->perm_pending = !ident;
    std::memcpy(st->perm_map, map, sizeof(int32_t) * st->n);
    return 0;
}

int qsim_validate_ops(int num_qubits, const qsim_op *ops, int nops, uint64_t mat_count)
{
    return check_ops(num_qubits, ops, nops, mat_count, false);
}

int qsim_apply_ops(qsim_state *st, int mode, const qsim_op *ops, int nops, const void *mats, uint64_t mat_count)
{
    if (!st) return fail("null state");
    if (check_ops(st->n, ops, nops, mat_count, false)) return -1;
    if (st->dtype == QSIM_C64) return apply_ops_t<float>(st, mode, ops, nops, mats, mat_count);
    return apply_ops_t<double>(st, mode, ops, nops, mats, mat_count);
}

int qsim_apply_ops_permute(qsim_state *st, int mode, const qsim_op *ops, int nops, const void *mats,
                           uint64_t mat_count, const int32_t *layout_old, const int32_t *layout_new, int num_qubits)
{
    if (!st) return fail("null state");
    if (check_ops(st->n, ops, nops, mat_count, false)) return -1;
    int32_t map[64];
    if (num_qubits != st->n || !layout_old || !layout_new) {
        // as apply then permute: the op list runs, the layout change is refused
        if (qsim_apply_ops(st, mode, ops, nops, mats, mat_count)) return -1;
        return fail("layout_old/layout_new must have length num_qubits");
    }
    if (permute_map_impl(st->n, layout_old, layout_new, map)) {
        const std::string msg = g_err;
        if (qsim_apply_ops(st, mode, ops, nops, mats, mat_count)) return -1;
        return fail(msg);
    }
    if (st->dtype == QSIM_C64) return apply_ops_t<float>(st, mode, ops, nops, mats, mat_count, map);
    return apply_ops_t<double>(st, mode, ops, nops, mats, mat_count, map);
}

int qsim_apply_1q(qsim_state *st, int target_qubit, const void *U)
{
    if (!st) return fail("null state");
    if (target_qubit < 0 || target_qubit >= st->n) return fail("Invalid target qubit");
    if (flush_perm(st)) return -1;
    if (st->dtype == QSIM_C64) {
        Mat1<float> M;
        std::memcpy(M.u, U, sizeof(M));
        HIP_TRY(launch_1q<float>(static_cast<cx<float> *>(st->psi), st->n, target_qubit, M, st->stream));
    } else {
        Mat1<double> M;
        std::memcpy(M.u, U, sizeof(M));
        HIP_TRY(launch_1q<double>(static_cast<cx<double> *>(st->psi), st->n, target_qubit, M, st->stream));
    }
    st->counters[0]++;
    return 0;
}

int qsim_apply_2q(qsim_state *st, int q0, int q1, const void *U)
{
    if (!st) return fail("null state");
    if (q0 < 0 || q0 >= st->n || q1 < 0 || q1 >= st->n) return fail("Invalid qubit indices");
    if (q0 == q1) return fail("q0 and q1 must be different");
    if (flush_perm(st)) return -1;
    if (st->dtype == QSIM_C64) {
        Mat2<float> M;
        std::memcpy(M.u, U, sizeof(M));
        HIP_TRY(launch_2q<float>(static_cast<cx<float> *>(st->psi), st->n, q0, q1, M, st->stream));
    } else {
        Mat2<double> M;
        std::memcpy(M.u, U, sizeof(M));
        HIP_TRY(launch_2q<double>(static_cast<cx<double> *>(st->psi), st->n, q0, q1, M, st->stream));
    }
    st->counters[0]++;
    return 0;
}

int qsim_apply_diagonal_1q(qsim_state *st, int q, double p0r, double p0i, double p1r, double p1i)
{
    if (!st) return fail("null state");
    if (q < 0 || q >= st->n) return fail("Invalid target qubit");
    if (flush_perm(st)) return -1;
    if (st->dtype == QSIM_C64)
        HIP_TRY(launch_diag1q<float>(static_cast<cx<float> *>(st->psi), st->n, q, cx<float>{(float)p0r, (float)p0i},
                                     cx<float>{(float)p1r, (float)p1i}, st->stream));
    else
        HIP_TRY(launch_diag1q<double>(static_cast<cx<double> *>(st->psi), st->n, q, cx<double>{p0r, p0i},
                                      cx<double>{p1r, p1i}, st->stream));
    st->counters[0]++;
    return 0;
}

int qsim_apply_1q_raw(void *psi, int dtype, int n, int q, const void *U, void *stream)
{
    if (!psi) return fail("device array has null pointer");
    if (n < 1 || n > 40) return fail("num_qubits must be in [1, 40]");
    if (q < 0 || q >= n) return fail("Invalid target qubit");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == QSIM_C64) {
        Mat1<float> M;
        std::memcpy(M.u, U, sizeof(M));
        HIP_TRY(launch_1q<float>(static_cast<cx<float> *>(psi), n, q, M, s));
    } else if (dtype == QSIM_C128) {
        Mat1<double> M;
        std::memcpy(M.u, U, sizeof(M));
        HIP_TRY(launch_1q<double>(static_cast<cx<double> *>(psi), n, q, M, s));
    } else {
        return fail("unknown dtype");
    }
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int qsim_apply_2q_raw(void *psi, int dtype, int n, int q0, int q1, const void *U, void *stream)
{
    if (!psi) return fail("device array has null pointer");
    if (n < 2 || n > 40) return fail("num_qubits must be in [2, 40]");
    if (q0 < 0 || q0 >= n || q1 < 0 || q1 >= n) return fail("Invalid qubit indices");
    if (q0 == q1) return fail("q0 and q1 must be different");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == QSIM_C64) {
        Mat2<float> M;
        std::memcpy(M.u, U, sizeof(M));
        HIP_TRY(launch_2q<float>(static_cast<cx<float> *>(psi), n, q0, q1, M, s));
    } else if (dtype == QSIM_C128) {
        Mat2<double> M;
        std::memcpy(M.u, U, sizeof(M));
        HIP_TRY(launch_2q<double>(static_cast<cx<double> *>(psi), n, q0, q1, M, s));
    } else {
        return fail("unknown dtype");
    }
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int qsim_norm2(qsim_state *st, double *out)
{
    if (!st || !out) return fail("null argument");
    // the norm does not depend on the stored bit order: a pending layout
    // change stays pending (no permutation pass)
    if (materialize_zero(st)) return -1;
    if (st->dtype == QSIM_C64) return dot_t<float>(st, st->psi, st->psi, out, nullptr);
    return dot_t<double>(st, st->psi, st->psi, out, nullptr);
}

int qsim_overlap(qsim_state *a, qsim_state *b, double *re, double *im)
{
    if (!a || !b || !re || !im) return fail("null argument");
    if (a->n != b->n) return fail("overlap: states differ in size");
    if (a->device != b->device) return fail("overlap: states live on different devices");
    if (flush_perm(a) || flush_perm(b)) return -1;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (a->dtype == QSIM_C64)
        return b->dtype == QSIM_C64 ? dot_t<float>(a, a->psi, b->psi, re, im)
                                    : dot_t<float, double>(a, a->psi, b->psi, re, im);
    return b->dtype == QSIM_C64 ? dot_t<double, float>(a, a->psi, b->psi, re, im)
                                : dot_t<double>(a, a->psi, b->psi, re, im);
}

int qsim_scale(qsim_state *st, double factor)
{
    if (!st) return fail("null state");
    if (flush_perm(st)) return -1;
    if (st->dtype == QSIM_C64)
        HIP_TRY(launch_scale<float>(static_cast<cx<float> *>(st->psi), st->dim, factor, st->stream));
    else
        HIP_TRY(launch_scale<double>(static_cast<cx<double> *>(st->psi), st->dim, factor, st->stream));
    st->counters[0]++;
    return 0;
}

int qsim_expect_pauli(qsim_state *st, const qsim_pauli *terms, int nterms, double *re, double *im)
{
    if (!st) return fail("null state");
    if (nterms < 0 || (nterms > 0 && !terms)) return fail("expect_pauli: bad term array");
    if (st->n < 6) return fail("expect_pauli: needs at least 6 qubits");
    if (flush_perm(st)) return -1;
    if (st->dtype == QSIM_C64) return pauli_t<float>(st, terms, nterms, nullptr, re, im);
    return pauli_t<double>(st, terms, nterms, nullptr, re, im);
}

int qsim_adjoint(qsim_state *st, const qsim_pauli *terms, int nterms, const qsim_op *rev_ops, int nops,
                 const void *mats, uint64_t mat_count, int nprobes, double *probes_out, double *expect_out)
{
    if (!st) return fail("null state");
    if (st->n < 6) return fail("adjoint: needs at least 6 qubits");
    if (nprobes < 0 || (nprobes > 0 && !probes_out)) return fail("adjoint: bad probe output");
    if (flush_perm(st)) return -1;
    if (check_ops(st->n, rev_ops, nops, mat_count, true)) return -1;
    for (int i = 0; i < nops; ++i)
        if (rev_ops[i].kind == QSIM_OP_PROBE && (rev_ops[i].q1 < 0 || rev_ops[i].q1 >= nprobes))
            return fail("adjoint: probe slot out of range");
    if (st->dtype == QSIM_C64)
        return adjoint_t<float>(st, terms, nterms, rev_ops, nops, mats, mat_count, nprobes, probes_out, expect_out);
    return adjoint_t<double>(st, terms, nterms, rev_ops, nops, mats, mat_count, nprobes, probes_out, expect_out);
}

int qsim_set_tile_params(qsim_state *st, int K, int cmin)
{
    if (!st) return fail("null state");
    const int kmax = st->dtype == QSIM_C64 ? 13 : 12;
    if (K < 6 || K > kmax) return fail("K out of range");
    if (cmin < 0 || cmin > K - 2) return fail("cmin out of range");
    st->K = K;
    st->cmin = cmin;
    return 0;
}

int qsim_set_jit(qsim_state *st, int enable)
{
    if (!st) return fail("null state");
    st->jit = enable != 0;
    return 0;
}

int qsim_set_fusion(qsim_state *st, int mode)
{
    if (!st) return fail("null state");
    if (mode < 0 || mode > 4) return fail("fusion mode must be between 0 and 4");
    st->fuse = mode;
    return 0;
}

int qsim_set_fusion_bars(qsim_state *st, double bar, double bar5)
{
    if (!st) return fail("null state");
    if (!(bar >= 0.0 && bar <= 4096.0) || !(bar5 >= 0.0 && bar5 <= 4096.0))
        return fail("fusion bars must be between 0 and 4096");
    st->fuse_bar = bar;
    st->fuse_bar5 = bar5;
    return 0;
}

int qsim_jit_compile_check(int num_qubits, int dtype, const qsim_op *ops, int nops, const void *mats,
                           uint64_t mat_count, int K, int cmin, int fuse, const int32_t *in_map,
                           const int32_t *out_map, int *out_folded)
{
    if (out_folded) *out_folded = 0;
    if (dtype != QSIM_C64 && dtype != QSIM_C128) return fail("unknown dtype");
    if (check_ops(num_qubits, ops, nops, mat_count, true)) return -1;
    std::vector<PlanOp> pops = to_plan_ops(ops, nops);
    int nv = 1;
    for (int i = 0; i < nops; ++i)
        if (ops[i].kind == QSIM_OP_PROBE) nv = 2;
    Plan plan;
    std::string err;
    if (!plan_passes(num_qubits, pops.data(), nops, mats, (int)amp_bytes(dtype), K, cmin,
                     nv == 2 ? kMaxProbes : 1, plan, err, nv == 1 ? fuse : 0,
                     nv == 1 ? in_map : nullptr, nv == 1 ? out_map : nullptr))
        return fail(err);
    if (!plan.out_map.empty() && !jit_mapped_last(plan, dtype == QSIM_C128, nv)) plan.out_map.clear();
    if (out_folded) *out_folded = plan.out_map.empty() ? 0 : 1;
    if (!jit_compile_only(plan, dtype == QSIM_C128, nv, "gfx950", err))
        return fail("run-time kernel compile: " + err);
    return (int)plan.passes.size();
}

int qsim_choose_layout(int num_qubits, int dtype, int cmin, const int32_t *perm_map, const qsim_op *ops, int nops,
                       int32_t *tgt_out)
{
    if (num_qubits < 1 || num_qubits > 63) return fail("num_qubits must be between 1 and 63");
    if (dtype != QSIM_C64 && dtype != QSIM_C128) return fail("unknown dtype");
    if (cmin < 0 || cmin > num_qubits) return fail("cmin out of range");
    if (nops < 0 || (nops > 0 && !ops) || !perm_map || !tgt_out) return fail("null argument");
    uint64_t seen = 0;
    for (int p = 0; p < num_qubits; ++p) {
        if (perm_map[p] < 0 || perm_map[p] >= num_qubits || ((seen >> perm_map[p]) & 1))
            return fail("perm_map is not a permutation");
        seen |= 1ull << perm_map[p];
    }
    for (int i = 0; i < nops; ++i) {
        const qsim_op &o = ops[i];
        const bool two = o.kind == QSIM_OP_2Q;
        if (o.kind == QSIM_OP_3Q || o.kind == QSIM_OP_4Q) {
            if (check_ops(num_qubits, &o, 1, UINT64_MAX, false)) return fail("invalid op");
            continue;
        }
        if ((o.kind != QSIM_OP_1Q && !two && o.kind != QSIM_OP_PROBE) || o.q0 < 0 || o.q0 >= num_qubits ||
            (two && (o.q1 < 0 || o.q1 >= num_qubits || o.q1 == o.q0)))
            return fail("invalid op");
    }
    choose_layout_impl(num_qubits, cmin, dtype == QSIM_C64 ? 18 : 17, perm_map, ops, nops, tgt_out);
    return 0;
}

int qsim_sample(qsim_state *st, uint64_t shots, uint64_t seed, uint64_t *indices_out)
{
    return qsim::sample_impl(st, shots, seed, indices_out, nullptr, 0);
}

int qsim_device_memory(uint64_t *free_bytes, uint64_t *total_bytes)
{
    if (!free_bytes || !total_bytes) return fail("null argument");
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return 0;
}

int qsim_profile(qsim_state *st, int enable)
{
    if (!st) return fail("null state");
    st->prof = enable != 0;
    return 0;
}

int qsim_profile_read(qsim_state *st, double *ms_out, int64_t *count_out, int nclasses)
{
    if (!st || !ms_out || !count_out) return fail("null argument");
    HIP_TRY(hipStreamSynchronize(st->stream));
    const int m = std::min(nclasses, kProfClasses);
    for (int c = 0; c < m; ++c) {
        ms_out[c] = 0.0;
        count_out[c] = 0;
    }
    for (auto &r : st->recs) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
        if (r.cls < m) {
            ms_out[r.cls] += ms;
            count_out[r.cls] += 1;
        }
        st->pool.push_back(r.a);
        st->pool.push_back(r.b);
    }
    st->recs.clear();
    return m;
}

int qsim_get_counters(const qsim_state *st, int64_t *out, int n)
{
    if (!st || !out) return fail("null argument");
    const int m = std::min(n, 7);
    for (int i = 0; i < m; ++i) out[i] = st->counters[i];
    return m;
}

int qsim_plan_describe(int num_qubits, int dtype, const qsim_op *ops, int nops, const void *mats,
                       uint64_t mat_count, int K, int cmin, const int32_t *in_map, const int32_t *out_map,
                       int32_t *desc, int desc_cap, int32_t *order)
{
    if (dtype != QSIM_C64 && dtype != QSIM_C128) return fail("unknown dtype");
    if (check_ops(num_qubits, ops, nops, mat_count, true)) return -1;
    std::vector<PlanOp> pops = to_plan_ops(ops, nops);
    Plan plan;
    std::string err;
    const bool maps = in_map || out_map;
    if (!plan_passes(num_qubits, pops.data(), nops, mats, (int)amp_bytes(dtype), K, cmin, maps ? 1 : kMaxProbes,
                     plan, err, 0, in_map, out_map))
        return fail(err);
    int w = 0;
    for (const PassDev &p : plan.passes) {
        if (w + 3 + p.K + 6 * p.nsteps > desc_cap) return fail("plan_describe: descriptor buffer too small");
        desc[w++] = p.K;
        desc[w++] = p.cfast;
        desc[w++] = p.nsteps;
        for (int k = 0; k < p.K; ++k) desc[w++] = p.pos[k];
        for (int s = p.step_off; s < p.step_off + p.nsteps; ++s) {
            const StepDev &sd = plan.steps[s];
            for (int t = 0; t < kG; ++t) desc[w++] = sd.bits[t];
            desc[w++] = sd.nrounds;
            desc[w++] = sd.dense >= 0 ? 1 : 0;
        }
    }
    for (size_t i = 0; i < plan.order.size(); ++i) order[i] = plan.order[i];
    return (int)plan.passes.size();
}

}  // extern "C"
