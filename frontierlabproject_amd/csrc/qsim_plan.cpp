// Chunk-pass planner (host only).  See qsim_plan.h.
#include "qsim_plan.h"

#include <algorithm>
#include <cstring>
#include <utility>

namespace qsim {

namespace {

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

struct Cd {
    double re, im;
};

Cd mat_at(const void *mats, int elem_bytes, int64_t idx)
{
    if (elem_bytes == 8) {
        const float *f = static_cast<const float *>(mats);
        return {f[2 * idx], f[2 * idx + 1]};
    }
    const double *d = static_cast<const double *>(mats);
    return {d[2 * idx], d[2 * idx + 1]};
}

inline bool is0(Cd c) { return c.re == 0.0 && c.im == 0.0; }
inline bool is1(Cd c) { return c.re == 1.0 && c.im == 0.0; }

uint64_t fill_low(uint64_t set, int width, int target)
{
    for (int b = 0; b < width && popc(set) < target; ++b) set |= 1ull << b;
    return set;
}

void make_pass(int n, int K, uint64_t P, PassDev &pd, int *loc)
{
    pd = PassDev{};
    pd.K = K;
    int k = 0, c = 0;
    for (int b = 0; b < n; ++b) {
        if (P >> b & 1) {
            loc[b] = k;
            pd.pos[k++] = b;
        } else {
            pd.comp[c++] = b;
        }
    }
    pd.ncomp = c;
    pd.cfast = 0;
    while (pd.cfast < K && pd.pos[pd.cfast] == pd.cfast) ++pd.cfast;
}

// pair slot index of register bits x < y, in round order
int pair_index(int x, int y)
{
    static const int tab[4][4] = {{-1, 0, 2, 4}, {-1, -1, 5, 3}, {-1, -1, -1, 1}, {-1, -1, -1, -1}};
    return tab[x][y];
}

// first slot index > last with index % 10 == target, or -1
int next_slot(int last, int target)
{
    int j = last + 1;
    j += ((target - j % kSlotsPerRound) + kSlotsPerRound) % kSlotsPerRound;
    return j < kSlots ? j : -1;
}

struct StepBuilder {
    int nreg = 0;
    int regbit[kG];        // register t -> chunk-local bit
    int last[kG];          // last slot used per register bit
    StepDev sd;
    std::vector<std::pair<int, int>> placed;  // (slot * 2 + sub, op index)

    StepBuilder()
    {
        std::memset(&sd, 0, sizeof(sd));
        for (int t = 0; t < kG; ++t) last[t] = -1, regbit[t] = -1;
        for (int t = 0; t < k1qSlots; ++t) sd.probe[t] = -1;
    }
    int reg_of(int chunk_bit) const
    {
        for (int t = 0; t < nreg; ++t)
            if (regbit[t] == chunk_bit) return t;
        return -1;
    }
};

struct Ctx {
    const PlanOp *ops;
    const void *mats;
    int eb;
    Plan *plan;
    PassDev *pd;
    int max_probes;
};

int push_mats(Plan &p, const Cd *v, int count)
{
    const int off = (int)(p.mats.size() / 2);
    for (int i = 0; i < count; ++i) {
        p.mats.push_back(v[i].re);
        p.mats.push_back(v[i].im);
    }
    return off;
}

// Place op `idx` (register bits rx, ry already assigned) into the step.
bool place(Ctx &c, StepBuilder &sb, int idx, int rx, int ry)
{
    const PlanOp &op = c.ops[idx];
    Plan &p = *c.plan;
    if (op.kind == OP_1Q || op.kind == OP_PROBE) {
        const int t = rx;
        Cd u[4];
        for (int k = 0; k < 4; ++k) u[k] = mat_at(c.mats, c.eb, (int64_t)op.mat_off + k);
        if (op.kind == OP_1Q && sb.last[t] >= 0 && sb.last[t] % kSlotsPerRound == t &&
            sb.sd.kind[sb.last[t]] == S1_PROBE) {
            // a gate right after a probe on the same bit shares the probe's slot
            const int j = sb.last[t];
            const int off = sb.sd.mat[j];
            for (int k = 0; k < 4; ++k) {
                p.mats[2 * (off + 4 + k)] = u[k].re;
                p.mats[2 * (off + 4 + k) + 1] = u[k].im;
            }
            sb.sd.kind[j] = S1_PROBE_GEN;
            sb.placed.push_back({2 * j + 1, idx});
            return true;
        }
        const int j = next_slot(sb.last[t], t);
        if (j < 0) return false;
        if (op.kind == OP_PROBE) {
            if (c.pd->nprobe >= c.max_probes) return false;
            Cd g8[8] = {u[0], u[1], u[2], u[3], {1, 0}, {0, 0}, {0, 0}, {1, 0}};
            sb.sd.kind[j] = S1_PROBE;
            sb.sd.mat[j] = push_mats(p, g8, 8);
            sb.sd.probe[(j / kSlotsPerRound) * 4 + t] = c.pd->nprobe++;
            p.probe_ids.push_back(op.probe);
        } else if (is0(u[1]) && is0(u[2])) {
            Cd d[2] = {u[0], u[3]};
            sb.sd.kind[j] = S1_DIAG;
            sb.sd.mat[j] = push_mats(p, d, 2);
        } else {
            sb.sd.kind[j] = S1_GEN;
            sb.sd.mat[j] = push_mats(p, u, 4);
        }
        sb.last[t] = j;
        sb.placed.push_back({2 * j, idx});
        return true;
    }
    // pair op: orient to (first, second) = (lower, higher register bit)
    const int lo = rx < ry ? rx : ry, hi = rx < ry ? ry : rx;
    const int j = next_slot(std::max(sb.last[rx], sb.last[ry]), 4 + pair_index(lo, hi));
    if (j < 0) return false;
    auto sw = [](int i) { return ((i & 1) << 1) | ((i >> 1) & 1); };
    Cd m[16];
    for (int r = 0; r < 4; ++r)
        for (int q = 0; q < 4; ++q) {
            // op index b_q0 + 2 b_q1; slot index b_first + 2 b_second
            const int rr = rx == lo ? r : sw(r), qq = rx == lo ? q : sw(q);
            m[4 * r + q] = mat_at(c.mats, c.eb, (int64_t)op.mat_off + 4 * rr + qq);
        }
    bool diag = true;
    for (int r = 0; r < 4; ++r)
        for (int q = 0; q < 4; ++q)
            if (r != q && !is0(m[4 * r + q])) diag = false;
    // control on the first bit: indices with b_first = 0 ({0, 2}) untouched
    bool cf = is1(m[0]) && is1(m[10]) && is0(m[2]) && is0(m[8]);
    for (int a : {0, 2})
        for (int b : {1, 3}) cf = cf && is0(m[4 * a + b]) && is0(m[4 * b + a]);
    // control on the second bit: indices {0, 1} untouched
    bool cs = is1(m[0]) && is1(m[5]) && is0(m[1]) && is0(m[4]);
    for (int a : {0, 1})
        for (int b : {2, 3}) cs = cs && is0(m[4 * a + b]) && is0(m[4 * b + a]);
    if (diag) {
        Cd d[4] = {m[0], m[5], m[10], m[15]};
        sb.sd.kind[j] = S2_DIAG;
        sb.sd.mat[j] = push_mats(p, d, 4);
    } else if (cf) {
        Cd u[4] = {m[5], m[7], m[13], m[15]};  // rows/cols 1, 3
        sb.sd.kind[j] = S2_CTRL_FIRST;
        sb.sd.mat[j] = push_mats(p, u, 4);
    } else if (cs) {
        Cd u[4] = {m[10], m[11], m[14], m[15]};  // rows/cols 2, 3
        sb.sd.kind[j] = S2_CTRL_SECOND;
        sb.sd.mat[j] = push_mats(p, u, 4);
    } else {
        sb.sd.kind[j] = S2_GEN;
        sb.sd.mat[j] = push_mats(p, m, 16);
    }
    sb.last[rx] = sb.last[ry] = j;
    sb.placed.push_back({2 * j, idx});
    return true;
}

}  // namespace

bool plan_passes(int n, const PlanOp *ops, int nops, const void *mats, int elem_bytes, int K, int cmin,
                 int max_probes, Plan &out, std::string &err)
{
    out = Plan();
    if (n < kG || n > 63) {
        err = "tile plan: state must have between 4 and 63 qubits";
        return false;
    }
    if (elem_bytes != 8 && elem_bytes != 16) {
        err = "tile plan: unknown matrix element size";
        return false;
    }
    if (K > kMaxK) K = kMaxK;
    if (K > n) K = n;
    if (K < kG) K = kG;
    if (n - K > kMaxComp) {
        err = "tile plan: too many qubits outside one chunk";
        return false;
    }
    if (cmin > K - 2) cmin = K - 2;
    if (cmin < 0) cmin = 0;
    if (max_probes < 1) max_probes = 1;
    for (int i = 0; i < nops; ++i) {
        const PlanOp &op = ops[i];
        if (op.kind != OP_1Q && op.kind != OP_2Q && op.kind != OP_PROBE) {
            err = "tile plan: unknown op kind";
            return false;
        }
        if (op.q0 < 0 || op.q0 >= n || (op.kind == OP_2Q && (op.q1 < 0 || op.q1 >= n))) {
            err = "tile plan: qubit out of range";
            return false;
        }
        if (op.kind == OP_2Q && op.q0 == op.q1) {
            err = "tile plan: q0 == q1";
            return false;
        }
    }

    std::vector<int> rem(nops);
    for (int i = 0; i < nops; ++i) rem[i] = i;
    const uint64_t low = cmin ? ((1ull << cmin) - 1) : 0;

    while (!rem.empty()) {
        // ---- pass: chunk bit set ----
        uint64_t P = low, blocked = 0;
        int nprobe = 0;
        std::vector<int> chosen, left;
        for (int idx : rem) {
            const PlanOp &op = ops[idx];
            uint64_t m = 1ull << op.q0;
            if (op.kind == OP_2Q) m |= 1ull << op.q1;
            const bool probe = op.kind == OP_PROBE;
            if ((m & blocked) || popc(P | m) > K || (probe && nprobe >= max_probes)) {
                blocked |= m;
                left.push_back(idx);
                continue;
            }
            P |= m;
            nprobe += probe;
            chosen.push_back(idx);
        }
        P = fill_low(P, n, K);
        PassDev pd;
        int loc[64];
        make_pass(n, K, P, pd, loc);
        pd.step_off = (int32_t)out.steps.size();
        pd.probe_off = (int32_t)out.probe_ids.size();
        pd.nprobe = 0;
        Ctx ctx{ops, mats, elem_bytes, &out, &pd, max_probes};

        // ---- register steps ----
        std::vector<int> srem = chosen;
        while (!srem.empty()) {
            StepBuilder sb;
            uint64_t sblk = 0;
            std::vector<int> sleft;
            for (int idx : srem) {
                const PlanOp &op = ops[idx];
                const int c0 = loc[op.q0];
                const int c1 = op.kind == OP_2Q ? loc[op.q1] : -1;
                uint64_t m = 1ull << c0;
                if (c1 >= 0) m |= 1ull << c1;
                if (m & sblk) {
                    sblk |= m;
                    sleft.push_back(idx);
                    continue;
                }
                StepBuilder save = sb;
                const size_t mats_before = out.mats.size(), ids_before = out.probe_ids.size();
                const int np_before = pd.nprobe;
                int rx = sb.reg_of(c0), ry = c1 >= 0 ? sb.reg_of(c1) : -1;
                bool ok = true;
                if (rx < 0) {
                    if (sb.nreg < kG) {
                        rx = sb.nreg;
                        sb.regbit[sb.nreg++] = c0;
                    } else {
                        ok = false;
                    }
                }
                if (ok && c1 >= 0 && ry < 0) {
                    if (sb.nreg < kG) {
                        ry = sb.nreg;
                        sb.regbit[sb.nreg++] = c1;
                    } else {
                        ok = false;
                    }
                }
                if (ok) ok = place(ctx, sb, idx, rx, ry);
                if (!ok) {
                    sb = save;
                    out.mats.resize(mats_before);
                    out.probe_ids.resize(ids_before);
                    pd.nprobe = np_before;
                    sblk |= m;
                    sleft.push_back(idx);
                }
            }
            // fill unused register bits with the lowest free chunk bits
            for (int b = 0; b < K && sb.nreg < kG; ++b)
                if (sb.reg_of(b) < 0) sb.regbit[sb.nreg++] = b;
            for (int t = 0; t < kG; ++t) sb.sd.bits[t] = sb.regbit[t];
            int sorted_bits[kG];
            std::memcpy(sorted_bits, sb.regbit, sizeof(sorted_bits));
            std::sort(sorted_bits, sorted_bits + kG);
            for (int t = 0; t < kG; ++t) sb.sd.sbits[t] = sorted_bits[t];
            std::sort(sb.placed.begin(), sb.placed.end());
            for (auto &pr : sb.placed) out.order.push_back(pr.second);
            out.steps.push_back(sb.sd);
            srem.swap(sleft);
        }
        pd.nsteps = (int32_t)out.steps.size() - pd.step_off;
        out.passes.push_back(pd);
        rem.swap(left);
    }
    return true;
}

bool plan_pauli(int n, const uint64_t *xmasks, int nterms, int K, int cmin, std::vector<PauliPass> &out,
                std::string &err)
{
    out.clear();
    if (n < kG || n > 63) {
        err = "pauli plan: state must have between 4 and 63 qubits";
        return false;
    }
    if (K > kMaxK) K = kMaxK;
    if (K > n) K = n;
    if (n - K > kMaxComp) {
        err = "pauli plan: too many qubits outside one chunk";
        return false;
    }
    if (cmin > K) cmin = K;
    const uint64_t full = (1ull << n) - 1;
    std::vector<int> rem;
    for (int t = 0; t < nterms; ++t) {
        if (xmasks[t] & ~full) {
            err = "pauli plan: term acts on a qubit out of range";
            return false;
        }
        rem.push_back(t);
    }
    const uint64_t low = cmin ? ((1ull << cmin) - 1) : 0;
    while (!rem.empty()) {
        uint64_t P = low;
        std::vector<int> chosen, left;
        for (int t : rem) {
            if (popc(P | xmasks[t]) <= K) {
                P |= xmasks[t];
                chosen.push_back(t);
            } else {
                left.push_back(t);
            }
        }
        if (chosen.empty()) {
            err = "pauli plan: a term's X support does not fit in one chunk";
            return false;
        }
        P = fill_low(P, n, K);
        PauliPass pp;
        int loc[64];
        make_pass(n, K, P, pp.pd, loc);
        for (int t : chosen) {
            int32_t xl = 0;
            for (int b = 0; b < n; ++b)
                if (xmasks[t] >> b & 1) xl |= 1 << loc[b];
            pp.terms.push_back(t);
            pp.xloc.push_back(xl);
        }
        out.push_back(pp);
        rem.swap(left);
    }
    return true;
}

}  // namespace qsim
