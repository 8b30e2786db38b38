// Chunk-pass planner (host only).  See qsim_plan.h.
#include "qsim_plan.h"

#include <algorithm>
#include <cmath>
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
inline bool isre(Cd c) { return c.im == 0.0; }
inline bool isim(Cd c) { return c.re == 0.0; }
inline bool iseq(Cd c, double re, double im) { return c.re == re && c.im == im; }
inline Cd cmulh(Cd a, Cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Cd cdivh(Cd a, Cd b)
{
    const double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

uint64_t fill_low(uint64_t set, int width, int target)
{
    for (int b = 0; b < width && popc(set) < target; ++b) set |= 1ull << b;
    return set;
}

// Ops of `masks` (qubit masks in program order) that one pass over chunk bits
// M runs: an op runs when its qubits lie in M (a diagonal one, isd[i], needs
// no chunk bits) and no earlier op left behind shares one of them.  Clears run
// entries of `live`; returns how many non-diagonal ops ran.
int run_ops(const std::vector<uint64_t> &masks, const std::vector<char> &isd, std::vector<char> &live, uint64_t M)
{
    uint64_t blocked = 0;
    int ran = 0;
    for (size_t i = 0; i < masks.size(); ++i) {
        if (!live[i]) continue;
        const uint64_t m = masks[i];
        if (isd[i]) {
            if (m & blocked)
                blocked |= m;
            else
                live[i] = 0;
            continue;
        }
        if ((m & blocked) || (m & ~M)) {
            blocked |= m;
            // no later op can run once every chunk bit is blocked
            if (!(M & ~blocked)) break;
            continue;
        }
        live[i] = 0;
        ++ran;
    }
    return ran;
}

// Passes the in-order rule needs for the live ops: each pass takes ops in
// program order while the chunk (the `low` bits plus the ops' qubits) has at
// most K bits.
int greedy_passes(const std::vector<uint64_t> &masks, const std::vector<char> &isd, std::vector<char> live,
                  uint64_t low, int K)
{
    int passes = 0;
    for (;;) {
        uint64_t P = low, blocked = 0;
        bool any = false;
        for (size_t i = 0; i < masks.size(); ++i) {
            if (!live[i]) continue;
            any = true;
            const uint64_t m = masks[i];
            if (isd[i]) {  // needs no chunk bits
                if (m & blocked)
                    blocked |= m;
                else
                    live[i] = 0;
                continue;
            }
            if ((m & blocked) || popc(P | m) > K) {
                blocked |= m;
                // a full chunk whose bits are all blocked takes no later op
                if (popc(P) >= K && !(P & ~blocked)) break;
                continue;
            }
            P |= m;
            live[i] = 0;
        }
        if (!any) return passes;
        ++passes;
    }
}

constexpr double kSearchMax = 20000;  // chunk-bit sets tried per pass
constexpr size_t kSearchDeep = 64;    // of which at most this many get the full look-ahead
// Source bits from log2(2 MiB / amplitude size) up (18 for complex64, 17
// for complex128) lie 2 MiB apart or more: a chunk gathering over them reads
// rows from separate address-translation pages.
inline int far_bit(int elem_bytes) { return elem_bytes == 8 ? 18 : 17; }

// Free chunk bits of one pass chosen by search when the candidates are few:
// over every set of `free` qubits among those the remaining ops touch, the
// one after which the in-order rule needs the fewest further passes (the
// in-order rule alone can strand an op that opens a third pass), then the one
// that leaves the op list's last pass room for the output order's row bits
// `out_rows` (when an output order is to be written), then the one with the
// fewest bits of `far` (bits the pass reads from far-apart rows: for a folded
// layout change, those stored high in the old order), then the one running
// the most ops now.  When the sets are many, only the kSearchDeep that
// run the most ops now get the look-ahead.  Returns 0 when the search does not
// apply (all qubits fit, or too many sets).
uint64_t search_chunk(const PlanOp *ops, const std::vector<int> &rem, uint64_t P0, int Kp, uint64_t low, int K,
                      uint64_t far, uint64_t out_rows, const std::vector<uint8_t> &diag)
{
    const int free = Kp - popc(P0);
    std::vector<uint64_t> masks;
    std::vector<char> isd;
    masks.reserve(rem.size());
    uint64_t Q = 0;
    for (int idx : rem) {
        const uint64_t m = op_mask(ops[idx]);
        masks.push_back(m);
        isd.push_back(diag[idx] != 0);
        if (!diag[idx]) Q |= m;  // diagonal gates need no chunk bits
    }
    Q &= ~P0;
    const int q = popc(Q);
    if (free <= 0 || q <= free) return 0;
    double combos = 1;
    for (int i = 0; i < free; ++i) combos = combos * (q - i) / (i + 1);
    if (combos > kSearchMax) return 0;
    int bit[64], nb = 0;
    for (int b = 0; b < 64; ++b)
        if ((Q >> b) & 1) bit[nb++] = b;
    // every candidate: how many ops it runs now (cheap); the full look-ahead
    // (passes the in-order rule needs afterwards) only for the best-ranked
    // kSearchDeep of them when there are more
    struct Cand {
        uint64_t M;
        int ran;
    };
    std::vector<Cand> cands;
    cands.reserve((size_t)combos);
    std::vector<char> live(masks.size());
    for (uint64_t sel = (1ull << free) - 1; sel < (1ull << q);) {
        uint64_t M = P0;
        for (int i = 0; i < q; ++i)
            if ((sel >> i) & 1) M |= 1ull << bit[i];
        std::fill(live.begin(), live.end(), 1);
        const int ran = run_ops(masks, isd, live, M);
        if (ran > 0) cands.push_back({M, ran});
        const uint64_t c = sel & (~sel + 1), r = sel + c;  // next set of the same size
        sel = (((r ^ sel) >> 2) / c) | r;
    }
    if (cands.size() > kSearchDeep)
        std::stable_sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.ran > b.ran; });
    uint64_t best = 0;
    int best_rest = 1 << 30, best_far = 1 << 30, best_ran = -1, best_nofold = 2;
    const size_t deep = std::min(cands.size(), kSearchDeep);
    for (size_t k = 0; k < deep; ++k) {
        const uint64_t M = cands[k].M;
        const int ran = cands[k].ran;
        std::fill(live.begin(), live.end(), 1);
        run_ops(masks, isd, live, M);
        const int rest = greedy_passes(masks, isd, live, low, K);
        const int nfar = popc(M & far);
        // with an output order to write (out_rows: its row bits), prefer a
        // last pass with room for them: this pass when nothing is left, else
        // the next one when it is the last
        int nofold = 0;
        if (out_rows && rest <= 1) {
            uint64_t U = 0;
            for (size_t i = 0; i < masks.size(); ++i)
                if (!isd[i] && (rest == 0 ? !live[i] : live[i])) U |= masks[i];
            nofold = rest == 0 ? popc(P0 | U | out_rows) > Kp : popc(low | U | out_rows) > K;
        }
        if (rest < best_rest ||
            (rest == best_rest &&
             (nofold < best_nofold ||
              (nofold == best_nofold && (nfar < best_far || (nfar == best_far && ran > best_ran)))))) {
            best = M;
            best_rest = rest;
            best_nofold = nofold;
            best_far = nfar;
            best_ran = ran;
        }
    }
    return best;
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

constexpr int kG5 = kG + 1;  // register bits of a 32x32 dense step

// Everything of a step under construction but its dense matrix (an op that
// does not fit leaves the matrix untouched, so a retry saves only this).
struct StepCore {
    int nreg = 0;
    int maxreg = kG;       // register bits this step may take (kG5: a 32x32 dense step)
    int regbit[kG5];       // register t -> chunk-local bit
    int last[kG5];         // last slot used per register bit
    StepDev sd;
    std::vector<std::pair<int, int>> placed;  // (slot * 2 + sub, op index)
    bool dense = false;    // the step applies one fused 16x16 matrix
    bool has_probe = false;
    Cd factor = {1.0, 0.0};  // scalar taken out of this step's slot gates
    double cost = 0.0;     // estimated multiply-adds per amplitude of the slot form
    uint64_t lock = 0;     // chunk bits that stay outside the registers (free diagonals)
    uint64_t endbits = 0;  // chunk bits of free diagonals: no later non-diagonal gate here
    std::vector<FreeDiag> fd;  // free diagonals, applied after the slot rounds

    StepCore()
    {
        std::memset(&sd, 0, sizeof(sd));
        sd.dense = -1;
        sd.bit4 = -1;
        for (int t = 0; t < kG5; ++t) last[t] = -1, regbit[t] = -1;
        for (int t = 0; t < k1qSlots; ++t) sd.probe[t] = -1;
    }
    int dim() const { return 1 << maxreg; }
    int reg_of(int chunk_bit) const
    {
        for (int t = 0; t < nreg; ++t)
            if (regbit[t] == chunk_bit) return t;
        return -1;
    }
};

struct StepBuilder : StepCore {
    Cd M[1024];            // dense matrix (row-major, dim() x dim()), valid when dense
};

struct Ctx {
    const PlanOp *ops;
    const void *mats;
    int eb;
    Plan *plan;
    PassDev *pd;
    int max_probes;
    const int *loc;      // state bit -> chunk-local bit
    bool real_scale;     // non-unit scalars may be factored out (no probes)
    bool generic;        // generic slot forms only (S1_GEN, S1_DIAG; S2_CPHASE, S2_CU_*)
    const std::vector<uint8_t> *diag;  // per op: a diagonal 1q/2q gate that may be a free diagonal
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

// u0 can be taken out of a 1q matrix without growing its other entries:
// |u0| is at least |u1| and |u2| (at least 1/sqrt2 for a unitary).
inline bool factorable(const Cd *u)
{
    const double a = u[0].re * u[0].re + u[0].im * u[0].im;
    return a > 0.0 && a >= u[1].re * u[1].re + u[1].im * u[1].im && a >= u[2].re * u[2].re + u[2].im * u[2].im;
}

// Slot form of the 1q gate u in slot j (not a probe): each gate is
// classified exactly (entries compared with 0, 1, i) and runs the cheapest
// arithmetic for its form; scalars taken out go to the step's factor.
void classify_1q(Ctx &c, StepBuilder &sb, const Cd *u, int j)
{
    Plan &p = *c.plan;
    if (c.generic) {  // diagonal or general, the matrix as given
        if (is0(u[1]) && is0(u[2])) {
            const Cd d[2] = {u[0], u[3]};
            sb.sd.kind[j] = S1_DIAG;
            sb.sd.mat[j] = push_mats(p, d, 2);
            sb.cost += 4.0;
        } else {
            sb.sd.kind[j] = S1_GEN;
            sb.sd.mat[j] = push_mats(p, u, 4);
            sb.cost += 8.0;
        }
        return;
    }
    if (is0(u[1]) && is0(u[2]) &&
        (c.real_scale ? !is0(u[0]) : std::fabs(u[0].re * u[0].re + u[0].im * u[0].im - 1.0) < 1e-12)) {
        // diag(u0, u3) = u0 * diag(1, u3/u0): the phase u0 is applied later
        const Cd d = cdivh(u[3], u[0]);
        sb.factor = cmulh(sb.factor, u[0]);
        if (iseq(d, 1, 0)) {
            sb.sd.kind[j] = S1_NOP;
        } else if (iseq(d, -1, 0)) {
            sb.sd.kind[j] = S1_Z;
            sb.cost += 0.5;
        } else if (iseq(d, 0, 1)) {
            sb.sd.kind[j] = S1_PI;
            sb.cost += 0.5;
        } else if (iseq(d, 0, -1)) {
            sb.sd.kind[j] = S1_MI;
            sb.cost += 0.5;
        } else {
            sb.sd.kind[j] = S1_PHASE;
            sb.sd.mat[j] = push_mats(p, &d, 1);
            sb.cost += 2.0;
        }
    } else if (is0(u[1]) && is0(u[2])) {
        Cd d[2] = {u[0], u[3]};
        sb.sd.kind[j] = S1_DIAG;
        sb.sd.mat[j] = push_mats(p, d, 2);
        sb.cost += 4.0;
    } else if (is0(u[0]) && is0(u[3]) && is1(u[1]) && is1(u[2])) {
        sb.sd.kind[j] = S1_X;
        sb.cost += 0.5;
    } else if (is0(u[0]) && is0(u[3]) && iseq(u[1], 0, -1) && iseq(u[2], 0, 1)) {
        sb.sd.kind[j] = S1_Y;
        sb.cost += 1.0;
    } else if (c.real_scale && isre(u[0]) && u[0].re != 0.0 && iseq(u[1], u[0].re, 0) &&
               iseq(u[2], u[0].re, 0) && iseq(u[3], -u[0].re, 0)) {
        // s * [[1, 1], [1, -1]] (H): butterfly, s applied later
        sb.factor = cmulh(sb.factor, u[0]);
        sb.sd.kind[j] = S1_BFLY;
        sb.cost += 2.0;
    } else if (isre(u[0]) && isre(u[1]) && isre(u[2]) && isre(u[3])) {
        if (c.real_scale && u[0].re == u[3].re && factorable(u)) {
            // u0 * [[1, a], [b, 1]] (RY): u0 applied later
            Cd ab = {u[1].re / u[0].re, u[2].re / u[0].re};
            sb.factor = cmulh(sb.factor, u[0]);
            sb.sd.kind[j] = S1_REAL1;
            sb.cost += 2.0;
            sb.sd.mat[j] = push_mats(p, &ab, 1);
        } else if (c.real_scale && factorable(u)) {
            // u0 * [[1, a], [b, d]], all real: u0 applied later
            Cd abd[2] = {{u[1].re / u[0].re, u[2].re / u[0].re}, {u[3].re / u[0].re, 0.0}};
            sb.factor = cmulh(sb.factor, u[0]);
            sb.sd.kind[j] = S1_REALG;
            sb.cost += 3.0;
            sb.sd.mat[j] = push_mats(p, abd, 2);
        } else {
            Cd r[2] = {{u[0].re, u[1].re}, {u[2].re, u[3].re}};
            sb.sd.kind[j] = S1_REAL;
            sb.cost += 4.0;
            sb.sd.mat[j] = push_mats(p, r, 2);
        }
    } else if (isre(u[0]) && isre(u[3]) && isim(u[1]) && isim(u[2])) {
        if (c.real_scale && u[0].re == u[3].re && factorable(u)) {
            // u0 * [[1, iq], [ir, 1]] (RX): u0 applied later
            Cd qr = {u[1].im / u[0].re, u[2].im / u[0].re};
            sb.factor = cmulh(sb.factor, u[0]);
            sb.sd.kind[j] = S1_RI1;
            sb.cost += 2.0;
            sb.sd.mat[j] = push_mats(p, &qr, 1);
        } else {
            Cd r[2] = {{u[0].re, u[1].im}, {u[2].im, u[3].re}};
            sb.sd.kind[j] = S1_RI;
            sb.cost += 4.0;
            sb.sd.mat[j] = push_mats(p, r, 2);
        }
    } else if (c.real_scale && factorable(u)) {
        // u0 * [[1, b], [c, d]]: u0 applied later
        Cd g[3] = {cdivh(u[1], u[0]), cdivh(u[2], u[0]), cdivh(u[3], u[0])};
        sb.factor = cmulh(sb.factor, u[0]);
        sb.sd.kind[j] = S1_GEN1;
        sb.cost += 6.0;
        sb.sd.mat[j] = push_mats(p, g, 3);
    } else {
        sb.sd.kind[j] = S1_GEN;
        sb.cost += 8.0;
        sb.sd.mat[j] = push_mats(p, u, 4);
    }
}

// Place op `idx` (register bits rx, ry already assigned) into the step.
// Returns 1 when placed, 0 when no slot is free, 2 when the op needs a dense step.
int place(Ctx &c, StepBuilder &sb, int idx, int rx, int ry)
{
    const PlanOp &op = c.ops[idx];
    Plan &p = *c.plan;
    if (op.kind == OP_3Q || op.kind == OP_4Q) return 2;  // a dense step
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
            sb.cost += 8.0;
            sb.placed.push_back({2 * j + 1, idx});
            return 1;
        }
        const int j = next_slot(sb.last[t], t);
        if (j < 0) return 0;
        if (op.kind == OP_PROBE) {
            if (c.pd->nprobe >= c.max_probes) return 0;
            Cd g8[8] = {u[0], u[1], u[2], u[3], {1, 0}, {0, 0}, {0, 0}, {1, 0}};
            sb.sd.kind[j] = S1_PROBE;
            sb.has_probe = true;
            sb.sd.mat[j] = push_mats(p, g8, 8);
            sb.sd.probe[(j / kSlotsPerRound) * 4 + t] = c.pd->nprobe++;
            p.probe_ids.push_back(op.probe);
        } else {
            // With u0 not the largest entry of its row, U = V X with V = U X
            // (columns exchanged), whose u0 is: the slot exchanges the pair
            // (a register rename in specialised kernels) and runs V's form.
            const Cd ux[4] = {u[1], u[0], u[3], u[2]};
            const bool swap = !c.generic && c.real_scale && !(is0(u[1]) && is0(u[2])) && !factorable(u) && factorable(ux) &&
                              !(is0(u[0]) && is0(u[3]) && ((is1(u[1]) && is1(u[2])) ||
                                                            (iseq(u[1], 0, -1) && iseq(u[2], 0, 1))));
            classify_1q(c, sb, swap ? ux : u, j);
            if (swap) sb.sd.kind[j] |= S1_SWAP;
        }
        sb.last[t] = j;
        sb.placed.push_back({2 * j, idx});
        return 1;
    }
    // pair op: orient to (first, second) = (lower, higher register bit)
    const int lo = rx < ry ? rx : ry, hi = rx < ry ? ry : rx;
    auto sw = [](int i) { return ((i & 1) << 1) | ((i >> 1) & 1); };
    Cd m[16];
    for (int r = 0; r < 4; ++r)
        for (int q = 0; q < 4; ++q) {
            // op index b_q0 + 2 b_q1; slot index b_first + 2 b_second
            const int rr = rx == lo ? r : sw(r), qq = rx == lo ? q : sw(q);
            m[4 * r + q] = mat_at(c.mats, c.eb, (int64_t)op.mat_off + 4 * rr + qq);
        }
    auto only = [&](std::initializer_list<int> ones, int skip) {
        for (int k = 0; k < 16; ++k) {
            if (k == skip) continue;
            bool one = false;
            for (int o : ones) one |= o == k;
            if (one ? !is1(m[k]) : !is0(m[k])) return false;
        }
        return true;
    };
    // controlled 2x2: identity on the indices where the control bit is 0,
    // U on the other two, nothing between the halves
    auto controlled = [&](int c0, int c1, int t0, int t1, Cd *u) {
        const int idx[4] = {c0, c1, t0, t1};
        for (int r = 0; r < 4; ++r)
            for (int q = 0; q < 4; ++q) {
                const Cd v = m[4 * idx[r] + idx[q]];
                if (r < 2 || q < 2) {
                    if (r == q ? !is1(v) : !is0(v)) return false;
                }
            }
        u[0] = m[4 * t0 + t0], u[1] = m[4 * t0 + t1], u[2] = m[4 * t1 + t0], u[3] = m[4 * t1 + t1];
        return true;
    };
    int kind;
    Cd cu[4];
    if (only({0, 5, 10}, 15)) kind = S2_CPHASE;                  // diag(1,1,1,d)
    else if (c.generic && controlled(0, 2, 1, 3, cu)) kind = S2_CU_FIRST;   // U on indices 1, 3
    else if (c.generic && controlled(0, 1, 2, 3, cu)) kind = S2_CU_SECOND;  // U on indices 2, 3
    else if (c.generic) return 2;                                  // needs a dense step
    else if (only({0, 10, 7, 13}, -1)) kind = S2_CX_FIRST;       // rows 1<->3
    else if (only({0, 5, 11, 14}, -1)) kind = S2_CX_SECOND;      // rows 2<->3
    else if (only({0, 6, 9, 15}, -1)) kind = S2_SWAP;            // rows 1<->2
    else return 2;                                                // needs a dense step
    const int j = next_slot(std::max(sb.last[rx], sb.last[ry]), 4 + pair_index(lo, hi));
    if (j < 0) return 0;
    sb.sd.kind[j] = kind;
    if (kind == S2_CPHASE) sb.sd.mat[j] = push_mats(p, &m[15], 1);
    if (kind == S2_CU_FIRST || kind == S2_CU_SECOND) sb.sd.mat[j] = push_mats(p, cu, 4);
    sb.cost += kind == S2_CPHASE ? 1.0 : (kind == S2_CU_FIRST || kind == S2_CU_SECOND) ? 4.0 : 0.5;
    sb.last[rx] = sb.last[ry] = j;
    sb.placed.push_back({2 * j, idx});
    return 1;
}

// Fold op `idx` (its qubits on register bits reg[0..k-1], in matrix-index
// order) into the dense matrix of the step: M <- U M, U acting on the rows'
// bits reg[] (rows with those bits zero taken 2^k at a time).
void fold_dense(Ctx &c, StepBuilder &sb, int idx, const int *reg)
{
    const PlanOp &op = c.ops[idx];
    int q[4];
    const int k = op_qubits(op, q);
    const int D = sb.dim(), K1 = 1 << k;
    int mask = 0;
    for (int j = 0; j < k; ++j) mask |= 1 << reg[j];
    Cd U[256];
    for (int i = 0; i < K1 * K1; ++i) U[i] = mat_at(c.mats, c.eb, (int64_t)op.mat_off + i);
    int rows[16];
    Cd tmp[16];
    for (int base = 0; base < D; ++base) {
        if (base & mask) continue;
        for (int ri = 0; ri < K1; ++ri) {
            int r = base;
            for (int j = 0; j < k; ++j) r |= ((ri >> j) & 1) << reg[j];
            rows[ri] = r;
        }
        for (int col = 0; col < D; ++col) {
            for (int ri = 0; ri < K1; ++ri) {
                double re = 0, im = 0;
                for (int ci = 0; ci < K1; ++ci) {
                    const Cd u = U[K1 * ri + ci], m = sb.M[D * rows[ci] + col];
                    re += u.re * m.re - u.im * m.im;
                    im += u.re * m.im + u.im * m.re;
                }
                tmp[ri] = {re, im};
            }
            for (int ri = 0; ri < K1; ++ri) sb.M[D * rows[ri] + col] = tmp[ri];
        }
    }
    sb.placed.push_back({(int)sb.placed.size(), idx});
}

// Place diagonal op `idx` as a free diagonal of the step.  cb[0..k-1] are its
// qubits' chunk-local bits, or -2 - m for a qubit outside the chunk (bit m of
// the block's chunk index).  At least one qubit must lie outside the
// registers; its chunk bits become locked out of the registers, provided
// enough unlocked bits remain to fill them.
bool place_free(Ctx &c, StepBuilder &sb, int idx, const int *cb, int k)
{
    int outside = 0;
    uint64_t lock = sb.lock, m = 0;
    for (int j = 0; j < k; ++j) {
        if (cb[j] < 0) {  // outside the chunk (FreeDiag encoding): fixed per block
            ++outside;
            continue;
        }
        m |= 1ull << cb[j];
        if (sb.reg_of(cb[j]) < 0) {
            ++outside;
            lock |= 1ull << cb[j];
        }
    }
    if (outside == 0) return false;
    int free_bits = 0;  // chunk bits neither registers nor locked
    for (int b = 0; b < c.pd->K; ++b) free_bits += !((lock >> b) & 1) && sb.reg_of(b) < 0;
    if (free_bits < sb.maxreg - sb.nreg) return false;
    const PlanOp &op = c.ops[idx];
    const int D = 1 << k;
    Cd d[4];
    for (int i = 0; i < D; ++i) d[i] = mat_at(c.mats, c.eb, (int64_t)op.mat_off + (D + 1) * i);
    sb.lock = lock;
    sb.endbits |= m;
    // complex64 phases (every entry of modulus 1) also as turns
    int32_t ro = -1;
    bool unit = c.eb == 8;
    for (int i = 0; i < D && unit; ++i) unit = std::fabs(d[i].re * d[i].re + d[i].im * d[i].im - 1.0) < 1e-12;
    if (unit) {
        ro = (int32_t)c.plan->raw.size();
        for (int i = 0; i < D; ++i) {
            const double t = std::atan2(d[i].im, d[i].re) / (2.0 * M_PI);  // (-1/2, 1/2]
            c.plan->raw.push_back((uint32_t)(int64_t)std::llround(t * 4294967296.0));
        }
    }
    sb.fd.push_back({cb[0], k > 1 ? cb[1] : -1, push_mats(*c.plan, d, D), ro});
    sb.placed.push_back({1000 + (int)sb.fd.size(), idx});  // after every slot
    sb.cost += 1.0;
    return true;
}

// A 1q or 2q op whose matrix is diagonal.
bool diagonal_op(const PlanOp &op, const void *mats, int eb)
{
    if (op.kind != OP_1Q && op.kind != OP_2Q) return false;
    const int D = op.kind == OP_1Q ? 2 : 4;
    for (int r = 0; r < D; ++r)
        for (int q = 0; q < D; ++q)
            if (r != q && !is0(mat_at(mats, eb, (int64_t)op.mat_off + D * r + q))) return false;
    return true;
}

// Vector arithmetic an op costs on its own, in real multiply-adds per
// amplitude: four per nonzero entry of an average matrix row.
double op_work(const Ctx &c, int idx)
{
    const PlanOp &op = c.ops[idx];
    int q[4];
    const int D = 1 << op_qubits(op, q);
    int nnz = 0;
    for (int i = 0; i < D * D; ++i) {
        const Cd m = mat_at(c.mats, c.eb, (int64_t)op.mat_off + i);
        nnz += m.re != 0.0 || m.im != 0.0;
    }
    return 4.0 * nnz / D;
}

// Turn a slot step into a dense one: multiply its placed ops in slot order.
void to_dense(Ctx &c, StepBuilder &sb)
{
    std::vector<std::pair<int, int>> seq = sb.placed;
    std::sort(seq.begin(), seq.end());
    const int D = sb.dim();
    for (int k = 0; k < D * D; ++k) sb.M[k] = {(k % (D + 1)) == 0 ? 1.0 : 0.0, 0.0};
    sb.placed.clear();
    sb.dense = true;
    sb.factor = {1.0, 0.0};
    for (auto &pr : seq) {
        const PlanOp &op = c.ops[pr.second];
        int q[4], reg[4];
        const int k = op_qubits(op, q);
        for (int j = 0; j < k; ++j) reg[j] = sb.reg_of(c.loc[q[j]]);
        fold_dense(c, sb, pr.second, reg);
    }
    for (int j = 0; j < kSlots; ++j) sb.sd.kind[j] = 0;
}

}  // namespace

bool plan_passes(int n, const PlanOp *ops, int nops, const void *mats, int elem_bytes, int K, int cmin,
                 int max_probes, Plan &out, std::string &err, int fuse, const int32_t *in_map,
                 const int32_t *out_map, bool generic, double bar, double bar5)
{
    out = Plan();
    out.generic = generic;
    // slot steps costlier than bar4 multiply-adds per amplitude become dense
    const double bar4 = (fuse == 2 || fuse == 4) ? (bar > 0.0 ? bar : kFuseBar) : 64.0;
    if (bar5 <= 0.0) bar5 = kFuseBar5;
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
        if (op.kind != OP_1Q && op.kind != OP_2Q && op.kind != OP_PROBE && op.kind != OP_3Q && op.kind != OP_4Q) {
            err = "tile plan: unknown op kind";
            return false;
        }
        int q[4];
        const int k = op_qubits(op, q);
        uint64_t seen = 0;
        for (int j = 0; j < k; ++j) {
            if (q[j] < 0 || q[j] >= n) {
                err = "tile plan: qubit out of range";
                return false;
            }
            if ((seen >> q[j]) & 1) {
                err = k == 2 ? "tile plan: q0 == q1" : "tile plan: repeated qubit";
                return false;
            }
            seen |= 1ull << q[j];
        }
        if (k > 2 && K < k) {
            err = "tile plan: chunk smaller than a multi-qubit op";
            return false;
        }
    }

    std::vector<int> rem(nops);
    for (int i = 0; i < nops; ++i) rem[i] = i;
    const uint64_t low = cmin ? ((1ull << cmin) - 1) : 0;
    bool any_probe = false;
    for (int i = 0; i < nops; ++i) any_probe |= ops[i].kind == OP_PROBE;
    // diagonal gates may become free diagonals (specialised kernels); a probe
    // is never moved past one on its qubits (it waits for the next step)
    std::vector<uint8_t> diag(nops, 0), zprobe(nops, 0);
    if (!generic && fuse != 1 && fuse != 3)
        for (int i = 0; i < nops; ++i) {
            diag[i] = diagonal_op(ops[i], mats, elem_bytes);
            // a probe of a diagonal generator commutes with free diagonals
            if (ops[i].kind == OP_PROBE)
                zprobe[i] = is0(mat_at(mats, elem_bytes, (int64_t)ops[i].mat_off + 1)) &&
                            is0(mat_at(mats, elem_bytes, (int64_t)ops[i].mat_off + 2));
        }
    out.fd_off.push_back(0);
    Cd pending = {1.0, 0.0};  // scalar factored out of finished steps, not yet applied

    if (in_map) {
        uint64_t seen = 0;
        for (int p = 0; p < n; ++p) {
            if (in_map[p] < 0 || in_map[p] >= n || ((seen >> in_map[p]) & 1)) {
                err = "tile plan: input map is not a permutation";
                return false;
            }
            seen |= 1ull << in_map[p];
        }
        out.in_map.assign(in_map, in_map + n);
    }
    uint64_t out_rows = 0;  // pass-layout positions that become the output's row bits
    if (out_map) {
        uint64_t seen = 0;
        for (int p = 0; p < n; ++p) {
            if (out_map[p] < 0 || out_map[p] >= n || ((seen >> out_map[p]) & 1)) {
                err = "tile plan: output map is not a permutation";
                return false;
            }
            seen |= 1ull << out_map[p];
            if (out_map[p] < std::max(cmin, 1)) out_rows |= 1ull << p;
        }
    }
    bool out_folded = false;
    bool first = true;  // a layout change with no op still gets its (stepless) pass
    while (!rem.empty() || (first && in_map)) {
        // ---- pass: chunk bit set ----
        uint64_t P = low, blocked = 0;
        // a folded layout change reserves the images of the low input bits;
        // that pass gets one more chunk bit for them (when the state has it)
        int Kp = K;
        if (first && in_map) {
            for (int k = 0; k < cmin; ++k) P |= 1ull << in_map[k];
            Kp = std::min(std::min(K + 1, n), kMaxKFold);
            if (Kp < K) Kp = K;
        }
        first = false;
        uint64_t far = 0;  // chunk bits this pass would gather from far-apart rows
        if (in_map && out.passes.empty())
            for (int p = far_bit(elem_bytes); p < n; ++p) far |= 1ull << in_map[p];
        const uint64_t fixed = any_probe ? 0 : search_chunk(ops, rem, P, Kp, low, K, far, out_folded ? 0 : out_rows, diag);
        if (fixed) P = fixed;
        int nprobe = 0;
        std::vector<int> chosen, left;
        for (int idx : rem) {
            const PlanOp &op = ops[idx];
            const uint64_t m = op_mask(op);
            const bool probe = op.kind == OP_PROBE;
            if (diag[idx]) {
                // a diagonal gate needs no chunk bits: a qubit outside the
                // chunk is fixed per block.  It keeps its place after the
                // ops left behind on its qubits.
                if (m & blocked) {
                    blocked |= m;
                    left.push_back(idx);
                } else {
                    chosen.push_back(idx);
                }
                continue;
            }
            if ((m & blocked) || popc(P | m) > Kp || (fixed && (m & ~fixed)) || (probe && nprobe >= max_probes)) {
                blocked |= m;
                left.push_back(idx);
                continue;
            }
            P |= m;
            nprobe += probe;
            chosen.push_back(idx);
        }
        if (chosen.empty() && !rem.empty()) {
            err = "tile plan: internal: a pass with no op";
            return false;
        }
        if (out_map && left.empty() && !out_folded && popc(P | out_rows) <= Kp) {
            // the last pass: it also writes the output order, in rows
            P |= out_rows;
            out_folded = true;
        }
        P = fill_low(P, n, Kp);
        PassDev pd;
        int loc[64], comp_index[64];
        make_pass(n, Kp, P, pd, loc);
        uint64_t pd_mask = 0;
        for (int k = 0; k < pd.K; ++k) pd_mask |= 1ull << pd.pos[k];
        for (int j = 0; j < pd.ncomp; ++j) comp_index[pd.comp[j]] = j;
        pd.step_off = (int32_t)out.steps.size();
        pd.probe_off = (int32_t)out.probe_ids.size();
        pd.nprobe = 0;
        Ctx ctx{ops, mats, elem_bytes, &out, &pd, max_probes, loc, !any_probe && !generic, generic, &diag};

        // ---- register steps ----
        // grow: place ops of srem into the step in order (an op whose bits
        // meet a skipped op's waits); the rest go to sleft.  work sums the
        // placed ops' own vector cost.
        auto grow = [&](StepBuilder &sb, const std::vector<int> &srem, std::vector<int> &sleft, double *work) {
            uint64_t sblk = 0;
            for (int idx : srem) {
                const PlanOp &op = ops[idx];
                int q[4], cb[4], reg[4];
                const int k = op_qubits(op, q);
                uint64_t m = 0;
                bool outside = false;  // a qubit outside the chunk (diagonal gates only)
                for (int j = 0; j < k; ++j) {
                    if ((pd_mask >> q[j]) & 1) {
                        cb[j] = loc[q[j]];
                        m |= 1ull << cb[j];
                    } else {
                        cb[j] = -2 - comp_index[q[j]];
                        outside = true;
                    }
                }
                // a gate that does not commute with diagonals waits for the
                // free diagonals on its bits
                const bool dg = diag[idx] != 0;
                if ((m & sblk) || (!dg && !zprobe[idx] && (m & sb.endbits)) || (outside && sb.dense)) {
                    sblk |= m;
                    sleft.push_back(idx);
                    continue;
                }
                if (outside) {
                    const size_t mats_before = out.mats.size();
                    if (place_free(ctx, sb, idx, cb, k)) {
                        if (work) *work += op_work(ctx, idx);
                    } else {
                        out.mats.resize(mats_before);
                        sblk |= m;
                        sleft.push_back(idx);
                    }
                    continue;
                }
                const StepCore save = sb;
                const size_t mats_before = out.mats.size(), ids_before = out.probe_ids.size();
                const int np_before = pd.nprobe;
                bool ok = true;
                for (int j = 0; j < k && ok; ++j) {
                    reg[j] = sb.reg_of(cb[j]);
                    if (reg[j] < 0) {
                        if (sb.nreg < sb.maxreg && !((sb.lock >> cb[j]) & 1)) {
                            reg[j] = sb.nreg;
                            sb.regbit[sb.nreg++] = cb[j];
                        } else {
                            ok = false;
                        }
                    }
                }
                if (ok) {
                    if (sb.dense) {
                        ok = op.kind != OP_PROBE;
                        if (ok) fold_dense(ctx, sb, idx, reg);
                    } else {
                        const int r = place(ctx, sb, idx, reg[0], k > 1 ? reg[1] : -1);
                        if (r == 2 && !sb.has_probe && sb.fd.empty()) {
                            to_dense(ctx, sb);
                            fold_dense(ctx, sb, idx, reg);
                        } else {
                            ok = r == 1;
                        }
                    }
                }
                if (!ok && dg && !sb.dense) {
                    // a diagonal gate that found no register room or slot
                    static_cast<StepCore &>(sb) = save;
                    out.mats.resize(mats_before);
                    ok = place_free(ctx, sb, idx, cb, k);
                }
                if (ok && work) *work += op_work(ctx, idx);
                if (!ok) {
                    static_cast<StepCore &>(sb) = save;
                    out.mats.resize(mats_before);
                    out.probe_ids.resize(ids_before);
                    pd.nprobe = np_before;
                    sblk |= m;
                    sleft.push_back(idx);
                }
            }
        };
        std::vector<int> srem = chosen;
        while (!srem.empty()) {
            StepBuilder sb;
            std::vector<int> sleft;
            bool five = false;
            if (fuse == 4 && !any_probe && !generic && elem_bytes == 8 && K >= kG5) {
                // one 32x32 step over five register bits when the vector work
                // of the ops it takes exceeds bar5
                StepBuilder b5;
                b5.maxreg = kG5;
                to_dense(ctx, b5);
                double work = 0.0;
                grow(b5, srem, sleft, &work);
                if (work > bar5) {
                    sb = b5;
                    five = true;
                } else {
                    sleft.clear();
                }
            }
            if (!five) {
                // fused plans: every step is one 16x16 (fuse 3: 32x32 over
                // five register bits) product of all its ops
                if (fuse == 3 && !any_probe && !generic && elem_bytes == 8 && K >= kG5) sb.maxreg = kG5;
                if ((fuse == 1 || fuse == 3) && !any_probe) to_dense(ctx, sb);
                grow(sb, srem, sleft, nullptr);
            }
            // a slot step whose arithmetic exceeds one dense product becomes dense
            if (!generic && !sb.dense && !sb.has_probe && sb.fd.empty() && sb.cost > bar4) to_dense(ctx, sb);
            // fill unused register bits with the lowest free chunk bits
            for (int b = 0; b < pd.K && sb.nreg < sb.maxreg; ++b)
                if (sb.reg_of(b) < 0 && !((sb.lock >> b) & 1)) sb.regbit[sb.nreg++] = b;
            for (int t = 0; t < kG; ++t) sb.sd.bits[t] = sb.regbit[t];
            if (sb.maxreg == kG5) sb.sd.bit4 = sb.regbit[kG];
            int sorted_bits[kG];
            std::memcpy(sorted_bits, sb.regbit, sizeof(sorted_bits));
            std::sort(sorted_bits, sorted_bits + kG);
            for (int t = 0; t < kG; ++t) sb.sd.sbits[t] = sorted_bits[t];
            if (sb.dense) {
                sb.sd.dense = push_mats(out, sb.M, sb.dim() * sb.dim());
                sb.sd.nrounds = 0;
            } else {
                int used = -1;
                for (int j = 0; j < kSlots; ++j)
                    if (sb.sd.kind[j]) used = j;
                sb.sd.nrounds = used / kSlotsPerRound + 1;
            }
            // keep every amplitude within 2^+-24 of its exact value
            pending = cmulh(pending, sb.factor);
            sb.sd.scale = -1;
            const double mag2 = pending.re * pending.re + pending.im * pending.im;
            if (mag2 > 0x1p48 || mag2 < 0x1p-48) {
                sb.sd.scale = push_mats(out, &pending, 1);
                pending = {1.0, 0.0};
            }
            if (sleft.size() == srem.size()) {
                err = "tile plan: internal: no op fits a register step";
                return false;
            }
            std::sort(sb.placed.begin(), sb.placed.end());
            for (auto &pr : sb.placed) out.order.push_back(pr.second);
            out.steps.push_back(sb.sd);
            out.fdiag.insert(out.fdiag.end(), sb.fd.begin(), sb.fd.end());
            out.fd_off.push_back((int32_t)out.fdiag.size());
            srem.swap(sleft);
        }
        pd.nsteps = (int32_t)out.steps.size() - pd.step_off;
        out.passes.push_back(pd);
        rem.swap(left);
    }
    if (out_folded) out.out_map.assign(out_map, out_map + n);
    if (out.fdiag.empty()) out.fd_off.clear();
    // the remaining factor goes to the last step
    if (!(pending.re == 1.0 && pending.im == 0.0) && !out.steps.empty()) {
        StepDev &last = out.steps.back();
        if (last.scale >= 0) {
            Cd v = {out.mats[2 * last.scale], out.mats[2 * last.scale + 1]};
            v = cmulh(v, pending);
            out.mats[2 * last.scale] = v.re;
            out.mats[2 * last.scale + 1] = v.im;
        } else {
            last.scale = push_mats(out, &pending, 1);
        }
    }
    // each pass's matrix entries, as one range of the arena
    for (PassDev &pd : out.passes) {
        int lo = INT32_MAX, hi = 0;
        auto cover = [&](int off, int count) {
            lo = std::min(lo, off);
            hi = std::max(hi, off + count);
        };
        for (int si = pd.step_off; si < pd.step_off + pd.nsteps; ++si) {
            const StepDev &sd = out.steps[si];
            if (sd.dense >= 0) cover(sd.dense, sd.bit4 >= 0 ? 1024 : 256);
            if (sd.scale >= 0) cover(sd.scale, 1);
            const FreeDiag *fd = nullptr;
            const int nfd = out.step_fdiag(si, &fd);
            for (int f = 0; f < nfd; ++f) cover(fd[f].mo, fd[f].b != -1 ? 4 : 2);
            for (int k = 0; k < kSlots; ++k)
                if (sd.kind[k] != S1_SKIP) cover(sd.mat[k], k % kSlotsPerRound < 4 ? 8 : 4);
        }
        pd.mat_lo = lo == INT32_MAX ? 0 : lo;
        pd.mat_hi = lo == INT32_MAX ? 0 : std::min(hi, (int)(out.mats.size() / 2));
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
