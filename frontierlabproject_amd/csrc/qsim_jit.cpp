// Run-time specialisation of chunk passes (see qsim_jit.h).
#include "qsim_jit.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <sstream>
#include <thread>
#include <utility>

namespace qsim {

namespace {

// qsim_dev_types.h + qsim_device.h as one string (written by build_native.py).
#include "qsim_jit_prelude.inc"

std::mutex g_mu;
std::map<std::pair<int, std::string>, hipFunction_t> g_fns;  // (device, source) -> function
std::vector<hipModule_t> g_modules;                           // never unloaded
std::atomic<long> g_compiled{0};

int swz_host(int e, bool dbl) { return dbl ? swz_d(e) : swz(e); }

uint64_t fnv1a(const std::string &s)
{
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

// Forward calls (one vector) never carry probes; the slot kernel keeps them.
bool skip_pass(const Plan &plan, size_t p, int nv) { return nv == 1 && plan.passes[p].nprobe > 0; }

// Pair slots of a round: register bits (first, second).
constexpr int kPairF[6] = {0, 2, 0, 1, 0, 1};
constexpr int kPairS[6] = {1, 3, 2, 3, 3, 2};

// The free diagonals of a step, after its slot rounds, with the step's
// factor: each thread's group bits outside the registers (gcb, the chunk
// index of its register index 0) and the block's chunk index c (bits outside
// the chunk) select every factor; a diagonal on a register bit t and an
// outside bit gives the pair (pt0, pt1) applied by that register bit's value.
// Entries exactly 1 are skipped: a controlled phase diag(1, 1, 1, d) is one
// conditional update (a scalar branch when its bits lie outside the chunk).
// complex64 phases (FreeDiag::ro) add up as 32-bit turns, exact modulo a
// turn, and become one complex factor per target (cis_turns); others multiply
// as complex factors.  The amplitudes then take the products from a table
// over the register bits used, skipping entries known to be 1.
void emit_free_diagonals(std::ostringstream &o, const Plan &plan, const StepDev &sd, const FreeDiag *fd, int nfd,
                         const char *T, int nv)
{
    auto reg = [&](int bit) {
        for (int t = 0; t < 4; ++t)
            if (bit >= 0 && sd.bits[t] == bit) return t;
        return -1;
    };
    // a chunk bit's value from the group; a bit outside the chunk (encoded
    // -2 - m) from bit m of the block's chunk index c
    auto bitv = [](int bit) {
        return bit >= 0 ? "((gcb >> " + std::to_string(bit) + ") & 1u)"
                        : "((uint32_t)(c >> " + std::to_string(-2 - bit) + ") & 1u)";
    };
    auto mat = [](int off) { return "mats[" + std::to_string(off) + "]"; };
    auto rw = [](int off) { return "raw[" + std::to_string(off) + "]"; };
    auto one = [&](int off) { return plan.mats[2 * off] == 1.0 && plan.mats[2 * off + 1] == 0.0; };
    const std::string C = std::string("cx<") + T + ">";
    // targets: 0 = ff, 1 + 2 t + s = p<t><s>; cplx: a complex factor, ang: turns
    std::string name[9] = {"ff"};
    for (int t = 0; t < 4; ++t)
        for (int b = 0; b < 2; ++b) name[1 + 2 * t + b] = "p" + std::to_string(t) + std::to_string(b);
    bool cplx[9] = {}, ang[9] = {};
    cplx[0] = sd.scale >= 0;
    // x <- x * (cond ? e1 : e0), entries at arena offsets o0, o1 (turns at r0, r1)
    auto need = [&](int x, int o0, int o1, int ro) {
        if (one(o0) && one(o1)) return;
        (ro >= 0 ? ang : cplx)[x] = true;
    };
    struct Up {
        int x, o0, o1, r0, r1;
        std::string cond;
    };
    std::vector<Up> ups;
    std::vector<std::string> quads;  // general 4-way factors of ff
    for (int f = 0; f < nfd; ++f) {
        const int mo = fd[f].mo, a = fd[f].a, b = fd[f].b, ro = fd[f].ro;
        const int ra = reg(a), rb = reg(b);
        auto R = [&](int i) { return ro >= 0 ? ro + i : -1; };
        if (b == -1) {  // 1q, its bit outside the registers
            need(0, mo, mo + 1, ro);
            ups.push_back({0, mo, mo + 1, R(0), R(1), bitv(a)});
        } else if (ra < 0 && rb < 0) {
            if (one(mo) && one(mo + 1) && one(mo + 2)) {
                need(0, mo, mo + 3, ro);
                ups.push_back({0, mo, mo + 3, R(0), R(3), bitv(a) + " & " + bitv(b)});
            } else {
                for (int e = 0; e < 4; ++e) (ro >= 0 ? ang : cplx)[0] |= !one(mo + e);
                std::ostringstream q;
                q << "      { const uint32_t cc = " << bitv(a) << " | (" << bitv(b) << " << 1);\n";
                if (ro >= 0)
                    q << "        aff += (cc & 2u) ? ((cc & 1u) ? " << rw(ro + 3) << " : " << rw(ro + 2) << ") : ((cc & 1u) ? "
                      << rw(ro + 1) << " : " << rw(ro) << "); }\n";
                else
                    q << "        ff = cmul(ff, (cc & 2u) ? ((cc & 1u) ? " << mat(mo + 3) << " : " << mat(mo + 2)
                      << ") : ((cc & 1u) ? " << mat(mo + 1) << " : " << mat(mo) << ")); }\n";
                quads.push_back(q.str());
            }
        } else if (ra >= 0) {  // matrix index m_a + 2 g_b
            need(1 + 2 * ra, mo, mo + 2, ro);
            need(2 + 2 * ra, mo + 1, mo + 3, ro);
            ups.push_back({1 + 2 * ra, mo, mo + 2, R(0), R(2), bitv(b)});
            ups.push_back({2 + 2 * ra, mo + 1, mo + 3, R(1), R(3), bitv(b)});
        } else {  // matrix index g_a + 2 m_b
            need(1 + 2 * rb, mo, mo + 1, ro);
            need(2 + 2 * rb, mo + 2, mo + 3, ro);
            ups.push_back({1 + 2 * rb, mo, mo + 1, R(0), R(1), bitv(a)});
            ups.push_back({2 + 2 * rb, mo + 2, mo + 3, R(2), R(3), bitv(a)});
        }
    }
    o << "    { const uint32_t gcb = ins0(ins0(ins0(ins0(g32, " << sd.sbits[0] << "), " << sd.sbits[1] << "), "
      << sd.sbits[2] << "), " << sd.sbits[3] << ");\n";
    o << "      (void)gcb;\n";
    bool anyang = false;
    for (int x = 0; x < 9; ++x) anyang |= ang[x];
    if (anyang)
        o << "      const uint32_t *raw = reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(mats) + "
          << plan.raw_offset(8) << ");\n";
    for (int x = 0; x < 9; ++x) {
        if (cplx[x]) o << "      " << C << " " << name[x] << " = " << (x == 0 && sd.scale >= 0 ? mat(sd.scale) : C + "{1, 0}") << ";\n";
        if (ang[x]) o << "      uint32_t a" << name[x] << " = 0u;\n";
    }
    for (const Up &u : ups) {
        if (one(u.o0) && one(u.o1)) continue;
        const std::string &x = name[u.x];
        if (u.r0 >= 0) {
            if (one(u.o0))
                o << "      if (" << u.cond << ") a" << x << " += " << rw(u.r1) << ";\n";
            else
                o << "      a" << x << " += (" << u.cond << ") ? " << rw(u.r1) << " : " << rw(u.r0) << ";\n";
        } else if (one(u.o0)) {
            o << "      if (" << u.cond << ") " << x << " = cmul(" << x << ", " << mat(u.o1) << ");\n";
        } else {
            o << "      " << x << " = cmul(" << x << ", (" << u.cond << ") ? " << mat(u.o1) << " : " << mat(u.o0) << ");\n";
        }
    }
    for (const std::string &q : quads) o << q;
    // one complex value per target
    for (int x = 0; x < 9; ++x) {
        if (!ang[x]) continue;
        if (cplx[x])
            o << "      " << name[x] << " = cmul(" << name[x] << ", cis_turns(a" << name[x] << "));\n";
        else
            o << "      const " << C << " " << name[x] << " = cis_turns(a" << name[x] << ");\n";
    }
    const bool ffu = cplx[0] || ang[0];
    bool u0[4], u1[4];
    for (int t = 0; t < 4; ++t) u0[t] = cplx[1 + 2 * t] || ang[1 + 2 * t], u1[t] = cplx[2 + 2 * t] || ang[2 + 2 * t];
    int U[4], nu = 0;
    for (int t = 0; t < 4; ++t)
        if (u0[t] || u1[t]) U[nu++] = t;
    // tb[e]: the factor of the amplitudes whose used register bits are e;
    // known[e]: identically 1
    std::vector<char> known(1 << nu, 0);
    known[0] = !ffu;
    o << "      " << C << " tb[" << (1 << nu) << "];\n      tb[0] = " << (ffu ? "ff" : C + "{1, 0}") << ";\n";
    for (int i = 0; i < nu; ++i) {
        const int t = U[i];
        for (int e = 0; e < (1 << i); ++e) {
            const int h = e + (1 << i);
            known[h] = known[e] && !u1[t];
            o << "      tb[" << h << "] = " << (u1[t] ? "cmul(tb[" + std::to_string(e) + "], p" + std::to_string(t) + "1)"
                                                      : "tb[" + std::to_string(e) + "]") << ";\n";
            if (u0[t]) o << "      tb[" << e << "] = cmul(tb[" << e << "], p" << t << "0);\n";
            known[e] = known[e] && !u0[t];
        }
    }
    const char *vec[2] = {"a", "l"};
    for (int v = 0; v < nv; ++v)
        for (int m = 0; m < 16; ++m) {
            int e = 0;
            for (int i = 0; i < nu; ++i) e |= ((m >> U[i]) & 1) << i;
            if (known[e]) continue;
            o << "      " << vec[v] << "[" << m << "] = cmul(" << vec[v] << "[" << m << "], tb[" << e << "]);\n";
        }
    o << "    }\n";
}

void emit_step(std::ostringstream &o, const StepDev &sd, int E, const char *T, int nv, int tpb,
               const Plan *plan = nullptr, const FreeDiag *fd = nullptr, int nfd = 0)
{
    const int NG = E >> kG;
    const bool dbl = std::string(T) == "double";
    int moff[16];
    for (int m = 0; m < 16; ++m)
        moff[m] = swz_host(((m & 1) << sd.bits[0]) | (((m >> 1) & 1) << sd.bits[1]) |
                               (((m >> 2) & 1) << sd.bits[2]) | (((m >> 3) & 1) << sd.bits[3]),
                           dbl);
    const bool loop = NG > tpb, guard = NG < tpb;
    if (sd.bit4 >= 0) {  // 32x32 on the matrix cores (complex64 one-vector passes only)
        int bits5[5] = {sd.bits[0], sd.bits[1], sd.bits[2], sd.bits[3], sd.bit4};
        int sorted5[5];
        std::copy(bits5, bits5 + 5, sorted5);
        std::sort(sorted5, sorted5 + 5);
        o << "  { const int moff[32] = {";
        for (int m = 0; m < 32; ++m) {
            int off = 0;
            for (int t = 0; t < 5; ++t) off |= ((m >> t) & 1) << bits5[t];
            o << swz_host(off, false) << (m < 31 ? ", " : "");
        }
        o << "};\n    dense32_mfma<" << tpb << ">(c0, mats + " << sd.dense << ", moff";
        for (int t = 0; t < 5; ++t) o << ", " << sorted5[t];
        o << ", " << (E >> 5) << ", "
          << (sd.scale >= 0 ? "mats + " + std::to_string(sd.scale) : std::string("(const cx<float> *)nullptr"))
          << ");\n  }\n";
        o << "  __syncthreads();\n";
        return;
    }
    if (sd.dense >= 0 && nv == 1 && std::string(T) == "float") {  // matrix cores
        o << "  { const int moff[16] = {";
        for (int m = 0; m < 16; ++m) o << moff[m] << (m < 15 ? ", " : "");
        o << "};\n    dense16_mfma<" << tpb << ">(c0, mats + " << sd.dense << ", moff, " << sd.sbits[0] << ", " << sd.sbits[1]
          << ", " << sd.sbits[2] << ", " << sd.sbits[3] << ", " << NG << ", "
          << (sd.scale >= 0 ? "mats + " + std::to_string(sd.scale) : std::string("(const cx<float> *)nullptr"))
          << ");\n  }\n";
        o << "  __syncthreads();\n";
        return;
    }
    o << "  {\n";
    if (loop) o << "  for (int gb = 0; gb < " << NG << "; gb += " << tpb << ") {\n";
    o << "    const int gi = " << (loop ? "gb + " : "") << "(int)threadIdx.x;\n";
    if (guard)
        o << "    const bool act = gi < " << NG << ";\n    const uint32_t g32 = (uint32_t)(act ? gi : " << NG - 1
          << ");\n";
    else
        o << "    const uint32_t g32 = (uint32_t)gi;\n";
    // complex64: with chunk bit 0 among the register bits the base slot is
    // even (swz keeps bit 0); saying so lets the compiler pair slots sb, sb ^ 1
    // into 16-byte LDS accesses.  (The complex128 swizzle feeds bit 0.)
    bool bit0 = false;
    for (int q = 0; q < 4; ++q) bit0 |= sd.bits[q] == 0 && !dbl;
    // LDS addressing in bytes (slot * amplitude size): the xor with each
    // register offset then needs no shift
    const int esz = std::string(T) == "float" ? 8 : 16;
    o << "    const int sb = (swzr<" << T << ">((int)ins0(ins0(ins0(ins0(g32, " << sd.sbits[0] << "), " << sd.sbits[1] << "), "
      << sd.sbits[2] << "), " << sd.sbits[3] << "))" << (bit0 ? " & ~1" : "") << ") * " << esz << ";\n";
    for (int m = 0; m < 16; ++m) moff[m] *= esz;
    // vectors: a (psi, in c0) and, with nv == 2, l (lam, in c1)
    const char *vec[2] = {"a", "l"}, *lds[2] = {"c0", "c1"};
    for (int v = 0; v < nv; ++v) {
        o << "    cx<" << T << "> " << vec[v] << "[16];\n";
        for (int m = 0; m < 16; ++m)
            o << "    " << vec[v] << "[" << m << "] = lds_at(" << lds[v] << ", sb ^ " << moff[m] << ");\n";
    }
    const char *store_guard = guard ? "if (act) " : "";
    if (sd.dense >= 0) {
        if (sd.scale >= 0) o << "    const cx<" << T << "> f = mats[" << sd.scale << "];\n";
        for (int v = 0; v < nv; ++v)
            for (int r = 0; r < 16; ++r) {
                const int M = sd.dense + 16 * r;
                o << "    { cx<" << T << "> acc = cmul(mats[" << M << "], " << vec[v] << "[0]);\n";
                o << "      #pragma unroll\n      for (int c = 1; c < 16; ++c) acc = cmad(mats[" << M
                  << " + c], " << vec[v] << "[c], acc);\n";
                if (sd.scale >= 0) o << "      acc = cmul(acc, f);\n";
                o << "      " << store_guard << "lds_at(" << lds[v] << ", sb ^ " << moff[r] << ") = acc; }\n";
            }
    } else {
        for (int r = 0; r < sd.nrounds; ++r) {
            const int b = r * kSlotsPerRound;
            for (int j = 0; j < 4; ++j) {
                const int k = sd.kind[b + j];
                const int mo = sd.mat[b + j];
                if (k == S1_SKIP) continue;
                if (k == S1_PROBE || k == S1_PROBE_GEN) {
                    // Im<lam|G|psi> over this step's pairs, one partial per wave
                    o << "    { const cx<" << T << "> g[4] = {mats[" << mo << "], mats[" << mo + 1 << "], mats["
                      << mo + 2 << "], mats[" << mo + 3 << "]};\n";
                    o << "      const double pv = wave_sum_rows(" << (guard ? "act ? " : "") << "probe1<" << T << ", "
                      << j << ">(a, l, g)" << (guard ? " : 0.0" : "") << ");\n";
                    o << "      if ((threadIdx.x & 63) == 0) pacc[" << sd.probe[4 * r + j]
                      << " * kWaves + (threadIdx.x >> 6)] += pv; }\n";
                    if (k == S1_PROBE_GEN) {
                        o << "    { const cx<" << T << "> u[4] = {mats[" << mo + 4 << "], mats[" << mo + 5
                          << "], mats[" << mo + 6 << "], mats[" << mo + 7 << "]};\n";
                        o << "      reg1<" << T << ", " << j << ">(a, u); reg1<" << T << ", " << j << ">(l, u); }\n";
                    }
                    continue;
                }
                for (int v = 0; v < nv; ++v)
                    o << "    apply1<" << T << ", " << j << ">(" << vec[v] << ", " << k << ", mats + " << mo << ");\n";
            }
            for (int j = 0; j < 6; ++j) {
                const int k = sd.kind[b + 4 + j];
                if (k == S2_SKIP) continue;
                for (int v = 0; v < nv; ++v)
                    o << "    apply2<" << T << ", " << kPairF[j] << ", " << kPairS[j] << ">(" << vec[v] << ", " << k
                      << ", mats + " << sd.mat[b + 4 + j] << ");\n";
            }
        }
        if (nfd > 0) {
            emit_free_diagonals(o, *plan, sd, fd, nfd, T, nv);
        } else if (sd.scale >= 0) {
            o << "    { const cx<" << T << "> f = mats[" << sd.scale << "];\n";
            for (int v = 0; v < nv; ++v)
                o << "      #pragma unroll\n      for (int m = 0; m < 16; ++m) " << vec[v] << "[m] = cmul(" << vec[v]
                  << "[m], f);\n";
            o << "    }\n";
        }
        for (int v = 0; v < nv; ++v)
            for (int m = 0; m < 16; ++m)
                o << "    " << store_guard << "lds_at(" << lds[v] << ", sb ^ " << moff[m] << ") = " << vec[v] << "[" << m
                  << "];\n";
    }
    if (loop) o << "  }\n";
    o << "  }\n  __syncthreads();\n";
}

void emit_plan_step(std::ostringstream &o, const Plan &plan, int s, int E, const char *T, int nv, int tpb)
{
    const FreeDiag *fd = nullptr;
    const int nfd = plan.step_fdiag(s, &fd);
    emit_step(o, plan.steps[s], E, T, nv, tpb, &plan, fd, nfd);
}

std::string kernel_name(const std::string &body) { return "qsim_jit_" + std::to_string(fnv1a(body)); }

std::string arch_flag(const std::string &arch) { return "--offload-arch=" + (arch.empty() ? "gfx950" : arch); }

// Compile one translation unit; returns the code object or sets err.
bool compile_unit(const std::string &src, const std::string &arch, std::vector<char> &code, std::string &err)
{
    hiprtcProgram prog = nullptr;
    hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "qsim_jit.hip", 0, nullptr, nullptr);
    if (r != HIPRTC_SUCCESS) {
        err = std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r);
        return false;
    }
    const std::string af = arch_flag(arch);
    std::vector<std::string> extra;  // QSIM_JIT_FLAGS: further compiler flags, space-separated
    if (const char *e = std::getenv("QSIM_JIT_FLAGS")) {
        std::string f = e, tok;
        for (char ch : f + " ") {
            if (ch == ' ') {
                if (!tok.empty()) extra.push_back(tok);
                tok.clear();
            } else {
                tok += ch;
            }
        }
    }
    std::vector<const char *> opts = {af.c_str(), "-O3", "-std=c++17", "-ffp-contract=fast",
                                      "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    for (auto &x : extra) opts.push_back(x.c_str());
    r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        err = std::string("hiprtcCompileProgram: ") + hiprtcGetErrorString(r) + "\n" + log.substr(0, 4000);
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    code.resize(n);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return true;
}

// Compile `bodies` as up to `threads` units in parallel.
bool compile_all(const std::vector<std::string> &bodies, const std::string &arch,
                 std::vector<std::vector<char>> &codes, std::vector<std::vector<int>> &members, std::string &err)
{
    int hw = (int)std::thread::hardware_concurrency();
    if (const char *e = std::getenv("QSIM_JIT_THREADS")) hw = std::atoi(e);
    const int nu = std::max(1, std::min({(int)bodies.size(), hw > 0 ? hw : 4, 16}));
    members.assign(nu, {});
    for (size_t i = 0; i < bodies.size(); ++i) members[i % nu].push_back((int)i);
    std::vector<std::string> srcs(nu);
    for (int u = 0; u < nu; ++u) {
        std::string s = kPrelude;
        s += "\nnamespace qsim {\n";
        for (int i : members[u]) s += bodies[i];
        s += "}  // namespace qsim\n";
        srcs[u] = std::move(s);
    }
    if (const char *dir = std::getenv("QSIM_JIT_DUMP")) {  // sources for inspection
        for (int u = 0; u < nu; ++u) {
            const std::string path = std::string(dir) + "/qsim_jit_" + std::to_string(fnv1a(srcs[u])) + ".hip";
            if (FILE *f = std::fopen(path.c_str(), "w")) {
                std::fwrite(srcs[u].data(), 1, srcs[u].size(), f);
                std::fclose(f);
            }
        }
    }
    codes.assign(nu, {});
    std::vector<std::string> errs(nu);
    std::vector<char> ok(nu, 0);
    std::vector<std::thread> th;
    for (int u = 0; u < nu; ++u)
        th.emplace_back([&, u] { ok[u] = compile_unit(srcs[u], arch, codes[u], errs[u]); });
    for (auto &t : th) t.join();
    for (int u = 0; u < nu; ++u)
        if (!ok[u]) {
            err = errs[u];
            return false;
        }
    g_compiled += (long)bodies.size();
    return true;
}

}  // namespace

namespace {

// Field moves "| (((x >> from) & mask) << to)" for a bit map src k -> dst[k],
// grouping runs of consecutive sources with consecutive destinations.
std::string bit_moves(const char *x, const int *dst, int nbits, const char *ty)
{
    std::ostringstream o;
    for (int k = 0; k < nbits;) {
        int len = 1;
        while (k + len < nbits && dst[k + len] == dst[k] + len) ++len;
        const unsigned long long mask = (1ull << len) - 1;
        o << " | ((((" << ty << ")" << x << " >> " << k << ") & " << mask << "ull) << " << dst[k] << ")";
        k += len;
    }
    return o.str();
}

// Per-lane constants of a pass kernel, computed here and compiled into the
// kernel as a table (one 64-entry row per item, read once per thread at the
// start): the lane part of each copy's global offset and LDS slot.  Offsets
// and slots are linear over GF(2) and the lane and wave parts of the thread
// index have disjoint bits, so the wave part (uniform) is added on the scalar
// unit.  (Per-step lane parts read the same way were measured slower than
// computing each step's base per thread: their loads land before the steps.)
struct LaneTable {
    std::vector<int32_t> v;  // rows of 64
    int add_row(const std::vector<int32_t> &r)
    {
        v.insert(v.end(), r.begin(), r.end());
        return (int)(v.size() / 64) - 1;
    }
};

uint64_t deposit(uint64_t x, const int *to, int nbits)
{
    uint64_t r = 0;
    for (int k = 0; k < nbits; ++k)
        if ((x >> k) & 1) r |= 1ull << to[k];
    return r;
}

// Specialised chunk copies.  The read index r of a chunk element splits into
// a lane part (bits below log2(per * 64); per = 2 complex64 amplitudes per
// 16-byte access, 1 complex128), a wave part (bits below log2(per * TPB))
// and a batch part j * per * TPB; each access is then the scalar address
// (base + C_j + wave part) plus the lane's 32-bit byte offset, into LDS slot
// (lane slot ^ wave slot ^ S_j), with C_j and S_j compile-time constants.
// src[k]: state position of read bit k; chunk[k]: chunk bit of read bit k
// (the identity unless a layout change is folded in).
struct CopyPlan {
    bool ok = false;
    int per = 1, tb = 0, lb = 0, nj = 0;
    int row = -1;  // lane table rows: offsets (bytes) at row, slots at row + 1
};

CopyPlan copy_plan(LaneTable &lt, int K, bool dbl, int tpb, const int *src, const int *chunk, int cf)
{
    CopyPlan c;
    c.per = dbl ? 1 : 2;
    int lt2 = 0;
    while ((1 << lt2) < tpb) ++lt2;
    c.tb = lt2 + (c.per == 2 ? 1 : 0);
    c.lb = 6 + (c.per == 2 ? 1 : 0);
    if (K < c.tb || (!dbl && cf < 1)) return c;  // guards, or unpaired complex64 reads
    int hi = 0;
    for (int k = 0; k < c.lb; ++k) hi = std::max(hi, src[k]);
    if (hi + 5 >= 31) return c;  // lane offsets must fit in 32 bits
    c.nj = 1 << (K - c.tb);
    std::vector<int32_t> off(64), slot(64);
    const int esz = dbl ? 16 : 8;
    for (uint32_t l = 0; l < 64; ++l) {
        const uint64_t r = (uint64_t)l * c.per;
        off[l] = (int32_t)(deposit(r, src, c.lb) * esz);
        slot[l] = swz_host((int)deposit(r, chunk, c.lb), dbl) * esz;  // LDS bytes
    }
    c.row = lt.add_row(off);
    lt.add_row(slot);
    c.ok = true;
    return c;
}

// Declares the thread's lane offset Tl<tag>, wave offset Tw<tag> (scalar,
// amplitudes) and slot st<tag>.
void emit_copy_thread(std::ostringstream &o, const CopyPlan &c, const char *tag, const int *src, const int *chunk)
{
    const std::string t(tag);
    o << "  const uint32_t Tl" << t << " = (uint32_t)qtab[" << c.row * 64 << " + lane];\n";
    o << "  const uint32_t rw" << t << " = wv << " << c.lb << ";\n";
    o << "  const uint64_t Tw" << t << " = 0ull" << bit_moves(("rw" + t).c_str(), src, c.tb, "uint64_t") << ";\n";
    o << "  const int st" << t << " = qtab[" << (c.row + 1) * 64 << " + lane] ^ (" << (c.per == 1 ? "swz_d" : "swz")
      << "(0"
      << bit_moves(("rw" + t).c_str(), chunk, c.tb, "int") << ") * " << (c.per == 2 ? 8 : 16) << ");\n";
}

// Copy between global vector `g` and LDS `lds` over the K read bits: loads
// (global -> LDS, `chunk` may reorder the bits: a folded layout change) or
// stores (LDS -> global, the pass layout).  Accesses are issued in batches of
// eight per thread, all loads of a batch before any of its LDS writes.
void emit_copy(std::ostringstream &o, const CopyPlan &c, const char *tag, const char *g, const char *lds,
               const char *base, const int *src, const int *chunk, int K, bool load, bool dbl, bool zeros = false)
{
    const int B = 8;
    const std::string t(tag);
    const char *T = dbl ? "double" : "float";
    bool ident = true;
    for (int k = 0; k < K; ++k) ident &= chunk[k] == k;
    const int esz = dbl ? 16 : 8;
    const int s1 = swz_host(1 << chunk[0], dbl) * esz;  // complex64: a pair's second amplitude (LDS bytes, xor)
    for (int j0 = 0; j0 < c.nj; j0 += B) {
        const int nb = std::min(B, c.nj - j0);
        o << "  {\n    v4f b[" << nb << "];\n";
        for (int pass = 0; pass < 2; ++pass)
            for (int q = 0; q < nb; ++q) {
                const std::string bq = "b[" + std::to_string(q) + "]";
                const uint64_t rj = (uint64_t)(j0 + q) << c.tb;
                const uint64_t Cj = deposit(rj, src, K);
                const int Sj = swz_host((int)deposit(rj, chunk, K), dbl) * esz;
                std::ostringstream a;
                a << "reinterpret_cast<v4f *>(reinterpret_cast<char *>(const_cast<cx<" << T << "> *>(" << g
                  << ") + uniform_u64(" << base << " + Tw" << t << " + " << Cj << "ull)) + Tl" << t << ")";
                const std::string slot = "st" + t + " ^ " + std::to_string(Sj);
                const std::string la = "reinterpret_cast<v4f *>(reinterpret_cast<char *>(" + std::string(lds) + ") + (" +
                                       slot + "))";
                if (load && pass == 0) {
                    o << "    " << bq << " = __builtin_nontemporal_load(" << a.str() << ");\n";
                } else if (load) {
                    if (dbl || ident)
                        o << "    *" << la << " = " << bq << ";\n";
                    else
                        o << "    { const int sl = " << slot << "; lds_at(" << lds << ", sl) = cx<float>{" << bq
                          << ".x, " << bq << ".y}; lds_at(" << lds << ", sl ^ " << s1 << ") = cx<float>{" << bq
                          << ".z, " << bq << ".w}; }\n";
                } else if (pass == 0) {
                    if (zeros)
                        o << "    " << bq << " = v4f{0.f, 0.f, 0.f, 0.f};\n";
                    else if (dbl || ident)
                        o << "    " << bq << " = *" << la << ";\n";
                    else
                        o << "    { const int sl = " << slot << "; const cx<float> x0 = lds_at(" << lds
                          << ", sl), x1 = lds_at(" << lds << ", sl ^ " << s1 << "); " << bq
                          << " = v4f{x0.re, x0.im, x1.re, x1.im}; }\n";
                } else {
                    o << "    __builtin_nontemporal_store(" << bq << ", " << a.str() << ");\n";
                }
            }
        o << "  }\n";
    }
}

// The lane table as a device array (named after the kernel) and the per-thread
// lane / wave indices.
std::string lane_table_source(const LaneTable &lt, const std::string &name)
{
    std::ostringstream o;
    o << "static __device__ const int32_t " << name << "_t[" << std::max<size_t>(1, lt.v.size()) << "] = {";
    for (size_t i = 0; i < lt.v.size(); ++i) o << (i ? "," : "") << lt.v[i];
    if (lt.v.empty()) o << "0";
    o << "};\n";
    return o.str();
}

void emit_lane_prologue(std::ostringstream &o, const std::string &name)
{
    o << "  const int32_t *__restrict__ qtab = " << name << "_t;\n";
    o << "  const int lane = (int)(threadIdx.x & 63);\n";
    o << "  const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)threadIdx.x) >> 6;\n";
}

// First pass of a plan with a folded layout change: reads `psi` in the
// source layout, writes `out` in the pass layout (complex64, one vector).
// Chunk bits of pass `pd` sorted by their position under `map` (state bit
// position -> position in the other layout): positions ascending (pos) and
// the chunk bit of each (chunk); comp[m]: position of block-index bit m.
void remap_bits(const PassDev &pd, const int *map, int *pos, int *chunk, int *comp)
{
    std::vector<std::pair<int, int>> v;
    for (int j = 0; j < pd.K; ++j) v.push_back({map[pd.pos[j]], j});
    std::sort(v.begin(), v.end());
    for (int k = 0; k < pd.K; ++k) pos[k] = v[k].first, chunk[k] = v[k].second;
    for (int m = 0; m < pd.ncomp; ++m) comp[m] = map[pd.comp[m]];
}

// A pass that reads its input in another layout (a folded layout change,
// Plan::in_map, first pass) and/or writes its output in another layout (the
// output order, Plan::out_map, last pass), from `psi` into `out`, through the
// specialised copies on both sides.  Empty when they do not apply.
std::string remapped_pass_source(const Plan &plan, size_t p, bool dbl, const std::string &name)
{
    const PassDev &pd = plan.passes[p];
    const char *T = dbl ? "double" : "float";
    const int n = pd.K + pd.ncomp;
    const int K = pd.K, E = 1 << K;
    const int TPB = jit_block_threads(K, 1);
    int ident[64], inv[64], Pin[64], Jin[64], Cin[64], Pout[64], Jout[64], Cout[64];
    for (int k = 0; k < 64; ++k) ident[k] = inv[k] = k;
    if (p == 0 && !plan.in_map.empty())
        for (int q = 0; q < n; ++q) inv[plan.in_map[q]] = q;
    remap_bits(pd, inv, Pin, Jin, Cin);
    const bool wmap = p + 1 == plan.passes.size() && !plan.out_map.empty();
    remap_bits(pd, wmap ? plan.out_map.data() : ident, Pout, Jout, Cout);
    int cfi = 0, cfo = 0;
    while (cfi < K && Pin[cfi] == cfi) ++cfi;
    while (cfo < K && Pout[cfo] == cfo) ++cfo;
    LaneTable lt;
    const CopyPlan ci = copy_plan(lt, K, dbl, TPB, Pin, Jin, cfi), co = copy_plan(lt, K, dbl, TPB, Pout, Jout, cfo);
    if (!ci.ok || !co.ok) return std::string();
    std::ostringstream o, b;
    o << "extern \"C\" __global__ __launch_bounds__(" << TPB << ")"
      << (dbl ? "" : " __attribute__((amdgpu_waves_per_eu(4)))") << "\nvoid " << name << "(const cx<" << T
      << "> *__restrict__ psi, cx<" << T << "> *__restrict__ out, const cx<" << T << "> *__restrict__ mats)\n{\n";
    o << "  __shared__ __align__(16) cx<" << T << "> c0[" << E << "];\n";
    emit_lane_prologue(b, name);
    emit_copy_thread(b, ci, "i", Pin, Jin);
    emit_copy_thread(b, co, "o", Pout, Jout);
    b << "  const uint64_t c = xcd_slab(blockIdx.x, " << pd.ncomp << ");\n";
    b << "  const uint64_t base_in = 0ull" << bit_moves("c", Cin, pd.ncomp, "uint64_t") << ";\n";
    b << "  const uint64_t base_out = 0ull" << bit_moves("c", Cout, pd.ncomp, "uint64_t") << ";\n";
    emit_copy(b, ci, "i", "psi", "c0", "base_in", Pin, Jin, K, true, dbl);
    b << "  __syncthreads();\n";
    for (int s = 0; s < pd.nsteps; ++s) emit_plan_step(b, plan, pd.step_off + s, E, T, 1, TPB);
    emit_copy(b, co, "o", "out", "c0", "base_out", Pout, Jout, K, false, dbl);
    return lane_table_source(lt, name) + o.str() + b.str() + "}\n\n";
}

std::string mapped_pass_source(const Plan &plan, bool dbl, const std::string &name)
{
    const std::string fast = remapped_pass_source(plan, 0, dbl, name);
    if (!fast.empty()) return fast;
    const PassDev &pd = plan.passes[0];
    const char *T = dbl ? "double" : "float";
    const int n = (int)plan.in_map.size();
    const int K = pd.K, CF = pd.cfast, E = 1 << K, NROW = 1 << (K - CF);
    const int TPB = jit_block_threads(K, 1);
    int inv[64];
    for (int q = 0; q < n; ++q) inv[plan.in_map[q]] = q;
    // chunk bits in source order: source position, chunk-local index
    std::vector<std::pair<int, int>> src;
    for (int j = 0; j < K; ++j) src.push_back({inv[pd.pos[j]], j});
    std::sort(src.begin(), src.end());
    int J[64], Pold[64];
    for (int k = 0; k < K; ++k) Pold[k] = src[k].first, J[k] = src[k].second;
    int CFR = 0;
    while (CFR < K && Pold[CFR] == CFR) ++CFR;
    const int NROWIN = 1 << (K - CFR);
    int comp_old[64];
    for (int m = 0; m < pd.ncomp; ++m) comp_old[m] = inv[pd.comp[m]];
    std::ostringstream o;
    o << "extern \"C\" __global__ __launch_bounds__(" << TPB << ")"
      << (dbl ? "" : " __attribute__((amdgpu_waves_per_eu(4)))") << "\nvoid " << name << "(const cx<" << T
      << "> *__restrict__ psi, cx<" << T << "> *__restrict__ out, const cx<" << T << "> *__restrict__ mats)\n{\n";
    o << "  __shared__ __align__(16) cx<" << T << "> c0[" << E << "];\n";
    o << "  __shared__ uint64_t rowoff[" << NROW << "];\n  __shared__ uint64_t rowin[" << NROWIN << "];\n";
    o << "  for (int t = threadIdx.x; t < " << NROW << "; t += " << TPB << ")\n    rowoff[t] = 0ull"
      << bit_moves("t", pd.pos + CF, K - CF, "uint64_t") << ";\n";
    o << "  for (int t = threadIdx.x; t < " << NROWIN << "; t += " << TPB << ")\n    rowin[t] = 0ull"
      << bit_moves("t", Pold + CFR, K - CFR, "uint64_t") << ";\n";
    o << "  const uint64_t c = xcd_slab(blockIdx.x, " << pd.ncomp << ");\n";
    o << "  const uint64_t base = 0ull" << bit_moves("c", pd.comp, pd.ncomp, "uint64_t") << ";\n";
    o << "  const uint64_t base_in = 0ull" << bit_moves("c", comp_old, pd.ncomp, "uint64_t") << ";\n";
    o << "  auto sigma = [](int r) -> int { return 0" << bit_moves("r", J, K, "unsigned") << "; };\n";
    o << "  __syncthreads();\n";
    if (dbl)
        o << "  chunk_copy_in_mapped_d<" << TPB << ">(psi, c0, base_in, rowin, " << E << ", " << CFR << ", sigma);\n";
    else
        o << "  chunk_copy_in_mapped<" << TPB << ">(psi, c0, base_in, rowin, " << E << ", " << CFR << ", sigma, "
          << J[0] << ");\n";
    o << "  __syncthreads();\n";
    for (int s = 0; s < pd.nsteps; ++s) emit_plan_step(o, plan, pd.step_off + s, E, T, 1, TPB);
    o << "  chunk_copy<" << T << ", false, true, " << TPB << ">(out, c0, base, rowoff, " << E << ", " << CF
      << ");\n}\n\n";
    return o.str();
}

}  // namespace

bool jit_mapped_last(const Plan &plan, bool dbl, int nv)
{
    if (plan.out_map.empty() || nv != 1 || plan.passes.empty()) return false;
    const size_t last = plan.passes.size() - 1;
    if (last == 0 && !plan.in_map.empty() && !jit_mapped_first(plan, dbl, nv)) return false;
    if (last == 0 && plan.zero_input) return false;
    return !remapped_pass_source(plan, last, dbl, "@").empty();
}

bool jit_mapped_first(const Plan &plan, bool dbl, int nv)
{
    if (plan.in_map.empty() || nv != 1 || plan.passes.empty()) return false;
    if (dbl) return plan.passes[0].K >= 2;  // one amplitude per 16-byte read
    const PassDev &pd = plan.passes[0];
    int inv[64];
    for (size_t q = 0; q < plan.in_map.size(); ++q) inv[plan.in_map[q]] = (int)q;
    bool has0 = false;  // source bit 0 is a chunk bit (16-byte reads)
    for (int j = 0; j < pd.K; ++j) has0 |= inv[pd.pos[j]] == 0;
    return has0 && pd.K >= 2;
}

std::string jit_pass_source(const Plan &plan, size_t p, bool dbl, int nv, const std::string &name)
{
    if (p + 1 == plan.passes.size() && jit_mapped_last(plan, dbl, nv))
        return remapped_pass_source(plan, p, dbl, name);
    if (p == 0 && jit_mapped_first(plan, dbl, nv)) return mapped_pass_source(plan, dbl, name);
    const PassDev &pd = plan.passes[p];
    const char *T = dbl ? "double" : "float";
    const int K = pd.K, CF = pd.cfast, E = 1 << K, NROW = 1 << (K - CF);
    const int NP = nv == 2 ? std::max(1, pd.nprobe) : 0;
    const int TPB = jit_block_threads(K, nv);
    const char *occ = dbl || nv == 2 ? "" : " __attribute__((amdgpu_waves_per_eu(4)))";
    std::ostringstream o;
    o << "extern \"C\" __global__ __launch_bounds__(" << TPB << ")" << occ << "\nvoid " << name << "(cx<" << T
      << "> *__restrict__ psi, ";
    if (nv == 2) o << "cx<" << T << "> *__restrict__ lam, double *__restrict__ probe_partial, ";
    o << "const cx<" << T << "> *__restrict__ mats)\n{\n";
    o << "  __shared__ __align__(16) cx<" << T << "> c0[" << E << "];\n";
    if (nv == 2) {
        o << "  __shared__ __align__(16) cx<" << T << "> c1[" << E << "];\n";
        o << "  __shared__ double pacc[" << NP << " * kWaves];\n";
        o << "  for (int t = threadIdx.x; t < " << NP << " * kWaves; t += kTPB) pacc[t] = 0.0;\n";
    }
    int ident[64];
    for (int k = 0; k < 64; ++k) ident[k] = k;
    LaneTable lt;
    const CopyPlan cp = copy_plan(lt, K, dbl, TPB, pd.pos, ident, CF);
    if (cp.ok) {
        std::ostringstream b;
        emit_lane_prologue(b, name);
        emit_copy_thread(b, cp, "", pd.pos, ident);
        b << "  const uint64_t c = xcd_slab(blockIdx.x, " << pd.ncomp << ");\n";
        b << "  const uint64_t base = 0ull" << bit_moves("c", pd.comp, pd.ncomp, "uint64_t") << ";\n";
        if (p == 0 && plan.zero_input && nv == 1) {
            // the input is |0...0>: the chunk holding amplitude 0 starts as
            // e_0, every other as zeros and stays zero under the pass's gates
            // (they act within the chunk), so those blocks only store zeros
            b << "  if (base != 0) {\n";
            emit_copy(b, cp, "", "psi", "c0", "base", pd.pos, ident, K, false, dbl, true);
            b << "    return;\n  }\n";
            b << "  for (int t = threadIdx.x; t < " << E << "; t += " << TPB << ") c0[t] = cx<" << T << ">{0, 0};\n";
            b << "  __syncthreads();\n";
            b << "  if (base == 0 && threadIdx.x == 0) c0[swz(0)] = cx<" << T << ">{1, 0};\n";
        } else {
            emit_copy(b, cp, "", "psi", "c0", "base", pd.pos, ident, K, true, dbl);
        }
        if (nv == 2) emit_copy(b, cp, "", "lam", "c1", "base", pd.pos, ident, K, true, dbl);
        b << "  __syncthreads();\n";
        for (int s = 0; s < pd.nsteps; ++s) emit_plan_step(b, plan, pd.step_off + s, E, T, nv, TPB);
        emit_copy(b, cp, "", "psi", "c0", "base", pd.pos, ident, K, false, dbl);
        if (nv == 2) {
            emit_copy(b, cp, "", "lam", "c1", "base", pd.pos, ident, K, false, dbl);
            if (pd.nprobe > 0) {
                b << "  for (int q = threadIdx.x; q < " << pd.nprobe << "; q += kTPB) {\n";
                b << "    double v = 0.0;\n    for (int w = 0; w < kWaves; ++w) v += pacc[q * kWaves + w];\n";
                b << "    probe_partial[c * " << pd.nprobe << " + q] = v;\n  }\n";
            }
        }
        return lane_table_source(lt, name) + o.str() + b.str() + "}\n\n";
    }
    o << "  __shared__ uint64_t rowoff[" << NROW << "];\n";
    o << "  for (int t = threadIdx.x; t < " << NROW << "; t += " << TPB << ") {\n    uint64_t r = 0;\n";
    for (int k = CF; k < K; ++k)
        o << "    r |= (uint64_t)((t >> " << k - CF << ") & 1) << " << pd.pos[k] << ";\n";
    o << "    rowoff[t] = r;\n  }\n";
    // chunk index -> base address: runs of consecutive positions move as one field
    std::ostringstream dep;
    for (int m = 0; m < pd.ncomp;) {
        int len = 1;
        while (m + len < pd.ncomp && pd.comp[m + len] == pd.comp[m] + len) ++len;
        dep << " | (((c >> " << m << ") & " << ((1ull << len) - 1) << "ull) << " << pd.comp[m] << ")";
        m += len;
    }
    o << "  const uint64_t c = xcd_slab(blockIdx.x, " << pd.ncomp << ");\n";
    o << "  const uint64_t base = 0ull" << dep.str() << ";\n";
    o << "  __syncthreads();\n";
    const std::string ntf = ", true, " + std::to_string(TPB);  // non-temporal (streamed once), block size
    if (p == 0 && plan.zero_input && nv == 1) {
        // the input is |0...0>: the chunk holding amplitude 0 starts as e_0,
        // every other as zeros and stays zero (those blocks store zeros only)
        o << "  for (int t = threadIdx.x; t < " << E << "; t += " << TPB << ") c0[t] = cx<" << T << ">{0, 0};\n";
        o << "  __syncthreads();\n";
        o << "  if (base != 0) {\n    chunk_copy<" << T << ", false" << ntf << ">(psi, c0, base, rowoff, " << E
          << ", " << CF << ");\n    return;\n  }\n";
        o << "  if (threadIdx.x == 0) c0[swz(0)] = cx<" << T << ">{1, 0};\n";
    } else {
        o << "  chunk_copy<" << T << ", true" << ntf << ">(psi, c0, base, rowoff, " << E << ", " << CF << ");\n";
    }
    if (nv == 2) o << "  chunk_copy<" << T << ", true" << ntf << ">(lam, c1, base, rowoff, " << E << ", " << CF << ");\n";
    o << "  __syncthreads();\n";
    for (int s = 0; s < pd.nsteps; ++s) emit_plan_step(o, plan, pd.step_off + s, E, T, nv, TPB);
    o << "  chunk_copy<" << T << ", false" << ntf << ">(psi, c0, base, rowoff, " << E << ", " << CF << ");\n";
    if (nv == 2) {
        o << "  chunk_copy<" << T << ", false" << ntf << ">(lam, c1, base, rowoff, " << E << ", " << CF << ");\n";
        if (pd.nprobe > 0) {
            o << "  for (int q = threadIdx.x; q < " << pd.nprobe << "; q += kTPB) {\n";
            o << "    double v = 0.0;\n    for (int w = 0; w < kWaves; ++w) v += pacc[q * kWaves + w];\n";
            o << "    probe_partial[c * " << pd.nprobe << " + q] = v;\n  }\n";
        }
    }
    o << "}\n\n";
    return o.str();
}

bool jit_functions(const Plan &plan, bool dbl, int nv, int device, const std::string &arch,
                   std::vector<hipFunction_t> &out, std::string &err)
{
    const size_t np = plan.passes.size();
    out.assign(np, nullptr);
    std::vector<std::string> keys(np);
    std::vector<std::string> missing_bodies;
    std::vector<std::string> missing_keys;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        std::map<std::string, int> seen;
        for (size_t p = 0; p < np; ++p) {
            if (skip_pass(plan, p, nv)) continue;
            keys[p] = jit_pass_source(plan, p, dbl, nv, "@");
            if (const char *e = std::getenv("QSIM_JIT_FLAGS")) keys[p] = std::string("// ") + e + "\n" + keys[p];
            auto it = g_fns.find({device, keys[p]});
            if (it != g_fns.end()) {
                out[p] = it->second;
            } else if (!seen.count(keys[p])) {
                seen[keys[p]] = 1;
                missing_keys.push_back(keys[p]);
                missing_bodies.push_back(jit_pass_source(plan, p, dbl, nv, kernel_name(keys[p])));
            }
        }
    }
    if (!missing_bodies.empty()) {
        std::vector<std::vector<char>> codes;
        std::vector<std::vector<int>> members;
        if (!compile_all(missing_bodies, arch, codes, members, err)) return false;
        std::lock_guard<std::mutex> lk(g_mu);
        for (size_t u = 0; u < codes.size(); ++u) {
            hipModule_t mod = nullptr;
            hipError_t e = hipModuleLoadData(&mod, codes[u].data());
            if (e != hipSuccess) {
                err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
                return false;
            }
            g_modules.push_back(mod);
            for (int i : members[u]) {
                hipFunction_t f = nullptr;
                e = hipModuleGetFunction(&f, mod, kernel_name(missing_keys[i]).c_str());
                if (e != hipSuccess) {
                    err = std::string("hipModuleGetFunction: ") + hipGetErrorString(e);
                    return false;
                }
                g_fns[{device, missing_keys[i]}] = f;
            }
        }
        for (size_t p = 0; p < np; ++p)
            if (!keys[p].empty() && !out[p]) out[p] = g_fns[{device, keys[p]}];
    }
    return true;
}

bool jit_compile_only(const Plan &plan, bool dbl, int nv, const std::string &arch, std::string &err)
{
    std::vector<std::string> bodies;
    for (size_t p = 0; p < plan.passes.size(); ++p) {
        if (skip_pass(plan, p, nv)) continue;
        bodies.push_back(jit_pass_source(plan, p, dbl, nv, kernel_name(jit_pass_source(plan, p, dbl, nv, "@"))));
    }
    if (bodies.empty()) return true;
    std::vector<std::vector<char>> codes;
    std::vector<std::vector<int>> members;
    return compile_all(bodies, arch, codes, members, err);
}

long jit_compiled_count() { return g_compiled.load(); }

int jit_block_threads(int K, int nv)
{
    // one 16-amplitude register group per thread per step: 2^K / 16 threads,
    // at least 256; two-vector passes keep 256 (their probe partials are per wave)
    return (nv == 1 && K >= 13) ? 512 : kTPB;
}

}  // namespace qsim
