// Volume probes on the device: VOL_AVERAGE / VOL_MIN / VOL_MAX of an FCT expression of the cell
// state and position (prb_zone_vol.f90:34-51, PROBECALC.f90:41-252).
//
// The expression string is parsed on the host with the FCT module's rules (FCT_PARSER.f90:113-340):
// spaces removed, lower case; an expression wholly inside brackets is its inside; a real constant;
// otherwise the binary operators are tried from the lowest priority (&&, ||, !=, ==, <, >, <=, >=, +,
// -, *, /, ^) and the string is split at the LAST occurrence of the first one found outside brackets
// (index_oper, FCT_PARSER.f90:350-405, which also skips the sign of a number's exponent), a leading +
// is dropped and a leading - is the opposite operator; then the functions inv sqr sqrt exp ln log sin
// cos tan sinh cosh tanh asin acos atan asinh acosh atanh erf erfc abs sign step ramp written
// "name(...)"; then the constants _pi and _e; then variables.  The tree is compiled to a postfix
// program that every work-item runs on its own cell (the same instruction stream for all, so no
// divergence), with the variables of prb_vol_calc (PROBECALC.f90:200-233): the primitive fields RHO,
// PS, the velocity V (its norm) and V_X, V_Y, V_Z, the cell centre X, Y, Z and the time T.  The
// cell results are reduced in a fixed order (wavefront butterflies, the block's wavefronts in
// order, then the blocks in order), so every call gives the same bits.
#include <algorithm>
#include <cctype>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "ty_probe.h"

// fixed workspace of the allocation-free probe evaluation (one copy per GPU, host code reaches it
// only through the device address probe_scratch_setup looks up)
static __device__ __align__(256) double scratch[ty::PROBE_SCRATCH_DOUBLES];

namespace ty {

namespace {

// ---------------------------------------------------------------------------
// host: FCT parse tree (FCT_NODE.f90) and compilation

enum BinOp { B_AND = 1, B_OR, B_NEQ, B_EQ, B_LT, B_GT, B_LEQ, B_GEQ, B_ADD, B_SUB, B_MUL, B_DIV, B_POW };
const char* const BIN_NAME[] = {"", "&&", "||", "!=", "==", "<", ">", "<=", ">=", "+", "-", "*", "/", "^"};
// FCT_NODE.f90:31-55 (fct_opp = 0 is not looked for by name)
const char* const FCT_NAME[] = {"", "inv", "sqr", "sqrt", "exp", "ln", "log", "sin", "cos", "tan", "sinh",
                                "cosh", "tanh", "asin", "acos", "atan", "asinh", "acosh", "atanh", "erf", "erfc",
                                "abs", "sign", "step", "ramp"};
constexpr int NFCT = 24;
const char* const VAR_NAME[] = {"rho", "ps", "v", "v_x", "v_y", "v_z", "x", "y", "z", "t"};
constexpr int NVAR = 10;

struct Node {
    int kind = 0;  // 0 constant, 1 variable, 2 unary (fct id, 0 = opposite), 3 binary
    int id = 0;
    double value = 0.0;
    std::unique_ptr<Node> left, right;
};

struct Parser {
    std::string err;

    // FCT_PARSER.f90:350-405: FIRST index (1-based, 0 = none) of op in s outside brackets, not the
    // sign of a number's exponent
    int index_oper(const std::string& s, const std::string& op)
    {
        const int ls = (int)s.size();
        const size_t p = s.find(op);
        const int iop = p == std::string::npos ? 0 : (int)p + 1;
        const size_t b = s.find('(');
        const int ibk1 = b == std::string::npos ? ls + 1 : (int)b + 1;
        if (iop == 0) return 0;
        if (iop == ibk1) return iop;
        if (iop < ibk1) {
            if (iop >= 3 && iop < ls && strchr("-+", s[iop - 1]) && strchr("eE", s[iop - 2]) &&
                strchr(".0123456789", s[iop - 3]) && strchr("0123456789", s[iop])) {
                const int r = index_oper(s.substr(iop), op);
                return r == 0 ? 0 : iop + r;
            }
            return iop;
        }
        const int rb = index_oper(s.substr(ibk1), ")");
        if (rb == 0) {
            err = "did not find ending bracket in \"" + s + "\"";
            return 0;
        }
        const int ibk2 = ibk1 + rb;
        const int r = index_oper(s.substr(ibk2), op);
        return r == 0 ? 0 : ibk2 + r;
    }

    // STRING.f90:265-316
    static bool all_figures(const std::string& s)
    {
        for (char ch : s)
            if (!isdigit((unsigned char)ch)) return false;
        return true;
    }
    static bool is_real(const std::string& w)
    {
        const int nch = (int)w.size();
        if (nch == 0) return false;
        size_t pe = w.find_first_of("dDeE");
        int iexp = pe == std::string::npos ? nch + 1 : (int)pe + 1;
        size_t pd = w.find('.');
        int idot = pd == std::string::npos ? iexp : (int)pd + 1;
        bool afterdot = false;
        if (idot + 1 <= iexp - 1) {
            if (!all_figures(w.substr(idot, iexp - 1 - idot))) return false;
            afterdot = true;
        }
        if (iexp == nch) return false;
        if (iexp + 1 <= nch) {
            int is = iexp + 1;
            if (is != nch && strchr("+-", w[is - 1])) ++is;
            if (!all_figures(w.substr(is - 1))) return false;
        }
        int is = 1;
        if (strchr("+-", w[0])) ++is;
        if (idot - 1 >= is) {
            if (!all_figures(w.substr(is - 1, idot - is))) return false;
        } else if (!afterdot) {
            return false;
        }
        return true;
    }

    // FCT_PARSER.f90:113-340
    std::unique_ptr<Node> parse(const std::string& s)
    {
        if (!err.empty()) return nullptr;
        const int ls = (int)s.size();
        if (ls == 0) {
            err = "empty expression";
            return nullptr;
        }
        int depth = 0;
        for (char ch : s) {
            depth += ch == '(' ? 1 : ch == ')' ? -1 : 0;
            if (depth < 0) break;
        }
        if (depth != 0) {
            err = std::string("unmatched '") + (depth > 0 ? "(" : ")") + "' in \"" + s + "\"";
            return nullptr;
        }
        if (s[0] == '(') {
            const int r = index_oper(s.substr(1), ")");
            if (r == 0) {
                err = "unmatched '(' in \"" + s + "\"";
                return nullptr;
            }
            if (r == ls - 1) return parse(s.substr(1, ls - 2));
        }
        if (is_real(s)) {
            std::string t = s;
            for (char& ch : t)
                if (ch == 'd' || ch == 'D') ch = 'e';
            auto n = std::make_unique<Node>();
            n->kind = 0;
            n->value = strtod(t.c_str(), nullptr);
            return n;
        }
        for (int op = B_AND; op <= B_POW; ++op) {
            const std::string name = BIN_NAME[op];
            // the LAST occurrence outside brackets
            int istr = 0, pos = ls;
            while (pos != 0) {
                pos = index_oper(s.substr(istr), name);
                if (!err.empty()) return nullptr;
                istr += pos;
            }
            if (istr == 0) continue;
            if (istr == ls) {
                err = "binary operator at the end of \"" + s + "\"";
                return nullptr;
            }
            if (istr == 1) {
                if (op == B_ADD) return parse(s.substr(1));
                if (op == B_SUB) {
                    auto n = std::make_unique<Node>();
                    n->kind = 2;
                    n->id = 0;
                    n->left = parse(s.substr(1));
                    return n;
                }
                err = "binary operator at the start of \"" + s + "\"";
                return nullptr;
            }
            if (op < B_ADD) {
                // FCT_EVAL.f90:219-230 evaluates only + - * / ^
                err = "operator " + name + " is not implemented in FCT evaluation";
                return nullptr;
            }
            auto n = std::make_unique<Node>();
            n->kind = 3;
            n->id = op;
            n->left = parse(s.substr(0, istr - 1));
            n->right = parse(s.substr(istr - 1 + name.size()));
            return n;
        }
        for (int f = 1; f <= NFCT; ++f) {
            const std::string key = std::string(FCT_NAME[f]) + "(";
            if (s.compare(0, key.size(), key) == 0) {
                auto n = std::make_unique<Node>();
                n->kind = 2;
                n->id = f;
                n->left = parse(s.substr(key.size() - 1));
                return n;
            }
        }
        if (s.compare(0, 3, "_pi") == 0 || s.compare(0, 2, "_e") == 0) {
            auto n = std::make_unique<Node>();
            n->kind = 0;
            n->value = s[1] == 'p' ? 3.14159265358979323846 : 2.71828182845904523536;
            return n;
        }
        if (!isalpha((unsigned char)s[0]) && s[0] != '_') {
            err = "bad variable name \"" + s + "\"";
            return nullptr;
        }
        for (int v = 0; v < NVAR; ++v)
            if (s == VAR_NAME[v]) {
                auto n = std::make_unique<Node>();
                n->kind = 1;
                n->id = v;
                return n;
            }
        err = "unknown variable \"" + s + "\" (probes read RHO, PS, V, V_X, V_Y, V_Z, X, Y, Z, T)";
        return nullptr;
    }
};

// post-order emission; returns the stack depth the subtree needs
int emit(const Node* n, ProbeProg& p, std::string& err)
{
    if (!n || !err.empty()) return 0;
    auto push = [&](int op, int arg) {
        if (p.nop >= PROBE_MAXOP) {
            err = "expression too long";
            return;
        }
        p.op[p.nop] = (unsigned char)op;
        p.arg[p.nop] = (unsigned char)arg;
        ++p.nop;
    };
    switch (n->kind) {
    case 0: {
        if (p.ncst >= PROBE_MAXCST) {
            err = "too many constants";
            return 0;
        }
        p.cst[p.ncst] = n->value;
        push(PO_CST, p.ncst++);
        return 1;
    }
    case 1:
        push(PO_VAR, n->id);
        return 1;
    case 2: {
        const int d = emit(n->left.get(), p, err);
        push(PO_FN, n->id);
        return d;
    }
    default: {
        const int dl = emit(n->left.get(), p, err);
        const int dr = emit(n->right.get(), p, err);
        push(n->id, 0);
        return std::max(dl, 1 + dr);
    }
    }
}

}  // namespace

int probe_compile(const char* expr, int type, ProbeProg* prog, std::string* err)
{
    std::string s;
    for (const char* q = expr; *q; ++q)
        if (!isspace((unsigned char)*q)) s += (char)tolower((unsigned char)*q);
    Parser ps;
    std::unique_ptr<Node> root = ps.parse(s);
    if (!ps.err.empty() || !root) {
        *err = ps.err.empty() ? "cannot parse the expression" : ps.err;
        return 1;
    }
    memset(prog, 0, sizeof(*prog));
    prog->type = type;
    std::string e;
    const int depth = emit(root.get(), *prog, e);
    if (!e.empty()) {
        *err = e;
        return 1;
    }
    if (depth > PROBE_STACK) {
        *err = "expression nests too deeply";
        return 1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// device

namespace {

__device__ __noinline__ double probe_fn(int f, double x)
{
    switch (f) {
    case 0: return -x;
    case 1: return 1.0 / x;
    case 2: return x * x;
    case 3: return sqrt(x);
    case 4: return exp(x);
    case 5: return log(x);
    case 6: return log10(x);
    case 7: return sin(x);
    case 8: return cos(x);
    case 9: return tan(x);
    case 10: return sinh(x);
    case 11: return cosh(x);
    case 12: return tanh(x);
    case 13: return asin(x);
    case 14: return acos(x);
    case 15: return atan(x);
    case 16: return asinh(x);
    case 17: return acosh(x);
    case 18: return atanh(x);
    case 19: return erf(x);
    case 20: return erfc(x);
    case 21: return fabs(x);
    case 22: return 1.0 / x;  // FCT_EVAL.f90:297-298: SIGN evaluates as INV
    case 23: return x >= 0.0 ? 1.0 : 0.0;
    default: return x >= 0.0 ? x : 0.0;  // ramp
    }
}

__device__ __noinline__ double probe_pow(double a, double b) { return pow(a, b); }

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wmin(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = v < o ? v : o;
    }
    return v;
}
__device__ __forceinline__ double wmax(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = v > o ? v : o;
    }
    return v;
}

constexpr int PB = 256;

__global__ __launch_bounds__(PB) void k_probe(ProbeProg p, ProbeData d, double* part)
{
    __shared__ double sa[PB / 64], sb[PB / 64];
    __shared__ double stk[PROBE_STACK - 1][PB];
    const int c = blockIdx.x * PB + threadIdx.x;
    double f = 0.0, w = 0.0;
    const bool avg = p.type == PROBE_AVG;
    if (c < d.n) {
        // calc_varprim_ns (SOURCE/EQNS/calc_varprim_ns.f90:47-52) on the FP64 state
        const int n = d.n;
        const double rho = d.u0[c];
        double vel[3] = {0.0, 0.0, 0.0}, v2 = 0.0;
        for (int e = 0; e < d.dim; ++e) {
            vel[e] = d.u0[(size_t)(2 + e) * n + c] / rho;
            v2 += vel[e] * vel[e];
        }
        const double ps = d.g1 * (d.u0[(size_t)n + c] - .5 * rho * v2);
        const double vn = sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
        // postfix evaluation with the top of the stack in a register and the rest in LDS (a
        // column per work-item); the program is the same for every lane, so every branch is uniform
        double top = 0.0;
        int sp = 0;  // entries below the top
        for (int i = 0; i < p.nop; ++i) {
            const int op = p.op[i], arg = p.arg[i];
            if (op == PO_CST || op == PO_VAR) {
                if (i > 0) stk[sp++][threadIdx.x] = top;
                if (op == PO_CST) {
                    top = p.cst[arg];
                } else {
                    switch (arg) {
                    case 0: top = rho; break;
                    case 1: top = ps; break;
                    case 2: top = vn; break;
                    case 3: top = vel[0]; break;
                    case 4: top = vel[1]; break;
                    case 5: top = vel[2]; break;
                    case 6: top = d.cc[c]; break;
                    case 7: top = d.cc[(size_t)n + c]; break;
                    case 8: top = d.cc[(size_t)2 * n + c]; break;
                    default: top = d.t; break;
                    }
                }
            } else if (op == PO_FN) {
                top = probe_fn(arg, top);
            } else {
                const double a = stk[--sp][threadIdx.x];
                switch (op) {
                case 9: top = a + top; break;
                case 10: top = a - top; break;
                case 11: top = a * top; break;
                case 12: top = a / top; break;
                default: top = probe_pow(a, top); break;
                }
            }
        }
        f = top;
        if (avg) {
            w = d.vol[c];
            f = f * w;
        }
    } else if (!avg) {
        f = p.type == PROBE_MIN ? DBL_MAX : -DBL_MAX;
    }
    f = avg ? wsum(f) : (p.type == PROBE_MIN ? wmin(f) : wmax(f));
    w = wsum(w);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sa[wid] = f;
        sb[wid] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = sa[0], b = sb[0];
        for (int i = 1; i < PB / 64; ++i) {
            a = avg ? a + sa[i] : (p.type == PROBE_MIN ? (a < sa[i] ? a : sa[i]) : (a > sa[i] ? a : sa[i]));
            b += sb[i];
        }
        part[2 * (size_t)blockIdx.x] = a;
        part[2 * (size_t)blockIdx.x + 1] = b;
    }
}

__global__ __launch_bounds__(1024) void k_probe_final(int type, const double* part, int nb, double* out)
{
    __shared__ double sa[16], sb[16];
    const bool avg = type == PROBE_AVG;
    double a = avg ? 0.0 : (type == PROBE_MIN ? DBL_MAX : -DBL_MAX), b = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        const double x = part[2 * (size_t)i];
        a = avg ? a + x : (type == PROBE_MIN ? (a < x ? a : x) : (a > x ? a : x));
        b += part[2 * (size_t)i + 1];
    }
    a = avg ? wsum(a) : (type == PROBE_MIN ? wmin(a) : wmax(a));
    b = wsum(b);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sa[wid] = a;
        sb[wid] = b;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    a = sa[0];
    b = sb[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
        a = avg ? a + sa[i] : (type == PROBE_MIN ? (a < sa[i] ? a : sa[i]) : (a > sa[i] ? a : sa[i]));
        b += sb[i];
    }
    *out = avg ? a / b : a;
}

}  // namespace

int probe_blocks(int n) { return (n + PB - 1) / PB; }

hipError_t launch_probe(const ProbeProg& p, const ProbeData& d, double* part, double* out, hipStream_t s)
{
    const int nb = probe_blocks(d.n);
    hipLaunchKernelGGL(k_probe, dim3(nb), dim3(PB), 0, s, p, d, part);
    hipLaunchKernelGGL(k_probe_final, dim3(1), dim3(1024), 0, s, p.type, (const double*)part, nb, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// fixed workspace of the allocation-free evaluation: the block partials and results of
// ty_probe_eval_noalloc.  Every GPU holds one copy for the life of the process; each call
// overwrites every partial it reads, so nothing in it has to start at zero.

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch[MAX_DEVICES] = {};

}  // namespace

static_assert(PROBE_SCRATCH_DOUBLES >= 2 * ((PROBE_SCRATCH_MAX_CELLS + PB - 1) / PB) + PROBE_SCRATCH_MAX_PROBES,
              "probe workspace too small for its limits");

hipError_t probe_scratch_setup(int device, double** ptr)
{
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        // the device address of this GPU's copy of the array, looked up with the device current
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<double*>(p);
    }
    *ptr = g_scratch[device];
    return hipSuccess;
}

}  // namespace ty
