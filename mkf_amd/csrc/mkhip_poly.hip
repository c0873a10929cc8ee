// MKF CKKS polynomial evaluation and the public linear operations around it
// (software/mkf/mkf_cpu/true_mkf_ckks_polynomial.go:18-273, true_mkf_ckks.go:176-303).
//
// The evaluator's schedule depends only on levels, scales and the coefficients,
// so a call is planned first on metadata (a list of steps over values, as the
// reference's metadata-only level test runs it), then laid out in device memory
// (values reuse the buffers of values no later step reads) and emitted as one
// plan: the linear steps as k_linear launches, the products as the stages of the
// MKF mul+relin plan (each operand folded at its own level, no rescale), each
// rescale as an inverse transform of the last limb and k_rescale_ntt.  The plan
// replays without host work and can be captured as one hipGraph.
#include "mkhip_internal.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>

using namespace mkhip;

namespace {

typedef std::complex<double> cd;
typedef unsigned __int128 u128;
constexpr double kNegligible = 1e-14;  // Lattigo v2.3.0 ckks.IsNegligbleThreshold

int bitlen(u64 x) {
    int n = 0;
    for (; x; x >>= 1) ++n;
    return n;
}
bool not_negligible(cd v) { return std::fabs(v.real()) > kNegligible || std::fabs(v.imag()) > kNegligible; }
u64 cred(u64 a, u64 q) { return a >= q ? a - q : a; }

u64 pow2_mod(int e, u64 q) {
    u64 r = 1 % q, b = 2 % q;
    for (; e > 0; e >>= 1, b = h_mulmod(b, b, q))
        if (e & 1) r = h_mulmod(r, b, q);
    return r;
}

// |x| = m 2^e with m an integer below 2^53 (x finite, nonzero)
void split_double(double x, u64 &m, int &e) {
    int E;
    const double f = std::frexp(std::fabs(x), &E);
    m = (u64)std::ldexp(f, 53);
    e = E - 53;
}

u64 floor_mod(double y, u64 q) {  // floor(y) mod q, y >= 0
    if (y < 18446744073709551616.0) return (u64)y % q;
    u64 m;
    int e;
    split_double(y, m, e);
    return h_mulmod(m % q, pow2_mod(e, q), q);
}

// scaleUpExact (mkckks/utils.go:59-84): round |value| n half up in float64, reduce,
// q - r for a negative value (q itself when r = 0, as the reference writes it).
u64 scale_up_exact(double value, double n, u64 q) {
    const bool neg = value < 0;
    const double x = neg ? -n * value : n * value;
    const u64 r = floor_mod(x + 0.5, q);
    return neg ? q - r : r;
}

// trunc(a b) mod q with the product exact (a 128-bit big.Float holds it) and the
// Euclidean residue of big.Int.Mod (true_mkf_ckks_polynomial.go:258-264).
u64 trunc_product_mod(double a, double b, u64 q) {
    if (a == 0 || b == 0) return 0;
    u64 ma, mb;
    int ea, eb;
    split_double(a, ma, ea);
    split_double(b, mb, eb);
    const u128 p = (u128)ma * mb;
    const int e = ea + eb;
    u64 r;
    if (e >= 0)
        r = h_mulmod((u64)(p % q), pow2_mod(e, q), q);
    else
        r = (u64)((-e >= 128 ? (u128)0 : p >> -e) % q);
    return ((a < 0) != (b < 0)) && r ? q - r : r;
}

// Go's x - float64(int64(x)) != 0 (int64 of an out-of-range float is the minimum
// int64 on amd64)
bool has_fraction(double x) {
    if (x == 0) return false;
    if (std::fabs(x) < 9.2233720368547758e18) return x - (double)(int64_t)x != 0;
    return x != -9223372036854775808.0;
}

struct Poly {
    std::vector<cd> c;
    int maxdeg = 0;
    bool lead = true;
    int basis = 0;
    int degree() const { return (int)c.size() - 1; }
};

struct Val {
    int id = -1;
    int level = 0;
    double scale = 0;
};

enum { S_LIN = 0, S_MUL = 1, S_RESC = 2 };

struct Step {
    int kind = S_LIN, op = LIN_COPY;
    int a = -1, b = -1, out = -1, tmp = -1;
    int level = 0;         // S_LIN: limbs 0..level take op; above, up to out's level, copied from b
    bool c0only = false;   // LIN_ADDC: components 1..k copied from a
    std::vector<u64> lo, hi;
};

class Prog {
  public:
    mkhip_ctx *c;
    int group, k, nct, N;
    int status = MKHIP_OK;
    int mul_calls = 0;
    struct V {
        int level = 0;
        std::vector<u64 *> ext;  // per-ciphertext pointers of an input / output value
        size_t off = 0;
    };
    std::vector<V> vals;
    std::vector<Step> steps;

    Prog(mkhip_ctx *ctx, int g, int kk, int n) : c(ctx), group(g), k(kk), nct(n), N(ctx->N) {}

    int value(int level) {
        vals.push_back(V());
        vals.back().level = level;
        return (int)vals.size() - 1;
    }
    Val input(int level, double scale, const std::vector<u64 *> &ptrs) {
        Val v{value(level), level, scale};
        vals[v.id].ext = ptrs;
        return v;
    }
    bool fail(int st) {
        if (status == MKHIP_OK) status = st;
        return false;
    }
    bool ok() const { return status == MKHIP_OK; }
    size_t words(int v) const { return (size_t)(1 + k) * (vals[v].level + 1) * N; }

    // ---------------------------------------------------------- operations
    Val zero(int level, double scale) {
        Val v{value(level), level, scale};
        Step s;
        s.op = LIN_ZERO;
        s.out = v.id;
        s.level = level;
        steps.push_back(s);
        return v;
    }

    // AddConst / MultByConst constants (mkckks/evaluator.go:118-160), limbs 0..level
    void const_halves(int level, double re, double im, double scale, bool mform, Step &s) const {
        s.lo.resize(level + 1);
        s.hi.resize(level + 1);
        for (int i = 0; i <= level; ++i) {
            const u64 q = c->Q[i], w = c->psi_half[i];
            u64 sr = 0, si = 0, sc = 0;
            if (re != 0) sc = sr = scale_up_exact(re, scale, q);
            if (im != 0) {
                si = h_mulmod(scale_up_exact(im, scale, q) % q, w, q);
                sc = cred(sc + si, q);
            }
            u64 lo = sc % q, hi = (im != 0 ? cred(sr + (q - si), q) : sc) % q;
            s.lo[i] = mform ? h_mform(lo, q) : lo;
            s.hi[i] = mform ? h_mform(hi, q) : hi;
        }
    }

    Val add_const(Val a, double re, double im) {
        Val v{value(a.level), a.level, a.scale};
        Step s;
        s.op = LIN_ADDC;
        s.a = a.id;
        s.out = v.id;
        s.level = a.level;
        s.c0only = true;
        const_halves(a.level, re, im, a.scale, false, s);
        steps.push_back(s);
        return v;
    }

    Val mult_by_const(Val a, double re, double im) {
        // getConstAndScale (mkckks/evaluator.go:41-95)
        double scale = 1;
        if (has_fraction(re) || has_fraction(im)) scale = (double)c->Q[a.level];
        Val v{value(a.level), a.level, a.scale * scale};
        Step s;
        s.op = LIN_MULC;
        s.a = a.id;
        s.out = v.id;
        s.level = a.level;
        const_halves(a.level, re, im, scale, true, s);
        steps.push_back(s);
        return v;
    }

    // AddNew / SubNew over evaluateInPlace (mkckks/evaluator.go:201-330)
    Val add_sub(Val a, Val b, bool sub) {
        const int level = std::min(a.level, b.level);
        Val ta = a, tb = b;
        if (b.scale > a.scale && std::floor(b.scale / a.scale) > 1)
            ta = mult_by_const(a, std::floor(b.scale / a.scale), 0);
        else if (a.scale > b.scale && std::floor(a.scale / b.scale) > 1)
            tb = mult_by_const(b, std::floor(a.scale / b.scale), 0);
        Val v{value(level), level, std::max(a.scale, b.scale)};
        Step s;
        s.op = sub ? LIN_SUB : LIN_ADD;
        s.a = ta.id;
        s.b = tb.id;
        s.out = v.id;
        s.level = level;
        steps.push_back(s);
        return v;
    }

    // MultByGaussianIntegerAndAdd: target + a (re + im i), re / im residues per limb
    Val gauss_add(Val a, const std::vector<u64> &rr, const std::vector<u64> &ri, Val target) {
        const int level = std::min(a.level, target.level);
        Val v{value(target.level), target.level, target.scale};
        Step s;
        s.op = LIN_MULCADD;
        s.a = a.id;
        s.b = target.id;
        s.out = v.id;
        s.level = level;
        s.lo.resize(level + 1);
        s.hi.resize(level + 1);
        for (int i = 0; i <= level; ++i) {
            const u64 q = c->Q[i];
            const u64 t = h_mulmod(ri[i], c->psi_half[i], q);
            s.lo[i] = h_mform(cred(rr[i] + t, q), q);
            s.hi[i] = h_mform(cred(rr[i] + (q - t), q) % q, q);
        }
        steps.push_back(s);
        return v;
    }

    Val drop(Val a, int levels) {
        const int lv = a.level - levels;
        Val v{value(lv), lv, a.scale};
        Step s;
        s.op = LIN_COPY;
        s.a = a.id;
        s.out = v.id;
        s.level = lv;
        steps.push_back(s);
        return v;
    }

    Val mulrelin(Val a, Val b) {
        const int level = std::min(a.level, b.level);
        Val v{value(level), level, a.scale * b.scale};
        Step s;
        s.kind = S_MUL;
        s.a = a.id;
        s.b = b.id;
        s.out = v.id;
        s.level = level;
        steps.push_back(s);
        ++mul_calls;
        return v;
    }

    // Evaluator.Rescale (mkckks/evaluator.go:360-407): one pass per divided prime
    Val rescale(Val a, double min_scale) {
        if (min_scale <= 0 || a.scale == 0 || a.level == 0) {
            fail(MKHIP_ERR_SCALE);
            return a;
        }
        double sc = a.scale;
        int nb = 0;
        while (a.level - nb >= 0 && sc / (double)c->Q[a.level - nb] >= min_scale / 2) {
            sc /= (double)c->Q[a.level - nb];
            ++nb;
        }
        if (nb > a.level) {  // the reference indexes below Q[0] here
            fail(MKHIP_ERR_SCALE);
            return a;
        }
        if (nb == 0) return a;
        Val cur = a;
        for (int r = 0; r < nb; ++r) {
            Val v{value(cur.level - 1), cur.level - 1, cur.scale};
            Step s;
            s.kind = S_RESC;
            s.a = cur.id;
            s.out = v.id;
            s.tmp = value(0);
            s.level = cur.level;
            steps.push_back(s);
            cur = v;
        }
        cur.scale = sc;
        return cur;
    }

    void copy_to(Val a, int out) {
        Step s;
        s.op = LIN_COPY;
        s.a = a.id;
        s.out = out;
        s.level = a.level;
        steps.push_back(s);
    }

    // ---------------------------------------------------------- EvaluatePoly
    std::map<int, Val> powers;

    void power_basis(int degree, double scale, int basis) {
        if (!ok() || powers.count(degree)) return;
        int left, right, subd = 0;
        if ((degree & (degree - 1)) == 0) {
            left = right = degree / 2;
        } else {
            const int kk = (int)std::ceil(std::log2((double)degree)) - 1;
            left = (1 << kk) - 1;
            right = degree + 1 - (1 << kk);
            if (basis == MKHIP_BASIS_CHEBYSHEV) subd = (int)std::fabs((double)left - (double)right);
        }
        power_basis(left, scale, basis);
        power_basis(right, scale, basis);
        if (!ok()) return;
        Val p = rescale(mulrelin(powers[left], powers[right]), scale);
        if (basis == MKHIP_BASIS_CHEBYSHEV) {
            p = add_sub(p, p, false);
            if (subd == 0) {
                p = add_const(p, -1, 0);
            } else {
                power_basis(subd, scale, basis);
                if (!ok()) return;
                p = add_sub(p, powers[subd], true);
            }
        }
        powers[degree] = p;
    }

    static void split(const Poly &p, int sp, Poly &quo, Poly &rem) {
        rem = Poly();
        rem.c.assign(p.c.begin(), p.c.begin() + sp);
        rem.basis = p.basis;
        rem.lead = false;
        rem.maxdeg = p.maxdeg == p.degree() ? sp - 1 : p.maxdeg - (p.degree() - sp + 1);
        quo = Poly();
        quo.c.assign(p.degree() - sp + 1, cd(0, 0));
        quo.maxdeg = p.maxdeg;
        quo.basis = p.basis;
        quo.c[0] = p.c[sp];
        if (p.basis == MKHIP_BASIS_STANDARD) {
            for (int i = sp + 1; i <= p.degree(); ++i) quo.c[i - sp] = p.c[i];
        } else {
            for (int i = sp + 1, off = 1; i <= p.degree(); ++i, ++off) {
                quo.c[i - sp] = 2.0 * p.c[i];
                rem.c[sp - off] -= p.c[i];
            }
        }
        quo.lead = p.lead;
    }

    Val recurse(double target, int lsplit, int ldeg, const Poly &p) {
        if (!ok()) return Val();
        if (p.degree() < (1 << lsplit)) {
            if (p.lead && lsplit > 1 && p.maxdeg % (1 << (lsplit + 1)) > (1 << (lsplit - 1))) {
                const int nd = bitlen((u64)p.degree());
                return recurse(target, nd >> 1, nd, p);
            }
            return from_power_basis(target, p);
        }
        int next = 1 << lsplit;
        while (next < (p.degree() >> 1) + 1) next <<= 1;
        Poly quo, rem;
        split(p, next, quo, rem);
        if (!powers.count(next)) {
            fail(MKHIP_ERR_ARG);
            return Val();
        }
        int level = powers[next].level - 1;
        if (quo.maxdeg >= (1 << (ldeg - 1)) && quo.lead) ++level;
        const double qi = (double)c->Q[level];
        Val left = recurse(target * qi / powers[next].scale, lsplit, ldeg, quo);
        Val right = recurse(target, lsplit, ldeg, rem);
        if (!ok()) return Val();
        if (left.level > right.level)
            while (left.level != right.level + 1) left = drop(left, 1);
        left = mulrelin(left, powers[next]);
        if (left.level > right.level) {
            left = rescale(left, target);
            return add_sub(left, right, false);
        }
        left = add_sub(left, right, false);
        return rescale(left, target);
    }

    Val from_power_basis(double target, const Poly &p) {
        int mind = 0;
        for (int d = p.degree(); d > 0; --d)
            if (not_negligible(p.c[d])) {
                mind = d;
                break;
            }
        const cd constant = p.c[0];
        if (mind == 0) {
            Val r = zero(powers[1].level, target);
            if (not_negligible(constant)) return add_const(r, constant.real(), constant.imag());
            return r;
        }
        mind = p.degree();
        if (!powers.count(mind)) {
            fail(MKHIP_ERR_ARG);
            return Val();
        }
        const double qi = (double)c->Q[powers[mind].level];
        Val r = zero(powers[mind].level, target * qi);
        if (not_negligible(constant)) r = add_const(r, constant.real(), constant.imag());
        for (int d = p.degree(); d > 0; --d) {
            const cd cf = p.c[d];
            if (!not_negligible(cf)) continue;
            if (!powers.count(d)) {
                fail(MKHIP_ERR_ARG);
                return Val();
            }
            const double cs = target * qi / powers[d].scale;
            const int lv = std::min(powers[d].level, r.level);
            std::vector<u64> rr(lv + 1), ri(lv + 1);
            for (int i = 0; i <= lv; ++i) {
                rr[i] = trunc_product_mod(cf.real(), cs, c->Q[i]);
                ri[i] = trunc_product_mod(cf.imag(), cs, c->Q[i]);
            }
            r = gauss_add(powers[d], rr, ri, r);
        }
        return rescale(r, target);
    }

    // EvaluatePoly (true_mkf_ckks_polynomial.go:79-111)
    Val evaluate(Val x, Poly poly, double target) {
        if (poly.c.empty()) {
            fail(MKHIP_ERR_ARG);
            return x;
        }
        if (x.level < bitlen((u64)poly.degree())) {
            fail(MKHIP_ERR_LEVEL);
            return x;
        }
        powers.clear();
        powers[1] = x;
        const int ldeg = bitlen((u64)poly.degree());
        const int lsplit = ldeg >> 1;
        // paperMKFOddOrEven (:41-62)
        bool odd = true, even = true;
        for (int i = 0; i < (int)poly.c.size(); ++i) {
            const bool nz = not_negligible(poly.c[i]);
            odd = odd && !((i & 1) == 0 && nz);
            even = even && !((i & 1) == 1 && nz);
            if (!odd && !even) break;
        }
        if (odd || even)
            for (int i = even ? 1 : 0; i < (int)poly.c.size(); i += 2) poly.c[i] = cd(0, 0);
        for (int d = 2; d < (1 << lsplit); ++d)
            if (!(even || odd) || ((d & 1) == 0 && even) || ((d & 1) == 1 && odd)) power_basis(d, target, poly.basis);
        for (int i = lsplit; i < ldeg; ++i) power_basis(1 << i, target, poly.basis);
        Val r = recurse(target, lsplit, ldeg, poly);
        r.scale = target;
        return r;
    }

    // ---------------------------------------------------------- layout and emission
    int build(int use_graph, mkhip_plan **out);
};

u64 *base_of(const Prog &P, int v, int i, char *arena) {
    const Prog::V &V = P.vals[v];
    return V.ext.empty() ? (u64 *)(arena + V.off) + (size_t)i * P.words(v) : V.ext[i];
}

int Prog::build(int use_graph, mkhip_plan **out) {
    *out = nullptr;
    if (!ok()) return status;
    const int nv = (int)vals.size(), ns = (int)steps.size();
    // last step reading each value
    std::vector<int> last(nv, -1);
    for (int s = 0; s < ns; ++s) {
        const Step &S = steps[s];
        for (int v : {S.a, S.b, S.tmp})
            if (v >= 0) last[v] = std::max(last[v], s);
    }
    // buffers: a value takes the buffer of a dead value of the same size, else a new one
    std::map<size_t, std::vector<size_t>> free_by_size;
    Carve cv;
    auto place = [&](int v) {
        if (!vals[v].ext.empty()) return;
        const size_t bytes = (size_t)nct * words(v) * 8;
        auto &fl = free_by_size[bytes];
        if (!fl.empty()) {
            vals[v].off = fl.back();
            fl.pop_back();
        } else {
            vals[v].off = cv.take(bytes);
        }
    };
    std::vector<char> placed(nv, 0), freed(nv, 0);
    auto release = [&](int v) {
        if (v < 0 || freed[v] || !placed[v] || !vals[v].ext.empty()) return;
        freed[v] = 1;
        free_by_size[(size_t)nct * words(v) * 8].push_back(vals[v].off);
    };
    for (int s = 0; s < ns; ++s) {
        const Step &S = steps[s];
        for (int v : {S.out, S.tmp})
            if (v >= 0 && !placed[v]) {
                place(v);
                placed[v] = 1;
            }
        // operands read for the last time here, and a result nothing reads, free
        // their buffers once this step is done (a value is released once)
        for (int v : {S.a, S.b, S.tmp})
            if (v >= 0 && last[v] == s && v != S.out) release(v);
        if (S.out >= 0 && last[S.out] < 0) release(S.out);
    }
    // invariant: values live at the same time never share memory
    {
        std::vector<int> first(nv, -1);
        for (int s = 0; s < ns; ++s)
            for (int v : {steps[s].out, steps[s].tmp})
                if (v >= 0 && first[v] < 0) first[v] = s;
        for (int u = 0; u < nv; ++u) {
            if (first[u] < 0 || !vals[u].ext.empty()) continue;
            const int eu = std::max(first[u], last[u]);
            const size_t bu = (size_t)nct * words(u) * 8;
            for (int v = u + 1; v < nv; ++v) {
                if (first[v] < 0 || !vals[v].ext.empty()) continue;
                const int ev = std::max(first[v], last[v]);
                if (first[u] > ev || first[v] > eu) continue;
                const size_t bv = (size_t)nct * words(v) * 8;
                if (vals[u].off < vals[v].off + bv && vals[v].off < vals[u].off + bu) return MKHIP_ERR_ARG;
            }
        }
    }
    // products: workspace and descriptor sizes
    std::vector<size_t> sub_desc(ns, 0);
    size_t ws_need = 0, sub_total = 0;
    std::vector<mkhip_mkf_op> ops(nct);
    for (int s = 0; s < ns; ++s) {
        const Step &S = steps[s];
        if (S.kind != S_MUL) continue;
        for (int i = 0; i < nct; ++i) {
            ops[i] = mkhip_mkf_op();
            ops[i].left = ops[i].right = (const uint64_t *)(uintptr_t)256;
            ops[i].out = (uint64_t *)(uintptr_t)256;
            ops[i].group = group;
            ops[i].level = S.level;
            ops[i].nres = 0;
        }
        size_t w = 0, d = 0;
        PlanMem q;
        q.query = true;
        q.ws_need = &w;
        q.desc_need = &d;
        mkhip_plan *tmp = nullptr;
        int st = plan_mkf_ckks_lv(c, nct, ops.data(), vals[S.a].level, vals[S.b].level, 0, &tmp, q);
        if (st != MKHIP_OK) return st;
        ws_need = std::max(ws_need, w);
        sub_desc[s] = (d + 255) & ~(size_t)255;
        sub_total += sub_desc[s];
    }
    const size_t ws_off = cv.take(std::max<size_t>(ws_need, 256));
    mkhip_plan *pl = new mkhip_plan();
    pl->ctx = c;
    pl->T = c->T;
    pl->bytes = cv.off;
    hipError_t e = hipMalloc(&pl->d_mem, std::max<size_t>(cv.off, 256));
    if (e != hipSuccess) {
        delete pl;
        return hip_status(e);
    }
    char *arena = (char *)pl->d_mem;
    // own stages (linear steps and rescales) into a host blob; product stages later
    DescBlob blob;
    std::vector<PlanStage> own;
    std::vector<int> own_at;  // step of each own stage
    Finish fin(c);
    for (int s = 0; s < ns; ++s) {
        const Step &S = steps[s];
        if (S.kind == S_LIN) {
            std::vector<LinDesc> d;
            const int olv = vals[S.out].level;
            for (int i = 0; i < nct; ++i)
                for (int comp = 0; comp <= k; ++comp)
                    for (int l = 0; l <= olv; ++l) {
                        LinDesc x = {};
                        auto row = [&](int v) {
                            return base_of(*this, v, i, arena) + ((size_t)comp * (vals[v].level + 1) + l) * N;
                        };
                        x.mod = l;
                        x.out = row(S.out);
                        if (l > S.level) {  // limbs of the target above the operand's level
                            x.op = LIN_COPY;
                            x.a = row(S.b);
                        } else {
                            x.op = (S.c0only && comp > 0) ? LIN_COPY : S.op;
                            if (S.a >= 0) x.a = row(S.a);
                            if (S.b >= 0) x.b = row(S.b);
                            if (!S.lo.empty()) {
                                x.c_lo = S.lo[l];
                                x.c_hi = S.hi[l];
                            }
                        }
                        d.push_back(x);
                    }
            own.push_back(make_stage(KLINEAR, blob.add(d), (int)d.size(), 0, 0, 1, "linear"));
            own_at.push_back(s);
        } else if (S.kind == S_RESC) {
            const int la = S.level;
            std::vector<NttJob> jobs;
            std::vector<RescaleNttDesc> rd;
            for (int i = 0; i < nct; ++i) {
                u64 *src = base_of(*this, S.a, i, arena), *dst = base_of(*this, S.out, i, arena);
                u64 *lst = base_of(*this, S.tmp, i, arena);
                for (int comp = 0; comp <= k; ++comp) {
                    NttJob j;
                    j.src = src + ((size_t)comp * (la + 1) + la) * N;
                    j.dst = lst + (size_t)comp * N;
                    j.mod = la;
                    j.reduce = 0;
                    jobs.push_back(j);
                    RescaleNttDesc r = {};
                    r.x = src + (size_t)comp * (la + 1) * N;
                    r.xlast = j.dst;
                    r.level = la;
                    r.nres = 1;
                    r.out[0] = dst + (size_t)comp * la * N;
                    r.nout = 1;
                    rd.push_back(r);
                }
            }
            own.push_back(make_stage(KNTT_INV, blob.add(jobs), (int)jobs.size(), 0, 0, 1, "rescale_intt"));
            own_at.push_back(s);
            fin.jobs(jobs);
            const size_t before = own.size();
            fin.emit(own, blob, "rescale_intt");
            for (size_t t = before; t < own.size(); ++t) own_at.push_back(s);
            own.push_back(make_stage(KRESCALE_NTT, blob.add(rd), (int)rd.size(), la, 0, 1, "rescale_div"));
            own_at.push_back(s);
        }
    }
    const size_t own_bytes = (std::max<size_t>(blob.host.size(), 256) + 255) & ~(size_t)255;
    e = hipMalloc(&pl->d_desc, own_bytes + sub_total);
    if (e != hipSuccess) {
        mkhip_plan_destroy(pl);
        return hip_status(e);
    }
    pl->host_desc = blob.host;
    pl->desc_bytes = own_bytes + sub_total;
    pl->bytes += own_bytes + sub_total;
    e = hipMemcpyAsync(pl->d_desc, pl->host_desc.data(), pl->host_desc.size(), hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) {
        mkhip_plan_destroy(pl);
        return hip_status(e);
    }
    for (PlanStage &st : own) st.descs = (const char *)pl->d_desc + reinterpret_cast<size_t>(st.descs);
    // assemble in step order
    std::vector<mkhip_plan *> subs;
    size_t sub_off = own_bytes;
    int st_status = MKHIP_OK;
    size_t oi = 0;
    for (int s = 0; s < ns && st_status == MKHIP_OK; ++s) {
        const Step &S = steps[s];
        if (S.kind == S_MUL) {
            for (int i = 0; i < nct; ++i) {
                ops[i] = mkhip_mkf_op();
                ops[i].left = base_of(*this, S.a, i, arena);
                ops[i].right = base_of(*this, S.b, i, arena);
                ops[i].out = base_of(*this, S.out, i, arena);
                ops[i].group = group;
                ops[i].level = S.level;
                ops[i].nres = 0;
            }
            PlanMem pm;
            pm.ws = arena + ws_off;
            pm.ws_bytes = std::max<size_t>(ws_need, 256);
            pm.desc = (char *)pl->d_desc + sub_off;
            pm.desc_bytes = sub_desc[s];
            pm.stream = c->stream;
            mkhip_plan *sp = nullptr;
            st_status = plan_mkf_ckks_lv(c, nct, ops.data(), vals[S.a].level, vals[S.b].level, 0, &sp, pm);
            if (st_status != MKHIP_OK) break;
            subs.push_back(sp);
            for (const PlanStage &x : sp->stages) pl->stages.push_back(x);
            sub_off += sub_desc[s];
        } else {
            while (oi < own.size() && own_at[oi] == s) pl->stages.push_back(own[oi++]);
        }
    }
    // the sub-plans' descriptor uploads read their host copies: finish them first
    hipError_t se = hipStreamSynchronize(c->stream);
    for (mkhip_plan *sp : subs) mkhip_plan_destroy(sp);
    if (st_status == MKHIP_OK && se != hipSuccess) st_status = hip_status(se);
    if (st_status == MKHIP_OK && use_graph) st_status = plan_capture(pl);
    if (st_status != MKHIP_OK) {
        mkhip_plan_destroy(pl);
        return st_status;
    }
    *out = pl;
    return MKHIP_OK;
}

int read_poly(const mkhip_polynomial *p, Poly &out) {
    if (!p || p->ncoeffs <= 0 || !p->coeffs_re) return MKHIP_ERR_ARG;
    if (p->basis != MKHIP_BASIS_STANDARD && p->basis != MKHIP_BASIS_CHEBYSHEV) return MKHIP_ERR_ARG;
    out.c.resize(p->ncoeffs);
    for (int i = 0; i < p->ncoeffs; ++i) out.c[i] = cd(p->coeffs_re[i], p->coeffs_im ? p->coeffs_im[i] : 0.0);
    out.maxdeg = p->max_deg;
    out.lead = p->lead != 0;
    out.basis = p->basis;
    return MKHIP_OK;
}

int mkf_k(mkhip_ctx *c, int group, int *k) {
    if (!c) return MKHIP_ERR_ARG;
    if (c->gamma != 1 || c->bfv) return MKHIP_ERR_SCHEME;
    auto it = c->mkf_groups.find(group);
    if (it == c->mkf_groups.end()) return MKHIP_ERR_NO_KEY;
    *k = it->second.k;
    return MKHIP_OK;
}

// Device staging for the host-buffer entry points.
struct DevBufs {
    std::vector<void *> p;
    ~DevBufs() {
        for (void *x : p) (void)hipFree(x);
    }
    u64 *alloc(size_t words) {
        void *x = nullptr;
        if (hipMalloc(&x, std::max<size_t>(words, 1) * 8) != hipSuccess) return nullptr;
        p.push_back(x);
        return (u64 *)x;
    }
};

// Build, launch once and release a program whose input / output values are the
// device buffers of `ins` / `outs` (ciphertext 0 of each).
int run_once(Prog &P) {
    mkhip_plan *pl = nullptr;
    int st = P.build(0, &pl);
    if (st != MKHIP_OK) return st;
    st = plan_launch_raw(pl);
    hipError_t e = hipStreamSynchronize(P.c->stream);
    mkhip_plan_destroy(pl);
    if (st != MKHIP_OK) return st;
    return hip_status(e);
}

}  // namespace

extern "C" int mkhip_plan_mkf_poly(mkhip_ctx *c, int group, int nct, const uint64_t *const *in, int level,
                                   double scale, const mkhip_polynomial *poly, double target_scale,
                                   uint64_t *const *out, int use_graph, mkhip_plan **plan, int *level_out,
                                   double *scale_out) {
    if (!c || nct <= 0 || !in || !out || !plan) return MKHIP_ERR_ARG;
    *plan = nullptr;
    int k = 0;
    int st = mkf_k(c, group, &k);
    if (st != MKHIP_OK) return st;
    if (level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
    if (!(scale > 0)) return MKHIP_ERR_SCALE;
    for (int i = 0; i < nct; ++i)
        if (!in[i] || !out[i]) return MKHIP_ERR_ARG;
    Poly p;
    if ((st = read_poly(poly, p)) != MKHIP_OK) return st;
    Prog P(c, group, k, nct);
    std::vector<u64 *> ip(nct);
    for (int i = 0; i < nct; ++i) ip[i] = const_cast<u64 *>(in[i]);
    Val x = P.input(level, scale, ip);
    Val r = P.evaluate(x, p, target_scale);
    if (!P.ok()) return P.status;
    const int o = P.value(r.level);
    P.vals[o].ext.assign(out, out + nct);
    P.copy_to(r, o);
    st = P.build(use_graph, plan);
    if (st != MKHIP_OK) return st;
    if (level_out) *level_out = r.level;
    if (scale_out) *scale_out = r.scale;
    return MKHIP_OK;
}

extern "C" int mkhip_mkf_poly_query(mkhip_ctx *c, int group, int level, double scale, const mkhip_polynomial *poly,
                                    double target_scale, int *level_out, double *scale_out, int *mulrelins) {
    int k = 0;
    int st = mkf_k(c, group, &k);
    if (st != MKHIP_OK) return st;
    if (level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
    if (!(scale > 0)) return MKHIP_ERR_SCALE;
    Poly p;
    if ((st = read_poly(poly, p)) != MKHIP_OK) return st;
    Prog P(c, group, k, 1);
    Val r = P.evaluate(P.input(level, scale, {nullptr}), p, target_scale);
    if (!P.ok()) return P.status;
    if (level_out) *level_out = r.level;
    if (scale_out) *scale_out = r.scale;
    if (mulrelins) *mulrelins = P.mul_calls;
    return MKHIP_OK;
}

extern "C" int mkhip_mkf_ckks_evaluate_poly(mkhip_ctx *c, int group, const uint64_t *in, int level, double scale,
                                            const mkhip_polynomial *poly, double target_scale, uint64_t *out,
                                            int *level_out, double *scale_out) {
    if (!c || !in || !out) return MKHIP_ERR_ARG;
    int lo = 0, k = 0;
    double so = 0;
    int st = mkhip_mkf_poly_query(c, group, level, scale, poly, target_scale, &lo, &so, nullptr);
    if (st != MKHIP_OK) return st;
    mkf_k(c, group, &k);
    const size_t iw = (size_t)(1 + k) * (level + 1) * c->N, ow = (size_t)(1 + k) * (lo + 1) * c->N;
    DevBufs B;
    u64 *di = B.alloc(iw), *dout = B.alloc(ow);
    if (!di || !dout) return MKHIP_ERR_OOM;
    HIP_TRY(hipMemcpyAsync(di, in, iw * 8, hipMemcpyHostToDevice, c->stream));
    mkhip_plan *pl = nullptr;
    const uint64_t *ins[1] = {di};
    uint64_t *outs[1] = {dout};
    st = mkhip_plan_mkf_poly(c, group, 1, ins, level, scale, poly, target_scale, outs, 0, &pl, &lo, &so);
    if (st != MKHIP_OK) return st;
    st = plan_launch_raw(pl);
    hipError_t e = hipMemcpyAsync(out, dout, ow * 8, hipMemcpyDeviceToHost, c->stream);
    hipError_t e2 = hipStreamSynchronize(c->stream);
    mkhip_plan_destroy(pl);
    if (st != MKHIP_OK) return st;
    HIP_TRY(e);
    HIP_TRY(e2);
    if (level_out) *level_out = lo;
    if (scale_out) *scale_out = so;
    return MKHIP_OK;
}

// ---------------------------------------------------------------------------
// public linear operations of the paper evaluator on host buffers [1+k][level+1][N]

namespace {
int lin_args(mkhip_ctx *c, int k, int level) {
    if (!c) return MKHIP_ERR_ARG;
    if (c->gamma != 1 || c->bfv) return MKHIP_ERR_SCHEME;
    if (k < 1 || k > kMaxParties) return MKHIP_ERR_ARG;
    if (level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
    return MKHIP_OK;
}

// one host operand to a fresh device value of program P
Val stage_in(Prog &P, DevBufs &B, const uint64_t *h, int level, double scale, int *st) {
    const size_t w = (size_t)(1 + P.k) * (level + 1) * P.N;
    u64 *d = B.alloc(w);
    if (!d) {
        *st = MKHIP_ERR_OOM;
        return Val();
    }
    hipError_t e = hipMemcpyAsync(d, h, w * 8, hipMemcpyHostToDevice, P.c->stream);
    if (e != hipSuccess) *st = hip_status(e);
    return P.input(level, scale, {d});
}

int finish_out(Prog &P, DevBufs &B, Val r, uint64_t *h) {
    const size_t w = (size_t)(1 + P.k) * (r.level + 1) * P.N;
    u64 *d = B.alloc(w);
    if (!d) return MKHIP_ERR_OOM;
    const int o = P.value(r.level);
    P.vals[o].ext = {d};
    P.copy_to(r, o);
    int st = run_once(P);
    if (st != MKHIP_OK) return st;
    HIP_TRY(hipMemcpy(h, d, w * 8, hipMemcpyDeviceToHost));
    return MKHIP_OK;
}
}  // namespace

extern "C" int mkhip_mkf_ckks_add(mkhip_ctx *c, int k, const uint64_t *a, int level_a, double scale_a,
                                  const uint64_t *b, int level_b, double scale_b, int sub, uint64_t *out,
                                  int *level_out, double *scale_out) {
    int st = lin_args(c, k, std::max(level_a, level_b));
    if (st != MKHIP_OK) return st;
    if (std::min(level_a, level_b) < 0) return MKHIP_ERR_LEVEL;
    if (!a || !b || !out) return MKHIP_ERR_ARG;
    if (!(scale_a > 0) || !(scale_b > 0)) return MKHIP_ERR_SCALE;
    Prog P(c, -1, k, 1);
    DevBufs B;
    Val x = stage_in(P, B, a, level_a, scale_a, &st);
    Val y = stage_in(P, B, b, level_b, scale_b, &st);
    if (st != MKHIP_OK) return st;
    Val r = P.add_sub(x, y, sub != 0);
    if ((st = finish_out(P, B, r, out)) != MKHIP_OK) return st;
    if (level_out) *level_out = r.level;
    if (scale_out) *scale_out = r.scale;
    return MKHIP_OK;
}

extern "C" int mkhip_mkf_ckks_add_const(mkhip_ctx *c, int k, const uint64_t *in, int level, double scale, double re,
                                        double im, uint64_t *out) {
    int st = lin_args(c, k, level);
    if (st != MKHIP_OK) return st;
    if (!in || !out) return MKHIP_ERR_ARG;
    if (!(scale > 0)) return MKHIP_ERR_SCALE;
    Prog P(c, -1, k, 1);
    DevBufs B;
    Val x = stage_in(P, B, in, level, scale, &st);
    if (st != MKHIP_OK) return st;
    return finish_out(P, B, P.add_const(x, re, im), out);
}

extern "C" int mkhip_mkf_ckks_mult_by_const(mkhip_ctx *c, int k, const uint64_t *in, int level, double scale,
                                            double re, double im, uint64_t *out, double *scale_out) {
    int st = lin_args(c, k, level);
    if (st != MKHIP_OK) return st;
    if (!in || !out) return MKHIP_ERR_ARG;
    if (!(scale > 0)) return MKHIP_ERR_SCALE;
    Prog P(c, -1, k, 1);
    DevBufs B;
    Val x = stage_in(P, B, in, level, scale, &st);
    if (st != MKHIP_OK) return st;
    Val r = P.mult_by_const(x, re, im);
    if ((st = finish_out(P, B, r, out)) != MKHIP_OK) return st;
    if (scale_out) *scale_out = r.scale;
    return MKHIP_OK;
}

extern "C" int mkhip_mkf_ckks_mult_by_gaussian_integer_and_add(mkhip_ctx *c, int k, const uint64_t *in,
                                                               int level_in, int64_t re, int64_t im, uint64_t *acc,
                                                               int level_acc) {
    int st = lin_args(c, k, std::max(level_in, level_acc));
    if (st != MKHIP_OK) return st;
    if (std::min(level_in, level_acc) < 0) return MKHIP_ERR_LEVEL;
    if (!in || !acc) return MKHIP_ERR_ARG;
    Prog P(c, -1, k, 1);
    DevBufs B;
    Val x = stage_in(P, B, in, level_in, 1.0, &st);
    Val t = stage_in(P, B, acc, level_acc, 1.0, &st);
    if (st != MKHIP_OK) return st;
    const int lv = std::min(level_in, level_acc);
    std::vector<u64> rr(lv + 1), ri(lv + 1);
    auto res = [](int64_t v, u64 q) {  // Euclidean residue (big.Int.Mod)
        if (v >= 0) return (u64)v % q;
        const u64 r = ((u64)(-(v + 1)) % q + 1) % q;
        return r ? q - r : (u64)0;
    };
    for (int i = 0; i <= lv; ++i) {
        rr[i] = res(re, c->Q[i]);
        ri[i] = res(im, c->Q[i]);
    }
    return finish_out(P, B, P.gauss_add(x, rr, ri, t), acc);
}

extern "C" int mkhip_mkf_ckks_rescale(mkhip_ctx *c, int k, const uint64_t *in, int level, double scale,
                                      double min_scale, uint64_t *out, int *level_out, double *scale_out) {
    int st = lin_args(c, k, level);
    if (st != MKHIP_OK) return st;
    if (!in || !out) return MKHIP_ERR_ARG;
    Prog P(c, -1, k, 1);
    DevBufs B;
    Val x = stage_in(P, B, in, level, scale, &st);
    if (st != MKHIP_OK) return st;
    Val r = P.rescale(x, min_scale);
    if (!P.ok()) return P.status;
    if ((st = finish_out(P, B, r, out)) != MKHIP_OK) return st;
    if (level_out) *level_out = r.level;
    if (scale_out) *scale_out = r.scale;
    return MKHIP_OK;
}

extern "C" int mkhip_mkf_ckks_mulrelin_levels(mkhip_ctx *c, int group, const uint64_t *left, int level_l,
                                              const uint64_t *right, int level_r, uint64_t *out) {
    int k = 0;
    int st = mkf_k(c, group, &k);
    if (st != MKHIP_OK) return st;
    if (!left || !right || !out) return MKHIP_ERR_ARG;
    if (level_l < 0 || level_r < 0 || level_l >= c->L || level_r >= c->L) return MKHIP_ERR_LEVEL;
    Prog P(c, group, k, 1);
    DevBufs B;
    Val x = stage_in(P, B, left, level_l, 1.0, &st);
    Val y = stage_in(P, B, right, level_r, 1.0, &st);
    if (st != MKHIP_OK) return st;
    return finish_out(P, B, P.mulrelin(x, y), out);
}
