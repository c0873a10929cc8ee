// Key generation and public-key encryption on the device (the step before the
// mul+relin path, feeding party registration): software/kklss/mkrlwe/keygen.go
// and encryptor.go.  The reference draws secrets, errors and the encryption's v
// from its PRNG inside these calls; here they are inputs (small signed
// polynomials), so every output is a deterministic function of its arguments.
// Keys are full-level ([beta][L+K][N] words, NTT + Montgomery form), secrets
// [L+K][N] (MForm(NTT(s))), as the reference stores them.
#include "mkhip_internal.hpp"

#include <algorithm>

using namespace mkhip;

namespace {

struct Scratch {
    void *p = nullptr;
    ~Scratch() {
        if (p) (void)hipFree(p);
    }
};

// Device copy of a host descriptor table; the launch runs on the context stream,
// which is synchronized before the table is released.
template <class D>
int run(mkhip_ctx *c, int what, const std::vector<D> &d) {
    if (d.empty()) return MKHIP_OK;
    Scratch s;
    HIP_TRY(hipMalloc(&s.p, d.size() * sizeof(D)));
    HIP_TRY(hipMemcpyAsync(s.p, d.data(), d.size() * sizeof(D), hipMemcpyHostToDevice, c->stream));
    const int n = (int)d.size();
    hipError_t e;
    if (what == KSMALL_NTT || what == KKEYOP)
        e = launch_keygen(what, c->T, s.p, n, c->stream);
    else if (what == KAUTO)
        e = launch_galois(what, c->T, s.p, n, c->stream);
    else
        e = launch_kernel(what, c->T, s.p, n, 0, 0, c->stream);
    HIP_TRY(e);
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

// P mod q for every ring modulus of Q u P
std::vector<u64> p_mod(const mkhip_ctx *c) {
    std::vector<u64> QP(c->Q);
    QP.insert(QP.end(), c->P.begin(), c->P.end());
    std::vector<u64> out(QP.size());
    for (size_t r = 0; r < QP.size(); ++r) {
        u64 p = 1 % QP[r];
        for (u64 pk : c->P) p = h_mulmod(p, pk % QP[r], QP[r]);
        out[r] = p;
    }
    return out;
}

u64 gal_pow(u64 b, u64 e, u64 mask) {
    u64 r = 1;
    b &= mask;
    for (; e; e >>= 1, b = (b * b) & mask)
        if (e & 1) r = (r * b) & mask;
    return r;
}

// small polynomials -> MForm(NTT) (or NTT) over rows 0..nrows-1 of each output
void small_ntt_jobs(std::vector<SmallNttDesc> &v, const int64_t *src, u64 *dst, int nrows, int N, int mform) {
    for (int r = 0; r < nrows; ++r) {
        SmallNttDesc d;
        d.src = src;
        d.dst = dst + (size_t)r * N;
        d.mod = r;
        d.mform = mform;
        v.push_back(d);
    }
}

// Is QP row rr one of digit i's Q limbs (GenSwitchingKey's gadget rows, keygen.go:311-324)?
bool digit_row(const mkhip_ctx *c, int i, int rr) { return rr < c->L && rr / c->alpha == i; }

// Switching keys over the context's full level: for each of nkeys keys, digit i and
// row rr, out = MForm(NTT(e)) [+ g G_i,rr] -/+ MRed(ka_i, kb) (op), where the
// gadget multiplier G_i,rr is P mod q_rr on digit i's Q rows (GenSwitchingKey,
// keygen.go:270-327) or, with gm, gm[i * (L+K) + rr] (0: no gadget term).
struct KeySpec {
    const u64 *g;    // gadget secret (device, [L+K][N]) or null
    const u64 *ka;   // CRS-like key [beta][L+K][N]
    const u64 *kb;   // secret [L+K][N]
    u64 *out;        // [beta][L+K][N], digit stride ostride words when set
    int op;
    const u64 *gm = nullptr;  // host [beta][L+K] gadget multipliers, or null
    size_t ostride = 0;
};

int gen_keys(mkhip_ctx *c, const int64_t *errs_host, const std::vector<KeySpec> &keys) {
    const int N = c->N, R = c->L + c->K, beta = c->beta_max;
    const size_t row = (size_t)N;
    const size_t ne = keys.size() * beta;
    Scratch s;
    HIP_TRY(hipMalloc(&s.p, ne * N * sizeof(int64_t) + ne * R * row * 8));
    int64_t *derr = (int64_t *)s.p;
    u64 *E = (u64 *)(derr + ne * N);
    HIP_TRY(hipMemcpyAsync(derr, errs_host, ne * N * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    std::vector<SmallNttDesc> sd;
    for (size_t k = 0; k < ne; ++k) small_ntt_jobs(sd, derr + k * N, E + k * R * row, R, N, 1);
    int st = run(c, KSMALL_NTT, sd);
    if (st != MKHIP_OK) return st;
    const std::vector<u64> pm = p_mod(c);
    std::vector<KeyOpDesc> ops;
    for (size_t k = 0; k < keys.size(); ++k)
        for (int i = 0; i < beta; ++i)
            for (int rr = 0; rr < R; ++rr) {
                const KeySpec &ks = keys[k];
                KeyOpDesc d = {};
                const size_t o = ((size_t)i * R + rr) * row;
                const size_t oo = (ks.ostride ? (size_t)i * ks.ostride : (size_t)i * R * row) + (size_t)rr * row;
                const u64 gmul = ks.gm ? ks.gm[(size_t)i * R + rr] : (digit_row(c, i, rr) ? pm[rr] : 0);
                d.x = E + (k * beta + i) * R * row + (size_t)rr * row;
                d.g = ks.g && gmul ? ks.g + (size_t)rr * row : nullptr;
                d.pmod = gmul;
                d.ka = ks.ka + o;
                d.kb = ks.kb + (size_t)rr * row;
                d.out = ks.out + oo;
                d.mod = rr;
                d.op = ks.op;
                ops.push_back(d);
            }
    return run(c, KKEYOP, ops);
}

int check_ctx(const mkhip_ctx *c) {
    if (!c) return MKHIP_ERR_ARG;
    if (c->alpha == 0 || c->K == 0) return MKHIP_ERR_GADGET;
    return MKHIP_OK;
}

}  // namespace

extern "C" int mkhip_gen_secret(mkhip_ctx *c, const int64_t *s, uint64_t *sk) {
    int st = check_ctx(c);
    if (st != MKHIP_OK) return st;
    if (!s || !sk) return MKHIP_ERR_ARG;
    const int N = c->N, R = c->L + c->K;
    Scratch m;
    HIP_TRY(hipMalloc(&m.p, (size_t)N * 8 + (size_t)R * N * 8));
    int64_t *ds = (int64_t *)m.p;
    u64 *dk = (u64 *)(ds + N);
    HIP_TRY(hipMemcpyAsync(ds, s, (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
    std::vector<SmallNttDesc> sd;
    small_ntt_jobs(sd, ds, dk, R, N, 1);
    if ((st = run(c, KSMALL_NTT, sd)) != MKHIP_OK) return st;
    HIP_TRY(hipMemcpyAsync(sk, dk, (size_t)R * N * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

extern "C" int mkhip_gen_public_key(mkhip_ctx *c, const uint64_t *sk, const int64_t *e, const uint64_t *a0,
                                    uint64_t *pk0, uint64_t *pk1) {
    int st = check_ctx(c);
    if (st != MKHIP_OK) return st;
    if (!sk || !e || !a0 || !pk0 || !pk1) return MKHIP_ERR_ARG;
    const int N = c->N, R = c->L + c->K;
    const size_t w = (size_t)R * N;
    Scratch m;
    HIP_TRY(hipMalloc(&m.p, (size_t)N * 8 + 4 * w * 8));
    int64_t *de = (int64_t *)m.p;
    u64 *dsk = (u64 *)(de + N), *da = dsk + w, *dt = da + w, *dout = dt + w;
    HIP_TRY(hipMemcpyAsync(de, e, (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dsk, sk, w * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(da, a0, w * 8, hipMemcpyHostToDevice, c->stream));
    std::vector<SmallNttDesc> sd;
    small_ntt_jobs(sd, de, dt, R, N, 0);
    if ((st = run(c, KSMALL_NTT, sd)) != MKHIP_OK) return st;
    std::vector<KeyOpDesc> ops;
    for (int r = 0; r < R; ++r) {
        KeyOpDesc d = {};
        d.x = dt + (size_t)r * N;
        d.ka = dsk + (size_t)r * N;
        d.kb = da + (size_t)r * N;
        d.out = dout + (size_t)r * N;
        d.mod = r;
        d.op = 0;
        ops.push_back(d);
    }
    if ((st = run(c, KKEYOP, ops)) != MKHIP_OK) return st;
    HIP_TRY(hipMemcpyAsync(pk0, dout, w * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::copy(a0, a0 + w, pk1);
    return MKHIP_OK;
}

namespace {

// b, d, v of GenRelinearizationKey into device buffers (kw words each)
int relin_on_device(mkhip_ctx *c, const uint64_t *sk, const uint64_t *r, const int64_t *errs, u64 *b, u64 *d,
                    u64 *v) {
    auto ai = c->crs.find(0);
    if (ai == c->crs.end() || !c->d_crs_u) return MKHIP_ERR_NO_KEY;
    const int N = c->N, R = c->L + c->K;
    const size_t w = (size_t)R * N;
    Scratch m;
    HIP_TRY(hipMalloc(&m.p, 2 * w * 8));
    u64 *dsk = (u64 *)m.p, *dr = dsk + w;
    HIP_TRY(hipMemcpyAsync(dsk, sk, w * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dr, r, w * 8, hipMemcpyHostToDevice, c->stream));
    std::vector<KeySpec> keys(3);
    keys[0] = KeySpec{nullptr, ai->second, dsk, b, 0};       // b = MForm(NTT e) - MRed(a, s)
    keys[1] = KeySpec{dsk, ai->second, dr, d, 0};            // d = swk(s) - MRed(a, r)
    keys[2] = KeySpec{dr, c->d_crs_u, dsk, v, 1};            // v = -(swk(r) + MRed(u, s))
    return gen_keys(c, errs, keys);
}

}  // namespace

extern "C" int mkhip_gen_relin_key(mkhip_ctx *c, const uint64_t *sk, const uint64_t *r, const int64_t *errs,
                                   uint64_t *b, uint64_t *d, uint64_t *v) {
    int st = check_ctx(c);
    if (st != MKHIP_OK) return st;
    if (!sk || !r || !errs || !b || !d || !v) return MKHIP_ERR_ARG;
    const size_t kw = c->swk_words;
    Scratch m;
    HIP_TRY(hipMalloc(&m.p, 3 * kw * 8));
    u64 *db = (u64 *)m.p, *dd = db + kw, *dv = dd + kw;
    if ((st = relin_on_device(c, sk, r, errs, db, dd, dv)) != MKHIP_OK) return st;
    HIP_TRY(hipMemcpyAsync(b, db, kw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(d, dd, kw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(v, dv, kw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

extern "C" int mkhip_register_party_keygen(mkhip_ctx *c, int party, const uint64_t *sk, const uint64_t *r,
                                           const int64_t *errs) {
    int st = check_ctx(c);
    if (st != MKHIP_OK) return st;
    if (!sk || !r || !errs || party < 0) return MKHIP_ERR_ARG;
    if (c->bfv) return MKHIP_ERR_SCHEME;
    const size_t bytes = c->swk_words * 8;
    PartyKeys k;
    hipError_t e = hipMalloc(&k.b, bytes);
    if (e == hipSuccess) e = hipMalloc(&k.d, bytes);
    if (e == hipSuccess) e = hipMalloc(&k.v, bytes);
    if (e == hipSuccess) {
        st = relin_on_device(c, sk, r, errs, k.b, k.d, k.v);
    } else {
        st = hip_status(e);
    }
    if (st != MKHIP_OK) {
        if (k.b) (void)hipFree(k.b);
        if (k.d) (void)hipFree(k.d);
        if (k.v) (void)hipFree(k.v);
        return st;
    }
    if (c->parties.count(party)) mkhip_unregister_party(c, party);
    c->parties[party] = k;
    return MKHIP_OK;
}

namespace {

// permuted secret (NTT slots) into a device buffer
int permute_secret(mkhip_ctx *c, const u64 *dsk, u64 gal, u64 *out) {
    AutoDesc a = {};
    a.src = dsk;
    a.dst = out;
    a.gal = gal;
    a.gal_inv = gal;  // unused in the NTT domain
    a.nl = c->L + c->K;
    a.ntt = 1;
    return run(c, KAUTO, std::vector<AutoDesc>{a});
}

int galois_key(mkhip_ctx *c, const uint64_t *sk, const int64_t *errs, int crs_index, u64 gal, bool conj,
               uint64_t *out) {
    auto ai = c->crs.find(crs_index);
    if (ai == c->crs.end()) return MKHIP_ERR_NO_KEY;
    const int N = c->N, R = c->L + c->K;
    const size_t w = (size_t)R * N, kw = c->swk_words;
    Scratch m;
    HIP_TRY(hipMalloc(&m.p, (2 * w + kw) * 8));
    u64 *dsk = (u64 *)m.p, *dperm = dsk + w, *dout = dperm + w;
    HIP_TRY(hipMemcpyAsync(dsk, sk, w * 8, hipMemcpyHostToDevice, c->stream));
    int st = permute_secret(c, dsk, gal, dperm);
    if (st != MKHIP_OK) return st;
    // rotation (keygen.go:213-228): gadget of s, minus a * sigma^-1(s);
    // conjugation (:250-266): gadget of sigma(s), minus a * s
    std::vector<KeySpec> keys(1);
    keys[0] = conj ? KeySpec{dperm, ai->second, dsk, dout, 0} : KeySpec{dsk, ai->second, dperm, dout, 0};
    if ((st = gen_keys(c, errs, keys)) != MKHIP_OK) return st;
    HIP_TRY(hipMemcpyAsync(out, dout, kw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

}  // namespace

extern "C" int mkhip_gen_rotation_key(mkhip_ctx *c, int rotidx, const uint64_t *sk, const int64_t *errs,
                                      uint64_t *rk) {
    int st = check_ctx(c);
    if (st != MKHIP_OK) return st;
    if (!sk || !errs || !rk) return MKHIP_ERR_ARG;
    const int half = c->N / 2;
    const int r = ((rotidx % half) + half) % half;
    if (r == 0) return MKHIP_ERR_ARG;
    const u64 mask = 2 * (u64)c->N - 1;
    const u64 g = gal_pow(5, (u64)r, mask);
    const u64 ginv = gal_pow(g, mask, mask);  // InverseGaloisElement: g^(2N-1)
    return galois_key(c, sk, errs, r, ginv, false, rk);
}

extern "C" int mkhip_gen_conjugation_key(mkhip_ctx *c, const uint64_t *sk, const int64_t *errs, uint64_t *ck) {
    int st = check_ctx(c);
    if (st != MKHIP_OK) return st;
    if (!sk || !errs || !ck) return MKHIP_ERR_ARG;
    return galois_key(c, sk, errs, -2, 2 * (u64)c->N - 1, true, ck);
}

// Encryptor.Encrypt (encryptor.go:55-119) at `level`; pk0 / pk1 full-level QP
// (their Q limbs are used); pt level+1 limbs.
extern "C" int mkhip_encrypt(mkhip_ctx *c, const uint64_t *pk0, const uint64_t *pk1, int level, const int64_t *v,
                             const int64_t *e0, const int64_t *e1, const uint64_t *pt, int pt_ntt, int ntt_out,
                             uint64_t *c0, uint64_t *c1) {
    if (!c) return MKHIP_ERR_ARG;
    if (!pk0 || !pk1 || !v || !e0 || !e1 || !pt || !c0 || !c1) return MKHIP_ERR_ARG;
    if (level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
    const int N = c->N, nl = level + 1;
    const size_t pw = (size_t)nl * N;
    Scratch m;
    HIP_TRY(hipMalloc(&m.p, 3 * (size_t)N * 8 + 9 * pw * 8));
    int64_t *dv = (int64_t *)m.p, *de0 = dv + N, *de1 = de0 + N;
    u64 *dp0 = (u64 *)(de1 + N), *dp1 = dp0 + pw, *dpt = dp1 + pw, *vm = dpt + pw, *t0 = vm + pw, *t1 = t0 + pw,
        *t2 = t1 + pw, *o0 = t2 + pw, *o1 = o0 + pw;
    HIP_TRY(hipMemcpyAsync(dv, v, (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(de0, e0, (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(de1, e1, (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dp0, pk0, pw * 8, hipMemcpyHostToDevice, c->stream));  // Q limbs lead a QP poly
    HIP_TRY(hipMemcpyAsync(dp1, pk1, pw * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dpt, pt, pw * 8, hipMemcpyHostToDevice, c->stream));
    std::vector<int> mods(nl);
    for (int j = 0; j < nl; ++j) mods[j] = j;
    // v -> MForm(NTT(v)) (:70-73)
    std::vector<SmallNttDesc> sd;
    small_ntt_jobs(sd, dv, vm, nl, N, 1);
    int st = run(c, KSMALL_NTT, sd);
    if (st != MKHIP_OK) return st;
    auto mul = [&](u64 *dst, const u64 *pk) {  // dst = MRed(v, pk) (:76-78)
        std::vector<KeyOpDesc> ops;
        for (int j = 0; j < nl; ++j) {
            KeyOpDesc d = {};
            d.ka = vm + (size_t)j * N;
            d.kb = pk + (size_t)j * N;
            d.out = dst + (size_t)j * N;
            d.mod = j;
            d.op = 2;
            ops.push_back(d);
        }
        return run(c, KKEYOP, ops);
    };
    auto ntt = [&](u64 *src, u64 *dst, bool inverse) {
        std::vector<NttJob> jobs;
        for (int j = 0; j < nl; ++j) {
            NttJob jb;
            jb.src = src + (size_t)j * N;
            jb.dst = dst + (size_t)j * N;
            jb.mod = j;
            jb.reduce = 0;
            jobs.push_back(jb);
        }
        int s2 = run(c, inverse ? KNTT_INV : KNTT_FWD, jobs);
        if (s2 == MKHIP_OK && inverse) s2 = finish_rows(c, dst, mods.data(), nl, 0);
        return s2;
    };
    auto add = [&](u64 *dst, const u64 *x, const u64 *y, const int64_t *e) {
        std::vector<KeyOpDesc> ops;
        for (int j = 0; j < nl; ++j) {
            KeyOpDesc d = {};
            d.x = x ? x + (size_t)j * N : nullptr;
            d.y = y ? y + (size_t)j * N : nullptr;
            d.e = e;
            d.out = dst + (size_t)j * N;
            d.mod = j;
            d.op = 2;
            ops.push_back(d);
        }
        return run(c, KKEYOP, ops);
    };
    if ((st = mul(t0, dp0)) != MKHIP_OK || (st = mul(t1, dp1)) != MKHIP_OK) return st;
    if (ntt_out) {  // (:80-100)
        std::vector<SmallNttDesc> e1d;
        small_ntt_jobs(e1d, de1, t2, nl, N, 0);
        if ((st = run(c, KSMALL_NTT, e1d)) != MKHIP_OK) return st;
        if ((st = add(o1, t1, t2, nullptr)) != MKHIP_OK) return st;
        if (!pt_ntt) {  // NTT(e0 + pt)
            if ((st = add(vm, dpt, nullptr, de0)) != MKHIP_OK) return st;
            if ((st = ntt(vm, t2, false)) != MKHIP_OK) return st;
            if ((st = add(o0, t0, t2, nullptr)) != MKHIP_OK) return st;
        } else {
            std::vector<SmallNttDesc> e0d;
            small_ntt_jobs(e0d, de0, t2, nl, N, 0);
            if ((st = run(c, KSMALL_NTT, e0d)) != MKHIP_OK) return st;
            if ((st = add(vm, t0, t2, nullptr)) != MKHIP_OK) return st;
            if ((st = add(o0, vm, dpt, nullptr)) != MKHIP_OK) return st;
        }
    } else {  // (:102-117)
        if ((st = ntt(t0, o0, true)) != MKHIP_OK || (st = ntt(t1, o1, true)) != MKHIP_OK) return st;
        if (pt_ntt) {
            if ((st = ntt(dpt, t2, true)) != MKHIP_OK) return st;
        } else {
            HIP_TRY(hipMemcpyAsync(t2, dpt, pw * 8, hipMemcpyDeviceToDevice, c->stream));
        }
        if ((st = add(o0, o0, t2, de0)) != MKHIP_OK || (st = add(o1, o1, nullptr, de1)) != MKHIP_OK) return st;
    }
    HIP_TRY(hipMemcpyAsync(c0, o0, pw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(c1, o1, pw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

// ---------------------------------------------------------------------------
// KKLSS MK-BFV relinearization keys: two gadget branches (mkbfv/keygen.go:24-162)

namespace {

// The BFV gadget multipliers G_br,i mod q_r of GenBFVSwitchingKey (mkbfv/keygen.go:
// 91-162) for one prime per digit (the KKLSS MK-BFV shape), over the rows of Q u P:
//   branch 1, digit i (prime q_i of Q):  G = (Q QMul / q_i) t T P / QMul with
//     T = (Q QMul / q_i)^-1 mod q_i; (Q QMul / q_i) / QMul = Q / q_i, so
//     G = (Q / q_i) t T P exactly: 0 mod every other prime of Q u P, and
//     t P QMul^-1 mod q_i on row i;
//   branch 2, digit i (prime m = qm_i of QMul):  G = floor(Y / m) with
//     Y = Q t T' P, T' = (Q QMul / m)^-1 mod m; Y = 0 mod every prime of Q u P,
//     so G = (0 - (Y mod m)) m^-1 mod q_r on every row.
std::vector<u64> bfv_gadget(const mkhip_ctx *c, int branch) {
    const int L = c->L, R = c->L + c->K;
    std::vector<u64> QP(c->Q);
    QP.insert(QP.end(), c->P.begin(), c->P.end());
    auto prod_mod = [](const std::vector<u64> &v, u64 m, int skip) {
        u64 r = 1 % m;
        for (int j = 0; j < (int)v.size(); ++j)
            if (j != skip) r = h_mulmod(r, v[j] % m, m);
        return r;
    };
    std::vector<u64> gm((size_t)L * R, 0);
    for (int i = 0; i < L; ++i) {
        if (branch == 0) {
            const u64 q = c->Q[i];
            const u64 g = h_mulmod(h_mulmod(c->t % q, prod_mod(c->P, q, -1), q), h_inv(prod_mod(c->QMul, q, -1), q), q);
            gm[(size_t)i * R + i] = g;
        } else {
            const u64 m = c->QMul[i];
            // T' = (Q QMul / m)^-1 mod m ; Y mod m = Q t T' P mod m
            const u64 tp = h_inv(h_mulmod(prod_mod(c->Q, m, -1), prod_mod(c->QMul, m, i), m), m);
            const u64 ym = h_mulmod(h_mulmod(h_mulmod(prod_mod(c->Q, m, -1), c->t % m, m), tp, m),
                                    prod_mod(c->P, m, -1), m);
            for (int r = 0; r < R; ++r) {
                const u64 q = QP[r];
                const u64 neg = (q - ym % q) % q;
                gm[(size_t)i * R + r] = h_mulmod(neg, h_inv(m % q, q), q);
            }
        }
    }
    return gm;
}

// b = [b_1; b_2], d = [d_1; d_2] (2 beta digits each, branch 1 then 2), v (beta
// digits) of GenRelinearizationKey into device buffers
int bfv_relin_on_device(mkhip_ctx *c, const uint64_t *sk, const uint64_t *r, const int64_t *errs, u64 *b, u64 *d,
                        u64 *v) {
    auto a1 = c->crs.find(0), a2 = c->crs.find(-3);
    if (a1 == c->crs.end() || a2 == c->crs.end() || !c->d_crs_u) return MKHIP_ERR_NO_KEY;
    const int N = c->N, R = c->L + c->K;
    const size_t w = (size_t)R * N, kw = c->swk_words;
    Scratch m;
    HIP_TRY(hipMalloc(&m.p, 2 * w * 8));
    u64 *dsk = (u64 *)m.p, *dr = dsk + w;
    HIP_TRY(hipMemcpyAsync(dsk, sk, w * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dr, r, w * 8, hipMemcpyHostToDevice, c->stream));
    const std::vector<u64> g1 = bfv_gadget(c, 0), g2 = bfv_gadget(c, 1);
    std::vector<KeySpec> keys(5);
    keys[0] = KeySpec{nullptr, a1->second, dsk, b, 0};             // b_1 = MForm(e) - MRed(a_1, s)  (:46-55)
    keys[1] = KeySpec{nullptr, a2->second, dsk, b + kw, 0};        // b_2 with a_2 = CRS[-3]      (:57-62)
    keys[2] = KeySpec{dsk, a1->second, dr, d, 0, g1.data()};       // d_1 = MForm(G_1 s + e) - MRed(a_1, r)
    keys[3] = KeySpec{dsk, a2->second, dr, d + kw, 0, g2.data()};  // d_2 (:65-73)
    keys[4] = KeySpec{dr, c->d_crs_u, dsk, v, 1};                  // v = -(swk(r) + MRed(u, s))  (:75-85)
    return gen_keys(c, errs, keys);
}

int check_bfv(const mkhip_ctx *c) {
    int st = check_ctx(c);
    if (st != MKHIP_OK) return st;
    if (!c->bfv || c->gamma != 2 || c->alpha != 1) return MKHIP_ERR_SCHEME;
    return MKHIP_OK;
}

}  // namespace

extern "C" int mkhip_gen_bfv_relin_key(mkhip_ctx *c, const uint64_t *sk, const uint64_t *r, const int64_t *errs,
                                       uint64_t *b, uint64_t *d, uint64_t *v) {
    int st = check_bfv(c);
    if (st != MKHIP_OK) return st;
    if (!sk || !r || !errs || !b || !d || !v) return MKHIP_ERR_ARG;
    const size_t kw = c->swk_words;
    Scratch m;
    HIP_TRY(hipMalloc(&m.p, 5 * kw * 8));
    u64 *db = (u64 *)m.p, *dd = db + 2 * kw, *dv = dd + 2 * kw;
    if ((st = bfv_relin_on_device(c, sk, r, errs, db, dd, dv)) != MKHIP_OK) return st;
    HIP_TRY(hipMemcpyAsync(b, db, 2 * kw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(d, dd, 2 * kw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(v, dv, kw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

extern "C" int mkhip_register_bfv_party_keygen(mkhip_ctx *c, int party, const uint64_t *sk, const uint64_t *r,
                                               const int64_t *errs) {
    int st = check_bfv(c);
    if (st != MKHIP_OK) return st;
    if (!sk || !r || !errs || party < 0) return MKHIP_ERR_ARG;
    const size_t bytes = c->swk_words * 8;
    PartyKeys k;
    k.bfv = true;
    hipError_t e = hipMalloc(&k.b, 2 * bytes);
    if (e == hipSuccess) e = hipMalloc(&k.d, 2 * bytes);
    if (e == hipSuccess) e = hipMalloc(&k.v, bytes);
    st = e == hipSuccess ? bfv_relin_on_device(c, sk, r, errs, k.b, k.d, k.v) : hip_status(e);
    if (st != MKHIP_OK) {
        if (k.b) (void)hipFree(k.b);
        if (k.d) (void)hipFree(k.d);
        if (k.v) (void)hipFree(k.v);
        return st;
    }
    if (c->parties.count(party)) mkhip_unregister_party(c, party);
    c->parties[party] = k;
    return MKHIP_OK;
}

// ---------------------------------------------------------------------------
// MKF keys (mkf_cpu/common.go:32-51): Lattigo switching keys s_in -> s_out,
// [beta][2][L+K][N] = (MForm(e + P s_in g_i) - MRed(a_i, s_out), a_i), a_i uniform
// (sampled as already in NTT and Montgomery form); the relinearization key is the
// switching key s^2 -> s; the fold keys switch each party's s_i to s = sum s_i.

namespace {

// swk of s_in -> s_out (device secrets) into dst (device, mkf_swk_words); a and
// errs are device / host inputs
int mkf_swk_on_device(mkhip_ctx *c, const u64 *dsin, const u64 *dsout, const u64 *da, const int64_t *errs, u64 *dst) {
    const int R = c->L + c->K, N = c->N, beta = c->beta_max;
    const size_t w = (size_t)R * N;
    std::vector<KeySpec> keys(1);
    keys[0] = KeySpec{dsin, da, dsout, dst, 0, nullptr, 2 * w};
    int st = gen_keys(c, errs, keys);
    if (st != MKHIP_OK) return st;
    for (int i = 0; i < beta; ++i)  // the second component is a itself
        HIP_TRY(hipMemcpyAsync(dst + (2 * (size_t)i + 1) * w, da + (size_t)i * w, w * 8, hipMemcpyDeviceToDevice,
                               c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

// out = x + y (+ MRed(ka, kb)) row by row over Q u P, device buffers
int keyop_rows(mkhip_ctx *c, const u64 *x, const u64 *y, const u64 *ka, const u64 *kb, u64 *out) {
    const int R = c->L + c->K, N = c->N;
    std::vector<KeyOpDesc> ops;
    for (int r = 0; r < R; ++r) {
        KeyOpDesc d = {};
        d.x = x ? x + (size_t)r * N : nullptr;
        d.y = y ? y + (size_t)r * N : nullptr;
        d.ka = ka ? ka + (size_t)r * N : nullptr;
        d.kb = kb ? kb + (size_t)r * N : nullptr;
        d.out = out + (size_t)r * N;
        d.mod = r;
        d.op = 2;
        ops.push_back(d);
    }
    return run(c, KKEYOP, ops);
}

int check_mkf(const mkhip_ctx *c) {
    int st = check_ctx(c);
    if (st != MKHIP_OK) return st;
    if (c->gamma != 1) return MKHIP_ERR_SCHEME;
    return MKHIP_OK;
}

}  // namespace

extern "C" int mkhip_gen_mkf_switching_key(mkhip_ctx *c, const uint64_t *s_in, const uint64_t *s_out,
                                           const uint64_t *a, const int64_t *errs, uint64_t *swk) {
    int st = check_mkf(c);
    if (st != MKHIP_OK) return st;
    if (!s_in || !s_out || !a || !errs || !swk) return MKHIP_ERR_ARG;
    const size_t w = (size_t)(c->L + c->K) * c->N, kw = c->swk_words;
    Scratch m;
    HIP_TRY(hipMalloc(&m.p, (2 * w + kw + c->mkf_swk_words) * 8));
    u64 *din = (u64 *)m.p, *dout = din + w, *da = dout + w, *dk = da + kw;
    HIP_TRY(hipMemcpyAsync(din, s_in, w * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dout, s_out, w * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(da, a, kw * 8, hipMemcpyHostToDevice, c->stream));
    if ((st = mkf_swk_on_device(c, din, dout, da, errs, dk)) != MKHIP_OK) return st;
    HIP_TRY(hipMemcpyAsync(swk, dk, c->mkf_swk_words * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

extern "C" int mkhip_gen_mkf_relin_key(mkhip_ctx *c, const uint64_t *s, const uint64_t *a, const int64_t *errs,
                                       uint64_t *swk) {
    int st = check_mkf(c);
    if (st != MKHIP_OK) return st;
    if (!s || !a || !errs || !swk) return MKHIP_ERR_ARG;
    const size_t w = (size_t)(c->L + c->K) * c->N, kw = c->swk_words;
    Scratch m;
    HIP_TRY(hipMalloc(&m.p, (2 * w + kw + c->mkf_swk_words) * 8));
    u64 *ds = (u64 *)m.p, *ds2 = ds + w, *da = ds2 + w, *dk = da + kw;
    HIP_TRY(hipMemcpyAsync(ds, s, w * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(da, a, kw * 8, hipMemcpyHostToDevice, c->stream));
    if ((st = keyop_rows(c, nullptr, nullptr, ds, ds, ds2)) != MKHIP_OK) return st;  // MForm(s^2)
    if ((st = mkf_swk_on_device(c, ds2, ds, da, errs, dk)) != MKHIP_OK) return st;
    HIP_TRY(hipMemcpyAsync(swk, dk, c->mkf_swk_words * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

extern "C" int mkhip_register_mkf_group_keygen(mkhip_ctx *c, int group, int k, const uint64_t *sks,
                                               const uint64_t *a, const int64_t *errs) {
    int st = check_mkf(c);
    if (st != MKHIP_OK) return st;
    if (!sks || !a || !errs || k <= 0 || group < 0) return MKHIP_ERR_ARG;
    if (k > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
    const int N = c->N, beta = c->beta_max;
    const size_t w = (size_t)(c->L + c->K) * N, kw = c->swk_words, mw = c->mkf_swk_words;
    Scratch m;
    HIP_TRY(hipMalloc(&m.p, ((size_t)k * w + 2 * w + (size_t)(k + 1) * kw) * 8));
    u64 *dsk = (u64 *)m.p, *dagg = dsk + (size_t)k * w, *ds2 = dagg + w, *da = ds2 + w;
    HIP_TRY(hipMemcpyAsync(dsk, sks, (size_t)k * w * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(da, a, (size_t)(k + 1) * kw * 8, hipMemcpyHostToDevice, c->stream));
    // aggregateSecret (common.go:32-42): s = sum s_i over Q u P
    if ((st = keyop_rows(c, dsk, nullptr, nullptr, nullptr, dagg)) != MKHIP_OK) return st;
    for (int i = 1; i < k; ++i)
        if ((st = keyop_rows(c, dagg, dsk + (size_t)i * w, nullptr, nullptr, dagg)) != MKHIP_OK) return st;
    if ((st = keyop_rows(c, nullptr, nullptr, dagg, dagg, ds2)) != MKHIP_OK) return st;  // MForm(s^2)
    MkfGroup g;
    g.k = k;
    for (int i = 0; i <= k && st == MKHIP_OK; ++i) {
        void *p = nullptr;
        const hipError_t e = hipMalloc(&p, mw * 8);
        if (e != hipSuccess) {
            st = hip_status(e);
            break;
        }
        if (i < k)
            g.fold.push_back((u64 *)p);
        else
            g.relin = (u64 *)p;
        // generateOnlineKeys (common.go:44-51): fold keys s_i -> s, then the relinearization key s^2 -> s
        st = mkf_swk_on_device(c, i < k ? dsk + (size_t)i * w : ds2, dagg, da + (size_t)i * kw,
                               errs + (size_t)i * beta * N, (u64 *)p);
    }
    if (st != MKHIP_OK) {
        for (u64 *p : g.fold) (void)hipFree(p);
        if (g.relin) (void)hipFree(g.relin);
        return st;
    }
    mkhip_unregister_mkf_group(c, group);
    c->mkf_groups[group] = g;
    return MKHIP_OK;
}
