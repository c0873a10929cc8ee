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
// row rr, out = MForm(NTT(e)) [+ g P on the digit rows] -/+ MRed(ka_i, kb) (op).
struct KeySpec {
    const u64 *g;    // gadget secret (device, [L+K][N]) or null
    const u64 *ka;   // CRS-like key [beta][L+K][N]
    const u64 *kb;   // secret [L+K][N]
    u64 *out;        // [beta][L+K][N]
    int op;
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
                KeyOpDesc d = {};
                const size_t o = ((size_t)i * R + rr) * row;
                d.x = E + (k * beta + i) * R * row + (size_t)rr * row;
                d.g = keys[k].g && digit_row(c, i, rr) ? keys[k].g + (size_t)rr * row : nullptr;
                d.pmod = pm[rr];
                d.ka = keys[k].ka + o;
                d.kb = keys[k].kb + (size_t)rr * row;
                d.out = keys[k].out + o;
                d.mod = rr;
                d.op = keys[k].op;
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
