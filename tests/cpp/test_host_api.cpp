// GPU test of the C++ host mirror (mkf_amd/host/mkhip.hpp): the reference's
// evaluator calls on map-based ciphertexts and key sets, each result compared
// bit for bit with the CPU oracle (oracle/mkref.h), plus the reference's error
// behaviour (KKLSS panics on a missing key, MKF returns an error).  Mirrors the
// shape of the reference's own tests (mkckks_test.go:352-358, mkbfv_test.go:
// 407-413, mkf_cpu/ckks_test.go, bfv_test.go): build parameters and keys, call
// MulRelinNew, check the output.  Exit status 0 = every check passed.
#include "../../mkf_amd/host/mkhip.hpp"
#include "../../oracle/mkref.h"

#include <cstdio>
#include <cstdlib>

using namespace mkhip_host;

static int failures = 0;
#define CHECK(cond, what)                                        \
    do {                                                         \
        if (!(cond)) {                                           \
            std::printf("FAIL %s (%s:%d)\n", what, __FILE__, __LINE__); \
            ++failures;                                          \
        }                                                        \
    } while (0)

static std::vector<uint64_t> primes(int bits, int count, const std::vector<uint64_t> &exclude = {}) {
    std::vector<uint64_t> out(count);
    if (mkref_gen_primes(bits, 15, count, 1, out.data(), exclude.empty() ? nullptr : exclude.data(),
                         (int)exclude.size()) != 0) {
        std::printf("FAIL prime generation\n");
        std::exit(2);
    }
    return out;
}

static std::vector<uint64_t> uniform(mkref_prng &g, const std::vector<uint64_t> &mods, int N) {
    std::vector<uint64_t> v((size_t)mods.size() * N);
    mkref_sample_uniform(&g, mods.data(), (int)mods.size(), N, v.data());
    return v;
}

// [digits][rows][N] words -> beta PolyQP (L Q rows, K P rows)
static std::vector<PolyQP> to_qp(const std::vector<uint64_t> &w, int digits, int L, int K, int N) {
    std::vector<PolyQP> out(digits);
    for (int d = 0; d < digits; ++d) {
        out[d].Q = Poly::from(w.data() + (size_t)d * (L + K) * N, N, L, true);
        out[d].P = Poly::from(w.data() + ((size_t)d * (L + K) + L) * N, N, K, true);
    }
    return out;
}

static bool same(const Poly &p, const uint64_t *w, int nl, int N) {
    if ((int)p.Coeffs.size() != nl) return false;
    for (int l = 0; l < nl; ++l)
        for (int c = 0; c < N; ++c)
            if (p.Coeffs[l][c] != w[(size_t)l * N + c]) return false;
    return true;
}

static void test_kklss_ckks() {
    const int logN = 13, N = 1 << logN;
    mkckks::Parameters params;
    params.LogN = logN;
    params.Q = primes(50, 4);
    params.P = primes(60, 2);
    params.Gamma = 2;
    params.DefaultScale = (double)params.Q.back();
    const int L = 4, K = 2, beta = params.Beta(L - 1);
    std::vector<uint64_t> QP = params.Q;
    QP.insert(QP.end(), params.P.begin(), params.P.end());
    mkref_prng g;
    mkref_prng_seed(&g, 101);
    std::vector<uint64_t> u;
    for (int d = 0; d < beta; ++d) {
        auto r = uniform(g, QP, N);
        u.insert(u.end(), r.begin(), r.end());
    }
    params.CRS[-1].Value = to_qp(u, beta, L, K, N);
    const std::vector<std::string> names = {"alice", "bob", "carol"};
    mkrlwe::RelinearizationKeySet rlkSet;
    std::vector<std::vector<uint64_t>> kb, kd, kv;
    for (const auto &name : names) {
        mkrlwe::RelinearizationKey rlk;
        rlk.ID = name;
        for (int j = 0; j < 3; ++j) {
            std::vector<uint64_t> w;
            for (int d = 0; d < beta; ++d) {
                auto r = uniform(g, QP, N);
                w.insert(w.end(), r.begin(), r.end());
            }
            rlk.Value[j].Value = to_qp(w, beta, L, K, N);
            (j == 0 ? kb : j == 1 ? kd : kv).push_back(w);
        }
        rlkSet.AddRelinearizationKey(rlk);
    }
    auto make_ct = [&](const std::vector<std::string> &ids, std::vector<uint64_t> &flat) {
        mkckks::Ciphertext ct;
        ct.Scale = params.DefaultScale;
        auto c0 = uniform(g, params.Q, N);
        flat.insert(flat.end(), c0.begin(), c0.end());
        ct.Value["0"] = Poly::from(c0.data(), N, L, false);
        for (const auto &id : ids) {  // ids ascending (the map order the flat buffer uses)
            auto c = uniform(g, params.Q, N);
            flat.insert(flat.end(), c.begin(), c.end());
            ct.Value[id] = Poly::from(c.data(), N, L, false);
        }
        return ct;
    };
    std::vector<uint64_t> f0, f1;
    mkckks::Ciphertext op0 = make_ct({"alice", "bob"}, f0), op1 = make_ct({"bob", "carol"}, f1);

    mkckks::Evaluator eval(params);
    mkckks::Ciphertext out = eval.MulRelinNew(op0, op1, rlkSet);

    // oracle: MulRelinNew without rescale, then the rescales of the scale loop
    mkref_ctx ref;
    mkref_ctx_init(&ref, logN, params.Q.data(), L, params.P.data(), K, 2, 0);
    const int ids0[2] = {0, 1}, ids1[2] = {1, 2};
    std::vector<const uint64_t *> pb, pd, pv;
    for (int p = 0; p < 3; ++p) {
        pb.push_back(kb[p].data());
        pd.push_back(kd[p].data());
        pv.push_back(kv[p].data());
    }
    std::vector<uint64_t> exp((size_t)4 * L * N);
    int ids_out[4];
    // the oracle returns the number of output components, 1 + ku
    const int ku = mkref_kklss_mulrelin(&ref, L - 1, f0.data(), ids0, 2, f1.data(), ids1, 2, pb.data(), pd.data(),
                                        pv.data(), u.data(), exp.data(), ids_out) - 1;
    double ns = 0;
    const int nres = mkref_rescale_count(params.Q.data(), L - 1, params.DefaultScale * params.DefaultScale,
                                         params.DefaultScale, &ns);
    for (int c = 0; c <= ku; ++c) mkref_div_round_last(&ref, L - 1, nres, exp.data() + (size_t)c * L * N);
    const int nl = L - nres;
    // the oracle's rescaled polys keep their limb stride L; compare the first nl limbs
    CHECK(ku == 3 && out.Value.size() == 4, "kklss_ckks component count");
    CHECK(out.Level() == L - 1 - nres && out.Scale == ns, "kklss_ckks level / scale");
    CHECK(same(out.Value["0"], exp.data(), nl, N), "kklss_ckks c0");
    for (int c = 0; c < ku; ++c)
        CHECK(same(out.Value[names[ids_out[c]]], exp.data() + (size_t)(1 + c) * L * N, nl, N),
              "kklss_ckks party component");
    // inputs are not modified (MulRelinNew allocates a fresh output)
    CHECK(same(op0.Value["0"], f0.data(), L, N), "kklss_ckks input untouched");
    // a party without a relinearization key panics with the reference's message
    mkckks::Ciphertext op2 = op0;
    op2.Value["dave"] = op2.Value["alice"];
    bool panicked = false;
    try {
        eval.MulRelinNew(op2, op1, rlkSet);
    } catch (const Panic &e) {
        panicked = std::string(e.what()).find("there is no relinearization key with given id") != std::string::npos;
    }
    CHECK(panicked, "kklss_ckks missing key panics");
    mkref_ctx_free(&ref);
    std::printf("ok kklss_ckks MulRelinNew (N=2^13, L=4, K=2, parties alice+bob x bob+carol, %d rescale)\n", nres);
}

static void test_kklss_bfv() {
    const int logN = 13, N = 1 << logN, L = 3, K = 2;
    mkbfv::Parameters params;
    params.LogN = logN;
    params.Q = primes(50, L);
    params.QMul = primes(50, L, params.Q);
    params.P = primes(60, K);
    params.T = 65537;
    params.Gamma = 2;
    std::vector<uint64_t> QP = params.Q;
    QP.insert(QP.end(), params.P.begin(), params.P.end());
    const int beta = L;  // alpha = 1
    mkref_prng g;
    mkref_prng_seed(&g, 202);
    auto digits = [&](int n) {
        std::vector<uint64_t> w;
        for (int d = 0; d < n; ++d) {
            auto r = uniform(g, QP, N);
            w.insert(w.end(), r.begin(), r.end());
        }
        return w;
    };
    std::vector<uint64_t> u = digits(beta);
    params.CRS[-1].Value = to_qp(u, beta, L, K, N);
    mkbfv::RelinearizationKeySet rlkSet;
    std::vector<std::vector<uint64_t>> kb, kd, kv;
    const std::vector<std::string> names = {"p0", "p1"};
    for (const auto &name : names) {
        const std::vector<uint64_t> b = digits(2 * beta), d = digits(2 * beta), v = digits(beta), v1 = digits(beta);
        mkbfv::RelinearizationKey rlk;
        rlk.ID = name;
        const size_t half = (size_t)beta * (L + K) * N;
        for (int br = 0; br < 2; ++br) {
            rlk.Value[br].ID = name;
            rlk.Value[br].Value[0].Value = to_qp(std::vector<uint64_t>(b.begin() + br * half, b.begin() + (br + 1) * half), beta, L, K, N);
            rlk.Value[br].Value[1].Value = to_qp(std::vector<uint64_t>(d.begin() + br * half, d.begin() + (br + 1) * half), beta, L, K, N);
            rlk.Value[br].Value[2].Value = to_qp(br == 0 ? v : v1, beta, L, K, N);
        }
        rlkSet.AddRelinearizationKey(rlk);
        kb.push_back(b);
        kd.push_back(d);
        kv.push_back(v);
    }
    auto make_ct = [&](const std::vector<std::string> &ids, std::vector<uint64_t> &flat) {
        mkbfv::Ciphertext ct;
        for (int c = 0; c <= (int)ids.size(); ++c) {
            auto p = uniform(g, params.Q, N);
            flat.insert(flat.end(), p.begin(), p.end());
            ct.Value[c == 0 ? "0" : ids[c - 1]] = Poly::from(p.data(), N, L, false);
        }
        return ct;
    };
    std::vector<uint64_t> f0, f1;
    mkbfv::Ciphertext op0 = make_ct({"p0", "p1"}, f0), op1 = make_ct({"p1"}, f1);
    mkbfv::Evaluator eval(params);
    mkbfv::Ciphertext out = eval.MulRelinNew(op0, op1, rlkSet);

    mkref_bfv *ref = mkref_bfv_new(logN, params.Q.data(), L, params.QMul.data(), params.P.data(), K, params.T, 2, 0);
    const int ids0[2] = {0, 1}, ids1[1] = {1};
    const uint64_t *pb[2] = {kb[0].data(), kb[1].data()}, *pd[2] = {kd[0].data(), kd[1].data()},
                   *pv[2] = {kv[0].data(), kv[1].data()};
    std::vector<uint64_t> exp((size_t)3 * L * N);
    int ids_out[3];
    const int ku = mkref_kklss_bfv_mulrelin(ref, f0.data(), ids0, 2, f1.data(), ids1, 1, pb, pd, pv, u.data(),
                                            exp.data(), ids_out) - 1;
    CHECK(ku == 2 && out.Value.size() == 3, "kklss_bfv component count");
    CHECK(same(out.Value["0"], exp.data(), L, N), "kklss_bfv c0");
    for (int c = 0; c < ku; ++c)
        CHECK(same(out.Value[names[ids_out[c]]], exp.data() + (size_t)(1 + c) * L * N, L, N),
              "kklss_bfv party component");
    mkref_bfv_delete(ref);
    std::printf("ok kklss_bfv MulRelinNew (N=2^13, L=3 + 3 QMul, K=2, t=65537)\n");
}

// Lattigo switching key [beta][2][L+K][N] -> mkf::SwitchingKey
static mkf::SwitchingKey to_swk(const std::vector<uint64_t> &w, int beta, int L, int K, int N) {
    mkf::SwitchingKey k;
    k.Value.resize(beta);
    for (int d = 0; d < beta; ++d)
        for (int o = 0; o < 2; ++o) {
            const uint64_t *base = w.data() + ((size_t)d * 2 + o) * (L + K) * N;
            k.Value[d][o].Q = Poly::from(base, N, L, true);
            k.Value[d][o].P = Poly::from(base + (size_t)L * N, N, K, true);
        }
    return k;
}

static void test_mkf(bool bfv) {
    const int logN = 13, N = 1 << logN, k = 2;
    const int L = bfv ? 3 : 4, K = bfv ? 2 : 1;
    mkbfv::Parameters params;  // the BFV fields are ignored by the CKKS evaluator
    params.LogN = logN;
    params.Q = primes(50, L);
    params.P = primes(60, K);
    params.Gamma = 1;
    if (bfv) {
        params.QMul = primes(50, L, params.Q);
        params.T = 65537;
    }
    std::vector<uint64_t> QP = params.Q;
    QP.insert(QP.end(), params.P.begin(), params.P.end());
    const int alpha = K, beta = (L + alpha - 1) / alpha;
    mkref_prng g;
    mkref_prng_seed(&g, bfv ? 303 : 404);
    auto swk_words = [&]() {
        std::vector<uint64_t> w;
        for (int d = 0; d < 2 * beta; ++d) {
            auto r = uniform(g, QP, N);
            w.insert(w.end(), r.begin(), r.end());
        }
        return w;
    };
    std::vector<std::vector<uint64_t>> fold;
    mkf::OnlineKeys keys;
    for (int p = 0; p < k; ++p) {
        fold.push_back(swk_words());
        keys.Fold.push_back(to_swk(fold.back(), beta, L, K, N));
    }
    const std::vector<uint64_t> relin = swk_words();
    keys.Relin = to_swk(relin, beta, L, K, N);
    auto make = [&](std::vector<uint64_t> &flat) {
        mkf::MultiCiphertext m;
        m.Level = L - 1;
        m.Scale = (double)params.Q.back();
        for (int c = 0; c <= k; ++c) {
            auto p = uniform(g, params.Q, N);
            flat.insert(flat.end(), p.begin(), p.end());
            (c == 0 ? m.C0 : (m.Components.emplace_back(), m.Components.back())) = Poly::from(p.data(), N, L, !bfv);
        }
        return m;
    };
    std::vector<uint64_t> fl, fr;
    const mkf::MultiCiphertext left = make(fl), right = make(fr);
    std::vector<const uint64_t *> fp = {fold[0].data(), fold[1].data()};
    std::vector<uint64_t> e0((size_t)L * N), e1((size_t)L * N);
    std::pair<mkf::OutputCiphertext, std::string> res;
    int nl = L;
    if (bfv) {
        mkf::BFVEvaluator eval(params);
        CHECK(eval.SetKeys(keys).empty(), "mkf_bfv keys");
        res = eval.MulRelin(left, right);
        mkref_bfv *ref = mkref_bfv_new(logN, params.Q.data(), L, params.QMul.data(), params.P.data(), K, params.T, 1, 0);
        mkref_mkf_bfv_mulrelin(ref, k, fl.data(), fr.data(), fp.data(), relin.data(), e0.data(), e1.data());
        mkref_bfv_delete(ref);
        // error value, not a panic: a group of the wrong size
        mkf::MultiCiphertext bad = left;
        bad.Components.pop_back();
        CHECK(!eval.MulRelin(bad, right).second.empty(), "mkf_bfv error value");
    } else {
        mkf::CKKSEvaluator eval(params, (double)params.Q.back());
        CHECK(eval.SetKeys(keys).empty(), "mkf_ckks keys");
        res = eval.MulRelin(left, right);
        mkref_ctx ref;
        mkref_ctx_init(&ref, logN, params.Q.data(), L, params.P.data(), K, 1, 0);
        mkref_mkf_ckks_mulrelin(&ref, L - 1, k, fl.data(), fr.data(), fp.data(), relin.data(), e0.data(), e1.data());
        mkref_div_round_last_ntt(&ref, L - 1, 1, e0.data());
        mkref_div_round_last_ntt(&ref, L - 1, 1, e1.data());
        mkref_ctx_free(&ref);
        nl = L - 1;
    }
    const char *tag = bfv ? "mkf_bfv" : "mkf_ckks";
    CHECK(res.second.empty(), tag);
    CHECK(res.first.Level == nl - 1, "mkf level");
    CHECK(same(res.first.C0, e0.data(), nl, N), "mkf c0");
    CHECK(same(res.first.Gamma, e1.data(), nl, N), "mkf Gamma");
    CHECK((int)res.first.Components.size() == k, "mkf component count");
    for (const Poly &c : res.first.Components) CHECK(same(c, e1.data(), nl, N), "mkf materialized component");
    std::printf("ok %s MulRelin (N=2^13, L=%d, K=%d, k=%d)\n", tag, L, K, k);
}


// Rotations and conjugation through the mirror: mkckks Evaluator RotateNew (the
// power-of-two fallback), RotateHoistedNew's panic, ConjugateNew, and the MKF
// party-specific RotateNew with its error value (mkckks_test.go:443-458,
// :493-505; true_mkf_ckks.go:323-353).
static void test_rotations() {
    const int logN = 13, N = 1 << logN, L = 4, K = 2;
    mkckks::Parameters params;
    params.LogN = logN;
    params.Q = primes(50, L);
    params.P = primes(60, K);
    params.Gamma = 2;
    params.DefaultScale = (double)params.Q.back();
    const int beta = params.Beta(L - 1);
    std::vector<uint64_t> QP = params.Q;
    QP.insert(QP.end(), params.P.begin(), params.P.end());
    mkref_prng g;
    mkref_prng_seed(&g, 505);
    auto swk = [&]() {
        std::vector<uint64_t> w;
        for (int d = 0; d < beta; ++d) {
            auto r = uniform(g, QP, N);
            w.insert(w.end(), r.begin(), r.end());
        }
        return w;
    };
    std::map<int, std::vector<uint64_t>> crs;
    for (int idx : {1, 2, -2}) {
        crs[idx] = swk();
        params.CRS[idx].Value = to_qp(crs[idx], beta, L, K, N);
    }
    const std::vector<std::string> names = {"alice", "bob"};
    mkrlwe::RotationKeySet rks;
    mkrlwe::ConjugationKeySet cks;
    std::map<std::pair<int, int>, std::vector<uint64_t>> rk;
    std::vector<std::vector<uint64_t>> ck;
    for (int p = 0; p < 2; ++p) {
        for (unsigned r : {1u, 2u}) {
            rk[{p, (int)r}] = swk();
            mkrlwe::RotationKey k;
            k.ID = names[p];
            k.RotIdx = r;
            k.Value.Value = to_qp(rk[{p, (int)r}], beta, L, K, N);
            rks.AddRotationKey(k);
        }
        ck.push_back(swk());
        mkrlwe::ConjugationKey c;
        c.ID = names[p];
        c.Value.Value = to_qp(ck.back(), beta, L, K, N);
        cks.AddConjugationKey(c);
    }
    std::vector<uint64_t> flat;
    mkckks::Ciphertext ct;
    ct.Scale = params.DefaultScale;
    for (int c = 0; c < 3; ++c) {
        auto x = uniform(g, params.Q, N);
        flat.insert(flat.end(), x.begin(), x.end());
        ct.Value[c == 0 ? "0" : names[c - 1]] = Poly::from(x.data(), N, L, false);
    }
    mkckks::Evaluator eval(params);
    // rotation by 3 = rotation by 1 then by 2 (no CRS[3])
    mkckks::Ciphertext rot = eval.RotateNew(ct, 3, rks);
    mkref_ctx ref;
    mkref_ctx_init(&ref, logN, params.Q.data(), L, params.P.data(), K, 2, 0);
    std::vector<uint64_t> t1(flat.size()), t2(flat.size());
    const uint64_t *k1[2] = {rk[{0, 1}].data(), rk[{1, 1}].data()};
    const uint64_t *k2[2] = {rk[{0, 2}].data(), rk[{1, 2}].data()};
    mkref_kklss_rotate(&ref, L - 1, 2, flat.data(), k1, crs[1].data(), mkref_galois_element(logN, 1), t1.data());
    mkref_kklss_rotate(&ref, L - 1, 2, t1.data(), k2, crs[2].data(), mkref_galois_element(logN, 2), t2.data());
    CHECK(rot.Scale == ct.Scale && rot.Value.size() == 3, "rotate shape");
    CHECK(same(rot.Value["0"], t2.data(), L, N), "rotate c0");
    for (int p = 0; p < 2; ++p)
        CHECK(same(rot.Value[names[p]], t2.data() + (size_t)(1 + p) * L * N, L, N), "rotate component");
    bool panicked = false;
    try {
        eval.RotateHoistedNew(ct, 3, rks);
    } catch (const Panic &e) {
        panicked = std::string(e.what()).find("Hoisted rotation only works") != std::string::npos;
    }
    CHECK(panicked, "hoisted rotation without a precomputed key panics");
    mkckks::Ciphertext hr = eval.RotateHoistedNew(ct, 1, rks);
    CHECK(same(hr.Value["0"], t1.data(), L, N), "hoisted rotate c0");
    // conjugation
    mkckks::Ciphertext cj = eval.ConjugateNew(ct, cks);
    std::vector<uint64_t> ec(flat.size());
    const uint64_t *cp[2] = {ck[0].data(), ck[1].data()};
    mkref_kklss_conjugate(&ref, L - 1, 2, flat.data(), cp, crs[-2].data(), ec.data());
    for (int c = 0; c < 3; ++c)
        CHECK(same(cj.Value[c == 0 ? "0" : names[c - 1]], ec.data() + (size_t)c * L * N, L, N), "conjugate");
    // a party without a rotation key panics with the reference's message
    mkckks::Ciphertext ct3 = ct;
    ct3.Value["carol"] = ct.Value["alice"];
    panicked = false;
    try {
        eval.RotateNew(ct3, 1, rks);
    } catch (const Panic &e) {
        panicked = std::string(e.what()).find("there is no rotation key with given id") != std::string::npos;
    }
    CHECK(panicked, "rotate missing key panics");
    mkref_ctx_free(&ref);
    std::printf("ok kklss RotateNew / RotateHoistedNew / ConjugateNew (N=2^13, L=4, K=2, alice+bob)\n");

    // MKF party-specific rotation (gamma = 1 context)
    const int Km = 1, betam = L;
    std::vector<uint64_t> Pm = primes(60, Km);
    std::vector<uint64_t> QPm = params.Q;
    QPm.insert(QPm.end(), Pm.begin(), Pm.end());
    mkrlwe::Parameters pm;
    pm.LogN = logN;
    pm.Q = params.Q;
    pm.P = Pm;
    pm.Gamma = 1;
    auto mswk = [&]() {
        std::vector<uint64_t> w;
        for (int d = 0; d < 2 * betam; ++d) {
            auto r = uniform(g, QPm, N);
            w.insert(w.end(), r.begin(), r.end());
        }
        return w;
    };
    mkf::OnlineKeys keys;
    for (int p = 0; p < 2; ++p) keys.Fold.push_back(to_swk(mswk(), betam, L, Km, N));
    keys.Relin = to_swk(mswk(), betam, L, Km, N);
    std::vector<std::vector<uint64_t>> rot1 = {mswk(), mswk()};
    const uint64_t gal1 = mkref_galois_element(logN, 1);
    for (const auto &w : rot1) keys.Rotations[gal1].push_back(to_swk(w, betam, L, Km, N));
    mkf::CKKSEvaluator me(pm, (double)params.Q.back());
    CHECK(me.SetKeys(keys).empty(), "mkf keys with rotations");
    mkf::MultiCiphertext in;
    std::vector<uint64_t> fin;
    in.Level = L - 1;
    in.Scale = 3.0;
    for (int c = 0; c < 3; ++c) {
        auto x = uniform(g, params.Q, N);
        fin.insert(fin.end(), x.begin(), x.end());
        (c == 0 ? in.C0 : (in.Components.emplace_back(), in.Components.back())) = Poly::from(x.data(), N, L, true);
    }
    auto mr = me.RotateNew(in, 1);
    CHECK(mr.second.empty(), "mkf rotate");
    mkref_ctx mref;
    mkref_ctx_init(&mref, logN, params.Q.data(), L, Pm.data(), Km, 1, 0);
    std::vector<uint64_t> o0((size_t)L * N), o1((size_t)L * N);
    const uint64_t *rp[2] = {rot1[0].data(), rot1[1].data()};
    mkref_mkf_ckks_rotate(&mref, L - 1, 2, fin.data(), rp, gal1, o0.data(), o1.data());
    mkref_ctx_free(&mref);
    CHECK(same(mr.first.C0, o0.data(), L, N), "mkf rotate C0");
    CHECK(mr.first.Components.size() == 2 && mr.first.Scale == 3.0, "mkf rotate shape");
    for (const Poly &c : mr.first.Components) CHECK(same(c, o1.data(), L, N), "mkf rotate component");
    CHECK(!me.RotateNew(in, 2).second.empty(), "mkf rotate missing keys is an error value");
    std::printf("ok mkf RotateNew (N=2^13, L=4, K=1, k=2)\n");
}

// The MKF evaluator's public operations through the mirror: MulRelinNew of operands
// at different levels and RescaleNew against the C oracle, AddNew against the sum,
// EvaluatePoly against the C entry point, and the error values
// (true_mkf_ckks.go:187-303, true_mkf_ckks_polynomial.go:79-111).
static void test_mkf_poly() {
    const int logN = 13, N = 1 << logN, k = 2, K = 2;
    mkbfv::Parameters params;
    params.LogN = logN;
    params.Q = primes(60, 1);
    const std::vector<uint64_t> q40 = primes(40, 4, params.Q);
    params.Q.insert(params.Q.end(), q40.begin(), q40.end());
    params.P = primes(60, K, params.Q);
    params.Gamma = 1;
    const int L = (int)params.Q.size(), beta = (L + K - 1) / K;
    std::vector<uint64_t> QP = params.Q;
    QP.insert(QP.end(), params.P.begin(), params.P.end());
    mkref_prng g;
    mkref_prng_seed(&g, 505);
    auto swk_words = [&]() {
        std::vector<uint64_t> w;
        for (int d = 0; d < 2 * beta; ++d) {
            auto r = uniform(g, QP, N);
            w.insert(w.end(), r.begin(), r.end());
        }
        return w;
    };
    std::vector<std::vector<uint64_t>> fold = {swk_words(), swk_words()};
    const std::vector<uint64_t> relin = swk_words();
    mkf::OnlineKeys keys;
    for (const auto &f : fold) keys.Fold.push_back(to_swk(f, beta, L, K, N));
    keys.Relin = to_swk(relin, beta, L, K, N);
    const double D = (double)(1ull << 40);
    mkf::CKKSEvaluator eval(params, D);
    CHECK(eval.SetKeys(keys).empty(), "mkf poly keys");
    auto make = [&](int level, std::vector<uint64_t> &flat) {
        mkf::MultiCiphertext m;
        m.Level = level;
        m.Scale = D;
        const std::vector<uint64_t> Ql(params.Q.begin(), params.Q.begin() + level + 1);
        for (int c = 0; c <= k; ++c) {
            auto p = uniform(g, Ql, N);
            flat.insert(flat.end(), p.begin(), p.end());
            (c == 0 ? m.C0 : (m.Components.emplace_back(), m.Components.back())) = Poly::from(p.data(), N, level + 1, true);
        }
        return m;
    };
    std::vector<uint64_t> fx, fy;
    const mkf::MultiCiphertext x = make(L - 1, fx), y = make(L - 2, fy);
    mkref_ctx ref;
    mkref_ctx_init(&ref, logN, params.Q.data(), L, params.P.data(), K, 1, 0);
    // MulRelinNew: x folded at level L-1, y at L-2, output at L-2 (c0, k copies of Gamma)
    auto m = eval.MulRelinNew(x, y);
    CHECK(m.second.empty() && m.first.Level == L - 2 && m.first.Scale == D * D, "mkf MulRelinNew");
    const int nl = L - 1;
    std::vector<uint64_t> e0((size_t)nl * N), e1((size_t)nl * N);
    const uint64_t *fp[2] = {fold[0].data(), fold[1].data()};
    mkref_mkf_ckks_mulrelin2(&ref, L - 1, L - 2, k, fx.data(), fy.data(), fp, relin.data(), e0.data(), e1.data());
    CHECK(same(m.first.C0, e0.data(), nl, N), "mkf MulRelinNew c0");
    for (const Poly &c : m.first.Components) CHECK(same(c, e1.data(), nl, N), "mkf MulRelinNew component");
    // RescaleNew(D): one prime
    auto r = eval.RescaleNew(m.first, D);
    CHECK(r.second.empty() && r.first.Level == L - 3, "mkf RescaleNew");
    mkref_div_round_last_ntt(&ref, L - 2, 1, e0.data());
    CHECK(same(r.first.C0, e0.data(), nl - 1, N), "mkf RescaleNew c0");
    // AddNew at equal scales: the sum mod q on the smaller level
    auto s = eval.AddNew(x, y);
    CHECK(s.second.empty() && s.first.Level == L - 2, "mkf AddNew");
    bool ok = true;
    for (int l = 0; l <= L - 2; ++l)
        for (int i = 0; i < N; i += 97) {
            const uint64_t q = params.Q[l];
            const uint64_t a = fx[(size_t)l * N + i], b = fy[(size_t)l * N + i];
            ok = ok && s.first.C0.Coeffs[l][i] == (a + b) % q;
        }
    CHECK(ok, "mkf AddNew values");
    // EvaluatePoly through the mirror = the C entry point on the same words
    mkf::CKKSEvaluator::Polynomial poly;
    poly.Coeffs = {0.1, 0.2, -0.3, 0.4};
    poly.MaxDeg = 3;
    auto ep = eval.EvaluatePoly(x, poly, D);
    CHECK(ep.second.empty() && ep.first.Level == L - 3, "mkf EvaluatePoly");
    const double re[4] = {0.1, 0.2, -0.3, 0.4};
    mkhip_polynomial mp = {re, nullptr, 4, 3, 1, MKHIP_BASIS_STANDARD};
    std::vector<uint64_t> out((size_t)(1 + k) * (L - 2) * N);
    int lo = 0;
    double so = 0;
    CHECK(mkhip_mkf_ckks_evaluate_poly(eval.handle(), 0, fx.data(), L - 1, D, &mp, D, out.data(), &lo, &so) ==
              MKHIP_OK && lo == L - 3 && so == D, "mkf evaluate_poly C entry");
    CHECK(same(ep.first.C0, out.data(), L - 2, N), "mkf EvaluatePoly words");
    // error values
    poly.Coeffs.assign(9, 0.0);
    poly.Coeffs[8] = 1.0;
    poly.MaxDeg = 8;
    CHECK(!eval.EvaluatePoly(y, poly, D).second.empty(), "mkf EvaluatePoly depth above level is an error value");
    CHECK(!eval.DropLevelNew(x, L).second.empty(), "mkf DropLevelNew error value");
    mkref_ctx_free(&ref);
    std::printf("ok mkf polynomial evaluator (N=2^13, L=%d, K=%d, k=%d)\n", L, K, k);
}

int main() {
    std::setvbuf(stdout, nullptr, _IONBF, 0);  // every line reaches the log as it is printed
    try {
        test_kklss_ckks();
        test_kklss_bfv();
        test_mkf(false);
        test_mkf(true);
        test_rotations();
        test_mkf_poly();
    } catch (const std::exception &e) {
        std::printf("FAIL unexpected exception: %s\n", e.what());
        return 1;
    }
    std::printf("%s\n", failures ? "FAILED" : "all host-API checks passed");
    return failures ? 1 : 0;
}
