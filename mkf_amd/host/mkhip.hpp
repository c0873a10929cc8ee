// mkhip.hpp -- C++ host mirror of the reference's Go evaluator surface for
// multi-key mul + relin, over the C ABI of include/mkhip.h (header only; link
// libmkhip.so).
//
// Types and calls keep the reference's names, argument meaning and error
// behaviour:
//   mkrlwe::Ciphertext{Value map[string]Poly}          mkrlwe/elements.go:17-19
//   mkrlwe::SwitchingKey{Value []PolyQP}               mkrlwe/keys.go:23-25
//   mkrlwe::RelinearizationKey{Value [3], ID}          mkrlwe/keys.go:34-37
//   mkrlwe::RelinearizationKeySet                      mkrlwe/keys.go:53-57, 165-198
//   mkckks::Evaluator::MulRelinNew(op0, op1, rlkSet)   mkckks/evaluator.go:419
//   mkbfv::RelinearizationKey{Value [2], ID}           mkbfv/keys.go:6-9
//   mkbfv::Evaluator::MulRelinNew(op0, op1, rlkSet)    mkbfv/evaluator.go:85
//   mkrlwe::RotationKeySet / ConjugationKeySet         mkrlwe/keys.go:39-67, 125-160, 201-225
//   mkckks / mkbfv Evaluator::RotateNew, RotateHoistedNew, ConjugateNew
//                                                      mkckks/evaluator.go:488-545, 590-624;
//                                                      mkbfv/evaluator.go:156-210
//   mkf::CKKSEvaluator / mkf::BFVEvaluator             mkf_cpu/ckks.go:102-166, bfv.go:212-252
//   mkf::CKKSEvaluator::RotateNew                      mkf_cpu/true_mkf_ckks.go:323-353
//   mkf::CKKSEvaluator::EvaluatePoly, AddNew, SubNew, AddConstNew, MultByConstNew,
//     RescaleNew, DropLevelNew, MulRelinNew           true_mkf_ckks_polynomial.go, true_mkf_ckks.go:187-303
// The KKLSS evaluators panic on bad input (keyswitch_hoisted.go:53-59, keys.go:
// 190-198): here they throw mkhip_host::Panic with the reference's message.  The
// MKF calls return an error value: here a non-empty std::string.
//
// Ring elements are ring.Poly-like: Coeffs[limb][coeff], canonical residues.
// KKLSS ciphertexts are coefficient domain; MKF-CKKS ones are NTT domain (as in
// the reference).  Keys are NTT + Montgomery form.  An evaluator owns one mkhip
// context (device tables + resident keys); like the reference's evaluators it is
// not re-entrant.
#pragma once

#include "../../include/mkhip.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <complex>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mkhip_host {

struct Panic : std::runtime_error {
    int status;
    Panic(const std::string &what, int st) : std::runtime_error(what), status(st) {}
};

inline std::string status_text(int st) { return mkhip_status_string(st); }

// ring.Poly
struct Poly {
    std::vector<std::vector<uint64_t>> Coeffs;
    bool IsNTT = false;
    Poly() = default;
    Poly(int N, int nlimbs) : Coeffs(nlimbs, std::vector<uint64_t>(N, 0)) {}
    int Level() const { return (int)Coeffs.size() - 1; }
    int N() const { return Coeffs.empty() ? 0 : (int)Coeffs[0].size(); }
    void append_to(std::vector<uint64_t> &out, int nlimbs) const {
        for (int l = 0; l < nlimbs; ++l) out.insert(out.end(), Coeffs[l].begin(), Coeffs[l].end());
    }
    static Poly from(const uint64_t *p, int N, int nlimbs, bool ntt) {
        Poly r(N, nlimbs);
        for (int l = 0; l < nlimbs; ++l) std::copy(p + (size_t)l * N, p + (size_t)(l + 1) * N, r.Coeffs[l].begin());
        r.IsNTT = ntt;
        return r;
    }
};

// rlwe.PolyQP
struct PolyQP {
    Poly Q, P;
};

namespace detail {
// [digit][Q limbs, P limbs][N] (the reference's fixture layout, mkckks/fpga_fixture_test.go:125-148)
inline void flatten_qp(const std::vector<PolyQP> &v, std::vector<uint64_t> &out) {
    for (const PolyQP &d : v) {
        d.Q.append_to(out, (int)d.Q.Coeffs.size());
        d.P.append_to(out, (int)d.P.Coeffs.size());
    }
}
struct Ctx {
    mkhip_ctx *c = nullptr;
    Ctx() = default;
    Ctx(const Ctx &) = delete;
    Ctx &operator=(const Ctx &) = delete;
    ~Ctx() {
        if (c) mkhip_ctx_destroy(c);
    }
};
}  // namespace detail

namespace mkrlwe {

// mkrlwe.Parameters (params.go:8-12): ring degree, moduli, gamma and the CRS
// (index -1 is u, the one the relinearization uses).
struct SwitchingKey {
    std::vector<PolyQP> Value;
};

struct Parameters {
    int LogN = 0;
    std::vector<uint64_t> Q, P;
    int Gamma = 2;
    std::map<int, SwitchingKey> CRS;
    int N() const { return 1 << LogN; }
    int MaxLevel() const { return (int)Q.size() - 1; }
    int Alpha() const { return (int)P.size() / Gamma; }
    int Beta(int levelQ) const { return (levelQ + Alpha()) / Alpha(); }
};

struct RelinearizationKey {
    std::array<SwitchingKey, 3> Value;  // b, d, v
    std::string ID;
};

struct RelinearizationKeySet {
    std::map<std::string, RelinearizationKey> Value;
    void AddRelinearizationKey(const RelinearizationKey &rlk) { Value[rlk.ID] = rlk; }
    void DelRelinearizationKey(const std::string &id) { Value.erase(id); }
    const RelinearizationKey &GetRelinearizationKey(const std::string &id) const {
        auto it = Value.find(id);
        if (it == Value.end())
            throw Panic("cannot GetRelinearizationKey: there is no relinearization key with given id",
                        MKHIP_ERR_UNKNOWN_PARTY);
        return it->second;
    }
};

// RotationKey / RotationKeySet (keys.go:39-44, 60-62, 125-160)
struct RotationKey {
    SwitchingKey Value;
    std::string ID;
    unsigned RotIdx = 0;
};

struct RotationKeySet {
    std::map<std::string, std::map<unsigned, RotationKey>> Value;
    void AddRotationKey(const RotationKey &rk) { Value[rk.ID][rk.RotIdx] = rk; }
    void DelRotationKey(const std::string &id, unsigned rotidx) {
        auto it = Value.find(id);
        if (it != Value.end()) it->second.erase(rotidx);
    }
    const RotationKey &GetRotationKey(const std::string &id, unsigned rotidx) const {
        auto it = Value.find(id);
        if (it != Value.end()) {
            auto jt = it->second.find(rotidx);
            if (jt != it->second.end()) return jt->second;
        }
        throw Panic("cannot GetRotationKeys: there is no rotation key with given id", MKHIP_ERR_NO_KEY);
    }
};

// ConjugationKey / ConjugationKeySet (keys.go:46-50, 64-67, 201-225)
struct ConjugationKey {
    SwitchingKey Value;
    std::string ID;
};

struct ConjugationKeySet {
    std::map<std::string, ConjugationKey> Value;
    void AddConjugationKey(const ConjugationKey &ck) { Value[ck.ID] = ck; }
    void DelConjugationKey(const std::string &id) { Value.erase(id); }
    const ConjugationKey &GetConjugationKey(const std::string &id) const {
        auto it = Value.find(id);
        if (it == Value.end())
            throw Panic("cannot GetConjugationKey: there is no conjugation key with given id", MKHIP_ERR_NO_KEY);
        return it->second;
    }
};

struct Ciphertext {
    std::map<std::string, Poly> Value;  // "0" is c0, the other keys are party ids
    int Level() const { return Value.empty() ? -1 : Value.begin()->second.Level(); }
    std::vector<std::string> IDSet() const {
        std::vector<std::string> ids;
        for (const auto &kv : Value)
            if (kv.first != "0") ids.push_back(kv.first);
        return ids;
    }
};

}  // namespace mkrlwe

namespace detail {
// Party ids are registered with the device context in first-seen order; the
// ABI orders output components by these integer ids.
struct PartyTable {
    std::map<std::string, int> gid;
    std::map<int, std::string> name;
    int id(const std::string &s) {
        auto it = gid.find(s);
        if (it != gid.end()) return it->second;
        const int g = (int)gid.size();
        gid[s] = g;
        name[g] = s;
        return g;
    }
};
inline void flatten_ct(const mkrlwe::Ciphertext &ct, const std::vector<std::string> &ids, int nlimbs,
                       std::vector<uint64_t> &out) {
    ct.Value.at("0").append_to(out, nlimbs);
    for (const std::string &id : ids) ct.Value.at(id).append_to(out, nlimbs);
}

// Rotation / conjugation keys of a KKLSS evaluator: the CRS of every rotation
// and of conjugation is uploaded with the evaluator; party keys on first use.
struct GaloisKeys {
    std::set<std::pair<std::string, unsigned>> rot;
    std::set<std::string> conj;

    static void upload_crs(mkhip_ctx *c, const mkrlwe::Parameters &params) {
        for (const auto &kv : params.CRS) {
            const int idx = kv.first;
            if (idx != -2 && (idx < 1 || idx >= params.N() / 2)) continue;
            std::vector<uint64_t> a;
            flatten_qp(kv.second.Value, a);
            const int st = mkhip_set_crs(c, idx, a.data());
            if (st != MKHIP_OK) throw Panic("cannot NewEvaluator: " + status_text(st), st);
        }
    }

    // Evaluator.rotate / rotateHoisted / conjugate (mkckks/evaluator.go:496-545,
    // 598-624): the keys each step needs are fetched (panicking like the reference
    // when one is missing) and uploaded, then the device runs the steps.
    mkrlwe::Ciphertext run(mkhip_ctx *c, PartyTable &parties, const mkrlwe::Parameters &params,
                           const mkrlwe::Ciphertext &ct, int rotidx, bool hoisted, bool conjugate,
                           const mkrlwe::RotationKeySet *rks, const mkrlwe::ConjugationKeySet *cks) {
        const std::vector<std::string> ids = ct.IDSet();
        std::vector<int> g;
        for (const auto &id : ids) g.push_back(parties.id(id));
        const int N = params.N(), half = N / 2;
        if (conjugate) {
            for (size_t i = 0; i < ids.size(); ++i) {
                if (conj.count(ids[i])) continue;
                std::vector<uint64_t> k;
                flatten_qp(cks->GetConjugationKey(ids[i]).Value.Value, k);
                const int st = mkhip_add_conjugation_key(c, g[i], k.data());
                if (st != MKHIP_OK) throw Panic("AddConjugationKey: " + status_text(st), st);
                conj.insert(ids[i]);
            }
        } else {
            const int r = ((rotidx % half) + half) % half;
            std::vector<unsigned> steps;
            if (r != 0) {
                if (params.CRS.count(r)) {
                    steps.push_back((unsigned)r);
                } else if (hoisted) {
                    throw Panic("Hoisted rotation only works for precomputed rotation keys", MKHIP_ERR_NO_KEY);
                } else {
                    for (unsigned kk = 1, rr = (unsigned)r; rr > 0; kk *= 2, rr /= 2)
                        if (rr % 2) steps.push_back(kk);
                }
            }
            for (unsigned st_r : steps)
                for (size_t i = 0; i < ids.size(); ++i) {
                    if (rot.count({ids[i], st_r})) continue;
                    std::vector<uint64_t> k;
                    flatten_qp(rks->GetRotationKey(ids[i], st_r).Value.Value, k);
                    const int st = mkhip_add_rotation_key(c, g[i], (int)st_r, k.data());
                    if (st != MKHIP_OK) throw Panic("AddRotationKey: " + status_text(st), st);
                    rot.insert({ids[i], st_r});
                }
        }
        const int nl = ct.Level() + 1;
        std::vector<uint64_t> in, out;
        flatten_ct(ct, ids, nl, in);
        out.resize(in.size());
        const int st = conjugate ? mkhip_kklss_conjugate(c, in.data(), nl - 1, g.data(), (int)g.size(), out.data())
                                 : mkhip_kklss_rotate(c, in.data(), nl - 1, g.data(), (int)g.size(), rotidx,
                                                      hoisted ? 1 : 0, out.data());
        if (st != MKHIP_OK) throw Panic(std::string(conjugate ? "Conjugate: " : "Rotate: ") + status_text(st), st);
        mkrlwe::Ciphertext res;
        res.Value["0"] = Poly::from(out.data(), N, nl, false);
        for (size_t i = 0; i < ids.size(); ++i)
            res.Value[ids[i]] = Poly::from(out.data() + (1 + i) * (size_t)nl * N, N, nl, false);
        return res;
    }
};
}  // namespace detail

namespace mkckks {

// mkckks.Parameters (params.go:8-13): mkrlwe parameters with gamma = 2 and the default scale.
struct Parameters : mkrlwe::Parameters {
    double DefaultScale = 1.0;
    double Scale() const { return DefaultScale; }
};

struct Ciphertext : mkrlwe::Ciphertext {
    double Scale = 1.0;
};

class Evaluator {
  public:
    explicit Evaluator(const Parameters &params) : params_(params) {
        int st = mkhip_ctx_create(&ctx_.c, params.LogN, params.Q.data(), (int)params.Q.size(), params.P.data(),
                                  (int)params.P.size(), params.Gamma, 0, nullptr, 0);
        if (st != MKHIP_OK) throw Panic("cannot NewEvaluator: " + status_text(st), st);
        auto it = params.CRS.find(-1);
        if (it != params.CRS.end()) {
            std::vector<uint64_t> u;
            detail::flatten_qp(it->second.Value, u);
            st = mkhip_set_crs(ctx_.c, -1, u.data());
            if (st != MKHIP_OK) throw Panic("cannot NewEvaluator: " + status_text(st), st);
        }
        detail::GaloisKeys::upload_crs(ctx_.c, params);
    }

    // RotateNew (evaluator.go:488-529): rotidx normalised into [0, N/2); a rotation
    // without a CRS entry runs as rotations by 1, 2, 4, ...
    Ciphertext RotateNew(const Ciphertext &ct0, int rotidx, const mkrlwe::RotationKeySet &rkSet) {
        return wrap(galois_.run(ctx_.c, parties_, params_, ct0, rotidx, false, false, &rkSet, nullptr), ct0);
    }
    // RotateHoistedNew (evaluator.go:590-624); the hoisted form is computed on the
    // device inside the call (HoistedForm :546-556 has no separate host object here).
    Ciphertext RotateHoistedNew(const Ciphertext &ct0, int rotidx, const mkrlwe::RotationKeySet &rkSet) {
        return wrap(galois_.run(ctx_.c, parties_, params_, ct0, rotidx, true, false, &rkSet, nullptr), ct0);
    }
    // ConjugateNew (evaluator.go:533-545)
    Ciphertext ConjugateNew(const Ciphertext &ct0, const mkrlwe::ConjugationKeySet &ckSet) {
        return wrap(galois_.run(ctx_.c, parties_, params_, ct0, 0, false, true, nullptr, &ckSet), ct0);
    }

    // MulRelinNew (evaluator.go:419-446): tensor + relinearization + rescale by the
    // float64 scale loop; keys of the ids involved come from rlkSet (registered with
    // the device on first use of each id).  Inputs are not modified.
    Ciphertext MulRelinNew(const Ciphertext &op0, const Ciphertext &op1, const mkrlwe::RelinearizationKeySet &rlkSet) {
        const std::vector<std::string> i0 = op0.IDSet(), i1 = op1.IDSet();
        std::vector<int> g0, g1;
        for (const auto &id : i0) g0.push_back(ensure(id, rlkSet));
        for (const auto &id : i1) g1.push_back(ensure(id, rlkSet));
        const int l0 = op0.Level(), l1 = op1.Level(), N = params_.N();
        std::vector<uint64_t> b0, b1;
        detail::flatten_ct(op0, i0, l0 + 1, b0);
        detail::flatten_ct(op1, i1, l1 + 1, b1);
        std::set<std::string> uni(i0.begin(), i0.end());
        uni.insert(i1.begin(), i1.end());
        const int level = std::min(l0, l1);
        std::vector<uint64_t> out((size_t)(1 + uni.size()) * (level + 1) * N);
        std::vector<int> ids_out(uni.size() + 1);
        int ku = 0, level_out = 0;
        double scale_out = 0;
        const int st = mkhip_kklss_ckks_mulrelin(ctx_.c, b0.data(), l0, op0.Scale, g0.data(), (int)g0.size(),
                                                 b1.data(), l1, op1.Scale, g1.data(), (int)g1.size(),
                                                 params_.Scale(), out.data(), ids_out.data(), &ku, &level_out,
                                                 &scale_out);
        if (st != MKHIP_OK) throw Panic("MulRelinNew: " + status_text(st), st);
        Ciphertext ct;
        ct.Scale = scale_out;
        const int nl = level_out + 1;
        ct.Value["0"] = Poly::from(out.data(), N, nl, false);
        for (int u = 0; u < ku; ++u)
            ct.Value[parties_.name[ids_out[u]]] = Poly::from(out.data() + (size_t)(1 + u) * nl * N, N, nl, false);
        return ct;
    }

    mkhip_ctx *handle() const { return ctx_.c; }

  private:
    static Ciphertext wrap(mkrlwe::Ciphertext &&c, const Ciphertext &like) {
        Ciphertext out;
        out.Value = std::move(c.Value);
        out.Scale = like.Scale;
        return out;
    }
    int ensure(const std::string &id, const mkrlwe::RelinearizationKeySet &rlkSet) {
        const int g = parties_.id(id);
        if (relin_.count(id)) return g;
        const mkrlwe::RelinearizationKey &rlk = rlkSet.GetRelinearizationKey(id);
        std::vector<uint64_t> b, d, v;
        detail::flatten_qp(rlk.Value[0].Value, b);
        detail::flatten_qp(rlk.Value[1].Value, d);
        detail::flatten_qp(rlk.Value[2].Value, v);
        const int st = mkhip_register_party(ctx_.c, g, b.data(), d.data(), v.data());
        if (st != MKHIP_OK) throw Panic("AddRelinearizationKey: " + status_text(st), st);
        relin_.insert(id);
        return g;
    }
    Parameters params_;
    detail::Ctx ctx_;
    detail::PartyTable parties_;
    std::set<std::string> relin_;
    detail::GaloisKeys galois_;
};

}  // namespace mkckks

namespace mkbfv {

// mkbfv.ParametersLiteral (params.go:12-22): Q, QMul (the auxiliary chain of
// R = Q u QMul), P and the plaintext modulus T.
struct Parameters : mkrlwe::Parameters {
    std::vector<uint64_t> QMul;
    uint64_t T = 0;
};

// mkbfv.RelinearizationKey (keys.go:6-9): one mkrlwe key per gadget branch.
struct RelinearizationKey {
    std::array<mkrlwe::RelinearizationKey, 2> Value;
    std::string ID;
};

struct RelinearizationKeySet {
    std::map<std::string, RelinearizationKey> Value;
    void AddRelinearizationKey(const RelinearizationKey &rlk) { Value[rlk.ID] = rlk; }
    void DelRelinearizationKey(const std::string &id) { Value.erase(id); }
    const RelinearizationKey &GetRelinearizationKey(const std::string &id) const {
        auto it = Value.find(id);
        if (it == Value.end())
            throw Panic("cannot GetRelinearizationKey: there is no relinearization key with given id",
                        MKHIP_ERR_UNKNOWN_PARTY);
        return it->second;
    }
};

struct Ciphertext : mkrlwe::Ciphertext {};

class Evaluator {
  public:
    explicit Evaluator(const Parameters &params) : params_(params) {
        int st = mkhip_ctx_create(&ctx_.c, params.LogN, params.Q.data(), (int)params.Q.size(), params.P.data(),
                                  (int)params.P.size(), 2, params.T, params.QMul.data(), 0);
        if (st != MKHIP_OK) throw Panic("cannot NewEvaluator: " + status_text(st), st);
        auto it = params.CRS.find(-1);
        if (it != params.CRS.end()) {
            std::vector<uint64_t> u;
            detail::flatten_qp(it->second.Value, u);
            st = mkhip_set_crs(ctx_.c, -1, u.data());
            if (st != MKHIP_OK) throw Panic("cannot NewEvaluator: " + status_text(st), st);
        }
        detail::GaloisKeys::upload_crs(ctx_.c, params);
    }

    // RotateNew / ConjugateNew (evaluator.go:156-210): the mkrlwe key switches over Q
    Ciphertext RotateNew(const Ciphertext &ct0, int rotidx, const mkrlwe::RotationKeySet &rkSet) {
        return wrap(galois_.run(ctx_.c, parties_, params_, ct0, rotidx, false, false, &rkSet, nullptr));
    }
    Ciphertext ConjugateNew(const Ciphertext &ct0, const mkrlwe::ConjugationKeySet &ckSet) {
        return wrap(galois_.run(ctx_.c, parties_, params_, ct0, 0, false, true, nullptr, &ckSet));
    }

    // MulRelinNew (evaluator.go:85-91): operands at the top level, coefficient domain.
    Ciphertext MulRelinNew(const Ciphertext &op0, const Ciphertext &op1, const RelinearizationKeySet &rlkSet) {
        const std::vector<std::string> i0 = op0.IDSet(), i1 = op1.IDSet();
        std::vector<int> g0, g1;
        for (const auto &id : i0) g0.push_back(ensure(id, rlkSet));
        for (const auto &id : i1) g1.push_back(ensure(id, rlkSet));
        const int L = (int)params_.Q.size(), N = params_.N();
        if (op0.Level() != L - 1 || op1.Level() != L - 1)
            throw Panic("MulRelinNew: operands must be at the top level", MKHIP_ERR_LEVEL);
        std::vector<uint64_t> b0, b1;
        detail::flatten_ct(op0, i0, L, b0);
        detail::flatten_ct(op1, i1, L, b1);
        std::set<std::string> uni(i0.begin(), i0.end());
        uni.insert(i1.begin(), i1.end());
        std::vector<uint64_t> out((size_t)(1 + uni.size()) * L * N);
        std::vector<int> ids_out(uni.size() + 1);
        int ku = 0;
        const int st = mkhip_kklss_bfv_mulrelin(ctx_.c, b0.data(), g0.data(), (int)g0.size(), b1.data(), g1.data(),
                                                (int)g1.size(), out.data(), ids_out.data(), &ku);
        if (st != MKHIP_OK) throw Panic("MulRelinNew: " + status_text(st), st);
        Ciphertext ct;
        ct.Value["0"] = Poly::from(out.data(), N, L, false);
        for (int u = 0; u < ku; ++u)
            ct.Value[parties_.name[ids_out[u]]] = Poly::from(out.data() + (size_t)(1 + u) * L * N, N, L, false);
        return ct;
    }

    mkhip_ctx *handle() const { return ctx_.c; }

  private:
    static Ciphertext wrap(mkrlwe::Ciphertext &&c) {
        Ciphertext out;
        out.Value = std::move(c.Value);
        return out;
    }
    int ensure(const std::string &id, const RelinearizationKeySet &rlkSet) {
        if (relin_.count(id)) return parties_.gid[id];
        const RelinearizationKey &rlk = rlkSet.GetRelinearizationKey(id);
        std::vector<uint64_t> b0, d0, v0, b1, d1;
        detail::flatten_qp(rlk.Value[0].Value[0].Value, b0);
        detail::flatten_qp(rlk.Value[0].Value[1].Value, d0);
        detail::flatten_qp(rlk.Value[0].Value[2].Value, v0);
        detail::flatten_qp(rlk.Value[1].Value[0].Value, b1);
        detail::flatten_qp(rlk.Value[1].Value[1].Value, d1);
        const int g = parties_.id(id);
        const int st = mkhip_register_bfv_party(ctx_.c, g, b0.data(), d0.data(), v0.data(), b1.data(), d1.data());
        if (st != MKHIP_OK) throw Panic("AddRelinearizationKey: " + status_text(st), st);
        relin_.insert(id);
        return g;
    }
    Parameters params_;
    detail::Ctx ctx_;
    detail::PartyTable parties_;
    std::set<std::string> relin_;
    detail::GaloisKeys galois_;
};

}  // namespace mkbfv

namespace mkf {

// rlwe.SwitchingKey of Lattigo (the MKF construction): per digit, two PolyQP outputs.
struct SwitchingKey {
    std::vector<std::array<PolyQP, 2>> Value;
};

// onlineKeys (common.go:27-30): fold keys s_i -> s_aggregate and the
// relinearization key of s_aggregate.
struct OnlineKeys {
    std::vector<SwitchingKey> Fold;
    SwitchingKey Relin;
    // party-specific Galois keys sigma(s_i) -> s_aggregate by Galois element
    // (paperMKFCKKSOnlineKeys.Rotations, true_mkf_ckks.go:30-34, 370-392)
    std::map<uint64_t, std::vector<SwitchingKey>> Rotations;
};

// multiCiphertext (common.go:10-15) and outputCiphertext (:19-25).
struct MultiCiphertext {
    Poly C0;
    std::vector<Poly> Components;
    int Level = 0;
    double Scale = 1.0;
};
struct OutputCiphertext {
    Poly C0, Gamma;
    std::vector<Poly> Components;
    int Level = 0;
    double Scale = 1.0;
};

namespace detail2 {
inline std::vector<uint64_t> flatten_swk(const SwitchingKey &k) {
    std::vector<uint64_t> out;
    for (const auto &d : k.Value)
        for (int o = 0; o < 2; ++o) {
            d[o].Q.append_to(out, (int)d[o].Q.Coeffs.size());
            d[o].P.append_to(out, (int)d[o].P.Coeffs.size());
        }
    return out;
}
inline std::vector<uint64_t> flatten_mc(const MultiCiphertext &m, int nlimbs) {
    std::vector<uint64_t> out;
    m.C0.append_to(out, nlimbs);
    for (const Poly &p : m.Components) p.append_to(out, nlimbs);
    return out;
}
inline OutputCiphertext unflatten(const std::vector<uint64_t> &out, int k, int N, int nl, bool ntt, int level,
                                  double scale) {
    OutputCiphertext o;
    o.C0 = Poly::from(out.data(), N, nl, ntt);
    o.Gamma = Poly::from(out.data() + (size_t)nl * N, N, nl, ntt);
    for (int p = 0; p < k; ++p) o.Components.push_back(Poly::from(out.data() + (size_t)(1 + p) * nl * N, N, nl, ntt));
    o.Level = level;
    o.Scale = scale;
    return o;
}
}  // namespace detail2

// CKKSCase.Run (ckks.go:102-151) for one party group: the arithmetic of the
// fold / Karatsuba / relinearization / rescale / materialize pipeline.  Errors
// are returned (empty string = nil), as the reference returns `error`.
class CKKSEvaluator {
  public:
    CKKSEvaluator(const mkrlwe::Parameters &params, double scale) : params_(params), scale_(scale) {
        const int st = mkhip_ctx_create(&ctx_.c, params.LogN, params.Q.data(), (int)params.Q.size(),
                                        params.P.data(), (int)params.P.size(), 1, 0, nullptr, 0);
        if (st != MKHIP_OK) err_ = status_text(st);
    }
    std::string Err() const { return err_; }
    // keys of a party group (one per party set, common.go:44-51)
    std::string SetKeys(const OnlineKeys &keys) {
        if (!err_.empty()) return err_;
        std::vector<std::vector<uint64_t>> fold;
        std::vector<const uint64_t *> fp;
        for (const auto &f : keys.Fold) fold.push_back(detail2::flatten_swk(f));
        for (const auto &f : fold) fp.push_back(f.data());
        const std::vector<uint64_t> relin = detail2::flatten_swk(keys.Relin);
        mkhip_unregister_mkf_group(ctx_.c, 0);
        int st = mkhip_register_mkf_group(ctx_.c, 0, (int)fp.size(), fp.data(), relin.data());
        if (st != MKHIP_OK) return status_text(st);
        k_ = (int)fp.size();
        rotations_.clear();
        for (const auto &kv : keys.Rotations) {
            if ((int)kv.second.size() != k_) continue;  // RotateNew reports the set as missing
            std::vector<std::vector<uint64_t>> rk;
            std::vector<const uint64_t *> rp;
            for (const auto &f : kv.second) rk.push_back(detail2::flatten_swk(f));
            for (const auto &f : rk) rp.push_back(f.data());
            const int rot = rotation_of(kv.first);
            if (rot < 0) continue;
            st = mkhip_register_mkf_rotation(ctx_.c, 0, rot, rp.data());
            if (st != MKHIP_OK) return status_text(st);
            rotations_.insert(kv.first);
        }
        return "";
    }
    // RotateNew (true_mkf_ckks.go:323-353): NTT-domain input, output C0 and k
    // copies of the second component at the input's level and scale.
    std::pair<MultiCiphertext, std::string> RotateNew(const MultiCiphertext &in, int rotation) {
        if (!err_.empty()) return {MultiCiphertext(), err_};
        if ((int)in.Components.size() != k_) return {MultiCiphertext(), "component count does not match the keys"};
        const uint64_t gal = galois_element(rotation);
        if (!rotations_.count(gal))
            return {MultiCiphertext(), "party-specific Galois keys missing for rotation " + std::to_string(rotation) +
                                           " (element " + std::to_string(gal) + ")"};
        const int N = params_.N(), nl = in.Level + 1;
        const std::vector<uint64_t> x = detail2::flatten_mc(in, nl);
        std::vector<uint64_t> out(x.size());
        const int st = mkhip_mkf_ckks_rotate(ctx_.c, 0, x.data(), in.Level, rotation, out.data());
        if (st != MKHIP_OK) return {MultiCiphertext(), status_text(st)};
        MultiCiphertext r;
        r.C0 = Poly::from(out.data(), N, nl, true);
        for (int p = 0; p < k_; ++p) r.Components.push_back(Poly::from(out.data() + (size_t)(1 + p) * nl * N, N, nl, true));
        r.Level = in.Level;
        r.Scale = in.Scale;
        return {r, ""};
    }
    std::pair<OutputCiphertext, std::string> MulRelin(const MultiCiphertext &left, const MultiCiphertext &right) {
        if (!err_.empty()) return {OutputCiphertext(), err_};
        if ((int)left.Components.size() != k_ || (int)right.Components.size() != k_)
            return {OutputCiphertext(), "party count does not match the registered keys"};
        const int N = params_.N(), level = std::min(left.Level, right.Level);
        const std::vector<uint64_t> l = detail2::flatten_mc(left, level + 1), r = detail2::flatten_mc(right, level + 1);
        std::vector<uint64_t> out((size_t)(1 + k_) * (level + 1) * N);
        int level_out = 0;
        double scale_out = 0;
        const int st = mkhip_mkf_ckks_mulrelin(ctx_.c, 0, l.data(), r.data(), level, left.Scale, right.Scale, scale_,
                                               out.data(), &level_out, &scale_out);
        if (st != MKHIP_OK) return {OutputCiphertext(), status_text(st)};
        return {detail2::unflatten(out, k_, N, level_out + 1, true, level_out, scale_out), ""};
    }
    // ---- paperMKFCKKSEvaluator's public operations (true_mkf_ckks.go:187-303) and
    // EvaluatePoly (true_mkf_ckks_polynomial.go:79-111); NTT-domain ciphertexts,
    // errors returned as the reference returns them.
    struct Polynomial {  // ckks.Polynomial: Coeffs, MaxDeg, Lead, Basis (A, B: the caller's map)
        std::vector<std::complex<double>> Coeffs;
        int MaxDeg = 0;
        bool Lead = true;
        int Basis = MKHIP_BASIS_STANDARD;
    };
    std::pair<MultiCiphertext, std::string> EvaluatePoly(const MultiCiphertext &in, const Polynomial &p,
                                                         double targetScale) {
        if (!err_.empty()) return {MultiCiphertext(), err_};
        if ((int)in.Components.size() != k_) return {MultiCiphertext(), "component count mismatch"};
        std::vector<double> re, im;
        for (const auto &c : p.Coeffs) {
            re.push_back(c.real());
            im.push_back(c.imag());
        }
        mkhip_polynomial mp = {re.data(), im.data(), (int)re.size(), p.MaxDeg, p.Lead ? 1 : 0, p.Basis};
        int lo = 0;
        double so = 0;
        int st = mkhip_mkf_poly_query(ctx_.c, 0, in.Level, in.Scale, &mp, targetScale, &lo, &so, nullptr);
        if (st != MKHIP_OK) return {MultiCiphertext(), status_text(st)};
        const std::vector<uint64_t> x = detail2::flatten_mc(in, in.Level + 1);
        std::vector<uint64_t> out((size_t)(1 + k_) * (lo + 1) * params_.N());
        st = mkhip_mkf_ckks_evaluate_poly(ctx_.c, 0, x.data(), in.Level, in.Scale, &mp, targetScale, out.data(), &lo,
                                          &so);
        if (st != MKHIP_OK) return {MultiCiphertext(), status_text(st)};
        return {state(out, lo, so), ""};
    }
    std::pair<MultiCiphertext, std::string> AddNew(const MultiCiphertext &a, const MultiCiphertext &b) {
        return add(a, b, 0);
    }
    std::pair<MultiCiphertext, std::string> SubNew(const MultiCiphertext &a, const MultiCiphertext &b) {
        return add(a, b, 1);
    }
    std::pair<MultiCiphertext, std::string> AddConstNew(const MultiCiphertext &a, std::complex<double> c) {
        if (!err_.empty()) return {MultiCiphertext(), err_};
        const std::vector<uint64_t> x = detail2::flatten_mc(a, a.Level + 1);
        std::vector<uint64_t> out(x.size());
        const int st = mkhip_mkf_ckks_add_const(ctx_.c, (int)a.Components.size(), x.data(), a.Level, a.Scale,
                                                c.real(), c.imag(), out.data());
        if (st != MKHIP_OK) return {MultiCiphertext(), status_text(st)};
        return {state(out, a.Level, a.Scale), ""};
    }
    std::pair<MultiCiphertext, std::string> MultByConstNew(const MultiCiphertext &a, std::complex<double> c) {
        if (!err_.empty()) return {MultiCiphertext(), err_};
        const std::vector<uint64_t> x = detail2::flatten_mc(a, a.Level + 1);
        std::vector<uint64_t> out(x.size());
        double so = 0;
        const int st = mkhip_mkf_ckks_mult_by_const(ctx_.c, (int)a.Components.size(), x.data(), a.Level, a.Scale,
                                                    c.real(), c.imag(), out.data(), &so);
        if (st != MKHIP_OK) return {MultiCiphertext(), status_text(st)};
        return {state(out, a.Level, so), ""};
    }
    std::pair<MultiCiphertext, std::string> RescaleNew(const MultiCiphertext &a, double targetScale) {
        if (!err_.empty()) return {MultiCiphertext(), err_};
        const std::vector<uint64_t> x = detail2::flatten_mc(a, a.Level + 1);
        std::vector<uint64_t> out(x.size());
        int lo = 0;
        double so = 0;
        const int st = mkhip_mkf_ckks_rescale(ctx_.c, (int)a.Components.size(), x.data(), a.Level, a.Scale,
                                              targetScale, out.data(), &lo, &so);
        if (st != MKHIP_OK) return {MultiCiphertext(), status_text(st)};
        out.resize((size_t)(1 + a.Components.size()) * (lo + 1) * params_.N());  // packed at the new level
        return {state(out, lo, so), ""};
    }
    std::pair<MultiCiphertext, std::string> DropLevelNew(const MultiCiphertext &a, int levels) {
        if (levels < 0 || a.Level - levels < 0)
            return {MultiCiphertext(), "invalid level drop " + std::to_string(levels) + " from level " +
                                           std::to_string(a.Level)};
        const std::vector<uint64_t> x = detail2::flatten_mc(a, a.Level - levels + 1);
        return {state(x, a.Level - levels, a.Scale), ""};
    }
    // the evaluator's MulRelinNew: no rescale, operands at their own levels
    std::pair<MultiCiphertext, std::string> MulRelinNew(const MultiCiphertext &a, const MultiCiphertext &b) {
        if (!err_.empty()) return {MultiCiphertext(), err_};
        if ((int)a.Components.size() != k_ || (int)b.Components.size() != k_)
            return {MultiCiphertext(), "component count mismatch"};
        const std::vector<uint64_t> x = detail2::flatten_mc(a, a.Level + 1), y = detail2::flatten_mc(b, b.Level + 1);
        const int lo = std::min(a.Level, b.Level);
        std::vector<uint64_t> out((size_t)(1 + k_) * (lo + 1) * params_.N());
        const int st = mkhip_mkf_ckks_mulrelin_levels(ctx_.c, 0, x.data(), a.Level, y.data(), b.Level, out.data());
        if (st != MKHIP_OK) return {MultiCiphertext(), status_text(st)};
        return {state(out, lo, a.Scale * b.Scale), ""};
    }
    mkhip_ctx *handle() const { return ctx_.c; }

  private:
    MultiCiphertext state(const std::vector<uint64_t> &w, int level, double scale) const {
        const int N = params_.N(), nl = level + 1;
        const int k = (int)(w.size() / ((size_t)nl * N)) - 1;
        MultiCiphertext r;
        r.C0 = Poly::from(w.data(), N, nl, true);
        for (int p = 0; p < k; ++p) r.Components.push_back(Poly::from(w.data() + (size_t)(1 + p) * nl * N, N, nl, true));
        r.Level = level;
        r.Scale = scale;
        return r;
    }
    std::pair<MultiCiphertext, std::string> add(const MultiCiphertext &a, const MultiCiphertext &b, int sub) {
        if (!err_.empty()) return {MultiCiphertext(), err_};
        if (a.Components.size() != b.Components.size()) return {MultiCiphertext(), "component count mismatch"};
        const std::vector<uint64_t> x = detail2::flatten_mc(a, a.Level + 1), y = detail2::flatten_mc(b, b.Level + 1);
        const int lv = std::min(a.Level, b.Level);
        std::vector<uint64_t> out((size_t)(1 + a.Components.size()) * (lv + 1) * params_.N());
        int lo = 0;
        double so = 0;
        const int st = mkhip_mkf_ckks_add(ctx_.c, (int)a.Components.size(), x.data(), a.Level, a.Scale, y.data(),
                                          b.Level, b.Scale, sub, out.data(), &lo, &so);
        if (st != MKHIP_OK) return {MultiCiphertext(), status_text(st)};
        return {state(out, lo, so), ""};
    }
    // GaloisElementForColumnRotationBy: 5^(k mod 2N) mod 2N
    uint64_t galois_element(long long k) const {
        const uint64_t mask = 2 * (uint64_t)params_.N() - 1;
        uint64_t r = 1, b = 5, e = (uint64_t)k & mask;
        for (; e; e >>= 1, b = (b * b) & mask)
            if (e & 1) r = (r * b) & mask;
        return r;
    }
    // a rotation in [0, N/2) with this Galois element, or -1
    int rotation_of(uint64_t gal) const {
        for (int r = 0; r < params_.N() / 2; ++r)
            if (galois_element(r) == gal) return r;
        return -1;
    }
    mkrlwe::Parameters params_;
    double scale_;
    mkhip_host::detail::Ctx ctx_;
    std::string err_;
    int k_ = 0;
    std::set<uint64_t> rotations_;
};

// BFVCase.Run (bfv.go:212-252): coefficient-domain operands at the top level.
class BFVEvaluator {
  public:
    explicit BFVEvaluator(const mkbfv::Parameters &params) : params_(params) {
        const int st = mkhip_ctx_create(&ctx_.c, params.LogN, params.Q.data(), (int)params.Q.size(),
                                        params.P.data(), (int)params.P.size(), 1, params.T, params.QMul.data(), 0);
        if (st != MKHIP_OK) err_ = status_text(st);
    }
    std::string Err() const { return err_; }
    std::string SetKeys(const OnlineKeys &keys) {
        if (!err_.empty()) return err_;
        std::vector<std::vector<uint64_t>> fold;
        std::vector<const uint64_t *> fp;
        for (const auto &f : keys.Fold) fold.push_back(detail2::flatten_swk(f));
        for (const auto &f : fold) fp.push_back(f.data());
        const std::vector<uint64_t> relin = detail2::flatten_swk(keys.Relin);
        mkhip_unregister_mkf_group(ctx_.c, 0);
        const int st = mkhip_register_mkf_group(ctx_.c, 0, (int)fp.size(), fp.data(), relin.data());
        if (st != MKHIP_OK) return status_text(st);
        k_ = (int)fp.size();
        return "";
    }
    std::pair<OutputCiphertext, std::string> MulRelin(const MultiCiphertext &left, const MultiCiphertext &right) {
        if (!err_.empty()) return {OutputCiphertext(), err_};
        if ((int)left.Components.size() != k_ || (int)right.Components.size() != k_)
            return {OutputCiphertext(), "party count does not match the registered keys"};
        const int N = params_.N(), L = (int)params_.Q.size();
        const std::vector<uint64_t> l = detail2::flatten_mc(left, L), r = detail2::flatten_mc(right, L);
        std::vector<uint64_t> out((size_t)(1 + k_) * L * N);
        const int st = mkhip_mkf_bfv_mulrelin(ctx_.c, 0, l.data(), r.data(), out.data());
        if (st != MKHIP_OK) return {OutputCiphertext(), status_text(st)};
        return {detail2::unflatten(out, k_, N, L, false, L - 1, 1.0), ""};
    }
    mkhip_ctx *handle() const { return ctx_.c; }

  private:
    mkbfv::Parameters params_;
    mkhip_host::detail::Ctx ctx_;
    std::string err_;
    int k_ = 0;
};

}  // namespace mkf

}  // namespace mkhip_host
