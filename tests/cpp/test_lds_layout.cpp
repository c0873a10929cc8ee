// Exhaustive host-side check of the transforms' padded LDS layout: for every pass
// of every ring degree the kernels use, every lane's 16 element positions split
// exactly into a lane base plus a compile-time offset, that split survives the
// padding (one pad word after every 16), the padded words stay inside the
// workgroup's allocation, and no two (lane, value) pairs share a padded word.
// The kernels assert only the closed form of this condition (pad_ok).
#include <cstdio>
#include <vector>

#include "../../mkf_amd/csrc/ntt.hpp"

using namespace mkhip;

template <class P>
static bool check(const char *name) {
    bool ok = P::pad_ok;
    std::vector<int> seen(lds_words(P::N), 0);
    for (int tid = 0; tid < P::T; ++tid)
        for (int i = 0; i < 16; ++i) {
            const int e = P::elem(tid, i);
            if (e != P::base(tid) + P::off(i)) ok = false;
            if (lds_pad(e) != lds_pad(P::base(tid)) + lds_pad(P::off(i))) ok = false;
            const int w = lds_pad(P::base(tid)) + lds_pad(P::off(i));
            if (w < 0 || w >= lds_words(P::N) || seen[w]++) ok = false;
        }
    std::printf("%-24s %s\n", name, ok ? "ok" : "FAILED");
    return ok;
}

template <int LOGN>
static bool check_plan() {
    typedef NttPlan<LOGN> P;
    bool ok = true;
    ok &= check<typename P::F0>("F0");
    ok &= check<typename P::F1>("F1");
    ok &= check<typename P::F2>("F2");
    ok &= check<typename P::F3>("F3");
    ok &= check<typename P::I0>("I0");
    ok &= check<typename P::I1>("I1");
    ok &= check<typename P::I2>("I2");
    ok &= check<typename P::I3>("I3");
    return ok;
}

int main() {
    bool ok = true;
    std::printf("logN = 13 (and the halves of 2^14 / 2^15 limbs)\n");
    ok &= check_plan<13>();
    std::printf("logN = 14 (and the halves of 2^15 limbs)\n");
    ok &= check_plan<14>();
    std::printf(ok ? "all LDS layout checks passed\n" : "LDS layout check FAILED\n");
    return ok ? 0 : 1;
}
