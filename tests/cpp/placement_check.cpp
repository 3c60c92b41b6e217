// Host check of the LDS placements of Pow2Line exchanges: for every
// placement (padding, the three block swizzles, the per-line key) and every
// radix a plan exchanges after, element m of a line must land in its own
// slot, inside the line's LS_RAW slots.  Prints one line per failure and
// "placements ok" when none fail.
#include <cstdio>
#include <vector>

#include "qsim_fft_core.h"

using namespace qsim;

static int failures = 0;

template <class L, int R>
static void check_radix(const char *name)
{
    if constexpr (R > 1) {
        for (int key = 0; key < 8; ++key) {
            std::vector<int> seen(L::LS_RAW, -1);
            for (int m = 0; m < L::E * L::TPL; ++m) {
                const int s = L::template pad<R>(m, key);
                if (s < 0 || s >= L::LS_RAW) {
                    std::printf("%s R=%d key=%d: element %d -> slot %d outside [0, %d)\n", name, R,
                                key, m, s, L::LS_RAW);
                    ++failures;
                    return;
                }
                if (seen[s] >= 0) {
                    std::printf("%s R=%d key=%d: elements %d and %d share slot %d\n", name, R, key,
                                seen[s], m, s);
                    ++failures;
                    return;
                }
                seen[s] = m;
            }
        }
    }
}

template <class B>
static void check_line(const char *name)
{
    // every placement of this plan; radices RA..RD are the ones exchanges follow
    auto each = [&](auto line, const char *what) {
        using L = decltype(line);
        char tag[128];
        std::snprintf(tag, sizeof tag, "%s %s", name, what);
        check_radix<L, L::RA_>(tag);
        check_radix<L, L::RB_>(tag);
        check_radix<L, L::RC_>(tag);
        check_radix<L, L::RD_>(tag);
    };
    each(typename B::template WithSwz<0>{}, "pad");
    each(typename B::template WithSwz<1>{}, "swz1");
    each(typename B::template WithSwz<2>{}, "swz2");
    if constexpr (B::E * B::TPL >= 16) each(typename B::template WithSwz<3>{}, "swz3");
    each(typename B::template WithSwz<4>{}, "key");
}

int main()
{
    check_line<Pow2Line<double, 16, 4, 4, 1>>("16 (4,4)");
    check_line<Pow2Line<double, 128, 8, 8, 2>>("128 (8,8,2)");
    check_line<Pow2Line<float, 128, 16, 8, 1>>("128 (16,8)");
    check_line<Pow2Line<double, 512, 8, 8, 8>>("512 (8,8,8)");
    check_line<Pow2Line<double, 1024, 8, 8, 16>>("1024 (8,8,16)");
    check_line<Pow2Line<double, 1024, 8, 8, 4, 4>>("1024 (8,8,4,4)");
    check_line<Pow2Line<double, 1024, 32, 32, 1>>("1024 (32,32)");
    check_line<Pow2Line<double, 2048, 16, 16, 8>>("2048 (16,16,8)");
    check_line<Pow2Line<double, 4096, 8, 8, 8, 8>>("4096 (8,8,8,8)");
    check_line<Pow2Line<float, 8192, 16, 8, 8, 8>>("8192 (16,8,8,8)");
    if (failures == 0) std::printf("placements ok\n");
    return failures == 0 ? 0 : 1;
}
