// Host check of map_stedc_info (imago_amd/csrc/hetrd_launch.hpp): the
// divide-and-conquer info of one half of a batch mapped to per-matrix
// codes.  Prints a line per case; exits 1 on the first mismatch.
#include "../../imago_amd/csrc/hetrd_launch.hpp"

#include <cstdio>
#include <vector>

using imago::map_stedc_info;

static int check(const char *name, int n, int nh, int info, std::vector<int> want)
{
    std::vector<int> got(nh, -7);
    map_stedc_info(n, nh, info, got.data());
    bool ok = got == want;
    std::printf("%-34s n=%d nh=%d info=%d ->", name, n, nh, info);
    for (int v : got) std::printf(" %d", v);
    std::printf("  %s\n", ok ? "ok" : "MISMATCH");
    return ok ? 0 : 1;
}

int main()
{
    const int n = 10, nh = 3, N = n * nh;   // rows 1..30: m0 1-10, m1 11-20, m2 21-30
    int bad = 0;
    bad |= check("success", n, nh, 0, {0, 0, 0});
    // submatrix rows 12..18 lie in the second matrix only
    bad |= check("inside matrix 2", n, nh, 12 * (N + 1) + 18, {0, n, 0});
    // rows 5..25 cross all three
    bad |= check("across all three", n, nh, 5 * (N + 1) + 25, {n, n, n});
    // rows 21..30: the last matrix
    bad |= check("last matrix", n, nh, 21 * (N + 1) + 30, {0, 0, n});
    // rows 1..10: the first matrix
    bad |= check("first matrix", n, nh, 1 * (N + 1) + 10, {n, 0, 0});
    // rows 9..11 straddle the first boundary
    bad |= check("straddles a boundary", n, nh, 9 * (N + 1) + 11, {n, n, 0});
    // a code naming no valid range (r0 = 0) marks the whole half
    bad |= check("no row range", n, nh, 7, {n, n, n});
    bad |= check("negative (argument error)", n, nh, -4, {n, n, n});
    // a single matrix half
    bad |= check("one matrix", 896, 1, 3 * 897 + 40, {896});
    return bad;
}
