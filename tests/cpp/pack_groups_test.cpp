// The ensemble's launch-group packing (uws_host.h: pack_groups) on the host:
//  - every item is in exactly one group, groups list their items in
//    increasing order, no group is empty;
//  - a group's weight stays within the capacity unless it holds one item
//    alone (an item above the capacity);
//  - the group count is never below the lower bound ceil(sum min(w, cap) / cap)
//    and, on small random instances, never above the optimum found by
//    exhaustive search plus one (the count it reaches is printed);
//  - the 24-case ensemble's weights (two-sweep workgroups at 16-plane chunks
//    and 13-row tiles) pack into 6 groups of 240.
// Exit status 0 when every case passes; the first failure is printed.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "uws_host.h"

using namespace uws;

namespace {

int g_fail = 0;
std::string g_case;

#define CHECK(cond)                                                                             \
    do {                                                                                        \
        if (!(cond)) {                                                                          \
            if (!g_fail) std::printf("FAIL [%s] line %d: %s\n", g_case.c_str(), __LINE__, #cond); \
            g_fail = 1;                                                                         \
        }                                                                                       \
    } while (0)

// fewest groups of weight <= cap (an oversized item alone), by trying every
// assignment of items to groups in order (first-empty-group symmetry cut)
int optimum(const std::vector<int>& w, int cap) {
    const int n = (int)w.size();
    int best = n;
    std::vector<int64_t> load;
    std::function<void(int)> rec = [&](int i) {
        if ((int)load.size() >= best) return;
        if (i == n) {
            best = (int)load.size();
            return;
        }
        for (size_t k = 0; k < load.size(); ++k) {
            if (load[k] + w[i] <= cap && load[k] <= cap) {
                load[k] += w[i];
                rec(i + 1);
                load[k] -= w[i];
            }
        }
        load.push_back(w[i]);
        rec(i + 1);
        load.pop_back();
    };
    rec(0);
    return best;
}

void check_packing(const std::vector<int>& w, int cap, int* groups_out) {
    const auto g = pack_groups(w, cap);
    std::vector<int> seen(w.size(), 0);
    int64_t lb = 0;
    for (int x : w) lb += std::min(x, cap);
    lb = (lb + cap - 1) / cap;
    for (const auto& grp : g) {
        CHECK(!grp.empty());
        int64_t s = 0;
        for (size_t j = 0; j < grp.size(); ++j) {
            CHECK(grp[j] >= 0 && grp[j] < (int)w.size());
            if (grp[j] < 0 || grp[j] >= (int)w.size()) return;
            seen[grp[j]]++;
            s += w[grp[j]];
            if (j) CHECK(grp[j - 1] < grp[j]);
        }
        CHECK(s <= cap || grp.size() == 1);
    }
    for (int c : seen) CHECK(c == 1);
    CHECK((int64_t)g.size() >= (w.empty() ? 0 : lb));
    *groups_out = (int)g.size();
}

}  // namespace

int main() {
    int cases = 0, at_optimum = 0;
    // the 24-case ensemble: 7 cases of 192 x 64 x 96 and 5 of 160 x 64 x 96
    // (90 workgroups each), 6 of 128 x 64 x 64 (40), 6 of 64^3 (20)
    {
        g_case = "ensemble24";
        std::vector<int> w;
        for (int i = 0; i < 7; ++i) w.push_back(90);
        for (int i = 0; i < 6; ++i) w.push_back(40);
        for (int i = 0; i < 5; ++i) w.push_back(90);
        for (int i = 0; i < 6; ++i) w.push_back(20);
        int ng = 0;
        check_packing(w, 256, &ng);
        CHECK(ng == 6);
        for (const auto& grp : pack_groups(w, 256)) {
            int s = 0;
            for (int c : grp) s += w[c];
            CHECK(s == 240);
        }
        ++cases;
    }
    // edge cases: nothing, one item, every item above the capacity, capacity 1
    {
        g_case = "empty";
        CHECK(pack_groups({}, 256).empty());
        g_case = "oversize";
        int ng = 0;
        check_packing({300, 500, 10}, 256, &ng);
        CHECK(ng == 3);
        g_case = "cap1";
        check_packing({1, 1, 1, 2}, 1, &ng);
        CHECK(ng == 4);
        g_case = "single";
        check_packing({7}, 256, &ng);
        CHECK(ng == 1);
        cases += 4;
    }
    // random small instances against the exhaustive optimum
    std::mt19937 rng(12345);
    for (int t = 0; t < 3000 && !g_fail; ++t) {
        const int n = 1 + (int)(rng() % 9);
        const int cap = 4 + (int)(rng() % 60);
        std::vector<int> w(n);
        for (int& x : w) x = 1 + (int)(rng() % (cap + cap / 4));
        g_case = "random " + std::to_string(t);
        int ng = 0;
        check_packing(w, cap, &ng);
        const int opt = optimum(w, cap);
        CHECK(ng >= opt);
        CHECK(ng <= opt + 1);
        at_optimum += ng == opt;
        ++cases;
    }
    if (g_fail) return 1;
    std::printf("%d cases passed (%d of the 3000 random ones at the exhaustive optimum)\n", cases, at_optimum);
    return 0;
}
