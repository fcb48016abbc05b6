// CPU check of the exchange planner (turbulent_lbm_multigpu_amd/csrc/exchange_plan.hpp).
//
// For block decompositions built like CManager::initSimulation
// (src/CManager.hpp:78-204: one comm per neighbour face, send origin on the
// inner layer, receive origin on the ghost layer, the whole face), every
// subdomain slot gets a distinct value.  The reference's order — all x
// copies, then all y copies, then all z copies, each phase reading the state
// the previous one left (src/CController.hpp:265-383) — is played out and
// compared with the merged plan, whose copies all read the state before the
// exchange.  Prints one line per case: "<case> merged|phased ok" or a
// mismatch; exit status 1 on any mismatch.  The chain planning (joins, lines
// and rings, tile shapes) is checked against known cases first, one
// "chains <case> ok|FAIL" line each.
#include <array>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "exchange_plan.hpp"

using lbm::HOp;

// D3Q19 directions (src/CLbmSolver.hpp:42-48 order; only the sets matter here)
static const int E[19][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {1, 1, 0},   {-1, -1, 0}, {1, -1, 0},
                             {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {0, 1, 1},   {0, -1, -1},
                             {0, 1, -1}, {0, -1, 1}, {0, 0, 1},  {0, 0, -1}, {0, 0, 0}};

static uint32_t mask_dot(const int *n, int sign) {  // populations with sign * (n . e) > 0
    uint32_t m = 0;
    for (int i = 0; i < 19; ++i)
        if (sign * (n[0] * E[i][0] + n[1] * E[i][1] + n[2] * E[i][2]) > 0) m |= 1u << i;
    return m;
}

struct Case {
    int nums[3], s[3];
    bool periodic[3];
    int mode;        // 0: all 19 after alpha, 1: the 5 an inner cell reads
    bool after_beta;
    bool fused_x;    // x faces done before the copies (the step kernel's chain tiles)
};

static std::string name(const Case &c) {
    char b[160];
    std::snprintf(b, sizeof b, "nums=%dx%dx%d s=%dx%dx%d per=%d%d%d mode=%d %s%s", c.nums[0], c.nums[1], c.nums[2],
                  c.s[0], c.s[1], c.s[2], c.periodic[0], c.periodic[1], c.periodic[2], c.mode,
                  c.after_beta ? "beta" : "alpha", c.fused_x ? " fused-x" : "");
    return b;
}

// Returns 0 when both orders give the same memory; sets merged.
static int run(const Case &c, bool &merged) {
    const int nm = c.nums[0] * c.nums[1] * c.nums[2];
    std::vector<HOp> phase[3];
    for (int id = 0; id < nm; ++id) {
        const int co[3] = {id % c.nums[0], (id / c.nums[0]) % c.nums[1], id / (c.nums[0] * c.nums[1])};
        for (int a = 0; a < 3; ++a)
            for (int side = 0; side < 2; ++side) {
                int nc[3] = {co[0], co[1], co[2]};
                nc[a] += side ? 1 : -1;
                if (nc[a] < 0 || nc[a] >= c.nums[a]) {
                    if (!c.periodic[a]) continue;
                    nc[a] = (nc[a] + c.nums[a]) % c.nums[a];
                }
                const int dst = nc[0] + c.nums[0] * (nc[1] + c.nums[1] * nc[2]);
                int normal[3] = {0, 0, 0};
                normal[a] = side ? -1 : 1;  // inward normal of my face
                // I receive from `dst`: after alpha its inner layer -> my ghost layer,
                // after beta its ghost layer -> my inner layer (n.e > 0)
                HOp h{dst, id, {0, 0, 0}, {0, 0, 0}, {c.s[0], c.s[1], c.s[2]}, 0u};
                h.sz[a] = 1;
                const int my_ghost = side ? c.s[a] - 1 : 0, my_inner = side ? c.s[a] - 2 : 1;
                const int its_inner = side ? 1 : c.s[a] - 2, its_ghost = side ? 0 : c.s[a] - 1;
                if (!c.after_beta) {
                    h.so[a] = its_inner;
                    h.dof[a] = my_ghost;
                    h.mask = c.mode == 1 ? mask_dot(normal, -1) : (1u << 19) - 1;
                } else {
                    h.so[a] = its_ghost;
                    h.dof[a] = my_inner;
                    h.mask = mask_dot(normal, 1);
                }
                if (a == 0 && c.fused_x) continue;
                phase[a].push_back(h);
            }
    }
    const size_t cells = (size_t)c.s[0] * c.s[1] * c.s[2];
    auto at = [&](int x, int y, int z) { return (size_t)x + c.s[0] * ((size_t)y + c.s[1] * z); };
    std::vector<std::vector<unsigned>> st(nm, std::vector<unsigned>(19 * cells));
    unsigned v = 1;
    for (auto &m : st)
        for (auto &e : m) e = v++;
    auto apply = [&](const std::vector<HOp> &ops, const std::vector<std::vector<unsigned>> &src,
                     std::vector<std::vector<unsigned>> &dst) {
        for (const HOp &o : ops)
            for (int i = 0; i < 19; ++i) {
                if (!(o.mask & (1u << i))) continue;
                for (int z = 0; z < o.sz[2]; ++z)
                    for (int y = 0; y < o.sz[1]; ++y)
                        for (int x = 0; x < o.sz[0]; ++x)
                            dst[o.dm][i * cells + at(o.dof[0] + x, o.dof[1] + y, o.dof[2] + z)] =
                                src[o.sm][i * cells + at(o.so[0] + x, o.so[1] + y, o.so[2] + z)];
            }
    };
    std::vector<std::vector<unsigned>> seq = st;
    for (int p = 0; p < 3; ++p) {
        const std::vector<std::vector<unsigned>> before = seq;
        apply(phase[p], before, seq);
    }
    std::vector<HOp> plan;
    merged = lbm::merge_phases(phase, plan);
    std::vector<std::vector<unsigned>> one = st;
    if (merged) {
        apply(plan, st, one);
    } else {
        for (int p = 0; p < 3; ++p) {
            const std::vector<std::vector<unsigned>> before = one;
            apply(phase[p], before, one);
        }
    }
    for (int m = 0; m < nm; ++m)
        for (size_t e = 0; e < st[m].size(); ++e)
            if (one[m][e] != seq[m][e]) {
                std::printf("%s MISMATCH member %d pop %zu cell %zu: %u vs %u\n", name(c).c_str(), m, e / cells,
                            e % cells, one[m][e], seq[m][e]);
                return 1;
            }
    return 0;
}

// ---- chains of x-joined members ----------------------------------------------
static int check(bool ok, const char *what) {
    std::printf("chains %s %s\n", what, ok ? "ok" : "FAIL");
    return ok ? 0 : 1;
}

// joins from pairs (a receives on its side s from b): both directions listed
static std::vector<lbm::JoinCandidate> cands(std::initializer_list<std::array<int, 3>> l) {
    std::vector<lbm::JoinCandidate> v;
    for (const auto &t : l) v.push_back({t[0], t[1], t[2], 0x1fu});
    return v;
}

static int run_chains() {
    int f = 0;
    using V = std::vector<std::vector<int>>;
    {  // a line 0 - 1 - 2 along x: 1's x- from 0, 0's x+ from 1, ...
        auto j = lbm::resolve_joins(3, cands({{1, 0, 0}, {0, 1, 1}, {2, 0, 1}, {1, 1, 2}}));
        auto ch = lbm::find_chains(j, {128, 128, 128}, 512);
        f += check(ch == V{{0, 1, 2}} && j[0].nb[0] == -1 && j[0].nb[1] == 1 && j[2].nb[1] == -1, "line of three");
    }
    {  // a ring of two (x periodic across two members): each feeds both faces of the other
        auto j = lbm::resolve_joins(2, cands({{0, 0, 1}, {0, 1, 1}, {1, 0, 0}, {1, 1, 0}}));
        auto ch = lbm::find_chains(j, {100, 100}, 256);
        f += check(ch == V{{0, 1}} && j[0].nb[0] == 1 && j[0].nb[1] == 1, "ring of two");
    }
    {  // one member periodic in x: a ring of one
        auto j = lbm::resolve_joins(1, cands({{0, 0, 0}, {0, 1, 0}}));
        auto ch = lbm::find_chains(j, {24}, 256);
        f += check(ch == V{{0}} && j[0].nb[0] == 0 && j[0].nb[1] == 0, "ring of one");
    }
    {  // two comms on member 1's x- face: that face joins nothing, nor do the partners' faces
        auto j = lbm::resolve_joins(3, cands({{1, 0, 0}, {1, 0, 2}, {0, 1, 1}, {2, 1, 1}}));
        auto ch = lbm::find_chains(j, {64, 64, 64}, 256);
        f += check(ch.empty() && j[1].nb[0] == -1 && j[0].nb[1] == -1 && j[2].nb[1] == -1, "conflicting face");
    }
    {  // a join one side does not name back is dropped
        auto j = lbm::resolve_joins(3, cands({{1, 0, 0}, {0, 1, 2}}));
        auto ch = lbm::find_chains(j, {64, 64, 64}, 256);
        f += check(ch.empty() && j[1].nb[0] == -1 && j[1].mask[0] == 0u && j[0].nb[1] == -1, "one-sided join");
    }
    {  // nine members in a line: longer than kMaxChain, no chain
        std::vector<std::array<int, 3>> l;
        std::vector<lbm::JoinCandidate> c;
        for (int i = 1; i < 9; ++i) {
            c.push_back({i, 0, i - 1, 1u});
            c.push_back({i - 1, 1, i, 1u});
        }
        auto j = lbm::resolve_joins(9, c);
        auto ch = lbm::find_chains(j, std::vector<int>(9, 16), 256);
        bool cleared = true;
        for (const auto &x : j) cleared = cleared && x.nb[0] == -1 && x.nb[1] == -1;
        f += check(ch.empty() && cleared, "nine in a line");
    }
    {  // tile shapes
        lbm::ChainShape a = lbm::chain_shape(2, 128, 256);  // the benchmark: one row of both, 128 lanes each
        f += check(a.fits && a.rows == 1 && a.lanes == 128 && a.nedge == 0, "shape 2 x 128, wg 256");
        lbm::ChainShape b = lbm::chain_shape(2, 14, 256);   // 8 rows of 14 per member -> 112 -> 128 lanes
        f += check(b.fits && b.rows == 8 && b.lanes == 128, "shape 2 x 14, wg 256");
        lbm::ChainShape c = lbm::chain_shape(2, 128, 128);  // rows too wide, face runs {0, 1} x 2 > 2 waves
        f += check(!c.fits, "shape 2 x 128, wg 128");
        lbm::ChainShape d = lbm::chain_shape(1, 384, 256);  // runs {0, 5}, 2 rows per handoff tile, runs 1..4
        f += check(d.fits && d.nedge == 2 && d.edge[0] == 0 && d.edge[1] == 5 && d.rows == 2 && d.ninner == 4 &&
                       d.ipasses == 2,
                   "shape 1 x 384, wg 256");
        lbm::ChainShape e = lbm::chain_shape(2, 130, 256);  // runs {0, 2}: one row per handoff tile, run 1 inside
        f += check(e.fits && e.nedge == 2 && e.rows == 1 && e.ninner == 1 && e.ipasses == 1, "shape 2 x 130, wg 256");
        lbm::ChainShape g = lbm::chain_shape(1, 129, 256);  // one row on 192 lanes
        f += check(g.fits && g.nedge == 0 && g.rows == 1 && g.lanes == 192, "shape 1 x 129, wg 256");
        lbm::ChainShape k = lbm::chain_shape(1, 321, 256);  // X - 2 = 319 and X - 1 = 320 in runs 4 and 5
        f += check(k.fits && k.nedge == 3 && k.edge[0] == 0 && k.edge[1] == 4 && k.edge[2] == 5 && k.rows == 1 &&
                       k.ninner == 3 && k.ipasses == 1,
                   "shape 1 x 321, wg 256");
        lbm::ChainShape h = lbm::chain_shape(9, 16, 1024);
        f += check(!h.fits, "shape 9 members");
    }
    return f;
}

int main() {
    int fails = run_chains();
    const int nums_list[][3] = {{2, 2, 2}, {3, 1, 2}, {1, 3, 3}, {2, 2, 1}, {3, 3, 3}, {4, 1, 1}, {1, 1, 1}};
    const bool per_list[][3] = {{false, false, false}, {true, false, false}, {true, true, true}, {false, true, false}};
    for (const auto &nums : nums_list)
        for (const auto &per : per_list)
            for (int mode = 0; mode < 2; ++mode)
                for (int beta = 0; beta < 2; ++beta)
                    for (int fx = 0; fx < 2; ++fx) {
                        Case c{{nums[0], nums[1], nums[2]}, {6, 5, 7}, {per[0], per[1], per[2]}, mode, beta != 0,
                               fx != 0};
                        bool merged = false;
                        const int rc = run(c, merged);
                        fails += rc;
                        if (!rc) std::printf("%s %s ok\n", name(c).c_str(), merged ? "merged" : "phased");
                    }
    return fails ? 1 : 0;
}
