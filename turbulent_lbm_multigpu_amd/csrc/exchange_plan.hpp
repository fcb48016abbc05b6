// Host-side exchange plan: the chains of x-joined members the step kernel
// exchanges in LDS, the ghost-layer copies of one half-step as boxes, and the
// merge of the reference's sequential x, y, z copy phases into one set of
// independent copies.  Plain C++ (no HIP), so the planner is also tested on
// the CPU.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <vector>

namespace lbm {

// ---- chains of x-joined members ---------------------------------------------
// Same-sized members joined across x faces form chains (lines, or rings where
// x is periodic); the step kernel's chain tiles hand the face values over in
// LDS (see chain_tile in lbm_device.hpp).
constexpr int kMaxChain = 8;    // members of one chain
constexpr int kChainRows = 16;  // members x rows of one chain tile (LDS capacity)

// Joins of one member: nb[side] = the member whose inner column feeds my x-
// (side 0) / x+ (side 1) ghost layer inside the step kernel, or -1; mask[side]
// = the slots that ghost layer takes after alpha.
struct XJoin {
    int nb[2];
    uint32_t mask[2];
};

// A fusable x comm: `member` receives across its face `side` from `dst`.
struct JoinCandidate {
    int member, side, dst;
    uint32_t mask;
};

// The joins the candidates make: a face named by two candidates joins
// nothing, and a join holds only when both members name each other.
inline std::vector<XJoin> resolve_joins(int nm, const std::vector<JoinCandidate> &cand) {
    std::vector<XJoin> j((size_t)nm, XJoin{{-1, -1}, {0u, 0u}});
    for (const JoinCandidate &c : cand) {
        XJoin &f = j[c.member];
        if (f.nb[c.side] != -1) {
            f.nb[c.side] = -2;  // conflict
            continue;
        }
        f.nb[c.side] = c.dst;
        f.mask[c.side] = c.mask;
    }
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < nm; ++i)
            for (int side = 0; side < 2; ++side) {
                const int k = j[i].nb[side];
                if (k == -2 || (k >= 0 && j[k].nb[1 - side] != i)) j[i].nb[side] = -1;
            }
    for (XJoin &f : j)
        for (int side = 0; side < 2; ++side)
            if (f.nb[side] < 0) f.mask[side] = 0u;
    return j;
}

// 64-cell runs of an x row that hold columns 0, 1, X-2 and X-1 (distinct).
inline int chain_edge_runs(int X, int *edge) {
    int n = 0;
    const int cand[3] = {0, (X - 2) / 64, (X - 1) / 64};
    for (int c : cand) {
        bool seen = false;
        for (int k = 0; k < n; ++k) seen |= edge[k] == c;
        if (!seen) edge[n++] = c;
    }
    return n;
}

// How a chain of M members with rows of X cells is tiled for work-groups of
// wg work-items:
//   whole rows: `rows` rows of every member, `lanes` work-items per member
//     (rows * X rounded up to whole wavefronts), all in one work-group;
//   run tiles (rows too wide): the 64-cell runs holding columns 0, 1, X-2,
//     X-1 of `rows` rows of every member per handoff tile, their interior
//     runs 1 .. ninner in `ipasses` tiles of their own.
struct ChainShape {
    bool fits;
    int rows, lanes, nedge, edge[3], ninner, ipasses;
};

inline ChainShape chain_shape(int M, int X, int wg) {
    ChainShape c{false, 0, 0, 0, {0, 0, 0}, 0, 0};
    if (M < 1 || M > kMaxChain || X < 4) return c;
    const int waves = wg / 64;
    auto lanes_of = [X](int r) { return (r * X + 63) / 64 * 64; };
    if (M * lanes_of(1) <= wg) {
        int R = kChainRows / M;
        while (R > 1 && M * lanes_of(R) > wg) --R;
        c.fits = true;
        c.rows = R;
        c.lanes = lanes_of(R);
        return c;
    }
    c.nedge = chain_edge_runs(X, c.edge);
    if (M * c.nedge > waves) {
        c.nedge = 0;
        return c;
    }
    c.fits = true;
    c.rows = std::max(1, std::min(waves / (M * c.nedge), kChainRows / M));
    c.ninner = std::max(0, (X - 2) / 64 - 1);
    c.ipasses = (M * c.ninner * c.rows + waves - 1) / waves;
    return c;
}

// The chains the joins form, each in x order: lines from the member with no
// left partner, then rings.  A chain that does not fit (chain_shape with its
// members' X) is left out and its members' joins are cleared.
// A periodic ring of one member in a kernel of packed pairs (two-byte storage,
// two cells per work-item) keeps its seam inside one wavefront per row
// (seam_pairs in lbm_device.hpp): rows of an even length of at least 128 cells.
inline bool seam_ring_ok(bool pairs, int X) { return pairs && X % 2 == 0 && X >= 128; }
// wavefronts per row of such a member: the seam wavefront, then runs of 64 pairs
inline int seam_waves_per_row(int X) { return 1 + (X / 2 - 64 + 63) / 64; }

inline std::vector<std::vector<int>> find_chains(std::vector<XJoin> &j, const std::vector<int> &X, int wg,
                                                 bool pairs = false) {
    const int nm = (int)j.size();
    std::vector<char> seen((size_t)nm, 0);
    auto walk = [&](int start) {
        std::vector<int> ch;
        for (int cur = start; cur >= 0 && !seen[cur]; cur = j[cur].nb[1]) {
            seen[cur] = 1;
            ch.push_back(cur);
        }
        return ch;
    };
    std::vector<std::vector<int>> found, kept;
    for (int i = 0; i < nm; ++i)
        if (!seen[i] && j[i].nb[0] < 0 && j[i].nb[1] >= 0) found.push_back(walk(i));
    for (int i = 0; i < nm; ++i)
        if (!seen[i] && j[i].nb[0] >= 0) found.push_back(walk(i));
    for (auto &ch : found) {
        const bool ring_of_one = ch.size() == 1 && j[ch[0]].nb[0] == ch[0];
        if (chain_shape((int)ch.size(), X[ch[0]], wg).fits || (ring_of_one && seam_ring_ok(pairs, X[ch[0]]))) {
            kept.push_back(ch);
            continue;
        }
        for (int m : ch) j[m] = XJoin{{-1, -1}, {0u, 0u}};
    }
    return kept;
}

// One copy of the exchange plan on the host: the box `sz` at `so` of member
// sm to `dof` of member dm, populations in `mask`.
struct HOp {
    int sm, dm;
    int so[3], dof[3], sz[3];
    uint32_t mask;
};

inline bool box_isect(const int *ao, const int *as, const int *bo, const int *bs, int *io, int *is) {
    for (int a = 0; a < 3; ++a) {
        const int lo = std::max(ao[a], bo[a]), hi = std::min(ao[a] + as[a], bo[a] + bs[a]);
        if (hi <= lo) return false;
        io[a] = lo;
        is[a] = hi - lo;
    }
    return true;
}

// Box A without its sub-box I, as up to six disjoint boxes.
inline void box_minus(const int *ao, const int *as, const int *io, const int *is, std::vector<std::array<int, 6>> &out) {
    int lo[3] = {ao[0], ao[1], ao[2]}, hi[3] = {ao[0] + as[0], ao[1] + as[1], ao[2] + as[2]};
    for (int a = 0; a < 3; ++a) {
        const int cuts[2][2] = {{lo[a], io[a]}, {io[a] + is[a], hi[a]}};
        for (const auto &c : cuts) {
            if (c[1] <= c[0]) continue;
            std::array<int, 6> b{lo[0], lo[1], lo[2], hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
            b[a] = c[0];
            b[3 + a] = c[1] - c[0];
            out.push_back(b);
        }
        lo[a] = io[a];
        hi[a] = io[a] + is[a];
    }
}

// o restricted to the sub-box (origin bo, size bs) of its destination / of its source
inline HOp op_at_dst(const HOp &o, const int *bo, const int *bs) {
    HOp r = o;
    for (int a = 0; a < 3; ++a) {
        r.so[a] = o.so[a] + (bo[a] - o.dof[a]);
        r.dof[a] = bo[a];
        r.sz[a] = bs[a];
    }
    return r;
}
inline HOp op_at_src(const HOp &o, const int *bo, const int *bs) {
    HOp r = o;
    for (int a = 0; a < 3; ++a) {
        r.dof[a] = o.dof[a] + (bo[a] - o.so[a]);
        r.so[a] = bo[a];
        r.sz[a] = bs[a];
    }
    return r;
}

// The x, y and z copy phases of one exchange (src/CController.hpp:265-383
// run them in that order) as one set of independent copies: a copy that
// reads a cell an earlier phase writes takes that copy's source instead, and
// a write a later phase repeats is dropped.  Returns false, leaving the
// phases as they are, if a copy would still read a cell another one writes
// or two copies of one phase write the same slot.
inline bool merge_phases(const std::vector<HOp> phase[3], std::vector<HOp> &out) {
    std::vector<HOp> done;  // destinations (cell, population) disjoint
    int io[3], is[3];
    for (int p = 0; p < 3; ++p) {
        for (size_t a = 0; a < phase[p].size(); ++a)
            for (size_t b = a + 1; b < phase[p].size(); ++b) {
                const HOp &x = phase[p][a], &y = phase[p][b];
                if (x.dm == y.dm && (x.mask & y.mask) && box_isect(x.dof, x.sz, y.dof, y.sz, io, is)) return false;
            }
        std::vector<HOp> added;
        for (const HOp &o0 : phase[p]) {
            std::vector<HOp> work{o0};
            while (!work.empty()) {
                const HOp o = work.back();
                work.pop_back();
                bool split = false;
                for (const HOp &q : done) {
                    if (q.dm != o.sm || !(q.mask & o.mask) || !box_isect(o.so, o.sz, q.dof, q.sz, io, is)) continue;
                    HOp r = op_at_src(o, io, is);
                    HOp via = r;  // the slots q wrote: read where q reads them
                    via.sm = q.sm;
                    for (int a = 0; a < 3; ++a) via.so[a] = q.so[a] + (io[a] - q.dof[a]);
                    via.mask = o.mask & q.mask;
                    added.push_back(via);
                    if (o.mask & ~q.mask) {
                        r.mask = o.mask & ~q.mask;
                        work.push_back(r);
                    }
                    std::vector<std::array<int, 6>> rest;
                    box_minus(o.so, o.sz, io, is, rest);
                    for (const auto &b : rest) work.push_back(op_at_src(o, b.data(), b.data() + 3));
                    split = true;
                    break;
                }
                if (!split) added.push_back(o);
            }
        }
        for (const HOp &o : added) {  // this phase's writes replace earlier ones
            std::vector<HOp> next;
            for (const HOp &q : done) {
                if (q.dm != o.dm || !(q.mask & o.mask) || !box_isect(q.dof, q.sz, o.dof, o.sz, io, is)) {
                    next.push_back(q);
                    continue;
                }
                HOp in = op_at_dst(q, io, is);
                in.mask = q.mask & ~o.mask;
                if (in.mask) next.push_back(in);
                std::vector<std::array<int, 6>> rest;
                box_minus(q.dof, q.sz, io, is, rest);
                for (const auto &b : rest) next.push_back(op_at_dst(q, b.data(), b.data() + 3));
            }
            done.swap(next);
        }
        done.insert(done.end(), added.begin(), added.end());
    }
    for (const HOp &o : done)
        for (const HOp &q : done)
            if (q.dm == o.sm && (q.mask & o.mask) && box_isect(o.so, o.sz, q.dof, q.sz, io, is)) return false;
    out.swap(done);
    return true;
}

}  // namespace lbm
