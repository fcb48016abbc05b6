// Host-side exchange plan: the ghost-layer copies of one half-step as boxes,
// and the merge of the reference's sequential x, y, z copy phases into one
// set of independent copies.  Plain C++ (no HIP), so the planner is also
// tested on the CPU.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <vector>

namespace lbm {

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
