#include "plan.hpp"

#include "host.hpp"

#include <algorithm>
#include <atomic>
#include <exception>
#include <memory>
#include <stdexcept>
#include <thread>

namespace lsc {

namespace {

// Geometric nested dissection (reference: src/cpp/circuit/system.cpp:41-93).
// The far half of a box of 2^18 cells or more is cut on a host thread of its
// own (up to the host thread count in all) into a Dissection of its own,
// whose cells and fronts are then appended after the near half's: the
// layout of the serial recursion (near half, far half, separator).
struct Dissection {
    const std::uint8_t* mark;
    i64 stride;
    i64 leaf;
    std::vector<i32> cells;
    std::vector<i32> first, npiv;
    std::vector<std::vector<i32>> kids;
    std::atomic<int>* helpers = nullptr;  // threads running far halves (shared)
    int max_helpers = 0;

    i32 add(i32 f, std::vector<i32> ch) {
        first.push_back(f);
        npiv.push_back(static_cast<i32>(cells.size()) - f);
        kids.push_back(std::move(ch));
        return static_cast<i32>(first.size()) - 1;
    }

    // appends the cells and fronts of o; returns o's front index shift
    i32 append(Dissection& o) {
        const i32 shift = static_cast<i32>(first.size()), base = static_cast<i32>(cells.size());
        cells.insert(cells.end(), o.cells.begin(), o.cells.end());
        for (i32 f : o.first) first.push_back(f + base);
        npiv.insert(npiv.end(), o.npiv.begin(), o.npiv.end());
        for (auto& k : o.kids) {
            for (i32& c : k) c += shift;
            kids.push_back(std::move(k));
        }
        return shift;
    }

    i32 cut(i64 r0, i64 r1, i64 c0, i64 c1) {
        i64 top = r1, bottom = r0 - 1, left = c1, right = c0 - 1;
        for (i64 r = r0; r < r1; ++r) {
            const std::uint8_t* row = mark + r * stride;
            for (i64 c = c0; c < c1; ++c)
                if (row[c]) {
                    if (r < top) top = r;
                    bottom = r;
                    if (c < left) left = c;
                    if (c > right) right = c;
                }
        }
        if (bottom < top) return -1;
        r0 = top, r1 = bottom + 1, c0 = left, c1 = right + 1;
        const i64 h = r1 - r0, w = c1 - c0;
        if (h * w <= leaf) {
            const i32 f = static_cast<i32>(cells.size());
            for (i64 r = r0; r < r1; ++r)
                for (i64 c = c0; c < c1; ++c)
                    if (mark[r * stride + c]) cells.push_back(static_cast<i32>(r * stride + c));
            return add(f, {});
        }
        std::vector<i32> ch;
        ch.reserve(2);
        const bool by_row = h >= w;
        const i64 mid = by_row ? r0 + h / 2 : c0 + w / 2;
        const auto near = [&] { return by_row ? cut(r0, mid, c0, c1) : cut(r0, r1, c0, mid); };
        const auto far = [&](Dissection& d) { return by_row ? d.cut(mid + 1, r1, c0, c1) : d.cut(r0, r1, mid + 1, c1); };
        i32 a, b;
        if (helpers && h * w >= (i64(1) << 18) && helpers->fetch_add(1) < max_helpers) {
            Dissection d2{mark, stride, leaf, {}, {}, {}, {}, helpers, max_helpers};
            i32 rb = -1;
            std::exception_ptr err;
            std::thread t([&] {
                try {
                    rb = far(d2);
                } catch (...) {
                    err = std::current_exception();
                }
            });
            try {
                a = near();
            } catch (...) {
                t.join();
                helpers->fetch_sub(1);
                throw;
            }
            t.join();
            helpers->fetch_sub(1);
            if (err) std::rethrow_exception(err);
            const i32 shift = append(d2);
            b = rb >= 0 ? rb + shift : -1;
        } else {
            if (helpers && h * w >= (i64(1) << 18)) helpers->fetch_sub(1);
            a = near();
            b = far(*this);
        }
        if (a >= 0) ch.push_back(a);
        if (b >= 0) ch.push_back(b);
        const i32 sep = static_cast<i32>(cells.size());
        if (by_row) {
            for (i64 c = c0; c < c1; ++c)
                if (mark[mid * stride + c]) cells.push_back(static_cast<i32>(mid * stride + c));
        } else {
            for (i64 r = r0; r < r1; ++r)
                if (mark[r * stride + mid]) cells.push_back(static_cast<i32>(r * stride + mid));
        }
        if (static_cast<i32>(cells.size()) == sep && ch.size() == 1) return ch[0];
        return add(sep, std::move(ch));
    }
};

// Nested dissection of a graph without geometry (George's level-structure
// scheme; the reference orders networks by the same kind of method,
// src/cpp/circuit/system.cpp:96-270). A connected set is cut at one level of
// a breadth-first search from a pseudo-peripheral node, where edges only
// join consecutive levels: the smallest level that leaves at least a quarter
// of the nodes on each side (else the median level), keeping in the
// separator only the nodes that touch the far side. A disconnected set gives
// one subtree per component. Fronts are laid out in post-order.
struct GraphDissection {
    // fronts of one subtree in post-order: pivots in elimination order,
    // each front's first pivot, pivot count and children (front indices)
    struct Out {
        std::vector<i32> order, first, npiv;
        std::vector<std::vector<i32>> kids;
    };
    const std::vector<i64>& ptr;
    const i32* adj;
    std::size_t leaf;
    // stamp[v]: the search that owns v (unique per call, so a search only
    // ever matches its own nodes, also while other threads re-stamp theirs);
    // level[v] is touched only by the owner of v
    std::unique_ptr<std::atomic<i32>[]> stamp;
    std::vector<i32> level;
    std::atomic<i32> stamps{0};
    std::atomic<int> helpers{0};  // threads running the far halves of large sets
    int max_helpers;
    Out out;

    GraphDissection(i32 n, const std::vector<i64>& p, const i32* a, int lf)
        : ptr(p), adj(a), leaf(static_cast<std::size_t>(lf)), stamp(new std::atomic<i32>[n > 0 ? n : 1]),
          level(n, -1),
          max_helpers(host_threads() - 1) {
        for (i32 v = 0; v < n; ++v) stamp[v].store(-1, std::memory_order_relaxed);
    }

    static i32 add(Out& o, std::vector<i32>& piv, std::vector<i32> ch) {
        std::sort(piv.begin(), piv.end());
        o.first.push_back(static_cast<i32>(o.order.size()));
        o.npiv.push_back(static_cast<i32>(piv.size()));
        o.order.insert(o.order.end(), piv.begin(), piv.end());
        o.kids.push_back(std::move(ch));
        return static_cast<i32>(o.first.size()) - 1;
    }
    // appends subtree b after o's fronts; returns the index shift of b's fronts
    static i32 append(Out& o, Out& b) {
        const i32 shift = static_cast<i32>(o.first.size()), base = static_cast<i32>(o.order.size());
        for (i32 f : b.first) o.first.push_back(f + base);
        o.npiv.insert(o.npiv.end(), b.npiv.begin(), b.npiv.end());
        o.order.insert(o.order.end(), b.order.begin(), b.order.end());
        for (auto& k : b.kids) {
            for (i32& c : k) c += shift;
            o.kids.push_back(std::move(k));
        }
        return shift;
    }
    i32 owner(i32 u) const { return stamp[u].load(std::memory_order_relaxed); }

    // Breadth-first search over the nodes stamped s: levels, nodes in visit order.
    void bfs(i32 root, i32 s, std::vector<i32>& seen) {
        seen.clear();
        seen.push_back(root);
        level[root] = 0;
        for (std::size_t h = 0; h < seen.size(); ++h) {
            const i32 v = seen[h];
            for (i64 e = ptr[v]; e < ptr[v + 1]; ++e) {
                const i32 u = adj[e];
                if (u < 0 || owner(u) != s || level[u] >= 0) continue;
                level[u] = level[v] + 1;
                seen.push_back(u);
            }
        }
    }
    void unlevel(const std::vector<i32>& vs) {
        for (i32 v : vs) level[v] = -1;
    }

    // The subtree roots of `set` appended to `roots`, its fronts to `o`.
    void dissect(std::vector<i32> set, Out& o, std::vector<i32>& roots) {
        if (set.empty()) return;
        if (set.size() <= leaf) {
            roots.push_back(add(o, set, {}));
            return;
        }
        const i32 s = stamps.fetch_add(1, std::memory_order_relaxed);
        for (i32 v : set) stamp[v].store(s, std::memory_order_relaxed);
        std::vector<i32> seen;
        bfs(*std::min_element(set.begin(), set.end()), s, seen);
        if (seen.size() < set.size()) {
            std::vector<std::vector<i32>> comps{seen};
            for (i32 v : set)
                if (level[v] < 0) {
                    std::vector<i32> c;
                    bfs(v, s, c);
                    comps.push_back(std::move(c));
                }
            for (const auto& c : comps) unlevel(c);
            for (auto& c : comps) dissect(std::move(c), o, roots);
            return;
        }
        // pseudo-peripheral root: restart from the smallest node of the last
        // level while the depth grows (at most three searches)
        i32 depth = level[seen.back()];
        for (int it = 0; it < 2; ++it) {
            i32 far = seen.back();
            for (i32 v : seen)
                if (level[v] == depth && v < far) far = v;
            std::vector<i32> again;
            unlevel(seen);
            bfs(far, s, again);
            const i32 d = level[again.back()];
            seen.swap(again);
            if (d <= depth) break;
            depth = d;
        }
        depth = level[seen.back()];
        if (depth < 2) {  // no level has nodes on both sides
            unlevel(seen);
            roots.push_back(add(o, set, {}));
            return;
        }
        std::vector<i64> count(depth + 1, 0);
        for (i32 v : seen) ++count[level[v]];
        const i64 n = static_cast<i64>(seen.size());
        i32 cut = -1, median = -1;
        i64 before = count[0];
        for (i32 l = 1; l < depth; ++l) {
            const i64 after = n - before - count[l];
            if (median < 0 && before + count[l] >= n / 2) median = l;
            if (4 * before >= n && 4 * after >= n && (cut < 0 || count[l] < count[cut])) cut = l;
            before += count[l];
        }
        if (cut < 0) cut = median >= 0 ? median : depth / 2;
        std::vector<i32> a, b, sep;
        for (i32 v : seen) {
            const i32 l = level[v];
            if (l < cut) {
                a.push_back(v);
            } else if (l > cut) {
                b.push_back(v);
            } else {
                bool far = false;
                for (i64 e = ptr[v]; e < ptr[v + 1] && !far; ++e) {
                    const i32 u = adj[e];
                    far = u >= 0 && owner(u) == s && level[u] == cut + 1;
                }
                (far ? sep : a).push_back(v);
            }
        }
        unlevel(seen);
        std::vector<i32>().swap(set);
        std::vector<i32> ch;
        // the far half of a large set on a thread of its own; the fronts are
        // laid out as the serial recursion lays them out (a, then b)
        if (b.size() >= (std::size_t(1) << 16) && helpers.fetch_add(1) < max_helpers) {
            Out ob;
            std::vector<i32> rb;
            std::exception_ptr err;
            std::thread t([&] {
                try {
                    dissect(std::move(b), ob, rb);
                } catch (...) {
                    err = std::current_exception();
                }
            });
            try {
                dissect(std::move(a), o, ch);
            } catch (...) {
                t.join();
                helpers.fetch_sub(1);
                throw;
            }
            t.join();
            helpers.fetch_sub(1);
            if (err) std::rethrow_exception(err);
            const i32 shift = append(o, ob);
            for (i32 r : rb) ch.push_back(r + shift);
        } else {
            if (b.size() >= (std::size_t(1) << 16)) helpers.fetch_sub(1);
            dissect(std::move(a), o, ch);
            dissect(std::move(b), o, ch);
        }
        if (sep.empty() && ch.size() == 1) {
            roots.push_back(ch[0]);
            return;
        }
        roots.push_back(add(o, sep, std::move(ch)));
    }
};

}  // namespace

namespace {

// Fronts, levels and layouts of a plan whose unknowns (elimination order),
// fronts (first, npiv, children) and neighbour lists (nptr / nidx: unknown ->
// neighbouring unknowns, nsrc: the source id of each entry) are set.
void finish_plan(Plan& p, const std::vector<std::vector<i32>>& kids, const i64* nptr, const i32* nidx,
                 const i64* nsrc) {
    const i32 n = p.n, nf = p.nfront;
    progress("plan: fronts of %d unknowns", n);
    p.nupd.assign(nf, 0), p.upd.assign(nf, 0), p.cb.assign(nf, 0), p.ce.assign(nf, 0), p.height.assign(nf, 0);
    p.kids.reserve(nf);
    // 1. Update rows, children before parents: the rows past a front's pivots
    // that its pivots' entries or its children's update rows reach, ascending.
    // Fronts are in post-order, so the subtree of t is the fronts
    // t - size[t] + 1 .. t: subtrees of at most `grain` fronts are tasks for
    // host threads, each writing its fronts' rows into a buffer of its own,
    // and the fronts above them follow in order on this thread. A stamp per
    // row skips most repeats; subtrees running at once may stamp one
    // ancestor row in turn, so each list is also sorted and made unique.
    std::vector<i32> size(nf, 1), parent(nf, -1);
    for (i32 t = 0; t < nf; ++t) {
        p.cb[t] = static_cast<i32>(p.kids.size());
        i32 h = 0;
        for (i32 c : kids[t]) {
            p.kids.push_back(c);
            h = std::max(h, p.height[c] + 1);
            size[t] += size[c];
            parent[c] = t;
        }
        p.height[t] = h;
        p.ce[t] = static_cast<i32>(p.kids.size());
    }
    const i32 grain = std::max<i32>(1024, nf / (8 * host_threads()));
    std::vector<i32> tasks, top;
    {
        std::vector<i32> st;
        for (i32 t = 0; t < nf; ++t)
            if (parent[t] < 0) st.push_back(t);
        while (!st.empty()) {
            const i32 t = st.back();
            st.pop_back();
            if (size[t] <= grain) {
                tasks.push_back(t);
                continue;
            }
            top.push_back(t);
            for (i32 ci = p.cb[t]; ci < p.ce[t]; ++ci) st.push_back(p.kids[ci]);
        }
        std::sort(top.begin(), top.end());
        // larger subtrees first (ties by index), for balance
        std::sort(tasks.begin(), tasks.end(), [&](i32 a, i32 b) { return size[a] != size[b] ? size[a] > size[b] : a < b; });
    }
    std::unique_ptr<std::atomic<i32>[]> stamp(new std::atomic<i32>[n > 0 ? n : 1]);
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 v = lo; v < hi; ++v) stamp[v].store(-1, std::memory_order_relaxed);
    });
    std::vector<std::vector<i32>> buf(tasks.size() + 1);
    std::vector<i32> buf_of(nf, -1);
    std::vector<i64> off(nf, 0);
    const auto rows_of = [&](std::vector<i32>& out, i32 b, i32 t, std::vector<i32>& upd) {
        const i32 end = p.first[t] + p.npiv[t];
        upd.clear();
        for (i32 q = p.first[t]; q < end; ++q)
            for (i64 e = nptr[q]; e < nptr[q + 1]; ++e) {
                const i32 r = nidx[e];
                if (r >= end && stamp[r].load(std::memory_order_relaxed) != t) {
                    stamp[r].store(t, std::memory_order_relaxed);
                    upd.push_back(r);
                }
            }
        for (i32 ci = p.cb[t]; ci < p.ce[t]; ++ci) {
            const i32 c = p.kids[ci];
            const i32* cr = buf[buf_of[c]].data() + off[c];
            for (i32 i = 0; i < p.nupd[c]; ++i) {
                const i32 r = cr[i];
                if (r >= end && stamp[r].load(std::memory_order_relaxed) != t) {
                    stamp[r].store(t, std::memory_order_relaxed);
                    upd.push_back(r);
                }
            }
        }
        std::sort(upd.begin(), upd.end());
        upd.erase(std::unique(upd.begin(), upd.end()), upd.end());
        buf_of[t] = b;
        off[t] = static_cast<i64>(out.size());
        p.nupd[t] = static_cast<i32>(upd.size());
        out.insert(out.end(), upd.begin(), upd.end());
    };
    host_tasks(static_cast<i64>(tasks.size()), [&](i64 k) {
        std::vector<i32> upd;
        const i32 r = tasks[k];
        for (i32 t = r - size[r] + 1; t <= r; ++t) rows_of(buf[k], static_cast<i32>(k), t, upd);
    });
    {
        std::vector<i32> upd;
        for (i32 t : top) rows_of(buf[tasks.size()], static_cast<i32>(tasks.size()), t, upd);
    }
    i64 nrows = 0;
    for (i32 t = 0; t < nf; ++t) {
        if (nrows >= (i64(1) << 31)) throw std::invalid_argument("too many update rows for a 32-bit plan");
        p.upd[t] = static_cast<i32>(nrows);
        nrows += p.nupd[t];
        const i32 m = p.npiv[t] + p.nupd[t];
        p.max_front = std::max(p.max_front, m);
        const double np = p.npiv[t], nu = p.nupd[t];
        p.flops += np * np * np / 3 + np * np * nu + np * nu * nu;
    }
    p.rows.resize(static_cast<std::size_t>(nrows));
    host_chunks(nf, [&](i64, i64 lo, i64 hi) {
        for (i64 t = lo; t < hi; ++t)
            std::copy_n(buf[buf_of[t]].data() + off[t], p.nupd[t], p.rows.data() + p.upd[t]);
    });
    progress("plan: %zu update rows", p.rows.size());
    // 2. Over host threads, per front: the rows of its children in its own
    // front (relmap; every child has one parent) and the entries (q, r),
    // r > q, of its pivots at their front-local (row, column). A row r of
    // front t sits at r - first[t] when it is a pivot, else at npiv[t] plus
    // its place among t's sorted update rows.
    const auto local_row = [&](i32 t, i32 r) {
        const i32 end = p.first[t] + p.npiv[t];
        if (r < end) return r - p.first[t];
        const i32* u = p.rows.data() + p.upd[t];
        return p.npiv[t] + static_cast<i32>(std::lower_bound(u, u + p.nupd[t], r) - u);
    };
    p.relmap.assign(p.rows.size(), -1);
    std::vector<i64> fcount(static_cast<std::size_t>(nf), 0);
    host_chunks(nf, [&](i64, i64 lo, i64 hi) {
        for (i64 tt = lo; tt < hi; ++tt) {
            const i32 t = static_cast<i32>(tt), end = p.first[t] + p.npiv[t];
            for (i32 ci = p.cb[t]; ci < p.ce[t]; ++ci) {
                const i32 c = p.kids[ci];
                for (i32 i = 0; i < p.nupd[c]; ++i) p.relmap[p.upd[c] + i] = local_row(t, p.rows[p.upd[c] + i]);
            }
            i64 k = 0;
            for (i32 q = p.first[t]; q < end; ++q)
                for (i64 e = nptr[q]; e < nptr[q + 1]; ++e) k += nidx[e] > q;
            fcount[t] = k;
        }
    });
    p.fe_ptr.assign(static_cast<std::size_t>(nf) + 1, 0);
    {
        i64 tot = 0;
        for (i32 t = 0; t < nf; ++t) {
            tot += fcount[t];
            if (tot >= (i64(1) << 31)) throw std::invalid_argument("too many matrix entries for a 32-bit plan");
            p.fe_ptr[t + 1] = static_cast<i32>(tot);
        }
    }
    const std::size_t nent = static_cast<std::size_t>(p.fe_ptr[nf]);
    p.fe_row.resize(nent), p.fe_col.resize(nent), p.fe_p.resize(nent), p.fe_q.resize(nent), p.fe_src.resize(nent);
    host_chunks(nf, [&](i64, i64 lo, i64 hi) {
        for (i64 tt = lo; tt < hi; ++tt) {
            const i32 t = static_cast<i32>(tt), end = p.first[t] + p.npiv[t];
            i64 k = p.fe_ptr[t];
            for (i32 q = p.first[t]; q < end; ++q)
                for (i64 e = nptr[q]; e < nptr[q + 1]; ++e) {
                    const i32 r = nidx[e];
                    if (r <= q) continue;
                    p.fe_row[k] = local_row(t, r);
                    p.fe_col[k] = q - p.first[t];
                    p.fe_p[k] = q;
                    p.fe_q[k] = r;
                    p.fe_src[k] = nsrc[e];
                    ++k;
                }
        }
    });
    progress("plan: update rows and entries");
    // Levels by height; within a level smaller fronts first (the kernels give
    // small fronts a wavefront each and large ones a workgroup or more).
    i32 H = 0;
    for (i32 t = 0; t < nf; ++t) H = std::max(H, p.height[t] + 1);
    p.lev_ptr.assign(H + 1, 0);
    for (i32 t = 0; t < nf; ++t) ++p.lev_ptr[p.height[t] + 1];
    for (i32 h = 0; h < H; ++h) p.lev_ptr[h + 1] += p.lev_ptr[h];
    p.order.assign(nf, 0);
    {
        std::vector<i32> fill(p.lev_ptr.begin(), p.lev_ptr.end() - 1);
        for (i32 t = 0; t < nf; ++t) p.order[fill[p.height[t]]++] = t;
    }
    for (i32 h = 0; h < H; ++h)
        std::stable_sort(p.order.begin() + p.lev_ptr[h], p.order.begin() + p.lev_ptr[h + 1], [&](i32 a, i32 b) {
            return p.npiv[a] + p.nupd[a] < p.npiv[b] + p.nupd[b];
        });
    // L level by level, every block on 4 entries (16-byte loads).
    p.lpos.assign(nf, 0), p.upos.assign(nf, 0), p.vpos.assign(nf, 0);
    i64 l = 0, u = 0, v = 0;
    for (i32 k = 0; k < nf; ++k) {
        const i32 t = p.order[k];
        const i64 np = p.npiv[t], m = np + p.nupd[t], nu = p.nupd[t];
        p.lpos[t] = l;
        l += np * m - np * (np - 1) / 2;
        l = (l + 3) & ~i64(3);
        p.upos[t] = u;
        u += nu * nu;
        p.vpos[t] = v;
        v += nu;
    }
    p.lsize = l, p.usize = u, p.vsize = v;
}

}  // namespace

Plan make_plan(i64 bx, i64 by, const std::uint8_t* mark, bool four, int leaf) {
    if (bx * by >= (i64(1) << 31)) throw std::invalid_argument("box too large for a 32-bit plan");
    Plan p;
    p.bx = bx, p.by = by, p.four = four, p.nd = four ? 4 : 8;
    Dissection d{mark, bx, leaf, {}, {}, {}, {}, nullptr, 0};
    d.cut(0, by, 0, bx);
    p.cell = std::move(d.cells);
    p.n = static_cast<i32>(p.cell.size());
    p.nfront = static_cast<i32>(d.first.size());
    p.first = std::move(d.first);
    p.npiv = std::move(d.npiv);
    const i32 n = p.n, nd = p.nd;
    p.unk_of.assign(bx * by, -1);
    for (i32 k = 0; k < n; ++k) p.unk_of[p.cell[k]] = k;
    const Offset* nb = four ? kNb4 : kNb8;
    p.nbr.assign(i64(n) * nd, -1);
    std::vector<i64> nptr(n + 1, 0), nsrc;
    std::vector<i32> nidx;
    for (i32 k = 0; k < n; ++k) {
        const i64 y = p.cell[k] / bx, x = p.cell[k] % bx;
        for (int e = 0; e < nd; ++e) {
            const i64 yy = y + nb[e].dy, xx = x + nb[e].dx;
            if (yy < 0 || yy >= by || xx < 0 || xx >= bx) continue;
            const i32 q = p.unk_of[yy * bx + xx];
            p.nbr[i64(k) * nd + e] = q;
            if (q < 0) continue;
            nidx.push_back(q);
            nsrc.push_back(i64(k) * nd + e);
        }
        nptr[k + 1] = static_cast<i64>(nidx.size());
    }
    finish_plan(p, d.kids, nptr.data(), nidx.data(), nsrc.data());
    return p;
}

namespace {

// The rest of a graph plan once the elimination order (perm) and the fronts
// are set: each unknown's neighbours (its node's adjacency renumbered, in
// adjacency order), over host threads.
void finish_graph(Plan& p, i32 n, i64 bx, i64 by, const std::vector<std::vector<i32>>& kids,
                  const std::vector<i64>& ptr, const i32* adj) {
    p.n = n;
    p.nfront = static_cast<i32>(p.first.size());
    p.iperm.resize(n);
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 k = lo; k < hi; ++k) p.iperm[p.perm[k]] = static_cast<i32>(k);
    });
    p.unk_of.assign(bx * by, -1);
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 k = lo; k < hi; ++k)
            if (p.cell[k] >= 0) p.unk_of[p.cell[k]] = static_cast<i32>(k);
    });
    std::vector<i64> nptr(static_cast<std::size_t>(n) + 1, 0);
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 k = lo; k < hi; ++k) {
            const i32 c = p.perm[k];
            i64 m = 0;
            for (i64 e = ptr[c]; e < ptr[c + 1]; ++e) m += adj[e] >= 0;
            nptr[k + 1] = m;
        }
    });
    host_prefix(n, nptr.data() + 1, nptr.data());
    RawVec<i32> nidx(static_cast<std::size_t>(nptr[n]));
    RawVec<i64> nsrc(static_cast<std::size_t>(nptr[n]));
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 k = lo; k < hi; ++k) {
            const i32 c = p.perm[k];
            i64 m = nptr[k];
            for (i64 e = ptr[c]; e < ptr[c + 1]; ++e) {
                if (adj[e] < 0) continue;
                nidx[m] = p.iperm[adj[e]];
                nsrc[m++] = e;
            }
        }
    });
    progress("plan: neighbour lists of %lld entries", static_cast<long long>(nptr[n]));
    finish_plan(p, kids, nptr.data(), nidx.data(), nsrc.data());
}

}  // namespace

Plan graph_plan(i64 bx, i64 by, const std::vector<i64>& cell, const std::vector<i64>& ptr, const i32* adj,
                int leaf) {
    if (bx * by >= (i64(1) << 31)) throw std::invalid_argument("box too large for a 32-bit plan");
    const i32 n = static_cast<i32>(cell.size());
    Plan p;
    p.bx = bx, p.by = by, p.nd = 0;
    if (std::none_of(cell.begin(), cell.end(), [](i64 c) { return c >= 0; })) {
        GraphDissection d(n, ptr, adj, leaf);
        std::vector<i32> all(n), roots;
        for (i32 c = 0; c < n; ++c) all[c] = c;
        d.dissect(std::move(all), d.out, roots);
        p.perm = std::move(d.out.order);
        p.first = std::move(d.out.first);
        p.npiv = std::move(d.out.npiv);
        p.cell.assign(n, -1);
        std::vector<std::vector<i32>> kids = std::move(d.out.kids);
        // several components: the last root (the last front) adopts the
        // others, which share no entry with it, so the tree has one root
        if (roots.empty() || roots.back() != static_cast<i32>(p.first.size()) - 1)
            throw std::logic_error("graph dissection: the last root is not the last front");
        for (std::size_t r = 0; r + 1 < roots.size(); ++r) kids[roots.back()].push_back(roots[r]);
        finish_graph(p, n, bx, by, kids, ptr, adj);
        return p;
    }
    if (std::any_of(cell.begin(), cell.end(), [&](i64 c) { return c >= bx * by; }))
        throw std::invalid_argument("graph node outside its box");
    RawVec<std::uint8_t> mark(static_cast<std::size_t>(bx * by));
    RawVec<i32> node_of(static_cast<std::size_t>(bx * by));
    host_chunks(bx * by, [&](i64, i64 lo, i64 hi) {
        std::fill(mark.begin() + lo, mark.begin() + hi, 0);
        std::fill(node_of.begin() + lo, node_of.begin() + hi, -1);
    });
    host_chunks(n, [&](i64, i64 lo, i64 hi) {  // cells are distinct: one writer per slot
        for (i64 c = lo; c < hi; ++c)
            if (cell[c] >= 0) {
                mark[cell[c]] = 1;
                node_of[cell[c]] = static_cast<i32>(c);
            }
    });
    progress("plan: %d cell nodes marked", n);
    std::atomic<int> helpers{0};
    Dissection d{mark.data(), bx, leaf, {}, {}, {}, {}, &helpers, host_threads() - 1};
    const i32 root = d.cut(0, by, 0, bx);
    progress("plan: dissection of %zu fronts", d.first.size());
    p.cell = std::move(d.cells);
    p.first = std::move(d.first);
    p.npiv = std::move(d.npiv);
    p.perm.reserve(n);
    for (i32 b : p.cell) p.perm.push_back(node_of[b]);
    const i32 first_merged = static_cast<i32>(p.perm.size());
    for (i32 c = 0; c < n; ++c)
        if (cell[c] < 0) {
            p.perm.push_back(c);
            p.cell.push_back(-1);
        }
    std::vector<std::vector<i32>> kids = std::move(d.kids);
    if (static_cast<i32>(p.perm.size()) > first_merged) {
        p.first.push_back(first_merged);
        p.npiv.push_back(static_cast<i32>(p.perm.size()) - first_merged);
        kids.push_back(root >= 0 ? std::vector<i32>{root} : std::vector<i32>{});
    }
    finish_graph(p, n, bx, by, kids, ptr, adj);
    return p;
}


Plan window_plan(int D, bool disk, bool four, int leaf) {
    const i64 W = 2 * i64(D) + 1;
    std::vector<std::uint8_t> mark(W * W, 0);
    for (i64 y = 0; y < W; ++y)
        for (i64 x = 0; x < W; ++x) {
            const i64 dy = y - D, dx = x - D;
            mark[y * W + x] = !disk || dy * dy + dx * dx <= i64(D) * D;
        }
    mark[D * W + D] = 0;
    return make_plan(W, W, mark.data(), four, leaf);
}

}  // namespace lsc
