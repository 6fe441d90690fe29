#include "plan.hpp"

#include <algorithm>
#include <atomic>
#include <exception>
#include <memory>
#include <stdexcept>
#include <thread>

namespace lsc {

namespace {

// Geometric nested dissection (reference: src/cpp/circuit/system.cpp:41-93).
struct Dissection {
    const std::uint8_t* mark;
    i64 stride;
    i64 leaf;
    std::vector<i32> cells;
    std::vector<i32> first, npiv;
    std::vector<std::vector<i32>> kids;

    i32 add(i32 f, std::vector<i32> ch) {
        first.push_back(f);
        npiv.push_back(static_cast<i32>(cells.size()) - f);
        kids.push_back(std::move(ch));
        return static_cast<i32>(first.size()) - 1;
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
        const i32 a = by_row ? cut(r0, mid, c0, c1) : cut(r0, r1, c0, mid);
        const i32 b = by_row ? cut(mid + 1, r1, c0, c1) : cut(r0, r1, mid + 1, c1);
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
    const std::vector<i32>& adj;
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

    GraphDissection(i32 n, const std::vector<i64>& p, const std::vector<i32>& a, int lf)
        : ptr(p), adj(a), leaf(static_cast<std::size_t>(lf)), stamp(new std::atomic<i32>[n > 0 ? n : 1]),
          level(n, -1),
          max_helpers(static_cast<int>(std::min<unsigned>(15, std::max(1u, std::thread::hardware_concurrency()) - 1))) {
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
void finish_plan(Plan& p, const std::vector<std::vector<i32>>& kids, const std::vector<i64>& nptr,
                 const std::vector<i32>& nidx, const std::vector<i64>& nsrc) {
    const i32 n = p.n, nf = p.nfront;
    p.nupd.assign(nf, 0), p.upd.assign(nf, 0), p.cb.assign(nf, 0), p.ce.assign(nf, 0), p.height.assign(nf, 0);
    std::vector<i32> stamp(n, -1), loc(n, -1), upd;
    p.fe_ptr.assign(nf + 1, 0);
    // every entry (q, r) with r > q is assembled into q's front: at most half
    // the neighbour lists
    const std::size_t nent = static_cast<std::size_t>(nptr[n] / 2 + 1);
    p.fe_row.reserve(nent), p.fe_col.reserve(nent), p.fe_p.reserve(nent), p.fe_q.reserve(nent);
    p.fe_src.reserve(nent);
    p.kids.reserve(nf);
    for (i32 t = 0; t < nf; ++t) {
        p.cb[t] = static_cast<i32>(p.kids.size());
        i32 h = 0;
        for (i32 c : kids[t]) {
            p.kids.push_back(c);
            h = std::max(h, p.height[c] + 1);
        }
        p.height[t] = h;
        p.ce[t] = static_cast<i32>(p.kids.size());
        const i32 end = p.first[t] + p.npiv[t];
        upd.clear();
        for (i32 q = p.first[t]; q < end; ++q)
            for (i64 e = nptr[q]; e < nptr[q + 1]; ++e) {
                const i32 r = nidx[e];
                if (r >= end && stamp[r] != t) stamp[r] = t, upd.push_back(r);
            }
        for (i32 ci = p.cb[t]; ci < p.ce[t]; ++ci) {
            const i32 c = p.kids[ci];
            for (i32 i = 0; i < p.nupd[c]; ++i) {
                const i32 r = p.rows[p.upd[c] + i];
                if (r >= end && stamp[r] != t) stamp[r] = t, upd.push_back(r);
            }
        }
        std::sort(upd.begin(), upd.end());
        p.upd[t] = static_cast<i32>(p.rows.size());
        p.nupd[t] = static_cast<i32>(upd.size());
        if (p.rows.capacity() < p.rows.size() + upd.size()) p.rows.reserve(2 * (p.rows.size() + upd.size()));
        p.rows.insert(p.rows.end(), upd.begin(), upd.end());
        if (p.relmap.capacity() < p.rows.size()) p.relmap.reserve(2 * p.rows.size());
        p.relmap.resize(p.rows.size(), -1);
        for (i32 q = p.first[t]; q < end; ++q) loc[q] = q - p.first[t];
        for (i32 i = 0; i < p.nupd[t]; ++i) loc[upd[i]] = p.npiv[t] + i;
        for (i32 ci = p.cb[t]; ci < p.ce[t]; ++ci) {
            const i32 c = p.kids[ci];
            for (i32 i = 0; i < p.nupd[c]; ++i) p.relmap[p.upd[c] + i] = loc[p.rows[p.upd[c] + i]];
        }
        // entries (q, r), r > q, of this front's pivots
        for (i32 q = p.first[t]; q < end; ++q)
            for (i64 e = nptr[q]; e < nptr[q + 1]; ++e) {
                const i32 r = nidx[e];
                if (r <= q) continue;
                p.fe_row.push_back(loc[r]);
                p.fe_col.push_back(q - p.first[t]);
                p.fe_p.push_back(q);
                p.fe_q.push_back(r);
                p.fe_src.push_back(nsrc[e]);
            }
        p.fe_ptr[t + 1] = static_cast<i32>(p.fe_row.size());
        const i32 m = p.npiv[t] + p.nupd[t];
        p.max_front = std::max(p.max_front, m);
        const double np = p.npiv[t], nu = p.nupd[t];
        p.flops += np * np * np / 3 + np * np * nu + np * nu * nu;
    }
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
    Dissection d{mark, bx, leaf, {}, {}, {}, {}};
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
    finish_plan(p, d.kids, nptr, nidx, nsrc);
    return p;
}

namespace {

// The rest of a graph plan once the elimination order (perm) and the fronts are set.
void finish_graph(Plan& p, i32 n, i64 bx, i64 by, const std::vector<std::vector<i32>>& kids,
                  const std::vector<i64>& ptr, const std::vector<i32>& adj) {
    p.n = n;
    p.nfront = static_cast<i32>(p.first.size());
    p.iperm.assign(n, -1);
    for (i32 k = 0; k < n; ++k) p.iperm[p.perm[k]] = k;
    p.unk_of.assign(bx * by, -1);
    for (i32 k = 0; k < n; ++k)
        if (p.cell[k] >= 0) p.unk_of[p.cell[k]] = k;
    std::vector<i64> nptr(n + 1, 0), nsrc;
    std::vector<i32> nidx;
    nidx.reserve(static_cast<std::size_t>(ptr[n])), nsrc.reserve(static_cast<std::size_t>(ptr[n]));
    for (i32 k = 0; k < n; ++k) {
        const i32 c = p.perm[k];
        for (i64 e = ptr[c]; e < ptr[c + 1]; ++e) {
            if (adj[e] < 0) continue;
            nidx.push_back(p.iperm[adj[e]]);
            nsrc.push_back(e);
        }
        nptr[k + 1] = static_cast<i64>(nidx.size());
    }
    finish_plan(p, kids, nptr, nidx, nsrc);
}

}  // namespace

Plan graph_plan(i64 bx, i64 by, const std::vector<i64>& cell, const std::vector<i64>& ptr,
                const std::vector<i32>& adj, int leaf) {
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
    std::vector<std::uint8_t> mark(bx * by, 0);
    std::vector<i32> node_of(bx * by, -1);
    for (i32 c = 0; c < n; ++c)
        if (cell[c] >= 0) {
            if (cell[c] >= bx * by) throw std::invalid_argument("graph node outside its box");
            mark[cell[c]] = 1;
            node_of[cell[c]] = c;
        }
    Dissection d{mark.data(), bx, leaf, {}, {}, {}, {}};
    const i32 root = d.cut(0, by, 0, bx);
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
