#include "plan.hpp"

#include <algorithm>
#include <stdexcept>

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
    for (i32 k = 0; k < n; ++k) {
        const i64 y = p.cell[k] / bx, x = p.cell[k] % bx;
        for (int e = 0; e < nd; ++e) {
            const i64 yy = y + nb[e].dy, xx = x + nb[e].dx;
            if (yy < 0 || yy >= by || xx < 0 || xx >= bx) continue;
            p.nbr[i64(k) * nd + e] = p.unk_of[yy * bx + xx];
        }
    }
    // Fronts: update rows (pivot neighbours past the front, and the children's
    // update rows past it), sorted; child rows mapped into the parent.
    const i32 nf = p.nfront;
    p.nupd.assign(nf, 0), p.upd.assign(nf, 0), p.cb.assign(nf, 0), p.ce.assign(nf, 0), p.height.assign(nf, 0);
    p.amap.assign(p.nbr.size(), -1);
    std::vector<i32> stamp(n, -1), loc(n, -1), upd;
    for (i32 t = 0; t < nf; ++t) {
        p.cb[t] = static_cast<i32>(p.kids.size());
        i32 h = 0;
        for (i32 c : d.kids[t]) {
            p.kids.push_back(c);
            h = std::max(h, p.height[c] + 1);
        }
        p.height[t] = h;
        p.ce[t] = static_cast<i32>(p.kids.size());
        const i32 end = p.first[t] + p.npiv[t];
        upd.clear();
        for (i32 q = p.first[t]; q < end; ++q)
            for (int e = 0; e < nd; ++e) {
                const i32 r = p.nbr[i64(q) * nd + e];
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
        p.rows.insert(p.rows.end(), upd.begin(), upd.end());
        p.relmap.resize(p.rows.size(), -1);
        for (i32 q = p.first[t]; q < end; ++q) loc[q] = q - p.first[t];
        for (i32 i = 0; i < p.nupd[t]; ++i) loc[upd[i]] = p.npiv[t] + i;
        for (i32 ci = p.cb[t]; ci < p.ce[t]; ++ci) {
            const i32 c = p.kids[ci];
            for (i32 i = 0; i < p.nupd[c]; ++i) p.relmap[p.upd[c] + i] = loc[p.rows[p.upd[c] + i]];
        }
        for (i32 q = p.first[t]; q < end; ++q)
            for (int e = 0; e < nd; ++e) {
                const i32 r = p.nbr[i64(q) * nd + e];
                if (r > q) p.amap[i64(q) * nd + e] = loc[r];
            }
        const i32 m = p.npiv[t] + p.nupd[t];
        p.max_front = std::max(p.max_front, m);
        const double np = p.npiv[t], nu = p.nupd[t];
        p.flops += np * np * np / 3 + np * np * nu + np * nu * nu;
    }
    // Levels by height; L laid out level by level (packed columns).
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
    // Within a level, smaller fronts first (the kernels give small fronts a
    // wavefront each and large ones the whole workgroup).
    for (i32 h = 0; h < H; ++h)
        std::stable_sort(p.order.begin() + p.lev_ptr[h], p.order.begin() + p.lev_ptr[h + 1], [&](i32 a, i32 b) {
            return p.npiv[a] + p.nupd[a] < p.npiv[b] + p.nupd[b];
        });
    p.lpos.assign(nf, 0), p.upos.assign(nf, 0), p.vpos.assign(nf, 0);
    i64 l = 0, u = 0, v = 0;
    for (i32 k = 0; k < nf; ++k) {
        const i32 t = p.order[k];
        const i64 np = p.npiv[t], m = np + p.nupd[t], nu = p.nupd[t];
        p.lpos[t] = l;
        l += np * m - np * (np - 1) / 2;
        l = (l + 3) & ~i64(3);  // every block starts on 4 entries (16-byte loads)
        p.upos[t] = u;
        u += nu * nu;
        p.vpos[t] = v;
        v += nu;
    }
    p.lsize = l, p.usize = u, p.vsize = v;
    // The off-diagonal entries of every front, front by front.
    p.fe_ptr.assign(nf + 1, 0);
    for (i32 t = 0; t < nf; ++t) {
        for (i32 q = p.first[t]; q < p.first[t] + p.npiv[t]; ++q)
            for (int e = 0; e < nd; ++e) {
                const i32 a = p.amap[i64(q) * nd + e];
                if (a < 0) continue;
                p.fe_row.push_back(a);
                p.fe_col.push_back(q - p.first[t]);
                p.fe_p.push_back(q);
                p.fe_d.push_back(e);
            }
        p.fe_ptr[t + 1] = static_cast<i32>(p.fe_row.size());
    }
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
